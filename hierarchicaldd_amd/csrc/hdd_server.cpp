#include "hdd_server.h"
#include <omp.h>
#include <algorithm>
#include <stdexcept>

namespace hdd {

Server::Server(const ServerParams& p) : prm(p), t0(std::chrono::steady_clock::now()) {
    if (prm.setup_threads < 1 || prm.groups < 1 || prm.max_batch < 1) throw std::runtime_error("server: thread, group and batch counts must be >= 1");
    budget_ = prm.hbm_budget;
    if (budget_ <= 0) {
        size_t free_b = 0, total_b = 0;
        hdd_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        budget_ = 0.9 * (double)free_b;
    }
    // the set-up workers share the host's threads: each runs its OpenMP work on its share
    const int per_worker = std::max(1, omp_get_max_threads() / prm.setup_threads);
    for (int i = 0; i < prm.setup_threads; ++i)
        threads.emplace_back([this, per_worker] {
            omp_set_num_threads(per_worker);
            setup_worker();
        });
    threads.emplace_back([this] { builder(); });
    threads.emplace_back([this] { driver(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu);
        stopping = true;
    }
    cv_setup.notify_all();
    cv_ready.notify_all();
    cv_driver.notify_all();
    for (auto& t : threads) t.join();
}

double Server::now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

int64_t Server::submit(const SolverParams& P, int n_pages, const uint64_t* offsets, const uint32_t* flags, const float* b, double tol,
                       int max_it) {
    auto r = std::make_shared<ServerRequest>();
    r->P = P;
    r->offsets.assign(offsets, offsets + n_pages);
    r->flags.assign(flags, flags + (size_t)n_pages * kPageCells);
    r->b.assign(b, b + (size_t)n_pages * kPageCells);
    r->tol = tol;
    r->max_it = max_it;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (stopping) throw std::runtime_error("server is shutting down");
        r->id = next_id++;
        r->t_submit = now();
        all[r->id] = r;
        submitted.push_back(r);
    }
    cv_setup.notify_one();
    return r->id;
}

std::shared_ptr<ServerRequest> Server::wait(int64_t id) {
    std::unique_lock<std::mutex> lk(mu);
    auto it = all.find(id);
    if (it == all.end()) throw std::runtime_error("no such request");
    std::shared_ptr<ServerRequest> r = it->second;
    cv_done.wait(lk, [&] { return r->done; });
    return r;
}

bool Server::cancel(int64_t id) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = all.find(id);
    if (it == all.end() || it->second->done) return false;
    it->second->cancel = true;
    return true;
}

void Server::forget(int64_t id) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = all.find(id);
    if (it != all.end() && it->second->done) all.erase(it);
}

void Server::finish_cancelled(const std::shared_ptr<ServerRequest>& r) {
    r->cancelled = true;
    r->done = true;
    r->t_done = now();
    r->host.reset();
}

bool Server::drained_for_builder() const { return stopping && ready.empty() && submitted.empty() && setting_up_ == 0; }

void Server::setup_worker() {
    for (;;) {
        std::shared_ptr<ServerRequest> r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_setup.wait(lk, [&] { return stopping || !submitted.empty(); });
            if (submitted.empty()) return;
            r = submitted.front();
            submitted.pop_front();
            if (r->cancel) {
                finish_cancelled(r);
                lk.unlock();
                cv_done.notify_all();
                cv_ready.notify_all();
                continue;
            }
            ++setting_up_;
        }
        std::string err;
        try {
            auto h = std::make_shared<ProblemHost>(r->P, (int)r->offsets.size(), r->offsets.data(), r->flags.data(), r->b.data());
            h->factorize();
            r->bytes = h->device_bytes();
            r->host = h;
            std::vector<uint32_t>().swap(r->flags);
            std::vector<float>().swap(r->b);
        } catch (std::exception& e) {
            err = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            --setting_up_;
            r->t_ready = now();
            if (err.empty() && (double)r->bytes > budget_) err = "problem needs more device memory than the server's HBM budget";
            if (!err.empty()) {
                r->failed = true; r->error = err; r->done = true; r->t_done = r->t_ready;
                r->host.reset();
            } else {
                ready.push_back(r);
            }
        }
        cv_done.notify_all();
        cv_ready.notify_all();
    }
}

static bool same_schedule(const SolverParams& a, const SolverParams& b) {
    return a.vlevels == b.vlevels && a.sigma_levels == b.sigma_levels && a.nV == b.nV && a.int_s == b.int_s && a.bnd_s == b.bnd_s &&
           a.sig_s == b.sig_s && a.sig_mg == b.sig_mg;
}

// Forms a unit from the ready requests and builds its device state; the driver
// picks it up at its next round.
void Server::builder() {
    for (;;) {
        std::unique_ptr<Unit> u(new Unit());
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                // requests cancelled while they waited never reach the device
                for (auto it = ready.begin(); it != ready.end();) {
                    if ((*it)->cancel) { finish_cancelled(*it); it = ready.erase(it); cv_done.notify_all(); }
                    else ++it;
                }
                if (!ready.empty() && units_ < prm.groups && in_use_ + (double)ready.front()->bytes <= budget_) break;
                if (ready.empty() && drained_for_builder()) {
                    builder_done_ = true;
                    lk.unlock();
                    cv_driver.notify_all();
                    return;
                }
                cv_ready.wait(lk);
            }
            // arrival order; problems with another V-cycle/sigma schedule wait for a later unit
            const SolverParams lead = ready.front()->P;
            for (auto it = ready.begin(); it != ready.end() && (int)u->reqs.size() < prm.max_batch;) {
                const double b = (double)(*it)->bytes;
                if (!same_schedule((*it)->P, lead)) { ++it; continue; }
                if (in_use_ + u->bytes + b > budget_) break;
                u->bytes += b;
                u->reqs.push_back(*it);
                it = ready.erase(it);
            }
            in_use_ += u->bytes;
            ++units_;
            for (auto& r : u->reqs) r->batch_size = (int)u->reqs.size();
        }
        try {
            std::vector<std::shared_ptr<ProblemHost>> hosts;
            for (auto& r : u->reqs) hosts.push_back(r->host);
            u->S.reset(new Solver(hosts));
        } catch (std::exception& e) {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& r : u->reqs) { r->failed = true; r->error = e.what(); r->done = true; r->t_done = now(); r->host.reset(); }
            in_use_ -= u->bytes;
            --units_;
            cv_done.notify_all();
            continue;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            joining.push_back(std::move(u));
        }
        cv_driver.notify_all();
    }
}

// Advances every resident unit by one PCG iteration per round.
void Server::driver() {
    std::vector<std::unique_ptr<Unit>> active;
    for (;;) {
        std::vector<std::unique_ptr<Unit>> joined;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_driver.wait(lk, [&] { return !joining.empty() || !active.empty() || (builder_done_ && joining.empty()); });
            while (!joining.empty()) { joined.push_back(std::move(joining.front())); joining.pop_front(); }
            if (joined.empty() && active.empty() && builder_done_) return;
        }
        for (auto& u : joined) {
            std::vector<double> tol;
            std::vector<int> mit;
            for (auto& r : u->reqs) { tol.push_back(r->tol); mit.push_back(r->max_it); }
            bool ok = true;
            try {
                u->S->pcg_start(u->run, tol, mit);
            } catch (std::exception& e) {
                ok = false;
                std::lock_guard<std::mutex> lk(mu);
                for (auto& r : u->reqs) { r->failed = true; r->error = e.what(); r->done = true; r->t_done = now(); r->host.reset(); }
            }
            if (!ok) {
                u->S.reset();
                std::lock_guard<std::mutex> lk(mu);
                in_use_ -= u->bytes;
                --units_;
                cv_done.notify_all();
                cv_ready.notify_all();
                continue;
            }
            active.push_back(std::move(u));
            std::lock_guard<std::mutex> lk(mu);
            const double t = now();
            for (auto& r : active.back()->reqs) { r->t_admit = t; r->concurrent = (int)active.size(); }
        }
        for (auto& up : active) {
            Unit& u = *up;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (size_t p = 0; p < u.reqs.size(); ++p) u.run.cancel[p] = u.reqs[p]->cancel ? 1 : 0;
            }
            try {
                u.S->pcg_advance(u.run, [&](int p) {
                    ServerRequest& r = *u.reqs[p];
                    const ProblemHost& H = *r.host;
                    std::vector<float> xp((size_t)H.lv[0].npages() * kPageCells);
                    u.S->download_x(p, xp.data());
                    std::vector<float> xin(H.input_page.size() * kPageCells, 0.f);
                    for (size_t i = 0; i < H.input_page.size(); ++i)
                        std::copy(xp.begin() + (size_t)H.input_page[i] * kPageCells, xp.begin() + (size_t)(H.input_page[i] + 1) * kPageCells,
                                  xin.begin() + i * kPageCells);
                    const PcgResult& res = u.run.out[p];
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        r.x.swap(xin);
                        r.iterations = res.iterations;
                        r.converged = res.converged;
                        r.cancelled = res.cancelled;
                        r.final_norm = res.hist.empty() ? 0.0 : res.hist.back();
                        r.t_done = now();
                        r.done = true;
                    }
                    cv_done.notify_all();
                });
            } catch (std::exception& e) {
                std::lock_guard<std::mutex> lk(mu);
                for (auto& r : u.reqs)
                    if (!r->done) { r->failed = true; r->error = e.what(); r->done = true; r->t_done = now(); }
                u.run.n_active = 0;
                cv_done.notify_all();
            }
        }
        // retire the units whose problems have all stopped: their device memory and slot go to the next unit
        for (auto& up : active) {
            if (up->run.n_active > 0) continue;
            up->run.skip.reset();
            up->S.reset();
            std::lock_guard<std::mutex> lk(mu);
            in_use_ -= up->bytes;
            --units_;
            for (auto& r : up->reqs) r->host.reset();
            up.reset();
        }
        active.erase(std::remove(active.begin(), active.end(), nullptr), active.end());
        cv_ready.notify_all();
    }
}

}  // namespace hdd
