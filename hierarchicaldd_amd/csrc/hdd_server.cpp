#include "hdd_server.h"
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
    for (int i = 0; i < prm.setup_threads; ++i) threads.emplace_back([this] { setup_worker(); });
    for (int i = 0; i < prm.groups; ++i) threads.emplace_back([this] { group_worker(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu);
        stopping = true;
    }
    cv_setup.notify_all();
    cv_ready.notify_all();
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

void Server::forget(int64_t id) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = all.find(id);
    if (it != all.end() && it->second->done) all.erase(it);
}

void Server::setup_worker() {
    for (;;) {
        std::shared_ptr<ServerRequest> r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_setup.wait(lk, [&] { return stopping || !submitted.empty(); });
            if (submitted.empty()) return;
            r = submitted.front();
            submitted.pop_front();
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

void Server::group_worker() {
    for (;;) {
        std::vector<std::shared_ptr<ServerRequest>> grp;
        double bytes = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_ready.wait(lk, [&] {
                if (!ready.empty() && in_use_ + (double)ready.front()->bytes <= budget_) return true;
                return stopping && ready.empty() && submitted.empty() && setting_up_ == 0;
            });
            if (ready.empty()) return;
            // arrival order; problems with another V-cycle/sigma schedule wait for a later group
            const SolverParams lead = ready.front()->P;
            for (auto it = ready.begin(); it != ready.end() && (int)grp.size() < prm.max_batch;) {
                const double b = (double)(*it)->bytes;
                if (!same_schedule((*it)->P, lead)) { ++it; continue; }
                if (in_use_ + bytes + b > budget_) break;
                bytes += b;
                grp.push_back(*it);
                it = ready.erase(it);
            }
            in_use_ += bytes;
            const double t = now();
            for (auto& r : grp) { r->t_admit = t; r->batch_size = (int)grp.size(); }
        }
        try {
            std::vector<std::shared_ptr<ProblemHost>> hosts;
            std::vector<double> tol;
            std::vector<int> mit;
            for (auto& r : grp) { hosts.push_back(r->host); tol.push_back(r->tol); mit.push_back(r->max_it); }
            Solver S(hosts);
            std::vector<PcgResult> res;
            S.pcg_batch(tol, mit, res, [&](int p) {
                ServerRequest& r = *grp[p];
                const ProblemHost& H = *hosts[p];
                std::vector<float> xp((size_t)H.lv[0].npages() * kPageCells);
                S.download_x(p, xp.data());
                r.x.assign(H.input_page.size() * kPageCells, 0.f);
                for (size_t i = 0; i < H.input_page.size(); ++i)
                    std::copy(xp.begin() + (size_t)H.input_page[i] * kPageCells, xp.begin() + (size_t)(H.input_page[i] + 1) * kPageCells,
                              r.x.begin() + i * kPageCells);
                {
                    std::lock_guard<std::mutex> lk(mu);
                    r.iterations = res[p].iterations;
                    r.final_norm = res[p].hist.empty() ? 0.0 : res[p].hist.back();
                    r.t_done = now();
                    r.done = true;
                }
                cv_done.notify_all();
            });
        } catch (std::exception& e) {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& r : grp)
                if (!r->done) { r->failed = true; r->error = e.what(); r->done = true; r->t_done = now(); }
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            in_use_ -= bytes;
            for (auto& r : grp) r->host.reset();
        }
        cv_done.notify_all();
        cv_ready.notify_all();
    }
}

}  // namespace hdd
