// Serving loop of the product: Poisson solves of different sizes arrive one
// at a time and are solved with continuous batching.
//
//   setup workers  host setup and factorisation of each request (ProblemHost)
//   builder        forms a unit from the requests that are ready (arrival
//                  order, same V-cycle/sigma schedule, up to max_batch, within
//                  the HBM budget) and builds its device state: one batched
//                  Solver, every stage one launch for all its problems, on its
//                  own non-blocking stream
//   driver         keeps every resident unit iterating: one PCG iteration per
//                  unit per round (Solver::pcg_advance), each on its stream, so
//                  the units' kernels overlap on the device. A unit built while
//                  others are mid-solve joins at the next round; a problem that
//                  stops has its x copied out and its result returned at once;
//                  a unit whose problems have all stopped is destroyed, which
//                  frees its device memory and its slot for the next unit.
//
// Admission: a problem's device footprint is estimated from its host setup
// (ProblemHost::device_bytes); a unit is admitted while the sum over resident
// units stays within the HBM budget and there are fewer than `groups` units.
// A problem larger than the whole budget is rejected.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "hdd_runtime.h"

namespace hdd {

struct ServerParams {
    int setup_threads = 2;   // host setup workers (each setup also uses OpenMP)
    int groups = 4;          // units resident and iterating on the device at once
    int max_batch = 16;      // problems per unit
    double hbm_budget = 0;   // device bytes for resident problems; 0: 90% of the free memory at start
};

struct ServerRequest {
    int64_t id = 0;
    SolverParams P;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> flags;
    std::vector<float> b;
    double tol = 0;
    int max_it = 0;
    std::shared_ptr<ProblemHost> host;
    size_t bytes = 0;
    // result
    std::vector<float> x;    // input page order
    int iterations = 0;
    double final_norm = 0;
    int batch_size = 0;      // problems in its unit
    int concurrent = 0;      // units iterating on the device when its unit started (its own included)
    double t_submit = 0, t_ready = 0, t_admit = 0, t_done = 0;
    bool done = false, failed = false, converged = false;
    bool cancel = false;     // requested by Server::cancel
    bool cancelled = false;  // it stopped because of that request
    std::string error;
};

class Server {
public:
    explicit Server(const ServerParams& p);
    ~Server();  // finishes the submitted work, then joins the workers
    int64_t submit(const SolverParams& P, int n_pages, const uint64_t* offsets, const uint32_t* flags, const float* b, double tol,
                   int max_it);
    std::shared_ptr<ServerRequest> wait(int64_t id);  // blocks until the request is done (or failed)
    // asks a request to stop: a queued one is dropped before admission, a running one stops at
    // its unit's next residual test; false if the request is unknown or already done
    bool cancel(int64_t id);
    void forget(int64_t id);                          // drops a finished request's record
    double now() const;
    double budget() const { return budget_; }

private:
    struct Unit {
        std::vector<std::shared_ptr<ServerRequest>> reqs;
        std::unique_ptr<Solver> S;
        PcgRun run;
        double bytes = 0;
    };
    void setup_worker();
    void builder();
    void driver();
    void finish_cancelled(const std::shared_ptr<ServerRequest>& r);  // mu held
    bool drained_for_builder() const;                                // mu held
    ServerParams prm;
    double budget_ = 0, in_use_ = 0;
    int setting_up_ = 0, units_ = 0;  // units being built or resident
    bool builder_done_ = false;
    std::mutex mu;
    std::condition_variable cv_setup, cv_ready, cv_done, cv_driver;
    std::deque<std::shared_ptr<ServerRequest>> submitted, ready;
    std::deque<std::unique_ptr<Unit>> joining;
    std::map<int64_t, std::shared_ptr<ServerRequest>> all;
    int64_t next_id = 1;
    bool stopping = false;
    std::vector<std::thread> threads;
    std::chrono::steady_clock::time_point t0;
};

}  // namespace hdd
