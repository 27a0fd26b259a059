// Serving loop of the product: Poisson solves of different sizes arrive one
// at a time, are set up on host worker threads, and are admitted to the GPU
// in groups as device memory frees up. A group is one batched Solver (every
// stage one launch for all its problems) on its own non-blocking stream;
// several groups run side by side, so requests that arrive while others are
// on the GPU overlap with them. Each problem's state and Krylov vectors stay
// resident until it has converged; its result is copied out as soon as it
// stops, while the rest of its group continues.
//
// Admission: a problem's device footprint is estimated from its host setup
// (ProblemHost::device_bytes); a group is formed from the queued problems, in
// arrival order, while the sum over all groups in flight stays within the HBM
// budget. A problem larger than the whole budget is rejected.
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
    int groups = 2;          // groups solved side by side on the device
    int max_batch = 16;      // problems per group
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
    int batch_size = 0;
    double t_submit = 0, t_ready = 0, t_admit = 0, t_done = 0;
    bool done = false, failed = false;
    std::string error;
};

class Server {
public:
    explicit Server(const ServerParams& p);
    ~Server();  // finishes the submitted work, then joins the workers
    int64_t submit(const SolverParams& P, int n_pages, const uint64_t* offsets, const uint32_t* flags, const float* b, double tol,
                   int max_it);
    std::shared_ptr<ServerRequest> wait(int64_t id);  // blocks until the request is done (or failed)
    void forget(int64_t id);                          // drops a finished request's record
    double now() const;
    double budget() const { return budget_; }

private:
    void setup_worker();
    void group_worker();
    ServerParams prm;
    double budget_ = 0, in_use_ = 0;
    int setting_up_ = 0;
    std::mutex mu;
    std::condition_variable cv_setup, cv_ready, cv_done;
    std::deque<std::shared_ptr<ServerRequest>> submitted, ready;
    std::map<int64_t, std::shared_ptr<ServerRequest>> all;
    int64_t next_id = 1;
    bool stopping = false;
    std::vector<std::thread> threads;
    std::chrono::steady_clock::time_point t0;
};

}  // namespace hdd
