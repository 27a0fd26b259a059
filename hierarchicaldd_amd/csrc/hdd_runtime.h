// Product runtime: the host facade that mirrors the reference's
// SPGrid_Domain_Decomposition_Solver (DDS/SPGrid_Domain_Decomposition_Solver.h:24-198)
// and runs every hot-path stage on the GPU.
//
// A Solver holds one or more independent problems in one device layout, so
// that every stage of the preconditioner is one launch for all of them:
//  * V-cycle levels: the problems' pages side by side in one paged grid
//    (problem p translated by xoff[p] cells along x at level 0, xoff[p] >> l
//    at level l; offsets are multiples of 4 * 2^(vlevels-1), so every level's
//    page boundaries and the coarse map c = (f+1)/2 are preserved, and a gap
//    keeps the problems' pages apart). No stencil, restriction or
//    prolongation reads across problems: face bits are per problem, and the
//    transfer stencils of Active cells stay inside the problem's own box.
//  * Interface solver: the problems' operators stacked block-diagonally.
//  * PCG: per-problem scalars (rho, alpha, beta), dot products and L-infinity
//    norms reduced per problem, and a per-problem stopping test; a converged
//    problem's x and r are frozen while the others continue.
// One problem is the special case of a batch of one.
#pragma once
#include <chrono>
#include <functional>
#include <memory>
#include <vector>
#include "hdd_device.h"
#include "hdd_setup.h"
#include "hdd_sigma.h"

namespace hdd {

struct SolverParams {
    Int3 size, sub;
    int vlevels = 3, sigma_levels = 3, nV = 1, int_s = 1, bnd_s = 3, sig_s = 3, sig_mg = 1;
};

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() {}
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    ~DBuf() { if (p) (void)hipFree(p); }
    // The clear and the upload go through the null stream, which does not order
    // the contexts' non-blocking streams: both wait for completion before returning.
    void alloc(size_t count) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        if (count) {
            hdd_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
            hdd_check(hipMemset(p, 0, count * sizeof(T)), "hipMemset");
            hdd_check(hipStreamSynchronize(nullptr), "hipMemset");
        }
    }
    void upload(const T* h, size_t count) {
        alloc(count);
        if (count) {
            hdd_check(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice), "upload");
            hdd_check(hipStreamSynchronize(nullptr), "upload");
        }
    }
    void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
    std::vector<T> download() const {
        std::vector<T> v(n);
        if (n) hdd_check(hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "download");
        return v;
    }
};

struct SigmaFactors;  // hdd_sigma_device.h

// Host half of Read_SPGrid + Initialize_Preconditioner for one problem
// (DDS/SPGrid_Domain_Decomposition_Solver.h:73-161): page pool, subdomain
// division, flags of every V-cycle level, interface operators, and (after
// factorize()) the factors of the interface solver's exact solves.
struct ProblemHost {
    SolverParams P;
    std::vector<PagedLevel> lv;        // V-cycle levels; lv[0] is the input grid (Morton page order)
    SigmaHost sigma;
    std::vector<int> input_page;       // level-0 page of each input page
    std::vector<float> b;              // level-0 right-hand side, npages*256 in page order
    std::unique_ptr<SigmaFactors> factors;
    double setup_seconds = 0, factor_seconds = 0;

    ProblemHost(const SolverParams& p, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, const float* b);
    ~ProblemHost();
    // idempotent; with wait = false the bottom level's factor may still be in progress on
    // return (factors->wait_bottom() before reading it)
    void factorize(bool wait = true);
    // device bytes of the problem inside a batch (pages of every level, vectors, interface plans)
    size_t device_bytes() const;
};

struct LevelBuffers {
    PagedLevel host;
    DBuf<uint32_t> flags;
    DBuf<int> nbr, base, lut;
    DBuf<int> prob;                 // problem of each page (zero page: -1)
    const int* active = nullptr;    // set while a batch PCG runs: its per-problem flags
    DBuf<float> u[2], b, r;
    DBuf<__half> hu[2], hb, hr;     // FP16-storage variant of u, b, r (allocated when it is selected)
    int cur = 0;  // which u buffer holds the current iterate
    DevLevel view() const;
    size_t cells() const { return (size_t)(host.npages() + 1) * kPageCells; }
};

// Measurement: mean duration (ms) of one level-0 damped-Jacobi sweep (k_smooth, FP32) of a
// page set on its own (the given flags plus their face bits, no set-up of the solver), over
// `reps` back-to-back sweeps timed with HIP events; bytes = 16 B per page slot
double profile_sweep_pages(Int3 size, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, int reps, double* bytes);

struct PcgResult {
    int iterations = 0;                // negative: not converged within max_it
    bool converged = false;            // the stopping test passed (iterations alone cannot say so at 0)
    bool cancelled = false;            // stopped by PcgRun::cancel before its test passed
    double seconds = 0;                // from the start of the batch solve to this problem's last test
    std::vector<double> hist;          // max|r| per iteration
};

// Pinned host memory (hipHostMalloc): the PCG loop's per-iteration scalar
// copies then run asynchronously on the stream.
template <class T>
struct PinnedBuf {
    T* p = nullptr;
    size_t n = 0;
    PinnedBuf() {}
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    void alloc(size_t count) {
        if (count <= n) return;
        if (p) { (void)hipHostFree(p); p = nullptr; n = 0; }
        hdd_check(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
        n = count;
    }
};

// A PCG solve of every problem of a Solver, advanced one iteration at a time:
// Solver::pcg_start queues the set-up and the first residual test, and each
// Solver::pcg_advance waits for the test, records it, stops the problems that
// are done and queues the next iteration, so one host thread can keep several
// Solvers' iterations in flight on their own streams (the serving loop).
// pcg_batch is pcg_start followed by pcg_advance until every problem stops.
struct PcgRun {
    std::vector<double> tol;
    std::vector<int> max_it;
    std::vector<PcgResult> out;
    std::vector<int> act;              // problems still iterating
    std::vector<char> cancel;          // set by the owner between advances: stop that problem at the next test
    int it = 0, n_active = 0;
    bool fp64 = false, started = false;
    std::chrono::steady_clock::time_point t0;
    std::shared_ptr<void> skip;        // while it lives, the stencil and solve kernels skip stopped problems
    bool done() const { return started && n_active == 0; }
};

class Solver {
public:
    SolverParams P;                    // schedule shared by every problem
    std::vector<std::shared_ptr<ProblemHost>> prob;
    std::vector<int> xoff;             // level-0 x offset of each problem in the shared grid
    std::vector<std::vector<int>> pg0; // [level][problem] first page; pg0[l][nprob] = pages of level l
    std::vector<LevelBuffers> lv;      // V-cycle levels of the shared grid; lv[0] is the main grid
    DBuf<float> x, bvec, q, s;         // second-allocator channels (x, b, q, s)
    DBuf<double> dpart, dscal;
    DBuf<float> fpart, scal;           // scal: 8 per problem (rho, alpha, beta, rho_old, max|r|)
    DBuf<int> seg_pg0, seg_parts, page_prob, active;
    int max_parts = 1;
    std::unique_ptr<SigmaDevice> sigma;
    DBuf<int> if_index;                // level-0 cell index of each interface dof
    hipStream_t stream = nullptr;
    double setup_seconds = 0;
    bool host_only = false;
    bool use_graph = false;            // replay the preconditioner apply from captured HIP graphs (opt-in)
    bool half_storage = false;         // V-cycles store u, b, r in FP16 (FP32 arithmetic); z then lives in zf
    DBuf<float> zf;                    // preconditioner output of the FP16-storage variant
    // the coarsest level's 200 sweeps run in one launch, a workgroup per four
    // page slots of whole page clusters (its Active cells couple only within
    // clusters of at most four pages; HDD_BOTTOM_LOCAL=0 turns it off)
    bool bottom_local = false;
    DBuf<int> bottom_wg_page, bottom_nb_slot;

    // one problem; host_only: run the host setup and skip every device allocation (introspection)
    Solver(const SolverParams& p, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, const float* b,
           bool host_only = false);
    // a batch of problems set up on the host (factorised here if they are not yet)
    explicit Solver(const std::vector<std::shared_ptr<ProblemHost>>& problems);
    ~Solver();

    int nprob() const { return (int)prob.size(); }
    const SigmaHost& sigma_host() const { return prob[0]->sigma; }
    // level-0 channel access in page order (c: 0=z 1=r 2=k 3=x 4=b 5=q 6=s)
    float* channel(int c);
    void vcycle();
    void smooth0(int mode);
    void apply_preconditioner();  // z <- M^{-1} r
    int pcg(double tol_abs, int max_it, std::vector<double>& hist, double* seconds, bool* converged = nullptr);
    // every problem from x = 0; tol[p], max_it[p] per problem. on_stop(p), when
    // given, runs on this thread as soon as problem p has stopped (its x is final
    // and the stream is idle), while the others continue.
    void pcg_batch(const std::vector<double>& tol, const std::vector<int>& max_it, std::vector<PcgResult>& out,
                   const std::function<void(int)>& on_stop = nullptr);
    // the same loop, one iteration per call (see PcgRun); pcg_advance returns true when every problem has stopped
    void pcg_start(PcgRun& R, const std::vector<double>& tol, const std::vector<int>& max_it);
    bool pcg_advance(PcgRun& R, const std::function<void(int)>& on_stop = nullptr);
    // x of problem p (its level-0 pages in its own page order), n_pages(p)*256 floats
    void download_x(int p, float* host);
    // FP64 outer PCG (opt-in; the reference keeps rho, alpha, beta in float): the
    // scalars in double, x accumulated in double (xd; the x channel gets its
    // float copy when the problem stops), and the true residual b - L x of the
    // double iterate replacing r every replace_every iterations and whenever the
    // recursive residual passes the test, which is then taken on the true one
    bool serial_dots = false;          // inner products summed in the reference's serial order (slow; for parity checks)
    bool outer_fp64 = false;
    int replace_every = 25;
    DBuf<double> xd;
    void download_x_f64(int p, double* host);
    // V-cycle storage of the preconditioner: FP32 (the reference's) or FP16 with FP32 arithmetic
    void set_half_storage(bool on);
    // steps of iterative refinement after every exact solve of the interface solver (0: none)
    void set_sigma_refinement(int steps);

    // Krylov plugin operations on level-0 channels (CG_SYSTEM / CG_VECTOR,
    // DDS/CG_SYSTEM.cpp:29-144, DDS/CG_VECTOR.cpp:18-74), enqueued on the stream;
    // the reductions wait for their results (one value per problem)
    DBuf<float>& channel_buffer(int c);
    void multiply(int c_in, int c_out);
    void inner_product(int a, int b, double* out);
    void convergence_norm(int c, float* out);
    void project(int c);
    void vector_op(int op, int c_out, int c_a, int c_b, float scalar);
    // z = M^{-1} rhs then Project(z), into the result channel; tmp is scratch; no other channel changes
    void apply_preconditioner_channels(int c_result, int c_rhs, int c_tmp);
    // the stream every later call is enqueued on (nullptr: the context's own)
    void set_stream(hipStream_t s);
    hipStream_t own_stream = nullptr;
    DevSegments segments();
    PinnedBuf<float> hs_pin;           // per-problem float scalars read back every iteration
    PinnedBuf<int> act_pin;            // per-problem iterate flags sent to the device

private:
    struct Graph {
        std::vector<int> key;          // ping-pong state of every level, skip mode
        std::vector<int> cur_out;      // ping-pong state after the apply
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    void record_apply();
    void build_device();
    void merge_levels();
    void upload_levels();
    void sweeps(int l, int count, bool interior, bool boundary);
    template <class S> void sweeps_t(int l, int count, bool interior, bool boundary);
    template <class S> void vcycle_t();
    void interface_gather();
    void interface_scatter();
};

}  // namespace hdd
