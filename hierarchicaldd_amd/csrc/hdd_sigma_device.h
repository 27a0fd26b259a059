// Device-side sigma (interface Schur complement) multigrid.
#pragma once
#include <exception>
#include <memory>
#include <thread>
#include <vector>
#include "hdd_device.h"
#include "hdd_factor.h"
#include "hdd_runtime.h"
#include "hdd_sigma.h"
#include "hdd_spectral.h"
#include "hdd_triplan.h"

namespace hdd {

struct CsrDev {
    int rows = 0, cols = 0;
    DBuf<int> p, ix;
    DBuf<float> v;
    void upload(const Csr& m);
    size_t bytes() const { return p.n * 4 + ix.n * 4 + v.n * 4; }
};

// Exact solve x = (A + eps I)^{-1} b for a batch of independent matrices,
// from their supernodal factors (plan layout in hdd_triplan.h).
struct TriPlan {
    int n = 0, H = -1;
    size_t factor_nnz = 0;
    // per-height ranges and launch parameters (host copies)
    std::vector<int> sn_begin, sl_begin, sl_rows, row_begin, ci_begin, bi_begin, max_sz, big, fused, G, Gc, Gf, Gb, run_begin, run_max;
    std::vector<int> wave;  // per fused height: 1 = a wavefront per run (k_tri_*_wave), 0 = a workgroup per run (same results)
    bool wave_forced = false;  // HDD_FUSED_WAVE set: the tuner leaves `wave` alone
    // per fused height: a cap on the workgroups of its launch, each looping over runs (default:
    // no cap, a workgroup per run; the tuner tries caps; results do not depend on it)
    std::vector<int> pgrid;
    static constexpr int kFusedGroups = 1 << 30;
    DBuf<int> run_q, run_n, orig_of, s0, sz, nr, roff, srows, sn_of, sn_prob, sn_list, sl_sn, sl_r0, rowlist, inc_ptr, ci_sn, ci_a0, ci_a1, bi_sn, bi_c, bi_j0;
    DBuf<long long> loff, doff, cofs, cpos, bofs, woff;
    DBuf<float> diag, diagT, l21, c, bpart, t, y, w21;
    // fused heights and heights with big supernodes use W = L21 Linv (one phase less per
    // fused workgroup, one launch less per big height and pass; float rounding differs from
    // the L21 form): on unless a fused form is forced (HDD_FUSED_WAVE) or the solves are
    // refined (the exact mode keeps the L21 form)
    bool wform = false;
    bool use_w() const { return wform && refine == 0; }
    // prob_of[m]: problem of factor m (empty: all problem 0)
    void build(const std::vector<const SupernodalFactor*>& F, const std::vector<int>& offsets, const std::vector<int>& prob_of = {});
    // per-problem flags (device, 1 = iterating): supernodes of stopped problems are skipped; nullptr: none skipped
    const int* active = nullptr;
    // x = (A + eps I)^{-1} b; with refine > 0, each solve is followed by `refine`
    // steps of iterative refinement, x += (A + eps I)^{-1} (b - (A + eps I) x), the
    // residual accumulated in double from the plan's matrix refA (its own numbering)
    void solve(const float* b_orig, float* x_orig, hipStream_t s);
    void solve_once(const float* b_orig, float* x_orig, hipStream_t s);
    int refine = 0;
    double refine_eps = 1e-5;
    CsrDev refA;
    DBuf<float> ref_r, ref_d;
    void fwd_height(int h, const float* b_orig, hipStream_t s);
    void bwd_height(int h, float* x_orig, hipStream_t s);
    // launch-shape autotuner (slice rows, contribution items per height); returns the choices changed
    int tune(hipStream_t s);
    size_t bytes() const;
    std::vector<int> ci_entries, h_sn_list, h_sz, h_nr;  // host copies the tuner rebuilds the lists from
    static constexpr int kTuneMinRows = 20000;           // plans below this size keep their default shapes
    void rebuild_lists(hipStream_t s);
    // dense tops (hdd_triplan.h): padded top rows, their tops, the Linv blocks
    int top_nq = 0, Gt = 4;
    long long top_dense_rows = 0;                        // rows of the distinct dense blocks
    DBuf<int> top_rows, top_of, top_inc_ptr, top_base, top_ld;
    DBuf<long long> top_cpos, top_woff;
    DBuf<float> top_W, top_t;
    void top_solve(const float* b_orig, float* x_orig, hipStream_t s);
};

struct SigmaLevel {
    int n_if = 0, n_int = 0;
    int n_distinct = 0;  // distinct interior matrices (subdomains of the same shape share one factor)
    CsrDev Air, AirT, Arr, P, PT;  // P: this level -> next coarser (rows = this level), PT its transpose
    TriPlan Lii, Lrr;
    DBuf<float> u, b, r, vb, vx, tmp;
};

// Host factors of one problem's interface solver: the exact solves of
// SIGMA_SMOOTHER (IFS/SIGMA_SMOOTHER.cpp:115-288, PARDISO phases 11/22 in the
// reference) and the bottom level's combined system and slow eigenmodes.
struct SigmaFactors {
    struct Level {
        std::vector<int> rep;                  // subdomain -> first subdomain with an identical Aii
        std::vector<SupernodalFactor> Aii;     // factors of the representatives (others empty)
        SupernodalFactor Arr;
    };
    std::vector<Level> lv;
    // Bottom level: the 200 smooths of the reference as one exact solve of the
    // combined system [[Aii+eI, Air],[Air^T, Arr+eI]] plus a low-rank
    // correction on the eigenmodes of the smoother that survive 200 sweeps.
    bool bottom_direct = false;                // the combined solve replaces the sweeps
    int bottom_n_int = 0;                      // interior rows of the combined system
    SupernodalFactor bottom;                   // its factor
    BottomSpectrum spec;                       // kept modes (none when the correction is off)
    double bottom_rho = -1;                    // largest eigenvalue of the bottom smoother (-1: not computed)
    int bottom_steps = 0;
    double bottom_seconds = 0;
    double seconds = 0;
    // factor_sigma(..., wait = false) leaves the bottom level's factor and spectrum to a
    // worker thread: wait_bottom() joins it (rethrowing its error) before they are read.
    // Joining is not thread-safe: one thread waits (the one that builds the device plans).
    void wait_bottom() const;
    SigmaFactors() = default;
    SigmaFactors(const SigmaFactors&) = delete;
    SigmaFactors& operator=(const SigmaFactors&) = delete;
    ~SigmaFactors();
    mutable std::thread bottom_worker;
    mutable std::exception_ptr bottom_error;
};
// F must be freshly constructed; the worker (wait = false) reads H until it is joined
void factor_sigma(const SigmaHost& H, SigmaFactors& F, bool wait = true);

struct SigmaDevice {
    int levels = 0, smoothing = 3, n_mg = 1, nprob = 1;
    // Bottom level: the reference runs 200 smooths, a stationary iteration
    // u <- (Arr+eI)^-1 (b + K u) whose iteration matrix M has eigenvalues in
    // [0, 1). Its result is x* + M^200 (u0 - x*) with x* the solution of the
    // bottom Schur system, computed here by one solve of the combined system
    // [[Aii+eI, Air], [Air^T, Arr+eI]], and M^200 kept on the k eigenmodes whose
    // 200th power is above 1e-9 (V, S V and their weights, found at setup).
    // In a batch every problem's combined system is one block of the plan and
    // every problem keeps its own modes; the shortcut is used only when every
    // problem of the batch resolved its spectrum.
    bool bottom_direct = false;
    double bottom_rho = -1;        // largest eigenvalue of M over the batch (-1: not computed)
    int bottom_k = 0;              // eigenmodes kept (all problems)
    int bottom_steps = 0;          // block Krylov steps of the eigen-solves (largest)
    double bottom_seconds = 0;     // host time of the eigen-solves
    double bottom_check = -1;      // max |direct+correction - 200 sweeps| / max |200 sweeps| on a random rhs
    int tuned_choices = 0;         // launch shapes the autotuner changed from the plans' defaults
    long long dense_top_rows = 0;  // rows of the plans' dense top blocks (all plans)
    TriPlan Lbottom;
    DBuf<float> bot_rhs, bot_sol, spec_V, spec_SV, spec_coef;
    DBuf<double> spec_w1, spec_w2, spec_part;
    DBuf<int> bot_pos;             // bottom interface dof -> row of the combined systems
    DBuf<int> row_prob;            // bottom interface dof -> problem
    DBuf<int> mode_row0, mode_len, prob_m0, prob_row0, prob_n;
    DBuf<long long> mode_vofs, prob_vofs;
    int spec_nch = 0;              // chunks of the longest problem
    void bottom(hipStream_t s, bool u_zero);
    std::vector<SigmaLevel> lv;
    float* b0 = nullptr;   // level-0 interface rhs (lv[0].b)
    float* u0 = nullptr;   // level-0 interface solution (lv[0].u)
    double factor_seconds = 0;
    size_t plan_bytes = 0;
    ~SigmaDevice();
    void smooth(int l, hipStream_t s);
    void residual(int l, hipStream_t s);
    void vcycle(hipStream_t s, bool first);
    // per-problem flags of the batch PCG (device): exact solves skip stopped problems; nullptr: skip nothing
    void set_active(const int* active);
    // steps of iterative refinement after every exact solve (0: none, the default)
    int refine_steps = 0;
    void set_refinement(const std::vector<const SigmaHost*>& hosts, int steps);
};

// Device interface solver of a batch: the problems' operators stacked
// block-diagonally (problem order), from their host operators and factors.
// The setup's own device work (the bottom self-check) runs on `stream`.
std::unique_ptr<SigmaDevice> make_sigma_device(const std::vector<const SigmaHost*>& hosts, const std::vector<const SigmaFactors*>& factors,
                                               int smoothing, int n_mg, hipStream_t stream);
void sigma_vcycles(SigmaDevice& S, hipStream_t s);
void launch_interface_gather(SigmaDevice& S, const DevLevel& L0, const int* if_index, const float* z, const float* r, hipStream_t s);
void launch_interface_gather(SigmaDevice& S, const DevLevel& L0, const int* if_index, const __half* z, const __half* r, hipStream_t s);
void launch_interface_scatter(SigmaDevice& S, const int* if_index, float* z, hipStream_t s);
void launch_interface_scatter(SigmaDevice& S, const int* if_index, __half* z, hipStream_t s);
void sigma_vcycle_host(SigmaDevice& S, const float* b, float* u, hipStream_t s);
void sigma_smooth_host(SigmaDevice& S, int level, const float* b, float* u, hipStream_t s);

// interface gather (Laplace_ri + Copy_To_Eigen_Array) and scatter (Copy_From_Eigen_Array), in hdd_kernels.hip
void launch_if_gather(const DevLevel& L0, const int* if_index, int n, const float* z, const float* r, float* b0, hipStream_t s);
void launch_if_gather(const DevLevel& L0, const int* if_index, int n, const __half* z, const __half* r, float* b0, hipStream_t s);
void launch_if_scatter(const int* if_index, int n, const float* u0, float* z, hipStream_t s);
void launch_if_scatter(const int* if_index, int n, const float* u0, __half* z, hipStream_t s);

}  // namespace hdd
