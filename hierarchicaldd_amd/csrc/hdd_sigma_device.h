// Device-side sigma (interface Schur complement) multigrid.
#pragma once
#include <memory>
#include <vector>
#include "hdd_device.h"
#include "hdd_factor.h"
#include "hdd_runtime.h"
#include "hdd_sigma.h"

namespace hdd {

struct CsrDev {
    int rows = 0, cols = 0;
    DBuf<int> p, ix;
    DBuf<float> v;
    void upload(const Csr& m);
    size_t bytes() const { return p.n * 4 + ix.n * 4 + v.n * 4; }
};

// Exact solve x = (A + eps I)^{-1} b for a batch of independent matrices,
// from their supernodal factors (see hdd_sigma_device.hip for the layout).
struct TriPlan {
    int n = 0, H = -1;
    std::vector<int> lev_begin;                 // rows of height h: rowlist[lev_begin[h] .. lev_begin[h+1])
    std::vector<int> lanes;                     // lanes per row, [height][fwd gather, fwd diag, bwd gather, bwd diag]
    DBuf<int> orig_of, sn_of, s0, sz, nr, roff, srows, rowlist, lp, li;
    DBuf<long long> doff, loff;
    DBuf<float> diag, diagT, l21, lv, y, t;
    void build(const std::vector<const SupernodalFactor*>& F, const std::vector<int>& offsets);
    void solve(const float* b_orig, float* x_orig, hipStream_t s);
    size_t bytes() const;
};

struct SigmaLevel {
    int n_if = 0, n_int = 0;
    CsrDev Air, AirT, Arr, P, PT;  // P: this level -> next coarser (rows = this level), PT its transpose
    TriPlan Lii, Lrr;
    DBuf<float> u, b, r, vb, vx, tmp;
};

struct SigmaDevice {
    int levels = 0, smoothing = 3, n_mg = 1;
    // Bottom level: the reference runs 200 smooths from u = 0, a fixed-point
    // iteration with contraction rho = spectral radius of (Arr+eI)^-1 K. When
    // rho^200 is far below float precision the result equals, to rounding, the
    // solution of the bottom Schur system, computed here by one solve of the
    // combined system [[Aii+eI, Air], [Air^T, Arr+eI]].
    bool bottom_direct = false;
    double bottom_rho = -1;
    TriPlan Lbottom;
    DBuf<float> bot_rhs, bot_sol;
    std::vector<SigmaLevel> lv;
    float* b0 = nullptr;   // level-0 interface rhs (lv[0].b)
    float* u0 = nullptr;   // level-0 interface solution (lv[0].u)
    double factor_seconds = 0;
    size_t plan_bytes = 0;
    ~SigmaDevice();
    void smooth(int l, hipStream_t s);
    void residual(int l, hipStream_t s);
    void vcycle(hipStream_t s);
};

std::unique_ptr<SigmaDevice> make_sigma_device(const SigmaHost& host, int smoothing, int n_mg);
void sigma_vcycles(SigmaDevice& S, hipStream_t s);
void launch_interface_gather(SigmaDevice& S, const DevLevel& L0, const int* if_index, const float* z, const float* r, hipStream_t s);
void launch_interface_scatter(SigmaDevice& S, const int* if_index, float* z, hipStream_t s);
void sigma_vcycle_host(SigmaDevice& S, const float* b, float* u, hipStream_t s);
void sigma_smooth_host(SigmaDevice& S, int level, const float* b, float* u, hipStream_t s);

// interface gather (Laplace_ri + Copy_To_Eigen_Array) and scatter (Copy_From_Eigen_Array), in hdd_kernels.hip
void launch_if_gather(const DevLevel& L0, const int* if_index, int n, const float* z, const float* r, float* b0, hipStream_t s);
void launch_if_scatter(const int* if_index, int n, const float* u0, float* z, hipStream_t s);

}  // namespace hdd
