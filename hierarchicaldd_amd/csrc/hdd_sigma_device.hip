// Sigma (interface Schur complement) multigrid on the GPU.
//
//   SIGMA_MULTIGRID_SOLVER::V_Cycle   IFS/SIGMA_MULTIGRID_SOLVER.cpp:36-47
//   SIGMA_SMOOTHER::Smooth            IFS/SIGMA_SMOOTHER.cpp:292-324
//   SIGMA_SMOOTHER::Compute_Residual  IFS/SIGMA_SMOOTHER.cpp:328-352
//
// Exact solves with (A + 1e-5 I) use supernodal Cholesky factors (host,
// double) turned into float triangular-solve plans (hdd_triplan.h). A solve
// walks the supernode tree by height (leaves first forward, root first
// backward); each height is two launches: wavefronts over bounded gather
// items of the sub-diagonal blocks, then workgroups over row slices of each
// supernode that add the items' partial sums and apply the explicit inverse
// of the diagonal block from LDS. All subdomains' Aii of one sigma level form
// one batch, so each height of all their trees is one pair of launches.
//
// Sparse products keep Eigen's accumulation order (rows in column order, the
// transposed products in source-row order), built as gathers so no atomics
// are needed and results are deterministic.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <algorithm>
#include <memory>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <map>
#include <thread>
#include <omp.h>
#include "hdd_factor.h"
#include "hdd_parallel.h"
#include "hdd_runtime.h"
#include "hdd_sigma_device.h"
#include "hdd_spectral.h"

namespace hdd {

// set-up phase timings on stderr when HDD_FACTOR_VERBOSE is set
static void setup_phase(const char* what, int level, std::chrono::steady_clock::time_point& t, const char* unit = "level") {
    static const bool verbose = std::getenv("HDD_FACTOR_VERBOSE") != nullptr;
    const auto now = std::chrono::steady_clock::now();
    if (verbose) std::fprintf(stderr, "[setup] %s (%s %d): %.3fs\n", what, unit, level, std::chrono::duration<double>(now - t).count());
    t = now;
}


namespace {

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_down(v, o, G);
    return v;
}

// In a batch, supernodes of problems that have stopped iterating are skipped
// (active == nullptr: nothing is skipped). Every skip is uniform over the
// workgroup, wavefront or lane group that owns the supernode or row.
__device__ __forceinline__ bool sn_stopped(const int* __restrict__ sn_prob, const int* __restrict__ active, int g) {
    return active && !active[sn_prob[g]];
}

// forward reduce (heights with big supernodes): t_i = b_i - sum of the
// contributions at row i's incidences, G lanes per row
template <int G>
__global__ void __launch_bounds__(256) k_tri_fwd_reduce(const int* __restrict__ rows, int nrows, const int* __restrict__ inc_ptr,
                                                        const long long* __restrict__ cpos, const float* __restrict__ c,
                                                        const int* __restrict__ orig_of, const float* __restrict__ b, float* __restrict__ t,
                                                        const int* __restrict__ sn_of, const int* __restrict__ sn_prob,
                                                        const int* __restrict__ active) {
    const int w = (blockIdx.x * 256 + threadIdx.x) / G, lane = threadIdx.x & (G - 1);
    if (w >= nrows) return;
    const int i = rows[w];
    if (sn_stopped(sn_prob, active, sn_of[i])) return;
    float s = 0.f;
#pragma unroll 8
    for (int p = inc_ptr[i] + lane; p < inc_ptr[i + 1]; p += G) s += c[cpos[p]];
    s = group_sum<G>(s);
    if (lane == 0) t[i] = b[orig_of[i]] - s;
}

// forward diagonal step: a workgroup per row slice [r0, r0 + slice_rows) of a
// supernode. t of rows [0, rend) into LDS (from the reduce kernel's output on
// big heights, else summed here), then y[r] = sum_{k<=r} Linv(r,k) t[k] with
// G lanes per row.
template <int G>
__device__ __forceinline__ void fwd_sn_slice(int q, int slice_rows, const int* __restrict__ sl_sn, const int* __restrict__ sl_r0,
                                             const int* __restrict__ s0, const int* __restrict__ sz,
                                             const long long* __restrict__ doff, const float* __restrict__ diag,
                                             const float* __restrict__ tg, const int* __restrict__ inc_ptr,
                                             const long long* __restrict__ cpos, const float* __restrict__ c,
                                             const int* __restrict__ orig_of, const float* __restrict__ b, float* __restrict__ y,
                                             const int* __restrict__ sn_prob, const int* __restrict__ active, float* ts) {
    const int g = sl_sn[q], r0 = sl_r0[q], n = sz[g], base = s0[g];
    if (sn_stopped(sn_prob, active, g)) return;
    const int rend = min(n, r0 + slice_rows);
    for (int k = threadIdx.x; k < rend; k += 256) {
        const int i = base + k;
        if (tg) {
            ts[k] = tg[i];
        } else {
            float sum = 0.f;
#pragma unroll 8
            for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) sum += c[cpos[p]];
            ts[k] = b[orig_of[i]] - sum;
        }
    }
    __syncthreads();
    const float* D = diag + doff[g];
    const int lane = threadIdx.x & (G - 1);
    for (int r = r0 + (int)threadIdx.x / G; r < rend; r += 256 / G) {
        const float* Dr = D + (long long)r * n;
        float acc = 0.f;
#pragma unroll 8
        for (int k = lane; k <= r; k += G) acc += Dr[k] * ts[k];
        acc = group_sum<G>(acc);
        if (lane == 0) y[base + r] = acc;
    }
}

template <int G>
__global__ void __launch_bounds__(256) k_tri_fwd_sn(int q0, int slice_rows, const int* __restrict__ sl_sn, const int* __restrict__ sl_r0,
                                                    const int* __restrict__ s0, const int* __restrict__ sz,
                                                    const long long* __restrict__ doff, const float* __restrict__ diag,
                                                    const float* __restrict__ tg, const int* __restrict__ inc_ptr,
                                                    const long long* __restrict__ cpos, const float* __restrict__ c,
                                                    const int* __restrict__ orig_of, const float* __restrict__ b, float* __restrict__ y,
                                                    const int* __restrict__ sn_prob, const int* __restrict__ active) {
    extern __shared__ float ts[];
    fwd_sn_slice<G>(q0 + blockIdx.x, slice_rows, sl_sn, sl_r0, s0, sz, doff, diag, tg, inc_ptr, cpos, c, orig_of, b, y, sn_prob, active, ts);
}

// contributions c_g[a] = L_g[a, :] . v_g for the rows [a0, a1) of an item, G lanes per row:
// L21 and v = y, or (W form) W = L21 Linv and v = t
template <int G>
__device__ __forceinline__ void fwd_contrib_item(int it, const int* __restrict__ ci_sn, const int* __restrict__ ci_a0,
                                                 const int* __restrict__ ci_a1, const int* __restrict__ s0, const int* __restrict__ sz,
                                                 const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                 const float* __restrict__ l21, const float* __restrict__ v, float* __restrict__ c,
                                                 const int* __restrict__ sn_prob, const int* __restrict__ active) {
    // a workgroup per item; its rows go to the 256/G lane groups in turn
    const int g = ci_sn[it], n = sz[g], a1 = ci_a1[it];
    if (sn_stopped(sn_prob, active, g)) return;
    const float* L = l21 + loff[g];
    const float* yy = v + s0[g];
    float* cg = c + cofs[g];
    const int lane = threadIdx.x & (G - 1);
    for (int a = ci_a0[it] + (int)threadIdx.x / G; a < a1; a += 256 / G) {
        const float* La = L + (long long)a * n;
        float acc = 0.f;
#pragma unroll 8
        for (int j = lane; j < n; j += G) acc += La[j] * yy[j];
        acc = group_sum<G>(acc);
        if (lane == 0) cg[a] = acc;
    }
}

template <int G>
__global__ void __launch_bounds__(256) k_tri_fwd_contrib(int i0, int i1, const int* __restrict__ ci_sn, const int* __restrict__ ci_a0,
                                                         const int* __restrict__ ci_a1, const int* __restrict__ s0, const int* __restrict__ sz,
                                                         const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                         const float* __restrict__ l21, const float* __restrict__ y, float* __restrict__ c,
                                                         const int* __restrict__ sn_prob, const int* __restrict__ active) {
    const int it = i0 + blockIdx.x;
    if (it >= i1) return;
    fwd_contrib_item<G>(it, ci_sn, ci_a0, ci_a1, s0, sz, loff, cofs, l21, y, c, sn_prob, active);
}

// W form of a height with big supernodes, one launch after the reduce kernel's t:
// workgroups [0, nq) are the diagonal step's slices (y = Linv t), the others the
// contribution items (c = W t); neither waits for the other.
template <int G1, int G2>
__global__ void __launch_bounds__(256) k_tri_fwd_sn_contrib_w(int q0, int nq, int slice_rows, const int* __restrict__ sl_sn,
                                                              const int* __restrict__ sl_r0, const int* __restrict__ s0,
                                                              const int* __restrict__ sz, const long long* __restrict__ doff,
                                                              const float* __restrict__ diag, const float* __restrict__ t,
                                                              float* __restrict__ y, int i0, const int* __restrict__ ci_sn,
                                                              const int* __restrict__ ci_a0, const int* __restrict__ ci_a1,
                                                              const long long* __restrict__ woff, const long long* __restrict__ cofs,
                                                              const float* __restrict__ w21, float* __restrict__ c,
                                                              const int* __restrict__ sn_prob, const int* __restrict__ active) {
    extern __shared__ float ts[];
    if ((int)blockIdx.x < nq)
        fwd_sn_slice<G1>(q0 + blockIdx.x, slice_rows, sl_sn, sl_r0, s0, sz, doff, diag, t, nullptr, nullptr, nullptr, nullptr, nullptr, y,
                         sn_prob, active, ts);
    else
        fwd_contrib_item<G2>(i0 + (int)blockIdx.x - nq, ci_sn, ci_a0, ci_a1, s0, sz, woff, cofs, w21, t, c, sn_prob, active);
}

// backward tiles: one wavefront per (supernode g, row chunk cc, column block j0);
// lane j accumulates L_g[a, j] z[rows_g[a]] over the chunk's rows (z: the
// ancestors' backward results, final; L: L21, or W in the W form)
__device__ __forceinline__ void bwd_tile_item(int it, const int* __restrict__ bi_sn, const int* __restrict__ bi_c,
                                              const int* __restrict__ bi_j0, const int* __restrict__ sz, const int* __restrict__ nr,
                                              const long long* __restrict__ loff, const int* __restrict__ roff,
                                              const int* __restrict__ srows, const long long* __restrict__ bofs,
                                              const float* __restrict__ l21, const float* __restrict__ z, float* __restrict__ bpart,
                                              const int* __restrict__ sn_prob, const int* __restrict__ active) {
    const int lane = threadIdx.x & 63;
    const int g = bi_sn[it], n = sz[g], cc = bi_c[it], j = bi_j0[it] + lane;
    if (sn_stopped(sn_prob, active, g)) return;
    const int a0 = cc * TriPlanHost::kTile, m = min(nr[g] - a0, TriPlanHost::kTile);
    // lane q holds z at the tile's row q; all the tile's loads are issued
    // before the first product (one memory round trip), then summed in row order
    const float zq = lane < m ? z[srows[roff[g] + a0 + lane]] : 0.f;
    const float* L = l21 + loff[g] + (long long)a0 * n + min(j, n - 1);
    float lv[TriPlanHost::kTile];
#pragma unroll
    for (int q = 0; q < TriPlanHost::kTile; ++q) lv[q] = q < m ? L[(long long)q * n] : 0.f;
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < TriPlanHost::kTile; ++q)
        if (q < m) acc += lv[q] * __shfl(zq, q);
    if (j < n) bpart[bofs[g] + (long long)cc * n + j] = acc;
}

__global__ void __launch_bounds__(256) k_tri_bwd_tiles(int i0, int i1, const int* __restrict__ bi_sn, const int* __restrict__ bi_c,
                                                       const int* __restrict__ bi_j0, const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const long long* __restrict__ loff, const int* __restrict__ roff,
                                                       const int* __restrict__ srows, const long long* __restrict__ bofs,
                                                       const float* __restrict__ l21, const float* __restrict__ z, float* __restrict__ bpart,
                                                       const int* __restrict__ sn_prob, const int* __restrict__ active) {
    const int it = i0 + blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (it >= i1) return;
    bwd_tile_item(it, bi_sn, bi_c, bi_j0, sz, nr, loff, roff, srows, bofs, l21, z, bpart, sn_prob, active);
}

// backward reduce (heights with big supernodes): t_j = y_j - sum over the row
// chunks of the tile partials of column j of its supernode, G lanes per row.
// W form (W set): v_j = t_j - (the partials), where t holds Linv^T y, is the
// result: written to z and to x in the original numbering.
template <int G>
__global__ void __launch_bounds__(256) k_tri_bwd_reduce(const int* __restrict__ rows, int nrows, const int* __restrict__ sn_of,
                                                        const int* __restrict__ s0, const int* __restrict__ sz, const int* __restrict__ nr,
                                                        const long long* __restrict__ bofs, const float* __restrict__ bpart,
                                                        const float* __restrict__ y, float* __restrict__ t,
                                                        const int* __restrict__ sn_prob, const int* __restrict__ active) {
    const int w = (blockIdx.x * 256 + threadIdx.x) / G, lane = threadIdx.x & (G - 1);
    if (w >= nrows) return;
    const int i = rows[w], g = sn_of[i], n = sz[g], k = i - s0[g];
    if (sn_stopped(sn_prob, active, g)) return;
    const int nch = (nr[g] + TriPlanHost::kTile - 1) / TriPlanHost::kTile;
    const float* P = bpart + bofs[g] + k;
    float s = 0.f;
#pragma unroll 8
    for (int cc = lane; cc < nch; cc += G) s += P[(long long)cc * n];
    s = group_sum<G>(s);
    if (lane == 0) t[i] = y[i] - s;
}

template <int G>
__global__ void __launch_bounds__(256) k_tri_bwd_reduce_w(const int* __restrict__ rows, int nrows, const int* __restrict__ sn_of,
                                                          const int* __restrict__ s0, const int* __restrict__ sz, const int* __restrict__ nr,
                                                          const long long* __restrict__ bofs, const float* __restrict__ bpart,
                                                          const float* __restrict__ t, const int* __restrict__ orig_of,
                                                          float* __restrict__ z, float* __restrict__ x,
                                                          const int* __restrict__ sn_prob, const int* __restrict__ active) {
    const int w = (blockIdx.x * 256 + threadIdx.x) / G, lane = threadIdx.x & (G - 1);
    if (w >= nrows) return;
    const int i = rows[w], g = sn_of[i], n = sz[g], k = i - s0[g];
    if (sn_stopped(sn_prob, active, g)) return;
    const int nch = (nr[g] + TriPlanHost::kTile - 1) / TriPlanHost::kTile;
    const float* P = bpart + bofs[g] + k;
    float s = 0.f;
#pragma unroll 8
    for (int cc = lane; cc < nch; cc += G) s += P[(long long)cc * n];
    s = group_sum<G>(s);
    if (lane == 0) {
        const float v = t[i] - s;
        z[i] = v;
        x[orig_of[i]] = v;
    }
}

// backward diagonal step: t of rows [r0, n) into LDS (reduce output on big
// heights, else summed here), then z[r] = sum_{k>=r} Linv(k,r) t[k]; y holds
// the forward result and z the backward one (separate arrays: the slices of
// one supernode read rows that other slices write); also writes x in the
// original numbering. With zout set (W form) the input is y itself and the
// result, Linv^T y, goes to zout only.
template <int G>
__device__ __forceinline__ void bwd_sn_slice(int q, int slice_rows, const int* __restrict__ sl_sn, const int* __restrict__ sl_r0,
                                             const int* __restrict__ s0, const int* __restrict__ sz, const int* __restrict__ nr,
                                             const long long* __restrict__ doff, const float* __restrict__ diagT,
                                             const float* __restrict__ tg, const long long* __restrict__ bofs,
                                             const float* __restrict__ bpart, const int* __restrict__ orig_of,
                                             const float* __restrict__ y, float* __restrict__ z, float* __restrict__ x,
                                             float* __restrict__ zout, const int* __restrict__ sn_prob, const int* __restrict__ active,
                                             float* ts) {
    const int g = sl_sn[q], r0 = sl_r0[q], n = sz[g], base = s0[g];
    if (sn_stopped(sn_prob, active, g)) return;
    const int rend = min(n, r0 + slice_rows);
    const int nch = (nr[g] + TriPlanHost::kTile - 1) / TriPlanHost::kTile;
    for (int k = r0 + (int)threadIdx.x; k < n; k += 256) {
        if (tg) {
            ts[k] = tg[base + k];
        } else {
            const float* P = bpart + bofs[g] + k;
            float sum = 0.f;
#pragma unroll 8
            for (int cc = 0; cc < nch; ++cc) sum += P[(long long)cc * n];
            ts[k] = y[base + k] - sum;
        }
    }
    __syncthreads();
    const float* DT = diagT + doff[g];
    const int lane = threadIdx.x & (G - 1);
    for (int r = r0 + (int)threadIdx.x / G; r < rend; r += 256 / G) {
        const float* Dr = DT + (long long)r * n;
        float acc = 0.f;
#pragma unroll 8
        for (int k = r + lane; k < n; k += G) acc += Dr[k] * ts[k];
        acc = group_sum<G>(acc);
        if (lane == 0) {
            if (zout) {
                zout[base + r] = acc;
            } else {
                z[base + r] = acc;
                x[orig_of[base + r]] = acc;
            }
        }
    }
}

template <int G>
__global__ void __launch_bounds__(256) k_tri_bwd_sn(int q0, int slice_rows, const int* __restrict__ sl_sn, const int* __restrict__ sl_r0,
                                                    const int* __restrict__ s0, const int* __restrict__ sz, const int* __restrict__ nr,
                                                    const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                    const float* __restrict__ tg, const long long* __restrict__ bofs,
                                                    const float* __restrict__ bpart, const int* __restrict__ orig_of,
                                                    const float* __restrict__ y, float* __restrict__ z, float* __restrict__ x,
                                                    const int* __restrict__ sn_prob, const int* __restrict__ active) {
    extern __shared__ float ts[];
    bwd_sn_slice<G>(q0 + blockIdx.x, slice_rows, sl_sn, sl_r0, s0, sz, nr, doff, diagT, tg, bofs, bpart, orig_of, y, z, x, nullptr,
                    sn_prob, active, ts);
}

// W form of a height with big supernodes, backward, one launch before the W reduce:
// workgroups [0, nq) are the diagonal step's slices on y (Linv^T y into zout), the
// others the tiles of W^T z (four per workgroup); neither waits for the other.
template <int G>
__global__ void __launch_bounds__(256) k_tri_bwd_sn_tiles_w(int q0, int nq, int slice_rows, const int* __restrict__ sl_sn,
                                                            const int* __restrict__ sl_r0, const int* __restrict__ s0,
                                                            const int* __restrict__ sz, const int* __restrict__ nr,
                                                            const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                            const float* __restrict__ y, float* __restrict__ zout, int i0, int i1,
                                                            const int* __restrict__ bi_sn, const int* __restrict__ bi_c,
                                                            const int* __restrict__ bi_j0, const long long* __restrict__ woff,
                                                            const int* __restrict__ roff, const int* __restrict__ srows,
                                                            const long long* __restrict__ bofs, const float* __restrict__ w21,
                                                            const float* __restrict__ z, float* __restrict__ bpart,
                                                            const int* __restrict__ sn_prob, const int* __restrict__ active) {
    extern __shared__ float ts[];
    if ((int)blockIdx.x < nq) {
        bwd_sn_slice<G>(q0 + blockIdx.x, slice_rows, sl_sn, sl_r0, s0, sz, nr, doff, diagT, y, bofs, nullptr, nullptr, nullptr, nullptr,
                        nullptr, zout, sn_prob, active, ts);
        return;
    }
    const int it = i0 + ((int)blockIdx.x - nq) * 4 + (int)(threadIdx.x >> 6);
    if (it >= i1) return;
    bwd_tile_item(it, bi_sn, bi_c, bi_j0, sz, nr, woff, roff, srows, bofs, w21, z, bpart, sn_prob, active);
}

// Row products of the fused kernels: rows r = threadIdx.x / G + i 256/G below
// nrows of the row-major block A (row stride n), each row's G lanes summing
// A[r][k] v[k][j] over k = k0(r) + lane + u G (u ascending, k < kend(r)) for
// the members j < cnt, then the lane-group tree; out(r, j, sum, leader) takes
// the sums (S: the kernel's largest supernode, 64 or 128 rows). Each round
// issues the loads of kRows rows before its first
// product, so a lane group waits on memory once per kRows rows; every row's
// sum keeps its order, so the round size changes no bit. Two rows per round
// for G >= 8 and one member (one for G = 4, whose rows hold 32 loads per
// lane) took the 256^3 level-0 Arr solve from 1,142 to 918 us; four rows
// (fewer waves per SIMD) read 963 us, two rows at G = 4 1,113 us. Runs of
// four members keep one row per round (their accumulators already fill the
// registers): level-0 Aii 759 us with two rows, 717 us with one.
template <int G, int M, int S, class K0, class KEnd, class Out>
__device__ __forceinline__ void row_products(const float* __restrict__ A, int n, int nrows, int cnt,
                                             const float (*__restrict__ v)[M], K0 k0, KEnd kend, Out out) {
    constexpr int kU = S / G, kStep = 256 / G;
    constexpr int kRows = G >= 8 && M == 1 ? 2 : 1;
    const int lane = threadIdx.x & (G - 1);
    for (int r = (int)threadIdx.x / G; r < nrows; r += kStep * kRows) {
        float d[kRows][kU];
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            const int ri = r + i * kStep, b = k0(ri), e = kend(ri);
            const float* Ar = A + (long long)ri * n;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int k = b + lane + u * G;
                d[i][u] = (ri < nrows && k < e) ? Ar[k] : 0.f;
            }
        }
#pragma unroll
        for (int i = 0; i < kRows; ++i) {
            const int ri = r + i * kStep, b = k0(ri), e = kend(ri);
            float acc[M];
#pragma unroll
            for (int j = 0; j < M; ++j) acc[j] = 0.f;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int k = b + lane + u * G;
                if (ri < nrows && k < e) {
                    float vk[M];
#pragma unroll
                    for (int j = 0; j < M; ++j) vk[j] = v[k][j];
#pragma unroll
                    for (int j = 0; j < M; ++j)
                        if (j < cnt) acc[j] += d[i][u] * vk[j];
                }
            }
            if (kRows > 1 && ri >= nrows) break;  // uniform over the lane group
#pragma unroll
            for (int j = 0; j < M; ++j)
                if (j < cnt) out(ri, j, group_sum<G>(acc[j]), lane == 0);
        }
    }
}

// Heights whose supernodes all have <= kSliceRows (128) rows: one workgroup per
// run of up to M supernodes that share their blocks (members of identical
// factors; runs of one otherwise) does the whole height's work for them in one
// launch. Every block entry is loaded once and applied to each member in turn,
// and each member's sums run in exactly the order of a one-member run.
// forward: t = b - (incidence sums) -> y = Linv t (LDS) -> c_g = L21_g y.
// With WF the block passed as l21 is W_g = L21_g Linv_g (formed in double at
// set-up), so c_g = W_g t needs only t: both products run in one phase.
template <int G, int M, bool WF, int S>
__device__ __forceinline__ void fwd_small_run(int run, int r0, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                       const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                       const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const long long* __restrict__ doff, const float* __restrict__ diag,
                                                       const int* __restrict__ inc_ptr, const long long* __restrict__ cpos,
                                                       const int* __restrict__ orig_of, const float* __restrict__ b,
                                                       const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                       const float* __restrict__ l21, float* __restrict__ c, float* __restrict__ y,
                                                       const int* __restrict__ sn_prob, const int* __restrict__ active) {
    // member-minor vectors: one row's values of all members are one (M = 8: two 16-byte) LDS read
    __shared__ __align__(16) float ts[S][M];
    __shared__ __align__(16) float ys[S][M];
    __shared__ int base_s[M];
    const int q0 = run_q[r0 + run], cnt = run_n[r0 + run];
    const int g0 = sn_list[q0], n = sz[g0], m = nr[g0];
    bool any = false;
    for (int j = 0; j < cnt; ++j) any = any || !sn_stopped(sn_prob, active, sn_list[q0 + j]);
    if (!any) return;
    if (threadIdx.x < cnt) base_s[threadIdx.x] = s0[sn_list[q0 + threadIdx.x]];
    __syncthreads();
    for (int idx = threadIdx.x; idx < M * n; idx += 256) {
        const int j = idx / n, k = idx - j * n;
        float v = 0.f;  // members past the run's count stay zero
        if (j < cnt) {
            const int i = base_s[j] + k;
            float sum = 0.f;
#pragma unroll 8
            for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) sum += c[cpos[p]];
            v = b[orig_of[i]] - sum;
        }
        ts[k][j] = v;
    }
    __syncthreads();
    // y = Linv t (entries k <= r); a member past the run's count is never read
    row_products<G, M, S>(
        diag + doff[g0], n, n, cnt, ts, [](int) { return 0; }, [](int r) { return r + 1; },
        [&](int r, int j, float v, bool leader) {
            if (leader) {
                if (!WF) ys[r][j] = v;
                y[base_s[j] + r] = v;
            }
        });
    if (!WF) __syncthreads();  // WF: the contributions read t, complete since the gather
    row_products<G, M, S>(
        l21 + loff[g0], n, m, cnt, WF ? ts : ys, [](int) { return 0; }, [n](int) { return n; },
        [&](int a, int j, float v, bool leader) {
            if (leader) c[cofs[sn_list[q0 + j]] + a] = v;
        });
}

template <int G, int M, bool WF, int S>
__global__ void __launch_bounds__(256) k_tri_fwd_small(int r0, int nruns, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                       const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                       const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const long long* __restrict__ doff, const float* __restrict__ diag,
                                                       const int* __restrict__ inc_ptr, const long long* __restrict__ cpos,
                                                       const int* __restrict__ orig_of, const float* __restrict__ b,
                                                       const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                       const float* __restrict__ l21, float* __restrict__ c, float* __restrict__ y,
                                                       const int* __restrict__ sn_prob, const int* __restrict__ active) {
    for (int run = blockIdx.x; run < nruns; run += gridDim.x) {
        fwd_small_run<G, M, WF, S>(run, r0, run_q, run_n, sn_list, s0, sz, nr, doff, diag, inc_ptr, cpos, orig_of, b, loff, cofs, l21, c, y, sn_prob, active);
        __syncthreads();  // the next run reuses the LDS vectors
    }
}

// backward: t_j = y_j - sum_a L21_g[a, j] z[rows_g[a]] (four wavefronts split
// the rows: wavefront w takes rows a = w mod 4 in ascending order; lanes own
// columns j and j + 64; LDS reduction) -> z = Linv^T t. With WF the block is
// W_g = L21_g Linv_g and z = Linv^T y - W^T z[rows]: the diagonal product needs
// no ancestor value and runs beside the tile sums.
template <int G, int M, bool WF, int S>
__device__ __forceinline__ void bwd_small_run(int run, int r0, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                       const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                       const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const int* __restrict__ roff, const int* __restrict__ srows,
                                                       const long long* __restrict__ loff, const float* __restrict__ l21,
                                                       const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                       const int* __restrict__ orig_of, const float* __restrict__ y,
                                                       float* __restrict__ z, float* __restrict__ x,
                                                       const int* __restrict__ sn_prob, const int* __restrict__ active) {
    constexpr int kChunk = 128;  // a multiple of 4: the rows' order per wavefront does not depend on it
    __shared__ __align__(16) float zr[kChunk][M];  // member-minor, as ts
    __shared__ float part[4][M][S];
    __shared__ __align__(16) float ts[S][M];
    __shared__ __align__(16) float z1[WF ? S : 1][M];  // WF: Linv^T y
    __shared__ int base_s[M], roff_s[M];
    const int q0 = run_q[r0 + run], cnt = run_n[r0 + run];
    const int g0 = sn_list[q0], n = sz[g0], m = nr[g0];
    bool any = false;
    for (int j = 0; j < cnt; ++j) any = any || !sn_stopped(sn_prob, active, sn_list[q0 + j]);
    if (!any) return;
    if (threadIdx.x < cnt) { base_s[threadIdx.x] = s0[sn_list[q0 + threadIdx.x]]; roff_s[threadIdx.x] = roff[sn_list[q0 + threadIdx.x]]; }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float* L = l21 + loff[g0];
    const int j0 = min(lane, n - 1), j1 = min(lane + 64, n - 1);
    float acc0[M], acc1[M];
#pragma unroll
    for (int j = 0; j < M; ++j) { acc0[j] = 0.f; acc1[j] = 0.f; }
    __syncthreads();
    const float* DT = diagT + doff[g0];
    // WF: Linv^T y into z1, rows and sums as in the diagonal step below
    auto diag_of_y = [&]() {
        row_products<G, M, S>(
            DT, n, n, cnt, ts, [](int r) { return r; }, [n](int) { return n; },
            [&](int r, int j, float v, bool leader) {
                if (leader) z1[r][j] = v;
            });
    };
    if (WF) {
        // t = y, visible after the first chunk's barrier (or the one below without sub-diagonal rows)
        for (int idx = threadIdx.x; idx < M * n; idx += 256) {
            const int j = idx / n, k = idx - j * n;
            ts[k][j] = j < cnt ? y[base_s[j] + k] : 0.f;
        }
        if (m == 0) {
            __syncthreads();
            diag_of_y();
        }
    }
    for (int a0 = 0; a0 < m; a0 += kChunk) {
        const int cm = min(kChunk, m - a0);
        for (int idx = threadIdx.x; idx < M * cm; idx += 256) {
            const int j = idx / cm, q = idx - j * cm;
            zr[q][j] = j < cnt ? z[srows[roff_s[j] + a0 + q]] : 0.f;
        }
        __syncthreads();
        if (WF && a0 == 0) diag_of_y();
#pragma unroll 4
        for (int q = w; q < cm; q += 4) {
            const float* La = L + (long long)(a0 + q) * n;
            const float l0 = La[j0], l1 = La[j1];
            float zq[M];
#pragma unroll
            for (int j = 0; j < M; ++j) zq[j] = zr[q][j];
#pragma unroll
            for (int j = 0; j < M; ++j)
                if (j < cnt) {
                    acc0[j] += l0 * zq[j];
                    acc1[j] += l1 * zq[j];
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < M; ++j)
        if (j < cnt) {
            if (lane < n) part[w][j][lane] = acc0[j];
            if (lane + 64 < n) part[w][j][lane + 64] = acc1[j];
        }
    __syncthreads();
    if (WF) {
        for (int idx = threadIdx.x; idx < M * n; idx += 256) {
            const int j = idx / n, k = idx - j * n;
            if (j < cnt) {
                const float v = z1[k][j] - ((part[0][j][k] + part[1][j][k]) + (part[2][j][k] + part[3][j][k]));
                z[base_s[j] + k] = v;
                x[orig_of[base_s[j] + k]] = v;
            }
        }
        return;
    }
    for (int idx = threadIdx.x; idx < M * n; idx += 256) {
        const int j = idx / n, k = idx - j * n;
        ts[k][j] = j < cnt ? y[base_s[j] + k] - ((part[0][j][k] + part[1][j][k]) + (part[2][j][k] + part[3][j][k])) : 0.f;
    }
    __syncthreads();
    row_products<G, M, S>(
        DT, n, n, cnt, ts, [](int r) { return r; }, [n](int) { return n; },
        [&](int r, int j, float v, bool leader) {
            if (leader) {
                z[base_s[j] + r] = v;
                x[orig_of[base_s[j] + r]] = v;
            }
        });
}

template <int G, int M, bool WF, int S>
__global__ void __launch_bounds__(256) k_tri_bwd_small(int r0, int nruns, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                       const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                       const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const int* __restrict__ roff, const int* __restrict__ srows,
                                                       const long long* __restrict__ loff, const float* __restrict__ l21,
                                                       const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                       const int* __restrict__ orig_of, const float* __restrict__ y,
                                                       float* __restrict__ z, float* __restrict__ x,
                                                       const int* __restrict__ sn_prob, const int* __restrict__ active) {
    for (int run = blockIdx.x; run < nruns; run += gridDim.x) {
        bwd_small_run<G, M, WF, S>(run, r0, run_q, run_n, sn_list, s0, sz, nr, roff, srows, loff, l21, doff, diagT, orig_of, y, z, x, sn_prob, active);
        __syncthreads();  // the next run reuses the LDS vectors
    }
}

// Wavefront-per-run forms of the two kernels above, four runs per 256-thread
// workgroup: heights of many tiny supernodes leave most of a workgroup idle and
// pay a workgroup barrier per phase, while a wavefront needs none. Every row's
// sums are formed exactly as above (G lanes per row, entries k = lane + u G in
// ascending u, then the lane-group tree; backward, the four row classes a mod 4
// in ascending order, combined as ((0 + 1) + (2 + 3))), so both forms give the
// same bits and the tuner may pick either per height.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G, int M>
__device__ __forceinline__ void fwd_wave_run(int run, int r0, int nruns, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                      const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                      const int* __restrict__ sz, const int* __restrict__ nr,
                                                      const long long* __restrict__ doff, const float* __restrict__ diag,
                                                      const int* __restrict__ inc_ptr, const long long* __restrict__ cpos,
                                                      const int* __restrict__ orig_of, const float* __restrict__ b,
                                                      const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                      const float* __restrict__ l21, float* __restrict__ c, float* __restrict__ y,
                                                      const int* __restrict__ sn_prob, const int* __restrict__ active) {
    __shared__ __align__(16) float ts_all[4][TriPlanHost::kSliceRows][M];
    __shared__ __align__(16) float ys_all[4][TriPlanHost::kSliceRows][M];
    __shared__ int base_all[4][M];
    const int w = threadIdx.x >> 6, lane64 = threadIdx.x & 63;
    float(*ts)[M] = ts_all[w];
    float(*ys)[M] = ys_all[w];
    int* base_s = base_all[w];
    const int q0 = run_q[r0 + run], cnt = run_n[r0 + run];
    const int g0 = sn_list[q0], n = sz[g0], m = nr[g0];
    bool any = false;
    for (int j = 0; j < cnt; ++j) any = any || !sn_stopped(sn_prob, active, sn_list[q0 + j]);
    if (!any) return;
    if (lane64 < cnt) base_s[lane64] = s0[sn_list[q0 + lane64]];
    wave_sync();
    for (int idx = lane64; idx < M * n; idx += 64) {
        const int j = idx / n, k = idx - j * n;
        float v = 0.f;
        if (j < cnt) {
            const int i = base_s[j] + k;
            float sum = 0.f;
#pragma unroll 8
            for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) sum += c[cpos[p]];
            v = b[orig_of[i]] - sum;
        }
        ts[k][j] = v;
    }
    wave_sync();
    const float* D = diag + doff[g0];
    const int lane = lane64 & (G - 1);
    constexpr int kU = TriPlanHost::kSliceRows / G;
    for (int r = lane64 / G; r < n; r += 64 / G) {
        const float* Dr = D + (long long)r * n;
        float acc[M];
#pragma unroll
        for (int j = 0; j < M; ++j) acc[j] = 0.f;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = lane + u * G;
            if (k <= r) {
                const float d = Dr[k];
                float tk[M];
#pragma unroll
                for (int j = 0; j < M; ++j) tk[j] = ts[k][j];
#pragma unroll
                for (int j = 0; j < M; ++j)
                    if (j < cnt) acc[j] += d * tk[j];
            }
        }
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j < cnt) {
                const float v = group_sum<G>(acc[j]);
                if (lane == 0) { ys[r][j] = v; y[base_s[j] + r] = v; }
            } else if (lane == 0) {
                ys[r][j] = 0.f;
            }
    }
    wave_sync();
    const float* L = l21 + loff[g0];
    for (int a = lane64 / G; a < m; a += 64 / G) {
        const float* La = L + (long long)a * n;
        float acc[M];
#pragma unroll
        for (int j = 0; j < M; ++j) acc[j] = 0.f;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = lane + u * G;
            if (k < n) {
                const float l = La[k];
                float yk[M];
#pragma unroll
                for (int j = 0; j < M; ++j) yk[j] = ys[k][j];
#pragma unroll
                for (int j = 0; j < M; ++j)
                    if (j < cnt) acc[j] += l * yk[j];
            }
        }
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j < cnt) {
                const float v = group_sum<G>(acc[j]);
                if (lane == 0) c[cofs[sn_list[q0 + j]] + a] = v;
            }
    }
}

template <int G, int M>
__global__ void __launch_bounds__(256) k_tri_fwd_wave(int r0, int nruns, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                      const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                      const int* __restrict__ sz, const int* __restrict__ nr,
                                                      const long long* __restrict__ doff, const float* __restrict__ diag,
                                                      const int* __restrict__ inc_ptr, const long long* __restrict__ cpos,
                                                      const int* __restrict__ orig_of, const float* __restrict__ b,
                                                      const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                      const float* __restrict__ l21, float* __restrict__ c, float* __restrict__ y,
                                                      const int* __restrict__ sn_prob, const int* __restrict__ active) {
    for (int run = blockIdx.x * 4 + (int)(threadIdx.x >> 6); run < nruns; run += gridDim.x * 4) {
        fwd_wave_run<G, M>(run, r0, nruns, run_q, run_n, sn_list, s0, sz, nr, doff, diag, inc_ptr, cpos, orig_of, b, loff, cofs, l21, c, y, sn_prob, active);
        wave_sync();  // the next run reuses this wavefront's LDS vectors
    }
}

template <int G, int M>
__device__ __forceinline__ void bwd_wave_run(int run, int r0, int nruns, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                      const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                      const int* __restrict__ sz, const int* __restrict__ nr,
                                                      const int* __restrict__ roff, const int* __restrict__ srows,
                                                      const long long* __restrict__ loff, const float* __restrict__ l21,
                                                      const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                      const int* __restrict__ orig_of, const float* __restrict__ y,
                                                      float* __restrict__ z, float* __restrict__ x,
                                                      const int* __restrict__ sn_prob, const int* __restrict__ active) {
    constexpr int kChunk = 128;
    __shared__ __align__(16) float zr_all[4][kChunk][M];
    __shared__ __align__(16) float ts_all[4][TriPlanHost::kSliceRows][M];
    __shared__ int base_all[4][M], roff_all[4][M];
    const int w = threadIdx.x >> 6, lane64 = threadIdx.x & 63;
    float(*zr)[M] = zr_all[w];
    float(*ts)[M] = ts_all[w];
    int* base_s = base_all[w];
    int* roff_s = roff_all[w];
    const int q0 = run_q[r0 + run], cnt = run_n[r0 + run];
    const int g0 = sn_list[q0], n = sz[g0], m = nr[g0];
    bool any = false;
    for (int j = 0; j < cnt; ++j) any = any || !sn_stopped(sn_prob, active, sn_list[q0 + j]);
    if (!any) return;
    if (lane64 < cnt) { base_s[lane64] = s0[sn_list[q0 + lane64]]; roff_s[lane64] = roff[sn_list[q0 + lane64]]; }
    const float* L = l21 + loff[g0];
    const int j0 = min(lane64, n - 1), j1 = min(lane64 + 64, n - 1);
    // acc0[c][j] / acc1[c][j]: the sums over the rows a = c mod 4 of columns j0 / j1
    float acc0[4][M], acc1[4][M];
#pragma unroll
    for (int cl = 0; cl < 4; ++cl)
#pragma unroll
        for (int j = 0; j < M; ++j) { acc0[cl][j] = 0.f; acc1[cl][j] = 0.f; }
    wave_sync();
    for (int a0 = 0; a0 < m; a0 += kChunk) {
        const int cm = min(kChunk, m - a0);
        for (int idx = lane64; idx < M * cm; idx += 64) {
            const int j = idx / cm, q = idx - j * cm;
            zr[q][j] = j < cnt ? z[srows[roff_s[j] + a0 + q]] : 0.f;
        }
        wave_sync();
        for (int q = 0; q < cm; q += 4) {
#pragma unroll
            for (int cl = 0; cl < 4; ++cl) {
                if (q + cl >= cm) break;
                const float* La = L + (long long)(a0 + q + cl) * n;
                const float l0 = La[j0], l1 = La[j1];
                float zq[M];
#pragma unroll
                for (int j = 0; j < M; ++j) zq[j] = zr[q + cl][j];
#pragma unroll
                for (int j = 0; j < M; ++j)
                    if (j < cnt) {
                        acc0[cl][j] += l0 * zq[j];
                        acc1[cl][j] += l1 * zq[j];
                    }
            }
        }
        wave_sync();
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (lane64 < n) ts[lane64][j] = j < cnt ? y[base_s[j] + lane64] - ((acc0[0][j] + acc0[1][j]) + (acc0[2][j] + acc0[3][j])) : 0.f;
        if (lane64 + 64 < n)
            ts[lane64 + 64][j] = j < cnt ? y[base_s[j] + lane64 + 64] - ((acc1[0][j] + acc1[1][j]) + (acc1[2][j] + acc1[3][j])) : 0.f;
    }
    wave_sync();
    const float* DT = diagT + doff[g0];
    const int gl = lane64 & (G - 1);
    constexpr int kU = TriPlanHost::kSliceRows / G;
    for (int r = lane64 / G; r < n; r += 64 / G) {
        const float* Dr = DT + (long long)r * n;
        float acc[M];
#pragma unroll
        for (int j = 0; j < M; ++j) acc[j] = 0.f;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = r + gl + u * G;
            if (k < n) {
                const float d = Dr[k];
                float tk[M];
#pragma unroll
                for (int j = 0; j < M; ++j) tk[j] = ts[k][j];
#pragma unroll
                for (int j = 0; j < M; ++j)
                    if (j < cnt) acc[j] += d * tk[j];
            }
        }
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j < cnt) {
                const float v = group_sum<G>(acc[j]);
                if (gl == 0) { z[base_s[j] + r] = v; x[orig_of[base_s[j] + r]] = v; }
            }
    }
}

template <int G, int M>
__global__ void __launch_bounds__(256) k_tri_bwd_wave(int r0, int nruns, const int* __restrict__ run_q, const int* __restrict__ run_n,
                                                      const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                      const int* __restrict__ sz, const int* __restrict__ nr,
                                                      const int* __restrict__ roff, const int* __restrict__ srows,
                                                      const long long* __restrict__ loff, const float* __restrict__ l21,
                                                      const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                      const int* __restrict__ orig_of, const float* __restrict__ y,
                                                      float* __restrict__ z, float* __restrict__ x,
                                                      const int* __restrict__ sn_prob, const int* __restrict__ active) {
    for (int run = blockIdx.x * 4 + (int)(threadIdx.x >> 6); run < nruns; run += gridDim.x * 4) {
        bwd_wave_run<G, M>(run, r0, nruns, run_q, run_n, sn_list, s0, sz, nr, roff, srows, loff, l21, doff, diagT, orig_of, y, z, x, sn_prob, active);
        wave_sync();  // the next run reuses this wavefront's LDS vectors
    }
}

// y = A x for CSR A (one thread per row, column order)
__global__ void k_csr_mv(int rows, const int* __restrict__ p, const int* __restrict__ ix, const float* __restrict__ v,
                         const float* __restrict__ x, float* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float acc = 0.f;
#pragma unroll 8
    for (int e = p[i]; e < p[i + 1]; ++e) acc += v[e] * x[ix[e]];
    y[i] = acc;
}

// out = b + sum (transposed product as a gather), optionally - A u (Compute_Residual)
__global__ void k_schur_rhs(int rows, const float* __restrict__ b, const int* __restrict__ tp, const int* __restrict__ ti,
                            const float* __restrict__ tv, const float* __restrict__ vx, const int* __restrict__ ap,
                            const int* __restrict__ ai, const float* __restrict__ av, const float* __restrict__ u, float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= rows) return;
    float acc = b[c];
#pragma unroll 8
    for (int e = tp[c]; e < tp[c + 1]; ++e) acc += tv[e] * vx[ti[e]];
    if (ap) {
#pragma unroll 8
        for (int e = ap[c]; e < ap[c + 1]; ++e) acc += av[e] * (-u[ai[e]]);
    }
    out[c] = acc;
}

// u[i] += sum_c P(i,c) uc[c]
__global__ void k_prolong_add(int rows, const int* __restrict__ p, const int* __restrict__ ix, const float* __restrict__ v,
                              const float* __restrict__ uc, float* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float acc = 0.f;
#pragma unroll 8
    for (int e = p[i]; e < p[i + 1]; ++e) acc += v[e] * uc[ix[e]];
    u[i] += acc;
}

// Low-rank part of the bottom solve. Modes of all problems are numbered
// together; mode j belongs to the problem whose bottom interface dofs are
// [mode_row0[j], mode_row0[j] + mode_len[j]), and its vectors V_j and (S V)_j
// hold mode_len[j] entries at mode_vofs[j].
// part[j][c] = (V_j . b, (S V)_j . u) over rows [c*kSpecChunk, (c+1)*kSpecChunk)
// of the mode's problem, in double; u == nullptr means u = 0.
constexpr int kSpecChunk = 4096;
__global__ void __launch_bounds__(256) k_spec_dot(int nch, const int* __restrict__ mode_row0, const int* __restrict__ mode_len,
                                                  const long long* __restrict__ mode_vofs, const float* __restrict__ V,
                                                  const float* __restrict__ SV, const float* __restrict__ b, const float* __restrict__ u,
                                                  double* __restrict__ part) {
    const int c = blockIdx.x, j = blockIdx.y;
    const int n = mode_len[j];
    const int r0 = c * kSpecChunk, r1 = min(n, r0 + kSpecChunk);
    if (r0 >= n) return;
    const float* vj = V + mode_vofs[j];
    const float* bj = b + mode_row0[j];
    double sb = 0.0, su = 0.0;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) sb += (double)vj[i] * (double)bj[i];
    if (u) {
        const float* sj = SV + mode_vofs[j];
        const float* uj = u + mode_row0[j];
        for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) su += (double)sj[i] * (double)uj[i];
    }
    for (int o = 32; o > 0; o >>= 1) { sb += __shfl_down(sb, o); su += __shfl_down(su, o); }
    __shared__ double red[2][4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][wv] = sb; red[1][wv] = su; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[((size_t)j * nch + c) * 2] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        part[((size_t)j * nch + c) * 2 + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// coef_j = l_j^200 (v_j^T S u - v_j^T b / (1 - l_j)) = w2_j * su_j - w1_j * sb_j
__global__ void k_spec_coef(int k, int nch, const int* __restrict__ mode_len, const double* __restrict__ part, const double* __restrict__ w1,
                            const double* __restrict__ w2, float* __restrict__ coef) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const int ncj = (mode_len[j] + kSpecChunk - 1) / kSpecChunk;
    double sb = 0.0, su = 0.0;
    for (int c = 0; c < ncj; ++c) { sb += part[((size_t)j * nch + c) * 2]; su += part[((size_t)j * nch + c) * 2 + 1]; }
    coef[j] = (float)(w2[j] * su - w1[j] * sb);
}

// u = x* + V coef: bottom interface dof i of problem p takes x* from row
// bot_pos[i] of the combined solution and the problem's modes [prob_m0[p], prob_m0[p+1])
__global__ void k_spec_apply(int n, const int* __restrict__ row_prob, const int* __restrict__ prob_row0, const int* __restrict__ prob_n,
                             const int* __restrict__ prob_m0, const long long* __restrict__ prob_vofs, const float* __restrict__ V,
                             const float* __restrict__ coef, const float* __restrict__ sol, const int* __restrict__ bot_pos,
                             float* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = row_prob[i], r = i - prob_row0[p], np = prob_n[p], m0 = prob_m0[p], m1 = prob_m0[p + 1];
    const float* Vp = V + prob_vofs[p];
    float acc = sol[bot_pos[i]];
    for (int j = m0; j < m1; ++j) acc += Vp[(size_t)(j - m0) * np + r] * coef[j];
    u[i] = acc;
}

// Dense top of a tree, right-hand side: t_q = b_i - (contributions of the
// supernodes outside the top to row i = rows[q]), G lanes per row. Padding
// rows (-1) keep t = 0 from the set-up.
template <int G>
__global__ void __launch_bounds__(256) k_top_rhs(int nq, const int* __restrict__ rows, const int* __restrict__ inc_ptr,
                                                 const long long* __restrict__ cpos, const float* __restrict__ c,
                                                 const int* __restrict__ orig_of, const float* __restrict__ b, float* __restrict__ t,
                                                 const int* __restrict__ sn_of, const int* __restrict__ sn_prob,
                                                 const int* __restrict__ active) {
    const int q = (blockIdx.x * 256 + threadIdx.x) / G, lane = threadIdx.x & (G - 1);
    if (q >= nq) return;
    const int i = rows[q];
    if (i < 0 || sn_stopped(sn_prob, active, sn_of[i])) return;
    float s = 0.f;
#pragma unroll 8
    for (int p = inc_ptr[q] + lane; p < inc_ptr[q + 1]; p += G) s += c[cpos[p]];
    s = group_sum<G>(s);
    if (lane == 0) t[q] = b[orig_of[i]] - s;
}

// Dense top, both passes over one ld x ld row-major block P holding
// Linv = L_SS^{-1} twice: P[i][k] = Linv(i,k) for k <= i (rows of Linv) and
// P[j][i] = Linv(i,j) for i > j (columns of Linv), padding zero. A wavefront
// per row, float4 loads, fixed-order sums.
//   forward:  y_i = sum_{k <= i} P[i][k] t_k       (row i, columns [0, i])
//   backward: z_j = sum_{i >= j} P[j][i] y_i       (row j, columns [j, ld))
// z goes to the backward pass's input (zb) and to x.
__device__ __forceinline__ float dot4_masked(const float4 w, const float4 v, int c0, int lo, int hi) {
    // sum of w[c] v[c] over the columns c0 + e in [lo, hi]
    float a = 0.f;
    if (c0 >= lo && c0 <= hi) a += w.x * v.x;
    if (c0 + 1 >= lo && c0 + 1 <= hi) a += w.y * v.y;
    if (c0 + 2 >= lo && c0 + 2 <= hi) a += w.z * v.z;
    if (c0 + 3 >= lo && c0 + 3 <= hi) a += w.w * v.w;
    return a;
}

template <bool kForward>
__global__ void __launch_bounds__(256) k_top_tri(int nq, const int* __restrict__ rows, const int* __restrict__ top_of,
                                                 const int* __restrict__ top_base, const int* __restrict__ top_ld,
                                                 const long long* __restrict__ top_woff, const float* __restrict__ P,
                                                 const float* __restrict__ v, float* __restrict__ out, const int* __restrict__ orig_of,
                                                 float* __restrict__ zb, float* __restrict__ x, const int* __restrict__ sn_of,
                                                 const int* __restrict__ sn_prob, const int* __restrict__ active) {
    const int qq = blockIdx.x * 4 + (int)threadIdx.x / 64, lane = threadIdx.x & 63;
    if (qq >= nq) return;
    const int q = kForward ? nq - 1 - qq : qq;  // longest rows first (forward rows grow with r, backward rows shrink)
    const int i = rows[q];
    if (i < 0 || sn_stopped(sn_prob, active, sn_of[i])) return;
    const int k = top_of[q], base = top_base[k], ld = top_ld[k], r = q - base;
    const float4* Pr = reinterpret_cast<const float4*>(P + top_woff[k] + (long long)r * ld);
    const float4* vv = reinterpret_cast<const float4*>(v + base);
    const int lo = kForward ? 0 : r, hi = kForward ? r : ld - 1;
    float acc = 0.f;
    for (int j = lo / 4 + lane; j <= hi / 4; j += 64) {
        const float4 w = Pr[j], t = vv[j];
        if (4 * j >= lo && 4 * j + 3 <= hi) acc += (w.x * t.x + w.y * t.y) + (w.z * t.z + w.w * t.w);
        else acc += dot4_masked(w, t, 4 * j, lo, hi);
    }
    acc = group_sum<64>(acc);
    if (lane == 0) {
        if (kForward) {
            out[q] = acc;
        } else {
            zb[i] = acc;
            x[orig_of[i]] = acc;
        }
    }
}

// P (ld x ld row-major float, see k_top_tri) from Linv (m x m column-major double, lower)
__global__ void k_top_pack(int m, int ld, const double* __restrict__ Li, float* __restrict__ P) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)m * ld) return;
    const int i = (int)(e / ld), j = (int)(e % ld);
    float val = 0.f;
    if (j <= i) val = (float)Li[(long long)j * m + i];        // Linv(i, j)
    else if (j < m) val = (float)Li[(long long)i * m + j];    // Linv(j, i)
    P[(long long)i * ld + j] = val;
}

// dst[pos[i]] = src[i] and its inverse, for the combined systems' interface rows
__global__ void k_pos_scatter(int n, const int* __restrict__ pos, const float* __restrict__ src, float* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[pos[i]] = src[i];
}
__global__ void k_pos_gather(int n, const int* __restrict__ pos, const float* __restrict__ src, float* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[pos[i]];
}

}  // namespace

static inline int blocks_for(int n, int per) { return (n + per - 1) / per; }

void CsrDev::upload(const Csr& m) {
    rows = m.rows; cols = m.cols;
    p.upload(m.ptr); ix.upload(m.idx); v.upload(m.val);
}

static Csr transpose(const Csr& A) {
    Csr T; T.rows = A.cols; T.cols = A.rows;
    std::vector<int> cnt(A.cols + 1, 0);
    for (int e = 0; e < (int)A.nnz(); ++e) cnt[A.idx[e] + 1]++;
    for (int c = 0; c < A.cols; ++c) cnt[c + 1] += cnt[c];
    T.ptr = cnt;
    T.idx.resize(A.nnz()); T.val.resize(A.nnz());
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < A.rows; ++r)  // source rows ascending -> each transposed row in source-row order
        for (int e = A.ptr[r]; e < A.ptr[r + 1]; ++e) { const int k = fill[A.idx[e]]++; T.idx[k] = r; T.val[k] = A.val[e]; }
    return T;
}

static void rocblas_check(rocblas_status st, const char* what) {
    if (st != rocblas_status_success) throw std::runtime_error(std::string("rocBLAS ") + what + ": " + rocblas_status_to_string(st));
}

// The block P of a dense top (see k_top_tri) from L_SS: the triangular inverse
// by rocBLAS in double on the null stream, then rounded to float. One block at
// a time (the device buffers are 2 m^2 doubles), serialised over the process's
// threads.
static void dense_top_inverse(const TriPlanHost::TopBlock& B, float* P) {
    const int m = B.m;
    DBuf<double> dL, dI;
    {
        // filled in pinned host memory: one DMA copy of up to 1.2 GB
        double* L = nullptr;
        hdd_check(hipHostMalloc((void**)&L, (size_t)m * m * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        try {
            dense_top_factor(B, L);
            dL.alloc((size_t)m * m);
            hdd_check(hipMemcpy(dL.p, L, (size_t)m * m * sizeof(double), hipMemcpyHostToDevice), "dense top upload");
        } catch (...) {
            (void)hipHostFree(L);
            throw;
        }
        hdd_check(hipHostFree(L), "hipHostFree");
    }
    dI.alloc((size_t)m * m);
    static std::mutex mu;
    static std::map<int, rocblas_handle> handles;  // one per device: a handle belongs to the device current at its creation
    std::lock_guard<std::mutex> lock(mu);
    int dev = 0;
    hdd_check(hipGetDevice(&dev), "hipGetDevice");
    rocblas_handle& handle = handles[dev];
    if (!handle) rocblas_check(rocblas_create_handle(&handle), "create handle");
    rocblas_check(rocblas_set_stream(handle, nullptr), "set stream");
    rocblas_check(rocblas_dtrtri(handle, rocblas_fill_lower, rocblas_diagonal_non_unit, m, dL.p, m, dI.p, m), "dtrtri");
    const long long cells = (long long)m * B.ld;
    hipLaunchKernelGGL(k_top_pack, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, nullptr, m, B.ld, (const double*)dI.p, P);
    hdd_check(hipGetLastError(), "dense top pack");
    hdd_check(hipStreamSynchronize(nullptr), "dense top");
}

void TriPlan::build(const std::vector<const SupernodalFactor*>& F, const std::vector<int>& offsets, const std::vector<int>& prob_of) {
    TriPlanHost P;
    auto tp = std::chrono::steady_clock::now();
    build_triplan(F, offsets, P, dense_top_min_rows());
    setup_phase("  plan lists and float blocks", (int)F.size(), tp, "factors");
    n = P.n; H = P.H; factor_nnz = P.factor_nnz;
    if (n == 0) return;
    {
        // problem of each supernode (supernodes are numbered factor by factor, in order)
        std::vector<int> sp(P.s0.size(), 0);
        size_t g = 0;
        for (size_t m = 0; m < F.size(); ++m)
            for (size_t k = 0; k < F[m]->sn.size(); ++k, ++g) sp[g] = prob_of.empty() ? 0 : prob_of[m];
        if (g != sp.size()) throw std::runtime_error("triangular-solve plan: supernode count mismatch");
        sn_prob.upload(sp);
    }
    sn_begin = P.sn_begin; sl_begin = P.sl_begin; sl_rows = P.sl_rows; row_begin = P.row_begin; ci_begin = P.ci_begin; bi_begin = P.bi_begin;
    ci_entries = P.ci_entries; h_sn_list = P.sn_list; h_sz = P.sz; h_nr = P.nr;
    run_begin = P.run_begin; run_max = P.run_max;
    // HDD_FUSED_WAVE=1 / 0 forces the wavefront / workgroup form on every fused height (same
    // results); unset, the autotuner picks per height
    {
        const char* e = std::getenv("HDD_FUSED_WAVE");
        wave_forced = e && (std::string(e) == "0" || std::string(e) == "1");
        wave.assign(P.H + 1, wave_forced && std::string(e) == "1" ? 1 : 0);
        wform = !wave_forced && !P.w21.empty();
    }
    pgrid.assign(P.H + 1, kFusedGroups);
    run_q.upload(P.run_q); run_n.upload(P.run_n);
    max_sz = P.max_sz; big = P.big; fused = P.fused; G = P.G; Gc = P.Gc; Gf = P.Gf; Gb = P.Gb;
    orig_of.upload(P.orig_of); s0.upload(P.s0); sz.upload(P.sz); nr.upload(P.nr); roff.upload(P.roff); srows.upload(P.srows);
    sn_of.upload(P.sn_of); sn_list.upload(P.sn_list); sl_sn.upload(P.sl_sn); sl_r0.upload(P.sl_r0); rowlist.upload(P.rowlist); inc_ptr.upload(P.inc_ptr);
    ci_sn.upload(P.ci_sn); ci_a0.upload(P.ci_a0); ci_a1.upload(P.ci_a1); bi_sn.upload(P.bi_sn); bi_c.upload(P.bi_c); bi_j0.upload(P.bi_j0);
    loff.upload(P.loff); doff.upload(P.doff); cofs.upload(P.cofs); cpos.upload(P.cpos); bofs.upload(P.bofs);
    diag.upload(P.diag); diagT.upload(P.diagT); l21.upload(P.l21);
    if (wform) { woff.upload(P.woff); w21.upload(P.w21); }
    c.alloc(std::max<long long>(1, P.c_size)); bpart.alloc(std::max<long long>(1, P.bpart_size));
    t.alloc(n); y.alloc(2 * (size_t)n);  // forward result, then backward result
    setup_phase("  plan uploaded", (int)F.size(), tp, "factors");
    if (!P.tops.empty()) {
        top_nq = (int)P.top_rows.size(); Gt = P.Gt;
        std::vector<long long> block_off, woff;
        long long wsize = 0;
        top_dense_rows = 0;
        for (const TriPlanHost::TopBlock& B : P.top_blocks) {
            block_off.push_back(wsize);
            wsize += (long long)B.ld * B.ld;
            top_dense_rows += B.m;
        }
        std::vector<int> base, ld;
        for (const TriPlanHost::Top& T : P.tops) { base.push_back(T.base); ld.push_back(T.ld); woff.push_back(block_off[T.w]); }
        top_rows.upload(P.top_rows); top_of.upload(P.top_of); top_inc_ptr.upload(P.top_inc_ptr);
        top_cpos.upload(P.top_cpos.empty() ? std::vector<long long>(1, 0) : P.top_cpos);
        top_base.upload(base); top_ld.upload(ld); top_woff.upload(woff);
        top_t.alloc(2 * (size_t)top_nq);  // t, then y (padding stays zero)
        top_W.alloc(wsize);
        for (size_t k = 0; k < P.top_blocks.size(); ++k) dense_top_inverse(P.top_blocks[k], top_W.p + block_off[k]);
        setup_phase("  dense tops inverted", (int)P.top_blocks.size(), tp, "blocks");
    }
}

size_t TriPlan::bytes() const {
    return (run_q.n + run_n.n + orig_of.n + s0.n + sz.n + nr.n + roff.n + srows.n + sn_of.n + sn_prob.n + sn_list.n + sl_sn.n + sl_r0.n + rowlist.n + inc_ptr.n + ci_sn.n + ci_a0.n +
            ci_a1.n + bi_sn.n + bi_c.n + bi_j0.n) * 4 +
           (loff.n + doff.n + cofs.n + cpos.n + bofs.n + woff.n) * 8 + (diag.n + diagT.n + l21.n + c.n + bpart.n + t.n + y.n + w21.n) * 4 +
           (top_rows.n + top_of.n + top_inc_ptr.n + top_base.n + top_ld.n + top_W.n + top_t.n) * 4 + (top_cpos.n + top_woff.n) * 8;
}

#define HDD_SN_DISPATCH(G, KERNEL, BLOCKS, SHMEM, ...)                                                           \
    switch (G) {                                                                                                \
        case 4: hipLaunchKernelGGL(KERNEL<4>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;              \
        case 8: hipLaunchKernelGGL(KERNEL<8>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;              \
        case 16: hipLaunchKernelGGL(KERNEL<16>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;            \
        case 32: hipLaunchKernelGGL(KERNEL<32>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;            \
        default: hipLaunchKernelGGL(KERNEL<64>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;            \
    }

static constexpr int kRunM = TriPlanHost::kRunMembers;
#define HDD_RUN_DISPATCH(G, M, KERNEL, BLOCKS, STREAM, ...)                                                       \
    do {                                                                                                        \
        switch (G) {                                                                                            \
            case 4: hipLaunchKernelGGL((KERNEL<4, M>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
            case 8: hipLaunchKernelGGL((KERNEL<8, M>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
            case 16: hipLaunchKernelGGL((KERNEL<16, M>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
            case 32: hipLaunchKernelGGL((KERNEL<32, M>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((KERNEL<64, M>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
        }                                                                                                       \
    } while (0)

#define HDD_RUN_WF_DISPATCH(G, M, WF, S64, KERNEL, BLOCKS, STREAM, ...)                                             \
    do {                                                                                                            \
        switch (G) {                                                                                                \
            case 4:                                                                                                 \
                if (S64) hipLaunchKernelGGL((KERNEL<4, M, WF, 64>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); \
                else hipLaunchKernelGGL((KERNEL<4, M, WF, 128>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__);    \
                break;                                                                                              \
            case 8:                                                                                                 \
                if (S64) hipLaunchKernelGGL((KERNEL<8, M, WF, 64>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); \
                else hipLaunchKernelGGL((KERNEL<8, M, WF, 128>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__);    \
                break;                                                                                              \
            case 16: hipLaunchKernelGGL((KERNEL<16, M, WF, 128>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
            case 32: hipLaunchKernelGGL((KERNEL<32, M, WF, 128>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((KERNEL<64, M, WF, 128>), dim3(BLOCKS), dim3(256), 0, STREAM, __VA_ARGS__); break; \
        }                                                                                                           \
    } while (0)
// the fused workgroup kernels: run length (runs of one or of up to kRunM) x block form (L21 or W);
// at 4 and 8 lanes per row, heights whose supernodes have at most 64 rows take the 64-row
// form (half the fixed trip count and registers, the same sums: level-0 Aii 711 -> 691 us)
#define HDD_FUSED_DISPATCH(H, WF, KERNEL, BLOCKS, STREAM, ...)                                                       \
    do {                                                                                                            \
        if (run_max[H] > 1) {                                                                                       \
            if (WF) HDD_RUN_WF_DISPATCH(G[H], kRunM, true, max_sz[H] <= 64, KERNEL, BLOCKS, STREAM, __VA_ARGS__);    \
            else HDD_RUN_WF_DISPATCH(G[H], kRunM, false, max_sz[H] <= 64, KERNEL, BLOCKS, STREAM, __VA_ARGS__);      \
        } else {                                                                                                    \
            if (WF) HDD_RUN_WF_DISPATCH(G[H], 1, true, max_sz[H] <= 64, KERNEL, BLOCKS, STREAM, __VA_ARGS__);        \
            else HDD_RUN_WF_DISPATCH(G[H], 1, false, max_sz[H] <= 64, KERNEL, BLOCKS, STREAM, __VA_ARGS__);          \
        }                                                                                                           \
    } while (0)

#define HDD_ROWS_DISPATCH(G, KERNEL, ROWS, ...)                                                                \
    switch (G) {                                                                                                \
        case 4: hipLaunchKernelGGL(KERNEL<4>, dim3(blocks_for((ROWS) * 4, 256)), dim3(256), 0, __VA_ARGS__); break;   \
        case 8: hipLaunchKernelGGL(KERNEL<8>, dim3(blocks_for((ROWS) * 8, 256)), dim3(256), 0, __VA_ARGS__); break;   \
        case 16: hipLaunchKernelGGL(KERNEL<16>, dim3(blocks_for((ROWS) * 16, 256)), dim3(256), 0, __VA_ARGS__); break; \
        case 32: hipLaunchKernelGGL(KERNEL<32>, dim3(blocks_for((ROWS) * 32, 256)), dim3(256), 0, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERNEL<64>, dim3(blocks_for((ROWS) * 64, 256)), dim3(256), 0, __VA_ARGS__); break; \
    }

// k_tri_fwd_sn_contrib_w with the diagonal step's and the contributions' lanes per row
template <int G1, class... A>
static void launch_fwd_w_g2(int g2, unsigned grid, size_t shmem, hipStream_t s, A... a) {
    switch (g2) {
        case 4: hipLaunchKernelGGL((k_tri_fwd_sn_contrib_w<G1, 4>), dim3(grid), dim3(256), shmem, s, a...); break;
        case 8: hipLaunchKernelGGL((k_tri_fwd_sn_contrib_w<G1, 8>), dim3(grid), dim3(256), shmem, s, a...); break;
        case 16: hipLaunchKernelGGL((k_tri_fwd_sn_contrib_w<G1, 16>), dim3(grid), dim3(256), shmem, s, a...); break;
        case 32: hipLaunchKernelGGL((k_tri_fwd_sn_contrib_w<G1, 32>), dim3(grid), dim3(256), shmem, s, a...); break;
        default: hipLaunchKernelGGL((k_tri_fwd_sn_contrib_w<G1, 64>), dim3(grid), dim3(256), shmem, s, a...); break;
    }
}
template <class... A>
static void launch_fwd_w(int g1, int g2, unsigned grid, size_t shmem, hipStream_t s, A... a) {
    switch (g1) {
        case 4: launch_fwd_w_g2<4>(g2, grid, shmem, s, a...); break;
        case 8: launch_fwd_w_g2<8>(g2, grid, shmem, s, a...); break;
        case 16: launch_fwd_w_g2<16>(g2, grid, shmem, s, a...); break;
        case 32: launch_fwd_w_g2<32>(g2, grid, shmem, s, a...); break;
        default: launch_fwd_w_g2<64>(g2, grid, shmem, s, a...); break;
    }
}

void TriPlan::fwd_height(int h, const float* b, hipStream_t s) {
    float* yf = y.p;
    const int* SP = sn_prob.p;
    {
        if (fused[h]) {
            const int r0 = run_begin[h], nr_ = run_begin[h + 1] - run_begin[h];
            if (nr_ && wave[h]) {
                const int grid = std::min((nr_ + 3) / 4, pgrid[h]);
                if (run_max[h] > 1)
                    HDD_RUN_DISPATCH(G[h], kRunM, k_tri_fwd_wave, grid, s, r0, nr_, (const int*)run_q.p, (const int*)run_n.p,
                                     (const int*)sn_list.p, (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)doff.p,
                                     (const float*)diag.p, (const int*)inc_ptr.p, (const long long*)cpos.p, (const int*)orig_of.p, b,
                                     (const long long*)loff.p, (const long long*)cofs.p, (const float*)l21.p, c.p, yf, SP, active);
                else
                    HDD_RUN_DISPATCH(G[h], 1, k_tri_fwd_wave, grid, s, r0, nr_, (const int*)run_q.p, (const int*)run_n.p,
                                     (const int*)sn_list.p, (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)doff.p,
                                     (const float*)diag.p, (const int*)inc_ptr.p, (const long long*)cpos.p, (const int*)orig_of.p, b,
                                     (const long long*)loff.p, (const long long*)cofs.p, (const float*)l21.p, c.p, yf, SP, active);
            } else if (nr_) {
                const int grid = std::min(nr_, pgrid[h]);
                const bool wf = use_w();
                const long long* blo = wf ? woff.p : loff.p;
                const float* blk = wf ? w21.p : l21.p;
                HDD_FUSED_DISPATCH(h, wf, k_tri_fwd_small, grid, s, r0, nr_, (const int*)run_q.p, (const int*)run_n.p, (const int*)sn_list.p,
                                   (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)doff.p, (const float*)diag.p,
                                   (const int*)inc_ptr.p, (const long long*)cpos.p, (const int*)orig_of.p, b, blo, (const long long*)cofs.p,
                                   blk, c.p, yf, SP, active);
            }
            return;
        }
        const int nrows = row_begin[h + 1] - row_begin[h];
        const int* rows = rowlist.p + row_begin[h];
        if (big[h] && nrows)
            HDD_ROWS_DISPATCH(Gf[h], k_tri_fwd_reduce, nrows, s, rows, nrows, (const int*)inc_ptr.p, (const long long*)cpos.p,
                              (const float*)c.p, (const int*)orig_of.p, b, t.p, (const int*)sn_of.p, SP, active);
        const int q0 = sl_begin[h], nq = sl_begin[h + 1] - sl_begin[h];
        if (big[h] && use_w()) {
            // W form: the slices (y = Linv t) and the contribution items (c = W t) in one launch
            const int i0 = ci_begin[h], ni = ci_begin[h + 1] - ci_begin[h];
            if (nq + ni)
                launch_fwd_w(G[h], Gc[h], (unsigned)(nq + ni), (size_t)max_sz[h] * sizeof(float), s, q0, nq, sl_rows[h], (const int*)sl_sn.p,
                             (const int*)sl_r0.p, (const int*)s0.p, (const int*)sz.p, (const long long*)doff.p, (const float*)diag.p,
                             (const float*)t.p, yf, i0, (const int*)ci_sn.p, (const int*)ci_a0.p, (const int*)ci_a1.p,
                             (const long long*)woff.p, (const long long*)cofs.p, (const float*)w21.p, c.p, SP, active);
            return;
        }
        if (nq)
            HDD_SN_DISPATCH(G[h], k_tri_fwd_sn, nq, (size_t)max_sz[h] * sizeof(float), s, q0, sl_rows[h], (const int*)sl_sn.p, (const int*)sl_r0.p,
                            (const int*)s0.p, (const int*)sz.p, (const long long*)doff.p, (const float*)diag.p,
                            big[h] ? (const float*)t.p : (const float*)nullptr, (const int*)inc_ptr.p, (const long long*)cpos.p,
                            (const float*)c.p, (const int*)orig_of.p, b, yf, SP, active);
        const int i0 = ci_begin[h], ni = ci_begin[h + 1] - ci_begin[h];
        if (ni)
            HDD_SN_DISPATCH(Gc[h], k_tri_fwd_contrib, ni, 0, s, i0, i0 + ni, (const int*)ci_sn.p, (const int*)ci_a0.p, (const int*)ci_a1.p,
                            (const int*)s0.p, (const int*)sz.p, (const long long*)loff.p, (const long long*)cofs.p, (const float*)l21.p,
                            (const float*)yf, c.p, SP, active);
    }
}

void TriPlan::bwd_height(int h, float* x, hipStream_t s) {
    float* yf = y.p;
    float* zb = y.p + n;
    const int* SP = sn_prob.p;
    {
        if (fused[h]) {
            const int r0 = run_begin[h], nr_ = run_begin[h + 1] - run_begin[h];
            if (nr_ && wave[h]) {
                const int grid = std::min((nr_ + 3) / 4, pgrid[h]);
                if (run_max[h] > 1)
                    HDD_RUN_DISPATCH(G[h], kRunM, k_tri_bwd_wave, grid, s, r0, nr_, (const int*)run_q.p, (const int*)run_n.p,
                                     (const int*)sn_list.p, (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const int*)roff.p,
                                     (const int*)srows.p, (const long long*)loff.p, (const float*)l21.p, (const long long*)doff.p,
                                     (const float*)diagT.p, (const int*)orig_of.p, (const float*)yf, zb, x, SP, active);
                else
                    HDD_RUN_DISPATCH(G[h], 1, k_tri_bwd_wave, grid, s, r0, nr_, (const int*)run_q.p, (const int*)run_n.p,
                                     (const int*)sn_list.p, (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const int*)roff.p,
                                     (const int*)srows.p, (const long long*)loff.p, (const float*)l21.p, (const long long*)doff.p,
                                     (const float*)diagT.p, (const int*)orig_of.p, (const float*)yf, zb, x, SP, active);
            } else if (nr_) {
                const int grid = std::min(nr_, pgrid[h]);
                const bool wf = use_w();
                const long long* blo = wf ? woff.p : loff.p;
                const float* blk = wf ? w21.p : l21.p;
                HDD_FUSED_DISPATCH(h, wf, k_tri_bwd_small, grid, s, r0, nr_, (const int*)run_q.p, (const int*)run_n.p, (const int*)sn_list.p,
                                   (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const int*)roff.p, (const int*)srows.p, blo, blk,
                                   (const long long*)doff.p, (const float*)diagT.p, (const int*)orig_of.p, (const float*)yf, zb, x, SP, active);
            }
            return;
        }
        const int i0 = bi_begin[h], ni = bi_begin[h + 1] - bi_begin[h];
        if (big[h] && use_w()) {
            // W form: the slices (Linv^T y into t) and the tiles of W^T z in one launch, then
            // z = t - (the tile partials) row by row
            const int q0 = sl_begin[h], nq = sl_begin[h + 1] - sl_begin[h];
            const int grid = nq + (int)blocks_for(ni, 4);
            if (grid)
                HDD_SN_DISPATCH(G[h], k_tri_bwd_sn_tiles_w, grid, (size_t)max_sz[h] * sizeof(float), s, q0, nq, sl_rows[h], (const int*)sl_sn.p,
                                (const int*)sl_r0.p, (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)doff.p,
                                (const float*)diagT.p, (const float*)yf, t.p, i0, i0 + ni, (const int*)bi_sn.p, (const int*)bi_c.p,
                                (const int*)bi_j0.p, (const long long*)woff.p, (const int*)roff.p, (const int*)srows.p, (const long long*)bofs.p,
                                (const float*)w21.p, (const float*)zb, bpart.p, SP, active);
            const int nrows = row_begin[h + 1] - row_begin[h];
            const int* rows = rowlist.p + row_begin[h];
            if (nrows)
                HDD_ROWS_DISPATCH(Gb[h], k_tri_bwd_reduce_w, nrows, s, rows, nrows, (const int*)sn_of.p, (const int*)s0.p, (const int*)sz.p,
                                  (const int*)nr.p, (const long long*)bofs.p, (const float*)bpart.p, (const float*)t.p, (const int*)orig_of.p,
                                  zb, x, SP, active);
            return;
        }
        if (ni)
            hipLaunchKernelGGL(k_tri_bwd_tiles, dim3(blocks_for(ni, 4)), dim3(256), 0, s, i0, i0 + ni, (const int*)bi_sn.p, (const int*)bi_c.p,
                               (const int*)bi_j0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)loff.p, (const int*)roff.p,
                               (const int*)srows.p, (const long long*)bofs.p, (const float*)l21.p, (const float*)zb, bpart.p, SP, active);
        const int nrows = row_begin[h + 1] - row_begin[h];
        const int* rows = rowlist.p + row_begin[h];
        if (big[h] && nrows)
            HDD_ROWS_DISPATCH(Gb[h], k_tri_bwd_reduce, nrows, s, rows, nrows, (const int*)sn_of.p, (const int*)s0.p, (const int*)sz.p,
                              (const int*)nr.p, (const long long*)bofs.p, (const float*)bpart.p, (const float*)yf, t.p, SP, active);
        const int q0 = sl_begin[h], nq = sl_begin[h + 1] - sl_begin[h];
        if (nq)
            HDD_SN_DISPATCH(G[h], k_tri_bwd_sn, nq, (size_t)max_sz[h] * sizeof(float), s, q0, sl_rows[h], (const int*)sl_sn.p, (const int*)sl_r0.p,
                            (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)doff.p, (const float*)diagT.p,
                            big[h] ? (const float*)t.p : (const float*)nullptr, (const long long*)bofs.p, (const float*)bpart.p,
                            (const int*)orig_of.p, (const float*)yf, zb, x, SP, active);
    }
}

void TriPlan::top_solve(const float* b, float* x, hipStream_t s) {
    float* tt = top_t.p;
    float* ty = top_t.p + top_nq;
    HDD_ROWS_DISPATCH(Gt, k_top_rhs, top_nq, s, top_nq, (const int*)top_rows.p, (const int*)top_inc_ptr.p, (const long long*)top_cpos.p,
                      (const float*)c.p, (const int*)orig_of.p, b, tt, (const int*)sn_of.p, (const int*)sn_prob.p, active);
    hipLaunchKernelGGL(k_top_tri<true>, dim3(blocks_for(top_nq, 4)), dim3(256), 0, s, top_nq, (const int*)top_rows.p, (const int*)top_of.p,
                       (const int*)top_base.p, (const int*)top_ld.p, (const long long*)top_woff.p, (const float*)top_W.p,
                       (const float*)tt, ty, (const int*)orig_of.p, y.p + n, x, (const int*)sn_of.p, (const int*)sn_prob.p, active);
    hipLaunchKernelGGL(k_top_tri<false>, dim3(blocks_for(top_nq, 4)), dim3(256), 0, s, top_nq, (const int*)top_rows.p, (const int*)top_of.p,
                       (const int*)top_base.p, (const int*)top_ld.p, (const long long*)top_woff.p, (const float*)top_W.p,
                       (const float*)ty, (float*)nullptr, (const int*)orig_of.p, y.p + n, x, (const int*)sn_of.p, (const int*)sn_prob.p, active);
}

void TriPlan::solve_once(const float* b, float* x, hipStream_t s) {
    if (n == 0) return;
    for (int h = 0; h <= H; ++h) fwd_height(h, b, s);
    if (top_nq) top_solve(b, x, s);
    for (int h = H; h >= 0; --h) bwd_height(h, x, s);
}

// r = b - (A + eps I) x, accumulated in double (one thread per row, column order)
__global__ void k_refine_residual(int rows, const int* __restrict__ p, const int* __restrict__ ix, const float* __restrict__ v,
                                  double eps, const float* __restrict__ b, const float* __restrict__ x, float* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    double acc = (double)b[i] - eps * (double)x[i];
    for (int e = p[i]; e < p[i + 1]; ++e) acc -= (double)v[e] * (double)x[ix[e]];
    r[i] = (float)acc;
}

__global__ void k_add_to(int n, const float* __restrict__ d, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

void TriPlan::solve(const float* b, float* x, hipStream_t s) {
    solve_once(b, x, s);
    if (n == 0 || refine <= 0) return;
    if (refA.rows != n || ref_r.n < (size_t)n || ref_d.n < (size_t)n) throw std::runtime_error("refined solve: refinement operator not set up");
    for (int k = 0; k < refine; ++k) {
        hipLaunchKernelGGL(k_refine_residual, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, (const int*)refA.p.p, (const int*)refA.ix.p,
                           (const float*)refA.v.p, refine_eps, b, (const float*)x, ref_r.p);
        solve_once(ref_r.p, ref_d.p, s);
        hipLaunchKernelGGL(k_add_to, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, (const float*)ref_d.p, x);
    }
}

// Rebuilds the slice and item lists into the tuner's fixed-capacity device arrays.
void TriPlan::rebuild_lists(hipStream_t s) {
    std::vector<int> a, b_, c_, d, e;
    build_slices_items(h_sn_list, sn_begin, h_sz, h_nr, sl_rows, ci_entries, a, b_, sl_begin, c_, d, e, ci_begin);
    if (a.size() > sl_sn.n || c_.size() > ci_sn.n) throw std::runtime_error("triangular-solve tuner: list capacity exceeded");
    hdd_check(hipMemcpyAsync(sl_sn.p, a.data(), a.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hdd_check(hipMemcpyAsync(sl_r0.p, b_.data(), b_.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hdd_check(hipMemcpyAsync(ci_sn.p, c_.data(), c_.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hdd_check(hipMemcpyAsync(ci_a0.p, d.data(), d.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hdd_check(hipMemcpyAsync(ci_a1.p, e.data(), e.size() * sizeof(int), hipMemcpyHostToDevice, s));
    hdd_check(hipStreamSynchronize(s));  // the host vectors go out of scope
}

// Launch-shape autotuner: for every height that is not done by per-run
// workgroups, the rows per diagonal-step slice and the L21 entries per
// contribution item are chosen by timing the height's forward and backward
// launches on the device (HIP events, the plan's own scratch as input). Neither
// choice changes the order of any sum, so the solve's results do not depend on
// what the timings pick.
int TriPlan::tune(hipStream_t s) {
    if (n < kTuneMinRows) return 0;  // small plans: every launch is latency-bound, little to choose
    // capacity for the largest lists any candidate produces (8-row slices, 1024-entry items)
    {
        const std::vector<int> keep_sl = sl_rows, keep_ci = ci_entries;
        for (int h = 0; h <= H; ++h) { sl_rows[h] = 8; ci_entries[h] = 1024; }
        std::vector<int> a, b_, c_, d, e, sb, cb;
        build_slices_items(h_sn_list, sn_begin, h_sz, h_nr, sl_rows, ci_entries, a, b_, sb, c_, d, e, cb);
        sl_rows = keep_sl; ci_entries = keep_ci;
        sl_sn.alloc(a.size()); sl_r0.alloc(a.size()); ci_sn.alloc(std::max<size_t>(1, c_.size()));
        ci_a0.alloc(std::max<size_t>(1, c_.size())); ci_a1.alloc(std::max<size_t>(1, c_.size()));
        rebuild_lists(s);
    }
    DBuf<float> bx, xx;
    bx.alloc(n); xx.alloc(n);
    hipEvent_t e0, e1;
    hdd_check(hipEventCreate(&e0)); hdd_check(hipEventCreate(&e1));
    auto time_height = [&](int h) {
        constexpr int kReps = 8;
        fwd_height(h, bx.p, s); bwd_height(h, xx.p, s);
        hdd_check(hipEventRecord(e0, s));
        for (int r = 0; r < kReps; ++r) { fwd_height(h, bx.p, s); bwd_height(h, xx.p, s); }
        hdd_check(hipEventRecord(e1, s));
        hdd_check(hipEventSynchronize(e1));
        float ms = 0.f;
        hdd_check(hipEventElapsedTime(&ms, e0, e1));
        return ms / kReps;
    };
    int changed = 0;
    for (int h = 0; h <= H; ++h) {
        if (fused[h]) {
            // a workgroup or a wavefront per run (L21 form): the same bits, whichever is faster
            if (run_begin[h + 1] == run_begin[h]) continue;
            // workgroups of the persistent launch (each loops over runs): a cap, or one per run
            float best = time_height(h);
            const int p0 = pgrid[h];
            for (int cand : {1024, 2048, 4096}) {
                const int keep = pgrid[h];
                pgrid[h] = cand;
                const float t = time_height(h);
                if (t < 0.97f * best) best = t;
                else pgrid[h] = keep;
            }
            changed += pgrid[h] != p0;
            // the wavefront form has the L21 blocks only: not a candidate beside the W form
            if (wave_forced || wform) continue;
            wave[h] = 1;
            const float t_wave = time_height(h);
            if (!(t_wave < 0.97f * best)) wave[h] = 0;
            changed += wave[h];
            continue;
        }
        if (sl_begin[h + 1] == sl_begin[h]) continue;
        const int sl0 = sl_rows[h], ci0 = ci_entries[h];
        float best = time_height(h);
        for (int cand : {8, 16, 32, 64, 128}) {
            if (cand == sl_rows[h] || (cand > 8 && cand / 2 >= max_sz[h])) continue;
            const int keep = sl_rows[h];
            sl_rows[h] = cand;
            rebuild_lists(s);
            const float t = time_height(h);
            if (t < 0.97f * best) best = t;
            else { sl_rows[h] = keep; rebuild_lists(s); }
        }
        for (int cand : {1024, 2048, 4096, 8192, 16384}) {
            if (cand == ci_entries[h] || ci_begin[h + 1] == ci_begin[h]) continue;
            const int keep = ci_entries[h];
            ci_entries[h] = cand;
            rebuild_lists(s);
            const float t = time_height(h);
            if (t < 0.97f * best) best = t;
            else { ci_entries[h] = keep; rebuild_lists(s); }
        }
        changed += (sl_rows[h] != sl0) + (ci_entries[h] != ci0);
    }
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    hdd_check(hipStreamSynchronize(s));
    return changed;
}

SigmaDevice::~SigmaDevice() {}

static uint64_t csr_hash(const Csr& m) {
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&](const void* p, size_t bytes) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < bytes; ++i) { h ^= c[i]; h *= 1099511628211ULL; }
    };
    mix(&m.rows, sizeof(m.rows));
    mix(m.ptr.data(), m.ptr.size() * sizeof(int));
    mix(m.idx.data(), m.idx.size() * sizeof(int));
    mix(m.val.data(), m.val.size() * sizeof(float));
    return h;
}
static bool csr_equal(const Csr& a, const Csr& b) { return a.rows == b.rows && a.ptr == b.ptr && a.idx == b.idx && a.val == b.val; }

// For each matrix, the index of the first matrix equal to it (same rows,
// pattern and values), or its own index.
static std::vector<int> identical_matrices(const std::vector<Csr>& A) {
    const int n = (int)A.size();
    std::vector<uint64_t> h(n);
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < n; ++s) h[s] = csr_hash(A[s]);
    std::vector<int> rep(n);
    std::map<uint64_t, std::vector<int>> seen;
    for (int s = 0; s < n; ++s) {
        rep[s] = s;
        for (int c : seen[h[s]])
            if (csr_equal(A[c], A[s])) { rep[s] = c; break; }
        if (rep[s] == s) seen[h[s]].push_back(s);
    }
    return rep;
}

// 200 sweeps leave l^200 of an eigenmode with eigenvalue l; modes with l^200 < 1e-9 are dropped
static const double kBottomLambdaCut = std::pow(1e-9, 1.0 / 200.0);
static const size_t kBottomBasisBytes = size_t(2) << 30;

// HDD_AUTOTUNE=0 keeps the plans' default launch shapes
static bool autotune() {
    const char* v = std::getenv("HDD_AUTOTUNE");
    return !(v && std::string(v) == "0");
}

static std::string bottom_mode() {
    const char* mode = std::getenv("HDD_SIGMA_BOTTOM");
    const std::string m = mode ? mode : "auto";
    if (m != "loop" && m != "direct" && m != "auto") throw std::runtime_error("HDD_SIGMA_BOTTOM must be loop, direct or auto");
    return m;
}

void SigmaFactors::wait_bottom() const {
    if (bottom_worker.joinable()) bottom_worker.join();
    if (bottom_error) std::rethrow_exception(bottom_error);
}

SigmaFactors::~SigmaFactors() {
    if (bottom_worker.joinable()) bottom_worker.join();
}

void factor_sigma(const SigmaHost& H, SigmaFactors& F, bool wait) {
    auto t0 = std::chrono::steady_clock::now();
    F.lv.assign(H.levels, SigmaFactors::Level());
    // The factorisations are independent: the bottom level's combined system (then its
    // spectrum) and the level-0 interface factor each run on a thread of their own while
    // this thread factors the interior matrices and the other interface matrices. Each
    // factor is computed as it would be alone (the work inside one factor does not depend
    // on what runs beside it).
    const int lb = H.levels - 1;
    const std::string mode = bottom_mode();
    const bool bottom = H.levels > 0 && H.n_if[lb] > 0 && mode != "loop";
    std::exception_ptr err_arr0;
    std::thread t_arr0;
    // the threads use this thread's device (dense fronts of the factorisation may run there)
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    // ... and this thread's OpenMP thread count (a server's set-up worker runs on its share)
    const int omp_threads = omp_get_max_threads();
    auto same_device = [dev, omp_threads] {
        if (dev >= 0) (void)hipSetDevice(dev);
        omp_set_num_threads(omp_threads);
    };
    auto arr_factor = [&](int l) {
        auto tp = std::chrono::steady_clock::now();
        std::vector<Int3> rc(H.if_coords[l].size());
        for (size_t i = 0; i < rc.size(); ++i) rc[i] = {H.if_coords[l][i].x << l, H.if_coords[l][i].y << l, H.if_coords[l][i].z << l};
        factor_spd(H.Arr[l], rc, 1e-5, F.lv[l].Arr);
        setup_phase("Arr factor", l, tp);
    };
    if (bottom)
        F.bottom_worker = std::thread([&H, &F, lb, mode, same_device] {
            try {
                same_device();
                auto tp = std::chrono::steady_clock::now();
                Csr C; std::vector<Int3> xyz;
                combined_bottom(H, lb, C, xyz);
                factor_spd(C, xyz, 1e-5, F.bottom);
                F.bottom_n_int = C.rows - H.n_if[lb];
                setup_phase("combined bottom factor", lb, tp);
                if (mode == "direct") {
                    F.bottom_direct = true;  // no low-rank correction: exact only when M^200 is negligible
                } else {
                    bottom_spectrum(H.Arr[lb], 1e-5, F.bottom, F.bottom_n_int, kBottomLambdaCut, kBottomBasisBytes, F.spec);
                    F.bottom_rho = F.spec.lambda_max; F.bottom_steps = F.spec.steps; F.bottom_seconds = F.spec.seconds;
                    F.bottom_direct = F.spec.resolved;
                    if (!F.spec.resolved) F.spec = BottomSpectrum();
                    setup_phase("bottom spectrum", lb, tp);
                }
                if (!F.bottom_direct) F.bottom = SupernodalFactor();
            } catch (...) {
                F.bottom_error = std::current_exception();
            }
        });
    if (H.levels > 0)
        t_arr0 = std::thread([&] {
            try {
                same_device();
                arr_factor(0);
            } catch (...) {
                err_arr0 = std::current_exception();
            }
        });
    std::exception_ptr err_main;
    try {
        auto tp = t0;
        for (int l = 0; l < H.levels; ++l) {
            SigmaFactors::Level& L = F.lv[l];
            // subdomains of the same shape have identical Aii: factor each distinct matrix once
            L.rep = identical_matrices(H.Aii[l]);
            L.Aii.resize(H.nsub);
            ParallelErrors errs;
#pragma omp parallel for schedule(dynamic)
            for (int s = 0; s < H.nsub; ++s)
                if (H.Aii[l][s].rows > 0 && L.rep[s] == s) errs.run([&] { factor_spd(H.Aii[l][s], H.Aii_coords[l][s], 1e-5, L.Aii[s]); });
            errs.rethrow();
            setup_phase("Aii factors", l, tp);
        }
        for (int l = 1; l < H.levels; ++l) arr_factor(l);
    } catch (...) {
        err_main = std::current_exception();
    }
    if (t_arr0.joinable()) t_arr0.join();
    if (wait || err_main || err_arr0) F.wait_bottom();
    if (err_main) std::rethrow_exception(err_main);
    if (err_arr0) std::rethrow_exception(err_arr0);
    F.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Compares the bottom solve with the reference's 200 sweeps on one random
// right-hand side (level-0 buffers of the bottom level are scratch here).
static double bottom_self_check(SigmaDevice& S, hipStream_t s) {
    SigmaLevel& B = S.lv[S.levels - 1];
    const int n = B.n_if;
    std::vector<float> b(n);
    uint64_t st = 0x243f6a8885a308d3ULL;
    for (int i = 0; i < n; ++i) { st = st * 6364136223846793005ULL + 1442695040888963407ULL; b[i] = (float)((double)(st >> 40) / 16777216.0 - 0.5); }
    hdd_check(hipMemcpyAsync(B.b.p, b.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    S.bottom(s, true);
    std::vector<float> ud(n), ul(n);
    hdd_check(hipMemcpyAsync(ud.data(), B.u.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    launch_fill(B.u.p, 0.f, n, s);
    for (int i = 0; i < 200; ++i) S.smooth(S.levels - 1, s);
    hdd_check(hipMemcpyAsync(ul.data(), B.u.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    launch_fill(B.u.p, 0.f, n, s);
    launch_fill(B.b.p, 0.f, n, s);
    hdd_check(hipStreamSynchronize(s));
    double dmax = 0, umax = 0;
    for (int i = 0; i < n; ++i) { dmax = std::max(dmax, (double)std::fabs(ud[i] - ul[i])); umax = std::max(umax, (double)std::fabs(ul[i])); }
    return umax > 0 ? dmax / umax : dmax;
}

// rows of A appended to M with columns shifted by col_off
static void append_rows(Csr& M, const Csr& A, int col_off) {
    for (int r = 0; r < A.rows; ++r) {
        for (int e = A.ptr[r]; e < A.ptr[r + 1]; ++e) { M.idx.push_back(A.idx[e] + col_off); M.val.push_back(A.val[e]); }
        M.ptr.push_back((int)M.idx.size());
    }
    M.rows += A.rows;
}

std::unique_ptr<SigmaDevice> make_sigma_device(const std::vector<const SigmaHost*>& hosts, const std::vector<const SigmaFactors*>& factors,
                                               int smoothing, int n_mg, hipStream_t stream) {
    const int NP = (int)hosts.size();
    if (NP == 0 || (int)factors.size() != NP) throw std::runtime_error("make_sigma_device: one host operator set and factor set per problem");
    std::unique_ptr<SigmaDevice> S(new SigmaDevice());
    const int NL = hosts[0]->levels;
    for (int p = 0; p < NP; ++p)
        if (hosts[p]->levels != NL || (int)factors[p]->lv.size() != NL)
            throw std::runtime_error("make_sigma_device: every problem of a batch needs the same number of sigma levels");
    S->levels = NL; S->smoothing = smoothing; S->n_mg = n_mg; S->nprob = NP;
    S->lv.resize(NL);
    auto t0 = std::chrono::steady_clock::now();
    auto tp = t0;
    // identical interior matrices across the whole batch share one factor
    std::map<uint64_t, std::vector<std::pair<const Csr*, const SupernodalFactor*>>> canon;
    auto canonical = [&](const Csr& A, const SupernodalFactor* f) {
        auto& bucket = canon[csr_hash(A)];
        for (auto& e : bucket) if (csr_equal(*e.first, A)) return e.second;
        bucket.push_back({&A, f});
        return f;
    };
    std::vector<std::vector<int>> ifoff(NL + 1, std::vector<int>(NP + 1, 0));
    for (int l = 0; l < NL; ++l)
        for (int p = 0; p < NP; ++p) ifoff[l][p + 1] = ifoff[l][p] + hosts[p]->n_if[l];
    for (int l = 0; l < NL; ++l) {
        SigmaLevel& L = S->lv[l];
        L.n_if = ifoff[l][NP];
        Csr air, arr, pr;
        air.cols = L.n_if; arr.cols = L.n_if;
        std::vector<const SupernodalFactor*> fi, fr;
        std::vector<int> oi, orr, pi, prr;
        int n_int = 0;
        for (int p = 0; p < NP; ++p) {
            const SigmaHost& H = *hosts[p];
            const SigmaFactors::Level& F = factors[p]->lv[l];
            for (int s = 0; s < H.nsub; ++s) {
                append_rows(air, H.Air[l][s], ifoff[l][p]);
                if (H.Aii[l][s].rows > 0) {
                    const int rp = F.rep[s];
                    fi.push_back(canonical(H.Aii[l][rp], &F.Aii[rp]));
                    oi.push_back(n_int);
                    pi.push_back(p);
                }
                n_int += H.Aii[l][s].rows;
            }
            append_rows(arr, H.Arr[l], ifoff[l][p]);
            if (H.n_if[l] > 0) { fr.push_back(&F.Arr); orr.push_back(ifoff[l][p]); prr.push_back(p); }
            if (l + 1 < NL) append_rows(pr, H.Prr[l], ifoff[l + 1][p]);
        }
        L.n_int = n_int;
        std::map<const SupernodalFactor*, int> distinct;
        for (auto f : fi) distinct[f] = 1;
        L.n_distinct = (int)distinct.size();
        L.Air.upload(air);
        L.AirT.upload(transpose(air));
        L.Arr.upload(arr);
        if (l + 1 < NL) { pr.cols = ifoff[l + 1][NP]; L.P.upload(pr); L.PT.upload(transpose(pr)); }
        setup_phase("interface operators uploaded", l, tp);
        L.Lii.build(fi, oi, pi);
        setup_phase("Aii plan", l, tp);
        L.Lrr.build(fr, orr, prr);
        setup_phase("Arr plan", l, tp);
        S->dense_top_rows += L.Lii.top_dense_rows + L.Lrr.top_dense_rows;
        L.u.alloc(std::max(1, L.n_if)); L.b.alloc(std::max(1, L.n_if)); L.r.alloc(std::max(1, L.n_if)); L.tmp.alloc(std::max(1, L.n_if));
        L.vb.alloc(std::max(1, L.n_int)); L.vx.alloc(std::max(1, L.n_int));
        S->plan_bytes += L.Air.bytes() + L.AirT.bytes() + L.Arr.bytes() + L.P.bytes() + L.PT.bytes();
    }
    // bottom level: one block of the combined plan and one set of modes per problem
    // bottom factors and spectra still being computed (factor_sigma without wait) are awaited here
    for (const SigmaFactors* f : factors) f->wait_bottom();
    setup_phase("bottom factors and spectra awaited", NL - 1, tp);
    // launch shapes are timed once no set-up work (dense fronts of the bottom factors) shares the GPU
    for (int l = 0; l < NL; ++l) {
        SigmaLevel& L = S->lv[l];
        if (autotune()) S->tuned_choices += L.Lii.tune(stream) + L.Lrr.tune(stream);
        S->plan_bytes += L.Lii.bytes() + L.Lrr.bytes();  // with the tuner's list capacity
        setup_phase("plans tuned", l, tp);
    }
    const int lb = NL - 1;
    bool any_if = false, all_direct = true;
    for (int p = 0; p < NP; ++p) {
        const SigmaFactors& F = *factors[p];
        S->bottom_rho = std::max(S->bottom_rho, F.bottom_rho);
        S->bottom_steps = std::max(S->bottom_steps, F.bottom_steps);
        S->bottom_seconds += F.bottom_seconds;
        if (hosts[p]->n_if[lb] == 0) continue;
        any_if = true;
        all_direct = all_direct && F.bottom_direct;
    }
    if (NL > 0 && any_if && all_direct) {
        const int n = ifoff[lb][NP];
        std::vector<const SupernodalFactor*> fb;
        std::vector<int> ob, pb, pos(n), rprob(n), m0(NP + 1, 0), prow0(NP), pn(NP), mrow0, mlen;
        std::vector<long long> pvofs(NP, 0), mvofs;
        std::vector<float> V, SV;
        std::vector<double> w1, w2;
        const bool keep_sv = NL == 1;
        int boff = 0;
        for (int p = 0; p < NP; ++p) {
            const SigmaFactors& F = *factors[p];
            const int np = hosts[p]->n_if[lb], r0 = ifoff[lb][p];
            prow0[p] = r0; pn[p] = np; pvofs[p] = (long long)V.size();
            const int k = np > 0 ? (int)F.spec.lambda.size() : 0;
            m0[p + 1] = m0[p] + k;
            if (np == 0) continue;
            fb.push_back(&F.bottom); ob.push_back(boff); pb.push_back(p);
            for (int i = 0; i < np; ++i) { pos[r0 + i] = boff + F.bottom_n_int + i; rprob[r0 + i] = p; }
            for (int j = 0; j < k; ++j) {
                mrow0.push_back(r0); mlen.push_back(np); mvofs.push_back((long long)V.size() + (long long)j * np);
                const double lam = F.spec.lambda[j];
                w2.push_back(std::pow(lam, 200.0)); w1.push_back(std::pow(lam, 200.0) / (1.0 - lam));
            }
            V.insert(V.end(), F.spec.V.begin(), F.spec.V.begin() + (size_t)np * k);
            if (keep_sv) SV.insert(SV.end(), F.spec.SV.begin(), F.spec.SV.begin() + (size_t)np * k);
            S->spec_nch = std::max(S->spec_nch, (np + kSpecChunk - 1) / kSpecChunk);
            boff += F.bottom_n_int + np;
        }
        setup_phase("bottom modes gathered", lb, tp);
        S->Lbottom.build(fb, ob, pb);
        setup_phase("bottom plan", lb, tp);
        if (autotune()) S->tuned_choices += S->Lbottom.tune(stream);
        setup_phase("bottom plan tuned", lb, tp);
        S->dense_top_rows += S->Lbottom.top_dense_rows;
        S->bot_rhs.alloc(boff); S->bot_sol.alloc(boff);
        S->bot_pos.upload(pos);
        S->bottom_k = m0[NP];
        if (S->bottom_k > 0) {
            S->row_prob.upload(rprob); S->prob_m0.upload(m0); S->prob_row0.upload(prow0); S->prob_n.upload(pn); S->prob_vofs.upload(pvofs);
            S->mode_row0.upload(mrow0); S->mode_len.upload(mlen); S->mode_vofs.upload(mvofs);
            S->spec_V.upload(V);
            if (keep_sv) S->spec_SV.upload(SV);
            S->spec_w1.upload(w1); S->spec_w2.upload(w2);
            S->spec_part.alloc((size_t)2 * S->bottom_k * S->spec_nch);
            S->spec_coef.alloc(S->bottom_k);
        }
        S->bottom_direct = true;
        S->plan_bytes += S->Lbottom.bytes() + S->spec_V.n * 4 + S->spec_SV.n * 4;
    }
    if (S->bottom_direct) {
        S->bottom_check = bottom_self_check(*S, stream);
        setup_phase("bottom self-check", lb, tp);
        // the reference's sweeps stay in charge if the shortcut disagrees with them
        if (bottom_mode() != "direct" && !(S->bottom_check < 1e-3)) S->bottom_direct = false;
    }
    double fs = 0;
    for (int p = 0; p < NP; ++p) fs += factors[p]->seconds;
    S->factor_seconds = fs + std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (NL > 0) { S->b0 = S->lv[0].b.p; S->u0 = S->lv[0].u.p; }
    return S;
}

// tmp = b + sum_s Air_s^T Aii_s^{-1} Air_s u, then u = Arr^{-1} tmp
void SigmaDevice::smooth(int l, hipStream_t s) {
    SigmaLevel& L = lv[l];
    if (L.n_if == 0) return;
    if (L.n_int) {
        hipLaunchKernelGGL(k_csr_mv, dim3(blocks_for(L.n_int, 256)), dim3(256), 0, s, L.n_int, L.Air.p.p, L.Air.ix.p, L.Air.v.p, L.u.p, L.vb.p);
        L.Lii.solve(L.vb.p, L.vx.p, s);
    }
    hipLaunchKernelGGL(k_schur_rhs, dim3(blocks_for(L.n_if, 256)), dim3(256), 0, s, L.n_if, L.b.p, L.AirT.p.p, L.AirT.ix.p, L.AirT.v.p,
                       (const float*)L.vx.p, (const int*)nullptr, (const int*)nullptr, (const float*)nullptr, (const float*)nullptr, L.tmp.p);
    L.Lrr.solve(L.tmp.p, L.u.p, s);
}

// r = b + sum_s Air_s^T Aii_s^{-1} Air_s u - Arr u
void SigmaDevice::residual(int l, hipStream_t s) {
    SigmaLevel& L = lv[l];
    if (L.n_if == 0) return;
    if (L.n_int) {
        hipLaunchKernelGGL(k_csr_mv, dim3(blocks_for(L.n_int, 256)), dim3(256), 0, s, L.n_int, L.Air.p.p, L.Air.ix.p, L.Air.v.p, L.u.p, L.vb.p);
        L.Lii.solve(L.vb.p, L.vx.p, s);
    }
    hipLaunchKernelGGL(k_schur_rhs, dim3(blocks_for(L.n_if, 256)), dim3(256), 0, s, L.n_if, L.b.p, L.AirT.p.p, L.AirT.ix.p, L.AirT.v.p,
                       (const float*)L.vx.p, (const int*)L.Arr.p.p, (const int*)L.Arr.ix.p, (const float*)L.Arr.v.p, (const float*)L.u.p, L.r.p);
}

// x* from the combined systems [[Aii+eI, Air],[Air^T, Arr+eI]] [y; x*] = [0; b]
// (one block per problem; the interior rows of bot_rhs stay zero), then u = x* + V coef.
void SigmaDevice::bottom(hipStream_t s, bool u_zero) {
    SigmaLevel& B = lv[levels - 1];
    const int n = B.n_if;
    if (!u_zero && bottom_k && !spec_SV.p) throw std::runtime_error("bottom solve from a nonzero start needs S V");
    hipLaunchKernelGGL(k_pos_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, (const int*)bot_pos.p, (const float*)B.b.p, bot_rhs.p);
    Lbottom.solve(bot_rhs.p, bot_sol.p, s);
    if (bottom_k == 0) {
        hipLaunchKernelGGL(k_pos_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, (const int*)bot_pos.p, (const float*)bot_sol.p, B.u.p);
        return;
    }
    hipLaunchKernelGGL(k_spec_dot, dim3(spec_nch, bottom_k), dim3(256), 0, s, spec_nch, (const int*)mode_row0.p, (const int*)mode_len.p,
                       (const long long*)mode_vofs.p, (const float*)spec_V.p, (const float*)spec_SV.p, (const float*)B.b.p,
                       u_zero ? (const float*)nullptr : (const float*)B.u.p, spec_part.p);
    hipLaunchKernelGGL(k_spec_coef, dim3(blocks_for(bottom_k, 64)), dim3(64), 0, s, bottom_k, spec_nch, (const int*)mode_len.p,
                       (const double*)spec_part.p, (const double*)spec_w1.p, (const double*)spec_w2.p, spec_coef.p);
    hipLaunchKernelGGL(k_spec_apply, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, (const int*)row_prob.p, (const int*)prob_row0.p,
                       (const int*)prob_n.p, (const int*)prob_m0.p, (const long long*)prob_vofs.p, (const float*)spec_V.p,
                       (const float*)spec_coef.p, (const float*)bot_sol.p, (const int*)bot_pos.p, B.u.p);
}

void SigmaDevice::vcycle(hipStream_t s, bool first) {
    const int NL = levels;
    for (int l = 0; l + 1 < NL; ++l) {
        for (int i = 0; i < smoothing; ++i) smooth(l, s);
        residual(l, s);
        SigmaLevel& C = lv[l + 1];
        launch_fill(C.u.p, 0.f, C.u.n, s);
        if (C.n_if)
            hipLaunchKernelGGL(k_csr_mv, dim3(blocks_for(C.n_if, 256)), dim3(256), 0, s, C.n_if, lv[l].PT.p.p, lv[l].PT.ix.p, lv[l].PT.v.p,
                               (const float*)lv[l].r.p, C.b.p);
    }
    if (bottom_direct) bottom(s, NL > 1 || first);
    else for (int i = 0; i < 200; ++i) smooth(NL - 1, s);
    for (int l = NL - 2; l >= 0; --l) {
        SigmaLevel& F = lv[l];
        if (F.n_if)
            hipLaunchKernelGGL(k_prolong_add, dim3(blocks_for(F.n_if, 256)), dim3(256), 0, s, F.n_if, F.P.p.p, F.P.ix.p, F.P.v.p,
                               (const float*)lv[l + 1].u.p, F.u.p);
        for (int i = 0; i < smoothing; ++i) smooth(l, s);
    }
}

// Iterative refinement of every exact solve (an accuracy mode; the reference's
// single-precision PARDISO solves have no counterpart): the plans' matrices are
// rebuilt from the host operators in the plans' numbering, problems and
// subdomains in the order make_sigma_device lays them out.
void SigmaDevice::set_refinement(const std::vector<const SigmaHost*>& hosts, int steps) {
    if (steps < 0 || steps > 3) throw std::runtime_error("sigma refinement: steps must be 0..3");
    const int NP = (int)hosts.size();
    if (NP != nprob) throw std::runtime_error("sigma refinement: one host operator set per problem");
    auto block_append = [](Csr& M, const Csr& A, int off) {
        for (int r = 0; r < A.rows; ++r) {
            for (int e = A.ptr[r]; e < A.ptr[r + 1]; ++e) { M.idx.push_back(A.idx[e] + off); M.val.push_back(A.val[e]); }
            M.ptr.push_back((int)M.idx.size());
        }
        M.rows += A.rows;
    };
    if (steps > 0 && refine_steps == 0) {
        for (int l = 0; l < levels; ++l) {
            SigmaLevel& L = lv[l];
            Csr ai, ar;
            for (int p = 0; p < NP; ++p) {
                const SigmaHost& H = *hosts[p];
                for (int s = 0; s < H.nsub; ++s) block_append(ai, H.Aii[l][s], ai.rows);
                block_append(ar, H.Arr[l], ar.rows);
            }
            ai.cols = ai.rows; ar.cols = ar.rows;
            if (ai.rows != L.n_int || ar.rows != L.n_if) throw std::runtime_error("sigma refinement: operator sizes differ from the plans");
            if (L.n_int) { L.Lii.refA.upload(ai); L.Lii.ref_r.alloc(L.n_int); L.Lii.ref_d.alloc(L.n_int); }
            if (L.n_if) { L.Lrr.refA.upload(ar); L.Lrr.ref_r.alloc(L.n_if); L.Lrr.ref_d.alloc(L.n_if); }
        }
        if (bottom_direct && Lbottom.n > 0) {
            const int lb = levels - 1;
            Csr cb;
            for (int p = 0; p < NP; ++p) {
                if (hosts[p]->n_if[lb] == 0) continue;
                Csr C; std::vector<Int3> xyz;
                combined_bottom(*hosts[p], lb, C, xyz);
                block_append(cb, C, cb.rows);
            }
            cb.cols = cb.rows;
            if (cb.rows != Lbottom.n) throw std::runtime_error("sigma refinement: bottom system size differs from its plan");
            Lbottom.refA.upload(cb); Lbottom.ref_r.alloc(cb.rows); Lbottom.ref_d.alloc(cb.rows);
        }
    }
    refine_steps = steps;
    for (auto& L : lv) { L.Lii.refine = steps; L.Lrr.refine = steps; }
    Lbottom.refine = steps;
}

void SigmaDevice::set_active(const int* a) {
    for (auto& L : lv) { L.Lii.active = a; L.Lrr.active = a; }
    Lbottom.active = a;
}

void sigma_vcycles(SigmaDevice& S, hipStream_t s) {
    if (S.levels == 0) return;
    launch_fill(S.u0, 0.f, S.lv[0].u.n, s);
    for (int i = 0; i < S.n_mg; ++i) S.vcycle(s, i == 0);
}

template <class S_>
static void gather_t(SigmaDevice& S, const DevLevel& L0, const int* if_index, const S_* z, const S_* r, hipStream_t s) {
    if (S.levels == 0 || S.lv[0].n_if == 0) return;
    launch_if_gather(L0, if_index, S.lv[0].n_if, z, r, S.b0, s);
}
template <class S_>
static void scatter_t(SigmaDevice& S, const int* if_index, S_* z, hipStream_t s) {
    if (S.levels == 0 || S.lv[0].n_if == 0) return;
    launch_if_scatter(if_index, S.lv[0].n_if, S.u0, z, s);
}
void launch_interface_gather(SigmaDevice& S, const DevLevel& L0, const int* if_index, const float* z, const float* r, hipStream_t s) {
    gather_t(S, L0, if_index, z, r, s);
}
void launch_interface_gather(SigmaDevice& S, const DevLevel& L0, const int* if_index, const __half* z, const __half* r, hipStream_t s) {
    gather_t(S, L0, if_index, z, r, s);
}
void launch_interface_scatter(SigmaDevice& S, const int* if_index, float* z, hipStream_t s) { scatter_t(S, if_index, z, s); }
void launch_interface_scatter(SigmaDevice& S, const int* if_index, __half* z, hipStream_t s) { scatter_t(S, if_index, z, s); }

void sigma_vcycle_host(SigmaDevice& S, const float* b, float* u, hipStream_t s) {
    if (S.levels == 0) throw std::runtime_error("no sigma levels");
    const size_t n = S.lv[0].n_if;
    hdd_check(hipMemcpyAsync(S.b0, b, n * sizeof(float), hipMemcpyHostToDevice, s));
    sigma_vcycles(S, s);
    hdd_check(hipMemcpyAsync(u, S.u0, n * sizeof(float), hipMemcpyDeviceToHost, s));
    hdd_check(hipStreamSynchronize(s));
}

void sigma_smooth_host(SigmaDevice& S, int l, const float* b, float* u, hipStream_t s) {
    SigmaLevel& L = S.lv.at(l);
    const size_t n = L.n_if;
    hdd_check(hipMemcpyAsync(L.b.p, b, n * sizeof(float), hipMemcpyHostToDevice, s));
    hdd_check(hipMemcpyAsync(L.u.p, u, n * sizeof(float), hipMemcpyHostToDevice, s));
    S.smooth(l, s);
    hdd_check(hipMemcpyAsync(u, L.u.p, n * sizeof(float), hipMemcpyDeviceToHost, s));
    hdd_check(hipStreamSynchronize(s));
}

}  // namespace hdd
