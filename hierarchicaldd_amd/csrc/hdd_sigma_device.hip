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
#include <algorithm>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <string>
#include <map>
#include <omp.h>
#include "hdd_factor.h"
#include "hdd_parallel.h"
#include "hdd_runtime.h"
#include "hdd_sigma_device.h"
#include "hdd_spectral.h"

namespace hdd {

namespace {

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_down(v, o, G);
    return v;
}

// forward reduce (heights with big supernodes): t_i = b_i - sum of the
// contributions at row i's incidences, G lanes per row
template <int G>
__global__ void __launch_bounds__(256) k_tri_fwd_reduce(const int* __restrict__ rows, int nrows, const int* __restrict__ inc_ptr,
                                                        const long long* __restrict__ cpos, const float* __restrict__ c,
                                                        const int* __restrict__ orig_of, const float* __restrict__ b, float* __restrict__ t) {
    const int w = (blockIdx.x * 256 + threadIdx.x) / G, lane = threadIdx.x & (G - 1);
    if (w >= nrows) return;
    const int i = rows[w];
    float s = 0.f;
    for (int p = inc_ptr[i] + lane; p < inc_ptr[i + 1]; p += G) s += c[cpos[p]];
    s = group_sum<G>(s);
    if (lane == 0) t[i] = b[orig_of[i]] - s;
}

// forward diagonal step: a workgroup per row slice [r0, r0 + slice_rows) of a
// supernode. t of rows [0, rend) into LDS (from the reduce kernel's output on
// big heights, else summed here), then y[r] = sum_{k<=r} Linv(r,k) t[k] with
// G lanes per row.
template <int G>
__global__ void __launch_bounds__(256) k_tri_fwd_sn(int q0, int slice_rows, const int* __restrict__ sl_sn, const int* __restrict__ sl_r0,
                                                    const int* __restrict__ s0, const int* __restrict__ sz,
                                                    const long long* __restrict__ doff, const float* __restrict__ diag,
                                                    const float* __restrict__ tg, const int* __restrict__ inc_ptr,
                                                    const long long* __restrict__ cpos, const float* __restrict__ c,
                                                    const int* __restrict__ orig_of, const float* __restrict__ b, float* __restrict__ y) {
    extern __shared__ float ts[];
    const int g = sl_sn[q0 + blockIdx.x], r0 = sl_r0[q0 + blockIdx.x], n = sz[g], base = s0[g];
    const int rend = min(n, r0 + slice_rows);
    for (int k = threadIdx.x; k < rend; k += 256) {
        const int i = base + k;
        if (tg) {
            ts[k] = tg[i];
        } else {
            float sum = 0.f;
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
        for (int k = lane; k <= r; k += G) acc += Dr[k] * ts[k];
        acc = group_sum<G>(acc);
        if (lane == 0) y[base + r] = acc;
    }
}

// contributions c_g[a] = L21_g[a, :] . y_g for the rows [a0, a1) of an item, G lanes per row
template <int G>
__global__ void __launch_bounds__(256) k_tri_fwd_contrib(int i0, int i1, const int* __restrict__ ci_sn, const int* __restrict__ ci_a0,
                                                         const int* __restrict__ ci_a1, const int* __restrict__ s0, const int* __restrict__ sz,
                                                         const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                         const float* __restrict__ l21, const float* __restrict__ y, float* __restrict__ c) {
    // a workgroup per item; its rows go to the 256/G lane groups in turn
    const int it = i0 + blockIdx.x;
    if (it >= i1) return;
    const int g = ci_sn[it], n = sz[g], a1 = ci_a1[it];
    const float* L = l21 + loff[g];
    const float* yy = y + s0[g];
    float* cg = c + cofs[g];
    const int lane = threadIdx.x & (G - 1);
    for (int a = ci_a0[it] + (int)threadIdx.x / G; a < a1; a += 256 / G) {
        const float* La = L + (long long)a * n;
        float acc = 0.f;
        for (int j = lane; j < n; j += G) acc += La[j] * yy[j];
        acc = group_sum<G>(acc);
        if (lane == 0) cg[a] = acc;
    }
}

// backward tiles: one wavefront per (supernode g, row chunk cc, column block j0);
// lane j accumulates L21_g[a, j] z[rows_g[a]] over the chunk's rows (z: the
// ancestors' backward results, final)
__global__ void __launch_bounds__(256) k_tri_bwd_tiles(int i0, int i1, const int* __restrict__ bi_sn, const int* __restrict__ bi_c,
                                                       const int* __restrict__ bi_j0, const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const long long* __restrict__ loff, const int* __restrict__ roff,
                                                       const int* __restrict__ srows, const long long* __restrict__ bofs,
                                                       const float* __restrict__ l21, const float* __restrict__ z, float* __restrict__ bpart) {
    const int it = i0 + blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (it >= i1) return;
    const int g = bi_sn[it], n = sz[g], cc = bi_c[it], j = bi_j0[it] + lane;
    const int a0 = cc * TriPlanHost::kTile, m = min(nr[g] - a0, TriPlanHost::kTile);
    // lane q holds z at the tile's row q; the row loop then issues only independent L21 loads
    const float zq = lane < m ? z[srows[roff[g] + a0 + lane]] : 0.f;
    const float* L = l21 + loff[g] + (long long)a0 * n + min(j, n - 1);
    float acc = 0.f;
#pragma unroll 8
    for (int q = 0; q < m; ++q) acc += L[(long long)q * n] * __shfl(zq, q);
    if (j < n) bpart[bofs[g] + (long long)cc * n + j] = acc;
}

// backward reduce (heights with big supernodes): t_j = y_j - sum over the row
// chunks of the tile partials of column j of its supernode, G lanes per row
template <int G>
__global__ void __launch_bounds__(256) k_tri_bwd_reduce(const int* __restrict__ rows, int nrows, const int* __restrict__ sn_of,
                                                        const int* __restrict__ s0, const int* __restrict__ sz, const int* __restrict__ nr,
                                                        const long long* __restrict__ bofs, const float* __restrict__ bpart,
                                                        const float* __restrict__ y, float* __restrict__ t) {
    const int w = (blockIdx.x * 256 + threadIdx.x) / G, lane = threadIdx.x & (G - 1);
    if (w >= nrows) return;
    const int i = rows[w], g = sn_of[i], n = sz[g], k = i - s0[g];
    const int nch = (nr[g] + TriPlanHost::kTile - 1) / TriPlanHost::kTile;
    const float* P = bpart + bofs[g] + k;
    float s = 0.f;
    for (int cc = lane; cc < nch; cc += G) s += P[(long long)cc * n];
    s = group_sum<G>(s);
    if (lane == 0) t[i] = y[i] - s;
}

// backward diagonal step: t of rows [r0, n) into LDS (reduce output on big
// heights, else summed here), then z[r] = sum_{k>=r} Linv(k,r) t[k]; y holds
// the forward result and z the backward one (separate arrays: the slices of
// one supernode read rows that other slices write); also writes x in the
// original numbering
template <int G>
__global__ void __launch_bounds__(256) k_tri_bwd_sn(int q0, int slice_rows, const int* __restrict__ sl_sn, const int* __restrict__ sl_r0,
                                                    const int* __restrict__ s0, const int* __restrict__ sz, const int* __restrict__ nr,
                                                    const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                    const float* __restrict__ tg, const long long* __restrict__ bofs,
                                                    const float* __restrict__ bpart, const int* __restrict__ orig_of,
                                                    const float* __restrict__ y, float* __restrict__ z, float* __restrict__ x) {
    extern __shared__ float ts[];
    const int g = sl_sn[q0 + blockIdx.x], r0 = sl_r0[q0 + blockIdx.x], n = sz[g], base = s0[g];
    const int rend = min(n, r0 + slice_rows);
    const int nch = (nr[g] + TriPlanHost::kTile - 1) / TriPlanHost::kTile;
    for (int k = r0 + (int)threadIdx.x; k < n; k += 256) {
        if (tg) {
            ts[k] = tg[base + k];
        } else {
            const float* P = bpart + bofs[g] + k;
            float sum = 0.f;
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
        for (int k = r + lane; k < n; k += G) acc += Dr[k] * ts[k];
        acc = group_sum<G>(acc);
        if (lane == 0) { z[base + r] = acc; x[orig_of[base + r]] = acc; }
    }
}

// Heights whose supernodes all have <= kSliceRows (128) rows: one workgroup per
// supernode does the whole height's work for it in one launch.
// forward: t = b - (incidence sums) -> y = Linv t (LDS) -> c_g = L21_g y
template <int G>
__global__ void __launch_bounds__(256) k_tri_fwd_small(int q0, const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                       const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const long long* __restrict__ doff, const float* __restrict__ diag,
                                                       const int* __restrict__ inc_ptr, const long long* __restrict__ cpos,
                                                       const int* __restrict__ orig_of, const float* __restrict__ b,
                                                       const long long* __restrict__ loff, const long long* __restrict__ cofs,
                                                       const float* __restrict__ l21, float* __restrict__ c, float* __restrict__ y) {
    __shared__ float ts[TriPlanHost::kSliceRows], ys[TriPlanHost::kSliceRows];
    const int g = sn_list[q0 + blockIdx.x], n = sz[g], base = s0[g], m = nr[g];
    for (int k = threadIdx.x; k < n; k += 256) {
        const int i = base + k;
        float sum = 0.f;
        for (int p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) sum += c[cpos[p]];
        ts[k] = b[orig_of[i]] - sum;
    }
    __syncthreads();
    const float* D = diag + doff[g];
    const int lane = threadIdx.x & (G - 1);
    constexpr int kU = TriPlanHost::kSliceRows / G;  // fixed trip count: every load of a row is issued at once
    for (int r = (int)threadIdx.x / G; r < n; r += 256 / G) {
        const float* Dr = D + (long long)r * n;
        float acc = 0.f;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = lane + u * G;
            if (k <= r) acc += Dr[k] * ts[k];
        }
        acc = group_sum<G>(acc);
        if (lane == 0) { ys[r] = acc; y[base + r] = acc; }
    }
    __syncthreads();
    const float* L = l21 + loff[g];
    float* cg = c + cofs[g];
    for (int a = (int)threadIdx.x / G; a < m; a += 256 / G) {
        const float* La = L + (long long)a * n;
        float acc = 0.f;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int j = lane + u * G;
            if (j < n) acc += La[j] * ys[j];
        }
        acc = group_sum<G>(acc);
        if (lane == 0) cg[a] = acc;
    }
}

// backward: t_j = y_j - sum_a L21_g[a, j] z[rows_g[a]] (four wavefronts split
// the rows, lanes own columns j and j + 64, LDS reduction) -> z = Linv^T t
template <int G>
__global__ void __launch_bounds__(256) k_tri_bwd_small(int q0, const int* __restrict__ sn_list, const int* __restrict__ s0,
                                                       const int* __restrict__ sz, const int* __restrict__ nr,
                                                       const int* __restrict__ roff, const int* __restrict__ srows,
                                                       const long long* __restrict__ loff, const float* __restrict__ l21,
                                                       const long long* __restrict__ doff, const float* __restrict__ diagT,
                                                       const int* __restrict__ orig_of, const float* __restrict__ y,
                                                       float* __restrict__ z, float* __restrict__ x) {
    constexpr int kChunk = 1024;
    __shared__ float zr[kChunk];
    __shared__ float part[4][TriPlanHost::kSliceRows];
    __shared__ float ts[TriPlanHost::kSliceRows];
    const int g = sn_list[q0 + blockIdx.x], n = sz[g], base = s0[g], m = nr[g];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int* R = srows + roff[g];
    const float* L = l21 + loff[g];
    const int j0 = min(lane, n - 1), j1 = min(lane + 64, n - 1);
    float acc0 = 0.f, acc1 = 0.f;
    for (int a0 = 0; a0 < m; a0 += kChunk) {
        const int cm = min(kChunk, m - a0);
        for (int q = threadIdx.x; q < cm; q += 256) zr[q] = z[R[a0 + q]];
        __syncthreads();
#pragma unroll 4
        for (int q = w; q < cm; q += 4) {
            const float* La = L + (long long)(a0 + q) * n;
            const float zq = zr[q];
            acc0 += La[j0] * zq;
            acc1 += La[j1] * zq;
        }
        __syncthreads();
    }
    if (lane < n) part[w][lane] = acc0;
    if (lane + 64 < n) part[w][lane + 64] = acc1;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += 256)
        ts[k] = y[base + k] - ((part[0][k] + part[1][k]) + (part[2][k] + part[3][k]));
    __syncthreads();
    const float* DT = diagT + doff[g];
    const int gl = threadIdx.x & (G - 1);
    constexpr int kU = TriPlanHost::kSliceRows / G;
    for (int r = (int)threadIdx.x / G; r < n; r += 256 / G) {
        const float* Dr = DT + (long long)r * n;
        float acc = 0.f;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int k = r + gl + u * G;
            if (k < n) acc += Dr[k] * ts[k];
        }
        acc = group_sum<G>(acc);
        if (gl == 0) { z[base + r] = acc; x[orig_of[base + r]] = acc; }
    }
}

// y = A x for CSR A (one thread per row, column order)
__global__ void k_csr_mv(int rows, const int* __restrict__ p, const int* __restrict__ ix, const float* __restrict__ v,
                         const float* __restrict__ x, float* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float acc = 0.f;
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
    for (int e = tp[c]; e < tp[c + 1]; ++e) acc += tv[e] * vx[ti[e]];
    if (ap)
        for (int e = ap[c]; e < ap[c + 1]; ++e) acc += av[e] * (-u[ai[e]]);
    out[c] = acc;
}

// u[i] += sum_c P(i,c) uc[c]
__global__ void k_prolong_add(int rows, const int* __restrict__ p, const int* __restrict__ ix, const float* __restrict__ v,
                              const float* __restrict__ uc, float* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float acc = 0.f;
    for (int e = p[i]; e < p[i + 1]; ++e) acc += v[e] * uc[ix[e]];
    u[i] += acc;
}

// Low-rank part of the bottom solve: part[j][c] = (V_j . b, (S V)_j . u) over
// rows [c*kSpecChunk, (c+1)*kSpecChunk), in double; u == nullptr means u = 0.
constexpr int kSpecChunk = 4096;
__global__ void __launch_bounds__(256) k_spec_dot(int n, int nch, const float* __restrict__ V, const float* __restrict__ SV,
                                                  const float* __restrict__ b, const float* __restrict__ u, double* __restrict__ part) {
    const int c = blockIdx.x, j = blockIdx.y;
    const int r0 = c * kSpecChunk, r1 = min(n, r0 + kSpecChunk);
    const float* vj = V + (size_t)j * n;
    double sb = 0.0, su = 0.0;
    for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) sb += (double)vj[i] * (double)b[i];
    if (u) {
        const float* sj = SV + (size_t)j * n;
        for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) su += (double)sj[i] * (double)u[i];
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
__global__ void k_spec_coef(int k, int nch, const double* __restrict__ part, const double* __restrict__ w1, const double* __restrict__ w2,
                            float* __restrict__ coef) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    double sb = 0.0, su = 0.0;
    for (int c = 0; c < nch; ++c) { sb += part[((size_t)j * nch + c) * 2]; su += part[((size_t)j * nch + c) * 2 + 1]; }
    coef[j] = (float)(w2[j] * su - w1[j] * sb);
}

// u = x* + V coef
__global__ void k_spec_apply(int n, int k, const float* __restrict__ V, const float* __restrict__ coef, const float* __restrict__ xs,
                             float* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = xs[i];
    for (int j = 0; j < k; ++j) acc += V[(size_t)j * n + i] * coef[j];
    u[i] = acc;
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

void TriPlan::build(const std::vector<const SupernodalFactor*>& F, const std::vector<int>& offsets) {
    TriPlanHost P;
    build_triplan(F, offsets, P);
    n = P.n; H = P.H; factor_nnz = P.factor_nnz;
    if (n == 0) return;
    sn_begin = P.sn_begin; sl_begin = P.sl_begin; sl_rows = P.sl_rows; row_begin = P.row_begin; ci_begin = P.ci_begin; bi_begin = P.bi_begin;
    max_sz = P.max_sz; big = P.big; fused = P.fused; G = P.G; Gc = P.Gc; Gf = P.Gf; Gb = P.Gb;
    orig_of.upload(P.orig_of); s0.upload(P.s0); sz.upload(P.sz); nr.upload(P.nr); roff.upload(P.roff); srows.upload(P.srows);
    sn_of.upload(P.sn_of); sn_list.upload(P.sn_list); sl_sn.upload(P.sl_sn); sl_r0.upload(P.sl_r0); rowlist.upload(P.rowlist); inc_ptr.upload(P.inc_ptr);
    ci_sn.upload(P.ci_sn); ci_a0.upload(P.ci_a0); ci_a1.upload(P.ci_a1); bi_sn.upload(P.bi_sn); bi_c.upload(P.bi_c); bi_j0.upload(P.bi_j0);
    loff.upload(P.loff); doff.upload(P.doff); cofs.upload(P.cofs); cpos.upload(P.cpos); bofs.upload(P.bofs);
    diag.upload(P.diag); diagT.upload(P.diagT); l21.upload(P.l21);
    c.alloc(std::max<long long>(1, P.c_size)); bpart.alloc(std::max<long long>(1, P.bpart_size));
    t.alloc(n); y.alloc(2 * (size_t)n);  // forward result, then backward result
}

size_t TriPlan::bytes() const {
    return (orig_of.n + s0.n + sz.n + nr.n + roff.n + srows.n + sn_of.n + sn_list.n + sl_sn.n + sl_r0.n + rowlist.n + inc_ptr.n + ci_sn.n + ci_a0.n +
            ci_a1.n + bi_sn.n + bi_c.n + bi_j0.n) * 4 +
           (loff.n + doff.n + cofs.n + cpos.n + bofs.n) * 8 + (diag.n + diagT.n + l21.n + c.n + bpart.n + t.n + y.n) * 4;
}

#define HDD_SN_DISPATCH(G, KERNEL, BLOCKS, SHMEM, ...)                                                           \
    switch (G) {                                                                                                \
        case 4: hipLaunchKernelGGL(KERNEL<4>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;              \
        case 8: hipLaunchKernelGGL(KERNEL<8>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;              \
        case 16: hipLaunchKernelGGL(KERNEL<16>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;            \
        case 32: hipLaunchKernelGGL(KERNEL<32>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;            \
        default: hipLaunchKernelGGL(KERNEL<64>, dim3(BLOCKS), dim3(256), SHMEM, __VA_ARGS__); break;            \
    }

#define HDD_ROWS_DISPATCH(G, KERNEL, ROWS, ...)                                                                \
    switch (G) {                                                                                                \
        case 4: hipLaunchKernelGGL(KERNEL<4>, dim3(blocks_for((ROWS) * 4, 256)), dim3(256), 0, __VA_ARGS__); break;   \
        case 8: hipLaunchKernelGGL(KERNEL<8>, dim3(blocks_for((ROWS) * 8, 256)), dim3(256), 0, __VA_ARGS__); break;   \
        case 16: hipLaunchKernelGGL(KERNEL<16>, dim3(blocks_for((ROWS) * 16, 256)), dim3(256), 0, __VA_ARGS__); break; \
        case 32: hipLaunchKernelGGL(KERNEL<32>, dim3(blocks_for((ROWS) * 32, 256)), dim3(256), 0, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERNEL<64>, dim3(blocks_for((ROWS) * 64, 256)), dim3(256), 0, __VA_ARGS__); break; \
    }

void TriPlan::solve(const float* b, float* x, hipStream_t s) {
    if (n == 0) return;
    float* yf = y.p;
    float* zb = y.p + n;
    for (int h = 0; h <= H; ++h) {
        if (fused[h]) {
            const int q0 = sn_begin[h], nq = sn_begin[h + 1] - sn_begin[h];
            if (nq)
                HDD_SN_DISPATCH(G[h], k_tri_fwd_small, nq, 0, s, q0, (const int*)sn_list.p, (const int*)s0.p, (const int*)sz.p, (const int*)nr.p,
                                (const long long*)doff.p, (const float*)diag.p, (const int*)inc_ptr.p, (const long long*)cpos.p,
                                (const int*)orig_of.p, b, (const long long*)loff.p, (const long long*)cofs.p, (const float*)l21.p, c.p, yf);
            continue;
        }
        const int nrows = row_begin[h + 1] - row_begin[h];
        const int* rows = rowlist.p + row_begin[h];
        if (big[h] && nrows)
            HDD_ROWS_DISPATCH(Gf[h], k_tri_fwd_reduce, nrows, s, rows, nrows, (const int*)inc_ptr.p, (const long long*)cpos.p,
                              (const float*)c.p, (const int*)orig_of.p, b, t.p);
        const int q0 = sl_begin[h], nq = sl_begin[h + 1] - sl_begin[h];
        if (nq)
            HDD_SN_DISPATCH(G[h], k_tri_fwd_sn, nq, (size_t)max_sz[h] * sizeof(float), s, q0, sl_rows[h], (const int*)sl_sn.p, (const int*)sl_r0.p,
                            (const int*)s0.p, (const int*)sz.p, (const long long*)doff.p, (const float*)diag.p,
                            big[h] ? (const float*)t.p : (const float*)nullptr, (const int*)inc_ptr.p, (const long long*)cpos.p,
                            (const float*)c.p, (const int*)orig_of.p, b, yf);
        const int i0 = ci_begin[h], ni = ci_begin[h + 1] - ci_begin[h];
        if (ni)
            HDD_SN_DISPATCH(Gc[h], k_tri_fwd_contrib, ni, 0, s, i0, i0 + ni, (const int*)ci_sn.p, (const int*)ci_a0.p, (const int*)ci_a1.p,
                            (const int*)s0.p, (const int*)sz.p, (const long long*)loff.p, (const long long*)cofs.p, (const float*)l21.p,
                            (const float*)yf, c.p);
    }
    for (int h = H; h >= 0; --h) {
        if (fused[h]) {
            const int q0 = sn_begin[h], nq = sn_begin[h + 1] - sn_begin[h];
            if (nq)
                HDD_SN_DISPATCH(G[h], k_tri_bwd_small, nq, 0, s, q0, (const int*)sn_list.p, (const int*)s0.p, (const int*)sz.p, (const int*)nr.p,
                                (const int*)roff.p, (const int*)srows.p, (const long long*)loff.p, (const float*)l21.p, (const long long*)doff.p,
                                (const float*)diagT.p, (const int*)orig_of.p, (const float*)yf, zb, x);
            continue;
        }
        const int i0 = bi_begin[h], ni = bi_begin[h + 1] - bi_begin[h];
        if (ni)
            hipLaunchKernelGGL(k_tri_bwd_tiles, dim3(blocks_for(ni, 4)), dim3(256), 0, s, i0, i0 + ni, (const int*)bi_sn.p, (const int*)bi_c.p,
                               (const int*)bi_j0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)loff.p, (const int*)roff.p,
                               (const int*)srows.p, (const long long*)bofs.p, (const float*)l21.p, (const float*)zb, bpart.p);
        const int nrows = row_begin[h + 1] - row_begin[h];
        const int* rows = rowlist.p + row_begin[h];
        if (big[h] && nrows)
            HDD_ROWS_DISPATCH(Gb[h], k_tri_bwd_reduce, nrows, s, rows, nrows, (const int*)sn_of.p, (const int*)s0.p, (const int*)sz.p,
                              (const int*)nr.p, (const long long*)bofs.p, (const float*)bpart.p, (const float*)yf, t.p);
        const int q0 = sl_begin[h], nq = sl_begin[h + 1] - sl_begin[h];
        if (nq)
            HDD_SN_DISPATCH(G[h], k_tri_bwd_sn, nq, (size_t)max_sz[h] * sizeof(float), s, q0, sl_rows[h], (const int*)sl_sn.p, (const int*)sl_r0.p,
                            (const int*)s0.p, (const int*)sz.p, (const int*)nr.p, (const long long*)doff.p, (const float*)diagT.p,
                            big[h] ? (const float*)t.p : (const float*)nullptr, (const long long*)bofs.p, (const float*)bpart.p,
                            (const int*)orig_of.p, (const float*)yf, zb, x);
    }
}

SigmaDevice::~SigmaDevice() {}

// For each matrix, the index of the first matrix equal to it (same rows,
// pattern and values), or its own index.
static std::vector<int> identical_matrices(const std::vector<Csr>& A) {
    auto hash = [](const Csr& m) {
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
    };
    const int n = (int)A.size();
    std::vector<uint64_t> h(n);
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < n; ++s) h[s] = hash(A[s]);
    std::vector<int> rep(n);
    std::map<uint64_t, std::vector<int>> seen;
    for (int s = 0; s < n; ++s) {
        rep[s] = s;
        for (int c : seen[h[s]])
            if (A[c].rows == A[s].rows && A[c].ptr == A[s].ptr && A[c].idx == A[s].idx && A[c].val == A[s].val) { rep[s] = c; break; }
        if (rep[s] == s) seen[h[s]].push_back(s);
    }
    return rep;
}

// 200 sweeps leave l^200 of an eigenmode with eigenvalue l; modes with l^200 < 1e-9 are dropped
static const double kBottomLambdaCut = std::pow(1e-9, 1.0 / 200.0);
static const size_t kBottomBasisBytes = size_t(2) << 30;

static void upload_spectrum(SigmaDevice& S, const BottomSpectrum& sp, int n, bool keep_sv) {
    const int k = (int)sp.lambda.size();
    S.bottom_k = k;
    if (k == 0) return;
    std::vector<float> v(sp.V.begin(), sp.V.end());
    S.spec_V.upload(v);
    if (keep_sv) { std::vector<float> sv(sp.SV.begin(), sp.SV.end()); S.spec_SV.upload(sv); }
    std::vector<double> w1(k), w2(k);
    for (int j = 0; j < k; ++j) { w2[j] = std::pow(sp.lambda[j], 200.0); w1[j] = w2[j] / (1.0 - sp.lambda[j]); }
    S.spec_w1.upload(w1); S.spec_w2.upload(w2);
    S.spec_nch = (n + kSpecChunk - 1) / kSpecChunk;
    S.spec_part.alloc((size_t)2 * k * S.spec_nch);
    S.spec_coef.alloc(k);
    S.plan_bytes += S.spec_V.n * 4 + S.spec_SV.n * 4;
}

// Compares the bottom solve with the reference's 200 sweeps on one random
// right-hand side (level-0 buffers of the bottom level are scratch here).
static double bottom_self_check(SigmaDevice& S) {
    SigmaLevel& B = S.lv[S.levels - 1];
    const int n = B.n_if;
    std::vector<float> b(n);
    uint64_t st = 0x243f6a8885a308d3ULL;
    for (int i = 0; i < n; ++i) { st = st * 6364136223846793005ULL + 1442695040888963407ULL; b[i] = (float)((double)(st >> 40) / 16777216.0 - 0.5); }
    hdd_check(hipMemcpy(B.b.p, b.data(), n * sizeof(float), hipMemcpyHostToDevice));
    S.bottom(0, true);
    std::vector<float> ud(n), ul(n);
    hdd_check(hipMemcpy(ud.data(), B.u.p, n * sizeof(float), hipMemcpyDeviceToHost));
    launch_fill(B.u.p, 0.f, n, 0);
    for (int i = 0; i < 200; ++i) S.smooth(S.levels - 1, 0);
    hdd_check(hipMemcpy(ul.data(), B.u.p, n * sizeof(float), hipMemcpyDeviceToHost));
    double dmax = 0, umax = 0;
    for (int i = 0; i < n; ++i) { dmax = std::max(dmax, (double)std::fabs(ud[i] - ul[i])); umax = std::max(umax, (double)std::fabs(ul[i])); }
    launch_fill(B.u.p, 0.f, n, 0);
    launch_fill(B.b.p, 0.f, n, 0);
    hdd_check(hipDeviceSynchronize());
    return umax > 0 ? dmax / umax : dmax;
}

std::unique_ptr<SigmaDevice> make_sigma_device(const SigmaHost& H, int smoothing, int n_mg) {
    std::unique_ptr<SigmaDevice> S(new SigmaDevice());
    S->levels = H.levels; S->smoothing = smoothing; S->n_mg = n_mg;
    S->lv.resize(H.levels);
    auto t0 = std::chrono::steady_clock::now();
    for (int l = 0; l < H.levels; ++l) {
        SigmaLevel& L = S->lv[l];
        L.n_if = H.n_if[l];
        // concatenate the subdomains' interior systems
        std::vector<int> off(H.nsub + 1, 0);
        for (int s = 0; s < H.nsub; ++s) off[s + 1] = off[s] + H.Aii[l][s].rows;
        L.n_int = off[H.nsub];
        Csr air; air.rows = L.n_int; air.cols = L.n_if; air.ptr.assign(1, 0);
        for (int s = 0; s < H.nsub; ++s) {
            const Csr& A = H.Air[l][s];
            for (int r = 0; r < A.rows; ++r) {
                for (int e = A.ptr[r]; e < A.ptr[r + 1]; ++e) { air.idx.push_back(A.idx[e]); air.val.push_back(A.val[e]); }
                air.ptr.push_back((int)air.idx.size());
            }
        }
        L.Air.upload(air);
        L.AirT.upload(transpose(air));
        L.Arr.upload(H.Arr[l]);
        if (l + 1 < H.levels) { L.P.upload(H.Prr[l]); L.PT.upload(transpose(H.Prr[l])); }
        // subdomains of the same shape have identical Aii: factor each distinct matrix once
        const std::vector<int> rep = identical_matrices(H.Aii[l]);
        std::vector<SupernodalFactor> fi(H.nsub);
        ParallelErrors errs;
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < H.nsub; ++s)
            if (H.Aii[l][s].rows > 0 && rep[s] == s) errs.run([&] { factor_spd(H.Aii[l][s], H.Aii_coords[l][s], 1e-5, fi[s]); });
        errs.rethrow();
        std::vector<const SupernodalFactor*> fp;
        std::vector<int> fo;
        for (int s = 0; s < H.nsub; ++s) if (H.Aii[l][s].rows > 0) { fp.push_back(&fi[rep[s]]); fo.push_back(off[s]); }
        L.Lii.build(fp, fo);
        L.n_distinct = 0;
        for (int s = 0; s < H.nsub; ++s) L.n_distinct += H.Aii[l][s].rows > 0 && rep[s] == s;
        std::vector<Int3> rc(H.if_coords[l].size());
        for (size_t i = 0; i < rc.size(); ++i) rc[i] = {H.if_coords[l][i].x << l, H.if_coords[l][i].y << l, H.if_coords[l][i].z << l};
        SupernodalFactor fr;
        factor_spd(H.Arr[l], rc, 1e-5, fr);
        L.Lrr.build({&fr}, {0});
        if (l == H.levels - 1 && L.n_if > 0) {
            const char* mode = std::getenv("HDD_SIGMA_BOTTOM");
            const std::string m = mode ? mode : "auto";
            if (m != "loop" && m != "direct" && m != "auto") throw std::runtime_error("HDD_SIGMA_BOTTOM must be loop, direct or auto");
            if (m != "loop") {
                Csr C; std::vector<Int3> xyz;
                combined_bottom(H, l, C, xyz);
                SupernodalFactor fc;
                factor_spd(C, xyz, 1e-5, fc);
                if (m == "direct") {
                    S->bottom_direct = true;  // no low-rank correction: exact only when M^200 is negligible
                } else {
                    BottomSpectrum sp;
                    bottom_spectrum(H.Arr[l], 1e-5, fc, L.n_int, kBottomLambdaCut, kBottomBasisBytes, sp);
                    S->bottom_rho = sp.lambda_max; S->bottom_steps = sp.steps; S->bottom_seconds = sp.seconds;
                    if (sp.resolved) {
                        S->bottom_direct = true;
                        upload_spectrum(*S, sp, L.n_if, H.levels == 1);
                    }
                }
                if (S->bottom_direct) {
                    S->Lbottom.build({&fc}, {0});
                    S->bot_rhs.alloc(C.rows); S->bot_sol.alloc(C.rows);
                    S->plan_bytes += S->Lbottom.bytes();
                }
            }
        }
        std::vector<SupernodalFactor>().swap(fi);
        L.u.alloc(std::max(1, L.n_if)); L.b.alloc(std::max(1, L.n_if)); L.r.alloc(std::max(1, L.n_if)); L.tmp.alloc(std::max(1, L.n_if));
        L.vb.alloc(std::max(1, L.n_int)); L.vx.alloc(std::max(1, L.n_int));
        S->plan_bytes += L.Lii.bytes() + L.Lrr.bytes() + L.Air.bytes() + L.AirT.bytes() + L.Arr.bytes() + L.P.bytes() + L.PT.bytes();
    }
    if (S->bottom_direct) {
        S->bottom_check = bottom_self_check(*S);
        const char* mode = std::getenv("HDD_SIGMA_BOTTOM");
        // the reference's sweeps stay in charge if the shortcut disagrees with them
        if (!(mode && std::string(mode) == "direct") && !(S->bottom_check < 1e-3)) S->bottom_direct = false;
    }
    S->factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (H.levels > 0) { S->b0 = S->lv[0].b.p; S->u0 = S->lv[0].u.p; }
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

// x* from the combined system [[Aii+eI, Air],[Air^T, Arr+eI]] [y; x*] = [0; b]
// (the interior part of bot_rhs stays zero), then u = x* + V coef.
void SigmaDevice::bottom(hipStream_t s, bool u_zero) {
    SigmaLevel& B = lv[levels - 1];
    const int ni = B.n_int, n = B.n_if;
    if (!u_zero && bottom_k && !spec_SV.p) throw std::runtime_error("bottom solve from a nonzero start needs S V");
    launch_copy(bot_rhs.p + ni, B.b.p, n, s);
    Lbottom.solve(bot_rhs.p, bot_sol.p, s);
    if (bottom_k == 0) { launch_copy(B.u.p, bot_sol.p + ni, n, s); return; }
    hipLaunchKernelGGL(k_spec_dot, dim3(spec_nch, bottom_k), dim3(256), 0, s, n, spec_nch, (const float*)spec_V.p,
                       (const float*)spec_SV.p, (const float*)B.b.p, u_zero ? (const float*)nullptr : (const float*)B.u.p, spec_part.p);
    hipLaunchKernelGGL(k_spec_coef, dim3(blocks_for(bottom_k, 64)), dim3(64), 0, s, bottom_k, spec_nch, (const double*)spec_part.p,
                       (const double*)spec_w1.p, (const double*)spec_w2.p, spec_coef.p);
    hipLaunchKernelGGL(k_spec_apply, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, bottom_k, (const float*)spec_V.p, (const float*)spec_coef.p,
                       (const float*)(bot_sol.p + ni), B.u.p);
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

void sigma_vcycles(SigmaDevice& S, hipStream_t s) {
    if (S.levels == 0) return;
    launch_fill(S.u0, 0.f, S.lv[0].u.n, s);
    for (int i = 0; i < S.n_mg; ++i) S.vcycle(s, i == 0);
}

void launch_interface_gather(SigmaDevice& S, const DevLevel& L0, const int* if_index, const float* z, const float* r, hipStream_t s) {
    if (S.levels == 0 || S.lv[0].n_if == 0) return;
    launch_if_gather(L0, if_index, S.lv[0].n_if, z, r, S.b0, s);
}

void launch_interface_scatter(SigmaDevice& S, const int* if_index, float* z, hipStream_t s) {
    if (S.levels == 0 || S.lv[0].n_if == 0) return;
    launch_if_scatter(if_index, S.lv[0].n_if, S.u0, z, s);
}

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
