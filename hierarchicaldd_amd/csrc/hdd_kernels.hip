// HIP kernels for gfx950: matrix-free V-cycle on paged SPGrid levels and the
// outer PCG vector operations.
//
// Launch shapes: the smoothing sweep and the residual take one 64-lane
// wavefront per 4x8x8 page, four cells per lane with 16-byte accesses (see
// Stencil4); the other page kernels one 256-thread workgroup per page, thread
// t owning cell t (t = x*64 + y*8 + z), so every channel access of a wavefront
// is a contiguous, coalesced segment. Face neighbours inside a page are at
// cell offsets +-64, +-8, +-1; across the page face they come from the page's
// 6-entry neighbour table (the zero page `np` stands in for unallocated pages).
//
// Floating-point order reproduces the reference kernels operation by
// operation (the library is compiled with -ffp-contract=off):
//   residual   r = b - sum_f -(u_n - u_c)   LIB/SPGrid_Residual_Legacy.h:52-61
//   correction u += (omega/fc) * r          LIB/SPGrid_Correction_Legacy.h:60-77
//   restriction / prolongation weights      LIB/SPGrid_Restriction_Legacy.h:50-181,
//                                           LIB/SPGrid_Prolongation_Legacy.h:48-123
//   PCG Laplace (double accumulation)       DDS/SPGrid_Gemini_Arithmetic_Helpers.h:240-262
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <mutex>
#include "hdd_device.h"

namespace hdd {

__constant__ float c_omega_fc[8];   // (float)(2/3) / (float)fc, fc = 0..6
__constant__ float c_rweights[64];  // restriction weights, [i][j][k] with w = {1,3,3,1}/8
__constant__ float c_pweights[8];   // prolongation weights 4 * {3/4,1/4}^3, [i][j][k]

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i mod 8), and each XCD has its own L2. Giving XCD x the contiguous page range
// [x*chunk, (x+1)*chunk) keeps Morton neighbours, whose halo cells a stencil
// re-reads, in the same L2. Launch with kXcds * chunk workgroups; ids past np
// return.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_chunk(int np) { return (np + kXcds - 1) / kXcds; }
__device__ __forceinline__ int xcd_page(int np) {
    return (int)(blockIdx.x % kXcds) * xcd_chunk(np) + (int)(blockIdx.x / kXcds);
}

__device__ __forceinline__ bool page_stopped(const DevLevel& L, int p) { return L.active && !L.active[L.prob[p]]; }

__device__ __forceinline__ int nb_index(const int* __restrict__ nbr, int p, int c, int f) {
    switch (f) {
        case 0: return (c >= 64) ? p * 256 + c - 64 : nbr[p * 6 + 0] * 256 + c + 192;
        case 1: return (c < 192) ? p * 256 + c + 64 : nbr[p * 6 + 1] * 256 + c - 192;
        case 2: return ((c & 56) != 0) ? p * 256 + c - 8 : nbr[p * 6 + 2] * 256 + c + 56;
        case 3: return ((c & 56) != 56) ? p * 256 + c + 8 : nbr[p * 6 + 3] * 256 + c - 56;
        case 4: return ((c & 7) != 0) ? p * 256 + c - 1 : nbr[p * 6 + 4] * 256 + c + 7;
        default: return ((c & 7) != 7) ? p * 256 + c + 1 : nbr[p * 6 + 5] * 256 + c - 7;
    }
}

// V-cycle storage: float, or half (FP16 storage, FP32 arithmetic: values are
// widened on load and rounded to nearest on store).
__device__ __forceinline__ float ld(const float* __restrict__ p, size_t i) { return p[i]; }
__device__ __forceinline__ float ld(const __half* __restrict__ p, size_t i) { return __half2float(p[i]); }
__device__ __forceinline__ void st(float* __restrict__ p, size_t i, float v) { p[i] = v; }
__device__ __forceinline__ void st(__half* __restrict__ p, size_t i, float v) { p[i] = __float2half_rn(v); }

// Stencil sweeps: one 64-lane wavefront per page, lane l owns the four cells
// 4l .. 4l+3 (z = 4(l&1) .. +3 of row x = l>>4, y = (l>>1)&7) and moves them
// with one 16-byte access per channel (8 bytes with FP16 storage). Face
// neighbours inside the page are the vectors of lanes l+-16 (x) and l+-2 (y)
// and the neighbouring elements (z); across a page face they come from the
// page's 6-entry neighbour table (the zero page `np` stands in for unallocated
// pages). Four pages per 256-thread workgroup (the sweep: eight per 512).
struct V4 { float v[4]; };
__device__ __forceinline__ V4 ld4(const float* __restrict__ p, size_t i4) {
    const float4 t = reinterpret_cast<const float4*>(p)[i4];
    return {{t.x, t.y, t.z, t.w}};
}
__device__ __forceinline__ V4 ld4(const __half* __restrict__ p, size_t i4) {
    const uint2 t = reinterpret_cast<const uint2*>(p)[i4];
    __half2 a, b;
    __builtin_memcpy(&a, &t.x, 4);
    __builtin_memcpy(&b, &t.y, 4);
    const float2 fa = __half22float2(a), fb = __half22float2(b);
    return {{fa.x, fa.y, fb.x, fb.y}};
}
__device__ __forceinline__ void st4(float* __restrict__ p, size_t i4, const V4& v) {
    reinterpret_cast<float4*>(p)[i4] = make_float4(v.v[0], v.v[1], v.v[2], v.v[3]);
}
__device__ __forceinline__ void st4(__half* __restrict__ p, size_t i4, const V4& v) {
    const __half2 a = __floats2half2_rn(v.v[0], v.v[1]), b = __floats2half2_rn(v.v[2], v.v[3]);
    uint2 t;
    __builtin_memcpy(&t.x, &a, 4);
    __builtin_memcpy(&t.y, &b, 4);
    reinterpret_cast<uint2*>(p)[i4] = t;
}

// The four cells of lane l of page p with their six face neighbours, and the
// float residual of each, in the reference operation order
// (r = b - sum over set face bits (-x,+x,-y,+y,-z,+z) of -(u_n - u_c),
// LIB/SPGrid_Residual_Legacy.h:52-61). Neighbour values are loaded whether or
// not their face bit is set (every address is valid) and selected afterwards.
template <class S>
struct Stencil4 {
    V4 uc, r;
    uint32_t f[4];
    __device__ __forceinline__ Stencil4(const DevLevel& L, int p, int l, const S* __restrict__ u, const S* __restrict__ b) {
        const int x = l >> 4, y = (l >> 1) & 7, zh = l & 1;
        const int* nb = L.nbr + p * 6;
        const size_t base = (size_t)p * 64;
        const uint4 fv = reinterpret_cast<const uint4*>(L.flags)[base + l];
        f[0] = fv.x; f[1] = fv.y; f[2] = fv.z; f[3] = fv.w;
        uc = ld4(u, base + l);
        const V4 xm = ld4(u, x > 0 ? base + l - 16 : (size_t)nb[0] * 64 + l + 48);
        const V4 xp = ld4(u, x < 3 ? base + l + 16 : (size_t)nb[1] * 64 + l - 48);
        const V4 ym = ld4(u, y > 0 ? base + l - 2 : (size_t)nb[2] * 64 + l + 14);
        const V4 yp = ld4(u, y < 7 ? base + l + 2 : (size_t)nb[3] * 64 + l - 14);
        const float zm = ld(u, zh ? (size_t)p * 256 + 4 * l - 1 : (size_t)nb[4] * 256 + 4 * l + 7);
        const float zp = ld(u, zh ? (size_t)nb[5] * 256 + 4 * l - 4 : (size_t)p * 256 + 4 * l + 4);
        const V4 bv = ld4(b, base + l);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float cv = uc.v[k];
            const float nv[6] = {xm.v[k], xp.v[k], ym.v[k], yp.v[k], k > 0 ? uc.v[k - 1] : zm, k < 3 ? uc.v[k + 1] : zp};
            float result = 0.f;
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (f[k] & (kFaceMX << q)) result -= (nv[q] - cv);
            r.v[k] = bv.v[k] - result;
        }
    }
};

__host__ __device__ inline int page_groups(int np) { return (np + 3) >> 2; }
// the sweep takes eight pages per 512-thread workgroup (measured 1.7% faster than four at 256^3)
constexpr int kSweepPages = 8;
__host__ __device__ inline int sweep_groups(int np) { return (np + kSweepPages - 1) / kSweepPages; }

// One damped-Jacobi sweep (residual + correction fused, ping-pong buffers);
// u += (omega / face count) r on the selected cells
// (LIB/SPGrid_Correction_Legacy.h:60-77), other cells copied unchanged.
template <class S>
__global__ void __launch_bounds__(64 * kSweepPages) k_smooth(DevLevel L, const S* __restrict__ u_in, S* __restrict__ u_out,
                                                             const S* __restrict__ b, int write_interior, int write_boundary) {
    const int ng = sweep_groups(L.np), gi = xcd_page(ng), l = threadIdx.x & 63;
    const int p = gi * kSweepPages + (int)(threadIdx.x >> 6);
    if (gi >= ng || p >= L.np || page_stopped(L, p)) return;
    const Stencil4<S> st(L, p, l, u_in, b);
    V4 out;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t f = st.f[k], sel = f & (kActive | kBoundary);
        const bool w = (sel == kActive && write_interior) || (sel == (kActive | kBoundary) && write_boundary);
        float u = st.uc.v[k];
        if (w) u += c_omega_fc[__popc(f & (0x3fu * kFaceMX))] * st.r.v[k];
        out.v[k] = u;
    }
    st4(u_out, (size_t)p * 64 + l, out);
}

// All `count` sweeps (interior and boundary cells) of a level whose Active
// cells couple only within small clusters of pages (the bottom of the V-cycle:
// each subdomain's coarsest block lies in one or two pages). A workgroup holds
// up to four pages, whole clusters (wg_page, -1 for an empty slot), 256
// threads per page, thread t owning cell t, and sweeps them in LDS with the
// operations of Stencil4 and k_smooth, storing every iterate in the level's
// storage type, so the result is bit-identical to `count` launches of
// k_smooth. A face neighbour page of the same cluster is read from its LDS
// slot (nb_slot[p*6+f], -1: none); other neighbour pages hold no cell a
// coupled Active cell reads that a sweep writes, so they are read once. Every
// thread takes part in the barriers; only those of a live page compute. The
// result goes to u_out, a buffer other than u_in (other workgroups read u_in's
// pages).
template <class S>
__global__ void __launch_bounds__(1024) k_smooth_local(DevLevel L, const int* __restrict__ wg_page, const int* __restrict__ nb_slot,
                                                       const S* __restrict__ u_in, S* __restrict__ u_out, const S* __restrict__ b,
                                                       int count) {
    __shared__ S us[2][4][kPageCells];
    const int c = threadIdx.x & 255, w = threadIdx.x >> 8;
    const int p = wg_page[blockIdx.x * 4 + w];
    const bool live = p >= 0 && !page_stopped(L, p);
    const int x = c >> 6, y = (c >> 3) & 7, z = c & 7;
    // per face: the in-page neighbour cell, or (at the page face) the
    // neighbour page's cell, read from its LDS slot or once from u_in
    const int nc[6] = {x > 0 ? c - 64 : c + 192, x < 3 ? c + 64 : c - 192, y > 0 ? c - 8 : c + 56,
                       y < 7 ? c + 8 : c - 56, z > 0 ? c - 1 : c + 7, z < 7 ? c + 1 : c - 7};
    const bool inside[6] = {x > 0, x < 3, y > 0, y < 7, z > 0, z < 7};
    int slot[6];
    float cst[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t f = 0u;
    float bv = 0.f, wgt = 0.f;
    if (live) {
        f = L.flags[(size_t)p * 256 + c];
        const uint32_t sel = f & (kActive | kBoundary);
        if (sel == kActive || sel == (kActive | kBoundary)) wgt = c_omega_fc[__popc(f & (0x3fu * kFaceMX))];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            slot[q] = inside[q] ? w : nb_slot[p * 6 + q];
            if (slot[q] < 0) cst[q] = ld(u_in, (size_t)L.nbr[p * 6 + q] * 256 + nc[q]);
        }
        bv = ld(b, (size_t)p * 256 + c);
        us[0][w][c] = u_in[(size_t)p * 256 + c];
    }
    __syncthreads();
    int cur = 0;
    for (int it = 0; it < count; ++it) {
        if (live) {
            const float cv = ld(us[cur][w], c);
            float nv[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) nv[q] = slot[q] >= 0 ? ld(us[cur][slot[q]], nc[q]) : cst[q];
            float result = 0.f;
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (f & (kFaceMX << q)) result -= (nv[q] - cv);
            const float r = bv - result;
            const uint32_t sel = f & (kActive | kBoundary);
            float u = cv;
            if (sel == kActive || sel == (kActive | kBoundary)) u += wgt * r;
            st(us[1 - cur][w], c, u);
        }
        __syncthreads();
        cur = 1 - cur;
    }
    if (live) u_out[(size_t)p * 256 + c] = us[cur][w][c];
}

// residual at Active cells; other cells of r are written 0 (the value they
// always hold: only this kernel writes r, and the reference clears it before
// each V-cycle, DDS/SPGrid_Domain_Decomposition_Preconditioner.h:93,120)
template <class S>
__global__ void __launch_bounds__(256) k_residual(DevLevel L, const S* __restrict__ u, const S* __restrict__ b,
                                                  S* __restrict__ r) {
    const int ng = page_groups(L.np), gi = xcd_page(ng), l = threadIdx.x & 63;
    const int p = gi * 4 + (int)(threadIdx.x >> 6);
    if (gi >= ng || p >= L.np || page_stopped(L, p)) return;
    const Stencil4<S> st(L, p, l, u, b);
    V4 out;
#pragma unroll
    for (int k = 0; k < 4; ++k) out.v[k] = (st.f[k] & kActive) ? st.r.v[k] : 0.f;
    st4(r, (size_t)p * 64 + l, out);
}

template <class S>
__global__ void k_fill(S* __restrict__ a, float v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) st(a, i, v);
}
__global__ void k_copy(float* __restrict__ d, const float* __restrict__ s, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i];
}

// page of block (bx, by, bz) of a level (the zero page outside the block box or where none is allocated)
__device__ __forceinline__ int lut_block(const DevLevel& L, int bx, int by, int bz) {
    if (bx < 0 || by < 0 || bz < 0 || bx >= L.nbx || by >= L.nby || bz >= L.nbz) return L.np;
    const int q = L.lut[(bx * L.nby + by) * L.nbz + bz];
    return q < 0 ? L.np : q;
}

// coarse b = sum_{4x4x4} w r_fine over fine cells 2c-2 .. 2c+1 (Active coarse cells only).
// The stencils of a coarse page (origin o) reach the fine region 2o-2 .. 2o+7
// in x and 2o-2 .. 2o+15 in y and z (10 x 18 x 18 cells in 3 x 3 x 3 fine
// blocks): its pages are looked up once and the region is staged in LDS
// once, so each fine value is read from memory once instead of by up to 8
// stencils; every stencil then sums the same values in the same order.
template <class S>
__global__ void __launch_bounds__(256) k_restrict(DevLevel F, DevLevel C, const S* __restrict__ r_f, S* __restrict__ b_c) {
    constexpr int RX = 10, RY = 18, RZ = 18;
    __shared__ int tab[27];
    __shared__ float sr[RX * RY * RZ];
    const int p = xcd_page(C.np), c = threadIdx.x, e = p * 256 + c;
    if (p >= C.np || page_stopped(C, p)) return;  // uniform over the workgroup
    const uint32_t fl = C.flags[e];  // the cell's own load goes out before the staging's
    const int ox = C.base[p * 3 + 0], oy = C.base[p * 3 + 1], oz = C.base[p * 3 + 2];
    const int X0 = 2 * ox - 2, Y0 = 2 * oy - 2, Z0 = 2 * oz - 2;
    const int bx0 = X0 >> 2, by0 = Y0 >> 3, bz0 = Z0 >> 3;
    if (c < 27) tab[c] = lut_block(F, bx0 + c / 9, by0 + (c / 3) % 3, bz0 + c % 3);
    __syncthreads();
    for (int t = c; t < RX * RY * RZ; t += 256) {
        const int X = X0 + t / (RY * RZ), Y = Y0 + (t / RZ) % RY, Z = Z0 + t % RZ;
        const int q = tab[((X >> 2) - bx0) * 9 + ((Y >> 3) - by0) * 3 + ((Z >> 3) - bz0)];
        sr[t] = ld(r_f, q * 256 + ((X & 3) << 6 | (Y & 7) << 3 | (Z & 7)));
    }
    __syncthreads();
    if (!(fl & kActive)) return;
    const int lx = 2 * (c >> 6), ly = 2 * ((c >> 3) & 7), lz = 2 * (c & 7);  // stencil origin in the region
    float acc = 0.f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k)
                acc += c_rweights[i * 16 + j * 4 + k] * sr[((lx + i) * RY + (ly + j)) * RZ + (lz + k)];
    st(b_c, e, acc);
}

// fine u += sum_{2x2x2} w u_coarse (Active fine cells only). The coarse cells
// a fine page (origin o) reads form the region o/2 .. o/2+2 (x) and
// o/2 .. o/2+4 (y, z), 3 x 5 x 5 cells in 2 x 2 x 2 coarse blocks: its pages
// are looked up once and the region is staged in LDS once per workgroup.
template <class S>
__global__ void __launch_bounds__(256) k_prolong(DevLevel F, DevLevel C, S* __restrict__ u_f, const S* __restrict__ u_c) {
    constexpr int RX = 3, RY = 5, RZ = 5;
    __shared__ int tab[8];
    __shared__ float su[RX * RY * RZ];
    const int p = xcd_page(F.np), c = threadIdx.x, e = p * 256 + c;
    if (p >= F.np || page_stopped(F, p)) return;  // uniform over the workgroup
    const uint32_t fl = F.flags[e];  // the cell's own loads go out before the staging's
    float u = ld(u_f, e);
    const int ox = F.base[p * 3 + 0], oy = F.base[p * 3 + 1], oz = F.base[p * 3 + 2];
    const int X0 = ox >> 1, Y0 = oy >> 1, Z0 = oz >> 1;
    const int bx0 = X0 >> 2, by0 = Y0 >> 3, bz0 = Z0 >> 3;
    if (c < 8) tab[c] = lut_block(C, bx0 + (c >> 2), by0 + ((c >> 1) & 1), bz0 + (c & 1));
    __syncthreads();
    if (c < RX * RY * RZ) {
        const int X = X0 + c / (RY * RZ), Y = Y0 + (c / RZ) % RY, Z = Z0 + c % RZ;
        const int q = tab[((X >> 2) - bx0) * 4 + ((Y >> 3) - by0) * 2 + ((Z >> 3) - bz0)];
        su[c] = ld(u_c, q * 256 + ((X & 3) << 6 | (Y & 7) << 3 | (Z & 7)));
    }
    __syncthreads();
    if (!(fl & kActive)) return;
    const int x = ox + (c >> 6), y = oy + ((c >> 3) & 7), z = oz + (c & 7);
    const int cx = (x + 1) >> 1, cy = (y + 1) >> 1, cz = (z + 1) >> 1;
    const int px = 1 - (x - (2 * cx - 1)), py = 1 - (y - (2 * cy - 1)), pz = 1 - (z - (2 * cz - 1));
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const int X = cx + i - px, Y = cy + j - py, Z = cz + k - pz;
                u += c_pweights[((i ^ px) << 2) | ((j ^ py) << 1) | (k ^ pz)] * su[((X - X0) * RY + (Y - Y0)) * RZ + (Z - Z0)];
            }
    st(u_f, e, u);
}

// PCG operator: out = L v on Active|Interface cells (double accumulation); kZero: 0 elsewhere
// (the PCG's Multiply fused with Project), else other cells are left as they are (SPGrid_Laplace)
template <bool kZero>
__global__ void __launch_bounds__(256) k_laplace_pcg(DevLevel L, const float* __restrict__ v, float* __restrict__ out) {
    const int p = xcd_page(L.np), c = threadIdx.x, e = p * 256 + c;
    if (p >= L.np || page_stopped(L, p)) return;
    const uint32_t f = L.flags[e];
    float res = 0.f;
    if (f & (kActive | kInterface)) {
        const double cv = (double)v[e];
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (f & (kFaceMX << k)) acc -= ((double)v[nb_index(L.nbr, p, c, k)] - cv);
        res = (float)acc;
    } else if (!kZero) {
        return;
    }
    out[e] = res;
}

// CG_VECTOR operations over the cells of the real pages (SPGrid_Gemini_Arithmetic_Helpers.h):
// 0 Copy out = a*c, 1 Copy2 out = a*c + b, 2 Plus_Equal out += a, 3 Minus_Equal out -= a, 4 Scale out *= c
__global__ void __launch_bounds__(256) k_vector_op(int op, float* __restrict__ out, const float* __restrict__ a, const float* __restrict__ b,
                                                   float c, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        switch (op) {
            case 0: out[i] = a[i] * c; break;
            case 1: out[i] = a[i] * c + b[i]; break;
            case 2: out[i] += a[i]; break;
            case 3: out[i] -= a[i]; break;
            default: out[i] *= c; break;
        }
    }
}

// Project: zero values outside Active|Interface
__global__ void __launch_bounds__(256) k_project(DevLevel L, float* __restrict__ v) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (!(L.flags[e] & (kActive | kInterface))) v[e] = 0.f;
}

// r = b - q, then Project
__global__ void __launch_bounds__(256) k_sub_project(DevLevel L, const float* __restrict__ b, const float* __restrict__ q, float* __restrict__ r) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    r[e] = (L.flags[e] & (kActive | kInterface)) ? b[e] - q[e] : 0.f;
}

// ---------------------------------------------------------------------------
// PCG vector operations, per problem. Problem p of a batch owns the level-0
// pages [pg0[p], pg0[p+1]); page_prob maps a page to its problem (-1 for the
// zero page, which is never written). Each problem's reductions are laid out
// exactly as a one-problem reduction over its own (pages + 1) * 256 cells
// (parts[p] blocks striding over them; the zero page adds nothing), so a
// problem's sums and scalars do not depend on the batch it is solved in.
// Scalars: 8 floats per problem (rho, alpha, beta, rho_old, max|r|), 4 doubles.

// out = in1 * scal[slot] + in2 (SPGrid_Copy2 with the problem's scalar), active problems only
__global__ void __launch_bounds__(256) k_seg_copy2(float* __restrict__ out, const float* __restrict__ in1, const float* __restrict__ in2,
                                                   const float* __restrict__ scal, int slot, const int* __restrict__ page_prob,
                                                   const int* __restrict__ active, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int pr = page_prob[i >> 8];
        if (pr < 0 || !active[pr]) continue;
        out[i] = in1[i] * scal[pr * 8 + slot] + in2[i];
    }
}

// x = s*alpha + x ; r = q*(-alpha) + r  (two SPGrid_Copy2 calls fused), active problems only
__global__ void __launch_bounds__(256) k_seg_xr_update(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ s,
                                                       const float* __restrict__ q, const float* __restrict__ scal,
                                                       const int* __restrict__ page_prob, const int* __restrict__ active, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int pr = page_prob[i >> 8];
        if (pr < 0 || !active[pr]) continue;
        const float a = scal[pr * 8 + 1], na = -a;
        x[i] = s[i] * a + x[i];
        r[i] = q[i] * na + r[i];
    }
}

// FP64 outer variant of k_seg_copy2: out = (float)(in1 * dscal[slot] + in2) in double
__global__ void __launch_bounds__(256) k_seg_copy2_d(float* __restrict__ out, const float* __restrict__ in1, const float* __restrict__ in2,
                                                     const double* __restrict__ dscal, int slot, const int* __restrict__ page_prob,
                                                     const int* __restrict__ active, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int pr = page_prob[i >> 8];
        if (pr < 0 || !active[pr]) continue;
        out[i] = (float)((double)in1[i] * dscal[pr * 8 + slot] + (double)in2[i]);
    }
}

// FP64 outer variant of k_seg_xr_update: x (double) += alpha s; r = (float)(r - alpha q), alpha double
__global__ void __launch_bounds__(256) k_seg_xr_update_d(double* __restrict__ x, float* __restrict__ r, const float* __restrict__ s,
                                                         const float* __restrict__ q, const double* __restrict__ dscal,
                                                         const int* __restrict__ page_prob, const int* __restrict__ active, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int pr = page_prob[i >> 8];
        if (pr < 0 || !active[pr]) continue;
        const double a = dscal[pr * 8 + 3];
        x[i] += a * (double)s[i];
        r[i] = (float)((double)r[i] - a * (double)q[i]);
    }
}

// True residual r = b - L x from the FP64 iterate (double accumulation), 0 off Active|Interface
__global__ void __launch_bounds__(256) k_true_residual(DevLevel L, const double* __restrict__ x, const float* __restrict__ b,
                                                       float* __restrict__ r) {
    const int p = xcd_page(L.np), c = threadIdx.x, e = p * 256 + c;
    if (p >= L.np || page_stopped(L, p)) return;
    const uint32_t f = L.flags[e];
    float res = 0.f;
    if (f & (kActive | kInterface)) {
        const double cv = x[e];
        double acc = (double)b[e];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (f & (kFaceMX << k)) acc += (x[nb_index(L.nbr, p, c, k)] - cv);
        res = (float)acc;
    }
    r[e] = res;
}

__global__ void k_cast_to_float(const double* __restrict__ x, float* __restrict__ y, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) y[i] = (float)x[i];
}

// block-partial double dot products of float products (a*b rounded to float, then summed in double);
// block (c, p) = partial c of problem p
__global__ void __launch_bounds__(256) k_seg_dot_partial(const float* __restrict__ a, const float* __restrict__ b, double* __restrict__ part,
                                                         const int* __restrict__ pg0, const int* __restrict__ parts, int max_parts) {
    const int p = blockIdx.y, c = blockIdx.x, g = parts[p];
    if (c >= g) return;
    const size_t base = (size_t)pg0[p] * 256, n = (size_t)(pg0[p + 1] - pg0[p]) * 256;
    __shared__ double sh[4];
    double s = 0.0;
    for (size_t i = (size_t)c * 256 + threadIdx.x; i < n; i += (size_t)g * 256) s += (double)(a[base + i] * b[base + i]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[(size_t)p * max_parts + c] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}
// The reference's order of CG_SYSTEM::Inner_Product (DDS/CG_SYSTEM.cpp:62-72): one
// running double sum over the problem's pages in page order and their cells in
// cell order, each term the float product. One workgroup per problem stages 1024
// products at a time in LDS and one lane adds them in sequence; the sum goes to
// the problem's first partial and its other partials are zero, so k_seg_finish
// gives it unchanged. For checking against the reference's rounding, not speed.
__global__ void __launch_bounds__(256) k_seg_dot_serial(const float* __restrict__ a, const float* __restrict__ b, double* __restrict__ part,
                                                        const int* __restrict__ pg0, const int* __restrict__ parts, int max_parts) {
    const int p = blockIdx.x;
    const size_t base = (size_t)pg0[p] * 256, n = (size_t)(pg0[p + 1] - pg0[p]) * 256;
    __shared__ double buf[1024];
    double s = 0.0;
    for (size_t c0 = 0; c0 < n; c0 += 1024) {
        const size_t cnt = n - c0 < 1024 ? n - c0 : 1024;
        for (size_t t = threadIdx.x; t < cnt; t += 256) buf[t] = (double)(a[base + c0 + t] * b[base + c0 + t]);
        __syncthreads();
        if (threadIdx.x == 0)
            for (size_t t = 0; t < cnt; ++t) s += buf[t];
        __syncthreads();
    }
    for (int c = threadIdx.x; c < parts[p]; c += 256) part[(size_t)p * max_parts + c] = c == 0 ? s : 0.0;
}

__global__ void __launch_bounds__(256) k_seg_maxabs_partial(const float* __restrict__ a, float* __restrict__ part,
                                                            const int* __restrict__ pg0, const int* __restrict__ parts, int max_parts) {
    const int p = blockIdx.y, c = blockIdx.x, g = parts[p];
    if (c >= g) return;
    const size_t base = (size_t)pg0[p] * 256, n = (size_t)(pg0[p + 1] - pg0[p]) * 256;
    __shared__ float sh[4];
    float m = 0.f;
    for (size_t i = (size_t)c * 256 + threadIdx.x; i < n; i += (size_t)g * 256) m = fmaxf(m, fabsf(a[base + i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[(size_t)p * max_parts + c] = fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// Final reduction of one problem's partials (workgroup p) and the PCG scalar
// logic of CONJUGATE_GRADIENT_DUMPING::Solve (:193-205) in float:
//   mode 0: rho = (float)dot -> scal[0]; beta = rho/rho_old -> scal[2]
//   mode 1: s_dot_q = (float)dot; alpha = rho/s_dot_q -> scal[1]; rho_old = rho -> scal[3]
//   mode 2: max-abs norm -> scal[4] (every problem; modes 0 and 1 active problems only)
//   mode 3: the plain sum -> dscal[2] (every problem)
//   modes 5 / 6: as 0 / 1 with the scalars in double (dscal: 0 rho, 1 s_dot_q,
//   3 alpha, 4 beta, 5 rho_old), active problems only
__global__ void __launch_bounds__(256) k_seg_finish(const double* __restrict__ dpart, const float* __restrict__ fpart,
                                                    const int* __restrict__ parts, int max_parts, int mode, const int* __restrict__ active,
                                                    float* __restrict__ scal_all, double* __restrict__ dscal_all) {
    const int p = blockIdx.x;
    if ((mode == 0 || mode == 1 || mode == 5 || mode == 6) && !active[p]) return;
    const int nparts = parts[p];
    float* scal = scal_all + p * 8;
    double* dscal = dscal_all + p * 8;
    __shared__ double sd[256];
    __shared__ float sf[256];
    double s = 0.0;
    float m = 0.f;
    for (int i = threadIdx.x; i < nparts; i += 256) {
        if (mode != 2) s += dpart[(size_t)p * max_parts + i]; else m = fmaxf(m, fpart[(size_t)p * max_parts + i]);
    }
    sd[threadIdx.x] = s; sf[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sd[threadIdx.x] += sd[threadIdx.x + o]; sf[threadIdx.x] = fmaxf(sf[threadIdx.x], sf[threadIdx.x + o]); }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == 0) {
            const float rho = (float)sd[0];
            scal[0] = rho;
            scal[2] = rho / scal[3];
            dscal[0] = sd[0];
        } else if (mode == 3) {
            dscal[2] = sd[0];  // a plain inner product (Inner_Product), no PCG scalar touched
        } else if (mode == 5) {
            // FP64 outer scalars: rho and beta = rho / rho_old in double
            dscal[0] = sd[0];
            dscal[4] = sd[0] / dscal[5];
        } else if (mode == 6) {
            // FP64 outer scalars: alpha = rho / s_dot_q, rho_old = rho
            dscal[1] = sd[0];
            dscal[3] = sd[0] != 0.0 ? dscal[0] / sd[0] : 3.40282347e+38;
            dscal[5] = dscal[0];
        } else if (mode == 1) {
            const float sdq = (float)sd[0];
            const float rho = scal[0];
            scal[1] = sdq != 0.f ? rho / sdq : 3.40282347e+38f;
            scal[3] = rho;
            dscal[1] = sd[0];
        } else {
            scal[4] = sf[0];
        }
    }
}
// FP16 storage: the preconditioner is linear, so it runs on r / a and its
// result is scaled back by a, with a = max|r| of the cell's problem (slot 4 of
// its scalars; 1 where r = 0). Values then stay O(1), inside half's normal range.
__global__ void __launch_bounds__(256) k_scale_to_half(const float* __restrict__ r, __half* __restrict__ hb, const int* __restrict__ page_prob,
                                                       const float* __restrict__ scal, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int pr = page_prob[i >> 8];
        if (pr < 0) continue;
        const float a = scal[pr * 8 + 4];
        hb[i] = __float2half_rn(a > 0.f ? r[i] / a : 0.f);
    }
}
// z = a * z16 on Active|Interface cells, 0 elsewhere (Project)
__global__ void __launch_bounds__(256) k_unscale_project(DevLevel L, const __half* __restrict__ hz, float* __restrict__ z,
                                                         const int* __restrict__ page_prob, const float* __restrict__ scal) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int pr = page_prob[e >> 8];
    const float a = pr >= 0 ? scal[pr * 8 + 4] : 0.f;
    z[e] = (pr >= 0 && (L.flags[e] & (kActive | kInterface))) ? a * __half2float(hz[e]) : 0.f;
}

// ---------------------------------------------------------------------------
// Inner product that allocates nothing (hdd_inner_product_noalloc): the block
// partials of k_seg_dot_partial for one problem of n_pages pages go to the
// file-scope array `scratch` (one copy per device, held until the process
// ends), and the final sum of k_seg_finish (mode 3) to the caller's device
// double. The blocks stride over the cells, and both reductions add, exactly
// as those kernels do for a one-problem context, so the value has the bits of
// hdd_inner_product on a context holding the same pages.
constexpr int kNoAllocParts = 1024;  // reduce_parts never exceeds it
static __device__ __align__(256) double scratch[kNoAllocParts];

__global__ void __launch_bounds__(256) k_dot_partial_flat(const float* __restrict__ a, const float* __restrict__ b, size_t n, int g,
                                                          double* __restrict__ part) {
    const int c = blockIdx.x;
    __shared__ double sh[4];
    double s = 0.0;
    for (size_t i = (size_t)c * 256 + threadIdx.x; i < n; i += (size_t)g * 256) s += (double)(a[i] * b[i]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[c] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void __launch_bounds__(256) k_dot_finish_flat(const double* __restrict__ part, int nparts, double* __restrict__ out) {
    __shared__ double sd[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
    sd[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sd[threadIdx.x] += sd[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sd[0];
}

// ---------------------------------------------------------------------------
// host launchers
static inline int grid_for(size_t n) { size_t g = (n + 255) / 256; return (int)(g > 4096 ? 4096 : (g == 0 ? 1 : g)); }

static void upload_constants_once() {
    float om[8];
    const float omega = (float)(2. / 3.);
    om[0] = 0.f;
    for (int k = 1; k < 8; ++k) om[k] = omega / (float)k;
    float s1[4] = {(float)(1. / 8.), (float)(3. / 8.), (float)(3. / 8.), (float)(1. / 8.)};
    float rw[64];
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) for (int k = 0; k < 4; ++k) rw[i * 16 + j * 4 + k] = s1[i] * s1[j] * s1[k];
    float s2[2] = {0.75f, 0.25f};
    float pw[8];
    for (int i = 0; i < 2; ++i) for (int j = 0; j < 2; ++j) for (int k = 0; k < 2; ++k) pw[i * 4 + j * 2 + k] = s2[i] * s2[j] * s2[k] * 4.0f;
    hdd_check(hipMemcpyToSymbol(HIP_SYMBOL(c_omega_fc), om, sizeof(om)));
    hdd_check(hipMemcpyToSymbol(HIP_SYMBOL(c_rweights), rw, sizeof(rw)));
    hdd_check(hipMemcpyToSymbol(HIP_SYMBOL(c_pweights), pw, sizeof(pw)));
}

// the stencil constants are the same for every context: uploaded once per process
void upload_constants() {
    static std::once_flag once;
    std::call_once(once, upload_constants_once);
}

template <class S>
static void smooth_t(const DevLevel& L, const S* u_in, S* u_out, const S* b, bool interior, bool boundary, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_smooth<S>, dim3(kXcds * xcd_chunk(sweep_groups(L.np))), dim3(64 * kSweepPages), 0, s, L, u_in, u_out, b, (int)interior,
                                 (int)boundary);
}
template <class S>
static void smooth_local_t(const DevLevel& L, const int* wg_page, int nwg, const int* nb_slot, const S* u_in, S* u_out, const S* b,
                           int count, hipStream_t s) {
    if (nwg) hipLaunchKernelGGL(k_smooth_local<S>, dim3(nwg), dim3(1024), 0, s, L, wg_page, nb_slot, u_in, u_out, b, count);
}
template <class S>
static void residual_t(const DevLevel& L, const S* u, const S* b, S* r, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_residual<S>, dim3(kXcds * xcd_chunk(page_groups(L.np))), dim3(256), 0, s, L, u, b, r);
}
template <class S>
static void fill_t(S* a, float v, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill<S>, dim3(grid_for(n)), dim3(256), 0, s, a, v, n);
}
template <class S>
static void restrict_t(const DevLevel& F, const DevLevel& C, const S* r_f, S* b_c, hipStream_t s) {
    if (C.np) hipLaunchKernelGGL(k_restrict<S>, dim3(kXcds * xcd_chunk(C.np)), dim3(256), 0, s, F, C, r_f, b_c);
}
template <class S>
static void prolong_t(const DevLevel& F, const DevLevel& C, S* u_f, const S* u_c, hipStream_t s) {
    if (F.np) hipLaunchKernelGGL(k_prolong<S>, dim3(kXcds * xcd_chunk(F.np)), dim3(256), 0, s, F, C, u_f, u_c);
}
void launch_smooth(const DevLevel& L, const float* u_in, float* u_out, const float* b, bool interior, bool boundary, hipStream_t s) {
    smooth_t(L, u_in, u_out, b, interior, boundary, s);
}
void launch_smooth(const DevLevel& L, const __half* u_in, __half* u_out, const __half* b, bool interior, bool boundary, hipStream_t s) {
    smooth_t(L, u_in, u_out, b, interior, boundary, s);
}
void launch_smooth_local(const DevLevel& L, const int* wg_page, int nwg, const int* nb_slot, const float* u_in, float* u_out,
                         const float* b, int count, hipStream_t s) {
    smooth_local_t(L, wg_page, nwg, nb_slot, u_in, u_out, b, count, s);
}
void launch_smooth_local(const DevLevel& L, const int* wg_page, int nwg, const int* nb_slot, const __half* u_in, __half* u_out,
                         const __half* b, int count, hipStream_t s) {
    smooth_local_t(L, wg_page, nwg, nb_slot, u_in, u_out, b, count, s);
}
void launch_residual(const DevLevel& L, const float* u, const float* b, float* r, hipStream_t s) { residual_t(L, u, b, r, s); }
void launch_residual(const DevLevel& L, const __half* u, const __half* b, __half* r, hipStream_t s) { residual_t(L, u, b, r, s); }
void launch_fill(float* a, float v, size_t n, hipStream_t s) { fill_t(a, v, n, s); }
void launch_fill(__half* a, float v, size_t n, hipStream_t s) { fill_t(a, v, n, s); }
void launch_copy(float* d, const float* src, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, s, d, src, n);
}
void launch_restrict(const DevLevel& F, const DevLevel& C, const float* r_f, float* b_c, hipStream_t s) { restrict_t(F, C, r_f, b_c, s); }
void launch_restrict(const DevLevel& F, const DevLevel& C, const __half* r_f, __half* b_c, hipStream_t s) { restrict_t(F, C, r_f, b_c, s); }
void launch_prolong(const DevLevel& F, const DevLevel& C, float* u_f, const float* u_c, hipStream_t s) { prolong_t(F, C, u_f, u_c, s); }
void launch_prolong(const DevLevel& F, const DevLevel& C, __half* u_f, const __half* u_c, hipStream_t s) { prolong_t(F, C, u_f, u_c, s); }
void launch_scale_to_half(const float* r, __half* hb, const int* page_prob, const float* scal, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_scale_to_half, dim3(grid_for(n)), dim3(256), 0, s, r, hb, page_prob, scal, n);
}
void launch_unscale_project(const DevLevel& L, const __half* hz, float* z, const int* page_prob, const float* scal, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_unscale_project, dim3(L.np + 1), dim3(256), 0, s, L, hz, z, page_prob, scal);
}
void launch_laplace_pcg(const DevLevel& L, const float* v, float* out, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_laplace_pcg<true>, dim3(kXcds * xcd_chunk(L.np)), dim3(256), 0, s, L, v, out);
}
void launch_laplace(const DevLevel& L, const float* v, float* out, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_laplace_pcg<false>, dim3(kXcds * xcd_chunk(L.np)), dim3(256), 0, s, L, v, out);
}
void launch_vector_op(int op, float* out, const float* a, const float* b, float c, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_vector_op, dim3(grid_for(n)), dim3(256), 0, s, op, out, a, b, c, n);
}
void launch_project(const DevLevel& L, float* v, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_project, dim3(L.np), dim3(256), 0, s, L, v);
}
void launch_sub_project(const DevLevel& L, const float* b, const float* q, float* r, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_sub_project, dim3(L.np), dim3(256), 0, s, L, b, q, r);
}
int reduce_parts(size_t n) { return grid_for(n) > 1024 ? 1024 : grid_for(n); }

// Device address of `scratch` on the current device: hipGetSymbolAddress once
// per device, with that device current, kept for later calls (each device
// holds its own copy of the array at its own address; the array's name in
// host code is not a device address).
static double* noalloc_scratch() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* ptr[kMaxDevices] = {};
    int dev = 0;
    hdd_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("inner product (no allocation): device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        hdd_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptr[dev] = static_cast<double*>(p);
    }
    return ptr[dev];
}

void launch_inner_product_noalloc(const float* a, const float* b, int64_t n_pages, double* out, hipStream_t s) {
    const size_t n = (size_t)n_pages * kPageCells;
    const int g = reduce_parts((size_t)(n_pages + 1) * kPageCells);  // the partition of a context of n_pages pages
    if (g > kNoAllocParts) throw std::runtime_error("inner product (no allocation): more partial sums than the scratch holds");
    double* part = noalloc_scratch();
    hipLaunchKernelGGL(k_dot_partial_flat, dim3(g), dim3(256), 0, s, a, b, n, g, part);
    hipLaunchKernelGGL(k_dot_finish_flat, dim3(1), dim3(256), 0, s, (const double*)part, g, out);
    hdd_check(hipGetLastError(), "inner product (no allocation)");
}
void launch_seg_copy2(const DevSegments& S, float* out, const float* in1, const float* in2, int slot, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_copy2, dim3(grid_for(S.ncells)), dim3(256), 0, s, out, in1, in2, (const float*)S.scal, slot, S.page_prob,
                       S.active, S.ncells);
}
void launch_seg_xr_update(const DevSegments& S, float* x, float* r, const float* sv, const float* q, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_xr_update, dim3(grid_for(S.ncells)), dim3(256), 0, s, x, r, sv, q, (const float*)S.scal, S.page_prob,
                       S.active, S.ncells);
}
void launch_seg_dot(const DevSegments& S, const float* a, const float* b, int mode, hipStream_t s) {
    if (S.serial_dots)
        hipLaunchKernelGGL(k_seg_dot_serial, dim3(S.nprob), dim3(256), 0, s, a, b, S.dpart, S.pg0, S.parts, S.max_parts);
    else
        hipLaunchKernelGGL(k_seg_dot_partial, dim3(S.max_parts, S.nprob), dim3(256), 0, s, a, b, S.dpart, S.pg0, S.parts, S.max_parts);
    hipLaunchKernelGGL(k_seg_finish, dim3(S.nprob), dim3(256), 0, s, (const double*)S.dpart, (const float*)nullptr, S.parts, S.max_parts,
                       mode, S.active, S.scal, S.dscal);
}
void launch_seg_copy2_d(const DevSegments& S, float* out, const float* in1, const float* in2, int slot, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_copy2_d, dim3(grid_for(S.ncells)), dim3(256), 0, s, out, in1, in2, (const double*)S.dscal, slot, S.page_prob,
                       S.active, S.ncells);
}
void launch_seg_xr_update_d(const DevSegments& S, double* x, float* r, const float* sv, const float* q, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_xr_update_d, dim3(grid_for(S.ncells)), dim3(256), 0, s, x, r, sv, q, (const double*)S.dscal, S.page_prob,
                       S.active, S.ncells);
}
void launch_true_residual(const DevLevel& L, const double* x, const float* b, float* r, hipStream_t s) {
    if (L.np) hipLaunchKernelGGL(k_true_residual, dim3(kXcds * xcd_chunk(L.np)), dim3(256), 0, s, L, x, b, r);
}
void launch_cast_to_float(const double* x, float* y, size_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_cast_to_float, dim3(grid_for(n)), dim3(256), 0, s, x, y, n);
}
void launch_seg_maxabs(const DevSegments& S, const float* a, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_maxabs_partial, dim3(S.max_parts, S.nprob), dim3(256), 0, s, a, S.fpart, S.pg0, S.parts, S.max_parts);
    hipLaunchKernelGGL(k_seg_finish, dim3(S.nprob), dim3(256), 0, s, (const double*)nullptr, (const float*)S.fpart, S.parts, S.max_parts,
                       2, S.active, S.scal, S.dscal);
}

}  // namespace hdd

namespace hdd {

// b0[i] = r_e - (float)(-sum over Active face neighbours of z, double accumulation)
// for each interface dof i at level-0 cell e (Laplace_ri + Copy_To_Eigen_Array,
// LIB/SPGrid_Laplace_Legacy.h:95-116, IFS/EIGEN_INTERFACE.cpp:18-30)
template <class S>
__global__ void k_if_gather(DevLevel L, const int* __restrict__ if_index, int n, const S* __restrict__ z,
                            const S* __restrict__ r, float* __restrict__ b0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = if_index[i];
    if (e < 0) { b0[i] = 0.f; return; }
    const int p = e >> 8, c = e & 255;
    const uint32_t f = L.flags[e];
    double res = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (f & (kFaceMX << k)) {
            const int q = nb_index(L.nbr, p, c, k);
            if (L.flags[q] & kActive) res -= (double)ld(z, q);
        }
    b0[i] = ld(r, e) - (float)res;
}

template <class S>
__global__ void k_if_scatter(const int* __restrict__ if_index, int n, const float* __restrict__ u0, S* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = if_index[i];
    if (e >= 0) st(z, e, u0[i]);
}

void launch_if_gather(const DevLevel& L0, const int* if_index, int n, const float* z, const float* r, float* b0, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_if_gather<float>, dim3((n + 255) / 256), dim3(256), 0, s, L0, if_index, n, z, r, b0);
}
void launch_if_gather(const DevLevel& L0, const int* if_index, int n, const __half* z, const __half* r, float* b0, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_if_gather<__half>, dim3((n + 255) / 256), dim3(256), 0, s, L0, if_index, n, z, r, b0);
}
void launch_if_scatter(const int* if_index, int n, const float* u0, float* z, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_if_scatter<float>, dim3((n + 255) / 256), dim3(256), 0, s, if_index, n, u0, z);
}
void launch_if_scatter(const int* if_index, int n, const float* u0, __half* z, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_if_scatter<__half>, dim3((n + 255) / 256), dim3(256), 0, s, if_index, n, u0, z);
}

}  // namespace hdd

