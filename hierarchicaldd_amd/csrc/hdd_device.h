// Device-side views shared by the HIP kernels and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <stdexcept>
#include <string>
#include "hdd_grid.h"

namespace hdd {

inline void hdd_check(hipError_t e, const char* what = "HIP call") {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One paged level on the device. flags has (np+1)*256 entries (zero page last),
// nbr np*6 face-neighbour pages (np = zero page), base np*3 cell origins,
// lut the block lookup table over nbx*nby*nbz blocks (-1 = unallocated).
// A stopped problem's pages keep stale values: no kernel reads them for another
// problem (no stencil crosses problems), and the problem's result is final.
struct DevLevel {
    int np = 0;
    int nbx = 0, nby = 0, nbz = 0;
    const uint32_t* flags = nullptr;
    const int* nbr = nullptr;
    const int* base = nullptr;
    const int* lut = nullptr;
    // batch: problem of each page, and per-problem flags (1 = iterating); pages of
    // stopped problems are skipped by the stencil kernels. active == nullptr: none skipped.
    const int* prob = nullptr;
    const int* active = nullptr;
};

void upload_constants();
// V-cycle kernels, float storage and half (FP16) storage with float arithmetic
void launch_smooth(const DevLevel& L, const float* u_in, float* u_out, const float* b, bool interior, bool boundary, hipStream_t s);
void launch_smooth(const DevLevel& L, const __half* u_in, __half* u_out, const __half* b, bool interior, bool boundary, hipStream_t s);
// `count` sweeps (interior and boundary) of a level whose Active cells couple only within clusters of at most
// four pages, in one launch: nwg workgroups of four page slots (wg_page, -1: empty), nb_slot[p*6+f] the slot
// of page p's face-f neighbour when it is in p's cluster, else -1
void launch_smooth_local(const DevLevel& L, const int* wg_page, int nwg, const int* nb_slot, const float* u_in, float* u_out,
                         const float* b, int count, hipStream_t s);
void launch_smooth_local(const DevLevel& L, const int* wg_page, int nwg, const int* nb_slot, const __half* u_in, __half* u_out,
                         const __half* b, int count, hipStream_t s);
void launch_residual(const DevLevel& L, const float* u, const float* b, float* r, hipStream_t s);
void launch_residual(const DevLevel& L, const __half* u, const __half* b, __half* r, hipStream_t s);
void launch_fill(float* a, float v, size_t n, hipStream_t s);
void launch_fill(__half* a, float v, size_t n, hipStream_t s);
void launch_copy(float* d, const float* src, size_t n, hipStream_t s);
void launch_restrict(const DevLevel& F, const DevLevel& C, const float* r_f, float* b_c, hipStream_t s);
void launch_restrict(const DevLevel& F, const DevLevel& C, const __half* r_f, __half* b_c, hipStream_t s);
void launch_prolong(const DevLevel& F, const DevLevel& C, float* u_f, const float* u_c, hipStream_t s);
void launch_prolong(const DevLevel& F, const DevLevel& C, __half* u_f, const __half* u_c, hipStream_t s);
// FP16 storage: hb = r / max|r| per problem (scalar slot 4), and z = max|r| * hz projected
void launch_scale_to_half(const float* r, __half* hb, const int* page_prob, const float* scal, size_t n, hipStream_t s);
void launch_unscale_project(const DevLevel& L, const __half* hz, float* z, const int* page_prob, const float* scal, hipStream_t s);
void launch_laplace_pcg(const DevLevel& L, const float* v, float* out, hipStream_t s);
// out = L v on Active|Interface cells only, other cells untouched (SPGrid_Laplace)
void launch_laplace(const DevLevel& L, const float* v, float* out, hipStream_t s);
// CG_VECTOR operations: 0 out = a*c, 1 out = a*c + b, 2 out += a, 3 out -= a, 4 out *= c
void launch_vector_op(int op, float* out, const float* a, const float* b, float c, size_t n, hipStream_t s);
void launch_project(const DevLevel& L, float* v, hipStream_t s);
void launch_sub_project(const DevLevel& L, const float* b, const float* q, float* r, hipStream_t s);
int reduce_parts(size_t n);
// hdd_inner_product_noalloc: partial sums in the library's file-scope scratch array, no allocation
void launch_inner_product_noalloc(const float* a, const float* b, int64_t n_pages, double* out, hipStream_t s);

// Per-problem view of the level-0 cells of a batch (see the PCG kernels in hdd_kernels.hip).
struct DevSegments {
    int nprob = 0, max_parts = 1;
    size_t ncells = 0;                // cells of the real pages (the zero page excluded)
    const int* pg0 = nullptr;         // [nprob+1] first page of each problem
    const int* parts = nullptr;       // [nprob] reduction blocks of each problem
    const int* page_prob = nullptr;   // [pages+1] problem of each page, -1 for the zero page
    const int* active = nullptr;      // [nprob] 1 while the problem iterates
    double* dpart = nullptr;          // [nprob*max_parts]
    float* fpart = nullptr;           // [nprob*max_parts]
    float* scal = nullptr;            // [nprob*8]
    double* dscal = nullptr;          // [nprob*8]
    bool serial_dots = false;         // inner products in the reference's serial order (k_seg_dot_serial)
};
void launch_seg_copy2(const DevSegments& S, float* out, const float* in1, const float* in2, int slot, hipStream_t s);
void launch_seg_xr_update(const DevSegments& S, float* x, float* r, const float* sv, const float* q, hipStream_t s);
void launch_seg_dot(const DevSegments& S, const float* a, const float* b, int mode, hipStream_t s);
void launch_seg_maxabs(const DevSegments& S, const float* a, hipStream_t s);
// FP64 outer PCG: s-type update and x/r update with the double scalars, the true
// residual b - L x from the double iterate, and the float copy of x
void launch_seg_copy2_d(const DevSegments& S, float* out, const float* in1, const float* in2, int slot, hipStream_t s);
void launch_seg_xr_update_d(const DevSegments& S, double* x, float* r, const float* sv, const float* q, hipStream_t s);
void launch_true_residual(const DevLevel& L, const double* x, const float* b, float* r, hipStream_t s);
void launch_cast_to_float(const double* x, float* y, size_t n, hipStream_t s);

}  // namespace hdd
