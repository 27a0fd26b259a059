// Device-side views shared by the HIP kernels and the host runtime.
#pragma once
#include <hip/hip_runtime.h>
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
struct DevLevel {
    int np = 0;
    int nbx = 0, nby = 0, nbz = 0;
    const uint32_t* flags = nullptr;
    const int* nbr = nullptr;
    const int* base = nullptr;
    const int* lut = nullptr;
};

void upload_constants();
void launch_smooth(const DevLevel& L, const float* u_in, float* u_out, const float* b, bool interior, bool boundary, hipStream_t s);
void launch_residual(const DevLevel& L, const float* u, const float* b, float* r, hipStream_t s);
void launch_fill(float* a, float v, size_t n, hipStream_t s);
void launch_copy(float* d, const float* src, size_t n, hipStream_t s);
void launch_restrict(const DevLevel& F, const DevLevel& C, const float* r_f, float* b_c, hipStream_t s);
void launch_prolong(const DevLevel& F, const DevLevel& C, float* u_f, const float* u_c, hipStream_t s);
void launch_laplace_pcg(const DevLevel& L, const float* v, float* out, hipStream_t s);
void launch_project(const DevLevel& L, float* v, hipStream_t s);
void launch_sub_project(const DevLevel& L, const float* b, const float* q, float* r, hipStream_t s);
void launch_copy2(float* out, const float* in1, const float* in2, const float* cptr, float csign, size_t n, hipStream_t s);
void launch_xr_update(float* x, float* r, const float* sv, const float* q, const float* alpha, size_t n, hipStream_t s);
int reduce_parts(size_t n);
void launch_dot(const float* a, const float* b, size_t n, double* part, int mode, float* scal, double* dscal, hipStream_t s);
void launch_maxabs(const float* a, size_t n, float* part, float* scal, hipStream_t s);

}  // namespace hdd
