// Dense fronts of the multifrontal factorisation on the GPU.
//
// The interface solver's factors replace the reference's PARDISO phases 11/22
// (IFS/SIGMA_SMOOTHER.cpp:115-288). Near the root of a nested-dissection tree
// the fronts are dense matrices of thousands of rows, and their partial
// Cholesky is most of the set-up time on the host. A front of at least
// kDeviceFrontMin rows goes to the GPU: the pivot block by rocSOLVER's dpotrf,
// the sub-diagonal block by rocBLAS dtrsm, the Schur update by dsyrk, all in
// FP64 on the null stream, with the front copied in and out whole. The
// arithmetic is the same partial Cholesky as the host's (only the rounding
// order differs); the factors are then laid out on the host as before.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include "hdd_factor.h"

namespace hdd {

namespace {

void rb_check(rocblas_status st, const char* what) {
    if (st != rocblas_status_success) throw std::runtime_error(std::string("device front: ") + what + " failed (rocBLAS status " + std::to_string((int)st) + ")");
}
void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("device front: ") + what + ": " + hipGetErrorString(e));
}

struct DeviceFrontState {
    rocblas_handle handle = nullptr;
    double* buf = nullptr;
    size_t cap = 0;  // doubles
    rocblas_int* info = nullptr;
};

}  // namespace

bool device_fronts_enabled() {
    static const bool have_device = [] {
        int n = 0;
        return hipGetDeviceCount(&n) == hipSuccess && n > 0;
    }();
    const char* e = std::getenv("HDD_DEVICE_FRONTS");
    return have_device && !(e && std::string(e) == "0");
}

int device_front_min_rows() {
    const char* e = std::getenv("HDD_DEVICE_FRONT_MIN");
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : kDeviceFrontMin;
}

// Same contract as the host partial_cholesky: F is the m x m lower front,
// column-major; on return its first ns columns hold L11 and L21 and its
// trailing block the update F22 - L21 L21^T (lower triangles).
void partial_cholesky_device(double* F, int m, int ns) {
    static std::mutex mu;
    static std::map<int, DeviceFrontState> states;  // per device: a handle belongs to the device current at its creation
    std::lock_guard<std::mutex> lock(mu);
    int dev = 0;
    hip_ok(hipGetDevice(&dev), "hipGetDevice");
    DeviceFrontState& S = states[dev];
    if (!S.handle) {
        rb_check(rocblas_create_handle(&S.handle), "create handle");
        hip_ok(hipMalloc(&S.info, sizeof(rocblas_int)), "hipMalloc");
    }
    rb_check(rocblas_set_stream(S.handle, nullptr), "set stream");
    const size_t cells = (size_t)m * m;
    if (S.cap < cells) {
        if (S.buf) hip_ok(hipFree(S.buf), "hipFree");
        S.buf = nullptr;
        hip_ok(hipMalloc(&S.buf, cells * sizeof(double)), "hipMalloc");
        S.cap = cells;
    }
    hip_ok(hipMemcpy(S.buf, F, cells * sizeof(double), hipMemcpyHostToDevice), "copy in");
    rb_check(rocsolver_dpotrf(S.handle, rocblas_fill_lower, ns, S.buf, m, S.info), "dpotrf");
    const int nr = m - ns;
    if (nr > 0) {
        const double one = 1.0, minus_one = -1.0;
        // L21 = F21 L11^{-T}
        rb_check(rocblas_dtrsm(S.handle, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose, rocblas_diagonal_non_unit, nr, ns,
                               &one, S.buf, m, S.buf + ns, m),
                 "dtrsm");
        // F22 -= L21 L21^T (lower triangle)
        rb_check(rocblas_dsyrk(S.handle, rocblas_fill_lower, rocblas_operation_none, nr, ns, &minus_one, S.buf + ns, m, &one,
                               S.buf + ns + (size_t)ns * m, m),
                 "dsyrk");
    }
    rocblas_int info = 0;
    hip_ok(hipMemcpy(&info, S.info, sizeof(rocblas_int), hipMemcpyDeviceToHost), "copy info");
    if (info != 0) throw std::runtime_error("supernodal Cholesky (device front): matrix is not positive definite");
    hip_ok(hipMemcpy(F, S.buf, cells * sizeof(double), hipMemcpyDeviceToHost), "copy out");
}

}  // namespace hdd
