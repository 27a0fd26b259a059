// Test driver of include/hdd_krylov.hpp: the reference's PCG loop
// (CONJUGATE_GRADIENT_DUMPING::Solve, restated in the header) run through the
// CG_SYSTEM / CG_VECTOR-shaped classes on a context's channels, with the
// vectors the reference's Solve passes (DDS/SPGrid_Domain_Decomposition_Solver.h:165-195).
// Built by hierarchicaldd_amd/build.py into _build/libhdd_krylov_check.so and
// called from tests/test_gpu_krylov.py.
#include "hdd_krylov.hpp"
#include <cstdio>

extern "C" int hdd_krylov_check_solve(hdd_ctx* ctx, float tolerance, int max_iterations, double* history, int history_len,
                                      int* iterations, char* err, int err_len) {
    using namespace hdd_krylov;
    try {
        CgVector x(ctx, HDD_CH_X), b(ctx, HDD_CH_B), q(ctx, HDD_CH_Q), s(ctx, HDD_CH_S), r(ctx, HDD_CH_R), k(ctx, HDD_CH_K),
            z(ctx, HDD_CH_Z);
        CgSystem system(ctx, HDD_CH_K);
        system.use_preconditioner = true;
        const bool ok = cg_dumping_solve(system, x, b, q, s, r, k, z, tolerance, 0, max_iterations, iterations, history, history_len);
        check(ctx, hdd_synchronize(ctx));
        return ok ? 1 : 0;
    } catch (std::exception& e) {
        if (err && err_len > 0) std::snprintf(err, err_len, "%s", e.what());
        return -1;
    }
}
