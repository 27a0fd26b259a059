/*
 * hdd_krylov.hpp — header-only C++ classes shaped like the reference's Krylov
 * plugin, over the device-channel entry points of hdd.h.
 *
 *   CgVector   ~ CG_VECTOR<T_STRUCT,float,3> (DDS/CG_VECTOR.h:25-70,
 *                DDS/CG_VECTOR.cpp:18-74): one SPGrid channel of a context
 *   CgSystem   ~ CG_SYSTEM (DDS/CG_SYSTEM.h:23-54, DDS/CG_SYSTEM.cpp:29-144)
 *                with the KRYLOV_SYSTEM_BASE interface
 *                (PhysBAM_Tools/Krylov_Solvers/KRYLOV_SYSTEM_BASE.h:15-43)
 *   cg_dumping_solve ~ CONJUGATE_GRADIENT_DUMPING<float,3>::Solve
 *                (DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212), written against
 *                the two classes only (its visualisation calls are no-ops in
 *                the reference, :36-43, and are left out)
 *
 * Every vector stays in device memory; the only values that reach the host
 * are the inner products and norms the loop branches on. Errors throw
 * hdd_krylov::Error with hdd_last_error's message. A one-problem context
 * (hdd_create) is assumed: the reductions return problem 0's value.
 */
#ifndef HDD_KRYLOV_HPP
#define HDD_KRYLOV_HPP
#include <limits>
#include <stdexcept>
#include <string>
#include "hdd.h"

namespace hdd_krylov {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline void check(hdd_ctx* ctx, hdd_status st) {
    if (st != HDD_OK) throw Error(std::string("hdd: ") + hdd_last_error(ctx));
}

class CgVector {
public:
    CgVector(hdd_ctx* ctx, int channel) : ctx_(ctx), ch_(channel) {}
    CgVector(const CgVector&) = delete;
    // KRYLOV_VECTOR_BASE::operator= copies values (Copy(1, v))
    CgVector& operator=(const CgVector& v) {
        if (this != &v) Copy(1.0f, v);
        return *this;
    }
    CgVector& operator+=(const CgVector& v) { check(ctx_, hdd_vector_plus_equal(ctx_, ch_, v.ch_)); return *this; }
    CgVector& operator-=(const CgVector& v) { check(ctx_, hdd_vector_minus_equal(ctx_, ch_, v.ch_)); return *this; }
    CgVector& operator*=(float a) { check(ctx_, hdd_vector_scale(ctx_, ch_, a)); return *this; }
    // this = c * v
    void Copy(float c, const CgVector& v) { check(ctx_, hdd_vector_copy(ctx_, ch_, c, v.ch_)); }
    // this = c * v1 + v2
    void Copy(float c, const CgVector& v1, const CgVector& v2) { check(ctx_, hdd_vector_copy2(ctx_, ch_, c, v1.ch_, v2.ch_)); }
    int channel() const { return ch_; }
    hdd_ctx* context() const { return ctx_; }

private:
    hdd_ctx* ctx_;
    int ch_;
};

class CgSystem {
public:
    bool use_preconditioner = true;  // the reference's Solve sets it (DDS/SPGrid_Domain_Decomposition_Solver.h:176)

    // tmp_channel: the preconditioner's scratch field (the reference passes ch2 = k)
    explicit CgSystem(hdd_ctx* ctx, int tmp_channel = HDD_CH_K) : ctx_(ctx), tmp_(tmp_channel) {}

    void Multiply(const CgVector& x, CgVector& result) const { check(ctx_, hdd_multiply(ctx_, x.channel(), result.channel())); }
    double Inner_Product(const CgVector& x, const CgVector& y) const {
        double v = 0;
        check(ctx_, hdd_inner_product(ctx_, x.channel(), y.channel(), &v));
        return v;
    }
    float Convergence_Norm(const CgVector& x) const {
        float v = 0;
        check(ctx_, hdd_convergence_norm(ctx_, x.channel(), &v));
        return v;
    }
    void Project(CgVector& x) const { check(ctx_, hdd_project(ctx_, x.channel())); }
    void Set_Boundary_Conditions(CgVector&) const {}
    void Project_Nullspace(CgVector&) const {}
    // KRYLOV_SYSTEM_BASE::Precondition: Apply_Preconditioner, then Project (the
    // DD preconditioner does not commute with the projection); the device call
    // applies the projection itself
    const CgVector& Precondition(const CgVector& r, CgVector& z) const {
        if (!use_preconditioner) return r;
        check(ctx_, hdd_apply_preconditioner_channels(ctx_, z.channel(), r.channel(), tmp_));
        return z;
    }

private:
    hdd_ctx* ctx_;
    int tmp_;
};

// Returns true when converged (convergence_norm <= tolerance); *iterations and
// the history (max|r| per iteration, when not null, up to history_len values)
// are filled as the reference's print_residuals output would be.
inline bool cg_dumping_solve(const CgSystem& system, CgVector& x, const CgVector& b, CgVector& q, CgVector& s, CgVector& r,
                             CgVector& k, CgVector& z, float tolerance, int min_iterations, int max_iterations, int* iterations_used,
                             double* history = nullptr, int history_len = 0) {
    (void)k;  // the preconditioner's tmp field belongs to the system (CG_SYSTEM's constructor argument)
    const float small_number = std::numeric_limits<float>::epsilon();
    CgVector& xs = x;
    system.Set_Boundary_Conditions(xs);
    float rho_old = std::numeric_limits<float>::max();
    float convergence_norm = 0;
    int iterations;
    for (iterations = 0;; iterations++) {
        const bool restart = !iterations;
        if (restart) {
            r = b;
            system.Multiply(x, q);
            r -= q;
            system.Project(r);
        }
        system.Project_Nullspace(r);
        convergence_norm = system.Convergence_Norm(r);
        if (history && iterations < history_len) history[iterations] = convergence_norm;
        if (convergence_norm <= tolerance && (iterations >= min_iterations || convergence_norm < small_number)) {
            if (iterations_used) *iterations_used = iterations;
            return true;
        }
        if (iterations == max_iterations) break;
        const CgVector& mr = system.Precondition(r, z);
        const float rho = (float)system.Inner_Product(mr, r);
        if (restart) s = mr;
        else s.Copy(rho / rho_old, s, mr);
        system.Multiply(s, q);
        system.Project(q);
        const float s_dot_q = (float)system.Inner_Product(s, q);
        const float alpha = s_dot_q ? rho / s_dot_q : std::numeric_limits<float>::max();
        x.Copy(alpha, s, x);
        r.Copy(-alpha, q, r);
        rho_old = rho;
    }
    if (iterations_used) *iterations_used = iterations;
    return false;
}

}  // namespace hdd_krylov
#endif
