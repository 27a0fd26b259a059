// ORACLE — test infrastructure only. CPU restatement of the reference's
// Domain_Decomposition dd_pcg path (demos/Domain_Decomposition_CPU_Only),
// used only as the checker by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg. The product never links, loads or calls it.
//
//   Solver facade         DDS/SPGrid_Domain_Decomposition_Solver.h:73-197
//   DD preconditioner     DDS/SPGrid_Domain_Decomposition_Preconditioner.h:81-128
//   PCG loop              DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212
//   CG system ops         DDS/CG_SYSTEM.cpp:29-144
//   CG vector ops         DDS/SPGrid_Gemini_Arithmetic_Helpers.h:31-263
//   Laplace_ri            LIB/SPGrid_Laplace_Legacy.h:71-118
//   Eigen gather/scatter  IFS/EIGEN_INTERFACE.cpp:18-47
// Main-grid channels: 0=z 1=r 2=k (first allocator), 3=x 4=b 5=q 6=s (second).
#include "dd_sigma.h"
#include "dd_vcycle.h"
#include <chrono>
#include <omp.h>
#include <cstdlib>
#include <cstdio>
#include <memory>
#include <string>

namespace ora {

enum { CZ = 0, CRr = 1, CK = 2, CX = 3, CB = 4, CQ = 5, CS = 6 };

struct Params {
    int size[3], sub[3];
    int vlevels, sigma_levels, nV, int_s, bnd_s, sig_s, sig_mg;
    bool sigma_single;  // single-precision sigma solves (factor rounded to float, float substitution)
};

template <class T>
struct Solver {
    Params P;
    I3 size, sub;
    Grid<T> g;  // main grid, 7 channels
    SigmaSetup<T> S;
    SigmaMG<T> mg;
    VTopology<T> topo;
    std::vector<T> Vbuf[3];  // not used, kept for layout clarity
    double t_setup = 0;
    std::string err;

    // Read_SPGrid (DDS/SPGRID_READ_WRITE.h:31-106) from in-memory arrays.
    void load(int nb, const uint64_t* page_offsets, const uint32_t* flags, const float* b) {
        g.init(0, 0, 0, size.x + BX, size.y + BY, size.z + BZ, 7);
        for (int i = 0; i < nb; ++i) {
            int bx, by, bz;
            morton_decode(page_offsets[i] >> 12, bx, by, bz);
            int p = g.activate_block(bx, by, bz);
            for (int c = 0; c < BCELLS; ++c) {
                g.flags[(size_t)p * BCELLS + c] = flags[(size_t)i * BCELLS + c];
                g.ch[CB][(size_t)p * BCELLS + c] = (T)b[(size_t)i * BCELLS + c];
            }
        }
        g.refresh();
    }

    // Initialize_Preconditioner (DDS/SPGrid_Domain_Decomposition_Solver.h:121-161)
    void setup() {
        auto t0 = std::chrono::steady_clock::now();
        const bool verbose = std::getenv("HDD_ORACLE_VERBOSE") != nullptr;
        auto lap = [&](const char* what) {
            if (verbose) std::fprintf(stderr, "[oracle] %s %.2fs\n", what, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        };
        split_and_generate(g, size, sub, S, P.sigma_levels);
        lap("interface operators");
        domain_divider(g, size, sub);
        flaging(g);
        mg.initialize(S, P.sig_s, P.sigma_single);
        lap("factorisation");
        build_topology(topo, g, size, P.vlevels);
        lap("topology");
        t_setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void clear(int c) { std::fill(g.ch[c].begin(), g.ch[c].end(), T(0)); }
    void copy(int dst, int src) { g.ch[dst] = g.ch[src]; }

    // Laplace_ri: on Interface cells k -= (T)(-sum of Active face neighbours of z), double accumulation
    void laplace_ri() {
        for (int p : g.blocks) {
            int X, Y, Z; g.base(p, X, Y, Z);
            for (int c = 0; c < BCELLS; ++c) {
                size_t e = (size_t)p * BCELLS + c;
                uint32_t f = g.flags[e];
                if (!(f & Interface)) continue;
                int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
                double result = 0.;
                for (int fc = 0; fc < 6; ++fc)
                    if (f & (FaceMX << fc)) {
                        uint32_t nf = g.flag_at(x + FDX[fc], y + FDY[fc], z + FDZ[fc]);
                        if (nf & Active) result -= (double)g.val(CZ, x + FDX[fc], y + FDY[fc], z + FDZ[fc]);
                    }
                g.ch[CK][e] -= (T)result;
            }
        }
    }

    // Apply_Preconditioner (DDS/SPGrid_Domain_Decomposition_Preconditioner.h:88-128):
    // result z, rhs r, tmp k. The diagnostic Compute_Residual calls do not change results.
    void apply_preconditioner() {
        clear(CZ); clear(CK);
        for (int i = 0; i < P.nV; ++i) v_cycle(topo, P.int_s, P.bnd_s);
        copy(CK, CRr);
        laplace_ri();
        auto& b0 = mg.b[0];
        std::fill(b0.begin(), b0.end(), T(0));
        std::fill(mg.r[0].begin(), mg.r[0].end(), T(0));
        for (size_t i = 0; i < S.interface_map.size(); ++i) {
            const I3& c = S.interface_map[i];
            size_t e = g.idx(c.x, c.y, c.z);
            if (e != (size_t)-1 && (g.flags[e] & Interface)) b0[i] = g.ch[CK][e];
        }
        std::fill(mg.u[0].begin(), mg.u[0].end(), T(0));
        for (int i = 0; i < P.sig_mg; ++i) mg.v_cycle();
        clear(CZ); clear(CK);
        for (size_t i = 0; i < S.interface_map.size(); ++i) {
            const I3& c = S.interface_map[i];
            size_t e = g.idx(c.x, c.y, c.z);
            if (e != (size_t)-1 && (g.flags[e] & Interface)) g.ch[CZ][e] = mg.u[0][i];
        }
        for (int i = 0; i < P.nV; ++i) v_cycle(topo, P.int_s, P.bnd_s);
    }

    // CG_SYSTEM::Project: zero cells that are neither Active nor Interface
    void project(int c) {
        for (int p : g.blocks)
            for (int k = 0; k < BCELLS; ++k) {
                size_t e = (size_t)p * BCELLS + k;
                if (!(g.flags[e] & (Active | Interface))) g.ch[c][e] = T(0);
            }
    }
    // SPGrid_Laplace with active_flag Active|Interface, double accumulation
    void multiply(int v, int res) {
        const int n = (int)g.blocks.size();
#pragma omp parallel for schedule(static)
        for (int ib = 0; ib < n; ++ib) {
            int p = g.blocks[ib];
            int X, Y, Z; g.base(p, X, Y, Z);
            for (int c = 0; c < BCELLS; ++c) {
                size_t e = (size_t)p * BCELLS + c;
                uint32_t f = g.flags[e];
                if (!(f & (Active | Interface))) continue;
                int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
                double cv = (double)g.ch[v][e], result = 0.;
                for (int fc = 0; fc < 6; ++fc)
                    if (f & (FaceMX << fc)) result -= ((double)g.val(v, x + FDX[fc], y + FDY[fc], z + FDZ[fc]) - cv);
                g.ch[res][e] = (T)result;
            }
        }
    }
    // Inner_Product: serial double sum in block order of float products
    double inner(int a, int b) const {
        double s = 0;
        for (int p : g.blocks)
            for (int k = 0; k < BCELLS; ++k) {
                size_t e = (size_t)p * BCELLS + k;
                s += g.ch[a][e] * g.ch[b][e];
            }
        return s;
    }
    T max_abs(int a) const {
        T m = 0;
        for (int p : g.blocks)
            for (int k = 0; k < BCELLS; ++k) {
                T v = std::fabs(g.ch[a][(size_t)p * BCELLS + k]);
                m = v > m ? v : m;
            }
        return m;
    }
    // y = c*x1 + x2  (SPGrid_Copy2: out = in1*c + in2)
    void copy2(int out, T c, int in1, int in2) {
        for (int p : g.blocks)
            for (int k = 0; k < BCELLS; ++k) {
                size_t e = (size_t)p * BCELLS + k;
                g.ch[out][e] = g.ch[in1][e] * c + g.ch[in2][e];
            }
    }

    // CONJUGATE_GRADIENT_DUMPING::Solve with the stopping rule max|r| <= tol_abs
    int solve(T tol_abs, int max_it, std::vector<double>& hist) {
        hist.clear();
        const T small_number = std::numeric_limits<T>::epsilon();
        T rho_old = (T)3.40282347e+38f;
        int it;
        for (it = 0;; ++it) {
            bool restart = (it == 0);
            if (restart) { copy(CRr, CB); multiply(CX, CQ); for (int p : g.blocks) for (int k = 0; k < BCELLS; ++k) { size_t e = (size_t)p * BCELLS + k; g.ch[CRr][e] -= g.ch[CQ][e]; } project(CRr); }
            T conv = max_abs(CRr);
            hist.push_back((double)conv);
            if (conv <= tol_abs && (it >= 0 || conv < small_number)) return it;
            if (it == max_it) break;
            apply_preconditioner();
            project(CZ);
            T rho = (T)inner(CZ, CRr);
            if (restart) copy(CS, CZ);
            else copy2(CS, rho / rho_old, CS, CZ);
            multiply(CS, CQ);
            project(CQ);
            T sdq = (T)inner(CS, CQ);
            T alpha = sdq ? rho / sdq : (T)3.40282347e+38f;
            copy2(CX, alpha, CS, CX);
            copy2(CRr, -alpha, CQ, CRr);
            rho_old = rho;
        }
        return -it;
    }
};

}  // namespace ora

using namespace ora;

struct OraHandle {
    int is_double;
    std::unique_ptr<Solver<float>> f;
    std::unique_ptr<Solver<double>> d;
    std::string err;
};

#define ORA_DISPATCH(h, expr) ((h)->is_double ? ([&](auto& s) { return expr; })(*(h)->d) : ([&](auto& s) { return expr; })(*(h)->f))

extern "C" {

// Builds the solver and runs the full setup. Returns nullptr on failure
// (message in *err_out if given). Arrays follow the reference file format:
// page_offsets are SPGrid linear offsets of 4 KiB pages; flags / b hold 256
// values per page in page-local (x,y,z) order. sigma_single selects the
// single-precision sigma solves (dd_sigma.h), 0 the exact (double) ones.
void* ora_create(int is_double, int nb, const uint64_t* page_offsets, const uint32_t* flags, const float* b,
                 const int* size3, const int* sub3, int vlevels, int sigma_levels, int nV, int int_s, int bnd_s,
                 int sig_s, int sig_mg, int sigma_single, char* err_out, int err_len) {
    auto* h = new OraHandle();
    h->is_double = is_double;
    try {
        auto init = [&](auto& s) {
            s.P = Params{{size3[0], size3[1], size3[2]}, {sub3[0], sub3[1], sub3[2]}, vlevels, sigma_levels, nV, int_s, bnd_s, sig_s, sig_mg,
                         sigma_single != 0};
            s.size = {size3[0], size3[1], size3[2]};
            s.sub = {sub3[0], sub3[1], sub3[2]};
            s.load(nb, page_offsets, flags, b);
            s.setup();
        };
        if (is_double) { h->d.reset(new Solver<double>()); init(*h->d); }
        else { h->f.reset(new Solver<float>()); init(*h->f); }
    } catch (std::exception& e) {
        if (err_out && err_len > 0) { std::snprintf(err_out, err_len, "%s", e.what()); }
        delete h;
        return nullptr;
    }
    return h;
}

void ora_destroy(void* hp) { delete (OraHandle*)hp; }

// OpenMP threads of later oracle calls (the reference's -threads N)
void ora_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int ora_max_threads() { return omp_get_max_threads(); }

double ora_setup_seconds(void* hp) { auto* h = (OraHandle*)hp; return ORA_DISPATCH(h, s.t_setup); }

int ora_num_pages(void* hp) { auto* h = (OraHandle*)hp; return ORA_DISPATCH(h, (int)s.g.blocks.size()); }

// page offsets of the main grid in block order
void ora_page_offsets(void* hp, uint64_t* out) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { for (size_t i = 0; i < s.g.blocks.size(); ++i) { int p = s.g.blocks[i]; out[i] = morton_page(s.g.pbx[p], s.g.pby[p], s.g.pbz[p]) << 12; } return 0; })());
}

// level count / page count / flags (block order) of V-cycle level l
int ora_level_pages(void* hp, int l) { auto* h = (OraHandle*)hp; return ORA_DISPATCH(h, (int)s.topo.lv[l]->blocks.size()); }
void ora_level_data(void* hp, int l, uint64_t* offsets, uint32_t* flags) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] {
        auto& G = *s.topo.lv[l];
        for (size_t i = 0; i < G.blocks.size(); ++i) {
            int p = G.blocks[i];
            if (offsets) offsets[i] = morton_page(G.pbx[p], G.pby[p], G.pbz[p]) << 12;
            if (flags) std::memcpy(flags + i * BCELLS, &G.flags[(size_t)p * BCELLS], BCELLS * sizeof(uint32_t));
        }
        return 0; })());
}

// channel access on the main grid (block order); ch 0..6 as documented above
void ora_get_channel(void* hp, int c, double* out) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { size_t n = 0; for (int p : s.g.blocks) for (int k = 0; k < BCELLS; ++k) out[n++] = (double)s.g.ch[c][(size_t)p * BCELLS + k]; return 0; })());
}
void ora_set_channel(void* hp, int c, const double* in) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { size_t n = 0; for (int p : s.g.blocks) for (int k = 0; k < BCELLS; ++k) s.g.ch[c][(size_t)p * BCELLS + k] = (typename std::remove_reference<decltype(s.g.ch[0][0])>::type)in[n++]; return 0; })());
}

// one V-cycle on the main grid: u=z(ch0) from b=r(ch1), tmp k(ch2); u and k are cleared first
void ora_vcycle(void* hp) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { s.clear(CZ); s.clear(CK); v_cycle(s.topo, s.P.int_s, s.P.bnd_s); return 0; })());
}
// one smoothing sweep on level 0: mode 0 interior, 1 boundary, 2 all Active
void ora_smooth0(void* hp, int mode) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { smooth(s.g, mode == 0 || mode == 2, mode == 1 || mode == 2); return 0; })());
}
// z = M^{-1} r on the main grid (r in ch1, z out in ch0), followed by Project(z)
void ora_apply_preconditioner(void* hp) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { s.apply_preconditioner(); s.project(CZ); return 0; })());
}

int ora_sigma_levels(void* hp) { auto* h = (OraHandle*)hp; return ORA_DISPATCH(h, s.S.mg_levels); }
int ora_num_subdomains(void* hp) { auto* h = (OraHandle*)hp; return ORA_DISPATCH(h, s.S.nsub); }
int ora_interface_dofs(void* hp, int l) { auto* h = (OraHandle*)hp; return ORA_DISPATCH(h, s.S.n_if[l]); }
// level-l interface coordinates (level units), 3 ints per dof
void ora_interface_coords(void* hp, int l, int* xyz) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { auto& v = s.S.icoord[l]; for (size_t i = 0; i < v.size(); ++i) { xyz[3 * i] = v[i].x; xyz[3 * i + 1] = v[i].y; xyz[3 * i + 2] = v[i].z; } return 0; })());
}
// matrix export: which 0=Arr 1=Aii 2=Air 3=Prr; returns nnz, rows/cols via out params
int ora_matrix_shape(void* hp, int which, int l, int sd, int* rows, int* cols) {
    auto* h = (OraHandle*)hp;
    return ORA_DISPATCH(h, ([&] {
        auto& M = which == 0 ? s.S.Arr[l] : which == 1 ? s.S.Aii[l][sd] : which == 2 ? s.S.Air[l][sd] : s.S.Prr[l];
        *rows = M.rows; *cols = M.cols; return (int)M.idx.size(); })());
}
void ora_matrix_data(void* hp, int which, int l, int sd, int* ptr, int* idx, double* val) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] {
        auto& M = which == 0 ? s.S.Arr[l] : which == 1 ? s.S.Aii[l][sd] : which == 2 ? s.S.Air[l][sd] : s.S.Prr[l];
        for (size_t i = 0; i < M.ptr.size(); ++i) ptr[i] = M.ptr[i];
        for (size_t i = 0; i < M.idx.size(); ++i) { idx[i] = M.idx[i]; val[i] = (double)M.val[i]; }
        return 0; })());
}
// sigma V-cycle on a level-0 interface vector b (n_if[0]), result u
void ora_sigma_vcycle(void* hp, const double* b, double* u) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] {
        auto& mg = s.mg;
        using TT = typename std::remove_reference<decltype(mg.b[0][0])>::type;
        for (size_t i = 0; i < mg.b[0].size(); ++i) mg.b[0][i] = (TT)b[i];
        std::fill(mg.u[0].begin(), mg.u[0].end(), TT(0));
        mg.v_cycle();
        for (size_t i = 0; i < mg.u[0].size(); ++i) u[i] = (double)mg.u[0][i];
        return 0; })());
}
// one sigma smooth on level l: u <- Smooth(u, b)
void ora_sigma_smooth(void* hp, int l, const double* b, double* u) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] {
        auto& mg = s.mg;
        using TT = typename std::remove_reference<decltype(mg.b[0][0])>::type;
        size_t n = mg.b[l].size();
        std::vector<TT> bb(n), uu(n), tmp(n);
        for (size_t i = 0; i < n; ++i) { bb[i] = (TT)b[i]; uu[i] = (TT)u[i]; }
        mg.smooth(l, uu, tmp, bb);
        for (size_t i = 0; i < n; ++i) u[i] = (double)uu[i];
        return 0; })());
}

// STRIDE_GENERATOR::Generate_Adaptive_Stride for one axis over [lo, hi]:
// q[0 .. hi-lo] receives the strides, *level the adaptation level.
void ora_adaptive_stride(int lo, int hi, int* q, int* level) {
    std::vector<int> v;
    *level = adaptive_stride_axis(lo, hi, v);
    for (size_t i = 0; i < v.size(); ++i) q[i] = v[i];
}

// channel c (0 u, 1 b, 2 r) of V-cycle level l in block order (level 0: the
// main grid's z, r, k)
void ora_level_get(void* hp, int l, int c, double* out) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] { auto& G = *s.topo.lv[l]; size_t n = 0; for (int p : G.blocks) for (int k = 0; k < BCELLS; ++k) out[n++] = (double)G.ch[c][(size_t)p * BCELLS + k]; return 0; })());
}
void ora_level_set(void* hp, int l, int c, const double* in) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] {
        auto& G = *s.topo.lv[l];
        using TT = typename std::remove_reference<decltype(G.ch[0][0])>::type;
        size_t n = 0;
        for (int p : G.blocks) for (int k = 0; k < BCELLS; ++k) G.ch[c][(size_t)p * BCELLS + k] = (TT)in[n++];
        return 0; })());
}
// single V-cycle kernels on level l: 0 residual, 1 correction (mode 0 interior,
// 1 boundary, 2 all Active), 2 restriction l -> l+1, 3 prolongation l+1 -> l
void ora_level_kernel(void* hp, int l, int kernel, int mode) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] {
        auto& T = s.topo;
        if (kernel == 0) vc_residual(*T.lv[l]);
        else if (kernel == 1) vc_correction(*T.lv[l], mode == 0 || mode == 2, mode == 1 || mode == 2);
        else if (kernel == 2) vc_restriction(*T.lv[l], *T.lv[l + 1]);
        else vc_prolongation(*T.lv[l], *T.lv[l + 1]);
        return 0; })());
}
// x = (A + 1e-5 I)^{-1} b with the sigma solver's factor of level l
// (which 0: Arr, 1: Aii of subdomain sd), in the oracle's solve precision
void ora_sigma_factor_solve(void* hp, int l, int which, int sd, const double* b, double* x) {
    auto* h = (OraHandle*)hp;
    ORA_DISPATCH(h, ([&] {
        const Cholesky& F = which == 0 ? s.mg.Lrr[l] : s.mg.Lii[l][sd];
        std::vector<double> bb(b, b + F.n), xx(F.n);
        F.solve(bb.data(), xx.data());
        for (int i = 0; i < F.n; ++i) x[i] = xx[i];
        return 0; })());
}

// Full PCG solve from x=0. tol_abs on max|r|. Returns iterations (negative if
// not converged within max_it); hist receives max|r| per iteration (size max_it+1).
int ora_solve(void* hp, double tol_abs, int max_it, double* hist, int hist_len, double* seconds) {
    auto* h = (OraHandle*)hp;
    return ORA_DISPATCH(h, ([&] {
        using TT = typename std::remove_reference<decltype(s.g.ch[0][0])>::type;
        s.clear(CX);
        std::vector<double> hv;
        auto t0 = std::chrono::steady_clock::now();
        int it = s.solve((TT)tol_abs, max_it, hv);
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < (int)hv.size() && i < hist_len; ++i) hist[i] = hv[i];
        return it; })());
}

}  // extern "C"
