#include "hdd_runtime.h"
#include "hdd_sigma_device.h"
#include <chrono>
#include <cmath>

namespace hdd {

DevLevel LevelBuffers::view() const {
    DevLevel v;
    v.np = host.npages();
    v.nbx = host.nbx; v.nby = host.nby; v.nbz = host.nbz;
    v.flags = flags.p; v.nbr = nbr.p; v.base = base.p; v.lut = lut.p;
    return v;
}

void Solver::build_levels(int n_pages, const uint64_t* page_offsets, const uint32_t* flags) {
    lv.clear();
    lv.resize(P.vlevels);
    PagedLevel& L0 = lv[0].host;
    L0.init(P.size);
    for (int i = 0; i < n_pages; ++i) {
        int bx, by, bz;
        page_coords(page_offsets[i] >> 12, bx, by, bz);
        const int p = L0.add_page(bx, by, bz);
        std::copy(flags + (size_t)i * kPageCells, flags + (size_t)(i + 1) * kPageCells, L0.flags.begin() + (size_t)p * kPageCells);
    }
    L0.sort_pages();
}

Solver::Solver(const SolverParams& p, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, const float* b, bool host)
    : P(p), host_only(host) {
    auto t0 = std::chrono::steady_clock::now();
    if (P.vlevels < 1) throw std::runtime_error("vlevels must be >= 1");
    build_levels(n_pages, page_offsets, flags);
    // the interface operators are generated from the input flags, before division
    // (Initialize_Preconditioner: Split precedes Domain_Divider, DDS/...Solver.h:122-133)
    generate_interface_operators(lv[0].host, P.size, P.sub, P.sigma_levels, sigma_host);
    divide_subdomains(lv[0].host, P.size, P.sub);
    set_face_flags(lv[0].host);
    for (int l = 0; l + 1 < P.vlevels; ++l) {
        const Int3 s = lv[l].host.size;
        lv[l + 1].host.init({s.x / 2 + 1, s.y / 2 + 1, s.z / 2 + 1});
        downsample_level(lv[l].host, lv[l + 1].host);
    }
    for (int l = 0; l < P.vlevels; ++l) {
        set_face_flags(lv[l].host);
        set_boundary_flags(lv[l].host);
    }
    if (host_only) {
        setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return;
    }
    hdd_check(hipStreamCreate(&stream), "hipStreamCreate");
    upload_constants();
    upload_levels();
    // rhs b in page order of level 0 (input order may differ: remap through the LUT)
    {
        const PagedLevel& L0 = lv[0].host;
        std::vector<float> hb((size_t)(L0.npages() + 1) * kPageCells, 0.f);
        for (int i = 0; i < n_pages; ++i) {
            int bx, by, bz;
            page_coords(page_offsets[i] >> 12, bx, by, bz);
            const int q = L0.page_at_block(bx, by, bz);
            std::copy(b + (size_t)i * kPageCells, b + (size_t)(i + 1) * kPageCells, hb.begin() + (size_t)q * kPageCells);
        }
        bvec.upload(hb);
    }
    const size_t n0 = lv[0].cells();
    x.alloc(n0); q.alloc(n0); s.alloc(n0);
    dpart.alloc(1024); dscal.alloc(4); fpart.alloc(1024); scal.alloc(8);
    // interface dof -> level-0 cell index
    {
        const PagedLevel& L0 = lv[0].host;
        std::vector<int> idx(sigma_host.interface_map.size());
        for (size_t i = 0; i < idx.size(); ++i) {
            const Int3 c = sigma_host.interface_map[i];
            const int pg = L0.page_at(c.x, c.y, c.z);
            const bool ok = pg >= 0 && (L0.flags[(size_t)pg * kPageCells + page_cell(c.x, c.y, c.z)] & kInterface);
            idx[i] = ok ? pg * kPageCells + page_cell(c.x, c.y, c.z) : -1;
        }
        if_index.upload(idx);
    }
    sigma = make_sigma_device(sigma_host, P.sig_s, P.sig_mg);
    hdd_check(hipDeviceSynchronize(), "setup sync");
    setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Solver::~Solver() {
    sigma.reset();
    if (stream) (void)hipStreamDestroy(stream);
}

void Solver::upload_levels() {
    static const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (auto& L : lv) {
        const PagedLevel& H = L.host;
        const int np = H.npages();
        std::vector<uint32_t> f(H.flags);
        f.resize((size_t)(np + 1) * kPageCells, 0u);
        L.flags.upload(f);
        std::vector<int> nb((size_t)np * 6), base((size_t)np * 3);
        for (int p = 0; p < np; ++p) {
            const Int3 o = H.page_base[p];
            base[p * 3] = o.x; base[p * 3 + 1] = o.y; base[p * 3 + 2] = o.z;
            for (int k = 0; k < 6; ++k) {
                const int q = H.page_at(o.x + d[k][0] * kPX, o.y + d[k][1] * kPY, o.z + d[k][2] * kPZ);
                nb[p * 6 + k] = q < 0 ? np : q;
            }
        }
        L.nbr.upload(nb);
        L.base.upload(base);
        L.lut.upload(H.lut);
        L.u[0].alloc(L.cells()); L.u[1].alloc(L.cells());
        L.b.alloc(L.cells()); L.r.alloc(L.cells());
        L.cur = 0;
    }
}

float* Solver::channel(int c) {
    LevelBuffers& L0 = lv[0];
    switch (c) {
        case 0: return L0.u[L0.cur].p;
        case 1: return L0.b.p;
        case 2: return L0.r.p;
        case 3: return x.p;
        case 4: return bvec.p;
        case 5: return q.p;
        case 6: return s.p;
    }
    throw std::runtime_error("bad channel id");
}

void Solver::sweeps(int l, int count, bool interior, bool boundary) {
    LevelBuffers& L = lv[l];
    const DevLevel v = L.view();
    for (int i = 0; i < count; ++i) {
        launch_smooth(v, L.u[L.cur].p, L.u[1 - L.cur].p, L.b.p, interior, boundary, stream);
        L.cur = 1 - L.cur;
    }
}

void Solver::smooth0(int mode) { sweeps(0, 1, mode == 0 || mode == 2, mode == 1 || mode == 2); }

// V_Cycle_Helper::v_cycle (LIB/SPGrid_V_Cycle_Helper.h:156-192). The two
// bottom-level residual passes only write the residual channel, which nothing
// reads before it is recomputed, so they are not launched.
void Solver::vcycle() {
    const int L = (int)lv.size();
    for (int l = 0; l < L - 1; ++l) {
        sweeps(l, P.bnd_s, false, true);
        sweeps(l, P.int_s, true, false);
        sweeps(l, P.bnd_s, false, true);
        launch_residual(lv[l].view(), lv[l].u[lv[l].cur].p, lv[l].b.p, lv[l].r.p, stream);
        launch_fill(lv[l + 1].u[lv[l + 1].cur].p, 0.f, lv[l + 1].cells(), stream);
        launch_restrict(lv[l].view(), lv[l + 1].view(), lv[l].r.p, lv[l + 1].b.p, stream);
    }
    sweeps(L - 1, 200, true, true);
    for (int l = L - 2; l >= 0; --l) {
        launch_prolong(lv[l].view(), lv[l + 1].view(), lv[l].u[lv[l].cur].p, lv[l + 1].u[lv[l + 1].cur].p, stream);
        sweeps(l, P.bnd_s, false, true);
        sweeps(l, P.int_s, true, false);
        sweeps(l, P.bnd_s, false, true);
    }
}

// Apply_Preconditioner (DDS/SPGrid_Domain_Decomposition_Preconditioner.h:88-128),
// z = channel 0, rhs = channel 1 (PCG residual r), tmp k = channel 2.
void Solver::apply_preconditioner() {
    LevelBuffers& L0 = lv[0];
    const size_t n0 = L0.cells();
    launch_fill(L0.u[L0.cur].p, 0.f, n0, stream);
    for (int i = 0; i < P.nV; ++i) vcycle();
    interface_gather();
    sigma_vcycles(*sigma, stream);
    launch_fill(L0.u[L0.cur].p, 0.f, n0, stream);
    interface_scatter();
    for (int i = 0; i < P.nV; ++i) vcycle();
    launch_project(L0.view(), L0.u[L0.cur].p, stream);
}

void Solver::interface_gather() {
    LevelBuffers& L0 = lv[0];
    launch_interface_gather(*sigma, L0.view(), if_index.p, L0.u[L0.cur].p, L0.b.p, stream);
}
void Solver::interface_scatter() {
    LevelBuffers& L0 = lv[0];
    launch_interface_scatter(*sigma, if_index.p, L0.u[L0.cur].p, stream);
}

// CONJUGATE_GRADIENT_DUMPING::Solve (DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212)
// with the stopping rule max|r| <= tol_abs, x starting at zero.
int Solver::pcg(double tol_abs, int max_it, std::vector<double>& hist, double* seconds) {
    LevelBuffers& L0 = lv[0];
    const DevLevel v0 = L0.view();
    const size_t n0 = L0.cells();
    float* r = L0.b.p;
    hist.clear();
    const float flt_max = 3.40282347e+38f;
    float init[8] = {0, 0, 0, flt_max, 0, 0, 0, 0};
    hdd_check(hipMemcpyAsync(scal.p, init, sizeof(init), hipMemcpyHostToDevice, stream));
    launch_fill(x.p, 0.f, n0, stream);
    hdd_check(hipStreamSynchronize(stream));
    auto t0 = std::chrono::steady_clock::now();
    const float small_number = 1.1920929e-07f;
    int it = 0;
    for (it = 0;; ++it) {
        const bool restart = (it == 0);
        if (restart) {
            launch_laplace_pcg(v0, x.p, q.p, stream);
            launch_sub_project(v0, bvec.p, q.p, r, stream);
        }
        launch_maxabs(r, n0, fpart.p, scal.p, stream);
        float conv = 0.f;
        hdd_check(hipMemcpyAsync(&conv, scal.p + 4, sizeof(float), hipMemcpyDeviceToHost, stream));
        hdd_check(hipStreamSynchronize(stream));
        hist.push_back((double)conv);
        if (conv <= (float)tol_abs && (it >= 0 || conv < small_number)) break;
        if (it == max_it) { it = -it; break; }
        apply_preconditioner();
        float* z = L0.u[L0.cur].p;
        launch_dot(z, r, n0, dpart.p, 0, scal.p, dscal.p, stream);          // rho, beta
        if (restart) launch_copy(s.p, z, n0, stream);
        else launch_copy2(s.p, s.p, z, scal.p + 2, 1.f, n0, stream);         // s = s*beta + z
        launch_laplace_pcg(v0, s.p, q.p, stream);                          // q = Project(A s)
        launch_dot(s.p, q.p, n0, dpart.p, 1, scal.p, dscal.p, stream);      // alpha, rho_old
        launch_xr_update(x.p, r, s.p, q.p, scal.p + 1, n0, stream);
    }
    hdd_check(hipStreamSynchronize(stream));
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return it;
}

}  // namespace hdd
