#include "hdd_runtime.h"
#include "hdd_sigma_device.h"
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <omp.h>

namespace hdd {

// set-up phase timings on stderr when HDD_FACTOR_VERBOSE is set
static void host_phase(const char* what, std::chrono::steady_clock::time_point& t) {
    static const bool verbose = std::getenv("HDD_FACTOR_VERBOSE") != nullptr;
    const auto now = std::chrono::steady_clock::now();
    if (verbose) std::fprintf(stderr, "[setup] %s: %.3fs\n", what, std::chrono::duration<double>(now - t).count());
    t = now;
}

DevLevel LevelBuffers::view() const {
    DevLevel v;
    v.np = host.npages();
    v.nbx = host.nbx; v.nby = host.nby; v.nbz = host.nbz;
    v.flags = flags.p; v.nbr = nbr.p; v.base = base.p; v.lut = lut.p;
    v.prob = prob.p; v.active = active;
    return v;
}

ProblemHost::ProblemHost(const SolverParams& p, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, const float* bin)
    : P(p) {
    auto t0 = std::chrono::steady_clock::now();
    if (P.vlevels < 1) throw std::runtime_error("vlevels must be >= 1");
    lv.resize(P.vlevels);
    PagedLevel& L0 = lv[0];
    L0.init(P.size);
    for (int i = 0; i < n_pages; ++i) {
        int bx, by, bz;
        page_coords(page_offsets[i] >> 12, bx, by, bz);
        const int q = L0.add_page(bx, by, bz);
        std::copy(flags + (size_t)i * kPageCells, flags + (size_t)(i + 1) * kPageCells, L0.flags.begin() + (size_t)q * kPageCells);
    }
    L0.sort_pages();
    // rhs b in page order of level 0 (input order may differ: remap through the LUT)
    input_page.resize(n_pages);
    b.assign((size_t)L0.npages() * kPageCells, 0.f);
    for (int i = 0; i < n_pages; ++i) {
        int bx, by, bz;
        page_coords(page_offsets[i] >> 12, bx, by, bz);
        const int q = L0.page_at_block(bx, by, bz);
        input_page[i] = q;
        std::copy(bin + (size_t)i * kPageCells, bin + (size_t)(i + 1) * kPageCells, b.begin() + (size_t)q * kPageCells);
    }
    // the interface operators are generated from the input flags, before division
    // (Initialize_Preconditioner: Split precedes Domain_Divider, DDS/...Solver.h:122-133)
    auto tp = std::chrono::steady_clock::now();
    host_phase("pages sorted", tp);
    // The interface operators are generated from the input flags, the divider and the
    // V-cycle levels rewrite them: the generator reads a copy on this thread while the
    // levels are built on another (neither reads what the other writes).
    {
        const PagedLevel input = L0;
        std::exception_ptr err;
        const int omp_threads = omp_get_max_threads();
        std::thread levels([&, omp_threads] {
            try {
                omp_set_num_threads(omp_threads);
                auto tl = std::chrono::steady_clock::now();
                divide_subdomains(L0, P.size, P.sub);
                set_face_flags(L0);
                for (int l = 0; l + 1 < P.vlevels; ++l) {
                    const Int3 s = lv[l].size;
                    lv[l + 1].init({s.x / 2 + 1, s.y / 2 + 1, s.z / 2 + 1});
                    downsample_level(lv[l], lv[l + 1]);
                }
                for (int l = 0; l < P.vlevels; ++l) {
                    set_face_flags(lv[l]);
                    set_boundary_flags(lv[l]);
                }
                host_phase("divider, flags and coarse levels", tl);
            } catch (...) {
                err = std::current_exception();
            }
        });
        try {
            generate_interface_operators(input, P.size, P.sub, P.sigma_levels, sigma);
        } catch (...) {
            levels.join();
            throw;
        }
        host_phase("interface operators generated", tp);
        levels.join();
        if (err) std::rethrow_exception(err);
    }
    host_phase("both done", tp);
    setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

ProblemHost::~ProblemHost() {}

void ProblemHost::factorize(bool wait) {
    if (factors) {
        if (wait) factors->wait_bottom();
        return;
    }
    auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<SigmaFactors> f(new SigmaFactors());
    factor_sigma(sigma, *f, wait);
    factors = std::move(f);
    factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

size_t ProblemHost::device_bytes() const {
    // every level: flags, two smoothing buffers, b, r (4 B each per slot), neighbour and origin tables
    double bytes = 0;
    for (const PagedLevel& L : lv) bytes += (L.npages() + 1.0) * kPageCells * 20.0 + L.npages() * 36.0;
    bytes += (lv[0].npages() + 1.0) * kPageCells * 16.0;  // x, b, q, s
    // interface solver: float plans of the factors (diagonal inverse twice, L21 once), operators, vectors
    for (int l = 0; l < sigma.levels; ++l) {
        for (const Csr& A : sigma.Air[l]) bytes += 16.0 * A.nnz();
        bytes += 8.0 * sigma.Arr[l].nnz() + 64.0 * sigma.n_if[l];
        if (l + 1 < sigma.levels) bytes += 16.0 * sigma.Prr[l].nnz();
        if (factors && l < (int)factors->lv.size()) {
            const SigmaFactors::Level& F = factors->lv[l];
            for (const SupernodalFactor& f : F.Aii) bytes += 8.0 * f.nnz();
            bytes += 8.0 * F.Arr.nnz();
            // dense tops (m x m floats each); a batch gives them to fewer rows, so this bounds them
            const int min_rows = dense_top_min_rows();
            int members = 0, m = 0;
            for (int s = 0; s < (int)F.rep.size(); ++s) members += min_rows > 0 && F.Aii[F.rep[s]].n >= min_rows;
            for (const SupernodalFactor& f : F.Aii)
                if (dense_top_cut(f, min_rows, members, &m) != INT_MAX) bytes += 4.0 * m * m;
            if (dense_top_cut(F.Arr, min_rows, 1, &m) != INT_MAX) bytes += 4.0 * m * m;
        } else {
            for (const Csr& A : sigma.Aii[l]) bytes += 64.0 * A.nnz();
            bytes += 64.0 * sigma.Arr[l].nnz();
        }
    }
    if (factors) {
        bytes += 8.0 * factors->bottom.nnz() + 8.0 * factors->spec.V.size();
        int m = 0;
        if (dense_top_cut(factors->bottom, dense_top_min_rows(), 1, &m) != INT_MAX) bytes += 4.0 * m * m;
    }
    return (size_t)(bytes * 1.1) + (size_t(4) << 20);
}

Solver::Solver(const SolverParams& p, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, const float* b, bool host)
    : P(p), host_only(host) {
    auto t0 = std::chrono::steady_clock::now();
    prob.push_back(std::make_shared<ProblemHost>(p, n_pages, page_offsets, flags, b));
    if (host_only) merge_levels();
    else build_device();
    setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Solver::Solver(const std::vector<std::shared_ptr<ProblemHost>>& problems) : prob(problems) {
    auto t0 = std::chrono::steady_clock::now();
    if (prob.empty()) throw std::runtime_error("a batch needs at least one problem");
    P = prob[0]->P;
    for (auto& q : prob) {
        const SolverParams& o = q->P;
        if (o.vlevels != P.vlevels || o.sigma_levels != P.sigma_levels || o.nV != P.nV || o.int_s != P.int_s || o.bnd_s != P.bnd_s ||
            o.sig_s != P.sig_s || o.sig_mg != P.sig_mg)
            throw std::runtime_error("every problem of a batch needs the same V-cycle and sigma schedule (levels and smoothing counts)");
    }
    build_device();
    setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Problems side by side along x: offsets are multiples of one coarsest-level
// page (4 * 2^(vlevels-1) cells) with at least that much empty space between
// consecutive boxes, so at every level no page holds cells of two problems.
void Solver::merge_levels() {
    const int NP = nprob(), NL = P.vlevels;
    const int align = kPX << (NL - 1);
    xoff.assign(NP, 0);
    for (int p = 1; p < NP; ++p) {
        const int end = xoff[p - 1] + prob[p - 1]->lv[0].size.x + align;
        xoff[p] = (end + align - 1) / align * align;
    }
    lv.clear();
    lv.resize(NL);
    pg0.assign(NL, std::vector<int>(NP + 1, 0));
    for (int l = 0; l < NL; ++l) {
        Int3 ext{0, 0, 0};
        for (int p = 0; p < NP; ++p) {
            const PagedLevel& S = prob[p]->lv[l];
            ext.x = std::max(ext.x, (xoff[p] >> l) + S.size.x);
            ext.y = std::max(ext.y, S.size.y);
            ext.z = std::max(ext.z, S.size.z);
        }
        PagedLevel& M = lv[l].host;
        M.init(ext);
        for (int p = 0; p < NP; ++p) {
            pg0[l][p] = M.npages();
            const PagedLevel& S = prob[p]->lv[l];
            const int dx = xoff[p] >> l;
            for (int q = 0; q < S.npages(); ++q) {
                const Int3 o = S.page_base[q];
                const int m = M.add_page((o.x + dx) >> 2, o.y >> 3, o.z >> 3);
                if (m != M.npages() - 1) throw std::runtime_error("batch layout: pages of two problems overlap");
                std::copy(S.flags.begin() + (size_t)q * kPageCells, S.flags.begin() + (size_t)(q + 1) * kPageCells,
                          M.flags.begin() + (size_t)m * kPageCells);
            }
        }
        pg0[l][NP] = M.npages();
    }
}

// Page-by-page sweeps of a level (the V-cycle bottom): pages are linked when
// an Active cell has a set face bit to an Active cell of the other page;
// linked pages form clusters. When no cluster has more than four pages, whole
// clusters are packed into workgroups of four page slots (wg_page, -1 for an
// empty slot) and nb_slot[p*6+f] gives the slot of page p's face-f neighbour
// page when it is in p's cluster (else -1). Every other face neighbour then
// holds no cell that a coupled Active cell reads and a sweep writes. Returns
// false (sweeps stay global) when a cluster is larger.
static bool plan_local_sweeps(const PagedLevel& H, std::vector<int>& wg_page, std::vector<int>& nb_slot) {
    static const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    const int np = H.npages();
    std::vector<int> parent(np);
    for (int p = 0; p < np; ++p) parent[p] = p;
    auto find = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    for (int p = 0; p < np; ++p) {
        const Int3 o = H.page_base[p];
        for (int c = 0; c < kPageCells; ++c) {
            const uint32_t f = H.flags[(size_t)p * kPageCells + c];
            if (!(f & kActive)) continue;
            const int x = o.x + (c >> 6), y = o.y + ((c >> 3) & 7), z = o.z + (c & 7);
            for (int q = 0; q < 6; ++q) {
                if (!(f & (kFaceMX << q))) continue;
                const int nx = x + d[q][0], ny = y + d[q][1], nz = z + d[q][2];
                const int pn = H.page_at(nx, ny, nz);
                if (pn == p || pn < 0 || !(H.flag(nx, ny, nz) & kActive)) continue;
                parent[find(pn)] = find(p);
            }
        }
    }
    std::vector<std::vector<int>> members(np);
    for (int p = 0; p < np; ++p) members[find(p)].push_back(p);  // ascending pages per cluster
    for (int r = 0; r < np; ++r)
        if (members[r].size() > 4) return false;
    wg_page.clear();
    std::vector<int> slot(np, -1), wg_of(np, -1);
    int used = 4;
    for (int p = 0; p < np; ++p) {
        const std::vector<int>& m = members[find(p)];
        if (m.front() != p) continue;  // clusters in the order of their first page (Morton order keeps them local)
        if (used + (int)m.size() > 4) { wg_page.insert(wg_page.end(), 4, -1); used = 0; }
        for (int q : m) {
            slot[q] = used;
            wg_of[q] = (int)wg_page.size() / 4 - 1;
            wg_page[wg_page.size() - 4 + used] = q;
            ++used;
        }
    }
    nb_slot.assign((size_t)np * 6, -1);
    for (int p = 0; p < np; ++p) {
        const Int3 o = H.page_base[p];
        for (int q = 0; q < 6; ++q) {
            const int pn = H.page_at(o.x + d[q][0] * kPX, o.y + d[q][1] * kPY, o.z + d[q][2] * kPZ);
            if (pn >= 0 && find(pn) == find(p) && wg_of[pn] == wg_of[p]) nb_slot[(size_t)p * 6 + q] = slot[pn];
        }
    }
    return true;
}

void Solver::build_device() {
    const int NP = nprob();
    auto tp = std::chrono::steady_clock::now();
    // the bottom levels' factors and spectra finish while the levels and plans above them are built
    for (auto& q : prob) q->factorize(false);
    host_phase("factorisations (bottom levels still running)", tp);
    // non-blocking: contexts of a server run side by side without the null stream serialising them
    hdd_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    upload_constants();
    merge_levels();
    upload_levels();
    {
        const char* e = std::getenv("HDD_BOTTOM_LOCAL");
        std::vector<int> wp, ns;
        bottom_local = !(e && std::string(e) == "0") && plan_local_sweeps(lv.back().host, wp, ns);
        if (bottom_local) { bottom_wg_page.upload(wp); bottom_nb_slot.upload(ns); }
    }
    for (int l = 0; l < (int)lv.size(); ++l) {
        std::vector<int> pp((size_t)lv[l].host.npages() + 1, -1);
        for (int p = 0; p < NP; ++p)
            for (int g = pg0[l][p]; g < pg0[l][p + 1]; ++g) pp[g] = p;
        lv[l].prob.upload(pp);
    }
    const PagedLevel& L0 = lv[0].host;
    const int np0 = L0.npages();
    {
        std::vector<float> hb((size_t)(np0 + 1) * kPageCells, 0.f);
        for (int p = 0; p < NP; ++p) std::copy(prob[p]->b.begin(), prob[p]->b.end(), hb.begin() + (size_t)pg0[0][p] * kPageCells);
        bvec.upload(hb);
    }
    const size_t n0 = lv[0].cells();
    x.alloc(n0); q.alloc(n0); s.alloc(n0);
    // per-problem reductions: each problem's partials as for a one-problem grid of its own size
    {
        std::vector<int> parts(NP), pp((size_t)np0 + 1, -1);
        max_parts = 1;
        for (int p = 0; p < NP; ++p) {
            parts[p] = reduce_parts((size_t)(pg0[0][p + 1] - pg0[0][p] + 1) * kPageCells);
            max_parts = std::max(max_parts, parts[p]);
            for (int g = pg0[0][p]; g < pg0[0][p + 1]; ++g) pp[g] = p;
        }
        seg_pg0.upload(pg0[0]); seg_parts.upload(parts); page_prob.upload(pp);
        active.upload(std::vector<int>(NP, 1));
        dpart.alloc((size_t)NP * max_parts); fpart.alloc((size_t)NP * max_parts);
        scal.alloc((size_t)NP * 8); dscal.alloc((size_t)NP * 8);
    }
    // interface dof -> level-0 cell index (problems' dofs concatenated in problem order)
    {
        std::vector<int> idx;
        for (int p = 0; p < NP; ++p) {
            for (const Int3& c : prob[p]->sigma.interface_map) {
                const int X = c.x + xoff[p];
                const int pg = L0.page_at(X, c.y, c.z);
                const bool ok = pg >= 0 && (L0.flags[(size_t)pg * kPageCells + page_cell(X, c.y, c.z)] & kInterface);
                idx.push_back(ok ? pg * kPageCells + page_cell(X, c.y, c.z) : -1);
            }
        }
        if_index.upload(idx);
    }
    std::vector<const SigmaHost*> hosts;
    std::vector<const SigmaFactors*> facs;
    for (auto& q : prob) { hosts.push_back(&q->sigma); facs.push_back(q->factors.get()); }
    host_phase("V-cycle levels uploaded", tp);
    sigma = make_sigma_device(hosts, facs, P.sig_s, P.sig_mg, stream);
    host_phase("interface solver on the device", tp);
    hdd_check(hipStreamSynchronize(stream), "setup sync");
}

Solver::~Solver() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    if (stream) (void)hipStreamSynchronize(stream);
    sigma.reset();
    // a caller's stream (set_stream) is the caller's to destroy
    hipStream_t mine = own_stream ? own_stream : stream;
    if (mine) (void)hipStreamDestroy(mine);
}

// one level's flags, face-neighbour and origin tables and block lookup table on the device,
// with its two smoothing buffers, rhs and residual
static void upload_level(LevelBuffers& L) {
    static const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    {
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

void Solver::upload_levels() {
    for (auto& L : lv) upload_level(L);
}

double profile_sweep_pages(Int3 size, int n_pages, const uint64_t* page_offsets, const uint32_t* flags, int reps, double* bytes) {
    LevelBuffers L;
    L.host.init(size);
    for (int i = 0; i < n_pages; ++i) {
        int bx, by, bz;
        page_coords(page_offsets[i] >> 12, bx, by, bz);
        const int q = L.host.add_page(bx, by, bz);
        std::copy(flags + (size_t)i * kPageCells, flags + (size_t)(i + 1) * kPageCells, L.host.flags.begin() + (size_t)q * kPageCells);
    }
    L.host.sort_pages();
    set_face_flags(L.host);
    upload_level(L);
    upload_constants();
    hipStream_t s = nullptr;
    hdd_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    hipEvent_t a = nullptr, b = nullptr;
    float ms = 0.f;
    try {
        hdd_check(hipEventCreate(&a), "hipEventCreate");
        hdd_check(hipEventCreate(&b), "hipEventCreate");
        const DevLevel v = L.view();
        launch_smooth(v, L.u[0].p, L.u[1].p, L.b.p, true, true, s);  // warm
        hdd_check(hipEventRecord(a, s), "hipEventRecord");
        for (int i = 0; i < reps; ++i) launch_smooth(v, L.u[i & 1].p, L.u[1 - (i & 1)].p, L.b.p, true, true, s);
        hdd_check(hipEventRecord(b, s), "hipEventRecord");
        hdd_check(hipEventSynchronize(b), "hipEventSynchronize");
        hdd_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    } catch (...) {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        (void)hipStreamDestroy(s);
        throw;
    }
    (void)hipEventDestroy(a); (void)hipEventDestroy(b);
    (void)hipStreamDestroy(s);
    if (bytes) *bytes = 16.0 * kPageCells * L.host.npages();
    return ms / reps;
}

float* Solver::channel(int c) {
    LevelBuffers& L0 = lv[0];
    switch (c) {
        case 0: return half_storage ? zf.p : L0.u[L0.cur].p;
        case 1: return L0.b.p;
        case 2: return L0.r.p;
        case 3: return x.p;
        case 4: return bvec.p;
        case 5: return q.p;
        case 6: return s.p;
    }
    throw std::runtime_error("bad channel id");
}

namespace {
// V-cycle buffers of one level by storage type
template <class S> struct VStore;
template <> struct VStore<float> {
    static float* u(LevelBuffers& L, int i) { return L.u[i].p; }
    static float* b(LevelBuffers& L) { return L.b.p; }
    static float* r(LevelBuffers& L) { return L.r.p; }
};
template <> struct VStore<__half> {
    static __half* u(LevelBuffers& L, int i) { return L.hu[i].p; }
    static __half* b(LevelBuffers& L) { return L.hb.p; }
    static __half* r(LevelBuffers& L) { return L.hr.p; }
};
}  // namespace

template <class S>
void Solver::sweeps_t(int l, int count, bool interior, bool boundary) {
    LevelBuffers& L = lv[l];
    const DevLevel v = L.view();
    for (int i = 0; i < count; ++i) {
        launch_smooth(v, VStore<S>::u(L, L.cur), VStore<S>::u(L, 1 - L.cur), VStore<S>::b(L), interior, boundary, stream);
        L.cur = 1 - L.cur;
    }
}

void Solver::sweeps(int l, int count, bool interior, bool boundary) { sweeps_t<float>(l, count, interior, boundary); }

void Solver::smooth0(int mode) {
    const bool in = mode == 0 || mode == 2, bd = mode == 1 || mode == 2;
    if (half_storage) sweeps_t<__half>(0, 1, in, bd);
    else sweeps(0, 1, in, bd);
}

// V_Cycle_Helper::v_cycle (LIB/SPGrid_V_Cycle_Helper.h:156-192). The two
// bottom-level residual passes only write the residual channel, which nothing
// reads before it is recomputed, so they are not launched.
template <class S>
void Solver::vcycle_t() {
    using V = VStore<S>;
    const int L = (int)lv.size();
    for (int l = 0; l < L - 1; ++l) {
        sweeps_t<S>(l, P.bnd_s, false, true);
        sweeps_t<S>(l, P.int_s, true, false);
        sweeps_t<S>(l, P.bnd_s, false, true);
        launch_residual(lv[l].view(), V::u(lv[l], lv[l].cur), V::b(lv[l]), V::r(lv[l]), stream);
        launch_fill(V::u(lv[l + 1], lv[l + 1].cur), 0.f, lv[l + 1].cells(), stream);
        launch_restrict(lv[l].view(), lv[l + 1].view(), V::r(lv[l]), V::b(lv[l + 1]), stream);
    }
    if (bottom_local) {
        // one launch; the result goes to the other buffer (other workgroups read this one's neighbour pages)
        LevelBuffers& B = lv[L - 1];
        launch_smooth_local(B.view(), bottom_wg_page.p, (int)(bottom_wg_page.n / 4), bottom_nb_slot.p, V::u(B, B.cur), V::u(B, 1 - B.cur),
                            V::b(B), 200, stream);
        B.cur = 1 - B.cur;
    } else {
        sweeps_t<S>(L - 1, 200, true, true);
    }
    for (int l = L - 2; l >= 0; --l) {
        launch_prolong(lv[l].view(), lv[l + 1].view(), V::u(lv[l], lv[l].cur), V::u(lv[l + 1], lv[l + 1].cur), stream);
        sweeps_t<S>(l, P.bnd_s, false, true);
        sweeps_t<S>(l, P.int_s, true, false);
        sweeps_t<S>(l, P.bnd_s, false, true);
    }
}

void Solver::vcycle() { vcycle_t<float>(); }

void Solver::set_half_storage(bool on) {
    if (on && !zf.p) {
        for (auto& L : lv) {
            L.hu[0].alloc(L.cells()); L.hu[1].alloc(L.cells());
            L.hb.alloc(L.cells()); L.hr.alloc(L.cells());
        }
        zf.alloc(lv[0].cells());
    }
    half_storage = on;
}

void Solver::set_sigma_refinement(int steps) {
    if (!sigma) throw std::runtime_error("no interface solver on this context");
    std::vector<const SigmaHost*> hosts;
    for (auto& q : prob) hosts.push_back(&q->sigma);
    sigma->set_refinement(hosts, steps);
    hdd_check(hipStreamSynchronize(stream), "sigma refinement setup");
    // captured applies hold the previous launch sequence
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

DBuf<float>& Solver::channel_buffer(int c) {
    LevelBuffers& L0 = lv[0];
    switch (c) {
        case 0: return half_storage ? zf : L0.u[L0.cur];
        case 1: return L0.b;
        case 2: return L0.r;
        case 3: return x;
        case 4: return bvec;
        case 5: return q;
        case 6: return s;
    }
    throw std::runtime_error("bad channel id");
}

void Solver::multiply(int c_in, int c_out) {
    launch_laplace(lv[0].view(), channel(c_in), channel(c_out), stream);
}

void Solver::inner_product(int a, int b, double* out) {
    const DevSegments sg = segments();
    launch_seg_dot(sg, channel(a), channel(b), 3, stream);
    std::vector<double> d((size_t)nprob() * 8);
    hdd_check(hipMemcpyAsync(d.data(), dscal.p, d.size() * sizeof(double), hipMemcpyDeviceToHost, stream), "inner product");
    hdd_check(hipStreamSynchronize(stream), "inner product");
    for (int p = 0; p < nprob(); ++p) out[p] = d[(size_t)p * 8 + 2];
}

void Solver::convergence_norm(int c, float* out) {
    launch_seg_maxabs(segments(), channel(c), stream);
    std::vector<float> f((size_t)nprob() * 8);
    hdd_check(hipMemcpyAsync(f.data(), scal.p, f.size() * sizeof(float), hipMemcpyDeviceToHost, stream), "convergence norm");
    hdd_check(hipStreamSynchronize(stream), "convergence norm");
    for (int p = 0; p < nprob(); ++p) out[p] = f[(size_t)p * 8 + 4];
}

void Solver::project(int c) { launch_project(lv[0].view(), channel(c), stream); }

void Solver::vector_op(int op, int c_out, int c_a, int c_b, float scalar) {
    const size_t n = (size_t)lv[0].host.npages() * kPageCells;
    launch_vector_op(op, channel(c_out), op == 4 ? nullptr : channel(c_a), op == 1 ? channel(c_b) : nullptr, scalar, n, stream);
}

// The apply reads its right-hand side from channel 1 and leaves its result in
// channel 0, using channel 2 and the other smoothing buffer as scratch. Other
// channels are lent to it by exchanging buffers (all channels have the same
// size and a zero last page), so no channel but result and tmp changes.
void Solver::apply_preconditioner_channels(int c_result, int c_rhs, int c_tmp) {
    for (int c : {c_result, c_rhs, c_tmp})
        if (c < 0 || c > 6) throw std::runtime_error("bad channel id");
    if (c_result == c_rhs || c_result == c_tmp || c_rhs == c_tmp) throw std::runtime_error("result, rhs and tmp must be three different channels");
    if (c_result == 0 && c_rhs == 1 && c_tmp == 2) { apply_preconditioner(); return; }
    // the exchanges below move buffers that captured applies have recorded
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
    LevelBuffers& L0 = lv[0];
    auto swap_buf = [](DBuf<float>& a, DBuf<float>& b) { std::swap(a.p, b.p); };
    // lend rhs as channel 1 and tmp as channel 2; a channel among {0, 1, 2} that is
    // not in its role is parked in the buffer of the channel that takes its role
    std::vector<std::pair<int, int>> done;
    auto lend = [&](int role, int c) {
        if (c == role) return;
        swap_buf(channel_buffer(role), channel_buffer(c));
        done.push_back({role, c});
    };
    lend(1, c_rhs);
    lend(2, c_tmp == 1 ? c_rhs : c_tmp);
    if (c_result != 0) swap_buf(channel_buffer(0), channel_buffer(c_result));
    try {
        record_apply();
    } catch (...) {
        if (c_result != 0) swap_buf(channel_buffer(0), channel_buffer(c_result));
        for (auto it = done.rbegin(); it != done.rend(); ++it) swap_buf(channel_buffer(it->first), channel_buffer(it->second));
        throw;
    }
    // the result is in channel 0's buffer now; give it to the result channel, channel 0 its own contents back
    if (c_result != 0) swap_buf(channel_buffer(0), channel_buffer(c_result));
    for (auto it = done.rbegin(); it != done.rend(); ++it) swap_buf(channel_buffer(it->first), channel_buffer(it->second));
}

void Solver::set_stream(hipStream_t s_new) {
    hdd_check(hipStreamSynchronize(stream), "set_stream");
    if (!own_stream) own_stream = stream;
    stream = s_new ? s_new : own_stream;
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

DevSegments Solver::segments() {
    DevSegments sg;
    sg.nprob = nprob(); sg.max_parts = max_parts; sg.ncells = (size_t)lv[0].host.npages() * kPageCells;
    sg.pg0 = seg_pg0.p; sg.parts = seg_parts.p; sg.page_prob = page_prob.p; sg.active = active.p;
    sg.dpart = dpart.p; sg.fpart = fpart.p; sg.scal = scal.p; sg.dscal = dscal.p;
    sg.serial_dots = serial_dots;
    return sg;
}

// Apply_Preconditioner (DDS/SPGrid_Domain_Decomposition_Preconditioner.h:88-128),
// z = channel 0, rhs = channel 1 (PCG residual r), tmp k = channel 2.
// The apply is a fixed sequence of launches for a given state of the
// smoothing ping-pong buffers and of the skip mode, so with use_graph it is
// captured once into a HIP graph per such state and replayed: one graph launch
// instead of ~2000 kernel launches per PCG iteration. Nothing in it allocates,
// copies or clears memory through the runtime (clears are fill kernels).
// Off by default: the replay is bit-identical but measured no faster (the
// iteration is bound by the GPU's dispatch of dependent kernels, not by host
// launches), and a capture fails if another thread of the process uses the
// null stream meanwhile (the server's setup workers do).
void Solver::apply_preconditioner() {
    if (!use_graph) { record_apply(); return; }
    std::vector<int> key;
    for (auto& L : lv) key.push_back(L.cur);
    key.push_back(lv[0].active != nullptr);
    key.push_back(half_storage);
    for (auto& g : graphs)
        if (g.key == key) {
            hdd_check(hipGraphLaunch(g.exec, stream), "hipGraphLaunch");
            for (size_t l = 0; l < lv.size(); ++l) lv[l].cur = g.cur_out[l];
            return;
        }
    Graph g;
    g.key = key;
    hipGraph_t graph = nullptr;
    hdd_check(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    try {
        record_apply();
    } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    hdd_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    hdd_check(e, "hipGraphInstantiate");
    for (auto& L : lv) g.cur_out.push_back(L.cur);
    for (size_t l = 0; l < lv.size(); ++l) lv[l].cur = key[l];  // capture recorded the launches without running them
    graphs.push_back(g);
    hdd_check(hipGraphLaunch(g.exec, stream), "hipGraphLaunch");
    for (size_t l = 0; l < lv.size(); ++l) lv[l].cur = g.cur_out[l];
}

void Solver::record_apply() {
    LevelBuffers& L0 = lv[0];
    const size_t n0 = L0.cells();
    if (half_storage) {
        // FP16 storage: the V-cycles and the interface exchange run on r / a per
        // problem (a = max|r|, recomputed here), the interface solver in FP32
        launch_seg_maxabs(segments(), L0.b.p, stream);
        launch_scale_to_half(L0.b.p, L0.hb.p, page_prob.p, scal.p, (size_t)L0.host.npages() * kPageCells, stream);
        launch_fill(L0.hu[L0.cur].p, 0.f, n0, stream);
        for (int i = 0; i < P.nV; ++i) vcycle_t<__half>();
        launch_interface_gather(*sigma, L0.view(), if_index.p, L0.hu[L0.cur].p, L0.hb.p, stream);
        sigma_vcycles(*sigma, stream);
        launch_fill(L0.hu[L0.cur].p, 0.f, n0, stream);
        launch_interface_scatter(*sigma, if_index.p, L0.hu[L0.cur].p, stream);
        for (int i = 0; i < P.nV; ++i) vcycle_t<__half>();
        launch_unscale_project(L0.view(), L0.hu[L0.cur].p, zf.p, page_prob.p, scal.p, stream);
        return;
    }
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

// While a batch PCG runs, the stencil kernels and exact solves skip the
// problems that have stopped (their pages and supernodes are left untouched).
namespace {
struct SkipStopped {
    Solver& S;
    explicit SkipStopped(Solver& s) : S(s) { set(S.active.p); }
    ~SkipStopped() {
        set(nullptr);
        std::vector<int> ones(S.nprob(), 1);
        (void)hipMemcpyAsync(S.active.p, ones.data(), ones.size() * sizeof(int), hipMemcpyHostToDevice, S.stream);
        (void)hipStreamSynchronize(S.stream);
    }
    void set(const int* a) {
        for (auto& L : S.lv) L.active = a;
        if (S.sigma) S.sigma->set_active(a);
    }
};
}  // namespace

// CONJUGATE_GRADIENT_DUMPING::Solve (DDS/CONJUGATE_GRADIENT_DUMPING.cpp:146-212)
// for every problem of the batch, x starting at zero, with the stopping rule
// max|r| <= tol[p]. Each problem has its own scalars; once it stops, its x
// and r are frozen (the shared launches still visit its pages). The loop is
// split at its one host synchronisation per iteration (the residual test):
// pcg_start queues the set-up and the first test, pcg_advance waits for a
// test, records it and queues the next iteration.
void Solver::pcg_start(PcgRun& R, const std::vector<double>& tol, const std::vector<int>& max_it) {
    const int NP = nprob();
    if ((int)tol.size() != NP || (int)max_it.size() != NP) throw std::runtime_error("pcg: one tolerance and iteration limit per problem");
    LevelBuffers& L0 = lv[0];
    const DevLevel v0 = L0.view();
    const size_t n0 = L0.cells();
    R.tol = tol; R.max_it = max_it;
    R.out.assign(NP, PcgResult());
    R.act.assign(NP, 1);
    R.cancel.assign(NP, 0);
    R.it = 0; R.n_active = NP; R.started = true;
    const float flt_max = 3.40282347e+38f;
    hs_pin.alloc((size_t)NP * 8);
    act_pin.alloc((size_t)NP);
    for (size_t i = 0; i < (size_t)NP * 8; ++i) hs_pin.p[i] = 0.f;
    for (int p = 0; p < NP; ++p) { hs_pin.p[(size_t)p * 8 + 3] = flt_max; act_pin.p[p] = 1; }  // rho_old
    R.fp64 = outer_fp64;
    if (R.fp64) {
        if (xd.n != n0) xd.alloc(n0);
        std::vector<double> hd((size_t)NP * 8, 0.0);
        for (int p = 0; p < NP; ++p) hd[(size_t)p * 8 + 5] = flt_max;  // rho_old
        hdd_check(hipMemcpyAsync(dscal.p, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        hdd_check(hipMemsetAsync(xd.p, 0, n0 * sizeof(double), stream));
        hdd_check(hipStreamSynchronize(stream));  // hd goes out of scope
    }
    hdd_check(hipMemcpyAsync(scal.p, hs_pin.p, (size_t)NP * 8 * sizeof(float), hipMemcpyHostToDevice, stream));
    hdd_check(hipMemcpyAsync(active.p, act_pin.p, NP * sizeof(int), hipMemcpyHostToDevice, stream));
    R.skip.reset();
    if (NP > 1) R.skip = std::shared_ptr<void>(new SkipStopped(*this), [](void* q) { delete static_cast<SkipStopped*>(q); });
    launch_fill(x.p, 0.f, n0, stream);
    hdd_check(hipStreamSynchronize(stream));
    R.t0 = std::chrono::steady_clock::now();
    // iteration 0 (a restart): r = Project(b - L x), then its test
    launch_laplace_pcg(v0, x.p, q.p, stream);
    launch_sub_project(v0, bvec.p, q.p, L0.b.p, stream);
    launch_seg_maxabs(segments(), L0.b.p, stream);
    hdd_check(hipMemcpyAsync(hs_pin.p, scal.p, (size_t)NP * 8 * sizeof(float), hipMemcpyDeviceToHost, stream));
}

bool Solver::pcg_advance(PcgRun& R, const std::function<void(int)>& on_stop) {
    if (!R.started) throw std::runtime_error("pcg_advance before pcg_start");
    if (R.n_active == 0) return true;
    const int NP = nprob();
    LevelBuffers& L0 = lv[0];
    DevLevel v0 = L0.view();
    v0.active = nullptr;  // the PCG operator visits every page; only the PCG vector updates skip stopped problems
    const size_t n0 = L0.cells();
    float* r = L0.b.p;
    const DevSegments sg = segments();
    float* hs = hs_pin.p;
    const int it = R.it;
    const bool restart = (it == 0);
    hdd_check(hipStreamSynchronize(stream));
    if (R.fp64 && it > 0) {
        // residual replacement: every replace_every iterations, and whenever a problem's
        // recursive residual passes its test, r becomes the true residual b - L x of the
        // double iterate, and the test is taken on that
        bool replace = replace_every > 0 && it % replace_every == 0;
        for (int p = 0; p < NP && !replace; ++p) replace = R.act[p] && hs[(size_t)p * 8 + 4] <= (float)R.tol[p];
        if (replace) {
            launch_true_residual(v0, xd.p, bvec.p, r, stream);
            launch_seg_maxabs(sg, r, stream);
            hdd_check(hipMemcpyAsync(hs, scal.p, (size_t)NP * 8 * sizeof(float), hipMemcpyDeviceToHost, stream));
            hdd_check(hipStreamSynchronize(stream));
        }
    }
    const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - R.t0).count();
    bool changed = false;
    for (int p = 0; p < NP; ++p) {
        if (!R.act[p]) continue;
        const float conv = hs[(size_t)p * 8 + 4];
        PcgResult& o = R.out[p];
        o.hist.push_back((double)conv);
        if (conv <= (float)R.tol[p]) { o.iterations = it; o.converged = true; }
        else if (R.cancel[p]) { o.iterations = -it; o.cancelled = true; }
        else if (it >= R.max_it[p]) o.iterations = -it;
        else continue;
        o.seconds = now;
        R.act[p] = 0;
        --R.n_active;
        changed = true;
        if (R.fp64) {
            // the float x channel of the stopped problem from its double iterate
            const size_t first = (size_t)pg0[0][p] * kPageCells, cnt = (size_t)(pg0[0][p + 1] - pg0[0][p]) * kPageCells;
            launch_cast_to_float(xd.p + first, x.p + first, cnt, stream);
            hdd_check(hipStreamSynchronize(stream));
        }
        if (on_stop) on_stop(p);
    }
    if (R.n_active == 0) {
        hdd_check(hipStreamSynchronize(stream));
        R.skip.reset();
        return true;
    }
    // act_pin is not modified again before the next synchronisation
    if (changed) {
        for (int p = 0; p < NP; ++p) act_pin.p[p] = R.act[p];
        hdd_check(hipMemcpyAsync(active.p, act_pin.p, NP * sizeof(int), hipMemcpyHostToDevice, stream));
    }
    apply_preconditioner();
    float* z = channel(0);
    if (R.fp64) {
        launch_seg_dot(sg, z, r, 5, stream);                   // rho, beta (double)
        if (restart) launch_copy(s.p, z, n0, stream);
        else launch_seg_copy2_d(sg, s.p, s.p, z, 4, stream);   // s = s*beta + z
        launch_laplace_pcg(v0, s.p, q.p, stream);              // q = Project(A s)
        launch_seg_dot(sg, s.p, q.p, 6, stream);               // alpha, rho_old (double)
        launch_seg_xr_update_d(sg, xd.p, r, s.p, q.p, stream); // x (double) += alpha s; r -= alpha q
    } else {
        launch_seg_dot(sg, z, r, 0, stream);                   // rho, beta
        if (restart) launch_copy(s.p, z, n0, stream);
        else launch_seg_copy2(sg, s.p, s.p, z, 2, stream);     // s = s*beta + z
        launch_laplace_pcg(v0, s.p, q.p, stream);              // q = Project(A s)
        launch_seg_dot(sg, s.p, q.p, 1, stream);               // alpha, rho_old
        launch_seg_xr_update(sg, x.p, r, s.p, q.p, stream);
    }
    R.it = it + 1;
    // the next iteration's test
    launch_seg_maxabs(sg, r, stream);
    hdd_check(hipMemcpyAsync(hs, scal.p, (size_t)NP * 8 * sizeof(float), hipMemcpyDeviceToHost, stream));
    return false;
}

void Solver::pcg_batch(const std::vector<double>& tol, const std::vector<int>& max_it, std::vector<PcgResult>& out,
                       const std::function<void(int)>& on_stop) {
    PcgRun R;
    pcg_start(R, tol, max_it);
    while (!pcg_advance(R, on_stop)) {
    }
    out = std::move(R.out);
}

int Solver::pcg(double tol_abs, int max_it, std::vector<double>& hist, double* seconds, bool* converged) {
    std::vector<PcgResult> res;
    pcg_batch(std::vector<double>(nprob(), tol_abs), std::vector<int>(nprob(), max_it), res);
    hist = res[0].hist;
    if (converged) *converged = res[0].converged;
    double worst = 0;
    for (auto& r : res) worst = std::max(worst, r.seconds);
    if (seconds) *seconds = worst;
    return res[0].iterations;
}

void Solver::download_x_f64(int p, double* host) {
    if (p < 0 || p >= nprob()) throw std::runtime_error("no such problem in the batch");
    if (!xd.p) throw std::runtime_error("no FP64 iterate: the last solve ran with float outer scalars");
    hdd_check(hipStreamSynchronize(stream), "sync");
    const size_t first = (size_t)pg0[0][p] * kPageCells, n = (size_t)(pg0[0][p + 1] - pg0[0][p]) * kPageCells;
    hdd_check(hipMemcpyAsync(host, xd.p + first, n * sizeof(double), hipMemcpyDeviceToHost, stream), "download x");
    hdd_check(hipStreamSynchronize(stream), "download x");
}

void Solver::download_x(int p, float* host) {
    if (p < 0 || p >= nprob()) throw std::runtime_error("no such problem in the batch");
    hdd_check(hipStreamSynchronize(stream), "sync");
    const size_t first = (size_t)pg0[0][p] * kPageCells, n = (size_t)(pg0[0][p + 1] - pg0[0][p]) * kPageCells;
    hdd_check(hipMemcpyAsync(host, x.p + first, n * sizeof(float), hipMemcpyDeviceToHost, stream), "download x");
    hdd_check(hipStreamSynchronize(stream), "download x");
}

}  // namespace hdd
