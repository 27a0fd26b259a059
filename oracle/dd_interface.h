// ORACLE — test infrastructure only (see spgrid.h header).
// Interface (Schur) operator generator of the reference, restated:
//   SPGrid_Subdomain_Splitter::Split      DDS/SPGrid_Subdomain_Splitter.h:43-158 (#else branch)
//   Subdomain_Rasterizer / _Masked_...    DDS/SPGrid_Subdomain_Rasterizer.h:26-140
//   STRIDE_GENERATOR                      IFS/STRIDE_GENERATOR.h:23-54
//   CELL_PROLONGATION_BAKER::Bake         IFS/CELL_PROLONGATION_BAKER.h:22-51
//   GALERKIN_PROCESS::Galerkin_Coarsen    IFS/GALERKIN_PROCESS.h:27-68
//   INTERFACE_MATRIX_GENERATOR            IFS/INTERFACE_MATRIX_GENERATOR.h:81-734
//   PROLONGATION_MATRIX_HELPER            IFS/PROLONGATION_MATRIX_HELPER.cpp:13-40
// (IFS = demos/Domain_Decomposition_CPU_Only/Interface_Solver)
// STENCIL (Common_Tools/Grids_Uniform_PDE_Linear/STENCIL.h) is an insertion-
// ordered array, so every accumulation order below is the reference's.
// The 8x8 Eigen products of Galerkin_Coarsen are evaluated here as plain
// ascending-k loops; Eigen's internal order is not pinned (all operands are
// dyadic rationals, so the sums are exact unless they exceed 24 bits).
#pragma once
#include "dd_setup.h"
#include <unordered_map>
#include <map>

namespace ora {

// adaptive hierarchy flags (IFS/ADAPTIVE_SUBDOMAIN_POISSON_FLAGS.h)
enum : uint32_t {
    UNodeActive = 0x10u, UNodeDirichlet = 0x20u, UNodeInterface = 0x40u,
    CellTraversed = 0x01000000u, CellInterior = 0x1u, NodeDegree = 0x10000000u,
    NodeTJunction = 0x100u, NodeCoarseShared = 0x200u, NodeActive = 0x400u,
};

template <class T>
struct Stencil1 {  // STENCIL<T,1>
    std::vector<std::pair<int, T>> e;
    T& get_or_insert(int k) {
        for (auto& p : e) if (p.first == k) return p.second;
        e.push_back({k, T(0)});
        return e.back().second;
    }
};

template <class T>
struct CSR {
    int rows = 0, cols = 0;
    std::vector<int> ptr{0}, idx;
    std::vector<T> val;
};

template <class T>
CSR<T> stencils_to_csr(const std::vector<Stencil1<T>>& rows, int ncols) {
    CSR<T> m; m.rows = (int)rows.size(); m.cols = ncols; m.ptr.assign(1, 0);
    for (auto& r : rows) {
        std::vector<std::pair<int, T>> v;
        for (auto& p : r.e) if (p.second != T(0)) v.push_back(p);
        std::sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (auto& p : v) { m.idx.push_back(p.first); m.val.push_back(p.second); }
        m.ptr.push_back((int)m.idx.size());
    }
    return m;
}

// STRIDE_GENERATOR::Generate_Adaptive_Stride for one axis over [lo, hi]
inline int adaptive_stride_axis(int lo, int hi, std::vector<int>& q) {
    int counts = hi - lo, level = 0;
    q.assign(hi - lo + 1, 0);
    std::vector<std::pair<int, int>> ad;
    ad.push_back({2, 1});
    int cur = 1, dom = 2;
    while (dom < counts) {
        int quo = std::min(2, (counts - dom) / cur);
        ad.push_back({quo, cur}); dom += quo * cur;
        if ((counts - dom) % (2 * cur) == 0) { cur *= 2; ++level; }
    }
    int cnt = 0;
    std::vector<int> st;
    for (size_t i = 0; i < ad.size(); ++i) {
        for (int k = 0; k < ad[i].second; ++k) q[cnt++] = ad[i].second;
        if (ad[i].first == 1) q[cnt++] = ad[i].second;
        else { if (i + 1 == ad.size()) q[cnt++] = ad[i].second; st.push_back(ad[i].second); }
    }
    while (!st.empty()) { int c = st.back(); st.pop_back(); for (int i = 0; i < c; ++i) q[cnt++] = c; }
    return level;
}

struct Stride3 { int lo[3]; std::vector<int> q[3]; int at(int a, int i) const { return q[a][i - lo[a]]; } };
inline int adaptive_stride(const int lo[3], const int hi[3], Stride3& s) {
    int lev = 1 << 30;
    for (int a = 0; a < 3; ++a) { s.lo[a] = lo[a]; lev = std::min(lev, adaptive_stride_axis(lo[a], hi[a], s.q[a])); }
    return lev;
}

typedef float M8[8][8];
// CELL_PROLONGATION_BAKER::Bake; index_map(i,j,k) = 4i+2j+k (RANGE order)
struct Baker {
    float P[8][8][8];  // [child][fine local][coarse local]
    Baker() {
        for (int mc = 0; mc < 8; ++mc) {
            int ox = mc >> 2, oy = (mc >> 1) & 1, oz = mc & 1;
            for (int a = 0; a < 8; ++a) for (int b = 0; b < 8; ++b) P[mc][a][b] = 0;
            for (int fi = 0; fi < 8; ++fi) {
                int f[3] = {(fi >> 2) + ox, ((fi >> 1) & 1) + oy, (fi & 1) + oz};
                int cn[3] = {f[0] / 2, f[1] / 2, f[2] / 2};
                for (int ni = 0; ni < 8; ++ni) {
                    int c[3] = {cn[0] + (ni >> 2), cn[1] + ((ni >> 1) & 1), cn[2] + (ni & 1)};
                    float w = 1.f;
                    for (int v = 0; v < 3; ++v) w *= (float)((double)(2.f - (float)std::abs(c[v] * 2 - f[v])) / 2.0);
                    if (w == 0) continue;
                    P[mc][fi][(c[0] << 2) | (c[1] << 1) | c[2]] = w;
                }
            }
        }
    }
};

template <class T>
struct SigmaSetup {
    int mg_levels = 0, nsub = 0;
    std::vector<int> n_if;
    std::vector<I3> interface_map;                       // level-0 interface cells in id order
    std::vector<std::unordered_map<uint64_t, int>> ihash;  // per level
    std::vector<std::vector<I3>> icoord;                 // per level: id -> level coordinates
    std::vector<std::vector<Stencil1<T>*>> Arr_s;
    std::vector<std::vector<std::vector<Stencil1<T>>>> Aii_s, Air_s;  // [sub][level][row]
    std::vector<CSR<T>> Arr, Prr;
    std::vector<std::vector<CSR<T>>> Aii, Air;  // [level][sub]
    ~SigmaSetup() { for (auto& v : Arr_s) for (auto* p : v) delete p; }
};

inline uint64_t key3(int x, int y, int z) { return ((uint64_t)(uint32_t)x << 42) | ((uint64_t)(uint32_t)y << 21) | (uint64_t)(uint32_t)z; }

template <class T>
struct SubdomainHierarchy {
    // hierarchy[m-1] -> levels grids (index L-1), flags + node_index ints
    std::vector<std::vector<Grid<T>>> h;
};

template <class T>
class InterfaceGenerator {
public:
    SigmaSetup<T>& S;
    Baker baker;
    I3 size;
    explicit InterfaceGenerator(SigmaSetup<T>& s) : S(s) {}

    void initialize(int mg_levels) {
        S.mg_levels = mg_levels;
        S.Arr_s.assign(mg_levels, {});
        S.ihash.assign(mg_levels, {});
        S.icoord.assign(mg_levels, {});
        S.n_if.assign(mg_levels, 0);
        S.interface_map.clear();
        S.Aii_s.clear(); S.Air_s.clear();
    }

    int get_if_id(int level, int x, int y, int z, bool push_map) {
        auto& H = S.ihash[level];
        uint64_t k = key3(x, y, z);
        auto it = H.find(k);
        if (it != H.end()) return it->second;
        int id = S.n_if[level]++;
        H[k] = id;
        S.icoord[level].push_back({x, y, z});
        if (push_map) S.interface_map.push_back({x, y, z});
        S.Arr_s[level].push_back(new Stencil1<T>());
        return id;
    }

    // Embedding_Helper (IFS/INTERFACE_MATRIX_GENERATOR.h:81-100)
    void embed(int ix, int iy, int iz, std::vector<std::pair<I3, T>>& parents, T weight, const Stride3& q) {
        int idx[3] = {ix, iy, iz};
        int qs = std::min(q.at(0, ix), std::min(q.at(1, iy), q.at(2, iz)));
        for (int a = 0; a < 3; ++a) {
            int cur = idx[a] - q.lo[a];
            if (cur % qs != 0) {
                int i1[3] = {ix, iy, iz}, i2[3] = {ix, iy, iz};
                i1[a] = (cur / qs) * qs + q.lo[a];
                i2[a] = (cur / qs + 1) * qs + q.lo[a];
                T w1 = weight * (T)(i2[a] - idx[a]) / (T)(i2[a] - i1[a]);
                T w2 = weight * (T)(idx[a] - i1[a]) / (T)(i2[a] - i1[a]);
                embed(i1[0], i1[1], i1[2], parents, w1, q);
                embed(i2[0], i2[1], i2[2], parents, w2, q);
                return;
            }
        }
        for (auto& p : parents) if (p.first.x == ix && p.first.y == iy && p.first.z == iz) { p.second += weight; return; }
        parents.push_back({{ix, iy, iz}, weight});
    }

    // GALERKIN_PROCESS::Galerkin_Coarsen, memoised per (level, cell)
    std::vector<std::unordered_map<uint64_t, std::array<T, 64>>> memo;
    const Grid<T>* subset = nullptr;

    std::array<T, 64> galerkin(int cx, int cy, int cz, int level) {
        auto& mm = memo[level];
        uint64_t k = key3(cx, cy, cz);
        auto it = mm.find(k);
        if (it != mm.end()) return it->second;
        std::array<T, 64> M; M.fill(T(0));
        if (level == 1) {
            const T ew = (T)(1.0 / (1 << 2));
            uint32_t nf[8];
            for (int n = 0; n < 8; ++n) {
                int x = cx - 1 + (n >> 2), y = cy - 1 + ((n >> 1) & 1), z = cz - 1 + (n & 1);
                nf[n] = 0;
                if (subset->is_set(x, y, z, Active | Interface)) nf[n] = Active;
                else if (subset->is_set(x, y, z, Dirichlet)) nf[n] = Dirichlet;
            }
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        // face_iterator over the two remaining axes in order, lower/upper along v
                        int lo[3], up[3];
                        int rem[2], r = 0;
                        for (int t = 0; t < 3; ++t) if (t != v) rem[r++] = t;
                        lo[rem[0]] = a; lo[rem[1]] = b; up[rem[0]] = a; up[rem[1]] = b;
                        lo[v] = 0; up[v] = 1;
                        int li = (lo[0] << 2) | (lo[1] << 1) | lo[2], ui = (up[0] << 2) | (up[1] << 1) | up[2];
                        if ((nf[li] & Active) && (nf[ui] & Active)) {
                            M[li * 8 + ui] -= ew; M[li * 8 + li] += ew; M[ui * 8 + ui] += ew; M[ui * 8 + li] -= ew;
                        }
                        if ((nf[li] & Active) && (nf[ui] & Dirichlet)) M[li * 8 + li] += ew;
                        if ((nf[li] & Dirichlet) && (nf[ui] & Active)) M[ui * 8 + ui] += ew;
                    }
        } else {
            for (int ch = 0; ch < 8; ++ch) {
                auto C = galerkin(2 * cx - 1 + (ch >> 2), 2 * cy - 1 + ((ch >> 1) & 1), 2 * cz - 1 + (ch & 1), level - 1);
                const auto& P = baker.P[ch];
                T tmp[64], prod[64];
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j) {
                        T s = 0;
                        for (int kk = 0; kk < 8; ++kk) s += (T)P[kk][i] * C[kk * 8 + j];
                        tmp[i * 8 + j] = s;
                    }
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j) {
                        T s = 0;
                        for (int kk = 0; kk < 8; ++kk) s += tmp[i * 8 + kk] * (T)P[kk][j];
                        prod[i * 8 + j] = s;
                    }
                for (int e = 0; e < 64; ++e) M[e] += prod[e];
            }
        }
        mm[k] = M;
        return M;
    }

    // cells of a grid with a flag, in SPGrid offset order (Morton pages, in-page x,y,z)
    static void cells_with(const Grid<T>& g, uint32_t flag, std::vector<I3>& out) {
        out.clear();
        for (int p : g.blocks) {
            int X, Y, Z; g.base(p, X, Y, Z);
            for (int c = 0; c < BCELLS; ++c)
                if (g.flags[(size_t)p * BCELLS + c] & flag) out.push_back({X + (c >> 6), Y + ((c >> 3) & 7), Z + (c & 7)});
        }
    }

    // INTERFACE_MATRIX_GENERATOR::Create_Subdomain_Hierarchy (IFS/INTERFACE_MATRIX_GENERATOR.h:129-659)
    void create_subdomain_hierarchy(const Grid<T>& set, I3 lo, I3 hi) {
        const int mgl = S.mg_levels;
        S.Aii_s.emplace_back(mgl);
        S.Air_s.emplace_back(mgl);
        auto& Aii_sub = S.Aii_s.back();
        auto& Air_sub = S.Air_s.back();
        int lo3[3] = {lo.x, lo.y, lo.z}, hi3[3] = {hi.x, hi.y, hi.z};
        Stride3 qtop;
        int levels = adaptive_stride(lo3, hi3, qtop);
        if (mgl > levels) throw std::runtime_error("sigma_levels exceeds adaptive levels");
        std::vector<Stride3> q(levels + 1);
        for (int L = 1; L <= levels; ++L) {
            int l3[3], h3[3];
            for (int a = 0; a < 3; ++a) { l3[a] = lo3[a] >> (L - 1); h3[a] = hi3[a] >> (L - 1); }
            adaptive_stride(l3, h3, q[L]);
        }
        SubdomainHierarchy<T> H;
        H.h.assign(mgl, std::vector<Grid<T>>(levels));
        for (int m = 0; m < mgl; ++m)
            for (int L = 1; L <= levels; ++L) {
                int s = L - 1;
                H.h[m][L - 1].init((lo.x >> s) - 8, (lo.y >> s) - 8, (lo.z >> s) - 8, (hi.x >> s) + 9, (hi.y >> s) + 9, (hi.z >> s) + 9, 0, true);
            }
        auto inside = [&](int x, int y, int z) { return x >= lo.x && y >= lo.y && z >= lo.z && x <= hi.x && y <= hi.y && z <= hi.z; };
        // contains(range thickened by -t, [mn, mx])
        auto contains = [&](int t, const int mn[3], const int mx[3]) {
            for (int a = 0; a < 3; ++a) if (!(lo3[a] + t <= mn[a] && hi3[a] - t >= mx[a])) return false;
            return true;
        };
        // --- activation of adaptive cells (:184-239)
        for (int p : set.blocks) {
            int X, Y, Z; set.base(p, X, Y, Z);
            for (int c = 0; c < BCELLS; ++c) {
                if (set.flags[(size_t)p * BCELLS + c] == 0) continue;
                int nx = X + (c >> 6), ny = Y + ((c >> 3) & 7), nz = Z + (c & 7);
                for (int cc = 0; cc < 8; ++cc) {
                    int cx = nx + (cc >> 2), cy = ny + ((cc >> 1) & 1), cz = nz + (cc & 1);
                    if (!inside(cx - 1, cy - 1, cz - 1) || !inside(cx, cy, cz)) continue;
                    Grid<T>& h1 = H.h[0][0];
                    if (h1.is_set(cx, cy, cz, CellTraversed)) continue;
                    h1.mask(cx, cy, cz, CellTraversed);
                    bool go_next = false;
                    int cell[3] = {cx, cy, cz};
                    for (int v = 0; v < 3 && !go_next; ++v) {
                        int rem[2], r = 0;
                        for (int t = 0; t < 3; ++t) if (t != v) rem[r++] = t;
                        for (int a = 0; a < 2 && !go_next; ++a)
                            for (int b = 0; b < 2 && !go_next; ++b) {
                                int l[3], u[3];
                                l[rem[0]] = cell[rem[0]] - 1 + a; l[rem[1]] = cell[rem[1]] - 1 + b; l[v] = cell[v] - 1;
                                u[rem[0]] = l[rem[0]]; u[rem[1]] = l[rem[1]]; u[v] = cell[v];
                                unsigned lt = 0, ut = 0;
                                if (set.is_set(l[0], l[1], l[2], Active | Interface)) lt |= UNodeActive;
                                if (set.is_set(l[0], l[1], l[2], Dirichlet)) lt |= UNodeDirichlet;
                                if (set.is_set(u[0], u[1], u[2], Active | Interface)) ut |= UNodeActive;
                                if (set.is_set(u[0], u[1], u[2], Dirichlet)) ut |= UNodeDirichlet;
                                if ((lt & ut & UNodeActive) || (((lt | ut) & UNodeActive) && ((lt | ut) & UNodeDirichlet))) {
                                    int ci[3] = {cx, cy, cz};
                                    int mn[3] = {cx - 1, cy - 1, cz - 1}, mx[3] = {cx, cy, cz};
                                    int cl = 1;
                                    do {
                                        for (int mg = 1; mg <= mgl; ++mg) {
                                            bool act = (cl == mg && !contains(1 << cl, mn, mx)) ||
                                                       (cl > mg && contains(1 << (cl - 1), mn, mx) && !contains(1 << cl, mn, mx));
                                            if (act) H.h[mg - 1][cl - 1].mask(ci[0], ci[1], ci[2], CellInterior);
                                        }
                                        for (int t = 0; t < 3; ++t) { mn[t] = (mn[t] / (1 << cl)) * (1 << cl); mx[t] = mn[t] + (1 << cl); ci[t] = (ci[t] + 1) / 2; }
                                        ++cl;
                                    } while (cl <= levels);
                                    go_next = true;
                                }
                            }
                    }
                }
            }
        }
        for (int m = 0; m < mgl; ++m) for (auto& g : H.h[m]) g.refresh();
        // --- DOF numbering per sigma level (:241-381)
        std::vector<int> node_count_lv(mgl, 0);
        for (int level = 1; level <= mgl; ++level) {
            auto& HL = H.h[level - 1];
            Grid<T>& G = HL[level - 1];
            if (level == 1) {
                for (int axis = 0; axis < 3; ++axis)
                    for (int face = 0; face < 2; ++face) {
                        int rem[2], r = 0;
                        for (int t = 0; t < 3; ++t) if (t != axis) rem[r++] = t;
                        for (int a = lo3[rem[0]]; a <= hi3[rem[0]]; ++a)
                            for (int b = lo3[rem[1]]; b <= hi3[rem[1]]; ++b) {
                                int n[3]; n[rem[0]] = a; n[rem[1]] = b; n[axis] = face == 0 ? lo3[axis] : hi3[axis];
                                if (set.is_set(n[0], n[1], n[2], Interface)) {
                                    int id = get_if_id(0, n[0], n[1], n[2], true);
                                    G.mask(n[0], n[1], n[2], UNodeInterface | NodeDegree);
                                    G.ivals[G.idx(n[0], n[1], n[2])] = id;
                                }
                                if (set.is_set(n[0], n[1], n[2], Active)) throw std::runtime_error("Active cell on a subdomain face");
                            }
                    }
            } else {
                int rl[3], rh[3];
                for (int a = 0; a < 3; ++a) { rl[a] = lo3[a] >> (level - 1); rh[a] = hi3[a] >> (level - 1); }
                std::vector<I3> cells;
                cells_with(G, CellInterior, cells);
                for (auto& cI : cells)
                    for (int nn = 0; nn < 8; ++nn) {
                        int n[3] = {cI.x - 1 + (nn >> 2), cI.y - 1 + ((nn >> 1) & 1), cI.z - 1 + (nn & 1)};
                        bool is_if = false;
                        for (int a = 0; a < 3; ++a) if (n[a] == rl[a] || n[a] == rh[a]) is_if = true;
                        if (!is_if) continue;
                        int id = get_if_id(level - 1, n[0], n[1], n[2], false);
                        G.mask(n[0], n[1], n[2], UNodeInterface | NodeDegree);
                        G.ivals[G.idx(n[0], n[1], n[2])] = id;
                    }
            }
            for (auto& g : HL) g.refresh();
            int node_count = 0;
            for (int hl = level; hl <= levels; ++hl) {
                Grid<T>& Gh = HL[hl - 1];
                std::vector<I3> cells;
                cells_with(Gh, CellInterior, cells);
                const Stride3& qh = q[hl];
                for (auto& cI : cells)
                    for (int nn = 0; nn < 8; ++nn) {
                        int n[3] = {cI.x - 1 + (nn >> 2), cI.y - 1 + ((nn >> 1) & 1), cI.z - 1 + (nn & 1)};
                        int fn[3] = {2 * n[0], 2 * n[1], 2 * n[2]};
                        if (hl == 1 && !set.is_set(n[0], n[1], n[2], Active)) continue;
                        if (hl == level && Gh.is_set(n[0], n[1], n[2], UNodeInterface)) continue;
                        int qs = std::min(qh.at(0, n[0]), std::min(qh.at(1, n[1]), qh.at(2, n[2])));
                        bool tj = false;
                        for (int a = 0; a < 3; ++a) if ((n[a] - qh.lo[a]) % qs != 0) tj = true;
                        if (!tj) {
                            if (hl > level) {
                                Grid<T>& Gf = HL[hl - 2];
                                if (Gf.is_set(fn[0], fn[1], fn[2], NodeDegree)) {
                                    int fid = Gf.ivals[Gf.idx(fn[0], fn[1], fn[2])];
                                    Gh.mask(n[0], n[1], n[2], NodeDegree | NodeCoarseShared);
                                    Gh.ivals[Gh.idx(n[0], n[1], n[2])] = fid;
                                }
                            }
                            if (!Gh.is_set(n[0], n[1], n[2], NodeDegree)) {
                                Gh.mask(n[0], n[1], n[2], NodeDegree | NodeActive);
                                Gh.ivals[Gh.idx(n[0], n[1], n[2])] = node_count++;
                            }
                        } else {
                            Gh.mask(n[0], n[1], n[2], NodeTJunction);
                            std::vector<std::pair<I3, T>> parents;
                            embed(n[0], n[1], n[2], parents, T(1), qh);
                            for (auto& pp : parents) {
                                const I3& pi = pp.first;
                                if (!Gh.is_set(pi.x, pi.y, pi.z, NodeDegree)) {
                                    Gh.mask(pi.x, pi.y, pi.z, NodeDegree | NodeActive);
                                    Gh.ivals[Gh.idx(pi.x, pi.y, pi.z)] = node_count++;
                                }
                            }
                        }
                    }
            }
            for (auto& g : HL) g.refresh();
            node_count_lv[level - 1] = node_count;
            Aii_sub[level - 1].assign(node_count, Stencil1<T>());
            Air_sub[level - 1].assign(node_count, Stencil1<T>());
        }
        // --- Galerkin assembly (:554-655)
        subset = &set;
        memo.assign(levels + 1, {});
        for (int level = 1; level <= mgl; ++level) {
            auto& HL = H.h[level - 1];
            for (int hl = level; hl <= levels; ++hl) {
                Grid<T>& Gh = HL[hl - 1];
                std::vector<I3> cells;
                cells_with(Gh, CellInterior, cells);
                for (auto& cI : cells) {
                    auto M = galerkin(cI.x, cI.y, cI.z, hl);
                    std::vector<std::pair<I3, T>> parents[8];
                    for (int nn = 0; nn < 8; ++nn)
                        embed(cI.x - 1 + (nn >> 2), cI.y - 1 + ((nn >> 1) & 1), cI.z - 1 + (nn & 1), parents[nn], T(1), q[hl]);
                    for (int n1 = 0; n1 < 8; ++n1)
                        for (int n2 = 0; n2 < 8; ++n2) {
                            T entry = M[n1 * 8 + n2];
                            if (entry == T(0)) continue;
                            for (auto& p1 : parents[n1])
                                for (auto& p2 : parents[n2]) {
                                    int id1, id2; uint32_t f1, f2;
                                    resolve(HL, hl, p1.first, id1, f1);
                                    resolve(HL, hl, p2.first, id2, f2);
                                    T add = p1.second * p2.second * entry;
                                    if ((f1 & UNodeInterface) && (f2 & UNodeInterface)) S.Arr_s[level - 1][id1]->get_or_insert(id2) += add;
                                    if ((f1 & (NodeActive | NodeCoarseShared)) && (f2 & (NodeActive | NodeCoarseShared)))
                                        Aii_sub[level - 1][id1].get_or_insert(id2) += add;
                                    if ((f1 & (NodeActive | NodeCoarseShared)) && (f2 & UNodeInterface))
                                        Air_sub[level - 1][id1].get_or_insert(id2) += add;
                                    if (!(f1 & NodeDegree) || !(f2 & NodeDegree)) throw std::runtime_error("unresolved adaptive node");
                                }
                        }
                }
            }
        }
        subset = nullptr;
        memo.clear();
    }

    // climb coarser levels until a node carries the degree marker (:598-628)
    void resolve(std::vector<Grid<T>>& HL, int hl, I3 n, int& id, uint32_t& f) {
        const Grid<T>* g = &HL[hl - 1];
        f = g->flag_at(n.x, n.y, n.z);
        int cl = hl;
        if (f & NodeDegree) { id = g->ivals[g->idx(n.x, n.y, n.z)]; return; }
        id = 0;
        while (!(f & NodeDegree)) {
            ++cl; n.x /= 2; n.y /= 2; n.z /= 2;
            if (cl > (int)HL.size()) throw std::runtime_error("adaptive node resolution ran past the coarsest level");
            g = &HL[cl - 1];
            size_t e = g->idx(n.x, n.y, n.z);
            id = e == (size_t)-1 ? 0 : g->ivals[e];
            f = g->flag_at(n.x, n.y, n.z);
        }
    }

    // INTERFACE_MATRIX_GENERATOR::Convert_To_Eigen (IFS/INTERFACE_MATRIX_GENERATOR.h:660-734)
    void convert() {
        const int mgl = S.mg_levels, ns = (int)S.Aii_s.size();
        S.nsub = ns;
        S.Arr.assign(mgl, {}); S.Aii.assign(mgl, std::vector<CSR<T>>(ns)); S.Air.assign(mgl, std::vector<CSR<T>>(ns));
        for (int l = 0; l < mgl; ++l) {
            std::vector<Stencil1<T>> rows(S.n_if[l]);
            for (int i = 0; i < S.n_if[l]; ++i) rows[i] = *S.Arr_s[l][i];
            S.Arr[l] = stencils_to_csr(rows, S.n_if[l]);
            for (int s = 0; s < ns; ++s) {
                S.Aii[l][s] = stencils_to_csr(S.Aii_s[s][l], (int)S.Aii_s[s][l].size());
                S.Air[l][s] = stencils_to_csr(S.Air_s[s][l], S.n_if[l]);
            }
        }
        for (int i = 0; i < S.n_if[0]; ++i) {
            bool diag = false;
            for (int k = S.Arr[0].ptr[i]; k < S.Arr[0].ptr[i + 1]; ++k) if (S.Arr[0].idx[k] == i) diag = true;
            if (!diag) throw std::runtime_error("Arr has a zero diagonal entry");
        }
        S.Prr.assign(std::max(0, mgl - 1), {});
        for (int l = 0; l + 1 < mgl; ++l) {
            std::vector<Stencil1<T>> rows(S.n_if[l]);
            for (int i = 0; i < S.n_if[l]; ++i) {
                const I3 f = S.icoord[l][i];
                int fc[3] = {f.x, f.y, f.z};
                int cn[3] = {f.x / 2, f.y / 2, f.z / 2};
                for (int nb = 0; nb < 8; ++nb) {
                    int c[3] = {cn[0] + (nb >> 2), cn[1] + ((nb >> 1) & 1), cn[2] + (nb & 1)};
                    T w = T(1);
                    for (int v = 0; v < 3; ++v) w *= (T)((double)(T(2) - (T)std::abs(c[v] * 2 - fc[v])) / 2.0);
                    if (w > 0) {
                        auto it = S.ihash[l + 1].find(key3(c[0], c[1], c[2]));
                        if (it == S.ihash[l + 1].end()) continue;
                        rows[i].get_or_insert(it->second) += w;
                    }
                }
            }
            S.Prr[l] = stencils_to_csr(rows, S.n_if[l + 1]);
        }
    }
};

// Split (DDS/SPGrid_Subdomain_Splitter.h:43-158, #else branch): per subdomain
// in RANGE order, rasterise the closed box, mark its Active cells, keep cells
// with an Active face neighbour, divide, and build the adaptive hierarchy.
template <class T>
void split_and_generate(const Grid<T>& main, I3 size, I3 sub, SigmaSetup<T>& S, int sigma_levels) {
    InterfaceGenerator<T> gen(S);
    gen.size = size;
    gen.initialize(sigma_levels);
    I3 ns = number_of_subdomains(size, sub);
    for (int sx = 0; sx < ns.x; ++sx)
        for (int sy = 0; sy < ns.y; ++sy)
            for (int sz = 0; sz < ns.z; ++sz) {
                I3 lo = {sx * sub.x, sy * sub.y, sz * sub.z}, hi = {lo.x + sub.x, lo.y + sub.y, lo.z + sub.z};
                auto in_box = [&](int x, int y, int z) { return x >= lo.x && y >= lo.y && z >= lo.z && x <= hi.x && y <= hi.y && z <= hi.z; };
                Grid<T> sset;
                sset.init(lo.x - 8, lo.y - 8, lo.z - 8, hi.x + 9, hi.y + 9, hi.z + 9, 0);
                // Subdomain_Rasterizer (DDS/SPGrid_Subdomain_Rasterizer.h:37-55)
                for (int bx = lo.x / BX; bx <= hi.x / BX; ++bx)
                    for (int by = lo.y / BY; by <= hi.y / BY; ++by)
                        for (int bz = lo.z / BZ; bz <= hi.z / BZ; ++bz) {
                            int p = main.page_of_block(bx, by, bz);
                            if (p < 0) continue;
                            for (int c = 0; c < BCELLS; ++c) {
                                int x = bx * BX + (c >> 6), y = by * BY + ((c >> 3) & 7), z = bz * BZ + (c & 7);
                                if (in_box(x, y, z)) sset.mask(x, y, z, main.flags[(size_t)p * BCELLS + c]);
                            }
                        }
                sset.refresh();
                for (auto& f : sset.flags) if (f & Active) f = Boundary;
                // Subdomain_Masked_Rasterizer (DDS/SPGrid_Subdomain_Rasterizer.h:87-138)
                Grid<T> sub2;
                sub2.init(lo.x - 8, lo.y - 8, lo.z - 8, hi.x + 9, hi.y + 9, hi.z + 9, 0);
                for (int bx = lo.x / BX; bx <= hi.x / BX; ++bx)
                    for (int by = lo.y / BY; by <= hi.y / BY; ++by)
                        for (int bz = lo.z / BZ; bz <= hi.z / BZ; ++bz) {
                            if (sset.page_of_block(bx, by, bz) < 0) continue;
                            const int X = bx * BX, Y = by * BY, Z = bz * BZ;
                            uint32_t fl[6][10][10];
                            for (int i = -1; i <= BX; ++i)
                                for (int j = -1; j <= BY; ++j)
                                    for (int k = -1; k <= BZ; ++k) {
                                        int x = X + i, y = Y + j, z = Z + k;
                                        uint32_t v = 0;
                                        if (in_box(x, y, z) && sset.page_of(x, y, z) >= 0) {
                                            v = sset.flag_at(x, y, z);
                                            if (v & Boundary) v = Active;
                                            else if (v & Active) v = 0;
                                        }
                                        fl[i + 1][j + 1][k + 1] = v;
                                    }
                            bool found = false;
                            for (int i = 0; i < BX; ++i)
                                for (int j = 0; j < BY; ++j)
                                    for (int k = 0; k < BZ; ++k) {
                                        uint32_t& v = fl[i + 1][j + 1][k + 1];
                                        if (v & Active) found = true;
                                        if ((v & Dirichlet) || v) {
                                            bool an = (fl[i + 2][j + 1][k + 1] & Active) || (fl[i][j + 1][k + 1] & Active) ||
                                                      (fl[i + 1][j + 2][k + 1] & Active) || (fl[i + 1][j][k + 1] & Active) ||
                                                      (fl[i + 1][j + 1][k + 2] & Active) || (fl[i + 1][j + 1][k] & Active);
                                            if (an) found = true; else v = 0;
                                        }
                                    }
                            if (found) {
                                int p = sub2.activate_block(bx, by, bz);
                                for (int c = 0; c < BCELLS; ++c) sub2.flags[(size_t)p * BCELLS + c] = fl[(c >> 6) + 1][((c >> 3) & 7) + 1][(c & 7) + 1];
                            }
                        }
                sub2.refresh();
                domain_divider(sub2, size, sub);
                gen.create_subdomain_hierarchy(sub2, lo, hi);
            }
    gen.convert();
}

}  // namespace ora
