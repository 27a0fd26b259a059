// Interface (Schur complement) operator generator of the product.
//
// Builds, per sigma level, the global interface operator Arr, the per-
// subdomain interior operators Aii and couplings Air, and the interface
// prolongations Prr, with the semantics of
//   SPGrid_Subdomain_Splitter::Split   DDS/SPGrid_Subdomain_Splitter.h:43-158
//   Subdomain_{,Masked_}Rasterizer     DDS/SPGrid_Subdomain_Rasterizer.h:37-138
//   STRIDE_GENERATOR                   IFS/STRIDE_GENERATOR.h:23-54
//   CELL_PROLONGATION_BAKER            IFS/CELL_PROLONGATION_BAKER.h:22-51
//   GALERKIN_PROCESS::Galerkin_Coarsen IFS/GALERKIN_PROCESS.h:27-68
//   INTERFACE_MATRIX_GENERATOR         IFS/INTERFACE_MATRIX_GENERATOR.h:81-734
//   PROLONGATION_MATRIX_HELPER         IFS/PROLONGATION_MATRIX_HELPER.cpp:13-40
//
// Design: every subdomain works on dense local arrays over its own box
// (no global sparse hierarchy), so subdomains run in parallel. The only
// order-dependent global results — interface dof numbering and the float
// accumulation of Arr — are replayed serially in subdomain order afterwards,
// which reproduces the reference's sequential results exactly.
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include "hdd_parallel.h"
#include "hdd_setup.h"
#include "hdd_sigma.h"

namespace hdd {
namespace {

enum : uint8_t { TNone = 0, TActive = 1, TInterface = 2, TDirichlet = 3 };
// node flag bits inside one (sigma level, hierarchy level) array
enum : uint8_t { NIface = 1, NDegree = 2, NActive = 4, NShared = 8, NTJunction = 16 };

inline uint64_t pack3(int x, int y, int z) { return ((uint64_t)(uint32_t)x << 42) | ((uint64_t)(uint32_t)y << 21) | (uint64_t)(uint32_t)z; }

// STRIDE_GENERATOR for one axis over [lo, hi]
int stride_axis(int lo, int hi, std::vector<int>& q) {
    const int n = hi - lo;
    int level = 0;
    std::vector<std::pair<int, int>> seq{{2, 1}};
    int cur = 1, covered = 2;
    while (covered < n) {
        const int k = std::min(2, (n - covered) / cur);
        seq.push_back({k, cur});
        covered += k * cur;
        if ((n - covered) % (2 * cur) == 0) { cur *= 2; ++level; }
    }
    q.assign(n + 1, 0);
    int w = 0;
    std::vector<int> back;
    for (size_t i = 0; i < seq.size(); ++i) {
        for (int k = 0; k < seq[i].second; ++k) q.at(w++) = seq[i].second;
        if (seq[i].first == 1) q.at(w++) = seq[i].second;
        else { if (i + 1 == seq.size()) q.at(w++) = seq[i].second; back.push_back(seq[i].second); }
    }
    for (auto it = back.rbegin(); it != back.rend(); ++it)
        for (int k = 0; k < *it; ++k) q.at(w++) = *it;
    return level;
}

struct Strides {
    int lo[3];
    std::vector<int> q[3];
    int at(int a, int i) const { return q[a][i - lo[a]]; }
};

// Embedding_Helper (IFS/INTERFACE_MATRIX_GENERATOR.h:81-100): split a hanging
// node linearly along the first axis where it is off the local stride lattice.
void embed_node(int x, int y, int z, float w, const Strides& q, std::vector<std::pair<Int3, float>>& out) {
    const int idx[3] = {x, y, z};
    const int qs = std::min(q.at(0, x), std::min(q.at(1, y), q.at(2, z)));
    for (int a = 0; a < 3; ++a) {
        const int cur = idx[a] - q.lo[a];
        if (cur % qs) {
            int i1[3] = {x, y, z}, i2[3] = {x, y, z};
            i1[a] = (cur / qs) * qs + q.lo[a];
            i2[a] = (cur / qs + 1) * qs + q.lo[a];
            const float w1 = w * (float)(i2[a] - idx[a]) / (float)(i2[a] - i1[a]);
            const float w2 = w * (float)(idx[a] - i1[a]) / (float)(i2[a] - i1[a]);
            embed_node(i1[0], i1[1], i1[2], w1, q, out);
            embed_node(i2[0], i2[1], i2[2], w2, q, out);
            return;
        }
    }
    for (auto& p : out)
        if (p.first.x == x && p.first.y == y && p.first.z == z) { p.second += w; return; }
    out.push_back({{x, y, z}, w});
}

// cell prolongation matrices P[child][fine corner][coarse corner], corner id 4x+2y+z
struct CellProlongation {
    float P[8][8][8];
    CellProlongation() {
        std::memset(P, 0, sizeof(P));
        for (int ch = 0; ch < 8; ++ch)
            for (int fi = 0; fi < 8; ++fi) {
                const int f[3] = {(fi >> 2) + (ch >> 2), ((fi >> 1) & 1) + ((ch >> 1) & 1), (fi & 1) + (ch & 1)};
                for (int ci = 0; ci < 8; ++ci) {
                    const int c[3] = {f[0] / 2 + (ci >> 2), f[1] / 2 + ((ci >> 1) & 1), f[2] / 2 + (ci & 1)};
                    float w = 1.f;
                    for (int a = 0; a < 3; ++a) w *= (float)((double)(2.f - (float)std::abs(2 * c[a] - f[a])) / 2.0);
                    if (w != 0.f) P[ch][fi][(c[0] << 2) | (c[1] << 1) | c[2]] = w;
                }
            }
    }
};
const CellProlongation& cell_prolongation() { static CellProlongation cp; return cp; }

struct Box {  // dense box [lo, hi] inclusive
    Int3 lo, hi;
    int nx = 0, ny = 0, nz = 0;
    void set(Int3 l, Int3 h) { lo = l; hi = h; nx = h.x - l.x + 1; ny = h.y - l.y + 1; nz = h.z - l.z + 1; }
    bool inside(int x, int y, int z) const { return x >= lo.x && y >= lo.y && z >= lo.z && x <= hi.x && y <= hi.y && z <= hi.z; }
    size_t at(int x, int y, int z) const {
        if (!inside(x, y, z)) throw std::runtime_error("interface generator: index outside subdomain box");
        return ((size_t)(x - lo.x) * ny + (y - lo.y)) * nz + (z - lo.z);
    }
    size_t size() const { return (size_t)nx * ny * nz; }
};

struct LevelArrays {  // one hierarchy level of one sigma level
    Box box;
    std::vector<uint8_t> interior;  // cell flags (cell index = its max-corner node)
    std::vector<uint8_t> node;      // node flag bits
    std::vector<int> index;         // dof id (interior id, or interface id for interface nodes)
};

struct ArrContribution { int level, i, j; float v; };

struct Subdomain {
    Int3 lo, hi;
    int levels = 0;
    Box fbox;                        // [lo-1, hi+1]
    std::vector<uint8_t> type;       // TNone / TActive / TInterface / TDirichlet after split + divide
    std::vector<Strides> q;          // [1..levels]
    std::vector<std::vector<LevelArrays>> H;  // [sigma level 0..][hierarchy level 0..levels-1]
    std::vector<std::vector<Int3>> iface_order;  // per sigma level: interface nodes in first-visit order
    std::vector<int> n_interior;     // per sigma level
    std::vector<std::vector<Int3>> interior_coords;  // per sigma level: fine coordinates of each interior dof
    // assembled outputs (row stencils in insertion order)
    std::vector<std::vector<std::vector<std::pair<int, float>>>> Aii, Air;  // [level][row]
    std::vector<ArrContribution> arr;

    uint8_t t(int x, int y, int z) const { return fbox.inside(x, y, z) ? type[fbox.at(x, y, z)] : TNone; }
};

// --- split: flags of the subdomain set after rasterising, masking and dividing
void split_subdomain(const PagedLevel& in, Int3 sub, Subdomain& S) {
    S.fbox.set({S.lo.x - 1, S.lo.y - 1, S.lo.z - 1}, {S.hi.x + 1, S.hi.y + 1, S.hi.z + 1});
    S.type.assign(S.fbox.size(), TNone);
    auto in_box = [&](int x, int y, int z) { return x >= S.lo.x && y >= S.lo.y && z >= S.lo.z && x <= S.hi.x && y <= S.hi.y && z <= S.hi.z; };
    // masked flag of a cell before the per-page pass: Active for input Active cells,
    // the input flag otherwise; 0 outside the closed box or on unallocated pages
    auto masked = [&](int x, int y, int z) -> uint32_t {
        if (!in_box(x, y, z) || in.page_at(x, y, z) < 0) return 0u;
        const uint32_t f = in.flag(x, y, z);
        return (f & kActive) ? kActive : f;
    };
    for (int bx = S.lo.x / kPX; bx <= S.hi.x / kPX; ++bx)
        for (int by = S.lo.y / kPY; by <= S.hi.y / kPY; ++by)
            for (int bz = S.lo.z / kPZ; bz <= S.hi.z / kPZ; ++bz) {
                if (in.page_at_block(bx, by, bz) < 0) continue;
                const int X = bx * kPX, Y = by * kPY, Z = bz * kPZ;
                uint32_t w[kPX + 2][kPY + 2][kPZ + 2];
                for (int i = -1; i <= kPX; ++i)
                    for (int j = -1; j <= kPY; ++j)
                        for (int k = -1; k <= kPZ; ++k) w[i + 1][j + 1][k + 1] = masked(X + i, Y + j, Z + k);
                bool keep = false;
                for (int i = 1; i <= kPX; ++i)
                    for (int j = 1; j <= kPY; ++j)
                        for (int k = 1; k <= kPZ; ++k) {
                            uint32_t& v = w[i][j][k];
                            if (v & kActive) keep = true;
                            if (!v) continue;
                            const bool nb = ((w[i - 1][j][k] | w[i + 1][j][k] | w[i][j - 1][k] | w[i][j + 1][k] |
                                              w[i][j][k - 1] | w[i][j][k + 1]) & kActive) != 0;
                            if (nb) keep = true; else v = 0;
                        }
                if (!keep) continue;
                for (int i = 0; i < kPX; ++i)
                    for (int j = 0; j < kPY; ++j)
                        for (int k = 0; k < kPZ; ++k) {
                            const int x = X + i, y = Y + j, z = Z + k;
                            if (!S.fbox.inside(x, y, z)) continue;
                            const uint32_t v = w[i + 1][j + 1][k + 1];
                            uint8_t tt = TNone;
                            if (v & kActive) tt = (x % sub.x == 0 || y % sub.y == 0 || z % sub.z == 0) ? TInterface : TActive;
                            else if (v & kDirichlet) tt = TDirichlet;
                            S.type[S.fbox.at(x, y, z)] = tt;
                        }
            }
}

inline bool isA(uint8_t t) { return t == TActive || t == TInterface; }

// --- activation of the adaptive cells (:184-239), per-level dof numbering (:241-381)
void build_hierarchy(Subdomain& S, int mgl) {
    const int lo3[3] = {S.lo.x, S.lo.y, S.lo.z}, hi3[3] = {S.hi.x, S.hi.y, S.hi.z};
    S.levels = 1 << 30;
    S.q.assign(1, Strides());
    {
        Strides top;
        for (int a = 0; a < 3; ++a) S.levels = std::min(S.levels, stride_axis(lo3[a], hi3[a], top.q[a]));
    }
    if (mgl > S.levels) throw std::runtime_error("sigma_levels exceeds the adaptive levels of a subdomain");
    const int L = S.levels;
    S.q.resize(L + 1);
    for (int l = 1; l <= L; ++l)
        for (int a = 0; a < 3; ++a) { S.q[l].lo[a] = lo3[a] >> (l - 1); stride_axis(lo3[a] >> (l - 1), hi3[a] >> (l - 1), S.q[l].q[a]); }
    S.H.assign(mgl, std::vector<LevelArrays>(L));
    for (int m = 0; m < mgl; ++m)
        for (int l = 1; l <= L; ++l) {
            LevelArrays& A = S.H[m][l - 1];
            A.box.set({S.lo.x >> (l - 1), S.lo.y >> (l - 1), S.lo.z >> (l - 1)}, {S.hi.x >> (l - 1), S.hi.y >> (l - 1), S.hi.z >> (l - 1)});
            A.interior.assign(A.box.size(), 0);
            A.node.assign(A.box.size(), 0);
            A.index.assign(A.box.size(), 0);
        }
    auto contains = [&](int t, const int mn[3], const int mx[3]) {
        for (int a = 0; a < 3; ++a) if (!(lo3[a] + t <= mn[a] && hi3[a] - t >= mx[a])) return false;
        return true;
    };
    // a level-1 cell c (nodes c-1..c) is activated iff one of its 12 edges joins
    // two Active|Interface nodes, or an Active|Interface node and a Dirichlet node
    for (int cx = S.lo.x + 1; cx <= S.hi.x; ++cx)
        for (int cy = S.lo.y + 1; cy <= S.hi.y; ++cy)
            for (int cz = S.lo.z + 1; cz <= S.hi.z; ++cz) {
                uint8_t t8[8];
                for (int n = 0; n < 8; ++n) t8[n] = S.t(cx - 1 + (n >> 2), cy - 1 + ((n >> 1) & 1), cz - 1 + (n & 1));
                bool act = false;
                for (int e = 0; e < 8 && !act; ++e)
                    for (int a = 0; a < 3 && !act; ++a) {
                        const int bit = 4 >> a;
                        if (e & bit) continue;
                        const uint8_t u = t8[e], v = t8[e | bit];
                        if ((isA(u) && isA(v)) || (isA(u) && v == TDirichlet) || (u == TDirichlet && isA(v))) act = true;
                    }
                if (!act) continue;
                int ci[3] = {cx, cy, cz}, mn[3] = {cx - 1, cy - 1, cz - 1}, mx[3] = {cx, cy, cz};
                for (int cl = 1; cl <= L; ++cl) {
                    for (int m = 1; m <= mgl; ++m) {
                        const bool on = (cl == m && !contains(1 << cl, mn, mx)) ||
                                        (cl > m && contains(1 << (cl - 1), mn, mx) && !contains(1 << cl, mn, mx));
                        if (on) { LevelArrays& A = S.H[m - 1][cl - 1]; A.interior[A.box.at(ci[0], ci[1], ci[2])] = 1; }
                    }
                    for (int a = 0; a < 3; ++a) { mn[a] = (mn[a] / (1 << cl)) * (1 << cl); mx[a] = mn[a] + (1 << cl); ci[a] = (ci[a] + 1) / 2; }
                }
            }
    S.iface_order.assign(mgl, {});
    S.n_interior.assign(mgl, 0);
    S.interior_coords.assign(mgl, {});
    for (int m = 1; m <= mgl; ++m) {
        auto& HL = S.H[m - 1];
        LevelArrays& G = HL[m - 1];
        auto& order = S.iface_order[m - 1];
        auto mark_iface = [&](int x, int y, int z) {
            const size_t e = G.box.at(x, y, z);
            if (!(G.node[e] & NIface)) order.push_back({x, y, z});  // first visit in this subdomain
            G.node[e] |= NIface | NDegree;
        };
        if (m == 1) {
            for (int axis = 0; axis < 3; ++axis)
                for (int face = 0; face < 2; ++face) {
                    const int r0 = axis == 0 ? 1 : 0, r1 = axis == 2 ? 1 : 2;
                    for (int a = lo3[r0]; a <= hi3[r0]; ++a)
                        for (int b = lo3[r1]; b <= hi3[r1]; ++b) {
                            int n[3]; n[r0] = a; n[r1] = b; n[axis] = face ? hi3[axis] : lo3[axis];
                            const uint8_t tt = S.t(n[0], n[1], n[2]);
                            if (tt == TActive) throw std::runtime_error("Active cell on a subdomain face");
                            if (tt == TInterface) mark_iface(n[0], n[1], n[2]);
                        }
                }
        } else {
            std::vector<std::pair<uint64_t, Int3>> cells;
            for (size_t e = 0; e < G.interior.size(); ++e)
                if (G.interior[e]) {
                    const int x = G.box.lo.x + (int)(e / ((size_t)G.box.ny * G.box.nz)), y = G.box.lo.y + (int)((e / G.box.nz) % G.box.ny), z = G.box.lo.z + (int)(e % G.box.nz);
                    cells.push_back({spgrid_offset(x, y, z), {x, y, z}});
                }
            std::sort(cells.begin(), cells.end(), [](auto& a, auto& b) { return a.first < b.first; });
            for (auto& c : cells)
                for (int nn = 0; nn < 8; ++nn) {
                    const int n[3] = {c.second.x - 1 + (nn >> 2), c.second.y - 1 + ((nn >> 1) & 1), c.second.z - 1 + (nn & 1)};
                    bool on = false;
                    for (int a = 0; a < 3; ++a) if (n[a] == (lo3[a] >> (m - 1)) || n[a] == (hi3[a] >> (m - 1))) on = true;
                    if (on) mark_iface(n[0], n[1], n[2]);
                }
        }
        // interior dofs, hierarchy levels m..L
        int count = 0;
        auto& icoords = S.interior_coords[m - 1];
        for (int hl = m; hl <= L; ++hl) {
            LevelArrays& Gh = HL[hl - 1];
            const Strides& qh = S.q[hl];
            std::vector<std::pair<uint64_t, Int3>> cells;
            for (size_t e = 0; e < Gh.interior.size(); ++e)
                if (Gh.interior[e]) {
                    const int x = Gh.box.lo.x + (int)(e / ((size_t)Gh.box.ny * Gh.box.nz)), y = Gh.box.lo.y + (int)((e / Gh.box.nz) % Gh.box.ny), z = Gh.box.lo.z + (int)(e % Gh.box.nz);
                    cells.push_back({spgrid_offset(x, y, z), {x, y, z}});
                }
            std::sort(cells.begin(), cells.end(), [](auto& a, auto& b) { return a.first < b.first; });
            for (auto& c : cells)
                for (int nn = 0; nn < 8; ++nn) {
                    const int n[3] = {c.second.x - 1 + (nn >> 2), c.second.y - 1 + ((nn >> 1) & 1), c.second.z - 1 + (nn & 1)};
                    if (hl == 1 && S.t(n[0], n[1], n[2]) != TActive) continue;
                    const size_t e = Gh.box.at(n[0], n[1], n[2]);
                    if (hl == m && (Gh.node[e] & NIface)) continue;
                    const int qs = std::min(qh.at(0, n[0]), std::min(qh.at(1, n[1]), qh.at(2, n[2])));
                    bool tj = false;
                    for (int a = 0; a < 3; ++a) if ((n[a] - qh.lo[a]) % qs) tj = true;
                    if (!tj) {
                        if (hl > m) {
                            const LevelArrays& Gf = HL[hl - 2];
                            if (Gf.box.inside(2 * n[0], 2 * n[1], 2 * n[2])) {
                                const size_t fe = Gf.box.at(2 * n[0], 2 * n[1], 2 * n[2]);
                                if (Gf.node[fe] & NDegree) { Gh.index[e] = Gf.index[fe]; Gh.node[e] |= NDegree | NShared; }
                            }
                        }
                        if (!(Gh.node[e] & NDegree)) {
                            Gh.node[e] |= NDegree | NActive; Gh.index[e] = count++;
                            icoords.push_back({n[0] << (hl - 1), n[1] << (hl - 1), n[2] << (hl - 1)});
                        }
                    } else {
                        Gh.node[e] |= NTJunction;
                        // parents of a T-junction node (Embedding_Helper), registered in insertion order
                        std::vector<std::pair<Int3, float>> par;
                        embed_node(n[0], n[1], n[2], 1.f, qh, par);
                        for (auto& pp : par) {
                            const size_t pe = Gh.box.at(pp.first.x, pp.first.y, pp.first.z);
                            if (!(Gh.node[pe] & NDegree)) {
                                Gh.node[pe] |= NDegree | NActive; Gh.index[pe] = count++;
                                icoords.push_back({pp.first.x << (hl - 1), pp.first.y << (hl - 1), pp.first.z << (hl - 1)});
                            }
                        }
                    }
                }
        }
        S.n_interior[m - 1] = count;
    }
}

using Mat8 = std::array<float, 64>;

// Galerkin cell operator at hierarchy level `level` (memoised per subdomain)
struct Galerkin {
    const Subdomain& S;
    std::vector<std::unordered_map<uint64_t, Mat8>> memo;
    explicit Galerkin(const Subdomain& s) : S(s), memo(s.levels + 1) {}
    const Mat8& get(int cx, int cy, int cz, int level) {
        auto& mm = memo[level];
        const uint64_t k = pack3(cx, cy, cz);
        auto it = mm.find(k);
        if (it != mm.end()) return it->second;
        Mat8 M; M.fill(0.f);
        if (level == 1) {
            const float ew = (float)(1.0 / 4.0);
            uint8_t nf[8];  // 1 = Active|Interface, 2 = Dirichlet
            for (int n = 0; n < 8; ++n) {
                const uint8_t tt = S.t(cx - 1 + (n >> 2), cy - 1 + ((n >> 1) & 1), cz - 1 + (n & 1));
                nf[n] = isA(tt) ? 1 : (tt == TDirichlet ? 2 : 0);
            }
            // edges in the reference order: axis v, then the two other axes ascending
            for (int v = 0; v < 3; ++v) {
                const int bit = 4 >> v;
                const int s0 = (v == 0) ? 2 : 4, s1 = (v == 2) ? 2 : 1;  // corner bits of the two other axes
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        const int lo = a * s0 + b * s1, up = lo | bit;
                        if (nf[lo] == 1 && nf[up] == 1) { M[lo * 8 + up] -= ew; M[lo * 8 + lo] += ew; M[up * 8 + up] += ew; M[up * 8 + lo] -= ew; }
                        if (nf[lo] == 1 && nf[up] == 2) M[lo * 8 + lo] += ew;
                        if (nf[lo] == 2 && nf[up] == 1) M[up * 8 + up] += ew;
                    }
            }
        } else {
            const auto& CP = cell_prolongation();
            for (int ch = 0; ch < 8; ++ch) {
                const Mat8 C = get(2 * cx - 1 + (ch >> 2), 2 * cy - 1 + ((ch >> 1) & 1), 2 * cz - 1 + (ch & 1), level - 1);
                const auto& P = CP.P[ch];
                float T[64];
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j) {
                        float s = 0.f;
                        for (int k = 0; k < 8; ++k) s += P[k][i] * C[k * 8 + j];
                        T[i * 8 + j] = s;
                    }
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j) {
                        float s = 0.f;
                        for (int k = 0; k < 8; ++k) s += T[i * 8 + k] * P[k][j];
                        M[i * 8 + j] += s;
                    }
            }
        }
        return mm.emplace(k, M).first->second;
    }
};

// Galerkin assembly (:554-655) for one subdomain; needs global interface ids.
void assemble(Subdomain& S, int mgl, const std::vector<std::unordered_map<uint64_t, int>>& if_id) {
    Galerkin G(S);
    S.Aii.assign(mgl, {}); S.Air.assign(mgl, {});
    for (int m = 1; m <= mgl; ++m) {
        auto& HL = S.H[m - 1];
        auto& Aii = S.Aii[m - 1];
        auto& Air = S.Air[m - 1];
        Aii.assign(S.n_interior[m - 1], {});
        Air.assign(S.n_interior[m - 1], {});
        auto add = [](std::vector<std::pair<int, float>>& row, int col, float v) {
            for (auto& p : row) if (p.first == col) { p.second += v; return; }
            row.push_back({col, v});
        };
        // dof of a node at hierarchy level hl: climb to coarser levels until it is a dof
        // dof of a node at hierarchy level hl: climb to coarser levels until it is a
        // dof; fl = 0 when no level holds one (such nodes only meet zero entries)
        auto resolve = [&](int hl, Int3 n, int& id, uint8_t& fl) {
            for (int cl = hl; cl <= S.levels; ++cl) {
                const LevelArrays& A = HL[cl - 1];
                const size_t e = A.box.inside(n.x, n.y, n.z) ? A.box.at(n.x, n.y, n.z) : (size_t)-1;
                fl = e == (size_t)-1 ? 0 : A.node[e];
                id = e == (size_t)-1 ? 0 : A.index[e];
                if (fl & NDegree) {
                    if (fl & NIface) {
                        auto it = if_id[m - 1].find(pack3(n.x, n.y, n.z));
                        if (it == if_id[m - 1].end()) throw std::runtime_error("interface node without a global id");
                        id = it->second;
                    }
                    return;
                }
                n.x /= 2; n.y /= 2; n.z /= 2;
            }
            fl = 0; id = -1;
        };
        for (int hl = m; hl <= S.levels; ++hl) {
            const LevelArrays& Gh = HL[hl - 1];
            std::vector<std::pair<uint64_t, Int3>> cells;
            for (size_t e = 0; e < Gh.interior.size(); ++e)
                if (Gh.interior[e]) {
                    const int x = Gh.box.lo.x + (int)(e / ((size_t)Gh.box.ny * Gh.box.nz)), y = Gh.box.lo.y + (int)((e / Gh.box.nz) % Gh.box.ny), z = Gh.box.lo.z + (int)(e % Gh.box.nz);
                    cells.push_back({spgrid_offset(x, y, z), {x, y, z}});
                }
            std::sort(cells.begin(), cells.end(), [](auto& a, auto& b) { return a.first < b.first; });
            for (auto& c : cells) {
                const Int3 ci = c.second;
                const Mat8 M = G.get(ci.x, ci.y, ci.z, hl);
                std::vector<std::pair<Int3, float>> par[8];
                int pid[8][8]; uint8_t pfl[8][8];
                for (int nn = 0; nn < 8; ++nn) {
                    embed_node(ci.x - 1 + (nn >> 2), ci.y - 1 + ((nn >> 1) & 1), ci.z - 1 + (nn & 1), 1.f, S.q[hl], par[nn]);
                    if (par[nn].size() > 8) throw std::runtime_error("hanging node with more than 8 parents");
                    for (size_t k = 0; k < par[nn].size(); ++k) resolve(hl, par[nn][k].first, pid[nn][k], pfl[nn][k]);
                }
                for (int n1 = 0; n1 < 8; ++n1)
                    for (int n2 = 0; n2 < 8; ++n2) {
                        const float entry = M[n1 * 8 + n2];
                        if (entry == 0.f) continue;
                        for (size_t k1 = 0; k1 < par[n1].size(); ++k1)
                            for (size_t k2 = 0; k2 < par[n2].size(); ++k2) {
                                const float v = par[n1][k1].second * par[n2][k2].second * entry;
                                const uint8_t f1 = pfl[n1][k1], f2 = pfl[n2][k2];
                                if (!(f1 & NDegree) || !(f2 & NDegree))
                                    throw std::runtime_error("adaptive node without a dof at any coarser level (subdomain " +
                                                             std::to_string(S.lo.x) + "," + std::to_string(S.lo.y) + "," + std::to_string(S.lo.z) + ")");
                                const int i1 = pid[n1][k1], i2 = pid[n2][k2];
                                if ((f1 & NIface) && (f2 & NIface)) S.arr.push_back({m - 1, i1, i2, v});
                                if ((f1 & (NActive | NShared)) && (f2 & (NActive | NShared))) add(Aii[i1], i2, v);
                                if ((f1 & (NActive | NShared)) && (f2 & NIface)) add(Air[i1], i2, v);
                            }
                    }
            }
        }
    }
}

Csr rows_to_csr(const std::vector<std::vector<std::pair<int, float>>>& rows, int ncols) {
    Csr m; m.rows = (int)rows.size(); m.cols = ncols; m.ptr.assign(1, 0);
    for (auto& r : rows) {
        std::vector<std::pair<int, float>> v;
        for (auto& p : r) if (p.second != 0.f) v.push_back(p);
        std::sort(v.begin(), v.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (auto& p : v) { m.idx.push_back(p.first); m.val.push_back(p.second); }
        m.ptr.push_back((int)m.idx.size());
    }
    return m;
}

}  // namespace

void generate_interface_operators(const PagedLevel& input, Int3 size, Int3 sub, int mgl, SigmaHost& out) {
    // phase timings on stderr when HDD_FACTOR_VERBOSE is set
    static const bool verbose = std::getenv("HDD_FACTOR_VERBOSE") != nullptr;
    auto tp = std::chrono::steady_clock::now();
    auto phase = [&](const char* what) {
        const auto now = std::chrono::steady_clock::now();
        if (verbose) std::fprintf(stderr, "[setup]   interface generator, %s: %.3fs\n", what, std::chrono::duration<double>(now - tp).count());
        tp = now;
    };
    const Int3 ns = subdomain_counts(size, sub);
    const int nsub = ns.x * ns.y * ns.z;
    std::vector<Subdomain> subs(nsub);
    ParallelErrors errs;
    // phase 1 (parallel): split, activate, number local dofs
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; ++s)
        errs.run([&] {
            const int sx = s / (ns.y * ns.z), sy = (s / ns.z) % ns.y, sz = s % ns.z;
            Subdomain& S = subs[s];
            S.lo = {sx * sub.x, sy * sub.y, sz * sub.z};
            S.hi = {S.lo.x + sub.x, S.lo.y + sub.y, S.lo.z + sub.z};
            split_subdomain(input, sub, S);
            build_hierarchy(S, mgl);
        });
    errs.rethrow();
    phase("split and hierarchy");
    // phase 2 (serial, subdomain order): global interface numbering per level
    out.levels = mgl; out.nsub = nsub;
    out.n_if.assign(mgl, 0);
    out.if_coords.assign(mgl, {});
    out.interface_map.clear();
    std::vector<std::unordered_map<uint64_t, int>> if_id(mgl);
    for (int s = 0; s < nsub; ++s)
        for (int l = 0; l < mgl; ++l)
            for (const Int3& c : subs[s].iface_order[l]) {
                auto ins = if_id[l].emplace(pack3(c.x, c.y, c.z), out.n_if[l]);
                if (ins.second) {
                    out.n_if[l]++;
                    out.if_coords[l].push_back(c);
                    if (l == 0) out.interface_map.push_back(c);
                }
            }
    phase("interface numbering");
    // phase 3 (parallel): Galerkin assembly per subdomain
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; ++s) errs.run([&] { assemble(subs[s], mgl, if_id); });
    errs.rethrow();
    phase("Galerkin assembly");
    // phase 4 (serial, subdomain order): accumulate Arr in the reference's order
    std::vector<std::vector<std::vector<std::pair<int, float>>>> arr(mgl);
    for (int l = 0; l < mgl; ++l) arr[l].assign(out.n_if[l], {});
    for (int s = 0; s < nsub; ++s) {
        for (const ArrContribution& c : subs[s].arr) {
            auto& row = arr[c.level][c.i];
            bool found = false;
            for (auto& p : row) if (p.first == c.j) { p.second += c.v; found = true; break; }
            if (!found) row.push_back({c.j, c.v});
        }
        std::vector<ArrContribution>().swap(subs[s].arr);
    }
    out.Arr.assign(mgl, {});
    out.Aii.assign(mgl, std::vector<Csr>(nsub));
    out.Air.assign(mgl, std::vector<Csr>(nsub));
    out.Aii_coords.assign(mgl, std::vector<std::vector<Int3>>(nsub));
    phase("Arr accumulation");
    for (int l = 0; l < mgl; ++l) out.Arr[l] = rows_to_csr(arr[l], out.n_if[l]);
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < nsub; ++s)
        for (int l = 0; l < mgl; ++l) {
            out.Aii[l][s] = rows_to_csr(subs[s].Aii[l], subs[s].n_interior[l]);
            out.Air[l][s] = rows_to_csr(subs[s].Air[l], out.n_if[l]);
            out.Aii_coords[l][s].swap(subs[s].interior_coords[l]);
        }
    if (mgl > 0)
        for (int i = 0; i < out.n_if[0]; ++i) {
            bool diag = false;
            for (int k = out.Arr[0].ptr[i]; k < out.Arr[0].ptr[i + 1]; ++k) diag |= out.Arr[0].idx[k] == i;
            if (!diag) throw std::runtime_error("interface operator has a zero diagonal entry");
        }
    phase("operators to CSR");
    // Prr: trilinear prolongation between consecutive interface levels
    out.Prr.assign(std::max(0, mgl - 1), {});
    for (int l = 0; l + 1 < mgl; ++l) {
        std::vector<std::vector<std::pair<int, float>>> rows(out.n_if[l]);
        for (int i = 0; i < out.n_if[l]; ++i) {
            const Int3 f = out.if_coords[l][i];
            const int fc[3] = {f.x, f.y, f.z};
            for (int nb = 0; nb < 8; ++nb) {
                const int c[3] = {f.x / 2 + (nb >> 2), f.y / 2 + ((nb >> 1) & 1), f.z / 2 + (nb & 1)};
                float w = 1.f;
                for (int a = 0; a < 3; ++a) w *= (float)((double)(2.f - (float)std::abs(2 * c[a] - fc[a])) / 2.0);
                if (!(w > 0.f)) continue;
                auto it = if_id[l + 1].find(pack3(c[0], c[1], c[2]));
                if (it == if_id[l + 1].end()) continue;
                rows[i].push_back({it->second, w});
            }
        }
        out.Prr[l] = rows_to_csr(rows, out.n_if[l + 1]);
    }
}

}  // namespace hdd
