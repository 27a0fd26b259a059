#include "hdd_factor.h"
#include "hdd_parallel.h"
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <numeric>
#include <omp.h>
#include <stdexcept>

namespace hdd {

size_t SupernodalFactor::nnz() const {
    size_t s = 0;
    for (const auto& S : sn) s += (size_t)S.n * (S.n + 1) / 2 + (size_t)S.n * S.rows.size();
    return s;
}

namespace {

struct NDBuilder {
    const Csr& A;
    const std::vector<Int3>& xyz;
    int leaf;
    std::vector<int> owner;  // current subtree id of each dof
    int next_owner = 1;
    SupernodalFactor& F;
    int next_pos = 0;

    NDBuilder(const Csr& a, const std::vector<Int3>& c, int lf, SupernodalFactor& f)
        : A(a), xyz(c), leaf(lf), owner(a.rows, 0), F(f) {}

    static int coord(const Int3& p, int axis) { return axis == 0 ? p.x : (axis == 1 ? p.y : p.z); }

    int make_supernode(std::vector<int>& nodes, const std::vector<int>& children) {
        std::sort(nodes.begin(), nodes.end());
        Supernode S;
        S.s0 = next_pos;
        S.n = (int)nodes.size();
        for (int v : nodes) { F.perm[next_pos] = v; F.iperm[v] = next_pos; ++next_pos; }
        const int id = (int)F.sn.size();
        int h = 0;
        for (int c : children) { F.sn[c].parent = id; h = std::max(h, F.sn[c].height + 1); }
        S.height = h;
        F.sn.push_back(std::move(S));
        return id;
    }

    // returns the roots of the subtree built for X (X is consumed)
    std::vector<int> build(std::vector<int>& X) {
        if ((int)X.size() <= leaf) return {make_supernode(X, {})};
        const int me = next_owner++;
        for (int v : X) owner[v] = me;
        // try axes by decreasing extent; the separator for a cut c is the set of dofs
        // v with coord(v) >= c that have a neighbour w with coord(w) < c
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = 1 << 30; hi[a] = -(1 << 30); }
        for (int v : X) for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], coord(xyz[v], a)); hi[a] = std::max(hi[a], coord(xyz[v], a)); }
        int order[3] = {0, 1, 2};
        std::sort(order, order + 3, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
        for (int oi = 0; oi < 3; ++oi) {
            const int axis = order[oi];
            const int span = hi[axis] - lo[axis];
            if (span <= 0) break;
            std::vector<int> mn(X.size());
            std::vector<long> sep(span + 2, 0), cnt(span + 2, 0);
            for (size_t k = 0; k < X.size(); ++k) {
                const int v = X[k];
                const int cv = coord(xyz[v], axis);
                int m = cv;
                for (int e = A.ptr[v]; e < A.ptr[v + 1]; ++e) {
                    const int w = A.idx[e];
                    if (owner[w] == me) m = std::min(m, coord(xyz[w], axis));
                }
                mn[k] = m;
                cnt[cv - lo[axis]]++;
                if (m < cv) { sep[m - lo[axis] + 1]++; sep[cv - lo[axis] + 1]--; }
            }
            // cut c = lo + t, t in [1, span]; left = coord < c
            long run_sep = 0, left = 0, best = -1;
            int best_t = -1;
            const long N = (long)X.size();
            for (int t = 1; t <= span; ++t) {
                run_sep += sep[t];
                left += cnt[t - 1];
                if (left * 4 < N || left * 4 > 3 * N) continue;
                const long score = run_sep * 4 + std::abs(2 * left - N) / 8;  // separator size, mild balance
                if (best < 0 || score < best) { best = score; best_t = t; }
            }
            if (best_t < 0) continue;
            const int c = lo[axis] + best_t;
            std::vector<int> L, R, Sp;
            for (size_t k = 0; k < X.size(); ++k) {
                const int v = X[k], cv = coord(xyz[v], axis);
                if (cv < c) L.push_back(v);
                else if (mn[k] < c) Sp.push_back(v);
                else R.push_back(v);
            }
            if (L.empty() || (R.empty() && Sp.size() == X.size())) continue;
            std::vector<int>().swap(X);
            std::vector<int> roots = build(L);
            std::vector<int> rr = build(R);
            roots.insert(roots.end(), rr.begin(), rr.end());
            if (Sp.empty()) return roots;
            return {make_supernode(Sp, roots)};
        }
        return {make_supernode(X, {})};  // no useful geometric cut: one dense block
    }
};

// partial Cholesky of the first ns columns of the m x m lower front F (column-major)
void partial_cholesky(double* F, int m, int ns, bool parallel) {
    const int B = 48;
    for (int k0 = 0; k0 < ns; k0 += B) {
        const int kb = std::min(B, ns - k0), k1 = k0 + kb;
        for (int k = k0; k < k1; ++k) {
            double* ck = F + (size_t)k * m;
            for (int j = k0; j < k; ++j) {
                const double* cj = F + (size_t)j * m;
                const double a = cj[k];
                if (a != 0.0) for (int i = k; i < m; ++i) ck[i] -= cj[i] * a;
            }
            const double d = ck[k];
            if (!(d > 0.0)) throw std::runtime_error("supernodal Cholesky: matrix is not positive definite");
            const double s = std::sqrt(d);
            ck[k] = s;
            const double inv = 1.0 / s;
            for (int i = k + 1; i < m; ++i) ck[i] *= inv;
        }
        // trailing update of columns k1..m-1 (rows >= column)
        auto upd = [&](int j) {
            double* cj = F + (size_t)j * m;
            for (int k = k0; k < k1; ++k) {
                const double* ck = F + (size_t)k * m;
                const double a = ck[j];
                if (a == 0.0) continue;
                for (int i = j; i < m; ++i) cj[i] -= ck[i] * a;
            }
        };
        if (parallel && (m - k1) > 64) {
#pragma omp parallel for schedule(dynamic, 8)
            for (int j = k1; j < m; ++j) upd(j);
        } else {
            for (int j = k1; j < m; ++j) upd(j);
        }
    }
}

}  // namespace

void invert_lower(std::vector<double>& L, int n) {
    // column c of the inverse by forward substitution, each sum in ascending k; the rows
    // of L are read from a row-major copy so the inner loop is contiguous
    const bool par = n > 128 && !omp_in_parallel();
    std::vector<double> R((size_t)n * n);
#pragma omp parallel for schedule(static) if (par)
    for (int i = 0; i < n; ++i)
        for (int k = 0; k <= i; ++k) R[(size_t)i * n + k] = L[(size_t)k * n + i];
    std::vector<double> X((size_t)n * n, 0.0);
    auto col = [&](int c) {
        double* x = &X[(size_t)c * n];
        x[c] = 1.0 / R[(size_t)c * n + c];
        for (int i = c + 1; i < n; ++i) {
            const double* Ri = &R[(size_t)i * n];
            double s = 0.0;
            for (int k = c; k < i; ++k) s += Ri[k] * x[k];
            x[i] = -s / Ri[i];
        }
    };
#pragma omp parallel for schedule(dynamic, 4) if (par)
    for (int c = 0; c < n; ++c) col(c);
    L.swap(X);
}

void factor_spd(const Csr& A, const std::vector<Int3>& coords, double eps, SupernodalFactor& F, int leaf) {
    const int n = A.rows;
    F = SupernodalFactor();
    F.n = n;
    F.perm.assign(n, -1);
    F.iperm.assign(n, -1);
    if (n == 0) return;
    if ((int)coords.size() != n) throw std::runtime_error("factor_spd: coordinate count mismatch");
    {
        NDBuilder nd(A, coords, leaf, F);
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        nd.build(all);
        if (nd.next_pos != n) throw std::runtime_error("nested dissection lost dofs");
    }
    const int ns = (int)F.sn.size();
    // symbolic: rows below each diagonal block (children first: postorder)
    std::vector<int> sn_of(n);
    for (int s = 0; s < ns; ++s) for (int k = 0; k < F.sn[s].n; ++k) sn_of[F.sn[s].s0 + k] = s;
    std::vector<std::vector<int>> kids(ns);
    for (int s = 0; s < ns; ++s) if (F.sn[s].parent >= 0) kids[F.sn[s].parent].push_back(s);
    {
        std::vector<int> mark(n, -1);
        for (int s = 0; s < ns; ++s) {
            Supernode& S = F.sn[s];
            const int end = S.s0 + S.n;
            std::vector<int> rows;
            for (int j = S.s0; j < end; ++j) {
                const int oj = F.perm[j];
                for (int e = A.ptr[oj]; e < A.ptr[oj + 1]; ++e) {
                    const int i = F.iperm[A.idx[e]];
                    if (i >= end && mark[i] != s) { mark[i] = s; rows.push_back(i); }
                }
            }
            for (int c : kids[s])
                for (int i : F.sn[c].rows)
                    if (i >= end && mark[i] != s) { mark[i] = s; rows.push_back(i); }
            std::sort(rows.begin(), rows.end());
            S.rows.swap(rows);
            F.max_height = std::max(F.max_height, S.height);
        }
    }
    // numeric: multifrontal, one height level at a time
    F.L11.assign(ns, {});
    F.L21.assign(ns, {});
    std::vector<std::vector<double>> U(ns);
    std::vector<std::vector<int>> by_h(F.max_height + 1);
    for (int s = 0; s < ns; ++s) by_h[F.sn[s].height].push_back(s);
    const int nthreads = omp_get_max_threads();
    const bool use_device = device_fronts_enabled();
    const int front_min = device_front_min_rows();
    // which fronts go to the GPU (whose rounding differs from the host's) must not depend on
    // the thread count: big fronts of heights with fewer than kDeviceHeightFronts fronts
    constexpr int kDeviceHeightFronts = 16;
    auto front = [&](int s, bool inner_parallel, bool device_ok, std::vector<int>& pos) {
        const Supernode& S = F.sn[s];
        const int nS = S.n, nR = (int)S.rows.size(), m = nS + nR;
        std::vector<double> Fm((size_t)m * m, 0.0);
        for (int k = 0; k < nS; ++k) pos[S.s0 + k] = k;
        for (int r = 0; r < nR; ++r) pos[S.rows[r]] = nS + r;
        for (int j = S.s0; j < S.s0 + nS; ++j) {
            const int oj = F.perm[j], jl = j - S.s0;
            for (int e = A.ptr[oj]; e < A.ptr[oj + 1]; ++e) {
                const int i = F.iperm[A.idx[e]];
                if (i < j) continue;
                Fm[(size_t)jl * m + pos[i]] += (double)A.val[e];
            }
            Fm[(size_t)jl * m + jl] += eps;
        }
        for (int c : kids[s]) {
            const Supernode& C = F.sn[c];
            const int nc = (int)C.rows.size();
            const std::vector<double>& Uc = U[c];
            for (int b = 0; b < nc; ++b) {
                const int pb = pos[C.rows[b]];
                for (int a = b; a < nc; ++a) Fm[(size_t)pb * m + pos[C.rows[a]]] += Uc[(size_t)b * nc + a];
            }
            std::vector<double>().swap(U[c]);
        }
        if (device_ok && m >= front_min && use_device) partial_cholesky_device(Fm.data(), m, nS);
        else partial_cholesky(Fm.data(), m, nS, inner_parallel);
        std::vector<double> l11((size_t)nS * nS), l21((size_t)nR * nS);
        for (int j = 0; j < nS; ++j) {
            for (int i = j; i < nS; ++i) l11[(size_t)j * nS + i] = Fm[(size_t)j * m + i];
            for (int r = 0; r < nR; ++r) l21[(size_t)j * nR + r] = Fm[(size_t)j * m + nS + r];
        }
        std::vector<double> u((size_t)nR * nR);
        for (int b = 0; b < nR; ++b)
            for (int a = b; a < nR; ++a) u[(size_t)b * nR + a] = Fm[(size_t)(nS + b) * m + nS + a];
        F.L11[s].swap(l11);
        F.L21[s].swap(l21);
        U[s].swap(u);
    };
    const bool verbose = std::getenv("HDD_FACTOR_VERBOSE") != nullptr;
    for (int h = 0; h <= F.max_height; ++h) {
        const auto& list = by_h[h];
        auto th0 = std::chrono::steady_clock::now();
        const bool device_ok = (int)list.size() < kDeviceHeightFronts;
        if ((int)list.size() >= nthreads) {
            ParallelErrors errs;
#pragma omp parallel
            {
                std::vector<int> pos(n);
#pragma omp for schedule(dynamic)
                for (size_t k = 0; k < list.size(); ++k) errs.run([&] { front(list[k], false, device_ok, pos); });
            }
            errs.rethrow();
        } else {
            std::vector<int> pos(n);
            for (int s : list) front(s, true, device_ok, pos);
        }
        if (verbose && n > 20000) {
            size_t mmax = 0; double flops = 0;
            for (int s : list) { size_t m = F.sn[s].n + F.sn[s].rows.size(); mmax = std::max(mmax, m); flops += (double)F.sn[s].n * m * m; }
            std::fprintf(stderr, "[factor n=%d] height %d: %zu fronts, max front %zu, ~%.2e flop, %.3fs\n", n, h, list.size(), mmax, flops,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - th0).count());
        }
    }
}

void factor_solve(const SupernodalFactor& F, const double* b, double* x) {
    const int n = F.n;
    std::vector<double> y(n);
    for (int k = 0; k < n; ++k) y[k] = b[F.perm[k]];
    for (size_t s = 0; s < F.sn.size(); ++s) {
        const Supernode& S = F.sn[s];
        const int nS = S.n, nR = (int)S.rows.size();
        const double* L = F.L11[s].data();
        for (int j = 0; j < nS; ++j) {
            double v = y[S.s0 + j] / L[(size_t)j * nS + j];
            y[S.s0 + j] = v;
            for (int i = j + 1; i < nS; ++i) y[S.s0 + i] -= L[(size_t)j * nS + i] * v;
            for (int r = 0; r < nR; ++r) y[S.rows[r]] -= F.L21[s][(size_t)j * nR + r] * v;
        }
    }
    for (int s = (int)F.sn.size() - 1; s >= 0; --s) {
        const Supernode& S = F.sn[s];
        const int nS = S.n, nR = (int)S.rows.size();
        const double* L = F.L11[s].data();
        for (int j = nS - 1; j >= 0; --j) {
            double v = y[S.s0 + j];
            for (int r = 0; r < nR; ++r) v -= F.L21[s][(size_t)j * nR + r] * y[S.rows[r]];
            for (int i = j + 1; i < nS; ++i) v -= L[(size_t)j * nS + i] * y[S.s0 + i];
            y[S.s0 + j] = v / L[(size_t)j * nS + j];
        }
    }
    for (int k = 0; k < n; ++k) x[F.perm[k]] = y[k];
}

}  // namespace hdd
