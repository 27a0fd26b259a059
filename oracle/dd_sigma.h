// ORACLE — test infrastructure only (see spgrid.h header).
// Sigma (interface Schur complement) multigrid of the reference:
//   SIGMA_MULTIGRID_SOLVER::V_Cycle   IFS/SIGMA_MULTIGRID_SOLVER.cpp:36-47
//   SIGMA_SMOOTHER::Smooth            IFS/SIGMA_SMOOTHER.cpp:292-324
//   SIGMA_SMOOTHER::Compute_Residual  IFS/SIGMA_SMOOTHER.cpp:328-352
//   Eigen_To_Pardiso (+1e-5 diagonal) IFS/SIGMA_SMOOTHER.cpp:50-111
// The reference solves (A + 1e-5 I) x = b with MKL PARDISO (METIS ordering,
// single precision: iparm[27] = 1 for float, IFS/SIGMA_SMOOTHER.cpp:151). MKL is
// absent here; this oracle factors A + 1e-5 I with its own sparse Cholesky
// (nested-dissection ordering from BFS level structures) in double and has two
// precisions for the solves:
//   * double (default): the double factor, double substitution, the solution
//     rounded to T: the exact solve of the shifted system up to T rounding;
//   * single: the factor rounded to float and float substitution, the storage
//     and arithmetic precision of single-precision PARDISO. PARDISO's own
//     ordering and rounding are not reproduced ("parity unpinned" at the
//     rounding level); this mode measures how far single-precision solves move
//     results from the exact ones.
// Eigen SpMV orders follow Eigen's kernels: row-major products accumulate each
// row in column order; transposed (scatter) products accumulate into the
// destination in source-row order.
#pragma once
#include "dd_interface.h"
#include <cmath>
#include <numeric>

namespace ora {

// symmetric graph (pattern of A without diagonal), CSR
struct Graph { int n = 0; std::vector<int> ptr, adj; };

template <class T>
Graph graph_of(const CSR<T>& A) {
    Graph g; g.n = A.rows; g.ptr.assign(g.n + 1, 0);
    std::vector<std::vector<int>> nb(g.n);
    for (int i = 0; i < A.rows; ++i)
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            int j = A.idx[k];
            if (j == i) continue;
            nb[i].push_back(j); nb[j].push_back(i);
        }
    for (int i = 0; i < g.n; ++i) {
        std::sort(nb[i].begin(), nb[i].end());
        nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
        g.ptr[i + 1] = g.ptr[i] + (int)nb[i].size();
        g.adj.insert(g.adj.end(), nb[i].begin(), nb[i].end());
    }
    return g;
}

// Nested dissection on BFS level structures. Produces perm (new -> old).
inline void nd_order(const Graph& g, std::vector<int>& perm) {
    perm.clear();
    std::vector<int> mark(g.n, -1), level(g.n, -1);
    std::vector<int> all(g.n);
    std::iota(all.begin(), all.end(), 0);
    int stamp = 0;
    std::vector<std::vector<int>> stack;
    stack.push_back(all);
    // iterative post-order: each work item is either a set to split or an emit marker
    struct Item { std::vector<int> nodes; bool emit; };
    std::vector<Item> work;
    work.push_back({all, false});
    while (!work.empty()) {
        Item it = std::move(work.back()); work.pop_back();
        if (it.emit) { for (int v : it.nodes) perm.push_back(v); continue; }
        auto& S = it.nodes;
        if (S.size() <= 64) { for (int v : S) perm.push_back(v); continue; }
        ++stamp;
        for (int v : S) mark[v] = stamp;
        // BFS from the first node, then from the farthest (pseudo-peripheral)
        auto bfs = [&](int src, std::vector<int>& order) {
            order.clear();
            for (int v : S) level[v] = -1;
            level[src] = 0; order.push_back(src);
            for (size_t h = 0; h < order.size(); ++h) {
                int v = order[h];
                for (int k = g.ptr[v]; k < g.ptr[v + 1]; ++k) {
                    int w = g.adj[k];
                    if (mark[w] == stamp && level[w] < 0) { level[w] = level[v] + 1; order.push_back(w); }
                }
            }
        };
        std::vector<int> order;
        bfs(S[0], order);
        if (order.size() < S.size()) {
            // disconnected: split off the reached component
            std::vector<int> a(order), b;
            std::sort(a.begin(), a.end());
            for (int v : S) if (level[v] < 0) b.push_back(v);
            work.push_back({b, false});
            work.push_back({a, false});
            continue;
        }
        bfs(order.back(), order);
        int maxl = level[order.back()];
        if (maxl < 2) { for (int v : S) perm.push_back(v); continue; }
        // separator level: where the cumulative count crosses half
        std::vector<int> cnt(maxl + 1, 0);
        for (int v : S) cnt[level[v]]++;
        int acc = 0, sl = 1;
        for (int l = 0; l <= maxl; ++l) { acc += cnt[l]; if (acc * 2 >= (int)S.size()) { sl = std::max(1, std::min(l, maxl - 1)); break; } }
        std::vector<int> A, B, Sep;
        for (int v : S) (level[v] < sl ? A : (level[v] > sl ? B : Sep)).push_back(v);
        work.push_back({Sep, true});
        work.push_back({B, false});
        work.push_back({A, false});
    }
}

// Up-looking sparse Cholesky (column storage) of P (A + eps I) P^T in double.
struct Cholesky {
    int n = 0;
    std::vector<int> perm, iperm;      // perm: new -> old
    std::vector<int> Lp, Li;           // CSC, diagonal first in each column
    std::vector<double> Lx;
    bool single = false;               // factor rounded to float, float substitution

    template <class T>
    void factor(const CSR<T>& A, double eps) {
        n = A.rows;
        Graph g = graph_of(A);
        nd_order(g, perm);
        iperm.assign(n, 0);
        for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
        // C = upper part of permuted matrix, column-wise: column j holds rows i<=j
        std::vector<std::vector<std::pair<int, double>>> Cu(n);
        std::vector<double> diag(n, eps);
        for (int i = 0; i < n; ++i)
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
                int j = A.idx[k];
                int pi = iperm[i], pj = iperm[j];
                if (pi == pj) diag[pi] += (double)A.val[k];
                else if (pi < pj) Cu[pj].push_back({pi, (double)A.val[k]});
            }
        // elimination tree
        std::vector<int> parent(n, -1), anc(n, -1);
        for (int k = 0; k < n; ++k)
            for (auto& e : Cu[k]) {
                int i = e.first;
                while (i != -1 && i < k) { int nx = anc[i]; anc[i] = k; if (nx == -1) parent[i] = k; i = nx; }
            }
        // row patterns via ereach -> column counts
        std::vector<int> flag(n, -1), stackv(n), colcnt(n, 1);
        auto ereach = [&](int k, std::vector<int>& out) {
            out.clear();
            flag[k] = k;
            for (auto& e : Cu[k]) {
                int i = e.first;
                int len = 0;
                for (; flag[i] != k; i = parent[i]) { stackv[len++] = i; flag[i] = k; }
                while (len > 0) out.push_back(stackv[--len]);
            }
        };
        std::vector<int> pat;
        for (int k = 0; k < n; ++k) { ereach(k, pat); for (int j : pat) colcnt[j]++; }
        Lp.assign(n + 1, 0);
        for (int j = 0; j < n; ++j) Lp[j + 1] = Lp[j] + colcnt[j];
        Li.assign(Lp[n], 0); Lx.assign(Lp[n], 0.0);
        std::vector<int> fill(n);
        for (int j = 0; j < n; ++j) fill[j] = Lp[j] + 1;
        std::fill(flag.begin(), flag.end(), -1);
        std::vector<double> x(n, 0.0);
        for (int k = 0; k < n; ++k) {
            ereach(k, pat);
            x[k] = 0;
            for (auto& e : Cu[k]) x[e.first] = e.second;
            double d = diag[k];
            // pat is in topological order only per branch; process by sorted index (valid: L dependencies go low->high)
            std::sort(pat.begin(), pat.end());
            for (int j : pat) {
                double lkj = x[j] / Lx[Lp[j]];
                x[j] = 0;
                for (int p = Lp[j] + 1; p < fill[j]; ++p) x[Li[p]] -= Lx[p] * lkj;
                d -= lkj * lkj;
                Li[fill[j]] = k; Lx[fill[j]] = lkj; fill[j]++;
            }
            if (!(d > 0)) throw std::runtime_error("Cholesky: matrix not positive definite");
            Li[Lp[k]] = k; Lx[Lp[k]] = std::sqrt(d);
        }
        if (single) for (auto& v : Lx) v = (double)(float)v;
    }
    template <class T>
    void solve(const T* b, T* xout) const {
        if (single) { solve_single(b, xout); return; }
        std::vector<double> y(n);
        for (int k = 0; k < n; ++k) y[k] = (double)b[perm[k]];
        for (int j = 0; j < n; ++j) {
            y[j] /= Lx[Lp[j]];
            for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) y[Li[p]] -= Lx[p] * y[j];
        }
        for (int j = n - 1; j >= 0; --j) {
            double s = y[j];
            for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) s -= Lx[p] * y[Li[p]];
            y[j] = s / Lx[Lp[j]];
        }
        for (int k = 0; k < n; ++k) xout[perm[k]] = (T)y[k];
    }
    // single-precision substitution with the float-rounded factor
    template <class T>
    void solve_single(const T* b, T* xout) const {
        std::vector<float> y(n);
        for (int k = 0; k < n; ++k) y[k] = (float)b[perm[k]];
        for (int j = 0; j < n; ++j) {
            y[j] /= (float)Lx[Lp[j]];
            for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) y[Li[p]] -= (float)Lx[p] * y[j];
        }
        for (int j = n - 1; j >= 0; --j) {
            float s = y[j];
            for (int p = Lp[j] + 1; p < Lp[j + 1]; ++p) s -= (float)Lx[p] * y[Li[p]];
            y[j] = s / (float)Lx[Lp[j]];
        }
        for (int k = 0; k < n; ++k) xout[perm[k]] = (T)y[k];
    }
    size_t nnz() const { return Lx.size(); }
};

template <class T>
struct SigmaMG {
    SigmaSetup<T>* S = nullptr;
    int smoothing = 3;
    std::vector<Cholesky> Lrr;                 // per level
    std::vector<std::vector<Cholesky>> Lii;    // [level][sub]
    std::vector<std::vector<T>> u, r, b;

    void initialize(SigmaSetup<T>& s, int sigma_smoothing, bool single_precision = false) {
        S = &s; smoothing = sigma_smoothing;
        const int L = s.mg_levels;
        Lrr.assign(L, {}); Lii.assign(L, std::vector<Cholesky>(s.nsub));
        for (int l = 0; l < L; ++l) {
            Lrr[l].single = single_precision;
            Lrr[l].factor(s.Arr[l], 1e-5);
#pragma omp parallel for schedule(dynamic)
            for (int sd = 0; sd < s.nsub; ++sd) {
                Lii[l][sd].single = single_precision;
                if (s.Aii[l][sd].rows > 0) Lii[l][sd].factor(s.Aii[l][sd], 1e-5);
            }
        }
        u.assign(L, {}); r.assign(L, {}); b.assign(L, {});
        for (int l = 0; l < L; ++l) { u[l].assign(s.n_if[l], T(0)); r[l].assign(s.n_if[l], T(0)); b[l].assign(s.n_if[l], T(0)); }
    }

    // out += sum_s Air_s^T Aii_s^{-1} Air_s u  (Eigen order: per subdomain, scatter by row)
    void schur_interior(int l, const std::vector<T>& uu, std::vector<T>& out) const {
        const int ns = S->nsub;
        std::vector<std::vector<T>> vx(ns);
#pragma omp parallel for schedule(dynamic)
        for (int sd = 0; sd < ns; ++sd) {
            const CSR<T>& Air = S->Air[l][sd];
            if (S->Aii[l][sd].rows == 0) continue;
            std::vector<T> vb(Air.rows);
            for (int i = 0; i < Air.rows; ++i) {
                T acc = 0;
                for (int k = Air.ptr[i]; k < Air.ptr[i + 1]; ++k) acc += Air.val[k] * uu[Air.idx[k]];
                vb[i] = acc;
            }
            vx[sd].assign(Air.rows, T(0));
            Lii[l][sd].solve(vb.data(), vx[sd].data());
        }
        for (int sd = 0; sd < ns; ++sd) {
            if (S->Aii[l][sd].rows == 0) continue;
            const CSR<T>& Air = S->Air[l][sd];
            for (int i = 0; i < Air.rows; ++i)
                for (int k = Air.ptr[i]; k < Air.ptr[i + 1]; ++k) out[Air.idx[k]] += Air.val[k] * vx[sd][i];
        }
    }
    void smooth(int l, std::vector<T>& uu, std::vector<T>& tmp, const std::vector<T>& bb) const {
        tmp = bb;
        schur_interior(l, uu, tmp);
        Lrr[l].solve(tmp.data(), uu.data());
    }
    void compute_residual(int l, const std::vector<T>& uu, std::vector<T>& rr, const std::vector<T>& bb) const {
        rr = bb;
        schur_interior(l, uu, rr);
        const CSR<T>& A = S->Arr[l];
        for (int i = 0; i < A.rows; ++i) {
            T acc = rr[i];
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) acc += A.val[k] * (-uu[A.idx[k]]);
            rr[i] = acc;
        }
    }
    void v_cycle() {
        const int L = S->mg_levels;
        for (int l = 0; l < L - 1; ++l) {
            for (int i = 0; i < smoothing; ++i) smooth(l, u[l], r[l], b[l]);
            compute_residual(l, u[l], r[l], b[l]);
            std::fill(u[l + 1].begin(), u[l + 1].end(), T(0));
            // b[l+1] = Prr^T r  (scatter over fine rows)
            std::fill(b[l + 1].begin(), b[l + 1].end(), T(0));
            const CSR<T>& P = S->Prr[l];
            for (int i = 0; i < P.rows; ++i)
                for (int k = P.ptr[i]; k < P.ptr[i + 1]; ++k) b[l + 1][P.idx[k]] += P.val[k] * r[l][i];
        }
        for (int i = 0; i < 200; ++i) smooth(L - 1, u[L - 1], r[L - 1], b[L - 1]);
        for (int l = L - 2; l >= 0; --l) {
            const CSR<T>& P = S->Prr[l];
            for (int i = 0; i < P.rows; ++i) {
                T acc = 0;
                for (int k = P.ptr[i]; k < P.ptr[i + 1]; ++k) acc += P.val[k] * u[l + 1][P.idx[k]];
                u[l][i] += acc;
            }
            for (int i = 0; i < smoothing; ++i) smooth(l, u[l], r[l], b[l]);
        }
    }
};

}  // namespace ora
