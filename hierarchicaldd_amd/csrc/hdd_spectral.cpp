#include "hdd_spectral.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <omp.h>

namespace hdd {

// The host block solves and inner products of the bottom spectrum split their work
// as for this many threads, whichever threads run it, so that their sums (and the
// modes) are the same on any OpenMP thread count.
constexpr int kSpectralParts = 16;

void sym_eig(int n, std::vector<double>& A, std::vector<double>& w) {
    w.assign(n, 0.0);
    if (n == 0) return;
    auto a = [&](int i, int j) -> double& { return A[(size_t)j * n + i]; };
    std::vector<double> Z((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), q(n), s(n);
    // Householder: H_k = I - 2 v v^T zeroes column k below the subdiagonal; A <- H A H, Z <- Z H
    for (int k = 0; k + 2 < n; ++k) {
        double nrm = 0;
        for (int i = k + 1; i < n; ++i) nrm += a(i, k) * a(i, k);
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = a(k + 1, k) > 0 ? -nrm : nrm;
        double vn = 0;
        for (int i = k + 1; i < n; ++i) { v[i] = a(i, k); }
        v[k + 1] -= alpha;
        for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
        vn = std::sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = k + 1; i < n; ++i) v[i] /= vn;
        // p = A22 v (column sweep), K = v^T p, q = p - K v; A22 -= 2 (v q^T + q v^T)
        for (int i = k + 1; i < n; ++i) p[i] = 0.0;
        for (int j = k + 1; j < n; ++j) {
            const double vj = v[j];
            const double* col = &A[(size_t)j * n];
            for (int i = k + 1; i < n; ++i) p[i] += col[i] * vj;
        }
        double K = 0;
        for (int i = k + 1; i < n; ++i) K += v[i] * p[i];
        for (int i = k + 1; i < n; ++i) q[i] = p[i] - K * v[i];
        for (int j = k + 1; j < n; ++j) {
            double* col = &A[(size_t)j * n];
            const double vj = 2.0 * v[j], qj = 2.0 * q[j];
            for (int i = k + 1; i < n; ++i) col[i] -= v[i] * qj + q[i] * vj;
        }
        a(k + 1, k) = alpha; a(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) { a(i, k) = 0.0; a(k, i) = 0.0; }
        // Z[:, k+1:] -= 2 (Z[:, k+1:] v) v^T
        for (int r = 0; r < n; ++r) s[r] = 0.0;
        for (int j = k + 1; j < n; ++j) {
            const double vj = v[j];
            const double* col = &Z[(size_t)j * n];
            for (int r = 0; r < n; ++r) s[r] += col[r] * vj;
        }
        for (int j = k + 1; j < n; ++j) {
            double* col = &Z[(size_t)j * n];
            const double vj = 2.0 * v[j];
            for (int r = 0; r < n; ++r) col[r] -= s[r] * vj;
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = a(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
    // implicit QL on the tridiagonal (d, e), e[i] couples d[i] and d[i+1]
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m + 1 < n; ++m) {
                // negligible: below rounding of the diagonal pair (a threshold under one ulp, as
                // 1e-16 was, can stall when the spectrum spans many decades, as it does for a
                // nearly singular bottom system of a pure-Neumann domain)
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (m == l) break;
            if (++iter > 100) throw std::runtime_error("sym_eig: QL iteration did not converge");
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double sn = 1.0, c = 1.0, pp = 0.0;
            bool deflated = false;
            for (int i = m - 1; i >= l; --i) {
                const double f = sn * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= pp; e[m] = 0.0; deflated = true; break; }
                sn = f / r; c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * sn + 2.0 * c * b;
                pp = sn * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                double* zi = &Z[(size_t)i * n];
                double* zi1 = &Z[(size_t)(i + 1) * n];
                for (int k = 0; k < n; ++k) {
                    const double fz = zi1[k];
                    zi1[k] = sn * zi[k] + c * fz;
                    zi[k] = c * zi[k] - sn * fz;
                }
            }
            if (deflated) continue;
            d[l] -= pp; e[l] = g; e[m] = 0.0;
        }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return d[x] < d[y]; });
    for (int i = 0; i < n; ++i) {
        w[i] = d[ord[i]];
        std::copy(&Z[(size_t)ord[i] * n], &Z[(size_t)ord[i] * n] + n, &A[(size_t)i * n]);
    }
}

namespace {

// forward elimination of supernode s on y (row-interleaved, width P); updates of
// rows[split..] go to buf[slot[row]] instead of y
template <int P>
inline void forward_sn(const SupernodalFactor& F, int s, double* y, int split, double* buf, const int* slot) {
    double v[P];
    const Supernode& S = F.sn[s];
    const int nS = S.n, nR = (int)S.rows.size();
    const double* L = F.L11[s].data();
    const double* L2 = F.L21[s].data();
    const int* rows = S.rows.data();
    for (int j = 0; j < nS; ++j) {
        double* yj = y + (size_t)(S.s0 + j) * P;
        const double inv = 1.0 / L[(size_t)j * nS + j];
        for (int a = 0; a < P; ++a) { yj[a] *= inv; v[a] = yj[a]; }
        const double* lc = L + (size_t)j * nS;
        for (int i = j + 1; i < nS; ++i) {
            const double l = lc[i];
            double* yi = y + (size_t)(S.s0 + i) * P;
            for (int a = 0; a < P; ++a) yi[a] -= l * v[a];
        }
        const double* l2 = L2 + (size_t)j * nR;
        for (int r = 0; r < split; ++r) {
            const double l = l2[r];
            double* yr = y + (size_t)rows[r] * P;
            for (int a = 0; a < P; ++a) yr[a] -= l * v[a];
        }
        for (int r = split; r < nR; ++r) {
            const double l = l2[r];
            double* yr = buf + (size_t)slot[rows[r]] * P;
            for (int a = 0; a < P; ++a) yr[a] -= l * v[a];
        }
    }
}

template <int P>
inline void backward_sn(const SupernodalFactor& F, int s, double* y) {
    double v[P];
    const Supernode& S = F.sn[s];
    const int nS = S.n, nR = (int)S.rows.size();
    const double* L = F.L11[s].data();
    const double* L2 = F.L21[s].data();
    const int* rows = S.rows.data();
    for (int j = nS - 1; j >= 0; --j) {
        for (int a = 0; a < P; ++a) v[a] = y[(size_t)(S.s0 + j) * P + a];
        const double* l2 = L2 + (size_t)j * nR;
        for (int r = 0; r < nR; ++r) {
            const double l = l2[r];
            const double* yr = y + (size_t)rows[r] * P;
            for (int a = 0; a < P; ++a) v[a] -= l * yr[a];
        }
        const double* lc = L + (size_t)j * nS;
        for (int i = j + 1; i < nS; ++i) {
            const double l = lc[i];
            const double* yi = y + (size_t)(S.s0 + i) * P;
            for (int a = 0; a < P; ++a) v[a] -= l * yi[a];
        }
        const double inv = 1.0 / L[(size_t)j * nS + j];
        for (int a = 0; a < P; ++a) y[(size_t)(S.s0 + j) * P + a] = v[a] * inv;
    }
}

// Large supernodes (the shared top of the tree): the diagonal triangle on one
// thread, the sub-diagonal block split over threads by rows (forward) or by
// columns (backward).
template <int P>
void forward_sn_wide(const SupernodalFactor& F, int s, double* y) {
    const Supernode& S = F.sn[s];
    const int nS = S.n, nR = (int)S.rows.size();
    const double* L = F.L11[s].data();
    const double* L2 = F.L21[s].data();
    const int* rows = S.rows.data();
    for (int j = 0; j < nS; ++j) {
        double* yj = y + (size_t)(S.s0 + j) * P;
        const double inv = 1.0 / L[(size_t)j * nS + j];
        for (int a = 0; a < P; ++a) yj[a] *= inv;
        const double* lc = L + (size_t)j * nS;
        for (int i = j + 1; i < nS; ++i) {
            const double l = lc[i];
            double* yi = y + (size_t)(S.s0 + i) * P;
            for (int a = 0; a < P; ++a) yi[a] -= l * yj[a];
        }
    }
    const double* ys = y + (size_t)S.s0 * P;
#pragma omp parallel for schedule(static)
    for (int r0 = 0; r0 < nR; r0 += 64) {
        const int r1 = std::min(nR, r0 + 64);
        for (int j = 0; j < nS; ++j) {
            const double* vj = ys + (size_t)j * P;
            const double* l2 = L2 + (size_t)j * nR;
            for (int r = r0; r < r1; ++r) {
                const double l = l2[r];
                double* yr = y + (size_t)rows[r] * P;
                for (int a = 0; a < P; ++a) yr[a] -= l * vj[a];
            }
        }
    }
}

template <int P>
void backward_sn_wide(const SupernodalFactor& F, int s, double* y) {
    const Supernode& S = F.sn[s];
    const int nS = S.n, nR = (int)S.rows.size();
    const double* L = F.L11[s].data();
    const double* L2 = F.L21[s].data();
    const int* rows = S.rows.data();
    double* ys = y + (size_t)S.s0 * P;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nS; ++j) {
        double t[P] = {};
        const double* l2 = L2 + (size_t)j * nR;
        for (int r = 0; r < nR; ++r) {
            const double l = l2[r];
            const double* yr = y + (size_t)rows[r] * P;
            for (int a = 0; a < P; ++a) t[a] += l * yr[a];
        }
        for (int a = 0; a < P; ++a) ys[(size_t)j * P + a] -= t[a];
    }
    double v[P];
    for (int j = nS - 1; j >= 0; --j) {
        for (int a = 0; a < P; ++a) v[a] = ys[(size_t)j * P + a];
        const double* lc = L + (size_t)j * nS;
        for (int i = j + 1; i < nS; ++i) {
            const double l = lc[i];
            const double* yi = ys + (size_t)i * P;
            for (int a = 0; a < P; ++a) v[a] -= l * yi[a];
        }
        const double inv = 1.0 / L[(size_t)j * nS + j];
        for (int a = 0; a < P; ++a) ys[(size_t)j * P + a] = v[a] * inv;
    }
}

// y (n x P, row-interleaved: y[k * P + a]) <- L^{-T} L^{-1} y with P a compile-time width
template <int P>
void solve_interleaved(const SupernodalFactor& F, double* y) {
    const int ns = (int)F.sn.size();
    for (int s = 0; s < ns; ++s) forward_sn<P>(F, s, y, (int)F.sn[s].rows.size(), nullptr, nullptr);
    for (int s = ns - 1; s >= 0; --s) backward_sn<P>(F, s, y);
}

template <int P>
void solve_columns(const SupernodalFactor& F, const double* B, double* X) {
    const int n = F.n;
    std::vector<double> y((size_t)n * P);
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < P; ++a) y[(size_t)k * P + a] = B[(size_t)a * n + F.perm[k]];
    solve_interleaved<P>(F, y.data());
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < P; ++a) X[(size_t)a * n + F.perm[k]] = y[(size_t)k * P + a];
}

}  // namespace

void factor_solve_block(const SupernodalFactor& F, int p, const double* B, double* X) {
    const size_t n = F.n;
    for (int a0 = 0; a0 < p;) {
        const int left = p - a0;
        if (left >= 8) { solve_columns<8>(F, B + a0 * n, X + a0 * n); a0 += 8; }
        else if (left >= 4) { solve_columns<4>(F, B + a0 * n, X + a0 * n); a0 += 4; }
        else { solve_columns<1>(F, B + a0 * n, X + a0 * n); a0 += 1; }
    }
}

ParallelSolve8::ParallelSolve8(const SupernodalFactor& F_) : F(F_) {
    const int ns = (int)F.sn.size();
    if (ns == 0) return;
    std::vector<std::vector<int>> kids(ns);
    std::vector<int> cnt(ns, 1);
    std::vector<double> work(ns, 0.0);
    bool postorder = true;
    for (int s = 0; s < ns; ++s) {
        const Supernode& S = F.sn[s];
        work[s] += 0.5 * S.n * (S.n + 1.0) + (double)S.n * S.rows.size();
        if (S.parent >= 0) {
            if (S.parent <= s) postorder = false;
            else { kids[S.parent].push_back(s); cnt[S.parent] += cnt[s]; work[S.parent] += work[s]; }
        }
        if (s > 0 && F.sn[s - 1].s0 + F.sn[s - 1].n != S.s0) postorder = false;
    }
    // subtrees must be contiguous index ranges: children's ranges tile [s - cnt + 1, s - 1]
    for (int s = 0; s < ns && postorder; ++s) {
        int next = s - cnt[s] + 1;
        for (int c : kids[s]) {
            if (c - cnt[c] + 1 != next) { postorder = false; break; }
            next = c + 1;
        }
        if (postorder && next != s) postorder = false;
    }
    auto add_task = [&](int s) {
        task_first.push_back(s - cnt[s] + 1);
        task_last.push_back(s);
        task_end.push_back(F.sn[s].s0 + F.sn[s].n);
    };
    if (!postorder) { add_task(ns - 1); task_first.back() = 0; return; }  // one serial task over everything
    // subtrees for kSpectralParts threads, whatever runs them: the split, and so the order
    // in which the subtrees' buffers are summed, does not depend on the thread count
    const size_t target = (size_t)4 * kSpectralParts;
    double total = 0;
    std::vector<int> heap;
    for (int s = 0; s < ns; ++s) if (F.sn[s].parent < 0) { heap.push_back(s); total += work[s]; }
    auto cmp = [&](int a, int b) { return work[a] < work[b]; };
    std::make_heap(heap.begin(), heap.end(), cmp);
    std::vector<int> done;  // leaves that cannot be split further
    while (!heap.empty() && heap.size() + done.size() < target) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        const int s = heap.back();
        heap.pop_back();
        if (kids[s].empty() || work[s] < total / (2.0 * target)) { done.push_back(s); continue; }
        top.push_back(s);
        for (int c : kids[s]) { heap.push_back(c); std::push_heap(heap.begin(), heap.end(), cmp); }
    }
    std::vector<int> roots(heap.begin(), heap.end());
    roots.insert(roots.end(), done.begin(), done.end());
    std::sort(roots.begin(), roots.end(), [&](int a, int b) { return work[a] > work[b]; });
    for (int s : roots) add_task(s);
    std::sort(top.begin(), top.end());
    top_slot.assign(F.n, -1);
    for (int s : top)
        for (int j = 0; j < F.sn[s].n; ++j) top_slot[F.sn[s].s0 + j] = n_top++;
    split.assign(ns, 0);
    for (int s : top) split[s] = (int)F.sn[s].rows.size();
    for (size_t t = 0; t < task_first.size(); ++t)
        for (int s = task_first[t]; s <= task_last[t]; ++s) {
            const std::vector<int>& r = F.sn[s].rows;
            split[s] = (int)(std::lower_bound(r.begin(), r.end(), task_end[t]) - r.begin());
        }
}

void ParallelSolve8::solve(const double* B, double* X) const {
    constexpr int P = 8;
    const int n = F.n;
    std::vector<double> y((size_t)n * P);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < P; ++a) y[(size_t)k * P + a] = B[(size_t)a * n + F.perm[k]];
    const int nt = (int)task_first.size();
    if (n_top == 0) {
        solve_interleaved<P>(F, y.data());
    } else {
        // one buffer per subtree, summed in subtree order: the result does not
        // depend on which thread ran which subtree
        std::vector<double> bufs((size_t)nt * n_top * P, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < nt; ++t) {
            double* buf = &bufs[(size_t)t * n_top * P];
            for (int s = task_first[t]; s <= task_last[t]; ++s) forward_sn<P>(F, s, y.data(), split[s], buf, top_slot.data());
        }
#pragma omp parallel for schedule(static)
        for (int c = 0; c < n; ++c) {
            const int slot = top_slot[c];
            if (slot < 0) continue;
            for (int t = 0; t < nt; ++t)
                for (int a = 0; a < P; ++a) y[(size_t)c * P + a] += bufs[((size_t)t * n_top + slot) * P + a];
        }
        for (int s : top) forward_sn_wide<P>(F, s, y.data());
        for (auto it = top.rbegin(); it != top.rend(); ++it) backward_sn_wide<P>(F, *it, y.data());
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < nt; ++t)
            for (int s = task_last[t]; s >= task_first[t]; --s) backward_sn<P>(F, s, y.data());
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < P; ++a) X[(size_t)a * n + F.perm[k]] = y[(size_t)k * P + a];
}

void combined_bottom(const SigmaHost& H, int l, Csr& C, std::vector<Int3>& xyz) {
    std::vector<int> off(H.nsub + 1, 0);
    for (int s = 0; s < H.nsub; ++s) off[s + 1] = off[s] + H.Aii[l][s].rows;
    const int ni = off[H.nsub], nf = H.n_if[l], n = ni + nf;
    std::vector<std::vector<std::pair<int, float>>> rows(n);
    for (int s = 0; s < H.nsub; ++s) {
        const Csr& Ai = H.Aii[l][s];
        const Csr& Ar = H.Air[l][s];
        for (int i = 0; i < Ai.rows; ++i) {
            for (int e = Ai.ptr[i]; e < Ai.ptr[i + 1]; ++e) rows[off[s] + i].push_back({off[s] + Ai.idx[e], Ai.val[e]});
            for (int e = Ar.ptr[i]; e < Ar.ptr[i + 1]; ++e) {
                rows[off[s] + i].push_back({ni + Ar.idx[e], Ar.val[e]});
                rows[ni + Ar.idx[e]].push_back({off[s] + i, Ar.val[e]});
            }
        }
    }
    const Csr& R = H.Arr[l];
    for (int i = 0; i < nf; ++i)
        for (int e = R.ptr[i]; e < R.ptr[i + 1]; ++e) rows[ni + i].push_back({ni + R.idx[e], R.val[e]});
    C = Csr(); C.rows = C.cols = n; C.ptr.assign(1, 0);
    for (auto& r : rows) {
        std::sort(r.begin(), r.end(), [](const std::pair<int, float>& a, const std::pair<int, float>& b) { return a.first < b.first; });
        for (auto& e : r) { C.idx.push_back(e.first); C.val.push_back(e.second); }
        C.ptr.push_back((int)C.idx.size());
    }
    xyz.clear();
    for (int s = 0; s < H.nsub; ++s) xyz.insert(xyz.end(), H.Aii_coords[l][s].begin(), H.Aii_coords[l][s].end());
    for (int i = 0; i < nf; ++i) xyz.push_back({H.if_coords[l][i].x << l, H.if_coords[l][i].y << l, H.if_coords[l][i].z << l});
}

namespace {
constexpr int kDenseSpectrum = 600;  // bottoms up to this many dofs are decomposed densely

struct Lcg {
    uint64_t s;
    explicit Lcg(uint64_t seed) : s(seed) {}
    double next() { s = s * 6364136223846793005ULL + 1442695040888963407ULL; return (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5; }
};

// y = (A + eps I) x for p columns of length A.rows
void spmv_shift(const Csr& A, double eps, int p, const double* x, double* y) {
    const int n = A.rows;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < p; ++a) {
            const double* xa = x + (size_t)a * n;
            double acc = eps * xa[i];
            for (int e = A.ptr[i]; e < A.ptr[i + 1]; ++e) acc += (double)A.val[e] * xa[A.idx[e]];
            y[(size_t)a * n + i] = acc;
        }
}

// W = (S - K)^{-1} Y for p <= 8 columns: the interface block of Fcomb^{-1} [0; Y]
void solve_schur(const ParallelSolve8& Fs, const SupernodalFactor& Fc, int n_int, int n, int p, const double* Y, double* W) {
    const size_t N = (size_t)Fc.n;
    std::vector<double> B(N * 8, 0.0), X(N * 8);
    for (int a = 0; a < p; ++a) std::copy(Y + (size_t)a * n, Y + (size_t)(a + 1) * n, &B[a * N + n_int]);
    Fs.solve(B.data(), X.data());
    for (int a = 0; a < p; ++a) std::copy(&X[a * N + n_int], &X[a * N + n_int] + n, W + (size_t)a * n);
}

// C (cols x p) = A[:, :cols]^T W over rows, A and W column-major with leading dimension n
void gemm_tn(int n, int cols, int p, const double* A, const double* W, double* C) {
    // partial sums over kSpectralParts contiguous ranges of 1024-row chunks (the ranges a static
    // schedule gives that many threads), added in range order: the same sums on any thread count
    const int nt = kSpectralParts;
    const int nchunks = (n + 1023) / 1024, q = nchunks / nt, rem = nchunks % nt;
    std::vector<double> part((size_t)nt * cols * p, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nt; ++t) {
        double* P = &part[(size_t)t * cols * p];
        const int k0 = t * q + std::min(t, rem), k1 = k0 + q + (t < rem ? 1 : 0);
        for (int k = k0; k < k1; ++k) {
            const int r0 = k * 1024, r1 = std::min(n, r0 + 1024);
            for (int c = 0; c < cols; ++c) {
                const double* ac = A + (size_t)c * n;
                for (int a = 0; a < p; ++a) {
                    const double* wa = W + (size_t)a * n;
                    double s = 0;
                    for (int r = r0; r < r1; ++r) s += ac[r] * wa[r];
                    P[(size_t)c * p + a] += s;
                }
            }
        }
    }
    for (size_t i = 0; i < (size_t)cols * p; ++i) {
        double s = 0;
        for (int t = 0; t < nt; ++t) s += part[(size_t)t * cols * p + i];
        C[i] = s;
    }
}

// W (n x p) -= A[:, :cols] C (cols x p)
void gemm_nn_sub(int n, int cols, int p, const double* A, const double* C, double* W) {
#pragma omp parallel for schedule(static)
    for (int r0 = 0; r0 < n; r0 += 1024) {
        const int r1 = std::min(n, r0 + 1024);
        for (int c = 0; c < cols; ++c) {
            const double* ac = A + (size_t)c * n;
            for (int a = 0; a < p; ++a) {
                const double f = C[(size_t)c * p + a];
                double* wa = W + (size_t)a * n;
                for (int r = r0; r < r1; ++r) wa[r] -= ac[r] * f;
            }
        }
    }
}

// X (n x k) = A[:, :cols] Y[:, sel] with Y cols x cols column-major
void gemm_select(int n, int cols, const double* A, const std::vector<double>& Y, const std::vector<int>& sel, std::vector<double>& X) {
    const int k = (int)sel.size();
    X.assign((size_t)n * k, 0.0);
#pragma omp parallel for schedule(static)
    for (int r0 = 0; r0 < n; r0 += 1024) {
        const int r1 = std::min(n, r0 + 1024);
        for (int c = 0; c < cols; ++c) {
            const double* ac = A + (size_t)c * n;
            for (int j = 0; j < k; ++j) {
                const double f = Y[(size_t)sel[j] * cols + c];
                double* xj = &X[(size_t)j * n];
                for (int r = r0; r < r1; ++r) xj[r] += ac[r] * f;
            }
        }
    }
}

double dot(int n, const double* a, const double* b) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// Small systems: T~ = L^T (S - K)^{-1} L with S = L L^T is symmetric with the
// eigenvalues of T; v = L^{-T} u are the S-orthonormal eigenvectors.
void dense_spectrum(const Csr& Arr, double eps, const SupernodalFactor& Fc, int n_int, double lambda_cut, BottomSpectrum& out) {
    const int n = Arr.rows;
    std::vector<double> L((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        L[(size_t)i * n + i] += eps;
        for (int e = Arr.ptr[i]; e < Arr.ptr[i + 1]; ++e) L[(size_t)Arr.idx[e] * n + i] += Arr.val[e];
    }
    std::vector<double> S = L;
    for (int j = 0; j < n; ++j) {  // lower Cholesky in place (column-major)
        double* cj = &L[(size_t)j * n];
        for (int k = 0; k < j; ++k) {
            const double ljk = L[(size_t)k * n + j];
            const double* ck = &L[(size_t)k * n];
            for (int i = j; i < n; ++i) cj[i] -= ck[i] * ljk;
        }
        if (cj[j] <= 0.0) throw std::runtime_error("bottom spectrum: Arr + eps I is not positive definite");
        const double d = std::sqrt(cj[j]);
        for (int i = j; i < n; ++i) cj[i] /= d;
        for (int i = 0; i < j; ++i) cj[i] = 0.0;
    }
    std::vector<double> G((size_t)n * n);
    const ParallelSolve8 Fs(Fc);
    for (int c0 = 0; c0 < n; c0 += 8) {
        const int p = std::min(8, n - c0);
        solve_schur(Fs, Fc, n_int, n, p, &L[(size_t)c0 * n], &G[(size_t)c0 * n]);
    }
    std::vector<double> T((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) T[(size_t)j * n + i] = dot(n, &L[(size_t)i * n], &G[(size_t)j * n]);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) {
            const double m = 0.5 * (T[(size_t)j * n + i] + T[(size_t)i * n + j]);
            T[(size_t)j * n + i] = T[(size_t)i * n + j] = m;
        }
    std::vector<double> w;
    sym_eig(n, T, w);
    out.lambda_max = 1.0 - 1.0 / w[n - 1];
    const double mu_cut = 1.0 / (1.0 - lambda_cut);
    for (int i = n - 1; i >= 0 && w[i] >= mu_cut; --i) {
        std::vector<double> v(&T[(size_t)i * n], &T[(size_t)i * n] + n);
        for (int r = n - 1; r >= 0; --r) {  // L^T v = u
            double s = v[r];
            for (int k = r + 1; k < n; ++k) s -= L[(size_t)r * n + k] * v[k];
            v[r] = s / L[(size_t)r * n + r];
        }
        out.lambda.push_back(1.0 - 1.0 / w[i]);
        out.V.insert(out.V.end(), v.begin(), v.end());
    }
    const int k = (int)out.lambda.size();
    out.SV.assign((size_t)n * k, 0.0);
    for (int j = 0; j < k; ++j)
        for (int c = 0; c < n; ++c) {
            const double vc = out.V[(size_t)j * n + c];
            if (vc == 0.0) continue;
            const double* sc = &S[(size_t)c * n];
            for (int r = 0; r < n; ++r) out.SV[(size_t)j * n + r] += sc[r] * vc;
        }
    out.resolved = true;
}

// S-orthonormalise the p columns of W (SW = S W) against each other, two
// Gram-Schmidt passes; R (p x p, column-major, upper) gets the coefficients.
// A column that vanishes is replaced by a random direction orthogonal to the
// basis (R's diagonal entry 0).
void orth_in_block(const Csr& Arr, double eps, int n, int p, double* W, double* SW, double* R, const double* Q, const double* SQ, int cols, Lcg& rng) {
    std::fill(R, R + (size_t)p * p, 0.0);
    for (int a = 0; a < p; ++a) {
        double* wa = W + (size_t)a * n;
        double* swa = SW + (size_t)a * n;
        const double before = std::sqrt(std::max(0.0, dot(n, wa, swa)));
        for (int pass = 0; pass < 2; ++pass)
            for (int b = 0; b < a; ++b) {
                const double c = dot(n, SW + (size_t)b * n, wa);
                const double* wb = W + (size_t)b * n;
                const double* swb = SW + (size_t)b * n;
                for (int i = 0; i < n; ++i) { wa[i] -= c * wb[i]; swa[i] -= c * swb[i]; }
                R[(size_t)a * p + b] += c;
            }
        double nrm = std::sqrt(std::max(0.0, dot(n, wa, swa)));
        if (!(nrm > 1e-10 * before) || nrm == 0.0) {
            for (int i = 0; i < n; ++i) wa[i] = rng.next();
            for (int pass = 0; pass < 2; ++pass) {
                spmv_shift(Arr, eps, 1, wa, swa);
                for (int c = 0; c < cols; ++c) {
                    const double f = dot(n, SQ + (size_t)c * n, wa);
                    const double* qc = Q + (size_t)c * n;
                    for (int i = 0; i < n; ++i) wa[i] -= f * qc[i];
                }
                spmv_shift(Arr, eps, 1, wa, swa);
                for (int b = 0; b < a; ++b) {
                    const double f = dot(n, SW + (size_t)b * n, wa);
                    const double* wb = W + (size_t)b * n;
                    const double* swb = SW + (size_t)b * n;
                    for (int i = 0; i < n; ++i) { wa[i] -= f * wb[i]; swa[i] -= f * swb[i]; }
                }
            }
            nrm = std::sqrt(std::max(0.0, dot(n, wa, swa)));
            if (!(nrm > 0.0)) throw std::runtime_error("bottom spectrum: block Krylov breakdown");
            R[(size_t)a * p + a] = 0.0;
        } else {
            R[(size_t)a * p + a] = nrm;
        }
        for (int i = 0; i < n; ++i) { wa[i] /= nrm; swa[i] /= nrm; }
    }
}

void lanczos_spectrum(const Csr& Arr, double eps, const SupernodalFactor& Fc, int n_int, double lambda_cut, size_t max_bytes, BottomSpectrum& out) {
    const int n = Arr.rows, p = 8;
    const int kmax = 256;                 // more slow modes than this: keep the sweeps
    const double tol = 1e-12;             // tightest Ritz residual asked for, relative to the Ritz value
    const double mu_cut = 1.0 / (1.0 - lambda_cut);
    int maxv = (int)std::min<size_t>(max_bytes / (16 * (size_t)n), 1024);
    maxv -= maxv % p;
    maxv = std::min(maxv, (n / p - 1) * p);
    if (maxv < 6 * p) { out.resolved = false; return; }
    const int ld = maxv + p;
    const ParallelSolve8 Fs(Fc);
    std::vector<double> Q((size_t)n * ld), SQ((size_t)n * ld), H((size_t)ld * ld, 0.0);
    std::vector<double> W((size_t)n * p), SW((size_t)n * p), C((size_t)ld * p), R((size_t)p * p);
    Lcg rng(0x5eed5eed1234ULL);
    for (auto& x : W) x = rng.next();
    spmv_shift(Arr, eps, p, W.data(), SW.data());
    orth_in_block(Arr, eps, n, p, W.data(), SW.data(), R.data(), Q.data(), SQ.data(), 0, rng);
    std::copy(W.begin(), W.end(), Q.begin());
    std::copy(SW.begin(), SW.end(), SQ.begin());
    int nb = 1, prev_above = -1;
    std::vector<double> Y, theta, res;
    int next_check = 4;
    // time spent in the block solves, the orthogonalisation and the Rayleigh-Ritz steps
    // (on stderr with HDD_FACTOR_VERBOSE)
    double t_solve = 0, t_orth = 0, t_ritz = 0;
    auto tick = std::chrono::steady_clock::now();
    auto lap = [&tick](double& acc) {
        const auto now = std::chrono::steady_clock::now();
        acc += std::chrono::duration<double>(now - tick).count();
        tick = now;
    };
    struct Report {
        const double &s, &o, &r;
        ~Report() {
            if (std::getenv("HDD_FACTOR_VERBOSE")) std::fprintf(stderr, "[setup]   bottom spectrum: block solves %.3fs, orthogonalisation %.3fs, Rayleigh-Ritz %.3fs\n", s, o, r);
        }
    } report{t_solve, t_orth, t_ritz};
    while (nb * p + p <= ld) {
        const int j = nb - 1, cols = nb * p;
        lap(t_orth);
        solve_schur(Fs, Fc, n_int, n, p, &SQ[(size_t)j * p * n], W.data());  // W = T Q_j
        lap(t_solve);
        for (int pass = 0; pass < 2; ++pass) {
            gemm_tn(n, cols, p, SQ.data(), W.data(), C.data());
            gemm_nn_sub(n, cols, p, Q.data(), C.data(), W.data());
            for (int c = 0; c < cols; ++c)
                for (int a = 0; a < p; ++a) H[(size_t)(j * p + a) * ld + c] += C[(size_t)c * p + a];
        }
        spmv_shift(Arr, eps, p, W.data(), SW.data());
        orth_in_block(Arr, eps, n, p, W.data(), SW.data(), R.data(), Q.data(), SQ.data(), cols, rng);
        for (int a = 0; a < p; ++a)
            for (int b = 0; b < p; ++b) H[(size_t)(j * p + a) * ld + cols + b] = R[(size_t)a * p + b];
        std::copy(W.begin(), W.end(), Q.begin() + (size_t)cols * n);
        std::copy(SW.begin(), SW.end(), SQ.begin() + (size_t)cols * n);
        ++nb;
        const int m = nb - 1;  // blocks in the Rayleigh-Ritz space; block m is the residual direction
        const bool last = nb * p + p > ld;
        if (m < next_check && !last) continue;
        lap(t_orth);
        next_check = m + std::max(2, m / 4);
        const int mc = m * p;
        Y.assign((size_t)mc * mc, 0.0);
        for (int c = 0; c < mc; ++c)
            for (int r = 0; r < mc; ++r) Y[(size_t)c * mc + r] = 0.5 * (H[(size_t)c * ld + r] + H[(size_t)r * ld + c]);
        sym_eig(mc, Y, theta);
        res.assign(mc, 0.0);
        for (int i = 0; i < mc; ++i) {
            double s2 = 0;
            for (int b = 0; b < p; ++b) {
                double s = 0;
                for (int a = 0; a < p; ++a) s += H[(size_t)((m - 1) * p + a) * ld + mc + b] * Y[(size_t)i * mc + (m - 1) * p + a];
                s2 += s * s;
            }
            res[i] = std::sqrt(s2);
        }
        // a kept pair with eigenvalue l and vector error d moves the result by about
        // l^200 d |x*|; its residual bound scales with 1 / l^200 (1e-3 gap allowance)
        int above = 0;
        bool all_conv = res[mc - 1] <= 1e-9 * theta[mc - 1];  // the largest eigenvalue is reported
        for (int i = mc - 1; i >= 0 && theta[i] >= mu_cut; --i) {
            ++above;
            const double lam200 = std::pow(1.0 - 1.0 / theta[i], 200.0);
            if (res[i] > std::max(tol, 1e-10 / lam200) * theta[i]) all_conv = false;
        }
        out.lambda_max = 1.0 - 1.0 / theta[mc - 1];
        out.steps = m;
        if (above > kmax) { out.resolved = false; return; }
        if (all_conv && above == prev_above && m >= 4) {
            std::vector<int> sel;
            for (int i = mc - 1; i >= 0 && theta[i] >= mu_cut; --i) {
                sel.push_back(i);
                out.lambda.push_back(1.0 - 1.0 / theta[i]);
                out.max_residual = std::max(out.max_residual, res[i] / theta[i]);
            }
            gemm_select(n, mc, Q.data(), Y, sel, out.V);
            gemm_select(n, mc, SQ.data(), Y, sel, out.SV);
            out.resolved = true;
            return;
        }
        prev_above = all_conv ? above : -1;
        lap(t_ritz);
        if (last) break;
    }
    out.resolved = false;
}

}  // namespace

void bottom_spectrum(const Csr& Arr, double eps, const SupernodalFactor& Fcomb, int n_int, double lambda_cut, size_t max_basis_bytes,
                     BottomSpectrum& out) {
    out = BottomSpectrum();
    const auto t0 = std::chrono::steady_clock::now();
    if (Arr.rows > 0) {
        if (Arr.rows <= kDenseSpectrum) dense_spectrum(Arr, eps, Fcomb, n_int, lambda_cut, out);
        else lanczos_spectrum(Arr, eps, Fcomb, n_int, lambda_cut, max_basis_bytes, out);
    } else {
        out.resolved = true;
    }
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace hdd
