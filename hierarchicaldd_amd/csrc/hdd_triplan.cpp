#include "hdd_triplan.h"
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <omp.h>

namespace hdd {

namespace {
// lanes per row for a mean of `len` entries per row: 4 .. 64
int lanes_for(double len) {
    int g = 4;
    while (g < 64 && g * 3 < len) g *= 2;
    return g;
}

// A height of a plan costs about 60 us in launches (forward and backward, about
// 100 us measured on the big ones); a dense top of m rows reads m^2 floats per
// solve (both triangles of its block) at about 4.5 TB/s, so a height joins the
// tops while the blocks of all the plan's members grow by at most
// 60e-6 * 4.5e12 / 4 B.
constexpr long long kTopStepCells = 67500000;
constexpr int kTopDefaultMinRows = 4096;
constexpr int kTopMaxMembers = 64;  // plans of many factors share each height's launches: no dense tops
}  // namespace

int dense_top_min_rows() {
    const char* e = std::getenv("HDD_DENSE_TOP");
    if (!e || !*e) return kTopDefaultMinRows;
    return std::max(0, std::atoi(e));
}

int dense_top_cut(const SupernodalFactor& f, int top_min_rows, int members, int* m_out) {
    *m_out = 0;
    if (top_min_rows <= 0 || f.n < top_min_rows || f.sn.empty() || members < 1 || members > kTopMaxMembers) return INT_MAX;
    std::vector<long long> rows(f.max_height + 1, 0);
    for (const Supernode& S : f.sn) rows[S.height] += S.n;
    long long m = 0;
    int cut = INT_MAX;
    for (int h = f.max_height; h >= 0; --h) {
        const long long m2 = m + rows[h];
        if (m2 > TriPlanHost::kTopMaxRows || 2 * m2 > f.n) break;
        if ((m2 * m2 - m * m) * members > kTopStepCells) break;
        m = m2;
        cut = h;
    }
    *m_out = (int)m;
    return cut;
}

void dense_top_factor(const TriPlanHost::TopBlock& B, double* L) {
    const SupernodalFactor& f = *B.f;
    const int m = B.m;
    // S-local index of each permuted row of S: rows of S ascending
    std::vector<int> loc(f.n, -1);
    {
        std::vector<char> in(f.n, 0);
        for (const Supernode& S : f.sn)
            if (S.height >= B.cut) for (int k = 0; k < S.n; ++k) in[S.s0 + k] = 1;
        int q = 0;
        for (int i = 0; i < f.n; ++i) if (in[i]) loc[i] = q++;
        if (q != m) throw std::runtime_error("dense top: row count mismatch");
    }
    bool outside = false;  // a sub-diagonal row of a top supernode below the top (the cut is not ancestor-closed)
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (int j = 0; j < m; ++j) std::fill(L + (size_t)j * m, L + (size_t)(j + 1) * m, 0.0);
        // each column belongs to one supernode: the supernodes fill disjoint columns
#pragma omp for schedule(dynamic)
        for (size_t s = 0; s < f.sn.size(); ++s) {
            const Supernode& S = f.sn[s];
            if (S.height < B.cut) continue;
            const int n = S.n, nr = (int)S.rows.size();
            for (int j = 0; j < n; ++j) {
                double* col = L + (size_t)loc[S.s0 + j] * m;
                for (int r = j; r < n; ++r) col[loc[S.s0 + r]] = f.L11[s][(size_t)j * n + r];
                for (int a = 0; a < nr; ++a) {
                    const int li = loc[S.rows[a]];
                    if (li < 0) {
#pragma omp atomic write
                        outside = true;
                        continue;
                    }
                    col[li] = f.L21[s][(size_t)j * nr + a];
                }
            }
        }
    }
    if (outside) throw std::runtime_error("dense top: a row below the top is not in it");
}

void build_triplan(const std::vector<const SupernodalFactor*>& F, const std::vector<int>& offsets, TriPlanHost& P, int top_min_rows) {
    P = TriPlanHost();
    for (size_t m = 0; m < F.size(); ++m) {
        P.n = std::max(P.n, offsets[m] + F[m]->n);
        if (F[m]->n) P.H = std::max(P.H, F[m]->max_height);
        P.factor_nnz += F[m]->nnz();
    }
    if (P.n == 0) return;
    P.orig_of.assign(P.n, 0);
    P.sn_of.assign(P.n, 0);
    std::vector<int> height;
    struct Ref { int m, s; };
    std::vector<Ref> refs;
    std::vector<int> owner;  // per supernode: 1 if it holds the data (first member of its factor)
    std::map<const SupernodalFactor*, int> first_gid;  // factor -> global id of its first supernode, first member
    std::map<const SupernodalFactor*, int> block_of;   // factor -> its dense top block
    std::vector<char> in_top;                          // per supernode: 1 if in its factor's dense top
    std::vector<int> top_member;                       // per top: factor index m
    int top_members = 0;                               // factors big enough for a dense top
    for (const SupernodalFactor* f : F) top_members += top_min_rows > 0 && f->n >= top_min_rows;
    long long dsize = 0, lsize = 0;
    for (size_t m = 0; m < F.size(); ++m) {
        const SupernodalFactor& f = *F[m];
        const int off = offsets[m];
        const auto seen = first_gid.find(&f);
        const int shared = seen == first_gid.end() ? -1 : seen->second;
        if (shared < 0) first_gid[&f] = (int)P.s0.size();
        int cut = INT_MAX;
        {
            auto b = block_of.find(&f);
            if (b == block_of.end()) {
                int tm = 0;
                const int c = dense_top_cut(f, top_min_rows, top_members, &tm);
                int id = -1;
                if (tm > 0) {
                    id = (int)P.top_blocks.size();
                    TriPlanHost::TopBlock B;
                    B.f = &f; B.cut = c; B.m = tm;
                    B.ld = (tm + TriPlanHost::kTopLd - 1) / TriPlanHost::kTopLd * TriPlanHost::kTopLd;
                    P.top_blocks.push_back(B);
                }
                b = block_of.emplace(&f, id).first;
            }
            if (b->second >= 0) {
                const TriPlanHost::TopBlock& B = P.top_blocks[b->second];
                cut = B.cut;
                TriPlanHost::Top T;
                T.m = B.m; T.ld = B.ld; T.w = b->second;
                P.tops.push_back(T);
                top_member.push_back((int)m);
            }
        }
        for (int k = 0; k < f.n; ++k) P.orig_of[off + k] = off + f.perm[k];
        for (size_t s = 0; s < f.sn.size(); ++s) {
            const Supernode& S = f.sn[s];
            if (S.n > TriPlanHost::kMaxSupernode) throw std::runtime_error("triangular-solve plan: supernode larger than the LDS tile");
            const int g = (int)P.s0.size();
            P.s0.push_back(off + S.s0); P.sz.push_back(S.n); P.nr.push_back((int)S.rows.size());
            P.roff.push_back((int)P.srows.size());
            for (int r : S.rows) P.srows.push_back(off + r);
            for (int k = 0; k < S.n; ++k) P.sn_of[off + S.s0 + k] = g;
            if (shared < 0) {
                P.doff.push_back(dsize); dsize += (long long)S.n * S.n;
                P.loff.push_back(lsize); lsize += (long long)S.n * S.rows.size();
            } else {  // a repeated factor (identical matrix): its blocks are stored once
                P.doff.push_back(P.doff[shared + s]);
                P.loff.push_back(P.loff[shared + s]);
            }
            owner.push_back(shared < 0);
            height.push_back(S.height);
            in_top.push_back(S.height >= cut);
            refs.push_back({(int)m, (int)s});
        }
    }
    const int nsn = P.nsn = (int)P.s0.size();
    // heights done by the fused kernels and heights with big supernodes (the rules of P.fused
    // and P.big below): their supernodes get W
    std::vector<char> fused_h(P.H + 1, 0), big_h(P.H + 1, 0), w_h(P.H + 1, 0);
    {
        std::vector<int> count(P.H + 1, 0), mx(P.H + 1, 0);
        for (int g = 0; g < nsn; ++g)
            if (!in_top[g]) { ++count[height[g]]; mx[height[g]] = std::max(mx[height[g]], P.sz[g]); }
        for (int h = 0; h <= P.H; ++h) {
            big_h[h] = mx[h] > TriPlanHost::kSliceRows;
            fused_h[h] = !big_h[h] && count[h] >= TriPlanHost::kFusedMinSupernodes;
            w_h[h] = fused_h[h] || big_h[h];
        }
    }
    P.woff.assign(nsn, 0);
    long long wsize = 0;
    for (int g = 0; g < nsn; ++g) {
        if (in_top[g] || !w_h[height[g]]) continue;
        if (owner[g]) { P.woff[g] = wsize; wsize += (long long)P.sz[g] * P.nr[g]; }
        else P.woff[g] = P.woff[first_gid.at(F[refs[g].m]) + refs[g].s];
    }
    P.w21.assign(wsize, 0.f);
    P.diag.assign(dsize, 0.f); P.diagT.assign(dsize, 0.f); P.l21.assign(lsize, 0.f);
    // diagonal-block inverses and the sub-diagonal blocks in float: small supernodes one per
    // thread, then the big ones one at a time with the threads over their columns
    constexpr int kBigInverse = 256;
    auto fill_blocks = [&](int g) {
        const SupernodalFactor& f = *F[refs[g].m];
        const int s = refs[g].s, ns = P.sz[g], nr = P.nr[g];
        std::vector<double> inv(f.L11[s]);
        invert_lower(inv, ns);
        float* D = &P.diag[P.doff[g]];
        float* DT = &P.diagT[P.doff[g]];
        for (int r = 0; r < ns; ++r)
            for (int k = 0; k < ns; ++k) {
                D[(size_t)r * ns + k] = k <= r ? (float)inv[(size_t)k * ns + r] : 0.f;   // Linv(r,k)
                DT[(size_t)r * ns + k] = k >= r ? (float)inv[(size_t)r * ns + k] : 0.f;  // Linv(k,r)
            }
        float* L = &P.l21[P.loff[g]];
        const std::vector<double>& src = f.L21[s];  // column-major nr x ns
        for (int a = 0; a < nr; ++a)
            for (int j = 0; j < ns; ++j) L[(size_t)a * ns + j] = (float)src[(size_t)j * nr + a];
        if (!in_top[g] && w_h[height[g]]) {
            // W(a, j) = sum_{k >= j} L21(a, k) Linv(k, j), Linv(k, j) = inv[j * ns + k]; rows of
            // L21 gathered first (row-major) so the inner loop is contiguous; threads over the
            // rows for the big supernodes (done one at a time)
            float* W = &P.w21[P.woff[g]];
#pragma omp parallel for schedule(dynamic, 16) if (ns > kBigInverse && !omp_in_parallel())
            for (int a = 0; a < nr; ++a) {
                std::vector<double> la(ns);
                for (int k = 0; k < ns; ++k) la[k] = src[(size_t)k * nr + a];
                for (int j = 0; j < ns; ++j) {
                    const double* ij = &inv[(size_t)j * ns];
                    double acc = 0.0;
                    for (int k = j; k < ns; ++k) acc += la[k] * ij[k];
                    W[(size_t)a * ns + j] = (float)acc;
                }
            }
        }
    };
#pragma omp parallel for schedule(dynamic)
    for (int g = 0; g < nsn; ++g)
        if (owner[g] && P.sz[g] <= kBigInverse) fill_blocks(g);
    for (int g = 0; g < nsn; ++g)
        if (owner[g] && P.sz[g] > kBigInverse) fill_blocks(g);
    // contributions and tile partials per supernode
    P.cofs.assign(nsn, 0); P.bofs.assign(nsn, 0);
    for (int g = 0; g < nsn; ++g) {
        P.cofs[g] = P.c_size; P.c_size += P.nr[g];
        P.bofs[g] = P.bpart_size;
        P.bpart_size += (long long)((P.nr[g] + TriPlanHost::kTile - 1) / TriPlanHost::kTile) * P.sz[g];
    }
    // incidences of each row in sub-diagonal blocks, ascending supernode (descendants first)
    P.inc_ptr.assign(P.n + 1, 0);
    for (int g = 0; g < nsn; ++g)
        for (int a = 0; a < P.nr[g]; ++a) ++P.inc_ptr[P.srows[P.roff[g] + a] + 1];
    for (int i = 0; i < P.n; ++i) P.inc_ptr[i + 1] += P.inc_ptr[i];
    P.cpos.assign(P.inc_ptr[P.n], 0);
    std::vector<int> inc_src(P.inc_ptr[P.n], 0);  // incidence -> its supernode
    {
        std::vector<int> fill(P.inc_ptr.begin(), P.inc_ptr.end() - 1);
        for (int g = 0; g < nsn; ++g)
            for (int a = 0; a < P.nr[g]; ++a) {
                const int k = fill[P.srows[P.roff[g] + a]]++;
                P.cpos[k] = P.cofs[g] + a;
                inc_src[k] = g;
            }
    }
    // dense tops: padded rows (ascending plan rows of S) and their incidences from R
    {
        std::vector<int> first_sn(F.size() + 1, 0);  // global supernode range of each factor member
        for (size_t m = 0; m < F.size(); ++m) first_sn[m + 1] = first_sn[m] + (int)F[m]->sn.size();
        P.top_inc_ptr.assign(1, 0);
        double incs = 0, trows = 0;
        for (size_t k = 0; k < P.tops.size(); ++k) {
            TriPlanHost::Top& T = P.tops[k];
            const int m = top_member[k];
            T.base = (int)P.top_rows.size();
            std::vector<int> rows;
            for (int g = first_sn[m]; g < first_sn[m + 1]; ++g)
                if (in_top[g]) for (int r = 0; r < P.sz[g]; ++r) rows.push_back(P.s0[g] + r);
            std::sort(rows.begin(), rows.end());
            if ((int)rows.size() != T.m) throw std::runtime_error("dense top: row count mismatch");
            rows.resize(T.ld, -1);
            for (int i : rows) {
                P.top_rows.push_back(i);
                P.top_of.push_back((int)k);
                if (i >= 0) {
                    for (int p = P.inc_ptr[i]; p < P.inc_ptr[i + 1]; ++p)
                        if (!in_top[inc_src[p]]) P.top_cpos.push_back(P.cpos[p]);
                    trows += 1;
                }
                P.top_inc_ptr.push_back((int)P.top_cpos.size());
            }
        }
        incs = (double)P.top_cpos.size();
        if (trows > 0) P.Gt = lanes_for(incs / trows);
    }
    // per height: supernodes, slices, rows, items
    const int nh = P.H + 1;
    P.sn_begin.assign(nh + 1, 0); P.row_begin.assign(nh + 1, 0);
    P.bi_begin.assign(nh + 1, 0);
    P.ci_entries.assign(nh, TriPlanHost::kContribEntries);
    P.max_sz.assign(nh, 0); P.big.assign(nh, 0); P.fused.assign(nh, 0); P.G.assign(nh, 4); P.Gc.assign(nh, 4); P.Gf.assign(nh, 4); P.Gb.assign(nh, 4);
    P.sl_rows.assign(nh, TriPlanHost::kSliceRows);
    for (int h = 0; h < nh; ++h) {
        int mx = 0;
        for (int g = 0; g < nsn; ++g) if (height[g] == h && !in_top[g]) mx = std::max(mx, P.sz[g]);
        if (mx > TriPlanHost::kSliceRows) P.sl_rows[h] = TriPlanHost::kBigSliceRows;
    }
    for (int h = 0; h < nh; ++h) {
        P.sn_begin[h] = (int)P.sn_list.size();
        P.row_begin[h] = (int)P.rowlist.size();
        P.bi_begin[h] = (int)P.bi_sn.size();
        double rows = 0, diag_len = 0, contrib_rows = 0, contrib_len = 0, incs = 0, chunks = 0;
        std::vector<int> at_h;
        for (int g = 0; g < nsn; ++g) if (height[g] == h && !in_top[g]) at_h.push_back(g);
        // supernodes that share blocks run next to each other (the shared blocks stay in cache)
        std::stable_sort(at_h.begin(), at_h.end(), [&](int a, int b) { return P.loff[a] != P.loff[b] ? P.loff[a] < P.loff[b] : P.doff[a] < P.doff[b]; });
        for (int g : at_h) {
            const int n = P.sz[g], nr = P.nr[g];
            P.sn_list.push_back(g);
            P.max_sz[h] = std::max(P.max_sz[h], n);
            for (int k = 0; k < n; ++k) {
                P.rowlist.push_back(P.s0[g] + k);
                incs += P.inc_ptr[P.s0[g] + k + 1] - P.inc_ptr[P.s0[g] + k];
            }
            rows += n; diag_len += 0.5 * n * (n + 1.0);
            contrib_rows += nr; contrib_len += (double)nr * n;
            const int nch = (nr + TriPlanHost::kTile - 1) / TriPlanHost::kTile;
            for (int c = 0; c < nch; ++c)
                for (int j0 = 0; j0 < n; j0 += TriPlanHost::kTile) { P.bi_sn.push_back(g); P.bi_c.push_back(c); P.bi_j0.push_back(j0); }
            chunks += (double)nch * n;
        }
        P.big[h] = P.max_sz[h] > TriPlanHost::kSliceRows;
        P.fused[h] = !P.big[h] && (int)P.sn_list.size() - P.sn_begin[h] >= TriPlanHost::kFusedMinSupernodes;
        if (rows > 0) { P.G[h] = lanes_for(diag_len / rows); P.Gf[h] = lanes_for(incs / rows); P.Gb[h] = lanes_for(chunks / rows); }
        if (contrib_rows > 0) P.Gc[h] = lanes_for(contrib_len / contrib_rows);
    }
    P.sn_begin[nh] = (int)P.sn_list.size();
    for (int h = 0; h < nh; ++h)
        if ((bool)P.fused[h] != (bool)fused_h[h] || (bool)P.big[h] != (bool)big_h[h])
            throw std::runtime_error("triangular-solve plan: fused or big heights disagree");
    // runs of block-sharing supernodes (sn_list is ordered by block offsets within each height)
    P.run_begin.assign(nh + 1, 0);
    P.run_max.assign(nh, 1);
    for (int h = 0; h < nh; ++h) {
        P.run_begin[h] = (int)P.run_q.size();
        for (int q = P.sn_begin[h]; q < P.sn_begin[h + 1];) {
            const int g = P.sn_list[q];
            int k = 1;
            while (k < TriPlanHost::kRunMembers && q + k < P.sn_begin[h + 1] && P.doff[P.sn_list[q + k]] == P.doff[g] &&
                   P.loff[P.sn_list[q + k]] == P.loff[g])
                ++k;
            P.run_q.push_back(q);
            P.run_n.push_back(k);
            P.run_max[h] = std::max(P.run_max[h], k);
            q += k;
        }
    }
    P.run_begin[nh] = (int)P.run_q.size();
    P.row_begin[nh] = (int)P.rowlist.size();
    P.bi_begin[nh] = (int)P.bi_sn.size();
    build_slices_items(P.sn_list, P.sn_begin, P.sz, P.nr, P.sl_rows, P.ci_entries, P.sl_sn, P.sl_r0, P.sl_begin, P.ci_sn, P.ci_a0,
                       P.ci_a1, P.ci_begin);
}

void build_slices_items(const std::vector<int>& sn_list, const std::vector<int>& sn_begin, const std::vector<int>& sz,
                        const std::vector<int>& nr, const std::vector<int>& sl_rows, const std::vector<int>& ci_entries,
                        std::vector<int>& sl_sn, std::vector<int>& sl_r0, std::vector<int>& sl_begin, std::vector<int>& ci_sn,
                        std::vector<int>& ci_a0, std::vector<int>& ci_a1, std::vector<int>& ci_begin) {
    const int nh = (int)sn_begin.size() - 1;
    sl_sn.clear(); sl_r0.clear(); ci_sn.clear(); ci_a0.clear(); ci_a1.clear();
    sl_begin.assign(nh + 1, 0);
    ci_begin.assign(nh + 1, 0);
    for (int h = 0; h < nh; ++h) {
        sl_begin[h] = (int)sl_sn.size();
        ci_begin[h] = (int)ci_sn.size();
        for (int q = sn_begin[h]; q < sn_begin[h + 1]; ++q) {
            const int g = sn_list[q], n = sz[g], m = nr[g];
            for (int r0 = 0; r0 < n; r0 += sl_rows[h]) { sl_sn.push_back(g); sl_r0.push_back(r0); }
            const int crows = std::max(1, ci_entries[h] / std::max(1, n));
            for (int a0 = 0; a0 < m; a0 += crows) { ci_sn.push_back(g); ci_a0.push_back(a0); ci_a1.push_back(std::min(m, a0 + crows)); }
        }
    }
    sl_begin[nh] = (int)sl_sn.size();
    ci_begin[nh] = (int)ci_sn.size();
}

}  // namespace hdd
