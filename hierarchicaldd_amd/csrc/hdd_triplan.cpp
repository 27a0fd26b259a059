#include "hdd_triplan.h"
#include <algorithm>
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
}  // namespace

void build_triplan(const std::vector<const SupernodalFactor*>& F, const std::vector<int>& offsets, TriPlanHost& P) {
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
    long long dsize = 0, lsize = 0;
    for (size_t m = 0; m < F.size(); ++m) {
        const SupernodalFactor& f = *F[m];
        const int off = offsets[m];
        const auto seen = first_gid.find(&f);
        const int shared = seen == first_gid.end() ? -1 : seen->second;
        if (shared < 0) first_gid[&f] = (int)P.s0.size();
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
            refs.push_back({(int)m, (int)s});
        }
    }
    const int nsn = P.nsn = (int)P.s0.size();
    P.diag.assign(dsize, 0.f); P.diagT.assign(dsize, 0.f); P.l21.assign(lsize, 0.f);
#pragma omp parallel for schedule(dynamic)
    for (int g = 0; g < nsn; ++g) {
        if (!owner[g]) continue;
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
    }
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
    {
        std::vector<int> fill(P.inc_ptr.begin(), P.inc_ptr.end() - 1);
        for (int g = 0; g < nsn; ++g)
            for (int a = 0; a < P.nr[g]; ++a) P.cpos[fill[P.srows[P.roff[g] + a]]++] = P.cofs[g] + a;
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
        for (int g = 0; g < nsn; ++g) if (height[g] == h) mx = std::max(mx, P.sz[g]);
        if (mx > TriPlanHost::kSliceRows) P.sl_rows[h] = TriPlanHost::kBigSliceRows;
    }
    for (int h = 0; h < nh; ++h) {
        P.sn_begin[h] = (int)P.sn_list.size();
        P.row_begin[h] = (int)P.rowlist.size();
        P.bi_begin[h] = (int)P.bi_sn.size();
        double rows = 0, diag_len = 0, contrib_rows = 0, contrib_len = 0, incs = 0, chunks = 0;
        std::vector<int> at_h;
        for (int g = 0; g < nsn; ++g) if (height[g] == h) at_h.push_back(g);
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

void triplan_solve_host(const TriPlanHost& P, const float* b, float* x) {
    if (P.n == 0) return;
    std::vector<double> y(P.n, 0.0), z(P.n, 0.0), c(P.c_size, 0.0), bp(P.bpart_size, 0.0), t(P.n, 0.0);
    for (int h = 0; h <= P.H; ++h) {
        for (int q = P.row_begin[h]; q < P.row_begin[h + 1]; ++q) {
            const int i = P.rowlist[q];
            double s = 0;
            for (int p = P.inc_ptr[i]; p < P.inc_ptr[i + 1]; ++p) s += c[P.cpos[p]];
            t[i] = (double)b[P.orig_of[i]] - s;
        }
        for (int q = P.sn_begin[h]; q < P.sn_begin[h + 1]; ++q) {
            const int g = P.sn_list[q], n = P.sz[g], base = P.s0[g];
            const float* D = &P.diag[P.doff[g]];
            for (int r = 0; r < n; ++r) {
                double acc = 0;
                for (int k = 0; k <= r; ++k) acc += (double)D[(size_t)r * n + k] * t[base + k];
                y[base + r] = acc;
            }
        }
        for (int it = P.ci_begin[h]; it < P.ci_begin[h + 1]; ++it) {
            const int g = P.ci_sn[it], n = P.sz[g];
            for (int a = P.ci_a0[it]; a < P.ci_a1[it]; ++a) {
                double acc = 0;
                for (int j = 0; j < n; ++j) acc += (double)P.l21[P.loff[g] + (long long)a * n + j] * y[P.s0[g] + j];
                c[P.cofs[g] + a] = acc;
            }
        }
    }
    for (int h = P.H; h >= 0; --h) {
        for (int it = P.bi_begin[h]; it < P.bi_begin[h + 1]; ++it) {
            const int g = P.bi_sn[it], n = P.sz[g], cc = P.bi_c[it], j0 = P.bi_j0[it];
            const int a0 = cc * TriPlanHost::kTile, a1 = std::min(P.nr[g], a0 + TriPlanHost::kTile);
            for (int j = j0; j < std::min(n, j0 + TriPlanHost::kTile); ++j) {
                double acc = 0;
                for (int a = a0; a < a1; ++a) acc += (double)P.l21[P.loff[g] + (long long)a * n + j] * z[P.srows[P.roff[g] + a]];
                bp[P.bofs[g] + (long long)cc * n + j] = acc;
            }
        }
        for (int q = P.sn_begin[h]; q < P.sn_begin[h + 1]; ++q) {
            const int g = P.sn_list[q], n = P.sz[g], base = P.s0[g];
            const int nch = (P.nr[g] + TriPlanHost::kTile - 1) / TriPlanHost::kTile;
            for (int k = 0; k < n; ++k) {
                double s = 0;
                for (int cc = 0; cc < nch; ++cc) s += bp[P.bofs[g] + (long long)cc * n + k];
                t[base + k] = y[base + k] - s;
            }
            const float* DT = &P.diagT[P.doff[g]];
            for (int r = 0; r < n; ++r) {
                double acc = 0;
                for (int k = r; k < n; ++k) acc += (double)DT[(size_t)r * n + k] * t[base + k];
                z[base + r] = acc;
                x[P.orig_of[base + r]] = (float)acc;
            }
        }
    }
}

}  // namespace hdd
