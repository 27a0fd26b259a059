// Host construction of the GPU triangular-solve plan for a batch of supernodal
// factors: x = (A + eps I)^{-1} b = L^{-T} L^{-1} b for every factor of the
// batch at once (the exact solves of SIGMA_SMOOTHER::Smooth and
// Compute_Residual, IFS/SIGMA_SMOOTHER.cpp:292-352, which the reference hands
// to PARDISO).
//
// Layout. Supernodes of all factors are numbered globally and grouped by tree
// height (leaves 0). Per supernode g: the explicit inverse of its diagonal
// block, row-major (diag, lower triangle used) and transposed (diagT, upper
// triangle used), and its sub-diagonal block L21_g row-major (nr x n), the
// only copy of L21. Both passes read L21 along its rows, contiguously.
//
// Forward, per height h:
//   reduce  t_i = b_i - sum over the incidences of row i of c[pos]       (heights with big supernodes)
//   diag    y_g = Linv_g t_g, a workgroup per row slice                  (t inline for small supernodes)
//   contrib c_g = L21_g y_g, one value per sub-diagonal row of g         (consumed by ancestors' rows)
// Backward, per height h from the root:
//   tiles   p_g[c, j] = sum_{a in row chunk c} L21_g[a, j] z[rows_g[a]]  (64 x 64 tiles, one wavefront each)
//   reduce  t_j = y_j - sum_c p_g[c, j]                                   (heights with big supernodes)
//   diag    z_g = Linv_g^T t_g
// Every sum runs in a fixed order, so results are deterministic.
//
// A factor passed more than once (identical matrices: subdomains of the same
// shape) keeps one copy of its blocks; its members get their own rows and
// partial sums, and within a height the supernodes that share blocks are
// scheduled next to each other, so the shared blocks are read from cache.
#pragma once
#include <vector>
#include "hdd_factor.h"

namespace hdd {

struct TriPlanHost {
    static constexpr int kSliceRows = 128;       // rows of Linv t per workgroup (a supernode up to this size is one slice)
    static constexpr int kBigSliceRows = 16;     // rows per workgroup for heights with bigger supernodes
    static constexpr int kTile = 64;             // backward tile: rows x columns
    static constexpr int kContribEntries = 4096; // L21 entries per contribution item
    static constexpr int kMaxSupernode = 16384;  // the per-supernode kernels keep t in LDS
    int n = 0, H = -1, nsn = 0;
    std::vector<int> orig_of;                 // plan row -> row of the batch's original numbering
    std::vector<int> s0, sz, nr, roff, srows; // supernode first row, size, sub-diagonal rows (global plan rows)
    std::vector<int> sn_of;                   // plan row -> supernode
    std::vector<long long> loff, doff;        // offsets into l21 / diag
    std::vector<float> diag, diagT, l21;
    // supernodes of fused heights: W_g = L21_g Linv_g (nr x n, row-major like L21, formed
    // in double), so that the fused kernels need not wait for y (forward) or t (backward)
    std::vector<long long> woff;              // offsets into w21 (fused heights only)
    std::vector<float> w21;
    static constexpr int kFusedMinSupernodes = 1024;  // a height is done by one workgroup per supernode only with this many
    std::vector<int> sn_begin, max_sz, big;   // per height: supernode range in sn_list, largest supernode, 1 if > kSliceRows
    std::vector<int> fused;                   // per height: 1 if every supernode is <= kSliceRows and there are many
    // Runs: consecutive supernodes of sn_list at one height that share their
    // blocks (members of identical factors), at most kRunMembers each; a fused
    // height is done by one workgroup per run, which reads the shared blocks once.
    // Measured on the 256^3 level-0 interior plan: runs of 2 / 4 / 8 / 16 members
    // solve in 0.93 / 0.83 / 0.94 / 1.16 ms (more members: fewer, longer workgroups
    // with more LDS each; the results are the same bits for any run length).
    static constexpr int kRunMembers = 4;
    std::vector<int> run_q, run_n, run_begin; // run start in sn_list, members; runs by height
    std::vector<int> run_max;                 // per height: largest run
    std::vector<int> sn_list;
    std::vector<int> sl_sn, sl_r0, sl_begin;  // row slices (supernode, first row) by height
    std::vector<int> sl_rows;                 // rows per slice, per height (launch shape: does not change any sum)
    std::vector<int> ci_entries;              // L21 entries per contribution item, per height (launch shape)
    std::vector<int> rowlist, row_begin;      // plan rows by height (for the reduce kernels)
    std::vector<int> G;                       // lanes per row of the diag kernels, per height
    // forward
    std::vector<long long> cofs;              // per supernode: offset of its contributions c_g
    std::vector<int> inc_ptr;                 // per plan row: incidences [inc_ptr[i], inc_ptr[i+1])
    std::vector<long long> cpos;              // incidence -> position in c (ascending supernode)
    std::vector<int> ci_sn, ci_a0, ci_a1, ci_begin, Gc;  // contribution items by height, lanes per row
    std::vector<int> Gf;                      // lanes per row of the forward reduce, per height
    long long c_size = 0;
    // backward
    std::vector<long long> bofs;              // per supernode: offset of its tile partials (nchunk x n)
    std::vector<int> bi_sn, bi_c, bi_j0, bi_begin;      // tiles by height
    std::vector<int> Gb;                      // lanes per row of the backward reduce, per height
    long long bpart_size = 0;
    size_t factor_nnz = 0;
    // Dense tops. The top heights of a tall tree hold few supernodes each, so
    // every height costs a handful of launches with little parallelism. For a
    // factor of at least top_min_rows rows, the supernodes S of its top heights
    // (ancestor-closed, at most kTopMaxRows rows) leave the per-height lists and
    // their part of the solve becomes two dense triangular matvecs:
    //   t_S = b_S - L_SR y_R          (contributions of the rest, R, to S rows)
    //   y_S = Linv t_S,  z_S = Linv^T y_S,  Linv = L_SS^{-1}
    // run between the forward and the backward pass (z_S feeds the backward
    // tiles of R). Linv is formed on the device at set-up from L_SS in double
    // and kept in float, as the diagonal blocks' inverses are; the inverse of
    // L_SS L_SS^T itself would lose eps * cond instead of eps * sqrt(cond).
    static constexpr int kTopMaxRows = 12288;
    static constexpr int kTopLd = 64;         // row stride of the dense blocks and of t: a multiple of this
    struct Top {
        int m = 0, ld = 0;    // rows of S, padded row stride
        int base = 0;         // first entry of the top in top_rows (padded rows)
        int w = 0;            // dense block: tops of a repeated factor share one
    };
    struct TopBlock {
        const SupernodalFactor* f = nullptr;  // factor the block is assembled from (valid while F is)
        int cut = 0;                          // supernodes of f at heights >= cut form S
        int m = 0, ld = 0;
    };
    std::vector<Top> tops;
    std::vector<TopBlock> top_blocks;
    std::vector<int> top_rows;                // per padded top row: plan row (-1 for padding), tops in order
    std::vector<int> top_of;                  // per padded top row: its top
    std::vector<int> top_inc_ptr;             // per padded top row: incidences from R, [ptr[q], ptr[q+1])
    std::vector<long long> top_cpos;          // incidence -> position in c (ascending supernode)
    int Gt = 4;                               // lanes per row of the top's right-hand side
};

// top_min_rows: factors with at least this many rows get a dense top (0: none).
void build_triplan(const std::vector<const SupernodalFactor*>& F, const std::vector<int>& offsets, TriPlanHost& P, int top_min_rows = 0);

// The height from which the supernodes of f form its dense top, with the top's
// rows in *m (a height above f's tree and *m = 0 when f gets none), in a plan
// with `members` factors of at least top_min_rows rows (at most 64). Heights
// are taken from the root down while the top stays within kTopMaxRows rows and
// half of f, and while the dense blocks of all members grow by less than one
// height's launches cost.
int dense_top_cut(const SupernodalFactor& f, int top_min_rows, int members, int* m);

// Dense L_SS of a top block, m x m column-major in double (rows of S in
// ascending permuted order; S is ancestor-closed, so L_SS is lower triangular),
// written to L (m * m doubles, every entry set).
void dense_top_factor(const TriPlanHost::TopBlock& B, double* L);

// Rows a factor needs to have a dense top by default, or HDD_DENSE_TOP from the
// environment (0 turns dense tops off).
int dense_top_min_rows();

// (Re)builds the row slices of the diagonal steps and the items of the forward
// contributions from sn_list, sl_rows and ci_entries. Every row's sum is the
// same whatever the slice and item sizes, so they can be tuned freely.
void build_slices_items(const std::vector<int>& sn_list, const std::vector<int>& sn_begin, const std::vector<int>& sz,
                        const std::vector<int>& nr, const std::vector<int>& sl_rows, const std::vector<int>& ci_entries,
                        std::vector<int>& sl_sn, std::vector<int>& sl_r0, std::vector<int>& sl_begin, std::vector<int>& ci_sn,
                        std::vector<int>& ci_a0, std::vector<int>& ci_a1, std::vector<int>& ci_begin);

}  // namespace hdd
