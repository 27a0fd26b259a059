// Slow eigenmodes of the bottom sigma smoother, for replacing its 200
// stationary sweeps with one exact solve plus a low-rank correction.
//
// The bottom of SIGMA_MULTIGRID_SOLVER::V_Cycle (IFS/SIGMA_MULTIGRID_SOLVER.cpp:36-47)
// applies SIGMA_SMOOTHER::Smooth (IFS/SIGMA_SMOOTHER.cpp:292-324) 200 times,
//     u <- S^{-1} (b + K u),  S = Arr + eps I,  K = sum_s Air_s^T (Aii_s + eps I)^{-1} Air_s.
// With M = S^{-1} K and x* = (S - K)^{-1} b, after 200 sweeps from u0
//     u200 = x* + M^200 (u0 - x*).
// M is self-adjoint in the S inner product with eigenvalues in [0, 1). With
// S-orthonormal eigenvectors v_j (M v_j = l_j v_j) and v_j^T S x* = v_j^T b / (1 - l_j),
//     u200 = x* + sum_j l_j^200 v_j (v_j^T S u0 - v_j^T b / (1 - l_j)),
// and every mode with l_j^200 below float resolution can be dropped. This file
// finds the modes that cannot, in double precision on the host.
#pragma once
#include <vector>
#include "hdd_factor.h"
#include "hdd_sigma.h"

namespace hdd {

// Eigen-decomposition of a dense symmetric n x n matrix (column-major):
// Householder reduction to tridiagonal form, then implicit QL iterations with
// Wilkinson shifts (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11, 1968).
// On return A holds the orthonormal eigenvectors (columns) and w the
// eigenvalues, ascending.
void sym_eig(int n, std::vector<double>& A, std::vector<double>& w);

// x = (A + eps I)^{-1} b for p right-hand sides at once (columns of B and X, leading dimension F.n).
void factor_solve_block(const SupernodalFactor& F, int p, const double* B, double* X);

// Multi-threaded solve with a supernodal factor, 8 right-hand sides at a time.
// Disjoint subtrees of the elimination tree run in parallel; their updates of
// rows that belong to the shared ancestors go to per-subtree buffers that are
// summed in subtree order (so the result does not depend on the thread
// schedule) before the shared top of the tree is solved on one thread. The
// backward pass runs the top first, then the subtrees in parallel.
class ParallelSolve8 {
public:
    explicit ParallelSolve8(const SupernodalFactor& F);
    // X = (A + eps I)^{-1} B for 8 columns of length F.n (column-major, original order)
    void solve(const double* B, double* X) const;
    int tasks() const { return (int)task_first.size(); }

private:
    const SupernodalFactor& F;
    std::vector<int> task_first, task_last, task_end;  // supernode range and exclusive column end of each subtree
    std::vector<int> top;                              // shared ancestors, postorder
    std::vector<int> top_slot;                         // column -> row of the thread buffers, -1 if not top
    std::vector<int> split;                            // per supernode: rows[split..] are top rows
    int n_top = 0;
};

// Combined bottom system [[Aii (all subdomains, concatenated), Air], [Air^T, Arr]]
// of sigma level l (interior dofs first), with coordinates for the ordering.
void combined_bottom(const SigmaHost& H, int l, Csr& C, std::vector<Int3>& xyz);

struct BottomSpectrum {
    bool resolved = false;        // every eigenvalue >= lambda_cut was found
    int steps = 0;                // block Krylov steps (0: dense path)
    double lambda_max = -1;       // largest eigenvalue of M
    double max_residual = 0;      // largest relative residual of a kept pair
    double seconds = 0;
    std::vector<double> lambda;   // kept eigenvalues, descending
    std::vector<double> V;        // n x k, column-major, S-orthonormal
    std::vector<double> SV;       // S V
};

// Eigenpairs of M with eigenvalue >= lambda_cut. Fcomb is the factor of the
// combined system plus eps I, whose interface block is rows [n_int, n_int + n).
// Systems of up to 600 dofs are decomposed densely; larger ones by block Lanczos with full
// reorthogonalisation on T = (S - K)^{-1} S (same eigenvectors, eigenvalues
// 1 / (1 - l)), which spreads the slow end of the spectrum apart.
void bottom_spectrum(const Csr& Arr, double eps, const SupernodalFactor& Fcomb, int n_int, double lambda_cut,
                     size_t max_basis_bytes, BottomSpectrum& out);

}  // namespace hdd
