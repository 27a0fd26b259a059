// Sparse Cholesky of (A + eps I) for the interface solver's exact solves.
//
// The reference factors every Aii and Arr with MKL PARDISO (METIS ordering,
// IFS/SIGMA_SMOOTHER.cpp:115-288) after adding 1e-5 to the diagonal
// (Eigen_To_Pardiso, IFS/SIGMA_SMOOTHER.cpp:50-111). This is the product's own
// replacement: geometric nested dissection on the dofs' grid coordinates, a
// multifrontal supernodal factorisation in double precision, and dense
// per-supernode blocks laid out for the GPU triangular solves.
#pragma once
#include <vector>
#include "hdd_sigma.h"

namespace hdd {

struct Supernode {
    int s0 = 0, n = 0;       // permuted columns [s0, s0+n)
    int parent = -1;
    int height = 0;          // 0 for leaves
    std::vector<int> rows;   // permuted row indices below the diagonal block, ascending
};

struct SupernodalFactor {
    int n = 0;
    std::vector<int> perm;                 // permuted position -> original index
    std::vector<int> iperm;                // original index -> permuted position
    std::vector<Supernode> sn;             // postorder
    std::vector<std::vector<double>> L11;  // per supernode, n x n column-major, lower
    std::vector<std::vector<double>> L21;  // per supernode, rows x n column-major
    int max_height = 0;
    size_t nnz() const;
};

// Factor (A + eps I). coords[i] orders dof i geometrically; A is symmetric CSR.
void factor_spd(const Csr& A, const std::vector<Int3>& coords, double eps, SupernodalFactor& F, int leaf_size = 96);

// Host solve with the factor (double), for checks.
void factor_solve(const SupernodalFactor& F, const double* b, double* x);

// Dense fronts of at least device_front_min_rows() rows (kDeviceFrontMin, or
// HDD_DEVICE_FRONT_MIN) are factored on the GPU (hdd_factor_device.cpp) when a
// device is present and HDD_DEVICE_FRONTS is not 0.
constexpr int kDeviceFrontMin = 1024;
bool device_fronts_enabled();
int device_front_min_rows();
void partial_cholesky_device(double* F, int m, int ns);

// In-place inverse of a dense lower-triangular n x n column-major matrix.
void invert_lower(std::vector<double>& L, int n);

}  // namespace hdd
