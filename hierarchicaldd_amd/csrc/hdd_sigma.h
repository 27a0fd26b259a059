// Interface (Schur complement) solver of the product: host-side operator
// generation and factorisation, device-side sigma multigrid.
#pragma once
#include <cstdint>
#include <vector>
#include <hip/hip_runtime.h>
#include "hdd_grid.h"

namespace hdd {

struct Csr {
    int rows = 0, cols = 0;
    std::vector<int> ptr{0}, idx;
    std::vector<float> val;
    size_t nnz() const { return idx.size(); }
};

// Output of the interface operator generator (reference INTERFACE_SOLVER_DATA,
// IFS/INTERFACE_SOLVER_DATA.h): per sigma level the global Arr, per subdomain
// Aii / Air, and the level prolongations Prr.
struct SigmaHost {
    int levels = 0, nsub = 0;
    std::vector<int> n_if;                      // interface dofs per level
    std::vector<Int3> interface_map;            // level-0 interface cell of each dof
    std::vector<std::vector<Int3>> if_coords;   // per level, level coordinates
    std::vector<Csr> Arr, Prr;
    std::vector<std::vector<Csr>> Aii, Air;     // [level][subdomain]
    std::vector<std::vector<std::vector<Int3>>> Aii_coords;  // [level][subdomain][dof], fine coordinates
};

void generate_interface_operators(const struct PagedLevel& input_flags, Int3 size, Int3 sub, int sigma_levels, SigmaHost& out);

struct SigmaDevice;  // defined in hdd_sigma_device.h

}  // namespace hdd
