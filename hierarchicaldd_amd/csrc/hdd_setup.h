// Per-geometry setup of the product (host side): subdomain division, face and
// boundary flags, and the coarse V-cycle levels. Semantics follow
//   Domain_Divider       DDS/SPGrid_Domain_Divider_Helper.h:49-81
//   Flaging<3>           DDS/SPGrid_Flag_Helper.h:42-74
//   Flaging_Boundary<3>  DDS/SPGrid_Flag_Boundary_Helper.h:40-109
//   Topology_Downsample  DDS/SPGrid_V_Cycle_Topology.h:35-67 (DDS copy, no pre-marking)
//   V_Cycle_Topology     DDS/SPGrid_V_Cycle_Topology.h:102-177 (size_{l+1} = size_l/2+1)
// (DDS = demos/Domain_Decomposition_CPU_Only/SPGrid_Domain_Decomposition_Solver)
#pragma once
#include "hdd_grid.h"

namespace hdd {

inline Int3 subdomain_counts(Int3 size, Int3 sub) {
    auto f = [](int s, int q) { return (s - 1) / q + (((s - 1) % q) ? 1 : 0); };
    return {f(size.x, sub.x), f(size.y, sub.y), f(size.z, sub.z)};
}

// Interface marking and partition ids (origin 0)
void divide_subdomains(PagedLevel& L, Int3 size, Int3 sub);
// face bits of Active|Interface cells whose neighbour is Active|Interface|Dirichlet
void set_face_flags(PagedLevel& L);
// Boundary bit: Active cells within the reference's 3-sweep distance 3 of a
// flag-0, Interface or Dirichlet cell
void set_boundary_flags(PagedLevel& L);
// coarse level from a fine level
void downsample_level(const PagedLevel& fine, PagedLevel& coarse);

}  // namespace hdd
