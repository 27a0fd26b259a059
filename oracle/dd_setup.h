// ORACLE — test infrastructure only (see spgrid.h header).
// Setup passes of the reference solver, restated on the oracle's Grid:
//   Domain_Divider         DDS/SPGrid_Domain_Divider_Helper.h:19-81
//   Flaging<3>             DDS/SPGrid_Flag_Helper.h:24-74
//   Flaging_Boundary<3>    DDS/SPGrid_Flag_Boundary_Helper.h:40-109
//   Topology_Downsample    DDS/SPGrid_V_Cycle_Topology.h:35-67
// (DDS = demos/Domain_Decomposition_CPU_Only/SPGrid_Domain_Decomposition_Solver)
#pragma once
#include "spgrid.h"
#include <array>

namespace ora {

struct I3 { int x, y, z; };

// Domain_Divider::Number_of_Subdomains (DDS/SPGrid_Domain_Divider_Helper.h:31-35)
inline I3 number_of_subdomains(I3 size, I3 sub) {
    auto f = [](int s, int q) { int n = (s - 1) / q; if ((s - 1) % q) ++n; return n; };
    return {f(size.x, sub.x), f(size.y, sub.y), f(size.z, sub.z)};
}

// Domain_Divider::Run (DDS/SPGrid_Domain_Divider_Helper.h:49-81), origin = 0.
template <class T>
void domain_divider(Grid<T>& g, I3 size, I3 sub) {
    I3 ns = number_of_subdomains(size, sub);
    for (int p = 0; p < g.npages(); ++p) {
        int X, Y, Z; g.base(p, X, Y, Z);
        for (int c = 0; c < BCELLS; ++c) {
            int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
            uint32_t& f = g.flags[(size_t)p * BCELLS + c];
            if (!(f & Active)) continue;
            if (x % sub.x == 0 || y % sub.y == 0 || z % sub.z == 0) {
                f -= Active; f |= Interface;
            } else {
                int id = ((x / sub.x) * ns.y + (y / sub.y)) * ns.z + (z / sub.z);
                f |= (uint32_t)(id + 1);
            }
        }
    }
}

// Flaging<3>::Run (DDS/SPGrid_Flag_Helper.h:42-74)
template <class T>
void flaging(Grid<T>& g) {
    const uint32_t nb = Active | Interface | Dirichlet;
    for (int p = 0; p < g.npages(); ++p) {
        int X, Y, Z; g.base(p, X, Y, Z);
        for (int c = 0; c < BCELLS; ++c) {
            int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
            uint32_t& f = g.flags[(size_t)p * BCELLS + c];
            if (!(f & (Active | Interface))) continue;
            for (int fc = 0; fc < 6; ++fc)
                if (g.is_set(x + FDX[fc], y + FDY[fc], z + FDZ[fc], nb)) f |= (FaceMX << fc);
        }
    }
}

// Flaging_Boundary<3>::Run (DDS/SPGrid_Flag_Boundary_Helper.h:40-109).
// Restated literally: a (4+8)x(8+8)x(8+8) padded gather whose x-index 8 aliases
// cell x=0 of the +x neighbour block (cell_x=(index.x+4)%4), distance init,
// three in-place sweeps over [-3, size+3] in x,y,z lexicographic order, and
// the <=3 threshold on Active cells.
template <class T>
void flaging_boundary(Grid<T>& g) {
    const int NX = 13, NY = 17, NZ = 17;  // index ranges [-4,8] x [-4,12] x [-4,12]
    std::vector<uint32_t> newflags(g.flags);
    std::vector<uint32_t> fb(NX * NY * NZ), db(NX * NY * NZ);
    auto I = [&](int i, int j, int k) { return ((i + 4) * NY + (j + 4)) * NZ + (k + 4); };
    for (int p = 0; p < g.npages(); ++p) {
        int bx = g.pbx[p], by = g.pby[p], bz = g.pbz[p];
        int nbp[3][3][3];
        for (int i = -1; i < 2; ++i)
            for (int j = -1; j < 2; ++j)
                for (int k = -1; k < 2; ++k) nbp[i + 1][j + 1][k + 1] = g.page_of_block(bx + i, by + j, bz + k);
        for (int i = -4; i <= 8; ++i)
            for (int j = -4; j <= 12; ++j)
                for (int k = -4; k <= 12; ++k) {
                    int bi = i < 0 ? 0 : (i >= BX ? 2 : 1);
                    int bj = j < 0 ? 0 : (j >= BY ? 2 : 1);
                    int bk = k < 0 ? 0 : (k >= BZ ? 2 : 1);
                    int q = nbp[bi][bj][bk];
                    uint32_t v = 0;
                    if (q >= 0) {
                        int cx = (i + BX) % BX, cy = (j + BY) % BY, cz = (k + BZ) % BZ;
                        v = g.flags[(size_t)q * BCELLS + ((cx << 6) | (cy << 3) | cz)];
                    }
                    fb[I(i, j, k)] = v;
                }
        for (int n = 0; n < NX * NY * NZ; ++n)
            db[n] = (fb[n] == 0u || (fb[n] & (Interface | Dirichlet))) ? 0u : 0xffffu;
        for (int dis = 0; dis < 3; ++dis)
            for (int i = -3; i <= 7; ++i)
                for (int j = -3; j <= 11; ++j)
                    for (int k = -3; k <= 11; ++k) {
                        uint32_t& d = db[I(i, j, k)];
                        if (d == 0) continue;
                        d = std::min<uint32_t>(d, db[I(i + 1, j, k)] + 1);
                        d = std::min<uint32_t>(d, db[I(i, j + 1, k)] + 1);
                        d = std::min<uint32_t>(d, db[I(i, j, k + 1)] + 1);
                        d = std::min<uint32_t>(d, db[I(i - 1, j, k)] + 1);
                        d = std::min<uint32_t>(d, db[I(i, j - 1, k)] + 1);
                        d = std::min<uint32_t>(d, db[I(i, j, k - 1)] + 1);
                    }
        for (int c = 0; c < BCELLS; ++c) {
            int i = c >> 6, j = (c >> 3) & 7, k = c & 7;
            uint32_t& f = newflags[(size_t)p * BCELLS + c];
            if ((f & Active) && db[I(i, j, k)] <= 3) f |= Boundary;
        }
    }
    g.flags.swap(newflags);
}

// Topology_Downsample (DDS/SPGrid_V_Cycle_Topology.h:35-67): c=(f+1)/2;
// Dirichlet/Interface fine -> Dirichlet coarse (overwrite), else Active fine ->
// Active coarse if the coarse cell is not yet set.
template <class T>
void topology_downsample(const Grid<T>& fine, Grid<T>& coarse) {
    for (int bi : fine.blocks) {
        int X, Y, Z; fine.base(bi, X, Y, Z);
        for (int c = 0; c < BCELLS; ++c) {
            int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
            uint32_t ff = fine.flags[(size_t)bi * BCELLS + c];
            int cx = (x + 1) / 2, cy = (y + 1) / 2, cz = (z + 1) / 2;
            if (!coarse.is_set(cx, cy, cz, Dirichlet | Active | Interface)) {
                if (ff & Active) coarse.mask(cx, cy, cz, Active);
                else if (ff & (Dirichlet | Interface)) coarse.mask(cx, cy, cz, Dirichlet);
                continue;
            }
            if (ff & (Dirichlet | Interface)) coarse.flags[coarse.idx(cx, cy, cz)] = Dirichlet;
        }
    }
    coarse.refresh();
}

}  // namespace ora
