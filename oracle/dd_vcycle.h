// ORACLE — test infrastructure only (see spgrid.h header).
// Matrix-free V-cycle kernels of the reference, restated on the oracle Grid.
// Channel roles on every level: 0 = u (result z), 1 = b (rhs r), 2 = r (tmp k).
//   Residual      LIB/SPGrid_Residual_Legacy.h:17-62     (float accumulation)
//   Correction    LIB/SPGrid_Correction_Legacy.h:19-79   (omega = 2/3)
//   Restriction   LIB/SPGrid_Restriction_Legacy.h:20-187 (w = {1,3,3,1}/8)
//   Prolongation  LIB/SPGrid_Prolongation_Legacy.h:19-129 (w = 4*{3/4,1/4}^3)
//   v_cycle       LIB/SPGrid_V_Cycle_Helper.h:156-192
// (LIB = demos/Domain_Decomposition_CPU_Only/SPGrid_Library)
// Every kernel is a Jacobi-type pass (reads one channel, writes another, or
// updates a cell from its own values only), so the per-block loops may run in
// parallel without changing any result.
#pragma once
#include "dd_setup.h"
#include <cmath>

namespace ora {

enum { CH_U = 0, CH_B = 1, CH_R = 2 };

template <class T>
void vc_residual(Grid<T>& g) {
    const int n = (int)g.blocks.size();
#pragma omp parallel for schedule(static)
    for (int ib = 0; ib < n; ++ib) {
        int p = g.blocks[ib];
        int X, Y, Z; g.base(p, X, Y, Z);
        for (int c = 0; c < BCELLS; ++c) {
            size_t e = (size_t)p * BCELLS + c;
            uint32_t f = g.flags[e];
            if (!(f & Active)) continue;
            int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
            T cv = g.ch[CH_U][e], result = 0;
            for (int fc = 0; fc < 6; ++fc)
                if (f & (FaceMX << fc)) {
                    T nv = g.val(CH_U, x + FDX[fc], y + FDY[fc], z + FDZ[fc]);
                    result -= (nv - cv);
                }
            g.ch[CH_R][e] = g.ch[CH_B][e] - result;
        }
    }
}

// which: 0 = interior only (Active & !Boundary), 1 = boundary only, 2 = all Active
template <class T>
void vc_correction(Grid<T>& g, bool write_interior, bool write_boundary) {
    const T omega = (T)(2. / 3.);
    const int n = (int)g.blocks.size();
#pragma omp parallel for schedule(static)
    for (int ib = 0; ib < n; ++ib) {
        int p = g.blocks[ib];
        for (int c = 0; c < BCELLS; ++c) {
            size_t e = (size_t)p * BCELLS + c;
            uint32_t f = g.flags[e];
            uint32_t sel = f & (Active | Boundary);
            if (sel == Active) { if (!write_interior) continue; }
            else if (sel == (Active | Boundary)) { if (!write_boundary) continue; }
            else if (sel == Boundary) throw std::runtime_error("Boundary without Active");
            else continue;
            int fc = 0;
            for (int k = 0; k < 6; ++k) if (f & (FaceMX << k)) ++fc;
            if (fc <= 0) throw std::runtime_error("face_count == 0");
            g.ch[CH_U][e] += (omega / (T)fc) * g.ch[CH_R][e];
        }
    }
}

template <class T>
void vc_clear(Grid<T>& g, int chn) {
    for (int p : g.blocks) std::fill(g.ch[chn].begin() + (size_t)p * BCELLS, g.ch[chn].begin() + (size_t)(p + 1) * BCELLS, T(0));
}

template <class T>
void vc_restriction(const Grid<T>& fine, Grid<T>& coarse) {
    float s1[4] = {(float)(1. / 8.), (float)(3. / 8.), (float)(3. / 8.), (float)(1. / 8.)};
    float w[4][4][4];
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) for (int k = 0; k < 4; ++k) w[i][j][k] = s1[i] * s1[j] * s1[k];
    const int n = (int)coarse.blocks.size();
#pragma omp parallel for schedule(static)
    for (int ib = 0; ib < n; ++ib) {
        int p = coarse.blocks[ib];
        int X, Y, Z; coarse.base(p, X, Y, Z);
        for (int c = 0; c < BCELLS; ++c) {
            size_t e = (size_t)p * BCELLS + c;
            if (!(coarse.flags[e] & Active)) continue;
            int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
            int fx = 2 * x - 1, fy = 2 * y - 1, fz = 2 * z - 1;
            T acc = 0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    for (int k = 0; k < 4; ++k)
                        acc += (T)w[i][j][k] * fine.val(CH_R, fx + i - 1, fy + j - 1, fz + k - 1);
            coarse.ch[CH_B][e] = acc;
        }
    }
}

template <class T>
void vc_prolongation(Grid<T>& fine, const Grid<T>& coarse) {
    float s1[2] = {0.75f, 0.25f};
    float w[2][2][2];
    for (int i = 0; i < 2; ++i) for (int j = 0; j < 2; ++j) for (int k = 0; k < 2; ++k) w[i][j][k] = s1[i] * s1[j] * s1[k] * 4.0f;
    const int n = (int)fine.blocks.size();
#pragma omp parallel for schedule(static)
    for (int ib = 0; ib < n; ++ib) {
        int p = fine.blocks[ib];
        int X, Y, Z; fine.base(p, X, Y, Z);
        for (int c = 0; c < BCELLS; ++c) {
            size_t e = (size_t)p * BCELLS + c;
            if (!(fine.flags[e] & Active)) continue;
            int x = X + (c >> 6), y = Y + ((c >> 3) & 7), z = Z + (c & 7);
            int cx = (x + 1) / 2, cy = (y + 1) / 2, cz = (z + 1) / 2;
            int px = 1 - (x - (2 * cx - 1)), py = 1 - (y - (2 * cy - 1)), pz = 1 - (z - (2 * cz - 1));
            T u = fine.ch[CH_U][e];
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    for (int k = 0; k < 2; ++k)
                        u += (T)w[i ^ px][j ^ py][k ^ pz] * coarse.val(CH_U, cx + i - px, cy + j - py, cz + k - pz);
            fine.ch[CH_U][e] = u;
        }
    }
}

template <class T>
struct VTopology {
    std::vector<Grid<T>*> lv;  // lv[0] = main grid (not owned)
    std::vector<I3> sizes;
    ~VTopology() { for (size_t i = 1; i < lv.size(); ++i) delete lv[i]; }
};

// V_Cycle_Topology ctor + Initialize + Flaging (DDS/SPGrid_V_Cycle_Topology.h:102-177)
template <class T>
void build_topology(VTopology<T>& topo, Grid<T>& top, I3 size, int levels) {
    topo.lv.assign(levels, nullptr);
    topo.sizes.assign(levels, size);
    topo.lv[0] = &top;
    for (int l = 0; l + 1 < levels; ++l) {
        I3 s = topo.sizes[l];
        topo.sizes[l + 1] = {s.x / 2 + 1, s.y / 2 + 1, s.z / 2 + 1};
        Grid<T>* g = new Grid<T>();
        I3 t = topo.sizes[l + 1];
        g->init(0, 0, 0, t.x + BX, t.y + BY, t.z + BZ, 3);
        topo.lv[l + 1] = g;
    }
    for (int l = 0; l + 1 < levels; ++l) topology_downsample(*topo.lv[l], *topo.lv[l + 1]);
    for (int l = 0; l < levels; ++l) { flaging(*topo.lv[l]); flaging_boundary(*topo.lv[l]); }
}

template <class T>
void smooth(Grid<T>& g, bool interior, bool boundary) { vc_residual(g); vc_correction(g, interior, boundary); }

// V_Cycle_Helper::v_cycle (LIB/SPGrid_V_Cycle_Helper.h:156-192)
template <class T>
void v_cycle(VTopology<T>& topo, int interior_itr, int boundary_itr) {
    const int L = (int)topo.lv.size();
    for (int l = 0; l < L - 1; ++l) {
        Grid<T>& f = *topo.lv[l]; Grid<T>& c = *topo.lv[l + 1];
        for (int i = 0; i < boundary_itr; ++i) smooth(f, false, true);
        for (int i = 0; i < interior_itr; ++i) smooth(f, true, false);
        for (int i = 0; i < boundary_itr; ++i) smooth(f, false, true);
        vc_residual(f);
        vc_clear(c, CH_U);
        vc_restriction(f, c);
    }
    Grid<T>& bot = *topo.lv[L - 1];
    vc_residual(bot);
    for (int i = 0; i < 200; ++i) smooth(bot, true, true);
    vc_residual(bot);
    for (int l = L - 2; l >= 0; --l) {
        Grid<T>& f = *topo.lv[l]; Grid<T>& c = *topo.lv[l + 1];
        vc_prolongation(f, c);
        for (int i = 0; i < boundary_itr; ++i) smooth(f, false, true);
        for (int i = 0; i < interior_itr; ++i) smooth(f, true, false);
        for (int i = 0; i < boundary_itr; ++i) smooth(f, false, true);
    }
}

}  // namespace ora
