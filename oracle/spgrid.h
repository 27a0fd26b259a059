// ORACLE — test infrastructure only. CPU restatement of the reference's SPGrid
// page layout, used solely as the checker by tests/, smoke() and bench.py's
// cpu_baseline leg. Nothing in the product links or calls this code.
//
// Layout restated from SPGrid/Core/SPGrid_Mask.h:20-140 (log2_struct=4,
// 4-byte fields): a page (block) holds 4(x) x 8(y) x 8(z) cells, z fastest
// inside the block (element_zmask lowest bits), and pages are ordered by the
// Morton interleave of the page coordinates with x in the lowest page bit,
// then z, then y (page_xmask/page_zmask/page_ymask with lob=1).
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <stdexcept>

namespace ora {

// SPGrid_Fluids/Solvers/Domain_Decomposition/SPGRID_MULTIGRID_FLAGS.h:9-31
enum : uint32_t {
    Active = 0x10000u, Dirichlet = 0x20000u, Boundary = 0x40000u, Interface = 0x80000u,
    FaceMX = 0x100000u,  // faces in order (-x,+x,-y,+y,-z,+z): FaceMX<<f
    PartitionMask = 0xffffu,
};

constexpr int BX = 4, BY = 8, BZ = 8, BCELLS = 256;

inline int cell_in_block(int x, int y, int z) { return ((x & 3) << 6) | ((y & 7) << 3) | (z & 7); }

// Morton page index of block coordinates (SPGrid_Mask.h:34-38, lob=1 -> x,z,y).
inline uint64_t morton_page(uint32_t bx, uint32_t by, uint32_t bz) {
    uint64_t m = 0;
    for (int k = 0; k < 21; ++k) {
        m |= (uint64_t)((bx >> k) & 1u) << (3 * k);
        m |= (uint64_t)((bz >> k) & 1u) << (3 * k + 1);
        m |= (uint64_t)((by >> k) & 1u) << (3 * k + 2);
    }
    return m;
}
inline void morton_decode(uint64_t m, int& bx, int& by, int& bz) {
    bx = by = bz = 0;
    for (int k = 0; k < 21; ++k) {
        bx |= (int)((m >> (3 * k)) & 1u) << k;
        bz |= (int)((m >> (3 * k + 1)) & 1u) << k;
        by |= (int)((m >> (3 * k + 2)) & 1u) << k;
    }
}
// Full SPGrid linear offset of a 4-byte channel element (flags channel):
// page index << 12 | cell << 2.  Used for offset-order sorting.
inline uint64_t linear_offset(int x, int y, int z) {
    return (morton_page((uint32_t)x >> 2, (uint32_t)y >> 3, (uint32_t)z >> 3) << 12) | ((uint64_t)cell_in_block(x, y, z) << 2);
}

// A sparse paged grid: pool of active 4x8x8 pages with a dense block lookup
// over a block box, `nch` channels of T per cell plus a flags channel and an
// int channel. Page activation mirrors SPGrid_Set::Mask/MarkPageActive
// (SPGrid/Core/SPGrid_Set.h); the block list mirrors Refresh_Block_Offsets.
template <class T>
struct Grid {
    int bx0 = 0, by0 = 0, bz0 = 0, bnx = 0, bny = 0, bnz = 0;
    int nch = 0;
    std::vector<int> lut;
    std::vector<int> pbx, pby, pbz;
    std::vector<uint32_t> flags;
    std::vector<int> ivals;
    std::vector<T> data;            // [ch][page*256+cell], stored as separate vectors
    std::vector<std::vector<T>> ch;
    std::vector<int> blocks;        // pool ids sorted by Morton page index
    bool dirty = false;
    bool with_ints = false;

    Grid() {}
    // box in cell coordinates [lo, hi) (rounded out to blocks)
    void init(int lox, int loy, int loz, int hix, int hiy, int hiz, int nchan, bool ints = false) {
        auto fl = [](int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); };
        bx0 = std::max(0, fl(lox, BX)); by0 = std::max(0, fl(loy, BY)); bz0 = std::max(0, fl(loz, BZ));
        int bx1 = fl(hix - 1, BX) + 1, by1 = fl(hiy - 1, BY) + 1, bz1 = fl(hiz - 1, BZ) + 1;
        bnx = std::max(0, bx1 - bx0); bny = std::max(0, by1 - by0); bnz = std::max(0, bz1 - bz0);
        lut.assign((size_t)bnx * bny * bnz, -1);
        nch = nchan; with_ints = ints;
        ch.assign(nch, std::vector<T>());
        pbx.clear(); pby.clear(); pbz.clear(); flags.clear(); ivals.clear(); blocks.clear();
        dirty = false;
    }
    int npages() const { return (int)pbx.size(); }
    inline int page_of_block(int bx, int by, int bz) const {
        int ix = bx - bx0, iy = by - by0, iz = bz - bz0;
        if (ix < 0 || iy < 0 || iz < 0 || ix >= bnx || iy >= bny || iz >= bnz) return -1;
        return lut[((size_t)ix * bny + iy) * bnz + iz];
    }
    inline int page_of(int x, int y, int z) const {
        if (x < 0 || y < 0 || z < 0) return -1;
        return page_of_block(x >> 2, y >> 3, z >> 3);
    }
    int activate_block(int bx, int by, int bz) {
        int ix = bx - bx0, iy = by - by0, iz = bz - bz0;
        if (ix < 0 || iy < 0 || iz < 0 || ix >= bnx || iy >= bny || iz >= bnz)
            throw std::runtime_error("oracle grid: page activation outside the allocated box");
        int& e = lut[((size_t)ix * bny + iy) * bnz + iz];
        if (e >= 0) return e;
        e = (int)pbx.size();
        pbx.push_back(bx); pby.push_back(by); pbz.push_back(bz);
        flags.resize(flags.size() + BCELLS, 0u);
        if (with_ints) ivals.resize(ivals.size() + BCELLS, 0);
        for (auto& c : ch) c.resize(c.size() + BCELLS, T(0));
        dirty = true;
        return e;
    }
    int activate(int x, int y, int z) { return activate_block(x >> 2, y >> 3, z >> 3); }
    // SPGrid_Set::Mask(coord, mask): activate page and OR the mask into flags.
    void mask(int x, int y, int z, uint32_t m) {
        int p = activate(x, y, z);
        flags[(size_t)p * BCELLS + cell_in_block(x, y, z)] |= m;
    }
    // SPGrid_Set::Is_Set(coord, mask)
    inline bool is_set(int x, int y, int z, uint32_t m) const {
        int p = page_of(x, y, z);
        if (p < 0) return false;
        return (flags[(size_t)p * BCELLS + cell_in_block(x, y, z)] & m) != 0;
    }
    inline uint32_t flag_at(int x, int y, int z) const {
        int p = page_of(x, y, z);
        return p < 0 ? 0u : flags[(size_t)p * BCELLS + cell_in_block(x, y, z)];
    }
    inline size_t idx(int x, int y, int z) const {
        int p = page_of(x, y, z);
        return p < 0 ? (size_t)-1 : (size_t)p * BCELLS + cell_in_block(x, y, z);
    }
    // value of channel c at a coordinate; unallocated pages read zero (mmap'd zero page)
    inline T val(int c, int x, int y, int z) const {
        int p = page_of(x, y, z);
        return p < 0 ? T(0) : ch[c][(size_t)p * BCELLS + cell_in_block(x, y, z)];
    }
    void refresh() {
        std::vector<std::pair<uint64_t, int>> v(pbx.size());
        for (size_t i = 0; i < pbx.size(); ++i) v[i] = {morton_page(pbx[i], pby[i], pbz[i]), (int)i};
        std::sort(v.begin(), v.end());
        blocks.resize(v.size());
        for (size_t i = 0; i < v.size(); ++i) blocks[i] = v[i].second;
        dirty = false;
    }
    // base cell coordinate of pool page p
    inline void base(int p, int& x, int& y, int& z) const { x = pbx[p] * BX; y = pby[p] * BY; z = pbz[p] * BZ; }
};

// neighbor offsets for faces (-x,+x,-y,+y,-z,+z): GRID_TOPOLOGY_HELPER.h:126-134
constexpr int FDX[6] = {-1, 1, 0, 0, 0, 0};
constexpr int FDY[6] = {0, 0, -1, 1, 0, 0};
constexpr int FDZ[6] = {0, 0, 0, 0, -1, 1};

}  // namespace ora
