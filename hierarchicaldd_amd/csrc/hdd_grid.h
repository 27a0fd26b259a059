// Host-side SPGrid page pool of the product.
//
// Pages keep the reference's geometry (SPGrid/Core/SPGrid_Mask.h:20-140 with a
// 16-byte struct): 4(x) x 8(y) x 8(z) cells per page, z fastest inside a page,
// pages ordered by the Morton interleave of page coordinates (x lowest, then
// z, then y). On the device every channel is a separate array of npages*256
// values, page p's cell c at index p*256+c, and one extra all-zero page at
// index npages stands in for unallocated neighbours (the role the reference's
// CUDA linearizer gives its 0xdeadbeef page).
#pragma once
#include <cstdint>
#include <vector>
#include <algorithm>
#include <stdexcept>
#include <string>

namespace hdd {

enum : uint32_t {
    kActive = 0x10000u, kDirichlet = 0x20000u, kBoundary = 0x40000u, kInterface = 0x80000u,
    kFaceMX = 0x100000u,  // face f (order -x,+x,-y,+y,-z,+z) is kFaceMX << f
    kPartitionMask = 0xffffu,
};
constexpr int kPX = 4, kPY = 8, kPZ = 8, kPageCells = 256;

struct Int3 { int x = 0, y = 0, z = 0; };

inline int page_cell(int x, int y, int z) { return ((x & 3) << 6) | ((y & 7) << 3) | (z & 7); }

inline uint64_t spread3(uint32_t v) {  // bit k of v -> bit 3k
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}
inline uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffff;
    return (uint32_t)x;
}
inline uint64_t page_key(int bx, int by, int bz) {
    return spread3((uint32_t)bx) | (spread3((uint32_t)bz) << 1) | (spread3((uint32_t)by) << 2);
}
inline void page_coords(uint64_t key, int& bx, int& by, int& bz) {
    bx = (int)compact3(key); bz = (int)compact3(key >> 1); by = (int)compact3(key >> 2);
}
// SPGrid byte offset of a 4-byte channel element, used as the traversal order key
inline uint64_t spgrid_offset(int x, int y, int z) {
    return (page_key(x >> 2, y >> 3, z >> 3) << 12) | ((uint64_t)page_cell(x, y, z) << 2);
}

// One level of the paged grid: page list in Morton order, block lookup table
// over the level's bounding box, flags per cell.
struct PagedLevel {
    Int3 size;                    // cells per axis of this level
    int nbx = 0, nby = 0, nbz = 0;
    std::vector<int> lut;         // [nbx*nby*nbz] -> page or -1
    std::vector<Int3> page_base;  // cell coordinates of each page's origin
    std::vector<uint32_t> flags;  // [npages*256]

    int npages() const { return (int)page_base.size(); }
    void init(Int3 s) {
        size = s;
        nbx = (s.x + kPX - 1) / kPX + 1; nby = (s.y + kPY - 1) / kPY + 1; nbz = (s.z + kPZ - 1) / kPZ + 1;
        lut.assign((size_t)nbx * nby * nbz, -1);
        page_base.clear(); flags.clear();
    }
    int page_at_block(int bx, int by, int bz) const {
        if (bx < 0 || by < 0 || bz < 0 || bx >= nbx || by >= nby || bz >= nbz) return -1;
        return lut[((size_t)bx * nby + by) * nbz + bz];
    }
    int page_at(int x, int y, int z) const {
        if (x < 0 || y < 0 || z < 0) return -1;
        return page_at_block(x >> 2, y >> 3, z >> 3);
    }
    uint32_t flag(int x, int y, int z) const {
        int p = page_at(x, y, z);
        return p < 0 ? 0u : flags[(size_t)p * kPageCells + page_cell(x, y, z)];
    }
    int add_page(int bx, int by, int bz) {
        if (bx < 0 || by < 0 || bz < 0 || bx >= nbx || by >= nby || bz >= nbz)
            throw std::runtime_error("page outside the level's block box");
        int& e = lut[((size_t)bx * nby + by) * nbz + bz];
        if (e < 0) {
            e = (int)page_base.size();
            page_base.push_back({bx * kPX, by * kPY, bz * kPZ});
            flags.resize(flags.size() + kPageCells, 0u);
        }
        return e;
    }
    // reorder pages by Morton key, keeping flags attached
    void sort_pages() {
        const int n = npages();
        std::vector<std::pair<uint64_t, int>> k(n);
        for (int i = 0; i < n; ++i) k[i] = {page_key(page_base[i].x >> 2, page_base[i].y >> 3, page_base[i].z >> 3), i};
        std::sort(k.begin(), k.end());
        std::vector<Int3> nb(n);
        std::vector<uint32_t> nf((size_t)n * kPageCells);
        for (int i = 0; i < n; ++i) {
            int o = k[i].second;
            nb[i] = page_base[o];
            std::copy(flags.begin() + (size_t)o * kPageCells, flags.begin() + (size_t)(o + 1) * kPageCells, nf.begin() + (size_t)i * kPageCells);
        }
        page_base.swap(nb); flags.swap(nf);
        std::fill(lut.begin(), lut.end(), -1);
        for (int i = 0; i < n; ++i) lut[((size_t)(page_base[i].x >> 2) * nby + (page_base[i].y >> 3)) * nbz + (page_base[i].z >> 3)] = i;
    }
};

}  // namespace hdd
