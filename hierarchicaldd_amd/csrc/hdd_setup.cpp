#include "hdd_setup.h"
#include <cstring>

namespace hdd {

void divide_subdomains(PagedLevel& L, Int3 size, Int3 sub) {
    const Int3 ns = subdomain_counts(size, sub);
    for (int p = 0; p < L.npages(); ++p) {
        const Int3 o = L.page_base[p];
        uint32_t* f = &L.flags[(size_t)p * kPageCells];
        for (int c = 0; c < kPageCells; ++c) {
            if (!(f[c] & kActive)) continue;
            const int x = o.x + (c >> 6), y = o.y + ((c >> 3) & 7), z = o.z + (c & 7);
            const bool on_plane = (x % sub.x == 0) || (y % sub.y == 0) || (z % sub.z == 0);
            if (on_plane) {
                f[c] = (f[c] & ~kActive) | kInterface;
            } else {
                const int id = ((x / sub.x) * ns.y + y / sub.y) * ns.z + z / sub.z;
                f[c] |= (uint32_t)(id + 1);
            }
        }
    }
}

void set_face_flags(PagedLevel& L) {
    static const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    const uint32_t solid = kActive | kInterface | kDirichlet;
    for (int p = 0; p < L.npages(); ++p) {
        const Int3 o = L.page_base[p];
        for (int c = 0; c < kPageCells; ++c) {
            uint32_t& f = L.flags[(size_t)p * kPageCells + c];
            if (!(f & (kActive | kInterface))) continue;
            const int x = o.x + (c >> 6), y = o.y + ((c >> 3) & 7), z = o.z + (c & 7);
            for (int k = 0; k < 6; ++k)
                if (L.flag(x + d[k][0], y + d[k][1], z + d[k][2]) & solid) f |= kFaceMX << k;
        }
    }
}

// The reference computes, per page, a distance field over a padded window of
// x in [-4,8], y,z in [-4,12] (page-local coordinates), where window index x=8
// reads cell x=0 of the +x neighbour page (its cell index wraps with %4).
// Three in-place relaxation sweeps over [-3,7]x[-3,11]x[-3,11] (x outer, z
// inner) follow; an Active cell is Boundary when its value is <= 3.
void set_boundary_flags(PagedLevel& L) {
    constexpr int WX = 13, WY = 17, WZ = 17;
    std::vector<uint32_t> out(L.flags);
    std::vector<uint16_t> dist(WX * WY * WZ);
    auto at = [&](int i, int j, int k) -> uint16_t& { return dist[((i + 4) * WY + (j + 4)) * WZ + (k + 4)]; };
    for (int p = 0; p < L.npages(); ++p) {
        const int bx = L.page_base[p].x >> 2, by = L.page_base[p].y >> 3, bz = L.page_base[p].z >> 3;
        int nb[27];
        for (int a = 0; a < 27; ++a) nb[a] = L.page_at_block(bx + a / 9 - 1, by + (a / 3) % 3 - 1, bz + a % 3 - 1);
        for (int i = -4; i <= 8; ++i) {
            const int si = i < 0 ? 0 : (i < kPX ? 1 : 2), ci = (i + kPX) % kPX;
            for (int j = -4; j <= 12; ++j) {
                const int sj = j < 0 ? 0 : (j < kPY ? 1 : 2), cj = (j + kPY) % kPY;
                for (int k = -4; k <= 12; ++k) {
                    const int sk = k < 0 ? 0 : (k < kPZ ? 1 : 2), ck = (k + kPZ) % kPZ;
                    const int q = nb[si * 9 + sj * 3 + sk];
                    const uint32_t f = q < 0 ? 0u : L.flags[(size_t)q * kPageCells + ((ci << 6) | (cj << 3) | ck)];
                    at(i, j, k) = (f == 0u || (f & (kInterface | kDirichlet))) ? 0 : 0xffff;
                }
            }
        }
        for (int sweep = 0; sweep < 3; ++sweep)
            for (int i = -3; i <= 7; ++i)
                for (int j = -3; j <= 11; ++j)
                    for (int k = -3; k <= 11; ++k) {
                        uint16_t& v = at(i, j, k);
                        if (!v) continue;
                        unsigned m = v;
                        m = std::min<unsigned>(m, at(i + 1, j, k) + 1u);
                        m = std::min<unsigned>(m, at(i, j + 1, k) + 1u);
                        m = std::min<unsigned>(m, at(i, j, k + 1) + 1u);
                        m = std::min<unsigned>(m, at(i - 1, j, k) + 1u);
                        m = std::min<unsigned>(m, at(i, j - 1, k) + 1u);
                        m = std::min<unsigned>(m, at(i, j, k - 1) + 1u);
                        v = (uint16_t)m;
                    }
        uint32_t* f = &out[(size_t)p * kPageCells];
        for (int c = 0; c < kPageCells; ++c)
            if ((f[c] & kActive) && at(c >> 6, (c >> 3) & 7, c & 7) <= 3) f[c] |= kBoundary;
    }
    L.flags.swap(out);
}

void downsample_level(const PagedLevel& fine, PagedLevel& coarse) {
    const uint32_t any = kDirichlet | kActive | kInterface;
    for (int p = 0; p < fine.npages(); ++p) {
        const Int3 o = fine.page_base[p];
        for (int c = 0; c < kPageCells; ++c) {
            const uint32_t ff = fine.flags[(size_t)p * kPageCells + c];
            const int cx = (o.x + (c >> 6) + 1) / 2, cy = (o.y + ((c >> 3) & 7) + 1) / 2, cz = (o.z + (c & 7) + 1) / 2;
            const int cp = coarse.page_at(cx, cy, cz);
            uint32_t* cf = cp < 0 ? nullptr : &coarse.flags[(size_t)cp * kPageCells + page_cell(cx, cy, cz)];
            if (!cf || !(*cf & any)) {
                uint32_t m = (ff & kActive) ? kActive : ((ff & (kDirichlet | kInterface)) ? kDirichlet : 0u);
                if (m) {
                    const int q = coarse.add_page(cx >> 2, cy >> 3, cz >> 3);
                    coarse.flags[(size_t)q * kPageCells + page_cell(cx, cy, cz)] |= m;
                }
            } else if (ff & (kDirichlet | kInterface)) {
                *cf = kDirichlet;
            }
        }
    }
    coarse.sort_pages();
}

}  // namespace hdd
