// grid.cpp -- host-side uniform voxel grid of a triangulated scene.
//
// Stands in for Polygon_Scene.partition()'s Hare Voxel_Grid (Classes_Scenes.cs:
// 709-755): per-voxel triangle lists (exact triangle/box overlap), plus an L-inf
// distance field that lets the GPU walk skip empty space.
#include "grid.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <deque>
#include <thread>

namespace pach {

// Triangle / axis-aligned box overlap by the separating-axis theorem
// (Akenine-Moller 2001).  bc: box centre, bh: half extents, tv: 9 coordinates.
bool tri_box_overlap(const double *bc, const double *bh, const double *tv)
{
    double v[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) v[i][k] = tv[3 * i + k] - bc[k];
    const double e[3][3] = {{v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]},
                            {v[2][0] - v[1][0], v[2][1] - v[1][1], v[2][2] - v[1][2]},
                            {v[0][0] - v[2][0], v[0][1] - v[2][1], v[0][2] - v[2][2]}};
    // nine axes: box axis x triangle edge
    for (int i = 0; i < 3; ++i) {
        for (int a = 0; a < 3; ++a) {
            double u[3] = {0, 0, 0};
            u[a] = 1;
            const double ax[3] = {u[1] * e[i][2] - u[2] * e[i][1], u[2] * e[i][0] - u[0] * e[i][2],
                                  u[0] * e[i][1] - u[1] * e[i][0]};
            double mn = 1e300, mx = -1e300;
            for (int j = 0; j < 3; ++j) {
                const double p = v[j][0] * ax[0] + v[j][1] * ax[1] + v[j][2] * ax[2];
                mn = std::min(mn, p);
                mx = std::max(mx, p);
            }
            const double rad = bh[0] * std::fabs(ax[0]) + bh[1] * std::fabs(ax[1]) + bh[2] * std::fabs(ax[2]);
            if (mn > rad || mx < -rad) return false;
        }
    }
    // three box face normals
    for (int k = 0; k < 3; ++k) {
        const double mn = std::min({v[0][k], v[1][k], v[2][k]}), mx = std::max({v[0][k], v[1][k], v[2][k]});
        if (mn > bh[k] || mx < -bh[k]) return false;
    }
    // triangle plane
    const double n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                         e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    const double d = -(n[0] * v[0][0] + n[1] * v[0][1] + n[2] * v[0][2]);
    const double r = bh[0] * std::fabs(n[0]) + bh[1] * std::fabs(n[1]) + bh[2] * std::fabs(n[2]);
    return std::fabs(d) <= r;
}

// L-inf distance transform: breadth-first search from the voxels marked 0
// over the 26-neighbourhood; every other voxel gets its distance, capped at 255.
void linf_distance(const int G[3], std::vector<uint8_t> &dist)
{
    const size_t ncell = (size_t)G[0] * G[1] * G[2];
    std::deque<uint32_t> q;
    for (size_t i = 0; i < ncell; ++i) {
        if (dist[i] == 0) q.push_back((uint32_t)i);
        else dist[i] = 255;
    }
    while (!q.empty()) {
        const uint32_t i = q.front();
        q.pop_front();
        const int x = (int)(i % G[0]), y = (int)((i / G[0]) % G[1]), z = (int)(i / ((size_t)G[0] * G[1]));
        const int nd = dist[i] + 1;
        if (nd > 255) continue;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int X = x + dx, Y = y + dy, Z = z + dz;
                    if (X < 0 || Y < 0 || Z < 0 || X >= G[0] || Y >= G[1] || Z >= G[2]) continue;
                    const size_t jn = ((size_t)Z * G[1] + Y) * G[0] + X;
                    if (dist[jn] > nd) {
                        dist[jn] = (uint8_t)nd;
                        q.push_back((uint32_t)jn);
                    }
                }
    }
}

void receiver_cells(const Grid &g, const double *centers, int R, const int *bit, double radius,
                    std::vector<uint64_t> &mask, std::vector<uint8_t> &dist)
{
    // voxels whose box comes within radius (+ 1e-3 voxel) of a centre: the walk
    // collects their receiver bits, and they count as occupied for the jumps
    const size_t ncell = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    mask.assign(ncell, 0);
    dist.assign(g.dist.begin(), g.dist.end());
    for (size_t i = 0; i < ncell; ++i)
        if (g.ranges[2 * i + 1] == 0) dist[i] = 255;
    const double pad = 1e-3 * std::max({g.cell[0], g.cell[1], g.cell[2]});
    const double rr = radius + pad;
    for (int r = 0; r < R; ++r) {
        if (bit[r] < 0 || bit[r] > 63) continue;
        const double *C = centers + 3 * r;
        int lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::max(0, (int)std::floor((C[k] - rr - g.origin[k]) / g.cell[k]));
            hi[k] = std::min(g.dims[k] - 1, (int)std::floor((C[k] + rr - g.origin[k]) / g.cell[k]));
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const int idx[3] = {x, y, z};
                    double d2 = 0;
                    for (int k = 0; k < 3; ++k) {
                        const double b0 = g.origin[k] + idx[k] * g.cell[k], b1 = b0 + g.cell[k];
                        const double q = C[k] < b0 ? b0 - C[k] : (C[k] > b1 ? C[k] - b1 : 0.0);
                        d2 += q * q;
                    }
                    if (d2 <= rr * rr) {
                        const size_t i = ((size_t)z * g.dims[1] + y) * g.dims[0] + x;
                        mask[i] |= 1ull << bit[r];
                        dist[i] = 0;
                    }
                }
    }
    linf_distance(g.dims, dist);
}

void octant_cubes(const int dims[3], const std::vector<uint8_t> &dist, std::vector<uint8_t> &oct)
{
    const int gx = dims[0], gy = dims[1], gz = dims[2];
    const size_t ncell = (size_t)gx * gy * gz;
    oct.assign(ncell * 8, 0);
    for (int o = 0; o < 8; ++o) {
        const int sx = (o & 1) ? -1 : 1, sy = (o & 2) ? -1 : 1, sz = (o & 4) ? -1 : 1;
        // visit the cells farthest along the octant first: every cube a cell
        // builds on (one cell further along one, two or three axes) is done
        auto at = [&](int x, int y, int z) -> int {
            if (x < 0 || x >= gx || y < 0 || y >= gy || z < 0 || z >= gz) return 255;  // outside: open
            return oct[(((size_t)z * gy + y) * gx + x) * 8 + o];
        };
        for (int iz = 0; iz < gz; ++iz) {
            const int z = sz > 0 ? gz - 1 - iz : iz;
            for (int iy = 0; iy < gy; ++iy) {
                const int y = sy > 0 ? gy - 1 - iy : iy;
                for (int ix = 0; ix < gx; ++ix) {
                    const int x = sx > 0 ? gx - 1 - ix : ix;
                    const size_t c = ((size_t)z * gy + y) * gx + x;
                    if (dist[c] == 0) continue;  // occupied: 0
                    const int m = std::min({at(x + sx, y, z), at(x, y + sy, z), at(x, y, z + sz),
                                            at(x + sx, y + sy, z), at(x + sx, y, z + sz), at(x, y + sy, z + sz),
                                            at(x + sx, y + sy, z + sz)});
                    oct[c * 8 + o] = (uint8_t)std::min(m + 1, 255);
                }
            }
        }
    }
}

void octant_boxes(const int dims[3], const std::vector<uint8_t> &dist, const std::vector<uint8_t> &oct,
                  std::vector<uint8_t> &ext)
{
    const int gx = dims[0], gy = dims[1], gz = dims[2];
    const size_t ncell = (size_t)gx * gy * gz;
    // occupied voxels of every box [x0, x1) x [y0, y1) x [z0, z1) from one
    // table of prefix sums
    std::vector<int32_t> P((size_t)(gx + 1) * (gy + 1) * (gz + 1), 0);
    auto at = [&](int x, int y, int z) -> int32_t & { return P[((size_t)z * (gy + 1) + y) * (gx + 1) + x]; };
    for (int z = 1; z <= gz; ++z)
        for (int y = 1; y <= gy; ++y)
            for (int x = 1; x <= gx; ++x)
                at(x, y, z) = (dist[((size_t)(z - 1) * gy + (y - 1)) * gx + (x - 1)] == 0) + at(x - 1, y, z) +
                              at(x, y - 1, z) + at(x, y, z - 1) - at(x - 1, y - 1, z) - at(x - 1, y, z - 1) -
                              at(x, y - 1, z - 1) + at(x - 1, y - 1, z - 1);
    const int g[3] = {gx, gy, gz};
    // the cells c .. c + s (n - 1) along one axis as a half-open range clipped to the grid
    auto span = [&](int c, int s, int n, int a, int &lo, int &hi) {
        lo = s > 0 ? c : c - n + 1;
        hi = s > 0 ? c + n : c + 1;
        lo = std::max(lo, 0);
        hi = std::min(hi, g[a]);
    };
    auto empty = [&](const int c[3], const int s[3], const int n[3]) {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            span(c[a], s[a], n[a], a, lo[a], hi[a]);
            if (lo[a] >= hi[a]) return true;  // wholly outside the grid
        }
        const int32_t occ = at(hi[0], hi[1], hi[2]) - at(lo[0], hi[1], hi[2]) - at(hi[0], lo[1], hi[2]) -
                            at(hi[0], hi[1], lo[2]) + at(lo[0], lo[1], hi[2]) + at(lo[0], hi[1], lo[2]) +
                            at(hi[0], lo[1], lo[2]) - at(lo[0], lo[1], lo[2]);
        return occ == 0;
    };
    ext.assign(ncell * 24, 0);
    auto octant = [&](int o) {
        const int s[3] = {(o & 1) ? -1 : 1, (o & 2) ? -1 : 1, (o & 4) ? -1 : 1};
        for (int z = 0; z < gz; ++z)
            for (int y = 0; y < gy; ++y)
                for (int x = 0; x < gx; ++x) {
                    const size_t c = ((size_t)z * gy + y) * gx + x;
                    if (dist[c] == 0) continue;
                    const int k = oct[8 * c + o], cc[3] = {x, y, z};
                    // the longest empty box along axis a, the box n[] holding on the
                    // other axes (n[a] = k holds: a sub-box of an empty box)
                    auto longest = [&](int a, int n[3]) {
                        // cells to the grid's edge along a, the box's reach past it
                        const int edge = s[a] > 0 ? g[a] - cc[a] : cc[a] + 1;
                        if (k >= 255 || k >= edge) return 255;  // the cube already reaches past the grid
                        n[a] = edge;
                        if (empty(cc, s, n)) return 255;
                        // largest length in [k, edge) with an empty box: k holds, edge does not
                        int lo = k, hi = edge;
                        while (hi - lo > 1) {
                            const int m = (lo + hi) / 2;
                            n[a] = m;
                            if (empty(cc, s, n)) lo = m;
                            else hi = m;
                        }
                        return lo;
                    };
                    int e[3];
                    for (int a = 0; a < 3; ++a) {
                        int n[3] = {k, k, k};
                        e[a] = longest(a, n);
                    }
                    // the box: the axis with the longest single-axis box (the lowest
                    // axis on a tie), then the longer of the other two (the next
                    // axis on a tie) lengthened with the first at its length
                    int dom = 0;
                    for (int a = 1; a < 3; ++a)
                        if (e[a] > e[dom]) dom = a;
                    const int a1 = (dom + 1) % 3, a2 = (dom + 2) % 3;
                    const int sec = e[a1] >= e[a2] ? a1 : a2;
                    int n[3] = {k, k, k};
                    n[dom] = e[dom];
                    const int e2 = longest(sec, n);
                    ext[24 * c + o] = (uint8_t)e[dom];
                    ext[24 * c + 8 + o] = (uint8_t)e2;
                    ext[24 * c + 16 + o] = (uint8_t)(dom | sec << 2);
                }
    };
    std::vector<std::thread> th;
    for (int o = 0; o < 8; ++o) th.emplace_back(octant, o);
    for (auto &t : th) t.join();
}

void face_skip(const Grid &g, std::vector<uint32_t> &skip)
{
    const int gx = g.dims[0], gy = g.dims[1], gz = g.dims[2];
    const size_t ncell = (size_t)gx * gy * gz;
    skip.assign(ncell * 6, 0u);
    std::vector<uint32_t> nb;
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) {
                const size_t c = ((size_t)z * gy + y) * gx + x;
                const uint32_t first = g.ranges[2 * c], n = std::min<uint32_t>(g.ranges[2 * c + 1], 32);
                if (n == 0) continue;
                for (int f = 0; f < 6; ++f) {
                    int q[3] = {x, y, z};
                    q[f / 2] += (f & 1) ? 1 : -1;
                    if (q[f / 2] < 0 || q[f / 2] >= g.dims[f / 2]) continue;
                    const size_t cn = ((size_t)q[2] * gy + q[1]) * gx + q[0];
                    const uint32_t f0 = g.ranges[2 * cn], fn = g.ranges[2 * cn + 1];
                    if (fn == 0) continue;
                    nb.assign(g.tris.begin() + f0, g.tris.begin() + f0 + fn);
                    std::sort(nb.begin(), nb.end());
                    uint32_t m = 0;
                    for (uint32_t i = 0; i < n; ++i)
                        if (std::binary_search(nb.begin(), nb.end(), g.tris[first + i])) m |= 1u << i;
                    skip[6 * c + f] = m;
                }
            }
}

void cell_records(const Grid &g, const std::vector<uint32_t> &skip, const std::vector<uint64_t> &mask,
                  const std::vector<uint8_t> &dist, const std::vector<uint8_t> &oct, const std::vector<uint8_t> &ext,
                  std::vector<uint32_t> &rec)
{
    // word layout of one record: 0 first, 1 count, 2-7 skip masks by face,
    // 8-9 receiver mask (low, high), 10-11 octant edges (bytes 0-7), 12
    // distance, 13-15 unused.  An empty voxel has no list, skip masks or
    // receivers: words 0-1, 2-3 and 8-9 hold its boxes (ext [plane][octant],
    // one byte per octant: first axis length, second axis length, axes), so
    // the words the walk loads anyway carry them
    const size_t ncell = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    rec.assign(ncell * 16, 0u);
    for (size_t c = 0; c < ncell; ++c) {
        uint32_t *r = rec.data() + 16 * c;
        if (dist[c] == 0) {
            r[0] = g.ranges[2 * c];
            r[1] = g.ranges[2 * c + 1];
            for (int f = 0; f < 6; ++f) r[2 + f] = skip[6 * c + f];
            const uint64_t m = mask.empty() ? 0 : mask[c];
            r[8] = (uint32_t)m;
            r[9] = (uint32_t)(m >> 32);
        } else {
            static const int word[3] = {0, 2, 8};
            for (int a = 0; a < 3; ++a)
                for (int o = 0; o < 8; ++o)
                    r[word[a] + o / 4] |= (uint32_t)ext[24 * c + 8 * a + o] << (8 * (o % 4));
        }
        for (int o = 0; o < 8; ++o) r[10 + o / 4] |= (uint32_t)oct[8 * c + o] << (8 * (o % 4));
        r[12] = dist[c];
    }
}

int build_grid(const double *verts, int T, const int dims_in[3], Grid &g, std::string &err)
{
    if (!verts || T <= 0) { err = "grid: empty scene"; return -1; }
    double bmin[3] = {1e300, 1e300, 1e300}, bmax[3] = {-1e300, -1e300, -1e300};
    for (int t = 0; t < T; ++t)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                const double x = verts[9 * t + 3 * j + k];
                if (!std::isfinite(x)) { err = "grid: non-finite vertex"; return -1; }
                bmin[k] = std::min(bmin[k], x);
                bmax[k] = std::max(bmax[k], x);
            }
    double ext[3];
    for (int k = 0; k < 3; ++k) ext[k] = std::max(bmax[k] - bmin[k], 1e-6);
    int G[3] = {dims_in ? dims_in[0] : 0, dims_in ? dims_in[1] : 0, dims_in ? dims_in[2] : 0};
    if (G[0] <= 0 || G[1] <= 0 || G[2] <= 0) {
        // about 8 voxels per triangle, 16 from 8,192 triangles on (large scenes
        // trade more voxel steps for fewer triangle tests; small ones keep
        // their distance field small enough for LDS), shaped like the
        // bounding box
        const double per = T >= 8192 ? 16.0 : 8.0;
        const double target = std::min(std::max(per * T, 4096.0), 4.0e6);
        const double unit = std::cbrt(ext[0] * ext[1] * ext[2] / target);
        for (int k = 0; k < 3; ++k) G[k] = std::max(1, std::min(256, (int)std::lround(ext[k] / unit)));
    }
    for (int k = 0; k < 3; ++k) {
        if (G[k] > 1024) { err = "grid: dimension above 1024"; return -1; }
        const double pad = 1e-4 * std::max(ext[k], 1.0);
        g.dims[k] = G[k];
        g.origin[k] = bmin[k] - pad;
        g.cell[k] = (ext[k] + 2 * pad) / G[k];
    }
    const size_t ncell = (size_t)G[0] * G[1] * G[2];
    if (ncell > (size_t)1 << 27) { err = "grid: too many voxels"; return -1; }
    std::vector<std::vector<uint32_t>> lists(ncell);
    for (int t = 0; t < T; ++t) {
        const double *p = verts + 9 * t;
        int lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            const double mn = std::min({p[k], p[3 + k], p[6 + k]}), mx = std::max({p[k], p[3 + k], p[6 + k]});
            const double e = 1e-3 * g.cell[k];
            lo[k] = std::max(0, (int)std::floor((mn - e - g.origin[k]) / g.cell[k]));
            hi[k] = std::min(G[k] - 1, (int)std::floor((mx + e - g.origin[k]) / g.cell[k]));
        }
        // boxes are inflated by 0.1 % so a triangle grazing a voxel face is listed
        // in both voxels and no FP32 walk can slip between them
        const double bh[3] = {0.5 * g.cell[0] * (1 + 2e-3), 0.5 * g.cell[1] * (1 + 2e-3), 0.5 * g.cell[2] * (1 + 2e-3)};
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const double bc[3] = {g.origin[0] + (x + 0.5) * g.cell[0], g.origin[1] + (y + 0.5) * g.cell[1],
                                          g.origin[2] + (z + 0.5) * g.cell[2]};
                    if (tri_box_overlap(bc, bh, p)) lists[((size_t)z * G[1] + y) * G[0] + x].push_back((uint32_t)t);
                }
    }
    g.ranges.assign(2 * ncell, 0);
    g.tris.clear();
    for (size_t i = 0; i < ncell; ++i) {
        if (lists[i].size() > 0xFFFFFFFFu) { err = "grid: voxel list too long"; return -1; }
        g.ranges[2 * i] = (uint32_t)g.tris.size();
        g.ranges[2 * i + 1] = (uint32_t)lists[i].size();
        g.tris.insert(g.tris.end(), lists[i].begin(), lists[i].end());
    }
    g.dist.assign(ncell, 255);
    for (size_t i = 0; i < ncell; ++i)
        if (g.ranges[2 * i + 1] > 0) g.dist[i] = 0;
    linf_distance(g.dims, g.dist);
    // a grid with no occupied voxel keeps 255 everywhere: every walk is one jump
    g.dist.resize((ncell + 15) & ~(size_t)15, 255);
    return 0;
}

}  // namespace pach

extern "C" int pach_grid_build(const double *verts, int T, int dims[3], double origin[3], double cell[3],
                               uint32_t *ranges, uint32_t *tris, int64_t *n_list, uint8_t *dist)
{
    if (!verts || !dims || !n_list) return -1;
    pach::Grid g;
    std::string err;
    if (pach::build_grid(verts, T, dims, g, err) != 0) return -1;
    for (int k = 0; k < 3; ++k) {
        dims[k] = g.dims[k];
        if (origin) origin[k] = g.origin[k];
        if (cell) cell[k] = g.cell[k];
    }
    const size_t ncell = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    if (ranges) std::copy(g.ranges.begin(), g.ranges.end(), ranges);
    if (tris && *n_list >= (int64_t)g.tris.size()) std::copy(g.tris.begin(), g.tris.end(), tris);
    if (dist) std::copy(g.dist.begin(), g.dist.begin() + ncell, dist);
    *n_list = (int64_t)g.tris.size();
    return 0;
}

extern "C" int pach_grid_face_skip(const double *verts, int T, int dims[3], uint32_t *skip)
{
    if (!verts || !dims || !skip) return -1;
    pach::Grid g;
    std::string err;
    if (pach::build_grid(verts, T, dims, g, err) != 0) return -1;
    for (int k = 0; k < 3; ++k) dims[k] = g.dims[k];
    std::vector<uint32_t> m;
    pach::face_skip(g, m);
    std::copy(m.begin(), m.end(), skip);
    return 0;
}

extern "C" int pach_grid_octant_cubes(const int dims[3], const uint8_t *dist, uint8_t *oct)
{
    if (!dims || !dist || !oct || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return -1;
    const size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<uint8_t> d(dist, dist + ncell), o;
    pach::octant_cubes(dims, d, o);
    std::copy(o.begin(), o.end(), oct);
    return 0;
}

extern "C" int pach_grid_octant_boxes(const int dims[3], const uint8_t *dist, uint8_t *ext)
{
    if (!dims || !dist || !ext || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return -1;
    const size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<uint8_t> d(dist, dist + ncell), o, e;
    pach::octant_cubes(dims, d, o);
    pach::octant_boxes(dims, d, o, e);
    std::copy(e.begin(), e.end(), ext);
    return 0;
}
