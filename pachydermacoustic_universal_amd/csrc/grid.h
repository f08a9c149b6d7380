// grid.h -- host-side uniform voxel grid (no HIP dependency).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace pach {

struct Grid {
    int dims[3] = {0, 0, 0};
    double origin[3] = {0, 0, 0};
    double cell[3] = {1, 1, 1};
    std::vector<uint32_t> ranges;  // [ncell][2]: first list index, triangle count
    std::vector<uint32_t> tris;    // concatenated voxel lists
    std::vector<uint8_t> dist;     // [ncell rounded up to 16]: L-inf distance to an occupied voxel
};

bool tri_box_overlap(const double *bc, const double *bh, const double *tv);
// L-inf distance (voxels, capped at 255) to the nearest voxel whose dist is 0
void linf_distance(const int dims[3], std::vector<uint8_t> &dist);
// per-voxel bit masks of the receiver spheres touching each voxel (receiver r
// sets bit[r], 0..63), and the grid's distance field with those voxels counted
// as occupied
void receiver_cells(const Grid &g, const double *centers, int R, const int *bit, double radius,
                    std::vector<uint64_t> &mask, std::vector<uint8_t> &dist);
// For each empty voxel (dist != 0) and each octant o of ray directions
// (bit 0: d_x < 0, bit 1: d_y < 0, bit 2: d_z < 0), oct[8 voxel + o] = the
// edge K of the largest cube of empty voxels that has this voxel as its corner
// and extends K voxels along each axis in o's directions (voxels outside the
// grid count as empty; capped at 255); 0 for occupied voxels.  A ray in
// octant o inside the voxel crosses the whole cube in one move.
void octant_cubes(const int dims[3], const std::vector<uint8_t> &dist, std::vector<uint8_t> &oct);
// For each empty voxel and octant o, the empty box a ray in octant o crosses
// in one move, cornered at the voxel and extending in o's directions (voxels
// outside the grid count as empty; lengths capped at 255): with K =
// oct[8 voxel + o] and E_a the length of the longest empty box that is K long
// on the two axes other than a, the first axis A is the one with the largest
// E_a (the lowest on a tie), the second B the longer of the other two (the
// next after A on a tie); the box is E_A long on A, E2 (the longest empty
// length on B with A at E_A and the third axis at K) on B and K on the third.
// ext[24 voxel + o] = E_A, ext[24 voxel + 8 + o] = E2, ext[24 voxel + 16 + o]
// = A | B << 2; all 0 for occupied voxels.
void octant_boxes(const int dims[3], const std::vector<uint8_t> &dist, const std::vector<uint8_t> &oct,
                  std::vector<uint8_t> &ext);
// For each voxel and each of its faces f (0: -x, 1: +x, 2: -y, 3: +y, 4: -z,
// 5: +z), skip[6 voxel + f] has bit i set when entry i (i < 32) of the voxel's
// list is also listed in the neighbour across f: a walk that steps in through
// face f has already tested that triangle in the neighbour.
void face_skip(const Grid &g, std::vector<uint32_t> &skip);

// The per-cell records of the flattened walks (trace_types.h CellRec), 16
// uint32 words per cell: list range, face skip masks (face_skip; occupied
// voxels) or box lengths (octant_boxes; empty voxels), receiver mask (mask
// empty: none), octant cube edges and distance.
void cell_records(const Grid &g, const std::vector<uint32_t> &skip, const std::vector<uint64_t> &mask,
                  const std::vector<uint8_t> &dist, const std::vector<uint8_t> &oct, const std::vector<uint8_t> &ext,
                  std::vector<uint32_t> &rec);
// dims_in: requested voxel counts (any <= 0: chosen from the triangle count)
int build_grid(const double *verts, int T, const int dims_in[3], Grid &g, std::string &err);

}  // namespace pach
