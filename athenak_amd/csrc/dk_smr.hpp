// dk_smr.hpp -- statically refined meshes: the leaf that holds a position, and the neighbour a
// particle crosses to (device functions shared by the routing, the ghost values and the cloud
// traversal).
//
// Every block is a leaf of level 0..maxlev with its logical location at its level; the finest
// block grid (F[d] = root blocks << maxlev slots per direction) maps each slot to the leaf that
// covers it.  A leaf is found from a position by its slot, then one slot over in each direction
// where the position lies outside that leaf's box by round-off (the boxes are LeftEdgeX at each
// level, src/coordinates/cell_locations.hpp:24-27, not exact multiples of the slot length).
#ifndef ATHENAK_AMD_CSRC_DK_SMR_HPP_
#define ATHENAK_AMD_CSRC_DK_SMR_HPP_

#include "dk_internal.hpp"

namespace dk {

// p moved by whole mesh lengths into [min, max) of the mesh; k: the lengths taken off
__device__ inline void SmrWrap(const SmrTree &t, double p[3], int k[3]) {
#pragma unroll
  for (int d=0; d<3; ++d) {
    k[d] = 0;
    while (p[d] < t.lo[d]) {p[d] += t.len[d]; --k[d];}
    while (p[d] >= t.lo[d] + t.len[d]) {p[d] -= t.len[d]; ++k[d];}
  }
}

// the leaf holding p (inside the mesh); -1 if none does
__device__ inline int SmrLeafAt(const DevBlock *B, const SmrTree &t, const double p[3]) {
  int sl[3];
#pragma unroll
  for (int d=0; d<3; ++d) {
    sl[d] = static_cast<int>(floor((p[d] - t.lo[d])/t.len[d]*t.F[d]));
    sl[d] = min(max(sl[d], 0), t.F[d] - 1);
  }
  for (int it=0; it<4; ++it) {
    const int b = t.leaf[((long long)sl[2]*t.F[1] + sl[1])*t.F[0] + sl[0]];
    const DevBlock &bb = B[b];
    const double bmin[3] = {bb.x1min, bb.x2min, bb.x3min};
    const double bmax[3] = {bb.x1max, bb.x2max, bb.x3max};
    const int r = 1 << (t.maxlev - bb.level);
    const int l[3] = {bb.lx1, bb.lx2, bb.lx3};
    bool moved = false;
#pragma unroll
    for (int d=0; d<3; ++d) {
      if (p[d] < bmin[d] && l[d]*r > 0) {sl[d] = l[d]*r - 1; moved = true;}
      else if (p[d] >= bmax[d] && (l[d] + 1)*r < t.F[d]) {sl[d] = (l[d] + 1)*r; moved = true;}
    }
    if (!moved) return b;
  }
  return -1;
}

// SetNewPrtclGID's neighbour slot (src/bvals/bvals_part.cpp:308-390) on a refined mesh: the
// leaf holding the same-level region beside block m in the direction off of the crossing when
// it is of m's level or coarser (a coarser leaf also covers an empty edge or corner slot,
// coarse_cover), else the finer leaf beside that face, edge or corner in the half of m that x
// occupies along the other directions (the sub-slot of fx, fy, fz, :224-229); -1 when none
__device__ inline int SmrDest(const DevBlock *B, const SmrTree &t, const int m, const int off[3],
                              const double x[3]) {
  const DevBlock &b = B[m];
  const double bmin[3] = {b.x1min, b.x2min, b.x3min}, bmax[3] = {b.x1max, b.x2max, b.x3max};
  double q[3];
  int k[3];
#pragma unroll
  for (int d=0; d<3; ++d) q[d] = bmin[d] + (off[d] + 0.5)*(bmax[d] - bmin[d]);
  SmrWrap(t, q, k);
  const int n0 = SmrLeafAt(B, t, q);
  if (n0 < 0 || B[n0].level <= b.level) return n0;
#pragma unroll
  for (int d=0; d<3; ++d) {
    const double l = bmax[d] - bmin[d];
    const int f = (x[d] < 0.5*(bmin[d] + bmax[d])) ? 0 : 1;
    q[d] = (off[d] < 0) ? bmin[d] - 0.25*l
         : ((off[d] > 0) ? bmax[d] + 0.25*l : bmin[d] + (f ? 0.75 : 0.25)*l);
  }
  SmrWrap(t, q, k);
  const int n = SmrLeafAt(B, t, q);
  return (n >= 0 && B[n].level == b.level + 1) ? n : -1;
}

}  // namespace dk

#endif  // ATHENAK_AMD_CSRC_DK_SMR_HPP_
