// dk_cloudwalk.hpp -- the refinable cloud's shape and its walk over the leaves of a statically
// refined mesh (device code shared by the FP64 form's cloud kernels and the FP32 form's
// cross-level clouds).
#ifndef ATHENAK_AMD_CSRC_DK_CLOUDWALK_HPP_
#define ATHENAK_AMD_CSRC_DK_CLOUDWALK_HPP_

#include <cmath>

#include "dk_internal.hpp"
#include "dk_smr.hpp"

namespace dk {

// the cloud half-width (the root cell size) per direction, its reciprocal, the mesh lengths
struct CloudGeo {
  double h[3], ih[3], len[3];
};

__device__ inline double Sq(double x) { return x*x; }

// src/dust/dust.hpp:169-182 (CloudGeometry::HatCDF, HatMoment)
__device__ inline double HatCDF(const double s) {
  if (s <= -1.0) return 0.0;
  if (s >= 1.0) return 1.0;
  return (s <= 0.0) ? 0.5*Sq(1.0 + s) : 1.0 - 0.5*Sq(1.0 - s);
}
__device__ inline double HatMoment(const double s) {
  if (s <= -1.0 || s >= 1.0) return 1.0/6.0;
  return (s <= 0.0) ? 0.5*Sq(s) + Sq(s)*s/3.0 : 0.5*Sq(s) - Sq(s)*s/3.0;
}
// src/dust/dust.hpp:186-188
__device__ inline double MinImage(const double d, const double length) {
  return d - length*round(d/length);
}

// LeafTraversal3D::ForEach (src/dust/dust_leaf3d.hpp:383-517) on a statically refined mesh:
// the actual leaves under the cloud of level lev at x (half-width root dx 2^-lev) at their own
// resolution, f(b, k, j, i, w, m1x, m1y, m1z) (padded indices of leaf b) with moments,
// f(b, k, j, i, w) without, per candidate leaf the reference's loop order (x runs of 4, then
// x3, x2, x1) and skips.  The candidates are the leaves holding the support's 8 corners (a leaf
// at least as long as the support in every direction meets it only if it holds a corner), in
// the order of the corners that found them (x1 fastest, low side first); false when a corner
// lies in no leaf or a candidate is shorter than the support.
template <bool kMoments, class F>
__device__ bool SmrCloudForEach(const Geo &g, const DevBlock *B, const SmrTree &t,
                                const CloudGeo &cg, const double x[3], const int lev,
                                const F &f) {
  double h[3], ih[3], lo[3], hi[3];
#pragma unroll
  for (int d=0; d<3; ++d) {
    h[d] = ldexp(cg.h[d], -lev);
    ih[d] = ldexp(cg.ih[d], lev);
    lo[d] = x[d] - h[d];
    hi[d] = x[d] + h[d];
  }
  int cb[8], cs[8][3], nc = 0;
  for (int c=0; c<8; ++c) {
    double p[3];
    int k[3];
#pragma unroll
    for (int d=0; d<3; ++d) p[d] = ((c >> d) & 1) ? nextafter(hi[d], -INFINITY) : lo[d];
    SmrWrap(t, p, k);
    const int b = SmrLeafAt(B, t, p);
    if (b < 0) return false;
    bool dup = false;
    for (int q=0; q<nc; ++q)
      if (cb[q] == b && cs[q][0] == k[0] && cs[q][1] == k[1] && cs[q][2] == k[2]) dup = true;
    if (dup) continue;
    cb[nc] = b; cs[nc][0] = k[0]; cs[nc][1] = k[1]; cs[nc][2] = k[2]; ++nc;
  }
  for (int q=0; q<nc; ++q) {
    const DevBlock &bb = B[cb[q]];
    if (bb.x1max - bb.x1min < 2.0*h[0] || bb.x2max - bb.x2min < 2.0*h[1] ||
        bb.x3max - bb.x3min < 2.0*h[2])
      return false;
  }
  constexpr int kRun = 4;
  for (int q=0; q<nc; ++q) {
    const int b = cb[q];
    const DevBlock &bb = B[b];
    const double dx = bb.dx1, dy = bb.dx2, dz = bb.dx3;
    const double bx0 = bb.x1min + cs[q][0]*cg.len[0], by0 = bb.x2min + cs[q][1]*cg.len[1];
    const double bz0 = bb.x3min + cs[q][2]*cg.len[2];
    const double idx = 1.0/dx, idy = 1.0/dy, idz = 1.0/dz;
    const int i0 = max(static_cast<int>(floor((lo[0] - bx0)*idx)), 0);
    const int i1 = min(static_cast<int>(ceil((hi[0] - bx0)*idx)) - 1, g.nx1 - 1);
    const int j0 = max(static_cast<int>(floor((lo[1] - by0)*idy)), 0);
    const int j1 = min(static_cast<int>(ceil((hi[1] - by0)*idy)) - 1, g.nx2 - 1);
    const int k0 = max(static_cast<int>(floor((lo[2] - bz0)*idz)), 0);
    const int k1 = min(static_cast<int>(ceil((hi[2] - bz0)*idz)) - 1, g.nx3 - 1);
    for (int ic=i0; ic<=i1; ic+=kRun) {
      const int nr = (i1 - ic + 1 < kRun) ? i1 - ic + 1 : kRun;
      double w1[kRun], g1[kRun];
      double fc = HatCDF((bx0 + ic*dx - x[0])*ih[0]);
      double fh = kMoments ? HatMoment((bx0 + ic*dx - x[0])*ih[0]) : 0.0;
#pragma unroll
      for (int a=0; a<kRun; ++a) {
        w1[a] = 0.0; g1[a] = 0.0;
        if (a < nr) {
          const double sz = (bx0 + (ic + a + 1)*dx - x[0])*ih[0];
          const double nc1 = HatCDF(sz);
          w1[a] = nc1 - fc;
          fc = nc1;
          if (kMoments) {
            const double nh = HatMoment(sz);
            g1[a] = h[0]*(nh - fh);
            fh = nh;
          }
        }
      }
      double f3c = HatCDF((bz0 + k0*dz - x[2])*ih[2]);
      double f3h = kMoments ? HatMoment((bz0 + k0*dz - x[2])*ih[2]) : 0.0;
      for (int k=k0; k<=k1; ++k) {
        const double sz3 = (bz0 + (k + 1)*dz - x[2])*ih[2];
        const double n3c = HatCDF(sz3);
        const double w3 = n3c - f3c;
        double g3 = 0.0;
        f3c = n3c;
        if (kMoments) {const double n3h = HatMoment(sz3); g3 = h[2]*(n3h - f3h); f3h = n3h;}
        if (w3 == 0.0 && (!kMoments || g3 == 0.0)) continue;
        double f2c = HatCDF((by0 + j0*dy - x[1])*ih[1]);
        double f2h = kMoments ? HatMoment((by0 + j0*dy - x[1])*ih[1]) : 0.0;
        for (int j=j0; j<=j1; ++j) {
          const double sz2 = (by0 + (j + 1)*dy - x[1])*ih[1];
          const double n2c = HatCDF(sz2);
          const double w2 = n2c - f2c;
          double g2 = 0.0;
          f2c = n2c;
          if (kMoments) {const double n2h = HatMoment(sz2); g2 = h[1]*(n2h - f2h); f2h = n2h;}
          if (w2 == 0.0 && (!kMoments || g2 == 0.0)) continue;
          const double w23 = w2*w3;
          for (int a=0; a<nr; ++a) {
            const double wa = w1[a];
            const double w = w23*wa;
            if constexpr (kMoments) {
              const double ga = g1[a];
              const double m1x = w23*ga;
              const double m1y = wa*g2*w3, m1z = wa*w2*g3;
              if (w == 0.0 && m1x == 0.0 && m1y == 0.0 && m1z == 0.0) continue;
              f(b, k + g.ks, j + g.js, ic + a + g.is, w, m1x, m1y, m1z);
            } else {
              if (w == 0.0) continue;
              f(b, k + g.ks, j + g.js, ic + a + g.is, w);
            }
          }
        }
      }
    }
  }
  return true;
}

inline CloudGeo MakeCloudGeo(const dk_ctx *c) {
  CloudGeo cg;
  const dk_mesh &M = c->M;
  // the root cell size (src/dust/dust_cloud.cpp:135, dust_cloud3d.cpp:591-593)
  cg.h[0] = (M.x1max - M.x1min)/static_cast<double>(M.nmbx1*M.nx1);
  cg.h[1] = (M.x2max - M.x2min)/static_cast<double>(M.nmbx2*M.nx2);
  cg.h[2] = (M.x3max - M.x3min)/static_cast<double>(M.nmbx3*M.nx3);
  for (int d=0; d<3; ++d) cg.ih[d] = 1.0/cg.h[d];
  cg.len[0] = M.x1max - M.x1min;
  cg.len[1] = M.x2max - M.x2min;
  cg.len[2] = M.x3max - M.x3min;
  return cg;
}

}  // namespace dk

#endif  // ATHENAK_AMD_CSRC_DK_CLOUDWALK_HPP_
