// dk_clouds.hip -- refinable clouds at the mesh level for the pc2 coupling (FP64 form).
//
// 3-D plain cloud codes (the cloud level of a same-level mesh, no transition lobes), meshes
// periodic in x1, x2 and x3.  A cloud of half-width h = root dx is traversed in the block that
// holds its centre, ghost cells included (it reaches at most one cell beyond that block): its
// ghost deposits go through the same fold as every deposit, and its gathers read the filled
// ghosts, which sums what LeafTraversal3D::ForEach (src/dust/dust_leaf3d.hpp:368-517) visits
// on the neighbour blocks themselves.  Thread per particle, FP64 atomics, no FMA contraction:
// the arithmetic is the reference's, in its order.
//
// The reference keeps the gather's (cell, weight) pairs for the deposit (CloudGeometry
// kCloudWeightCache); here the deposit walks the cloud again without moments, which visits the
// same cells with nonzero weight in the same order with the same weights (the weights do not
// depend on the moments).
#include <hip/hip_runtime.h>
#include <cmath>
#include "dk_internal.hpp"
#include "dk_smr.hpp"
#include "dk_cloudwalk.hpp"

using dk::DevBlock;
using dk::Dir;
using dk::Geo;
using dk::Idx;
using dk::kBlock;
using dk::CloudGeo;
using dk::HatCDF;
using dk::HatMoment;
using dk::MakeCloudGeo;
using dk::MinImage;
using dk::SmrCloudForEach;
using dk::Sq;

namespace {

#define GRID_STRIDE(t, n) \
  for (long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x; t < (n); \
       t += (long long)gridDim.x*blockDim.x)
#define PR(c, p) r[(long long)(c)*cap + (p)]
#define PI(c, p) iv[(long long)(c)*cap + (p)]

inline int Grid(long long n) {
  long long b = (n + kBlock - 1)/kBlock;
  if (b < 1) b = 1;
  return static_cast<int>(std::min<long long>(b, 65536));
}

// One direction of the cloud (CloudShape1D::MakeShape / FaceAt / Cell of dust.hpp:240-298,
// lobe 0): the first cell i0 of the support [x - h, x + h] in a block starting at b0 with cells
// of size d, and the weight w and first moment g (about x, length units) of up to 3 cells.
// Returns the cell count, or 0 when the support reaches beyond the first ghost layer.
template <bool kMoments>
__device__ inline int CloudAxis(const double x, const double h, const double ih, const double b0,
                                const double d, const int nx, int &i0, double w[3],
                                double g[3]) {
  const double lo = x - h, hi = x + h;
  const double id = 1.0/d;
  i0 = static_cast<int>(floor((lo - b0)*id));
  const int i1 = static_cast<int>(ceil((hi - b0)*id)) - 1;
  const int n = i1 - i0 + 1;
  if (i0 < -1 || i1 > nx || n > 3) return 0;
  double fc = HatCDF((b0 + i0*d - x)*ih);
  double fh = kMoments ? HatMoment((b0 + i0*d - x)*ih) : 0.0;
#pragma unroll
  for (int a=0; a<3; ++a) {
    w[a] = 0.0; g[a] = 0.0;
    if (a < n) {
      const double sz = (b0 + (i0 + a + 1)*d - x)*ih;
      const double nc = HatCDF(sz);
      w[a] = nc - fc;
      fc = nc;
      if (kMoments) {
        const double nh = HatMoment(sz);
        g[a] = h*(nh - fh);
        fh = nh;
      }
    }
  }
  return n;
}

// The block that holds x (block m or a same-level neighbour; x is moved by a mesh length across
// a periodic face into that block's frame); -1 when x is farther away.
__device__ inline int CloudAnchor(const DevBlock *B, const CloudGeo &cg, const int m,
                                  double x[3]) {
  const DevBlock &b = B[m];
  const double lo[3] = {b.x1min, b.x2min, b.x3min}, hi[3] = {b.x1max, b.x2max, b.x3max};
  int o[3];
#pragma unroll
  for (int d=0; d<3; ++d) o[d] = (x[d] < lo[d]) ? -1 : ((x[d] >= hi[d]) ? 1 : 0);
  if (o[0] == 0 && o[1] == 0 && o[2] == 0) return m;
  const int nb = b.nghbr[Dir(o[0], o[1], o[2])];
  if (nb < 0) return -1;
  const DevBlock &n = B[nb];
  const double nlo[3] = {n.x1min, n.x2min, n.x3min}, nhi[3] = {n.x1max, n.x2max, n.x3max};
#pragma unroll
  for (int d=0; d<3; ++d) {
    if (o[d] == 0) continue;
    if (x[d] < nlo[d]) x[d] += cg.len[d];
    else if (x[d] >= nhi[d]) x[d] -= cg.len[d];
  }
  return nb;
}

// the three directions of the cloud at x in block am; false when it reaches too far
template <bool kMoments>
struct Cloud {
  int i0, j0, k0, nx, ny, nz;
  double wx[3], wy[3], wz[3], gx[3], gy[3], gz[3];
  __device__ bool Make(const Geo &g, const DevBlock &b, const CloudGeo &cg, const double x[3]) {
    nx = CloudAxis<kMoments>(x[0], cg.h[0], cg.ih[0], b.x1min, b.dx1, g.nx1, i0, wx, gx);
    ny = CloudAxis<kMoments>(x[1], cg.h[1], cg.ih[1], b.x2min, b.dx2, g.nx2, j0, wy, gy);
    nz = CloudAxis<kMoments>(x[2], cg.h[2], cg.ih[2], b.x3min, b.dx3, g.nx3, k0, wz, gz);
    return nx > 0 && ny > 0 && nz > 0;
  }
  // LeafTraversal3D::ForEach's loop order (x3, x2, then x1) and skips: f(k, j, i, w, m1x, m1y,
  // m1z) with moments, f(k, j, i, w) without (padded indices)
  template <class F>
  __device__ void ForEach(const Geo &g, const F &f) const {
    // (fixed trip counts with the cell counts as predicates: the arrays stay in registers)
#pragma unroll
    for (int c=0; c<3; ++c) {
      if (c >= nz) continue;
      const double w3 = wz[c], g3 = gz[c];
      if (w3 == 0.0 && (!kMoments || g3 == 0.0)) continue;
#pragma unroll
      for (int bb=0; bb<3; ++bb) {
        if (bb >= ny) continue;
        const double w2 = wy[bb], g2 = gy[bb];
        if (w2 == 0.0 && (!kMoments || g2 == 0.0)) continue;
        const double w23 = w2*w3;
#pragma unroll
        for (int a=0; a<3; ++a) {
          if (a >= nx) continue;
          const double wa = wx[a];
          const double w = w23*wa;
          const int k = k0 + c + g.ks, j = j0 + bb + g.js, i = i0 + a + g.is;
          if constexpr (kMoments) {
            const double ga = gx[a];
            const double m1x = w23*ga;
            const double m1y = wa*g2*w3, m1z = wa*w2*g3;
            if (w == 0.0 && m1x == 0.0 && m1y == 0.0 && m1z == 0.0) continue;
            f(k, j, i, w, m1x, m1y, m1z);
          } else {
            if (w == 0.0) continue;
            f(k, j, i, w);
          }
        }
      }
    }
  }
};

// LeafApi<LeafTraversal3D>::GatherLinear (src/dust/dust_leaf_api.hpp:202-220)
__device__ inline void GatherLinear(const Geo &g, const DevBlock &b, const CloudGeo &cg,
                                    const double *gp, const int m, const int k, const int j,
                                    const int i, const double w, const double m1x,
                                    const double m1y, const double m1z, const double x[3],
                                    double u[3]) {
  const double dx1 = b.dx1, dx2 = b.dx2, dx3 = b.dx3;
  const double xc = b.x1min + (i - g.is + 0.5)*dx1;
  const double yc = b.x2min + (j - g.js + 0.5)*dx2;
  const double zc = b.x3min + (k - g.ks + 0.5)*dx3;
  const double cx = 0.5*(m1x + MinImage(x[0] - xc, cg.len[0])*w)/dx1;
  const double cy = 0.5*(m1y + MinImage(x[1] - yc, cg.len[1])*w)/dx2;
  const double cz = 0.5*(m1z + MinImage(x[2] - zc, cg.len[2])*w)/dx3;
  const double *q = gp + (((((long long)m*g.n3 + k)*g.n2 + j)*g.n1 + i)*12);
  u[0] += w*q[0] + cx*q[3] + cy*q[6] + cz*q[9];
  u[1] += w*q[1] + cx*q[4] + cy*q[7] + cz*q[10];
  u[2] += w*q[2] + cx*q[5] + cy*q[8] + cz*q[11];
}

// src/dust/dust_cloud3d.cpp:1342-1360 (PackGasGather3D, dust_pc2_predictor_gasvel3d)
__global__ void k_gas_velocity(Geo g, const double *u, double *vel) {
  const long long n = g.ncb*g.nmb;
  GRID_STRIDE(c, n) {
    const long long m = c/g.ncb, rr = c - m*g.ncb;
    const double invrho = 1.0/u[(m*DK_NGAS + DK_IDN)*g.ncb + rr];
    for (int v=0; v<3; ++v)
      vel[(m*3 + v)*g.ncb + rr] = u[(m*DK_NGAS + DK_IM1 + v)*g.ncb + rr]*invrho;
  }
}

// src/dust/dust_cloud3d.cpp:1319-1340 (PackGatherField3D, dust_gather_pack3d): velocity and
// its central differences of every cell with both neighbours in each direction
__global__ void k_pack_gather(Geo g, const double *vel, double *gp) {
  const long long ni = g.n1 - 2, nj = g.n2 - 2, nk = g.n3 - 2;
  const long long n = (long long)g.nmb*nk*nj*ni;
  GRID_STRIDE(t, n) {
    const int i = 1 + static_cast<int>(t % ni);
    long long r = t/ni;
    const int j = 1 + static_cast<int>(r % nj);
    r /= nj;
    const int k = 1 + static_cast<int>(r % nk);
    const int m = static_cast<int>(r/nk);
    double *q = gp + (((((long long)m*g.n3 + k)*g.n2 + j)*g.n1 + i)*12);
    for (int c=0; c<3; ++c) {
      q[c] = vel[Idx(g,m,3,c,k,j,i)];
      q[3+c] = vel[Idx(g,m,3,c,k,j,i+1)] - vel[Idx(g,m,3,c,k,j,i-1)];
      q[6+c] = vel[Idx(g,m,3,c,k,j+1,i)] - vel[Idx(g,m,3,c,k,j-1,i)];
      q[9+c] = vel[Idx(g,m,3,c,k+1,j,i)] - vel[Idx(g,m,3,c,k-1,j,i)];
    }
  }
}

// src/dust/dust_cloud3d.cpp:1369-1437 (CloudPredictorT, 3-D): the linear gather of the
// step-start gas velocity over the cloud, the drag impulse deposited into qdep 1-3 (and 4)
__global__ void __launch_bounds__(kBlock) k_cloud_predictor(Geo g, const DevBlock *B, CloudGeo cg, const double *r,
                                  const int32_t *iv, long long cap, long long np,
                                  const double *gp, double *qdep, double dt, int rate_dep,
                                  dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    double x[3] = {PR(DK_IPX,p), PR(DK_IPY,p), PR(DK_IPZ,p)};
    const int am = CloudAnchor(B, cg, PI(DK_PGID,p), x);
    if (am < 0) {dk::RecordError(err, DK_EHALO, p); continue;}
    const DevBlock &b = B[am];
    Cloud<true> cl;
    if (!cl.Make(g, b, cg, x)) {dk::RecordError(err, DK_EHALO, p); continue;}
    double ug[3] = {0.0, 0.0, 0.0};
    cl.ForEach(g, [&](int k, int j, int i, double w, double m1x, double m1y, double m1z) {
      GatherLinear(g, b, cg, gp, am, k, j, i, w, m1x, m1y, m1z, x, ug);
    });
    const double mass_rate = PR(DK_IPM,p)/PR(DK_IPTS,p);
    const double gx = -dt*mass_rate*(ug[0] - PR(DK_IPVX,p));
    const double gy = -dt*mass_rate*(ug[1] - PR(DK_IPVY,p));
    const double gz = -dt*mass_rate*(ug[2] - PR(DK_IPVZ,p));
    const double inv_volume = 1.0/(b.dx1*b.dx2*b.dx3);
    Cloud<false> cw;
    cw.Make(g, b, cg, x);
    cw.ForEach(g, [&](int k, int j, int i, double w) {
      unsafeAtomicAdd(&qdep[Idx(g,am,5,1,k,j,i)], inv_volume*w*gx);
      unsafeAtomicAdd(&qdep[Idx(g,am,5,2,k,j,i)], inv_volume*w*gy);
      unsafeAtomicAdd(&qdep[Idx(g,am,5,3,k,j,i)], inv_volume*w*gz);
      if (rate_dep) unsafeAtomicAdd(&qdep[Idx(g,am,5,4,k,j,i)], inv_volume*w*mass_rate);
    });
  }
}

// src/dust/dust_cloud3d.cpp:1449-1536 (CloudAdvanceT, 3-D) with CloudPC2Push<LeafTraversal3D>
// (:118-158): the linear gather of u* at the midpoint, the implicit-midpoint kick and drift, the
// impulse deposited into dmom
__global__ void __launch_bounds__(kBlock) k_cloud_advance(Geo g, const DevBlock *B, CloudGeo cg, double *r,
                                const int32_t *iv, long long cap, long long np,
                                const double *gp, double *dmom, double dt, int br, int sbox,
                                int strat, double qs, double om, dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const double x0[3] = {PR(DK_IPX1,p), PR(DK_IPY1,p), PR(DK_IPZ1,p)};
    const double v0[3] = {PR(DK_IPVX1,p), PR(DK_IPVY1,p), PR(DK_IPVZ1,p)};
    double xh[3], xg[3];
    xh[0] = x0[0] + 0.5*dt*v0[0];
    xh[1] = x0[1] + 0.5*dt*(v0[1] - (sbox ? qs*om*x0[0] : 0.0));
    xh[2] = x0[2] + 0.5*dt*v0[2];
    xg[0] = xh[0];
    xg[1] = xh[1] + (sbox ? qs*om*xh[0]*0.5*dt : 0.0);
    xg[2] = xh[2];
    const int am = CloudAnchor(B, cg, PI(DK_PGID,p), xg);
    if (am < 0) {dk::RecordError(err, DK_EHALO, p); continue;}
    const DevBlock &b = B[am];
    Cloud<true> cl;
    if (!cl.Make(g, b, cg, xg)) {dk::RecordError(err, DK_EHALO, p); continue;}
    double ug[3] = {0.0, 0.0, 0.0};
    cl.ForEach(g, [&](int k, int j, int i, double w, double m1x, double m1y, double m1z) {
      GatherLinear(g, b, cg, gp, am, k, j, i, w, m1x, m1y, m1z, xg, ug);
    });
    const double h = 0.5*dt;
    const double lambda = h/PR(DK_IPTS,p);
    const double aa = 1.0 + lambda;
    double rr[3] = {v0[0] + lambda*ug[0], v0[1] + lambda*ug[1], v0[2] + lambda*ug[2]};
    double vh[3];
    if (sbox) {
      const double cxy = 2.0*h*om;
      const double cyx = h*(2.0-qs)*om;
      const double det = aa*aa + cxy*cyx;
      vh[0] = (aa*rr[0] + cxy*rr[1])/det;
      vh[1] = (aa*rr[1] - cyx*rr[0])/det;
      if (strat) rr[2] -= h*Sq(om)*xh[2];
      vh[2] = rr[2]/aa;
    } else {
      vh[0] = rr[0]/aa; vh[1] = rr[1]/aa; vh[2] = rr[2]/aa;
    }
    PR(DK_IPVX,p) = 2.0*vh[0] - v0[0];
    PR(DK_IPVY,p) = 2.0*vh[1] - v0[1];
    PR(DK_IPVZ,p) = 2.0*vh[2] - v0[2];
    PR(DK_IPX,p) = x0[0] + dt*vh[0];
    PR(DK_IPY,p) = x0[1] + dt*(vh[1] - (sbox ? qs*om*xh[0] : 0.0));
    PR(DK_IPZ,p) = x0[2] + dt*vh[2];
    PR(DK_IPX1,p) = xh[0]; PR(DK_IPY1,p) = xh[1]; PR(DK_IPZ1,p) = xh[2];
    const double ts = PR(DK_IPTS,p);
    const double dvx = dt*(ug[0]-vh[0])/ts;
    const double dvy = dt*(ug[1]-vh[1])/ts;
    const double dvz = dt*(ug[2]-vh[2])/ts;
    PR(DK_IPRX,p) = dvx; PR(DK_IPRY,p) = dvy; PR(DK_IPRZ,p) = dvz;
    if (!br) continue;
    const double mass = -PR(DK_IPM,p);
    const double measure = b.dx1*b.dx2*b.dx3;
    Cloud<false> cw;
    cw.Make(g, b, cg, xg);
    cw.ForEach(g, [&](int k, int j, int i, double w) {
      const double dw = mass*w/measure;
      unsafeAtomicAdd(&dmom[Idx(g,am,3,0,k,j,i)], dw*dvx);
      unsafeAtomicAdd(&dmom[Idx(g,am,3,1,k,j,i)], dw*dvy);
      unsafeAtomicAdd(&dmom[Idx(g,am,3,2,k,j,i)], dw*dvz);
    });
  }
}

// CloudPredictorT on a statically refined mesh: the gather and the deposit over the leaves of
// every level (the deposit walks the cloud again without moments: the same cells and weights in
// the same order), each leaf cell with its own volume
__global__ void __launch_bounds__(kBlock) k_smr_cloud_predictor(
    Geo g, const DevBlock *B, dk::SmrTree t, CloudGeo cg, const double *r, const int32_t *iv,
    long long cap, long long np, const double *gp, double *qdep, double dt, int rate_dep,
    dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const double x[3] = {PR(DK_IPX,p), PR(DK_IPY,p), PR(DK_IPZ,p)};
    const int lev = PI(DK_PLEV,p);
    double ug[3] = {0.0, 0.0, 0.0};
    const bool ok = SmrCloudForEach<true>(g, B, t, cg, x, lev,
        [&](int b, int k, int j, int i, double w, double m1x, double m1y, double m1z) {
          GatherLinear(g, B[b], cg, gp, b, k, j, i, w, m1x, m1y, m1z, x, ug);
        });
    if (!ok) {dk::RecordError(err, DK_EHALO, p); continue;}
    const double mass_rate = PR(DK_IPM,p)/PR(DK_IPTS,p);
    const double gx = -dt*mass_rate*(ug[0] - PR(DK_IPVX,p));
    const double gy = -dt*mass_rate*(ug[1] - PR(DK_IPVY,p));
    const double gz = -dt*mass_rate*(ug[2] - PR(DK_IPVZ,p));
    SmrCloudForEach<false>(g, B, t, cg, x, lev, [&](int b, int k, int j, int i, double w) {
      const DevBlock &bb = B[b];
      const double inv_volume = 1.0/(bb.dx1*bb.dx2*bb.dx3);
      unsafeAtomicAdd(&qdep[Idx(g,b,5,1,k,j,i)], inv_volume*w*gx);
      unsafeAtomicAdd(&qdep[Idx(g,b,5,2,k,j,i)], inv_volume*w*gy);
      unsafeAtomicAdd(&qdep[Idx(g,b,5,3,k,j,i)], inv_volume*w*gz);
      if (rate_dep) unsafeAtomicAdd(&qdep[Idx(g,b,5,4,k,j,i)], inv_volume*w*mass_rate);
    });
  }
}

// CloudAdvanceT on a statically refined mesh (periodic, no shearing box)
__global__ void __launch_bounds__(kBlock) k_smr_cloud_advance(
    Geo g, const DevBlock *B, dk::SmrTree t, CloudGeo cg, double *r, const int32_t *iv,
    long long cap, long long np, const double *gp, double *dmom, double dt, int br,
    dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const double x0[3] = {PR(DK_IPX1,p), PR(DK_IPY1,p), PR(DK_IPZ1,p)};
    const double v0[3] = {PR(DK_IPVX1,p), PR(DK_IPVY1,p), PR(DK_IPVZ1,p)};
    double xh[3];
    xh[0] = x0[0] + 0.5*dt*v0[0];
    xh[1] = x0[1] + 0.5*dt*v0[1];
    xh[2] = x0[2] + 0.5*dt*v0[2];
    const int lev = PI(DK_PLEV,p);
    double ug[3] = {0.0, 0.0, 0.0};
    const bool ok = SmrCloudForEach<true>(g, B, t, cg, xh, lev,
        [&](int b, int k, int j, int i, double w, double m1x, double m1y, double m1z) {
          GatherLinear(g, B[b], cg, gp, b, k, j, i, w, m1x, m1y, m1z, xh, ug);
        });
    if (!ok) {dk::RecordError(err, DK_EHALO, p); continue;}
    const double h = 0.5*dt;
    const double lambda = h/PR(DK_IPTS,p);
    const double aa = 1.0 + lambda;
    const double rr[3] = {v0[0] + lambda*ug[0], v0[1] + lambda*ug[1], v0[2] + lambda*ug[2]};
    const double vh[3] = {rr[0]/aa, rr[1]/aa, rr[2]/aa};
    PR(DK_IPVX,p) = 2.0*vh[0] - v0[0];
    PR(DK_IPVY,p) = 2.0*vh[1] - v0[1];
    PR(DK_IPVZ,p) = 2.0*vh[2] - v0[2];
    PR(DK_IPX,p) = x0[0] + dt*vh[0];
    PR(DK_IPY,p) = x0[1] + dt*vh[1];
    PR(DK_IPZ,p) = x0[2] + dt*vh[2];
    PR(DK_IPX1,p) = xh[0]; PR(DK_IPY1,p) = xh[1]; PR(DK_IPZ1,p) = xh[2];
    const double ts = PR(DK_IPTS,p);
    const double dvx = dt*(ug[0]-vh[0])/ts;
    const double dvy = dt*(ug[1]-vh[1])/ts;
    const double dvz = dt*(ug[2]-vh[2])/ts;
    PR(DK_IPRX,p) = dvx; PR(DK_IPRY,p) = dvy; PR(DK_IPRZ,p) = dvz;
    if (!br) continue;
    const double mass = -PR(DK_IPM,p);
    SmrCloudForEach<false>(g, B, t, cg, xh, lev, [&](int b, int k, int j, int i, double w) {
      const DevBlock &bb = B[b];
      const double dw = mass*w/(bb.dx1*bb.dx2*bb.dx3);
      unsafeAtomicAdd(&dmom[Idx(g,b,3,0,k,j,i)], dw*dvx);
      unsafeAtomicAdd(&dmom[Idx(g,b,3,1,k,j,i)], dw*dvy);
      unsafeAtomicAdd(&dmom[Idx(g,b,3,2,k,j,i)], dw*dvz);
    });
  }
}

}  // namespace

namespace dk {

bool CloudsSupported(const dk_mesh *mesh, const dk_physics *phys) {
  return phys->coupling == DK_PC2 && mesh->bc_x1 == DK_X1_PERIODIC &&
         mesh->bc_x3 == DK_X3_PERIODIC && mesh->ng >= 2;
}

int CloudPredictor(dk_ctx *c, double dt, int rate_dep) {
  const Geo &g = c->g;
  const long long nc = g.ncb*g.nmb;
  const long long nin = (long long)g.nmb*(g.n1 - 2)*(g.n2 - 2)*(g.n3 - 2);
  {const int t = ProfBegin(c, "cloud_pack");
  k_gas_velocity<<<Grid(nc), kBlock, 0, c->stream>>>(g, c->u1, c->cvel);
  k_pack_gather<<<Grid(nin), kBlock, 0, c->stream>>>(g, c->cvel, c->gpack);
  ProfEnd(c, t);}
  if (c->np == 0) return Check(hipGetLastError());
  if (c->smr) {
    const int t = ProfBegin(c, "cloud_predictor");
    k_smr_cloud_predictor<<<Grid(c->np), kBlock, 0, c->stream>>>(
        g, c->d_blocks, c->tree, MakeCloudGeo(c), c->rdata, c->idata, (long long)c->cap,
        (long long)c->np, c->gpack, c->qdep, dt, rate_dep, c->d_err);
    ProfEnd(c, t);
    return Check(hipGetLastError());
  }
  {const int t = ProfBegin(c, "cloud_predictor");
  k_cloud_predictor<<<Grid(c->np), kBlock, 0, c->stream>>>(
      g, c->d_blocks, MakeCloudGeo(c), c->rdata, c->idata, (long long)c->cap,
      (long long)c->np, c->gpack, c->qdep, dt, rate_dep, c->d_err);
  ProfEnd(c, t);}
  return Check(hipGetLastError());
}

int CloudAdvance(dk_ctx *c, double dt) {
  const Geo &g = c->g;
  const long long nin = (long long)g.nmb*(g.n1 - 2)*(g.n2 - 2)*(g.n3 - 2);
  {const int t = ProfBegin(c, "cloud_pack");
  k_pack_gather<<<Grid(nin), kBlock, 0, c->stream>>>(g, c->ustar, c->gpack);
  ProfEnd(c, t);}
  if (c->np == 0) return Check(hipGetLastError());
  if (c->smr) {
    const int t = ProfBegin(c, "cloud_advance");
    k_smr_cloud_advance<<<Grid(c->np), kBlock, 0, c->stream>>>(
        g, c->d_blocks, c->tree, MakeCloudGeo(c), c->rdata, c->idata, (long long)c->cap,
        (long long)c->np, c->gpack, c->dmom, dt, c->P.back_reaction, c->d_err);
    ProfEnd(c, t);
    return Check(hipGetLastError());
  }
  {const int t = ProfBegin(c, "cloud_advance");
  k_cloud_advance<<<Grid(c->np), kBlock, 0, c->stream>>>(
      g, c->d_blocks, MakeCloudGeo(c), c->rdata, c->idata, (long long)c->cap,
      (long long)c->np, c->gpack, c->dmom, dt, c->P.back_reaction, c->P.shearing_box,
      c->P.stratified, c->P.qshear, c->P.omega0, c->d_err);
  ProfEnd(c, t);}
  return Check(hipGetLastError());
}

}  // namespace dk
