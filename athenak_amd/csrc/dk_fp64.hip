// dk_fp64.hip -- context management and the plain FP64 form of the drag coupling.
//
// One kernel per reference kernel, thread per particle or per cell, deposits with hardware
// FP64 atomics, operating in place on the caller's particle SoA and gas arrays.  This is the
// baseline form the FP32 cell-sorted form is measured against.  Each kernel cites the
// reference lambda it computes; the arithmetic is written in the reference's operation order.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <utility>

#include "dk_internal.hpp"
#include "dk_smr.hpp"

using dk::Geo;
using dk::DevBlock;
using dk::Idx;
using dk::Dir;
using dk::kBlock;

namespace dk {

int Check(hipError_t e) { return (e == hipSuccess) ? DK_OK : DK_EHIP; }

long long NCells(const dk_ctx *c) { return c->g.ncb*c->g.nmb; }

namespace {

inline int Grid(long long n, int cap = 65536) {
  long long b = (n + kBlock - 1)/kBlock;
  if (b < 1) b = 1;
  return static_cast<int>(std::min<long long>(b, cap));
}

__global__ void k_fill(double *a, long long n, double v) {
  for (long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x*blockDim.x)
    a[t] = v;
}

}  // namespace

int Zero(dk_ctx *c, double *a, long long n) {
  // a fill kernel rather than hipMemsetAsync, so that the sequence can be graph-captured
  const int t = ProfBegin(c, "fill");
  k_fill<<<Grid(n), kBlock, 0, c->stream>>>(a, n, 0.0);
  ProfEnd(c, t);
  return Check(hipGetLastError());
}

}  // namespace dk

namespace {

#define GRID_STRIDE(t, n) \
  for (long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x; t < (n); \
       t += (long long)gridDim.x*blockDim.x)

inline int Grid(long long n) {
  long long b = (n + kBlock - 1)/kBlock;
  if (b < 1) b = 1;
  return static_cast<int>(std::min<long long>(b, 65536));
}

struct Stencil {
  int ip, jp, kp;
  double wx[3], wy[3], wz[3];
};

__device__ inline void MakeStencil(const Geo &g, const DevBlock &b, double x, double y,
                                   double z, Stencil &w) {
  dk::PMWeights(x, b.x1min, b.x1max, g.nx1, g.is, g.deposit, w.ip, w.wx);
  dk::PMWeights(y, b.x2min, b.x2max, g.nx2, g.js, g.deposit, w.jp, w.wy);
  dk::PMWeights(z, b.x3min, b.x3max, g.nx3, g.ks, g.deposit, w.kp, w.wz);
}

__device__ inline bool HaloOK(const Geo &g, const Stencil &w) {   // dust_pm.cpp:241-242
  return !(w.ip < (g.is-1) || w.ip > (g.is+g.nx1) || w.jp < (g.js-1) || w.jp > (g.js+g.nx2) ||
           w.kp < (g.ks-1) || w.kp > (g.ks+g.nx3));
}

// a stencil footprint must lie inside the padded block (guards every gather/deposit)
__device__ inline bool InPad(const Geo &g, const Stencil &w) {
  return w.ip >= 1 && w.ip <= g.n1 - 2 && w.jp >= 1 && w.jp <= g.n2 - 2 && w.kp >= 1 &&
         w.kp <= g.n3 - 2;
}

__device__ inline void AtomicAdd(double *p, double v) { unsafeAtomicAdd(p, v); }

#define PR(c, p) r[(long long)(c)*cap + (p)]
#define PI(c, p) iv[(long long)(c)*cap + (p)]

//----------------------------------------------------------------------------------------
// dust_push.cpp:69-77 (dust_reg_copy)
__global__ void k_reg_copy(double *r, long long cap, long long np) {
  GRID_STRIDE(p, np) {
    PR(DK_IPX1,p) = PR(DK_IPX,p);
    PR(DK_IPVX1,p) = PR(DK_IPVX,p);
    PR(DK_IPY1,p) = PR(DK_IPY,p);
    PR(DK_IPVY1,p) = PR(DK_IPVY,p);
    PR(DK_IPZ1,p) = PR(DK_IPZ,p);
    PR(DK_IPVZ1,p) = PR(DK_IPVZ,p);
  }
}

// dust_push.cpp:53-59 (dust_gas_midpoint), ghosts included
__global__ void k_gas_midpoint(Geo g, const double *u0, const double *u1, double *us) {
  const long long n = g.ncb*g.nmb;
  GRID_STRIDE(c, n) {
    const long long m = c/g.ncb, rr = c - m*g.ncb;
    const double inv_rho_sum = 1.0/(u1[(m*DK_NGAS + DK_IDN)*g.ncb + rr] +
                                    u0[(m*DK_NGAS + DK_IDN)*g.ncb + rr]);
    for (int v=0; v<3; ++v)
      us[(m*3 + v)*g.ncb + rr] = (u1[(m*DK_NGAS + DK_IM1 + v)*g.ncb + rr] +
                                  u0[(m*DK_NGAS + DK_IM1 + v)*g.ncb + rr])*inv_rho_sum;
  }
}

// dust_push.cpp:262-297 (dust_push)
__global__ void k_push_imex(double *r, long long cap, long long np, double gam0, double gam1,
                            double beta_dt, double atw_dt, int sbox, int strat, double qs,
                            double om) {
  GRID_STRIDE(p, np) {
    const double x_old = PR(DK_IPX,p), y_old = PR(DK_IPY,p), z_old = PR(DK_IPZ,p);
    const double vx_old = PR(DK_IPVX,p), vy_old = PR(DK_IPVY,p), vz_old = PR(DK_IPVZ,p);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (sbox) {
      fx = 2.0*om*vy_old;
      fy = -(2.0-qs)*om*vx_old;
      if (strat) fz = -(om*om)*z_old;
    }
    PR(DK_IPVX,p) = gam0*vx_old + gam1*PR(DK_IPVX1,p) + beta_dt*fx + atw_dt*PR(DK_IPRX,p);
    PR(DK_IPVY,p) = gam0*vy_old + gam1*PR(DK_IPVY1,p) + beta_dt*fy + atw_dt*PR(DK_IPRY,p);
    PR(DK_IPVZ,p) = gam0*vz_old + gam1*PR(DK_IPVZ1,p) + beta_dt*fz + atw_dt*PR(DK_IPRZ,p);
    double ydot = vy_old;
    if (sbox) ydot -= qs*om*x_old;
    PR(DK_IPX,p) = gam0*x_old + gam1*PR(DK_IPX1,p) + beta_dt*vx_old;
    PR(DK_IPY,p) = gam0*y_old + gam1*PR(DK_IPY1,p) + beta_dt*ydot;
    PR(DK_IPZ,p) = gam0*z_old + gam1*PR(DK_IPZ1,p) + beta_dt*vz_old;
  }
}

// dust_push.cpp:132-245 (dust_pc2_advance)
__global__ void k_pc2_advance(Geo g, const DevBlock *B, double *r, const int32_t *iv,
                              long long cap, long long np, const double *us, double *dmom,
                              double dt, int br, int sbox, int strat, double qs, double om,
                              dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    const double x0 = PR(DK_IPX1,p), y0 = PR(DK_IPY1,p), z0 = PR(DK_IPZ1,p);
    const double vx0 = PR(DK_IPVX1,p), vy0 = PR(DK_IPVY1,p), vz0 = PR(DK_IPVZ1,p);
    const double xh = x0 + 0.5*dt*vx0;
    const double yh = y0 + 0.5*dt*(vy0 - (sbox ? qs*om*x0 : 0.0));
    const double zh = z0 + 0.5*dt*vz0;
    const double yg = yh + (sbox ? qs*om*xh*0.5*dt : 0.0);
    Stencil w;
    MakeStencil(g, b, xh, yg, zh, w);
    if (!InPad(g, w)) {dk::RecordError(err, DK_EHALO, p); continue;}
    double ugx = 0.0, ugy = 0.0, ugz = 0.0;
    for (int c=0; c<=2; ++c)
      for (int bb=0; bb<3; ++bb) {
        const double wcb = w.wz[c]*w.wy[bb];
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
          ugx += ww*us[Idx(g,m,3,0,kk,jj,ii)];
          ugy += ww*us[Idx(g,m,3,1,kk,jj,ii)];
          ugz += ww*us[Idx(g,m,3,2,kk,jj,ii)];
        }
      }
    const double h = 0.5*dt;
    const double lambda = h/PR(DK_IPTS,p);
    const double aa = 1.0 + lambda;
    const double rx = vx0 + lambda*ugx;
    const double ry = vy0 + lambda*ugy;
    double rz = vz0 + lambda*ugz;
    double vxh, vyh, vzh;
    if (sbox) {
      const double cxy = 2.0*h*om;
      const double cyx = h*(2.0-qs)*om;
      const double det = aa*aa + cxy*cyx;
      vxh = (aa*rx + cxy*ry)/det;
      vyh = (aa*ry - cyx*rx)/det;
      if (strat) rz -= h*(om*om)*zh;
      vzh = rz/aa;
    } else {
      vxh = rx/aa; vyh = ry/aa; vzh = rz/aa;
    }
    PR(DK_IPVX,p) = 2.0*vxh - vx0;
    PR(DK_IPVY,p) = 2.0*vyh - vy0;
    PR(DK_IPVZ,p) = 2.0*vzh - vz0;
    PR(DK_IPX,p) = x0 + dt*vxh;
    PR(DK_IPY,p) = y0 + dt*(vyh - (sbox ? qs*om*xh : 0.0));
    PR(DK_IPZ,p) = z0 + dt*vzh;
    PR(DK_IPX1,p) = xh; PR(DK_IPY1,p) = yh; PR(DK_IPZ1,p) = zh;
    const double ts = PR(DK_IPTS,p);
    const double dvx = dt*(ugx-vxh)/ts, dvy = dt*(ugy-vyh)/ts, dvz = dt*(ugz-vzh)/ts;
    PR(DK_IPRX,p) = dvx; PR(DK_IPRY,p) = dvy; PR(DK_IPRZ,p) = dvz;
    if (br) {
      const double vol = b.dx1*b.dx2*b.dx3;
      const double fac = -PR(DK_IPM,p)/vol;
      for (int c=0; c<=2; ++c)
        for (int bb=0; bb<3; ++bb) {
          const double wcb = w.wz[c]*w.wy[bb]*fac;
          if (wcb == 0.0) continue;
          for (int a=0; a<3; ++a) {
            const double ww = wcb*w.wx[a];
            const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
            AtomicAdd(&dmom[Idx(g,m,3,0,kk,jj,ii)], ww*dvx);
            AtomicAdd(&dmom[Idx(g,m,3,1,kk,jj,ii)], ww*dvy);
            AtomicAdd(&dmom[Idx(g,m,3,2,kk,jj,ii)], ww*dvz);
          }
        }
    }
  }
}

// dust.cpp:1678-1702 (dust_shear_route_time)
__global__ void k_shift_shear_crossings(const DevBlock *B, double *r, const int32_t *iv,
                                        long long cap, long long np, double dsh, double xa,
                                        double xb) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const double b0 = B[m].x1min, b1 = B[m].x1max, lx = b1 - b0;
    const int ix = static_cast<int>((PR(DK_IPX,p) - b0 + lx)/lx) - 1;
    const bool lo1 = (ix < 0) && (b0 < xa + 0.5*lx);
    const bool hi1 = (ix > 0) && (b1 > xb - 0.5*lx);
    if (lo1 || hi1) {
      const double d = hi1 ? dsh : -dsh;
      PR(DK_IPY,p) += d;
      PR(DK_IPY1,p) += d;
    }
  }
}

// bvals_part.cpp:164-429 (part_update) on a uniform mesh of one rank
__global__ void k_part_update(Geo g, const DevBlock *B, double *r, int32_t *iv, long long cap,
                              long long np, int shear, double ysh, double *x3rem,
                              dk::SmrTree tree, dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    const double z3a = g.x3min, z3b = g.x3max;
    if (g.bc_x3 != 0 && (PR(DK_IPZ,p) < z3a || PR(DK_IPZ,p) >= z3b)) {
      const bool lo = (PR(DK_IPZ,p) < z3a);
      if (g.bc_x3 == DK_X3_REFLECT) {
        const double zf = lo ? z3a : z3b;
        double z = 2.0*zf - PR(DK_IPZ,p);
        if (z >= z3b) z = z3b - 1.0e-12*(z3b - z3a);
        if (z < z3a) z = z3a;
        PR(DK_IPZ,p) = z;
        PR(DK_IPVZ,p) = -PR(DK_IPVZ,p);
        PR(DK_IPZ1,p) = 2.0*zf - PR(DK_IPZ1,p);
        PR(DK_IPVZ1,p) = -PR(DK_IPVZ1,p);
      } else {
        AtomicAdd(&x3rem[0], PR(DK_IPM,p));
        AtomicAdd(&x3rem[1], 1.0);
        PR(DK_IPM,p) = 0.0;
        const double dz = b.dx3;
        const double z = lo ? (b.x3min + 0.5*dz) : (b.x3max - 0.5*dz);
        PR(DK_IPZ,p) = z;
        PR(DK_IPVX,p) = 0.0; PR(DK_IPVY,p) = 0.0; PR(DK_IPVZ,p) = 0.0;
        PR(DK_IPX1,p) = PR(DK_IPX,p); PR(DK_IPY1,p) = PR(DK_IPY,p); PR(DK_IPZ1,p) = z;
        PR(DK_IPVX1,p) = 0.0; PR(DK_IPVY1,p) = 0.0; PR(DK_IPVZ1,p) = 0.0;
      }
    }
    double x1 = PR(DK_IPX,p), x2 = PR(DK_IPY,p), x3 = PR(DK_IPZ,p);
    const double lx = b.x1max - b.x1min, ly = b.x2max - b.x2min, lz = b.x3max - b.x3min;
    const int ix = static_cast<int>((x1 - b.x1min + lx)/lx) - 1;
    const int iy = static_cast<int>((x2 - b.x2min + ly)/ly) - 1;
    const int iz = static_cast<int>((x3 - b.x3min + lz)/lz) - 1;
    // snap onto [xmin, xmax) of the block when the offset keeps the particle there
    auto snap = [&](const int off, double &x, const double xa, const double xb, const int ip,
                    const int ir) {
      if (off != 0) return;
      double xs = x;
      if (x < xa) {
        xs = xa;
      } else if (x >= xb) {
        xs = nextafter(xb, xa);
      }
      if (xs != x) {
        PR(ir,p) += xs - x;
        PR(ip,p) = xs;
        x = xs;
      }
    };
    snap(ix, x1, b.x1min, b.x1max, DK_IPX, DK_IPX1);
    snap(iy, x2, b.x2min, b.x2max, DK_IPY, DK_IPY1);
    snap(iz, x3, b.x3min, b.x3max, DK_IPZ, DK_IPZ1);
    if (abs(ix) > 1 || abs(iy) > 1 || abs(iz) > 1) {
      dk::RecordError(err, DK_EROUTE, p);
      continue;
    }
    if ((abs(ix) + abs(iy) + abs(iz)) == 0) continue;
    const bool lo1 = (ix < 0) && (b.x1min < g.x1min + 0.5*lx);
    const bool hi1 = (ix > 0) && (b.x1max > g.x1max - 0.5*lx);
    const bool lo2 = (iy < 0) && (b.x2min < g.x2min + 0.5*ly);
    const bool hi2 = (iy > 0) && (b.x2max > g.x2max - 0.5*ly);
    const bool lo3 = (iz < 0) && (b.x3min < g.x3min + 0.5*lz);
    const bool hi3 = (iz > 0) && (b.x3max > g.x3max - 0.5*lz);
    const bool cross_in = shear && lo1, cross_out = shear && hi1;
    if (cross_in || cross_out) {
      const double lxm = g.x1max - g.x1min;
      PR(DK_IPX,p) += cross_in ? lxm : -lxm;
      PR(DK_IPX1,p) += cross_in ? lxm : -lxm;
      const double lym = g.x2max - g.x2min;
      double ynew = PR(DK_IPY,p) + (cross_in ? -ysh : ysh);
      if (ynew < g.x2min) ynew += lym;
      if (ynew >= g.x2max) ynew -= lym;
      const double dy = ynew - PR(DK_IPY,p);
      PR(DK_IPY,p) = ynew;
      PR(DK_IPY1,p) += dy;
      const double lzm = g.x3max - g.x3min;
      if (lo3) {
        PR(DK_IPZ,p) += lzm; PR(DK_IPZ1,p) += lzm;
      } else if (hi3) {
        PR(DK_IPZ,p) -= lzm; PR(DK_IPZ1,p) -= lzm;
      }
      int b2 = static_cast<int>(((PR(DK_IPY,p) - g.x2min)/lym)*static_cast<double>(g.nmbx2));
      b2 = (b2 < 0) ? 0 : ((b2 > g.nmbx2-1) ? (g.nmbx2-1) : b2);
      int b3 = static_cast<int>(((PR(DK_IPZ,p) - g.x3min)/lzm)*static_cast<double>(g.nmbx3));
      b3 = (b3 < 0) ? 0 : ((b3 > g.nmbx3-1) ? (g.nmbx3-1) : b3);
      const int b1 = cross_in ? (g.nmbx1 - 1) : 0;
      PI(DK_PGID,p) = b1 + g.nmbx1*(b2 + g.nmbx2*b3);
      continue;
    }
    int nb;
    if (tree.leaf != nullptr) {   // a refined mesh: the neighbour of any level
      const int off[3] = {ix, iy, iz};
      const double xs[3] = {x1, x2, x3};
      nb = dk::SmrDest(B, tree, m, off, xs);
    } else {
      nb = b.nghbr[Dir(ix, iy, iz)];
    }
    if (nb < 0) {dk::RecordError(err, DK_EROUTE, p); continue;}
    PI(DK_PGID,p) = nb;
    const double lxm = g.x1max - g.x1min, lym = g.x2max - g.x2min, lzm = g.x3max - g.x3min;
    if (lo1) {PR(DK_IPX,p) += lxm; PR(DK_IPX1,p) += lxm;}
    else if (hi1) {PR(DK_IPX,p) -= lxm; PR(DK_IPX1,p) -= lxm;}
    if (lo2) {PR(DK_IPY,p) += lym; PR(DK_IPY1,p) += lym;}
    else if (hi2) {PR(DK_IPY,p) -= lym; PR(DK_IPY1,p) -= lym;}
    if (lo3) {PR(DK_IPZ,p) += lzm; PR(DK_IPZ1,p) += lzm;}
    else if (hi3) {PR(DK_IPZ,p) -= lzm; PR(DK_IPZ1,p) -= lzm;}
  }
}

// dust_pm.cpp:225-274 (dust_scatter)
__global__ void k_scatter_imex(Geo g, const DevBlock *B, const double *r, const int32_t *iv,
                               long long cap, long long np, double *qdep, double a_dt,
                               double sh, dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    Stencil w;
    MakeStencil(g, b, PR(DK_IPX,p), PR(DK_IPY,p) + sh*PR(DK_IPX,p), PR(DK_IPZ,p), w);
    if (!HaloOK(g, w)) {dk::RecordError(err, DK_EHALO, p); continue;}
    const double vol = b.dx1*b.dx2*b.dx3;
    const double cj = a_dt/(PR(DK_IPTS,p) + a_dt);
    const double muc = PR(DK_IPM,p)*cj/vol;
    const double vx = PR(DK_IPVX,p), vy = PR(DK_IPVY,p), vz = PR(DK_IPVZ,p);
    for (int c=0; c<=2; ++c)
      for (int bb=0; bb<3; ++bb) {
        const double wcb = w.wz[c]*w.wy[bb]*muc;
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
          AtomicAdd(&qdep[Idx(g,m,5,0,kk,jj,ii)], ww);
          AtomicAdd(&qdep[Idx(g,m,5,1,kk,jj,ii)], ww*vx);
          AtomicAdd(&qdep[Idx(g,m,5,2,kk,jj,ii)], ww*vy);
          AtomicAdd(&qdep[Idx(g,m,5,3,kk,jj,ii)], ww*vz);
        }
      }
  }
}

// dust_pm.cpp:141-193 (dust_pc2_predictor)
__global__ void k_pc2_predictor(Geo g, const DevBlock *B, const double *r, const int32_t *iv,
                                long long cap, long long np, const double *u1, double *qdep,
                                double dt, int rate_dep, dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    Stencil w;
    MakeStencil(g, b, PR(DK_IPX,p), PR(DK_IPY,p), PR(DK_IPZ,p), w);
    if (!InPad(g, w)) {dk::RecordError(err, DK_EHALO, p); continue;}
    double ugx = 0.0, ugy = 0.0, ugz = 0.0;
    for (int c=0; c<=2; ++c)
      for (int bb=0; bb<3; ++bb) {
        const double wcb = w.wz[c]*w.wy[bb];
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
          const double invrho = 1.0/u1[Idx(g,m,DK_NGAS,DK_IDN,kk,jj,ii)];
          ugx += ww*u1[Idx(g,m,DK_NGAS,DK_IM1,kk,jj,ii)]*invrho;
          ugy += ww*u1[Idx(g,m,DK_NGAS,DK_IM2,kk,jj,ii)]*invrho;
          ugz += ww*u1[Idx(g,m,DK_NGAS,DK_IM3,kk,jj,ii)]*invrho;
        }
      }
    const double vol = b.dx1*b.dx2*b.dx3;
    const double rate = PR(DK_IPM,p)/(PR(DK_IPTS,p)*vol);
    const double gx = -dt*rate*(ugx-PR(DK_IPVX,p));
    const double gy = -dt*rate*(ugy-PR(DK_IPVY,p));
    const double gz = -dt*rate*(ugz-PR(DK_IPVZ,p));
    for (int c=0; c<=2; ++c)
      for (int bb=0; bb<3; ++bb) {
        const double wcb = w.wz[c]*w.wy[bb];
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
          AtomicAdd(&qdep[Idx(g,m,5,1,kk,jj,ii)], ww*gx);
          AtomicAdd(&qdep[Idx(g,m,5,2,kk,jj,ii)], ww*gy);
          AtomicAdd(&qdep[Idx(g,m,5,3,kk,jj,ii)], ww*gz);
          if (rate_dep) AtomicAdd(&qdep[Idx(g,m,5,4,kk,jj,ii)], ww*rate);
        }
      }
  }
}

//----------------------------------------------------------------------------------------
// dust_pm.cpp:426-436 / dust_solver.cpp:643-649 (dust_ustar, dust_solver_local_guess)
__global__ void k_local_guess(Geo g, const double *u0, const double *qdep, double *us, int br) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double qd0 = br ? qdep[Idx(g,m,5,0,k,j,i)] : 0.0;
    const double denom = 1.0/(u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] + qd0);
    for (int c=0; c<3; ++c) {
      const double qd = br ? qdep[Idx(g,m,5,1+c,k,j,i)] : 0.0;
      us[Idx(g,m,3,c,k,j,i)] = (u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] + qd)*denom;
    }
  }
}

// dust_solver.cpp:151-201 (dust_applya_gtsg)
__global__ void k_applya_gtsg(Geo g, const DevBlock *B, const double *r, const int32_t *iv,
                              long long cap, long long np, const double *field, double *res,
                              double a_dt, double sh, dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    Stencil w;
    MakeStencil(g, b, PR(DK_IPX,p), PR(DK_IPY,p) + sh*PR(DK_IPX,p), PR(DK_IPZ,p), w);
    if (!InPad(g, w)) {dk::RecordError(err, DK_EHALO, p); continue;}
    double gth[3] = {0.0, 0.0, 0.0};
    for (int c=0; c<=2; ++c)
      for (int bb=0; bb<3; ++bb) {
        const double wcb = w.wz[c]*w.wy[bb];
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
          gth[0] += ww*field[Idx(g,m,3,0,kk,jj,ii)];
          gth[1] += ww*field[Idx(g,m,3,1,kk,jj,ii)];
          gth[2] += ww*field[Idx(g,m,3,2,kk,jj,ii)];
        }
      }
    const double vol = b.dx1*b.dx2*b.dx3;
    const double cj = a_dt/(PR(DK_IPTS,p) + a_dt);
    const double fac = PR(DK_IPM,p)*cj/vol;
    for (int c=0; c<=2; ++c)
      for (int bb=0; bb<3; ++bb) {
        const double wcb = w.wz[c]*w.wy[bb]*fac;
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
          AtomicAdd(&res[Idx(g,m,3,0,kk,jj,ii)], ww*gth[0]);
          AtomicAdd(&res[Idx(g,m,3,1,kk,jj,ii)], ww*gth[1]);
          AtomicAdd(&res[Idx(g,m,3,2,kk,jj,ii)], ww*gth[2]);
        }
      }
  }
}

// dust_solver.cpp:243-260 (dust_applya_dc / dust_applya_rho)
__global__ void k_applya_cell(Geo g, const double *u0, const double *qdep, double *field,
                              double *res, int correct) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double rho = u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
    if (correct) {
      const double pinv = 1.0/(rho + qdep[Idx(g,m,5,0,k,j,i)]);
      for (int d=0; d<3; ++d) {
        const double av = res[Idx(g,m,3,d,k,j,i)] + rho*field[Idx(g,m,3,d,k,j,i)];
        const double rv = u0[Idx(g,m,DK_NGAS,DK_IM1+d,k,j,i)] + qdep[Idx(g,m,5,1+d,k,j,i)] - av;
        field[Idx(g,m,3,d,k,j,i)] += pinv*rv;
      }
    } else {
      for (int d=0; d<3; ++d) res[Idx(g,m,3,d,k,j,i)] += rho*field[Idx(g,m,3,d,k,j,i)];
    }
  }
}

// UpdateStiffness (dust_pm.cpp:288-342): the largest pc2 rate/rho over the active cells (q: the
// rate component comp of an nv-component deposit field), one partial per workgroup
template <typename T>
__global__ void k_stiff_partial(Geo g, const T *q, int nv, int comp, const double *u0,
                                double *partial) {
  __shared__ double sh[kBlock];
  double mx = 0.0;
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    mx = fmax(mx, static_cast<double>(q[Idx(g,m,nv,comp,k,j,i)])/u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)]);
  }
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int h = kBlock/2; h > 0; h >>= 1) {
    if (threadIdx.x < h) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// chi = dt max, the running maximum and the rate maximum (rec[0..2]; dust_pm.cpp:315-331)
__global__ void k_stiff_record(const double *partial, int nblk, double dt, double *rec) {
  if (threadIdx.x != 0) return;
  double r = 0.0;
  for (int b=0; b<nblk; ++b) r = fmax(r, partial[b]);
  const double chi = dt*r;
  rec[0] = chi;
  rec[1] = (rec[1] < chi) ? chi : rec[1];
  rec[2] = r;
  // the interval records of the particle history (since its last row; -1: none yet)
  rec[5] = (rec[5] < chi) ? chi : rec[5];
  if (r > 0.0) rec[6] = (rec[6] < r) ? r : rec[6];
}

// RefreshStoppingTimeMaximum (dust_newdt.cpp:34-98): per workgroup the largest and smallest
// stopping time, an invalid one (not finite, not positive, species index negative) recorded
__global__ void k_stop_partial(const double *r, const int32_t *iv, long long cap, long long np,
                               double *partial, dk::ErrWord *err) {
  __shared__ double smax[kBlock], smin[kBlock];
  double mx = 0.0, mn = 3.4028234663852886e38;
  GRID_STRIDE(p, np) {
    const double ts = PR(DK_IPTS,p);
    if (!isfinite(ts) || !(ts > 0.0) || PI(DK_PSP,p) < 0) {dk::RecordError(err, DK_ESTOP, p); continue;}
    mx = fmax(mx, ts);
    mn = fmin(mn, ts);
  }
  smax[threadIdx.x] = mx; smin[threadIdx.x] = mn;
  __syncthreads();
  for (int h = kBlock/2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + h]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {partial[2*blockIdx.x] = smax[0]; partial[2*blockIdx.x + 1] = smin[0];}
}

// dust_pm.cpp:387-402 (dust_pc2_apply_predictor)
__global__ void k_pc2_apply_predictor(Geo g, double *u0, double *qdep, int damp, double dt) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    if (damp) {
      const double chi = dt*qdep[Idx(g,m,5,4,k,j,i)]/u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
      if (chi > 1.0) {
        const double s = 1.0/chi;
        for (int c=1; c<=3; ++c) qdep[Idx(g,m,5,c,k,j,i)] *= s;
      }
    }
    for (int c=0; c<3; ++c) u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] += qdep[Idx(g,m,5,1+c,k,j,i)];
  }
}

// dust_pm.cpp:500-564 (dust_gather)
__global__ void k_gather_kick(Geo g, const DevBlock *B, double *r, const int32_t *iv,
                              long long cap, long long np, const double *us, double *dmom,
                              double a_dt, double sh, int br, dk::ErrWord *err) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    Stencil w;
    MakeStencil(g, b, PR(DK_IPX,p), PR(DK_IPY,p) + sh*PR(DK_IPX,p), PR(DK_IPZ,p), w);
    if (!InPad(g, w)) {dk::RecordError(err, DK_EHALO, p); continue;}
    double ux = 0.0, uy = 0.0, uz = 0.0;
    for (int c=0; c<=2; ++c)
      for (int bb=0; bb<3; ++bb) {
        const double wcb = w.wz[c]*w.wy[bb];
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
          ux += ww*us[Idx(g,m,3,0,kk,jj,ii)];
          uy += ww*us[Idx(g,m,3,1,kk,jj,ii)];
          uz += ww*us[Idx(g,m,3,2,kk,jj,ii)];
        }
      }
    const double cj = a_dt/(PR(DK_IPTS,p) + a_dt);
    const double dvx = cj*(ux - PR(DK_IPVX,p));
    const double dvy = cj*(uy - PR(DK_IPVY,p));
    const double dvz = cj*(uz - PR(DK_IPVZ,p));
    PR(DK_IPVX,p) += dvx;
    PR(DK_IPVY,p) += dvy;
    PR(DK_IPVZ,p) += dvz;
    PR(DK_IPRX,p) = dvx/a_dt;
    PR(DK_IPRY,p) = dvy/a_dt;
    PR(DK_IPRZ,p) = dvz/a_dt;
    if (br) {
      const double vol = b.dx1*b.dx2*b.dx3;
      const double fac = -PR(DK_IPM,p)/vol;
      for (int c=0; c<=2; ++c)
        for (int bb=0; bb<3; ++bb) {
          const double wcb = w.wz[c]*w.wy[bb]*fac;
          if (wcb == 0.0) continue;
          for (int a=0; a<3; ++a) {
            const double ww = wcb*w.wx[a];
            const int kk = w.kp+c-1, jj = w.jp+bb-1, ii = w.ip+a-1;
            AtomicAdd(&dmom[Idx(g,m,3,0,kk,jj,ii)], ww*dvx);
            AtomicAdd(&dmom[Idx(g,m,3,1,kk,jj,ii)], ww*dvy);
            AtomicAdd(&dmom[Idx(g,m,3,2,kk,jj,ii)], ww*dvz);
          }
        }
    }
  }
}

// dust_pm.cpp:660-667 (dust_imex_point_center_mass)
__global__ void k_center_mass(Geo g, const DevBlock *B, const double *r, const int32_t *iv,
                              long long cap, long long np, double *res, double a_dt) {
  GRID_STRIDE(p, np) {
    if (PR(DK_IPM,p) <= 0.0) continue;
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    Stencil w;
    MakeStencil(g, b, PR(DK_IPX,p), PR(DK_IPY,p), PR(DK_IPZ,p), w);
    if (!InPad(g, w)) continue;
    const double vol = b.dx1*b.dx2*b.dx3;
    AtomicAdd(&res[Idx(g,m,4,3,w.kp,w.jp,w.ip)], PR(DK_IPM,p)*(a_dt/(PR(DK_IPTS,p) + a_dt))/vol);
  }
}

// dust_pm.cpp:670-685 (dust_imex_point_consistent_gas)
__global__ void k_consistent_gas(Geo g, double *u0, const double *us, const double *qdep,
                                 double *dmom, double *res, double inv_adt) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double rho = u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)], qt = res[Idx(g,m,4,3,k,j,i)];
    const double q = qdep[Idx(g,m,5,0,k,j,i)];
    for (int c=0; c<3; ++c) {
      double dgas = dmom[Idx(g,m,3,c,k,j,i)];
      if (qt > 0.0) {
        const double gg = rho*us[Idx(g,m,3,c,k,j,i)] - u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)];
        const double rr2 = (dmom[Idx(g,m,3,c,k,j,i)] - gg)/(rho + q);
        dgas = gg + rho*rr2;
        res[Idx(g,m,4,c,k,j,i)] = rr2*q/qt;
      }
      u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] += dgas;
      dmom[Idx(g,m,3,c,k,j,i)] = dgas*inv_adt;
    }
  }
}

// dust_pm.cpp:686-696 (dust_imex_point_consistent_dust)
__global__ void k_consistent_dust(Geo g, const DevBlock *B, double *r, const int32_t *iv,
                                  long long cap, long long np, const double *res, double a_dt) {
  GRID_STRIDE(p, np) {
    if (PR(DK_IPM,p) <= 0.0) continue;
    const int m = PI(DK_PGID,p);
    const DevBlock &b = B[m];
    Stencil w;
    MakeStencil(g, b, PR(DK_IPX,p), PR(DK_IPY,p), PR(DK_IPZ,p), w);
    if (!InPad(g, w)) continue;
    const double cj = a_dt/(PR(DK_IPTS,p) + a_dt);
    for (int c=0; c<3; ++c) {
      const double dv = cj*res[Idx(g,m,4,c,w.kp,w.jp,w.ip)];
      PR(DK_IPVX+2*c,p) += dv;
      PR(DK_IPRX+c,p) += dv/a_dt;
    }
  }
}

// dust_pm.cpp:700-708 (dust_pmbr), 596-601 (dust_pc2_commit), 734-739 (dust_gatwid)
__global__ void k_cell_update(Geo g, double *u0, double *dmom, const double *qdep, int mode,
                              double c0) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    for (int c=0; c<3; ++c) {
      double &u = u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)];
      double &d = dmom[Idx(g,m,3,c,k,j,i)];
      if (mode == 0) {         // pmbr: u0 += dmom, dmom *= 1/(a dt)
        u += d;
        d *= c0;
      } else if (mode == 1) {  // pc2 commit
        u += d - c0*qdep[Idx(g,m,5,1+c,k,j,i)];
      } else {                 // drag history
        u += c0*d;
      }
    }
  }
}

// ---- the explicit gas update: isothermal fluxes (dk_gas_src.hydro) ----
//
// reconstruct/plm.hpp:13-28: the limited slope of q (van Leer's harmonic mean on a uniform
// mesh); lo = the right state of the cell's lower face, hi = the left state of its upper face
__device__ inline void PlmFace(double qm, double q0, double qp, double &lo, double &hi) {
  const double dql = q0 - qm, dqr = qp - q0;
  const double dq2 = dql*dqr;
  double dqm = dq2/(dql + dqr);
  if (dq2 <= 0.0) dqm = 0.0;
  hi = q0 + dqm;
  lo = q0 - dqm;
}

// src/hydro/rsolvers/hlle_hyd.hpp:22-150 with the isothermal EOS; states and flux in the face's
// frame (d, v_n, v_t1, v_t2)
__device__ inline void HlleIsoFlux(const double wl[4], const double wr[4], double cs,
                                   double f[4]) {
  const double sdl = sqrt(wl[0]), sdr = sqrt(wr[0]);
  const double isd = 1.0/(sdl + sdr);
  const double vroe = (sdl*wl[1] + sdr*wr[1])*isd;
  const double al = fmin(vroe - cs, wl[1] - cs);
  const double ar = fmax(vroe + cs, wr[1] + cs);
  const double bp = (ar > 0.0) ? ar : 1.0e-20;
  const double bm = (al < 0.0) ? al : -1.0e-20;
  const double qa = wl[1] - bm, qb = wr[1] - bp;
  double fl[4], fr[4];
  fl[0] = wl[0]*qa;
  fr[0] = wr[0]*qb;
#pragma unroll
  for (int c=1; c<4; ++c) {
    fl[c] = wl[0]*wl[c]*qa;
    fr[c] = wr[0]*wr[c]*qb;
  }
  fl[1] += (cs*cs)*wl[0];
  fr[1] += (cs*cs)*wr[0];
  double h = 0.0;
  if (bp != bm) h = 0.5*(bp + bm)/(bp - bm);
#pragma unroll
  for (int c=0; c<4; ++c) f[c] = 0.5*(fl[c] + fr[c]) + h*(fl[c] - fr[c]);
}

// IsothermalHydro::ConsToPrim of one cell (src/eos/isothermal_hyd.cpp:34-48: density floor,
// then the velocities by the reciprocal density); at = the cell's density index
__device__ inline void PrimOf(const double *u, long long at, long long ncb, double dfloor,
                              double w[4]) {
  double d = u[at];
  if (d < dfloor) d = dfloor;
  const double di = 1.0/d;
  w[0] = d;
  w[1] = di*u[at + ncb];
  w[2] = di*u[at + 2*ncb];
  w[3] = di*u[at + 3*ncb];
}

// Hydro::Fluxes (src/hydro/hydro_fluxes.cpp:20-140, plm + hlle) and the flux divergence of
// Hydro::RKUpdate (src/hydro/hydro_update.cpp:55-77): one thread per active cell computes the
// six face fluxes of its cell from the primitives of the 13 cells on the three lines through
// it (u0 ghosts filled).  Both cells beside a face compute its flux with the same operations
// on the same values, so the divergence telescopes as the reference's stored face fluxes do;
// the axes are added in the reference's order x1, x2, x3.
__global__ __launch_bounds__(kBlock, 4) void k_hydro_divf(Geo g, const DevBlock *B,
                                                       const double *u0, double *divf,
                                                       double cs, double dfloor) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  const long long ncb = g.ncb;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const long long at = Idx(g,m,DK_NGAS,DK_IDN,k,j,i);
    const DevBlock &b = B[m];
    double wc[4];
    PrimOf(u0, at, ncb, dfloor, wc);
    double dv[4] = {0.0, 0.0, 0.0, 0.0};
    // one axis at a time (not unrolled: the live state stays within the register budget)
#pragma unroll 1
    for (int d=0; d<3; ++d) {
      const long long sd = (d == 0) ? 1 : ((d == 1) ? (long long)g.n1 : (long long)g.n1*g.n2);
      const double dx = (d == 0) ? b.dx1 : ((d == 1) ? b.dx2 : b.dx3);
      // primitives of the cells -2..2 along the axis in the face frame (d, v_n, v_t1, v_t2)
      double w[5][4];
#pragma unroll
      for (int o=0; o<5; ++o) {
        double p[4];
        if (o == 2) {
#pragma unroll
          for (int v=0; v<4; ++v) p[v] = wc[v];
        } else {
          PrimOf(u0, at + (o - 2)*sd, ncb, dfloor, p);
        }
        w[o][0] = p[0];
        w[o][1] = (d == 0) ? p[1] : ((d == 1) ? p[2] : p[3]);
        w[o][2] = (d == 0) ? p[2] : ((d == 1) ? p[3] : p[1]);
        w[o][3] = (d == 0) ? p[3] : ((d == 1) ? p[1] : p[2]);
      }
      // face states: the lower face (cells -1 | 0) and the upper face (cells 0 | +1)
      double wlm[4], wrm[4], wlp[4], wrp[4];
#pragma unroll
      for (int v=0; v<4; ++v) {
        double unused;
        PlmFace(w[0][v], w[1][v], w[2][v], unused, wlm[v]);
        PlmFace(w[1][v], w[2][v], w[3][v], wrm[v], wlp[v]);
        PlmFace(w[2][v], w[3][v], w[4][v], wrp[v], unused);
      }
      double fm[4], fp[4];
      HlleIsoFlux(wlm, wrm, cs, fm);
      HlleIsoFlux(wlp, wrp, cs, fp);
      double r[4];
#pragma unroll
      for (int v=0; v<4; ++v) r[v] = (fp[v] - fm[v])/dx;
      // back to the gas components; the axes add in the order x1, x2, x3
      dv[0] += r[0];
      dv[1] += (d == 0) ? r[1] : ((d == 1) ? r[3] : r[2]);
      dv[2] += (d == 0) ? r[2] : ((d == 1) ? r[1] : r[3]);
      dv[3] += (d == 0) ? r[3] : ((d == 1) ? r[2] : r[1]);
    }
#pragma unroll
    for (int c=0; c<4; ++c) divf[at + c*ncb] = dv[c];
  }
}

// Hydro::RKUpdate (zero fluxes, or minus beta dt divf when divf is given) + HydroSrcTerms
// (see dustk.h dk_gas_src); with dmom, then AddDragHistoryGas's u0.M += atw dmom
// (dust_pm.cpp:721-742), for a caller whose drag history follows the gas update directly
__global__ void k_gas_hook(Geo g, const DevBlock *B, double *u0, const double *u1, double gam0,
                           double gam1, double bdt, int use_src, dk_gas_src src,
                           const double *divf, const double *dmom, double atw) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    double w[4];
    double den = u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
    if (divf != nullptr && den < src.dfloor) {
      den = src.dfloor;
      u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] = den;
    }
    w[0] = den;
    for (int c=1; c<4; ++c) w[c] = u0[Idx(g,m,DK_NGAS,c,k,j,i)]/den;
    for (int v=0; v<DK_NGAS; ++v) {
      double &u = u0[Idx(g,m,DK_NGAS,v,k,j,i)];
      if (divf != nullptr)
        u = gam0*u + gam1*u1[Idx(g,m,DK_NGAS,v,k,j,i)] - bdt*divf[Idx(g,m,DK_NGAS,v,k,j,i)];
      else
        u = gam0*u + gam1*u1[Idx(g,m,DK_NGAS,v,k,j,i)];
    }
    if (use_src && src.const_accel != 0.0)
      u0[Idx(g,m,DK_NGAS,src.const_accel_dir,k,j,i)] += bdt*src.const_accel*w[0];
    if (use_src && src.shearing_box) {
      const double coef1 = 2.0*bdt*src.omega0;
      const double coef2 = (2.0-src.qshear)*bdt*src.omega0;
      const double coef3 = bdt*(src.omega0*src.omega0);
      const double mom1 = w[0]*w[1], mom2 = w[0]*w[2];
      u0[Idx(g,m,DK_NGAS,DK_IM1,k,j,i)] += coef1*mom2;
      u0[Idx(g,m,DK_NGAS,DK_IM2,k,j,i)] -= coef2*mom1;
      if (src.stratified) {
        const DevBlock &b = B[m];
        const double xx = (static_cast<double>(k - g.ks) + 0.5)/g.nx3;
        const double x3v = (xx*b.x3max - xx*b.x3min) - (0.5*b.x3max - 0.5*b.x3min) +
                           (0.5*b.x3min + 0.5*b.x3max);
        u0[Idx(g,m,DK_NGAS,DK_IM3,k,j,i)] -= coef3*w[0]*x3v;
      }
    }
    if (dmom != nullptr)
      for (int c=0; c<3; ++c) u0[Idx(g,m,DK_NGAS,1+c,k,j,i)] += atw*dmom[Idx(g,m,3,c,k,j,i)];
  }
}

// FP64 sums of the particles and active cells; fixed grid, partials reduced in order
constexpr int kDiagGrid = 512;
__global__ void k_diag_part(Geo g, const DevBlock *B, const double *r, long long cap,
                            long long np, const double *u0, double *partial) {
  __shared__ double sh[kBlock][9];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  GRID_STRIDE(p, np) {
    const double m = PR(DK_IPM,p);
    acc[0] += m*PR(DK_IPVX,p);
    acc[1] += m*PR(DK_IPVY,p);
    acc[2] += m*PR(DK_IPVZ,p);
    acc[6] += m;
    acc[8] += m*(fabs(PR(DK_IPVX,p)) + fabs(PR(DK_IPVY,p)) + fabs(PR(DK_IPVZ,p)));
  }
  const long long nc = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, nc) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double vol = B[m].dx1*B[m].dx2*B[m].dx3;
    acc[7] += vol*u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
    for (int c=0; c<3; ++c) {
      const double mom = u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)];
      acc[3+c] += vol*mom;
      acc[8] += vol*fabs(mom);
    }
  }
  for (int c=0; c<9; ++c) sh[threadIdx.x][c] = acc[c];
  __syncthreads();
  for (int s=kBlock/2; s>0; s>>=1) {
    if (threadIdx.x < s)
      for (int c=0; c<9; ++c) sh[threadIdx.x][c] += sh[threadIdx.x + s][c];
    __syncthreads();
  }
  if (threadIdx.x < 9) partial[blockIdx.x*9 + threadIdx.x] = sh[0][threadIdx.x];
}

__global__ void k_reduce_partials(const double *partial, int nblk, int ncomp, double *out) {
  const int c = threadIdx.x;
  if (c >= ncomp) return;
  double s = 0.0;
  for (int b=0; b<nblk; ++b) s += partial[b*ncomp + c];
  out[c] = s;
}

// dust_newdt.cpp:135-145: min over particles of dx/max(|v|, 1e-30), the azimuthal transport
// velocity with the background shear in a 3-D shearing box; stopping times validated
// (dust_newdt.cpp:34-99)
__global__ void k_newdt(const DevBlock *B, const double *r, const int32_t *iv, long long cap,
                        long long np, double qo, double *partial, dk::ErrWord *err) {
  __shared__ double sh[kBlock];
  double dtmin = 3.4028234663852886e38;   // the largest float, the reference's start value
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    const double ts = PR(DK_IPTS,p);
    if (!isfinite(ts) || !(ts > 0.0)) dk::RecordError(err, DK_ESTOP, p);
    const double vy = PR(DK_IPVY,p) - qo*PR(DK_IPX,p);
    dtmin = fmin(B[m].dx1/fmax(fabs(PR(DK_IPVX,p)), 1.0e-30), dtmin);
    dtmin = fmin(B[m].dx2/fmax(fabs(vy), 1.0e-30), dtmin);
    dtmin = fmin(B[m].dx3/fmax(fabs(PR(DK_IPVZ,p)), 1.0e-30), dtmin);
  }
  sh[threadIdx.x] = dtmin;
  __syncthreads();
  for (int s=kBlock/2; s>0; s>>=1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// cell sort (dust_cloud_device.cpp:1903-1950): key = block*ncell + cell of the position
__global__ void k_sort_keys(Geo g, const DevBlock *B, const double *r, const int32_t *iv,
                            long long cap, long long np, unsigned long long *keys, int *idx) {
  GRID_STRIDE(p, np) {
    const int m = PI(DK_PGID,p);
    Stencil w;
    MakeStencil(g, B[m], PR(DK_IPX,p), PR(DK_IPY,p), PR(DK_IPZ,p), w);
    const long long cell = ((long long)(w.kp - g.ks)*g.nx2 + (w.jp - g.js))*g.nx1 + (w.ip - g.is);
    keys[p] = (unsigned long long)m*g.nx1*g.nx2*g.nx3 + (unsigned long long)cell;
    idx[p] = static_cast<int>(p);
  }
}

__global__ void k_permute(const double *r, const int32_t *iv, long long cap, long long np,
                          const int *order, double *rbuf, int32_t *ibuf) {
  GRID_STRIDE(q, np) {
    const long long p = order[q];
    for (int c=0; c<DK_NRDATA; ++c) rbuf[(long long)c*np + q] = r[(long long)c*cap + p];
    for (int c=0; c<DK_NIDATA; ++c) ibuf[(long long)c*np + q] = iv[(long long)c*cap + p];
  }
}

__global__ void k_unpermute_copy(double *r, int32_t *iv, long long cap, long long np,
                                 const double *rbuf, const int32_t *ibuf) {
  GRID_STRIDE(q, np) {
    for (int c=0; c<DK_NRDATA; ++c) r[(long long)c*cap + q] = rbuf[(long long)c*np + q];
    for (int c=0; c<DK_NIDATA; ++c) iv[(long long)c*cap + q] = ibuf[(long long)c*np + q];
  }
}

#undef PR
#undef PI

//----------------------------------------------------------------------------------------
// host side

bool ActiveStage(const dk_ctx *c, const dk_stage *st) {   // dust.cpp:1646-1649
  if (c->P.coupling == DK_PC2) return true;
  return !(st->nexp_stages == 3 && st->stage == st->nexp_stages);
}

bool MigrationActive(const dk_ctx *c, const dk_stage *st) {   // dust.cpp:1651-1654
  if (c->P.coupling == DK_IMEX) return ActiveStage(c, st);
  return st->stage == 2;
}

double DragStep(const dk_ctx *c, const dk_stage *st) {   // dust.cpp:1704-1707
  if (c->P.coupling == DK_PC2) return st->dt;
  return st->a_impl*st->dt;
}

double ImplicitShearShift(const dk_ctx *c, const dk_stage *st) {   // dust.hpp:1699-1702
  if (!c->P.shearing_box || c->P.coupling != DK_IMEX) return 0.0;
  double cc = 0.0;
  for (int k=0; k<st->stage; ++k) cc = st->gam0[k]*cc + st->beta[k];
  const double timp = st->time + cc*st->dt;
  return c->P.qshear*c->P.omega0*(timp - st->time);
}

double ShearRouteTime(const dk_ctx *c, const dk_stage *st) {   // dust.cpp:1668-1676
  const double t = st->time, dt = st->dt;
  if (c->P.coupling != DK_IMEX) return t + dt;
  double w = 0.0;
  for (int k = st->stage + 1; k <= st->nexp_stages; ++k) w = st->gam0[k-1]*w + st->beta[k-1];
  return t + (1.0 - w)*dt;
}

int Launched(dk_ctx *c) { return dk::Check(hipGetLastError()); }

int Ready(const dk_ctx *c, bool need_particles = true) {
  if (c == nullptr) return DK_EINVAL;
  if (c->form != DK_FORM_FP64) return DK_EUNSUPPORTED;
  if (need_particles && c->rdata == nullptr) return DK_ESTATE;
  if (c->u0 == nullptr) return DK_ESTATE;
  return DK_OK;
}

#define PARGS c->rdata, c->idata, (long long)c->cap, (long long)c->np
#define PGRID Grid(c->np), kBlock, 0, c->stream
#define CGRID(n) Grid(n), kBlock, 0, c->stream

long long ActiveCells(const dk_ctx *c) {
  return (long long)c->g.nmb*c->g.nx3*c->g.nx2*c->g.nx1;
}

int FillGhosts(dk_ctx *c, double *a, int nvar, bool velocity, double t) {
  return dk::FillGhostsT<double>(c, a, nvar, velocity, t);
}

int Fold(dk_ctx *c, double *a, int nvar, double t) { return dk::FoldT<double>(c, a, nvar, t); }

// ApplyCoupledOperator (dust_solver.cpp:218-263) on field into out: out = A field, or with
// correct one defect-correction sweep of field
int ApplyOperatorTo(dk_ctx *c, double *field, double *out, double a_dt, double sh, bool correct,
                    double t) {
  int e;
  if ((e = FillGhosts(c, field, 3, true, t))) return e;
  if ((e = dk::Zero(c, out, 3*dk::NCells(c)))) return e;
  {const int _pt = dk::ProfBegin(c, "applya_gtsg");
  k_applya_gtsg<<<PGRID>>>(c->g, c->d_blocks, PARGS, field, out, a_dt, sh, c->d_err);
   dk::ProfEnd(c, _pt);}
  if ((e = Fold(c, out, 3, t))) return e;
  {const int _pt = dk::ProfBegin(c, "applya_cell");
  k_applya_cell<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, field, out, correct ? 1 : 0);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

int ApplyCoupledOperator(dk_ctx *c, double a_dt, double sh, bool correct, double t) {
  return ApplyOperatorTo(c, c->ustar, c->ap, a_dt, sh, correct, t);
}

// ---- the strict solver (dust_solver.cpp:268-625) ------------------------------------------
// Per-workgroup partial sums over the active cells (three per workgroup, in a fixed tree order);
// the host adds the kDiagGrid partials in order.  mode 0: vol a.b (GlobalDot), 1: vol a^2/(rho
// + Q) (GlobalPreconditionedNorm), 2: the same of b = M + P (the right-hand side), 3: [0] max
// Q/rho, [1] vol rho |a|^2 (the gas state scale)
__global__ void k_solver_sums(Geo g, const DevBlock *B, const double *u0, const double *qdep,
                              const double *a, const double *b, int mode, double *partial) {
  __shared__ double sh[3][kBlock];
  double acc[3] = {0.0, 0.0, 0.0};
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double vol = (B[m].dx1*B[m].dx2)*B[m].dx3;
    const double rho = u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
    const double pd = rho + qdep[Idx(g,m,5,0,k,j,i)];
    if (mode == 3) {
      acc[0] = fmax(acc[0], qdep[Idx(g,m,5,0,k,j,i)]/rho);
      double u2 = 0.0;
      for (int d=0; d<3; ++d) {const double v = a[Idx(g,m,3,d,k,j,i)]; u2 += v*v;}
      acc[1] += vol*rho*u2;
      continue;
    }
    for (int d=0; d<3; ++d) {
      const long long id = Idx(g,m,3,d,k,j,i);
      if (mode == 0) {
        acc[d] += vol*a[id]*b[id];
      } else {
        const double v = (mode == 1) ? a[id] : u0[Idx(g,m,DK_NGAS,DK_IM1+d,k,j,i)] +
                                                qdep[Idx(g,m,5,1+d,k,j,i)];
        acc[d] += vol/pd*v*v;
      }
    }
  }
  for (int d=0; d<3; ++d) sh[d][threadIdx.x] = acc[d];
  __syncthreads();
  for (int s=kBlock/2; s>0; s>>=1) {
    if (threadIdx.x < s)
      for (int d=0; d<3; ++d)
        sh[d][threadIdx.x] = (mode == 3 && d == 0) ? fmax(sh[d][threadIdx.x], sh[d][threadIdx.x + s])
                                                   : sh[d][threadIdx.x] + sh[d][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[3*blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

#define PR(c, p) r[(long long)(c)*cap + (p)]
#define PI(c, p) iv[(long long)(c)*cap + (p)]
// the particle side of the adaptive bound: [0] max a_dt/(t_s + a_dt), [1] sum m |v + c (G x -
// v)|^2 with the gather at the particle's position (dust_solver.cpp:548-591)
__global__ void k_solver_particles(Geo g, const DevBlock *B, const double *r, const int32_t *iv,
                                   long long cap, long long np, const double *x, double a_dt,
                                   double *partial) {
  __shared__ double sh[2][kBlock];
  double acc[2] = {0.0, 0.0};
  GRID_STRIDE(p, np) {
    const double ts = PR(DK_IPTS,p);
    acc[0] = fmax(acc[0], a_dt/(ts + a_dt));
    const int m = PI(DK_PGID,p);
    Stencil w;
    MakeStencil(g, B[m], PR(DK_IPX,p), PR(DK_IPY,p), PR(DK_IPZ,p), w);
    double gu[3] = {0.0, 0.0, 0.0};
    if (InPad(g, w))
      for (int cc=0; cc<3; ++cc)
        for (int bb=0; bb<3; ++bb)
          for (int a=0; a<3; ++a) {
            const double ww = w.wz[cc]*w.wy[bb]*w.wx[a];
            for (int d=0; d<3; ++d)
              gu[d] += ww*x[Idx(g,m,3,d,w.kp+cc-1,w.jp+bb-1,w.ip+a-1)];
          }
    const double cj = a_dt/(ts + a_dt);
    double v2 = 0.0;
    for (int d=0; d<3; ++d) {
      const double v = PR(DK_IPVX + 2*d,p) + cj*(gu[d] - PR(DK_IPVX + 2*d,p));
      v2 += v*v;
    }
    acc[1] += PR(DK_IPM,p)*v2;
  }
  sh[0][threadIdx.x] = acc[0];
  sh[1][threadIdx.x] = acc[1];
  __syncthreads();
  for (int s=kBlock/2; s>0; s>>=1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) partial[2*blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}
#undef PR
#undef PI

// mode 0: r = b - ap, p = r/(rho + Q); 1: p = r/(rho + Q); 2: r = b - ap
__global__ void k_pcg_residual(Geo g, const double *u0, const double *qdep, const double *ap,
                               double *r, double *p, int mode) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double pinv = 1.0/(u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] + qdep[Idx(g,m,5,0,k,j,i)]);
    for (int d=0; d<3; ++d) {
      const long long id = Idx(g,m,3,d,k,j,i);
      if (mode != 1)
        r[id] = u0[Idx(g,m,DK_NGAS,DK_IM1+d,k,j,i)] + qdep[Idx(g,m,5,1+d,k,j,i)] - ap[id];
      if (mode != 2) p[id] = pinv*r[id];
    }
  }
}

// x += alpha p, r -= alpha ap (per component)
__global__ void k_pcg_xr(Geo g, double *x, double *r, const double *p, const double *ap,
                         double a0, double a1, double a2) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double al[3] = {a0, a1, a2};
    for (int d=0; d<3; ++d) {
      const long long id = Idx(g,m,3,d,k,j,i);
      x[id] += al[d]*p[id];
      r[id] -= al[d]*ap[id];
    }
  }
}

// p = r/(rho + Q) + beta p, or 0 for a converged component
__global__ void k_pcg_p(Geo g, const double *u0, const double *qdep, const double *r, double *p,
                        double b0, double b1, double b2, int done) {
  const long long n = (long long)g.nmb*g.nx3*g.nx2*g.nx1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = g.is + static_cast<int>(rr % g.nx1); rr /= g.nx1;
    const int j = g.js + static_cast<int>(rr % g.nx2); rr /= g.nx2;
    const int k = g.ks + static_cast<int>(rr % g.nx3);
    const int m = static_cast<int>(rr/g.nx3);
    const double pinv = 1.0/(u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] + qdep[Idx(g,m,5,0,k,j,i)]);
    const double be[3] = {b0, b1, b2};
    for (int d=0; d<3; ++d) {
      const long long id = Idx(g,m,3,d,k,j,i);
      p[id] = ((done >> d) & 1) ? 0.0 : pinv*r[id] + be[d]*p[id];
    }
  }
}

// the kDiagGrid partials of k_solver_sums summed on the host in order
int SolverSums(dk_ctx *c, const double *a, const double *b, int mode, double out[3]) {
  k_solver_sums<<<kDiagGrid, kBlock, 0, c->stream>>>(c->g, c->d_blocks, c->u0, c->qdep, a, b,
                                                     mode, c->d_partial);
  std::vector<double> h(3*kDiagGrid);
  int e = dk::Check(hipMemcpyAsync(h.data(), c->d_partial, sizeof(double)*h.size(),
                                   hipMemcpyDeviceToHost, c->stream));
  if (!e) e = dk::Check(hipStreamSynchronize(c->stream));
  if (e) return e;
  out[0] = out[1] = out[2] = 0.0;
  for (int q=0; q<kDiagGrid; ++q)
    for (int d=0; d<3; ++d)
      out[d] = (mode == 3 && d == 0) ? std::fmax(out[d], h[3*q + d]) : out[d] + h[3*q + d];
  return DK_OK;
}

// StrictPCG (dust_solver.cpp:350-526): the completed updates in *iters, DK_ESOLVER for the
// reference's SolverFatal
int StrictPCG(dk_ctx *c, double a_dt, double sh, double t, bool residual_is_current,
              int *iters) {
  double *x = c->ustar, *r = c->solver_r, *p = c->solver_p, *ap = c->ap;
  int e;
  if (!residual_is_current) {
    if ((e = ApplyOperatorTo(c, x, ap, a_dt, sh, false, t))) return e;
    k_pcg_residual<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, ap, r, p, 0);
  } else {
    k_pcg_residual<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, ap, r, p, 1);
  }
  double rz[3], bn2[3], target[3];
  bool conv[3];
  if ((e = SolverSums(c, r, p, 0, rz))) return e;
  if ((e = SolverSums(c, nullptr, nullptr, 2, bn2))) return e;
  for (int d=0; d<3; ++d) {
    target[d] = c->drag_atol + c->drag_rtol*std::sqrt(std::max(bn2[d], 0.0));
    conv[d] = std::sqrt(std::max(rz[d], 0.0)) <= target[d];
  }
  for (int it=0; it<=c->drag_iter_max; ++it) {
    if (conv[0] && conv[1] && conv[2]) {
      if ((e = ApplyOperatorTo(c, x, ap, a_dt, sh, false, t))) return e;
      k_pcg_residual<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, ap, r, p, 2);
      double tn2[3];
      if ((e = SolverSums(c, r, nullptr, 1, tn2))) return e;
      bool all = true;
      c->solver_last[2] = 0.0;
      for (int d=0; d<3; ++d) {
        const double nrm = std::sqrt(std::max(tn2[d], 0.0));
        c->solver_last[2] = std::max(c->solver_last[2], nrm);
        conv[d] = nrm <= target[d];
        all = all && conv[d];
        rz[d] = tn2[d];
      }
      if (all) {*iters = it; return DK_OK;}
      k_pcg_residual<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, ap, r, p, 1);
    }
    if (it == c->drag_iter_max) break;
    if ((e = ApplyOperatorTo(c, p, ap, a_dt, sh, false, t))) return e;
    double pap[3], alpha[3];
    if ((e = SolverSums(c, p, ap, 0, pap))) return e;
    for (int d=0; d<3; ++d) {
      if (!conv[d] && (!(pap[d] > 0.0) || !std::isfinite(pap[d]))) return DK_ESOLVER;
      alpha[d] = conv[d] ? 0.0 : rz[d]/pap[d];
    }
    k_pcg_xr<<<CGRID(ActiveCells(c))>>>(c->g, x, r, p, ap, alpha[0], alpha[1], alpha[2]);
    double rzn[3], beta[3];
    int done = 0;
    if ((e = SolverSums(c, r, nullptr, 1, rzn))) return e;
    bool now[3];
    for (int d=0; d<3; ++d) {
      now[d] = std::sqrt(std::max(rzn[d], 0.0)) <= target[d];
      beta[d] = (conv[d] || now[d]) ? 0.0 : rzn[d]/rz[d];
      if (now[d]) done |= 1 << d;
    }
    k_pcg_p<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, r, p, beta[0], beta[1], beta[2],
                                       done);
    for (int d=0; d<3; ++d) {conv[d] = now[d]; rz[d] = rzn[d];}
  }
  return DK_ESOLVER;
}

// the adaptive path after its DC1 sweep (dust_solver.cpp:664-686, 529-624): the true residual,
// the error bound and the acceptance target; PCG from the residual when the bound exceeds it
int AdaptiveSolve(dk_ctx *c, double a_dt, double sh, double outer_dt, double t) {
  int e;
  if ((e = ApplyOperatorTo(c, c->ustar, c->ap, a_dt, sh, false, t))) return e;
  k_pcg_residual<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, c->ap, c->solver_r,
                                            c->solver_p, 2);
  double n2[3];
  if ((e = SolverSums(c, c->solver_r, nullptr, 1, n2))) return e;
  const double res = std::sqrt(std::max(n2[0] + n2[1] + n2[2], 0.0));
  double gas[3];
  if ((e = SolverSums(c, c->ustar, nullptr, 3, gas))) return e;
  const double eps = gas[0];
  k_solver_particles<<<kDiagGrid, kBlock, 0, c->stream>>>(c->g, c->d_blocks, PARGS, c->ustar,
                                                          a_dt, c->d_partial);
  std::vector<double> h(2*kDiagGrid);
  if ((e = dk::Check(hipMemcpyAsync(h.data(), c->d_partial, sizeof(double)*h.size(),
                                    hipMemcpyDeviceToHost, c->stream))))
    return e;
  if ((e = dk::Check(hipStreamSynchronize(c->stream)))) return e;
  double cmax = 0.0, part = 0.0;
  for (int q=0; q<kDiagGrid; ++q) {cmax = std::fmax(cmax, h[2*q]); part += h[2*q + 1];}
  const double scale = std::sqrt(std::max(gas[1] + part, 0.0));
  const double theta = eps/(1.0 + eps);
  const double response = (1.0 + eps)*std::sqrt((1.0 + eps)*theta*theta + cmax*theta);
  const double rtol_step = std::min(c->adaptive_rtol_max,
                                    c->adaptive_order_c*std::pow(outer_dt/c->adaptive_tref, 3));
  const double target = c->drag_atol + rtol_step*std::max(scale, c->adaptive_state_floor);
  const double bound = response*res;
  c->solver_last[2] = res; c->solver_last[3] = bound; c->solver_last[4] = target;
  c->solver_last[5] = scale; c->solver_last[6] = eps;
  if (bound <= target) {c->solver_last[1] = 1.0; return DK_OK;}
  int it = 0;
  if ((e = StrictPCG(c, a_dt, sh, t, true, &it))) return e;
  c->solver_last[0] = it;
  return DK_OK;
}

}  // namespace

namespace dk {

// The options the explicit gas update supports (dk_gas_src in dustk.h)
int HydroSupported(const dk_ctx *c, const dk_gas_src *src) {
  if (!(src->iso_cs > 0.0)) return DK_EINVAL;
  if (c->smr || c->g.ng < 2 || c->g.bc_x1 != DK_X1_PERIODIC) return DK_EUNSUPPORTED;
  return DK_OK;
}

// the flux divergence field (allocated once, outside any stream capture)
int HydroField(dk_ctx *c) {
  if (c->hdivf != nullptr) return DK_OK;
  if (hipMalloc(&c->hdivf, sizeof(double)*DK_NGAS*NCells(c)) != hipSuccess) {
    (void)hipGetLastError();
    c->hdivf = nullptr;
    return DK_ENOMEM;
  }
  return DK_OK;
}

int GasHook(dk_ctx *c, const dk_stage *st, const dk_gas_src *src, const double *dmom,
            double atw) {
  const int s = st->stage - 1;
  const double gam0 = st->gam0[s], gam1 = st->gam1[s], bdt = st->beta[s]*st->dt;
  const bool hydro = (src != nullptr && src->hydro);
  int e;
  if (hydro && (e = HydroSupported(c, src))) return e;
  if (bdt == 0.0 && gam0 == 1.0 && gam1 == 0.0) return DK_OK;   // identity stage
  dk_gas_src z;
  std::memset(&z, 0, sizeof(z));
  if (hydro) {
    // the u0 ghosts of the stage start (the previous stage's Hydro::SendU/RecvU and
    // ApplyPhysicalBCs; x1 is periodic, so the time is not read), then the flux divergence
    if ((e = HydroField(c))) return e;
    if ((e = FillGhosts(c, c->u0, DK_NGAS, true, st->time))) return e;
    {const int _pt = dk::ProfBegin(c, "hydro_divf");
    k_hydro_divf<<<CGRID(ActiveCells(c))>>>(c->g, c->d_blocks, c->u0, c->hdivf, src->iso_cs,
                                            src->dfloor);
     dk::ProfEnd(c, _pt);}
    if ((e = Launched(c))) return e;
  }
  {const int _pt = dk::ProfBegin(c, "gas_hook");
  k_gas_hook<<<CGRID(ActiveCells(c))>>>(c->g, c->d_blocks, c->u0, c->u1, gam0, gam1, bdt,
                                        src ? 1 : 0, src ? *src : z,
                                        hydro ? c->hdivf : nullptr, dmom, atw);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

template <typename T>
int StiffnessRecord(dk_ctx *c, const T *q, int nv, int comp, double dt) {
  const int t = ProfBegin(c, "stiffness");
  k_stiff_partial<T><<<kDiagGrid, kBlock, 0, c->stream>>>(c->g, q, nv, comp, c->u0,
                                                          c->d_partial);
  k_stiff_record<<<1, 64, 0, c->stream>>>(c->d_partial, kDiagGrid, dt, c->d_scalar + 9);
  ProfEnd(c, t);
  return Check(hipGetLastError());
}
template int StiffnessRecord<float>(dk_ctx *, const float *, int, int, double);
template int StiffnessRecord<double>(dk_ctx *, const double *, int, int, double);

}  // namespace dk

//========================================================================================
// C-ABI

extern "C" {

const char *dk_strerror(int code) {
  switch (code) {
    case DK_OK: return "ok";
    case DK_EINVAL: return "invalid argument";
    case DK_EHIP: return "HIP runtime error";
    case DK_ENOMEM: return "device allocation failed";
    case DK_EHALO: return "particle outside the deposit halo of its block";
    case DK_EROUTE: return "particle moved more than one block";
    case DK_ESTOP: return "non-positive stopping time";
    case DK_EUNSUPPORTED: return "option not provided";
    case DK_ECAPACITY: return "size beyond a fixed maximum";
    case DK_ESTATE: return "call order violated";
    case DK_ENOTFINITE: return "non-finite value";
    case DK_ESOLVER: return "drag solver failed";
    default: return "unknown error";
  }
}

namespace {
int ScratchInit(int device);
}  // namespace

int dk_create(const dk_mesh *mesh, const dk_block *blocks, const dk_physics *phys,
              int32_t form, void *stream, dk_ctx **out) {
  if (mesh == nullptr || blocks == nullptr || phys == nullptr || out == nullptr)
    return DK_EINVAL;
  if (mesh->nmb < 1 || mesh->nx1 < 2 || mesh->nx2 < 2 || mesh->nx3 < 2 || mesh->ng < 1 ||
      mesh->nmbx1 < 1 || mesh->nmbx2 < 1 || mesh->nmbx3 < 1)
    return DK_EINVAL;
  if (form != DK_FORM_FP64 && form != DK_FORM_FP32) return DK_EINVAL;
  // a statically refined mesh: refinable clouds with pc2 (their traversal of the actual leaves;
  // point-particle deposits would need the cross-level fold)
  int maxlev = 0;
  for (int m=0; m<mesh->nmb; ++m) maxlev = std::max(maxlev, blocks[m].level);
  const bool smr = maxlev > 0 || mesh->nmb != mesh->nmbx1*mesh->nmbx2*mesh->nmbx3;
  if (smr && !phys->clouds) return DK_EUNSUPPORTED;
  if (maxlev > 20) return DK_EINVAL;
  // the mesh passes launch one grid row per block (at most 65535), and index a block's padded
  // cells in 32 bits
  if (mesh->nmb > 65535 ||
      (long long)(mesh->nx1 + 2*mesh->ng)*(mesh->nx2 + 2*mesh->ng)*(mesh->nx3 + 2*mesh->ng) >
          INT32_MAX)
    return DK_ECAPACITY;
  if (phys->solver < DK_LOCAL || phys->solver > DK_ADAPTIVE) return DK_EUNSUPPORTED;
  if (phys->clouds && !dk::CloudsSupported(mesh, phys))
    return DK_EUNSUPPORTED;
  dk_ctx *c = new dk_ctx();
  c->M = *mesh;
  c->P = *phys;
  c->B.assign(blocks, blocks + mesh->nmb);
  c->form = form;
  c->stream = static_cast<hipStream_t>(stream);
  (void)hipGetDevice(&c->device);
  // the device address of this GPU's sort scratch (dk_sort_particles_noalloc); a failure is
  // kept for that entry point, which returns it, and leaves every other entry point usable
  c->noalloc_err = ScratchInit(c->device);
  Geo &g = c->g;
  g.nmb = mesh->nmb; g.nx1 = mesh->nx1; g.nx2 = mesh->nx2; g.nx3 = mesh->nx3; g.ng = mesh->ng;
  g.n1 = g.nx1 + 2*g.ng; g.n2 = g.nx2 + 2*g.ng; g.n3 = g.nx3 + 2*g.ng;
  g.is = g.ng; g.ie = g.ng + g.nx1 - 1;
  g.js = g.ng; g.je = g.ng + g.nx2 - 1;
  g.ks = g.ng; g.ke = g.ng + g.nx3 - 1;
  g.ncb = (long long)g.n1*g.n2*g.n3;
  g.nmbx1 = mesh->nmbx1; g.nmbx2 = mesh->nmbx2; g.nmbx3 = mesh->nmbx3;
  g.bc_x1 = mesh->bc_x1; g.bc_x3 = mesh->bc_x3; g.x3_dep_drop = mesh->x3_dep_drop;
  g.x1min = mesh->x1min; g.x1max = mesh->x1max;
  g.x2min = mesh->x2min; g.x2max = mesh->x2max;
  g.x3min = mesh->x3min; g.x3max = mesh->x3max;
  g.deposit = phys->deposit;
  std::vector<DevBlock> hb(mesh->nmb);
  for (int m=0; m<mesh->nmb; ++m) {
    const dk_block &b = blocks[m];
    DevBlock &d = hb[m];
    d.x1min = b.x1min; d.x1max = b.x1max; d.x2min = b.x2min; d.x2max = b.x2max;
    d.x3min = b.x3min; d.x3max = b.x3max;
    d.dx1 = (b.x1max - b.x1min)/static_cast<double>(g.nx1);   // meshblock.cpp:113-118
    d.dx2 = (b.x2max - b.x2min)/static_cast<double>(g.nx2);
    d.dx3 = (b.x3max - b.x3min)/static_cast<double>(g.nx3);
    c->dx1.push_back(d.dx1);
    c->dx2.push_back(d.dx2);
    c->dx3.push_back(d.dx3);
    d.sx1 = static_cast<double>(g.nx1)/(b.x1max - b.x1min);
    d.sx2 = static_cast<double>(g.nx2)/(b.x2max - b.x2min);
    d.sx3 = static_cast<double>(g.nx3)/(b.x3max - b.x3min);
    d.ivol = 1.0/(d.dx1*d.dx2*d.dx3);
    d.lx1 = b.lx1; d.lx2 = b.lx2; d.lx3 = b.lx3;
    d.level = b.level;
    d.xmask = 0;
    {
      const double lx = b.x1max - b.x1min, ly = b.x2max - b.x2min, lz = b.x3max - b.x3min;
      d.face = ((b.x1min < mesh->x1min + 0.5*lx) ? 1 : 0) | ((b.x1max > mesh->x1max - 0.5*lx) ? 2 : 0) |
               ((b.x2min < mesh->x2min + 0.5*ly) ? 4 : 0) | ((b.x2max > mesh->x2max - 0.5*ly) ? 8 : 0) |
               ((b.x3min < mesh->x3min + 0.5*lz) ? 16 : 0) | ((b.x3max > mesh->x3max - 0.5*lz) ? 32 : 0);
    }
    for (int q=0; q<27; ++q) d.nghbr[q] = b.nghbr[q];
  }
  // the same-level fold slots in NeighborIndex order (src/mesh/nghbr_index.hpp:19-54)
  std::vector<std::pair<int,int>> order;
  for (int o3=-1; o3<=1; ++o3)
    for (int o2=-1; o2<=1; ++o2)
      for (int o1=-1; o1<=1; ++o1) {
        if (o1 == 0 && o2 == 0 && o3 == 0) continue;
        int slot;
        if (o3 == 0) slot = (o1*o2 == 0) ? (std::abs(o1)*2*(o1 + 1) + std::abs(o2)*2*(o2 + 5))
                                         : (16 + (o1 + 1) + 2*(o2 + 1));
        else slot = (o1*o2 == 0) ? (24 + std::abs(o1)*(o1 + 9) + std::abs(o2)*(o2 + 17) +
                                    2*(o3 + 1))
                                 : (48 + (o1 + 1)/2 + (o2 + 1) + 2*(o3 + 1));
        order.push_back({slot, Dir(o1, o2, o3)});
      }
  std::sort(order.begin(), order.end());
  int sd[26];
  for (int q=0; q<26; ++q) sd[q] = order[q].second;
  const long long nc = dk::NCells(c);
  int e = DK_OK;
  auto alloc = [&](void **p, size_t bytes) {
    if (e) return;
    if (hipMalloc(p, bytes) != hipSuccess) {*p = nullptr; e = DK_ENOMEM;}
  };
  alloc((void**)&c->d_blocks, sizeof(DevBlock)*mesh->nmb);
  alloc((void**)&c->d_slot_dir, sizeof(int)*26);
  alloc((void**)&c->d_err, sizeof(dk::ErrWord));
  alloc((void**)&c->d_partial, sizeof(double)*9*4096);
  alloc((void**)&c->d_scalar, sizeof(double)*16);
  if (form == DK_FORM_FP64) {
    alloc((void**)&c->qdep, sizeof(double)*5*nc);
    alloc((void**)&c->ustar, sizeof(double)*3*nc);
    alloc((void**)&c->dmom, sizeof(double)*3*nc);
    alloc((void**)&c->ap, sizeof(double)*3*nc);
    alloc((void**)&c->resid, sizeof(double)*4*nc);
    if (phys->clouds) {
      alloc((void**)&c->gpack, sizeof(double)*12*nc);
      alloc((void**)&c->cvel, sizeof(double)*3*nc);
    }
    if (phys->solver == DK_PCG || phys->solver == DK_ADAPTIVE) {
      alloc((void**)&c->solver_r, sizeof(double)*3*nc);
      alloc((void**)&c->solver_p, sizeof(double)*3*nc);
    }
  }
  if (!e) e = dk::Check(hipMemcpy(c->d_blocks, hb.data(), sizeof(DevBlock)*mesh->nmb,
                                  hipMemcpyHostToDevice));
  if (!e) e = dk::Check(hipMemcpy(c->d_slot_dir, sd, sizeof(sd), hipMemcpyHostToDevice));
  if (!e) e = dk::Check(hipMemset(c->d_err, 0, sizeof(dk::ErrWord)));
  if (!e) e = dk::Check(hipMemset(c->d_scalar, 0, sizeof(double)*16));
  if (!e) {   // d_scalar[14..15]: the particle history's interval records, none yet
    const double none[2] = {-1.0, -1.0};
    e = dk::Check(hipMemcpy(c->d_scalar + 14, none, sizeof(none), hipMemcpyHostToDevice));
  }
  if (!e && form == DK_FORM_FP64) {
    for (double *p : {c->qdep, c->ustar, c->dmom, c->ap, c->resid}) {
      const long long nv = (p == c->qdep) ? 5 : ((p == c->resid) ? 4 : 3);
      if (!e) e = dk::Check(hipMemset(p, 0, sizeof(double)*nv*nc));
    }
    if (!e && phys->clouds) e = dk::Check(hipMemset(c->gpack, 0, sizeof(double)*12*nc));
    for (double *p : {c->solver_r, c->solver_p})
      if (!e && p) e = dk::Check(hipMemset(p, 0, sizeof(double)*3*nc));
  }
  if (!e && smr) {
    // the finest block grid: every slot covered by exactly one leaf
    const int F[3] = {mesh->nmbx1 << maxlev, mesh->nmbx2 << maxlev, mesh->nmbx3 << maxlev};
    std::vector<int> leaf((size_t)F[0]*F[1]*F[2], -1);
    for (int m=0; m<mesh->nmb && !e; ++m) {
      const dk_block &b = blocks[m];
      if (b.level < 0) {e = DK_EINVAL; break;}
      const int r = 1 << (maxlev - b.level);
      const int l[3] = {b.lx1, b.lx2, b.lx3};
      for (int d=0; d<3; ++d)
        if (l[d] < 0 || (long long)(l[d] + 1)*r > F[d]) e = DK_EINVAL;
      for (int k=b.lx3*r; k<(b.lx3 + 1)*r && !e; ++k)
        for (int j=b.lx2*r; j<(b.lx2 + 1)*r && !e; ++j)
          for (int i=b.lx1*r; i<(b.lx1 + 1)*r; ++i) {
            int &slot = leaf[((size_t)k*F[1] + j)*F[0] + i];
            if (slot >= 0) {e = DK_EINVAL; break;}
            slot = m;
          }
    }
    for (size_t q=0; q<leaf.size() && !e; ++q)
      if (leaf[q] < 0) e = DK_EINVAL;
    // per block, the directions whose same-level region is not one leaf of its level (a
    // region aligned at the block's level holds one leaf of that level exactly when the leaf
    // at its first finest slot is of that level), and the same-level neighbours from the leaf
    // map (the FP32 form's fold reads them; the caller's nghbr[] is not used on a refined mesh)
    for (int m=0; m<mesh->nmb && !e; ++m) {
      DevBlock &d = hb[m];
      const int r = 1 << (maxlev - d.level);
      const int l[3] = {d.lx1, d.lx2, d.lx3};
      for (int o3=-1; o3<=1; ++o3)
        for (int o2=-1; o2<=1; ++o2)
          for (int o1=-1; o1<=1; ++o1) {
            if (o1 == 0 && o2 == 0 && o3 == 0) continue;
            const int o[3] = {o1, o2, o3};
            int s[3];
            for (int a=0; a<3; ++a) s[a] = (((l[a] + o[a])*r) % F[a] + F[a]) % F[a];
            const int n = leaf[((size_t)s[2]*F[1] + s[1])*F[0] + s[0]];
            if (blocks[n].level != d.level) d.xmask |= 1 << Dir(o1, o2, o3);
            d.nghbr[Dir(o1, o2, o3)] = (blocks[n].level == d.level) ? n : -1;
          }
    }
    if (!e) e = dk::Check(hipMemcpy(c->d_blocks, hb.data(), sizeof(DevBlock)*mesh->nmb,
                                    hipMemcpyHostToDevice));
    alloc((void**)&c->d_leafmap, sizeof(int)*leaf.size());
    if (!e) e = dk::Check(hipMemcpy(c->d_leafmap, leaf.data(), sizeof(int)*leaf.size(),
                                    hipMemcpyHostToDevice));
    if (!e) {
      c->smr = 1;
      dk::SmrTree &t = c->tree;
      for (int d=0; d<3; ++d) t.F[d] = F[d];
      t.maxlev = maxlev;
      t.leaf = c->d_leafmap;
      t.lo[0] = mesh->x1min; t.lo[1] = mesh->x2min; t.lo[2] = mesh->x3min;
      t.len[0] = mesh->x1max - mesh->x1min; t.len[1] = mesh->x2max - mesh->x2min;
      t.len[2] = mesh->x3max - mesh->x3min;
    }
  }
  if (!e && form == DK_FORM_FP32) e = dk::FastCreate(c);
  if (e) {dk_destroy(c); return e;}
  *out = c;
  return DK_OK;
}

int dk_destroy(dk_ctx *c) {
  if (c == nullptr) return DK_OK;
  (void)hipStreamSynchronize(c->stream);
  if (c->fast) dk::FastDestroy(c);
  for (auto &x : c->gexec)
    if (x) {(void)hipGraphExecDestroy(x); x = nullptr;}
  dk::ProfDestroy(c);
  for (void *p : {(void*)c->d_blocks, (void*)c->d_slot_dir, (void*)c->d_err,
                  (void*)c->d_partial, (void*)c->d_scalar, (void*)c->qdep, (void*)c->ustar,
                  (void*)c->dmom, (void*)c->ap, (void*)c->resid, c->sort_tmp,
                  (void*)c->sort_keys, (void*)c->sort_idx, (void*)c->sort_rbuf,
                  (void*)c->sort_ibuf, (void*)c->gpack, (void*)c->cvel, (void*)c->solver_r,
                  (void*)c->solver_p, (void*)c->d_leafmap, (void*)c->d_block_off,
                  (void*)c->hdivf})
    if (p) (void)hipFree(p);
  delete c;
  return DK_OK;
}

int dk_bind_particles(dk_ctx *c, double *rdata, int32_t *idata, int64_t capacity,
                      int64_t npart) {
  if (c == nullptr || rdata == nullptr || idata == nullptr || capacity < npart || npart < 0 ||
      npart > INT32_MAX)
    return DK_EINVAL;
  c->rdata = rdata; c->idata = idata; c->cap = capacity; c->np = npart;
  if (c->form == DK_FORM_FP32) return dk::FastBind(c);
  return DK_OK;
}

int dk_bind_gas(dk_ctx *c, double *u0, double *u1) {
  if (c == nullptr || u0 == nullptr || u1 == nullptr) return DK_EINVAL;
  c->u0 = u0; c->u1 = u1;
  return DK_OK;
}

int dk_export_particles(dk_ctx *c) {
  if (c == nullptr) return DK_EINVAL;
  if (c->form == DK_FORM_FP32) return dk::FastExport(c);
  return DK_OK;
}

int64_t dk_npart(const dk_ctx *c) { return c ? c->np : 0; }

// src/dust/dust_push.cpp:35-80
int dk_first_two_imp_rk(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c);
  if (e) return e;
  if (c->P.coupling == DK_PC2 && st->stage == 2) {
    {const int _pt = dk::ProfBegin(c, "gas_midpoint");
    k_gas_midpoint<<<CGRID(dk::NCells(c))>>>(c->g, c->u0, c->u1, c->ustar);
     dk::ProfEnd(c, _pt);}
    return Launched(c);
  }
  if (st->stage != 1) return DK_OK;
  e = dk::Check(hipMemcpyAsync(c->u1, c->u0, sizeof(double)*DK_NGAS*dk::NCells(c),
                               hipMemcpyDeviceToDevice, c->stream));
  if (e) return e;
  {const int _pt = dk::ProfBegin(c, "reg_copy");
  k_reg_copy<<<PGRID>>>(c->rdata, (long long)c->cap, (long long)c->np);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

// src/dust/dust_push.cpp:91-300
int dk_explicit_push(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c);
  if (e || !ActiveStage(c, st)) return e;
  const int sbox = c->P.shearing_box, strat = c->P.stratified;
  const double qs = c->P.qshear, om = c->P.omega0;
  if (c->P.coupling == DK_PC2) {
    if (st->stage != 2) return DK_OK;
    if (c->P.back_reaction && (e = dk::Zero(c, c->dmom, 3*dk::NCells(c)))) return e;
    if (c->P.clouds) {   // dust_push.cpp:124-129
      if ((e = dk::CloudAdvance(c, st->dt))) return e;
      return Launched(c);
    }
    {const int _pt = dk::ProfBegin(c, "pc2_advance");
    k_pc2_advance<<<PGRID>>>(c->g, c->d_blocks, PARGS, c->ustar, c->dmom, st->dt,
                             c->P.back_reaction, sbox, strat, qs, om, c->d_err);
     dk::ProfEnd(c, _pt);}
    return Launched(c);
  }
  const int s = st->stage - 1;
  const double atw_dt = (st->stage == 2) ? st->a_twid22*st->dt : 0.0;
  {const int _pt = dk::ProfBegin(c, "push_imex");
  k_push_imex<<<PGRID>>>(c->rdata, (long long)c->cap, (long long)c->np, st->gam0[s],
                         st->gam1[s], st->beta[s]*st->dt, atw_dt, sbox, strat, qs, om);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

// src/dust/dust_tasks.cpp:149-162, dust.cpp:1678-1702, bvals_part.cpp:115-471
int dk_update_particle_gids(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c);
  if (e || !MigrationActive(c, st)) return e;
  const bool shear = (c->M.bc_x1 == DK_X1_SHEAR);
  double ysh = 0.0;
  if (shear) {
    const double tau = ShearRouteTime(c, st);
    const double dsh = c->P.qshear*c->P.omega0*(c->M.x1max - c->M.x1min)*(tau - st->time);
    if (dsh != 0.0)
      {const int _pt = dk::ProfBegin(c, "shift_shear_crossings");
      k_shift_shear_crossings<<<PGRID>>>(c->d_blocks, PARGS, dsh, c->M.x1min, c->M.x1max);
       dk::ProfEnd(c, _pt);}
    const double lx = c->M.x1max - c->M.x1min, ly = c->M.x2max - c->M.x2min;
    ysh = std::fmod(c->P.qshear*c->P.omega0*lx*st->time, ly);
  }
  {const int _pt = dk::ProfBegin(c, "part_update");
  k_part_update<<<PGRID>>>(c->g, c->d_blocks, PARGS, shear ? 1 : 0, ysh, c->d_scalar + 12,
                           c->tree, c->d_err);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

// src/dust/dust_pm.cpp:104-277
int dk_deposit_drag(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c);
  if (e || !ActiveStage(c, st) || !c->P.back_reaction) return e;
  if (c->P.coupling == DK_PC2) {
    if (st->stage != 1) return DK_OK;
    if ((e = dk::Zero(c, c->qdep, 5*dk::NCells(c)))) return e;
    if (c->P.clouds) {   // dust_pm.cpp:130-137
      if ((e = dk::CloudPredictor(c, st->dt, dk::StiffnessActive(c->P)))) return e;
      return Launched(c);
    }
    {const int _pt = dk::ProfBegin(c, "pc2_predictor");
    k_pc2_predictor<<<PGRID>>>(c->g, c->d_blocks, PARGS, c->u1, c->qdep, st->dt,
                               dk::StiffnessActive(c->P) ? 1 : 0, c->d_err);
     dk::ProfEnd(c, _pt);}
    return Launched(c);
  }
  if ((e = dk::Zero(c, c->qdep, 5*dk::NCells(c)))) return e;
  {const int _pt = dk::ProfBegin(c, "scatter_imex");
  k_scatter_imex<<<PGRID>>>(c->g, c->d_blocks, PARGS, c->qdep, DragStep(c, st),
                            ImplicitShearShift(c, st), c->d_err);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

// src/dust/dust_tasks.cpp:249-263
int dk_fold_deposit_qp(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c, false);
  const bool use = (c && c->P.coupling == DK_IMEX) || st->stage == 1;
  // (a refined mesh: the clouds deposited into the leaves themselves, no ghost holds any)
  if (e || !ActiveStage(c, st) || !c->P.back_reaction || !use || c->smr) return e;
  return Fold(c, c->qdep, 5, st->time);
}

// src/dust/dust_pm.cpp:372-442, src/dust/dust_solver.cpp:630-724
int dk_gas_implicit_solve(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c);
  if (e || !ActiveStage(c, st)) return e;
  if (c->P.coupling == DK_PC2) {
    if (st->stage == 1 && c->P.back_reaction) {
      if (dk::StiffnessActive(c->P) && (e = dk::StiffnessRecord<double>(c, c->qdep, 5, 4, st->dt))) return e;
      {const int _pt = dk::ProfBegin(c, "pc2_apply_predictor");
      k_pc2_apply_predictor<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, c->P.pc2_damping,
                                                       st->dt);
       dk::ProfEnd(c, _pt);}
      return Launched(c);
    }
    return DK_OK;
  }
  const int br = c->P.back_reaction;
  {const int _pt = dk::ProfBegin(c, "local_guess");
  k_local_guess<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->qdep, c->ustar, br);
   dk::ProfEnd(c, _pt);}
  if ((e = Launched(c))) return e;
  if (!br || c->P.solver == DK_LOCAL) return DK_OK;
  const double a_dt = DragStep(c, st), sh = ImplicitShearShift(c, st);
  for (int q=0; q<7; ++q) c->solver_last[q] = (q < 2) ? 0.0 : -1.0;
  if (c->P.solver == DK_APPLYA) return ApplyCoupledOperator(c, a_dt, sh, false, st->time);
  if (c->P.solver == DK_PCG) {
    int it = 0;
    if ((e = StrictPCG(c, a_dt, sh, st->time, false, &it))) return e;
    c->solver_last[0] = it;
    return DK_OK;
  }
  const int nsweeps = (c->P.solver == DK_DC2) ? 2 : 1;
  for (int s=0; s<nsweeps; ++s)
    if ((e = ApplyCoupledOperator(c, a_dt, sh, true, st->time))) return e;
  if (c->P.solver == DK_ADAPTIVE) return AdaptiveSolve(c, a_dt, sh, st->dt, st->time);
  return DK_OK;
}

// src/dust/dust_tasks.cpp:265-312
int dk_fill_ustar(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c, false);
  if (e || !ActiveStage(c, st) || c->P.coupling != DK_IMEX) return e;
  return FillGhosts(c, c->ustar, 3, true, st->time);
}

// src/dust/dust_pm.cpp:450-568
int dk_gather_kick_pmbr(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c);
  if (e || !ActiveStage(c, st) || c->P.coupling == DK_PC2) return e;
  if (c->P.back_reaction && (e = dk::Zero(c, c->dmom, 3*dk::NCells(c)))) return e;
  {const int _pt = dk::ProfBegin(c, "gather_kick");
  k_gather_kick<<<PGRID>>>(c->g, c->d_blocks, PARGS, c->ustar, c->dmom, DragStep(c, st),
                           ImplicitShearShift(c, st), c->P.back_reaction, c->d_err);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

// src/dust/dust_tasks.cpp:314-328
int dk_fold_pmbr(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c, false);
  const bool use = (c && c->P.coupling == DK_IMEX) || st->stage == 2;
  if (e || !ActiveStage(c, st) || !c->P.back_reaction || !use || c->smr) return e;
  return Fold(c, c->dmom, 3, st->time);
}

// src/dust/dust_pm.cpp:576-711
int dk_apply_pmbr(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c);
  if (e || !ActiveStage(c, st) || !c->P.back_reaction) return e;
  if (c->P.coupling == DK_PC2) {
    if (st->stage != 2) return DK_OK;
    {const int _pt = dk::ProfBegin(c, "cell_update");
    k_cell_update<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->dmom, c->qdep, 1,
                                             st->gam1[st->nexp_stages-1]);
     dk::ProfEnd(c, _pt);}
    return Launched(c);
  }
  const double inv_adt = 1.0/(st->a_impl*st->dt);
  if (c->P.consistent_gas) {
    const double a_dt = DragStep(c, st);
    if ((e = dk::Zero(c, c->resid, 4*dk::NCells(c)))) return e;
    {const int _pt = dk::ProfBegin(c, "center_mass");
    k_center_mass<<<PGRID>>>(c->g, c->d_blocks, c->rdata, c->idata, (long long)c->cap,
                             (long long)c->np, c->resid, a_dt);
     dk::ProfEnd(c, _pt);}
    {const int _pt = dk::ProfBegin(c, "consistent_gas");
    k_consistent_gas<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->ustar, c->qdep, c->dmom,
                                                 c->resid, inv_adt);
     dk::ProfEnd(c, _pt);}
    {const int _pt = dk::ProfBegin(c, "consistent_dust");
    k_consistent_dust<<<PGRID>>>(c->g, c->d_blocks, PARGS, c->resid, a_dt);
     dk::ProfEnd(c, _pt);}
    return Launched(c);
  }
  {const int _pt = dk::ProfBegin(c, "cell_update");
  k_cell_update<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->dmom, c->qdep, 0, inv_adt);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

// src/dust/dust_pm.cpp:721-742
int dk_add_drag_history_gas(dk_ctx *c, const dk_stage *st) {
  int e = Ready(c, false);
  if (e || c->P.coupling != DK_IMEX || st->stage != 2 || !c->P.back_reaction) return e;
  {const int _pt = dk::ProfBegin(c, "cell_update");
  k_cell_update<<<CGRID(ActiveCells(c))>>>(c->g, c->u0, c->dmom, c->qdep, 2,
                                           st->a_twid22*st->dt);
   dk::ProfEnd(c, _pt);}
  return Launched(c);
}

int dk_gas_hook(dk_ctx *c, const dk_stage *st, const dk_gas_src *src) {
  if (c == nullptr || c->u0 == nullptr) return DK_ESTATE;
  return dk::GasHook(c, st, src);
}

int dk_fill_gas_ghosts(dk_ctx *c, const dk_stage *st) {
  if (c == nullptr || c->u0 == nullptr) return DK_ESTATE;
  return FillGhosts(c, c->u0, DK_NGAS, true, st->time + st->dt);
}

namespace {

int StageBody(dk_ctx *c, const dk_stage *st, const dk_gas_src *src) {
  if (c->form == DK_FORM_FP32) return dk::FastStage(c, st, src);
  int e;
  if ((e = dk_first_two_imp_rk(c, st))) return e;
  if ((e = dk_gas_hook(c, st, src))) return e;
  if ((e = dk_add_drag_history_gas(c, st))) return e;
  if ((e = dk_explicit_push(c, st))) return e;
  if ((e = dk_update_particle_gids(c, st))) return e;
  if ((e = dk_deposit_drag(c, st))) return e;
  if ((e = dk_fold_deposit_qp(c, st))) return e;
  if ((e = dk_gas_implicit_solve(c, st))) return e;
  if ((e = dk_fill_ustar(c, st))) return e;
  if ((e = dk_gather_kick_pmbr(c, st))) return e;
  if ((e = dk_fold_pmbr(c, st))) return e;
  if ((e = dk_apply_pmbr(c, st))) return e;
  if (ActiveStage(c, st) && (e = dk_fill_gas_ghosts(c, st))) return e;
  return DK_OK;
}

// The stage captured from the context's stream into a graph, which then updates the stage's
// executable graph (same topology every time: only kernel arguments change) and is launched as
// one unit.  Not used on the null stream, while the caller captures the stream itself, or while
// per-kernel timing is on (its events would enter the graph).
int StageGraph(dk_ctx *c, const dk_stage *st, const dk_gas_src *src) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (c->stream == nullptr || dk::ProfOn(c) ||
      hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return StageBody(c, st, src);
  const int slot = std::min(std::max(st->stage, 1), 3) - 1 + (src ? 3 : 0);
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return DK_EHIP;
  const int e = StageBody(c, st, src);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(c->stream, &g);
  if (e || ce != hipSuccess || g == nullptr) {
    if (g) (void)hipGraphDestroy(g);
    return e ? e : DK_EHIP;
  }
  hipGraphExec_t &x = c->gexec[slot];
  if (x != nullptr) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
    if (hipGraphExecUpdate(x, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
  }
  int r = DK_OK;
  if (x == nullptr && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
    x = nullptr;
    r = DK_EHIP;
  }
  (void)hipGraphDestroy(g);
  if (r) return r;
  r = dk::Check(hipGraphLaunch(x, c->stream));
  if (!r) ++c->graph_launches;
  return r;
}

}  // namespace

int dk_stage_run(dk_ctx *c, const dk_stage *st, const dk_gas_src *src) {
  if (c == nullptr || st == nullptr) return DK_EINVAL;
  // a stage with isothermal fluxes is checked before any of its work and allocates its field
  // before a graph capture can begin
  int e = DK_OK;
  if (src != nullptr && src->hydro && c->u0 != nullptr && !(e = dk::HydroSupported(c, src)))
    e = dk::HydroField(c);
  if (!e && c->fast) e = dk::FastTuneBegin(c, st);
  if (!e) e = c->graphs ? StageGraph(c, st, src) : StageBody(c, st, src);
  if (!e && c->fast) e = dk::FastTuneEnd(c, st);
  return e;
}

// the strict solver's parameters (src/dust/dust.cpp:467-474)
int dk_set_solver(dk_ctx *c, double rtol, double atol, int32_t iter_max, double order_c,
                  double rtol_max, double tref, double state_floor) {
  if (c == nullptr || !(rtol > 0.0) || atol < 0.0 || iter_max < 1 || !(order_c > 0.0) ||
      !(rtol_max > 0.0) || !(tref > 0.0) || !(state_floor > 0.0))
    return DK_EINVAL;
  c->drag_rtol = rtol; c->drag_atol = atol; c->drag_iter_max = iter_max;
  c->adaptive_order_c = order_c; c->adaptive_rtol_max = rtol_max; c->adaptive_tref = tref;
  c->adaptive_state_floor = state_floor;
  return DK_OK;
}

int dk_solver_stats(dk_ctx *c, double out[7]) {
  if (c == nullptr || out == nullptr) return DK_EINVAL;
  for (int q=0; q<7; ++q) out[q] = c->solver_last[q];
  return DK_OK;
}

int dk_set_graphs(dk_ctx *c, int32_t enable) {
  if (c == nullptr) return DK_EINVAL;
  c->graphs = enable ? 1 : 0;
  return DK_OK;
}

int64_t dk_graph_launches(const dk_ctx *c) { return c ? c->graph_launches : 0; }

// src/dust/dust_cloud_device.cpp:1881-1980: stable order by (block, cell).  The FP32 form
// keeps its store sorted by (block, tile, cell) every stage, so it has nothing to do here.
namespace {
int SortScratch(dk_ctx *c, long long n);
}  // namespace

namespace {

// The buffers of one cell sort of n particles: keys and ranks twice (radix sort input and
// output), every column once (the permuted copy), the radix sort's temporary storage.
struct SortBufs {
  unsigned long long *keys;
  int *idx;
  double *rbuf;
  int32_t *ibuf;
  void *tmp;
  size_t tmp_bytes;
};

int SortRadixBits(const dk_ctx *c) {
  return 64 - __builtin_clzll((unsigned long long)c->g.nmb*c->g.nx1*c->g.nx2*c->g.nx3);
}

// the radix sort's temporary bytes for n pairs (a size query: no work, no allocation)
int SortTmpBytes(const dk_ctx *c, const SortBufs &b, long long n, size_t &bytes) {
  bytes = 0;
  return dk::Check(hipcub::DeviceRadixSort::SortPairs(
      nullptr, bytes, b.keys, b.keys + n, b.idx, b.idx + n, (int)n, 0, SortRadixBits(c),
      c->stream));
}

// keys of (block, cell), their stable radix sort, the columns gathered into the sorted order
// and copied back (timed: per-kernel event timing, which the allocation-free entry leaves out)
int SortInto(dk_ctx *c, const SortBufs &b, long long n, bool timed) {
  int e;
  {const int _pt = timed ? dk::ProfBegin(c, "sort_keys") : -1;
  k_sort_keys<<<PGRID>>>(c->g, c->d_blocks, PARGS, b.keys, b.idx);
   if (timed) dk::ProfEnd(c, _pt);}
  size_t bytes = b.tmp_bytes;
  if ((e = dk::Check(hipcub::DeviceRadixSort::SortPairs(
           b.tmp, bytes, b.keys, b.keys + n, b.idx, b.idx + n, (int)n, 0, SortRadixBits(c),
           c->stream))))
    return e;
  {const int _pt = timed ? dk::ProfBegin(c, "permute") : -1;
  k_permute<<<PGRID>>>(c->rdata, c->idata, (long long)c->cap, n, b.idx + n, b.rbuf, b.ibuf);
   if (timed) dk::ProfEnd(c, _pt);}
  {const int _pt = timed ? dk::ProfBegin(c, "unpermute_copy") : -1;
  k_unpermute_copy<<<PGRID>>>(c->rdata, c->idata, (long long)c->cap, n, b.rbuf, b.ibuf);
   if (timed) dk::ProfEnd(c, _pt);}
  return Launched(c);
}

}  // namespace

int dk_sort_particles(dk_ctx *c) {
  if (c != nullptr && c->form == DK_FORM_FP32) return c->fast ? DK_OK : DK_ESTATE;
  int e = Ready(c);
  if (e) return e;
  const long long n = c->np;
  if (n == 0) return DK_OK;
  if ((e = SortScratch(c, n))) return e;
  SortBufs b{c->sort_keys, c->sort_idx, c->sort_rbuf, c->sort_ibuf, nullptr, 0};
  size_t bytes = 0;
  if ((e = SortTmpBytes(c, b, n, bytes))) return e;
  if (bytes > c->sort_tmp_bytes) {
    if (c->sort_tmp) (void)hipFree(c->sort_tmp);
    if (hipMalloc(&c->sort_tmp, bytes) != hipSuccess) {c->sort_tmp = nullptr; return DK_ENOMEM;}
    c->sort_tmp_bytes = bytes;
  }
  b.tmp = c->sort_tmp;
  b.tmp_bytes = bytes;
  return SortInto(c, b, n, true);
}

//----------------------------------------------------------------------------------------
// The allocation-free cell sort (dk_sort_particles_noalloc).  Its buffers live in one
// file-scope device array per GPU, held until the process ends: at most DK_SORT_NOALLOC_MAX
// particles (keys and ranks twice, the 17 + 5 columns once: 180 B per particle) and the radix
// sort's temporary storage in the rest.  In host code the array's name is not a device address:
// every host use goes through the pointer hipGetSymbolAddress returned for the device, taken
// once per device when a context is created on it.

namespace {
constexpr long long kSortNoAllocMax = DK_SORT_NOALLOC_MAX;
constexpr size_t kSortNoAllocTmp = size_t(8) << 20;
constexpr size_t kSortNoAllocBytes =
    (size_t)kSortNoAllocMax*(2*sizeof(unsigned long long) + 2*sizeof(int) +
                             DK_NRDATA*sizeof(double) + DK_NIDATA*sizeof(int32_t)) +
    4*256 + kSortNoAllocTmp;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kSortNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// with `device` current: the device address of its copy of scratch (once per device)
int ScratchInit(int device) {
  if (device < 0 || device >= kMaxDevices) return DK_ECAPACITY;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] != nullptr) return DK_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
    return DK_EHIP;
  g_scratch_ptr[device] = p;
  return DK_OK;
}

void *ScratchPtr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}
}  // namespace

int dk_sort_particles_noalloc(dk_ctx *c) {
  if (c != nullptr && c->form == DK_FORM_FP32) return c->fast ? DK_OK : DK_ESTATE;
  int e = Ready(c);
  if (e) return e;
  const long long n = c->np;
  if (n > kSortNoAllocMax) return DK_ECAPACITY;
  if (n == 0) return DK_OK;
  if (c->noalloc_err) return c->noalloc_err;   // (no address obtained at dk_create)
  char *base = static_cast<char*>(ScratchPtr(c->device));
  if (base == nullptr) return DK_EHIP;
  char *p = base;
  auto take = [&p](size_t bytes) {char *q = p; p += (bytes + 255)/256*256; return q;};
  SortBufs b;
  b.keys = reinterpret_cast<unsigned long long*>(take(2*sizeof(unsigned long long)*n));
  b.idx = reinterpret_cast<int*>(take(2*sizeof(int)*n));
  b.rbuf = reinterpret_cast<double*>(take(DK_NRDATA*sizeof(double)*n));
  b.ibuf = reinterpret_cast<int32_t*>(take(DK_NIDATA*sizeof(int32_t)*n));
  b.tmp = p;
  b.tmp_bytes = kSortNoAllocBytes - static_cast<size_t>(p - base);
  size_t bytes = 0;
  if ((e = SortTmpBytes(c, b, n, bytes))) return e;
  if (bytes > b.tmp_bytes) return DK_ECAPACITY;
  return SortInto(c, b, n, false);
}

// RemoveOutflowParticles (src/dust/dust_tasks.cpp:365-420): the keep flags (mass > 0), their
// exclusive scan, the kept particles copied in order into the scratch columns, then back
__global__ void k_keep_flags(const double *r, long long cap, long long np, int *flag) {
  GRID_STRIDE(p, np) flag[p] = (r[(long long)DK_IPM*cap + p] > 0.0) ? 1 : 0;
}
__global__ void k_compact_out(const double *r, const int32_t *iv, long long cap, long long np,
                              const int *flag, const int *pos, double *rb, int32_t *ib) {
  GRID_STRIDE(p, np) {
    if (!flag[p]) continue;
    const long long k = pos[p];
    for (int c=0; c<DK_NRDATA; ++c) rb[c*np + k] = r[c*cap + p];
    for (int c=0; c<DK_NIDATA; ++c) ib[c*np + k] = iv[c*cap + p];
  }
}
__global__ void k_compact_back(double *r, int32_t *iv, long long cap, long long np,
                               long long nkeep, const double *rb, const int32_t *ib) {
  GRID_STRIDE(k, nkeep) {
    for (int c=0; c<DK_NRDATA; ++c) r[c*cap + k] = rb[c*np + k];
    for (int c=0; c<DK_NIDATA; ++c) iv[c*cap + k] = ib[c*np + k];
  }
}

namespace {

// the sort scratch of the context for n particles (grown on demand)
int SortScratch(dk_ctx *c, long long n) {
  if (c->sort_cap >= n) return DK_OK;
  for (void *p : {(void*)c->sort_keys, (void*)c->sort_idx, (void*)c->sort_rbuf,
                  (void*)c->sort_ibuf})
    if (p) (void)hipFree(p);
  c->sort_keys = nullptr; c->sort_idx = nullptr; c->sort_rbuf = nullptr;
  c->sort_ibuf = nullptr;
  if (hipMalloc((void**)&c->sort_keys, sizeof(unsigned long long)*2*n) != hipSuccess ||
      hipMalloc((void**)&c->sort_idx, sizeof(int)*2*n) != hipSuccess ||
      hipMalloc((void**)&c->sort_rbuf, sizeof(double)*DK_NRDATA*n) != hipSuccess ||
      hipMalloc((void**)&c->sort_ibuf, sizeof(int32_t)*DK_NIDATA*n) != hipSuccess)
    return DK_ENOMEM;
  c->sort_cap = n;
  return DK_OK;
}

// the caller's particle columns without the particles of zero mass, in order; c->np updated
int CompactMassless(dk_ctx *c) {
  const long long n = c->np;
  if (n == 0) return DK_OK;
  int e;
  if ((e = SortScratch(c, n))) return e;
  int *flag = c->sort_idx, *pos = c->sort_idx + n;
  k_keep_flags<<<PGRID>>>(c->rdata, (long long)c->cap, n, flag);
  size_t bytes = 0;
  if ((e = dk::Check(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flag, pos, (int)n,
                                                      c->stream))))
    return e;
  if (bytes > c->sort_tmp_bytes) {
    if (c->sort_tmp) (void)hipFree(c->sort_tmp);
    if (hipMalloc(&c->sort_tmp, bytes) != hipSuccess) {c->sort_tmp = nullptr; return DK_ENOMEM;}
    c->sort_tmp_bytes = bytes;
  }
  if ((e = dk::Check(hipcub::DeviceScan::ExclusiveSum(c->sort_tmp, bytes, flag, pos, (int)n,
                                                      c->stream))))
    return e;
  k_compact_out<<<PGRID>>>(c->rdata, c->idata, (long long)c->cap, n, flag, pos, c->sort_rbuf,
                           c->sort_ibuf);
  int last[2] = {0, 0};
  if ((e = dk::Check(hipMemcpyAsync(&last[0], pos + n - 1, sizeof(int), hipMemcpyDeviceToHost,
                                    c->stream))) ||
      (e = dk::Check(hipMemcpyAsync(&last[1], flag + n - 1, sizeof(int), hipMemcpyDeviceToHost,
                                    c->stream))) ||
      (e = dk::Check(hipStreamSynchronize(c->stream))))
    return e;
  const long long nkeep = (long long)last[0] + last[1];
  k_compact_back<<<Grid(nkeep), kBlock, 0, c->stream>>>(c->rdata, c->idata, (long long)c->cap,
                                                        n, nkeep, c->sort_rbuf, c->sort_ibuf);
  c->np = nkeep;
  return Launched(c);
}

}  // namespace

int dk_remove_outflow_particles(dk_ctx *c, double out[4]) {
  if (c == nullptr || out == nullptr) return DK_EINVAL;
  if (c->rdata == nullptr || c->u0 == nullptr) return DK_ESTATE;
  int e;
  out[0] = out[1] = 0.0;
  if (c->M.bc_x3 == DK_X3_OUTFLOW) {
    double led[2] = {0.0, 0.0};
    if ((e = dk::Check(hipMemcpyAsync(led, c->d_scalar + 12, sizeof(led),
                                      hipMemcpyDeviceToHost, c->stream))) ||
        (e = dk::Check(hipStreamSynchronize(c->stream))))
      return e;
    if (led[1] != 0.0) {
      // the FP32 form: its store back into the caller's columns, compacted there, bound again
      if (c->form == DK_FORM_FP32 && (e = dk::FastExport(c))) return e;
      if ((e = CompactMassless(c))) return e;
      if (c->form == DK_FORM_FP32 && (e = dk::FastBind(c))) return e;
      out[0] = led[0]; out[1] = led[1];
      c->x3_out_mass += led[0];
      c->x3_out_count += led[1];
      if ((e = dk::Check(hipMemsetAsync(c->d_scalar + 12, 0, 2*sizeof(double), c->stream))))
        return e;
    }
  }
  out[2] = c->x3_out_mass;
  out[3] = c->x3_out_count;
  return dk::Check(hipStreamSynchronize(c->stream));
}

int dk_x3_removed(dk_ctx *c, double out[2]) {
  if (c == nullptr || out == nullptr) return DK_EINVAL;
  int e;
  if ((e = dk::Check(hipMemcpyAsync(out, c->d_scalar + 12, 2*sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream))))
    return e;
  return dk::Check(hipStreamSynchronize(c->stream));
}

int dk_new_time_step(dk_ctx *c, double dt_cfl, double *dt_out) {
  const int nb = 1024;
  if (c == nullptr) return DK_EINVAL;
  const double qo = c->P.shearing_box ? c->P.qshear*c->P.omega0 : 0.0;
  int e;
  if (c->form == DK_FORM_FP32) {
    if ((e = dk::FastNewDt(c, qo, c->d_partial, nb))) return e;
  } else {
    if ((e = Ready(c))) return e;
    k_newdt<<<nb, kBlock, 0, c->stream>>>(c->d_blocks, c->rdata, c->idata, (long long)c->cap,
                                          (long long)c->np, qo, c->d_partial, c->d_err);
  }
  std::vector<double> h(nb);
  if ((e = dk::Check(hipMemcpyAsync(h.data(), c->d_partial, sizeof(double)*nb,
                                    hipMemcpyDeviceToHost, c->stream))))
    return e;
  if ((e = dk::Check(hipStreamSynchronize(c->stream)))) return e;
  double m = 3.4028234663852886e38;
  for (double x : h) m = std::fmin(m, x);
  int32_t code = 0;
  int64_t part = -1;
  if ((e = dk_error(c, &code, &part))) return e;
  if (code == DK_ESTOP) return DK_ESTOP;
  double dt = dt_cfl*m;
  if (c->form == DK_FORM_FP32 && c->P.shearing_box) {
    // the FP32 form's shearing-box limit: a stage's orbital-advection frame shift
    // q Omega c dt |x1|max, with the stage abscissa c at most 1 + 1/sqrt(2) (imex2+), must stay
    // below half a cell in x2 (FastStage refuses a larger one), so no proposed dt reaches it
    const double xmax = std::max(std::fabs(c->M.x1min), std::fabs(c->M.x1max));
    const double dy = (c->M.x2max - c->M.x2min)/(c->g.nx2*c->g.nmbx2);
    const double cmax = 1.7071067811865475;
    if (qo*xmax > 0.0) dt = std::fmin(dt, (1.0 - 1.0e-9)*0.5*dy/(qo*cmax*xmax));
  }
  if (dt_out) *dt_out = dt;
  return DK_OK;
}

int dk_diag_sums(dk_ctx *c, double out[9]) {
  if (c == nullptr || out == nullptr) return DK_EINVAL;
  if (c->form == DK_FORM_FP32) return dk::FastDiag(c, out);
  if (c->u0 == nullptr || c->rdata == nullptr) return DK_ESTATE;
  k_diag_part<<<kDiagGrid, kBlock, 0, c->stream>>>(c->g, c->d_blocks, c->rdata,
                                                   (long long)c->cap, (long long)c->np, c->u0,
                                                   c->d_partial);
  k_reduce_partials<<<1, 64, 0, c->stream>>>(c->d_partial, kDiagGrid, 9, c->d_scalar);
  int e = dk::Check(hipMemcpyAsync(out, c->d_scalar, sizeof(double)*9, hipMemcpyDeviceToHost,
                                   c->stream));
  if (!e) e = dk::Check(hipStreamSynchronize(c->stream));
  return e;
}

// the UpdateStiffness records of the last pc2 cycle (dust_pm.cpp:288-342), both forms
int dk_stiffness(dk_ctx *c, double out[3]) {
  if (c == nullptr || out == nullptr) return DK_EINVAL;
  int e = dk::Check(hipMemcpyAsync(out, c->d_scalar + 9, sizeof(double)*3,
                                   hipMemcpyDeviceToHost, c->stream));
  if (!e) e = dk::Check(hipStreamSynchronize(c->stream));
  return e;
}

// RefreshStoppingTimeMaximum (src/dust/dust_newdt.cpp:34-98), both forms
int dk_stopping_times(dk_ctx *c, double out[2]) {
  if (c == nullptr || out == nullptr) return DK_EINVAL;
  int e = DK_OK;
  if (c->form == DK_FORM_FP32) {
    if ((e = dk::FastStoppingTimes(c, c->d_partial, kDiagGrid))) return e;
  } else {
    if ((e = Ready(c))) return e;
    k_stop_partial<<<kDiagGrid, kBlock, 0, c->stream>>>(c->rdata, c->idata, (long long)c->cap,
                                                        (long long)c->np, c->d_partial,
                                                        c->d_err);
    if ((e = dk::Check(hipGetLastError()))) return e;
  }
  std::vector<double> h(2*kDiagGrid);
  if ((e = dk::Check(hipMemcpyAsync(h.data(), c->d_partial, sizeof(double)*h.size(),
                                    hipMemcpyDeviceToHost, c->stream)))) return e;
  if ((e = dk::Check(hipStreamSynchronize(c->stream)))) return e;
  int32_t code = 0;
  int64_t part = -1;
  if ((e = dk_error(c, &code, &part))) return e;
  if (code == DK_ESTOP) return DK_ESTOP;
  double mx = 0.0, mn = 3.4028234663852886e38;
  for (int b=0; b<kDiagGrid; ++b) {mx = std::max(mx, h[2*b]); mn = std::min(mn, h[2*b + 1]);}
  out[0] = mx;
  out[1] = mn;
  return DK_OK;
}

int dk_error(dk_ctx *c, int32_t *code, int64_t *particle) {
  if (c == nullptr) return DK_EINVAL;
  dk::ErrWord h;
  int e = dk::Check(hipMemcpyAsync(&h, c->d_err, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  if (!e) e = dk::Check(hipStreamSynchronize(c->stream));
  if (e) return e;
  if (code) *code = h.code;
  if (particle) *particle = h.particle;
  if (h.code != 0) {
    e = dk::Check(hipMemsetAsync(c->d_err, 0, sizeof(dk::ErrWord), c->stream));
    if (!e) e = dk::Check(hipStreamSynchronize(c->stream));
  }
  return e;
}

void *dk_field(dk_ctx *c, int32_t which, int32_t *nvar) {
  if (c == nullptr) return nullptr;
  int nv = 0;
  void *p = nullptr;
  switch (which) {
    case 0: nv = 5; p = c->qdep; break;
    case 1: nv = 3; p = c->ustar; break;
    case 2: nv = 3; p = c->dmom; break;
    case 3: nv = 3; p = c->ap; break;
    case 4: nv = 4; p = c->resid; break;
    default: break;
  }
  if (nvar) *nvar = nv;
  return p;
}

}  // extern "C"
