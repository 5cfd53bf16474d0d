// dk_setup.hip -- the synthetic initial condition of include/dustk_setup.h on the device.
//
// Every value is a function of (seed, particle index) or (seed, global cell index): the
// per-particle hash draws of src/pgen/dust_amr/dust_setup.hpp:35-206 (HashUniform01,
// RootPlacement::Base3, NSHDustDrift) and the default masses of src/dust/dust.cpp:1716-1778.
// Large workloads are generated in place in HBM instead of being copied from the host.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "dk_internal.hpp"

using dk::DevBlock;
using dk::Geo;
using dk::Idx;
using dk::kBlock;

namespace {

// src/pgen/dust_amr/dust_setup.hpp:35-50
__host__ __device__ inline double HashUniform01(unsigned long long key) {
  key += 0x9e3779b97f4a7c15ULL;
  key = (key ^ (key >> 30)) * 0xbf58476d1ce4e5b9ULL;
  key = (key ^ (key >> 27)) * 0x94d049bb133111ebULL;
  key ^= key >> 31;
  return static_cast<double>(key >> 11) * (1.0/9007199254740992.0);
}

__device__ inline double Gauss(unsigned long long k) {
  const double u1 = fmax(HashUniform01(k), 1.0e-300);
  const double u2 = HashUniform01(k + 0x1d8e4e27c47d124fULL);
  return sqrt(-2.0*log(u1))*cos(2.0*M_PI*u2);
}

// the host's Gaussian draw of a block's count (the oracle draws it with the same host library)
double HostGauss(unsigned long long k) {
  const double u1 = std::fmax(HashUniform01(k), 1.0e-300);
  const double u2 = HashUniform01(k + 0x1d8e4e27c47d124fULL);
  return std::sqrt(-2.0*std::log(u1))*std::cos(2.0*M_PI*u2);
}

struct SetupArgs {
  dk_setup su;
  double wmean[DK_MAXSPEC];
  double gas_mass;
  long long npb, ntot;
  const long long *boff;   // uneven blocks: the first particle of every block (nmb + 1)
};

// the block of particle q from the blocks' offsets (the last block whose offset is <= q)
__device__ inline int BlockOf(const long long *off, int nmb, long long q) {
  int a = 0, b = nmb;   // off[a] <= q < off[b]
  while (b - a > 1) {
    const int mid = (a + b)/2;
    if (off[mid] <= q) a = mid; else b = mid;
  }
  return a;
}

__global__ void k_setup_particles(Geo g, const DevBlock *B, double *r, int32_t *iv,
                                  long long cap, SetupArgs a) {
  const dk_setup &su = a.su;
  const unsigned long long base = su.seed*0x9e3779b97f4a7c15ULL;
  const double lxm = g.x1max - g.x1min, lym = g.x2max - g.x2min, lzm = g.x3max - g.x3min;
  const unsigned long long saltv[3] = {DK_SALT_V0, DK_SALT_V1, DK_SALT_V2};
  const double ppc_s = static_cast<double>(su.ppc)/su.nspecies;
  for (long long q = blockIdx.x*(long long)blockDim.x + threadIdx.x; q < a.ntot;
       q += (long long)gridDim.x*blockDim.x) {
    const unsigned long long key = static_cast<unsigned long long>(q) + base;
    int m = a.boff ? BlockOf(a.boff, g.nmb, q) : static_cast<int>(q/a.npb);
    const int sp = static_cast<int>(q % su.nspecies);
    double x, y, z;
    if (su.nclump > 0) {
      if (HashUniform01(key + DK_SALT_CLUMP) < su.clump_frac) {
        const int ci = min(su.nclump - 1,
                           static_cast<int>(HashUniform01(key + DK_SALT_CLUMP + 1)*su.nclump));
        const unsigned long long ck = base + static_cast<unsigned long long>(ci)*0x2545f4914f6cdd1dULL;
        const double cx = g.x1min + HashUniform01(ck + 0x632be59bd9b4e019ULL)*lxm;
        const double cy = g.x2min + HashUniform01(ck + 0x8cb92baa3f3d8dd7ULL)*lym;
        const double cz = g.x3min + HashUniform01(ck + 0x58f38ded6f7c55b5ULL)*lzm;
        x = cx + su.clump_sigma*Gauss(key + 0x632be59bd9b4e019ULL);
        y = cy + su.clump_sigma*Gauss(key + 0x8cb92baa3f3d8dd7ULL);
        z = cz + su.clump_sigma*Gauss(key + 0x58f38ded6f7c55b5ULL);
        x -= lxm*floor((x - g.x1min)/lxm);
        y -= lym*floor((y - g.x2min)/lym);
        z -= lzm*floor((z - g.x3min)/lzm);
        if (x >= g.x1max) x = g.x1min;
        if (y >= g.x2max) y = g.x2min;
        if (z >= g.x3max) z = g.x3min;
      } else {
        x = g.x1min + HashUniform01(key + 0x632be59bd9b4e019ULL)*lxm;
        y = g.x2min + HashUniform01(key + 0x8cb92baa3f3d8dd7ULL)*lym;
        z = g.x3min + HashUniform01(key + 0x58f38ded6f7c55b5ULL)*lzm;
      }
      const int bx = min(g.nmbx1 - 1, static_cast<int>((x - g.x1min)/lxm*g.nmbx1));
      const int by = min(g.nmbx2 - 1, static_cast<int>((y - g.x2min)/lym*g.nmbx2));
      const int bz = min(g.nmbx3 - 1, static_cast<int>((z - g.x3min)/lzm*g.nmbx3));
      m = bx + g.nmbx1*(by + g.nmbx2*bz);
      const DevBlock &b = B[m];
      if (x < b.x1min) x = b.x1min;
      if (x >= b.x1max) x = nextafter(b.x1max, b.x1min);
      if (y < b.x2min) y = b.x2min;
      if (y >= b.x2max) y = nextafter(b.x2max, b.x2min);
      if (z < b.x3min) z = b.x3min;
      if (z >= b.x3max) z = nextafter(b.x3max, b.x3min);
    } else {
      const DevBlock &b = B[m];
      const double lx = b.x1max - b.x1min, ly = b.x2max - b.x2min, lz = b.x3max - b.x3min;
      x = b.x1min + HashUniform01(key + 0x632be59bd9b4e019ULL)*lx;
      y = b.x2min + HashUniform01(key + 0x8cb92baa3f3d8dd7ULL)*ly;
      if (su.gauss_z) {
        z = 0.0;
        for (int t=0; t<64; ++t) {
          const unsigned long long kz = key + 0x58f38ded6f7c55b5ULL +
                                        static_cast<unsigned long long>(t)*0x2545f4914f6cdd1dULL;
          const double u1 = fmax(HashUniform01(kz), 1.0e-300);
          const double u2 = HashUniform01(kz + 0x1d8e4e27c47d124fULL);
          const double zz = su.hp*sqrt(-2.0*log(u1))*cos(2.0*M_PI*u2);
          if (zz >= g.x3min && zz < g.x3max) {z = zz; break;}
        }
        const int bz = min(g.nmbx3 - 1, static_cast<int>((z - g.x3min)/lzm*g.nmbx3));
        m = b.lx1 + g.nmbx1*(b.lx2 + g.nmbx2*bz);
        const DevBlock &bb = B[m];
        if (z < bb.x3min) z = bb.x3min;
        if (z >= bb.x3max) z = nextafter(bb.x3max, bb.x3min);
      } else {
        z = b.x3min + HashUniform01(key + 0x58f38ded6f7c55b5ULL)*lz;
      }
    }
    double ts = su.taus[sp], wm = 1.0;
    if (su.continuous) {
      const double u = HashUniform01(key + DK_SALT_TAU);
      ts = exp(log(su.tau_lo[sp]) + u*log(su.tau_hi[sp]/su.tau_lo[sp]));
      if (su.mass_slope != 0.0) wm = pow(ts, su.mass_slope)/a.wmean[sp];
    }
    const DevBlock &b = B[m];
    const double vol = b.dx1*b.dx2*b.dx3;
    double mass;
    if (su.zmetal > 0.0) mass = wm*su.zmetal*a.gas_mass/(static_cast<double>(a.ntot));
    else mass = wm*su.eps[sp]*su.rho0*vol/ppc_s;
    double vm[3] = {su.gas_vx, su.gas_vy, su.gas_vz};
    if (su.nsh) {   // dust_setup.hpp:58-64
      const double t = su.omega0*ts, bq = 2.0 - su.qshear;
      vm[0] = (su.gas_vx + 2.0*t*su.gas_vy)/(1.0 + 2.0*bq*t*t);
      vm[1] = su.gas_vy - bq*t*vm[0];
      vm[2] = 0.0;
    }
    double v[3];
    for (int c=0; c<3; ++c) v[c] = vm[c] + su.dust_noise*Gauss(key + saltv[c]);
    r[DK_IPX*cap + q] = x; r[DK_IPY*cap + q] = y; r[DK_IPZ*cap + q] = z;
    r[DK_IPVX*cap + q] = v[0]; r[DK_IPVY*cap + q] = v[1]; r[DK_IPVZ*cap + q] = v[2];
    r[DK_IPX1*cap + q] = x; r[DK_IPY1*cap + q] = y; r[DK_IPZ1*cap + q] = z;
    r[DK_IPVX1*cap + q] = v[0]; r[DK_IPVY1*cap + q] = v[1]; r[DK_IPVZ1*cap + q] = v[2];
    r[DK_IPRX*cap + q] = 0.0; r[DK_IPRY*cap + q] = 0.0; r[DK_IPRZ*cap + q] = 0.0;
    r[DK_IPTS*cap + q] = ts; r[DK_IPM*cap + q] = mass;
    iv[DK_PGID*cap + q] = m; iv[DK_PTAG*cap + q] = static_cast<int32_t>(q);
    // (a refinable cloud of its block's level: half-width one cell of that level)
    iv[DK_PSP*cap + q] = sp; iv[DK_PLEV*cap + q] = b.level; iv[DK_PTAGH*cap + q] = 0;
  }
}

__global__ void k_setup_gas(Geo g, const DevBlock *B, double *u0, dk_setup su) {
  const unsigned long long base = su.seed*0x9e3779b97f4a7c15ULL;
  const unsigned long long saltg[3] = {DK_SALT_G0, DK_SALT_G1, DK_SALT_G2};
  const long long n = (long long)g.nmb*g.ncb;
  for (long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x*blockDim.x) {
    const int m = static_cast<int>(t/g.ncb);
    long long rr = t - (long long)m*g.ncb;
    const int i = static_cast<int>(rr % g.n1); rr /= g.n1;
    const int j = static_cast<int>(rr % g.n2);
    const int k = static_cast<int>(rr/g.n2);
    const bool act = (i >= g.is && i <= g.ie && j >= g.js && j <= g.je && k >= g.ks &&
                      k <= g.ke);
    if (!act) {
      for (int v=0; v<DK_NGAS; ++v) u0[Idx(g,m,DK_NGAS,v,k,j,i)] = 0.0;
      continue;
    }
    const DevBlock &b = B[m];
    // the global cell index of the block's level, plus level * 2^40
    const long long ny = (long long)g.nx2*(g.nmbx2 << b.level);
    const long long nxg = (long long)g.nx1*(g.nmbx1 << b.level);
    const long long gi = b.lx1*g.nx1 + (i - g.is), gj = b.lx2*g.nx2 + (j - g.js);
    const long long gk = b.lx3*g.nx3 + (k - g.ks);
    const unsigned long long ck = static_cast<unsigned long long>((gk*ny + gj)*nxg + gi) + base +
                                  (static_cast<unsigned long long>(b.level) << 40);
    const double z = b.x3min + (k - g.ks + 0.5)*b.dx3;
    const double rho = su.gas_strat ? su.rho0*exp(-0.5*z*z) : su.rho0;
    const double um[3] = {su.gas_vx, su.gas_vy, su.gas_vz};
    u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] = rho;
    for (int c=0; c<3; ++c)
      u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] = rho*(um[c] + su.gas_noise*Gauss(ck + saltg[c]));
  }
}

}  // namespace

extern "C" {

/* Fills the bound particle and gas arrays (dk_bind_particles/dk_bind_gas) with the problem of
 * su and sets the particle count; in DK_FORM_FP32 the particles are then rebound. */
int dk_setup_problem(dk_ctx *c, const dk_setup *su) {
  if (c == nullptr || su == nullptr) return DK_EINVAL;
  if (c->rdata == nullptr || c->u0 == nullptr) return DK_ESTATE;
  if (su->nspecies < 1 || su->nspecies > DK_MAXSPEC || su->ppc < 1) return DK_EINVAL;
  const Geo &g = c->g;
  if ((su->block_lognormal || c->smr) && (su->nclump > 0 || su->gauss_z)) return DK_EINVAL;
  SetupArgs a;
  a.su = *su;
  a.npb = (long long)su->ppc*g.nx1*g.nx2*g.nx3;
  a.ntot = a.npb*g.nmb;
  a.boff = nullptr;
  std::vector<long long> off;
  if (su->block_lognormal) {
    // every block's count, drawn on the host: round(ppc ncell exp(sigma z - sigma^2/2))
    off.assign(g.nmb + 1, 0);
    const unsigned long long base = su->seed*0x9e3779b97f4a7c15ULL;
    const double sg = su->block_sigma;
    for (int m=0; m<g.nmb; ++m) {
      const double z = HostGauss(base + static_cast<unsigned long long>(m)*0x2545f4914f6cdd1dULL +
                                 DK_SALT_BLOCK);
      off[m + 1] = off[m] + std::llround(static_cast<double>(a.npb)*std::exp(sg*z - 0.5*sg*sg));
    }
    a.ntot = off[g.nmb];
  }
  if (a.ntot > c->cap) return DK_ECAPACITY;
  if (su->block_lognormal) {
    if (c->d_block_off) (void)hipFree(c->d_block_off);
    c->d_block_off = nullptr;
    if (hipMalloc((void**)&c->d_block_off, sizeof(long long)*(g.nmb + 1)) != hipSuccess)
      return DK_ENOMEM;
    const int e0 = dk::Check(hipMemcpy(c->d_block_off, off.data(), sizeof(long long)*(g.nmb + 1),
                                       hipMemcpyHostToDevice));
    if (e0) return e0;
    a.boff = c->d_block_off;
  }
  // total gas mass in the oracle's order (blocks, k; every row alike)
  a.gas_mass = 0.0;
  for (int m=0; m<g.nmb; ++m) {
    const dk_block &b = c->B[m];
    const double dx1 = (b.x1max - b.x1min)/g.nx1, dx2 = (b.x2max - b.x2min)/g.nx2;
    const double dx3 = (b.x3max - b.x3min)/g.nx3;
    const double vol = dx1*dx2*dx3;
    for (int k=g.ks; k<=g.ke; ++k) {
      const double z = b.x3min + (k - g.ks + 0.5)*dx3;
      const double rho = su->gas_strat ? su->rho0*std::exp(-0.5*z*z) : su->rho0;
      a.gas_mass += rho*vol*g.nx1*g.nx2;
    }
  }
  for (int sp=0; sp<su->nspecies; ++sp) {
    a.wmean[sp] = 1.0;
    if (su->continuous && su->mass_slope != 0.0) {
      const double s = su->mass_slope, l = std::log(su->tau_hi[sp]/su->tau_lo[sp]);
      a.wmean[sp] = (std::pow(su->tau_hi[sp], s) - std::pow(su->tau_lo[sp], s))/(s*l);
    }
  }
  const int grid = static_cast<int>(std::min<long long>((a.ntot + kBlock - 1)/kBlock, 65536));
  k_setup_particles<<<grid, kBlock, 0, c->stream>>>(g, c->d_blocks, c->rdata, c->idata,
                                                     (long long)c->cap, a);
  const long long nc = (long long)g.nmb*g.ncb;
  const int gridc = static_cast<int>(std::min<long long>((nc + kBlock - 1)/kBlock, 65536));
  k_setup_gas<<<gridc, kBlock, 0, c->stream>>>(g, c->d_blocks, c->u0, *su);
  int e = dk::Check(hipGetLastError());
  if (e) return e;
  c->np = a.ntot;
  dk_stage st;
  std::memset(&st, 0, sizeof(st));
  st.stage = 1; st.nexp_stages = 3;
  // ghosts of u0 (copy, shear at t = 0, x3 values), then u1 = u0
  if ((e = dk_fill_gas_ghosts(c, &st))) return e;
  e = dk::Check(hipMemcpyAsync(c->u1, c->u0, sizeof(double)*DK_NGAS*nc,
                               hipMemcpyDeviceToDevice, c->stream));
  if (e) return e;
  if (c->form == DK_FORM_FP32) return dk_bind_particles(c, c->rdata, c->idata, c->cap, c->np);
  return DK_OK;
}

}  // extern "C"
