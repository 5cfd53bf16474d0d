// dk_fp32.hip -- the FP32 cell-sorted form of the imex2+ drag coupling.
//
// Precision: positions and RK position registers, the gas momentum u0, the kick's momentum
// exchange D and every global sum stay FP64; velocities, registers, drag rates, stopping times,
// masses and the mesh work fields (Q, P, u*, the dc1 operator result, the split residual) are
// FP32.  The particle impulse a kick hands to the gas is the exact difference of the FP32
// velocities before and after (exact in FP64), deposited with FP64 weights, and the consistent
// gas/dust split gives the gas D minus the dust's actual change: total momentum is conserved to
// FP64 round-off.
//
// Order: particles are kept in a store sorted by (block, cell of the deposit position).  Every
// stage a key pass computes each particle's new cell, a counting sort (atomic histogram,
// exclusive scan, placement, then every cell's list ordered by source slot so that the result is
// the stable order) gives the new order, and the push pass moves every particle into it.  Each
// deposit is atomic-free: a thread owns one cell and sums its particles' 27 stencil terms in
// registers in sorted order; the threads of a tile then add their sums into an LDS copy of the
// tile and its halo in 27 phases, one stencil offset per phase (no two threads of a phase touch
// the same cell); the tiles' halos are combined per cell in fixed tile order and the ghost
// deposits folded.  Results are bit-identical from run to run.
//
// The stage follows the task order of src/dust/dust_tasks.cpp:58-109: (pending consistent-dust
// correction of the previous stage, dust_pm.cpp:686-696) -> register copy / push / migration
// (dust_push.cpp:35-297, bvals_part.cpp:164-429) -> deposit Q, P (dust_pm.cpp:225-274) -> local
// guess + dc1 sweep(s) (dust_solver.cpp:126-263,630-724) -> gather-kick-PMBR (dust_pm.cpp:
// 500-564) -> consistent split (dust_pm.cpp:660-685).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "dk_internal.hpp"

using dk::DevBlock;
using dk::Dir;
using dk::Geo;
using dk::Idx;
using dk::kBlock;

namespace {

constexpr int kSegSmall = 32;       // lists longer than this are ordered by k_bigsort
constexpr int kBigMax = 16384;      // longest list of one cell in the FP32 form
constexpr int kTileThreads = 256;

struct Store {
  double *x[3], *x1[3], *r[3];   // positions, RK registers, drag rates (FP64: see k_p3)
  float *v[3], *v1[3];
  float *ts, *m;
  int *gid, *orig;
};

// Tile geometry: a tile is TX x TY x TZ active cells of one block, one thread per cell; its LDS
// region adds a halo of 1 cell (plus SY more in x2 when deposit centres may sit one cell beyond
// a block's x2 face: the orbital-advection frame shift of the 3-D shearing box).
struct Tiles {
  int TX, TY, TZ, SY, RX, RY, RZ, RN;
  int ntx, nty, ntz, ntb, nt;
};

}  // namespace

struct Fast {
  Tiles T;
  Store s[2];
  int cur = 0;
  long long np = 0, cap = 0, nkeys = 0;
  int *key = nullptr, *rank = nullptr, *count = nullptr, *start = nullptr, *perm = nullptr;
  int *big = nullptr;               // [0] = count of long lists, then their cells
  void *scan_tmp = nullptr;
  size_t scan_bytes = 0;
  int32_t *itag = nullptr;          // caller's PTAG, PSP, PLEV, PTAGH by original index
  // fields, gas layout with ghosts
  float *qdep = nullptr;            // Q, P1, P2, P3
  float *qt[3] = {nullptr, nullptr, nullptr};   // centre mass m c/V: own, from below, above
  float *ustar = nullptr, *ap = nullptr, *res = nullptr;
  double *dmom = nullptr;           // D of the kick, then R_g
  double *dc[3] = {nullptr, nullptr, nullptr};  // actual dust correction: own, below, above
  double *pres = nullptr;           // push rounding residue per cell (to the gas)
  float *scr_f = nullptr;
  double *scr_d = nullptr;
  int pending = 0;
  double a_prev = 0.0;
};

namespace {

#define GRID_STRIDE(t, n) \
  for (long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x; t < (n); \
       t += (long long)gridDim.x*blockDim.x)

inline int Grid(long long n) {
  long long b = (n + kBlock - 1)/kBlock;
  if (b < 1) b = 1;
  return static_cast<int>(std::min<long long>(b, 65536));
}

int Largest(int n, int cap) {
  for (int t = std::min(n, cap); t >= 1; --t)
    if (n % t == 0) return t;
  return 1;
}

// all numbers a tile kernel needs, by value
struct KArgs {
  Geo g;
  Tiles T;
  const DevBlock *B;
  Store src, dst;
  const int *perm, *start;
  int *key, *rank, *count;
  dk::ErrWord *err;
  // stage numbers
  double gam0, gam1, beta_dt, atw_dt, a_dt, sh, dt;
  double ysh, dsh;                  // shear crossing offsets (bvals_part.cpp, dust.cpp)
  int stage1, sbox, strat, shear;
  double qs, om;
  // pending correction
  int pending;
  double a_prev;
  const float *res;
  double *dc0, *dc1, *dc2;
};

__device__ inline void TileCell(const Tiles &T, const Geo &g, int tile, int t, int &m, int &ci,
                                int &cj, int &ck, int &lx, int &ly, int &lz) {
  m = tile/T.ntb;
  int rr = tile - m*T.ntb;
  const int tx = rr % T.ntx; rr /= T.ntx;
  const int ty = rr % T.nty;
  const int tz = rr/T.nty;
  lx = t % T.TX;
  ly = (t/T.TX) % T.TY;
  lz = t/(T.TX*T.TY);
  ci = tx*T.TX + lx; cj = ty*T.TY + ly; ck = tz*T.TZ + lz;
}

__device__ inline long long CellKey(const Geo &g, int m, int ci, int cj, int ck) {
  return (long long)m*g.nx1*g.nx2*g.nx3 + ((long long)ck*g.nx2 + cj)*g.nx1 + ci;
}

// the cell (active index, unclamped) of position x along one axis, as PMWeights
__device__ inline int CellOf(double x, double xmin, double xmax, int nx) {
#pragma clang fp contract(off)
  const double cell = (x - xmin)*(static_cast<double>(nx)/(xmax - xmin));
  return static_cast<int>(cell + 1.0) - 1;
}

// The push of one particle (dust_push.cpp:262-297) and its migration (bvals_part.cpp:164-429,
// dust.cpp:1678-1702), without the x3 outflow bookkeeping: new position, register shifts and
// block.  Returns false on a routing error.
struct PState {
  double x[3], x1[3];
  double r[3];
  float v[3], v1[3];
  double vres[3];   // exact FP64 push result minus its FP32 rounding
  int gid;
  int parked;       // left through an outflow x3 face: zero mass from now on
};

__device__ inline bool PushMigrate(const KArgs &A, PState &s, bool write_regs) {
#pragma clang fp contract(off)
  const Geo &g = A.g;
  const double xo = s.x[0], yo = s.x[1], zo = s.x[2];
  const float vxo = s.v[0], vyo = s.v[1], vzo = s.v[2];
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (A.sbox) {
    fx = static_cast<float>(2.0*A.om)*vyo;
    fy = static_cast<float>(-(2.0-A.qs)*A.om)*vxo;
    if (A.strat) fz = static_cast<float>(-(A.om*A.om)*zo);
  }
  if (write_regs) {   // stage 1: (x1, v1) <- (x, v) (dust_push.cpp:69-77)
    for (int d=0; d<3; ++d) {s.x1[d] = s.x[d]; s.v1[d] = s.v[d];}
  }
  // the new velocity in FP64 from the stored FP32 values, rounded once to FP32; the rounding
  // residue goes to the gas (momentum-conserving rounding)
  const double fxd = A.sbox ? 2.0*A.om*static_cast<double>(vyo) : 0.0;
  const double fyd = A.sbox ? -(2.0-A.qs)*A.om*static_cast<double>(vxo) : 0.0;
  const double fzd = (A.sbox && A.strat) ? -(A.om*A.om)*zo : 0.0;
  (void)fx; (void)fy; (void)fz;
  const double vn[3] = {
      A.gam0*vxo + A.gam1*s.v1[0] + A.beta_dt*fxd + A.atw_dt*s.r[0],
      A.gam0*vyo + A.gam1*s.v1[1] + A.beta_dt*fyd + A.atw_dt*s.r[1],
      A.gam0*vzo + A.gam1*s.v1[2] + A.beta_dt*fzd + A.atw_dt*s.r[2]};
  for (int d=0; d<3; ++d) {
    s.v[d] = static_cast<float>(vn[d]);
    s.vres[d] = vn[d] - static_cast<double>(s.v[d]);
  }
  double ydot = vyo;
  if (A.sbox) ydot -= A.qs*A.om*xo;
  s.x[0] = A.gam0*xo + A.gam1*s.x1[0] + A.beta_dt*static_cast<double>(vxo);
  s.x[1] = A.gam0*yo + A.gam1*s.x1[1] + A.beta_dt*ydot;
  s.x[2] = A.gam0*zo + A.gam1*s.x1[2] + A.beta_dt*static_cast<double>(vzo);
  // ShiftShearCrossings
  const DevBlock *B = A.B;
  if (A.shear && A.dsh != 0.0) {
    const double b0 = B[s.gid].x1min, b1 = B[s.gid].x1max, lx = b1 - b0;
    const int ix = static_cast<int>((s.x[0] - b0 + lx)/lx) - 1;
    const bool lo1 = (ix < 0) && (b0 < g.x1min + 0.5*lx);
    const bool hi1 = (ix > 0) && (b1 > g.x1max - 0.5*lx);
    if (lo1 || hi1) {
      const double d = hi1 ? A.dsh : -A.dsh;
      s.x[1] += d; s.x1[1] += d;
    }
  }
  // x3 physical faces
  if (g.bc_x3 != 0 && (s.x[2] < g.x3min || s.x[2] >= g.x3max)) {
    const bool lo = (s.x[2] < g.x3min);
    if (g.bc_x3 == DK_X3_REFLECT) {
      const double zf = lo ? g.x3min : g.x3max;
      double z = 2.0*zf - s.x[2];
      if (z >= g.x3max) z = g.x3max - 1.0e-12*(g.x3max - g.x3min);
      if (z < g.x3min) z = g.x3min;
      s.x[2] = z; s.v[2] = -s.v[2]; s.vres[2] = -s.vres[2];
      s.x1[2] = 2.0*zf - s.x1[2]; s.v1[2] = -s.v1[2];
    } else {
      // outflow: park the particle at rest in the boundary cell of its block with zero mass
      // (bvals_part.cpp:184-196); its momentum leaves with it
      const DevBlock &bo = A.B[s.gid];
      const double z = lo ? (bo.x3min + 0.5*bo.dx3) : (bo.x3max - 0.5*bo.dx3);
      s.x[2] = z;
      for (int d=0; d<3; ++d) {s.v[d] = 0.0f; s.vres[d] = 0.0; s.v1[d] = 0.0f;}
      s.x1[0] = s.x[0]; s.x1[1] = s.x[1]; s.x1[2] = z;
      s.parked = 1;
    }
  }
  const DevBlock &b = B[s.gid];
  const double lx = b.x1max - b.x1min, ly = b.x2max - b.x2min, lz = b.x3max - b.x3min;
  const int ix = static_cast<int>((s.x[0] - b.x1min + lx)/lx) - 1;
  const int iy = static_cast<int>((s.x[1] - b.x2min + ly)/ly) - 1;
  const int iz = static_cast<int>((s.x[2] - b.x3min + lz)/lz) - 1;
  const double bmin[3] = {b.x1min, b.x2min, b.x3min}, bmax[3] = {b.x1max, b.x2max, b.x3max};
  const int off[3] = {ix, iy, iz};
  for (int d=0; d<3; ++d) {
    if (off[d] != 0) continue;
    double xs = s.x[d];
    if (s.x[d] < bmin[d]) xs = bmin[d];
    else if (s.x[d] >= bmax[d]) xs = nextafter(bmax[d], bmin[d]);
    if (xs != s.x[d]) {s.x1[d] += xs - s.x[d]; s.x[d] = xs;}
  }
  if (abs(ix) > 1 || abs(iy) > 1 || abs(iz) > 1) return false;
  if ((abs(ix) + abs(iy) + abs(iz)) == 0) return true;
  const bool lo1 = (ix < 0) && (b.x1min < g.x1min + 0.5*lx);
  const bool hi1 = (ix > 0) && (b.x1max > g.x1max - 0.5*lx);
  const bool lo2 = (iy < 0) && (b.x2min < g.x2min + 0.5*ly);
  const bool hi2 = (iy > 0) && (b.x2max > g.x2max - 0.5*ly);
  const bool lo3 = (iz < 0) && (b.x3min < g.x3min + 0.5*lz);
  const bool hi3 = (iz > 0) && (b.x3max > g.x3max - 0.5*lz);
  const double lxm = g.x1max - g.x1min, lym = g.x2max - g.x2min, lzm = g.x3max - g.x3min;
  if (A.shear && (lo1 || hi1)) {
    const bool cin = lo1;
    s.x[0] += cin ? lxm : -lxm; s.x1[0] += cin ? lxm : -lxm;
    double ynew = s.x[1] + (cin ? -A.ysh : A.ysh);
    if (ynew < g.x2min) ynew += lym;
    if (ynew >= g.x2max) ynew -= lym;
    const double dy = ynew - s.x[1];
    s.x[1] = ynew; s.x1[1] += dy;
    if (lo3) {s.x[2] += lzm; s.x1[2] += lzm;} else if (hi3) {s.x[2] -= lzm; s.x1[2] -= lzm;}
    int b2 = static_cast<int>(((s.x[1] - g.x2min)/lym)*static_cast<double>(g.nmbx2));
    b2 = (b2 < 0) ? 0 : ((b2 > g.nmbx2-1) ? (g.nmbx2-1) : b2);
    int b3 = static_cast<int>(((s.x[2] - g.x3min)/lzm)*static_cast<double>(g.nmbx3));
    b3 = (b3 < 0) ? 0 : ((b3 > g.nmbx3-1) ? (g.nmbx3-1) : b3);
    s.gid = (cin ? (g.nmbx1 - 1) : 0) + g.nmbx1*(b2 + g.nmbx2*b3);
    return true;
  }
  const int nb = b.nghbr[Dir(ix, iy, iz)];
  if (nb < 0) return false;
  s.gid = nb;
  if (lo1) {s.x[0] += lxm; s.x1[0] += lxm;} else if (hi1) {s.x[0] -= lxm; s.x1[0] -= lxm;}
  if (lo2) {s.x[1] += lym; s.x1[1] += lym;} else if (hi2) {s.x[1] -= lym; s.x1[1] -= lym;}
  if (lo3) {s.x[2] += lzm; s.x1[2] += lzm;} else if (hi3) {s.x[2] -= lzm; s.x1[2] -= lzm;}
  return true;
}

// the deposit cell of a position (y shifted by sh x, the gas's t^n frame) clamped into the block
// (cj is returned unclamped too: a centre one cell beyond the x2 face is handled by the caller)
__device__ inline void DepositCell(const Geo &g, const DevBlock &b, const double x[3],
                                   double sh, int &ci, int &cj, int &ck, int &cju) {
#pragma clang fp contract(off)
  ci = CellOf(x[0], b.x1min, b.x1max, g.nx1);
  cju = CellOf(x[1] + sh*x[0], b.x2min, b.x2max, g.nx2);
  ck = CellOf(x[2], b.x3min, b.x3max, g.nx3);
  ci = min(max(ci, 0), g.nx1 - 1);
  ck = min(max(ck, 0), g.nx3 - 1);
  cj = min(max(cju, 0), g.nx2 - 1);
}

// The three weights of one axis (PMWeights, src/dust/dust.hpp:1088-1111) about a given
// centre cell c: del = cell - (c + 1/2).  With c = PMWeights' own cell this is PMWeights; with
// c one below it (a position that rounds onto the upper face, cell = nx, del = -1/2) it gives
// the same two non-zero weights one slot lower (del = +1/2), inside the block.
__device__ inline void AxisW(double x, double xmin, double xmax, int nx, int scheme, int c,
                             double w[3]) {
#pragma clang fp contract(off)
  const double cell = (x - xmin)*(static_cast<double>(nx)/(xmax - xmin));
  const double del = cell - (static_cast<double>(c) + 0.5);
  if (scheme == 0) {
    w[0] = 0.0; w[1] = 1.0; w[2] = 0.0;
  } else if (scheme == 1) {
    w[0] = fmax(0.0, -del); w[1] = 1.0 - fabs(del); w[2] = fmax(0.0, del);
  } else {
    w[0] = 0.5*(0.5 - del)*(0.5 - del);
    w[1] = 0.75 - del*del;
    w[2] = 0.5*(0.5 + del)*(0.5 + del);
  }
}

// the centre cell of the unshifted position (dust_pm.cpp:649-659); false when it rounds onto
// a ghost cell, where the reference's centre-mass and correction touch no active cell
__device__ inline bool CenterCell(const Geo &g, const DevBlock &b, const double x[3], int &ci,
                                  int &cj, int &ck) {
  ci = CellOf(x[0], b.x1min, b.x1max, g.nx1);
  cj = CellOf(x[1], b.x2min, b.x2max, g.nx2);
  ck = CellOf(x[2], b.x3min, b.x3max, g.nx3);
  return ci >= 0 && ci < g.nx1 && cj >= 0 && cj < g.nx2 && ck >= 0 && ck < g.nx3;
}

__device__ inline void LoadP(const Store &S, long long p, PState &s, bool regs, bool rates) {
  for (int d=0; d<3; ++d) {
    s.x[d] = S.x[d][p];
    s.v[d] = S.v[d][p];
    s.x1[d] = regs ? S.x1[d][p] : 0.0;
    s.v1[d] = regs ? S.v1[d][p] : 0.0f;
    s.r[d] = rates ? S.r[d][p] : 0.0;
  }
  s.gid = S.gid[p];
  s.parked = 0;
}

//----------------------------------------------------------------------------------------
// key pass: the pending consistent-dust correction of the previous stage (dust_pm.cpp:686-696;
// the actual change of every particle is recorded for the gas), then the new cell of every
// particle after this stage's push and migration, and its slot in the atomic histogram
__global__ void __launch_bounds__(kTileThreads)
k_key(KArgs A, int do_key) {
  const Geo &g = A.g;
  int m, ci, cj, ck, lx, ly, lz;
  TileCell(A.T, g, blockIdx.x, threadIdx.x, m, ci, cj, ck, lx, ly, lz);
  if (threadIdx.x >= A.T.TX*A.T.TY*A.T.TZ) return;
  const long long kc = CellKey(g, m, ci, cj, ck);
  const int s0 = A.start[kc], s1 = A.start[kc + 1];
  const DevBlock &b = A.B[m];
  const double vol = b.dx1*b.dx2*b.dx3;
  double dc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // [own, from below, from above][comp]
  for (int q=s0; q<s1; ++q) {
    PState s;
    LoadP(A.src, q, s, !A.stage1, false);
    if (A.pending) {
      const float mass = A.src.m[q];
      int ic_, jc, kc_;
      if (mass > 0.0f && CenterCell(g, b, s.x, ic_, jc, kc_)) {
        // the centre: the own cell or one cell beside it in x2
        const int dj = jc - cj;   // -1, 0 or +1
        if (ic_ != ci || kc_ != ck || dj < -1 || dj > 1) {
          dk::RecordError(A.err, DK_EHALO, q);
          continue;
        }
        const float cjf = static_cast<float>(A.a_prev/(static_cast<double>(A.src.ts[q]) +
                                                       A.a_prev));
        const long long ic = Idx(g, m, 3, 0, g.ks + ck, g.js + jc, g.is + ci);
        double rr[3];
        for (int d=0; d<3; ++d) {
          const float res = A.res[ic + d*g.ncb];
          const float dv = cjf*res;
          const float vo = s.v[d];
          const float vn = vo + dv;
          s.v[d] = vn;
          // the recorded rate takes the actual change, so that the history terms of dust and
          // gas stay opposite (dust_pm.cpp:694)
          const double delta = static_cast<double>(vn) - static_cast<double>(vo);
          rr[d] = A.src.r[d][q] + delta/A.a_prev;
          dc[dj + 1][d] += static_cast<double>(mass)*delta/vol;
        }
        for (int d=0; d<3; ++d) {A.src.v[d][q] = s.v[d]; A.src.r[d][q] = rr[d];}
      }
    }
    if (!do_key) continue;
    if (!PushMigrate(A, s, A.stage1 != 0)) {dk::RecordError(A.err, DK_EROUTE, q); continue;}
    int ni, nj, nk, nju;
    DepositCell(g, A.B[s.gid], s.x, A.sh, ni, nj, nk, nju);
    const long long key = CellKey(g, s.gid, ni, nj, nk);
    A.key[q] = static_cast<int>(key);
    A.rank[q] = atomicAdd(&A.count[key], 1);
  }
  if (A.pending) {
    // own cell, and the cells beside it in x2 (every target written by exactly one thread)
    const long long i0 = Idx(g, m, 3, 0, g.ks + ck, g.js + cj, g.is + ci);
    for (int d=0; d<3; ++d) {
      A.dc0[i0 + d*g.ncb] = dc[1][d];
      A.dc1[i0 + g.n1 + d*g.ncb] = dc[2][d];   // to the cell above: "from below" there
      A.dc2[i0 - g.n1 + d*g.ncb] = dc[0][d];   // to the cell below: "from above" there
    }
  }
}

__global__ void k_place(const int *key, const int *rank, const int *start, long long np,
                        int *perm) {
  GRID_STRIDE(p, np) perm[start[key[p]] + rank[p]] = static_cast<int>(p);
}

// every cell's list in ascending source slot (the stable order); long lists are queued
__global__ void k_segsort(const int *start, long long nkeys, int *perm, int *big) {
  GRID_STRIDE(c, nkeys) {
    const int s0 = start[c], s1 = start[c + 1], n = s1 - s0;
    if (n <= 1) continue;
    if (n > kSegSmall) {
      const int slot = atomicAdd(&big[0], 1);
      big[1 + slot] = static_cast<int>(c);
      continue;
    }
    for (int i=s0+1; i<s1; ++i) {
      const int v = perm[i];
      int j = i - 1;
      while (j >= s0 && perm[j] > v) {perm[j + 1] = perm[j]; --j;}
      perm[j + 1] = v;
    }
  }
}

// one workgroup per long list: bitonic sort in LDS (padded to a power of two)
__global__ void __launch_bounds__(1024) k_bigsort(const int *start, int *perm, const int *big,
                                                  dk::ErrWord *err) {
  __shared__ int sh[kBigMax];
  const int nbig = big[0];
  for (int e = blockIdx.x; e < nbig; e += gridDim.x) {
    const int c = big[1 + e];
    const int s0 = start[c], n = start[c + 1] - s0;
    if (n > kBigMax) {
      if (threadIdx.x == 0) dk::RecordError(err, DK_ECAPACITY, c);
      continue;
    }
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int i = threadIdx.x; i < np2; i += blockDim.x) sh[i] = (i < n) ? perm[s0 + i] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= np2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < np2; i += blockDim.x) {
          const int l = i ^ j;
          if (l > i) {
            const bool up = ((i & k) == 0);
            const int a = sh[i], bb = sh[l];
            if ((a > bb) == up) {sh[i] = bb; sh[l] = a;}
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[s0 + i] = sh[i];
    __syncthreads();
  }
}

//----------------------------------------------------------------------------------------
// add per-offset sums into the tile's LDS region, one stencil offset per phase
template <typename T, int NC>
__device__ inline void Phases(T (&acc)[27][NC], T *lds, const Tiles &Tl, int rx, int ry, int rz,
                              bool active) {
  for (int o=0; o<27; ++o) {
    const int a = o % 3, bb = (o/3) % 3, cz = o/9;
    if (active) {
      const int rc = ((rz + cz - 1)*Tl.RY + (ry + bb - 1))*Tl.RX + (rx + a - 1);
      for (int n=0; n<NC; ++n) lds[n*Tl.RN + rc] += acc[o][n];
    }
    __syncthreads();
  }
}

template <typename T>
__device__ inline void ZeroLDS(T *lds, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = T(0);
  __syncthreads();
}

template <typename T>
__device__ inline void FlushLDS(const T *lds, T *scr, long long tile, int n) {
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) scr[tile*n + i] = lds[i];
}

//----------------------------------------------------------------------------------------
// P1: push + migration into the new order, and the deposit of Q = sum m c W/V and
// P = sum m c W v/V (dust_pm.cpp:225-274) plus the centre mass m c/V (dust_pm.cpp:660-667)
__global__ void __launch_bounds__(kTileThreads)
k_p1(KArgs A, float *qt0, float *qt1, float *qt2, double *pres, float *scr) {
  extern __shared__ float ldsf[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  int m, ci, cj, ck, lx, ly, lz;
  TileCell(Tl, g, blockIdx.x, threadIdx.x, m, ci, cj, ck, lx, ly, lz);
  const bool act = threadIdx.x < Tl.TX*Tl.TY*Tl.TZ;
  ZeroLDS(ldsf, 4*Tl.RN);
  float acc[27][4];
  for (int o=0; o<27; ++o) for (int n=0; n<4; ++n) acc[o][n] = 0.0f;
  float q3[3] = {0.0f, 0.0f, 0.0f};
  double pr3[3] = {0.0, 0.0, 0.0};   // push rounding residue of this cell's particles
  int nshift = 0;   // particles whose deposit centre is one cell beyond the block in x2
  int s0 = 0, s1 = 0;
  const DevBlock &b = A.B[m];
  const double vol = b.dx1*b.dx2*b.dx3;
  if (act) {
    const long long kc = CellKey(g, m, ci, cj, ck);
    s0 = A.start[kc]; s1 = A.start[kc + 1];
  }
  for (int pass=0; pass<2; ++pass) {
    if (pass == 1) {
      // second pass only where some thread of the tile has shifted centres
      if (!__syncthreads_or(nshift > 0)) break;
      for (int o=0; o<27; ++o) for (int n=0; n<4; ++n) acc[o][n] = 0.0f;
    }
    int dshift = 0;
    for (int q=s0; q<s1; ++q) {
      const int p = A.perm[q];
      PState s;
      LoadP(A.src, p, s, true, !A.stage1);
      if (!PushMigrate(A, s, A.stage1 != 0)) {dk::RecordError(A.err, DK_EROUTE, p); continue;}
      const float ts = A.src.ts[p], mass = s.parked ? 0.0f : A.src.m[p];
      if (pass == 0) {
        for (int d=0; d<3; ++d) {
          A.dst.x[d][q] = s.x[d]; A.dst.x1[d][q] = s.x1[d];
          A.dst.v[d][q] = s.v[d]; A.dst.v1[d][q] = s.v1[d];
        }
        A.dst.ts[q] = ts; A.dst.m[q] = mass; A.dst.gid[q] = s.gid; A.dst.orig[q] = A.src.orig[p];
      }
      int ni, nj, nk, nju;
      const DevBlock &bn = A.B[s.gid];
      DepositCell(g, bn, s.x, A.sh, ni, nj, nk, nju);
      if (s.gid != m || ni != ci || nj != cj || nk != ck) {
        dk::RecordError(A.err, DK_EHALO, p); continue;
      }
      const double cjd = A.a_dt/(static_cast<double>(ts) + A.a_dt);
      const float muc = static_cast<float>(static_cast<double>(mass)*cjd/vol);
      if (pass == 0) {
        for (int d=0; d<3; ++d) pr3[d] += static_cast<double>(mass)*s.vres[d]/vol;
        int ic_, jc, kc_;
        if (mass > 0.0f && CenterCell(g, b, s.x, ic_, jc, kc_)) {
          // the centre of the unshifted position: this cell or one beside it in x2
          if (ic_ != ci || kc_ != ck || jc - cj < -1 || jc - cj > 1) {
            dk::RecordError(A.err, DK_EHALO, p);
          } else {
            q3[jc - cj + 1] += muc;
          }
        }
      }
      // 0, or +-1: the deposit centre beyond the block's x2 face (frame shift)
      const int sj = Tl.SY ? (nju - nj) : 0;
      if (pass == 0 && sj != 0) {++nshift; continue;}
      if (pass == 1 && sj == 0) continue;
      if (pass == 1) dshift = sj;
      double wx[3], wy[3], wz[3];
      AxisW(s.x[0], b.x1min, b.x1max, g.nx1, g.deposit, ci, wx);
      AxisW(s.x[1] + A.sh*s.x[0], b.x2min, b.x2max, g.nx2, g.deposit, cj + sj, wy);
      AxisW(s.x[2], b.x3min, b.x3max, g.nx3, g.deposit, ck, wz);
      const float vx = s.v[0], vy = s.v[1], vz = s.v[2];
      for (int c=0; c<3; ++c)
        for (int bb=0; bb<3; ++bb) {
          const float wcb = static_cast<float>(wz[c]*wy[bb])*muc;
          for (int a=0; a<3; ++a) {
            const float w = wcb*static_cast<float>(wx[a]);
            const int o = (c*3 + bb)*3 + a;
            acc[o][0] += w;
            acc[o][1] += w*vx;
            acc[o][2] += w*vy;
            acc[o][3] += w*vz;
          }
        }
    }
    Phases<float, 4>(acc, ldsf, Tl, lx + 1, ly + 1 + Tl.SY + dshift, lz + 1, act);
  }
  if (act) {
    const long long i0 = Idx(g, m, 1, 0, g.ks + ck, g.js + cj, g.is + ci);
    qt0[i0] = q3[1];
    qt1[i0 + g.n1] = q3[2];
    qt2[i0 - g.n1] = q3[0];
    const long long i3 = Idx(g, m, 3, 0, g.ks + ck, g.js + cj, g.is + ci);
    for (int d=0; d<3; ++d) pres[i3 + d*g.ncb] = pr3[d];
  }
  FlushLDS(ldsf, scr, blockIdx.x, 4*Tl.RN);
}

// P2: the particle part of the dc1 operator, result += G^T (m c/V) G u* (dust_solver.cpp:151-201)
__global__ void __launch_bounds__(kTileThreads)
k_p2(KArgs A, const float *us, float *scr) {
  extern __shared__ float ldsf[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  int m, ci, cj, ck, lx, ly, lz;
  TileCell(Tl, g, blockIdx.x, threadIdx.x, m, ci, cj, ck, lx, ly, lz);
  const bool act = threadIdx.x < Tl.TX*Tl.TY*Tl.TZ;
  ZeroLDS(ldsf, 3*Tl.RN);
  float acc[27][3];
  int nshift = 0, s0 = 0, s1 = 0;
  const DevBlock &b = A.B[m];
  const double vol = b.dx1*b.dx2*b.dx3;
  if (act) {
    const long long kc = CellKey(g, m, ci, cj, ck);
    s0 = A.start[kc]; s1 = A.start[kc + 1];
  }
  for (int pass=0; pass<2; ++pass) {
    if (pass == 1 && !__syncthreads_or(nshift > 0)) break;
    for (int o=0; o<27; ++o) for (int n=0; n<3; ++n) acc[o][n] = 0.0f;
    int dshift = 0;
    for (int q=s0; q<s1; ++q) {
      const double x[3] = {A.src.x[0][q], A.src.x[1][q], A.src.x[2][q]};
      const int jdu = CellOf(x[1] + A.sh*x[0], b.x2min, b.x2max, g.nx2);
      const int sj = Tl.SY ? (jdu - min(max(jdu, 0), g.nx2 - 1)) : 0;
      const int jd = cj + sj;
      if (pass == 0 && sj != 0) {++nshift; continue;}
      if (pass == 1 && sj == 0) continue;
      if (pass == 1) dshift = sj;
      double wx[3], wy[3], wz[3];
      AxisW(x[0], b.x1min, b.x1max, g.nx1, g.deposit, ci, wx);
      AxisW(x[1] + A.sh*x[0], b.x2min, b.x2max, g.nx2, g.deposit, jd, wy);
      AxisW(x[2], b.x3min, b.x3max, g.nx3, g.deposit, ck, wz);
      float ww[27];
      for (int c=0; c<3; ++c)
        for (int bb=0; bb<3; ++bb)
          for (int a=0; a<3; ++a)
            ww[(c*3 + bb)*3 + a] = static_cast<float>(wz[c]*wy[bb])*static_cast<float>(wx[a]);
      float gth[3] = {0.0f, 0.0f, 0.0f};
      const long long ic = Idx(g, m, 3, 0, g.ks + ck, g.js + jd, g.is + ci);
      for (int c=0; c<3; ++c)
        for (int bb=0; bb<3; ++bb)
          for (int a=0; a<3; ++a) {
            const long long id = ic + ((long long)(c - 1)*g.n2 + (bb - 1))*g.n1 + (a - 1);
            const float w = ww[(c*3 + bb)*3 + a];
            gth[0] += w*us[id];
            gth[1] += w*us[id + g.ncb];
            gth[2] += w*us[id + 2*g.ncb];
          }
      const double cjd = A.a_dt/(static_cast<double>(A.src.ts[q]) + A.a_dt);
      const float fac = static_cast<float>(static_cast<double>(A.src.m[q])*cjd/vol);
      for (int d=0; d<3; ++d) gth[d] *= fac;
      for (int o=0; o<27; ++o)
        for (int d=0; d<3; ++d) acc[o][d] += ww[o]*gth[d];
    }
    Phases<float, 3>(acc, ldsf, Tl, lx + 1, ly + 1 + Tl.SY + dshift, lz + 1, act);
  }
  FlushLDS(ldsf, scr, blockIdx.x, 3*Tl.RN);
}

// P3: gather u*, kick, drag rate, and the FP64 deposit of the actual impulse
// (dust_pm.cpp:500-564)
__global__ void __launch_bounds__(kTileThreads)
k_p3(KArgs A, const float *us, double *scr) {
  extern __shared__ double ldsd[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  int m, ci, cj, ck, lx, ly, lz;
  TileCell(Tl, g, blockIdx.x, threadIdx.x, m, ci, cj, ck, lx, ly, lz);
  const bool act = threadIdx.x < Tl.TX*Tl.TY*Tl.TZ;
  ZeroLDS(ldsd, 3*Tl.RN);
  double acc[27][3];
  int nshift = 0, s0 = 0, s1 = 0;
  const DevBlock &b = A.B[m];
  const double vol = b.dx1*b.dx2*b.dx3;
  if (act) {
    const long long kc = CellKey(g, m, ci, cj, ck);
    s0 = A.start[kc]; s1 = A.start[kc + 1];
  }
  for (int pass=0; pass<2; ++pass) {
    if (pass == 1 && !__syncthreads_or(nshift > 0)) break;
    for (int o=0; o<27; ++o) for (int n=0; n<3; ++n) acc[o][n] = 0.0;
    int dshift = 0;
    for (int q=s0; q<s1; ++q) {
      const double x[3] = {A.src.x[0][q], A.src.x[1][q], A.src.x[2][q]};
      const int jdu = CellOf(x[1] + A.sh*x[0], b.x2min, b.x2max, g.nx2);
      const int sj = Tl.SY ? (jdu - min(max(jdu, 0), g.nx2 - 1)) : 0;
      const int jd = cj + sj;
      if (pass == 0 && sj != 0) {++nshift; continue;}
      if (pass == 1 && sj == 0) continue;
      if (pass == 1) dshift = sj;
      double wx[3], wy[3], wz[3];
      AxisW(x[0], b.x1min, b.x1max, g.nx1, g.deposit, ci, wx);
      AxisW(x[1] + A.sh*x[0], b.x2min, b.x2max, g.nx2, g.deposit, jd, wy);
      AxisW(x[2], b.x3min, b.x3max, g.nx3, g.deposit, ck, wz);
      float u[3] = {0.0f, 0.0f, 0.0f};
      const long long ic = Idx(g, m, 3, 0, g.ks + ck, g.js + jd, g.is + ci);
      for (int c=0; c<3; ++c)
        for (int bb=0; bb<3; ++bb) {
          const float wcb = static_cast<float>(wz[c]*wy[bb]);
          for (int a=0; a<3; ++a) {
            const long long id = ic + ((long long)(c - 1)*g.n2 + (bb - 1))*g.n1 + (a - 1);
            const float w = wcb*static_cast<float>(wx[a]);
            u[0] += w*us[id];
            u[1] += w*us[id + g.ncb];
            u[2] += w*us[id + 2*g.ncb];
          }
        }
      const float cj32 = static_cast<float>(A.a_dt/(static_cast<double>(A.src.ts[q]) + A.a_dt));
      double delta[3];
      for (int d=0; d<3; ++d) {
        const float vo = A.src.v[d][q];
        const float dv = cj32*(u[d] - vo);
        const float vn = vo + dv;
        A.src.v[d][q] = vn;
        delta[d] = static_cast<double>(vn) - static_cast<double>(vo);
        A.src.r[d][q] = delta[d]/A.a_dt;
      }
      const double fac = -static_cast<double>(A.src.m[q])/vol;
      for (int c=0; c<3; ++c)
        for (int bb=0; bb<3; ++bb) {
          const double wcb = wz[c]*wy[bb]*fac;
          for (int a=0; a<3; ++a) {
            const double w = wcb*wx[a];
            const int o = (c*3 + bb)*3 + a;
            acc[o][0] += w*delta[0];
            acc[o][1] += w*delta[1];
            acc[o][2] += w*delta[2];
          }
        }
    }
    Phases<double, 3>(acc, ldsd, Tl, lx + 1, ly + 1 + Tl.SY + dshift, lz + 1, act);
  }
  FlushLDS(ldsd, scr, blockIdx.x, 3*Tl.RN);
}

// every padded cell sums the tile regions covering it, in ascending tile order
template <typename T>
__global__ void k_combine(Geo g, Tiles Tl, const T *scr, int nvar, T *out) {
  const long long n = g.ncb*g.nmb;
  GRID_STRIDE(c, n) {
    const int m = static_cast<int>(c/g.ncb);
    long long rr = c - (long long)m*g.ncb;
    const int i = static_cast<int>(rr % g.n1); rr /= g.n1;
    const int j = static_cast<int>(rr % g.n2);
    const int k = static_cast<int>(rr/g.n2);
    // region coordinates relative to tile 0: r = (i - is + 1) - tx*TX etc.
    const int ai = i - g.is + 1, aj = j - g.js + 1 + Tl.SY, ak = k - g.ks + 1;
    T sum[5] = {T(0), T(0), T(0), T(0), T(0)};
    const int tz0 = max(0, (ak - Tl.RZ + Tl.TZ)/Tl.TZ), tz1 = min(Tl.ntz - 1, ak/Tl.TZ);
    const int ty0 = max(0, (aj - Tl.RY + Tl.TY)/Tl.TY), ty1 = min(Tl.nty - 1, aj/Tl.TY);
    const int tx0 = max(0, (ai - Tl.RX + Tl.TX)/Tl.TX), tx1 = min(Tl.ntx - 1, ai/Tl.TX);
    if (ai >= 0 && aj >= 0 && ak >= 0) {
      for (int tz=tz0; tz<=tz1; ++tz) {
        const int rz = ak - tz*Tl.TZ;
        if (rz < 0 || rz >= Tl.RZ) continue;
        for (int ty=ty0; ty<=ty1; ++ty) {
          const int ry = aj - ty*Tl.TY;
          if (ry < 0 || ry >= Tl.RY) continue;
          for (int tx=tx0; tx<=tx1; ++tx) {
            const int rx = ai - tx*Tl.TX;
            if (rx < 0 || rx >= Tl.RX) continue;
            const long long tile = (long long)m*Tl.ntb + (tz*Tl.nty + ty)*Tl.ntx + tx;
            const long long base = tile*nvar*Tl.RN + (rz*Tl.RY + ry)*Tl.RX + rx;
            for (int v=0; v<nvar; ++v) sum[v] += scr[base + (long long)v*Tl.RN];
          }
        }
      }
    }
    for (int v=0; v<nvar; ++v) out[Idx(g,m,nvar,v,k,j,i)] = sum[v];
  }
}

//----------------------------------------------------------------------------------------
// cell passes (active cells)

#define ACTIVE_CELLS(t, m, k, j, i) \
  const long long nact_ = (long long)g.nmb*g.nx3*g.nx2*g.nx1; \
  GRID_STRIDE(t, nact_) { \
    long long rr_ = t; \
    const int i = g.is + static_cast<int>(rr_ % g.nx1); rr_ /= g.nx1; \
    const int j = g.js + static_cast<int>(rr_ % g.nx2); rr_ /= g.nx2; \
    const int k = g.ks + static_cast<int>(rr_ % g.nx3); \
    const int m = static_cast<int>(rr_/g.nx3);

// stage start: the gas share of the previous split, u0 += D - (actual dust correction),
// R_g = that/a (dust_pm.cpp:670-685); register copy u1 <- u0 (stage 1, dust_push.cpp:65);
// the harness's gas update (zero-flux RK combination + source terms); the drag history
// (stage 2, dust_pm.cpp:734-739)
__global__ void k_cell_start(Geo g, const DevBlock *B, double *u0, double *u1, double *dmom,
                             const double *dc0, const double *dc1, const double *dc2,
                             const double *pres,
                             int pending, double inv_aprev, int stage1, int hook, double gam0,
                             double gam1, double bdt, int use_src, dk_gas_src src,
                             double atw_dt) {
  ACTIVE_CELLS(t, m, k, j, i)
    double u[4];
    for (int v=0; v<4; ++v) u[v] = u0[Idx(g,m,DK_NGAS,v,k,j,i)];
    double rg[3] = {0.0, 0.0, 0.0};
    if (pending) {
      for (int c=0; c<3; ++c) {
        const long long id = Idx(g,m,3,c,k,j,i);
        const double dgas = dmom[id] - ((dc0[id] + dc1[id]) + dc2[id]);
        u[1+c] += dgas + pres[id];   // pres: what the FP32 rounding of the push withheld
        rg[c] = dgas*inv_aprev;
        dmom[id] = rg[c];
      }
    } else {
      for (int c=0; c<3; ++c) rg[c] = dmom[Idx(g,m,3,c,k,j,i)];
    }
    if (stage1) for (int v=0; v<4; ++v) u1[Idx(g,m,DK_NGAS,v,k,j,i)] = u[v];
    if (hook) {
      double w[4];
      w[0] = u[0];
      for (int c=1; c<4; ++c) w[c] = u[c]/u[0];
      for (int v=0; v<4; ++v) u[v] = gam0*u[v] + gam1*u1[Idx(g,m,DK_NGAS,v,k,j,i)];
      if (use_src) {
        if (src.const_accel != 0.0) u[src.const_accel_dir] += bdt*src.const_accel*w[0];
        if (src.shearing_box) {
          const double coef1 = 2.0*bdt*src.omega0;
          const double coef2 = (2.0-src.qshear)*bdt*src.omega0;
          const double coef3 = bdt*(src.omega0*src.omega0);
          const double mom1 = w[0]*w[1], mom2 = w[0]*w[2];
          u[1] += coef1*mom2;
          u[2] -= coef2*mom1;
          if (src.stratified) {
            const DevBlock &b = B[m];
            const double xx = (static_cast<double>(k - g.ks) + 0.5)/g.nx3;
            const double x3v = (xx*b.x3max - xx*b.x3min) - (0.5*b.x3max - 0.5*b.x3min) +
                               (0.5*b.x3min + 0.5*b.x3max);
            u[3] -= coef3*w[0]*x3v;
          }
        }
      }
    }
    if (atw_dt != 0.0) for (int c=0; c<3; ++c) u[1+c] += atw_dt*rg[c];
    for (int v=0; v<4; ++v) u0[Idx(g,m,DK_NGAS,v,k,j,i)] = u[v];
  }
}

// the local guess u* = (M + P)/(rho + Q) (dust_solver.cpp:643-649)
__global__ void k_guess(Geo g, const double *u0, const float *qd, float *us) {
  ACTIVE_CELLS(t, m, k, j, i)
    const double pinv = 1.0/(u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] +
                             static_cast<double>(qd[Idx(g,m,4,0,k,j,i)]));
    for (int c=0; c<3; ++c)
      us[Idx(g,m,3,c,k,j,i)] = static_cast<float>(
          (u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] + static_cast<double>(qd[Idx(g,m,4,1+c,k,j,i)]))*pinv);
  }
}

// one defect-correction sweep x += (rho + Q)^-1 (M + P - (Ax)) (dust_solver.cpp:243-252)
__global__ void k_dc(Geo g, const double *u0, const float *qd, const float *ap, float *us) {
  ACTIVE_CELLS(t, m, k, j, i)
    const double rho = u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
    const double pinv = 1.0/(rho + static_cast<double>(qd[Idx(g,m,4,0,k,j,i)]));
    for (int d=0; d<3; ++d) {
      const long long id = Idx(g,m,3,d,k,j,i);
      const double x = static_cast<double>(us[id]);
      const double av = static_cast<double>(ap[id]) + rho*x;
      const double rv = u0[Idx(g,m,DK_NGAS,DK_IM1+d,k,j,i)] +
                        static_cast<double>(qd[Idx(g,m,4,1+d,k,j,i)]) - av;
      us[id] = static_cast<float>(x + pinv*rv);
    }
  }
}

// the consistent split (dust_pm.cpp:670-685): the residual r = (D - G)/(rho + Q) of the
// dust per unit coupled mass res = r Q/qt for the particles centred here; the gas gets
// D - (the dust's actual change) at the next stage start
__global__ void k_split(Geo g, const double *u0, const float *us, const float *qd,
                        const float *qt0, const float *qt1, const float *qt2, const double *dmom,
                        float *res) {
  ACTIVE_CELLS(t, m, k, j, i)
    const double rho = u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
    const long long i1 = Idx(g,m,1,0,k,j,i);
    const double qt = static_cast<double>(qt0[i1]) + static_cast<double>(qt1[i1]) +
                      static_cast<double>(qt2[i1]);
    const double q = static_cast<double>(qd[Idx(g,m,4,0,k,j,i)]);
    for (int c=0; c<3; ++c) {
      const long long id = Idx(g,m,3,c,k,j,i);
      double rs = 0.0;
      if (qt > 0.0) {
        const double gg = rho*static_cast<double>(us[id]) - u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)];
        const double r = (dmom[id] - gg)/(rho + q);
        rs = r*q/qt;
      }
      res[id] = static_cast<float>(rs);
    }
  }
}

// bind: the caller's FP64 SoA into the store (original order)
__global__ void k_bind(const double *r, const int32_t *iv, long long cap, long long np, Store S,
                       int32_t *itag) {
  GRID_STRIDE(p, np) {
    for (int d=0; d<3; ++d) {
      S.x[d][p] = r[(long long)(DK_IPX + 2*d)*cap + p];
      S.v[d][p] = static_cast<float>(r[(long long)(DK_IPVX + 2*d)*cap + p]);
      S.x1[d][p] = r[(long long)(DK_IPX1 + 2*d)*cap + p];
      S.v1[d][p] = static_cast<float>(r[(long long)(DK_IPVX1 + 2*d)*cap + p]);
      S.r[d][p] = r[(long long)(DK_IPRX + d)*cap + p];
    }
    S.ts[p] = static_cast<float>(r[(long long)DK_IPTS*cap + p]);
    S.m[p] = static_cast<float>(r[(long long)DK_IPM*cap + p]);
    S.gid[p] = iv[(long long)DK_PGID*cap + p];
    S.orig[p] = static_cast<int>(p);
    for (int c=1; c<DK_NIDATA; ++c) itag[(long long)(c - 1)*np + p] = iv[(long long)c*cap + p];
  }
}

// the initial order: every particle's deposit cell at its present position
__global__ void k_key_static(Geo g, const DevBlock *B, Store S, long long np, double sh,
                             int *key, int *rank, int *count) {
  GRID_STRIDE(p, np) {
    const double x[3] = {S.x[0][p], S.x[1][p], S.x[2][p]};
    int ni, nj, nk, nju;
    DepositCell(g, B[S.gid[p]], x, sh, ni, nj, nk, nju);
    const long long k = CellKey(g, S.gid[p], ni, nj, nk);
    key[p] = static_cast<int>(k);
    rank[p] = atomicAdd(&count[k], 1);
  }
}

__global__ void k_permute_all(Store src, Store dst, const int *perm, long long np) {
  GRID_STRIDE(q, np) {
    const int p = perm[q];
    for (int d=0; d<3; ++d) {
      dst.x[d][q] = src.x[d][p]; dst.x1[d][q] = src.x1[d][p];
      dst.v[d][q] = src.v[d][p]; dst.v1[d][q] = src.v1[d][p]; dst.r[d][q] = src.r[d][p];
    }
    dst.ts[q] = src.ts[p]; dst.m[q] = src.m[p]; dst.gid[q] = src.gid[p];
    dst.orig[q] = src.orig[p];
  }
}

__global__ void k_export(Store S, long long np, const int32_t *itag, long long np0, double *r,
                         int32_t *iv, long long cap) {
  GRID_STRIDE(q, np) {
    for (int d=0; d<3; ++d) {
      r[(long long)(DK_IPX + 2*d)*cap + q] = S.x[d][q];
      r[(long long)(DK_IPVX + 2*d)*cap + q] = static_cast<double>(S.v[d][q]);
      r[(long long)(DK_IPX1 + 2*d)*cap + q] = S.x1[d][q];
      r[(long long)(DK_IPVX1 + 2*d)*cap + q] = static_cast<double>(S.v1[d][q]);
      r[(long long)(DK_IPRX + d)*cap + q] = S.r[d][q];
    }
    r[(long long)DK_IPTS*cap + q] = static_cast<double>(S.ts[q]);
    r[(long long)DK_IPM*cap + q] = static_cast<double>(S.m[q]);
    iv[(long long)DK_PGID*cap + q] = S.gid[q];
    const int o = S.orig[q];
    for (int c=1; c<DK_NIDATA; ++c) iv[(long long)c*cap + q] = itag[(long long)(c - 1)*np0 + o];
  }
}

// FP64 sums of the store (dust momentum, mass, scale) and the gas (as the FP64 form)
constexpr int kDiagGrid = 512;
__global__ void k_fdiag(Geo g, const DevBlock *B, Store S, long long np, const double *u0,
                        double *partial) {
  __shared__ double sh[kBlock][9];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  GRID_STRIDE(p, np) {
    const double m = S.m[p];
    const double vx = S.v[0][p], vy = S.v[1][p], vz = S.v[2][p];
    acc[0] += m*vx; acc[1] += m*vy; acc[2] += m*vz;
    acc[6] += m;
    acc[8] += m*(fabs(vx) + fabs(vy) + fabs(vz));
  }
  ACTIVE_CELLS(t, m, k, j, i)
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

__global__ void k_freduce(const double *partial, int nblk, double *out) {
  const int c = threadIdx.x;
  if (c >= 9) return;
  double s = 0.0;
  for (int b=0; b<nblk; ++b) s += partial[b*9 + c];
  out[c] = s;
}

template <typename T>
__global__ void k_fillv(T *a, long long n, T v) {
  GRID_STRIDE(t, n) a[t] = v;
}

//----------------------------------------------------------------------------------------
// host

template <typename T>
int FillV(dk_ctx *c, T *a, long long n, T v) {
  k_fillv<T><<<Grid(n), kBlock, 0, c->stream>>>(a, n, v);
  return dk::Check(hipGetLastError());
}

#define TIMED(label, launch) \
  do { const int tk_ = dk::ProfBegin(c, label); launch; dk::ProfEnd(c, tk_); } while (0)

bool ActiveStage(const dk_ctx *c, const dk_stage *st) {
  return !(st->nexp_stages == 3 && st->stage == st->nexp_stages);
}

double ImplicitShearShift(const dk_ctx *c, const dk_stage *st) {
  if (!c->P.shearing_box) return 0.0;
  double cc = 0.0;
  for (int k=0; k<st->stage; ++k) cc = st->gam0[k]*cc + st->beta[k];
  const double timp = st->time + cc*st->dt;
  return c->P.qshear*c->P.omega0*(timp - st->time);
}

double ShearRouteTime(const dk_stage *st) {
  double w = 0.0;
  for (int k = st->stage + 1; k <= st->nexp_stages; ++k) w = st->gam0[k-1]*w + st->beta[k-1];
  return st->time + (1.0 - w)*st->dt;
}

int Alloc(void **p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) {*p = nullptr; return DK_ENOMEM;}
  return DK_OK;
}

KArgs MakeArgs(dk_ctx *c) {
  Fast *f = c->fast;
  KArgs A;
  std::memset(&A, 0, sizeof(A));
  A.g = c->g; A.T = f->T; A.B = c->d_blocks;
  A.src = f->s[f->cur]; A.dst = f->s[1 - f->cur];
  A.perm = f->perm; A.start = f->start; A.key = f->key; A.rank = f->rank; A.count = f->count;
  A.err = c->d_err;
  A.sbox = c->P.shearing_box; A.strat = c->P.stratified;
  A.shear = (c->M.bc_x1 == DK_X1_SHEAR);
  A.qs = c->P.qshear; A.om = c->P.omega0;
  A.pending = f->pending; A.a_prev = f->a_prev;
  A.res = f->res; A.dc0 = f->dc[0]; A.dc1 = f->dc[1]; A.dc2 = f->dc[2];
  return A;
}

// counting sort of the store by the keys of the key pass: scan, placement, stable lists
int SortKeys(dk_ctx *c) {
  Fast *f = c->fast;
  int e;
  TIMED("scan", e = dk::Check(hipcub::DeviceScan::ExclusiveSum(
      f->scan_tmp, f->scan_bytes, f->count, f->start, (int)(f->nkeys + 1), c->stream)));
  if (e) return e;
  TIMED("place", (k_place<<<Grid(f->np), kBlock, 0, c->stream>>>(f->key, f->rank, f->start,
                                                                   f->np, f->perm)));
  if ((e = FillV<int>(c, f->big, 1, 0))) return e;
  TIMED("segsort", (k_segsort<<<Grid(f->nkeys), kBlock, 0, c->stream>>>(f->start, f->nkeys,
                                                                         f->perm, f->big)));
  TIMED("bigsort", (k_bigsort<<<256, 1024, 0, c->stream>>>(f->start, f->perm, f->big,
                                                            c->d_err)));
  return dk::Check(hipGetLastError());
}

// the pending consistent-dust correction alone (before an export or a diagnostic)
int Flush(dk_ctx *c) {
  Fast *f = c->fast;
  if (!f->pending) return DK_OK;
  KArgs A = MakeArgs(c);
  TIMED("key", (k_key<<<f->T.nt, kTileThreads, 0, c->stream>>>(A, 0)));
  dk_gas_src z;
  std::memset(&z, 0, sizeof(z));
  const long long nact = (long long)c->g.nmb*c->g.nx1*c->g.nx2*c->g.nx3;
  TIMED("cell_start", (k_cell_start<<<Grid(nact), kBlock, 0, c->stream>>>(
      c->g, c->d_blocks, c->u0, c->u1, f->dmom, f->dc[0], f->dc[1], f->dc[2], f->pres, 1,
      1.0/f->a_prev, 0, 0, 1.0, 0.0, 0.0, 0, z, 0.0)));
  f->pending = 0;
  return dk::Check(hipGetLastError());
}

}  // namespace

namespace dk {

int FastCreate(dk_ctx *c) {
  if (c->P.coupling != DK_IMEX) return DK_EUNSUPPORTED;          // pc2: the FP64 form
  if (c->P.solver == DK_APPLYA) return DK_EUNSUPPORTED;
  if (c->g.ng < 2 || !c->P.back_reaction) return DK_EUNSUPPORTED;
  Fast *f = new Fast();
  c->fast = f;
  Tiles &T = f->T;
  T.TX = Largest(c->g.nx1, 16);
  T.TY = Largest(c->g.nx2, 4);
  T.TZ = Largest(c->g.nx3, 4);
  T.SY = (c->P.shearing_box) ? 1 : 0;
  T.RX = T.TX + 2; T.RY = T.TY + 2 + 2*T.SY; T.RZ = T.TZ + 2;
  T.RN = T.RX*T.RY*T.RZ;
  T.ntx = c->g.nx1/T.TX; T.nty = c->g.nx2/T.TY; T.ntz = c->g.nx3/T.TZ;
  T.ntb = T.ntx*T.nty*T.ntz;
  T.nt = T.ntb*c->g.nmb;
  f->nkeys = (long long)c->g.nmb*c->g.nx1*c->g.nx2*c->g.nx3;
  if (f->nkeys >= INT32_MAX) return DK_ECAPACITY;
  const long long nc = NCells(c);
  int e = DK_OK;
  e = e ? e : Alloc((void**)&f->count, sizeof(int)*(f->nkeys + 1));
  e = e ? e : Alloc((void**)&f->start, sizeof(int)*(f->nkeys + 1));
  e = e ? e : Alloc((void**)&f->big, sizeof(int)*(f->nkeys + 1));
  e = e ? e : Alloc((void**)&f->qdep, sizeof(float)*4*nc);
  for (int q=0; q<3; ++q) e = e ? e : Alloc((void**)&f->qt[q], sizeof(float)*nc);
  e = e ? e : Alloc((void**)&f->ustar, sizeof(float)*3*nc);
  e = e ? e : Alloc((void**)&f->ap, sizeof(float)*3*nc);
  e = e ? e : Alloc((void**)&f->res, sizeof(float)*3*nc);
  e = e ? e : Alloc((void**)&f->dmom, sizeof(double)*3*nc);
  e = e ? e : Alloc((void**)&f->pres, sizeof(double)*3*nc);
  for (int q=0; q<3; ++q) e = e ? e : Alloc((void**)&f->dc[q], sizeof(double)*3*nc);
  e = e ? e : Alloc((void**)&f->scr_f, sizeof(float)*4*(size_t)T.RN*T.nt);
  e = e ? e : Alloc((void**)&f->scr_d, sizeof(double)*3*(size_t)T.RN*T.nt);
  if (e) return e;
  (void)hipMemset(f->qt[0], 0, sizeof(float)*nc);
  (void)hipMemset(f->qt[1], 0, sizeof(float)*nc);
  (void)hipMemset(f->qt[2], 0, sizeof(float)*nc);
  for (int q=0; q<3; ++q) (void)hipMemset(f->dc[q], 0, sizeof(double)*3*nc);
  (void)hipMemset(f->dmom, 0, sizeof(double)*3*nc);
  (void)hipMemset(f->pres, 0, sizeof(double)*3*nc);
  (void)hipMemset(f->ustar, 0, sizeof(float)*3*nc);
  (void)hipMemset(f->res, 0, sizeof(float)*3*nc);
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, f->scan_bytes, f->count, f->start,
                                       (int)(f->nkeys + 1)) != hipSuccess)
    return DK_EHIP;
  return Alloc(&f->scan_tmp, std::max<size_t>(f->scan_bytes, 16));
}

void FastDestroy(dk_ctx *c) {
  Fast *f = c->fast;
  if (f == nullptr) return;
  for (int b=0; b<2; ++b) {
    Store &S = f->s[b];
    for (int d=0; d<3; ++d)
      for (void *p : {(void*)S.x[d], (void*)S.x1[d], (void*)S.v[d], (void*)S.v1[d],
                      (void*)S.r[d]})
        if (p) (void)hipFree(p);
    for (void *p : {(void*)S.ts, (void*)S.m, (void*)S.gid, (void*)S.orig})
      if (p) (void)hipFree(p);
  }
  for (void *p : {(void*)f->key, (void*)f->rank, (void*)f->count, (void*)f->start,
                  (void*)f->perm, (void*)f->big, f->scan_tmp, (void*)f->itag, (void*)f->qdep,
                  (void*)f->ustar, (void*)f->ap, (void*)f->res, (void*)f->dmom, (void*)f->pres,
                  (void*)f->scr_f, (void*)f->scr_d})
    if (p) (void)hipFree(p);
  for (int q=0; q<3; ++q) {
    if (f->qt[q]) (void)hipFree(f->qt[q]);
    if (f->dc[q]) (void)hipFree(f->dc[q]);
  }
  delete f;
  c->fast = nullptr;
}

// copy the caller's particles in and sort them by cell
int FastBind(dk_ctx *c) {
  Fast *f = c->fast;
  if (f == nullptr) return DK_ESTATE;
  if (c->np > f->cap) {
    // (re)allocate the store for the new capacity
    const long long n = std::max<long long>(c->np, 1);
    for (int b=0; b<2; ++b) {
      Store &S = f->s[b];
      for (int d=0; d<3; ++d) {
        for (void **p : {(void**)&S.x[d], (void**)&S.x1[d], (void**)&S.r[d]}) {
          if (*p) (void)hipFree(*p);
          if (Alloc(p, sizeof(double)*n)) return DK_ENOMEM;
        }
        for (void **p : {(void**)&S.v[d], (void**)&S.v1[d]}) {
          if (*p) (void)hipFree(*p);
          if (Alloc(p, sizeof(float)*n)) return DK_ENOMEM;
        }
      }
      for (void **p : {(void**)&S.ts, (void**)&S.m, (void**)&S.gid, (void**)&S.orig}) {
        if (*p) (void)hipFree(*p);
        if (Alloc(p, 4*n)) return DK_ENOMEM;
      }
    }
    for (void **p : {(void**)&f->key, (void**)&f->rank, (void**)&f->perm}) {
      if (*p) (void)hipFree(*p);
      if (Alloc(p, sizeof(int)*n)) return DK_ENOMEM;
    }
    if (f->itag) (void)hipFree(f->itag);
    if (Alloc((void**)&f->itag, sizeof(int32_t)*(DK_NIDATA - 1)*n)) return DK_ENOMEM;
    f->cap = n;
  }
  f->np = c->np;
  f->cur = 0;
  f->pending = 0;
  if (f->np == 0) return DK_OK;
  k_bind<<<Grid(f->np), kBlock, 0, c->stream>>>(c->rdata, c->idata, (long long)c->cap, f->np,
                                                 f->s[0], f->itag);
  int e;
  if ((e = FillV<int>(c, f->count, f->nkeys + 1, 0))) return e;
  k_key_static<<<Grid(f->np), kBlock, 0, c->stream>>>(c->g, c->d_blocks, f->s[0], f->np, 0.0,
                                                      f->key, f->rank, f->count);
  if ((e = SortKeys(c))) return e;
  k_permute_all<<<Grid(f->np), kBlock, 0, c->stream>>>(f->s[0], f->s[1], f->perm, f->np);
  f->cur = 1;
  return dk::Check(hipGetLastError());
}

int FastExport(dk_ctx *c) {
  Fast *f = c->fast;
  if (f == nullptr) return DK_ESTATE;
  int e;
  if ((e = Flush(c))) return e;
  k_export<<<Grid(f->np), kBlock, 0, c->stream>>>(f->s[f->cur], f->np, f->itag, f->np,
                                                   c->rdata, c->idata, (long long)c->cap);
  return dk::Check(hipGetLastError());
}

int FastDiag(dk_ctx *c, double out[9]) {
  Fast *f = c->fast;
  if (f == nullptr || c->u0 == nullptr) return DK_ESTATE;
  int e;
  if ((e = Flush(c))) return e;
  k_fdiag<<<kDiagGrid, kBlock, 0, c->stream>>>(c->g, c->d_blocks, f->s[f->cur], f->np, c->u0,
                                                c->d_partial);
  k_freduce<<<1, 64, 0, c->stream>>>(c->d_partial, kDiagGrid, c->d_scalar);
  e = dk::Check(hipMemcpyAsync(out, c->d_scalar, sizeof(double)*9, hipMemcpyDeviceToHost,
                               c->stream));
  if (!e) e = dk::Check(hipStreamSynchronize(c->stream));
  return e;
}

int FastStage(dk_ctx *c, const dk_stage *st, const dk_gas_src *src) {
  Fast *f = c->fast;
  if (f == nullptr || c->u0 == nullptr) return DK_ESTATE;
  if (!ActiveStage(c, st)) return DK_OK;   // the identity stage of imex2+
  const Geo &g = c->g;
  const Tiles &T = f->T;
  const int s = st->stage - 1;
  const double a_dt = st->a_impl*st->dt;
  const double sh = ImplicitShearShift(c, st);
  // the orbital-advection frame shift must stay below half a cell (one-cell deposit slack)
  if (c->P.shearing_box) {
    const double xmax = std::max(std::fabs(c->M.x1min), std::fabs(c->M.x1max));
    const double dy = (c->M.x2max - c->M.x2min)/(g.nx2*g.nmbx2);
    if (sh*xmax > 0.5*dy) return DK_EUNSUPPORTED;
  }
  int e;
  const long long nact = (long long)g.nmb*g.nx1*g.nx2*g.nx3;
  KArgs A = MakeArgs(c);
  A.stage1 = (st->stage == 1);
  A.gam0 = st->gam0[s]; A.gam1 = st->gam1[s]; A.beta_dt = st->beta[s]*st->dt;
  A.atw_dt = (st->stage == 2) ? st->a_twid22*st->dt : 0.0;
  A.a_dt = a_dt; A.sh = sh; A.dt = st->dt;
  if (A.shear) {
    const double lx = c->M.x1max - c->M.x1min, ly = c->M.x2max - c->M.x2min;
    A.ysh = std::fmod(c->P.qshear*c->P.omega0*lx*st->time, ly);
    A.dsh = c->P.qshear*c->P.omega0*lx*(ShearRouteTime(st) - st->time);
  }
  // 1. key pass (with the pending correction), histogram
  if ((e = FillV<int>(c, f->count, f->nkeys + 1, 0))) return e;
  TIMED("key", (k_key<<<T.nt, kTileThreads, 0, c->stream>>>(A, 1)));
  // 2. stage start cell pass: previous split's gas share, u1 <- u0, gas update, drag history
  {
    const int hook = 1;
    dk_gas_src z;
    std::memset(&z, 0, sizeof(z));
    const double gam0 = st->gam0[s], gam1 = st->gam1[s], bdt = st->beta[s]*st->dt;
    TIMED("cell_start", (k_cell_start<<<Grid(nact), kBlock, 0, c->stream>>>(
        g, c->d_blocks, c->u0, c->u1, f->dmom, f->dc[0], f->dc[1], f->dc[2], f->pres, f->pending,
        f->pending ? 1.0/f->a_prev : 0.0, A.stage1, hook, gam0, gam1, bdt, src ? 1 : 0,
        src ? *src : z, (c->P.back_reaction && st->stage == 2) ? st->a_twid22*st->dt : 0.0)));
  }
  f->pending = 0;
  A.pending = 0;
  // 3. counting sort
  if ((e = SortKeys(c))) return e;
  // 4. push into the new order + deposit Q, P, qt
  const size_t lds4 = sizeof(float)*4*T.RN, lds3 = sizeof(float)*3*T.RN;
  const size_t ldsd = sizeof(double)*3*T.RN;
  TIMED("p1_push_deposit", (k_p1<<<T.nt, kTileThreads, lds4, c->stream>>>(
      A, f->qt[0], f->qt[1], f->qt[2], f->pres, f->scr_f)));
  f->cur = 1 - f->cur;
  A.src = f->s[f->cur];
  A.dst = f->s[1 - f->cur];
  TIMED("combine", (k_combine<float><<<Grid(NCells(c)), kBlock, 0, c->stream>>>(
      g, T, f->scr_f, 4, f->qdep)));
  if ((e = FoldT<float>(c, f->qdep, 4, st->time))) return e;
  // 5. local guess + dc1/dc2 sweeps
  TIMED("guess", (k_guess<<<Grid(nact), kBlock, 0, c->stream>>>(g, c->u0, f->qdep, f->ustar)));
  const int nsweeps = (c->P.solver == DK_DC2) ? 2 : ((c->P.solver == DK_DC1) ? 1 : 0);
  for (int sw=0; sw<nsweeps; ++sw) {
    if ((e = FillGhostsT<float>(c, f->ustar, 3, true, st->time))) return e;
    TIMED("p2_operator", (k_p2<<<T.nt, kTileThreads, lds3, c->stream>>>(A, f->ustar, f->scr_f)));
    TIMED("combine", (k_combine<float><<<Grid(NCells(c)), kBlock, 0, c->stream>>>(
        g, T, f->scr_f, 3, f->ap)));
    if ((e = FoldT<float>(c, f->ap, 3, st->time))) return e;
    TIMED("dc", (k_dc<<<Grid(nact), kBlock, 0, c->stream>>>(g, c->u0, f->qdep, f->ap, f->ustar)));
  }
  if ((e = FillGhostsT<float>(c, f->ustar, 3, true, st->time))) return e;
  // 6. gather-kick + impulse deposit
  TIMED("p3_gather_kick", (k_p3<<<T.nt, kTileThreads, ldsd, c->stream>>>(A, f->ustar,
                                                                         f->scr_d)));
  TIMED("combine", (k_combine<double><<<Grid(NCells(c)), kBlock, 0, c->stream>>>(
      g, T, f->scr_d, 3, f->dmom)));
  if ((e = FoldT<double>(c, f->dmom, 3, st->time))) return e;
  // 7. consistent split: the dust's residual per unit coupled mass
  TIMED("split", (k_split<<<Grid(nact), kBlock, 0, c->stream>>>(
      g, c->u0, f->ustar, f->qdep, f->qt[0], f->qt[1], f->qt[2], f->dmom, f->res)));
  f->pending = 1;
  f->a_prev = a_dt;
  return dk::Check(hipGetLastError());
}

}  // namespace dk
