// dk_fp32.hip -- the FP32 cell-sorted form of the imex2+ drag coupling.
//
// Precision: positions and RK position registers, the gas momentum u0, the kick's momentum
// exchange D and every global sum stay FP64; velocities, registers, drag rates, stopping times,
// masses and the mesh work fields (Q, P, u*, the dc1 operator result, the split residual) are
// FP32.  The particle impulse a kick hands to the gas is the exact difference of the FP32
// velocities before and after (exact in FP64); it is deposited in FP32, and every tile's defect
// (its particles' FP64 impulse sum minus the FP64 sum of its FP32 deposit) is added to one of
// its cells, so that the FP64 field D holds the particles' total impulse to FP64 round-off; the
// consistent gas/dust split gives the gas D minus the dust's actual change: total momentum is
// conserved to FP64 round-off.
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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "dk_internal.hpp"
#include "dk_smr.hpp"
#include "dk_cloudwalk.hpp"

using dk::DevBlock;
using dk::Dir;
using dk::Geo;
using dk::Idx;
using dk::kBlock;

namespace {

constexpr int kSegSmall = 32;       // lists longer than this are ordered by k_medsort
constexpr int kMedMax = 512;        // lists longer than this are ordered by k_bigsort
constexpr int kBigMax = 16384;      // longest list of one cell in the FP32 form
constexpr int kPartRank = 8192;     // a stayer's rank: part*kPartRank + its rank in the part
constexpr int kKeyThreads = 256;    // the key pass: one particle per thread per chunk (default)
constexpr int kKeyThreadsMax = 512;
constexpr int kOverGrid = 4096;     // workgroups of the launch over the extra parts of full tiles
constexpr int kCellThreads = 384;   // P1-P3: one thread per (cell, stencil slab), <= 128 cells
constexpr int kMaxTileCells = 128;

// The particle store: three buffers in groups of 64 particles (one wavefront), each group
// holding its columns one after another (double: x[3], x1[3], r[3]; float: v[3], v1[3], t_s, m;
// int: block, original index).  A wavefront reading one column of 64 consecutive particles
// reads one contiguous 256 B or 512 B piece, and every column of a particle is addressed from
// one base with a constant offset, so that a kernel holds few addresses in registers.
constexpr int kGrp = 64;
struct Store {
  double *d;
  float *f;
  int *i;
  long long cap;    // a multiple of kGrp
  __device__ static long long At(int ncol, int c, long long p) {
    return ((p >> 6)*ncol + c)*kGrp + (p & (kGrp - 1));
  }
  __device__ double &X(int c, long long p) const { return d[At(9, c, p)]; }
  __device__ double &X1(int c, long long p) const { return d[At(9, 3 + c, p)]; }
  __device__ double &R(int c, long long p) const { return d[At(9, 6 + c, p)]; }
  __device__ float &V(int c, long long p) const { return f[At(8, c, p)]; }
  __device__ float &V1(int c, long long p) const { return f[At(8, 3 + c, p)]; }
  __device__ float &TS(long long p) const { return f[At(8, 6, p)]; }
  __device__ float &M(long long p) const { return f[At(8, 7, p)]; }
  __device__ int &GID(long long p) const { return i[At(2, 0, p)]; }
  __device__ int &ORIG(long long p) const { return i[At(2, 1, p)]; }
};

// Tile geometry: a tile is TX x TY x TZ active cells of one block, one thread per cell; its LDS
// region adds a halo of 1 cell (plus SY more in x2 when deposit centres may sit one cell beyond
// a block's x2 face: the orbital-advection frame shift of the 3-D shearing box).
struct Tiles {
  int TX, TY, TZ, TC, SY, RX, RY, RZ, RN;
  int LX, LY, LZ;                   // log2 TX, TY, TZ (the tile sides are powers of 2)
  int ntx, nty, ntz, ntb, nt;
  float iRX, iRY, iRN;              // 1/RX, 1/RY, 1/RN for FDiv
  int xcd;                          // workgroups dealt to the 8 XCDs in runs of tiles (1) or not
  int wrows;                        // every wavefront of a tile kernel holds whole x1-x2 layers
};

}  // namespace

struct Fast {
  Tiles T;
  Store s[2];
  int cur = 0;
  long long np = 0, cap = 0, nkeys = 0;
  int *key = nullptr, *rank = nullptr, *count = nullptr, *start = nullptr, *perm = nullptr;
  int *stay = nullptr;              // per cell: the particles that stayed in its tile
  // [0], [1]: counts of medium and long lists; their cells from [2] up and from [1 + nkeys] down
  int *big = nullptr;
  // work items (KArgs): parts per tile, their scan, the tile of every item; overflow slots
  int *nparts = nullptr, *item_start = nullptr, *item_tile = nullptr;
  long long item_cap = 0, ov_cap = 0;
  void *iscan_tmp = nullptr;
  size_t iscan_bytes = 0;
  float *ov_scr = nullptr, *ov_q = nullptr, *ov_pr = nullptr;
  double *ov_def = nullptr, *ov_dc = nullptr;
  int *ov_cnt = nullptr;
  // shearing box: the items of the tiles inside a block's x2 range ([0]) and of the tiles on its
  // x2 faces ([1]), in item order, their counts (ilist_n), and the scan that splits them
  int *ilist[2] = {nullptr, nullptr};
  int *ilist_n = nullptr, *iflag = nullptr, *iflag_scan = nullptr;
  void *lscan_tmp = nullptr;
  size_t lscan_bytes = 0;
  void *scan_tmp = nullptr;
  size_t scan_bytes = 0;
  int32_t *itag = nullptr;          // caller's PTAG, PSP, PLEV, PTAGH by original index
  // fields, gas layout with ghosts
  float *qdep = nullptr;            // Q, P1, P2, P3
  float *qt[3] = {nullptr, nullptr, nullptr};   // centre mass m c/V: own, from below, above
  float *ustar = nullptr, *ap = nullptr, *res = nullptr;
  double *dmom = nullptr;           // D of the kick, then R_g
  // actual dust correction: own (with its tile neighbours' shares), from the tile below, from
  // the tile above
  double *dc[3] = {nullptr, nullptr, nullptr};
  double *pres[2] = {nullptr, nullptr};   // push rounding residue per cell (to the gas):
  int pres_rd = 0;                        // the buffer of the last push, read by the next stage
  float *scr_f = nullptr;
  float *cpack = nullptr;           // pc2 clouds: the 12-component gather pack (PackGatherField3D)
  int *xl_list = nullptr;           // refined meshes: the slots of the cross-level clouds and
  int *xl_n = nullptr;              // their count
  double *defect = nullptr;    // [nt][3] the FP64 defect of every tile's impulse deposit
  double *x3p = nullptr;       // [item_cap][2] outflow partials of the key pass (outflow x3 faces)
  int pending = 0;
  double a_prev = 0.0;
  int regs_dead = 0;                // the store holds no registers (after a step's last stage)
  int keep_regs = 0;                // every stage keeps the registers (dk_keep_registers)
  // pc2: R lives in the other store (after a predictor); the positions await the migration
  // that the next key pass applies (after an advance)
  int r_other = 0, unmigrated = 0;
  // launch parameters that leave every result unchanged (the autotuner's choices)
  int key_threads = kKeyThreads;
  // online autotuning: state 0 fixed, 1 tuning, 2 tuned; candidates run one after another
  // for kTuneStages active stages each, after as many warm-up stages
  int tune_state = 0, tune_cand = -1, tune_stage = 0, ncand = 0;
  int cand_xcd[8] = {0}, cand_kt[8] = {0};
  float cand_ms[8] = {0.0f};        // the fastest visit of each candidate
  float visit_ms = 0.0f;
  hipEvent_t tev[2] = {nullptr, nullptr};
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

// the largest power of two <= cap that divides n
int Largest(int n, int cap) {
  int t = 1;
  while (2*t <= cap && n % (2*t) == 0) t *= 2;
  return t;
}

// all numbers a tile kernel needs, by value
struct KArgs {
  Geo g;
  Tiles T;
  const DevBlock *B;
  Store src, dst;
  const int *perm, *start;
  int *key, *rank, *count, *stay;
  dk::ErrWord *err;
  // stage numbers
  double gam0, gam1, beta_dt, atw_dt, a_dt, sh, dt;
  double inv_a_dt, inv_a_prev;      // 1/a_dt, 1/a_prev (the drag rates R = delta/a)
  double ysh, dsh;                  // shear crossing offsets (bvals_part.cpp, dust.cpp)
  int stage1, sbox, strat, shear;
  int rate_dep;                     // pc2: the predictor also deposits the rate m/(t_s V)
  int deal;                         // one-item launches deal their items to the XCDs (FirstItem)
  int cell_sums;                    // the key pass sums corrections and push residues per cell
                                    // (0 with pc2, where both are zero)
  // the stage registers x1, v1 and R of this stage are read by a later stage (0 at the last
  // active stage: the next step's stage 1 sets them before any use, so they are not kept)
  int regs_live;
  double qs, om;
  double c_fx, c_fy, c_fz, c_sh;    // 2 om, -(2 - q) om, -om^2 (stratified), q om; 0 without
  double lxm, lym, lzm;             // mesh lengths
  // pending correction
  int pending;
  double a_prev;
  const float *res;
  double *dc0, *dc1, *dc2;
  // work items: a tile's particles in parts of at most W (item i: tile item_tile[i], part
  // i - item_start[tile]); parts after the first leave their partial sums in overflow slots
  // (slot of part p >= 1 of tile t: item_start[t] - t + p - 1)
  const int *item_start, *item_tile;
  int W;
  float *ov_scr;                    // [slot][4][RN] regions
  double *ov_def;                   // [slot][3] impulse defects
  float *ov_q;                      // [slot][3][TC] centre masses (own, to above, to below)
  double *ov_dc;                    // [slot][9][TC] dust corrections by centre row
  float *ov_pr;                     // [slot][3][TC] push residues
  int *ov_cnt;                      // [slot][TC] stayers counted by the part, then the offsets
  // outflow x3 faces: every work item's (mass, count) of the particles it parked this key pass
  // (bvals_part.cpp:184-186), summed in item order by k_x3sum (null without outflow faces)
  double *x3p;
  // the launch's items: ilist[0 .. *ilist_n) (all items when ilist is null), the first nfirst of
  // them one workgroup each
  const int *ilist, *ilist_n;
  int nfirst;
  // refined meshes (tree.leaf null otherwise): the routing across levels, every particle's cloud
  // level by original index, and the list of the store slots whose clouds the tile kernels
  // leave to the leaf traversal (xl_n: its length)
  dk::SmrTree tree;
  const int32_t *plev;
  int *xl_list, *xl_n;
};

// the cell (active index, unclamped) of position x along one axis, as PMWeights
__device__ inline int CellOf(double x, double xmin, double scale) {
#pragma clang fp contract(off)
  const double cell = (x - xmin)*scale;
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

// the block offset of x along one axis (bvals_part.cpp:164-429 form int((x - lo + l)/l) - 1),
// with the division skipped for a position inside [lo, hi)
__device__ inline int BlockOff(double x, double lo, double hi, double l) {
  if (x >= lo && x < hi) return 0;
  return static_cast<int>((x - lo + l)/l) - 1;
}

// VEL: also the new velocity (the key pass needs the position only).  b0: the particle's block
// (s.gid) before the push, passed in so that a caller whose particles share one block reads it
// once for all of them.  SMR: a refined mesh (the routing across levels, a separate
// instantiation: inlined into the uniform mesh's key pass it would cost registers there).
template <bool VEL = true, bool SMR = false>
__device__ inline bool PushMigrate(const KArgs &A, const DevBlock &b0, PState &s,
                                   bool write_regs) {
#pragma clang fp contract(off)
  const Geo &g = A.g;
  const double xo = s.x[0], yo = s.x[1], zo = s.x[2];
  const float vxo = s.v[0], vyo = s.v[1], vzo = s.v[2];
  if (write_regs) {   // stage 1: (x1, v1) <- (x, v) (dust_push.cpp:69-77)
    for (int d=0; d<3; ++d) {s.x1[d] = s.x[d]; s.v1[d] = s.v[d];}
  }
  // the new velocity in FP64 from the stored FP32 values, rounded once to FP32; the rounding
  // residue goes to the gas (momentum-conserving rounding).  Forces: Coriolis and tidal terms
  // of the shearing box (zero coefficients without it)
  if constexpr (VEL) {
    const double fxd = A.c_fx*static_cast<double>(vyo);
    const double fyd = A.c_fy*static_cast<double>(vxo);
    const double fzd = A.c_fz*zo;
    const double vn[3] = {
        A.gam0*vxo + A.gam1*s.v1[0] + A.beta_dt*fxd + A.atw_dt*s.r[0],
        A.gam0*vyo + A.gam1*s.v1[1] + A.beta_dt*fyd + A.atw_dt*s.r[1],
        A.gam0*vzo + A.gam1*s.v1[2] + A.beta_dt*fzd + A.atw_dt*s.r[2]};
    for (int d=0; d<3; ++d) {
      s.v[d] = static_cast<float>(vn[d]);
      s.vres[d] = vn[d] - static_cast<double>(s.v[d]);
    }
  }
  const double ydot = static_cast<double>(vyo) - A.c_sh*xo;
  s.x[0] = A.gam0*xo + A.gam1*s.x1[0] + A.beta_dt*static_cast<double>(vxo);
  s.x[1] = A.gam0*yo + A.gam1*s.x1[1] + A.beta_dt*ydot;
  s.x[2] = A.gam0*zo + A.gam1*s.x1[2] + A.beta_dt*static_cast<double>(vzo);
  // ShiftShearCrossings
  const DevBlock *B = A.B;
  if (A.shear && A.dsh != 0.0) {
    const DevBlock &bo = b0;
    const int ix = BlockOff(s.x[0], bo.x1min, bo.x1max, bo.x1max - bo.x1min);
    const bool lo1 = (ix < 0) && (bo.face & 1);
    const bool hi1 = (ix > 0) && (bo.face & 2);
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
      if (z >= g.x3max) z = g.x3max - 1.0e-12*A.lzm;
      if (z < g.x3min) z = g.x3min;
      s.x[2] = z; s.v[2] = -s.v[2]; s.vres[2] = -s.vres[2];
      s.x1[2] = 2.0*zf - s.x1[2]; s.v1[2] = -s.v1[2];
    } else {
      // outflow: park the particle at rest in the boundary cell of its block with zero mass
      // (bvals_part.cpp:184-196); its momentum leaves with it
      const DevBlock &bo = b0;
      const double z = lo ? (bo.x3min + 0.5*bo.dx3) : (bo.x3max - 0.5*bo.dx3);
      s.x[2] = z;
      for (int d=0; d<3; ++d) {s.v[d] = 0.0f; s.vres[d] = 0.0; s.v1[d] = 0.0f;}
      s.x1[0] = s.x[0]; s.x1[1] = s.x[1]; s.x1[2] = z;
      s.parked = 1;
    }
  }
  const DevBlock &b = b0;
  const int ix = BlockOff(s.x[0], b.x1min, b.x1max, b.x1max - b.x1min);
  const int iy = BlockOff(s.x[1], b.x2min, b.x2max, b.x2max - b.x2min);
  const int iz = BlockOff(s.x[2], b.x3min, b.x3max, b.x3max - b.x3min);
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
  // the block's faces within half a block of a mesh face (the reference's periodic test)
  const bool lo1 = (ix < 0) && (b.face & 1), hi1 = (ix > 0) && (b.face & 2);
  const bool lo2 = (iy < 0) && (b.face & 4), hi2 = (iy > 0) && (b.face & 8);
  const bool lo3 = (iz < 0) && (b.face & 16), hi3 = (iz > 0) && (b.face & 32);
  const double lxm = A.lxm, lym = A.lym, lzm = A.lzm;
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
  // (a refined mesh: the neighbour of any level, from the block and the position, s.gid
  // holding the block the particle leaves)
  int nb;
  if constexpr (SMR) nb = dk::SmrDest(A.B, A.tree, s.gid, off, s.x);
  else nb = b.nghbr[Dir(ix, iy, iz)];
  if (nb < 0) return false;
  s.gid = nb;
  if (lo1) {s.x[0] += lxm; s.x1[0] += lxm;} else if (hi1) {s.x[0] -= lxm; s.x1[0] -= lxm;}
  if (lo2) {s.x[1] += lym; s.x1[1] += lym;} else if (hi2) {s.x[1] -= lym; s.x1[1] -= lym;}
  if (lo3) {s.x[2] += lzm; s.x1[2] += lzm;} else if (hi3) {s.x[2] -= lzm; s.x1[2] -= lzm;}
  return true;
}

// the deposit cell of a position (y shifted by sh x, the gas's t^n frame) clamped into the block
// (cj is returned unclamped too: a centre one cell beyond the x2 face is handled by the caller)
// (the block's lower corner lo and cells per length sc)
__device__ inline void DepositCellOf(const Geo &g, const double lo[3], const double sc[3],
                                     const double x[3], double sh, int &ci, int &cj, int &ck,
                                     int &cju) {
#pragma clang fp contract(off)
  ci = CellOf(x[0], lo[0], sc[0]);
  cju = CellOf(x[1] + sh*x[0], lo[1], sc[1]);
  ck = CellOf(x[2], lo[2], sc[2]);
  ci = min(max(ci, 0), g.nx1 - 1);
  ck = min(max(ck, 0), g.nx3 - 1);
  cj = min(max(cju, 0), g.nx2 - 1);
}
__device__ inline void DepositCell(const Geo &g, const DevBlock &b, const double x[3],
                                   double sh, int &ci, int &cj, int &ck, int &cju) {
  const double lo[3] = {b.x1min, b.x2min, b.x3min}, sc[3] = {b.sx1, b.sx2, b.sx3};
  DepositCellOf(g, lo, sc, x, sh, ci, cj, ck, cju);
}

// pc2: the half-step position xh = x0 + dt v0/2 (dust_push.cpp:150-154), where the advance gathers
// and deposits (periodic meshes without the shearing box: no frame terms)
__device__ inline void Pc2Mid(const KArgs &A, const double x0[3], const float v0[3],
                              double xh[3]) {
#pragma clang fp contract(off)
  const double h = 0.5*A.dt;
  for (int d=0; d<3; ++d) xh[d] = x0[d] + h*static_cast<double>(v0[d]);
}

// pc2 stage 2: the particle goes to the block of its half-step position for the advance, its
// position moved by a mesh length across a periodic face with it (the advance recomputes xh from
// the moved position, the same arithmetic as here); x1 <- xh, v1 <- v0 (dust_push.cpp:240-242).
// The block of the advanced position is found by the next stage's key pass (or an export).
// Returns false on a routing error (the particle keeps its block); xh: the position keyed.
// On a refined mesh the particle keeps its block, as the reference's does until the migration
// after the advance (the routing across levels depends on the block left, so routing through
// the half-step block could end in another leaf); a half-step position outside the block is
// keyed into the block's edge cells and its cloud goes to the leaf walk (CrossLevel).
__device__ inline bool Pc2Route(const KArgs &A, const DevBlock &b, PState &s, double xh[3]) {
  Pc2Mid(A, s.x, s.v, xh);
  if (A.tree.leaf != nullptr) {
    for (int d=0; d<3; ++d) {s.x1[d] = xh[d]; s.v1[d] = s.v[d];}
    return true;
  }
  const int ix = BlockOff(xh[0], b.x1min, b.x1max, b.x1max - b.x1min);
  const int iy = BlockOff(xh[1], b.x2min, b.x2max, b.x2max - b.x2min);
  const int iz = BlockOff(xh[2], b.x3min, b.x3max, b.x3max - b.x3min);
  bool ok = true;
  if (ix != 0 || iy != 0 || iz != 0) {
    const int nb = (abs(ix) > 1 || abs(iy) > 1 || abs(iz) > 1) ? -1 : b.nghbr[Dir(ix, iy, iz)];
    if (nb < 0) {
      ok = false;
    } else {
      s.gid = nb;
      if ((ix < 0) && (b.face & 1)) s.x[0] += A.lxm; else if ((ix > 0) && (b.face & 2)) s.x[0] -= A.lxm;
      if ((iy < 0) && (b.face & 4)) s.x[1] += A.lym; else if ((iy > 0) && (b.face & 8)) s.x[1] -= A.lym;
      if ((iz < 0) && (b.face & 16)) s.x[2] += A.lzm; else if ((iz > 0) && (b.face & 32)) s.x[2] -= A.lzm;
      Pc2Mid(A, s.x, s.v, xh);
    }
  }
  for (int d=0; d<3; ++d) {s.x1[d] = xh[d]; s.v1[d] = s.v[d];}
  return ok;
}

// the centre cell of the unshifted position (dust_pm.cpp:649-659); false when it rounds onto
// a ghost cell, where the reference's centre-mass and correction touch no active cell
__device__ inline bool CenterCell(const Geo &g, const DevBlock &b, const double x[3], int &ci,
                                  int &cj, int &ck) {
  ci = CellOf(x[0], b.x1min, b.sx1);
  cj = CellOf(x[1], b.x2min, b.sx2);
  ck = CellOf(x[2], b.x3min, b.sx3);
  return ci >= 0 && ci < g.nx1 && cj >= 0 && cj < g.nx2 && ck >= 0 && ck < g.nx3;
}

__device__ inline void LoadP(const Store &S, long long p, PState &s, bool regs, bool rates,
                             bool vregs = true) {
  for (int d=0; d<3; ++d) {
    s.x[d] = S.X(d, p);
    s.v[d] = S.V(d, p);
    s.x1[d] = regs ? S.X1(d, p) : 0.0;
    s.v1[d] = (regs && vregs) ? S.V1(d, p) : 0.0f;
    s.r[d] = rates ? S.R(d, p) : 0.0;
    s.vres[d] = 0.0;
  }
  s.gid = -1;   // set by the caller (the tile's block)
  s.parked = 0;
}

//----------------------------------------------------------------------------------------
// Tile kernels.  A workgroup owns one tile; the tile's particles are one contiguous range of
// the store (the keys are tile-major).  A kernel walks the range in chunks of CH particles: a
// particle phase (each thread takes kPPT particles, coalesced loads and stores) computes each
// particle's work and leaves one record per particle in LDS; a cell phase (one thread per
// (cell, stencil slab), the slab being the x3 offset) sums the records of its cell, in store
// order, into its 9 stencil offsets; the sums are then added into the LDS copy of the tile
// region in phases, so that no two threads of a phase touch the same word (DepositSlabs).
// Thread (cell l, slab) forms cell l's sums.  At the end the region is written to the tile's
// scratch.

constexpr int kPPT = 2;             // particles per thread in the particle phase
// a work item (part) is at most PartChunks chunks of a tile's particles.  Without the shearing
// box the items of multi-part tiles run one workgroup each in their own launch (the crowded-cell
// split): small parts (2 chunks) spread a clump over many workgroups.  In the shearing box a
// tile's extra parts run in the looping launch, whose loop-carried state costs registers: large
// parts (8 chunks) keep the dense midplane tiles of a stratified box one item (C3: 64.1 -> 55.9
// ms per step; the clumped cases without the shearing box lose with 8: C5 12.0 -> 18.6 ms).  The
// part size is part of the summation order (a later part's partial sums are added in part
// order), so it is fixed by the mesh, not tuned.
constexpr int kPartChunksMax = 8;
static_assert(kPartChunksMax*kPPT*kCellThreads <= kPartRank,
              "a part's stayer ranks must fit below kPartRank");
__host__ inline int PartChunks(bool shear) { return shear ? 8 : 2; }

struct TileInfo {
  int m, tx, ty, tz, ci0, cj0, ck0;
  long long key0;
};

// workgroups are dealt to the 8 XCDs in turn: consecutive tiles go to one XCD so that
// neighbouring tiles (whose regions overlap) share its L2
__device__ inline int TileIndex(const Tiles &T) {
  const int b = blockIdx.x;
  if (!T.xcd || T.nt % 8 != 0) return b;
  return (b % 8)*(T.nt/8) + b/8;
}

// the same for a launch of n workgroups
__device__ inline int DealIndex(const Tiles &T, int n) {
  const int b = blockIdx.x;
  if (!T.xcd || n % 8 != 0) return b;
  return (b % 8)*(n/8) + b/8;
}

// the list index of a one-item workgroup: dealt to the XCDs in runs (DealIndex), except in a
// launch whose grid only bounds its list (A.deal = 0: the items of multi-part tiles fill a
// prefix of the grid, which runs of the grid would give to one XCD)
__device__ inline int FirstItem(const Tiles &T, const KArgs &A) {
  return A.deal ? DealIndex(T, A.nfirst) : static_cast<int>(blockIdx.x);
}

// a tile on an x2 face of its block (in the shearing box the only tiles whose records can sit
// one row off in x2: the deposit centre beyond the face in the shifted frame)
__host__ __device__ inline bool EdgeTile(const Tiles &T, int tile) {
  const int ty = ((tile % T.ntb)/T.ntx) % T.nty;
  return ty == 0 || ty == T.nty - 1;
}

__device__ inline TileInfo TileOf(const Tiles &T, int tile) {
  TileInfo t;
  t.m = tile/T.ntb;
  int r = tile - t.m*T.ntb;
  t.tx = r % T.ntx; r /= T.ntx;
  t.ty = r % T.nty;
  t.tz = r/T.nty;
  t.ci0 = t.tx*T.TX; t.cj0 = t.ty*T.TY; t.ck0 = t.tz*T.TZ;
  t.key0 = (long long)tile*T.TC;
  return t;
}

__host__ __device__ inline long long TileKey(const Tiles &T, int m, int ci, int cj, int ck) {
  const int tx = ci >> T.LX, ty = cj >> T.LY, tz = ck/T.TZ;
  const int lx = ci & (T.TX - 1), ly = cj & (T.TY - 1), lz = ck - tz*T.TZ;
  return ((long long)m*T.ntb + (tz*T.nty + ty)*T.ntx + tx)*T.TC +
         (((lz << T.LY) + ly) << T.LX) + lx;
}

// i/d for 0 <= i < 4096 and a divisor d < 64, or 0 <= i < 4 d with any d < 2^20, from the
// float reciprocal inv = 1/d: the error of (i + 1/2) inv (below 2^-22 (i + 1/2)/d) is below the
// distance 1/(2 d) of (i + 1/2)/d from an integer
__device__ inline int FDiv(int i, float inv) {
  return static_cast<int>((static_cast<float>(i) + 0.5f)*inv);
}

// the sum of v over the na active lanes of a wavefront (the first na), in a fixed order: a
// butterfly for a full wavefront, else lane order; the result is lane 0's
__device__ inline double WaveSum(double v, int na) {
  if (na >= 64) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
  }
  double s = 0.0;
  for (int j = 0; j < na; ++j) s += __shfl(v, j, 64);
  return s;
}

// the tile-local cell of store slot q: the last c with st[c] <= q (st[0] <= q < st[TC])
__device__ inline int SlotCell(const int *st, int TC, int q) {
  int lo = 0, hi = TC;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (st[mid] <= q) lo = mid; else hi = mid;
  }
  return lo;
}

// the three 1-D weights of PMWeights for a given del (S: 0 NGP, 1 CIC, 2 TSC)
template <int S, typename T>
__device__ inline void W3(T d, T w[3]) {
  if constexpr (S == 2) {
    const T h0 = T(0.5) - d, h2 = T(0.5) + d;
    w[0] = T(0.5)*h0*h0;
    w[1] = T(0.75) - d*d;
    w[2] = T(0.5)*h2*h2;
  } else if constexpr (S == 1) {
    w[0] = fmax(T(0), -d); w[1] = T(1) - fabs(d); w[2] = fmax(T(0), d);
  } else {
    w[0] = T(0); w[1] = T(1); w[2] = T(0);
  }
}

template <int S, typename T>
__device__ inline T W1(int a, T d) {
  if constexpr (S == 2) {
    // (only the weight asked for: 1/2 (1/2 - d)^2 as 1/2 (d - 1/2)^2, the same bits)
    const T h = d + ((a == 0) ? T(-0.5) : T(0.5));
    return (a == 1) ? T(0.75) - d*d : T(0.5)*h*h;
  } else {
    T w[3];
    W3<S, T>(d, w);
    return (a == 0) ? w[0] : ((a == 1) ? w[1] : w[2]);
  }
}

// One record's 9 terms of one slab, w_z w_y(b) w_x(a) val[n], into acc[b][a][n].  With SY (the
// shearing box) a record centred one cell off in x2 (sj = -1 or +1: the deposit centre beyond
// the block's x2 face in the shifted frame) moves its x2 weights by sj, and the term that leaves
// the thread's three rows goes to spl[a][n], deposited at x2 offset 2 sj.  Branch-free: the
// lanes of a wave may hold records of both kinds.
template <typename T, int NC, bool SY>
__device__ inline void AddSlab(T (&acc)[3][3][NC], T (&spl)[3][NC], int &spdir, int sj, T wz,
                               const T wy[3], const T wx[3], const T val[NC]) {
  T e0 = wy[0], e1 = wy[1], e2 = wy[2];
  if constexpr (SY) {
    const T z = T(0);
    const T sp = (sj < 0) ? wy[0] : ((sj > 0) ? wy[2] : z);
    e0 = (sj == 0) ? wy[0] : ((sj < 0) ? wy[1] : z);
    e1 = (sj == 0) ? wy[1] : ((sj < 0) ? wy[2] : wy[0]);
    e2 = (sj == 0) ? wy[2] : ((sj < 0) ? z : wy[1]);
    if (sj != 0) spdir = sj;
    const T wsp = wz*sp;
#pragma unroll
    for (int a=0; a<3; ++a) {
      const T w = wsp*wx[a];
#pragma unroll
      for (int n=0; n<NC; ++n) spl[a][n] += w*val[n];
    }
  }
  const T wzy[3] = {wz*e0, wz*e1, wz*e2};
#pragma unroll
  for (int b=0; b<3; ++b)
#pragma unroll
    for (int a=0; a<3; ++a) {
      const T w = wzy[b]*wx[a];
#pragma unroll
      for (int n=0; n<NC; ++n) acc[b][a][n] += w*val[n];
    }
}

// The deposit of one chunk's sums into the LDS tile region.  Each slab first adds into its own
// copy of the region (cp: 3 copies, placed in the record area, which the chunk no longer needs):
// 9 phases, one (x1, x2) offset per phase, all slabs at once, no two threads of a slab touching
// one word in a phase; the spill terms of cells centred one row off in x2 take 3 more phases.
// The copies are then added into the region in slab order.  Every word's sum has a fixed order.
// A thread writes only the x3 row (its layer + its slab) of its slab's copy, so when every
// wavefront holds whole x1-x2 layers of one slab (Tl.wrows) no two wavefronts touch one word in
// the 9 phases: the phases then need no workgroup barrier, only program order within the
// wavefront (its LDS operations complete in issue order; the empty asm keeps the compiler from
// moving one phase's LDS accesses across the next).  Starts and ends with a barrier.
// dep: the thread holds sums to deposit (in the grouped map of ChunkSums only the first lane of
// a group does); layers: the thread map is the one where every wavefront holds whole x1-x2
// layers of one slab (the 9 phases then need no barrier when Tl.wrows).
template <typename T, int NC, bool SY>
__device__ inline void DepositSlabs(const Tiles &Tl, T *reg, T *cp, int l, int slab,
                                    const T (&acc)[3][3][NC], const T (&spl)[3][NC], int spdir,
                                    bool dep = true, bool layers = true) {
  const int n3 = 3*NC*Tl.RN;
  __syncthreads();
  for (int i = threadIdx.x; i < n3; i += blockDim.x) cp[i] = T(0);
  __syncthreads();
  const int lx = l & (Tl.TX - 1), ly = (l >> Tl.LX) & (Tl.TY - 1), lz = l >> (Tl.LX + Tl.LY);
  const int rbase = ((lz + slab)*Tl.RY + (ly + Tl.SY))*Tl.RX + lx;
  T *mine = cp + slab*NC*Tl.RN;
  const bool wrows = Tl.wrows != 0 && layers;
#pragma unroll
  for (int pb=0; pb<3; ++pb)
#pragma unroll
    for (int pa=0; pa<3; ++pa) {
      const int rc = rbase + pb*Tl.RX + pa;
      if (dep) {
#pragma unroll
        for (int n=0; n<NC; ++n) mine[n*Tl.RN + rc] += acc[pb][pa][n];
      }
      if (wrows) asm volatile("" ::: "memory");
      else __syncthreads();
    }
  if (wrows) __syncthreads();
  if constexpr (SY) {
    if (__syncthreads_or(dep && spdir != 0)) {
#pragma unroll
      for (int pa=0; pa<3; ++pa) {
        if (dep && spdir != 0) {
          const int rc = rbase + (1 + 2*spdir)*Tl.RX + pa;
#pragma unroll
          for (int n=0; n<NC; ++n) mine[n*Tl.RN + rc] += spl[pa][n];
        }
        __syncthreads();
      }
    }
  }
  const int nr = NC*Tl.RN;
  for (int i = threadIdx.x; i < nr; i += blockDim.x) {
    T v = reg[i];
    v += cp[i];
    v += cp[nr + i];
    v += cp[2*nr + i];
    reg[i] = v;
  }
  __syncthreads();
}

// The cell phase of one chunk [c0, c1) outside the crowded-cell split: every (cell, slab)'s sums
// of the chunk's records of that cell.  The plain map gives thread (cell l, slab) all of l's
// records in store order.  A chunk whose records span at most half of the tile's cells (a dense
// tile: in a stratified box's midplane a chunk of 768 records covers ~27 cells) would leave most
// threads idle and a few with long loops; then a group of G = 2, 4 or 8 adjacent lanes takes one
// (cell, slab), lane k the records whose slot within the cell is k mod G, in store order, and a
// butterfly over the group's lanes adds their sums (a fixed order) into its first lane, which
// deposits them.  G depends only on the chunk's cell span.  add(r, slab, acc, spl, spdir, e)
// adds record r (chunk index); e: three extra sums per thread, added over the group too.
// Returns the thread's cell and slab and whether it deposits; layers: the plain map.
struct CellMap {
  int l, slab;
  bool dep, layers;
};
constexpr int kGroupMax = 8;

// The cells of the chunk's first and last slots into span[0], span[1] (threads of the first TC;
// the slots' cells are the non-empty cells whose range holds them); read after a barrier.
constexpr int kSpanSlot = 10;   // (two ints of the plan area that CellSums leaves unused)
__device__ inline void MarkChunkSpan(const Tiles &Tl, const int *st, int c0, int c1, int *span) {
  const int l = threadIdx.x;
  if (l < Tl.TC) {
    const int a = st[l], b = st[l + 1];
    if (a <= c0 && c0 < b) span[0] = l;
    if (a < c1 && c1 <= b) span[1] = l;
  }
}

template <int NC, bool SY, class Add>
__device__ inline CellMap ChunkSums(const Tiles &Tl, const int *st, int c0, int c1,
                                    const int *ends, const Add &add, float (&acc)[3][3][NC],
                                    float (&spl)[3][NC], int &spdir, float (&e)[3]) {
  const int tid = threadIdx.x;
  const int la = ends[0], lb = ends[1];
  const int span = lb - la + 1;
  int G = 1;
  while (G < kGroupMax && 2*G*span <= Tl.TC) G *= 2;
  CellMap m;
  int sub = 0;
  bool act = true;
  if (G == 1) {
    m.slab = tid/Tl.TC;
    m.l = tid - m.slab*Tl.TC;
  } else {
    const int grp = tid/G;
    sub = tid - grp*G;
    m.slab = grp/span;
    m.l = la + (grp - m.slab*span);
    act = m.slab < 3;
    if (!act) {m.slab = 0; m.l = la;}
  }
  m.dep = act && sub == 0;
  m.layers = (G == 1);
  const int a = st[m.l];
  int s0 = max(a, c0);
  const int s1 = act ? min(st[m.l + 1], c1) : s0;
  s0 += (sub - (s0 - a)) & (G - 1);
  for (int qq = s0; qq < s1; qq += G) add(qq - c0, m.slab, acc, spl, spdir, e);
  // (G is the same in the whole workgroup: every lane of a wavefront takes part)
  for (int k = 1; k < G; k <<= 1) {
#pragma unroll
    for (int bb=0; bb<3; ++bb)
#pragma unroll
      for (int aa=0; aa<3; ++aa)
#pragma unroll
        for (int n=0; n<NC; ++n) acc[bb][aa][n] += __shfl_xor(acc[bb][aa][n], k, 64);
    if constexpr (SY) {
#pragma unroll
      for (int aa=0; aa<3; ++aa)
#pragma unroll
        for (int n=0; n<NC; ++n) spl[aa][n] += __shfl_xor(spl[aa][n], k, 64);
      const int o = __shfl_xor(spdir, k, 64);
      if (spdir == 0) spdir = o;
    }
#pragma unroll
    for (int n=0; n<3; ++n) e[n] += __shfl_xor(e[n], k, 64);
  }
  return m;
}

// Crowded cells.  The cell phase gives each cell's records to one thread per slab, in store
// order; a clump cell holding hundreds of a chunk's records would serialise its workgroup.  When
// a cell of the chunk holds more than kSplitMin records, its records are cut into runs of T (the
// smallest power of two >= 16 for which the runs fit the idle threads and the record area), the
// first run stays with the cell's own thread and every further run goes to the thread of an
// empty cell (a helper, in cell order); the helpers' partial sums are then added to the owner's
// in run order.  Every sum has a fixed order (the split depends only on the counts).
constexpr int kSplitMin = 16;

// the plan area of a tile kernel, in ints after st[TC + 1]: 8 counters, 2 wave totals, and
// hcell, hsub, lhelp, hbase [TC] each
__host__ __device__ inline int PlanInts(int TC) { return 12 + 4*TC; }

// inclusive prefix sum over the 64 lanes of a wavefront
__device__ inline int WaveInclScan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// Sums the chunk's records of this thread's cell l (chunk slots [s0, s1)) into acc (and ex),
// splitting crowded cells over helper threads as above.  add(r, acc, ex) adds record r (chunk
// index) to the accumulators.  area: the record area (free once every record is summed), of
// area_words floats; plan: PlanInts(TC) ints.  All threads of the workgroup call it.
template <int NC, int NE, class Add>
__device__ inline void CellSums(const Tiles &Tl, const int *st, int c0, int c1, int l, int slab,
                                float *area, int area_words, int *plan, const Add &add,
                                float (&acc)[3][3][NC], float (&ex)[NE]) {
  const int s0 = max(st[l], c0), s1 = min(st[l + 1], c1);
  const int n = max(s1 - s0, 0);
  if (!__syncthreads_or(n > kSplitMin)) {
    for (int qq = s0; qq < s1; ++qq) add(qq - c0, acc, ex);
    return;
  }
  const int TC = Tl.TC, tid = threadIdx.x;
  int *cnt = plan;                 // [0, 7): helpers for T = 16 << i; [7]: empty cells
  int *wtot = plan + 8;            // [2] wave totals of the scans
  int *hcell = plan + 12, *hsub = hcell + TC, *lhelp = hsub + TC, *hbase = lhelp + TC;
  if (tid < 8) cnt[tid] = 0;
  __syncthreads();
  if (slab == 0) {
    for (int i = 0; i < 7; ++i) {
      const int T = kSplitMin << i;
      const int h = (n > 0) ? (n + T - 1)/T - 1 : 0;
      if (h) atomicAdd(&cnt[i], h);
    }
    if (n == 0) atomicAdd(&cnt[7], 1);
  }
  __syncthreads();
  const int slot_words = 9*NC + NE;
  const int hmax = min(cnt[7], area_words/(3*slot_words));
  int T = kSplitMin << 6, H = cnt[6];
  for (int i = 5; i >= 0; --i)
    if (cnt[i] <= hmax) {T = kSplitMin << i; H = cnt[i];}
  if (H == 0 || H > hmax) {   // (T = 1024 exceeds any chunk: no helpers)
    for (int qq = s0; qq < s1; ++qq) add(qq - c0, acc, ex);
    return;
  }
  // helpers per cell and their first slot, empty cells' ranks (exclusive scans over the cells,
  // run by the threads of the first ceil(TC/64) wavefronts)
  const int nh = (n > 0) ? (n + T - 1)/T - 1 : 0;
  int hb = 0, er = 0;
  if (tid < ((TC + 63) & ~63)) {
    const bool mine = tid < TC;
    const int a = WaveInclScan(mine ? nh : 0), e = WaveInclScan((mine && n == 0) ? 1 : 0);
    if ((tid & 63) == 63) wtot[tid >> 6] = a | (e << 16);
    hb = a - (mine ? nh : 0);
    er = e - ((mine && n == 0) ? 1 : 0);
  }
  __syncthreads();
  if (tid >= 64 && tid < TC) {hb += wtot[0] & 0xffff; er += wtot[0] >> 16;}
  if (slab == 0) {
    hbase[l] = hb;
    for (int k = 0; k < nh; ++k) {hcell[hb + k] = l; hsub[hb + k] = k + 1;}
    lhelp[l] = (n == 0 && er < H) ? er : -1;
  }
  __syncthreads();
  // this thread's run: the first of its own cell, or (a helper) run hsub of cell hcell
  const int h = lhelp[l];
  float hex[NE > 0 ? NE : 1];
  for (int e = 0; e < NE; ++e) hex[e] = 0.0f;
  if (h >= 0) {
    const int L = hcell[h], k = hsub[h];
    const int a0 = max(st[L], c0) + k*T, a1 = min(a0 + T, min(st[L + 1], c1));
    for (int qq = a0; qq < a1; ++qq) add(qq - c0, acc, hex);
  } else {
    for (int qq = s0; qq < min(s0 + T, s1); ++qq) add(qq - c0, acc, ex);
  }
  __syncthreads();   // every record summed: the record area takes the helpers' partials
  if (h >= 0) {
    float *o = area + (slab*H + h)*slot_words;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c) {o[(b*3 + a)*NC + c] = acc[b][a][c]; acc[b][a][c] = 0.0f;}
    for (int e = 0; e < NE; ++e) o[9*NC + e] = hex[e];
  }
  __syncthreads();
  if (nh > 0) {
    const int hb0 = hbase[l];
    for (int k = 0; k < nh; ++k) {
      const float *o = area + (slab*H + hb0 + k)*slot_words;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[b][a][c] += o[(b*3 + a)*NC + c];
        // (one row's loads at a time: the compiler would otherwise hoist all 9 NC of them)
        asm volatile("" ::: "memory");
      }
      for (int e = 0; e < NE; ++e) ex[e] += o[9*NC + e];
    }
  }
}

// the tile region of a padded cell field, fetched into registers (up to kRegionPer values per
// thread) so that its latency overlaps the LDS set-up, then stored into LDS
constexpr int kRegionPer = 6;

template <typename T>
struct RegionFetch {
  T val[kRegionPer];
  __device__ void Load(const Geo &g, const Tiles &Tl, const TileInfo &t, const T *f, int nvar) {
    const int n = nvar*Tl.RN;
#pragma unroll
    for (int k=0; k<kRegionPer; ++k) {
      const int i = threadIdx.x + k*blockDim.x;
      val[k] = T(0);
      if (i < n) {
        const int v = FDiv(i, Tl.iRN);
        const int rc = i - v*Tl.RN;
        const int rr = FDiv(rc, Tl.iRX);
        const int rx = rc - rr*Tl.RX;
        const int rz = FDiv(rr, Tl.iRY);
        const int ry = rr - rz*Tl.RY;
        val[k] = f[Idx(g, t.m, nvar, v, g.ks + t.ck0 - 1 + rz, g.js + t.cj0 - 1 - Tl.SY + ry,
                       g.is + t.ci0 - 1 + rx)];
      }
    }
  }
  __device__ void Store(const Tiles &Tl, int nvar, T *reg) const {
    const int n = nvar*Tl.RN;
#pragma unroll
    for (int k=0; k<kRegionPer; ++k) {
      const int i = threadIdx.x + k*blockDim.x;
      if (i < n) reg[i] = val[k];
    }
  }
};

// the whole tile region of an nvar-component padded cell field straight into LDS (for fields
// larger than RegionFetch holds)
template <typename T>
__device__ inline void RegionLoad(const Geo &g, const Tiles &Tl, const TileInfo &t, const T *f,
                                  int nvar, T *reg) {
  const int n = nvar*Tl.RN;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int v = FDiv(i, Tl.iRN);
    const int rc = i - v*Tl.RN;
    const int rr = FDiv(rc, Tl.iRX);
    const int rx = rc - rr*Tl.RX;
    const int rz = FDiv(rr, Tl.iRY);
    const int ry = rr - rz*Tl.RY;
    reg[i] = f[Idx(g, t.m, nvar, v, g.ks + t.ck0 - 1 + rz, g.js + t.cj0 - 1 - Tl.SY + ry,
                   g.is + t.ci0 - 1 + rx)];
  }
}

// Refinable clouds at the mesh level (half-width h = one cell): the first moments, in cells, of
// the hat over the three cells of the support about the particle, whose offset from the centre
// of its cell is del (CloudShape1D with HatMoment, src/dust/dust.hpp:169-182, 240-298: the
// faces sit at s = a - 3/2 - del in units of h); the weights are the TSC weights (the hat's CDF
// differences over the same cells)
__device__ inline float HatMomentF(float s) {
  if (s <= -1.0f || s >= 1.0f) return 1.0f/6.0f;
  const float s2 = s*s;
  return (s <= 0.0f) ? 0.5f*s2 + s2*s/3.0f : 0.5f*s2 - s2*s/3.0f;
}
__device__ inline void CloudMoments(float del, float hm[3]) {
  float f0 = HatMomentF(-1.5f - del);
#pragma unroll
  for (int a=0; a<3; ++a) {
    const float f1 = HatMomentF(-0.5f - del + static_cast<float>(a));
    hm[a] = f1 - f0;
    f0 = f1;
  }
}

// The linear gather of a cloud (LeafApi<3D>::GatherLinear, src/dust/dust_leaf_api.hpp:202-220)
// from the tile region of the 12-component pack (velocity, then its central differences along
// x1, x2, x3; PackGatherField3D): u += w q + cx dq/dx1 + cy dq/dx2 + cz dq/dx3 per cell, with
// c = (first moment + (x - x_cell) w)/(2 dx) in cells
__device__ inline void CloudGather(const Tiles &Tl, const float *pk, int rb, const float o[3],
                                   const float wx[3], const float wy[3], const float wz[3],
                                   float u[3]) {
  float hx[3], hy[3], hz[3];
  CloudMoments(o[0], hx);
  CloudMoments(o[1], hy);
  CloudMoments(o[2], hz);
#pragma unroll
  for (int c=0; c<3; ++c)
#pragma unroll
    for (int bb=0; bb<3; ++bb) {
      const int rr = rb + (c*Tl.RY + bb)*Tl.RX;
#pragma unroll
      for (int a=0; a<3; ++a) {
        const float w = wz[c]*wy[bb]*wx[a];
        const float cx = 0.5f*(wy[bb]*wz[c]*hx[a] + (o[0] + static_cast<float>(1 - a))*w);
        const float cy = 0.5f*(wx[a]*wz[c]*hy[bb] + (o[1] + static_cast<float>(1 - bb))*w);
        const float cz = 0.5f*(wx[a]*wy[bb]*hz[c] + (o[2] + static_cast<float>(1 - c))*w);
        const int q = rr + a;
#pragma unroll
        for (int d=0; d<3; ++d)
          u[d] += w*pk[d*Tl.RN + q] + cx*pk[(3 + d)*Tl.RN + q] + cy*pk[(6 + d)*Tl.RN + q] +
                  cz*pk[(9 + d)*Tl.RN + q];
      }
    }
}

// The tile scratch holds the regions row by row: one row of the scratch is the x1 rows (RX
// values each) of the regions of all tiles of a block that share (tz, ty, component, rz, ry), in
// tx order, so that the combine reads consecutive cells of a padded row from consecutive words.
__host__ __device__ inline long long ScrRow(const Tiles &T, int m, int tz, int ty, int nv, int v,
                                            int rz, int ry) {
  return ((((((long long)m*T.ntz + tz)*T.nty + ty)*nv + v)*T.RZ + rz)*T.RY + ry)*
         ((long long)T.ntx*T.RX);
}

// (sum: also the FP64 sum of the words this thread writes, per component)
template <typename T, int NV>
__device__ inline void FlushRegion(const Tiles &Tl, const TileInfo &t, const T *reg, T *scr,
                                   double *sum = nullptr) {
  __syncthreads();
#pragma unroll 1
  for (int v=0; v<NV; ++v) {
    double s = 0.0;
    for (int rc = threadIdx.x; rc < Tl.RN; rc += blockDim.x) {
      const int rr = FDiv(rc, Tl.iRX);
      const int rx = rc - rr*Tl.RX;
      const int rz = FDiv(rr, Tl.iRY);
      const int ry = rr - rz*Tl.RY;
      const T val = reg[v*Tl.RN + rc];
      scr[ScrRow(Tl, t.m, t.tz, t.ty, NV, v, rz, ry) + t.tx*Tl.RX + rx] = val;
      s += static_cast<double>(val);
    }
    if (sum != nullptr) sum[v] = s;
  }
}

// the record area of a tile kernel, in words: nrec words per particle of a chunk of ch, or the
// three slab copies of an nc-component region that the deposit places there, if larger
__host__ __device__ inline int RecArea(const Tiles &Tl, int ch, int nrec, int nc) {
  return (nrec*ch > 3*nc*Tl.RN) ? nrec*ch : 3*nc*Tl.RN;
}

// The key pass's record of a particle: the correction's m delta/V by component, the push
// residue m vres/V, the centre row + 2 (0: none); 48 bytes, read as three 16-byte words.
struct KeyRec {
  double dv[3];
  float pr[3];
  int code, pad;
};
static_assert(sizeof(KeyRec) == 48, "KeyRec is three 16-byte words");

// Every cell's record sums are split by the parity of the record's slot within the cell (two
// half sums, added at the end of the work item): two threads per cell sum a chunk's records,
// and the order of every sum is fixed whatever the chunk size.
constexpr int kKeyHalves = 2;

// the key pass's LDS before its per-cell sums: records, cell offsets, stayer counts
__host__ __device__ inline size_t LdsKeyBase(const Tiles &T, int threads) {
  return (size_t)threads*sizeof(KeyRec) + sizeof(int)*(2*T.TC + 1);
}

// the key pass's per-wavefront stayer counts of a chunk: [threads/64][TC] ints, each the chunk's
// stamp << 16 | the count (a stale entry has another stamp)
__host__ __device__ inline size_t LdsKeyWaves(const Tiles &T, int threads) {
  return sizeof(int)*(size_t)(threads/64)*T.TC;
}

// the overflow slot of part >= 1 of a tile
__device__ inline long long OvSlot(const int *item_start, int tile, int part) {
  return (long long)item_start[tile] - tile + part - 1;
}

// the region of a part after the first into its overflow slot, words in region order (sum: also
// the FP64 sum of the words this thread writes)
template <int NV>
__device__ inline void FlushOverflow(const Tiles &Tl, const float *reg, float *ov,
                                     double *sum = nullptr) {
  __syncthreads();
#pragma unroll 1
  for (int v=0; v<NV; ++v) {
    double s = 0.0;
    for (int rc = threadIdx.x; rc < Tl.RN; rc += blockDim.x) {
      const float val = reg[v*Tl.RN + rc];
      ov[v*Tl.RN + rc] = val;
      s += static_cast<double>(val);
    }
    if (sum != nullptr) sum[v] = s;
  }
}

// active cell indices of tile-local cell l
__device__ inline void CellIdx(const Tiles &Tl, const TileInfo &t, int l, int &ci, int &cj,
                               int &ck) {
  ci = t.ci0 + (l & (Tl.TX - 1));
  cj = t.cj0 + ((l >> Tl.LX) & (Tl.TY - 1));
  ck = t.ck0 + (l >> (Tl.LX + Tl.LY));
}

// Where a particle deposits (PMWeights of src/dust/dust.hpp:1088-1111 on each axis, x2 shifted
// by sh x1, dust.hpp:1699-1702): the deposit cell clamped into the block, the x2 shift of the
// centre (with SY: -1 or +1 when the shifted centre is one cell beyond the block's x2 face), del
// about that centre on each axis, and the cell of the unshifted position (the centre of the
// consistent split, dust_pm.cpp:649-659), which is an active cell when cen is true.
struct Loc {
  int ci, cj, ck, sj;
  int jc;                   // x2 cell of the unshifted position
  bool cen;                 // the unshifted position's cell is an active cell
  double dx, dy, dz;
};

template <bool SY>
__device__ inline Loc Locate(const Geo &g, const DevBlock &b, const double x[3], double sh) {
#pragma clang fp contract(off)
  Loc L;
  const double cx = (x[0] - b.x1min)*b.sx1;
  const double cy = ((x[1] + sh*x[0]) - b.x2min)*b.sx2;
  const double cz = (x[2] - b.x3min)*b.sx3;
  const double cu = (x[1] - b.x2min)*b.sx2;
  const int iu = static_cast<int>(cx + 1.0) - 1;
  const int ju = static_cast<int>(cy + 1.0) - 1;
  const int ku = static_cast<int>(cz + 1.0) - 1;
  L.jc = static_cast<int>(cu + 1.0) - 1;
  L.ci = min(max(iu, 0), g.nx1 - 1);
  L.cj = min(max(ju, 0), g.nx2 - 1);
  L.ck = min(max(ku, 0), g.nx3 - 1);
  L.sj = SY ? min(max(ju - L.cj, -1), 1) : 0;
  L.cen = (iu == L.ci) && (ku == L.ck) && (L.jc >= 0) && (L.jc < g.nx2);
  L.dx = cx - (static_cast<double>(L.ci) + 0.5);
  L.dy = cy - (static_cast<double>(L.cj + L.sj) + 0.5);
  L.dz = cz - (static_cast<double>(L.ck) + 0.5);
  return L;
}

// the same, with the cell clamped into the tile (LDS indices stay inside the region whatever the
// position; a particle outside its tile is an error the push pass has recorded)
template <bool SY>
__device__ inline Loc LocateInTile(const Geo &g, const Tiles &Tl, const TileInfo &t,
                                   const DevBlock &b, const double x[3], double sh) {
  Loc L = Locate<SY>(g, b, x, sh);
  L.ci = min(max(L.ci, t.ci0), t.ci0 + Tl.TX - 1);
  L.cj = min(max(L.cj, t.cj0), t.cj0 + Tl.TY - 1);
  L.ck = min(max(L.ck, t.ck0), t.ck0 + Tl.TZ - 1);
  return L;
}

// Refined meshes: a cloud that the tile kernels leave to the leaf traversal.  The tile kernels
// deposit and gather a cloud as the TSC stencil of its block's cells over the tile region
// (ghost layer included, folded to and filled from same-level neighbours), which is the cloud
// of half-width one cell of its level when that level is its block's and the cells it reaches
// beyond the block belong to one leaf of the same level.  Otherwise (its level is not its
// block's, or it lies in the first or last cell layer of an axis toward a region beside the
// block that xmask marks as not one same-level leaf, or it lies outside its block) it goes to
// the leaf traversal.  x: the cloud's centre, (ci, cj, ck): its active cell.
__device__ inline bool CrossLevel(const Geo &g, const DevBlock &b, int lev, const double x[3],
                                  int ci, int cj, int ck) {
  if (lev != b.level) return true;
  // (a position outside its block, which the routing across levels can leave near a corner)
  if (!(x[0] >= b.x1min && x[0] < b.x1max && x[1] >= b.x2min && x[1] < b.x2max &&
        x[2] >= b.x3min && x[2] < b.x3max))
    return true;
  if (b.xmask == 0) return false;
  // the offsets the support can reach per axis: bit (o + 1) for o = -1, 0, 1
  const int a1 = 2 | ((ci == 0) ? 1 : 0) | ((ci == g.nx1 - 1) ? 4 : 0);
  const int a2 = 2 | ((cj == 0) ? 1 : 0) | ((cj == g.nx2 - 1) ? 4 : 0);
  const int a3 = 2 | ((ck == 0) ? 1 : 0) | ((ck == g.nx3 - 1) ? 4 : 0);
  for (int o3 = 0; o3 < 3; ++o3)
    for (int o2 = 0; o2 < 3; ++o2)
      for (int o1 = 0; o1 < 3; ++o1)
        if (((a1 >> o1) & (a2 >> o2) & (a3 >> o3) & 1) &&
            ((b.xmask >> (o1 + 3*o2 + 9*o3)) & 1))
          return true;
  return false;
}

// What the key pass reads of one particle (store slot q): position, velocity, rate (when the
// push or the pending correction uses it), the stage registers x1, v1 (at stage 2 of imex: where
// the stage-1 key pass wrote them, the other store at the slot of the last sort's permutation),
// mass and, with a pending correction, stopping time, the slot's tile-local cell and the
// residual rows of that cell and the cells beside it in x2 (the centre is one of them).
struct KeyIn {
  double x[3], x1[3], r[3];
  float v[3], v1[3];
  float mass, ts;
  float rrow[3][3];
  int cell;
};

// pq: the slot in the other store where the stage-1 key pass left this particle's registers
// (the last sort's permutation at q, loaded a chunk earlier by the caller: a load that waits on
// one issued just before it would wait on every load in flight)
template <bool MID>
__device__ inline void KeyFetch(const KArgs &A, const TileInfo &t, const int *st, int q,
                                int do_key, int pq, KeyIn &k) {
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  const bool rates = A.atw_dt != 0.0 || !do_key;
  k.cell = 0;
  k.ts = 0.0f;
#pragma unroll
  for (int r=0; r<3; ++r)
#pragma unroll
    for (int d=0; d<3; ++d) k.rrow[r][d] = 0.0f;
  if (A.pending) {
    k.cell = SlotCell(st, Tl.TC, q);
    int ci, cj, ck;
    CellIdx(Tl, t, k.cell, ci, cj, ck);
    const long long i0 = Idx(g, t.m, 3, 0, g.ks + ck, g.js + cj, g.is + ci);
#pragma unroll
    for (int r=0; r<3; ++r)
#pragma unroll
      for (int d=0; d<3; ++d) k.rrow[r][d] = A.res[i0 + (r - 1)*g.n1 + d*g.ncb];
    k.ts = A.src.TS(q);
  }
  for (int d=0; d<3; ++d) {
    k.x[d] = A.src.X(d, q);
    k.v[d] = A.src.V(d, q);
    k.r[d] = rates ? A.src.R(d, q) : 0.0;
    k.x1[d] = 0.0;
    k.v1[d] = 0.0f;
  }
  if (!MID && do_key && !A.stage1)
    for (int d=0; d<3; ++d) {k.x1[d] = A.dst.X1(d, pq); k.v1[d] = A.dst.V1(d, pq);}
  k.mass = A.src.M(q);
}

//----------------------------------------------------------------------------------------
// key pass, one pass over the store in its present order: the pending consistent-dust
// correction of the previous stage (dust_pm.cpp:686-696; the actual change of every particle,
// summed per centre cell, is recorded for the gas), this stage's push and migration
// (dust_push.cpp:249-297, bvals_part.cpp:164-429) written back in place, the push rounding
// residue summed per cell for the gas, and every particle's new tile-major key and rank: a
// particle that stays in its tile takes its rank from a count in LDS (the tile owns those
// cells: stay[]), one that leaves it from a global count of arrivals (count[]).  With do_key = 0
// only the pending correction is applied (written to v and R).  MID: the pc2 stage-2 form, keyed
// by the half-step position (Pc2Route) instead of pushed.
// (Bk: the block table again, as a restrict pointer: the tile's block, read by every particle,
// is then read with scalar loads once per item instead of vector loads in every chunk)
template <bool OV, bool MID = false, bool SMR = false>
__global__ void __launch_bounds__(kKeyThreadsMax, (OV || SMR) ? 2 : 4)
k_key(KArgs A, int do_key, double *pres, const DevBlock *__restrict__ Bk) {
  extern __shared__ double lds_d[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  // the workgroup's work items in the launch's list: one of the first nfirst (a launch of
  // nfirst workgroups, !OV), or every gridDim.x-th after those (OV: a smaller launch over the
  // extra parts of full tiles)
  const int nlist = A.ilist ? *A.ilist_n : A.item_start[Tl.nt];
  for (int li = OV ? A.nfirst + static_cast<int>(blockIdx.x) : FirstItem(Tl, A);
       li < nlist; li += gridDim.x) {
    const int it = A.ilist ? A.ilist[li] : li;
    const int tile = A.item_tile[it];
    const int part = it - A.item_start[tile];
    const TileInfo t = TileOf(Tl, tile);
    const int tid = threadIdx.x, CH = blockDim.x;
    KeyRec *rec = reinterpret_cast<KeyRec*>(lds_d);        // [CH]
    double *rdc = lds_d;                                   // (after the chunks: wave sums)
    int *st = reinterpret_cast<int*>(rec + CH);            // [TC + 1]
    int *lcnt = st + Tl.TC + 1;                            // [TC]
    // every cell's half sums (the correction by centre row, the push residue) in LDS slots
    // that only one thread touches: they stay out of registers during the particle phase
    double *cacc = reinterpret_cast<double*>(lds_d) + (LdsKeyBase(Tl, CH) + 7)/8;   // [2][9][TC]
    float *pacc = reinterpret_cast<float*>(cacc + kKeyHalves*9*Tl.TC);        // [2][3][TC]
    int *wcnt = reinterpret_cast<int*>(pacc + kKeyHalves*3*Tl.TC);          // [CH/64][TC]
    for (int i = tid; i < (CH >> 6)*Tl.TC; i += CH) wcnt[i] = 0;
    for (int i = tid; i < kKeyHalves*Tl.TC; i += CH) {
      for (int k=0; k<9; ++k) cacc[k*Tl.TC + (i/Tl.TC)*9*Tl.TC + i % Tl.TC] = 0.0;
      for (int k=0; k<3; ++k) pacc[k*Tl.TC + (i/Tl.TC)*3*Tl.TC + i % Tl.TC] = 0.0f;
    }
    for (int i = tid; i <= Tl.TC; i += CH) st[i] = A.start[t.key0 + i];
    for (int i = tid; i < Tl.TC; i += CH) lcnt[i] = 0;
    __syncthreads();
    // this part's particles: at most W from the tile's range
    const int P0 = min(st[0] + part*A.W, st[Tl.TC]), P1 = min(P0 + A.W, st[Tl.TC]);
    const DevBlock &b = Bk[t.m];
    const double ivol = b.ivol;
    double pmass = 0.0;   // (outflow faces: the mass and count this thread parked, chunk order)
    int pcnt = 0;
    // every particle's inputs are fetched one chunk ahead: the next chunk's loads are in flight
    // during this chunk's per-cell sums and barriers
    const bool gather_regs = !MID && do_key && !A.stage1;
    int pq_n = 0;   // (the permutation at the next chunk's slot: stage registers' source)
    if (gather_regs && P0 + tid < P1) pq_n = A.perm[P0 + tid];
    KeyIn nx;
    if (P0 + tid < P1) KeyFetch<MID>(A, t, st, P0 + tid, do_key, pq_n, nx);
    if (gather_regs && P0 + CH + tid < P1) pq_n = A.perm[P0 + CH + tid];
    int stamp = 0;   // (the chunk's number within the item, from 1)
    int acount = -1, aq = 0;   // an arrival's place in its cell's count, its slot (stored a chunk
                               // later)
    for (int c0 = P0; c0 < P1; c0 += CH) {
      ++stamp;
      int lkey = -1;   // a stayer's cell in the tile
      // the previous chunk's arrival's rank, stored now: its atomic's return is not waited for
      // before the next chunk's loads are issued, and those loads are waited for here anyway
      if (acount >= 0) {A.rank[aq] = -1 - acount; acount = -1;}
      const int c1 = min(c0 + CH, P1);
      const int q = c0 + tid;
      if (q < c1) {
        const KeyIn in = nx;
        // the slot's cell (its residual rows: this cell and the cells beside it in x2)
        int ci = 0, cj = 0, ck = 0;
        if (A.pending) CellIdx(Tl, t, in.cell, ci, cj, ck);
        PState s;
        for (int d=0; d<3; ++d) {
          s.x[d] = in.x[d]; s.v[d] = in.v[d]; s.x1[d] = in.x1[d]; s.v1[d] = in.v1[d];
          s.r[d] = in.r[d]; s.vres[d] = 0.0;
        }
        s.parked = 0;
        s.gid = t.m;   // (the tile's particles are in its block: the store is tile-major)
        const float mass = in.mass;
        int code = 0;
        double dv[3] = {0.0, 0.0, 0.0};
        if (A.pending) {
          int ic_, jc, kc_;
          if (mass > 0.0f && CenterCell(g, b, s.x, ic_, jc, kc_)) {
            const int dj = jc - cj;   // the centre: this cell or one beside it in x2
            if (ic_ != ci || kc_ != ck || dj < -1 || dj > 1) {
              dk::RecordError(A.err, DK_EHALO, q);
            } else {
              const float apf = static_cast<float>(A.a_prev);
              const float cjf = apf/(in.ts + apf);
              for (int d=0; d<3; ++d) {
                const float vo = s.v[d];
                const float rd = (dj < 0) ? in.rrow[0][d]
                                          : ((dj > 0) ? in.rrow[2][d] : in.rrow[1][d]);
                const float vn = vo + cjf*rd;
                s.v[d] = vn;
                // the recorded rate takes the actual change, so that the history terms of dust
                // and gas stay opposite (dust_pm.cpp:694)
                const double delta = static_cast<double>(vn) - static_cast<double>(vo);
                s.r[d] += delta*A.inv_a_prev;
                dv[d] = static_cast<double>(mass)*delta*ivol;
              }
              code = dj + 2;
              if (!do_key)
                for (int d=0; d<3; ++d) {
                  A.src.V(d, q) = s.v[d];
                  if (A.regs_live) A.src.R(d, q) = s.r[d];
                }
            }
          }
        }
        float pr[3] = {0.0f, 0.0f, 0.0f};
        if (do_key) {
          double xh[3];   // (MID: the half-step position, which gives the key)
          if constexpr (MID) {
            if (!Pc2Route(A, b, s, xh)) dk::RecordError(A.err, DK_EROUTE, q);
          } else {
            if (!PushMigrate<true, SMR>(A, b, s, A.stage1 != 0))
              dk::RecordError(A.err, DK_EROUTE, q);
          }
          for (int d=0; d<3; ++d) {
            A.src.X(d, q) = s.x[d];
            A.src.V(d, q) = s.v[d];
            if (A.regs_live) {A.src.X1(d, q) = s.x1[d]; A.src.V1(d, q) = s.v1[d];}
            pr[d] = static_cast<float>(static_cast<double>(mass)*s.vres[d]*ivol);
          }
          if (s.parked) {
            pmass += static_cast<double>(mass);
            ++pcnt;
            A.src.M(q) = 0.0f;
          }
          // a particle that failed to route keeps its block: its key stays valid
          int ni, nj, nk, nju;
          // the tile's block (read once by the workgroup), or the block a particle migrated to
          // (its fields read and used in the branch, so that only a wavefront holding a migrated
          // particle waits for them)
          if (s.gid == t.m) {
            const double lo[3] = {b.x1min, b.x2min, b.x3min}, sc[3] = {b.sx1, b.sx2, b.sx3};
            DepositCellOf(g, lo, sc, MID ? xh : s.x, A.sh, ni, nj, nk, nju);
          } else {
            const DevBlock &bn = A.B[s.gid];
            const double lo[3] = {bn.x1min, bn.x2min, bn.x3min}, sc[3] = {bn.sx1, bn.sx2, bn.sx3};
            DepositCellOf(g, lo, sc, MID ? xh : s.x, A.sh, ni, nj, nk, nju);
          }
          const long long key = TileKey(Tl, s.gid, ni, nj, nk);
          A.key[q] = static_cast<int>(key);
          const long long lk = key - t.key0;
          // (a stayer's rank is taken below; an arrival's from the cell's global count)
          if (lk >= 0 && lk < Tl.TC) lkey = static_cast<int>(lk);
          else {acount = atomicAdd(&A.count[key], 1); aq = q;}
        }
        KeyRec kr;
        for (int d=0; d<3; ++d) {kr.dv[d] = dv[d]; kr.pr[d] = pr[d];}
        kr.code = code;
        kr.pad = 0;
        rec[tid] = kr;
      }
      // A stayer's rank is its rank among its part's stayers of its new cell in slot order (so
      // the stayers of a cell need no sort): the earlier chunks' count (lcnt), the earlier
      // wavefronts' counts in this chunk, the earlier lanes of its wavefront.  Within the
      // wavefront: one pass per distinct cell, its lanes found by a ballot.
      const int lane = tid & 63, wv = tid >> 6;
      int wrank = 0;
      if (do_key) {
        unsigned long long pend = __ballot(lkey >= 0);
        while (pend) {
          const int leader = __ffsll(static_cast<long long>(pend)) - 1;
          const int cell = __shfl(lkey, leader, 64);
          const unsigned long long peers = __ballot(lkey == cell);
          if (lkey == cell) wrank = __popcll(peers & ((1ull << lane) - 1ull));
          if (lane == leader) wcnt[wv*Tl.TC + cell] = (stamp << 16) | __popcll(peers);
          pend &= ~peers;
        }
      }
      // (the next chunk's slots: no particle of this chunk is written there)
      if (q + CH < P1) KeyFetch<MID>(A, t, st, q + CH, do_key, pq_n, nx);
      if (gather_regs && q + 2*CH < P1) pq_n = A.perm[q + 2*CH];
      __syncthreads();
      int ctot = 0;   // (thread l < TC: cell l's stayers of this chunk)
      if (do_key) {
        if (lkey >= 0) {
          int r = lcnt[lkey] + wrank;
          for (int w = 0; w < wv; ++w) {
            const int x = wcnt[w*Tl.TC + lkey];
            if ((x >> 16) == stamp) r += x & 0xffff;
          }
          A.rank[q] = part*kPartRank + r;
        }
        if (tid < Tl.TC)
          for (int w = 0; w < (CH >> 6); ++w) {
            const int x = wcnt[w*Tl.TC + tid];
            if ((x >> 16) == stamp) ctot += x & 0xffff;
          }
      }
      if (A.cell_sums) {
        // (cell l, half h): the records of l in this chunk whose slot has the parity h within
        // the cell, in slot order
        for (int i = tid; i < kKeyHalves*Tl.TC; i += CH) {
          const int l = i % Tl.TC, h = i/Tl.TC;
          double *ca = cacc + h*9*Tl.TC + l;
          float *pa = pacc + h*3*Tl.TC + l;
          double dcs[3][3];
          float prs[3];
          for (int k=0; k<9; ++k) dcs[k/3][k%3] = ca[k*Tl.TC];
          for (int k=0; k<3; ++k) prs[k] = pa[k*Tl.TC];
          const int a = st[l];
          int s0 = max(a, c0);
          s0 += (s0 - a + h) & 1;
          const int s1 = min(st[l + 1], c1);
          for (int qq = s0; qq < s1; qq += 2) {
            const KeyRec kr = rec[qq - c0];
            const int code = kr.code;
            if (code == 1) {dcs[0][0] += kr.dv[0]; dcs[0][1] += kr.dv[1]; dcs[0][2] += kr.dv[2];}
            else if (code == 2) {dcs[1][0] += kr.dv[0]; dcs[1][1] += kr.dv[1]; dcs[1][2] += kr.dv[2];}
            else if (code == 3) {dcs[2][0] += kr.dv[0]; dcs[2][1] += kr.dv[1]; dcs[2][2] += kr.dv[2];}
            prs[0] += kr.pr[0]; prs[1] += kr.pr[1]; prs[2] += kr.pr[2];
          }
          for (int k=0; k<9; ++k) ca[k*Tl.TC] = dcs[k/3][k%3];
          for (int k=0; k<3; ++k) pa[k*Tl.TC] = prs[k];
        }
      }
      __syncthreads();
      if (tid < Tl.TC) lcnt[tid] += ctot;   // (read again only after the next chunk's barrier)
    }
    if (acount >= 0) A.rank[aq] = -1 - acount;   // (the last chunk's arrival)
    double dcs[3][3];
    float prs[3];
    if (tid < Tl.TC) {   // (the two half sums in order)
      for (int k=0; k<9; ++k) dcs[k/3][k%3] = cacc[k*Tl.TC + tid] + cacc[(9 + k)*Tl.TC + tid];
      for (int k=0; k<3; ++k) prs[k] = pacc[k*Tl.TC + tid] + pacc[(3 + k)*Tl.TC + tid];
    }
    if (do_key && part == 0)
      for (int i = tid; i < Tl.TC; i += CH) A.stay[t.key0 + i] = lcnt[i];
    else if (do_key)
      for (int i = tid; i < Tl.TC; i += CH)
        A.ov_cnt[OvSlot(A.item_start, tile, part)*Tl.TC + i] = lcnt[i];
    // The correction's shares for the cells beside each cell in x2 (centre one row below: dcs[0],
    // one row above: dcs[2]) are exchanged in LDS (the records' area, free once the chunks are
    // done): a cell's dc0 holds its own share plus those its tile neighbours hold for it, and
    // only a tile's last and first x2 rows pass theirs on, to the neighbouring tiles' rows, in
    // dc1 (from below) and dc2 (from above).  Every target is written by exactly one thread.
    double *nb = lds_d;   // [2][3][TC]
    if (A.pending && part == 0) {
      if (tid < Tl.TC)
        for (int d=0; d<3; ++d) {nb[d*Tl.TC + tid] = dcs[0][d]; nb[(3 + d)*Tl.TC + tid] = dcs[2][d];}
      __syncthreads();
    }
    if (tid < Tl.TC && part == 0) {
      int ci, cj, ck;
      CellIdx(Tl, t, tid, ci, cj, ck);
      const long long i0 = Idx(g, t.m, 3, 0, g.ks + ck, g.js + cj, g.is + ci);
      if (do_key)
        for (int d=0; d<3; ++d) pres[i0 + d*g.ncb] = static_cast<double>(prs[d]);
      if (A.pending) {
        const int row = (tid >> Tl.LX) & (Tl.TY - 1);
        for (int d=0; d<3; ++d) {
          double own = dcs[1][d];
          if (row > 0) own += nb[(3 + d)*Tl.TC + tid - Tl.TX];
          if (row < Tl.TY - 1) own += nb[d*Tl.TC + tid + Tl.TX];
          A.dc0[i0 + d*g.ncb] = own;
          if (row == Tl.TY - 1) A.dc1[i0 + g.n1 + d*g.ncb] = dcs[2][d];
          if (row == 0) A.dc2[i0 - g.n1 + d*g.ncb] = dcs[0][d];
        }
      }
    } else if (tid < Tl.TC) {
      // a later part: its partial sums to its overflow slot (added by k_fix_key)
      const long long o = OvSlot(A.item_start, tile, part);
      if (do_key)
        for (int d=0; d<3; ++d) A.ov_pr[(o*3 + d)*Tl.TC + tid] = prs[d];
      if (A.pending)
        for (int k=0; k<3; ++k)
          for (int d=0; d<3; ++d) A.ov_dc[(o*9 + 3*k + d)*Tl.TC + tid] = dcs[k][d];
    }
    if (A.x3p != nullptr && do_key) {
      // the item's parked mass and count in a fixed order: lanes by a butterfly, the waves in
      // order (rdc is free once the chunks are done)
      if (__syncthreads_or(pcnt)) {
        const double wm = WaveSum(pmass, 64);
        const double wc = WaveSum(static_cast<double>(pcnt), 64);
        if ((tid & 63) == 0) {rdc[2*(tid >> 6)] = wm; rdc[2*(tid >> 6) + 1] = wc;}
        __syncthreads();
        if (tid == 0) {
          double sm = 0.0, sc = 0.0;
          for (int w = 0; w < CH/64; ++w) {sm += rdc[2*w]; sc += rdc[2*w + 1];}
          A.x3p[2*it] = sm; A.x3p[2*it + 1] = sc;
        }
      } else if (tid == 0) {
        A.x3p[2*it] = 0.0; A.x3p[2*it + 1] = 0.0;
      }
    }
    if constexpr (!OV) break;
    __syncthreads();   // (the next item reuses LDS)
  }
}

// the key pass's outflow partials summed over its items in a fixed order into the ledger (mass,
// count): kX3Groups workgroups each sum a contiguous run of items (a strided sum per thread,
// then a tree), one workgroup then adds the runs' sums by a tree
constexpr int kX3Groups = 128;
__global__ void __launch_bounds__(256) k_x3part(const int *item_start, int nt, const double *x3p,
                                                double *part) {
  __shared__ double sm[256], sc[256];
  const int n = item_start[nt];
  const int per = (n + kX3Groups - 1)/kX3Groups;
  const int i0 = blockIdx.x*per, i1 = min(n, i0 + per);
  double a = 0.0, b = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {a += x3p[2*i]; b += x3p[2*i + 1];}
  sm[threadIdx.x] = a; sc[threadIdx.x] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {sm[threadIdx.x] += sm[threadIdx.x + h]; sc[threadIdx.x] += sc[threadIdx.x + h];}
    __syncthreads();
  }
  if (threadIdx.x == 0) {part[2*blockIdx.x] = sm[0]; part[2*blockIdx.x + 1] = sc[0];}
}

__global__ void __launch_bounds__(kX3Groups) k_x3sum(const double *part, double *ledger) {
  __shared__ double sm[kX3Groups], sc[kX3Groups];
  sm[threadIdx.x] = part[2*threadIdx.x];
  sc[threadIdx.x] = part[2*threadIdx.x + 1];
  __syncthreads();
  for (int h = kX3Groups/2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {sm[threadIdx.x] += sm[threadIdx.x + h]; sc[threadIdx.x] += sc[threadIdx.x + h];}
    __syncthreads();
  }
  if (threadIdx.x == 0) {ledger[0] += sm[0]; ledger[1] += sc[0];}
}

// the parts of every tile: at most W particles each, at least one (the entry after the last tile
// is 0, for the scan)
__global__ void k_nparts(Tiles Tl, const int *start, int W, int *nparts) {
  GRID_STRIDE(t, (long long)Tl.nt + 1) {
    if (t == Tl.nt) {nparts[t] = 0; continue;}
    const long long k0 = t*(long long)Tl.TC;
    const int n = start[k0 + Tl.TC] - start[k0];
    nparts[t] = max(1, (n + W - 1)/W);
  }
}

// the tile of every work item
__global__ void k_item_tile(Tiles Tl, const int *item_start, int *item_tile) {
  GRID_STRIDE(t, (long long)Tl.nt) {
    const int i0 = item_start[t], i1 = item_start[t + 1];
    for (int i = i0; i < i1; ++i) item_tile[i] = static_cast<int>(t);
  }
}

// the class of every item slot (1: a tile on a block's x2 face; 0 beyond the last item)
// (multi: the items of tiles of more than one part, where crowded cells are, instead)
__global__ void k_item_flags(Tiles Tl, const int *item_start, const int *item_tile, long long n,
                             int *flag, int multi) {
  const int nitem = item_start[Tl.nt];
  GRID_STRIDE(i, n) {
    int f = 0;
    if (i < nitem) {
      const int t = item_tile[i];
      f = multi ? ((item_start[t + 1] - item_start[t] > 1) ? 1 : 0) : (EdgeTile(Tl, t) ? 1 : 0);
    }
    flag[i] = f;
  }
}

// the two item lists in item order (e: the exclusive scan of the flags over n + 1 slots)
__global__ void k_item_lists(Tiles Tl, const int *item_start, const int *flag, const int *e,
                             long long n, int *list0, int *list1, int *count) {
  const int nitem = item_start[Tl.nt];
  GRID_STRIDE(i, (long long)nitem) {
    if (flag[i]) list1[e[i]] = static_cast<int>(i);
    else list0[i - e[i]] = static_cast<int>(i);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    count[1] = e[n];
    count[0] = nitem - e[n];
  }
}

// the count of every cell: the particles that stay in its tile plus the arrivals
__global__ void k_total(int *count, const int *stay, long long nkeys) {
  GRID_STRIDE(c, nkeys) count[c] += stay[c];
}

// placement: stayers first (part*kPartRank + their rank in the part: a later part's stayers after
// the offset in its overflow slot), then arrivals (rank -1 - r); item_start: the items of the
// order the ranks were taken in.  Each thread places kPlaceU particles a block apart, their loads
// issued together: the kernel is bound by the latency of its dependent loads (key, then the
// cell's start), not by their bytes.
constexpr int kPlaceU = 4;
__device__ inline int PlaceOff(int k, int r, const int *stay, int TC, const int *item_start,
                               const int *ov_cnt) {
  if (r < 0) return stay[k] - 1 - r;
  if (r < kPartRank) return r;
  const int part = r/kPartRank, tile = k/TC;
  return ov_cnt[((long long)item_start[tile] - tile + part - 1)*TC + (k - tile*TC)] +
         (r - part*kPartRank);
}
__global__ void __launch_bounds__(kBlock)
k_place(const int *key, const int *rank, const int *start, const int *stay, long long np, int TC,
        const int *item_start, const int *ov_cnt, int *perm) {
  const long long span = (long long)gridDim.x*kBlock*kPlaceU;
  for (long long p0 = (long long)blockIdx.x*kBlock*kPlaceU + threadIdx.x; p0 < np; p0 += span) {
    int k[kPlaceU], r[kPlaceU], s[kPlaceU];
#pragma unroll
    for (int u=0; u<kPlaceU; ++u) {
      const long long p = p0 + u*kBlock;
      k[u] = -1; r[u] = 0;
      if (p < np) {k[u] = key[p]; r[u] = rank[p];}
    }
#pragma unroll
    for (int u=0; u<kPlaceU; ++u)
      if (k[u] >= 0) s[u] = start[k[u]] + PlaceOff(k[u], r[u], stay, TC, item_start, ov_cnt);
#pragma unroll
    for (int u=0; u<kPlaceU; ++u)
      if (k[u] >= 0) perm[s[u]] = static_cast<int>(p0 + u*kBlock);
  }
}

// every cell's arrivals (its list after its stay[] stayers, which the key pass ranked in slot
// order) in ascending source slot, a thread per key: a short list of arrivals (2..kSegSmall) is
// ordered in place in global memory by insertion sort; longer ones are queued for the medium
// and long sorts, with one global atomic per workgroup pass and class.  Stayers and lists of
// one arrival are not touched: arrivals are a small share of a stage's particles, so a pass
// reads the keys' starts and stayer counts and little else.
__global__ void __launch_bounds__(kBlock) k_segsort(const int *start, const int *stay,
                                                    long long nkeys, int *perm, int *big) {
  __shared__ int qn[2], qbase[2];   // the pass's medium / long lists and their queue base
  const int t = threadIdx.x;
  for (long long c0 = (long long)blockIdx.x*kBlock; c0 < nkeys;
       c0 += (long long)gridDim.x*kBlock) {
    const long long c = c0 + t;
    int n = 0, s0 = 0;
    if (c < nkeys) {
      s0 = start[c] + stay[c];
      n = start[c + 1] - s0;
    }
    const int cls = (n > kMedMax) ? 1 : ((n > kSegSmall) ? 0 : -1);
    if (t < 2) qn[t] = 0;
    __syncthreads();
    const int qs = (cls >= 0) ? atomicAdd(&qn[cls], 1) : 0;
    __syncthreads();
    if (t < 2) qbase[t] = qn[t] ? atomicAdd(&big[t], qn[t]) : 0;
    __syncthreads();
    if (cls == 1) big[1 + nkeys - (qbase[1] + qs)] = static_cast<int>(c);
    else if (cls == 0) big[2 + qbase[0] + qs] = static_cast<int>(c);
    if (n > 1 && n <= kSegSmall) {
      int *a = perm + s0;
      for (int i=1; i<n; ++i) {
        const int v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) {a[j + 1] = a[j]; --j;}
        a[j + 1] = v;
      }
    }
    __syncthreads();   // (qn and qbase are reset by the next pass)
  }
}

// one wavefront per medium list: an element's place is the count of smaller source slots in the
// list (the slots are distinct), the list staged in LDS
__global__ void __launch_bounds__(64) k_medsort(const int *start, const int *stay, int *perm,
                                                 const int *big) {
  constexpr int kPer = kMedMax/64;
  __shared__ int sh[kMedMax];
  const int nmed = big[0];
  const int lane = threadIdx.x;
  for (int e = blockIdx.x; e < nmed; e += gridDim.x) {
    const int c = big[2 + e];
    const int s0 = start[c] + stay[c], n = start[c + 1] - s0;
    int v[kPer], r[kPer];
#pragma unroll
    for (int t=0; t<kPer; ++t) {
      const int i = t*64 + lane;
      v[t] = (i < n) ? perm[s0 + i] : INT_MAX;
      r[t] = 0;
      sh[i] = v[t];
    }
    __syncthreads();
    for (int j=0; j<n; ++j) {
      const int x = sh[j];
#pragma unroll
      for (int t=0; t<kPer; ++t) r[t] += (x < v[t]) ? 1 : 0;
    }
#pragma unroll
    for (int t=0; t<kPer; ++t)
      if (t*64 + lane < n) perm[s0 + r[t]] = v[t];
    __syncthreads();
  }
}

// one workgroup per long list: bitonic sort in LDS (padded to a power of two)
__global__ void __launch_bounds__(1024) k_bigsort(const int *start, const int *stay, int *perm,
                                                  const int *big, long long nkeys,
                                                  dk::ErrWord *err) {
  __shared__ int sh[kBigMax];
  const int nbig = big[1];
  for (int e = blockIdx.x; e < nbig; e += gridDim.x) {
    const int c = big[1 + nkeys - e];
    const int s0 = start[c] + stay[c], n = start[c + 1] - s0;
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
// P1: the pushed particles (the key pass pushed them in place) gathered into the new order, and
// the deposit of Q = sum m c W/V and P = sum m c W v/V (dust_pm.cpp:225-274) and of the centre
// mass m c/V (dust_pm.cpp:660-667)
template <int S, bool SY, bool OV, bool SPLIT = false>
__global__ void __launch_bounds__(kCellThreads, (SY || OV || SPLIT) ? 4 : 6)
k_p1(KArgs A, float *qt0, float *qt1, float *qt2, float *scr, const DevBlock *__restrict__ Bk) {
  extern __shared__ double lds_d[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  // the workgroup's work items in the launch's list: one of the first nfirst (a launch of
  // nfirst workgroups, !OV), or every gridDim.x-th after those (OV: a smaller launch over the
  // extra parts of full tiles)
  const int nlist = A.ilist ? *A.ilist_n : A.item_start[Tl.nt];
  for (int li = OV ? A.nfirst + static_cast<int>(blockIdx.x) : FirstItem(Tl, A);
       li < nlist; li += gridDim.x) {
    const int it = A.ilist ? A.ilist[li] : li;
    const int tile = A.item_tile[it];
    const int part = it - A.item_start[tile];
    const TileInfo t = TileOf(Tl, tile);
    const int tid = threadIdx.x, NT = blockDim.x, CH = kPPT*NT;
    float *rd = reinterpret_cast<float*>(lds_d);           // [3][CH] offsets in the cell x, y, z
    float *rv = rd + 3*CH;                                 // [4][CH] mu, mu v
    int *rcode = reinterpret_cast<int*>(rv + 4*CH);        // [CH] centre | (sj + 1) << 2
    float *reg = rd + RecArea(Tl, CH, 8, 4);               // [4][RN]
    int *st = reinterpret_cast<int*>(reg + 4*Tl.RN);       // [TC + 1]
    [[maybe_unused]] int *plan = st + Tl.TC + 1;           // [PlanInts(TC)]
    float *q3s = reinterpret_cast<float*>(plan + PlanInts(Tl.TC));   // [3][TC] centre masses
    for (int i = tid; i <= Tl.TC; i += NT) st[i] = A.start[t.key0 + i];
    for (int i = tid; i < 4*Tl.RN; i += NT) reg[i] = 0.0f;
    for (int i = tid; i < 3*Tl.TC; i += NT) q3s[i] = 0.0f;
    __syncthreads();
    // this part's particles: at most W from the tile's range
    const int P0 = min(st[0] + part*A.W, st[Tl.TC]), P1 = min(P0 + A.W, st[Tl.TC]);
    const DevBlock &b = Bk[t.m];
    const double ivol = b.ivol;
    const int slab = tid/Tl.TC;
    const int l = tid - slab*Tl.TC;
    float q3[3] = {0.0f, 0.0f, 0.0f};
    for (int c0 = P0; c0 < P1; c0 += CH) {
      const int c1 = min(c0 + CH, P1);
      // the permutation of the chunk's slots, one particle ahead
      int pn = (c0 + tid < c1) ? A.perm[c0 + tid] : 0;
#pragma unroll 1
      for (int k=0; k<kPPT; ++k) {
        const int r = tid + k*NT;
        const int q = c0 + r;
        if (q >= c1) break;
        const int p = pn;
        if (k + 1 < kPPT && q + NT < c1) pn = A.perm[q + NT];
        const double x[3] = {A.src.X(0, p), A.src.X(1, p), A.src.X(2, p)};
        const float v[3] = {A.src.V(0, p), A.src.V(1, p), A.src.V(2, p)};
        const float ts = A.src.TS(p), mass = A.src.M(p);
        const int kp = A.key[p];
        for (int d=0; d<3; ++d) {A.dst.X(d, q) = x[d]; A.dst.V(d, q) = v[d];}
        // (x1, v1 stay in the source order: the next stage's key pass reads them there through
        // this sort's permutation)
        // (no block index: every particle of a tile is in the tile's block; an export reads it
        // from the key)
        A.dst.TS(q) = ts; A.dst.M(q) = mass; A.dst.ORIG(q) = A.src.ORIG(p);
        float ox = 0.0f, oy = 0.0f, oz = 0.0f;
        float mu = 0.0f;
        int code = 1 << 2;   // no centre, sj = 0
        const Loc L = Locate<SY>(g, b, x, A.sh);
        // the key pass placed the particle by this position: the cell must be its slot's
        if (TileKey(Tl, t.m, L.ci, L.cj, L.ck) != kp) {
          dk::RecordError(A.err, DK_EHALO, p);
        } else {
          const float af = static_cast<float>(A.a_dt);
          mu = mass*(af/(ts + af))*static_cast<float>(ivol);
          int cen = 0;
          if (mass > 0.0f && L.cen) {
            // the centre of the unshifted position: this cell or one beside it in x2
            if (L.jc - L.cj < -1 || L.jc - L.cj > 1) dk::RecordError(A.err, DK_EHALO, p);
            else cen = L.jc - L.cj + 2;
          }
          code = cen | ((L.sj + 1) << 2);
          ox = static_cast<float>(L.dx);
          oy = static_cast<float>(L.dy);
          oz = static_cast<float>(L.dz);
        }
        // (a record with mu = 0 deposits zeros whatever its weights)
        rd[r] = ox; rd[CH + r] = oy; rd[2*CH + r] = oz;
        rv[r] = mu;
        rv[CH + r] = mu*v[0]; rv[2*CH + r] = mu*v[1]; rv[3*CH + r] = mu*v[2];
        rcode[r] = code;
      }
      MarkChunkSpan(Tl, st, c0, c1, plan + kSpanSlot);
      __syncthreads();
      float acc[3][3][4], spl[3][4];
#pragma unroll
      for (int a=0; a<3; ++a)
#pragma unroll
        for (int n=0; n<4; ++n) {
          spl[a][n] = 0.0f;
#pragma unroll
          for (int bb=0; bb<3; ++bb) acc[bb][a][n] = 0.0f;
        }
      int spdir = 0;
      if constexpr (SPLIT) {
        const auto add = [&](int r, float (&ac)[3][3][4], float (&e)[3]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(slab, rd[2*CH + r]);
          const float val[4] = {rv[r], rv[CH + r], rv[2*CH + r], rv[3*CH + r]};
          const int code = rcode[r];
          AddSlab<float, 4, SY>(ac, spl, spdir, ((code >> 2) & 3) - 1, wz, wy, wx, val);
          if (slab == 1) {
            const int cen = code & 3;
            e[0] += (cen == 1) ? val[0] : 0.0f;
            e[1] += (cen == 2) ? val[0] : 0.0f;
            e[2] += (cen == 3) ? val[0] : 0.0f;
          }
        };
        CellSums<4, 3>(Tl, st, c0, c1, l, slab, rd, RecArea(Tl, CH, 8, 4), plan, add, acc, q3);
        DepositSlabs<float, 4, SY>(Tl, reg, rd, l, slab, acc, spl, spdir);
      } else {
        // (the centre masses: summed by the slab-1 threads, the chunk's sums of a cell added to
        // its running sums in LDS)
        float e[3] = {0.0f, 0.0f, 0.0f};
        const auto add = [&](int r, int sl, float (&ac)[3][3][4], float (&sp)[3][4], int &sd,
                             float (&ee)[3]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(sl, rd[2*CH + r]);
          const float val[4] = {rv[r], rv[CH + r], rv[2*CH + r], rv[3*CH + r]};
          const int code = rcode[r];
          AddSlab<float, 4, SY>(ac, sp, sd, ((code >> 2) & 3) - 1, wz, wy, wx, val);
          if (sl == 1) {
            const int cen = code & 3;
            ee[0] += (cen == 1) ? val[0] : 0.0f;
            ee[1] += (cen == 2) ? val[0] : 0.0f;
            ee[2] += (cen == 3) ? val[0] : 0.0f;
          }
        };
        const CellMap m = ChunkSums<4, SY>(Tl, st, c0, c1, plan + kSpanSlot, add, acc, spl,
                                                   spdir, e);
        if (m.dep && m.slab == 1)
          for (int k=0; k<3; ++k) q3s[k*Tl.TC + m.l] += e[k];
        DepositSlabs<float, 4, SY>(Tl, reg, rd, m.l, m.slab, acc, spl, spdir, m.dep,
                                      m.layers);
      }
    }
    if constexpr (!SPLIT) {   // (the last deposit ended with a barrier)
      for (int k=0; k<3; ++k) q3[k] = q3s[k*Tl.TC + l];
    }
    if (slab == 1 && part == 0) {
      int ci, cj, ck;
      CellIdx(Tl, t, l, ci, cj, ck);
      const long long i0 = Idx(g, t.m, 1, 0, g.ks + ck, g.js + cj, g.is + ci);
      qt0[i0] = q3[1];
      qt1[i0 + g.n1] = q3[2];
      qt2[i0 - g.n1] = q3[0];
    } else if (slab == 1) {
      // a later part: to its overflow slot (added by k_fix_qt)
      float *o = A.ov_q + OvSlot(A.item_start, tile, part)*3*Tl.TC;
      o[l] = q3[1]; o[Tl.TC + l] = q3[2]; o[2*Tl.TC + l] = q3[0];
    }
    if (part == 0) FlushRegion<float, 4>(Tl, t, reg, scr);
    else FlushOverflow<4>(Tl, reg, A.ov_scr + OvSlot(A.item_start, tile, part)*4*Tl.RN);
    if constexpr (!OV) break;
    __syncthreads();   // (the next item reuses LDS)
  }
}

// P2: the particle part of the dc1 operator, result += G^T (m c/V) G u* (dust_solver.cpp:151-201)
template <int S, bool SY, bool OV, bool SPLIT = false>
__global__ void __launch_bounds__(kCellThreads, (SY || OV || SPLIT) ? 4 : 8)
k_p2(KArgs A, const float *us, float *scr, const DevBlock *__restrict__ Bk) {
  extern __shared__ double lds_d[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  // the workgroup's work items in the launch's list: one of the first nfirst (a launch of
  // nfirst workgroups, !OV), or every gridDim.x-th after those (OV: a smaller launch over the
  // extra parts of full tiles)
  const int nlist = A.ilist ? *A.ilist_n : A.item_start[Tl.nt];
  for (int li = OV ? A.nfirst + static_cast<int>(blockIdx.x) : FirstItem(Tl, A);
       li < nlist; li += gridDim.x) {
    const int it = A.ilist ? A.ilist[li] : li;
    const int tile = A.item_tile[it];
    const int part = it - A.item_start[tile];
    const TileInfo t = TileOf(Tl, tile);
    const int tid = threadIdx.x, NT = blockDim.x, CH = kPPT*NT;
    float *dep = reinterpret_cast<float*>(lds_d);          // [3][RN] the deposit
    float *us_ = dep + 3*Tl.RN;                            // [3][RN] u*
    float *rd = us_ + 3*Tl.RN;                             // [3][CH] offsets in the cell x, y, z
    float *rg = rd + 3*CH;                                 // [3][CH] (m c/V) G u*
    int *rcode = reinterpret_cast<int*>(rg + 3*CH);        // [CH] sj + 1 (SY only)
    int *st = reinterpret_cast<int*>(rd + RecArea(Tl, CH, SY ? 7 : 6, 3));   // [TC + 1]
    [[maybe_unused]] int *plan = st + Tl.TC + 1;           // [PlanInts(TC)]
    RegionFetch<float> uf;
    uf.Load(g, Tl, t, us, 3);
    for (int i = tid; i <= Tl.TC; i += NT) st[i] = A.start[t.key0 + i];
    for (int i = tid; i < 3*Tl.RN; i += NT) dep[i] = 0.0f;
    __syncthreads();
    uf.Store(Tl, 3, us_);
    __syncthreads();
    // this part's particles: at most W from the tile's range
    const int P0 = min(st[0] + part*A.W, st[Tl.TC]), P1 = min(P0 + A.W, st[Tl.TC]);
    const DevBlock &b = Bk[t.m];
    const double ivol = b.ivol;
    const int slab = tid/Tl.TC;
    const int l = tid - slab*Tl.TC;
    for (int c0 = P0; c0 < P1; c0 += CH) {
      const int c1 = min(c0 + CH, P1);
      // each particle's inputs are loaded one particle ahead (in flight during the gather)
      double nx[3] = {0.0, 0.0, 0.0};
      float nts = 0.0f, nm = 0.0f;
      if (c0 + tid < c1) {
        for (int d=0; d<3; ++d) nx[d] = A.src.X(d, c0 + tid);
        nts = A.src.TS(c0 + tid); nm = A.src.M(c0 + tid);
      }
#pragma unroll 1
      for (int k=0; k<kPPT; ++k) {
        const int r = tid + k*NT;
        const int q = c0 + r;
        if (q >= c1) break;
        const double x[3] = {nx[0], nx[1], nx[2]};
        const float tsq = nts, mq = nm;
        if (k + 1 < kPPT && q + NT < c1) {
          for (int d=0; d<3; ++d) nx[d] = A.src.X(d, q + NT);
          nts = A.src.TS(q + NT); nm = A.src.M(q + NT);
        }
        const Loc L = LocateInTile<SY>(g, Tl, t, b, x, A.sh);
        const int sj = L.sj;
        float wx[3], wy[3], wz[3];
        W3<S, float>(static_cast<float>(L.dx), wx);
        W3<S, float>(static_cast<float>(L.dy), wy);
        W3<S, float>(static_cast<float>(L.dz), wz);
        const int rb = ((L.ck - t.ck0)*Tl.RY + (L.cj - t.cj0 + Tl.SY + sj))*Tl.RX + (L.ci - t.ci0);
        float gth[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c=0; c<3; ++c)
#pragma unroll
          for (int bb=0; bb<3; ++bb) {
            const float wcb = wz[c]*wy[bb];
            const int rr = rb + (c*Tl.RY + bb)*Tl.RX;
#pragma unroll
            for (int a=0; a<3; ++a) {
              const float w = wcb*wx[a];
              gth[0] += w*us_[rr + a];
              gth[1] += w*us_[Tl.RN + rr + a];
              gth[2] += w*us_[2*Tl.RN + rr + a];
            }
          }
        const float af = static_cast<float>(A.a_dt);
        const float fac = mq*(af/(tsq + af))*static_cast<float>(ivol);
        rd[r] = static_cast<float>(L.dx); rd[CH + r] = static_cast<float>(L.dy);
        rd[2*CH + r] = static_cast<float>(L.dz);
        rg[r] = fac*gth[0]; rg[CH + r] = fac*gth[1]; rg[2*CH + r] = fac*gth[2];
        if constexpr (SY) rcode[r] = sj + 1;
      }
      MarkChunkSpan(Tl, st, c0, c1, plan + kSpanSlot);
      __syncthreads();
      float acc[3][3][3], spl[3][3];
#pragma unroll
      for (int a=0; a<3; ++a)
#pragma unroll
        for (int n=0; n<3; ++n) {
          spl[a][n] = 0.0f;
#pragma unroll
          for (int bb=0; bb<3; ++bb) acc[bb][a][n] = 0.0f;
        }
      int spdir = 0;
      if constexpr (SPLIT) {
        float nex[1] = {0.0f};
        const auto add = [&](int r, float (&ac)[3][3][3], float (&)[1]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(slab, rd[2*CH + r]);
          const float val[3] = {rg[r], rg[CH + r], rg[2*CH + r]};
          AddSlab<float, 3, SY>(ac, spl, spdir, SY ? rcode[r] - 1 : 0, wz, wy, wx, val);
        };
        CellSums<3, 1>(Tl, st, c0, c1, l, slab, rd, RecArea(Tl, CH, 6, 3), plan, add, acc, nex);
        DepositSlabs<float, 3, SY>(Tl, dep, rd, l, slab, acc, spl, spdir);
      } else {
        float e[3] = {0.0f, 0.0f, 0.0f};
        const auto add = [&](int r, int sl, float (&ac)[3][3][3], float (&sp)[3][3], int &sd,
                             float (&)[3]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(sl, rd[2*CH + r]);
          const float val[3] = {rg[r], rg[CH + r], rg[2*CH + r]};
          AddSlab<float, 3, SY>(ac, sp, sd, SY ? rcode[r] - 1 : 0, wz, wy, wx, val);
        };
        const CellMap m = ChunkSums<3, SY>(Tl, st, c0, c1, plan + kSpanSlot, add, acc, spl,
                                                   spdir, e);
        DepositSlabs<float, 3, SY>(Tl, dep, rd, m.l, m.slab, acc, spl, spdir, m.dep,
                                      m.layers);
      }
    }
    if (part == 0) FlushRegion<float, 3>(Tl, t, dep, scr);
    else FlushOverflow<3>(Tl, dep, A.ov_scr + OvSlot(A.item_start, tile, part)*4*Tl.RN);
    if constexpr (!OV) break;
    __syncthreads();   // (the next item reuses LDS)
  }
}

// P3: gather u*, kick, drag rate, and the deposit of the actual impulse (dust_pm.cpp:500-564).
// The impulse is deposited in FP32; the tile's defect, the FP64 sum of its particles' impulses
// minus the FP64 sum of its FP32 region, goes to the combine, which adds it to the tile's first
// cell: the FP64 field then holds exactly (to FP64 round-off) the particles' total impulse.
// (REGS: the drag rate R is written, for a later stage; a template argument, as a run-time test
// around the store costs the kernel registers)
template <int S, bool SY, bool REGS, bool OV, bool SPLIT = false>
__device__ inline void P3Body(const KArgs &A, const float *us, float *scr, double *defect,
                              const DevBlock *__restrict__ Bk) {
  extern __shared__ double lds_d[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  // the workgroup's work items in the launch's list: one of the first nfirst (a launch of
  // nfirst workgroups, !OV), or every gridDim.x-th after those (OV: a smaller launch over the
  // extra parts of full tiles)
  const int nlist = A.ilist ? *A.ilist_n : A.item_start[Tl.nt];
  for (int li = OV ? A.nfirst + static_cast<int>(blockIdx.x) : FirstItem(Tl, A);
       li < nlist; li += gridDim.x) {
    const int it = A.ilist ? A.ilist[li] : li;
    const int tile = A.item_tile[it];
    const int part = it - A.item_start[tile];
    const TileInfo t = TileOf(Tl, tile);
    const int tid = threadIdx.x, NT = blockDim.x, CH = kPPT*NT;
    float *dep = reinterpret_cast<float*>(lds_d);          // [3][RN] the deposit
    float *us_ = dep + 3*Tl.RN;                            // [3][RN] u*
    float *rd = us_ + 3*Tl.RN;                             // [3][CH] offsets in the cell x, y, z
    float *rv = rd + 3*CH;                                 // [3][CH] -m delta/V
    int *rcode = reinterpret_cast<int*>(rv + 3*CH);        // [CH] sj + 1 (SY only)
    int *st = reinterpret_cast<int*>(rd + RecArea(Tl, CH, SY ? 7 : 6, 3));   // [TC + 1]
    [[maybe_unused]] int *plan = st + Tl.TC + 1;           // [PlanInts(TC)]
    RegionFetch<float> uf;
    uf.Load(g, Tl, t, us, 3);
    for (int i = tid; i <= Tl.TC; i += NT) st[i] = A.start[t.key0 + i];
    for (int i = tid; i < 3*Tl.RN; i += NT) dep[i] = 0.0f;
    __syncthreads();
    uf.Store(Tl, 3, us_);
    __syncthreads();
    // this part's particles: at most W from the tile's range
    const int P0 = min(st[0] + part*A.W, st[Tl.TC]), P1 = min(P0 + A.W, st[Tl.TC]);
    const DevBlock &b = Bk[t.m];
    const double ivol = b.ivol;
    const int slab = tid/Tl.TC;
    const int l = tid - slab*Tl.TC;
    // this thread's particles' impulses in FP64, in its fixed particle order
    double ex[3] = {0.0, 0.0, 0.0};
    for (int c0 = P0; c0 < P1; c0 += CH) {
      const int c1 = min(c0 + CH, P1);
#pragma unroll 1
      for (int k=0; k<kPPT; ++k) {
        const int r = tid + k*NT;
        const int q = c0 + r;
        if (q >= c1) break;
        const double x[3] = {A.src.X(0, q), A.src.X(1, q), A.src.X(2, q)};
        const float vk0[3] = {A.src.V(0, q), A.src.V(1, q), A.src.V(2, q)};
        const float tsq = A.src.TS(q), mq = A.src.M(q);
        const Loc L = LocateInTile<SY>(g, Tl, t, b, x, A.sh);
        const int sj = L.sj;
        float wx[3], wy[3], wz[3];
        W3<S, float>(static_cast<float>(L.dx), wx);
        W3<S, float>(static_cast<float>(L.dy), wy);
        W3<S, float>(static_cast<float>(L.dz), wz);
        const int rb = ((L.ck - t.ck0)*Tl.RY + (L.cj - t.cj0 + Tl.SY + sj))*Tl.RX + (L.ci - t.ci0);
        float u[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c=0; c<3; ++c)
#pragma unroll
          for (int bb=0; bb<3; ++bb) {
            const float wcb = wz[c]*wy[bb];
            const int rr = rb + (c*Tl.RY + bb)*Tl.RX;
#pragma unroll
            for (int a=0; a<3; ++a) {
              const float w = wcb*wx[a];
              u[0] += w*us_[rr + a];
              u[1] += w*us_[Tl.RN + rr + a];
              u[2] += w*us_[2*Tl.RN + rr + a];
            }
          }
        const float af = static_cast<float>(A.a_dt);
        const float cj32 = af/(tsq + af);
        const double fac = -static_cast<double>(mq)*ivol;
        for (int d=0; d<3; ++d) {
          const float vo = vk0[d];
          const float vn = vo + cj32*(u[d] - vo);
          A.src.V(d, q) = vn;
          const double delta = static_cast<double>(vn) - static_cast<double>(vo);
          if constexpr (REGS) A.src.R(d, q) = delta*A.inv_a_dt;
          const double imp = fac*delta;
          ex[d] += imp;
          rv[d*CH + r] = static_cast<float>(imp);
        }
        rd[r] = static_cast<float>(L.dx); rd[CH + r] = static_cast<float>(L.dy);
        rd[2*CH + r] = static_cast<float>(L.dz);
        if constexpr (SY) rcode[r] = sj + 1;
      }
      MarkChunkSpan(Tl, st, c0, c1, plan + kSpanSlot);
      __syncthreads();
      float acc[3][3][3], spl[3][3];
#pragma unroll
      for (int a=0; a<3; ++a)
#pragma unroll
        for (int n=0; n<3; ++n) {
          spl[a][n] = 0.0f;
#pragma unroll
          for (int bb=0; bb<3; ++bb) acc[bb][a][n] = 0.0f;
        }
      int spdir = 0;
      if constexpr (SPLIT) {
        float nex[1] = {0.0f};
        const auto add = [&](int r, float (&ac)[3][3][3], float (&)[1]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(slab, rd[2*CH + r]);
          const float val[3] = {rv[r], rv[CH + r], rv[2*CH + r]};
          AddSlab<float, 3, SY>(ac, spl, spdir, SY ? rcode[r] - 1 : 0, wz, wy, wx, val);
        };
        CellSums<3, 1>(Tl, st, c0, c1, l, slab, rd, RecArea(Tl, CH, 6, 3), plan, add, acc, nex);
        DepositSlabs<float, 3, SY>(Tl, dep, rd, l, slab, acc, spl, spdir);
      } else {
        float e[3] = {0.0f, 0.0f, 0.0f};
        const auto add = [&](int r, int sl, float (&ac)[3][3][3], float (&sp)[3][3], int &sd,
                             float (&)[3]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(sl, rd[2*CH + r]);
          const float val[3] = {rv[r], rv[CH + r], rv[2*CH + r]};
          AddSlab<float, 3, SY>(ac, sp, sd, SY ? rcode[r] - 1 : 0, wz, wy, wx, val);
        };
        const CellMap m = ChunkSums<3, SY>(Tl, st, c0, c1, plan + kSpanSlot, add, acc, spl,
                                                   spdir, e);
        DepositSlabs<float, 3, SY>(Tl, dep, rd, m.l, m.slab, acc, spl, spdir, m.dep,
                                      m.layers);
      }
    }
    // the flush, every thread summing (FP64) the words it writes; the defect: every thread's
    // impulses minus those words, summed over each wavefront, then over the wavefronts in order
    // (the record area is free after the last deposit's closing barrier)
    double ws[3];
    if (part == 0) FlushRegion<float, 3>(Tl, t, dep, scr, ws);
    else FlushOverflow<3>(Tl, dep, A.ov_scr + OvSlot(A.item_start, tile, part)*4*Tl.RN, ws);
    double *red = reinterpret_cast<double*>(rd);           // [NT/64][3]
    const int wv = tid >> 6;
    for (int d=0; d<3; ++d) {
      const double a = WaveSum(ex[d] - ws[d], min(64, NT - (wv << 6)));
      if ((tid & 63) == 0) red[3*wv + d] = a;
    }
    __syncthreads();
    if (tid < 3) {
      double a = 0.0;
      for (int w = 0; w < (NT + 63) >> 6; ++w) a += red[3*w + tid];
      if (part == 0) defect[3*tile + tid] = a;
      else A.ov_def[3*OvSlot(A.item_start, tile, part) + tid] = a;
    }
    if constexpr (!OV) break;
    __syncthreads();   // (the next item reuses LDS)
  }
}

template <int S, bool SY, bool OV, bool SPLIT = false>
__global__ void __launch_bounds__(kCellThreads, (SY || OV || SPLIT) ? 4 : 8)
k_p3(KArgs A, const float *us, float *scr, double *defect, const DevBlock *__restrict__ Bk) {
  P3Body<S, SY, true, OV, SPLIT>(A, us, scr, defect, Bk);
}

template <int S, bool SY, bool OV, bool SPLIT = false>
__global__ void __launch_bounds__(kCellThreads, (SY || OV || SPLIT) ? 4 : 8)
k_p3n(KArgs A, const float *us, float *scr, double *defect, const DevBlock *__restrict__ Bk) {
  P3Body<S, SY, false, OV, SPLIT>(A, us, scr, defect, Bk);
}

// pc2 predictor (dust_pm.cpp:141-193): the particles (routed in place by the key pass) moved into
// the new order, the gas velocity u = M/rho of u1 (uv, with ghosts) gathered at each particle,
// and the deposit of the rate m/(t_s V) (A.rate_dep) and of g = -dt (m/(t_s V)) (u - v) with the
// same weights: qdep components 0 (rate), 1-3 (g).  The registers x1, v1 and the rate R stay in
// the source order (the next key pass and an export read them through this sort's permutation).
// (CLOUD: refinable clouds at the mesh level, the linear gather from the 12-component pack uv)
template <int S, bool OV, bool SPLIT = false, bool CLOUD = false>
__global__ void __launch_bounds__(kCellThreads, (OV || SPLIT || CLOUD) ? 4 : 6)
k_pp(KArgs A, const float *uv, float *scr, const DevBlock *__restrict__ Bk) {
  extern __shared__ double lds_d[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  const int nlist = A.ilist ? *A.ilist_n : A.item_start[Tl.nt];
  for (int li = OV ? A.nfirst + static_cast<int>(blockIdx.x) : FirstItem(Tl, A);
       li < nlist; li += gridDim.x) {
    const int it = A.ilist ? A.ilist[li] : li;
    const int tile = A.item_tile[it];
    const int part = it - A.item_start[tile];
    const TileInfo t = TileOf(Tl, tile);
    const int tid = threadIdx.x, NT = blockDim.x, CH = kPPT*NT;
    constexpr int NR = CLOUD ? 12 : 3;
    float *dep = reinterpret_cast<float*>(lds_d);          // [4][RN] the deposit
    float *uv_ = dep + 4*Tl.RN;                            // [NR][RN] u = M/rho (or its pack)
    float *rd = uv_ + NR*Tl.RN;                            // [3][CH] offsets in the cell x, y, z
    float *rv = rd + 3*CH;                                 // [4][CH] rate, g
    int *st = reinterpret_cast<int*>(rd + RecArea(Tl, CH, 7, 4));   // [TC + 1]
    [[maybe_unused]] int *plan = st + Tl.TC + 1;           // [PlanInts(TC)]
    RegionFetch<float> uf;
    if constexpr (!CLOUD) uf.Load(g, Tl, t, uv, 3);
    for (int i = tid; i <= Tl.TC; i += NT) st[i] = A.start[t.key0 + i];
    for (int i = tid; i < 4*Tl.RN; i += NT) dep[i] = 0.0f;
    if constexpr (CLOUD) RegionLoad(g, Tl, t, uv, NR, uv_);
    __syncthreads();
    if constexpr (!CLOUD) uf.Store(Tl, 3, uv_);
    __syncthreads();
    const int P0 = min(st[0] + part*A.W, st[Tl.TC]), P1 = min(P0 + A.W, st[Tl.TC]);
    const DevBlock &b = Bk[t.m];
    const float ivolf = static_cast<float>(b.ivol);
    const float dtf = static_cast<float>(A.dt);
    const int slab = tid/Tl.TC;
    const int l = tid - slab*Tl.TC;
    for (int c0 = P0; c0 < P1; c0 += CH) {
      const int c1 = min(c0 + CH, P1);
      // the permutation of the chunk's slots, one particle ahead
      int pn = (c0 + tid < c1) ? A.perm[c0 + tid] : 0;
#pragma unroll 1
      for (int k=0; k<kPPT; ++k) {
        const int r = tid + k*NT;
        const int q = c0 + r;
        if (q >= c1) break;
        const int p = pn;
        if (k + 1 < kPPT && q + NT < c1) pn = A.perm[q + NT];
        const double x[3] = {A.src.X(0, p), A.src.X(1, p), A.src.X(2, p)};
        const float v[3] = {A.src.V(0, p), A.src.V(1, p), A.src.V(2, p)};
        const float ts = A.src.TS(p), mass = A.src.M(p);
        const int kp = A.key[p];
        const int orig = A.src.ORIG(p);
        for (int d=0; d<3; ++d) {A.dst.X(d, q) = x[d]; A.dst.V(d, q) = v[d];}
        A.dst.TS(q) = ts; A.dst.M(q) = mass; A.dst.ORIG(q) = orig;
        const Loc L = Locate<false>(g, b, x, 0.0);
        float ox = 0.0f, oy = 0.0f, oz = 0.0f;
        float val[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool xl = false;   // (refined meshes: a cross-level cloud, deposited by k_fxl_predictor)
        if constexpr (CLOUD)
          xl = A.xl_list != nullptr && CrossLevel(g, b, A.plev[orig], x, L.ci, L.cj, L.ck);
        if (TileKey(Tl, t.m, L.ci, L.cj, L.ck) != kp) {
          dk::RecordError(A.err, DK_EHALO, p);
        } else if (xl) {
          A.xl_list[atomicAdd(A.xl_n, 1)] = q;
        } else {
          ox = static_cast<float>(L.dx);
          oy = static_cast<float>(L.dy);
          oz = static_cast<float>(L.dz);
          float wx[3], wy[3], wz[3];
          W3<S, float>(ox, wx);
          W3<S, float>(oy, wy);
          W3<S, float>(oz, wz);
          const int rb = ((L.ck - t.ck0)*Tl.RY + (L.cj - t.cj0))*Tl.RX + (L.ci - t.ci0);
          float u[3] = {0.0f, 0.0f, 0.0f};
          if constexpr (CLOUD) {
            const float o[3] = {ox, oy, oz};
            CloudGather(Tl, uv_, rb, o, wx, wy, wz, u);
          } else {
#pragma unroll
            for (int c=0; c<3; ++c)
#pragma unroll
              for (int bb=0; bb<3; ++bb) {
                const float wcb = wz[c]*wy[bb];
                const int rr = rb + (c*Tl.RY + bb)*Tl.RX;
#pragma unroll
                for (int a=0; a<3; ++a) {
                  const float w = wcb*wx[a];
                  u[0] += w*uv_[rr + a];
                  u[1] += w*uv_[Tl.RN + rr + a];
                  u[2] += w*uv_[2*Tl.RN + rr + a];
                }
              }
          }
          const float rate = mass*ivolf/ts;
          val[0] = A.rate_dep ? rate : 0.0f;
          for (int d=0; d<3; ++d) val[1 + d] = -dtf*rate*(u[d] - v[d]);
        }
        rd[r] = ox; rd[CH + r] = oy; rd[2*CH + r] = oz;
        for (int n=0; n<4; ++n) rv[n*CH + r] = val[n];
      }
      MarkChunkSpan(Tl, st, c0, c1, plan + kSpanSlot);
      __syncthreads();
      float acc[3][3][4], spl[3][4];
#pragma unroll
      for (int a=0; a<3; ++a)
#pragma unroll
        for (int n=0; n<4; ++n) {
          spl[a][n] = 0.0f;
#pragma unroll
          for (int bb=0; bb<3; ++bb) acc[bb][a][n] = 0.0f;
        }
      int spdir = 0;
      if constexpr (SPLIT) {
        float nex[1] = {0.0f};
        const auto add = [&](int r, float (&ac)[3][3][4], float (&)[1]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(slab, rd[2*CH + r]);
          const float val[4] = {rv[r], rv[CH + r], rv[2*CH + r], rv[3*CH + r]};
          AddSlab<float, 4, false>(ac, spl, spdir, 0, wz, wy, wx, val);
        };
        CellSums<4, 1>(Tl, st, c0, c1, l, slab, rd, RecArea(Tl, CH, 7, 4), plan, add, acc, nex);
        DepositSlabs<float, 4, false>(Tl, dep, rd, l, slab, acc, spl, spdir);
      } else {
        float e[3] = {0.0f, 0.0f, 0.0f};
        const auto add = [&](int r, int sl, float (&ac)[3][3][4], float (&sp)[3][4], int &sd,
                             float (&)[3]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(sl, rd[2*CH + r]);
          const float val[4] = {rv[r], rv[CH + r], rv[2*CH + r], rv[3*CH + r]};
          AddSlab<float, 4, false>(ac, sp, sd, 0, wz, wy, wx, val);
        };
        const CellMap m = ChunkSums<4, false>(Tl, st, c0, c1, plan + kSpanSlot, add, acc, spl,
                                                   spdir, e);
        DepositSlabs<float, 4, false>(Tl, dep, rd, m.l, m.slab, acc, spl, spdir, m.dep,
                                      m.layers);
      }
    }
    if (part == 0) FlushRegion<float, 4>(Tl, t, dep, scr);
    else FlushOverflow<4>(Tl, dep, A.ov_scr + OvSlot(A.item_start, tile, part)*4*Tl.RN);
    if constexpr (!OV) break;
    __syncthreads();   // (the next item reuses LDS)
  }
}

// pc2 advance (dust_push.cpp:132-245, no frame terms): the particles moved into the new order,
// u* gathered at the half-step position xh (the key pass placed them by it), the implicit
// midpoint kick (1 + lambda) vh = v0 + lambda u*, lambda = dt/(2 t_s), v = 2 vh - v0,
// x = x0 + dt vh, and the deposit of the impulse -m (v - v0)/V at xh (dmom, dust_push.cpp:
// 226-244).  v - v0 = dt (u* - vh)/t_s in exact arithmetic; the impulse and R take the exact
// difference of the FP32 velocities, so that the gas receives the dust's actual change.  As P3:
// the impulse is deposited in FP32 and every tile's FP64 defect goes to the combine.
// (CLOUD: as k_pp, the linear gather from the 12-component pack us of u*)
template <int S, bool OV, bool SPLIT = false, bool CLOUD = false>
__global__ void __launch_bounds__(kCellThreads, 4)
k_pa(KArgs A, const float *us, float *scr, double *defect, const DevBlock *__restrict__ Bk) {
  extern __shared__ double lds_d[];
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  const int nlist = A.ilist ? *A.ilist_n : A.item_start[Tl.nt];
  for (int li = OV ? A.nfirst + static_cast<int>(blockIdx.x) : FirstItem(Tl, A);
       li < nlist; li += gridDim.x) {
    const int it = A.ilist ? A.ilist[li] : li;
    const int tile = A.item_tile[it];
    const int part = it - A.item_start[tile];
    const TileInfo t = TileOf(Tl, tile);
    const int tid = threadIdx.x, NT = blockDim.x, CH = kPPT*NT;
    constexpr int NR = CLOUD ? 12 : 3;
    float *dep = reinterpret_cast<float*>(lds_d);          // [3][RN] the deposit
    float *us_ = dep + 3*Tl.RN;                            // [NR][RN] u* (or its pack)
    float *rd = us_ + NR*Tl.RN;                            // [3][CH] offsets in the cell x, y, z
    float *rv = rd + 3*CH;                                 // [3][CH] -m (v - v0)/V
    int *st = reinterpret_cast<int*>(rd + RecArea(Tl, CH, 6, 3));   // [TC + 1]
    [[maybe_unused]] int *plan = st + Tl.TC + 1;           // [PlanInts(TC)]
    RegionFetch<float> uf;
    if constexpr (!CLOUD) uf.Load(g, Tl, t, us, 3);
    for (int i = tid; i <= Tl.TC; i += NT) st[i] = A.start[t.key0 + i];
    for (int i = tid; i < 3*Tl.RN; i += NT) dep[i] = 0.0f;
    if constexpr (CLOUD) RegionLoad(g, Tl, t, us, NR, us_);
    __syncthreads();
    if constexpr (!CLOUD) uf.Store(Tl, 3, us_);
    __syncthreads();
    const int P0 = min(st[0] + part*A.W, st[Tl.TC]), P1 = min(P0 + A.W, st[Tl.TC]);
    const DevBlock &b = Bk[t.m];
    const double ivol = b.ivol;
    const float hf = static_cast<float>(0.5*A.dt);
    const int slab = tid/Tl.TC;
    const int l = tid - slab*Tl.TC;
    double ex[3] = {0.0, 0.0, 0.0};
    for (int c0 = P0; c0 < P1; c0 += CH) {
      const int c1 = min(c0 + CH, P1);
      // the permutation of the chunk's slots, one particle ahead
      int pn = (c0 + tid < c1) ? A.perm[c0 + tid] : 0;
#pragma unroll 1
      for (int k=0; k<kPPT; ++k) {
        const int r = tid + k*NT;
        const int q = c0 + r;
        if (q >= c1) break;
        const int p = pn;
        if (k + 1 < kPPT && q + NT < c1) pn = A.perm[q + NT];
        double x[3] = {A.src.X(0, p), A.src.X(1, p), A.src.X(2, p)};
        float v[3] = {A.src.V(0, p), A.src.V(1, p), A.src.V(2, p)};
        const float ts = A.src.TS(p), mass = A.src.M(p);
        const int kp = A.key[p];
        const int orig = A.src.ORIG(p);
        double xh[3];
        Pc2Mid(A, x, v, xh);
        const Loc L = Locate<false>(g, b, xh, 0.0);
        float ox = 0.0f, oy = 0.0f, oz = 0.0f;
        double dv[3] = {0.0, 0.0, 0.0};
        bool xl = false;   // (refined meshes: a cross-level cloud, advanced by k_fxl_advance)
        if constexpr (CLOUD)
          xl = A.xl_list != nullptr && CrossLevel(g, b, A.plev[orig], xh, L.ci, L.cj, L.ck);
        if (TileKey(Tl, t.m, L.ci, L.cj, L.ck) != kp) {
          dk::RecordError(A.err, DK_EHALO, p);
        } else if (xl) {
          A.xl_list[atomicAdd(A.xl_n, 1)] = q;
        } else {
          ox = static_cast<float>(L.dx);
          oy = static_cast<float>(L.dy);
          oz = static_cast<float>(L.dz);
          float wx[3], wy[3], wz[3];
          W3<S, float>(ox, wx);
          W3<S, float>(oy, wy);
          W3<S, float>(oz, wz);
          const int rb = ((L.ck - t.ck0)*Tl.RY + (L.cj - t.cj0))*Tl.RX + (L.ci - t.ci0);
          float u[3] = {0.0f, 0.0f, 0.0f};
          if constexpr (CLOUD) {
            const float o[3] = {ox, oy, oz};
            CloudGather(Tl, us_, rb, o, wx, wy, wz, u);
          } else {
#pragma unroll
            for (int c=0; c<3; ++c)
#pragma unroll
              for (int bb=0; bb<3; ++bb) {
                const float wcb = wz[c]*wy[bb];
                const int rr = rb + (c*Tl.RY + bb)*Tl.RX;
#pragma unroll
                for (int a=0; a<3; ++a) {
                  const float w = wcb*wx[a];
                  u[0] += w*us_[rr + a];
                  u[1] += w*us_[Tl.RN + rr + a];
                  u[2] += w*us_[2*Tl.RN + rr + a];
                }
              }
          }
          const float lam = hf/ts;
          const float aa = 1.0f + lam;
          for (int d=0; d<3; ++d) {
            const float vo = v[d];
            const float vh = (vo + lam*u[d])/aa;
            const float vn = 2.0f*vh - vo;
            x[d] = x[d] + A.dt*static_cast<double>(vh);
            v[d] = vn;
            dv[d] = static_cast<double>(vn) - static_cast<double>(vo);
          }
        }
        for (int d=0; d<3; ++d) {
          A.dst.X(d, q) = x[d];
          A.dst.V(d, q) = v[d];
          A.dst.R(d, q) = dv[d];
        }
        A.dst.TS(q) = ts; A.dst.M(q) = mass; A.dst.ORIG(q) = orig;
        const double fac = -static_cast<double>(mass)*ivol;
        for (int d=0; d<3; ++d) {
          const double imp = fac*dv[d];
          ex[d] += imp;
          rv[d*CH + r] = static_cast<float>(imp);
        }
        rd[r] = ox; rd[CH + r] = oy; rd[2*CH + r] = oz;
      }
      MarkChunkSpan(Tl, st, c0, c1, plan + kSpanSlot);
      __syncthreads();
      float acc[3][3][3], spl[3][3];
#pragma unroll
      for (int a=0; a<3; ++a)
#pragma unroll
        for (int n=0; n<3; ++n) {
          spl[a][n] = 0.0f;
#pragma unroll
          for (int bb=0; bb<3; ++bb) acc[bb][a][n] = 0.0f;
        }
      int spdir = 0;
      if constexpr (SPLIT) {
        float nex[1] = {0.0f};
        const auto add = [&](int r, float (&ac)[3][3][3], float (&)[1]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(slab, rd[2*CH + r]);
          const float val[3] = {rv[r], rv[CH + r], rv[2*CH + r]};
          AddSlab<float, 3, false>(ac, spl, spdir, 0, wz, wy, wx, val);
        };
        CellSums<3, 1>(Tl, st, c0, c1, l, slab, rd, RecArea(Tl, CH, 6, 3), plan, add, acc, nex);
        DepositSlabs<float, 3, false>(Tl, dep, rd, l, slab, acc, spl, spdir);
      } else {
        float e[3] = {0.0f, 0.0f, 0.0f};
        const auto add = [&](int r, int sl, float (&ac)[3][3][3], float (&sp)[3][3], int &sd,
                             float (&)[3]) {
          float wx[3], wy[3];
          W3<S, float>(rd[r], wx);
          W3<S, float>(rd[CH + r], wy);
          const float wz = W1<S, float>(sl, rd[2*CH + r]);
          const float val[3] = {rv[r], rv[CH + r], rv[2*CH + r]};
          AddSlab<float, 3, false>(ac, sp, sd, 0, wz, wy, wx, val);
        };
        const CellMap m = ChunkSums<3, false>(Tl, st, c0, c1, plan + kSpanSlot, add, acc, spl,
                                                   spdir, e);
        DepositSlabs<float, 3, false>(Tl, dep, rd, m.l, m.slab, acc, spl, spdir, m.dep,
                                      m.layers);
      }
    }
    double ws[3];
    if (part == 0) FlushRegion<float, 3>(Tl, t, dep, scr, ws);
    else FlushOverflow<3>(Tl, dep, A.ov_scr + OvSlot(A.item_start, tile, part)*4*Tl.RN, ws);
    double *red = reinterpret_cast<double*>(rd);           // [NT/64][3]
    const int wv = tid >> 6;
    for (int d=0; d<3; ++d) {
      const double a = WaveSum(ex[d] - ws[d], min(64, NT - (wv << 6)));
      if ((tid & 63) == 0) red[3*wv + d] = a;
    }
    __syncthreads();
    if (tid < 3) {
      double a = 0.0;
      for (int w = 0; w < (NT + 63) >> 6; ++w) a += red[3*w + tid];
      if (part == 0) defect[3*tile + tid] = a;
      else A.ov_def[3*OvSlot(A.item_start, tile, part) + tid] = a;
    }
    if constexpr (!OV) break;
    __syncthreads();   // (the next item reuses LDS)
  }
}

// Refined meshes: the cross-level clouds the tile kernels listed (CrossLevel), in the store's
// new order, walked over the actual leaves of every level (LeafTraversal3D::ForEach,
// src/dust/dust_leaf3d.hpp:383-517, as the FP64 form walks them) with each leaf cell's own
// volume; the arithmetic of the particle is the tile kernels'.  Their deposits are atomic
// additions into the folded fields, so the order of these additions (and the last bits of the
// cells they touch) is not fixed.  A launch has a thread per particle of the store (up to the
// grid limit; the threads beyond the list's length exit at once): each cloud's walk is a chain
// of dependent loads, which only many threads in flight hide.  The walk would take 336 registers
// (one wavefront per SIMD); a bound of 3 wavefronts per SIMD spills some of them and runs
// C4's cross-level clouds 12 % faster (2 and 4: 11 % and 8 %).

// LeafApi<LeafTraversal3D>::GatherLinear (src/dust/dust_leaf_api.hpp:202-220) from the FP32
// 12-component pack of leaf b
__device__ inline void GatherLinearF(const Geo &g, const DevBlock &b, const dk::CloudGeo &cg,
                                     const float *gp, int m, int k, int j, int i, double w,
                                     double m1x, double m1y, double m1z, const double x[3],
                                     double u[3]) {
  const double xc = b.x1min + (i - g.is + 0.5)*b.dx1;
  const double yc = b.x2min + (j - g.js + 0.5)*b.dx2;
  const double zc = b.x3min + (k - g.ks + 0.5)*b.dx3;
  const double cx = 0.5*(m1x + dk::MinImage(x[0] - xc, cg.len[0])*w)/b.dx1;
  const double cy = 0.5*(m1y + dk::MinImage(x[1] - yc, cg.len[1])*w)/b.dx2;
  const double cz = 0.5*(m1z + dk::MinImage(x[2] - zc, cg.len[2])*w)/b.dx3;
#pragma unroll
  for (int c=0; c<3; ++c)
    u[c] += w*gp[Idx(g, m, 12, c, k, j, i)] + cx*gp[Idx(g, m, 12, 3 + c, k, j, i)] +
            cy*gp[Idx(g, m, 12, 6 + c, k, j, i)] + cz*gp[Idx(g, m, 12, 9 + c, k, j, i)];
}

// CloudPredictorT (src/dust/dust_cloud3d.cpp:1369-1437): the gather of the pack at x, the rate
// and the impulse -dt rate (u - v) into qdep 0 and 1-3 (k_pp's values, per leaf cell volume)
__global__ void __launch_bounds__(kBlock, 3) k_fxl_predictor(KArgs A, dk::CloudGeo cg,
                                                          const float *gp, float *qdep) {
  const Geo &g = A.g;
  const int n = *A.xl_n;
  for (int t = blockIdx.x*blockDim.x + threadIdx.x; t < n; t += gridDim.x*blockDim.x) {
    const int q = A.xl_list[t];
    const double x[3] = {A.dst.X(0, q), A.dst.X(1, q), A.dst.X(2, q)};
    const float v[3] = {A.dst.V(0, q), A.dst.V(1, q), A.dst.V(2, q)};
    const float ts = A.dst.TS(q), mass = A.dst.M(q);
    const int lev = A.plev[A.dst.ORIG(q)];
    double ug[3] = {0.0, 0.0, 0.0};
    const bool ok = dk::SmrCloudForEach<true>(g, A.B, A.tree, cg, x, lev,
        [&](int b, int k, int j, int i, double w, double m1x, double m1y, double m1z) {
          GatherLinearF(g, A.B[b], cg, gp, b, k, j, i, w, m1x, m1y, m1z, x, ug);
        });
    if (!ok) {dk::RecordError(A.err, DK_EHALO, q); continue;}
    const double mrate = static_cast<double>(mass)/static_cast<double>(ts);
    double gi[3];
    for (int d=0; d<3; ++d) gi[d] = -A.dt*mrate*(static_cast<double>(static_cast<float>(ug[d])) -
                                                 static_cast<double>(v[d]));
    dk::SmrCloudForEach<false>(g, A.B, A.tree, cg, x, lev,
        [&](int b, int k, int j, int i, double w) {
          const double wv = w*A.B[b].ivol;
          if (A.rate_dep)
            unsafeAtomicAdd(&qdep[Idx(g, b, 4, 0, k, j, i)], static_cast<float>(wv*mrate));
          for (int d=0; d<3; ++d)
            unsafeAtomicAdd(&qdep[Idx(g, b, 4, 1 + d, k, j, i)], static_cast<float>(wv*gi[d]));
        });
  }
}

// CloudAdvanceT (src/dust/dust_cloud3d.cpp:1449-1536) with CloudPC2Push (:118-158): the gather
// of u* at the half-step position, k_pa's kick and drift, R and the impulse -m (v - v0) into
// dmom (FP64, per leaf cell volume)
__global__ void __launch_bounds__(kBlock, 3) k_fxl_advance(KArgs A, dk::CloudGeo cg,
                                                        const float *gp, double *dmom) {
  const Geo &g = A.g;
  const int n = *A.xl_n;
  const float hf = static_cast<float>(0.5*A.dt);
  for (int t = blockIdx.x*blockDim.x + threadIdx.x; t < n; t += gridDim.x*blockDim.x) {
    const int q = A.xl_list[t];
    double x[3] = {A.dst.X(0, q), A.dst.X(1, q), A.dst.X(2, q)};
    float v[3] = {A.dst.V(0, q), A.dst.V(1, q), A.dst.V(2, q)};
    const float ts = A.dst.TS(q), mass = A.dst.M(q);
    const int lev = A.plev[A.dst.ORIG(q)];
    double xh[3];
    Pc2Mid(A, x, v, xh);
    double ug[3] = {0.0, 0.0, 0.0};
    const bool ok = dk::SmrCloudForEach<true>(g, A.B, A.tree, cg, xh, lev,
        [&](int b, int k, int j, int i, double w, double m1x, double m1y, double m1z) {
          GatherLinearF(g, A.B[b], cg, gp, b, k, j, i, w, m1x, m1y, m1z, xh, ug);
        });
    if (!ok) {dk::RecordError(A.err, DK_EHALO, q); continue;}
    const float lam = hf/ts;
    const float aa = 1.0f + lam;
    double dv[3];
    for (int d=0; d<3; ++d) {
      const float vo = v[d];
      const float vh = (vo + lam*static_cast<float>(ug[d]))/aa;
      const float vn = 2.0f*vh - vo;
      x[d] = x[d] + A.dt*static_cast<double>(vh);
      v[d] = vn;
      dv[d] = static_cast<double>(vn) - static_cast<double>(vo);
    }
    for (int d=0; d<3; ++d) {A.dst.X(d, q) = x[d]; A.dst.V(d, q) = v[d]; A.dst.R(d, q) = dv[d];}
    const double fac = -static_cast<double>(mass);
    dk::SmrCloudForEach<false>(g, A.B, A.tree, cg, xh, lev,
        [&](int b, int k, int j, int i, double w) {
          const double wv = fac*w*A.B[b].ivol;
          for (int d=0; d<3; ++d) unsafeAtomicAdd(&dmom[Idx(g, b, 3, d, k, j, i)], wv*dv[d]);
        });
  }
}

// the centre masses of a tile's later parts added in part order (thread per cell of the tiles
// with more than one part; every target has one writer: its own cell, the cell above, below)
__global__ void k_fix_qt(KArgs A, float *qt0, float *qt1, float *qt2) {
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  if (A.item_start[Tl.nt] == Tl.nt) return;   // every tile is one part
  GRID_STRIDE(u, (long long)Tl.nt*Tl.TC) {
    const int tile = static_cast<int>(u/Tl.TC), l = static_cast<int>(u - (long long)tile*Tl.TC);
    const int i0 = A.item_start[tile], np = A.item_start[tile + 1] - i0;
    if (np <= 1) continue;
    const TileInfo t = TileOf(Tl, tile);
    int ci, cj, ck;
    CellIdx(Tl, t, l, ci, cj, ck);
    const long long c = Idx(g, t.m, 1, 0, g.ks + ck, g.js + cj, g.is + ci);
    float a0 = qt0[c], a1 = qt1[c + g.n1], a2 = qt2[c - g.n1];
    for (int p = 1; p < np; ++p) {
      const float *o = A.ov_q + ((long long)i0 - tile + p - 1)*3*Tl.TC;
      a0 += o[l]; a1 += o[Tl.TC + l]; a2 += o[2*Tl.TC + l];
    }
    qt0[c] = a0; qt1[c + g.n1] = a1; qt2[c - g.n1] = a2;
  }
}

// the push residues and dust corrections of a tile's later parts added in part order (as
// k_fix_qt)
__global__ void k_fix_key(KArgs A, int do_key, double *pres) {
  const Geo &g = A.g;
  const Tiles &Tl = A.T;
  if (A.item_start[Tl.nt] == Tl.nt) return;   // every tile is one part
  GRID_STRIDE(u, (long long)Tl.nt*Tl.TC) {
    const int tile = static_cast<int>(u/Tl.TC), l = static_cast<int>(u - (long long)tile*Tl.TC);
    const int i0 = A.item_start[tile], np = A.item_start[tile + 1] - i0;
    if (np <= 1) continue;
    const TileInfo t = TileOf(Tl, tile);
    int ci, cj, ck;
    CellIdx(Tl, t, l, ci, cj, ck);
    const long long c = Idx(g, t.m, 3, 0, g.ks + ck, g.js + cj, g.is + ci);
    // the later parts' shares in the key pass's form (own cell plus the tile neighbours' shares
    // for it in dc0, the tile's edge rows passing theirs on in dc1 / dc2), part by part
    const int row = (l >> Tl.LX) & (Tl.TY - 1);
    for (int d=0; d<3; ++d) {
      double pr = do_key ? pres[c + d*g.ncb] : 0.0;
      double own = 0.0, up = 0.0, dn = 0.0;
      if (A.pending) {
        own = A.dc0[c + d*g.ncb];
        if (row == Tl.TY - 1) up = A.dc1[c + g.n1 + d*g.ncb];
        if (row == 0) dn = A.dc2[c - g.n1 + d*g.ncb];
      }
      for (int p = 1; p < np; ++p) {
        const long long o = (long long)i0 - tile + p - 1;
        if (do_key) pr += static_cast<double>(A.ov_pr[(o*3 + d)*Tl.TC + l]);
        if (A.pending) {
          const double *od = A.ov_dc + o*9*Tl.TC;
          own += od[(3 + d)*Tl.TC + l];
          if (row > 0) own += od[(6 + d)*Tl.TC + l - Tl.TX];
          if (row < Tl.TY - 1) own += od[d*Tl.TC + l + Tl.TX];
          if (row == Tl.TY - 1) up += od[(6 + d)*Tl.TC + l];
          if (row == 0) dn += od[d*Tl.TC + l];
        }
      }
      if (do_key) pres[c + d*g.ncb] = pr;
      if (A.pending) {
        A.dc0[c + d*g.ncb] = own;
        if (row == Tl.TY - 1) A.dc1[c + g.n1 + d*g.ncb] = up;
        if (row == 0) A.dc2[c - g.n1 + d*g.ncb] = dn;
      }
    }
    if (do_key) {
      // the stayers of each later part follow those of the parts before it: its count becomes
      // its offset, and the cell's stay count the total
      const long long k = (long long)tile*Tl.TC + l;
      int base = A.stay[k];
      for (int p = 1; p < np; ++p) {
        int &o = A.ov_cnt[((long long)i0 - tile + p - 1)*Tl.TC + l];
        const int n = o;
        o = base;
        base += n;
      }
      A.stay[k] = base;
    }
  }
}

// every padded cell sums the tile regions covering it, in ascending tile order and, within a tile,
// in part order (the later parts' regions in their overflow slots) (grid: x over the cells of a
// block, y over the blocks), in the output type TO; with defect, the first cell of every tile
// then adds the tile's defects (NV values per tile and part) in part order
template <typename T, int NV, typename TO = T>
__global__ void k_combine(Geo g, Tiles Tl, const T *scr, TO *out, const double *defect,
                          const int *item_start, const float *ov_scr, const double *ov_def) {
  const int m = blockIdx.y;
  const unsigned ncb = static_cast<unsigned>(g.ncb);
  // (any tile of more than one part: only then are the overflow slots read)
  const bool multi = item_start[Tl.nt] > Tl.nt;
  for (unsigned rr = blockIdx.x*blockDim.x + threadIdx.x; rr < ncb;
       rr += gridDim.x*blockDim.x) {
    const unsigned t2 = rr/static_cast<unsigned>(g.n1);
    const unsigned k_ = t2/static_cast<unsigned>(g.n2);
    const int i = static_cast<int>(rr - t2*g.n1);
    const int j = static_cast<int>(t2 - k_*g.n2);
    const int k = static_cast<int>(k_);
    // region coordinates relative to tile 0: r = (i - is + 1) - tx*TX etc.
    const int ai = i - g.is + 1, aj = j - g.js + 1 + Tl.SY, ak = k - g.ks + 1;
    TO sum[NV];
#pragma unroll
    for (int v=0; v<NV; ++v) sum[v] = TO(0);
    if (ai >= 0 && aj >= 0 && ak >= 0) {
      // (arithmetic shifts: a negative lower bound clamps to 0 either way)
      const int tz0 = max(0, (ak - Tl.RZ + Tl.TZ) >> Tl.LZ), tz1 = min(Tl.ntz - 1, ak >> Tl.LZ);
      const int ty0 = max(0, (aj - Tl.RY + Tl.TY) >> Tl.LY), ty1 = min(Tl.nty - 1, aj >> Tl.LY);
      const int tx0 = max(0, (ai - Tl.RX + Tl.TX) >> Tl.LX), tx1 = min(Tl.ntx - 1, ai >> Tl.LX);
      for (int tz=tz0; tz<=tz1; ++tz) {
        const int rz = ak - tz*Tl.TZ;
        if (rz < 0 || rz >= Tl.RZ) continue;
        for (int ty=ty0; ty<=ty1; ++ty) {
          const int ry = aj - ty*Tl.TY;
          if (ry < 0 || ry >= Tl.RY) continue;
          for (int tx=tx0; tx<=tx1; ++tx) {
            const int rx = ai - tx*Tl.TX;
            if (rx < 0 || rx >= Tl.RX) continue;
            const long long base = ScrRow(Tl, m, tz, ty, NV, 0, rz, ry) + tx*Tl.RX + rx;
            const long long vstride = (long long)Tl.RZ*Tl.RY*Tl.ntx*Tl.RX;
#pragma unroll
            for (int v=0; v<NV; ++v) sum[v] += static_cast<TO>(scr[base + v*vstride]);
            if (!multi) continue;
            const int tile = (m*Tl.ntz + tz)*Tl.nty*Tl.ntx + ty*Tl.ntx + tx;
            const int i0 = item_start[tile], np = item_start[tile + 1] - i0;
            for (int p = 1; p < np; ++p) {
              const float *o = ov_scr + ((long long)(i0 - tile + p - 1)*4*Tl.RN +
                                         (rz*Tl.RY + ry)*Tl.RX + rx);
#pragma unroll
              for (int v=0; v<NV; ++v) sum[v] += static_cast<TO>(o[v*Tl.RN]);
            }
          }
        }
      }
      // the tile whose first cell this is (active cells only)
      const int li = ai - 1, lj = aj - 1 - Tl.SY, lk = ak - 1;
      if (defect != nullptr && li < Tl.ntx*Tl.TX && lj >= 0 && lj < Tl.nty*Tl.TY &&
          lk < Tl.ntz*Tl.TZ && (li & (Tl.TX - 1)) == 0 && (lj & (Tl.TY - 1)) == 0 &&
          (lk & (Tl.TZ - 1)) == 0 && li >= 0 && lk >= 0) {
        const long long ct = (long long)m*Tl.ntb +
                             ((lk >> Tl.LZ)*Tl.nty + (lj >> Tl.LY))*Tl.ntx + (li >> Tl.LX);
#pragma unroll
        for (int v=0; v<NV; ++v) sum[v] += defect[NV*ct + v];
        const int i0 = item_start[ct], np = multi ? item_start[ct + 1] - i0 : 1;
        for (int p = 1; p < np; ++p)
#pragma unroll
          for (int v=0; v<NV; ++v) sum[v] += ov_def[NV*((long long)i0 - ct + p - 1) + v];
      }
    }
#pragma unroll
    for (int v=0; v<NV; ++v) out[Idx(g,m,NV,v,k,j,i)] = sum[v];
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

// the same over a 2-D grid (CellGrid: x over the active cells of a block, y over the blocks), with
// 32-bit index arithmetic
#define BLOCK_CELLS(m, k, j, i) \
  const int m = blockIdx.y; \
  const unsigned ncell_ = static_cast<unsigned>(g.nx1)*g.nx2*g.nx3; \
  for (unsigned t_ = blockIdx.x*blockDim.x + threadIdx.x; t_ < ncell_; \
       t_ += gridDim.x*blockDim.x) { \
    const unsigned q_ = t_/static_cast<unsigned>(g.nx1); \
    const unsigned k_ = q_/static_cast<unsigned>(g.nx2); \
    const int i = g.is + static_cast<int>(t_ - q_*g.nx1); \
    const int j = g.js + static_cast<int>(q_ - k_*g.nx2); \
    const int k = g.ks + static_cast<int>(k_);

// stage start: the gas share of the previous split, u0 += D - (actual dust correction),
// R_g = that/a (dust_pm.cpp:670-685); register copy u1 <- u0 (stage 1, dust_push.cpp:65);
// the harness's gas update (zero-flux RK combination + source terms); the drag history
// (stage 2, dust_pm.cpp:734-739).  pc2 (fout, before the gas update): fmode 1 the gas velocity
// M/rho of u1 (= u0 here, stage 1) that the predictor gathers (dust_pm.cpp:163-172), fmode 2
// u* = (M1 + M0)/(rho1 + rho0) that the advance gathers (dust_push.cpp:45-58)
__global__ void k_cell_start(Geo g, const DevBlock *B, double *u0, double *u1, double *dmom,
                             const double *dc0, const double *dc1, const double *dc2, int ty,
                             const double *pres,
                             int pending, double inv_aprev, int stage1, int hook, double gam0,
                             double gam1, double bdt, int use_src, dk_gas_src src,
                             double atw_dt, float *fout = nullptr, int fmode = 0) {
  BLOCK_CELLS(m, k, j, i)
    double u[4];
    for (int v=0; v<4; ++v) u[v] = u0[Idx(g,m,DK_NGAS,v,k,j,i)];
    if (fmode == 1) {
      const double irho = 1.0/u[0];
      for (int c=0; c<3; ++c) fout[Idx(g,m,3,c,k,j,i)] = static_cast<float>(u[1+c]*irho);
    } else if (fmode == 2) {
      const double irho = 1.0/(u1[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] + u[0]);
      for (int c=0; c<3; ++c)
        fout[Idx(g,m,3,c,k,j,i)] =
            static_cast<float>((u1[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] + u[1+c])*irho);
    }
    double rg[3] = {0.0, 0.0, 0.0};
    if (pending) {
      for (int c=0; c<3; ++c) {
        const long long id = Idx(g,m,3,c,k,j,i);
        // dc0: the cell's own share and its tile neighbours'; dc1 / dc2: the shares from the
        // tiles below / above (a tile's first / last x2 row only)
        double dc = dc0[id];
        if (((j - g.js) & (ty - 1)) == 0) dc += dc1[id];
        if (((j - g.js) & (ty - 1)) == ty - 1) dc += dc2[id];
        const double dgas = dmom[id] - dc;
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

// PackGatherField3D (src/dust/dust_cloud3d.cpp:1319-1340) in FP32: the velocity and its central
// differences along x1, x2, x3 of every cell with both neighbours in each direction (the
// active cells and the first ghost layer, which the tile regions read)
__global__ void k_fpack(Geo g, const float *vel, float *gp) {
  const long long ni = g.n1 - 2, nj = g.n2 - 2, nk = g.n3 - 2;
  const long long n = (long long)g.nmb*nk*nj*ni;
  GRID_STRIDE(t, n) {
    const int i = 1 + static_cast<int>(t % ni);
    long long r = t/ni;
    const int j = 1 + static_cast<int>(r % nj);
    r /= nj;
    const int k = 1 + static_cast<int>(r % nk);
    const int m = static_cast<int>(r/nk);
    for (int c=0; c<3; ++c) {
      gp[Idx(g,m,12,c,k,j,i)] = vel[Idx(g,m,3,c,k,j,i)];
      gp[Idx(g,m,12,3+c,k,j,i)] = vel[Idx(g,m,3,c,k,j,i+1)] - vel[Idx(g,m,3,c,k,j,i-1)];
      gp[Idx(g,m,12,6+c,k,j,i)] = vel[Idx(g,m,3,c,k,j+1,i)] - vel[Idx(g,m,3,c,k,j-1,i)];
      gp[Idx(g,m,12,9+c,k,j,i)] = vel[Idx(g,m,3,c,k+1,j,i)] - vel[Idx(g,m,3,c,k-1,j,i)];
    }
  }
}

// pc2 apply-predictor (dust_pm.cpp:387-402): with damping, the cells where chi = dt rate/rho
// exceeds 1 scale their g by 1/chi (kept: the commit removes the same values); u0 += g
__global__ void k_pc2_apply(Geo g, double *u0, float *qd, int damp, double dt) {
  BLOCK_CELLS(m, k, j, i)
    float q[3];
    for (int c=0; c<3; ++c) q[c] = qd[Idx(g,m,4,1+c,k,j,i)];
    if (damp) {
      const double chi = dt*static_cast<double>(qd[Idx(g,m,4,0,k,j,i)])/
                         u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)];
      if (chi > 1.0) {
        const double sc = 1.0/chi;
        for (int c=0; c<3; ++c) {
          q[c] = static_cast<float>(static_cast<double>(q[c])*sc);
          qd[Idx(g,m,4,1+c,k,j,i)] = q[c];
        }
      }
    }
    for (int c=0; c<3; ++c) u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] += static_cast<double>(q[c]);
  }
}

// pc2 commit (dust_pm.cpp:596-601): u0 += dmom - gam1 g (the predictor's share of the stage
// combination taken back, the actual impulse given)
__global__ void k_pc2_commit(Geo g, double *u0, const double *dmom, const float *qd,
                             double gam1) {
  BLOCK_CELLS(m, k, j, i)
    for (int c=0; c<3; ++c)
      u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] +=
          dmom[Idx(g,m,3,c,k,j,i)] - gam1*static_cast<double>(qd[Idx(g,m,4,1+c,k,j,i)]);
  }
}

// the local guess u* = (M + P)/(rho + Q) (dust_solver.cpp:643-649)
__global__ void k_guess(Geo g, const double *u0, const float *qd, float *us) {
  BLOCK_CELLS(m, k, j, i)
    const double pinv = 1.0/(u0[Idx(g,m,DK_NGAS,DK_IDN,k,j,i)] +
                             static_cast<double>(qd[Idx(g,m,4,0,k,j,i)]));
    for (int c=0; c<3; ++c)
      us[Idx(g,m,3,c,k,j,i)] = static_cast<float>(
          (u0[Idx(g,m,DK_NGAS,DK_IM1+c,k,j,i)] + static_cast<double>(qd[Idx(g,m,4,1+c,k,j,i)]))*pinv);
  }
}

// one defect-correction sweep x += (rho + Q)^-1 (M + P - (Ax)) (dust_solver.cpp:243-252)
__global__ void k_dc(Geo g, const double *u0, const float *qd, const float *ap, float *us) {
  BLOCK_CELLS(m, k, j, i)
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
  BLOCK_CELLS(m, k, j, i)
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
      S.X(d, p) = r[(long long)(DK_IPX + 2*d)*cap + p];
      S.V(d, p) = static_cast<float>(r[(long long)(DK_IPVX + 2*d)*cap + p]);
      S.X1(d, p) = r[(long long)(DK_IPX1 + 2*d)*cap + p];
      S.V1(d, p) = static_cast<float>(r[(long long)(DK_IPVX1 + 2*d)*cap + p]);
      S.R(d, p) = r[(long long)(DK_IPRX + d)*cap + p];
    }
    S.TS(p) = static_cast<float>(r[(long long)DK_IPTS*cap + p]);
    S.M(p) = static_cast<float>(r[(long long)DK_IPM*cap + p]);
    S.GID(p) = iv[(long long)DK_PGID*cap + p];
    S.ORIG(p) = static_cast<int>(p);
    for (int c=1; c<DK_NIDATA; ++c) itag[(long long)(c - 1)*np + p] = iv[(long long)c*cap + p];
  }
}

// the initial order: every particle's deposit cell at its present position
__global__ void k_key_static(Geo g, Tiles Tl, const DevBlock *B, Store S, long long np,
                             double sh, int *key, int *rank, int *count) {
  GRID_STRIDE(p, np) {
    const double x[3] = {S.X(0, p), S.X(1, p), S.X(2, p)};
    int ni, nj, nk, nju;
    DepositCell(g, B[S.GID(p)], x, sh, ni, nj, nk, nju);
    const long long k = TileKey(Tl, S.GID(p), ni, nj, nk);
    key[p] = static_cast<int>(k);
    rank[p] = -1 - atomicAdd(&count[k], 1);
  }
}

__global__ void k_permute_all(Store src, Store dst, const int *perm, long long np) {
  GRID_STRIDE(q, np) {
    const int p = perm[q];
    for (int d=0; d<3; ++d) {
      dst.X(d, q) = src.X(d, p); dst.X1(d, q) = src.X1(d, p);
      dst.V(d, q) = src.V(d, p); dst.V1(d, q) = src.V1(d, p); dst.R(d, q) = src.R(d, p);
    }
    dst.TS(q) = src.TS(p); dst.M(q) = src.M(p); dst.GID(q) = src.GID(p);
    dst.ORIG(q) = src.ORIG(p);
  }
}

// (regs_dead: the store keeps no registers; they are written as x1 = x, v1 = v, R = 0, the values
// the next stage 1 gives them before any use.  Otherwise x1, v1 are read where the stage-1 key
// pass wrote them: the other store O, at the slot of the last sort's permutation; r_other: R
// too (after a pc2 predictor, which leaves it there).  migrate: after a pc2 advance, whose
// positions the next key pass routes, the export applies that migration (bvals_part.cpp:164-429)
// to x, x1 and the block, as the reference's state holds them; A: its mesh numbers, with
// gam0 = 1 and gam1 = beta_dt = 0 so that PushMigrate only routes)
__global__ void k_export(KArgs A, Store S, Store O, const int *perm, const int *key, Tiles Tl,
                         long long np, const int32_t *itag, long long np0, double *r,
                         int32_t *iv, long long cap, int regs_dead, int r_other, int migrate) {
  GRID_STRIDE(q, np) {
    const int pq = perm[q];
    // the block: the store is tile-major by the keys of the last sort
    PState s;
    s.gid = key[pq]/Tl.TC/Tl.ntb;
    for (int d=0; d<3; ++d) {
      s.x[d] = S.X(d, q);
      s.v[d] = S.V(d, q);
      s.x1[d] = regs_dead ? s.x[d] : O.X1(d, pq);
      s.v1[d] = regs_dead ? s.v[d] : O.V1(d, pq);
      s.r[d] = regs_dead ? 0.0 : (r_other ? O.R(d, pq) : S.R(d, q));
      s.vres[d] = 0.0;
    }
    s.parked = 0;
    const bool ok = !migrate ||
        ((A.tree.leaf != nullptr) ? PushMigrate<false, true>(A, A.B[s.gid], s, false)
                                  : PushMigrate<false, false>(A, A.B[s.gid], s, false));
    if (!ok) dk::RecordError(A.err, DK_EROUTE, q);
    for (int d=0; d<3; ++d) {
      r[(long long)(DK_IPX + 2*d)*cap + q] = s.x[d];
      r[(long long)(DK_IPVX + 2*d)*cap + q] = static_cast<double>(s.v[d]);
      r[(long long)(DK_IPX1 + 2*d)*cap + q] = s.x1[d];
      r[(long long)(DK_IPVX1 + 2*d)*cap + q] = static_cast<double>(s.v1[d]);
      r[(long long)(DK_IPRX + d)*cap + q] = s.r[d];
    }
    r[(long long)DK_IPTS*cap + q] = static_cast<double>(S.TS(q));
    r[(long long)DK_IPM*cap + q] = static_cast<double>(S.M(q));
    iv[(long long)DK_PGID*cap + q] = s.gid;
    const int o = S.ORIG(q);
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
    const double m = S.M(p);
    const double vx = S.V(0, p), vy = S.V(1, p), vz = S.V(2, p);
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

// dynamic LDS of the tile kernels (NT = 3 TC threads, CH = kPPT NT records)
size_t LdsKey(const Tiles &T, int threads) {
  return (LdsKeyBase(T, threads) + 7)/8*8 +
         (size_t)kKeyHalves*T.TC*(9*sizeof(double) + 3*sizeof(float)) + LdsKeyWaves(T, threads);
}
size_t LdsTail(const Tiles &T) { return sizeof(int)*(T.TC + 1 + PlanInts(T.TC)); }
// (sy: the x2-shift code, whose P2/P3 records carry the shift in one more word)
size_t LdsP1(const Tiles &T, bool = true) {
  return sizeof(float)*(RecArea(T, kPPT*3*T.TC, 8, 4) + 4*T.RN + 3*T.TC) + LdsTail(T);
}
size_t LdsP2(const Tiles &T, bool sy = true) {
  return sizeof(float)*(6*T.RN + RecArea(T, kPPT*3*T.TC, sy ? 7 : 6, 3)) + LdsTail(T);
}
size_t LdsP3(const Tiles &T, bool sy = true) { return LdsP2(T, sy); }   // the same areas
// (cloud: the region holds the 12-component pack instead of the 3 velocity components)
size_t LdsPP(const Tiles &T, bool cloud = false) {
  return sizeof(float)*((cloud ? 16 : 7)*T.RN + RecArea(T, kPPT*3*T.TC, 7, 4)) + LdsTail(T);
}
size_t LdsPA(const Tiles &T, bool cloud = false) {
  return sizeof(float)*((cloud ? 15 : 6)*T.RN + RecArea(T, kPPT*3*T.TC, 6, 3)) + LdsTail(T);
}

// workgroups of the launch over the extra parts (they loop over the items; all exit at once when
// no tile has more than one part)
dim3 OverGrid(const dk_ctx *c) {
  return dim3(static_cast<unsigned>(std::min<long long>(c->fast->ov_cap, kOverGrid)));
}

KArgs ClassArgs(const dk_ctx *c, const KArgs &A0, int cls, bool &sy);

// (mid: the pc2 stage-2 form).  Without the shearing box over the two item classes of the tile
// kernels, one workgroup per item each (the lists of the last sort describe the present order);
// in the shearing box over all items, the extra parts in a looping launch.
void LaunchKey(dk_ctx *c, const KArgs &A, int do_key, double *pres, bool mid = false) {
  const Fast *f = c->fast;
  const size_t lds = LdsKey(f->T, f->key_threads);
  if (!f->T.SY) {
    bool sy;
    for (int cls = 0; cls < 2; ++cls) {
      const KArgs Ac = ClassArgs(c, A, cls, sy);
      if (mid) k_key<false, true><<<Ac.nfirst, f->key_threads, lds, c->stream>>>(Ac, do_key, pres, c->d_blocks);
      else if (c->smr)
        k_key<false, false, true><<<Ac.nfirst, f->key_threads, lds, c->stream>>>(Ac, do_key, pres, c->d_blocks);
      else k_key<false><<<Ac.nfirst, f->key_threads, lds, c->stream>>>(Ac, do_key, pres, c->d_blocks);
    }
    return;
  }
  if (mid) {
    k_key<false, true><<<f->T.nt, f->key_threads, lds, c->stream>>>(A, do_key, pres, c->d_blocks);
    k_key<true, true><<<OverGrid(c), f->key_threads, lds, c->stream>>>(A, do_key, pres, c->d_blocks);
    return;
  }
  k_key<false><<<f->T.nt, f->key_threads, lds, c->stream>>>(A, do_key, pres, c->d_blocks);
  k_key<true><<<OverGrid(c), f->key_threads, lds, c->stream>>>(A, do_key, pres, c->d_blocks);
}

// a deposit kernel of the scheme of the context with the x2-shift code SYV
#define SWITCH_S(KERNEL, SYV, OV, SPL, grid, ARGS) \
  switch (c->g.deposit) { \
    case 2: KERNEL<2, SYV, OV, SPL><<<grid, blk, lds, c->stream>>> ARGS; break; \
    case 1: KERNEL<1, SYV, OV, SPL><<<grid, blk, lds, c->stream>>> ARGS; break; \
    default: KERNEL<0, SYV, OV, SPL><<<grid, blk, lds, c->stream>>> ARGS; break; \
  }

// The tile kernels P1-P3 run over their items in two classes, each over its own list.  Without
// the shearing box: the items of single-part tiles (one workgroup each, the lean code) and the
// items of tiles of several parts, where crowded cells are (the extra-parts launch, whose code
// splits crowded cells: CellSums).  In the shearing box: the tiles inside a block's x2 range
// (their records never sit one row off in x2: the plain code, fewer registers) and the tiles on
// its x2 faces (the SY code).  Both codes give the same sums for a tile inside the range.
int TileClasses(const dk_ctx *) { return 2; }

KArgs ClassArgs(const dk_ctx *c, const KArgs &A0, int cls, bool &sy) {
  const Fast *f = c->fast;
  const Tiles &T = f->T;
  KArgs A = A0;
  A.ilist = f->ilist[cls]; A.ilist_n = f->ilist_n + cls;
  if (!T.SY) {
    // (class 1: one workgroup per item over a bound of its count, every multi-part tile having
    // at least two parts and at most n/W extra ones; the surplus workgroups exit at once)
    sy = false;
    A.nfirst = (cls == 0) ? T.nt : static_cast<int>(2*f->ov_cap);
    A.deal = (cls == 0) ? 1 : 0;
    return A;
  }
  const int nedge = (T.nt/T.ntb)*T.ntz*T.ntx*((T.nty > 1) ? 2 : 1);
  sy = (cls == 1);
  A.nfirst = sy ? nedge : T.nt - nedge;
  return A;
}

#define LAUNCH_CLASSES(KERNEL, LDS, ARGS) \
  for (int cls = 0; cls < TileClasses(c); ++cls) { \
    bool sy; \
    const KArgs A = ClassArgs(c, A0, cls, sy); \
    const size_t lds = LDS(T, sy); \
    if (!T.SY) { \
      if (cls == 0) {SWITCH_S(KERNEL, false, false, false, dim3(A.nfirst), ARGS)} \
      else {SWITCH_S(KERNEL, false, false, true, dim3(A.nfirst), ARGS)} \
      continue; \
    } \
    if (A.nfirst == 0) continue; \
    if (sy) { \
      SWITCH_S(KERNEL, true, false, false, dim3(A.nfirst), ARGS) \
      SWITCH_S(KERNEL, true, true, false, OverGrid(c), ARGS) \
    } else { \
      SWITCH_S(KERNEL, false, false, false, dim3(A.nfirst), ARGS) \
      SWITCH_S(KERNEL, false, true, false, OverGrid(c), ARGS) \
    } \
  }

void LaunchP1(dk_ctx *c, const KArgs &A0, float *qt0, float *qt1, float *qt2, float *scr) {
  const Tiles &T = c->fast->T;
  const dim3 blk(3*T.TC);
  LAUNCH_CLASSES(k_p1, LdsP1, (A, qt0, qt1, qt2, scr, c->d_blocks));
}

void LaunchP2(dk_ctx *c, const KArgs &A0, const float *us, float *scr) {
  const Tiles &T = c->fast->T;
  const dim3 blk(3*T.TC);
  LAUNCH_CLASSES(k_p2, LdsP2, (A, us, scr, c->d_blocks));
}

void LaunchP3(dk_ctx *c, const KArgs &A0, const float *us, float *scr, double *defect) {
  const Tiles &T = c->fast->T;
  const dim3 blk(3*T.TC);
  if (A0.regs_live) {LAUNCH_CLASSES(k_p3, LdsP3, (A, us, scr, defect, c->d_blocks));}
  else {LAUNCH_CLASSES(k_p3n, LdsP3, (A, us, scr, defect, c->d_blocks));}
}

// the pc2 tile kernels (one class: no shearing box), first parts then the extra parts
#define SWITCH_S2(KERNEL, OV, SPL, grid, ARGS) \
  switch (c->g.deposit) { \
    case 2: KERNEL<2, OV, SPL><<<grid, blk, lds, c->stream>>> ARGS; break; \
    case 1: KERNEL<1, OV, SPL><<<grid, blk, lds, c->stream>>> ARGS; break; \
    default: KERNEL<0, OV, SPL><<<grid, blk, lds, c->stream>>> ARGS; break; \
  }

// (the two classes of items as for P1-P3 without the shearing box; clouds: the cloud weights
// are the TSC weights, uv/us the pack)
void LaunchPP(dk_ctx *c, const KArgs &A0, const float *uv, float *scr, bool cloud = false) {
  const Tiles &T = c->fast->T;
  const dim3 blk(3*T.TC);
  const size_t lds = LdsPP(T, cloud);
  bool sy;
  KArgs A = ClassArgs(c, A0, 0, sy);
  if (cloud) k_pp<2, false, false, true><<<dim3(A.nfirst), blk, lds, c->stream>>>(A, uv, scr, c->d_blocks);
  else {SWITCH_S2(k_pp, false, false, dim3(A.nfirst), (A, uv, scr, c->d_blocks))}
  A = ClassArgs(c, A0, 1, sy);
  if (cloud) k_pp<2, false, true, true><<<dim3(A.nfirst), blk, lds, c->stream>>>(A, uv, scr, c->d_blocks);
  else {SWITCH_S2(k_pp, false, true, dim3(A.nfirst), (A, uv, scr, c->d_blocks))}
}

void LaunchPA(dk_ctx *c, const KArgs &A0, const float *us, float *scr, double *defect,
              bool cloud = false) {
  const Tiles &T = c->fast->T;
  const dim3 blk(3*T.TC);
  const size_t lds = LdsPA(T, cloud);
  bool sy;
  KArgs A = ClassArgs(c, A0, 0, sy);
  if (cloud) {
    k_pa<2, false, false, true><<<dim3(A.nfirst), blk, lds, c->stream>>>(A, us, scr, defect, c->d_blocks);
  } else {
    SWITCH_S2(k_pa, false, false, dim3(A.nfirst), (A, us, scr, defect, c->d_blocks))
  }
  A = ClassArgs(c, A0, 1, sy);
  if (cloud) {
    k_pa<2, false, true, true><<<dim3(A.nfirst), blk, lds, c->stream>>>(A, us, scr, defect, c->d_blocks);
  } else {
    SWITCH_S2(k_pa, false, true, dim3(A.nfirst), (A, us, scr, defect, c->d_blocks))
  }
}

// the cell passes: x over the active cells of a block, y over the blocks
dim3 CellGrid(const dk_ctx *c) {
  const long long b = ((long long)c->g.nx1*c->g.nx2*c->g.nx3 + kBlock - 1)/kBlock;
  return dim3(static_cast<unsigned>(std::min<long long>(b, 65535)), c->g.nmb);
}

dim3 CombineGrid(const dk_ctx *c) {
  const long long b = (c->g.ncb + kBlock - 1)/kBlock;
  return dim3(static_cast<unsigned>(std::min<long long>(b, 4096)), c->g.nmb);
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
  A.stay = f->stay;
  A.err = c->d_err;
  A.sbox = c->P.shearing_box; A.strat = c->P.stratified;
  A.shear = (c->M.bc_x1 == DK_X1_SHEAR);
  A.qs = c->P.qshear; A.om = c->P.omega0;
  if (A.sbox) {
    A.c_fx = 2.0*A.om;
    A.c_fy = -(2.0-A.qs)*A.om;
    A.c_fz = A.strat ? -(A.om*A.om) : 0.0;
    A.c_sh = A.qs*A.om;
  }
  A.lxm = c->M.x1max - c->M.x1min;
  A.lym = c->M.x2max - c->M.x2min;
  A.lzm = c->M.x3max - c->M.x3min;
  A.pending = f->pending; A.a_prev = f->a_prev;
  A.inv_a_prev = f->pending ? 1.0/f->a_prev : 0.0;
  A.regs_live = !f->regs_dead;
  A.tree = c->tree;
  if (c->smr) {
    A.plev = f->itag + (long long)(DK_PLEV - 1)*f->np;
    A.xl_list = f->xl_list; A.xl_n = f->xl_n;
  }
  A.res = f->res; A.dc0 = f->dc[0]; A.dc1 = f->dc[1]; A.dc2 = f->dc[2];
  A.item_start = f->item_start; A.item_tile = f->item_tile;
  A.W = PartChunks(f->T.SY)*kPPT*3*f->T.TC;
  A.ov_scr = f->ov_scr; A.ov_def = f->ov_def; A.ov_q = f->ov_q; A.ov_dc = f->ov_dc;
  A.ov_pr = f->ov_pr; A.ov_cnt = f->ov_cnt;
  A.x3p = f->x3p;
  A.ilist = nullptr; A.ilist_n = nullptr; A.nfirst = f->T.nt;
  A.deal = 1;
  A.cell_sums = 1;
  return A;
}

// counting sort of the store by the keys of the key pass: scan, placement, and the arrivals of
// every cell in ascending source slot after its stayers (whose ranks are in slot order): a fixed
// order
int SortKeys(dk_ctx *c) {
  Fast *f = c->fast;
  int e;
  TIMED("total", (k_total<<<Grid(f->nkeys), kBlock, 0, c->stream>>>(f->count, f->stay,
                                                                     f->nkeys)));
  TIMED("scan", e = dk::Check(hipcub::DeviceScan::ExclusiveSum(
      f->scan_tmp, f->scan_bytes, f->count, f->start, (int)(f->nkeys + 1), c->stream)));
  if (e) return e;
  TIMED("place", (k_place<<<Grid((f->np + kPlaceU - 1)/kPlaceU), kBlock, 0, c->stream>>>(
      f->key, f->rank, f->start, f->stay, f->np, f->T.TC, f->item_start, f->ov_cnt, f->perm)));
  if ((e = FillV<int>(c, f->big, 2, 0))) return e;
  TIMED("segsort", (k_segsort<<<Grid(f->nkeys), kBlock, 0, c->stream>>>(
                        f->start, f->stay, f->nkeys, f->perm, f->big)));
  TIMED("medsort", (k_medsort<<<4096, 64, 0, c->stream>>>(f->start, f->stay, f->perm, f->big)));
  TIMED("bigsort", (k_bigsort<<<256, 1024, 0, c->stream>>>(f->start, f->stay, f->perm, f->big,
                                                            f->nkeys, c->d_err)));
  return dk::Check(hipGetLastError());
}

// the work items of the present order: parts per tile, their scan, the tile of every item
int BuildItems(dk_ctx *c) {
  Fast *f = c->fast;
  const Tiles &T = f->T;
  const int W = PartChunks(T.SY)*kPPT*3*T.TC;
  TIMED("items", (k_nparts<<<Grid(T.nt + 1), kBlock, 0, c->stream>>>(T, f->start, W,
                                                                      f->nparts)));
  int e;
  TIMED("items", e = dk::Check(hipcub::DeviceScan::ExclusiveSum(
      f->iscan_tmp, f->iscan_bytes, f->nparts, f->item_start, T.nt + 1, c->stream)));
  if (e) return e;
  TIMED("items", (k_item_tile<<<Grid(T.nt), kBlock, 0, c->stream>>>(T, f->item_start,
                                                                     f->item_tile)));
  {
    TIMED("items", (k_item_flags<<<Grid(f->item_cap + 1), kBlock, 0, c->stream>>>(
        T, f->item_start, f->item_tile, f->item_cap + 1, f->iflag, T.SY ? 0 : 1)));
    TIMED("items", e = dk::Check(hipcub::DeviceScan::ExclusiveSum(
        f->lscan_tmp, f->lscan_bytes, f->iflag, f->iflag_scan, (int)(f->item_cap + 1),
        c->stream)));
    if (e) return e;
    TIMED("items", (k_item_lists<<<Grid(f->item_cap), kBlock, 0, c->stream>>>(
        T, f->item_start, f->iflag, f->iflag_scan, f->item_cap, f->ilist[0], f->ilist[1],
        f->ilist_n)));
  }
  return dk::Check(hipGetLastError());
}

// the overflow slots and item list for a store of n particles
int AllocItems(dk_ctx *c, long long n) {
  Fast *f = c->fast;
  const Tiles &T = f->T;
  const long long W = (long long)PartChunks(T.SY)*kPPT*3*T.TC;
  const long long ov = (n + W - 1)/W + 1;
  if (ov <= f->ov_cap && f->item_start) return DK_OK;
  for (void *p : {(void*)f->item_tile, (void*)f->ov_scr, (void*)f->ov_q, (void*)f->ov_pr,
                  (void*)f->ov_def, (void*)f->ov_dc, (void*)f->ov_cnt, (void*)f->ilist[0],
                  (void*)f->ilist[1], (void*)f->iflag, (void*)f->iflag_scan, f->lscan_tmp,
                  (void*)f->x3p})
    if (p) (void)hipFree(p);
  f->x3p = nullptr;
  f->item_tile = nullptr; f->ov_scr = nullptr; f->ov_q = nullptr; f->ov_pr = nullptr;
  f->ov_def = nullptr; f->ov_dc = nullptr; f->ov_cnt = nullptr;
  f->ilist[0] = nullptr; f->ilist[1] = nullptr; f->iflag = nullptr; f->iflag_scan = nullptr;
  f->lscan_tmp = nullptr;
  if (f->nparts == nullptr) {
    if (Alloc((void**)&f->nparts, sizeof(int)*(T.nt + 1))) return DK_ENOMEM;
    if (Alloc((void**)&f->item_start, sizeof(int)*(T.nt + 1))) return DK_ENOMEM;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, f->iscan_bytes, f->nparts, f->item_start,
                                         T.nt + 1) != hipSuccess)
      return DK_EHIP;
    if (Alloc(&f->iscan_tmp, std::max<size_t>(f->iscan_bytes, 16))) return DK_ENOMEM;
  }
  f->item_cap = T.nt + ov;
  f->ov_cap = ov;
  if (Alloc((void**)&f->item_tile, sizeof(int)*f->item_cap) ||
      Alloc((void**)&f->ov_scr, sizeof(float)*4*(size_t)T.RN*ov) ||
      Alloc((void**)&f->ov_q, sizeof(float)*3*(size_t)T.TC*ov) ||
      Alloc((void**)&f->ov_pr, sizeof(float)*3*(size_t)T.TC*ov) ||
      Alloc((void**)&f->ov_def, sizeof(double)*3*(size_t)ov) ||
      Alloc((void**)&f->ov_dc, sizeof(double)*9*(size_t)T.TC*ov) ||
      Alloc((void**)&f->ov_cnt, sizeof(int)*(size_t)T.TC*ov))
    return DK_ENOMEM;
  // (the items' partials, then the kX3Groups run sums of k_x3part)
  if (c->g.bc_x3 == DK_X3_OUTFLOW &&
      Alloc((void**)&f->x3p, sizeof(double)*2*((size_t)f->item_cap + kX3Groups)))
    return DK_ENOMEM;
  {
    if (f->ilist_n == nullptr && Alloc((void**)&f->ilist_n, sizeof(int)*2)) return DK_ENOMEM;
    if (Alloc((void**)&f->ilist[0], sizeof(int)*f->item_cap) ||
        Alloc((void**)&f->ilist[1], sizeof(int)*f->item_cap) ||
        Alloc((void**)&f->iflag, sizeof(int)*(f->item_cap + 1)) ||
        Alloc((void**)&f->iflag_scan, sizeof(int)*(f->item_cap + 1)))
      return DK_ENOMEM;
    f->lscan_bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, f->lscan_bytes, f->iflag, f->iflag_scan,
                                         (int)(f->item_cap + 1)) != hipSuccess)
      return DK_EHIP;
    if (Alloc(&f->lscan_tmp, std::max<size_t>(f->lscan_bytes, 16))) return DK_ENOMEM;
  }
  return DK_OK;
}

// the pending consistent-dust correction alone (before an export or a diagnostic)
int Flush(dk_ctx *c) {
  Fast *f = c->fast;
  if (!f->pending) return DK_OK;
  KArgs A = MakeArgs(c);
  TIMED("key", LaunchKey(c, A, 0, nullptr));
  TIMED("fix", (k_fix_key<<<Grid((long long)f->T.nt*f->T.TC), kBlock, 0, c->stream>>>(
      A, 0, nullptr)));
  dk_gas_src z;
  std::memset(&z, 0, sizeof(z));
  const long long nact = (long long)c->g.nmb*c->g.nx1*c->g.nx2*c->g.nx3;
  TIMED("cell_start", (k_cell_start<<<CellGrid(c), kBlock, 0, c->stream>>>(
      c->g, c->d_blocks, c->u0, c->u1, f->dmom, f->dc[0], f->dc[1], f->dc[2], f->T.TY,
      f->pres[f->pres_rd], 1,
      1.0/f->a_prev, 0, 0, 1.0, 0.0, 0.0, 0, z, 0.0)));
  f->pending = 0;
  return dk::Check(hipGetLastError());
}

}  // namespace

namespace dk {

int FastCreate(dk_ctx *c) {
  if (c->P.coupling == DK_PC2) {
    // pc2 of point particles and of clouds at the mesh level on periodic meshes without the
    // shearing box (the half-step position is keyed in the block that holds it)
    if (c->P.shearing_box || c->M.bc_x1 != DK_X1_PERIODIC || c->g.bc_x3 != DK_X3_PERIODIC)
      return DK_EUNSUPPORTED;
  } else if (c->P.coupling != DK_IMEX) {
    return DK_EUNSUPPORTED;
  } else if (c->P.solver == DK_APPLYA || c->P.solver == DK_PCG || c->P.solver == DK_ADAPTIVE) {
    return DK_EUNSUPPORTED;                                          // the FP64 form
  }
  if (c->g.ng < 2 || !c->P.back_reaction) return DK_EUNSUPPORTED;
  Fast *f = new Fast();
  c->fast = f;
  Tiles &T = f->T;
  // tile shape: 8 x 4 x 4 cells by default (or the largest powers of two dividing the block
  // below those); DK_TILE=AxBxC selects another shape whose sides are powers of two dividing the
  // block and whose cell count is at most kMaxTileCells
  T.TX = Largest(c->g.nx1, 8);
  T.TY = Largest(c->g.nx2, 4);
  T.TZ = Largest(c->g.nx3, 4);
  if (const char *e = std::getenv("DK_TILE")) {
    int a = 0, b = 0, d = 0;
    if (std::sscanf(e, "%dx%dx%d", &a, &b, &d) == 3 && a > 0 && b > 0 && d > 0 &&
        c->g.nx1 % a == 0 && c->g.nx2 % b == 0 && c->g.nx3 % d == 0 &&
        a*b*d <= kMaxTileCells && (a & (a - 1)) == 0 && (b & (b - 1)) == 0 &&
        (d & (d - 1)) == 0) {
      T.TX = a; T.TY = b; T.TZ = d;
    }
  }
  T.TC = T.TX*T.TY*T.TZ;
  T.SY = (c->P.shearing_box) ? 1 : 0;
  T.xcd = 1;
  // 3 TC threads, TC a power of 2 <= 128: a wavefront of 64 holds whole x1-x2 layers of one slab
  // (or whole slabs) unless a layer has more than 64 cells
  T.wrows = (T.TX*T.TY <= 64) ? 1 : 0;
  T.RX = T.TX + 2; T.RY = T.TY + 2 + 2*T.SY; T.RZ = T.TZ + 2;
  T.RN = T.RX*T.RY*T.RZ;
  T.LX = 0; while ((1 << T.LX) < T.TX) ++T.LX;
  T.LY = 0; while ((1 << T.LY) < T.TY) ++T.LY;
  T.LZ = 0; while ((1 << T.LZ) < T.TZ) ++T.LZ;
  T.iRX = 1.0f/static_cast<float>(T.RX);
  T.iRY = 1.0f/static_cast<float>(T.RY);
  T.iRN = 1.0f/static_cast<float>(T.RN);
  // the u* region of P2/P3 in kRegionPer values per thread
  if (3*T.RN > kRegionPer*3*T.TC) return DK_EUNSUPPORTED;
  // every tile kernel's LDS within 64 KiB
  const bool cl = c->P.clouds != 0;
  if (std::max(std::max(std::max(LdsP1(T), LdsP2(T)), LdsP3(T)),
               std::max(LdsPP(T, cl), LdsPA(T, cl))) > 65536 - 1024)
    return DK_EUNSUPPORTED;
  // a block's first and last x2 rows must differ (cells centred one row off in x2)
  if (T.SY && c->g.nx2 < 2) return DK_EUNSUPPORTED;
  T.ntx = c->g.nx1/T.TX; T.nty = c->g.nx2/T.TY; T.ntz = c->g.nx3/T.TZ;
  T.ntb = T.ntx*T.nty*T.ntz;
  T.nt = T.ntb*c->g.nmb;
  f->nkeys = (long long)c->g.nmb*c->g.nx1*c->g.nx2*c->g.nx3;
  if (f->nkeys >= INT32_MAX) return DK_ECAPACITY;
  const long long nc = NCells(c);
  int e = DK_OK;
  e = e ? e : Alloc((void**)&f->count, sizeof(int)*(f->nkeys + 1));
  e = e ? e : Alloc((void**)&f->start, sizeof(int)*(f->nkeys + 1));
  e = e ? e : Alloc((void**)&f->big, sizeof(int)*(f->nkeys + 2));
  e = e ? e : Alloc((void**)&f->stay, sizeof(int)*(f->nkeys + 1));
  e = e ? e : Alloc((void**)&f->qdep, sizeof(float)*4*nc);
  for (int q=0; q<3; ++q) e = e ? e : Alloc((void**)&f->qt[q], sizeof(float)*nc);
  e = e ? e : Alloc((void**)&f->ustar, sizeof(float)*3*nc);
  e = e ? e : Alloc((void**)&f->ap, sizeof(float)*3*nc);
  e = e ? e : Alloc((void**)&f->res, sizeof(float)*3*nc);
  e = e ? e : Alloc((void**)&f->dmom, sizeof(double)*3*nc);
  for (int q=0; q<2; ++q) e = e ? e : Alloc((void**)&f->pres[q], sizeof(double)*3*nc);
  for (int q=0; q<3; ++q) e = e ? e : Alloc((void**)&f->dc[q], sizeof(double)*3*nc);
  e = e ? e : Alloc((void**)&f->scr_f, sizeof(float)*4*(size_t)T.RN*T.nt);
  e = e ? e : Alloc((void**)&f->defect, sizeof(double)*3*(size_t)T.nt);
  if (cl) e = e ? e : Alloc((void**)&f->cpack, sizeof(float)*12*nc);
  if (c->smr) e = e ? e : Alloc((void**)&f->xl_n, sizeof(int));
  if (e) return e;
  if (cl) (void)hipMemset(f->cpack, 0, sizeof(float)*12*nc);
  (void)hipMemset(f->qt[0], 0, sizeof(float)*nc);
  (void)hipMemset(f->qt[1], 0, sizeof(float)*nc);
  (void)hipMemset(f->qt[2], 0, sizeof(float)*nc);
  for (int q=0; q<3; ++q) (void)hipMemset(f->dc[q], 0, sizeof(double)*3*nc);
  (void)hipMemset(f->dmom, 0, sizeof(double)*3*nc);
  for (int q=0; q<2; ++q) (void)hipMemset(f->pres[q], 0, sizeof(double)*3*nc);
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
  for (auto &ev : f->tev)
    if (ev) {(void)hipEventDestroy(ev); ev = nullptr;}
  for (void *p : {(void*)f->nparts, (void*)f->item_start, (void*)f->item_tile, f->iscan_tmp,
                  (void*)f->ov_scr, (void*)f->ov_q, (void*)f->ov_pr, (void*)f->ov_def,
                  (void*)f->ov_dc, (void*)f->ov_cnt, (void*)f->ilist[0], (void*)f->ilist[1],
                  (void*)f->ilist_n, (void*)f->iflag, (void*)f->iflag_scan, f->lscan_tmp,
                  (void*)f->x3p})
    if (p) (void)hipFree(p);
  for (int b=0; b<2; ++b) {
    Store &S = f->s[b];
    for (void *p : {(void*)S.d, (void*)S.f, (void*)S.i})
      if (p) (void)hipFree(p);
  }
  for (void *p : {(void*)f->key, (void*)f->rank, (void*)f->count, (void*)f->start,
                  (void*)f->perm, (void*)f->big, (void*)f->stay, f->scan_tmp, (void*)f->itag, (void*)f->qdep,
                  (void*)f->ustar, (void*)f->ap, (void*)f->res, (void*)f->dmom, (void*)f->pres[0], (void*)f->pres[1],
                  (void*)f->scr_f, (void*)f->defect, (void*)f->cpack, (void*)f->xl_list,
                  (void*)f->xl_n})
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
    const long long ng = (n + kGrp - 1)/kGrp*kGrp;
    for (int b=0; b<2; ++b) {
      Store &S = f->s[b];
      for (void **p : {(void**)&S.d, (void**)&S.f, (void**)&S.i}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
      }
      if (Alloc((void**)&S.d, sizeof(double)*9*ng)) return DK_ENOMEM;
      if (Alloc((void**)&S.f, sizeof(float)*8*ng)) return DK_ENOMEM;
      if (Alloc((void**)&S.i, sizeof(int)*2*ng)) return DK_ENOMEM;
      S.cap = ng;
    }
    for (void **p : {(void**)&f->key, (void**)&f->rank, (void**)&f->perm}) {
      if (*p) (void)hipFree(*p);
      if (Alloc(p, sizeof(int)*n)) return DK_ENOMEM;
    }
    if (c->smr) {
      if (f->xl_list) (void)hipFree(f->xl_list);
      if (Alloc((void**)&f->xl_list, sizeof(int)*n)) return DK_ENOMEM;
    }
    if (f->itag) (void)hipFree(f->itag);
    if (Alloc((void**)&f->itag, sizeof(int32_t)*(DK_NIDATA - 1)*n)) return DK_ENOMEM;
    f->cap = n;
  }
  f->np = c->np;
  f->cur = 0;
  f->pending = 0;
  f->regs_dead = 0;
  f->r_other = 0;
  f->unmigrated = 0;
  int ea;
  if ((ea = AllocItems(c, std::max<long long>(f->cap, f->np)))) return ea;
  // (with no particles the passes below run over empty sets: they still write the cell
  // offsets, start[] = 0, that every later pass reads)
  k_bind<<<Grid(f->np), kBlock, 0, c->stream>>>(c->rdata, c->idata, (long long)c->cap, f->np,
                                                 f->s[0], f->itag);
  int e;
  if ((e = FillV<int>(c, f->count, f->nkeys + 1, 0))) return e;
  if ((e = FillV<int>(c, f->stay, f->nkeys + 1, 0))) return e;
  k_key_static<<<Grid(f->np), kBlock, 0, c->stream>>>(c->g, f->T, c->d_blocks, f->s[0], f->np,
                                                      0.0, f->key, f->rank, f->count);
  if ((e = SortKeys(c))) return e;
  if ((e = BuildItems(c))) return e;
  k_permute_all<<<Grid(f->np), kBlock, 0, c->stream>>>(f->s[0], f->s[1], f->perm, f->np);
  f->cur = 1;
  return dk::Check(hipGetLastError());
}

int FastExport(dk_ctx *c) {
  Fast *f = c->fast;
  if (f == nullptr) return DK_ESTATE;
  int e;
  if ((e = Flush(c))) return e;
  KArgs A = MakeArgs(c);
  A.gam0 = 1.0; A.gam1 = 0.0; A.beta_dt = 0.0;
  k_export<<<Grid(f->np), kBlock, 0, c->stream>>>(A, f->s[f->cur], f->s[1 - f->cur], f->perm,
                                                   f->key, f->T, f->np, f->itag, f->np,
                                                   c->rdata, c->idata, (long long)c->cap,
                                                   f->regs_dead, f->r_other, f->unmigrated);
  return dk::Check(hipGetLastError());
}

// RefreshStoppingTimeMaximum over the store (dust_newdt.cpp:34-98, as the plain form's
// k_stop_partial): per workgroup the largest and smallest (FP32) stopping time, an invalid one
// recorded
__global__ void k_fstop(Store S, long long np, const int32_t *itag, double *partial,
                        dk::ErrWord *err) {
  __shared__ double smax[kBlock], smin[kBlock];
  double mx = 0.0, mn = 3.4028234663852886e38;
  GRID_STRIDE(q, np) {
    const double ts = static_cast<double>(S.TS(q));
    const int sp = itag[(long long)(DK_PSP - 1)*np + S.ORIG(q)];
    if (!isfinite(ts) || !(ts > 0.0) || sp < 0) {dk::RecordError(err, DK_ESTOP, q); continue;}
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

// NewTimeStep over the store (dust_newdt.cpp:135-145, as the plain form's k_newdt): min over
// particles of dx/max(|v|, 1e-30), the azimuthal velocity without the background shear in the
// 3-D shearing box, every stopping time validated (dust_newdt.cpp:34-99); the block of slot q
// from the last sort's key
__global__ void k_fnewdt(const DevBlock *B, Store S, const int *perm, const int *key, Tiles Tl,
                         long long np, double qo, double *partial, dk::ErrWord *err) {
  __shared__ double sh[kBlock];
  double dtmin = 3.4028234663852886e38;   // the largest float, the reference's start value
  GRID_STRIDE(q, np) {
    const int m = key[perm[q]]/Tl.TC/Tl.ntb;
    const double ts = static_cast<double>(S.TS(q));
    if (!isfinite(ts) || !(ts > 0.0)) dk::RecordError(err, DK_ESTOP, q);
    const double vy = static_cast<double>(S.V(1, q)) - qo*S.X(0, q);
    dtmin = fmin(B[m].dx1/fmax(fabs(static_cast<double>(S.V(0, q))), 1.0e-30), dtmin);
    dtmin = fmin(B[m].dx2/fmax(fabs(vy), 1.0e-30), dtmin);
    dtmin = fmin(B[m].dx3/fmax(fabs(static_cast<double>(S.V(2, q))), 1.0e-30), dtmin);
  }
  sh[threadIdx.x] = dtmin;
  __syncthreads();
  for (int s=kBlock/2; s>0; s>>=1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// the smallest particle time step over the store, after the pending correction (the velocities a
// step boundary holds); partial receives nb per-workgroup minima
int FastStoppingTimes(dk_ctx *c, double *partial, int nb) {
  Fast *f = c->fast;
  if (f == nullptr) return DK_ESTATE;
  k_fstop<<<nb, kBlock, 0, c->stream>>>(f->s[f->cur], f->np, f->itag, partial, c->d_err);
  return dk::Check(hipGetLastError());
}

int FastNewDt(dk_ctx *c, double qo, double *partial, int nb) {
  Fast *f = c->fast;
  if (f == nullptr || c->u0 == nullptr) return DK_ESTATE;
  int e;
  if ((e = Flush(c))) return e;
  k_fnewdt<<<nb, kBlock, 0, c->stream>>>(c->d_blocks, f->s[f->cur], f->perm, f->key, f->T,
                                          f->np, qo, partial, c->d_err);
  return dk::Check(hipGetLastError());
}

// the last sort's integer output (keys in the order before it, the permutation, the cell starts)
int FastSortState(dk_ctx *c, int which, int32_t *host, long long n) {
  Fast *f = c->fast;
  if (f == nullptr) return DK_ESTATE;
  const int *src = nullptr;
  long long want = 0;
  if (which == 0) {src = f->key; want = f->np;}
  else if (which == 1) {src = f->perm; want = f->np;}
  else if (which == 2) {src = f->start; want = f->nkeys + 1;}
  else return DK_EINVAL;
  if (n != want) return DK_EINVAL;
  if (n == 0) return DK_OK;
  int e = dk::Check(hipMemcpyAsync(host, src, sizeof(int)*n, hipMemcpyDeviceToHost, c->stream));
  return e ? e : dk::Check(hipStreamSynchronize(c->stream));
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

// One rk2 stage of the pc2 coupling (dust_tasks.cpp:58-109 with coupling = pc2).  Stage 1: u1 <- u0
// and the gas velocity M/rho for the predictor, the gas update, the key pass routing the last
// advance's positions (UpdateParticleGIDs of that stage, bvals_part.cpp:164-429, with x1 <- x,
// v1 <- v, dust_push.cpp:65-77), the sort, the predictor deposit, its fold and the predictor's
// application to the gas (dust_pm.cpp:387-402).  Stage 2: u* = (M1 + M0)/(rho1 + rho0), the gas
// update, the key pass placing every particle by its half-step position, the sort, the advance
// with its impulse deposit, the fold and the commit (dust_pm.cpp:596-601).
static int FastStagePc2(dk_ctx *c, const dk_stage *st, const dk_gas_src *src) {
  Fast *f = c->fast;
  const Geo &g = c->g;
  const Tiles &T = f->T;
  const int s = st->stage - 1;
  const bool s1 = (st->stage == 1);
  int e;
  KArgs A = MakeArgs(c);
  A.pending = 0;
  A.regs_live = 1;
  A.stage1 = s1 ? 1 : 0;
  // (the key pass's push with these coefficients leaves x and v as they are: it only routes)
  A.gam0 = 1.0; A.gam1 = 0.0; A.beta_dt = 0.0; A.atw_dt = 0.0;
  A.sh = 0.0; A.dt = st->dt;
  A.a_dt = st->dt; A.inv_a_dt = 1.0/st->dt;
  A.rate_dep = dk::StiffnessActive(c->P);
  A.cell_sums = 0;
  {
    // with isothermal fluxes the gas update runs after the cell pass (it reads u0's ghosts)
    const bool hydro = (src != nullptr && src->hydro);
    dk_gas_src z;
    std::memset(&z, 0, sizeof(z));
    TIMED("cell_start", (k_cell_start<<<CellGrid(c), kBlock, 0, c->stream>>>(
        g, c->d_blocks, c->u0, c->u1, f->dmom, f->dc[0], f->dc[1], f->dc[2], f->T.TY,
        f->pres[f->pres_rd], 0, 0.0, A.stage1, hydro ? 0 : 1, st->gam0[s], st->gam1[s],
        st->beta[s]*st->dt, src ? 1 : 0, src ? *src : z, 0.0, f->ustar, s1 ? 1 : 2)));
    if (hydro && (e = dk::GasHook(c, st, src))) return e;
  }
  if ((e = FillGhostsT<float>(c, f->ustar, 3, true, st->time))) return e;
  // clouds: the gather pack of the velocity (PackGatherField3D)
  const bool cl = c->P.clouds != 0;
  const float *gf = cl ? f->cpack : f->ustar;
  if (cl) {
    const Geo &gg = c->g;
    const long long nin = (long long)gg.nmb*(gg.n1 - 2)*(gg.n2 - 2)*(gg.n3 - 2);
    TIMED("cloud_pack", (k_fpack<<<Grid(nin), kBlock, 0, c->stream>>>(gg, f->ustar, f->cpack)));
  }
  if ((e = FillV<int>(c, f->count, f->nkeys + 1, 0))) return e;
  TIMED("key", LaunchKey(c, A, 1, f->pres[1 - f->pres_rd], !s1));
  TIMED("fix", (k_fix_key<<<Grid((long long)T.nt*T.TC), kBlock, 0, c->stream>>>(
      A, 1, f->pres[1 - f->pres_rd])));
  if ((e = SortKeys(c))) return e;
  if ((e = BuildItems(c))) return e;
  if (c->smr && (e = FillV<int>(c, f->xl_n, 1, 0))) return e;
  if (s1) {
    TIMED("pc2_predictor", LaunchPP(c, A, gf, f->scr_f, cl));
    TIMED("combine", (k_combine<float, 4><<<CombineGrid(c), kBlock, 0, c->stream>>>(
        g, T, f->scr_f, f->qdep, nullptr, f->item_start, f->ov_scr, f->ov_def)));
    if ((e = FoldT<float>(c, f->qdep, 4, st->time))) return e;
    if (c->smr)
      TIMED("cloud_xlevel", (k_fxl_predictor<<<Grid(f->np), kBlock, 0, c->stream>>>(
          A, dk::MakeCloudGeo(c), gf, f->qdep)));
    if (dk::StiffnessActive(c->P) && (e = StiffnessRecord<float>(c, f->qdep, 4, 0, st->dt))) return e;
    TIMED("pc2_apply", (k_pc2_apply<<<CellGrid(c), kBlock, 0, c->stream>>>(
        g, c->u0, f->qdep, c->P.pc2_damping, st->dt)));
  } else {
    TIMED("pc2_advance", LaunchPA(c, A, gf, f->scr_f, f->defect, cl));
    TIMED("combine", (k_combine<float, 3, double><<<CombineGrid(c), kBlock, 0, c->stream>>>(
        g, T, f->scr_f, f->dmom, f->defect, f->item_start, f->ov_scr, f->ov_def)));
    if ((e = FoldT<double>(c, f->dmom, 3, st->time))) return e;
    if (c->smr)
      TIMED("cloud_xlevel", (k_fxl_advance<<<Grid(f->np), kBlock, 0, c->stream>>>(
          A, dk::MakeCloudGeo(c), gf, f->dmom)));
    TIMED("pc2_commit", (k_pc2_commit<<<CellGrid(c), kBlock, 0, c->stream>>>(
        g, c->u0, f->dmom, f->qdep, st->gam1[st->nexp_stages - 1])));
  }
  f->cur = 1 - f->cur;
  f->regs_dead = 0;
  f->r_other = s1 ? 1 : 0;
  f->unmigrated = s1 ? 0 : 1;
  return dk::Check(hipGetLastError());
}

int FastStage(dk_ctx *c, const dk_stage *st, const dk_gas_src *src) {
  Fast *f = c->fast;
  if (f == nullptr || c->u0 == nullptr || f->item_start == nullptr) return DK_ESTATE;
  if (c->P.coupling == DK_PC2) return FastStagePc2(c, st, src);
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
  // the last active stage (stage 3 of a three-stage scheme is the identity)
  A.regs_live = f->keep_regs ||
                !(st->stage == st->nexp_stages || (st->nexp_stages == 3 && st->stage == 2));
  A.gam0 = st->gam0[s]; A.gam1 = st->gam1[s]; A.beta_dt = st->beta[s]*st->dt;
  A.atw_dt = (st->stage == 2) ? st->a_twid22*st->dt : 0.0;
  A.a_dt = a_dt; A.sh = sh; A.dt = st->dt;
  A.inv_a_dt = 1.0/a_dt;
  if (A.shear) {
    const double lx = c->M.x1max - c->M.x1min, ly = c->M.x2max - c->M.x2min;
    A.ysh = std::fmod(c->P.qshear*c->P.omega0*lx*st->time, ly);
    A.dsh = c->P.qshear*c->P.omega0*lx*(ShearRouteTime(st) - st->time);
  }
  // 1. key pass (with the pending correction), histogram
  if ((e = FillV<int>(c, f->count, f->nkeys + 1, 0))) return e;
  // the push of this stage writes its residue into the buffer the next stage's start reads
  TIMED("key", LaunchKey(c, A, 1, f->pres[1 - f->pres_rd]));
  if (f->x3p != nullptr)   // the outflow ledger (bvals_part.cpp:184-186) of the items just run
    TIMED("x3_ledger", (k_x3part<<<kX3Groups, 256, 0, c->stream>>>(
                            f->item_start, T.nt, f->x3p, f->x3p + 2*(size_t)f->item_cap),
                        k_x3sum<<<1, kX3Groups, 0, c->stream>>>(
                            f->x3p + 2*(size_t)f->item_cap, c->d_scalar + 12)));
  TIMED("fix", (k_fix_key<<<Grid((long long)T.nt*T.TC), kBlock, 0, c->stream>>>(
      A, 1, f->pres[1 - f->pres_rd])));
  // 2. stage start cell pass: previous split's gas share, u1 <- u0, gas update, drag history
  {
    // with isothermal fluxes the gas update and the drag history run after the cell pass, in
    // one pass (the fluxes read u0's ghosts, refilled once the pending share is in)
    const bool hydro = (src != nullptr && src->hydro);
    dk_gas_src z;
    std::memset(&z, 0, sizeof(z));
    const double gam0 = st->gam0[s], gam1 = st->gam1[s], bdt = st->beta[s]*st->dt;
    const double atw = (c->P.back_reaction && st->stage == 2) ? st->a_twid22*st->dt : 0.0;
    TIMED("cell_start", (k_cell_start<<<CellGrid(c), kBlock, 0, c->stream>>>(
        g, c->d_blocks, c->u0, c->u1, f->dmom, f->dc[0], f->dc[1], f->dc[2], f->T.TY,
        f->pres[f->pres_rd], f->pending,
        f->pending ? 1.0/f->a_prev : 0.0, A.stage1, hydro ? 0 : 1, gam0, gam1, bdt,
        src ? 1 : 0, src ? *src : z, hydro ? 0.0 : atw)));
    if (hydro && (e = dk::GasHook(c, st, src, atw != 0.0 ? f->dmom : nullptr, atw))) return e;
  }
  f->pending = 0;
  A.pending = 0;
  // 3. counting sort
  if ((e = SortKeys(c))) return e;
  if ((e = BuildItems(c))) return e;
  // 4. push into the new order + deposit Q, P, qt
  TIMED("p1_gather_deposit", LaunchP1(c, A, f->qt[0], f->qt[1], f->qt[2], f->scr_f));
  TIMED("fix", (k_fix_qt<<<Grid((long long)T.nt*T.TC), kBlock, 0, c->stream>>>(
      A, f->qt[0], f->qt[1], f->qt[2])));
  f->cur = 1 - f->cur;
  A.src = f->s[f->cur];
  A.dst = f->s[1 - f->cur];
  TIMED("combine", (k_combine<float, 4><<<CombineGrid(c), kBlock, 0, c->stream>>>(
      g, T, f->scr_f, f->qdep, nullptr, f->item_start, f->ov_scr, f->ov_def)));
  if ((e = FoldT<float>(c, f->qdep, 4, st->time))) return e;
  // 5. local guess + dc1/dc2 sweeps
  TIMED("guess", (k_guess<<<CellGrid(c), kBlock, 0, c->stream>>>(g, c->u0, f->qdep, f->ustar)));
  const int nsweeps = (c->P.solver == DK_DC2) ? 2 : ((c->P.solver == DK_DC1) ? 1 : 0);
  for (int sw=0; sw<nsweeps; ++sw) {
    if ((e = FillGhostsT<float>(c, f->ustar, 3, true, st->time))) return e;
    TIMED("p2_operator", LaunchP2(c, A, f->ustar, f->scr_f));
    TIMED("combine", (k_combine<float, 3><<<CombineGrid(c), kBlock, 0, c->stream>>>(
        g, T, f->scr_f, f->ap, nullptr, f->item_start, f->ov_scr, f->ov_def)));
    if ((e = FoldT<float>(c, f->ap, 3, st->time))) return e;
    TIMED("dc", (k_dc<<<CellGrid(c), kBlock, 0, c->stream>>>(g, c->u0, f->qdep, f->ap, f->ustar)));
  }
  if ((e = FillGhostsT<float>(c, f->ustar, 3, true, st->time))) return e;
  // 6. gather-kick + impulse deposit
  TIMED("p3_gather_kick", LaunchP3(c, A, f->ustar, f->scr_f, f->defect));
  TIMED("combine", (k_combine<float, 3, double><<<CombineGrid(c), kBlock, 0, c->stream>>>(
      g, T, f->scr_f, f->dmom, f->defect, f->item_start, f->ov_scr, f->ov_def)));
  if ((e = FoldT<double>(c, f->dmom, 3, st->time))) return e;
  // 7. consistent split: the dust's residual per unit coupled mass
  TIMED("split", (k_split<<<CellGrid(c), kBlock, 0, c->stream>>>(
      g, c->u0, f->ustar, f->qdep, f->qt[0], f->qt[1], f->qt[2], f->dmom, f->res)));
  f->pending = 1;
  f->regs_dead = !A.regs_live;
  f->pres_rd = 1 - f->pres_rd;
  f->a_prev = a_dt;
  return dk::Check(hipGetLastError());
}

}  // namespace dk

//----------------------------------------------------------------------------------------
// The autotuner.  Its choices are launch parameters that leave every result bit-identical: the
// dealing of tiles to the XCDs (runs of tiles on one XCD, so that neighbouring tiles share its
// L2, or the plain order) and the key pass's workgroup size (its per-cell sums are split by the
// slot's parity within the cell whatever the chunk, its stayers are ranked in slot order and its
// arrivals re-sorted by source slot).  The tile shape is not among them: it groups the cross-tile sums.  Tuning runs online,
// on the stages the caller runs anyway: after kTuneStages warm-up stages each candidate runs
// kTuneStages active stages, timed with events on the context's stream, in kTuneRounds rounds
// over the candidates, and the one with the fastest visit is kept (the smallest of its visits:
// a stage's time varies by about as much as the candidates differ); with DK_TUNE_CACHE naming a JSON file the choice is stored there under a key of the
// mesh, tile shape and particle count, and read back instead of tuning again.

namespace {

constexpr int kTuneStages = 2;
constexpr int kTuneRounds = 2;

std::string TuneKey(const dk_ctx *c) {
  const Fast *f = c->fast;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "fp32 nmb=%d n=%dx%dx%d tile=%dx%dx%d sy=%d np=%lld",
                c->g.nmb, c->g.nx1, c->g.nx2, c->g.nx3, f->T.TX, f->T.TY, f->T.TZ, f->T.SY,
                (long long)f->np);
  return buf;
}

bool ValidKeyThreads(const Fast *f, int kt) {
  return (kt == 128 || kt == 256 || kt == 512) && kt >= f->T.TC;
}

// the entry of this configuration in the cache file (one entry per line)
bool TuneCacheRead(const dk_ctx *c, int &xcd, int &kt) {
  const char *path = std::getenv("DK_TUNE_CACHE");
  if (path == nullptr || *path == 0) return false;
  FILE *fp = std::fopen(path, "r");
  if (fp == nullptr) return false;
  const std::string want = "\"" + TuneKey(c) + "\"";
  char line[1024];
  bool found = false;
  while (std::fgets(line, sizeof(line), fp)) {
    const char *k = std::strstr(line, want.c_str());
    if (k == nullptr) continue;
    const char *a = std::strstr(k, "\"xcd\":"), *b = std::strstr(k, "\"key_threads\":");
    if (a && b && std::sscanf(a, "\"xcd\": %d", &xcd) == 1 &&
        std::sscanf(b, "\"key_threads\": %d", &kt) == 1 && (xcd == 0 || xcd == 1) &&
        ValidKeyThreads(c->fast, kt))
      found = true;
    break;
  }
  std::fclose(fp);
  return found;
}

void TuneCacheWrite(const dk_ctx *c) {
  const char *path = std::getenv("DK_TUNE_CACHE");
  if (path == nullptr || *path == 0) return;
  const Fast *f = c->fast;
  const std::string key = "\"" + TuneKey(c) + "\"";
  std::vector<std::string> keep;
  if (FILE *fp = std::fopen(path, "r")) {
    char line[1024];
    while (std::fgets(line, sizeof(line), fp)) {
      std::string l(line);
      while (!l.empty() && (l.back() == '\n' || l.back() == ',' || l.back() == ' ')) l.pop_back();
      if (l.size() > 2 && l[0] == ' ' && l.find("\": {") != std::string::npos &&
          l.find(key) == std::string::npos)
        keep.push_back(l);
    }
    std::fclose(fp);
  }
  std::string e = "  " + key + ": {\"xcd\": " + std::to_string(f->T.xcd) +
                  ", \"key_threads\": " + std::to_string(f->key_threads) + ", \"ms\": [";
  for (int i=0; i<f->ncand; ++i) {
    char b[64];
    std::snprintf(b, sizeof(b), "%s[%d, %d, %.4f]", i ? ", " : "", f->cand_xcd[i],
                  f->cand_kt[i], f->cand_ms[i]);
    e += b;
  }
  e += "]}";
  keep.push_back(e);
  if (FILE *fp = std::fopen(path, "w")) {
    std::fputs("{\n", fp);
    for (size_t i=0; i<keep.size(); ++i)
      std::fprintf(fp, "%s%s\n", keep[i].c_str(), (i + 1 < keep.size()) ? "," : "");
    std::fputs("}\n", fp);
    std::fclose(fp);
  }
}

void ApplyCand(Fast *f, int i) {
  f->T.xcd = f->cand_xcd[i];
  f->key_threads = f->cand_kt[i];
}

}  // namespace

namespace dk {

int FastTuneBegin(dk_ctx *c, const dk_stage *st) {
  Fast *f = c->fast;
  if (f == nullptr || f->tune_state != 1 || !ActiveStage(c, st)) return DK_OK;
  if (f->tune_cand >= 0) return Check(hipEventRecord(f->tev[0], c->stream));
  return DK_OK;
}

int FastTuneEnd(dk_ctx *c, const dk_stage *st) {
  Fast *f = c->fast;
  if (f == nullptr || f->tune_state != 1 || !ActiveStage(c, st)) return DK_OK;
  int e = DK_OK;
  if (f->tune_cand >= 0) {
    float ms = 0.0f;
    e = Check(hipEventRecord(f->tev[1], c->stream));
    if (!e) e = Check(hipEventSynchronize(f->tev[1]));
    if (!e) e = Check(hipEventElapsedTime(&ms, f->tev[0], f->tev[1]));
    if (e) return e;
    f->visit_ms += ms;
  }
  if (++f->tune_stage < kTuneStages) return DK_OK;
  f->tune_stage = 0;
  if (f->tune_cand >= 0) {
    float &best_visit = f->cand_ms[f->tune_cand % f->ncand];
    if (f->tune_cand < f->ncand || f->visit_ms < best_visit) best_visit = f->visit_ms;
  }
  f->visit_ms = 0.0f;
  ++f->tune_cand;
  if (f->tune_cand < kTuneRounds*f->ncand) {
    ApplyCand(f, f->tune_cand % f->ncand);
    return DK_OK;
  }
  int best = 0;
  for (int i=1; i<f->ncand; ++i)
    if (f->cand_ms[i] < f->cand_ms[best]) best = i;
  ApplyCand(f, best);
  f->tune_state = 2;
  TuneCacheWrite(c);
  return DK_OK;
}

}  // namespace dk

extern "C" {

int dk_autotune(dk_ctx *c, int32_t mode) {
  if (c == nullptr || mode < 0 || mode > 1) return DK_EINVAL;
  Fast *f = c->fast;
  if (f == nullptr) return DK_OK;   // the plain FP64 form has nothing to tune
  if (mode == 0) {
    if (f->tune_state == 1) {   // stopped early: back to the defaults
      f->T.xcd = 1;
      f->key_threads = kKeyThreads;
    }
    f->tune_state = 0;
    return DK_OK;
  }
  int xcd = 1, kt = kKeyThreads;
  if (TuneCacheRead(c, xcd, kt)) {
    f->T.xcd = xcd;
    f->key_threads = kt;
    f->tune_state = 2;
    return DK_OK;
  }
  f->ncand = 0;
  for (int x : {1, 0})
    for (int k : {256, 128, 512})
      if (ValidKeyThreads(f, k) && f->ncand < 8) {
        f->cand_xcd[f->ncand] = x;
        f->cand_kt[f->ncand] = k;
        f->cand_ms[f->ncand] = 0.0f;
        ++f->ncand;
      }
  for (auto &ev : f->tev)
    if (ev == nullptr && hipEventCreate(&ev) != hipSuccess) return DK_EHIP;
  f->tune_state = 1;
  f->tune_cand = -1;
  f->tune_stage = 0;
  f->visit_ms = 0.0f;
  return DK_OK;
}

int dk_tune_config(dk_ctx *c, int32_t *state, int32_t *xcd, int32_t *key_threads) {
  if (c == nullptr || state == nullptr || xcd == nullptr || key_threads == nullptr)
    return DK_EINVAL;
  const Fast *f = c->fast;
  *state = f ? f->tune_state : 0;
  *xcd = f ? f->T.xcd : 0;
  *key_threads = f ? f->key_threads : 0;
  return DK_OK;
}

int dk_sort_state(dk_ctx *c, int32_t which, int32_t *host, int64_t n) {
  if (c == nullptr || host == nullptr) return DK_EINVAL;
  if (c->form != DK_FORM_FP32) return DK_EUNSUPPORTED;
  return dk::FastSortState(c, which, host, n);
}

int dk_keep_registers(dk_ctx *c, int32_t on) {
  if (c == nullptr || (on != 0 && on != 1)) return DK_EINVAL;
  if (c->fast != nullptr) c->fast->keep_regs = on;   // (the plain form always keeps them)
  return DK_OK;
}

int dk_set_tune_config(dk_ctx *c, int32_t xcd, int32_t key_threads) {
  if (c == nullptr) return DK_EINVAL;
  Fast *f = c->fast;
  if (f == nullptr) return DK_ESTATE;
  if ((xcd != 0 && xcd != 1) || !ValidKeyThreads(f, key_threads)) return DK_EINVAL;
  f->T.xcd = xcd;
  f->key_threads = key_threads;
  f->tune_state = 0;
  return DK_OK;
}

}  // extern "C"
