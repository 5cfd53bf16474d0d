// dust_oracle.cpp -- CPU FP64 oracle of the dust drag coupling.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h): a restatement, in the reference's operation order,
// of the AthenaK dust drag path for point particles on a uniform mesh of MeshBlocks.  Each
// function cites the reference file:line it follows.  With one thread every particle loop
// runs in index order and deposits are plain adds (Kokkos::Serial, src/dust/dust.hpp:986-998);
// with more threads deposits use omp atomics (Kokkos::OpenMP).
//
// Pinning of each function (tests/ hold the checks):
//   PMWeights                 partition of unity, closed-form TSC/CIC values, ip ranges
//   deposit / gather          brute-force dense G and G^T, lattice uniform deposit,
//                             sum Q V = sum m c
//   local u*                  single particle closed form; BS10 deceleration
//   imex2+ stage sequence     damping ODE vs RK4 (dust_damping.cpp:404-447), momentum
//                             conservation to round-off, NSH fixed point
//   dc1 sweep                 brute-force dense solve of A on a 4^3 mesh (error decreases)
//   consistent-gas split      momentum conservation; parity unpinned beyond that (no closed
//                             form; restated from src/dust/dust_pm.cpp:624-698)
//   pc2 predictor/advance     damping ODE (second order), momentum conservation
//   ghost copy / fold         brute force against an unpadded periodic mesh
//   shear wrap                position invariants (x in box, y offset = shear offset)

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

namespace {

constexpr int NG_VARS = DK_NGAS;

inline double SQR(double x) { return x*x; }

// src/dust/dust.hpp:1088-1111
inline void PMWeights(const double x, const double xmin, const double xmax, const int nx,
                      const int is, const int scheme, int &ip, double w[3]) {
  double cell = (x - xmin)*(static_cast<double>(nx)/(xmax - xmin));
  int ig = static_cast<int>(cell + 1.0) - 1;
  double del = cell - (static_cast<double>(ig) + 0.5);
  ip = ig + is;
  if (scheme == 0) {
    w[0] = 0.0; w[1] = 1.0; w[2] = 0.0;
  } else if (scheme == 1) {
    w[0] = std::fmax(0.0, -del);
    w[1] = 1.0 - std::fabs(del);
    w[2] = std::fmax(0.0, del);
  } else {
    w[0] = 0.5*SQR(0.5 - del);
    w[1] = 0.75 - SQR(del);
    w[2] = 0.5*SQR(0.5 + del);
  }
}

// src/pgen/dust_amr/dust_setup.hpp:35-50
inline double HashUniform01(std::uint64_t key) {
  key += UINT64_C(0x9e3779b97f4a7c15);
  key = (key ^ (key >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
  key = (key ^ (key >> 27)) * UINT64_C(0x94d049bb133111eb);
  key ^= key >> 31;
  return static_cast<double>(key >> 11) * (1.0/9007199254740992.0);
}

inline double Gauss(std::uint64_t k) {
  const double u1 = std::fmax(HashUniform01(k), 1.0e-300);
  const double u2 = HashUniform01(k + UINT64_C(0x1d8e4e27c47d124f));
  return std::sqrt(-2.0*std::log(u1))*std::cos(2.0*M_PI*u2);
}

// src/pgen/dust_amr/dust_setup.hpp:58-64
inline void NSHDustDrift(double omega0, double qshear, double ts, double ux, double uphi,
                         double &vx, double &vphi) {
  const double t = omega0*ts, b = 2.0 - qshear;
  vx = (ux + 2.0*t*uphi)/(1.0 + 2.0*b*t*t);
  vphi = uphi - b*t*vx;
}

// src/mesh/nghbr_index.hpp:29-54: the slot of a same-level neighbour (n1 = n2 = 0)
inline int NeighborIndex(int ix, int iy, int iz) {
  if (iz == 0) {
    if (ix*iy == 0) return std::abs(ix)*2*(ix + 1) + std::abs(iy)*2*(iy + 5);
    return 16 + (ix + 1) + 2*(iy + 1);
  }
  if (ix*iy == 0) return 24 + std::abs(ix)*(ix + 9) + std::abs(iy)*(iy + 17) + 2*(iz + 1);
  return 48 + (ix + 1)/2 + (iy + 1) + 2*(iz + 1);
}

inline int Dir(int ox1, int ox2, int ox3) { return (ox1 + 1) + 3*(ox2 + 1) + 9*(ox3 + 1); }

}  // namespace

struct or_sim {
  dk_mesh M;
  dk_physics P;
  std::vector<dk_block> B;
  int ng, nx1, nx2, nx3, n1, n2, n3;
  int is, ie, js, je, ks, ke;
  int64_t ncb;   // cells per block and variable, ghosts included
  std::vector<double> pr;
  std::vector<int32_t> pi;
  int64_t cap = 0, np = 0;
  std::vector<double> u0, u1, qdep, ustar, dmom, ap, resid;
  // the strict solver's fields and parameters (src/dust/dust.cpp:467-474 defaults) and the last
  // solve's figures: iterations, fast accept, true residual, error bound, target, state scale,
  // max Q/rho
  std::vector<double> sr, sp;
  double drag_rtol = 1.0e-11, drag_atol = 1.0e-14;
  int drag_iter_max = 200;
  double adaptive_order_c = 0.25, adaptive_rtol_max = 1.0e-3, adaptive_tref = 1.0,
         adaptive_state_floor = 1.0;
  double solver_last[7] = {0, 0, -1, -1, -1, -1, -1};
  // UpdateStiffness records (dust_pm.cpp:288-342): chi of the last pc2 cycle, its running max
  double stiff_chi_last = 0.0, stiff_chi_run_max = 0.0, stiff_rate_max = 0.0;
  std::vector<double> dx1, dx2, dx3;   // per block (meshblock.cpp:113-118)
  int nthreads = 1;
  int err = 0;
  int64_t err_p = -1;
  // the x3 outflow ledger since the last removal (bvals_part.cpp:184-186) and the run totals of
  // RemoveOutflowParticles (dust_tasks.cpp:407-408)
  double x3_removed_mass = 0.0, x3_removed_count = 0.0;
  double x3_outflow_mass = 0.0, x3_outflow_count = 0.0;
  // refinable clouds: the root cell size (the cloud half-width at the mesh level) and the
  // packed gather field (m, k, j, i, 12): u, and its central differences in x1, x2, x3
  double root_dx[3] = {0.0, 0.0, 0.0}, iroot_dx[3] = {0.0, 0.0, 0.0};
  std::vector<double> gpack;
  // fold order: the same-level neighbour slots in NeighborIndex order
  std::vector<int> slot_dir;   // dir index of each used slot, ordered
  // a statically refined mesh: every block a leaf of level 0..maxlev; F[d] finest-level block
  // slots per direction, leafmap the leaf covering each slot (x1 fastest)
  bool smr = false;
  int maxlev = 0;
  int F[3] = {0, 0, 0};
  std::vector<int> leafmap;

  int64_t Idx(int m, int nvar, int v, int k, int j, int i) const {
    return ((((int64_t)m*nvar + v)*n3 + k)*n2 + j)*(int64_t)n1 + i;
  }
  double &R(int c, int64_t p) { return pr[(int64_t)c*cap + p]; }
  int32_t &I(int c, int64_t p) { return pi[(int64_t)c*cap + p]; }
  void Fail(int code, int64_t p) {
#pragma omp critical(or_err)
    {
      if (err == 0) {err = code; err_p = p;}
    }
  }
};

namespace {

// deposit add: plain on one thread, atomic otherwise (src/dust/dust.hpp:986-998)
inline void DepositAdd(or_sim *s, double *dst, double v) {
  if (s->nthreads <= 1) {
    *dst += v;
  } else {
#pragma omp atomic
    *dst += v;
  }
}

bool ActiveStage(const or_sim *s, const dk_stage *st) {   // src/dust/dust.cpp:1646-1649
  if (s->P.coupling == DK_PC2) return true;
  return !(st->nexp_stages == 3 && st->stage == st->nexp_stages);
}

bool MigrationActive(const or_sim *s, const dk_stage *st) {   // src/dust/dust.cpp:1651-1654
  if (s->P.coupling == DK_IMEX) return ActiveStage(s, st);
  return st->stage == 2;
}

double DragStep(const or_sim *s, const dk_stage *st) {   // src/dust/dust.cpp:1704-1707
  if (s->P.coupling == DK_PC2) return st->dt;
  return st->a_impl*st->dt;
}

// src/dust/dust_imex3d.cpp:46-50 and src/dust/dust.hpp:1699-1702
double ImplicitShearShift(const or_sim *s, const dk_stage *st) {
  if (!s->P.shearing_box || s->P.coupling != DK_IMEX) return 0.0;
  double c = 0.0;
  for (int k=0; k<st->stage; ++k) c = st->gam0[k]*c + st->beta[k];
  const double timp = st->time + c*st->dt;
  return s->P.qshear*s->P.omega0*(timp - st->time);
}

// src/dust/dust.cpp:1668-1676
double ShearRouteTime(const or_sim *s, const dk_stage *st) {
  const double t = st->time, dt = st->dt;
  if (s->P.coupling != DK_IMEX) return t + dt;
  double w = 0.0;
  for (int k = st->stage + 1; k <= st->nexp_stages; ++k) w = st->gam0[k-1]*w + st->beta[k-1];
  return t + (1.0 - w)*dt;
}

void Zero(std::vector<double> &a) { std::fill(a.begin(), a.end(), 0.0); }

// The three per-axis weight sets of particle p at position (x, y, z) in block m.
struct Stencil {
  int ip, jp, kp;
  double wx[3], wy[3], wz[3];
};

inline void MakeStencil(or_sim *s, int m, double x, double y, double z, Stencil &w) {
  const dk_block &b = s->B[m];
  PMWeights(x, b.x1min, b.x1max, s->nx1, s->is, s->P.deposit, w.ip, w.wx);
  PMWeights(y, b.x2min, b.x2max, s->nx2, s->js, s->P.deposit, w.jp, w.wy);
  PMWeights(z, b.x3min, b.x3max, s->nx3, s->ks, s->P.deposit, w.kp, w.wz);
}

inline bool HaloOK(const or_sim *s, const Stencil &w) {   // src/dust/dust_pm.cpp:241-242
  return !(w.ip < (s->is-1) || w.ip > (s->is+s->nx1) || w.jp < (s->js-1) ||
           w.jp > (s->js+s->nx2) || w.kp < (s->ks-1) || w.kp > (s->ks+s->nx3));
}

//----------------------------------------------------------------------------------------
// Refinable clouds at the mesh level (3-D, plain codes: no lobes), on a same-level mesh
// periodic in every direction.  A cloud of half-width h = root dx is traversed in the block
// that holds its centre, ghost cells included (it reaches at most one cell beyond that block):
// its ghost deposits are folded like every deposit and its gathers read filled ghosts, which
// sums what LeafTraversal3D::ForEach (src/dust/dust_leaf3d.hpp:368-517) visits on the
// neighbour blocks themselves.

// src/dust/dust.hpp:169-182 (CloudGeometry::HatCDF, HatMoment)
inline double HatCDF(const double sv) {
  if (sv <= -1.0) return 0.0;
  if (sv >= 1.0) return 1.0;
  return (sv <= 0.0) ? 0.5*SQR(1.0 + sv) : 1.0 - 0.5*SQR(1.0 - sv);
}
inline double HatMoment(const double sv) {
  if (sv <= -1.0 || sv >= 1.0) return 1.0/6.0;
  return (sv <= 0.0) ? 0.5*SQR(sv) + SQR(sv)*sv/3.0 : 0.5*SQR(sv) - SQR(sv)*sv/3.0;
}
// src/dust/dust.hpp:186-188
inline double MinImage(const double d, const double length) {
  return d - length*std::round(d/length);
}

// CloudShape1D::Shape / FaceAt / Cell of a plain cloud (src/dust/dust.hpp:240-298, lobe 0)
struct CloudAxis { double x, ell, inv, lo, hi; };
inline CloudAxis MakeAxis(const double x, const double ell, const double inv) {
  CloudAxis a;
  a.ell = ell; a.lo = x - ell; a.hi = x + ell; a.x = x; a.inv = inv;
  return a;
}
struct CloudFace { double c = 0.0, h = 0.0; };
inline CloudFace FaceAt(const CloudAxis &a, const double z, const bool moments) {
  CloudFace f;
  const double sz = (z - a.x)*a.inv;
  f.c = HatCDF(sz);
  if (moments) f.h = HatMoment(sz);
  return f;
}
inline void CloudCell(const CloudAxis &a, const CloudFace &fa, const CloudFace &fb,
                      const bool moments, double &w, double &g) {
  w = fb.c - fa.c;
  g = moments ? a.ell*(fb.h - fa.h) : 0.0;
}

// The block that holds x (block m or a same-level neighbour; x is moved by a mesh length
// across a periodic face into that block's frame); -1 when x is farther away.
int CloudAnchor(const or_sim *s, int m, double x[3]) {
  const dk_block &b = s->B[m];
  const double lo[3] = {b.x1min, b.x2min, b.x3min}, hi[3] = {b.x1max, b.x2max, b.x3max};
  int o[3];
  for (int d=0; d<3; ++d) o[d] = (x[d] < lo[d]) ? -1 : ((x[d] >= hi[d]) ? 1 : 0);
  if (o[0] == 0 && o[1] == 0 && o[2] == 0) return m;
  const int nb = b.nghbr[Dir(o[0], o[1], o[2])];
  if (nb < 0) return -1;
  const dk_block &n = s->B[nb];
  const double nlo[3] = {n.x1min, n.x2min, n.x3min}, nhi[3] = {n.x1max, n.x2max, n.x3max};
  const double len[3] = {s->M.x1max - s->M.x1min, s->M.x2max - s->M.x2min,
                         s->M.x3max - s->M.x3min};
  for (int d=0; d<3; ++d) {
    if (o[d] == 0) continue;
    if (x[d] < nlo[d]) x[d] += len[d];
    else if (x[d] >= nhi[d]) x[d] -= len[d];
  }
  return nb;
}

// LeafTraversal3D::ForEach (src/dust/dust_leaf3d.hpp:383-517) for a plain cloud at the mesh
// level centred in block m: function(k, j, i, w, m1x, m1y, m1z) (padded indices) with moments,
// function(k, j, i, w) without, in the reference's loop order (x runs of 4, then x3, x2, x1).
// Returns false when the cloud reaches beyond the first ghost layer.
template <bool kMoments, class F>
bool CloudForEach(const or_sim *s, const int m, const double x[3], const F &function) {
  const CloudAxis s1 = MakeAxis(x[0], s->root_dx[0], s->iroot_dx[0]);
  const CloudAxis s2 = MakeAxis(x[1], s->root_dx[1], s->iroot_dx[1]);
  const CloudAxis s3 = MakeAxis(x[2], s->root_dx[2], s->iroot_dx[2]);
  const dk_block &b = s->B[m];
  const double dx = s->dx1[m], dy = s->dx2[m], dz = s->dx3[m];
  const double bx0 = b.x1min, by0 = b.x2min, bz0 = b.x3min;
  const double idx = 1.0/dx, idy = 1.0/dy, idz = 1.0/dz;
  const int i0 = static_cast<int>(std::floor((s1.lo - bx0)*idx));
  const int i1 = static_cast<int>(std::ceil((s1.hi - bx0)*idx)) - 1;
  const int j0 = static_cast<int>(std::floor((s2.lo - by0)*idy));
  const int j1 = static_cast<int>(std::ceil((s2.hi - by0)*idy)) - 1;
  const int k0 = static_cast<int>(std::floor((s3.lo - bz0)*idz));
  const int k1 = static_cast<int>(std::ceil((s3.hi - bz0)*idz)) - 1;
  if (i0 < -1 || j0 < -1 || k0 < -1 || i1 > s->nx1 || j1 > s->nx2 || k1 > s->nx3) return false;
  constexpr int kRun = 4;
  for (int ic=i0; ic<=i1; ic+=kRun) {
    const int nr = (i1 - ic + 1 < kRun) ? i1 - ic + 1 : kRun;
    double w1[kRun], g1[kRun];
    CloudFace f1 = FaceAt(s1, bx0 + ic*dx, kMoments);
    for (int a=0; a<kRun; ++a) {
      w1[a] = 0.0; g1[a] = 0.0;
      if (a < nr) {
        const CloudFace fn = FaceAt(s1, bx0 + (ic + a + 1)*dx, kMoments);
        CloudCell(s1, f1, fn, kMoments, w1[a], g1[a]);
        f1 = fn;
      }
    }
    CloudFace f3 = FaceAt(s3, bz0 + k0*dz, kMoments);
    for (int k=k0; k<=k1; ++k) {
      const CloudFace fn3 = FaceAt(s3, bz0 + (k + 1)*dz, kMoments);
      double w3, g3;
      CloudCell(s3, f3, fn3, kMoments, w3, g3);
      f3 = fn3;
      if (w3 == 0.0 && (!kMoments || g3 == 0.0)) continue;
      CloudFace f2 = FaceAt(s2, by0 + j0*dy, kMoments);
      for (int j=j0; j<=j1; ++j) {
        const CloudFace fn2 = FaceAt(s2, by0 + (j + 1)*dy, kMoments);
        double w2, g2;
        CloudCell(s2, f2, fn2, kMoments, w2, g2);
        f2 = fn2;
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
            function(k + s->ks, j + s->js, ic + a + s->is, w, m1x, m1y, m1z);
          } else {
            if (w == 0.0) continue;
            function(k + s->ks, j + s->js, ic + a + s->is, w);
          }
        }
      }
    }
  }
  return true;
}

// LeafApi<LeafTraversal3D>::GatherLinear (src/dust/dust_leaf_api.hpp:202-220): the cloud
// integral of the cell's linear reconstruction (12 packed components, PackGatherField3D)
inline void GatherLinear(const or_sim *s, const int m, const int k, const int j, const int i,
                         const double w, const double m1x, const double m1y, const double m1z,
                         const double x[3], double u[3]) {
  const double dx1 = s->dx1[m], dx2 = s->dx2[m], dx3 = s->dx3[m];
  const dk_block &b = s->B[m];
  const double xc = b.x1min + (i - s->is + 0.5)*dx1;
  const double yc = b.x2min + (j - s->js + 0.5)*dx2;
  const double zc = b.x3min + (k - s->ks + 0.5)*dx3;
  const double lx = s->M.x1max - s->M.x1min, ly = s->M.x2max - s->M.x2min;
  const double lz = s->M.x3max - s->M.x3min;
  const double cx = 0.5*(m1x + MinImage(x[0] - xc, lx)*w)/dx1;
  const double cy = 0.5*(m1y + MinImage(x[1] - yc, ly)*w)/dx2;
  const double cz = 0.5*(m1z + MinImage(x[2] - zc, lz)*w)/dx3;
  const double *gp = &s->gpack[((((int64_t)m*s->n3 + k)*s->n2 + j)*s->n1 + i)*12];
  u[0] += w*gp[0] + cx*gp[3] + cy*gp[6] + cz*gp[9];
  u[1] += w*gp[1] + cx*gp[4] + cy*gp[7] + cz*gp[10];
  u[2] += w*gp[2] + cx*gp[5] + cy*gp[8] + cz*gp[11];
}

//----------------------------------------------------------------------------------------
// Statically refined meshes (every block a leaf; levels differ by one at most between
// neighbours).  A leaf is found from a position through the finest-level block grid; the
// traversal of a cloud visits the actual leaves under its support at their own resolution, as
// LeafTraversal3D::ForEach (src/dust/dust_leaf3d.hpp:383-517) does, with the candidate blocks
// taken from the leaves holding the support's 8 corners (a leaf is at least as long as any
// support in each direction, so every leaf that meets the support holds one of its corners)
// instead of the anchor's neighbour list.

// p moved by whole mesh lengths into [min, max) of the mesh; k: the lengths taken off
void WrapPoint(const or_sim *s, double p[3], int k[3]) {
  const double lo[3] = {s->M.x1min, s->M.x2min, s->M.x3min};
  const double hi[3] = {s->M.x1max, s->M.x2max, s->M.x3max};
  for (int d=0; d<3; ++d) {
    const double len = hi[d] - lo[d];
    k[d] = 0;
    while (p[d] < lo[d]) {p[d] += len; --k[d];}
    while (p[d] >= hi[d]) {p[d] -= len; ++k[d];}
  }
}

// the leaf holding p (inside the mesh): its finest block slot, then one slot over in each
// direction where p lies outside that leaf's box by round-off; -1 if none holds it
int LeafAt(const or_sim *s, const double p[3]) {
  const double lo[3] = {s->M.x1min, s->M.x2min, s->M.x3min};
  const double len[3] = {s->M.x1max - s->M.x1min, s->M.x2max - s->M.x2min,
                         s->M.x3max - s->M.x3min};
  int sl[3];
  for (int d=0; d<3; ++d) {
    sl[d] = static_cast<int>(std::floor((p[d] - lo[d])/len[d]*s->F[d]));
    sl[d] = std::min(std::max(sl[d], 0), s->F[d] - 1);
  }
  for (int it=0; it<4; ++it) {
    const int b = s->leafmap[((int64_t)sl[2]*s->F[1] + sl[1])*s->F[0] + sl[0]];
    const dk_block &B = s->B[b];
    const double bmin[3] = {B.x1min, B.x2min, B.x3min}, bmax[3] = {B.x1max, B.x2max, B.x3max};
    const int r = 1 << (s->maxlev - B.level);
    const int l[3] = {B.lx1, B.lx2, B.lx3};
    bool moved = false;
    for (int d=0; d<3; ++d) {
      if (p[d] < bmin[d] && l[d]*r > 0) {sl[d] = l[d]*r - 1; moved = true;}
      else if (p[d] >= bmax[d] && (l[d] + 1)*r < s->F[d]) {sl[d] = (l[d] + 1)*r; moved = true;}
    }
    if (!moved) return b;
  }
  return -1;
}

// SetNewPrtclGID's choice of the neighbour slot (src/bvals/bvals_part.cpp:308-390) on a
// statically refined mesh: the leaf holding the same-level region beside block m in the
// direction (ix, iy, iz) of the crossing when it is of m's level or coarser (a coarser leaf
// also covers an empty edge or corner slot, coarse_cover), else the finer leaf beside that
// face, edge or corner in the half of m that x occupies along the other directions (the
// sub-slot of fx, fy, fz, :224-229); -1 when none (a mesh that breaks the 2:1 balance)
int SmrDest(const or_sim *s, const int m, const int off[3], const double x[3]) {
  const dk_block &b = s->B[m];
  const double bmin[3] = {b.x1min, b.x2min, b.x3min}, bmax[3] = {b.x1max, b.x2max, b.x3max};
  double q[3];
  int k[3];
  for (int d=0; d<3; ++d) q[d] = bmin[d] + (off[d] + 0.5)*(bmax[d] - bmin[d]);
  WrapPoint(s, q, k);
  const int n0 = LeafAt(s, q);
  if (n0 < 0 || s->B[n0].level <= b.level) return n0;
  for (int d=0; d<3; ++d) {
    const double l = bmax[d] - bmin[d];
    const int f = (x[d] < 0.5*(bmin[d] + bmax[d])) ? 0 : 1;
    q[d] = (off[d] < 0) ? bmin[d] - 0.25*l
         : ((off[d] > 0) ? bmax[d] + 0.25*l : bmin[d] + (f ? 0.75 : 0.25)*l);
  }
  WrapPoint(s, q, k);
  const int n = LeafAt(s, q);
  return (n >= 0 && s->B[n].level == b.level + 1) ? n : -1;
}

// LeafTraversal3D::ForEach for a plain cloud of level lev at x (half-width root dx 2^-lev):
// function(b, k, j, i, w, m1x, m1y, m1z) (padded indices of leaf b) with moments,
// function(b, k, j, i, w) without; per candidate leaf the reference's loop order (x runs of 4,
// then x3, x2, x1) and skips.  Candidates in the order of the support corners that found them
// (x1 fastest, low side first).  False when a corner lies in no leaf.
template <bool kMoments, class F>
bool LeafForEach(const or_sim *s, const double x[3], const int lev, const F &function) {
  CloudAxis sa[3];
  for (int d=0; d<3; ++d)
    sa[d] = MakeAxis(x[d], std::ldexp(s->root_dx[d], -lev), std::ldexp(s->iroot_dx[d], lev));
  int cb[8], cs[8][3], nc = 0;
  for (int c=0; c<8; ++c) {
    double p[3];
    int k[3];
    for (int d=0; d<3; ++d)
      p[d] = ((c >> d) & 1) ? std::nextafter(sa[d].hi, -std::numeric_limits<double>::infinity())
                            : sa[d].lo;
    WrapPoint(s, p, k);
    const int b = LeafAt(s, p);
    if (b < 0) return false;
    bool dup = false;
    for (int q=0; q<nc; ++q)
      if (cb[q] == b && cs[q][0] == k[0] && cs[q][1] == k[1] && cs[q][2] == k[2]) dup = true;
    if (dup) continue;
    cb[nc] = b; cs[nc][0] = k[0]; cs[nc][1] = k[1]; cs[nc][2] = k[2]; ++nc;
  }
  const double len[3] = {s->M.x1max - s->M.x1min, s->M.x2max - s->M.x2min,
                         s->M.x3max - s->M.x3min};
  // (a leaf shorter than the support in some direction might meet it without holding a
  // corner: refused)
  for (int q=0; q<nc; ++q) {
    const dk_block &B = s->B[cb[q]];
    if (B.x1max - B.x1min < 2.0*sa[0].ell || B.x2max - B.x2min < 2.0*sa[1].ell ||
        B.x3max - B.x3min < 2.0*sa[2].ell)
      return false;
  }
  constexpr int kRun = 4;
  for (int q=0; q<nc; ++q) {
    const int b = cb[q];
    const dk_block &B = s->B[b];
    const double dx = s->dx1[b], dy = s->dx2[b], dz = s->dx3[b];
    const double bx0 = B.x1min + cs[q][0]*len[0], by0 = B.x2min + cs[q][1]*len[1];
    const double bz0 = B.x3min + cs[q][2]*len[2];
    const double idx = 1.0/dx, idy = 1.0/dy, idz = 1.0/dz;
    int i0 = static_cast<int>(std::floor((sa[0].lo - bx0)*idx));
    int i1 = static_cast<int>(std::ceil((sa[0].hi - bx0)*idx)) - 1;
    int j0 = static_cast<int>(std::floor((sa[1].lo - by0)*idy));
    int j1 = static_cast<int>(std::ceil((sa[1].hi - by0)*idy)) - 1;
    int k0 = static_cast<int>(std::floor((sa[2].lo - bz0)*idz));
    int k1 = static_cast<int>(std::ceil((sa[2].hi - bz0)*idz)) - 1;
    i0 = std::max(i0, 0); i1 = std::min(i1, s->nx1 - 1);
    j0 = std::max(j0, 0); j1 = std::min(j1, s->nx2 - 1);
    k0 = std::max(k0, 0); k1 = std::min(k1, s->nx3 - 1);
    for (int ic=i0; ic<=i1; ic+=kRun) {
      const int nr = (i1 - ic + 1 < kRun) ? i1 - ic + 1 : kRun;
      double w1[kRun], g1[kRun];
      CloudFace f1 = FaceAt(sa[0], bx0 + ic*dx, kMoments);
      for (int a=0; a<kRun; ++a) {
        w1[a] = 0.0; g1[a] = 0.0;
        if (a < nr) {
          const CloudFace fn = FaceAt(sa[0], bx0 + (ic + a + 1)*dx, kMoments);
          CloudCell(sa[0], f1, fn, kMoments, w1[a], g1[a]);
          f1 = fn;
        }
      }
      CloudFace f3 = FaceAt(sa[2], bz0 + k0*dz, kMoments);
      for (int k=k0; k<=k1; ++k) {
        const CloudFace fn3 = FaceAt(sa[2], bz0 + (k + 1)*dz, kMoments);
        double w3, g3;
        CloudCell(sa[2], f3, fn3, kMoments, w3, g3);
        f3 = fn3;
        if (w3 == 0.0 && (!kMoments || g3 == 0.0)) continue;
        CloudFace f2 = FaceAt(sa[1], by0 + j0*dy, kMoments);
        for (int j=j0; j<=j1; ++j) {
          const CloudFace fn2 = FaceAt(sa[1], by0 + (j + 1)*dy, kMoments);
          double w2, g2;
          CloudCell(sa[1], f2, fn2, kMoments, w2, g2);
          f2 = fn2;
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
              function(b, k + s->ks, j + s->js, ic + a + s->is, w, m1x, m1y, m1z);
            } else {
              if (w == 0.0) continue;
              function(b, k + s->ks, j + s->js, ic + a + s->is, w);
            }
          }
        }
      }
    }
  }
  return true;
}

// The harness's gas ghost values on a statically refined mesh (a stand-in for the hydro
// boundary exchange of src/bvals/bvals_cc.cpp, which is not part of the drag path): every ghost
// cell takes the cell of the leaf holding its centre when that leaf is of the block's level
// (a copy) or coarser (the coarse value, injected), or the mean of the 8 cells of a finer leaf
// that make it up (the restriction of src/bvals/prolongation.cpp:104-111, same order)
void SmrGhosts(or_sim *s, std::vector<double> &a, int nvar) {
  const int nmb = s->M.nmb;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m) {
    const dk_block &b = s->B[m];
    for (int k=0; k<s->n3; ++k)
      for (int j=0; j<s->n2; ++j)
        for (int i=0; i<s->n1; ++i) {
          if (i >= s->is && i <= s->ie && j >= s->js && j <= s->je && k >= s->ks &&
              k <= s->ke)
            continue;
          double c[3] = {b.x1min + (i - s->is + 0.5)*s->dx1[m],
                         b.x2min + (j - s->js + 0.5)*s->dx2[m],
                         b.x3min + (k - s->ks + 0.5)*s->dx3[m]};
          int sh[3];
          WrapPoint(s, c, sh);
          const int n = LeafAt(s, c);
          if (n < 0) {s->Fail(DK_EROUTE, -1); continue;}
          const dk_block &nb = s->B[n];
          if (nb.level <= b.level) {
            const int in = static_cast<int>(std::floor((c[0] - nb.x1min)/s->dx1[n])) + s->is;
            const int jn = static_cast<int>(std::floor((c[1] - nb.x2min)/s->dx2[n])) + s->js;
            const int kn = static_cast<int>(std::floor((c[2] - nb.x3min)/s->dx3[n])) + s->ks;
            for (int v=0; v<nvar; ++v) a[s->Idx(m,nvar,v,k,j,i)] = a[s->Idx(n,nvar,v,kn,jn,in)];
          } else {
            const int fi = static_cast<int>(std::floor((c[0] - 0.5*s->dx1[m] - nb.x1min)/
                                                       s->dx1[n] + 0.5)) + s->is;
            const int fj = static_cast<int>(std::floor((c[1] - 0.5*s->dx2[m] - nb.x2min)/
                                                       s->dx2[n] + 0.5)) + s->js;
            const int fk = static_cast<int>(std::floor((c[2] - 0.5*s->dx3[m] - nb.x3min)/
                                                       s->dx3[n] + 0.5)) + s->ks;
            for (int v=0; v<nvar; ++v)
              a[s->Idx(m,nvar,v,k,j,i)] = 0.125*(
                  a[s->Idx(n,nvar,v,fk  ,fj  ,fi)] + a[s->Idx(n,nvar,v,fk  ,fj  ,fi+1)]
                + a[s->Idx(n,nvar,v,fk  ,fj+1,fi)] + a[s->Idx(n,nvar,v,fk  ,fj+1,fi+1)]
                + a[s->Idx(n,nvar,v,fk+1,fj  ,fi)] + a[s->Idx(n,nvar,v,fk+1,fj  ,fi+1)]
                + a[s->Idx(n,nvar,v,fk+1,fj+1,fi)] + a[s->Idx(n,nvar,v,fk+1,fj+1,fi+1)]);
          }
        }
  }
}

// src/dust/dust_cloud3d.cpp:1319-1340 (PackGatherField3D): velocity and its central
// differences of every cell with both neighbours in each direction; vel(m, c, k, j, i)
void PackGatherField(or_sim *s, const std::vector<double> &vel) {
  s->gpack.assign(s->ncb*s->M.nmb*12, 0.0);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int m=0; m<s->M.nmb; ++m)
    for (int k=1; k<s->n3-1; ++k)
      for (int j=1; j<s->n2-1; ++j)
        for (int i=1; i<s->n1-1; ++i) {
          double *gp = &s->gpack[((((int64_t)m*s->n3 + k)*s->n2 + j)*s->n1 + i)*12];
          for (int c=0; c<3; ++c) {
            gp[c] = vel[s->Idx(m,3,c,k,j,i)];
            gp[3+c] = vel[s->Idx(m,3,c,k,j,i+1)] - vel[s->Idx(m,3,c,k,j,i-1)];
            gp[6+c] = vel[s->Idx(m,3,c,k,j+1,i)] - vel[s->Idx(m,3,c,k,j-1,i)];
            gp[9+c] = vel[s->Idx(m,3,c,k+1,j,i)] - vel[s->Idx(m,3,c,k-1,j,i)];
          }
        }
}

// src/dust/dust_cloud3d.cpp:1342-1360 (PackGasGather3D): u = M/rho of every cell, then packed
void PackGasGather(or_sim *s, const std::vector<double> &u) {
  std::vector<double> vel(3*s->ncb*s->M.nmb);
  for (int m=0; m<s->M.nmb; ++m)
    for (int64_t r=0; r<s->ncb; ++r) {
      const double invrho = 1.0/u[((int64_t)m*NG_VARS + DK_IDN)*s->ncb + r];
      for (int c=0; c<3; ++c)
        vel[((int64_t)m*3 + c)*s->ncb + r] = u[((int64_t)m*NG_VARS + DK_IM1 + c)*s->ncb + r]*invrho;
    }
  PackGatherField(s, vel);
}

// src/dust/dust_cloud3d.cpp:1369-1437 (CloudPredictorT, 3-D, plain codes): the linear gather of
// the step-start gas velocity over the cloud and the drag impulse deposited with the gather's
// weights into qdep 1-3 (and 4)
void CloudPredictor(or_sim *s, const double dt, const bool rate_dep) {
  PackGasGather(s, s->u1);
  if (s->smr) {
    // the actual leaves under the cloud at their own resolution; the deposit walks the cloud
    // again without moments (the same cells and weights in the same order)
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t p=0; p<s->np; ++p) {
      const double x[3] = {s->R(DK_IPX,p), s->R(DK_IPY,p), s->R(DK_IPZ,p)};
      const int lev = s->I(DK_PLEV,p);
      double ug[3] = {0.0, 0.0, 0.0};
      const bool ok = LeafForEach<true>(s, x, lev,
          [&](int b, int k, int j, int i, double w, double m1x, double m1y, double m1z) {
            GatherLinear(s, b, k, j, i, w, m1x, m1y, m1z, x, ug);
          });
      if (!ok) {s->Fail(DK_EHALO, p); continue;}
      const double mass_rate = s->R(DK_IPM,p)/s->R(DK_IPTS,p);
      const double gx = -dt*mass_rate*(ug[0] - s->R(DK_IPVX,p));
      const double gy = -dt*mass_rate*(ug[1] - s->R(DK_IPVY,p));
      const double gz = -dt*mass_rate*(ug[2] - s->R(DK_IPVZ,p));
      LeafForEach<false>(s, x, lev, [&](int b, int k, int j, int i, double w) {
        const double inv_volume = 1.0/(s->dx1[b]*s->dx2[b]*s->dx3[b]);
        DepositAdd(s, &s->qdep[s->Idx(b,5,1,k,j,i)], inv_volume*w*gx);
        DepositAdd(s, &s->qdep[s->Idx(b,5,2,k,j,i)], inv_volume*w*gy);
        DepositAdd(s, &s->qdep[s->Idx(b,5,3,k,j,i)], inv_volume*w*gz);
        if (rate_dep) DepositAdd(s, &s->qdep[s->Idx(b,5,4,k,j,i)], inv_volume*w*mass_rate);
      });
    }
    return;
  }
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t p=0; p<s->np; ++p) {
    double x[3] = {s->R(DK_IPX,p), s->R(DK_IPY,p), s->R(DK_IPZ,p)};
    const int am = CloudAnchor(s, s->I(DK_PGID,p), x);
    double ug[3] = {0.0, 0.0, 0.0};
    int nt = 0, tk[27], tj[27], ti[27];
    double tw[27];
    const bool ok = (am >= 0) && CloudForEach<true>(s, am, x,
        [&](int k, int j, int i, double w, double m1x, double m1y, double m1z) {
          GatherLinear(s, am, k, j, i, w, m1x, m1y, m1z, x, ug);
          if (nt < 27) {tk[nt] = k; tj[nt] = j; ti[nt] = i; tw[nt] = w;}
          ++nt;
        });
    if (!ok || nt > 27) {s->Fail(DK_EHALO, p); continue;}
    const double mass_rate = s->R(DK_IPM,p)/s->R(DK_IPTS,p);
    const double gx = -dt*mass_rate*(ug[0] - s->R(DK_IPVX,p));
    const double gy = -dt*mass_rate*(ug[1] - s->R(DK_IPVY,p));
    const double gz = -dt*mass_rate*(ug[2] - s->R(DK_IPVZ,p));
    const double inv_volume = 1.0/(s->dx1[am]*s->dx2[am]*s->dx3[am]);
    for (int q=0; q<nt; ++q) {
      if (tw[q] == 0.0) continue;
      const double w = tw[q];
      DepositAdd(s, &s->qdep[s->Idx(am,5,1,tk[q],tj[q],ti[q])], inv_volume*w*gx);
      DepositAdd(s, &s->qdep[s->Idx(am,5,2,tk[q],tj[q],ti[q])], inv_volume*w*gy);
      DepositAdd(s, &s->qdep[s->Idx(am,5,3,tk[q],tj[q],ti[q])], inv_volume*w*gz);
      if (rate_dep) DepositAdd(s, &s->qdep[s->Idx(am,5,4,tk[q],tj[q],ti[q])], inv_volume*w*mass_rate);
    }
  }
}

// src/dust/dust_cloud3d.cpp:1449-1536 (CloudAdvanceT, 3-D, plain codes) with
// CloudPC2Push<LeafTraversal3D> (:118-158): the linear gather of u* at the midpoint, the
// implicit-midpoint kick and drift, the impulse deposited into dmom with the gather's weights
void CloudAdvance(or_sim *s, const double dt) {
  const bool br = s->P.back_reaction;
  const bool sbox = s->P.shearing_box, strat = s->P.stratified;
  const double qs = s->P.qshear, om = s->P.omega0;
  std::vector<double> vel(s->ustar);
  PackGatherField(s, vel);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t p=0; p<s->np; ++p) {
    const double x0[3] = {s->R(DK_IPX1,p), s->R(DK_IPY1,p), s->R(DK_IPZ1,p)};
    const double v0[3] = {s->R(DK_IPVX1,p), s->R(DK_IPVY1,p), s->R(DK_IPVZ1,p)};
    double xh[3], xg[3];
    xh[0] = x0[0] + 0.5*dt*v0[0];
    xh[1] = x0[1] + 0.5*dt*(v0[1] - (sbox ? qs*om*x0[0] : 0.0));
    xh[2] = x0[2] + 0.5*dt*v0[2];
    xg[0] = xh[0];
    xg[1] = xh[1] + (sbox ? qs*om*xh[0]*0.5*dt : 0.0);
    xg[2] = xh[2];
    const int lev = s->I(DK_PLEV,p);
    const int am = s->smr ? -1 : CloudAnchor(s, s->I(DK_PGID,p), xg);
    double ug[3] = {0.0, 0.0, 0.0};
    int nt = 0, tk[27], tj[27], ti[27];
    double tw[27];
    bool ok;
    if (s->smr) {
      ok = LeafForEach<true>(s, xg, lev,
          [&](int b, int k, int j, int i, double w, double m1x, double m1y, double m1z) {
            GatherLinear(s, b, k, j, i, w, m1x, m1y, m1z, xg, ug);
          });
    } else {
      ok = (am >= 0) && CloudForEach<true>(s, am, xg,
          [&](int k, int j, int i, double w, double m1x, double m1y, double m1z) {
            GatherLinear(s, am, k, j, i, w, m1x, m1y, m1z, xg, ug);
            if (nt < 27) {tk[nt] = k; tj[nt] = j; ti[nt] = i; tw[nt] = w;}
            ++nt;
          });
      ok = ok && nt <= 27;
    }
    if (!ok) {s->Fail(DK_EHALO, p); continue;}
    const double h = 0.5*dt;
    const double lambda = h/s->R(DK_IPTS,p);
    const double aa = 1.0 + lambda;
    double r[3] = {v0[0] + lambda*ug[0], v0[1] + lambda*ug[1], v0[2] + lambda*ug[2]};
    double vh[3];
    if (sbox) {
      const double cxy = 2.0*h*om;
      const double cyx = h*(2.0-qs)*om;
      const double det = aa*aa + cxy*cyx;
      vh[0] = (aa*r[0] + cxy*r[1])/det;
      vh[1] = (aa*r[1] - cyx*r[0])/det;
      if (strat) r[2] -= h*SQR(om)*xh[2];
      vh[2] = r[2]/aa;
    } else {
      vh[0] = r[0]/aa; vh[1] = r[1]/aa; vh[2] = r[2]/aa;
    }
    s->R(DK_IPVX,p) = 2.0*vh[0] - v0[0];
    s->R(DK_IPVY,p) = 2.0*vh[1] - v0[1];
    s->R(DK_IPVZ,p) = 2.0*vh[2] - v0[2];
    s->R(DK_IPX,p) = x0[0] + dt*vh[0];
    s->R(DK_IPY,p) = x0[1] + dt*(vh[1] - (sbox ? qs*om*xh[0] : 0.0));
    s->R(DK_IPZ,p) = x0[2] + dt*vh[2];
    s->R(DK_IPX1,p) = xh[0]; s->R(DK_IPY1,p) = xh[1]; s->R(DK_IPZ1,p) = xh[2];
    const double ts = s->R(DK_IPTS,p);
    const double dvx = dt*(ug[0]-vh[0])/ts;
    const double dvy = dt*(ug[1]-vh[1])/ts;
    const double dvz = dt*(ug[2]-vh[2])/ts;
    s->R(DK_IPRX,p) = dvx; s->R(DK_IPRY,p) = dvy; s->R(DK_IPRZ,p) = dvz;
    if (!br) continue;
    const double mass = -s->R(DK_IPM,p);
    if (s->smr) {
      LeafForEach<false>(s, xg, lev, [&](int b, int k, int j, int i, double w) {
        const double dw = mass*w/(s->dx1[b]*s->dx2[b]*s->dx3[b]);
        DepositAdd(s, &s->dmom[s->Idx(b,3,0,k,j,i)], dw*dvx);
        DepositAdd(s, &s->dmom[s->Idx(b,3,1,k,j,i)], dw*dvy);
        DepositAdd(s, &s->dmom[s->Idx(b,3,2,k,j,i)], dw*dvz);
      });
      continue;
    }
    const double measure = s->dx1[am]*s->dx2[am]*s->dx3[am];
    for (int q=0; q<nt; ++q) {
      if (tw[q] == 0.0) continue;
      const double dw = mass*tw[q]/measure;
      DepositAdd(s, &s->dmom[s->Idx(am,3,0,tk[q],tj[q],ti[q])], dw*dvx);
      DepositAdd(s, &s->dmom[s->Idx(am,3,1,tk[q],tj[q],ti[q])], dw*dvy);
      DepositAdd(s, &s->dmom[s->Idx(am,3,2,tk[q],tj[q],ti[q])], dw*dvz);
    }
  }
}

//----------------------------------------------------------------------------------------
// ghost copy exchange of a cell field (periodic same-level), src/bvals/bvals_cc.cpp:53-296;
// with shear-periodic x1 the x1 ghosts of the blocks on the mesh x1 faces are then replaced
// by the shear remap, and at physical x3 faces by UstarX3BCs (src/dust/dust_tasks.cpp:445-469)
void SmrGhosts(or_sim *s, std::vector<double> &a, int nvar);

void CopyGhosts(or_sim *s, std::vector<double> &a, int nvar) {
  if (s->smr) {SmrGhosts(s, a, nvar); return;}
  const int nmb = s->M.nmb;
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m) {
    for (int k=0; k<s->n3; ++k) {
      const int ox3 = (k < s->ks) ? -1 : ((k > s->ke) ? 1 : 0);
      for (int j=0; j<s->n2; ++j) {
        const int ox2 = (j < s->js) ? -1 : ((j > s->je) ? 1 : 0);
        for (int i=0; i<s->n1; ++i) {
          const int ox1 = (i < s->is) ? -1 : ((i > s->ie) ? 1 : 0);
          if (ox1 == 0 && ox2 == 0 && ox3 == 0) continue;
          const int nb = s->B[m].nghbr[Dir(ox1, ox2, ox3)];
          if (nb < 0) continue;   // physical face: set by the boundary conditions
          for (int v=0; v<nvar; ++v) {
            a[s->Idx(m,nvar,v,k,j,i)] =
                a[s->Idx(nb,nvar,v,k-ox3*s->nx3,j-ox2*s->nx2,i-ox1*s->nx1)];
          }
        }
      }
    }
  }
}

// src/dust/dust_tasks.cpp:445-469 (reflect: mirror with v_z negated; else zero gradient); for
// the gas (nvar = 4: density, momenta) the hydro's reflect negates IM3 the same way
// (src/bvals/physics/hydro_bcs.cpp:229-290).  velocity: the field's last of 3 components, or
// the gas's momentum IM3, is the normal one
void X3BCs(or_sim *s, std::vector<double> &a, int nvar, bool velocity) {
  const int vn = (nvar == NG_VARS) ? DK_IM3 : 2;
  if (s->M.bc_x3 == DK_X3_PERIODIC) return;
  const bool refl = (s->M.bc_x3 == DK_X3_REFLECT);
  const int ng = s->ng;
  for (int m=0; m<s->M.nmb; ++m) {
    const bool lo = (s->B[m].nghbr[Dir(0,0,-1)] < 0);
    const bool hi = (s->B[m].nghbr[Dir(0,0,1)] < 0);
    for (int g=0; g<ng; ++g) {
      for (int j=0; j<s->n2; ++j) {
        for (int i=0; i<s->n1; ++i) {
          if (lo) {
            const int kg = s->ks - 1 - g, kc = refl ? (s->ks + g) : s->ks;
            for (int v=0; v<nvar; ++v) {
              const double val = a[s->Idx(m,nvar,v,kc,j,i)];
              a[s->Idx(m,nvar,v,kg,j,i)] = (refl && velocity && v == vn) ? -val : val;
            }
          }
          if (hi) {
            const int kg = s->ke + 1 + g, kc = refl ? (s->ke - g) : s->ke;
            for (int v=0; v<nvar; ++v) {
              const double val = a[s->Idx(m,nvar,v,kc,j,i)];
              a[s->Idx(m,nvar,v,kg,j,i)] = (refl && velocity && v == vn) ? -val : val;
            }
          }
        }
      }
    }
  }
}

// The shear offset of the x1-face crossing at time t folded into [0, Ly)
// (src/bvals/bvals_part.cpp:138-142)
double ShearOffset(const or_sim *s, double t) {
  const double lx = s->M.x1max - s->M.x1min, ly = s->M.x2max - s->M.x2min;
  return std::fmod(s->P.qshear*s->P.omega0*lx*t, ly);
}

void ShearGhostsX1(or_sim *s, std::vector<double> &a, int nvar, double t);

// Additive fold of the ghost deposits (src/bvals/bvals_dep.cpp:320-562): every block sums the
// ghost strips of its same-level neighbours into its active edge strips, buffer by buffer in
// NeighborIndex order.  With shear-periodic x1 or physical x3 faces the general remap
// (src/bvals/bvals_dep_general.cpp) replaces the parts that cross those faces.
void FoldShear(or_sim *s, std::vector<double> &a, int nvar, double t);

void Fold(or_sim *s, std::vector<double> &a, int nvar, double t) {
  const int ng1 = s->ng - 1;
  const bool shear = (s->M.bc_x1 == DK_X1_SHEAR);
  // physical x3 faces first: the ghost deposits beyond the face are mirrored back into the
  // first ng layers (reflect) or dropped (outflow), over the whole (j, i) plane, so that the
  // x1/x2 ghost columns then carry them on to the neighbours
  // (src/bvals/bvals_dep_general.cpp:131-140)
  if (s->M.bc_x3 != DK_X3_PERIODIC && !s->M.x3_dep_drop) {
    for (int m=0; m<s->M.nmb; ++m) {
      const bool lo = (s->B[m].nghbr[Dir(0,0,-1)] < 0);
      const bool hi = (s->B[m].nghbr[Dir(0,0,1)] < 0);
      for (int v=0; v<nvar; ++v)
        for (int g=0; g<s->ng; ++g)
          for (int j=0; j<s->n2; ++j)
            for (int i=0; i<s->n1; ++i) {
              if (lo) a[s->Idx(m,nvar,v,s->ks+g,j,i)] += a[s->Idx(m,nvar,v,s->ks-1-g,j,i)];
              if (hi) a[s->Idx(m,nvar,v,s->ke-g,j,i)] += a[s->Idx(m,nvar,v,s->ke+1+g,j,i)];
            }
    }
  }
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int mv=0; mv<s->M.nmb*nvar; ++mv) {
    const int m = mv/nvar, v = mv - m*nvar;
    for (int d : s->slot_dir) {
      const int ox1 = d % 3 - 1, ox2 = (d/3) % 3 - 1, ox3 = d/9 - 1;
      const int nb = s->B[m].nghbr[d];
      if (nb < 0) continue;
      // the sender's strip crosses the shear-periodic x1 face: handled by the remap
      if (shear && ox1 != 0) {
        const bool crosses = (ox1 < 0) ? (s->B[m].lx1 == 0) : (s->B[m].lx1 == s->M.nmbx1 - 1);
        if (crosses) continue;
      }
      const int il = (ox1 > 0) ? (s->ie - ng1) : s->is, iu = (ox1 < 0) ? (s->is + ng1) : s->ie;
      const int jl = (ox2 > 0) ? (s->je - ng1) : s->js, ju = (ox2 < 0) ? (s->js + ng1) : s->je;
      const int kl = (ox3 > 0) ? (s->ke - ng1) : s->ks, ku = (ox3 < 0) ? (s->ks + ng1) : s->ke;
      for (int k=kl; k<=ku; ++k)
        for (int j=jl; j<=ju; ++j)
          for (int i=il; i<=iu; ++i)
            a[s->Idx(m,nvar,v,k,j,i)] +=
                a[s->Idx(nb,nvar,v,k-ox3*s->nx3,j-ox2*s->nx2,i-ox1*s->nx1)];
    }
  }
  if (shear) FoldShear(s, a, nvar, t);
}

}  // namespace

//========================================================================================
// shear-periodic x1 pieces


namespace {

// PLM remap "flux" through face j of a column u shifted by eps cells
// (src/shearing_box/remap_fluxes.hpp, PLM_RemapFlx)
template <class U>
inline double PLMRemapFlx(const U &u, int j, double eps) {
  if (eps > 0.0) {
    const double dql = u(j-1) - u(j-2);
    const double dqr = u(j  ) - u(j-1);
    const double dq2 = dql*dqr;
    double dqm = 2.0*dq2/(dql + dqr);
    if (dq2 <= 0.0) dqm = 0.0;
    return eps*(u(j-1) + 0.5*(1.0 - eps)*dqm);
  }
  const double dql = u(j  ) - u(j-1);
  const double dqr = u(j+1) - u(j  );
  const double dq2 = dql*dqr;
  double dqm = 2.0*dq2/(dql + dqr);
  if (dq2 <= 0.0) dqm = 0.0;
  return eps*(u(j) - 0.5*(1.0 + eps)*dqm);
}

// Shear-periodic x1 ghost fill of a cell field after the periodic copy
// (ShearingBoxCC::PackAndSend/RecvAndUnpack, src/shearing_box/shearing_box_cc.cpp:68-260, and
// ShearingBox::Pieces, src/shearing_box/shearing_box.cpp:260-330): the x1 ghost column of a
// block on a mesh x1 face, which the copy filled with the opposite face at the same x2, is
// remapped by the fractional offset eps = fmod(yshear, dy)/dy (negated on the outer face) with
// the PLM fluxes and moved by joffset = int(yshear/dy) cells: inner ghost at global x2 cell G
// takes the remapped value of cell G - joffset, outer ghost the one of G + joffset;
// yshear = q Omega Lx t (src/shearing_box/shearing_box_tasks.cpp:143-147).  All k rows,
// x3 ghosts included (ShearingBox::Set, shearing_box.cpp:236-250).
void ShearGhostsX1(or_sim *s, std::vector<double> &a, int nvar, double t) {
  const int ny = s->nx2*s->M.nmbx2, nzt = s->nx3*s->M.nmbx3;
  const double lx = s->M.x1max - s->M.x1min;
  const double yshear = (s->P.qshear*s->P.omega0)*lx*t;
  std::vector<double> col(ny);
  for (int m=0; m<s->M.nmb; ++m) {
    const dk_block &b = s->B[m];
    for (int side=0; side<2; ++side) {
      const bool on_face = (side == 0) ? (b.lx1 == 0) : (b.lx1 == s->M.nmbx1 - 1);
      if (!on_face) continue;
      const int bxs = (side == 0) ? s->M.nmbx1 - 1 : 0;   // the opposite face
      const double dy = s->dx2[m];
      const int joffset = static_cast<int>(yshear/dy);
      double eps = std::fmod(yshear, dy)/dy;
      if (side == 1) eps *= -1.0;
      for (int v=0; v<nvar; ++v) {
        for (int k=0; k<s->n3; ++k) {
          int gk = b.lx3*s->nx3 + (k - s->ks);
          if (s->M.bc_x3 == DK_X3_PERIODIC) gk = ((gk % nzt) + nzt) % nzt;
          else gk = std::min(std::max(gk, 0), nzt - 1);   // replaced by the x3 values later
          const int bz = gk/s->nx3, kl = s->ks + (gk - bz*s->nx3);
          for (int g=0; g<s->ng; ++g) {
            const int ig = (side == 0) ? (s->is - s->ng + g) : (s->ie + 1 + g);
            const int isrc = (side == 0) ? (ig + s->nx1) : (ig - s->nx1);
            for (int G=0; G<ny; ++G) {
              const int by = G/s->nx2;
              const int ms = bxs + s->M.nmbx1*(by + s->M.nmbx2*bz);
              col[G] = a[s->Idx(ms,nvar,v,kl,s->js + (G - by*s->nx2),isrc)];
            }
            auto u = [&](int G) {return col[((G % ny) + ny) % ny];};
            for (int j=0; j<s->n2; ++j) {
              const int Gt = b.lx2*s->nx2 + (j - s->js);
              const int G = (side == 0) ? (Gt - joffset) : (Gt + joffset);
              const double r = u(G) - (PLMRemapFlx(u, G+1, eps) - PLMRemapFlx(u, G, eps));
              a[s->Idx(m,nvar,v,k,j,ig)] = r;
            }
          }
        }
      }
    }
  }
}

// Ghost deposits beyond a shear-periodic x1 face: a ghost cell of a block on the outer face
// lies (after x -= Lx, y += ysh: bvals_part.cpp:274-283) on the inner face, shifted by +ysh in
// x2; it is given to the two active cells it overlaps by the overlap fractions (piecewise-
// constant remap, conservative: src/bvals/bvals_dep_general.cpp:1-30,100-140).
void FoldShear(or_sim *s, std::vector<double> &a, int nvar, double t) {
  const int ny = s->nx2*s->M.nmbx2;
  const double dy = (s->M.x2max - s->M.x2min)/ny;
  const double ysh = ShearOffset(s, t);
  const double shc = ysh/dy;
  const int ng = s->ng;
  // Snapshot of every source ghost cell on the two mesh x1 faces (all j, k of the ghost
  // slabs, including x2/x3 ghost corners), so that the sums read unmodified values.
  struct Src {int m, k, j, i; double val[8];};
  std::vector<Src> src;
  for (int m=0; m<s->M.nmb; ++m) {
    const dk_block &b = s->B[m];
    for (int side=0; side<2; ++side) {
      const bool on_face = (side == 0) ? (b.lx1 == 0) : (b.lx1 == s->M.nmbx1 - 1);
      if (!on_face) continue;
      for (int k=0; k<s->n3; ++k)
        for (int j=0; j<s->n2; ++j)
          for (int g=0; g<ng; ++g) {
            const int i = (side == 0) ? (s->is - 1 - g) : (s->ie + 1 + g);
            Src e{m, k, j, i, {0}};
            for (int v=0; v<nvar; ++v) e.val[v] = a[s->Idx(m,nvar,v,k,j,i)];
            src.push_back(e);
          }
    }
  }
  const int nx3t = s->nx3*s->M.nmbx3;
  for (const Src &e : src) {
    const dk_block &b = s->B[e.m];
    const int side = (e.i < s->is) ? 0 : 1;
    // global indices of the ghost cell
    int gk = b.lx3*s->nx3 + (e.k - s->ks);
    const int gj = b.lx2*s->nx2 + (e.j - s->js);
    // x1: the opposite face; inner ghost (x < x1min) -> outer active, shifted by -ysh;
    // outer ghost -> inner active, shifted by +ysh
    const int bx = (side == 0) ? s->M.nmbx1 - 1 : 0;
    const int it = (side == 0) ? (s->ie - (s->is - 1 - e.i)) : (s->is + (e.i - s->ie - 1));
    if (gk < 0 || gk >= nx3t) {
      // beyond a physical x3 face the first pass of the fold has already mirrored (or
      // dropped) this layer into the active layers of the same ghost column
      if (s->M.bc_x3 != DK_X3_PERIODIC) continue;
      gk = (gk + nx3t) % nx3t;
    }
    const int by3 = gk / s->nx3, kt = s->ks + (gk - by3*s->nx3);
    const double lo = gj + ((side == 0) ? -shc : shc);
    const int c0 = static_cast<int>(std::floor(lo));
    const double fr = lo - c0;
    const int cells[2] = {c0, c0 + 1};
    const double wts[2] = {1.0 - fr, fr};
    for (int q=0; q<2; ++q) {
      if (wts[q] <= 0.0) continue;
      const int gjt = ((cells[q] % ny) + ny) % ny;
      const int by = gjt / s->nx2, jt = s->js + (gjt - by*s->nx2);
      const int mt = bx + s->M.nmbx1*(by + s->M.nmbx2*by3);
      for (int v=0; v<nvar; ++v) a[s->Idx(mt,nvar,v,kt,jt,it)] += wts[q]*e.val[v];
    }
  }
}

}  // namespace

//========================================================================================
// construction and storage

extern "C" {

or_sim *or_create(const dk_mesh *mesh, const dk_block *blocks, const dk_physics *phys,
                  int64_t capacity) {
  if (mesh == nullptr || blocks == nullptr || phys == nullptr) return nullptr;
  if (mesh->nx3 < 2 || mesh->ng < 1) return nullptr;   // 3-D meshes only
  // refinable clouds: pc2 on a mesh periodic in x1, x2, x3, ghosts of 2
  if (phys->clouds && (phys->coupling != DK_PC2 || mesh->bc_x1 != DK_X1_PERIODIC ||
                       mesh->bc_x3 != DK_X3_PERIODIC || mesh->ng < 2))
    return nullptr;
  // a statically refined mesh (levels above 0): refinable clouds only (their traversal of the
  // actual leaves; point-particle deposits would need the cross-level fold)
  int maxlev = 0;
  for (int m=0; m<mesh->nmb; ++m) maxlev = std::max(maxlev, blocks[m].level);
  const bool smr = maxlev > 0 || mesh->nmb != mesh->nmbx1*mesh->nmbx2*mesh->nmbx3;
  if (smr && !phys->clouds) return nullptr;
  or_sim *s = new or_sim();
  s->M = *mesh;
  s->P = *phys;
  s->B.assign(blocks, blocks + mesh->nmb);
  s->ng = mesh->ng; s->nx1 = mesh->nx1; s->nx2 = mesh->nx2; s->nx3 = mesh->nx3;
  s->n1 = s->nx1 + 2*s->ng; s->n2 = s->nx2 + 2*s->ng; s->n3 = s->nx3 + 2*s->ng;
  s->is = s->ng; s->ie = s->ng + s->nx1 - 1;
  s->js = s->ng; s->je = s->ng + s->nx2 - 1;
  s->ks = s->ng; s->ke = s->ng + s->nx3 - 1;
  s->ncb = (int64_t)s->n1*s->n2*s->n3;
  const int64_t nc = s->ncb*mesh->nmb;
  s->cap = std::max<int64_t>(capacity, 1);
  s->pr.assign(DK_NRDATA*s->cap, 0.0);
  s->pi.assign(DK_NIDATA*s->cap, 0);
  s->u0.assign(NG_VARS*nc, 0.0);
  s->u1.assign(NG_VARS*nc, 0.0);
  s->qdep.assign(5*nc, 0.0);
  s->ustar.assign(3*nc, 0.0);
  s->dmom.assign(3*nc, 0.0);
  s->ap.assign(3*nc, 0.0);
  s->resid.assign(4*nc, 0.0);
  if (phys->solver == DK_PCG || phys->solver == DK_ADAPTIVE) {
    s->sr.assign(3*nc, 0.0);
    s->sp.assign(3*nc, 0.0);
  }
  s->dx1.resize(mesh->nmb); s->dx2.resize(mesh->nmb); s->dx3.resize(mesh->nmb);
  for (int m=0; m<mesh->nmb; ++m) {
    s->dx1[m] = (s->B[m].x1max - s->B[m].x1min)/static_cast<double>(s->nx1);
    s->dx2[m] = (s->B[m].x2max - s->B[m].x2min)/static_cast<double>(s->nx2);
    s->dx3[m] = (s->B[m].x3max - s->B[m].x3min)/static_cast<double>(s->nx3);
  }
  // the root cell size (src/dust/dust_cloud.cpp:135, dust_cloud3d.cpp:591-593)
  s->root_dx[0] = (mesh->x1max - mesh->x1min)/static_cast<double>(mesh->nmbx1*mesh->nx1);
  s->root_dx[1] = (mesh->x2max - mesh->x2min)/static_cast<double>(mesh->nmbx2*mesh->nx2);
  s->root_dx[2] = (mesh->x3max - mesh->x3min)/static_cast<double>(mesh->nmbx3*mesh->nx3);
  for (int d=0; d<3; ++d) s->iroot_dx[d] = 1.0/s->root_dx[d];
  // the 26 same-level slots in NeighborIndex order
  std::vector<std::pair<int,int>> order;
  for (int ox3=-1; ox3<=1; ++ox3)
    for (int ox2=-1; ox2<=1; ++ox2)
      for (int ox1=-1; ox1<=1; ++ox1) {
        if (ox1 == 0 && ox2 == 0 && ox3 == 0) continue;
        order.push_back({NeighborIndex(ox1, ox2, ox3), Dir(ox1, ox2, ox3)});
      }
  std::sort(order.begin(), order.end());
  for (auto &o : order) s->slot_dir.push_back(o.second);
  if (smr) {
    // the finest-level block grid: every slot covered by exactly one leaf
    s->smr = true;
    s->maxlev = maxlev;
    s->F[0] = mesh->nmbx1 << maxlev; s->F[1] = mesh->nmbx2 << maxlev;
    s->F[2] = mesh->nmbx3 << maxlev;
    s->leafmap.assign((size_t)s->F[0]*s->F[1]*s->F[2], -1);
    for (int m=0; m<mesh->nmb; ++m) {
      const dk_block &b = blocks[m];
      if (b.level < 0 || b.level > maxlev) {delete s; return nullptr;}
      const int r = 1 << (maxlev - b.level);
      const int l[3] = {b.lx1, b.lx2, b.lx3};
      for (int d=0; d<3; ++d)
        if (l[d] < 0 || (l[d] + 1)*r > s->F[d]) {delete s; return nullptr;}
      for (int k=b.lx3*r; k<(b.lx3 + 1)*r; ++k)
        for (int j=b.lx2*r; j<(b.lx2 + 1)*r; ++j)
          for (int i=b.lx1*r; i<(b.lx1 + 1)*r; ++i) {
            int &slot = s->leafmap[((size_t)k*s->F[1] + j)*s->F[0] + i];
            if (slot >= 0) {delete s; return nullptr;}
            slot = m;
          }
    }
    for (int v : s->leafmap)
      if (v < 0) {delete s; return nullptr;}
  }
  return s;
}

void or_destroy(or_sim *s) { delete s; }
void or_set_threads(or_sim *s, int n) { s->nthreads = std::max(1, n); }
double *or_rdata(or_sim *s) { return s->pr.data(); }
int32_t *or_idata(or_sim *s) { return s->pi.data(); }
int64_t or_capacity(or_sim *s) { return s->cap; }
int64_t or_npart(or_sim *s) { return s->np; }
void or_set_npart(or_sim *s, int64_t n) { s->np = std::min(n, s->cap); }
double *or_u0(or_sim *s) { return s->u0.data(); }
double *or_u1(or_sim *s) { return s->u1.data(); }
int64_t or_cells_per_block(or_sim *s) { return s->ncb; }

double *or_field(or_sim *s, int which, int *nvar) {
  switch (which) {
    case 0: if (nvar) *nvar = 5; return s->qdep.data();
    case 1: if (nvar) *nvar = 3; return s->ustar.data();
    case 2: if (nvar) *nvar = 3; return s->dmom.data();
    case 3: if (nvar) *nvar = 3; return s->ap.data();
    case 4: if (nvar) *nvar = 4; return s->resid.data();
    default: if (nvar) *nvar = 0; return nullptr;
  }
}

int or_last_error(or_sim *s, int64_t *particle) {
  const int e = s->err;
  if (particle) *particle = s->err_p;
  s->err = 0; s->err_p = -1;
  return e;
}

//========================================================================================
// problem set-up (dustk_setup.h)

// the particle count of every block of an uneven set-up (dk_setup.block_lognormal) and their
// prefix offsets (nmb + 1 entries)
std::vector<int64_t> BlockOffsets(const or_sim *s, const dk_setup *su) {
  const int nmb = s->M.nmb;
  const int64_t ncell_b = (int64_t)s->nx1*s->nx2*s->nx3;
  const uint64_t base = su->seed*UINT64_C(0x9e3779b97f4a7c15);
  std::vector<int64_t> off(nmb + 1, 0);
  for (int m=0; m<nmb; ++m) {
    int64_t n = (int64_t)su->ppc*ncell_b;
    if (su->block_lognormal) {
      const double z = Gauss(base + static_cast<uint64_t>(m)*UINT64_C(0x2545f4914f6cdd1d) +
                             DK_SALT_BLOCK);
      const double sg = su->block_sigma;
      n = std::llround(static_cast<double>(su->ppc*ncell_b)*std::exp(sg*z - 0.5*sg*sg));
    }
    off[m + 1] = off[m] + n;
  }
  return off;
}

int or_setup(or_sim *s, const dk_setup *su) {
  const int nmb = s->M.nmb;
  const int64_t ncell_b = (int64_t)s->nx1*s->nx2*s->nx3;
  const int64_t npb = (int64_t)su->ppc*ncell_b;
  if (su->nspecies < 1 || su->nspecies > DK_MAXSPEC) return DK_ECAPACITY;
  if ((su->block_lognormal || s->smr) && (su->nclump > 0 || su->gauss_z)) return DK_EINVAL;
  const std::vector<int64_t> boff = BlockOffsets(s, su);
  const int64_t ntot = boff[nmb];
  if (ntot > s->cap) return DK_ECAPACITY;
  s->np = ntot;
  const uint64_t base = su->seed*UINT64_C(0x9e3779b97f4a7c15);
  const double lxm = s->M.x1max - s->M.x1min, lym = s->M.x2max - s->M.x2min;
  const double lzm = s->M.x3max - s->M.x3min;
  const uint64_t saltv[3] = {DK_SALT_V0, DK_SALT_V1, DK_SALT_V2};
  const uint64_t saltg[3] = {DK_SALT_G0, DK_SALT_G1, DK_SALT_G2};
  // total gas mass (zmetal normalisation)
  double gas_mass = 0.0;
  for (int m=0; m<nmb; ++m) {
    const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
    for (int k=s->ks; k<=s->ke; ++k) {
      const double z = s->B[m].x3min + (k - s->ks + 0.5)*s->dx3[m];
      const double rho = su->gas_strat ? su->rho0*std::exp(-0.5*z*z) : su->rho0;
      gas_mass += rho*vol*s->nx1*s->nx2;
    }
  }
  // mean mass weight of each continuous bin: <tau^slope> over log-uniform tau (closed form)
  double wmean[DK_MAXSPEC];
  for (int sp=0; sp<su->nspecies; ++sp) {
    wmean[sp] = 1.0;
    if (su->continuous && su->mass_slope != 0.0) {
      const double a = su->mass_slope, l = std::log(su->tau_hi[sp]/su->tau_lo[sp]);
      wmean[sp] = (std::pow(su->tau_hi[sp], a) - std::pow(su->tau_lo[sp], a))/(a*l);
    }
  }
  const double ppc_s = static_cast<double>(su->ppc)/su->nspecies;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t q=0; q<ntot; ++q) {
    const uint64_t key = static_cast<uint64_t>(q) + base;
    int m = su->block_lognormal
          ? static_cast<int>(std::upper_bound(boff.begin(), boff.end(), q) - boff.begin()) - 1
          : static_cast<int>(q/npb);
    const int sp = static_cast<int>(q % su->nspecies);
    double x, y, z;
    if (su->nclump > 0) {
      if (HashUniform01(key + DK_SALT_CLUMP) < su->clump_frac) {
        const int ci = std::min(su->nclump - 1, static_cast<int>(
            HashUniform01(key + DK_SALT_CLUMP + 1)*su->nclump));
        const uint64_t ck = base + static_cast<uint64_t>(ci)*UINT64_C(0x2545f4914f6cdd1d);
        const double cx = s->M.x1min + HashUniform01(ck + UINT64_C(0x632be59bd9b4e019))*lxm;
        const double cy = s->M.x2min + HashUniform01(ck + UINT64_C(0x8cb92baa3f3d8dd7))*lym;
        const double cz = s->M.x3min + HashUniform01(ck + UINT64_C(0x58f38ded6f7c55b5))*lzm;
        x = cx + su->clump_sigma*Gauss(key + UINT64_C(0x632be59bd9b4e019));
        y = cy + su->clump_sigma*Gauss(key + UINT64_C(0x8cb92baa3f3d8dd7));
        z = cz + su->clump_sigma*Gauss(key + UINT64_C(0x58f38ded6f7c55b5));
        x -= lxm*std::floor((x - s->M.x1min)/lxm);
        y -= lym*std::floor((y - s->M.x2min)/lym);
        z -= lzm*std::floor((z - s->M.x3min)/lzm);
        if (x >= s->M.x1max) x = s->M.x1min;
        if (y >= s->M.x2max) y = s->M.x2min;
        if (z >= s->M.x3max) z = s->M.x3min;
      } else {
        x = s->M.x1min + HashUniform01(key + UINT64_C(0x632be59bd9b4e019))*lxm;
        y = s->M.x2min + HashUniform01(key + UINT64_C(0x8cb92baa3f3d8dd7))*lym;
        z = s->M.x3min + HashUniform01(key + UINT64_C(0x58f38ded6f7c55b5))*lzm;
      }
      // the block holding the position (uniform mesh, gid = bx + nx*(by + ny*bz))
      const int bx = std::min(s->M.nmbx1 - 1, static_cast<int>((x - s->M.x1min)/lxm*s->M.nmbx1));
      const int by = std::min(s->M.nmbx2 - 1, static_cast<int>((y - s->M.x2min)/lym*s->M.nmbx2));
      const int bz = std::min(s->M.nmbx3 - 1, static_cast<int>((z - s->M.x3min)/lzm*s->M.nmbx3));
      m = bx + s->M.nmbx1*(by + s->M.nmbx2*bz);
      // a position on a block face by round-off: keep it inside [xmin, xmax) of the block
      const dk_block &b = s->B[m];
      if (x < b.x1min) x = b.x1min;
      if (x >= b.x1max) x = std::nextafter(b.x1max, b.x1min);
      if (y < b.x2min) y = b.x2min;
      if (y >= b.x2max) y = std::nextafter(b.x2max, b.x2min);
      if (z < b.x3min) z = b.x3min;
      if (z >= b.x3max) z = std::nextafter(b.x3max, b.x3min);
    } else {
      const dk_block &b = s->B[m];
      const double lx = b.x1max - b.x1min, ly = b.x2max - b.x2min, lz = b.x3max - b.x3min;
      x = b.x1min + HashUniform01(key + UINT64_C(0x632be59bd9b4e019))*lx;
      y = b.x2min + HashUniform01(key + UINT64_C(0x8cb92baa3f3d8dd7))*ly;
      if (su->gauss_z) {
        // Gaussian heights over the whole mesh (dust_setup.hpp:157-171), then the block
        z = 0.0;
        for (int t=0; t<64; ++t) {
          const uint64_t kz = key + UINT64_C(0x58f38ded6f7c55b5) +
                              static_cast<uint64_t>(t)*UINT64_C(0x2545f4914f6cdd1d);
          const double u1 = std::fmax(HashUniform01(kz), 1.0e-300);
          const double u2 = HashUniform01(kz + UINT64_C(0x1d8e4e27c47d124f));
          const double zz = su->hp*std::sqrt(-2.0*std::log(u1))*std::cos(2.0*M_PI*u2);
          if (zz >= s->M.x3min && zz < s->M.x3max) {z = zz; break;}
        }
        const int bz = std::min(s->M.nmbx3 - 1,
                                static_cast<int>((z - s->M.x3min)/lzm*s->M.nmbx3));
        m = b.lx1 + s->M.nmbx1*(b.lx2 + s->M.nmbx2*bz);
        const dk_block &bb = s->B[m];
        if (z < bb.x3min) z = bb.x3min;
        if (z >= bb.x3max) z = std::nextafter(bb.x3max, bb.x3min);
      } else {
        z = b.x3min + HashUniform01(key + UINT64_C(0x58f38ded6f7c55b5))*lz;
      }
    }
    // stopping time and mass
    double ts = su->taus[sp], wm = 1.0;
    if (su->continuous) {
      const double u = HashUniform01(key + DK_SALT_TAU);
      ts = std::exp(std::log(su->tau_lo[sp]) + u*std::log(su->tau_hi[sp]/su->tau_lo[sp]));
      if (su->mass_slope != 0.0) wm = std::pow(ts, su->mass_slope)/wmean[sp];
    }
    const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
    double mass;
    if (su->zmetal > 0.0) {
      mass = wm*su->zmetal*gas_mass/(static_cast<double>(ntot));
    } else {
      mass = wm*su->eps[sp]*su->rho0*vol/ppc_s;
    }
    double vm[3] = {su->gas_vx, su->gas_vy, su->gas_vz};
    if (su->nsh) {
      NSHDustDrift(su->omega0, su->qshear, ts, su->gas_vx, su->gas_vy, vm[0], vm[1]);
      vm[2] = 0.0;
    }
    double v[3];
    for (int c=0; c<3; ++c) v[c] = vm[c] + su->dust_noise*Gauss(key + saltv[c]);
    s->R(DK_IPX,q) = x; s->R(DK_IPY,q) = y; s->R(DK_IPZ,q) = z;
    s->R(DK_IPVX,q) = v[0]; s->R(DK_IPVY,q) = v[1]; s->R(DK_IPVZ,q) = v[2];
    s->R(DK_IPX1,q) = x; s->R(DK_IPY1,q) = y; s->R(DK_IPZ1,q) = z;
    s->R(DK_IPVX1,q) = v[0]; s->R(DK_IPVY1,q) = v[1]; s->R(DK_IPVZ1,q) = v[2];
    s->R(DK_IPRX,q) = 0.0; s->R(DK_IPRY,q) = 0.0; s->R(DK_IPRZ,q) = 0.0;
    s->R(DK_IPTS,q) = ts; s->R(DK_IPM,q) = mass;
    s->I(DK_PGID,q) = m; s->I(DK_PTAG,q) = static_cast<int32_t>(q);
    // (a refinable cloud of its block's level: half-width one cell of that level)
    s->I(DK_PSP,q) = sp; s->I(DK_PLEV,q) = s->B[m].level; s->I(DK_PTAGH,q) = 0;
  }
  // gas (the global cell index of the block's level, plus level * 2^40)
  Zero(s->u0);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m) {
    const dk_block &b = s->B[m];
    const int64_t ny = (int64_t)s->nx2*(s->M.nmbx2 << b.level);
    const int64_t nxg = (int64_t)s->nx1*(s->M.nmbx1 << b.level);
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i) {
          const int64_t gi = b.lx1*s->nx1 + (i - s->is), gj = b.lx2*s->nx2 + (j - s->js);
          const int64_t gk = b.lx3*s->nx3 + (k - s->ks);
          const uint64_t ck = static_cast<uint64_t>((gk*ny + gj)*nxg + gi) + base +
                              (static_cast<uint64_t>(b.level) << 40);
          const double z = b.x3min + (k - s->ks + 0.5)*s->dx3[m];
          const double rho = su->gas_strat ? su->rho0*std::exp(-0.5*z*z) : su->rho0;
          const double um[3] = {su->gas_vx, su->gas_vy, su->gas_vz};
          s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)] = rho;
          for (int c=0; c<3; ++c) {
            const double u = um[c] + su->gas_noise*Gauss(ck + saltg[c]);
            s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)] = rho*u;
          }
        }
  }
  CopyGhosts(s, s->u0, NG_VARS);
  if (s->M.bc_x1 == DK_X1_SHEAR) ShearGhostsX1(s, s->u0, NG_VARS, 0.0);
  X3BCs(s, s->u0, NG_VARS, true);
  s->u1 = s->u0;
  return DK_OK;
}

//========================================================================================
// tasks

// src/dust/dust_push.cpp:35-80
int or_first_two_imp_rk(or_sim *s, const dk_stage *st) {
  if (s->P.coupling == DK_PC2 && st->stage == 2) {
    const int64_t nc = s->ncb*s->M.nmb;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t c=0; c<nc; ++c) {
      const int64_t m = c/s->ncb, r = c - m*s->ncb;
      auto U = [&](std::vector<double> &u, int v) -> double& {
        return u[(m*NG_VARS + v)*s->ncb + r];
      };
      double inv_rho_sum = 1.0/(U(s->u1,DK_IDN) + U(s->u0,DK_IDN));
      for (int v=0; v<3; ++v) {
        s->ustar[(m*3 + v)*s->ncb + r] = (U(s->u1,DK_IM1+v) + U(s->u0,DK_IM1+v))*inv_rho_sum;
      }
    }
    return DK_OK;
  }
  if (st->stage != 1) return DK_OK;
  s->u1 = s->u0;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t p=0; p<s->np; ++p) {
    s->R(DK_IPX1,p) = s->R(DK_IPX,p);
    s->R(DK_IPVX1,p) = s->R(DK_IPVX,p);
    s->R(DK_IPY1,p) = s->R(DK_IPY,p);
    s->R(DK_IPVY1,p) = s->R(DK_IPVY,p);
    s->R(DK_IPZ1,p) = s->R(DK_IPZ,p);
    s->R(DK_IPVZ1,p) = s->R(DK_IPVZ,p);
  }
  return DK_OK;
}

// src/dust/dust_push.cpp:91-300
int or_explicit_push(or_sim *s, const dk_stage *st) {
  if (!ActiveStage(s, st)) return DK_OK;
  const bool sbox = s->P.shearing_box, strat = s->P.stratified;
  const double qs = s->P.qshear, om = s->P.omega0;
  if (s->P.coupling == DK_PC2) {
    if (st->stage != 2) return DK_OK;
    const double dt = st->dt;
    const bool br = s->P.back_reaction;
    if (br) Zero(s->dmom);
    if (s->P.clouds) {   // src/dust/dust_push.cpp:124-129
      CloudAdvance(s, dt);
      return s->err;
    }
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t p=0; p<s->np; ++p) {
      const int m = s->I(DK_PGID,p);
      const double x0 = s->R(DK_IPX1,p), y0 = s->R(DK_IPY1,p), z0 = s->R(DK_IPZ1,p);
      const double vx0 = s->R(DK_IPVX1,p), vy0 = s->R(DK_IPVY1,p), vz0 = s->R(DK_IPVZ1,p);
      const double xh = x0 + 0.5*dt*vx0;
      const double yh = y0 + 0.5*dt*(vy0 - (sbox ? qs*om*x0 : 0.0));
      const double zh = z0 + 0.5*dt*vz0;
      const double yg = yh + (sbox ? qs*om*xh*0.5*dt : 0.0);
      Stencil w;
      MakeStencil(s, m, xh, yg, zh, w);
      double ugx = 0.0, ugy = 0.0, ugz = 0.0;
      for (int c=0; c<=2; ++c)
        for (int b=0; b<3; ++b) {
          const double wcb = w.wz[c]*w.wy[b];
          if (wcb == 0.0) continue;
          for (int a=0; a<3; ++a) {
            const double ww = wcb*w.wx[a];
            const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
            ugx += ww*s->ustar[s->Idx(m,3,0,kk,jj,ii)];
            ugy += ww*s->ustar[s->Idx(m,3,1,kk,jj,ii)];
            ugz += ww*s->ustar[s->Idx(m,3,2,kk,jj,ii)];
          }
        }
      const double h = 0.5*dt;
      const double lambda = h/s->R(DK_IPTS,p);
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
        if (strat) rz -= h*SQR(om)*zh;
        vzh = rz/aa;
      } else {
        vxh = rx/aa; vyh = ry/aa; vzh = rz/aa;
      }
      s->R(DK_IPVX,p) = 2.0*vxh - vx0;
      s->R(DK_IPVY,p) = 2.0*vyh - vy0;
      s->R(DK_IPVZ,p) = 2.0*vzh - vz0;
      s->R(DK_IPX,p) = x0 + dt*vxh;
      s->R(DK_IPY,p) = y0 + dt*(vyh - (sbox ? qs*om*xh : 0.0));
      s->R(DK_IPZ,p) = z0 + dt*vzh;
      s->R(DK_IPX1,p) = xh; s->R(DK_IPY1,p) = yh; s->R(DK_IPZ1,p) = zh;
      const double ts = s->R(DK_IPTS,p);
      const double dvx = dt*(ugx-vxh)/ts, dvy = dt*(ugy-vyh)/ts, dvz = dt*(ugz-vzh)/ts;
      s->R(DK_IPRX,p) = dvx; s->R(DK_IPRY,p) = dvy; s->R(DK_IPRZ,p) = dvz;
      if (br) {
        const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
        const double fac = -s->R(DK_IPM,p)/vol;
        for (int c=0; c<=2; ++c)
          for (int b=0; b<3; ++b) {
            const double wcb = w.wz[c]*w.wy[b]*fac;
            if (wcb == 0.0) continue;
            for (int a=0; a<3; ++a) {
              const double ww = wcb*w.wx[a];
              const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
              DepositAdd(s, &s->dmom[s->Idx(m,3,0,kk,jj,ii)], ww*dvx);
              DepositAdd(s, &s->dmom[s->Idx(m,3,1,kk,jj,ii)], ww*dvy);
              DepositAdd(s, &s->dmom[s->Idx(m,3,2,kk,jj,ii)], ww*dvz);
            }
          }
      }
    }
    return DK_OK;
  }
  const double dt = st->dt;
  const double gam0 = st->gam0[st->stage-1], gam1 = st->gam1[st->stage-1];
  const double beta_dt = st->beta[st->stage-1]*dt;
  const double atw_dt = (st->stage == 2) ? st->a_twid22*dt : 0.0;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t p=0; p<s->np; ++p) {
    const double x_old = s->R(DK_IPX,p), y_old = s->R(DK_IPY,p), z_old = s->R(DK_IPZ,p);
    const double vx_old = s->R(DK_IPVX,p), vy_old = s->R(DK_IPVY,p);
    const double vz_old = s->R(DK_IPVZ,p);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (sbox) {
      fx = 2.0*om*vy_old;
      fy = -(2.0-qs)*om*vx_old;
      if (strat) fz = -SQR(om)*z_old;
    }
    s->R(DK_IPVX,p) = gam0*vx_old + gam1*s->R(DK_IPVX1,p) + beta_dt*fx + atw_dt*s->R(DK_IPRX,p);
    s->R(DK_IPVY,p) = gam0*vy_old + gam1*s->R(DK_IPVY1,p) + beta_dt*fy + atw_dt*s->R(DK_IPRY,p);
    s->R(DK_IPVZ,p) = gam0*vz_old + gam1*s->R(DK_IPVZ1,p) + beta_dt*fz + atw_dt*s->R(DK_IPRZ,p);
    double ydot = vy_old;
    if (sbox) ydot -= qs*om*x_old;
    s->R(DK_IPX,p) = gam0*x_old + gam1*s->R(DK_IPX1,p) + beta_dt*vx_old;
    s->R(DK_IPY,p) = gam0*y_old + gam1*s->R(DK_IPY1,p) + beta_dt*ydot;
    s->R(DK_IPZ,p) = gam0*z_old + gam1*s->R(DK_IPZ1,p) + beta_dt*vz_old;
  }
  return DK_OK;
}

// src/dust/dust_tasks.cpp:149-162 -> src/dust/dust.cpp:1678-1702 and
// src/bvals/bvals_part.cpp:115-471 (uniform mesh, one rank)
int or_update_particle_gids(or_sim *s, const dk_stage *st) {
  if (!MigrationActive(s, st)) return DK_OK;
  const dk_mesh &ms = s->M;
  const bool shear = (ms.bc_x1 == DK_X1_SHEAR);
  if (shear) {   // ShiftShearCrossings(ShearRouteTime)
    const double tau = ShearRouteTime(s, st);
    const double dsh = s->P.qshear*s->P.omega0*(ms.x1max - ms.x1min)*(tau - st->time);
    if (dsh != 0.0) {
      const double xa = ms.x1min, xb = ms.x1max;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
      for (int64_t p=0; p<s->np; ++p) {
        const int m = s->I(DK_PGID,p);
        const double b0 = s->B[m].x1min, b1 = s->B[m].x1max, lx = b1 - b0;
        const int ix = static_cast<int>((s->R(DK_IPX,p) - b0 + lx)/lx) - 1;
        const bool lo1 = (ix < 0) && (b0 < xa + 0.5*lx);
        const bool hi1 = (ix > 0) && (b1 > xb - 0.5*lx);
        if (lo1 || hi1) {
          const double d = hi1 ? dsh : -dsh;
          s->R(DK_IPY,p) += d;
          s->R(DK_IPY1,p) += d;
        }
      }
    }
  }
  double ysh = 0.0;
  if (shear) ysh = ShearOffset(s, st->time);
  const int x3bc = ms.bc_x3;
  const double z3a = ms.x3min, z3b = ms.x3max;
  double rm_mass = 0.0, rm_count = 0.0;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(+:rm_mass,rm_count)
  for (int64_t p=0; p<s->np; ++p) {
    const int m = s->I(DK_PGID,p);
    const dk_block &b = s->B[m];
    if (x3bc != 0 && (s->R(DK_IPZ,p) < z3a || s->R(DK_IPZ,p) >= z3b)) {
      const bool lo = (s->R(DK_IPZ,p) < z3a);
      if (x3bc == DK_X3_REFLECT) {
        const double zf = lo ? z3a : z3b;
        double z = 2.0*zf - s->R(DK_IPZ,p);
        if (z >= z3b) z = z3b - 1.0e-12*(z3b - z3a);
        if (z < z3a) z = z3a;
        s->R(DK_IPZ,p) = z;
        s->R(DK_IPVZ,p) = -s->R(DK_IPVZ,p);
        s->R(DK_IPZ1,p) = 2.0*zf - s->R(DK_IPZ1,p);
        s->R(DK_IPVZ1,p) = -s->R(DK_IPVZ1,p);
      } else {
        rm_mass += s->R(DK_IPM,p);
        rm_count += 1.0;
        s->R(DK_IPM,p) = 0.0;
        const double dz = s->dx3[m];
        const double z = lo ? (b.x3min + 0.5*dz) : (b.x3max - 0.5*dz);
        s->R(DK_IPZ,p) = z;
        s->R(DK_IPVX,p) = 0.0; s->R(DK_IPVY,p) = 0.0; s->R(DK_IPVZ,p) = 0.0;
        s->R(DK_IPX1,p) = s->R(DK_IPX,p); s->R(DK_IPY1,p) = s->R(DK_IPY,p);
        s->R(DK_IPZ1,p) = z;
        s->R(DK_IPVX1,p) = 0.0; s->R(DK_IPVY1,p) = 0.0; s->R(DK_IPVZ1,p) = 0.0;
      }
    }
    double x1 = s->R(DK_IPX,p), x2 = s->R(DK_IPY,p), x3 = s->R(DK_IPZ,p);
    const double lx = b.x1max - b.x1min, ly = b.x2max - b.x2min, lz = b.x3max - b.x3min;
    const int ix = static_cast<int>((x1 - b.x1min + lx)/lx) - 1;
    const int iy = static_cast<int>((x2 - b.x2min + ly)/ly) - 1;
    const int iz = static_cast<int>((x3 - b.x3min + lz)/lz) - 1;
    auto snap = [&](const int off, double &x, const double xa, const double xb, const int ip,
                    const int ir) {
      if (off != 0) return;
      double xs = x;
      if (x < xa) {
        xs = xa;
      } else if (x >= xb) {
        xs = std::nextafter(xb, xa);
      }
      if (xs != x) {
        s->R(ir,p) += xs - x;
        s->R(ip,p) = xs;
        x = xs;
      }
    };
    snap(ix, x1, b.x1min, b.x1max, DK_IPX, DK_IPX1);
    snap(iy, x2, b.x2min, b.x2max, DK_IPY, DK_IPY1);
    snap(iz, x3, b.x3min, b.x3max, DK_IPZ, DK_IPZ1);
    if (std::abs(ix) > 1 || std::abs(iy) > 1 || std::abs(iz) > 1) {
      s->Fail(DK_EROUTE, p);
      continue;
    }
    if ((std::abs(ix) + std::abs(iy) + std::abs(iz)) != 0) {
      const bool lo1 = (ix < 0) && (b.x1min < ms.x1min + 0.5*lx);
      const bool hi1 = (ix > 0) && (b.x1max > ms.x1max - 0.5*lx);
      const bool lo2 = (iy < 0) && (b.x2min < ms.x2min + 0.5*ly);
      const bool hi2 = (iy > 0) && (b.x2max > ms.x2max - 0.5*ly);
      const bool lo3 = (iz < 0) && (b.x3min < ms.x3min + 0.5*lz);
      const bool hi3 = (iz > 0) && (b.x3max > ms.x3max - 0.5*lz);
      const bool cross_in = shear && lo1, cross_out = shear && hi1;
      if (cross_in || cross_out) {
        const double lxm = ms.x1max - ms.x1min;
        s->R(DK_IPX,p) += cross_in ? lxm : -lxm;
        s->R(DK_IPX1,p) += cross_in ? lxm : -lxm;
        const double lym = ms.x2max - ms.x2min;
        double ynew = s->R(DK_IPY,p) + (cross_in ? -ysh : ysh);
        if (ynew < ms.x2min) ynew += lym;
        if (ynew >= ms.x2max) ynew -= lym;
        const double dy = ynew - s->R(DK_IPY,p);
        s->R(DK_IPY,p) = ynew;
        s->R(DK_IPY1,p) += dy;
        const double lzm = ms.x3max - ms.x3min;
        if (lo3) {
          s->R(DK_IPZ,p) += lzm; s->R(DK_IPZ1,p) += lzm;
        } else if (hi3) {
          s->R(DK_IPZ,p) -= lzm; s->R(DK_IPZ1,p) -= lzm;
        }
        int b2 = static_cast<int>(((s->R(DK_IPY,p) - ms.x2min)/lym)*static_cast<double>(ms.nmbx2));
        b2 = (b2 < 0) ? 0 : ((b2 > ms.nmbx2-1) ? (ms.nmbx2-1) : b2);
        int b3 = static_cast<int>(((s->R(DK_IPZ,p) - ms.x3min)/lzm)*static_cast<double>(ms.nmbx3));
        b3 = (b3 < 0) ? 0 : ((b3 > ms.nmbx3-1) ? (ms.nmbx3-1) : b3);
        const int b1 = cross_in ? (ms.nmbx1 - 1) : 0;   // the boundary block on the other side
        s->I(DK_PGID,p) = b1 + ms.nmbx1*(b2 + ms.nmbx2*b3);
      } else {
        int nb;
        if (s->smr) {
          const int off[3] = {ix, iy, iz};
          const double xs[3] = {x1, x2, x3};
          nb = SmrDest(s, m, off, xs);
        } else {
          nb = b.nghbr[Dir(ix, iy, iz)];
        }
        if (nb < 0) {
          s->Fail(DK_EROUTE, p);
          continue;
        }
        s->I(DK_PGID,p) = nb;
        if (lo1) {
          s->R(DK_IPX,p) += (ms.x1max - ms.x1min); s->R(DK_IPX1,p) += (ms.x1max - ms.x1min);
        } else if (hi1) {
          s->R(DK_IPX,p) -= (ms.x1max - ms.x1min); s->R(DK_IPX1,p) -= (ms.x1max - ms.x1min);
        }
        if (lo2) {
          s->R(DK_IPY,p) += (ms.x2max - ms.x2min); s->R(DK_IPY1,p) += (ms.x2max - ms.x2min);
        } else if (hi2) {
          s->R(DK_IPY,p) -= (ms.x2max - ms.x2min); s->R(DK_IPY1,p) -= (ms.x2max - ms.x2min);
        }
        if (lo3) {
          s->R(DK_IPZ,p) += (ms.x3max - ms.x3min); s->R(DK_IPZ1,p) += (ms.x3max - ms.x3min);
        } else if (hi3) {
          s->R(DK_IPZ,p) -= (ms.x3max - ms.x3min); s->R(DK_IPZ1,p) -= (ms.x3max - ms.x3min);
        }
      }
    }
  }
  s->x3_removed_mass += rm_mass;
  s->x3_removed_count += rm_count;
  return s->err;
}

// src/dust/dust_pm.cpp:104-277
int or_deposit_drag(or_sim *s, const dk_stage *st) {
  if (!ActiveStage(s, st) || !s->P.back_reaction) return DK_OK;
  // StiffnessActive (src/dust/dust.hpp:1150-1152): the stiffness record or the damping
  const bool rate_dep = s->P.stiffness || s->P.pc2_damping;
  if (s->P.coupling == DK_PC2) {
    if (st->stage != 1) return DK_OK;
    Zero(s->qdep);
    const double dt = st->dt;
    if (s->P.clouds) {   // src/dust/dust_pm.cpp:130-137
      CloudPredictor(s, dt, rate_dep);
      return s->err;
    }
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t p=0; p<s->np; ++p) {
      const int m = s->I(DK_PGID,p);
      Stencil w;
      MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p), s->R(DK_IPZ,p), w);
      double ugx = 0.0, ugy = 0.0, ugz = 0.0;
      for (int c=0; c<=2; ++c)
        for (int b=0; b<3; ++b) {
          const double wcb = w.wz[c]*w.wy[b];
          if (wcb == 0.0) continue;
          for (int a=0; a<3; ++a) {
            const double ww = wcb*w.wx[a];
            const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
            const double invrho = 1.0/s->u1[s->Idx(m,NG_VARS,DK_IDN,kk,jj,ii)];
            ugx += ww*s->u1[s->Idx(m,NG_VARS,DK_IM1,kk,jj,ii)]*invrho;
            ugy += ww*s->u1[s->Idx(m,NG_VARS,DK_IM2,kk,jj,ii)]*invrho;
            ugz += ww*s->u1[s->Idx(m,NG_VARS,DK_IM3,kk,jj,ii)]*invrho;
          }
        }
      const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
      const double rate = s->R(DK_IPM,p)/(s->R(DK_IPTS,p)*vol);
      const double gx = -dt*rate*(ugx-s->R(DK_IPVX,p));
      const double gy = -dt*rate*(ugy-s->R(DK_IPVY,p));
      const double gz = -dt*rate*(ugz-s->R(DK_IPVZ,p));
      for (int c=0; c<=2; ++c)
        for (int b=0; b<3; ++b) {
          const double wcb = w.wz[c]*w.wy[b];
          if (wcb == 0.0) continue;
          for (int a=0; a<3; ++a) {
            const double ww = wcb*w.wx[a];
            const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
            DepositAdd(s, &s->qdep[s->Idx(m,5,1,kk,jj,ii)], ww*gx);
            DepositAdd(s, &s->qdep[s->Idx(m,5,2,kk,jj,ii)], ww*gy);
            DepositAdd(s, &s->qdep[s->Idx(m,5,3,kk,jj,ii)], ww*gz);
            if (rate_dep) DepositAdd(s, &s->qdep[s->Idx(m,5,4,kk,jj,ii)], ww*rate);
          }
        }
    }
    return DK_OK;
  }
  Zero(s->qdep);
  const double a_dt = DragStep(s, st);
  const double sh = ImplicitShearShift(s, st);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t p=0; p<s->np; ++p) {
    const int m = s->I(DK_PGID,p);
    Stencil w;
    MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p) + sh*s->R(DK_IPX,p), s->R(DK_IPZ,p), w);
    if (!HaloOK(s, w)) {s->Fail(DK_EHALO, p); continue;}
    const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
    const double cj = a_dt/(s->R(DK_IPTS,p) + a_dt);
    const double muc = s->R(DK_IPM,p)*cj/vol;
    const double vx = s->R(DK_IPVX,p), vy = s->R(DK_IPVY,p), vz = s->R(DK_IPVZ,p);
    for (int c=0; c<=2; ++c)
      for (int b=0; b<3; ++b) {
        const double wcb = w.wz[c]*w.wy[b]*muc;
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
          DepositAdd(s, &s->qdep[s->Idx(m,5,0,kk,jj,ii)], ww);
          DepositAdd(s, &s->qdep[s->Idx(m,5,1,kk,jj,ii)], ww*vx);
          DepositAdd(s, &s->qdep[s->Idx(m,5,2,kk,jj,ii)], ww*vy);
          DepositAdd(s, &s->qdep[s->Idx(m,5,3,kk,jj,ii)], ww*vz);
        }
      }
  }
  return s->err;
}

// src/dust/dust_tasks.cpp:249-263 (the shear part of the remap at t^n: InitRecvDep,
// dust_tasks.cpp:223-225)
int or_fold_deposit_qp(or_sim *s, const dk_stage *st) {
  const bool use = (s->P.coupling == DK_IMEX) || st->stage == 1;
  // (a refined mesh: the clouds deposited into the leaves themselves, no ghost holds any)
  if (!ActiveStage(s, st) || !s->P.back_reaction || !use || s->smr) return DK_OK;
  Fold(s, s->qdep, 5, st->time);
  return DK_OK;
}

namespace {

// src/dust/dust_solver.cpp:126-203 (PointOperatorParticles): result += G^T (m c/V) G field
void PointOperatorParticles(or_sim *s, const std::vector<double> &field,
                            std::vector<double> &result, double a_dt, double sh) {
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t p=0; p<s->np; ++p) {
    const int m = s->I(DK_PGID,p);
    Stencil w;
    MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p) + sh*s->R(DK_IPX,p), s->R(DK_IPZ,p), w);
    double g[3] = {0.0, 0.0, 0.0};
    for (int c=0; c<=2; ++c)
      for (int b=0; b<3; ++b) {
        const double wcb = w.wz[c]*w.wy[b];
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
          g[0] += ww*field[s->Idx(m,3,0,kk,jj,ii)];
          g[1] += ww*field[s->Idx(m,3,1,kk,jj,ii)];
          g[2] += ww*field[s->Idx(m,3,2,kk,jj,ii)];
        }
      }
    const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
    const double cj = a_dt/(s->R(DK_IPTS,p) + a_dt);
    const double fac = s->R(DK_IPM,p)*cj/vol;
    for (int c=0; c<=2; ++c)
      for (int b=0; b<3; ++b) {
        const double wcb = w.wz[c]*w.wy[b]*fac;
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
          DepositAdd(s, &result[s->Idx(m,3,0,kk,jj,ii)], ww*g[0]);
          DepositAdd(s, &result[s->Idx(m,3,1,kk,jj,ii)], ww*g[1]);
          DepositAdd(s, &result[s->Idx(m,3,2,kk,jj,ii)], ww*g[2]);
        }
      }
  }
}

// the u* ghost fill of the stage (periodic copy, shear remap at t^n, x3 values)
void FillVelocityGhosts(or_sim *s, std::vector<double> &f, double t) {
  CopyGhosts(s, f, 3);
  if (s->M.bc_x1 == DK_X1_SHEAR) ShearGhostsX1(s, f, 3, t);
  X3BCs(s, f, 3, true);
}

// src/dust/dust_solver.cpp:226-263 (ApplyCoupledOperatorImpl)
void ApplyCoupledOperatorImpl(or_sim *s, std::vector<double> &field, std::vector<double> &result,
                              double a_dt, double sh, bool correct, double t) {
  FillVelocityGhosts(s, field, t);
  Zero(result);
  PointOperatorParticles(s, field, result, a_dt, sh);
  Fold(s, result, 3, t);
  const int nmb = s->M.nmb;
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i) {
          const double rho = s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)];
          if (correct) {
            const double pinv = 1.0/(rho + s->qdep[s->Idx(m,5,0,k,j,i)]);
            for (int d=0; d<3; ++d) {
              const double av = result[s->Idx(m,3,d,k,j,i)] + rho*field[s->Idx(m,3,d,k,j,i)];
              const double rv = s->u0[s->Idx(m,NG_VARS,DK_IM1+d,k,j,i)] +
                                s->qdep[s->Idx(m,5,1+d,k,j,i)] - av;
              field[s->Idx(m,3,d,k,j,i)] += pinv*rv;
            }
          } else {
            for (int d=0; d<3; ++d)
              result[s->Idx(m,3,d,k,j,i)] += rho*field[s->Idx(m,3,d,k,j,i)];
          }
        }
}

// ---- the strict solver (src/dust/dust_solver.cpp:268-625) -------------------------------
extern "C++" {   // (templates, inside the C-linkage region)
// Sums over the active cells, serial and in cell order (m, k, j, i), of a quantity per velocity
// component; b = M + P is the right-hand side, P = diag(rho + Q) the row-sum preconditioner.

double CellVolume(const or_sim *s, int m) { return s->dx1[m]*s->dx2[m]*s->dx3[m]; }
double RhsB(const or_sim *s, int m, int d, int k, int j, int i) {
  return s->u0[s->Idx(m,NG_VARS,DK_IM1+d,k,j,i)] + s->qdep[s->Idx(m,5,1+d,k,j,i)];
}
double PrecondDiag(const or_sim *s, int m, int k, int j, int i) {
  return s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)] + s->qdep[s->Idx(m,5,0,k,j,i)];
}

template <class F>
void CellSums3(const or_sim *s, F f, double out[3]) {
  out[0] = out[1] = out[2] = 0.0;
  for (int m=0; m<s->M.nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i)
          for (int d=0; d<3; ++d) out[d] += f(m, d, k, j, i);
}

template <class F>
void ForActive3(or_sim *s, F f) {
  for (int m=0; m<s->M.nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i)
          for (int d=0; d<3; ++d) f(m, d, k, j, i);
}

// GlobalDot: sum of vol a b per component
void VolDot(const or_sim *s, const std::vector<double> &a, const std::vector<double> &b,
            double out[3]) {
  CellSums3(s, [&](int m, int d, int k, int j, int i) {
    const int64_t id = s->Idx(m,3,d,k,j,i);
    return CellVolume(s, m)*a[id]*b[id];
  }, out);
}

// GlobalPreconditionedNorm: sum of vol r^2/(rho + Q) per component
void PrecondNorm(const or_sim *s, const std::vector<double> &r, double out[3]) {
  CellSums3(s, [&](int m, int d, int k, int j, int i) {
    const double v = r[s->Idx(m,3,d,k,j,i)];
    return CellVolume(s, m)/PrecondDiag(s, m, k, j, i)*v*v;
  }, out);
}

// StrictPCG: three independent preconditioned CG recurrences sharing the operator passes;
// returns the completed updates, or -1 (the reference's SolverFatal)
int StrictPCG(or_sim *s, double a_dt, double sh, double t, bool residual_is_current) {
  std::vector<double> &x = s->ustar, &r = s->sr, &p = s->sp, &ap = s->ap;
  if (!residual_is_current) {
    ApplyCoupledOperatorImpl(s, x, ap, a_dt, sh, false, t);
    ForActive3(s, [&](int m, int d, int k, int j, int i) {
      const int64_t id = s->Idx(m,3,d,k,j,i);
      r[id] = RhsB(s, m, d, k, j, i) - ap[id];
      p[id] = (1.0/PrecondDiag(s, m, k, j, i))*r[id];
    });
  } else {
    ForActive3(s, [&](int m, int d, int k, int j, int i) {
      const int64_t id = s->Idx(m,3,d,k,j,i);
      p[id] = (1.0/PrecondDiag(s, m, k, j, i))*r[id];
    });
  }
  double rz[3], bn2[3], target[3];
  bool conv[3];
  VolDot(s, r, p, rz);
  CellSums3(s, [&](int m, int d, int k, int j, int i) {
    const double b = RhsB(s, m, d, k, j, i);
    return CellVolume(s, m)/PrecondDiag(s, m, k, j, i)*b*b;
  }, bn2);
  for (int d=0; d<3; ++d) {
    target[d] = s->drag_atol + s->drag_rtol*std::sqrt(std::max(bn2[d], 0.0));
    conv[d] = std::sqrt(std::max(rz[d], 0.0)) <= target[d];
  }
  for (int it=0; it<=s->drag_iter_max; ++it) {
    if (conv[0] && conv[1] && conv[2]) {
      // the recursive residual says converged: verify with the true residual
      ApplyCoupledOperatorImpl(s, x, ap, a_dt, sh, false, t);
      ForActive3(s, [&](int m, int d, int k, int j, int i) {
        const int64_t id = s->Idx(m,3,d,k,j,i);
        r[id] = RhsB(s, m, d, k, j, i) - ap[id];
      });
      double tn2[3];
      PrecondNorm(s, r, tn2);
      bool all = true;
      s->solver_last[2] = 0.0;
      for (int d=0; d<3; ++d) {
        const double nrm = std::sqrt(std::max(tn2[d], 0.0));
        s->solver_last[2] = std::max(s->solver_last[2], nrm);
        conv[d] = nrm <= target[d];
        all = all && conv[d];
        rz[d] = tn2[d];
      }
      if (all) return it;
      ForActive3(s, [&](int m, int d, int k, int j, int i) {
        const int64_t id = s->Idx(m,3,d,k,j,i);
        p[id] = (1.0/PrecondDiag(s, m, k, j, i))*r[id];
      });
    }
    if (it == s->drag_iter_max) break;
    ApplyCoupledOperatorImpl(s, p, ap, a_dt, sh, false, t);
    double pap[3], alpha[3];
    VolDot(s, p, ap, pap);
    for (int d=0; d<3; ++d) {
      if (!conv[d] && (!(pap[d] > 0.0) || !std::isfinite(pap[d]))) return -1;
      alpha[d] = conv[d] ? 0.0 : rz[d]/pap[d];
    }
    ForActive3(s, [&](int m, int d, int k, int j, int i) {
      const int64_t id = s->Idx(m,3,d,k,j,i);
      x[id] += alpha[d]*p[id];
      r[id] -= alpha[d]*ap[id];
    });
    double rzn[3], beta[3];
    bool now[3];
    PrecondNorm(s, r, rzn);
    for (int d=0; d<3; ++d) {
      now[d] = std::sqrt(std::max(rzn[d], 0.0)) <= target[d];
      beta[d] = (conv[d] || now[d]) ? 0.0 : rzn[d]/rz[d];
    }
    ForActive3(s, [&](int m, int d, int k, int j, int i) {
      const int64_t id = s->Idx(m,3,d,k,j,i);
      const double pinv = 1.0/PrecondDiag(s, m, k, j, i);
      p[id] = now[d] ? 0.0 : pinv*r[id] + beta[d]*p[id];
    });
    for (int d=0; d<3; ++d) {conv[d] = now[d]; rz[d] = rzn[d];}
  }
  return -1;
}

// AdaptiveErrorBound: the conservative bound response*residual on the error of the DC1 iterate,
// and the acceptance target from the mass-weighted state scale
double AdaptiveBound(or_sim *s, double a_dt, double outer_dt, double residual,
                     double &state_scale, double &target) {
  double eps = 0.0;
  for (int m=0; m<s->M.nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i)
          eps = std::fmax(eps, s->qdep[s->Idx(m,5,0,k,j,i)]/s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)]);
  double cmax = 0.0;
  for (int64_t p=0; p<s->np; ++p) cmax = std::fmax(cmax, a_dt/(s->R(DK_IPTS,p) + a_dt));
  s->solver_last[6] = eps;
  const std::vector<double> &x = s->ustar;
  double gas = 0.0;
  for (int m=0; m<s->M.nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i) {
          const double u2 = x[s->Idx(m,3,0,k,j,i)]*x[s->Idx(m,3,0,k,j,i)] +
                            x[s->Idx(m,3,1,k,j,i)]*x[s->Idx(m,3,1,k,j,i)] +
                            x[s->Idx(m,3,2,k,j,i)]*x[s->Idx(m,3,2,k,j,i)];
          gas += CellVolume(s, m)*s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)]*u2;
        }
  double part = 0.0;
  for (int64_t p=0; p<s->np; ++p) {
    const int m = s->I(DK_PGID,p);
    Stencil w;
    MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p), s->R(DK_IPZ,p), w);
    double gu[3] = {0.0, 0.0, 0.0};
    for (int c=0; c<3; ++c)
      for (int b=0; b<3; ++b)
        for (int a=0; a<3; ++a) {
          const double ww = w.wz[c]*w.wy[b]*w.wx[a];
          for (int d=0; d<3; ++d)
            gu[d] += ww*x[s->Idx(m,3,d,w.kp+c-1,w.jp+b-1,w.ip+a-1)];
        }
    const double cj = a_dt/(s->R(DK_IPTS,p) + a_dt);
    double v2 = 0.0;
    for (int d=0; d<3; ++d) {
      const double v = s->R(DK_IPVX + 2*d,p) + cj*(gu[d] - s->R(DK_IPVX + 2*d,p));
      v2 += v*v;
    }
    part += s->R(DK_IPM,p)*v2;
  }
  state_scale = std::sqrt(std::max(gas + part, 0.0));
  const double theta = eps/(1.0 + eps);
  const double response = (1.0 + eps)*std::sqrt((1.0 + eps)*theta*theta + cmax*theta);
  const double rtol_step = std::min(s->adaptive_rtol_max,
                                    s->adaptive_order_c*std::pow(outer_dt/s->adaptive_tref, 3));
  target = s->drag_atol + rtol_step*std::max(state_scale, s->adaptive_state_floor);
  return response*residual;
}

}  // extern "C++"

}  // namespace

// src/dust/dust_pm.cpp:372-442 and src/dust/dust_solver.cpp:630-724
int or_gas_implicit_solve(or_sim *s, const dk_stage *st) {
  if (!ActiveStage(s, st)) return DK_OK;
  const int nmb = s->M.nmb;
  if (s->P.coupling == DK_PC2) {
    if (st->stage == 1 && s->P.back_reaction) {
      const bool damp = s->P.pc2_damping;
      const double dt = st->dt;
      // UpdateStiffness (dust_pm.cpp:288-342): chi = dt max_cells qdep[4]/rho over the active
      // cells, before the predictor is applied; the running maximum kept beside it
      if (s->P.stiffness || s->P.pc2_damping) {   // StiffnessActive (dust_pm.cpp:377)
        double rate = 0.0;
        for (int m=0; m<nmb; ++m)
          for (int k=s->ks; k<=s->ke; ++k)
            for (int j=s->js; j<=s->je; ++j)
              for (int i=s->is; i<=s->ie; ++i)
                rate = std::fmax(rate, s->qdep[s->Idx(m,5,4,k,j,i)]/
                                       s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)]);
        s->stiff_rate_max = rate;
        s->stiff_chi_last = dt*rate;
        s->stiff_chi_run_max = std::max(s->stiff_chi_run_max, s->stiff_chi_last);
      }
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
      for (int m=0; m<nmb; ++m)
        for (int k=s->ks; k<=s->ke; ++k)
          for (int j=s->js; j<=s->je; ++j)
            for (int i=s->is; i<=s->ie; ++i) {
              if (damp) {
                const double chi = dt*s->qdep[s->Idx(m,5,4,k,j,i)]/
                                   s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)];
                if (chi > 1.0) {
                  const double sc = 1.0/chi;
                  for (int c=1; c<=3; ++c) s->qdep[s->Idx(m,5,c,k,j,i)] *= sc;
                }
              }
              for (int c=0; c<3; ++c)
                s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)] += s->qdep[s->Idx(m,5,1+c,k,j,i)];
            }
    }
    return DK_OK;
  }
  const bool br = s->P.back_reaction;
  const double a_dt = DragStep(s, st);
  const bool coupled = br && s->P.solver != DK_LOCAL;
  // the local guess (dust_pm.cpp:426-436; dust_solver.cpp:643-649: the same arithmetic)
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i) {
          const double qd0 = br ? s->qdep[s->Idx(m,5,0,k,j,i)] : 0.0;
          const double denom = 1.0/(s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)] + qd0);
          for (int c=0; c<3; ++c) {
            const double qd = br ? s->qdep[s->Idx(m,5,1+c,k,j,i)] : 0.0;
            s->ustar[s->Idx(m,3,c,k,j,i)] = (s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)] + qd)*denom;
          }
        }
  if (!coupled) return DK_OK;
  const double sh = ImplicitShearShift(s, st);
  for (int q=0; q<7; ++q) s->solver_last[q] = (q < 2) ? 0.0 : -1.0;
  if (s->P.solver == DK_APPLYA) {
    ApplyCoupledOperatorImpl(s, s->ustar, s->ap, a_dt, sh, false, st->time);
  } else if (s->P.solver == DK_PCG) {
    const int it = StrictPCG(s, a_dt, sh, st->time, false);
    if (it < 0) return DK_ESOLVER;
    s->solver_last[0] = it;
  } else {
    const int nsweeps = (s->P.solver == DK_DC2) ? 2 : 1;
    for (int sw=0; sw<nsweeps; ++sw)
      ApplyCoupledOperatorImpl(s, s->ustar, s->ap, a_dt, sh, true, st->time);
    if (s->P.solver == DK_ADAPTIVE) {
      // the true residual of the DC1 iterate, its error bound, and PCG from it if too large
      ApplyCoupledOperatorImpl(s, s->ustar, s->ap, a_dt, sh, false, st->time);
      ForActive3(s, [&](int m, int d, int k, int j, int i) {
        const int64_t id = s->Idx(m,3,d,k,j,i);
        s->sr[id] = RhsB(s, m, d, k, j, i) - s->ap[id];
      });
      double n2[3];
      PrecondNorm(s, s->sr, n2);
      const double res = std::sqrt(std::max(n2[0] + n2[1] + n2[2], 0.0));
      double scale, target;
      const double bound = AdaptiveBound(s, a_dt, st->dt, res, scale, target);
      s->solver_last[2] = res; s->solver_last[3] = bound; s->solver_last[4] = target;
      s->solver_last[5] = scale;
      if (bound <= target) {
        s->solver_last[1] = 1.0;
      } else {
        const int it = StrictPCG(s, a_dt, sh, st->time, true);
        if (it < 0) return DK_ESOLVER;
        s->solver_last[0] = it;
      }
    }
  }
  return DK_OK;
}

void or_set_solver(or_sim *s, double rtol, double atol, int iter_max, double order_c,
                   double rtol_max, double tref, double state_floor) {
  s->drag_rtol = rtol; s->drag_atol = atol; s->drag_iter_max = iter_max;
  s->adaptive_order_c = order_c; s->adaptive_rtol_max = rtol_max; s->adaptive_tref = tref;
  s->adaptive_state_floor = state_floor;
}

void or_stiffness(or_sim *s, double out[3]) {
  out[0] = s->stiff_chi_last;
  out[1] = s->stiff_chi_run_max;
  out[2] = s->stiff_rate_max;
}

void or_solver_stats(or_sim *s, double out[7]) {
  for (int q=0; q<7; ++q) out[q] = s->solver_last[q];
}

// src/dust/dust_tasks.cpp:265-312
int or_fill_ustar(or_sim *s, const dk_stage *st) {
  if (!ActiveStage(s, st) || s->P.coupling != DK_IMEX) return DK_OK;
  FillVelocityGhosts(s, s->ustar, st->time);
  return DK_OK;
}

// src/dust/dust_pm.cpp:450-568
int or_gather_kick_pmbr(or_sim *s, const dk_stage *st) {
  if (!ActiveStage(s, st) || s->P.coupling == DK_PC2) return DK_OK;
  const bool br = s->P.back_reaction;
  if (br) Zero(s->dmom);
  const double a_dt = DragStep(s, st);
  const double sh = ImplicitShearShift(s, st);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int64_t p=0; p<s->np; ++p) {
    const int m = s->I(DK_PGID,p);
    Stencil w;
    MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p) + sh*s->R(DK_IPX,p), s->R(DK_IPZ,p), w);
    double ux = 0.0, uy = 0.0, uz = 0.0;
    for (int c=0; c<=2; ++c)
      for (int b=0; b<3; ++b) {
        const double wcb = w.wz[c]*w.wy[b];
        if (wcb == 0.0) continue;
        for (int a=0; a<3; ++a) {
          const double ww = wcb*w.wx[a];
          const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
          ux += ww*s->ustar[s->Idx(m,3,0,kk,jj,ii)];
          uy += ww*s->ustar[s->Idx(m,3,1,kk,jj,ii)];
          uz += ww*s->ustar[s->Idx(m,3,2,kk,jj,ii)];
        }
      }
    const double cj = a_dt/(s->R(DK_IPTS,p) + a_dt);
    const double dvx = cj*(ux - s->R(DK_IPVX,p));
    const double dvy = cj*(uy - s->R(DK_IPVY,p));
    const double dvz = cj*(uz - s->R(DK_IPVZ,p));
    s->R(DK_IPVX,p) += dvx;
    s->R(DK_IPVY,p) += dvy;
    s->R(DK_IPVZ,p) += dvz;
    s->R(DK_IPRX,p) = dvx/a_dt;
    s->R(DK_IPRY,p) = dvy/a_dt;
    s->R(DK_IPRZ,p) = dvz/a_dt;
    if (br) {
      const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
      const double fac = -s->R(DK_IPM,p)/vol;
      for (int c=0; c<=2; ++c)
        for (int b=0; b<3; ++b) {
          const double wcb = w.wz[c]*w.wy[b]*fac;
          if (wcb == 0.0) continue;
          for (int a=0; a<3; ++a) {
            const double ww = wcb*w.wx[a];
            const int kk = w.kp+c-1, jj = w.jp+b-1, ii = w.ip+a-1;
            DepositAdd(s, &s->dmom[s->Idx(m,3,0,kk,jj,ii)], ww*dvx);
            DepositAdd(s, &s->dmom[s->Idx(m,3,1,kk,jj,ii)], ww*dvy);
            DepositAdd(s, &s->dmom[s->Idx(m,3,2,kk,jj,ii)], ww*dvz);
          }
        }
    }
  }
  return DK_OK;
}

// src/dust/dust_tasks.cpp:314-328
int or_fold_pmbr(or_sim *s, const dk_stage *st) {
  const bool use = (s->P.coupling == DK_IMEX) || st->stage == 2;
  if (!ActiveStage(s, st) || !s->P.back_reaction || !use || s->smr) return DK_OK;
  Fold(s, s->dmom, 3, st->time);
  return DK_OK;
}

// src/dust/dust_pm.cpp:576-711
int or_apply_pmbr(or_sim *s, const dk_stage *st) {
  if (!ActiveStage(s, st) || !s->P.back_reaction) return DK_OK;
  const int nmb = s->M.nmb;
  if (s->P.coupling == DK_PC2) {
    if (st->stage != 2) return DK_OK;
    const double wpred = st->gam1[st->nexp_stages-1];
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
    for (int m=0; m<nmb; ++m)
      for (int k=s->ks; k<=s->ke; ++k)
        for (int j=s->js; j<=s->je; ++j)
          for (int i=s->is; i<=s->ie; ++i)
            for (int c=0; c<3; ++c)
              s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)] +=
                  s->dmom[s->Idx(m,3,c,k,j,i)] - wpred*s->qdep[s->Idx(m,5,1+c,k,j,i)];
    return DK_OK;
  }
  const double inv_adt = 1.0/(st->a_impl*st->dt);
  if (s->P.consistent_gas) {
    const double a_dt = DragStep(s, st);
    Zero(s->resid);
    // dust_imex_point_center_mass (dust_pm.cpp:660-667); the centre uses the unshifted y
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t p=0; p<s->np; ++p) {
      if (s->R(DK_IPM,p) <= 0.0) continue;
      const int m = s->I(DK_PGID,p);
      Stencil w;
      MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p), s->R(DK_IPZ,p), w);
      const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
      DepositAdd(s, &s->resid[s->Idx(m,4,3,w.kp,w.jp,w.ip)],
                 s->R(DK_IPM,p)*(a_dt/(s->R(DK_IPTS,p) + a_dt))/vol);
    }
    // dust_imex_point_consistent_gas (dust_pm.cpp:670-685)
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
    for (int m=0; m<nmb; ++m)
      for (int k=s->ks; k<=s->ke; ++k)
        for (int j=s->js; j<=s->je; ++j)
          for (int i=s->is; i<=s->ie; ++i) {
            const double rho = s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)];
            const double qt = s->resid[s->Idx(m,4,3,k,j,i)];
            const double q = s->qdep[s->Idx(m,5,0,k,j,i)];
            for (int c=0; c<3; ++c) {
              double dgas = s->dmom[s->Idx(m,3,c,k,j,i)];
              if (qt > 0.0) {
                const double g = rho*s->ustar[s->Idx(m,3,c,k,j,i)] -
                                 s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)];
                const double r = (s->dmom[s->Idx(m,3,c,k,j,i)] - g)/(rho + q);
                dgas = g + rho*r;
                s->resid[s->Idx(m,4,c,k,j,i)] = r*q/qt;
              }
              s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)] += dgas;
              s->dmom[s->Idx(m,3,c,k,j,i)] = dgas*inv_adt;
            }
          }
    // dust_imex_point_consistent_dust (dust_pm.cpp:686-696)
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
    for (int64_t p=0; p<s->np; ++p) {
      if (s->R(DK_IPM,p) <= 0.0) continue;
      const int m = s->I(DK_PGID,p);
      Stencil w;
      MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p), s->R(DK_IPZ,p), w);
      const double cj = a_dt/(s->R(DK_IPTS,p) + a_dt);
      for (int c=0; c<3; ++c) {
        const double dv = cj*s->resid[s->Idx(m,4,c,w.kp,w.jp,w.ip)];
        s->R(DK_IPVX+2*c,p) += dv;
        s->R(DK_IPRX+c,p) += dv/a_dt;
      }
    }
    return DK_OK;
  }
  // dust_pmbr (dust_pm.cpp:700-708)
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i)
          for (int c=0; c<3; ++c) {
            s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)] += s->dmom[s->Idx(m,3,c,k,j,i)];
            s->dmom[s->Idx(m,3,c,k,j,i)] *= inv_adt;
          }
  return DK_OK;
}

// src/dust/dust_pm.cpp:721-742
int or_add_drag_history_gas(or_sim *s, const dk_stage *st) {
  if (s->P.coupling != DK_IMEX) return DK_OK;
  if (st->stage != 2 || !s->P.back_reaction) return DK_OK;
  const double atw_dt = st->a_twid22*st->dt;
  const int nmb = s->M.nmb;
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i)
          for (int c=0; c<3; ++c)
            s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)] += atw_dt*s->dmom[s->Idx(m,3,c,k,j,i)];
  return DK_OK;
}

namespace {

// reconstruct/plm.hpp:13-28: the limited slope of q_i (van Leer's harmonic mean on a uniform
// mesh) gives the left state of face i+1/2 and the right state of face i-1/2
void Plm(double qm, double q0, double qp, double &ql_next, double &qr_here) {
  const double dql = q0 - qm, dqr = qp - q0;
  const double dq2 = dql*dqr;
  double dqm = dq2/(dql + dqr);
  if (dq2 <= 0.0) dqm = 0.0;
  ql_next = q0 + dqm;
  qr_here = q0 - dqm;
}

// src/hydro/rsolvers/hlle_hyd.hpp:22-150 with the isothermal EOS: L/R states (d, v_n, v_t1,
// v_t2) in the face's frame, the flux in the same order
void HlleIso(const double wl[4], const double wr[4], double cs, double f[4]) {
  const double sdl = std::sqrt(wl[0]), sdr = std::sqrt(wr[0]);
  const double isd = 1.0/(sdl + sdr);
  const double vroe = (sdl*wl[1] + sdr*wr[1])*isd;
  const double al = std::fmin(vroe - cs, wl[1] - cs);
  const double ar = std::fmax(vroe + cs, wr[1] + cs);
  const double bp = (ar > 0.0) ? ar : 1.0e-20;
  const double bm = (al < 0.0) ? al : -1.0e-20;
  const double qa = wl[1] - bm, qb = wr[1] - bp;
  double fl[4], fr[4];
  fl[0] = wl[0]*qa;
  fr[0] = wr[0]*qb;
  for (int c=1; c<4; ++c) {
    fl[c] = wl[0]*wl[c]*qa;
    fr[c] = wr[0]*wr[c]*qb;
  }
  fl[1] += (cs*cs)*wl[0];
  fr[1] += (cs*cs)*wr[0];
  double h = 0.0;
  if (bp != bm) h = 0.5*(bp + bm)/(bp - bm);
  for (int c=0; c<4; ++c) f[c] = 0.5*(fl[c] + fr[c]) + h*(fl[c] - fr[c]);
}

// The options the explicit gas update supports (dk_gas_src in dustk.h)
int HydroSupported(const or_sim *s, const dk_gas_src *src) {
  if (!(src->iso_cs > 0.0)) return DK_EINVAL;
  if (s->smr || s->ng < 2 || s->M.bc_x1 != DK_X1_PERIODIC) return DK_EUNSUPPORTED;
  return DK_OK;
}

// Hydro::Fluxes (src/hydro/hydro_fluxes.cpp:20-140, plm + hlle) and the flux divergence of
// Hydro::RKUpdate (src/hydro/hydro_update.cpp:55-77) from the primitives of u0, ghosts
// included (IsothermalHydro::ConsToPrim, src/eos/isothermal_hyd.cpp:34-110; the density floor
// is applied where a density is read): divf (4 components, the gas layout, active cells)
void HydroDivf(or_sim *s, const dk_gas_src *src, std::vector<double> &divf) {
  const int nmb = s->M.nmb;
  const int64_t nc = (int64_t)nmb*NG_VARS*s->ncb;
  std::vector<double> w(nc);
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m)
    for (int k=0; k<s->n3; ++k)
      for (int j=0; j<s->n2; ++j)
        for (int i=0; i<s->n1; ++i) {
          double d = s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)];
          if (d < src->dfloor) d = src->dfloor;
          const double di = 1.0/d;
          w[s->Idx(m,NG_VARS,0,k,j,i)] = d;
          for (int c=1; c<4; ++c)
            w[s->Idx(m,NG_VARS,c,k,j,i)] = di*s->u0[s->Idx(m,NG_VARS,c,k,j,i)];
        }
  // face fluxes: flx[d] at face index f of axis d lies between cells f-1 and f
  std::vector<double> flx[3];
  const int64_t stride[3] = {1, s->n1, (int64_t)s->n1*s->n2};
  for (int d=0; d<3; ++d) {
    flx[d].assign(nc, 0.0);
    const int iu = s->ie + (d == 0), ju = s->je + (d == 1), ku = s->ke + (d == 2);
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
    for (int m=0; m<nmb; ++m)
      for (int k=s->ks; k<=ku; ++k)
        for (int j=s->js; j<=ju; ++j)
          for (int i=s->is; i<=iu; ++i) {
            double wl[4], wr[4], f[4], dummy;
            for (int v=0; v<4; ++v) {
              // the face frame's component v: density, then v_d, v_d+1, v_d+2
              const int c = (v == 0) ? 0 : 1 + (d + v - 1) % 3;
              const int64_t at = s->Idx(m,NG_VARS,c,k,j,i);
              const int64_t sd = stride[d];
              Plm(w[at - 2*sd], w[at - sd], w[at], wl[v], dummy);
              Plm(w[at - sd], w[at], w[at + sd], dummy, wr[v]);
            }
            HlleIso(wl, wr, src->iso_cs, f);
            for (int v=0; v<4; ++v) {
              const int c = (v == 0) ? 0 : 1 + (d + v - 1) % 3;
              flx[d][s->Idx(m,NG_VARS,c,k,j,i)] = f[v];
            }
          }
  }
  divf.assign(nc, 0.0);
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m) {
    const double dx[3] = {s->dx1[m], s->dx2[m], s->dx3[m]};
    for (int v=0; v<NG_VARS; ++v)
      for (int k=s->ks; k<=s->ke; ++k)
        for (int j=s->js; j<=s->je; ++j)
          for (int i=s->is; i<=s->ie; ++i) {
            const int64_t at = s->Idx(m,NG_VARS,v,k,j,i);
            double dv = (flx[0][at + stride[0]] - flx[0][at])/dx[0];
            dv += (flx[1][at + stride[1]] - flx[1][at])/dx[1];
            dv += (flx[2][at + stride[2]] - flx[2][at])/dx[2];
            divf[at] = dv;
          }
  }
}

}  // namespace

// Hydro::RKUpdate (src/hydro/hydro_update.cpp:23-83; zero fluxes unless src->hydro) and
// HydroSrcTerms (src/hydro/hydro_tasks.cpp:268-297: ConstantAccel srcterms.cpp:128-145,
// ShearingBoxCC::SourceTermsCC shearing_box_srcterms.cpp:26-60), primitives from u0 at the
// stage start
int or_gas_hook(or_sim *s, const dk_stage *st, const dk_gas_src *src) {
  const int sidx = st->stage - 1;
  const double gam0 = st->gam0[sidx], gam1 = st->gam1[sidx];
  const double bdt = st->beta[sidx]*st->dt;
  const bool hydro = (src != nullptr && src->hydro);
  if (hydro) {
    const int e = HydroSupported(s, src);
    if (e) return e;
  }
  if (bdt == 0.0 && gam0 == 1.0 && gam1 == 0.0) return DK_OK;   // identity stage
  std::vector<double> divf;
  if (hydro) {
    // the u0 ghosts of the stage start (the previous stage's Hydro::SendU/RecvU and
    // ApplyPhysicalBCs), then the flux divergence
    CopyGhosts(s, s->u0, NG_VARS);
    X3BCs(s, s->u0, NG_VARS, true);
    HydroDivf(s, src, divf);
  }
  const int nmb = s->M.nmb;
#pragma omp parallel for collapse(2) num_threads(s->nthreads) schedule(static)
  for (int m=0; m<nmb; ++m)
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i) {
          double w[4];
          double den = s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)];
          if (hydro && den < src->dfloor) {
            den = src->dfloor;
            s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)] = den;
          }
          w[0] = den;
          for (int c=1; c<4; ++c) w[c] = s->u0[s->Idx(m,NG_VARS,c,k,j,i)]/den;
          for (int v=0; v<NG_VARS; ++v) {
            double &u = s->u0[s->Idx(m,NG_VARS,v,k,j,i)];
            if (hydro)
              u = gam0*u + gam1*s->u1[s->Idx(m,NG_VARS,v,k,j,i)] -
                  bdt*divf[s->Idx(m,NG_VARS,v,k,j,i)];
            else
              u = gam0*u + gam1*s->u1[s->Idx(m,NG_VARS,v,k,j,i)];
          }
          if (src == nullptr) continue;
          if (src->const_accel != 0.0) {
            s->u0[s->Idx(m,NG_VARS,src->const_accel_dir,k,j,i)] += bdt*src->const_accel*w[0];
          }
          if (src->shearing_box) {
            const double coef1 = 2.0*bdt*src->omega0;
            const double coef2 = (2.0-src->qshear)*bdt*src->omega0;
            const double coef3 = bdt*SQR(src->omega0);
            const double mom1 = w[0]*w[1], mom2 = w[0]*w[2];
            s->u0[s->Idx(m,NG_VARS,DK_IM1,k,j,i)] += coef1*mom2;
            s->u0[s->Idx(m,NG_VARS,DK_IM2,k,j,i)] -= coef2*mom1;
            if (src->stratified) {
              const dk_block &b = s->B[m];
              const double xx = (static_cast<double>(k - s->ks) + 0.5)/s->nx3;
              const double x3v = (xx*b.x3max - xx*b.x3min) - (0.5*b.x3max - 0.5*b.x3min) +
                                 (0.5*b.x3min + 0.5*b.x3max);
              s->u0[s->Idx(m,NG_VARS,DK_IM3,k,j,i)] -= coef3*w[0]*x3v;
            }
          }
        }
  return DK_OK;
}

// Hydro::SendU/RecvU/SendU_Shr/RecvU_Shr + ApplyPhysicalBCs stand-in on u0
int or_fill_gas_ghosts(or_sim *s, const dk_stage *st) {
  CopyGhosts(s, s->u0, NG_VARS);
  if (s->M.bc_x1 == DK_X1_SHEAR) ShearGhostsX1(s, s->u0, NG_VARS, st->time + st->dt);
  X3BCs(s, s->u0, NG_VARS, true);
  return DK_OK;
}

// The stage in the order of the combined task graph (src/dust/dust_tasks.cpp:58-109)
int or_stage_run(or_sim *s, const dk_stage *st, const dk_gas_src *src) {
  int e;
  if ((e = or_first_two_imp_rk(s, st))) return e;
  if ((e = or_gas_hook(s, st, src))) return e;
  if ((e = or_add_drag_history_gas(s, st))) return e;
  if ((e = or_explicit_push(s, st))) return e;
  if ((e = or_update_particle_gids(s, st))) return e;
  if ((e = or_deposit_drag(s, st))) return e;
  if ((e = or_fold_deposit_qp(s, st))) return e;
  if ((e = or_gas_implicit_solve(s, st))) return e;
  if ((e = or_fill_ustar(s, st))) return e;
  if ((e = or_gather_kick_pmbr(s, st))) return e;
  if ((e = or_fold_pmbr(s, st))) return e;
  if ((e = or_apply_pmbr(s, st))) return e;
  // the hydro tail: u0 ghosts for the next stage (pc2 gathers read them)
  if (ActiveStage(s, st)) {
    if ((e = or_fill_gas_ghosts(s, st))) return e;
  }
  return DK_OK;
}

// src/dust/dust_cloud_device.cpp:1881-1980: stable order by (block, cell of the position)
// DustGasDrag::RemoveOutflowParticles (src/dust/dust_tasks.cpp:365-420): with an outflow x3
// face and a non-zero ledger, every particle of zero mass is removed, the others keep their order
// (the exclusive scan of the keep flags); the ledger goes to the run totals and is cleared.
// out: this call's removed mass and count, then the run totals
int or_remove_outflow_particles(or_sim *s, double out[4]) {
  out[0] = out[1] = 0.0;
  if (s->M.bc_x3 == DK_X3_OUTFLOW && s->x3_removed_count != 0.0) {
    int64_t k = 0;
    for (int64_t p=0; p<s->np; ++p) {
      if (!(s->R(DK_IPM,p) > 0.0)) continue;
      if (k != p) {
        for (int c=0; c<DK_NRDATA; ++c) s->R(c,k) = s->R(c,p);
        for (int c=0; c<DK_NIDATA; ++c) s->I(c,k) = s->I(c,p);
      }
      ++k;
    }
    s->np = k;
    out[0] = s->x3_removed_mass;
    out[1] = s->x3_removed_count;
    s->x3_outflow_mass += s->x3_removed_mass;
    s->x3_outflow_count += s->x3_removed_count;
    s->x3_removed_mass = 0.0;
    s->x3_removed_count = 0.0;
  }
  out[2] = s->x3_outflow_mass;
  out[3] = s->x3_outflow_count;
  return DK_OK;
}

// the ledger since the last removal (mass, count)
void or_x3_removed(or_sim *s, double out[2]) {
  out[0] = s->x3_removed_mass;
  out[1] = s->x3_removed_count;
}

int or_sort_particles(or_sim *s) {
  const int64_t n = s->np;
  std::vector<int64_t> key(n), ord(n);
  for (int64_t p=0; p<n; ++p) {
    const int m = s->I(DK_PGID,p);
    Stencil w;
    MakeStencil(s, m, s->R(DK_IPX,p), s->R(DK_IPY,p), s->R(DK_IPZ,p), w);
    const int64_t cell = ((int64_t)(w.kp - s->ks)*s->nx2 + (w.jp - s->js))*s->nx1 + (w.ip - s->is);
    key[p] = (int64_t)m*s->nx1*s->nx2*s->nx3 + cell;
    ord[p] = p;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {return key[a] < key[b];});
  std::vector<double> nr(s->pr.size());
  std::vector<int32_t> ni(s->pi.size());
  for (int c=0; c<DK_NRDATA; ++c)
    for (int64_t q=0; q<n; ++q) nr[(int64_t)c*s->cap + q] = s->pr[(int64_t)c*s->cap + ord[q]];
  for (int c=0; c<DK_NIDATA; ++c)
    for (int64_t q=0; q<n; ++q) ni[(int64_t)c*s->cap + q] = s->pi[(int64_t)c*s->cap + ord[q]];
  s->pr.swap(nr);
  s->pi.swap(ni);
  return DK_OK;
}

// src/dust/dust_newdt.cpp:108-149: dt_cfl min dx/max(|v|, 1e-30), with the background shear in
// the azimuthal transport velocity of the 3-D shearing box; stopping times validated
// (dust_newdt.cpp:34-99)
int or_new_time_step(or_sim *s, double dt_cfl, double *dt_out) {
  double dtp = static_cast<double>(std::numeric_limits<float>::max());
  const double qo = s->P.qshear*s->P.omega0;
  for (int64_t p=0; p<s->np; ++p) {
    const int m = s->I(DK_PGID,p);
    const double ts = s->R(DK_IPTS,p);
    if (!std::isfinite(ts) || !(ts > 0.0)) return DK_ESTOP;
    double vy = s->R(DK_IPVY,p);
    if (s->P.shearing_box) vy -= qo*s->R(DK_IPX,p);
    dtp = std::fmin(s->dx1[m]/std::fmax(std::fabs(s->R(DK_IPVX,p)), 1.0e-30), dtp);
    dtp = std::fmin(s->dx2[m]/std::fmax(std::fabs(vy), 1.0e-30), dtp);
    dtp = std::fmin(s->dx3[m]/std::fmax(std::fabs(s->R(DK_IPVZ,p)), 1.0e-30), dtp);
  }
  if (dt_out) *dt_out = dt_cfl*dtp;
  return DK_OK;
}

int or_apply_operator(or_sim *s, double a_dt, double sh, const double *x, double *y) {
  const int64_t nc = 3*s->ncb*s->M.nmb;
  std::vector<double> f(x, x + nc), r(nc, 0.0);
  CopyGhosts(s, f, 3);
  PointOperatorParticles(s, f, r, a_dt, sh);
  Fold(s, r, 3, 0.0);
  for (int m=0; m<s->M.nmb; ++m)
    for (int d=0; d<3; ++d)
      for (int k=s->ks; k<=s->ke; ++k)
        for (int j=s->js; j<=s->je; ++j)
          for (int i=s->is; i<=s->ie; ++i) {
            const int64_t id = s->Idx(m,3,d,k,j,i);
            y[id] = r[id] + s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)]*f[id];
          }
  return DK_OK;
}

int or_cloud_cells(or_sim *s, int m, const double x[3], int64_t *cell, double *w, double *m1) {
  int n = 0;
  const bool ok = CloudForEach<true>(s, m, x,
      [&](int k, int j, int i, double ww, double a, double b, double c) {
        if (n < 27) {
          cell[n] = ((int64_t)k*s->n2 + j)*s->n1 + i;
          w[n] = ww; m1[3*n] = a; m1[3*n+1] = b; m1[3*n+2] = c;
        }
        ++n;
      });
  return (ok && n <= 27) ? n : -1;
}

int or_cloud_gather(or_sim *s, const double *vel, int m, const double x[3], double out[3]) {
  std::vector<double> v(vel, vel + 3*s->ncb*s->M.nmb);
  PackGatherField(s, v);
  double u[3] = {0.0, 0.0, 0.0};
  const bool ok = CloudForEach<true>(s, m, x,
      [&](int k, int j, int i, double ww, double a, double b, double c) {
        GatherLinear(s, m, k, j, i, ww, a, b, c, x, u);
      });
  for (int d=0; d<3; ++d) out[d] = u[d];
  return ok ? 0 : -1;
}

int or_leaf_cells(or_sim *s, const double x[3], int lev, int32_t *blk, int64_t *cell,
                  double *w, double *m1, int cap) {
  int n = 0;
  const bool ok = LeafForEach<true>(s, x, lev,
      [&](int b, int k, int j, int i, double ww, double a, double bb, double c) {
        if (n < cap) {
          blk[n] = b;
          cell[n] = ((int64_t)k*s->n2 + j)*s->n1 + i;
          w[n] = ww; m1[3*n] = a; m1[3*n+1] = bb; m1[3*n+2] = c;
        }
        ++n;
      });
  return (ok && n <= cap) ? n : -1;
}

int or_leaf_gather(or_sim *s, const double *vel, const double x[3], int lev, double out[3]) {
  std::vector<double> v(vel, vel + 3*s->ncb*s->M.nmb);
  PackGatherField(s, v);
  double u[3] = {0.0, 0.0, 0.0};
  const bool ok = LeafForEach<true>(s, x, lev,
      [&](int b, int k, int j, int i, double ww, double a, double bb, double c) {
        GatherLinear(s, b, k, j, i, ww, a, bb, c, x, u);
      });
  for (int d=0; d<3; ++d) out[d] = u[d];
  return ok ? 0 : -1;
}

int or_leaf_at(or_sim *s, const double x[3]) {
  if (!s->smr) return -1;
  double p[3] = {x[0], x[1], x[2]};
  int k[3];
  WrapPoint(s, p, k);
  return LeafAt(s, p);
}

int or_diag_sums(or_sim *s, double out[9]) {
  for (int c=0; c<9; ++c) out[c] = 0.0;
  for (int64_t p=0; p<s->np; ++p) {
    const double m = s->R(DK_IPM,p);
    out[0] += m*s->R(DK_IPVX,p);
    out[1] += m*s->R(DK_IPVY,p);
    out[2] += m*s->R(DK_IPVZ,p);
    out[6] += m;
    out[8] += m*(std::fabs(s->R(DK_IPVX,p)) + std::fabs(s->R(DK_IPVY,p)) +
                 std::fabs(s->R(DK_IPVZ,p)));
  }
  for (int m=0; m<s->M.nmb; ++m) {
    const double vol = s->dx1[m]*s->dx2[m]*s->dx3[m];
    for (int k=s->ks; k<=s->ke; ++k)
      for (int j=s->js; j<=s->je; ++j)
        for (int i=s->is; i<=s->ie; ++i) {
          out[7] += vol*s->u0[s->Idx(m,NG_VARS,DK_IDN,k,j,i)];
          for (int c=0; c<3; ++c) {
            const double mom = s->u0[s->Idx(m,NG_VARS,DK_IM1+c,k,j,i)];
            out[3+c] += vol*mom;
            out[8] += vol*std::fabs(mom);
          }
        }
  }
  return DK_OK;
}

// src/driver/driver.cpp:195-203,278-300 (gamma passed as the literal 1.707106781186547)
void or_imex2p_coef(double time, double dt, dk_stage *st) {
  const double gamma = 1.707106781186547;
  dk_stage c;
  std::memset(&c, 0, sizeof(c));
  c.nexp_stages = 3; c.time = time; c.dt = dt;
  c.gam0[0] = 1.0; c.gam1[0] = 0.0; c.beta[0] = gamma;
  c.gam0[1] = (2.0*gamma-1.0)/(2.0*gamma*gamma);
  c.gam1[1] = (1.0-(2.0*gamma-1.0)/(2.0*gamma*gamma));
  c.beta[1] = 1.0/(2.0*gamma);
  c.gam0[2] = 1.0; c.gam1[2] = 0.0; c.beta[2] = 0.0;
  c.a_twid22 = (1.0-2.0*gamma*gamma)/2.0/gamma;
  c.a_impl = gamma;
  for (int k=0; k<3; ++k) {st[k] = c; st[k].stage = k + 1;}
}

// src/driver/driver.cpp:106-118
void or_rk2_coef(double time, double dt, dk_stage *st) {
  dk_stage c;
  std::memset(&c, 0, sizeof(c));
  c.nexp_stages = 2; c.time = time; c.dt = dt;
  c.gam0[0] = 0.0; c.gam1[0] = 1.0; c.beta[0] = 1.0;
  c.gam0[1] = 0.5; c.gam1[1] = 0.5; c.beta[1] = 0.5;
  c.a_impl = 0.0;
  for (int k=0; k<2; ++k) {st[k] = c; st[k].stage = k + 1;}
}

}  // extern "C"
