// dk_mesh.hip -- mesh-side exchanges of the drag path on one device, for FP64 and FP32 fields.
//
// Ghost copy of a cell field (src/bvals/bvals_cc.cpp:53-296), the physical x3 values of u*
// (src/dust/dust_tasks.cpp:445-469), the shear-periodic x1 ghost remap
// (src/shearing_box/shearing_box_cc.cpp:68-260, remap_fluxes.hpp, shearing_box.cpp:260-330) and
// the additive ghost-deposit fold (src/bvals/bvals_dep.cpp:320-562 and the shear / x3 parts of
// src/bvals/bvals_dep_general.cpp).  Every kernel is a gather: each cell it writes is written by
// one thread, summing its inputs in a fixed order, so the results do not depend on scheduling.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dk_internal.hpp"
#include "dk_smr.hpp"

using dk::DevBlock;
using dk::Dir;
using dk::Geo;
using dk::Idx;
using dk::kBlock;

namespace {

#define GRID_STRIDE(t, n) \
  for (long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x; t < (n); \
       t += (long long)gridDim.x*blockDim.x)

inline int Grid(long long n) {
  long long b = (n + kBlock - 1)/kBlock;
  if (b < 1) b = 1;
  return static_cast<int>(std::min<long long>(b, 65536));
}

// x over n cells of one block, y over the blocks (at most 65535: checked at creation)
inline dim3 BlockGrid(long long n, int nmb) {
  const long long b = std::max<long long>(1, (n + kBlock - 1)/kBlock);
  return dim3(static_cast<unsigned>(std::min<long long>(b, 65535)), static_cast<unsigned>(nmb));
}

#define BLOCK_STRIDE(t, n) \
  for (unsigned t = blockIdx.x*blockDim.x + threadIdx.x; t < (n); t += gridDim.x*blockDim.x)

// the padded cells outside the active box of a block: GhostCount of them, the x3 ghost layers
// (whole planes), then the x2 ghost rows of the active layers, then the x1 ghost cells of the
// active rows (32-bit index arithmetic)
__host__ __device__ inline unsigned GhostCount(const Geo &g) {
  return 2u*g.ng*g.n2*g.n1 + static_cast<unsigned>(g.nx3)*2u*g.ng*g.n1 +
         static_cast<unsigned>(g.nx3)*g.nx2*2u*g.ng;
}

__device__ inline void GhostCell(const Geo &g, unsigned s, int &i, int &j, int &k) {
  const unsigned a = 2u*g.ng*g.n2*g.n1, b = static_cast<unsigned>(g.nx3)*2u*g.ng*g.n1;
  const unsigned ng2 = 2u*g.ng;
  if (s < a) {
    const unsigned plane = static_cast<unsigned>(g.n2)*g.n1;
    const unsigned t = s/plane, r = s - t*plane;
    const unsigned jj = r/g.n1;
    k = (t < static_cast<unsigned>(g.ng)) ? static_cast<int>(t) : g.ke + 1 + static_cast<int>(t) - g.ng;
    j = static_cast<int>(jj);
    i = static_cast<int>(r - jj*g.n1);
  } else if (s < a + b) {
    const unsigned r0 = s - a, rows = ng2*g.n1;
    const unsigned t = r0/rows, r = r0 - t*rows;
    const unsigned jj = r/g.n1;
    k = g.ks + static_cast<int>(t);
    j = (jj < static_cast<unsigned>(g.ng)) ? static_cast<int>(jj) : g.je + 1 + static_cast<int>(jj) - g.ng;
    i = static_cast<int>(r - jj*g.n1);
  } else {
    const unsigned r0 = s - a - b, row = static_cast<unsigned>(g.nx2)*ng2;
    const unsigned t = r0/row, r = r0 - t*row;
    const unsigned jj = r/ng2, ii = r - jj*ng2;
    k = g.ks + static_cast<int>(t);
    j = g.js + static_cast<int>(jj);
    i = (ii < static_cast<unsigned>(g.ng)) ? static_cast<int>(ii) : g.ie + 1 + static_cast<int>(ii) - g.ng;
  }
}

template <typename T>
__global__ void k_copy_ghosts(Geo g, const DevBlock *B, T *a, int nvar) {
  const unsigned n = GhostCount(g);
  const int m = blockIdx.y;
  BLOCK_STRIDE(c, n) {
    int i, j, k;
    GhostCell(g, c, i, j, k);
    const int ox1 = (i < g.is) ? -1 : ((i > g.ie) ? 1 : 0);
    const int ox2 = (j < g.js) ? -1 : ((j > g.je) ? 1 : 0);
    const int ox3 = (k < g.ks) ? -1 : ((k > g.ke) ? 1 : 0);
    const int nb = B[m].nghbr[Dir(ox1, ox2, ox3)];
    if (nb < 0) continue;
    for (int v=0; v<nvar; ++v)
      a[Idx(g,m,nvar,v,k,j,i)] = a[Idx(g,nb,nvar,v,k-ox3*g.nx3,j-ox2*g.nx2,i-ox1*g.nx1)];
  }
}

// The harness's gas ghost values on a statically refined mesh (a stand-in for the hydro
// boundary exchange of src/bvals/bvals_cc.cpp, which is not part of the drag path): every ghost
// cell takes the cell of the leaf holding its centre when that leaf is of the block's level (a
// copy) or coarser (the coarse value, injected), or the mean of the 8 cells of a finer leaf
// that make it up (the restriction of src/bvals/prolongation.cpp:104-111, same order).  Reads
// active cells only, writes ghost cells only.
template <typename T>
__global__ void k_smr_ghosts(Geo g, const DevBlock *B, dk::SmrTree tree, T *a, int nvar,
                             dk::ErrWord *err) {
  const unsigned n = GhostCount(g);
  const int m = blockIdx.y;
  const DevBlock &b = B[m];
  BLOCK_STRIDE(s, n) {
    int i, j, k;
    GhostCell(g, s, i, j, k);
    double c[3] = {b.x1min + (i - g.is + 0.5)*b.dx1, b.x2min + (j - g.js + 0.5)*b.dx2,
                   b.x3min + (k - g.ks + 0.5)*b.dx3};
    int sh[3];
    dk::SmrWrap(tree, c, sh);
    const int nb = dk::SmrLeafAt(B, tree, c);
    if (nb < 0) {dk::RecordError(err, DK_EROUTE, -1); continue;}
    const DevBlock &q = B[nb];
    if (q.level <= b.level) {
      const int in = static_cast<int>(floor((c[0] - q.x1min)/q.dx1)) + g.is;
      const int jn = static_cast<int>(floor((c[1] - q.x2min)/q.dx2)) + g.js;
      const int kn = static_cast<int>(floor((c[2] - q.x3min)/q.dx3)) + g.ks;
      for (int v=0; v<nvar; ++v) a[Idx(g,m,nvar,v,k,j,i)] = a[Idx(g,nb,nvar,v,kn,jn,in)];
    } else {
      const int fi = static_cast<int>(floor((c[0] - 0.5*b.dx1 - q.x1min)/q.dx1 + 0.5)) + g.is;
      const int fj = static_cast<int>(floor((c[1] - 0.5*b.dx2 - q.x2min)/q.dx2 + 0.5)) + g.js;
      const int fk = static_cast<int>(floor((c[2] - 0.5*b.dx3 - q.x3min)/q.dx3 + 0.5)) + g.ks;
      for (int v=0; v<nvar; ++v)
        a[Idx(g,m,nvar,v,k,j,i)] = 0.125*(
            a[Idx(g,nb,nvar,v,fk  ,fj  ,fi)] + a[Idx(g,nb,nvar,v,fk  ,fj  ,fi+1)]
          + a[Idx(g,nb,nvar,v,fk  ,fj+1,fi)] + a[Idx(g,nb,nvar,v,fk  ,fj+1,fi+1)]
          + a[Idx(g,nb,nvar,v,fk+1,fj  ,fi)] + a[Idx(g,nb,nvar,v,fk+1,fj  ,fi+1)]
          + a[Idx(g,nb,nvar,v,fk+1,fj+1,fi)] + a[Idx(g,nb,nvar,v,fk+1,fj+1,fi+1)]);
    }
  }
}

// velocity: the field's normal component (the last of 3, or the gas's momentum IM3 of 4: the
// hydro's reflect, src/bvals/physics/hydro_bcs.cpp:229-290) is negated at a reflecting face
template <typename T>
__global__ void k_x3bcs(Geo g, const DevBlock *B, T *a, int nvar, int velocity) {
  const long long n = (long long)g.nmb*g.ng*g.n2*g.n1;
  const bool refl = (g.bc_x3 == DK_X3_REFLECT);
  const int vn = (nvar == DK_NGAS) ? DK_IM3 : 2;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = static_cast<int>(rr % g.n1); rr /= g.n1;
    const int j = static_cast<int>(rr % g.n2); rr /= g.n2;
    const int gg = static_cast<int>(rr % g.ng);
    const int m = static_cast<int>(rr/g.ng);
    if (B[m].nghbr[Dir(0,0,-1)] < 0) {
      const int kg = g.ks - 1 - gg, kc = refl ? (g.ks + gg) : g.ks;
      for (int v=0; v<nvar; ++v) {
        const T val = a[Idx(g,m,nvar,v,kc,j,i)];
        a[Idx(g,m,nvar,v,kg,j,i)] = (refl && velocity && v == vn) ? -val : val;
      }
    }
    if (B[m].nghbr[Dir(0,0,1)] < 0) {
      const int kg = g.ke + 1 + gg, kc = refl ? (g.ke - gg) : g.ke;
      for (int v=0; v<nvar; ++v) {
        const T val = a[Idx(g,m,nvar,v,kc,j,i)];
        a[Idx(g,m,nvar,v,kg,j,i)] = (refl && velocity && v == vn) ? -val : val;
      }
    }
  }
}

// PLM_RemapFlx of src/shearing_box/remap_fluxes.hpp on a periodic column u(G)
template <class Col>
__device__ inline double PLMRemapFlx(const Col &u, int j, double eps) {
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

// the x1 ghosts of the blocks on the mesh x1 faces: the opposite face remapped by the fractional
// offset eps = fmod(yshear, dy)/dy (negated on the outer face) and moved by joffset cells
template <typename T>
__global__ void k_shear_ghosts(Geo g, const DevBlock *B, T *a, int nvar, double yshear) {
  const int nface = g.nmbx2*g.nmbx3;
  const long long n = 2LL*nface*nvar*g.n3*g.ng*g.n2;
  const int ny = g.nx2*g.nmbx2, nzt = g.nx3*g.nmbx3;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int j = static_cast<int>(rr % g.n2); rr /= g.n2;
    const int gg = static_cast<int>(rr % g.ng); rr /= g.ng;
    const int k = static_cast<int>(rr % g.n3); rr /= g.n3;
    const int v = static_cast<int>(rr % nvar); rr /= nvar;
    const int f = static_cast<int>(rr % nface);
    const int side = static_cast<int>(rr/nface);
    const int by0 = f % g.nmbx2, bz0 = f/g.nmbx2;
    const int bx = (side == 0) ? 0 : g.nmbx1 - 1;
    const int m = bx + g.nmbx1*(by0 + g.nmbx2*bz0);
    const int bxs = (side == 0) ? g.nmbx1 - 1 : 0;
    const double dy = B[m].dx2;
    const int joffset = static_cast<int>(yshear/dy);
    double eps = fmod(yshear, dy)/dy;
    if (side == 1) eps *= -1.0;
    int gk = B[m].lx3*g.nx3 + (k - g.ks);
    if (g.bc_x3 == DK_X3_PERIODIC) gk = ((gk % nzt) + nzt) % nzt;
    else gk = min(max(gk, 0), nzt - 1);
    const int bz = gk/g.nx3, kl = g.ks + (gk - bz*g.nx3);
    const int ig = (side == 0) ? (g.is - g.ng + gg) : (g.ie + 1 + gg);
    const int isrc = (side == 0) ? (ig + g.nx1) : (ig - g.nx1);
    auto col = [&](int G) -> double {
      G = ((G % ny) + ny) % ny;
      const int by = G/g.nx2;
      const int ms = bxs + g.nmbx1*(by + g.nmbx2*bz);
      return static_cast<double>(a[Idx(g,ms,nvar,v,kl,g.js + (G - by*g.nx2),isrc)]);
    };
    const int Gt = B[m].lx2*g.nx2 + (j - g.js);
    const int G = (side == 0) ? (Gt - joffset) : (Gt + joffset);
    a[Idx(g,m,nvar,v,k,j,ig)] =
        static_cast<T>(col(G) - (PLMRemapFlx(col, G+1, eps) - PLMRemapFlx(col, G, eps)));
  }
}

// physical x3 faces: the ghost layers are mirrored into the first ng active layers over the
// whole padded plane (each thread owns one (m, v, g, j, i) pair of cells)
template <typename T>
__global__ void k_fold_x3(Geo g, const DevBlock *B, T *a, int nvar) {
  const long long n = (long long)g.nmb*nvar*g.ng*g.n2*g.n1;
  GRID_STRIDE(t, n) {
    long long rr = t;
    const int i = static_cast<int>(rr % g.n1); rr /= g.n1;
    const int j = static_cast<int>(rr % g.n2); rr /= g.n2;
    const int gg = static_cast<int>(rr % g.ng); rr /= g.ng;
    const int v = static_cast<int>(rr % nvar);
    const int m = static_cast<int>(rr/nvar);
    if (B[m].nghbr[Dir(0,0,-1)] < 0)
      a[Idx(g,m,nvar,v,g.ks+gg,j,i)] += a[Idx(g,m,nvar,v,g.ks-1-gg,j,i)];
    if (B[m].nghbr[Dir(0,0,1)] < 0)
      a[Idx(g,m,nvar,v,g.ke-gg,j,i)] += a[Idx(g,m,nvar,v,g.ke+1+gg,j,i)];
  }
}

// one of e = min(2 ng, n) edge layers of an axis of n cells: the first ng, then the last ones
__device__ inline int EdgeLayer(int t, int n, int ng, int e) {
  return (t < ng) ? t : (n - e + t);
}

// the s-th cell (0 <= s < ShellCount) of the active cells within ng of a block face: the x3
// edge layers, then the x2 edge rows of the other layers, then the x1 edge cells of the rest
__host__ __device__ inline long long ShellCount(const Geo &g) {
  const int e1 = min(2*g.ng, g.nx1), e2 = min(2*g.ng, g.nx2), e3 = min(2*g.ng, g.nx3);
  return (long long)e3*g.nx2*g.nx1 + (long long)(g.nx3 - e3)*e2*g.nx1 +
         (long long)(g.nx3 - e3)*(g.nx2 - e2)*e1;
}

__device__ inline void ShellCell(const Geo &g, unsigned s, int &i, int &j, int &k) {
  const int e1 = min(2*g.ng, g.nx1), e2 = min(2*g.ng, g.nx2), e3 = min(2*g.ng, g.nx3);
  const unsigned plane = static_cast<unsigned>(g.nx2)*g.nx1;
  const unsigned a = static_cast<unsigned>(e3)*plane;
  const unsigned b = static_cast<unsigned>(g.nx3 - e3)*e2*g.nx1;
  if (s < a) {
    const unsigned t = s/plane, r = s - t*plane;
    const unsigned jj = r/g.nx1;
    k = EdgeLayer(static_cast<int>(t), g.nx3, g.ng, e3);
    j = static_cast<int>(jj);
    i = static_cast<int>(r - jj*g.nx1);
  } else if (s < a + b) {
    const unsigned r0 = s - a, rows = static_cast<unsigned>(e2)*g.nx1;
    const unsigned t = r0/rows, r = r0 - t*rows;
    const unsigned jj = r/g.nx1;
    k = g.ng + static_cast<int>(t);
    j = EdgeLayer(static_cast<int>(jj), g.nx2, g.ng, e2);
    i = static_cast<int>(r - jj*g.nx1);
  } else {
    const unsigned r0 = s - a - b, row = static_cast<unsigned>(g.nx2 - e2)*e1;
    const unsigned t = r0/row, r = r0 - t*row;
    const unsigned jj = r/e1;
    k = g.ng + static_cast<int>(t);
    j = g.ng + static_cast<int>(jj);
    i = EdgeLayer(static_cast<int>(r - jj*e1), g.nx1, g.ng, e1);
  }
  i += g.is; j += g.js; k += g.ks;
}

// same-level neighbours: every active cell within ng of a block face sums the neighbours'
// ghost strips in NeighborIndex order (src/mesh/nghbr_index.hpp:19-54), one thread per such
// cell.  A strip of direction (ox1, ox2, ox3) covers the cell when each non-zero ox puts the cell
// in the first (ox < 0) or last (ox > 0) ng layers of that axis.  The strips that cross a
// shear-periodic x1 face are left to k_fold_shear.
constexpr int kFoldVars = 5;

template <typename T>
__global__ void k_fold(Geo g, const DevBlock *B, const int *slot_dir, T *a, int nvar) {
  __shared__ int sdir[26];
  if (threadIdx.x < 26) sdir[threadIdx.x] = slot_dir[threadIdx.x];
  __syncthreads();
  const unsigned nsh = static_cast<unsigned>(ShellCount(g));
  const bool shear = (g.bc_x1 == DK_X1_SHEAR);
  const int m = blockIdx.y;
  BLOCK_STRIDE(t, nsh) {
    int i, j, k;
    ShellCell(g, t, i, j, k);
    const DevBlock &bm = B[m];
    // the allowed offsets per axis: bit (o + 1) set when offset o covers this cell
    const int a1 = 2 | ((i < g.is + g.ng) ? 1 : 0) | ((i > g.ie - g.ng) ? 4 : 0);
    const int a2 = 2 | ((j < g.js + g.ng) ? 1 : 0) | ((j > g.je - g.ng) ? 4 : 0);
    const int a3 = 2 | ((k < g.ks + g.ng) ? 1 : 0) | ((k > g.ke - g.ng) ? 4 : 0);
    // every variable's sum in slot order (at most kFoldVars variables: the deposits fold 3-5)
    T sum[kFoldVars];
#pragma unroll
    for (int v=0; v<kFoldVars; ++v) sum[v] = (v < nvar) ? a[Idx(g, m, nvar, v, k, j, i)] : T(0);
    for (int s=0; s<26; ++s) {
      const int d = sdir[s];
      const int o1 = d % 3, o2 = (d/3) % 3, o3 = d/9;
      if (!((a1 >> o1) & 1) || !((a2 >> o2) & 1) || !((a3 >> o3) & 1)) continue;
      const int nb = bm.nghbr[d];
      // (on a refined mesh only same-level neighbours fold: no deposit of the FP32 tile
      // kernels reaches a region of another level, whose clouds go to the leaf traversal)
      if (nb < 0 || B[nb].level != bm.level) continue;
      if (shear && o1 != 1) {
        const bool crosses = (o1 == 0) ? (bm.lx1 == 0) : (bm.lx1 == g.nmbx1 - 1);
        if (crosses) continue;
      }
      const long long src = Idx(g, nb, nvar, 0, k - (o3 - 1)*g.nx3, j - (o2 - 1)*g.nx2,
                                i - (o1 - 1)*g.nx1);
#pragma unroll
      for (int v=0; v<kFoldVars; ++v)
        if (v < nvar) sum[v] += a[src + v*g.ncb];
    }
#pragma unroll
    for (int v=0; v<kFoldVars; ++v)
      if (v < nvar) a[Idx(g, m, nvar, v, k, j, i)] = sum[v];
  }
}

// The cells of a padded block field that hold global row r (0 <= r < nx*nb) along one axis, in
// fixed slots: [0] the active row of block r/nx, [1] the upper ghost copy in block r/nx - 1,
// [2] the lower ghost copy in block r/nx + 1 (periodic wrap, or none beyond a physical face).
// blk/loc receive block row and padded index; the bit s of the result is set when slot s holds a
// copy (fixed slots, so that callers' loops keep the arrays in registers).
__device__ inline int RowCopies(int r, int nx, int nb, int ng, int start, bool periodic,
                                int blk[3], int loc[3]) {
  int mask = 1;
  const int b = r/nx, l = r - b*nx;
  blk[0] = b; loc[0] = start + l;
  blk[1] = 0; loc[1] = 0; blk[2] = 0; loc[2] = 0;
  if (l < ng) {                 // the upper ghost rows of block b - 1
    int b2 = b - 1;
    if (b2 < 0) b2 = periodic ? b2 + nb : -1;
    if (b2 >= 0) {blk[1] = b2; loc[1] = start + nx + l; mask |= 2;}
  }
  if (l >= nx - ng) {           // the lower ghost rows of block b + 1
    int b2 = b + 1;
    if (b2 >= nb) b2 = periodic ? b2 - nb : -1;
    if (b2 >= 0) {blk[2] = b2; loc[2] = start - (nx - l); mask |= 4;}
  }
  return mask;
}

// The shear part as a gather: a ghost cell beyond the outer x1 face deposits into the inner
// face at x2 shifted by +ysh, beyond the inner face into the outer face at -ysh, split between
// the two cells its shifted interval overlaps (the piecewise-constant remap of
// src/bvals/bvals_dep_general.cpp:100-160).  Every target cell (thread) enumerates the source
// ghost cells of the opposite face (x2 and x3 ghost copies included) whose shifted interval can
// touch it and adds their shares in a fixed order; each share is computed from the source's
// own row index exactly as a scatter would.
template <typename T>
__global__ void k_fold_shear(Geo g, const DevBlock *B, T *a, int nvar, double shc) {
  const int nface = g.nmbx2*g.nmbx3;
  const int ny = g.nx2*g.nmbx2;
  const unsigned n = 2u*nface*g.nx3*g.nx2*g.ng;
  const bool perz = (g.bc_x3 == DK_X3_PERIODIC);
  BLOCK_STRIDE(t, n) {
    unsigned rr = t;
    unsigned q = rr/g.ng;
    const int gg = static_cast<int>(rr - q*g.ng); rr = q;
    q = rr/g.nx2;
    const int jt = g.js + static_cast<int>(rr - q*g.nx2); rr = q;
    q = rr/g.nx3;
    const int kt = g.ks + static_cast<int>(rr - q*g.nx3); rr = q;
    q = rr/nface;
    const int f = static_cast<int>(rr - q*nface);
    const int tside = static_cast<int>(q);   // 0: target on the inner face
    const int byt = f % g.nmbx2, bzt = f/g.nmbx2;
    const int mt = ((tside == 0) ? 0 : g.nmbx1 - 1) + g.nmbx1*(byt + g.nmbx2*bzt);
    const int it = (tside == 0) ? (g.is + gg) : (g.ie - gg);
    // the sources sit in the ghost column beyond the opposite face at the same depth gg
    const int sside = 1 - tside;
    const int bxs = (sside == 0) ? 0 : g.nmbx1 - 1;
    const int isrc = (sside == 0) ? (g.is - 1 - gg) : (g.ie + 1 + gg);
    const double sh = (sside == 0) ? -shc : shc;
    const int GJt = byt*g.nx2 + (jt - g.js);
    const int GKt = bzt*g.nx3 + (kt - g.ks);
    int kb[3], kl[3];
    const int kmask = RowCopies(GKt, g.nx3, g.nmbx3, g.ng, g.ks, perz, kb, kl);
    // rows base + d: a source row r sends to floor(r + sh) and the row above, so with ny >= 5
    // only d = 1 and d = 2 can reach the target (the other rows add nothing)
    const int d0 = (ny < 5) ? 0 : 1, d1 = (ny < 5) ? ny : 3;
    const int base = GJt - static_cast<int>(floor(sh)) - 2;
    for (int v=0; v<nvar; ++v) {
      T sum = a[Idx(g,mt,nvar,v,kt,jt,it)];
      // (fixed trip counts with predicates: the copy arrays stay in registers)
#pragma unroll
      for (int q=0; q<3; ++q) {
        if (!((kmask >> q) & 1)) continue;
#pragma unroll
        for (int d=0; d<5; ++d) {
          if (d < d0 || d >= d1) continue;
          const int r = (((base + d) % ny) + ny) % ny;
          int jb[3], jl[3];
          const int jmask = RowCopies(r, g.nx2, g.nmbx2, g.ng, g.js, true, jb, jl);
#pragma unroll
          for (int p=0; p<3; ++p) {
            if (!((jmask >> p) & 1)) continue;
            const int ms = bxs + g.nmbx1*(jb[p] + g.nmbx2*kb[q]);
            const int gj = jb[p]*g.nx2 + (jl[p] - g.js);    // the source's unwrapped row
            const double lo = gj + sh;
            const int c0 = static_cast<int>(floor(lo));
            const double fr = lo - c0;
            const T val = a[Idx(g,ms,nvar,v,kl[q],jl[p],isrc)];
            if ((((c0 % ny) + ny) % ny) == GJt && (1.0 - fr) > 0.0)
              sum += static_cast<T>(1.0 - fr)*val;
            if (((((c0 + 1) % ny) + ny) % ny) == GJt && fr > 0.0)
              sum += static_cast<T>(fr)*val;
          }
        }
      }
      a[Idx(g,mt,nvar,v,kt,jt,it)] = sum;
    }
  }
}

}  // namespace

namespace dk {

template <typename T>
int FillGhostsT(dk_ctx *c, T *a, int nvar, bool velocity, double t) {
  const Geo &g = c->g;
  if (c->smr) {   // (refined meshes are periodic in every direction)
    const int tk = ProfBegin(c, "smr_ghosts");
    k_smr_ghosts<T><<<BlockGrid(GhostCount(g), g.nmb), kBlock, 0, c->stream>>>(
        g, c->d_blocks, c->tree, a, nvar, c->d_err);
    ProfEnd(c, tk);
    return Check(hipGetLastError());
  }
  {const int tk = ProfBegin(c, "copy_ghosts");
   k_copy_ghosts<T><<<BlockGrid(GhostCount(g), g.nmb), kBlock, 0, c->stream>>>(
       g, c->d_blocks, a, nvar);
   ProfEnd(c, tk);}
  if (c->M.bc_x1 == DK_X1_SHEAR) {
    const double yshear = (c->P.qshear*c->P.omega0)*(c->M.x1max - c->M.x1min)*t;
    const long long n = 2LL*g.nmbx2*g.nmbx3*nvar*g.n3*g.ng*g.n2;
    const int tk = ProfBegin(c, "shear_ghosts");
    k_shear_ghosts<T><<<Grid(n), kBlock, 0, c->stream>>>(g, c->d_blocks, a, nvar, yshear);
    ProfEnd(c, tk);
  }
  if (c->M.bc_x3 != DK_X3_PERIODIC) {
    const long long n = (long long)g.nmb*g.ng*g.n2*g.n1;
    k_x3bcs<T><<<Grid(n), kBlock, 0, c->stream>>>(g, c->d_blocks, a, nvar, velocity ? 1 : 0);
  }
  return Check(hipGetLastError());
}

template <typename T>
int FoldT(dk_ctx *c, T *a, int nvar, double t) {
  const Geo &g = c->g;
  if (nvar < 1 || nvar > kFoldVars) return DK_EINVAL;
  if (c->M.bc_x3 != DK_X3_PERIODIC && !c->M.x3_dep_drop) {
    const long long n = (long long)g.nmb*nvar*g.ng*g.n2*g.n1;
    k_fold_x3<T><<<Grid(n), kBlock, 0, c->stream>>>(g, c->d_blocks, a, nvar);
  }
  {const int tk = ProfBegin(c, "fold");
   k_fold<T><<<BlockGrid(ShellCount(g), g.nmb), kBlock, 0, c->stream>>>(
       g, c->d_blocks, c->d_slot_dir, a, nvar);
   ProfEnd(c, tk);}
  if (c->M.bc_x1 == DK_X1_SHEAR) {
    const double lx = c->M.x1max - c->M.x1min, ly = c->M.x2max - c->M.x2min;
    const double ysh = std::fmod(c->P.qshear*c->P.omega0*lx*t, ly);
    const double dy = ly/(g.nx2*g.nmbx2);
    const long long n = 2LL*g.nmbx2*g.nmbx3*g.nx3*g.nx2*g.ng;
    const int tk = ProfBegin(c, "fold_shear");
    k_fold_shear<T><<<BlockGrid(n, 1), kBlock, 0, c->stream>>>(g, c->d_blocks, a, nvar, ysh/dy);
    ProfEnd(c, tk);
  }
  return Check(hipGetLastError());
}

template int FillGhostsT<double>(dk_ctx *, double *, int, bool, double);
template int FillGhostsT<float>(dk_ctx *, float *, int, bool, double);
template int FoldT<double>(dk_ctx *, double *, int, double);
template int FoldT<float>(dk_ctx *, float *, int, double);

}  // namespace dk
