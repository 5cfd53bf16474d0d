// dk_internal.hpp -- shared definitions of the gfx950 drag-coupling library.
#ifndef ATHENAK_AMD_CSRC_DK_INTERNAL_HPP_
#define ATHENAK_AMD_CSRC_DK_INTERNAL_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/dustk.h"
#include "../../include/dustk_setup.h"

namespace dk {

constexpr int kBlock = 256;   // threads per workgroup of the streaming kernels

// Mesh geometry handed to kernels by value.
struct Geo {
  int nmb, nx1, nx2, nx3, ng, n1, n2, n3;
  int is, ie, js, je, ks, ke;
  long long ncb;           // cells per block and variable, ghosts included
  int nmbx1, nmbx2, nmbx3;
  int bc_x1, bc_x3, x3_dep_drop;
  double x1min, x1max, x2min, x2max, x3min, x3max;
  int deposit;
};

// Block geometry on the device: the box and cell sizes (MeshBlock::mb_size) and neighbours.
struct DevBlock {
  double x1min, x1max, x2min, x2max, x3min, x3max;
  double dx1, dx2, dx3;
  double sx1, sx2, sx3;    // nx/(xmax - xmin) of each axis, as PMWeights forms it
  double ivol;             // 1/(dx1 dx2 dx3)
  int lx1, lx2, lx3;
  int level;               // refinement level (0 on a uniform mesh)
  int face;                // bits 1,2 / 4,8 / 16,32: lower/upper face within half a block of
                           // the mesh face in x1 / x2 / x3 (the test of bvals_part.cpp)
  int xmask;               // refined meshes: bit Dir(o) set when the same-level region beside
                           // the block in direction o is not one leaf of the block's level
  int nghbr[27];
};

// A statically refined mesh by value for kernels: F[d] finest-level block slots per direction,
// leaf[(s3*F[1] + s2)*F[0] + s1] the leaf covering slot (s1, s2, s3); the mesh's lower corner
// and lengths.  leaf is null on a uniform mesh.
struct SmrTree {
  int F[3];
  int maxlev;
  const int *leaf;
  double lo[3], len[3];
};

__host__ __device__ inline long long Idx(const Geo &g, int m, int nvar, int v, int k, int j,
                                         int i) {
  return ((((long long)m*nvar + v)*g.n3 + k)*g.n2 + j)*(long long)g.n1 + i;
}

__host__ __device__ inline int Dir(int ox1, int ox2, int ox3) {
  return (ox1 + 1) + 3*(ox2 + 1) + 9*(ox3 + 1);
}

// src/dust/dust.hpp:1088-1111 (same operations in the same order)
template <typename T>
__device__ inline void PMWeights(const double x, const double xmin, const double xmax,
                                 const int nx, const int is, const int scheme, int &ip, T w[3]) {
  double cell = (x - xmin)*(static_cast<double>(nx)/(xmax - xmin));
  int ig = static_cast<int>(cell + 1.0) - 1;
  double del = cell - (static_cast<double>(ig) + 0.5);
  ip = ig + is;
  if (scheme == 0) {
    w[0] = 0.0; w[1] = 1.0; w[2] = 0.0;
  } else if (scheme == 1) {
    w[0] = fmax(0.0, -del);
    w[1] = 1.0 - fabs(del);
    w[2] = fmax(0.0, del);
  } else {
    w[0] = 0.5*(0.5 - del)*(0.5 - del);
    w[1] = 0.75 - del*del;
    w[2] = 0.5*(0.5 + del)*(0.5 + del);
  }
}

// StiffnessActive (src/dust/dust.hpp:1150-1152): the pc2 predictor deposits the rate field
// qdep[4] and UpdateStiffness records chi whenever the stiffness record is asked for or the
// feedback damping of stiff cells needs it (phst_track, the particle history, is not built)
inline bool StiffnessActive(const dk_physics &P) { return P.stiffness != 0 || P.pc2_damping != 0; }

// First error wins: code and particle index.
struct ErrWord {
  int code;
  int pad;
  long long particle;
};

__device__ inline void RecordError(ErrWord *e, int code, long long p) {
  if (atomicCAS(&e->code, 0, code) == 0) e->particle = p;
}

}  // namespace dk

struct dk_ctx {
  dk_mesh M;
  dk_physics P;
  std::vector<dk_block> B;
  std::vector<double> dx1, dx2, dx3;   // per block (MeshBlock::mb_size dx1..dx3)
  int form = DK_FORM_FP64;
  hipStream_t stream = nullptr;
  int device = 0;
  dk::Geo g;
  dk::DevBlock *d_blocks = nullptr;
  int *d_slot_dir = nullptr;       // the 26 same-level fold slots in NeighborIndex order
  // caller-owned
  double *rdata = nullptr;
  int32_t *idata = nullptr;
  int64_t cap = 0, np = 0;
  double *u0 = nullptr, *u1 = nullptr;
  // library-owned work fields (layout of the gas, nvar components)
  double *qdep = nullptr, *ustar = nullptr, *dmom = nullptr, *ap = nullptr, *resid = nullptr;
  // the explicit gas update's flux divergence (4 components), allocated by the first stage
  // with isothermal fluxes (dk_gas_src.hydro)
  double *hdivf = nullptr;
  // refinable clouds: the packed gather field (m, k, j, i, 12) and a velocity field (3)
  double *gpack = nullptr, *cvel = nullptr;
  // the strict solver (pcg, adaptive): residual and search direction (3 components), the
  // parameters (the reference's <dust> defaults, dust.cpp:467-474) and the last solve's
  // iterations, fast accept, true residual, error bound, target, state scale, max Q/rho
  double *solver_r = nullptr, *solver_p = nullptr;
  double drag_rtol = 1.0e-11, drag_atol = 1.0e-14;
  int drag_iter_max = 200;
  double adaptive_order_c = 0.25, adaptive_rtol_max = 1.0e-3, adaptive_tref = 1.0,
         adaptive_state_floor = 1.0;
  double solver_last[7] = {0, 0, -1, -1, -1, -1, -1};
  dk::ErrWord *d_err = nullptr;
  double *d_partial = nullptr;     // reduction partials
  double *d_scalar = nullptr;
  // sort scratch (grown on demand)
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  unsigned long long *sort_keys = nullptr;
  int *sort_idx = nullptr;
  double *sort_rbuf = nullptr;
  int32_t *sort_ibuf = nullptr;
  int64_t sort_cap = 0;
  int noalloc_err = DK_OK;         // set up of the allocation-free sort's scratch (dk_create)
  // a statically refined mesh (levels above 0): the finest block grid on the device
  int smr = 0;
  dk::SmrTree tree{};
  int *d_leafmap = nullptr;
  long long *d_block_off = nullptr;   // uneven set-ups: particle offsets of the blocks
  // the FP32 form (dk_fp32.hip)
  struct Fast *fast = nullptr;
  int graphs = 0;
  long long graph_launches = 0;    // stage graphs launched (dk_graph_launches)
  // RemoveOutflowParticles run totals (x3_outflow_mass, x3_outflow_count, dust.hpp); the ledger
  // since the last removal is d_scalar[12..13] (bvals_part.cpp:184-186)
  double x3_out_mass = 0.0, x3_out_count = 0.0;
  hipGraphExec_t gexec[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // per-kernel event timing
  struct Prof *prof = nullptr;
};

namespace dk {

int Check(hipError_t e);
long long NCells(const dk_ctx *c);   // nmb * ncb
int Zero(dk_ctx *c, double *a, long long n);

// dk_mesh.hip: ghost fill (copy, shear remap, x3 values) and additive ghost fold
template <typename T> int FillGhostsT(dk_ctx *c, T *a, int nvar, bool velocity, double t);
template <typename T> int FoldT(dk_ctx *c, T *a, int nvar, double t);

// per-kernel event timing (no-ops unless enabled): Begin returns a token for End
int ProfBegin(dk_ctx *c, const char *label);
bool ProfOn(const dk_ctx *c);
void ProfEnd(dk_ctx *c, int token);
void ProfDestroy(dk_ctx *c);

// dk_fp32.hip
int FastCreate(dk_ctx *c);
// the FP32 form's online autotuner around an active stage (no-ops when not tuning)
int FastTuneBegin(dk_ctx *c, const dk_stage *st);
int FastTuneEnd(dk_ctx *c, const dk_stage *st);
// refinable clouds at the mesh level (dk_clouds.hip)
bool CloudsSupported(const dk_mesh *mesh, const dk_physics *phys);
int CloudPredictor(dk_ctx *c, double dt, int rate_dep);
int CloudAdvance(dk_ctx *c, double dt);
void FastDestroy(dk_ctx *c);
int FastBind(dk_ctx *c);
int FastExport(dk_ctx *c);
int FastStage(dk_ctx *c, const dk_stage *st, const dk_gas_src *src);
int FastDiag(dk_ctx *c, double out[9]);
int FastSortState(dk_ctx *c, int which, int32_t *host, long long n);
int FastNewDt(dk_ctx *c, double qo, double *partial, int nb);
// RefreshStoppingTimeMaximum over the FP32 store: nb partial (max, min) pairs
int FastStoppingTimes(dk_ctx *c, double *partial, int nb);
// UpdateStiffness (dust_pm.cpp:288-342): chi = dt max_cells q[comp]/rho and its running maximum,
// and the rate maximum, kept on the device (d_scalar[9..11]); q: the pc2 rate component of a
// deposit field
template <typename T> int StiffnessRecord(dk_ctx *c, const T *q, int nv, int comp, double dt);

// dk_fp64.hip pieces reused by the FP32 form
// the gas update of a stage; dmom: the drag history u0.M += atw dmom applied after it in the same
// pass (the FP32 form's stage 2 with isothermal fluxes)
int GasHook(dk_ctx *c, const dk_stage *st, const dk_gas_src *src, const double *dmom = nullptr,
            double atw = 0.0);
// the explicit gas update's options check and its flux divergence field (allocated on first use)
int HydroSupported(const dk_ctx *c, const dk_gas_src *src);
int HydroField(dk_ctx *c);

}  // namespace dk

#endif  // ATHENAK_AMD_CSRC_DK_INTERNAL_HPP_
