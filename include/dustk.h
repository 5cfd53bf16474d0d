/*
 * dustk.h -- C-ABI of the MI355X (gfx950) particle-gas drag coupling.
 *
 * Every entry point replaces one task method of the reference's DustGasDrag task list
 * (src/dust/dust_tasks.cpp:38-120).  A task there has the form
 *     TaskStatus DustGasDrag::X(Driver *pdrive, int stage)
 * (src/tasklist/task_list.hpp:180-199); here it is
 *     int dk_x(dk_ctx *ctx, const dk_stage *st)
 * where dk_stage carries the numbers the task read from Driver and Mesh (stage, dt, time,
 * the RK/IMEX tableau of src/driver/driver.cpp:106-118,278-300).  Every function returns 0 on
 * success or a positive DK_E* code; none aborts.  Calls are asynchronous and ordered on the
 * context's HIP stream; only dk_diag_sums, dk_error and dk_new_time_step wait for the stream.
 *
 * Layouts are the reference's:
 *   particles  rdata[c*capacity + p], c < DK_NRDATA (double), idata[c*capacity + p],
 *              c < DK_NIDATA (int32): Particles::prtcl_rdata/idata, LayoutRight
 *              (src/particles/particles.cpp:190-205, src/athena.hpp:67-80)
 *   gas        u[(((m*DK_NGAS + v)*n3 + k)*n2 + j)*n1 + i], n_d = nx_d + 2 ng, ghosts
 *              included: Hydro::u0/u1 (src/hydro/hydro.cpp:64)
 * All pointers passed to the library are device pointers unless stated otherwise.  The
 * library owns its work fields (qdep, ustar, dmom, solver scratch, sort keys, FP32 shadows).
 */
#ifndef DUSTK_H_
#define DUSTK_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* particle columns: src/athena.hpp:77-80 */
enum {
  DK_IPX = 0, DK_IPVX = 1, DK_IPY = 2, DK_IPVY = 3, DK_IPZ = 4, DK_IPVZ = 5,
  DK_IPX1 = 6, DK_IPVX1 = 7, DK_IPY1 = 8, DK_IPVY1 = 9, DK_IPZ1 = 10, DK_IPVZ1 = 11,
  DK_IPRX = 12, DK_IPRY = 13, DK_IPRZ = 14, DK_IPTS = 15, DK_IPM = 16
};
enum { DK_PGID = 0, DK_PTAG = 1, DK_PSP = 2, DK_PLEV = 3, DK_PTAGH = 4 };
#define DK_NRDATA 17
#define DK_NIDATA 5

/* isothermal gas conserved variables: src/athena.hpp:59 */
enum { DK_IDN = 0, DK_IM1 = 1, DK_IM2 = 2, DK_IM3 = 3 };
#define DK_NGAS 4

/* src/dust/dust.hpp:59 */
enum { DK_NGP = 0, DK_CIC = 1, DK_TSC = 2 };
/* src/dust/dust.hpp:74 */
enum { DK_IMEX = 0, DK_PC2 = 1 };
/* src/dust/dust.hpp:70 (pcg and adaptive: DK_FORM_FP64 only) */
enum { DK_LOCAL = 0, DK_APPLYA = 1, DK_DC1 = 2, DK_DC2 = 3, DK_PCG = 4, DK_ADAPTIVE = 5 };
/* mesh boundaries handled by the drag path */
enum { DK_X1_PERIODIC = 0, DK_X1_SHEAR = 1 };
enum { DK_X3_PERIODIC = 0, DK_X3_REFLECT = 1, DK_X3_OUTFLOW = 2 };

/* execution forms of a context */
enum {
  DK_FORM_FP64 = 0,   /* plain FP64: one kernel per reference kernel, atomic deposits,
                         operating directly on the caller's arrays */
  DK_FORM_FP32 = 1    /* FP32 bulk, FP64 positions and sums: cell-sorted particles,
                         atomic-free deposits, fused stage passes (dk_stage_run only).
                         imex with local/dc1/dc2 on any mesh of the FP64 form; pc2 with
                         point particles or clouds at the mesh level on periodic meshes
                         without the shearing box, clouds on statically refined periodic
                         meshes (atomic deposits for the clouds near a level change);
                         otherwise dk_create returns DK_EUNSUPPORTED */
};

/* error codes */
enum {
  DK_OK = 0,
  DK_EINVAL = 1,        /* bad argument */
  DK_EHIP = 2,          /* a HIP runtime call failed */
  DK_ENOMEM = 3,        /* device allocation failed */
  DK_EHALO = 4,         /* a particle lies outside the deposit halo of its block
                           (Kokkos::abort of src/dust/dust_pm.cpp:241-251) */
  DK_EROUTE = 5,        /* a particle moved more than one block
                           (Kokkos::abort of src/bvals/bvals_part.cpp:245-252) */
  DK_ESTOP = 6,         /* a non-positive stopping time (src/dust/dust_newdt.cpp:73-79) */
  DK_EUNSUPPORTED = 7,  /* option not provided by this build */
  DK_ECAPACITY = 8,     /* a size beyond a fixed maximum */
  DK_ESTATE = 9,        /* call order violated (e.g. particles not bound) */
  DK_ENOTFINITE = 10,   /* a non-finite value met by a guarded kernel */
  DK_ESOLVER = 11       /* the strict drag solver lost positive definiteness or did not reach
                           its tolerance (SolverFatal, src/dust/dust_solver.cpp:468-526) */
};

/* A MeshBlock: its box (RegionSize, src/mesh/mesh.hpp:25-29), logical location and level,
 * and the gid of the same-level neighbour in each direction
 * d = (ox1+1) + 3*(ox2+1) + 9*(ox3+1) (d = 13 is the block itself; -1 = none, a physical
 * face).  On a statically refined mesh (any level > 0) nghbr is not read: the library finds
 * neighbours of every level from the levels and logical locations.  Replaces
 * MeshBlock::mb_size/nghbr (src/mesh/meshblock.hpp). */
typedef struct dk_block {
  double x1min, x1max, x2min, x2max, x3min, x3max;
  int32_t lx1, lx2, lx3, level;
  int32_t nghbr[27];
} dk_block;

/* The mesh: block size, ghost depth, root box and the physical/shear boundaries
 * (Mesh::mb_indcs, mesh_size, mesh_bcs). */
typedef struct dk_mesh {
  int32_t nmb, nx1, nx2, nx3, ng;
  double x1min, x1max, x2min, x2max, x3min, x3max;
  int32_t nmbx1, nmbx2, nmbx3;   /* blocks per direction (uniform mesh) */
  int32_t bc_x1;                 /* DK_X1_PERIODIC or DK_X1_SHEAR */
  int32_t bc_x3;                 /* DK_X3_PERIODIC, DK_X3_REFLECT or DK_X3_OUTFLOW: the
                                    particle x3 boundary (bvals_part.cpp:159-197) and the u*
                                    ghost values (reflect: mirror, else zero gradient) */
  int32_t x3_dep_drop;           /* physical x3 faces: 0 fold the ghost deposits back (default
                                    <dust>/x3_ghost_deposit = fold), 1 drop them
                                    (src/dust/dust.cpp:1323-1333) */
} dk_mesh;

/* The drag options of the DustGasDrag constructor (src/dust/dust.cpp:60-450). */
typedef struct dk_physics {
  int32_t deposit;          /* DK_TSC (default), DK_CIC, DK_NGP */
  int32_t coupling;         /* DK_IMEX or DK_PC2 */
  int32_t solver;           /* DK_DC1 default for imex point particles, DK_LOCAL for pc2 */
  int32_t back_reaction;    /* 1: the gas feels the dust */
  int32_t consistent_gas;   /* imex point consistent-gas split (dust.cpp:359, default 1) */
  int32_t shearing_box;     /* 3-D shearing box forces and frames */
  int32_t stratified;       /* vertical gravity -Omega^2 z */
  double omega0, qshear;
  int32_t pc2_damping;      /* pc2 feedback damping of stiff cells (dust_pm.cpp:389-398) */
  int32_t stiffness;        /* record chi (UpdateStiffness); the pc2 rate field qdep[4] is
                               deposited and chi recorded when stiffness or pc2_damping is set
                               (StiffnessActive, src/dust/dust.hpp:1150-1152) */
  int32_t clouds;           /* pc2 with 3-D refinable clouds at the mesh level (plain codes:
                               CloudPredictorT / CloudAdvanceT of dust_cloud3d.cpp:1369-1536);
                               meshes periodic in x1, x2, x3 only, same-level or statically
                               refined (DK_EUNSUPPORTED otherwise); 0: point particles (refused
                               on refined meshes) */
} dk_physics;

/* What a task read from Driver and Mesh for one stage. */
typedef struct dk_stage {
  int32_t stage;            /* 1-based */
  int32_t nexp_stages;      /* 3 for imex2+, 2 for rk2 */
  double time, dt;          /* t^n and the step */
  double gam0[3], gam1[3], beta[3];
  double a_twid22;          /* imex2+ a_twid[2][2] */
  double a_impl;            /* imex2+ gamma */
} dk_stage;

/* The explicit gas update of a stage: Hydro::RKUpdate (src/hydro/hydro_update.cpp:23-83)
 * plus Hydro::HydroSrcTerms (src/hydro/hydro_tasks.cpp:268-297).  hydro = 0: zero fluxes (the
 * gas moves only by the source terms and the drag).  hydro = 1: the flux divergence of an
 * isothermal gas (Hydro::Fluxes, src/hydro/hydro_fluxes.cpp, with reconstruct = plm,
 * src/reconstruct/plm.hpp, and rsolver = hlle, src/hydro/rsolvers/hlle_hyd.hpp) from the
 * primitives of u0 at the stage start (IsothermalHydro::ConsToPrim with the density floor,
 * src/eos/isothermal_hyd.cpp:34-110); the stage first refills the u0 ghosts.  It needs at least
 * 2 ghost cells, a uniform mesh and periodic x1 (DK_EUNSUPPORTED otherwise; x3 may be
 * periodic, outflow or reflecting, the gas mirrored with IM3 negated) and iso_cs > 0 (DK_EINVAL); the first stage with hydro allocates the flux
 * divergence field (4 components, the gas layout). */
typedef struct dk_gas_src {
  int32_t shearing_box;     /* ShearingBoxCC::SourceTermsCC (3-D) */
  int32_t stratified;
  double omega0, qshear;
  double const_accel;       /* SourceTerms::ConstantAccel value (0 = off) */
  int32_t const_accel_dir;  /* 1..3 */
  int32_t hydro;            /* 1: PLM + HLLE fluxes of an isothermal gas */
  double iso_cs;            /* <hydro>/iso_sound_speed */
  double dfloor;            /* <hydro>/dfloor (the reference's default is FLT_MIN) */
} dk_gas_src;

typedef struct dk_ctx dk_ctx;

/* ---- context ------------------------------------------------------------------------- */

/* Creates a context on the current HIP device.  blocks: host array of mesh->nmb entries.
 * stream: a hipStream_t (NULL = the null stream).  Replaces the allocations of the
 * DustGasDrag constructor (src/dust/dust.cpp:1160-1300). */
int dk_create(const dk_mesh *mesh, const dk_block *blocks, const dk_physics *phys,
              int32_t form, void *stream, dk_ctx **out);
int dk_destroy(dk_ctx *ctx);

/* Binds the caller's particle SoA (Particles::prtcl_rdata/idata) and gas registers
 * (Hydro::u0/u1).  In DK_FORM_FP32 the particles are copied into the library's
 * mixed-precision cell-sorted store; dk_export_particles writes them back (in the store's
 * order; at a step boundary the stage registers x1, v1, R, which no stage reads before the
 * next stage 1 sets them, are written as x1 = x, v1 = v, R = 0). */
int dk_bind_particles(dk_ctx *ctx, double *rdata, int32_t *idata, int64_t capacity,
                      int64_t npart);
int dk_bind_gas(dk_ctx *ctx, double *u0, double *u1);
int dk_export_particles(dk_ctx *ctx);
/* The stage registers at step boundaries (no counterpart in the reference, whose particle
 * arrays always hold them; for callers that write restarts or pbin files from exported
 * particles).  on = 1: every FP32-form stage keeps x1, v1 and R, so that an export after a step
 * writes the reference's values (x1, v1: the stage-1 copies as the step's migration left them;
 * R: the last stage's drag rate with its consistent correction), at the cost of the last
 * stage's register writes (60 B per particle-step); on = 0 (the default): as above.  The plain
 * FP64 form always holds them. */
int dk_keep_registers(dk_ctx *ctx, int32_t on);
int64_t dk_npart(const dk_ctx *ctx);

/* ---- the task entry points (DK_FORM_FP64) ------------------------------------------ */

/* DustGasDrag::FirstTwoImpRK, src/dust/dust_push.cpp:35-80 */
int dk_first_two_imp_rk(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::ExplicitPush, src/dust/dust_push.cpp:91-300 */
int dk_explicit_push(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::UpdateParticleGIDs -> ShiftShearCrossings + SetNewPrtclGID,
 * src/dust/dust_tasks.cpp:149-162, src/dust/dust.cpp:1668-1702,
 * src/bvals/bvals_part.cpp:115-471 (one rank: no send list) */
int dk_update_particle_gids(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::DepositDrag, src/dust/dust_pm.cpp:104-277 */
int dk_deposit_drag(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::SendDepQP + RecvDepQP (additive ghost fold of qdep),
 * src/dust/dust_tasks.cpp:249-263, src/bvals/bvals_dep.cpp:320-562 */
int dk_fold_deposit_qp(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::GasImplicitSolve (local / dc1 / dc2 / applya, pc2 predictor apply),
 * src/dust/dust_pm.cpp:372-442, src/dust/dust_solver.cpp:126-263,630-724 */
int dk_gas_implicit_solve(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::SendUstar/RecvUstar/SendUstarShr/RecvUstarShr (u* ghost fill),
 * src/dust/dust_tasks.cpp:265-312,429-469 */
int dk_fill_ustar(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::GatherKickPMBR, src/dust/dust_pm.cpp:450-568 */
int dk_gather_kick_pmbr(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::SendPMBR + RecvPMBR (additive ghost fold of dmom),
 * src/dust/dust_tasks.cpp:314-328 */
int dk_fold_pmbr(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::ApplyPMBR, src/dust/dust_pm.cpp:576-711 */
int dk_apply_pmbr(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::AddDragHistoryGas, src/dust/dust_pm.cpp:721-742 */
int dk_add_drag_history_gas(dk_ctx *ctx, const dk_stage *st);
/* DustGasDrag::SortParticles (cell sort of point particles),
 * src/dust/dust_cloud_device.cpp:1881-1980: stable order by (block, cell), all columns.  The
 * FP32 form keeps its particles sorted by (block, tile, cell) every stage: a no-op there. */
int dk_sort_particles(dk_ctx *ctx);
/* No counterpart in the reference.  dk_sort_particles without any device allocation, for
 * callers that cannot allocate: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  The same stable (block, cell) order, bit for bit, with the sort's
 * buffers in a static device array of the library, one per GPU (about 53 MiB, held until the
 * process ends, its address taken at dk_create).  At most DK_SORT_NOALLOC_MAX particles
 * (DK_ECAPACITY beyond, nothing done).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (two contexts, two streams) race.  FP32 form: a no-op, as
 * dk_sort_particles. */
#define DK_SORT_NOALLOC_MAX 262144
int dk_sort_particles_noalloc(dk_ctx *ctx);
/* DustGasDrag::RemoveOutflowParticles, src/dust/dust_tasks.cpp:365-420 (both forms): with an
 * outflow x3 face (DK_X3_OUTFLOW) and a non-zero ledger of particles parked at the faces since
 * the last call (bvals_part.cpp:184-186), every particle of zero mass is removed from the
 * caller's arrays, the others keep their order, and dk_npart drops accordingly (the FP32 form
 * exports its store, compacts the caller's arrays and binds them again).  out[0], out[1]: the
 * mass and count removed by this call; out[2], out[3]: the run totals (x3_outflow_mass,
 * x3_outflow_count).  Call it between steps, as the reference's before_timeintegrator task;
 * waits for the stream. */
int dk_remove_outflow_particles(dk_ctx *ctx, double out[4]);
/* The outflow ledger since the last removal: out[0] mass, out[1] count (bvals_part.cpp:184-186,
 * BoundaryValuesPart::x3_removed).  The FP32 form sums it over its work items in a fixed order.
 * Waits for the stream. */
int dk_x3_removed(dk_ctx *ctx, double out[2]);
/* DustGasDrag::NewTimeStep, src/dust/dust_newdt.cpp:108-149: dt_out = dt_cfl min dx/|v|
 * (both forms; the FP32 form applies a pending correction first, from its FP32 velocities).
 * The FP32 form in the shearing box also caps dt_out where a stage's orbital-advection frame
 * shift q Omega c dt |x1|max (c <= 1 + 1/sqrt(2)) would reach half a cell in x2, the largest
 * shift its stages accept (DK_EUNSUPPORTED beyond); the reference has no such limit. */
int dk_new_time_step(dk_ctx *ctx, double dt_cfl, double *dt_out);

/* The strict drag solver (dk_physics.solver = DK_PCG or DK_ADAPTIVE, FP64 form): its
 * parameters, the <dust> inputs drag_rtol, drag_atol, drag_iter_max, drag_adaptive_order_c,
 * drag_adaptive_rtol_max, drag_adaptive_tref, drag_adaptive_state_floor
 * (src/dust/dust.cpp:467-474; defaults 1e-11, 1e-14, 200, 0.25, 1e-3, 1, 1), and the last
 * implicit solve's figures (the DUST_SOLVER_DIAG line, src/dust/dust_solver.cpp:709-718):
 * out[0] PCG iterations, [1] adaptive fast accept (0/1), [2] true residual, [3] error bound,
 * [4] acceptance target, [5] state scale, [6] max Q/rho (-1 where the path did not form it).
 * A solve that loses positive definiteness or exhausts drag_iter_max returns DK_ESOLVER from
 * dk_gas_implicit_solve (the reference aborts, dust_solver.cpp:468-526). */
int dk_set_solver(dk_ctx *ctx, double rtol, double atol, int32_t iter_max, double order_c,
                  double rtol_max, double tref, double state_floor);
int dk_solver_stats(dk_ctx *ctx, double out[7]);

/* ---- whole stage ------------------------------------------------------------------- */

/* All dust tasks of one stage in task-graph order (dust_tasks.cpp:58-109), gas hook
 * included when src != NULL.  The only stage entry of DK_FORM_FP32, where it runs the
 * fused cell-sorted passes; with graphs enabled (dk_set_graphs) a stage is captured on the
 * context's stream into a HIP graph and launched (the executable graph of the same stage
 * index is updated in place on later calls).  Not while the caller's stream is capturing. */
int dk_stage_run(dk_ctx *ctx, const dk_stage *st, const dk_gas_src *src);
int dk_set_graphs(dk_ctx *ctx, int32_t enable);
/* The number of stage graphs launched so far (no counterpart in the reference: it tells a
 * caller whether its stages ran as graphs, which needs a stream other than the null stream). */
int64_t dk_graph_launches(const dk_ctx *ctx);

/* The autotuner (no counterpart in the reference; the "work-tile and block sizes picked by a
 * project-written autotuner" of the FP32 form).  mode 1: take this configuration's entry from
 * the JSON file named by DK_TUNE_CACHE, else tune online over the next dk_stage_run calls
 * (2 warm-up stages, then 2 active stages per candidate, timed on the context's stream; the
 * fastest is kept and stored in the cache file).  The candidates are launch parameters that
 * leave every result bit-identical (the dealing of tiles to XCDs, the key pass's workgroup
 * size), so tuning on the caller's own stages changes no result.  mode 0: stop tuning.  The
 * plain FP64 form has nothing to tune (DK_OK). */
int dk_autotune(dk_ctx *ctx, int32_t mode);
/* state: 0 fixed, 1 tuning, 2 tuned; the current choices */
int dk_tune_config(dk_ctx *ctx, int32_t *state, int32_t *xcd, int32_t *key_threads);
/* sets the choices (xcd 0 or 1; key_threads 128, 256 or 512, at least the tile's cells) */
int dk_set_tune_config(dk_ctx *ctx, int32_t xcd, int32_t key_threads);

/* ---- gas stand-ins of the harness (not part of the reference's dust path) --------------- */

/* Hydro::RKUpdate (zero fluxes, or the isothermal PLM + HLLE fluxes when src->hydro) +
 * Hydro::HydroSrcTerms (see dk_gas_src). */
int dk_gas_hook(dk_ctx *ctx, const dk_stage *st, const dk_gas_src *src);
/* Hydro::SendU/RecvU: periodic (and shear-periodic) copy of the u0 ghost cells. */
int dk_fill_gas_ghosts(dk_ctx *ctx, const dk_stage *st);

/* ---- diagnostics ------------------------------------------------------------------- */

/* The FP32 form's counting sort of its last stage (no counterpart in the reference: it lets a
 * caller check the sort's integer output against an independent stable sort).  which 0: the
 * tile-major key of every particle in the order before the sort (n = dk_npart entries; key =
 * (block*tiles_per_block + tile)*cells_per_tile + cell in tile), 1: the permutation (slot of the
 * new order -> slot of the old order, n entries), 2: the first slot of every key in the new order
 * (nmb*nx1*nx2*nx3 + 1 entries).  host: n int32 of host memory; DK_EINVAL when n is not that size,
 * DK_EUNSUPPORTED in the FP64 form.  Waits for the stream. */
int dk_sort_state(dk_ctx *ctx, int32_t which, int32_t *host, int64_t n);

/* FP64 sums over particles and active cells (waits for the stream):
 * out[0..2] dust momentum sum m v, out[3..5] gas momentum sum V M, out[6] dust mass,
 * out[7] gas mass sum V rho, out[8] sum m |v| + sum V |M| (the conservation scale).
 * Cf. DustGasDrag::ParticleHistory, src/dust/dust_science.cpp:391-484. */
int dk_diag_sums(dk_ctx *ctx, double out[9]);
/* Per-kernel timing with HIP events recorded around the library's launches on the context's
 * stream (the counterpart of the reference's fenced dust ProfTimer, src/dust/dust.hpp:902-954,
 * without fences).  dk_prof_read returns entry idx (label, total milliseconds, launches) and
 * DK_EINVAL past the last entry; it waits for the stream. */
int dk_prof_enable(dk_ctx *ctx, int32_t on);
int dk_prof_reset(dk_ctx *ctx);
int dk_prof_read(dk_ctx *ctx, int32_t idx, char *label, int32_t label_len, double *ms,
                 int64_t *count);
/* UpdateStiffness / SyncStiffness (src/dust/dust_pm.cpp:288-359; dust.hpp stiff_chi_last,
 * stiff_chi_run_max, stiff_rate_max): with the stiffness field on (dk_physics.stiffness) every
 * pc2 cycle records chi = dt max_cells qdep[4]/rho before the predictor is applied; out[0] = chi
 * of the last cycle, out[1] = the running maximum, out[2] = the rate maximum max qdep[4]/rho
 * (NewTimeStep's stiff_thresh/stiff_rate_max limit, dust_newdt.cpp:114-119).  One rank: no
 * reduction over ranks.  Both forms. */
int dk_stiffness(dk_ctx *ctx, double out[3]);
/* RefreshStoppingTimeMaximum (src/dust/dust_newdt.cpp:34-98): validates every particle's
 * stopping time (finite, positive, species index not negative; DK_ESTOP otherwise) and returns
 * out[0] = the largest, out[1] = the smallest (taus_max, taus_min), which GammaSwitch
 * (dust_newdt.cpp:160-173) compares with dt on the host.  The FP32 form reduces its FP32 copies.
 * Waits for the stream. */
int dk_stopping_times(dk_ctx *ctx, double out[2]);
/* The first device-side error (code, particle index); clears it. */
int dk_error(dk_ctx *ctx, int32_t *code, int64_t *particle);
/* Device pointer to a work field: 0 qdep(5), 1 ustar(3), 2 dmom(3), 3 solver ap(3),
 * 4 imex residual(4); layout as the gas (nvar components).  NULL if absent. */
void *dk_field(dk_ctx *ctx, int32_t which, int32_t *nvar);
const char *dk_strerror(int code);

/* ---- science outputs: the particle file (pbin) and the particle history (phst) ---------- */

/* What the outputs record beside the particles: Mesh::time, Mesh::dt, Mesh::ncycle, the output
 * block's file_number and precision, the species table, and the run's strings (NULL: empty;
 * source_commit NULL: "unknown"; input_dump: the ParameterInput::ParameterDump text that a
 * pbin file carries as its [input] section). */
typedef struct dk_output_info {
  double time, dt;
  int64_t cycle;
  int32_t file_number;
  int32_t single_precision;   /* pbin <outputN>/precision = single: particle reals as float32 */
  int32_t stopping_time_mode; /* 0 species_fixed, 1 particle_static, 2 dynamic */
  int32_t nspecies;           /* 1..8 */
  double taus[8];             /* the species' stopping times */
  const char *basename, *output_block, *source_commit, *input_dump;
} dk_output_info;

/* ParticleBinaryOutput::WriteOutputFile (src/outputs/pbin.cpp:6-60 the format, :450-1026):
 * writes the active particles (mass != 0) and the blocks of the context to `path`, format
 * version 2 (raw columns; species, stopping time and mass as [constants] under the reference's
 * rules), as path.tmp renamed when complete.  One rank.  The FP32 form exports its particles
 * first.  Waits for the stream. */
int dk_write_pbin(dk_ctx *ctx, const char *path, const dk_output_info *info);
/* DustGasDrag::ParticleHistory (src/dust/dust_science.cpp:372-484): row[0..13] = n_particles,
 * mass, mom1..3 (sum m v), ekin, z_rms (the particle sums per chunk of 4096 rows, the chunks
 * added in order on the host), rho_max and the center and level of its cell (the mass deposit
 * of the point particles with the coupling's stencil and fold at `time` (with x3_dep_drop the
 * parts beyond physical x3 faces are dropped, as the coupling's; the reference's science deposit
 * folds them back), of clouds on a uniform mesh with TSC, their exact cell integrals; clouds on
 * a refined mesh: rho_max = -1, not built), chi_max and dt_stiff_min (the interval records of UpdateStiffness
 * since the previous call, -1 when none; dt_stiff_min needs stiff_thresh > 0).  Resets the
 * interval records.  One rank.  Waits for the stream. */
int dk_particle_history(dk_ctx *ctx, double time, double stiff_thresh, double row[14]);
/* ParticleHistoryOutput::WriteOutputFile (src/outputs/phst.cpp:80-163): appends time, dt and
 * row to `path` (%.16e columns), preceded by the header block when header != 0. */
int dk_write_phst(dk_ctx *ctx, const char *path, const dk_output_info *info, const double row[14],
                  int32_t header);

#ifdef __cplusplus
}
#endif

#endif  /* DUSTK_H_ */
