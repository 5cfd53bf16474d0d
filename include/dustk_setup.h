/*
 * dustk_setup.h -- deterministic synthetic initial conditions (the harness's problem
 * generator).  No counterpart in the reference's drag FFI: it restates the per-particle draws
 * of the reference's dust problem generators (src/pgen/dust_amr/dust_setup.hpp:35-206,
 * src/pgen/dust_amr/dust_setup.cpp:735-750, src/dust/dust.cpp:1716-1778) so that the HIP
 * library and the CPU oracle start from identical, index-keyed particle sets.
 *
 * Particle q of block m (q = m*npb + r, npb = ppc * nx1*nx2*nx3; with block_lognormal q runs
 * over the blocks' prefix offsets instead) gets PGID = m, PTAG = q and the draws below, each a
 * function of (seed, q) only:
 *   key  = q + seed*0x9e3779b97f4a7c15
 *   x    = x1min_m + U(key + 0x632be59bd9b4e019) * lx_m        (RootPlacement::Base3 mode 0)
 *   y    = x2min_m + U(key + 0x8cb92baa3f3d8dd7) * ly_m
 *   z    = x3min_m + U(key + 0x58f38ded6f7c55b5) * lz_m        (uniform), or a Gaussian of
 *          width hp redrawn until inside the mesh (Base3 with gauss, 64 tries)
 *   U    = HashUniform01 (SplitMix64 finaliser, 53 bits; dust_setup.hpp:35-50)
 *   N(s) = s*sqrt(-2 ln max(U(k),1e-300))*cos(2 pi U(k + 0x1d8e4e27c47d124f)), one key per
 *          velocity component: k = key + DK_SALT_V[c]
 * Species of particle q: q % nspecies (the legacy striping of SetDefaultMasses); mass
 * eps_s*rho0*V_cell*nspecies/ppc; stopping time taus_s, or for continuous bins a log-uniform
 * draw in [tau_lo_s, tau_hi_s] with mass weight tau^mass_slope / <tau^mass_slope>_bin.
 * Gas: rho = rho0 (times exp(-z^2/2) if stratified), momentum rho*(u_mean + N(gas_noise))
 * with the cell key (global cell index) + seed*0x9e3779b97f4a7c15 + DK_SALT_G[c]; on a
 * statically refined mesh the global cell index is that of the block's level, plus
 * level * 2^40.  Block counts (block_lognormal): z_m = N with the key
 * seed*0x9e3779b97f4a7c15 + m*0x2545f4914f6cdd1d + DK_SALT_BLOCK, evaluated on the host.
 */
#ifndef DUSTK_SETUP_H_
#define DUSTK_SETUP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DK_MAXSPEC 8

typedef struct dk_setup {
  uint64_t seed;
  int32_t ppc;                 /* particles per cell, all species together */
  int32_t nspecies;            /* striped over the particle index (q % nspecies) */
  double taus[DK_MAXSPEC];     /* stopping time of each species (discrete) */
  double eps[DK_MAXSPEC];      /* dust-to-gas mass ratio of each species */
  int32_t continuous;          /* 1: per-particle tau, log-uniform in [tau_lo, tau_hi] */
  double tau_lo[DK_MAXSPEC], tau_hi[DK_MAXSPEC];
  double mass_slope;           /* particle mass weight tau^mass_slope (continuous) */
  double rho0;                 /* gas density (midplane value if stratified) */
  double gas_vx, gas_vy, gas_vz;   /* mean gas velocity */
  double gas_noise;            /* Gaussian sigma per gas velocity component and cell */
  double dust_noise;           /* Gaussian sigma per dust velocity component */
  int32_t nsh;                 /* 1: dust mean velocity = NSHDustDrift(tau) of the gas mean */
  double omega0, qshear;       /* for nsh */
  int32_t gauss_z;             /* 1: dust heights ~ N(0, hp) inside the mesh */
  double hp;
  int32_t gas_strat;           /* 1: rho = rho0 exp(-z^2/2) (H = 1) */
  int32_t nclump;              /* clumps: a fraction clump_frac of all particles sits at */
  double clump_frac;           /*   N(centre, clump_sigma) around nclump centres hashed from */
  double clump_sigma;          /*   the seed; the rest is uniform over the mesh; PGID is then
                                    the block holding the position */
  double zmetal;               /* > 0: masses scaled so that total dust mass = zmetal * total
                                  gas mass, equal per species (stratified problems) */
  int32_t block_lognormal;     /* 1: uneven blocks: block m holds n_m = round(ppc * nx1*nx2*nx3 *
                                  exp(block_sigma z_m - block_sigma^2/2)) particles, z_m = N(1)
                                  keyed by (seed, m), uniform in the block (the particles of
                                  block m follow those of blocks 0..m-1); with refinable clouds
                                  PLEV is the block's level.  No clumps, no Gaussian heights. */
  double block_sigma;
} dk_setup;

/* salts of the velocity and gas draws */
#define DK_SALT_V0 0x3c6ef372fe94f82bULL
#define DK_SALT_V1 0xa54ff53a5f1d36f1ULL
#define DK_SALT_V2 0x510e527fade682d1ULL
#define DK_SALT_G0 0x9b05688c2b3e6c1fULL
#define DK_SALT_G1 0x1f83d9abfb41bd6bULL
#define DK_SALT_G2 0x5be0cd19137e2179ULL
#define DK_SALT_TAU 0xcbbb9d5dc1059ed8ULL
#define DK_SALT_CLUMP 0x629a292a367cd507ULL
#define DK_SALT_BLOCK 0x7137449123ef65cdULL

struct dk_ctx;

/* Fills the particle and gas arrays bound to ctx (dk_bind_particles, dk_bind_gas) with the
 * problem of su on the device and sets the particle count (capacity must hold
 * ppc * nmb * nx1*nx2*nx3); u0 ghosts are filled and u1 = u0. */
int dk_setup_problem(struct dk_ctx *ctx, const dk_setup *su);

#ifdef __cplusplus
}
#endif

#endif  /* DUSTK_SETUP_H_ */
