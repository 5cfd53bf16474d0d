"""The five drag-coupling workloads (mesh, drag options, seeded initial condition, dt, steps),
plus small variants of them for tests.  Each returns a Workload; nothing here allocates
device memory.

  c1  32^3 periodic, 1 block, ppc 1, tau_s 0.1, eps 1, imex2+/dc1, dt 1e-3, 100 steps
  c2  unstratified shearing box, 256^3 in 64 blocks of 64^3, ppc 4, St 0.1, eps 3,
      NSH equilibrium + 1e-4 noise, imex2+/dc1, dt 2e-4, 50 steps
  c3  stratified (0.2 H)^3 at 256^3, 7 log-uniform Stokes bins over [1e-3, 1],
      mass_slope 0.5, Z 0.01, 1 per bin per cell, H_p 0.02, x3 outflow, imex2+/dc1
  c4  batching over uneven blocks: a statically refined mesh (64^3 root in 16^3 blocks, levels
      1 and 2 in nested regions, 792 blocks), lognormal particle counts per block (sigma 1.5,
      mean 4 per cell), refinable clouds of their block's level, tau_s 0.05, eps 1, pc2
  c4u the root level of c4 alone (64^3 in 16^3 blocks, same-level clouds) with a quarter of the
      particles in 16 Gaussian clumps of width 2 cells; c4p: 128^3 in 512 blocks with point
      particles; c4b: 128^3 in 512 blocks of 16^3 with c4's lognormal block counts, point
      particles (the uneven-block case in a form both layouts run)
  c5  128^3 periodic in 8 blocks of 64^3, ppc 8, two species (tau_s 0.01 and 1, eps 0.5
      each), half the particles in 64 Gaussian clumps of width 2 cells, imex2+/dc1, 100 steps
  c1h c1's box at 256^3 in 64 blocks with the isothermal gas update (PLM + HLLE fluxes, cs 1)
"""

import dataclasses
import math

from . import abi


@dataclasses.dataclass
class Workload:
    name: str
    mesh: abi.dk_mesh
    blocks: object
    phys: abi.dk_physics
    setup: abi.dk_setup
    src: object          # abi.dk_gas_src or None
    dt: float
    steps: int

    @property
    def ncells(self):
        m = self.mesh
        return m.nmb * m.nx1 * m.nx2 * m.nx3

    @property
    def npart(self):
        if self.setup.block_lognormal:
            return int(sum(block_counts(self.mesh, self.setup)))
        return self.setup.ppc * self.ncells

    @property
    def ppc(self):
        return self.setup.ppc


_M64 = (1 << 64) - 1


def _hash_uniform01(key):
    """HashUniform01 (src/pgen/dust_amr/dust_setup.hpp:35-50): the SplitMix64 finaliser."""
    key = (key + 0x9e3779b97f4a7c15) & _M64
    key = ((key ^ (key >> 30)) * 0xbf58476d1ce4e5b9) & _M64
    key = ((key ^ (key >> 27)) * 0x94d049bb133111eb) & _M64
    key ^= key >> 31
    return (key >> 11) * (1.0 / 9007199254740992.0)


def block_counts(mesh, su):
    """The particle count of every block of an uneven set-up (dk_setup.block_lognormal), as the
    library and the oracle draw it on the host."""
    ncell = mesh.nx1 * mesh.nx2 * mesh.nx3
    if not su.block_lognormal:
        return [su.ppc * ncell] * mesh.nmb
    base = (su.seed * 0x9e3779b97f4a7c15) & _M64
    out = []
    for m in range(mesh.nmb):
        k = (base + m * 0x2545f4914f6cdd1d + 0x7137449123ef65cd) & _M64
        u1 = max(_hash_uniform01(k), 1e-300)
        u2 = _hash_uniform01((k + 0x1d8e4e27c47d124f) & _M64)
        z = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
        sg = su.block_sigma
        out.append(int(math.floor(su.ppc * ncell * math.exp(sg * z - 0.5 * sg * sg) + 0.5)))
    return out


def _setup(seed=1, ppc=1, taus=(0.1,), eps=(1.0,), rho0=1.0, gas_noise=0.0, dust_noise=0.0,
           **kw):
    su = abi.dk_setup(seed=seed, ppc=ppc, nspecies=len(taus), rho0=rho0,
                      gas_noise=gas_noise, dust_noise=dust_noise)
    for k, (t, e) in enumerate(zip(taus, eps)):
        su.taus[k], su.eps[k] = t, e
    for k, v in kw.items():
        if k in ("tau_lo", "tau_hi"):
            for i, x in enumerate(v):
                getattr(su, k)[i] = x
        else:
            setattr(su, k, v)
    return su


def _phys(coupling=abi.DK_IMEX, solver=abi.DK_DC1, sbox=False, strat=False, omega0=1.0,
          qshear=1.5, deposit=abi.DK_TSC, back_reaction=1, consistent_gas=1, clouds=0):
    return abi.dk_physics(deposit=deposit, coupling=coupling, solver=solver,
                          back_reaction=back_reaction, consistent_gas=consistent_gas,
                          shearing_box=int(sbox), stratified=int(strat), omega0=omega0,
                          qshear=qshear, pc2_damping=0, stiffness=0, clouds=int(clouds))


def nsh_gas(omega0, qshear, ax, eps, taus):
    """SolveNSHGas, src/pgen/dust_amr/dust_setup.cpp:186-200."""
    b = 2.0 - qshear
    sa = sb = 0.0
    for e, t in zip(eps, taus):
        tt = omega0 * t
        dd = 1.0 + 2.0 * b * tt * tt
        sa += e / dd
        sb += e * tt / dd
    den = (1.0 + sa) * (1.0 + sa) + 2.0 * b * sb * sb
    return (ax / omega0) * sb / den, -(ax / (2.0 * omega0)) * (1.0 + sa) / den


def c1(n=32, nb=1, ppc=1, steps=100, coupling=abi.DK_IMEX, solver=abi.DK_DC1, seed=1):
    """Plain method, FP64: [0,1)^3 periodic, tau_s 0.1, eps 1, 1e-3 noise on gas and dust."""
    nx = n // nb
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nb, nb, nb), ((0.0, 1.0),) * 3, ng=2)
    su = _setup(seed=seed, ppc=ppc, taus=(0.1,), eps=(1.0,), gas_noise=1e-3, dust_noise=1e-3)
    return Workload("c1", mesh, blocks, _phys(coupling=coupling, solver=solver), su, None,
                    1.0e-3, steps)


def c2(n=256, nbx=4, ppc=4, steps=50, seed=1):
    """Unstratified shearing box at the NSH equilibrium (dust_nsh, SolveNSHGas)."""
    nx = n // nbx
    L = 0.1
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nbx, nbx, nbx), ((-L, L),) * 3, ng=2,
                                    bc_x1=abi.DK_X1_SHEAR)
    omega0, qshear, etavk, eps, ts = 1.0, 1.5, 0.05, 3.0, 0.1
    ax = 2.0 * omega0 * etavk
    ugx, ugy = nsh_gas(omega0, qshear, ax, (eps,), (ts,))
    su = _setup(seed=seed, ppc=ppc, taus=(ts,), eps=(eps,), dust_noise=1e-4, gas_vx=ugx,
                gas_vy=ugy, nsh=1, omega0=omega0, qshear=qshear)
    src = abi.dk_gas_src(shearing_box=1, stratified=0, omega0=omega0, qshear=qshear,
                         const_accel=ax, const_accel_dir=1)
    return Workload("c2", mesh, blocks, _phys(sbox=True, omega0=omega0, qshear=qshear), su,
                    src, 2.0e-4, steps)


def c3(n=256, nbx=4, steps=50, seed=1):
    """Stratified box with a continuous size distribution (dust_si_strat, continuous)."""
    nx = n // nbx
    L = 0.1
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nbx, nbx, nbx), ((-L, L),) * 3, ng=2,
                                    bc_x1=abi.DK_X1_SHEAR, bc_x3=abi.DK_X3_OUTFLOW)
    nbin = 7
    lo, hi = math.log(1e-3), math.log(1.0)
    edges = [math.exp(lo + (hi - lo) * k / nbin) for k in range(nbin + 1)]
    omega0, qshear = 1.0, 1.5
    su = _setup(seed=seed, ppc=nbin, taus=tuple(edges[:-1]), eps=(0.0,) * nbin,
                continuous=1, tau_lo=edges[:-1], tau_hi=edges[1:], mass_slope=0.5,
                gauss_z=1, hp=0.02, gas_strat=1, zmetal=0.01, dust_noise=0.0)
    src = abi.dk_gas_src(shearing_box=1, stratified=1, omega0=omega0, qshear=qshear,
                         const_accel=0.1, const_accel_dir=1)
    return Workload("c3", mesh, blocks, _phys(sbox=True, strat=True, omega0=omega0,
                                              qshear=qshear), su, src, 2.0e-4, steps)


def c5(n=128, nbx=2, ppc=8, steps=100, seed=1):
    """Clumped two-species box: half of the particles in 64 Gaussian clumps of 2 cells."""
    nx = n // nbx
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nbx, nbx, nbx), ((0.0, 1.0),) * 3, ng=2)
    su = _setup(seed=seed, ppc=ppc, taus=(0.01, 1.0), eps=(0.5, 0.5), gas_noise=1e-3,
                dust_noise=1e-3, nclump=64, clump_frac=0.5, clump_sigma=2.0 / n)
    return Workload("c5", mesh, blocks, _phys(), su, None, 1.0e-3, steps)


FLT_MIN = 1.17549435e-38   # the density floor's default (<hydro>/dfloor)


def with_hydro(w, cs=1.0, dfloor=FLT_MIN):
    """w with the explicit gas update's isothermal PLM + HLLE fluxes (dk_gas_src.hydro, sound
    speed cs) on top of its gas source terms; its name gains an "h"."""
    src = abi.dk_gas_src()
    if w.src is not None:
        for f, _ in abi.dk_gas_src._fields_:
            setattr(src, f, getattr(w.src, f))
    src.hydro, src.iso_cs, src.dfloor = 1, cs, dfloor
    return dataclasses.replace(w, name=w.name + "h", src=src)


def c1h(n=256, nbx=4, ppc=1, steps=50, coupling=abi.DK_IMEX, solver=abi.DK_DC1, seed=1):
    """c1's box at n^3 in nbx^3 blocks with the isothermal gas update (cs = 1; dt 1e-3 is a
    Courant number of 0.256 at 256^3): the gas moves by its own fluxes as well as the drag."""
    return with_hydro(c1(n=n, nb=nbx, ppc=ppc, steps=steps, coupling=coupling, solver=solver,
                         seed=seed))


def c4u(n=64, nbx=4, ppc=4, steps=50, seed=1, clump_frac=0.25, clouds=True):
    """pc2 with refinable clouds (CloudPredictorT / CloudAdvanceT) on a uniform mesh (the root
    level of the SMR case alone); clumps make the per-block particle counts uneven.
    clouds=False: the same case with point particles (dust_pc2_predictor / dust_pc2_advance)."""
    nx = n // nbx
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nbx, nbx, nbx), ((0.0, 1.0),) * 3, ng=2)
    su = _setup(seed=seed, ppc=ppc, taus=(0.05,), eps=(1.0,), gas_noise=1e-3, dust_noise=1e-3,
                nclump=16 if clump_frac > 0 else 0, clump_frac=clump_frac,
                clump_sigma=2.0 / n)
    return Workload("c4u" if clouds else "c4p", mesh, blocks,
                    _phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=int(clouds)), su,
                    None, 1.0e-3, steps)


def c4p(n=128, nbx=8, ppc=4, steps=50, seed=1, clump_frac=0.25):
    """The c4 case with point particles (pc2 without clouds) at the SMR case's scale: 128^3 in
    512 blocks of 16^3, clumps making the per-block counts uneven (8,388,608 particles).  The
    clump centres reach dust-to-gas ratios of ~260 (chi = dt eps/t_s ~ 5), where the explicit pc2
    predictor is unstable: the rate field and the reference's damping of stiff cells are on
    (StiffnessActive, dust_pm.cpp:288-402)."""
    w = c4u(n=n, nbx=nbx, ppc=ppc, steps=steps, seed=seed, clump_frac=clump_frac, clouds=False)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    return w


# the refined regions of the SMR case in root-block units (root 4^3 blocks): level 1 over the
# middle half in x3 (all of x1, x2: periodic), level 2 inside it, one level-1 block from every
# level-1 face (the 2:1 balance)
C4_REGIONS = (((0.0, 4.0), (0.0, 4.0), (1.0, 3.0)), ((0.5, 3.5), (0.5, 3.5), (1.5, 2.5)))


def c4(n=64, nbx=4, ppc=4, steps=50, seed=1, sigma=1.5, regions=C4_REGIONS):
    """Batching over uneven blocks: a statically refined mesh, root n^3 in nbx^3 blocks of
    (n/nbx)^3 with levels 1 and 2 over C4_REGIONS (792 blocks at the defaults), periodic.
    Block m holds round(ppc (n/nbx)^3 exp(sigma z_m - sigma^2/2)) particles, z_m ~ N(0, 1) keyed
    by the block (lognormal, mean ppc per cell), uniform in the block, each a refinable cloud of
    its block's level (half-width one cell of that level); eps 1 (the mass of a cloud of level
    L is eps rho V_L / ppc), tau_s 0.05, pc2 with the rate field and the damping of stiff cells
    (the densest blocks reach eps ~ 30)."""
    nx = n // nbx
    mesh, blocks = abi.smr_mesh((nx, nx, nx), (nbx, nbx, nbx), ((0.0, 1.0),) * 3, regions)
    su = _setup(seed=seed, ppc=ppc, taus=(0.05,), eps=(1.0,), gas_noise=1e-3, dust_noise=1e-3,
                block_lognormal=1, block_sigma=sigma)
    w = Workload("c4", mesh, blocks,
                 _phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=1), su, None,
                 1.0e-3, steps)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    return w


def c4b(n=128, nbx=8, ppc=4, steps=50, seed=1, sigma=1.5):
    """Uneven blocks in one launch for both forms: pc2 point particles on 128^3 in 512 blocks of
    16^3 with the lognormal per-block counts of c4 (sigma 1.5, mean ppc per cell, uniform in each
    block), tau_s 0.05, eps 1; the densest blocks reach eps ~ 30, so the rate field and the
    damping of stiff cells are on (as c4p)."""
    nx = n // nbx
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nbx, nbx, nbx), ((0.0, 1.0),) * 3, ng=2)
    su = _setup(seed=seed, ppc=ppc, taus=(0.05,), eps=(1.0,), gas_noise=1e-3, dust_noise=1e-3,
                block_lognormal=1, block_sigma=sigma)
    w = Workload("c4b", mesh, blocks, _phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL), su,
                 None, 1.0e-3, steps)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    return w


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c4u": c4u, "c4p": c4p, "c4b": c4b,
             "c5": c5, "c1h": c1h}
