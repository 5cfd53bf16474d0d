"""The five drag-coupling workloads (mesh, drag options, seeded initial condition, dt, steps),
plus small variants of them for tests.  Each returns a Workload; nothing here allocates
device memory.

  c1  32^3 periodic, 1 block, ppc 1, tau_s 0.1, eps 1, imex2+/dc1, dt 1e-3, 100 steps
  c2  unstratified shearing box, 256^3 in 64 blocks of 64^3, ppc 4, St 0.1, eps 3,
      NSH equilibrium + 1e-4 noise, imex2+/dc1, dt 2e-4, 50 steps
  c3  stratified (0.2 H)^3 at 256^3, 7 log-uniform Stokes bins over [1e-3, 1],
      mass_slope 0.5, Z 0.01, 1 per bin per cell, H_p 0.02, x3 outflow, imex2+/dc1
  c4  refinable clouds with pc2 at the root level of the SMR case: 64^3 periodic in blocks
      of 16^3, tau_s 0.05, eps 1, ppc 4, a quarter of the particles in 16 Gaussian clumps of
      width 2 cells (the nested refined boxes of the SMR case are not built: same-level clouds)
  c5  128^3 periodic in 8 blocks of 64^3, ppc 8, two species (tau_s 0.01 and 1, eps 0.5
      each), half the particles in 64 Gaussian clumps of width 2 cells, imex2+/dc1, 100 steps
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
        return self.setup.ppc * self.ncells

    @property
    def ppc(self):
        return self.setup.ppc


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


def c4(n=64, nbx=4, ppc=4, steps=50, seed=1, clump_frac=0.25, clouds=True):
    """pc2 with same-level refinable clouds (CloudPredictorT / CloudAdvanceT) on the root level
    of the SMR case; clumps make the per-block particle counts uneven.  clouds=False: the same
    case with point particles (dust_pc2_predictor / dust_pc2_advance)."""
    nx = n // nbx
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nbx, nbx, nbx), ((0.0, 1.0),) * 3, ng=2)
    su = _setup(seed=seed, ppc=ppc, taus=(0.05,), eps=(1.0,), gas_noise=1e-3, dust_noise=1e-3,
                nclump=16 if clump_frac > 0 else 0, clump_frac=clump_frac,
                clump_sigma=2.0 / n)
    return Workload("c4" if clouds else "c4p", mesh, blocks,
                    _phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=int(clouds)), su,
                    None, 1.0e-3, steps)


def c4p(n=128, nbx=8, ppc=4, steps=50, seed=1, clump_frac=0.25):
    """The c4 case with point particles (pc2 without clouds) at the SMR case's scale: 128^3 in
    512 blocks of 16^3, clumps making the per-block counts uneven (8,388,608 particles).  The
    clump centres reach dust-to-gas ratios of ~260 (chi = dt eps/t_s ~ 5), where the explicit pc2
    predictor is unstable: the rate field and the reference's damping of stiff cells are on
    (StiffnessActive, dust_pm.cpp:288-402)."""
    w = c4(n=n, nbx=nbx, ppc=ppc, steps=steps, seed=seed, clump_frac=clump_frac, clouds=False)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    return w


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c4p": c4p, "c5": c5}
