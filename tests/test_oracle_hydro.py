"""The oracle's explicit gas update (isothermal PLM + HLLE fluxes, dk_gas_src.hydro), pinned by
what the mathematics fixes: a uniform flow is steady, the fluxes conserve mass and momentum,
a linear sound wave travels at cs with second-order errors, the x2 and x3 sweeps are the x1
sweep rotated, and the damped dusty sound wave (gas + drag) follows the dispersion relation
omega^2 (1 + eps/(1 - i omega ts)) = k^2 cs^2 (src/pgen/dust_amr/dusty_wave_amr.cpp:57-79)."""
import ctypes as C

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle, lib
from helpers import active
from test_oracle import place

FLT_MIN = 1.17549435e-38


def hydro_src(cs=1.0, **kw):
    return abi.dk_gas_src(hydro=1, iso_cs=cs, dfloor=FLT_MIN, **kw)


def gas_hook(o, st, src):
    return lib().or_gas_hook(o.h, C.byref(st), C.byref(src))


def euler_stage(dt):
    st = abi.dk_stage(stage=1, nexp_stages=2, time=0.0, dt=dt)
    st.gam0[0], st.gam1[0], st.beta[0] = 1.0, 0.0, 1.0
    return st


def wave_mesh(nx, axis=0, nb=1, bc_x3=abi.DK_X3_PERIODIC):
    """nx cells (nb blocks) along the wave's axis, 2 cells across, unit wavelength box."""
    dx = 1.0 / nx
    n = [2, 2, 2]
    n[axis] = nx // nb
    b = [1, 1, 1]
    b[axis] = nb
    box = [(0.0, 2 * dx)] * 3
    box[axis] = (0.0, 1.0)
    return abi.uniform_mesh(tuple(n), tuple(b), tuple(box), ng=2, bc_x3=bc_x3)


def set_wave(o, axis, amp, k=2 * np.pi, cs=1.0):
    """A right-going linear sound wave along axis: rho = 1 + amp cos(k x), v = amp cs cos(k x)."""
    mesh, ng = o.mesh, o.mesh.ng
    nxa = (mesh.nx1, mesh.nx2, mesh.nx3)[axis]
    dx = 1.0 / (nxa * (mesh.nmbx1, mesh.nmbx2, mesh.nmbx3)[axis])
    u0 = o.u0()
    for m in range(mesh.nmb):
        lb = (m % mesh.nmbx1, (m // mesh.nmbx1) % mesh.nmbx2, m // (mesh.nmbx1 * mesh.nmbx2))
        x = (lb[axis] * nxa + np.arange(-ng, nxa + ng) + 0.5) * dx
        d = 1.0 + amp * np.cos(k * x)
        shape = [1, 1, 1]
        shape[2 - axis] = x.size
        u0[m] = 0.0
        u0[m, 0] = d.reshape(shape)
        u0[m, 1 + axis] = (d * amp * cs * np.cos(k * x)).reshape(shape)


def run_wave(o, T, cs=1.0, cfl=0.4, nx=None):
    """Heun stages of the gas update alone (u1 = u0 at each step start, as FirstTwoImpRK)."""
    dt = cfl / nx / cs
    nst = int(np.ceil(T / dt))
    dt = T / nst
    src = hydro_src(cs)
    t = 0.0
    for _ in range(nst):
        o.u1()[:] = o.u0()
        for st in abi.rk2_stages(t, dt):
            assert gas_hook(o, st, src) == abi.DK_OK
        t += dt


def density_profile(o, axis):
    """The active density along the wave's axis (blocks in order), averaged across."""
    mesh = o.mesh
    ua = active(o.u0(), mesh.ng)
    prof = []
    nba = (mesh.nmbx1, mesh.nmbx2, mesh.nmbx3)[axis]
    for b in range(nba):
        m = (b, b * mesh.nmbx1, b * mesh.nmbx1 * mesh.nmbx2)[axis]
        d = ua[m, 0]
        prof.append(d.mean(axis=tuple(a for a in range(3) if a != 2 - axis)))
    return np.concatenate(prof)


def test_uniform_flow_is_steady():
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (2, 1, 1), ((0, 1),) * 3, ng=2)
    o = Oracle(mesh, blocks, workloads._phys(), 16)
    u0 = o.u0()
    u0[:, 0] = 1.3
    u0[:, 1], u0[:, 2], u0[:, 3] = 1.3 * 0.2, 1.3 * -0.1, 1.3 * 0.05
    before = u0.copy()
    assert gas_hook(o, euler_stage(0.01), hydro_src(0.7)) == abi.DK_OK
    np.testing.assert_array_equal(active(o.u0(), 2), active(before, 2))


def test_fluxes_conserve_mass_and_momentum():
    rng = np.random.default_rng(3)
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (2, 2, 2), ((0, 1),) * 3, ng=2)
    o = Oracle(mesh, blocks, workloads._phys(), 16)
    u0 = o.u0()
    d = 1.0 + 0.5 * rng.random(u0[:, 0].shape)
    u0[:, 0] = d
    for c in range(1, 4):
        u0[:, c] = d * rng.normal(scale=0.5, size=d.shape)
    before = active(o.u0(), 2).copy()
    tot0 = before.sum(axis=(0, 2, 3, 4))
    assert gas_hook(o, euler_stage(0.01), hydro_src(1.0)) == abi.DK_OK
    ua = active(o.u0(), 2)
    tot1 = ua.sum(axis=(0, 2, 3, 4))
    scale = np.abs(ua).sum(axis=(0, 2, 3, 4))
    assert np.all(np.abs(tot1 - tot0) < 1e-14 * scale)
    assert np.abs(ua - before).max() > 1e-3   # the fluxes moved the gas


def test_sound_wave_travels_at_cs_with_second_order_error():
    amp = 1e-6
    errs = []
    for nx in (32, 64):
        mesh, blocks = wave_mesh(nx)
        o = Oracle(mesh, blocks, workloads._phys(), 16)
        set_wave(o, 0, amp)
        run_wave(o, 1.0, nx=nx)
        x = (np.arange(nx) + 0.5) / nx
        exact = 1.0 + amp * np.cos(2 * np.pi * (x - 1.0))
        errs.append(np.abs(density_profile(o, 0) - exact).mean() / amp)
    assert errs[1] < 0.012
    assert errs[0] / errs[1] > 3.3


@pytest.mark.parametrize("axis", [1, 2])
def test_x2_and_x3_sweeps_are_the_x1_sweep_rotated(axis):
    nx = 32
    prof = []
    for a in (0, axis):
        mesh, blocks = wave_mesh(nx, axis=a)
        o = Oracle(mesh, blocks, workloads._phys(), 16)
        set_wave(o, a, 1e-3)
        run_wave(o, 0.3, nx=nx)
        prof.append(density_profile(o, a))
    np.testing.assert_allclose(prof[1], prof[0], rtol=0, atol=1e-15)


def test_blocks_exchange_ghosts_for_the_fluxes():
    nx = 32
    prof = []
    for nb in (1, 4):
        mesh, blocks = wave_mesh(nx, nb=nb)
        o = Oracle(mesh, blocks, workloads._phys(), 16)
        set_wave(o, 0, 1e-3)
        run_wave(o, 0.3, nx=nx)
        prof.append(density_profile(o, 0))
    np.testing.assert_allclose(prof[1], prof[0], rtol=0, atol=1e-15)


def test_unsupported_options_are_refused():
    st = euler_stage(0.01)
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (1, 1, 1), ((0, 1),) * 3, ng=2,
                                    bc_x1=abi.DK_X1_SHEAR)
    o = Oracle(mesh, blocks, workloads._phys(sbox=True), 16)
    assert gas_hook(o, st, hydro_src()) == abi.DK_EUNSUPPORTED
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (1, 1, 1), ((0, 1),) * 3, ng=1)
    o = Oracle(mesh, blocks, workloads._phys(), 16)
    assert gas_hook(o, st, hydro_src()) == abi.DK_EUNSUPPORTED
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (1, 1, 1), ((0, 1),) * 3, ng=2)
    o = Oracle(mesh, blocks, workloads._phys(), 16)
    assert gas_hook(o, st, hydro_src(0.0)) == abi.DK_EINVAL


def test_reflecting_x3_walls_hold_mass_and_tangential_momentum():
    """Reflecting x3 faces mirror the gas with IM3 negated (the hydro's reflect,
    src/bvals/physics/hydro_bcs.cpp:229-290): the face states at a wall are mirror images, so
    no mass and no x1/x2 momentum crosses it; the wall pushes only on IM3."""
    rng = np.random.default_rng(8)
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (1, 1, 2), ((0, 1), (0, 1), (0, 2)), ng=2,
                                    bc_x3=abi.DK_X3_REFLECT)
    o = Oracle(mesh, blocks, workloads._phys(), 16)
    u0 = o.u0()
    d = 1.0 + 0.5 * rng.random(u0[:, 0].shape)
    u0[:, 0] = d
    for c in range(1, 4):
        u0[:, c] = d * rng.normal(scale=0.3, size=d.shape)
    tot0 = active(o.u0(), 2).sum(axis=(0, 2, 3, 4))
    for s in range(4):
        o.u1()[:] = o.u0()
        for st in abi.rk2_stages(0.0, 0.01):
            assert gas_hook(o, st, hydro_src(1.0)) == abi.DK_OK
    ua = active(o.u0(), 2)
    tot1 = ua.sum(axis=(0, 2, 3, 4))
    scale = np.abs(ua).sum(axis=(0, 2, 3, 4))
    assert np.all(np.abs(tot1[:3] - tot0[:3]) < 1e-14 * scale[:3])
    assert abs(tot1[3] - tot0[3]) > 1e-6 * scale[3]     # (the walls push on IM3)
    # the ghosts (refilled, as a stage's tail does) mirror the first active layers, IM3 negated
    lib().or_fill_gas_ghosts(o.h, C.byref(euler_stage(0.0)))
    g = o.u0()
    np.testing.assert_array_equal(g[0, 0, 1], g[0, 0, 2])
    np.testing.assert_array_equal(g[0, 3, 1], -g[0, 3, 2])
    np.testing.assert_array_equal(g[1, 1, -2], g[1, 1, -3])


# ---- the damped dusty sound wave -----------------------------------------------------------

def dusty_wave_omega(k, cs, eps, ts):
    """The right-travelling damped root of omega^2 (1 + eps/(1 - i omega ts)) = k^2 cs^2 by
    Newton's method from the dusty sound speed cs/sqrt(1 + eps) (fields ~ exp(i(kx - omega t)))."""
    w = complex(k * cs / np.sqrt(1.0 + eps))
    for _ in range(100):
        den = 1.0 - 1j * w * ts
        s, ds = eps / den, eps * 1j * ts / den ** 2
        f = w * w * (1.0 + s) - (k * cs) ** 2
        fp = 2.0 * w * (1.0 + s) + w * w * ds
        d = f / fp
        w -= d
        if abs(d) < 1e-15 * abs(w):
            break
    return w


def dusty_wave_setup(nx, coupling, eps=1.0, ts=0.1, amp=1e-4, cs=1.0):
    """One species, one particle per cell on a lattice displaced by the eigenmode; returns the
    oracle, omega, and the gas and dust velocity amplitudes at t = 0."""
    k, rho0 = 2 * np.pi, 1.0
    w = dusty_wave_omega(k, cs, eps, ts)
    mesh, blocks = wave_mesh(nx)
    solver = abi.DK_DC1 if coupling == abi.DK_IMEX else abi.DK_LOCAL
    o = Oracle(mesh, blocks, workloads._phys(coupling=coupling, solver=solver), 4 * nx + 16)
    dx = 1.0 / nx
    du = amp * cs
    drho = (k / w) * rho0 * du
    dv = du / (1.0 - 1j * w * ts)
    xi = 1j * dv / w
    X, Y, Z = np.meshgrid((np.arange(nx) + 0.5) * dx, (np.arange(2) + 0.5) * dx,
                          (np.arange(2) + 0.5) * dx, indexing="ij")
    ph = np.exp(1j * k * X.ravel())
    xp = np.stack([(X.ravel() + (xi * ph).real) % 1.0, Y.ravel(), Z.ravel()], 1)
    vp = np.zeros_like(xp)
    vp[:, 0] = (dv * ph).real
    place(o, xp, v=vp, ts=ts, m=eps * rho0 * dx ** 3)
    ng = mesh.ng
    xg = (np.arange(-ng, nx + ng) + 0.5) * dx
    d = rho0 + (drho * np.exp(1j * k * xg)).real
    u = (du * np.exp(1j * k * xg)).real
    u0 = o.u0()
    u0[0] = 0.0
    u0[0, 0] = d[None, None, :]
    u0[0, 1] = (d * u)[None, None, :]
    o.u1()[:] = u0
    return o, w, du, dv


def dusty_wave_modes(o, nx):
    """The complex amplitudes of the gas and dust x velocities at k = 2 pi."""
    k = 2 * np.pi
    ua = active(o.u0(), o.mesh.ng)[0]
    vx = (ua[1] / ua[0]).mean(axis=(0, 1))
    x = (np.arange(nx) + 0.5) / nx
    cg = 2.0 * np.mean(vx * np.exp(-1j * k * x))
    r, n = o.rdata(), o.npart
    cd = 2.0 * np.mean(r[abi.IPVX, :n] * np.exp(-1j * k * r[abi.IPX, :n]))
    return cg, cd


def test_dusty_wave_dispersion_root():
    w = dusty_wave_omega(2 * np.pi, 1.0, 1.0, 0.1)
    lhs = w * w * (1.0 + 1.0 / (1.0 - 1j * w * 0.1))
    assert abs(lhs - (2 * np.pi) ** 2) < 1e-12 * (2 * np.pi) ** 2
    assert w.real > 0 and w.imag < 0
    # no drag coupling (eps = 0): the plain sound wave
    assert abs(dusty_wave_omega(2 * np.pi, 1.0, 0.0, 0.1) - 2 * np.pi) < 1e-12


@pytest.mark.parametrize("coupling", [abi.DK_IMEX, abi.DK_PC2])
def test_dusty_wave_follows_the_dispersion_relation(coupling):
    T = 0.5
    phase = []
    for nx in (32, 64):
        o, w, du, dv = dusty_wave_setup(nx, coupling)
        dt = 0.3 / nx
        nst = int(np.ceil(T / dt))
        src = hydro_src(1.0)
        for _ in range(nst):
            o.step(T / nst, src)
        cg, cd = dusty_wave_modes(o, nx)
        eg, ed = du * np.exp(-1j * w * T), dv * np.exp(-1j * w * T)
        # the damping exp(Im(omega) T) = 0.78 is matched to a fraction of a percent
        assert abs(abs(cg) / abs(eg) - 1.0) < (0.01 if nx == 32 else 0.003)
        assert abs(abs(cd) / abs(ed) - 1.0) < (0.015 if nx == 32 else 0.004)
        phase.append((abs(np.angle(cg / eg)), abs(np.angle(cd / ed))))
    assert phase[1][0] < 0.005 and phase[1][1] < 0.005
    assert phase[0][0] / phase[1][0] > 3.0 and phase[0][1] / phase[1][1] > 3.0


def test_streaming_instability_lina_eigenvalue():
    """The linear system of tests/si_linear.py gives the published linA growth rate (Youdin &
    Johansen 2007, Table 1: t_s = 0.1, eps = 3, K = 30, s = 0.4190 Omega) in the incompressible
    limit, and nearly the same with eta v_K = c_s/2 (the GPU test's case)."""
    import si_linear
    lam, _ = si_linear.growing_mode(30.0, 0.05, 0.1, 3.0)
    assert abs(lam.real - 0.4190) < 5e-4
    lam2, _ = si_linear.growing_mode(30.0, 0.5, 0.1, 3.0)
    assert 0.41 < lam2.real < lam.real
