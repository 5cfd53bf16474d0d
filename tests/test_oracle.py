"""Pins the CPU FP64 oracle against closed forms, brute force and invariants (no GPU).

Each check is chosen so that a plausible slip in the restatement (a weight, a sign, an index, a
dropped term) fails it: TSC weights against the B-spline closed form, the dc1 operator against
a dense matrix built from that closed form, the imex2+ stage sequence against the exact damping
solution and its convergence order, conservation to round-off, the NSH equilibrium held, and the
ghost fold against an unpartitioned mesh."""

import math

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle
from tests.helpers import active, bspline2, dense_gather_matrix, lattice_positions

GAMMA = abi.IMEX2P_GAMMA


def one_block(n=8, L=1.0, ng=2, coupling=abi.DK_IMEX, solver=abi.DK_DC1, cap=4096, nb=1,
              **phys_kw):
    nx = n // nb
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nb, nb, nb), ((0.0, L),) * 3, ng=ng)
    phys = workloads._phys(coupling=coupling, solver=solver, **phys_kw)
    return Oracle(mesh, blocks, phys, cap)


def place(o, xp, v=None, ts=0.1, m=1e-3, uniform_gas=(1.0, 0.0, 0.0, 0.0)):
    """Put particles at xp (npart x 3) into block 0.. by position and set a uniform gas."""
    r, ii = o.rdata(), o.idata()
    npart = xp.shape[0]
    from oracle.pyoracle import lib
    lib().or_set_npart(o.h, npart)
    r[:, :] = 0.0
    ii[:, :] = 0
    mesh = o.mesh
    for p in range(npart):
        x = xp[p]
        bx = min(mesh.nmbx1 - 1, int((x[0] - mesh.x1min) / (mesh.x1max - mesh.x1min) * mesh.nmbx1))
        by = min(mesh.nmbx2 - 1, int((x[1] - mesh.x2min) / (mesh.x2max - mesh.x2min) * mesh.nmbx2))
        bz = min(mesh.nmbx3 - 1, int((x[2] - mesh.x3min) / (mesh.x3max - mesh.x3min) * mesh.nmbx3))
        ii[abi.PGID, p] = bx + mesh.nmbx1 * (by + mesh.nmbx2 * bz)
        ii[abi.PTAG, p] = p
    r[abi.IPX, :npart], r[abi.IPY, :npart], r[abi.IPZ, :npart] = xp[:, 0], xp[:, 1], xp[:, 2]
    r[abi.IPX1, :npart], r[abi.IPY1, :npart], r[abi.IPZ1, :npart] = xp[:, 0], xp[:, 1], xp[:, 2]
    if v is not None:
        for c, col in enumerate((abi.IPVX, abi.IPVY, abi.IPVZ)):
            r[col, :npart] = v[:, c]
            r[col + 6, :npart] = v[:, c]
    r[abi.IPTS, :npart] = ts
    r[abi.IPM, :npart] = m
    u0 = o.u0()
    for c in range(4):
        u0[:, c] = uniform_gas[c]
    o.u1()[:] = u0


# ------------------------------------------------------------------------------------------


def test_tsc_deposit_matches_bspline_closed_form():
    rng = np.random.default_rng(3)
    n, dt = 8, 1e-3
    o = one_block(n)
    xp = rng.random((5, 3))
    ts = np.array([0.05, 0.1, 0.2, 0.3, 1.0])
    place(o, xp, v=rng.normal(size=(5, 3)))
    o.rdata()[abi.IPTS, :5] = ts
    st = abi.imex2p_stages(0.0, dt)[0]
    o.task("deposit_drag", st)
    o.task("fold_deposit_qp", st)
    q = active(o.field(0), 2)[0]
    a = GAMMA * dt
    c = a / (ts + a)
    V = (1.0 / n) ** 3
    G = dense_gather_matrix(xp, n)
    mu = 1e-3 * c / V
    v = o.rdata()[[abi.IPVX, abi.IPVY, abi.IPVZ], :5].T
    np.testing.assert_allclose(q[0].ravel(), G.T @ mu, rtol=1e-12, atol=1e-12)
    for d in range(3):
        np.testing.assert_allclose(q[1 + d].ravel(), G.T @ (mu * v[:, d]), rtol=1e-12,
                                   atol=1e-12)
    # sum Q V = sum m c (partition of unity)
    assert abs(q[0].sum() * V - (1e-3 * c).sum()) < 1e-15


def test_cic_and_ngp_partition_of_unity():
    rng = np.random.default_rng(4)
    for scheme in (abi.DK_CIC, abi.DK_NGP):
        o = one_block(8, deposit=scheme)
        xp = rng.random((7, 3))
        place(o, xp)
        st = abi.imex2p_stages(0.0, 1e-3)[0]
        o.task("deposit_drag", st)
        o.task("fold_deposit_qp", st)
        q = active(o.field(0), 2)[0, 0]
        a = GAMMA * 1e-3
        assert abs(q.sum() / 8 ** 3 - 7 * 1e-3 * a / (0.1 + a)) < 1e-15


def test_operator_matches_dense_matrix_and_is_spd():
    rng = np.random.default_rng(5)
    n = 4
    o = one_block(n)
    npart = 12
    xp = rng.random((npart, 3))
    place(o, xp, ts=0.1)
    rng_rho = 1.0 + 0.5 * rng.random((n, n, n))
    o.u0()[0, 0, 2:-2, 2:-2, 2:-2] = rng_rho
    o.rdata()[abi.IPTS, :npart] = 0.05 + rng.random(npart)
    ts = o.rdata()[abi.IPTS, :npart].copy()
    a = 0.7
    c = a / (ts + a)
    V = (1.0 / n) ** 3
    G = dense_gather_matrix(xp, n)
    A = np.diag(rng_rho.ravel()) + G.T @ np.diag(1e-3 * c / V) @ G
    assert np.all(np.linalg.eigvalsh(A) > 0)
    x = np.zeros_like(o.field(1))
    xa = rng.normal(size=(3, n, n, n))
    x[0, :, 2:-2, 2:-2, 2:-2] = xa
    y = o.apply_operator(a, x)
    ya = active(y, 2)[0]
    for d in range(3):
        np.testing.assert_allclose(ya[d].ravel(), A @ xa[d].ravel(), rtol=1e-12, atol=1e-14)


def test_dc1_sweep_is_the_defect_correction_of_the_dense_system():
    rng = np.random.default_rng(6)
    n, dt = 4, 0.05
    errs = {}
    for solver in (abi.DK_LOCAL, abi.DK_DC1, abi.DK_DC2):
        o = one_block(n, solver=solver)
        npart = 40
        xp = rng.random((npart, 3)) if solver == abi.DK_LOCAL else xp   # noqa: F821
        v = np.random.default_rng(7).normal(size=(npart, 3))
        place(o, xp, v=v, ts=0.02, m=2e-2)
        u = np.random.default_rng(8).normal(size=(3, n, n, n))
        for d in range(3):
            o.u0()[0, 1 + d, 2:-2, 2:-2, 2:-2] = u[d]
        st = abi.imex2p_stages(0.0, dt)[0]
        o.task("deposit_drag", st)
        o.task("fold_deposit_qp", st)
        o.task("gas_implicit_solve", st)
        xs = active(o.field(1), 2)[0].reshape(3, -1)
        a = GAMMA * dt
        c = a / (0.02 + a)
        V = (1.0 / n) ** 3
        G = dense_gather_matrix(xp, n)
        W = np.diag(2e-2 * c / V * np.ones(npart))
        A = np.eye(n ** 3) + G.T @ W @ G
        Q = (G.T @ (2e-2 * c / V * np.ones(npart)))
        D = 1.0 + Q
        b = u.reshape(3, -1) + (G.T @ W @ v).T
        x0 = b / D
        xstar = np.linalg.solve(A, b.T).T
        if solver == abi.DK_LOCAL:
            np.testing.assert_allclose(xs, x0, rtol=1e-12, atol=1e-13)
        elif solver == abi.DK_DC1:
            x1 = x0 + (b - (A @ x0.T).T) / D
            np.testing.assert_allclose(xs, x1, rtol=1e-12, atol=1e-13)
        else:
            x1 = x0 + (b - (A @ x0.T).T) / D
            x2 = x1 + (b - (A @ x1.T).T) / D
            np.testing.assert_allclose(xs, x2, rtol=1e-12, atol=1e-13)
        errs[solver] = np.abs(xs - xstar).max()
    assert errs[abi.DK_DC1] < 0.5 * errs[abi.DK_LOCAL]
    assert errs[abi.DK_DC2] < 0.5 * errs[abi.DK_DC1]


def _momentum_drift(w, steps, threads=1):
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=threads)
    o.setup(w.setup)
    d0 = o.diag()
    for _ in range(steps):
        o.step(w.dt, w.src)
    d1 = o.diag()
    p0 = d0[0:3] + d0[3:6]
    p1 = d1[0:3] + d1[3:6]
    return np.abs(p1 - p0).max() / d0[8], o


@pytest.mark.parametrize("coupling,solver", [(abi.DK_IMEX, abi.DK_DC1), (abi.DK_IMEX, abi.DK_LOCAL),
                                             (abi.DK_IMEX, abi.DK_DC2), (abi.DK_PC2, abi.DK_LOCAL)])
def test_momentum_conserved_to_roundoff(coupling, solver):
    w = workloads.c1(n=8, nb=2, ppc=2, coupling=coupling, solver=solver)
    # larger velocities so that particles cross blocks and the periodic boundary
    w.setup.dust_noise = 2.0
    w.setup.gas_noise = 0.5
    w.dt = 2e-3
    drift, o = _momentum_drift(w, 20)
    assert drift < 1e-13, drift
    # every particle inside its block
    r, ii = o.rdata(), o.idata()
    n = o.npart
    for p in range(0, n, 7):
        b = o.blocks[ii[abi.PGID, p]]
        assert b.x1min <= r[abi.IPX, p] < b.x1max
        assert b.x2min <= r[abi.IPY, p] < b.x2max
        assert b.x3min <= r[abi.IPZ, p] < b.x3max


def _damping_run(dt, T, coupling, eps=1.0, ts=0.1, v0=(1.0, -0.5, 0.25)):
    n = 4
    o = one_block(n, coupling=coupling,
                  solver=abi.DK_DC1 if coupling == abi.DK_IMEX else abi.DK_LOCAL)
    xp = lattice_positions(n, 1)
    V = (1.0 / n) ** 3
    place(o, xp, v=np.tile(v0, (xp.shape[0], 1)), ts=ts, m=eps * 1.0 * V)
    o.time = 0.0
    nsteps = int(round(T / dt))
    for _ in range(nsteps):
        o.step(dt)
    r = o.rdata()
    vd = np.array([r[c, :xp.shape[0]].mean() for c in (abi.IPVX, abi.IPVY, abi.IPVZ)])
    ug = np.array([active(o.u0(), 2)[0, 1 + c].mean() for c in range(3)])
    return vd, ug


def _damping_exact(t, eps=1.0, ts=0.1, v0=(1.0, -0.5, 0.25)):
    v0 = np.array(v0)
    vcm = eps * v0 / (1 + eps)             # u0 = 0, rho = 1
    w = v0 * np.exp(-(1 + eps) * t / ts)
    return vcm + w / (1 + eps), vcm - eps * w / (1 + eps)


@pytest.mark.parametrize("coupling", [abi.DK_IMEX, abi.DK_PC2])
def test_damping_ode_second_order(coupling):
    """dust_damping (src/pgen/tests/dust_damping.cpp:404-447): uniform lattice, the mean
    velocities follow du/dt = eps (v - u)/ts, dv/dt = -(v - u)/ts; the error against the exact
    solution drops by > 2.5 when dt halves (tst/test_suite/dust/check_cloud_imex3d.py:114-115)."""
    T = 0.2
    e = []
    for dt in (0.02, 0.01, 0.005):
        vd, ug = _damping_run(dt, T, coupling)
        ve, ue = _damping_exact(T)
        e.append(max(np.abs(vd - ve).max(), np.abs(ug - ue).max()))
    assert e[0] / e[1] > 2.5 and e[1] / e[2] > 2.5, e
    assert e[2] < 2e-3


def test_nsh_equilibrium_held():
    """dust_nsh (src/pgen/tests/dust_nsh.cpp:7-26): a lattice of dust at the NSH drift in gas at
    its NSH velocity stays there to round-off under the imex2+ coupling in a shearing box."""
    n = 8
    L = 0.1
    mesh, blocks = abi.uniform_mesh((n // 2,) * 3, (2, 2, 2), ((-L, L),) * 3, ng=2,
                                    bc_x1=abi.DK_X1_SHEAR)
    phys = workloads._phys(sbox=True)
    o = Oracle(mesh, blocks, phys, n ** 3 + 8)
    eps, ts, ax = 3.0, 0.1, 0.1
    ugx, ugy = workloads.nsh_gas(1.0, 1.5, ax, (eps,), (ts,))
    tq = 1.0 * ts
    b = 0.5
    vx = (ugx + 2.0 * tq * ugy) / (1.0 + 2.0 * b * tq * tq)
    vy = ugy - b * tq * vx
    xp = lattice_positions(n, 1, L=2 * L, x0=-L)
    V = (2 * L / n) ** 3
    place(o, xp, v=np.tile((vx, vy, 0.0), (xp.shape[0], 1)), ts=ts, m=eps * V)
    u0 = o.u0()
    u0[:, 1], u0[:, 2] = ugx, ugy
    o.u1()[:] = u0
    src = abi.dk_gas_src(shearing_box=1, omega0=1.0, qshear=1.5, const_accel=ax,
                         const_accel_dir=1)
    for _ in range(30):
        o.step(2e-3, src)
    r = o.rdata()
    npart = xp.shape[0]
    assert np.abs(r[abi.IPVX, :npart] - vx).max() < 1e-13
    assert np.abs(r[abi.IPVY, :npart] - vy).max() < 1e-13
    ua = active(o.u0(), 2)
    assert np.abs(ua[:, 1] - ugx).max() < 1e-13
    assert np.abs(ua[:, 2] - ugy).max() < 1e-13


def test_fold_of_blocks_equals_unpartitioned_mesh():
    rng = np.random.default_rng(11)
    n = 8
    xp = rng.random((200, 3))
    v = rng.normal(size=(200, 3))
    res = []
    for nb in (1, 2):
        o = one_block(n, nb=nb)
        place(o, xp, v=v)
        st = abi.imex2p_stages(0.0, 1e-3)[0]
        o.task("deposit_drag", st)
        o.task("fold_deposit_qp", st)
        q = active(o.field(0), 2)
        nx = n // nb
        full = np.zeros((5, n, n, n))
        for m in range(nb ** 3):
            bx, by, bz = m % nb, (m // nb) % nb, m // (nb * nb)
            full[:, bz * nx:(bz + 1) * nx, by * nx:(by + 1) * nx, bx * nx:(bx + 1) * nx] = q[m]
        res.append(full)
    np.testing.assert_allclose(res[0], res[1], rtol=1e-13, atol=1e-13)
