"""Pins the oracle's same-level refinable clouds (pc2, CloudPredictorT / CloudAdvanceT) against
what the mathematics fixes, no GPU.

- the cloud is the unit hat of half-width one root cell: its weight over a cell is the integral
  of the hat over that cell, which is the quadratic B-spline (TSC) at the cell centre
  (textbook closed form, tests.helpers.bspline2); its first moment about the particle is
  checked by quadrature of s * hat(s);
- the linear gather integrates the hat against the cells' linear reconstructions (central
  differences): it returns a linear field's value at the particle exactly;
- pc2 with clouds conserves total momentum to round-off, and with a uniform gas and no
  back-reaction the clouds' gathers return the gas velocity exactly, so the particles move as
  the point-particle pc2 moves them (pinned separately against the damping ODE).
A dropped moment, a wrong sign of a slope, a transposed axis or a wrong cell range fails one of
these."""

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle
from tests.helpers import bspline2


def cloud_oracle(n=8, nb=1, L=1.0, cap=64, back_reaction=1, clouds=1):
    nx = n // nb
    mesh, blocks = abi.uniform_mesh((nx, nx, nx), (nb, nb, nb), ((0.0, L),) * 3, ng=2)
    phys = workloads._phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=clouds,
                           back_reaction=back_reaction)
    return Oracle(mesh, blocks, phys, cap)


def unpack(cell, n1, n2):
    i = cell % n1
    j = (cell // n1) % n2
    k = cell // (n1 * n2)
    return k, j, i


def test_cloud_weights_are_the_tsc_bspline():
    rng = np.random.default_rng(11)
    n, ng = 8, 2
    o = cloud_oracle(n)
    dx = 1.0 / n
    n1 = n + 2 * ng
    for _ in range(40):
        x = rng.random(3)
        cell, w, _ = o.cloud_cells(0, x)
        assert abs(w.sum() - 1.0) < 1e-14
        k, j, i = unpack(cell, n1, n1)
        # cell centres of the padded indices
        xc = ((i - ng) + 0.5) * dx, ((j - ng) + 0.5) * dx, ((k - ng) + 0.5) * dx
        want = (bspline2((x[0] - xc[0]) / dx) * bspline2((x[1] - xc[1]) / dx) *
                bspline2((x[2] - xc[2]) / dx))
        np.testing.assert_allclose(w, want, rtol=0, atol=2e-15)
        # every cell with a nonzero TSC weight is visited
        assert (want > 1e-12).sum() == (w > 1e-12).sum()


def test_cloud_first_moments_by_quadrature():
    rng = np.random.default_rng(12)
    n, ng = 8, 2
    o = cloud_oracle(n)
    dx = 1.0 / n
    n1 = n + 2 * ng
    # 1-D integrals over a cell [a, a + dx] of hat((s - x)/dx)/dx and of (s - x) times it:
    # Gauss-Legendre on the pieces between the kinks x - dx, x, x + dx (the integrand is a
    # polynomial of degree <= 2 on each piece: exact to round-off)
    gx, gw = np.polynomial.legendre.leggauss(4)

    def integral(x, a, moment):
        cuts = sorted({a, a + dx} | {c for c in (x - dx, x, x + dx) if a < c < a + dx})
        tot = 0.0
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            s = 0.5 * (hi - lo) * gx + 0.5 * (hi + lo)
            h = np.maximum(0.0, 1.0 - np.abs(s - x) / dx) / dx
            f = (s - x) * h if moment else h
            tot += 0.5 * (hi - lo) * np.sum(gw * f)
        return tot

    def hat_moment_1d(x, a):
        return integral(x, a, True)

    def hat_weight_1d(x, a):
        return integral(x, a, False)

    for _ in range(6):
        x = rng.random(3)
        cell, w, m1 = o.cloud_cells(0, x)
        k, j, i = unpack(cell, n1, n1)
        lo = ((i - ng) * dx, (j - ng) * dx, (k - ng) * dx)
        for q in range(len(cell)):
            wx = hat_weight_1d(x[0], lo[0][q])
            wy = hat_weight_1d(x[1], lo[1][q])
            wz = hat_weight_1d(x[2], lo[2][q])
            want = (hat_moment_1d(x[0], lo[0][q]) * wy * wz,
                    wx * hat_moment_1d(x[1], lo[1][q]) * wz,
                    wx * wy * hat_moment_1d(x[2], lo[2][q]))
            np.testing.assert_allclose(m1[q], want, rtol=0, atol=1e-15 * dx)


def test_linear_gather_is_exact_for_a_linear_field():
    rng = np.random.default_rng(13)
    n, ng = 8, 2
    o = cloud_oracle(n)
    dx = 1.0 / n
    n1 = n + 2 * ng
    c = np.arange(n1)
    xc = (c - ng + 0.5) * dx
    Z, Y, X = np.meshgrid(xc, xc, xc, indexing="ij")
    a = np.array([0.3, -1.2, 2.0])
    B = rng.normal(size=(3, 3))
    vel = np.zeros((1, 3, n1, n1, n1))
    for comp in range(3):
        vel[0, comp] = a[comp] + B[comp, 0] * X + B[comp, 1] * Y + B[comp, 2] * Z
    for _ in range(30):
        x = rng.random(3)
        got = o.cloud_gather(vel, 0, x)
        np.testing.assert_allclose(got, a + B @ x, rtol=0, atol=1e-13)


def test_gather_of_a_quadratic_field_differs_from_the_point_value():
    # the linear reconstruction is exact only for linear fields: a quadratic field shows the
    # second moment of the hat (dx^2/6 times the curvature), so the test above is not vacuous
    n, ng = 8, 2
    o = cloud_oracle(n)
    dx = 1.0 / n
    n1 = n + 2 * ng
    xc = (np.arange(n1) - ng + 0.5) * dx
    Z, Y, X = np.meshgrid(xc, xc, xc, indexing="ij")
    vel = np.zeros((1, 3, n1, n1, n1))
    vel[0, 0] = X * X
    x = np.array([0.45, 0.5, 0.55])
    got = o.cloud_gather(vel, 0, x)
    # the excess is the hat's second moment dx^2/6 minus what the cells' linear pieces carry
    assert 0.05 * dx * dx < got[0] - x[0] ** 2 < dx * dx / 6.0


def test_cloud_pc2_conserves_momentum_across_blocks():
    w = workloads.c4u(n=16, nbx=2, ppc=2)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16)
    o.setup(w.setup)
    d0 = o.diag()
    for _ in range(10):
        o.step(w.dt, w.src)
    d1 = o.diag()
    drift = np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8]
    assert drift < 1e-13, drift
    # the particles moved: the test exercises the drift and the migration
    assert np.abs(o.rdata()[abi.IPX, :o.npart] - o.rdata()[abi.IPX1, :o.npart]).max() > 0


def test_clouds_in_a_uniform_gas_move_as_point_particles():
    w = workloads.c4u(n=16, nbx=2, ppc=2, clump_frac=0.0)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 1.0, 0.0, 2e-3
    w.setup.gas_vx, w.setup.gas_vy, w.setup.gas_vz = 0.3, -0.2, 0.1
    outs = []
    for clouds in (1, 0):
        phys = workloads._phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=clouds,
                               back_reaction=0)
        o = Oracle(w.mesh, w.blocks, phys, w.npart + 16)
        o.setup(w.setup)
        for _ in range(8):
            o.step(w.dt, w.src)
        n = o.npart
        outs.append(o.rdata()[:, :n].copy())
    a, b = outs
    for c in (abi.IPVX, abi.IPVY, abi.IPVZ, abi.IPX, abi.IPY, abi.IPZ):
        np.testing.assert_allclose(a[c], b[c], rtol=0, atol=1e-14 * max(1.0, np.abs(b[c]).max()))


def test_clouds_are_refused_where_not_built():
    from oracle.pyoracle import lib
    import ctypes as C
    nx = 8
    for bc_x1, bc_x3 in ((abi.DK_X1_SHEAR, abi.DK_X3_PERIODIC),
                         (abi.DK_X1_PERIODIC, abi.DK_X3_OUTFLOW)):
        mesh, blocks = abi.uniform_mesh((nx,) * 3, (1, 1, 1), ((0.0, 1.0),) * 3, ng=2,
                                        bc_x1=bc_x1, bc_x3=bc_x3)
        phys = workloads._phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=1)
        h = lib().or_create(C.byref(mesh), blocks, C.byref(phys), 16)
        assert not h
    # imex with clouds: not this path
    mesh, blocks = abi.uniform_mesh((nx,) * 3, (1, 1, 1), ((0.0, 1.0),) * 3, ng=2)
    phys = workloads._phys(clouds=1)
    assert not lib().or_create(C.byref(mesh), blocks, C.byref(phys), 16)
