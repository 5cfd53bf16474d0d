"""Pins the oracle's shear-periodic x1 pieces (no GPU).

u* ghost remap (ShearingBoxCC, src/shearing_box/shearing_box_cc.cpp:68-260): with a whole-cell
offset it is an exact shifted copy whose direction matches the particle crossing
(bvals_part.cpp:274-283: inward crossing y -> y - ysh); with a fractional offset it reproduces
the cell averages of a smooth field to second order.  Deposit fold (bvals_dep_general.cpp): it
conserves the deposited total, and with a whole-cell offset it moves every ghost deposit to the
shifted cell of the opposite face."""

import numpy as np

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle


def sbox(n=8, nb=2, L=0.1, bc_x3=abi.DK_X3_PERIODIC):
    mesh, blocks = abi.uniform_mesh((n // nb,) * 3, (nb, nb, nb), ((-L, L),) * 3, ng=2,
                                    bc_x1=abi.DK_X1_SHEAR, bc_x3=bc_x3)
    phys = workloads._phys(sbox=True)
    return Oracle(mesh, blocks, phys, 16), mesh


def global_field(o, mesh, f):
    """Sets every block's active ustar cells to f(gx, gy, gz) (global cell indices)."""
    us = o.field(1)
    nx = mesh.nx1
    ng = mesh.ng
    for m in range(mesh.nmb):
        b = o.blocks[m]
        gi = b.lx1 * nx + np.arange(nx)
        gj = b.lx2 * nx + np.arange(nx)
        gk = b.lx3 * nx + np.arange(nx)
        K, J, I = np.meshgrid(gk, gj, gi, indexing="ij")
        for c in range(3):
            us[m, c, ng:-ng, ng:-ng, ng:-ng] = f(I, J, K, c)


def test_whole_cell_offset_is_a_shifted_copy_in_the_crossing_direction():
    o, mesh = sbox()
    n, L = 8, 0.1
    dy = 2 * L / n
    lx = 2 * L
    # yshear = q Omega Lx t = 3 dy exactly: t = 3 dy/(1.5 lx)
    t = 3 * dy / (1.5 * lx)
    global_field(o, mesh, lambda I, J, K, c: 1000.0 * c + 100.0 * I + J + 0.01 * K)
    st = abi.imex2p_stages(t, 1e-3)[0]
    o.task("fill_ustar", st)
    us = o.field(1)
    ng = mesh.ng
    for m in range(mesh.nmb):
        b = o.blocks[m]
        if b.lx1 != 0:
            continue
        for j in range(ng, ng + 4):
            GJ = b.lx2 * 4 + (j - ng)
            for g in range(ng):
                # inner ghost i = g holds outer-face cell i + nx (global gi = 8 - ng + g) at
                # global y index GJ - 3 (the point (x1min - d, y) is (x1max - d, y - ysh))
                gi = n - ng + g
                want = 100.0 * gi + ((GJ - 3) % n) + 0.01 * (b.lx3 * 4 + 1 - ng + ng)
                got = us[m, 0, ng + 1, j, g]
                assert abs(got - want) < 1e-9, (m, j, g, got, want)


def test_fractional_offset_is_second_order_on_a_smooth_field():
    errs = []
    for n in (16, 32):
        o, mesh = sbox(n=n, nb=2)
        L = 0.1
        dy = 2 * L / n
        lx = 2 * L
        t = 0.37 * dy / (1.5 * lx) + 2 * dy / (1.5 * lx)   # 2.37 cells
        ysh = 1.5 * lx * t
        k = 2 * np.pi / (2 * L)

        def cellavg(J, shift):
            y0 = -L + J * dy - shift
            return (np.cos(k * y0) - np.cos(k * (y0 + dy))) / (k * dy)

        global_field(o, mesh, lambda I, J, K, c: cellavg(J, 0.0))
        o.task("fill_ustar", abi.imex2p_stages(t, 1e-3)[0])
        us = o.field(1)
        ng = mesh.ng
        e = 0.0
        nxb = n // 2
        for m in range(mesh.nmb):
            b = o.blocks[m]
            if b.lx1 != 0:
                continue
            J = b.lx2 * nxb + np.arange(nxb)
            got = us[m, 0, ng + 1, ng:ng + nxb, 0]
            e = max(e, np.abs(got - cellavg(J, ysh)).max())
        errs.append(e)
    assert errs[1] < errs[0] / 3.0, errs
    assert errs[1] < 5e-3


def test_deposit_fold_conserves_and_moves_whole_cell_offsets():
    o, mesh = sbox(n=8, nb=2)
    rng = np.random.default_rng(2)
    q = o.field(0)
    q[:] = rng.random(q.shape)
    ng = mesh.ng
    # keep only the first ghost layer and the active cells (as a TSC deposit does)
    q[:, :, :ng - 1] = 0.0
    q[:, :, -(ng - 1):] = 0.0
    q[:, :, :, :ng - 1] = 0.0
    q[:, :, :, -(ng - 1):] = 0.0
    q[:, :, :, :, :ng - 1] = 0.0
    q[:, :, :, :, -(ng - 1):] = 0.0
    total = q.sum(axis=(0, 2, 3, 4))
    n, L = 8, 0.1
    dy = 2 * L / n
    t = 2.43 * dy / (1.5 * 2 * L)
    o.task("fold_deposit_qp", abi.imex2p_stages(t, 1e-3)[0])
    act = q[:, :, ng:-ng, ng:-ng, ng:-ng].sum(axis=(0, 2, 3, 4))
    np.testing.assert_allclose(act, total, rtol=1e-13)
