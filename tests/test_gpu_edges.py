"""Edge cases on the GPU against the oracle, both forms: no particles, one particle, thousands
of particles in one cell (the long-list sort and tiles of many chunks), a cell beyond the FP32
form's list limit (reported, not fatal), and particles exactly on block and mesh faces.

Bars as elsewhere: the plain FP64 form to 1e-11 of the oracle (tests/test_gpu_parity.py), the
FP32 form to 1e-5 of the largest velocity with total momentum conserved to 1e-12."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle, lib as olib
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-edges] " + msg, flush=True)
    sys.stdout.flush()


def base():
    # 16^3 in 2x2x2 blocks of 8^3, tau_s 0.1, gas noise (the workload's gas; its particles are
    # replaced by the case's own)
    w = workloads.c1(n=16, nb=2, ppc=1)
    w.setup.gas_noise, w.dt = 0.3, 2e-3
    return w


def block_of(w, x):
    m = w.mesh
    b = [min(n - 1, max(0, int((x[d] - lo) / (hi - lo) * n)))
         for d, (lo, hi, n) in enumerate(((m.x1min, m.x1max, m.nmbx1),
                                          (m.x2min, m.x2max, m.nmbx2),
                                          (m.x3min, m.x3max, m.nmbx3)))]
    return b[0] + m.nmbx1 * (b[1] + m.nmbx2 * b[2])


def particles(w, xp, vp, ts=0.1, m=2e-4):
    n = xp.shape[0]
    r = np.zeros((abi.DK_NRDATA, n))
    iv = np.zeros((abi.DK_NIDATA, n), dtype=np.int32)
    for d in range(3):
        r[abi.IPX + 2 * d] = xp[:, d]
        r[abi.IPX1 + 2 * d] = xp[:, d]
        r[abi.IPVX + 2 * d] = vp[:, d]
        r[abi.IPVX1 + 2 * d] = vp[:, d]
    r[abi.IPTS] = ts
    r[abi.IPM] = m
    iv[abi.PGID] = [block_of(w, x) for x in xp]
    iv[abi.PTAG] = np.arange(n)
    return r, iv


def pair(w, xp, vp, form, m=2e-4):
    from athenak_amd.lib import DustContext
    n = xp.shape[0]
    cap = max(n, w.npart) + 16
    o = Oracle(w.mesh, w.blocks, w.phys, cap)
    o.setup(w.setup)                      # the gas
    r, iv = particles(w, xp, vp, m=m)
    o.rdata()[:, :n] = r
    o.idata()[:, :n] = iv
    olib().or_set_npart(o.h, n)
    g = DustContext(w.mesh, w.blocks, w.phys, cap, form=form)
    g.load(r, iv, o.u0().copy(), o.u1().copy())
    return o, g


def fp32_check(o, g, w, what):
    g.export()
    n = o.npart
    assert g.npart == n
    ro = o.rdata()[:, :n]
    oo = np.argsort(o.idata()[abi.PTAG, :n], kind="stable")
    rg = g.rdata[:, :n].cpu().numpy()
    og = np.argsort(g.idata[abi.PTAG, :n].cpu().numpy(), kind="stable")
    ro, rg = ro[:, oo], rg[:, og]
    v = [abi.IPVX, abi.IPVY, abi.IPVZ]
    err = np.abs(ro[v] - rg[v]).max() / np.abs(ro[v]).max()
    say("%s: fp32 max rel velocity error %.3e" % (what, err))
    assert err < 1e-5, (what, err)


def run_both(w, xp, vp, steps, what, m=2e-4):
    for form in (abi.DK_FORM_FP64, abi.DK_FORM_FP32):
        o, g = pair(w, xp, vp, form, m=m)
        d0 = g.diag()
        for _ in range(steps):
            o.step(w.dt, w.src)
            g.step(w.dt, w.src)
        g.check_device_error()
        d1 = g.diag()
        drift = np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8]
        if xp.shape[0] > 0 and form == abi.DK_FORM_FP64:
            compare(o, g, w, what=what)
            assert drift < 1e-13, drift
        elif xp.shape[0] > 0:
            fp32_check(o, g, w, what)
            assert drift < 1e-12, drift
        else:
            ng = w.mesh.ng
            np.testing.assert_allclose(g.u0.cpu().numpy()[:, :, ng:-ng, ng:-ng, ng:-ng],
                                       o.u0()[:, :, ng:-ng, ng:-ng, ng:-ng], rtol=0,
                                       atol=1e-14)
        say("%s: form %d, %d steps, drift %.2e" % (what, form, steps, drift))
        g.close()


def test_no_particles():
    w = base()
    run_both(w, np.zeros((0, 3)), np.zeros((0, 3)), 4, "empty")


def test_one_particle():
    w = base()
    run_both(w, np.array([[0.31, 0.52, 0.77]]), np.array([[0.4, -0.3, 0.2]]), 10, "one")


def test_many_particles_in_one_cell():
    # 6000 particles in one cell: a sorted list far above the short-list length, and a tile of
    # eight chunks in the FP32 form
    w = base()
    rng = np.random.default_rng(5)
    n = 6000
    xp = (np.array([5, 9, 2]) + rng.random((n, 3))) / 16.0
    vp = rng.normal(scale=0.2, size=(n, 3))
    # (the cell's dust mass about its gas mass)
    run_both(w, xp, vp, 4, "one-cell", m=(1.0 / 16) ** 3 / n)


def test_fp32_list_limit_is_reported_not_fatal():
    # more particles in one cell than the FP32 form's long-list sort holds: the bind reports
    # DK_ECAPACITY through the device error word and every kernel exits normally; the plain
    # FP64 form has no such limit and stays the oracle's
    from athenak_amd.lib import DustError
    w = base()
    rng = np.random.default_rng(6)
    n = 17000
    xp = (np.array([3, 3, 3]) + rng.random((n, 3))) / 16.0
    vp = rng.normal(scale=0.1, size=(n, 3))
    m = (1.0 / 16) ** 3 / n
    o, g = pair(w, xp, vp, abi.DK_FORM_FP32, m=m)
    with pytest.raises(DustError, match="fixed maximum"):
        g.check_device_error()
    g.close()
    o, g = pair(w, xp, vp, abi.DK_FORM_FP64, m=m)
    for _ in range(2):
        o.step(w.dt, w.src)
        g.step(w.dt, w.src)
    g.check_device_error()
    compare(o, g, w, what="17000 in one cell")
    g.close()


def test_particles_on_block_and_mesh_faces():
    w = base()
    h = 0.5
    eps = np.nextafter(1.0, 0.0)
    xp = np.array([
        [h, h, h],                              # the corner shared by eight blocks
        [np.nextafter(h, 0.0), 0.25, 0.25],     # one ulp below a block face
        [0.0, 0.0, 0.0],                        # the mesh's lower corner
        [eps, eps, eps],                        # one ulp below the mesh's upper faces
        [0.25, h, eps],
        [eps, 0.0, h],
    ])
    vp = np.array([[0.5, -0.5, 0.5], [0.8, 0.0, 0.0], [-0.6, -0.6, -0.6],
                   [0.6, 0.6, 0.6], [0.0, -0.4, 0.7], [0.7, -0.7, 0.1]])
    run_both(w, xp, vp, 12, "faces")
