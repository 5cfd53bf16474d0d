"""The workloads at their full sizes on the GPU, through properties that do not depend on size.

- C2 (256^3, 67,108,864 particles) and C3 (256^3, 117,440,512 particles, 7 continuous bins, x3
  outflow): the FP32 form against the plain FP64 form from the same state, velocities within
  1e-5 of the largest (the FP32 form's bar), the same particles parked at the x3 faces;
- C5 (128^3, 16,777,216 particles, half in clumps): two FP32 runs bit-identical in every
  exported array and the gas, total momentum conserved to 1e-12, and within 1e-5 of FP64;
- C4 (same-level clouds, 64^3 in 16^3 blocks, 1,048,576 particles): the FP64 cloud form against
  the oracle to 1e-11 and conserving momentum.
A few steps each (the FP64 form takes ~1 s per C2 step)."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-full] " + msg, flush=True)
    sys.stdout.flush()


def run_form(w, form, steps):
    import torch
    from athenak_amd.lib import DustContext
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=form)
    g.setup(w.setup)
    d0 = g.diag()
    for _ in range(steps):
        g.step(w.dt, w.src)
    g.check_device_error()
    d1 = g.diag()
    g.export()
    n = g.npart
    order = torch.argsort(g.idata[abi.PTAG, :n])
    r = g.rdata[:, :n][:, order].cpu().numpy()
    tag = g.idata[abi.PTAG, :n][order].cpu().numpy()
    u0 = g.u0.cpu().numpy()
    g.close()
    del g
    torch.cuda.empty_cache()
    return r, tag, u0, d0, d1


def velocity_error(ra, rb):
    v = [abi.IPVX, abi.IPVY, abi.IPVZ]
    return np.abs(ra[v] - rb[v]).max() / np.abs(rb[v]).max()


@pytest.mark.parametrize("name,steps", [("c2", 3), ("c3", 3)])
def test_full_size_fp32_against_fp64(name, steps):
    w = workloads.WORKLOADS[name]()
    say("%s: %d particles, fp64 %d steps" % (name, w.npart, steps))
    r64, t64, _, _, _ = run_form(w, abi.DK_FORM_FP64, steps)
    say("%s: fp32 %d steps" % (name, steps))
    r32, t32, _, _, _ = run_form(w, abi.DK_FORM_FP32, steps)
    np.testing.assert_array_equal(t64, t32)
    err = velocity_error(r32, r64)
    parked64 = int((r64[abi.IPM] == 0.0).sum())
    parked32 = int((r32[abi.IPM] == 0.0).sum())
    say("%s: max rel velocity difference %.3e, parked %d / %d" % (name, err, parked32, parked64))
    assert err < 1e-5, err
    assert parked32 == parked64


def test_full_size_c5_reruns_identical_conserving_and_near_fp64():
    w = workloads.c5()
    steps = 4
    say("c5: %d particles, two fp32 runs of %d steps" % (w.npart, steps))
    a = run_form(w, abi.DK_FORM_FP32, steps)
    b = run_form(w, abi.DK_FORM_FP32, steps)
    assert a[0].tobytes() == b[0].tobytes()
    assert a[2].tobytes() == b[2].tobytes()
    d0, d1 = a[3], a[4]
    drift = np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8]
    say("c5: reruns identical, momentum drift %.3e" % drift)
    assert drift < 1e-12, drift
    r64 = run_form(w, abi.DK_FORM_FP64, steps)[0]
    err = velocity_error(a[0], r64)
    say("c5: max rel velocity difference to fp64 %.3e" % err)
    assert err < 1e-5, err


def test_full_size_c4_clouds_against_the_oracle():
    from oracle.pyoracle import Oracle
    w = workloads.c4u()
    steps = 3
    say("c4: %d particles, oracle %d steps" % (w.npart, steps))
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=16)
    o.setup(w.setup)
    for _ in range(steps):
        o.step(w.dt, w.src)
    n = o.npart
    order = np.argsort(o.idata()[abi.PTAG, :n], kind="stable")
    ro = o.rdata()[:, :n][:, order]
    say("c4: gpu %d steps" % steps)
    rg, tg, ug, d0, d1 = run_form(w, abi.DK_FORM_FP64, steps)
    np.testing.assert_array_equal(o.idata()[abi.PTAG, :n][order], tg)
    for cols in ((abi.IPX, abi.IPY, abi.IPZ), (abi.IPVX, abi.IPVY, abi.IPVZ)):
        scale = np.abs(ro[list(cols)]).max()
        err = np.abs(ro[list(cols)] - rg[list(cols)]).max() / scale
        assert err < 1e-11, (cols, err)
    drift = np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8]
    say("c4: matches the oracle, momentum drift %.3e" % drift)
    assert drift < 1e-13, drift
