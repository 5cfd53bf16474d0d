"""The uneven-block case on a statically refined mesh (pc2 with refinable clouds of their block's
level, lognormal particle counts per block) in the plain FP64 form of the HIP library, against
the oracle on the same seeded inputs.

Tolerances as for the other FP64-form parity tests: velocities and gas momenta 1e-11 of their
maxima, positions 1e-12 of the box; block ids identical (integer work).  Both sides run the
reference's arithmetic in its order without fused multiply-adds and differ only in the order of
the atomic deposit additions."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

# the small refined mesh of the oracle's pins (tests/test_oracle_smr.py): root 2x2x4 blocks of
# 8^3, levels 1 and 2, 128 leaves
REGIONS = (((0.0, 2.0), (0.0, 2.0), (1.0, 3.0)), ((0.5, 1.5), (0.5, 1.5), (1.5, 2.5)))
BOX = ((0.0, 1.0), (0.0, 1.0), (0.0, 2.0))


def say(msg):
    print("[gpu-smr] " + msg, flush=True)
    sys.stdout.flush()


def small_case(back_reaction=1, damping=True):
    mesh, blocks = abi.smr_mesh((8, 8, 8), (2, 2, 4), BOX, REGIONS)
    phys = workloads._phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=1,
                           back_reaction=back_reaction)
    if damping:
        phys.stiffness, phys.pc2_damping = 1, 1
    su = workloads._setup(seed=4, ppc=2, taus=(0.05,), eps=(1.0,), gas_noise=1e-2,
                          dust_noise=1e-2, gas_vz=0.6, gas_vx=0.3, block_lognormal=1,
                          block_sigma=1.5)
    return mesh, blocks, phys, su


def pair(mesh, blocks, phys, su, cap):
    from athenak_amd.lib import DustContext
    o = Oracle(mesh, blocks, phys, cap)
    o.setup(su)
    g = DustContext(mesh, blocks, phys, cap, form=abi.DK_FORM_FP64)
    g.setup(su)
    return o, g


def by_tag(rdata, idata, n):
    order = np.argsort(idata[abi.PTAG, :n], kind="stable")
    return rdata[:, :n][:, order], idata[:, :n][:, order]


def compare(o, g, what):
    g.stream.synchronize()
    n = o.npart
    assert g.npart == n
    ro, io = by_tag(o.rdata(), o.idata(), n)
    rg, ig = by_tag(g.rdata.cpu().numpy(), g.idata.cpu().numpy(), n)
    np.testing.assert_array_equal(io, ig, err_msg=what + ": integer columns")
    for cols, scale, tol in (([abi.IPX, abi.IPY, abi.IPZ], 1.0, 1e-12),
                             ([abi.IPVX, abi.IPVY, abi.IPVZ], None, 1e-11)):
        s = scale if scale is not None else max(np.abs(ro[cols]).max(), 1e-300)
        err = np.abs(ro[cols] - rg[cols]).max() / s
        assert err < tol, "%s: columns %s differ by %.3e" % (what, cols, err)
    uo, ug = o.u0(), g.u0.cpu().numpy()
    ng = o.mesh.ng
    a = (slice(None), slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    err = np.abs(uo[a][:, 1:] - ug[a][:, 1:]).max() / np.abs(uo[a][:, 1:]).max()
    assert err < 1e-11, "%s: gas momentum differs by %.3e" % (what, err)


def test_refined_mesh_setup_equals_the_oracle():
    mesh, blocks, phys, su = small_case()
    cap = workloads.Workload("t", mesh, blocks, phys, su, None, 0.01, 1).npart + 16
    o, g = pair(mesh, blocks, phys, su, cap)
    say("setup: %d blocks, %d particles" % (mesh.nmb, o.npart))
    g.stream.synchronize()
    n = o.npart
    assert g.npart == n
    np.testing.assert_array_equal(o.idata()[:, :n], g.idata[:, :n].cpu().numpy())
    # (positions are exact; velocity draws carry the host/device log/cos last bits)
    np.testing.assert_array_equal(o.rdata()[[abi.IPX, abi.IPY, abi.IPZ], :n],
                                  g.rdata[[abi.IPX, abi.IPY, abi.IPZ], :n].cpu().numpy())
    np.testing.assert_allclose(o.u0(), g.u0.cpu().numpy(), rtol=1e-15, atol=1e-18)


def test_pc2_clouds_on_the_refined_mesh_match_the_oracle():
    mesh, blocks, phys, su = small_case()
    cap = workloads.Workload("t", mesh, blocks, phys, su, None, 0.01, 1).npart + 16
    o, g = pair(mesh, blocks, phys, su, cap)
    # start both from the oracle's state (the set-up's velocity draws differ in the last bits)
    g.load(o.rdata()[:, :o.npart].copy(), o.idata()[:, :o.npart].copy(), o.u0().copy())
    lev0 = o.idata()[abi.PGID, :o.npart].copy()
    d0 = g.diag()
    for s in range(10):
        say("step %d" % s)
        o.step(0.01)
        g.step(0.01)
        g.check_device_error()
    compare(o, g, "10 pc2 steps")
    moved = (o.idata()[abi.PGID, :o.npart] != lev0).sum()
    assert moved > 100   # particles changed blocks, across level faces
    d1 = g.diag()
    assert np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8] < 1e-13
    np.testing.assert_allclose(g.stiffness(), o.stiffness(), rtol=1e-11)


def test_refined_meshes_are_refused_where_unsupported():
    """The FP32 form and point particles on a refined mesh are refused at creation."""
    from athenak_amd.lib import DustContext, DustError
    mesh, blocks, phys, su = small_case()
    with pytest.raises(DustError, match="option not provided"):
        DustContext(mesh, blocks, phys, 64, form=abi.DK_FORM_FP32)
    phys.clouds = 0
    with pytest.raises(DustError, match="option not provided"):
        DustContext(mesh, blocks, phys, 64, form=abi.DK_FORM_FP64)


def test_full_size_c4_conserves_and_matches_the_oracle():
    """The uneven-block case at its size (792 blocks, 12,144,771 clouds): 2 steps in both, the
    oracle on 16 threads (atomic deposits, as Kokkos::OpenMP)."""
    import torch
    from athenak_amd.lib import DustContext
    w = workloads.c4()
    say("c4: %d blocks, %d particles" % (w.mesh.nmb, w.npart))
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=16)
    o.setup(w.setup)
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=abi.DK_FORM_FP64,
                    stream=torch.cuda.Stream())
    g.load(o.rdata()[:, :o.npart].copy(), o.idata()[:, :o.npart].copy(), o.u0().copy())
    d0 = g.diag()
    for s in range(2):
        say("c4 step %d" % s)
        o.step(w.dt)
        g.step(w.dt)
        g.check_device_error()
    compare(o, g, "c4, 2 steps")
    d1 = g.diag()
    assert np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8] < 1e-12
    g.close()
