"""The HIP forms through the C-ABI on the pieces that the oracle pins of test_oracle_pins.py fix:
the pc2 damping of stiff cells without the stiffness record, the x3 outflow ledger and the removal
of outflow particles (both forms against the oracle), and the FP32 form's counting sort, whose
integer output must be the stable sort of its keys, each key the tile-major cell of the particle's
position."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-pins] " + msg, flush=True)
    sys.stdout.flush()


def ctx(w, form):
    from athenak_amd.lib import DustContext
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=form)
    g.setup(w.setup)
    return g


def state(g):
    g.export()
    n = g.npart
    r = g.rdata[:, :n].cpu().numpy()
    i = g.idata[:, :n].cpu().numpy()
    order = np.argsort(i[abi.PTAG], kind="stable")
    return r[:, order], i[:, order], g.u0.cpu().numpy()


def ostate(o):
    n = o.npart
    order = np.argsort(o.idata()[abi.PTAG, :n], kind="stable")
    return o.rdata()[:, :n][:, order].copy(), o.idata()[:, :n][:, order].copy(), o.u0().copy()


def vel_err(ro, rg):
    cols = [abi.IPVX, abi.IPVY, abi.IPVZ]
    return np.abs(ro[cols] - rg[cols]).max() / np.abs(ro[cols]).max()


def test_pc2_damping_without_the_stiffness_record_both_forms():
    # the damping alone deposits the rate field and damps the stiff cells (StiffnessActive,
    # src/dust/dust.hpp:1150-1152): both forms against the oracle, and the chi record kept
    w = workloads.c1(n=16, nb=2, ppc=2, coupling=abi.DK_PC2, solver=abi.DK_LOCAL)
    w.phys.stiffness, w.phys.pc2_damping = 0, 1
    w.setup.taus[0], w.setup.dust_noise, w.dt = 2e-3, 0.5, 4e-3
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16)
    o.setup(w.setup)
    g64, g32 = ctx(w, abi.DK_FORM_FP64), ctx(w, abi.DK_FORM_FP32)
    for step in range(4):
        for x in (o, g64, g32):
            x.step(w.dt, w.src)
    so = o.stiffness()
    assert so[0] > 1.0, so          # stiff cells were damped
    ro, _, uo = ostate(o)
    for g, tol, name in ((g64, 1e-11, "fp64"), (g32, 1e-5, "fp32")):
        g.check_device_error()
        rg, _, ug = state(g)
        e = vel_err(ro, rg)
        say("damping only, %s: velocity %.2e, chi %s vs %s" % (name, e, g.stiffness(), so))
        assert e < tol
        assert np.all(np.abs(g.stiffness() - so) <= 10 * tol * so)
        ng = w.mesh.ng
        a = lambda u: u[:, 1:, ng:-ng, ng:-ng, ng:-ng]   # noqa: E731
        assert np.abs(a(uo) - a(ug)).max() <= tol * np.abs(a(uo)).max()
        g.close()


def test_outflow_ledger_and_removal_both_forms():
    # particles leave through the x3 faces; the ledger and RemoveOutflowParticles
    # (src/dust/dust_tasks.cpp:365-420) agree with the oracle, and the run continues from the
    # compacted particles
    w = workloads.c3(n=32, nbx=2)
    w.setup.hp, w.setup.dust_noise, w.dt = 0.05, 1.0, 2e-3
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=4)
    o.setup(w.setup)
    g64, g32 = ctx(w, abi.DK_FORM_FP64), ctx(w, abi.DK_FORM_FP32)
    for _ in range(8):
        for x in (o, g64, g32):
            x.step(w.dt, w.src)
    lo = o.x3_removed()
    say("oracle ledger: mass %.9e, count %d" % (lo[0], lo[1]))
    assert lo[1] > 0
    for g, tol in ((g64, 1e-14), (g32, 1e-6)):
        lg = g.x3_removed()
        assert lg[1] == lo[1], (lg, lo)
        assert abs(lg[0] - lo[0]) <= tol * lo[0], (lg, lo)
    oo = o.remove_outflow_particles()
    ro, io, _ = ostate(o)
    for g, tol, name in ((g64, 1e-14, "fp64"), (g32, 1e-6, "fp32")):
        out = g.remove_outflow_particles()
        say("%s removal: %s (oracle %s), npart %d (oracle %d)" % (name, out, oo, g.npart,
                                                                  o.npart))
        assert out[1] == oo[1] and out[3] == oo[3]
        assert abs(out[0] - oo[0]) <= tol * oo[0]
        assert g.npart == o.npart
        assert np.array_equal(g.x3_removed(), [0.0, 0.0])
        rg, ig, _ = state(g)
        np.testing.assert_array_equal(ig[abi.PTAG], io[abi.PTAG])
        assert (rg[abi.IPM] > 0).all()
        # a second call removes nothing
        out2 = g.remove_outflow_particles()
        assert out2[1] == 0.0 and out2[3] == oo[3] and g.npart == o.npart
    # the run goes on from the compacted particles
    for _ in range(3):
        for x in (o, g64, g32):
            x.step(w.dt, w.src)
    ro, _, _ = ostate(o)
    for g, tol, name in ((g64, 1e-11, "fp64"), (g32, 1e-5, "fp32")):
        g.check_device_error()
        rg, _, _ = state(g)
        e = vel_err(ro, rg)
        say("after removal, %s: velocity %.2e" % (name, e))
        assert e < tol
        g.close()


def _tile_keys(x, gid, mesh, blocks, tile=(8, 4, 4)):
    """The tile-major key of the cell holding each position (PMWeights' cell, clamped into the
    block), computed here from the block boxes."""
    TX, TY, TZ = tile
    nx = (mesh.nx1, mesh.nx2, mesh.nx3)
    ntx, nty, ntz = nx[0] // TX, nx[1] // TY, nx[2] // TZ
    TC = TX * TY * TZ
    cells = []
    for d, (lo, hi) in enumerate((("x1min", "x1max"), ("x2min", "x2max"), ("x3min", "x3max"))):
        a = np.array([getattr(blocks[m], lo) for m in gid])
        b = np.array([getattr(blocks[m], hi) for m in gid])
        c = ((x[d] - a) * (nx[d] / (b - a)) + 1.0).astype(np.int64) - 1
        cells.append(np.clip(c, 0, nx[d] - 1))
    ci, cj, ck = cells
    tilei = ((gid * ntz + ck // TZ) * nty + cj // TY) * ntx + ci // TX
    local = ((ck % TZ) * TY + cj % TY) * TX + ci % TX
    return tilei * TC + local


def test_fp32_sort_output_is_the_stable_counting_sort():
    """The FP32 sort's output against numpy: every cell's list holds first the particles that
    stay in their tile, then those that arrive from other tiles, each group in ascending source
    slot (numpy's lexicographic sort of (key, arrival, source slot))."""
    w = workloads.c1(n=16, nb=2, ppc=3)
    w.setup.dust_noise, w.dt = 3.0, 2e-3      # particles change cells, tiles and blocks
    g = ctx(w, abi.DK_FORM_FP32)
    st = g.stages(w.dt)
    g.step(w.dt)
    # the tile of every slot of the present order: the keys of the last sort, in its new order
    key0, perm0, _ = g.sort_state()
    TC = 8 * 4 * 4
    tile_of_slot = key0[perm0] // TC
    g.stage(g.stages(w.dt)[0])                # stage 1 of the next step: its sort is the last
    key, perm, start = g.sort_state()
    g.check_device_error()
    n = g.npart
    nk = w.mesh.nmb * w.mesh.nx1 * w.mesh.nx2 * w.mesh.nx3
    assert len(st) == 3 and len(key) == n and len(perm) == n and len(start) == nk + 1
    # a permutation, and the order of (key, stayer before arrival, source slot)
    assert np.array_equal(np.sort(perm), np.arange(n))
    arrival = (key // TC) != tile_of_slot
    np.testing.assert_array_equal(perm, np.lexsort((np.arange(n), arrival, key)))
    assert 0 < arrival.sum() < n              # both kinds occur
    # the cell starts are the exclusive scan of the key counts
    counts = np.bincount(key, minlength=nk)
    np.testing.assert_array_equal(start, np.concatenate([[0], np.cumsum(counts)]))
    moved = int((np.diff(key[perm]) < 0).sum())
    assert moved == 0
    # every key is the cell of the particle's position (the store holds the new order)
    g.export()
    r = g.rdata[:, :n].cpu().numpy()
    i = g.idata[:, :n].cpu().numpy()
    x = r[[abi.IPX, abi.IPY, abi.IPZ]]
    want = _tile_keys(x, i[abi.PGID].astype(np.int64), w.mesh, w.blocks)
    np.testing.assert_array_equal(key[perm], want)
    # particles did move between cells since the set-up's order (the sort is not trivial)
    assert (perm != np.arange(n)).sum() > n // 100
    g.close()
