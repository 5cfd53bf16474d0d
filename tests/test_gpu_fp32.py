"""The FP32 cell-sorted form through the C-ABI, against the FP64 oracle on the same inputs.

Bars (the targets of the FP32 form): particle velocities within 1e-5 of the FP64 oracle relative
to the largest velocity, total momentum (dust + gas) conserved to 1e-12 relative to
sum m|v| + sum V|M| on periodic meshes without external forces, and every exported array and the
gas bit-identical between two runs (the deposits are atomic-free and the sort is stable)."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-fp32] " + msg, flush=True)
    sys.stdout.flush()


def fp32_ctx(w, own_stream=False):
    """(own_stream: on a stream of its own, default priority; graphs are captured only off the
    null stream, which is torch's current stream here)"""
    import torch
    from athenak_amd.lib import DustContext
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=abi.DK_FORM_FP32,
                    stream=torch.cuda.Stream() if own_stream else None)
    g.setup(w.setup)
    return g


def exported(g):
    g.export()
    n = g.npart
    r = g.rdata[:, :n].cpu().numpy()
    i = g.idata[:, :n].cpu().numpy()
    order = np.argsort(i[abi.PTAG], kind="stable")
    return r[:, order], i[:, order], g.u0.cpu().numpy()


def oracle_run(w, steps):
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=4)
    o.setup(w.setup)
    for _ in range(steps):
        o.step(w.dt, w.src)
    n = o.npart
    order = np.argsort(o.idata()[abi.PTAG, :n], kind="stable")
    return o.rdata()[:, :n][:, order].copy(), o.idata()[:, :n][:, order].copy(), o.u0().copy()


def check_against_oracle(w, steps, vtol=1e-5):
    say("%s: oracle %d steps" % (w.name, steps))
    ro, io, uo = oracle_run(w, steps)
    say("%s: fp32 %d steps" % (w.name, steps))
    g = fp32_ctx(w)
    d0 = g.diag()
    for _ in range(steps):
        g.step(w.dt, w.src)
    g.check_device_error()
    d1 = g.diag()
    rg, ig, ug = exported(g)
    np.testing.assert_array_equal(io[abi.PTAG], ig[abi.PTAG])
    np.testing.assert_array_equal(io[abi.PGID], ig[abi.PGID])
    v = [abi.IPVX, abi.IPVY, abi.IPVZ]
    vmax = np.abs(ro[v]).max()
    verr = np.abs(ro[v] - rg[v]).max() / vmax
    # positions: the velocity tolerance over the distance travelled
    travel = vmax * w.dt * steps
    xerr = max(np.abs(ro[c] - rg[c]).max() for c in (abi.IPX, abi.IPY, abi.IPZ)) / travel
    ng = w.mesh.ng
    ua, ub = uo[:, 1:, ng:-ng, ng:-ng, ng:-ng], ug[:, 1:, ng:-ng, ng:-ng, ng:-ng]
    merr = np.abs(ua - ub).max() / np.abs(ua).max()
    drift = np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8]
    say("%s: max rel v err %.3e, x err/travel %.3e, gas mom err %.3e, momentum drift %.3e"
        % (w.name, verr, xerr, merr, drift))
    assert verr < vtol, verr
    assert merr < vtol, merr
    assert xerr < vtol, xerr
    # (no invariant in a shearing box, which has external forces, or with physical x3 faces,
    # where particles reflect or leave and the gas meets walls)
    if not w.phys.shearing_box and w.mesh.bc_x3 == abi.DK_X3_PERIODIC:
        assert drift < 1e-12, drift
    return g


def test_fp32_c1_100_steps_within_1e5_of_fp64_oracle():
    check_against_oracle(workloads.c1(), 100)


def test_fp32_particles_crossing_blocks():
    w = workloads.c1(n=16, nb=2, ppc=2)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    w.name = "fp32-cross"
    check_against_oracle(w, 20)


def test_fp32_dc2_and_local():
    for solver in (abi.DK_DC2, abi.DK_LOCAL):
        w = workloads.c1(n=16, nb=2, ppc=2, solver=solver)
        w.setup.dust_noise, w.setup.gas_noise, w.dt = 1.0, 0.3, 2e-3
        w.name = "fp32-solver%d" % solver
        check_against_oracle(w, 10)


def test_fp32_shearing_box_small():
    w = workloads.c2(n=32, nbx=2, ppc=2)
    w.setup.dust_noise, w.dt = 0.5, 1e-3
    w.name = "fp32-c2-small"
    check_against_oracle(w, 20)


def test_fp32_clumped_small():
    w = workloads.c5(n=32, nbx=2, ppc=8)
    w.name = "fp32-c5-small"
    check_against_oracle(w, 10)


def pc2_check(w, steps, stage1=False):
    """pc2 against the oracle after `steps` steps (and one more stage 1 with stage1): as
    check_against_oracle, plus the registers x1, v1 and R that the export assembles (x1 and the
    block after the migration the next key pass would apply, R from the store it was left in)."""
    say("%s: oracle %d steps%s" % (w.name, steps, " + stage 1" if stage1 else ""))
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=4)
    o.setup(w.setup)
    g = fp32_ctx(w)
    d0 = g.diag()
    for _ in range(steps):
        o.step(w.dt, w.src)
        g.step(w.dt, w.src)
    if stage1:
        o.stage(abi.rk2_stages(o.time, w.dt)[0], w.src)
        g.stage(g.stages(w.dt)[0], w.src)
    g.check_device_error()
    d1 = g.diag()
    n = o.npart
    oo = np.argsort(o.idata()[abi.PTAG, :n], kind="stable")
    ro, io, uo = o.rdata()[:, :n][:, oo], o.idata()[:, :n][:, oo], o.u0()
    rg, ig, ug = exported(g)
    np.testing.assert_array_equal(io[abi.PTAG], ig[abi.PTAG])
    np.testing.assert_array_equal(io[abi.PGID], ig[abi.PGID])
    v = [abi.IPVX, abi.IPVY, abi.IPVZ]
    vmax = np.abs(ro[v]).max()
    travel = vmax * w.dt * max(steps, 1)
    errs = {
        "v": max(np.abs(ro[c] - rg[c]).max() for c in v) / vmax,
        "v1": max(np.abs(ro[c] - rg[c]).max() for c in (abi.IPVX1, abi.IPVY1, abi.IPVZ1)) / vmax,
        "R": max(np.abs(ro[c] - rg[c]).max() for c in (abi.IPRX, abi.IPRY, abi.IPRZ)) / vmax,
        "x": max(np.abs(ro[c] - rg[c]).max() for c in (abi.IPX, abi.IPY, abi.IPZ)) / travel,
        "x1": max(np.abs(ro[c] - rg[c]).max() for c in (abi.IPX1, abi.IPY1, abi.IPZ1)) / travel,
    }
    ng = w.mesh.ng
    ua, ub = uo[:, 1:, ng:-ng, ng:-ng, ng:-ng], ug[:, 1:, ng:-ng, ng:-ng, ng:-ng]
    errs["gas"] = np.abs(ua - ub).max() / np.abs(ua).max()
    drift = np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8]
    say("%s: %s, momentum drift %.3e" % (w.name, ", ".join("%s %.3e" % kv for kv in
                                                             errs.items()), drift))
    for k, e in errs.items():
        assert e < 1e-5, (k, e)
    if not stage1:   # (the predictor's deposit is a transient of the cycle)
        assert drift < 1e-12, drift
    return g


def test_fp32_pc2_particles_crossing_blocks():
    w = workloads.c1(n=16, nb=2, ppc=2, coupling=abi.DK_PC2, solver=abi.DK_LOCAL)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    w.name = "fp32-pc2-cross"
    pc2_check(w, 20)
    pc2_check(w, 7, stage1=True)


def test_fp32_pc2_stiffness_and_damping():
    # t_s = 2e-3 at dt = 4e-3: chi = dt sum(rate)/rho reaches 2 in the cells with dust
    w = workloads.c1(n=16, nb=2, ppc=2, coupling=abi.DK_PC2, solver=abi.DK_LOCAL)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    w.setup.taus[0], w.setup.dust_noise, w.dt = 2e-3, 0.5, 4e-3
    w.name = "fp32-pc2-damped"
    pc2_check(w, 10)


def test_fp32_pc2_clumped_and_schemes():
    # clumps: tiles of more than one part; NGP and CIC as well as TSC
    for dep in (abi.DK_TSC, abi.DK_CIC, abi.DK_NGP):
        w = workloads.c4u(n=32, nbx=2, ppc=8, clouds=False)
        w.phys.deposit = dep
        w.setup.dust_noise = 0.3
        w.name = "fp32-pc2-c4p-dep%d" % dep
        pc2_check(w, 6)


def test_fp32_pc2_clouds_against_the_oracle():
    # refinable clouds at the mesh level (CloudPredictorT / CloudAdvanceT): the TSC weights and
    # the linear gather from the velocity pack, clouds reaching into neighbour blocks and across
    # periodic faces; then clumps with the stiffness field and damping (tiles of several parts)
    w = workloads.c4u(n=16, nbx=2, ppc=2)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    w.name = "fp32-clouds-cross"
    pc2_check(w, 20)
    pc2_check(w, 5, stage1=True)
    w = workloads.c4u(n=32, nbx=2, ppc=8)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    w.setup.dust_noise = 0.3
    w.name = "fp32-clouds-clumped"
    pc2_check(w, 6)


def test_fp32_pc2_uneven_blocks_against_the_oracle():
    # lognormal particle counts per block (the batching case: blocks of a few particles next to
    # blocks of thousands, one launch over all of them), point particles and clouds
    for clouds in (0, 1):
        w = workloads.c4b(n=32, nbx=4, ppc=4)
        w.phys.clouds = clouds
        w.setup.dust_noise = 0.3
        w.name = "fp32-pc2-uneven-clouds%d" % clouds
        pc2_check(w, 6)


def test_fp32_pc2_bit_identical_reruns():
    w = workloads.c4u(n=32, nbx=2, ppc=8, clouds=False)
    w.setup.dust_noise = 0.5
    outs = []
    for run in range(2):
        say("pc2 rerun %d" % run)
        g = fp32_ctx(w)
        for _ in range(4):
            g.step(w.dt, w.src)
        g.check_device_error()
        outs.append(exported(g))
        g.close()
    for a, b in zip(outs[0], outs[1]):
        assert a.tobytes() == b.tobytes()


def test_fp32_crowded_cells_split_over_helpers():
    # clumps a quarter cell wide: ~2000 particles in one cell, so that every chunk of a clump tile
    # is one cell and its records are split over up to 47 helper threads (runs of 16); imex2+/dc1
    # and pc2 against the oracle, and two runs bit-identical
    for coupling in (abi.DK_IMEX, abi.DK_PC2):
        w = workloads.c5(n=32, nbx=2, ppc=8)
        w.setup.clump_sigma = 0.25 / 32
        if coupling == abi.DK_PC2:
            w.phys.coupling, w.phys.solver = abi.DK_PC2, abi.DK_LOCAL
            w.phys.stiffness, w.phys.pc2_damping = 1, 1
        w.name = "fp32-crowded-%s" % ("pc2" if coupling == abi.DK_PC2 else "imex")
        g = check_against_oracle(w, 5)
        a = exported(g)
        g.close()
        g = fp32_ctx(w)
        for _ in range(5):
            g.step(w.dt, w.src)
        g.check_device_error()
        b = exported(g)
        g.close()
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes()


def test_fp32_bit_identical_reruns():
    w = workloads.c5(n=32, nbx=2, ppc=8)
    w.setup.dust_noise = 0.5
    outs = []
    for run in range(2):
        say("rerun %d" % run)
        g = fp32_ctx(w)
        for _ in range(5):
            g.step(w.dt, w.src)
        g.check_device_error()
        outs.append(exported(g))
        g.close()
    for a, b in zip(outs[0], outs[1]):
        assert a.tobytes() == b.tobytes()


def test_fp32_bind_export_roundtrip():
    w = workloads.c1(n=16, nb=2, ppc=2)
    from athenak_amd.lib import DustContext
    g64 = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16)
    g64.setup(w.setup)
    n = g64.npart
    r0 = g64.rdata[:, :n].cpu().numpy()
    i0 = g64.idata[:, :n].cpu().numpy()
    g = fp32_ctx(w)
    rg, ig, _ = exported(g)
    o = np.argsort(i0[abi.PTAG], kind="stable")
    r0, i0 = r0[:, o], i0[:, o]
    np.testing.assert_array_equal(i0, ig)
    for c in (abi.IPX, abi.IPY, abi.IPZ, abi.IPX1, abi.IPY1, abi.IPZ1):
        np.testing.assert_array_equal(r0[c], rg[c])
    for c in (abi.IPVX, abi.IPVY, abi.IPVZ, abi.IPTS, abi.IPM):
        np.testing.assert_array_equal(r0[c].astype(np.float32).astype(np.float64), rg[c])


def test_fp32_stratified_outflow_small():
    # particles leave through the physical x3 faces (outflow: parked at rest with zero mass)
    w = workloads.c3(n=32, nbx=2)
    w.setup.hp, w.setup.dust_noise, w.dt = 0.05, 1.0, 2e-3
    w.name = "fp32-c3-small"
    g = check_against_oracle(w, 15)
    rg, _, _ = exported(g)
    parked = int((rg[abi.IPM] == 0.0).sum())
    say("c3-small: %d particles parked at the x3 faces" % parked)
    assert parked > 0


def test_fp32_tile_shapes_agree():
    # the tile shape groups the cross-tile sums: every shape is deterministic, and two shapes
    # agree to FP32 round-off
    import os
    w = workloads.c5(n=32, nbx=2, ppc=8)
    w.setup.dust_noise = 0.5
    outs = []
    for shape in ("8x4x4", "4x4x8"):
        os.environ["DK_TILE"] = shape
        try:
            g = fp32_ctx(w)
        finally:
            del os.environ["DK_TILE"]
        for _ in range(5):
            g.step(w.dt, w.src)
        g.check_device_error()
        outs.append(exported(g))
        g.close()
    (ra, ia, ua), (rb, ib, ub) = outs
    np.testing.assert_array_equal(ia, ib)
    v = [abi.IPVX, abi.IPVY, abi.IPVZ]
    err = np.abs(ra[v] - rb[v]).max() / np.abs(ra[v]).max()
    say("tile shapes 8x4x4 vs 4x4x8: max rel v difference %.3e" % err)
    assert err < 1e-5


def test_fp32_graph_replays_match_eager():
    # each stage captured into a graph and replayed (the executable graph updated every stage),
    # with eager work and host synchronisations between the replays: every step identical to
    # the eager run
    for w in (workloads.c5(n=32, nbx=2, ppc=4), workloads.c4u(n=32, nbx=2, ppc=4, clouds=False)):
        w.setup.dust_noise = 0.5
        graph_replays(w)


def graph_replays(w):
    ge, gg = fp32_ctx(w), fp32_ctx(w, own_stream=True)
    gg.set_graphs(True)
    for step in range(4):
        ge.step(w.dt, w.src)
        gg.step(w.dt, w.src)
        ge.check_device_error()
        gg.check_device_error()
        de, dg = ge.diag(), gg.diag()          # eager reductions and a host sync in between
        assert np.array_equal(de, dg), (step, de, dg)
        a, b = exported(ge), exported(gg)
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes(), step
        say("%s: graph replay step %d identical" % (w.name, step))
    assert gg.graphs_captured() > 0
    ge.close()
    gg.close()


def test_fp32_registers_mid_step_and_at_step_end():
    # after stage 1 the registers x1, v1, R are live and exported as the oracle's; after the
    # step's last stage the store keeps none (the next stage 1 sets them before any use) and
    # export writes x1 = x, v1 = v, R = 0
    w = workloads.c1(n=16, nb=2, ppc=2)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 1.0, 0.3, 2e-3
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16)
    o.setup(w.setup)
    g = fp32_ctx(w)
    st = abi.imex2p_stages(0.0, w.dt)
    o.stage(st[0], w.src)
    g.stage(st[0], w.src)
    g.check_device_error()
    rg, ig, _ = exported(g)
    n = o.npart
    order = np.argsort(o.idata()[abi.PTAG, :n], kind="stable")
    ro = o.rdata()[:, :n][:, order]
    np.testing.assert_array_equal(o.idata()[abi.PTAG, :n][order], ig[abi.PTAG])
    for c in (abi.IPX1, abi.IPY1, abi.IPZ1):
        assert np.abs(ro[c] - rg[c]).max() <= 1e-12, c
    for cols in ((abi.IPVX1, abi.IPVY1, abi.IPVZ1), (abi.IPRX, abi.IPRY, abi.IPRZ)):
        scale = np.abs(ro[list(cols)]).max()
        err = np.abs(ro[list(cols)] - rg[list(cols)]).max() / scale
        say("stage 1 registers %s: max rel err %.3e" % (cols, err))
        assert err < 1e-5, (cols, err)
    for s in st[1:]:
        g.stage(s, w.src)
    g.check_device_error()
    rg, _, _ = exported(g)
    for a, b in ((abi.IPX1, abi.IPX), (abi.IPY1, abi.IPY), (abi.IPZ1, abi.IPZ),
                 (abi.IPVX1, abi.IPVX), (abi.IPVY1, abi.IPVY), (abi.IPVZ1, abi.IPVZ)):
        np.testing.assert_array_equal(rg[a], rg[b])
    for c in (abi.IPRX, abi.IPRY, abi.IPRZ):
        assert not rg[c].any()
    g.close()


def test_fp32_registers_kept_at_step_end():
    # with dk_keep_registers every stage keeps x1, v1 and R: the export after whole steps carries
    # the oracle's x1, v1 (the stage-1 copies as the step's migration across blocks and the
    # periodic faces left them) and R (the last stage's drag rate with its consistent correction)
    w = workloads.c1(n=16, nb=2, ppc=2)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 1.0, 0.3, 2e-3
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16)
    o.setup(w.setup)
    g = fp32_ctx(w)
    g.keep_registers(True)
    for _ in range(3):
        o.step(w.dt, w.src)
        g.step(w.dt, w.src)
    g.check_device_error()
    rg, ig, _ = exported(g)
    n = o.npart
    order = np.argsort(o.idata()[abi.PTAG, :n], kind="stable")
    ro = o.rdata()[:, :n][:, order]
    np.testing.assert_array_equal(o.idata()[abi.PTAG, :n][order], ig[abi.PTAG])
    vmax = np.abs(ro[[abi.IPVX, abi.IPVY, abi.IPVZ]]).max()
    travel = vmax * w.dt
    err = max(np.abs(ro[c] - rg[c]).max() for c in (abi.IPX1, abi.IPY1, abi.IPZ1)) / travel
    say("step-end x1: max err %.3e of a step's travel" % err)
    assert err < 1e-5, err
    for cols in ((abi.IPVX1, abi.IPVY1, abi.IPVZ1), (abi.IPRX, abi.IPRY, abi.IPRZ)):
        scale = np.abs(ro[list(cols)]).max()
        assert scale > 0.0
        err = np.abs(ro[list(cols)] - rg[list(cols)]).max() / scale
        say("step-end registers %s: max rel err %.3e" % (cols, err))
        assert err < 1e-5, (cols, err)
    g.close()


def test_fp32_autotuner_choices_are_bit_identical(tmp_path):
    # every choice the autotuner makes (tile dealing to XCDs, key-pass workgroup size) leaves
    # the results bit-identical; online tuning over the caller's stages and the cache file
    # therefore change nothing either
    import os
    w = workloads.c5(n=32, nbx=2, ppc=8)
    w.setup.dust_noise = 0.5
    outs = []
    for xcd, kt in ((1, 256), (0, 128), (1, 512)):
        g = fp32_ctx(w)
        g.set_tune_config(xcd, kt)
        for _ in range(5):
            g.step(w.dt, w.src)
        g.check_device_error()
        outs.append(exported(g))
        g.close()
        say("tune config xcd=%d key_threads=%d done" % (xcd, kt))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert a.tobytes() == b.tobytes()
    cache = tmp_path / "tune.json"
    os.environ["DK_TUNE_CACHE"] = str(cache)
    try:
        g = fp32_ctx(w)
        g.autotune(True)
        assert g.tune_config()[0] == 1
        steps = 0
        while g.tune_config()[0] == 1 and steps < 20:
            g.step(w.dt, w.src)
            steps += 1
        state, xcd, kt = g.tune_config()
        assert state == 2, state
        for _ in range(5 - steps if steps < 5 else 0):
            g.step(w.dt, w.src)
        say("online tuning: %d steps, chose xcd=%d key_threads=%d" % (steps, xcd, kt))
        assert cache.exists() and '"key_threads"' in cache.read_text()
        g.close()
        g2 = fp32_ctx(w)
        g2.autotune(True)
        assert g2.tune_config() == (2, xcd, kt)
        g2.close()
    finally:
        del os.environ["DK_TUNE_CACHE"]
    # the tuned run (its first steps ran under every candidate) against a plain run
    g = fp32_ctx(w)
    g.autotune(True)
    for _ in range(8):
        g.step(w.dt, w.src)
    g.check_device_error()
    tuned = exported(g)
    g.close()
    g = fp32_ctx(w)
    for _ in range(8):
        g.step(w.dt, w.src)
    plain = exported(g)
    g.close()
    for a, b in zip(tuned, plain):
        assert a.tobytes() == b.tobytes()


def test_fp32_new_time_step_and_sort():
    # NewTimeStep over the FP32 store (shearing box: the azimuthal velocity without the
    # background shear) against the oracle from the same state, at the start and after steps:
    # the FP32 velocities put it within 1e-5 relative; the cell sort is a no-op that succeeds
    import ctypes
    from oracle.pyoracle import lib as olib
    w = workloads.c2(n=16, nbx=2, ppc=2)
    w.setup.dust_noise, w.dt = 0.5, 1e-3
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=4)
    o.setup(w.setup)
    g = fp32_ctx(w)
    for steps in (0, 3):
        for _ in range(steps):
            o.step(w.dt, w.src)
            g.step(w.dt, w.src)
        out = ctypes.c_double(0.0)
        olib().or_new_time_step(o.h, 0.5, ctypes.byref(out))
        dt = g.new_time_step(0.5)
        say("new time step after %d steps: %.9e (oracle %.9e)" % (steps, dt, out.value))
        assert abs(dt - out.value) <= 1e-5 * out.value
    g.sort()
    g.check_device_error()
    g.close()


def test_fp32_shearing_box_new_time_step_stays_within_the_frame_limit():
    # the FP32 form's stages refuse an orbital-advection frame shift beyond half a cell in x2;
    # its new time step never proposes such a dt: a stage at the proposed cap runs, one 20 %
    # above it is refused before any work
    from athenak_amd.lib import DustError
    w = workloads.c2(n=16, nbx=2, ppc=2)
    g = fp32_ctx(w)
    dt = g.new_time_step(1.0e6)   # (no particle CFL limit: the frame limit alone)
    dy = (w.mesh.x2max - w.mesh.x2min) / (w.mesh.nx2 * w.mesh.nmbx2)
    xmax = max(abs(w.mesh.x1min), abs(w.mesh.x1max))
    cap = 0.5 * dy / (w.phys.qshear * w.phys.omega0 * (1.0 + 0.5 ** 0.5) * xmax)
    say("shearing-box dt cap %.9e (frame limit %.9e)" % (dt, cap))
    assert cap * (1 - 1e-8) <= dt <= cap
    g.step(dt, w.src)
    g.check_device_error()
    with pytest.raises(DustError, match="option not provided"):
        g.step(1.2 * dt, w.src)
    g.close()


def test_pc2_stiffness_record_both_forms():
    # UpdateStiffness (dust_pm.cpp:288-342): chi of the last cycle and its running maximum, in
    # the FP64 form to the oracle's round-off and in the FP32 form to FP32 round-off
    from athenak_amd.lib import DustContext
    w = workloads.c1(n=16, nb=2, ppc=2, coupling=abi.DK_PC2, solver=abi.DK_LOCAL)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    w.setup.taus[0], w.setup.dust_noise, w.dt = 2e-3, 0.5, 4e-3
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16)
    o.setup(w.setup)
    g64 = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=abi.DK_FORM_FP64)
    g64.setup(w.setup)
    g32 = fp32_ctx(w)
    for step in range(3):
        for x in (o, g64, g32):
            x.step(w.dt, w.src)
        so, s64, s32 = o.stiffness(), g64.stiffness(), g32.stiffness()
        say("stiffness step %d: oracle %s, fp64 %s, fp32 %s" % (step, so, s64, s32))
        assert so[0] > 1.0 and abs(so[2] * w.dt - so[0]) <= 1e-15 * so[0]
        assert np.all(np.abs(s64 - so) <= 1e-12 * so), (s64, so)
        assert np.all(np.abs(s32 - so) <= 1e-5 * so), (s32, so)
    # NewTimeStep's pc2 limit: stiff_thresh over the recorded rate
    dt_lim = g64.new_time_step_pc2(0.3, stiff_thresh=0.5)
    assert dt_lim == min(g64.new_time_step(0.3), 0.5 / s64[2])
    g64.check_device_error()
    g32.check_device_error()
    g64.close()
    g32.close()


def test_stopping_time_range_and_validation():
    # RefreshStoppingTimeMaximum: the largest and smallest stopping time of a continuous size
    # distribution (the FP32 form: of its FP32 copies), and an invalid one reported (DK_ESTOP);
    # GammaSwitch picks 1/2 only when dt exceeds the largest
    from athenak_amd.lib import DustContext, DustError
    w = workloads.c3(n=32, nbx=2)
    g64 = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=abi.DK_FORM_FP64)
    g64.setup(w.setup)
    g32 = fp32_ctx(w)
    ts = g64.rdata[abi.IPTS, :g64.npart].cpu().numpy()
    mx, mn = g64.stopping_times()
    say("stopping times: fp64 (%.6e, %.6e), numpy (%.6e, %.6e)" % (mx, mn, ts.max(), ts.min()))
    assert (mx, mn) == (ts.max(), ts.min())
    t32 = ts.astype(np.float32).astype(np.float64)
    assert g32.stopping_times() == (t32.max(), t32.min())
    assert abi.gamma_switch(2.0 * mx, mx) == 0.5
    assert abi.gamma_switch(mx, mx) == abi.IMEX2P_GAMMA
    assert abi.gamma_switch(2.0 * mx, mx, abi.DK_PC2) is None
    g64.rdata[abi.IPTS, 5] = -1.0
    g64.stream.synchronize()
    with pytest.raises(DustError, match="stopping"):
        g64.stopping_times()
    g64.close()
    g32.close()
