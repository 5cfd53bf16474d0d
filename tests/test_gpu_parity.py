"""Parity of the HIP library (plain FP64 form, through the C-ABI) with the CPU oracle on the
same seeded inputs.

Tolerances: positions 1e-12 of the box, velocities and gas momenta 1e-11 relative to their
maxima.  Both sides compute in FP64 with the reference's operation order and no fused
multiply-add; they differ only in the order in which atomic deposits add up (the reference's
GPU order is unspecified too, src/dust/dust.cpp:191-193), so agreement is at round-off.
Integer results (block ids, tags, sort order) must be identical."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-parity] " + msg, flush=True)
    sys.stdout.flush()


def make_pair(w, form=abi.DK_FORM_FP64, own_stream=False):
    """(own_stream: the context on a stream of its own, default priority: graphs are captured
    only off the null stream, which is torch's current stream here)"""
    import torch
    from athenak_amd.lib import DustContext
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16)
    o.setup(w.setup)
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=form,
                    stream=torch.cuda.Stream() if own_stream else None)
    g.setup(w.setup)
    return o, g


def gpu_state(g):
    g.stream.synchronize()
    n = g.npart
    r = g.rdata[:, :n].cpu().numpy()
    i = g.idata[:, :n].cpu().numpy()
    return r, i, g.u0.cpu().numpy()


def by_tag(r, i):
    order = np.argsort(i[abi.PTAG], kind="stable")
    return r[:, order], i[:, order]


def compare(o, g, w, vtol=1e-11, xtol=1e-12, what=""):
    n = o.npart
    assert g.npart == n
    ro, io = by_tag(o.rdata()[:, :n].copy(), o.idata()[:, :n].copy())
    rg, ig, u0g = gpu_state(g)
    rg, ig = by_tag(rg, ig)
    np.testing.assert_array_equal(io[abi.PTAG], ig[abi.PTAG])
    np.testing.assert_array_equal(io[abi.PGID], ig[abi.PGID], err_msg=what + " PGID")
    L = max(w.mesh.x1max - w.mesh.x1min, w.mesh.x3max - w.mesh.x3min)
    for c in (abi.IPX, abi.IPY, abi.IPZ, abi.IPX1, abi.IPY1, abi.IPZ1):
        assert np.abs(ro[c] - rg[c]).max() <= xtol * L, (what, c)
    for cols in ((abi.IPVX, abi.IPVY, abi.IPVZ), (abi.IPRX, abi.IPRY, abi.IPRZ),
                 (abi.IPVX1, abi.IPVY1, abi.IPVZ1)):
        scale = max(np.abs(ro[list(cols)]).max(), 1e-300)
        err = np.abs(ro[list(cols)] - rg[list(cols)]).max() / scale
        assert err <= vtol, (what, cols, err)
    # stopping times and masses: equal, except the continuous bins whose exp/log/pow draws may
    # round 1-2 ulp apart between the host and device math libraries
    if w.setup.continuous:
        np.testing.assert_allclose(ro[abi.IPTS], rg[abi.IPTS], rtol=5e-16, atol=0)
        np.testing.assert_allclose(ro[abi.IPM], rg[abi.IPM], rtol=1e-15, atol=0)
    else:
        np.testing.assert_array_equal(ro[abi.IPTS], rg[abi.IPTS])
        np.testing.assert_array_equal(ro[abi.IPM], rg[abi.IPM])
    ng = w.mesh.ng
    uo = o.u0()[:, :, ng:-ng, ng:-ng, ng:-ng]
    ug = u0g[:, :, ng:-ng, ng:-ng, ng:-ng]
    if w.src is not None and w.src.hydro:
        # with the gas fluxes on, the density follows the momenta, whose deposits sum in
        # another order: the same relative tolerance
        assert np.abs(uo[:, 0] - ug[:, 0]).max() <= vtol * np.abs(uo[:, 0]).max(), what
    else:
        np.testing.assert_array_equal(uo[:, 0], ug[:, 0])
    mscale = np.abs(uo[:, 1:]).max()
    assert np.abs(uo[:, 1:] - ug[:, 1:]).max() <= vtol * max(mscale, 1e-300), what


def run(w, steps, vtol=1e-11, check_every=None):
    o, g = make_pair(w)
    say("%s: set-up" % w.name)
    compare(o, g, w, vtol=1e-15, what="set-up")
    d0 = g.diag()
    for s in range(steps):
        o.step(w.dt, w.src)
        g.step(w.dt, w.src)
        if check_every and (s + 1) % check_every == 0:
            g.check_device_error()
            compare(o, g, w, vtol=vtol, what="step %d" % (s + 1))
    g.check_device_error()
    compare(o, g, w, vtol=vtol, what="end")
    d1 = g.diag()
    drift = np.abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() / d0[8]
    say("%s: %d steps, momentum drift %.3e" % (w.name, steps, drift))
    # (no invariant in a shearing box, which has external forces, or when gas flows through
    # open x3 faces)
    open_faces = (w.src is not None and w.src.hydro and w.mesh.bc_x3 != abi.DK_X3_PERIODIC)
    if not w.phys.shearing_box and not open_faces:
        assert drift < 1e-13
    return o, g


def test_setup_on_device_equals_oracle():
    for w in (workloads.c1(n=8, nb=2, ppc=2), workloads.c5(n=16, nbx=2, ppc=4),
              workloads.c3(n=16, nbx=2)):
        say("set-up %s" % w.name)
        o, g = make_pair(w)
        compare(o, g, w, vtol=1e-15, what="set-up " + w.name)


def test_c1_plain_fp64_100_steps():
    """The C1 workload itself: 32^3, one block, ppc 1, imex2+/dc1, dt 1e-3, 100 steps."""
    run(workloads.c1(), 100, check_every=25)


@pytest.mark.parametrize("solver", [abi.DK_DC1, abi.DK_LOCAL, abi.DK_DC2, abi.DK_APPLYA])
def test_imex_particles_crossing_blocks(solver):
    w = workloads.c1(n=16, nb=2, ppc=2, solver=solver)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    w.name = "imex-cross-%d" % solver
    run(w, 20, check_every=5)


@pytest.mark.parametrize("solver,params", [
    (abi.DK_PCG, {}),
    (abi.DK_ADAPTIVE, {}),                                   # the default target: PCG fallback
    (abi.DK_ADAPTIVE, {"order_c": 1e12, "rtol_max": 1.0})])  # a loose target: DC1 accepted
def test_strict_solver_matches_oracle(solver, params):
    """StrictPCG and the adaptive path (dust_solver.cpp:268-625) against the oracle: the same
    path every stage (fast accept or PCG, iterations within one: the reductions sum in another
    order) and the same state to 1e-11 after 10 steps of particles crossing blocks."""
    w = workloads.c1(n=16, nb=2, ppc=2, solver=solver)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    w.name = "strict-%d" % solver
    o, g = make_pair(w)
    o.set_solver(**params)
    g.set_solver(**params)
    for s in range(10):
        o.step(w.dt, w.src)
        g.step(w.dt, w.src)
        so, sg = o.solver_stats(), g.solver_stats()
        say("step %d: oracle %s, gpu %s" % (s, so[:2], sg[:2]))
        assert so[1] == sg[1] and abs(so[0] - sg[0]) <= 1, (so, sg)
    g.check_device_error()
    compare(o, g, w, vtol=1e-11, what="strict solver")


def test_strict_solver_budget_is_reported():
    from athenak_amd.lib import DustError
    w = workloads.c1(n=16, nb=2, ppc=2, solver=abi.DK_PCG)
    w.setup.dust_noise, w.setup.gas_noise = 0.5, 0.3
    _, g = make_pair(w)
    g.set_solver(rtol=1e-16, atol=0.0, iter_max=1)
    with pytest.raises(DustError, match="drag solver failed"):
        g.step(w.dt, w.src)
    g.close()


def test_pc2_particles_crossing_blocks():
    w = workloads.c1(n=16, nb=2, ppc=2, coupling=abi.DK_PC2, solver=abi.DK_LOCAL)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    w.name = "pc2-cross"
    run(w, 20, check_every=5)


def test_shearing_box_small():
    w = workloads.c2(n=16, nbx=2, ppc=2)
    w.setup.dust_noise, w.dt = 0.5, 1e-3
    w.name = "c2-small"
    run(w, 20, check_every=5)


def test_stratified_x3_outflow_small():
    w = workloads.c3(n=16, nbx=2)
    w.setup.dust_noise, w.dt = 0.3, 1e-3
    w.name = "c3-small"
    o, g = make_pair(w)
    for s in range(10):
        o.step(w.dt, w.src)
        g.step(w.dt, w.src)
    g.check_device_error()
    compare(o, g, w, what="c3-small")


def test_clumped_small():
    w = workloads.c5(n=16, nbx=2, ppc=8)
    w.name = "c5-small"
    run(w, 10)


def test_sort_by_cell_is_the_oracles_stable_order():
    w = workloads.c1(n=16, nb=2, ppc=3)
    o, g = make_pair(w)
    say("sort")
    o.sort()
    g.sort()
    n = o.npart
    # the permutation is exact; the values carry the set-up's last-bit differences of the
    # velocity draws (log/cos of the host and device libraries)
    np.testing.assert_array_equal(o.idata()[:, :n], g.idata[:, :n].cpu().numpy())
    np.testing.assert_allclose(o.rdata()[:, :n], g.rdata[:, :n].cpu().numpy(), rtol=1e-15,
                               atol=1e-18)


def test_new_time_step_matches_oracle():
    from oracle.pyoracle import lib as olib
    import ctypes
    w = workloads.c1(n=8, nb=2, ppc=2)
    o, g = make_pair(w)
    out = ctypes.c_double(0.0)
    olib().or_new_time_step(o.h, 0.5, ctypes.byref(out))
    assert g.new_time_step(0.5) == out.value


def test_halo_violation_is_reported_not_fatal():
    """A particle left far outside its block (a routing bug upstream) is reported by the
    deposit's own bounds check (dust_pm.cpp:241-251) and skipped; the kernel exits normally."""
    from athenak_amd.lib import DustError
    w = workloads.c1(n=8, nb=2, ppc=1)
    o, g = make_pair(w)
    say("halo error path")
    m = int(g.idata[abi.PGID, 5].item())
    g.rdata[abi.IPX, 5] = w.blocks[m].x1max + 2.5 / 8   # 2.5 cells beyond the block face
    st = abi.imex2p_stages(0.0, 1e-3)[0]
    g.task("deposit_drag", st)
    with pytest.raises(DustError, match="halo"):
        g.check_device_error()
    g.check_device_error()   # cleared


def test_graph_replays_of_the_fp64_form_match_the_oracle():
    # the FP64 form with every stage captured into a graph and replayed: still the oracle's
    # result (its deposits use atomics, so replays and eager runs agree to round-off)
    for coupling, solver in ((abi.DK_IMEX, abi.DK_DC1), (abi.DK_PC2, abi.DK_LOCAL)):
        w = workloads.c1(n=16, nb=2, ppc=2, coupling=coupling, solver=solver)
        w.setup.dust_noise, w.setup.gas_noise, w.dt = 1.0, 0.3, 2e-3
        w.name = "c1-graphs-%s" % ("pc2" if coupling == abi.DK_PC2 else "imex")
        o, g = make_pair(w, own_stream=True)
        g.set_graphs(True)
        for s in range(6):
            o.step(w.dt, w.src)
            g.step(w.dt, w.src)
            g.check_device_error()
            g.diag()                     # eager work and a host sync between the replays
            compare(o, g, w, what="graph step %d" % s)
        assert g.graphs_captured() > 0
        say("%s: 6 steps of graph replays (%d graph launches) match the oracle"
            % (w.name, g.graphs_captured()))
        g.close()


def test_clouds_pc2_particles_crossing_blocks():
    # same-level refinable clouds (CloudPredictorT / CloudAdvanceT): linear gathers and cloud
    # weights, clouds reaching into neighbour blocks and across periodic faces
    w = workloads.c4u(n=16, nbx=2, ppc=2)
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 2.0, 0.5, 2e-3
    w.name = "c4-clouds-cross"
    run(w, 20, check_every=5)


def test_clouds_pc2_with_the_stiffness_field_and_damping():
    w = workloads.c4u(n=16, nbx=2, ppc=4)
    w.phys.stiffness, w.phys.pc2_damping = 1, 1
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 1.0, 0.3, 5e-3
    w.name = "c4-clouds-stiff"
    o, g = run(w, 10, check_every=5)
    # the predictor's rate field (qdep 4, active cells after the fold) is the oracle's
    ng = w.mesh.ng
    qo = o.field(0)[:, 4, ng:-ng, ng:-ng, ng:-ng]
    qg = g.field(0)[:, 4, ng:-ng, ng:-ng, ng:-ng]
    assert np.abs(qo).max() > 0
    np.testing.assert_allclose(qg, qo, rtol=0, atol=1e-11 * np.abs(qo).max())
