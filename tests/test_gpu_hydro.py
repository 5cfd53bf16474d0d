"""The explicit gas update with isothermal PLM + HLLE fluxes (dk_gas_src.hydro) on the GPU,
through the C-ABI: both forms against the oracle (the FP64 form to the FP64 parity bar, the
FP32 form within 1e-5), the stage graphs against eager stages, the refused options, and the
damped dusty sound wave against its dispersion relation at resolutions beyond the oracle's
test sizes."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from test_gpu_parity import run as fp64_run, make_pair, compare
from test_gpu_fp32 import check_against_oracle, pc2_check, graph_replays
from test_oracle_hydro import dusty_wave_setup, dusty_wave_modes, hydro_src

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-hydro] " + msg, flush=True)
    sys.stdout.flush()


def crossing(coupling=abi.DK_IMEX, solver=abi.DK_DC1):
    """Particles and gas in motion across 8 blocks of 8^3, with the gas fluxes on."""
    w = workloads.with_hydro(workloads.c1(n=16, nb=2, ppc=2, coupling=coupling, solver=solver))
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 0.5, 0.2, 2e-3
    return w


def test_fp64_imex_with_fluxes_matches_the_oracle():
    w = crossing()
    w.name = "fp64-imex-hydro"
    fp64_run(w, 20, check_every=5)


def test_fp64_pc2_with_fluxes_matches_the_oracle():
    w = crossing(abi.DK_PC2, abi.DK_LOCAL)
    w.name = "fp64-pc2-hydro"
    fp64_run(w, 20, check_every=5)


@pytest.mark.parametrize("bc", [abi.DK_X3_OUTFLOW, abi.DK_X3_REFLECT])
def test_fp64_physical_x3_faces_with_fluxes_match_the_oracle(bc):
    w = crossing()
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (2, 2, 2), ((0.0, 1.0),) * 3, ng=2, bc_x3=bc)
    w.mesh, w.blocks, w.name = mesh, blocks, "fp64-x3-%d-hydro" % bc
    fp64_run(w, 10, check_every=5)


def test_fp32_reflecting_x3_with_fluxes_within_1e5_of_the_oracle():
    w = crossing()
    mesh, blocks = abi.uniform_mesh((8, 8, 8), (2, 2, 2), ((0.0, 1.0),) * 3, ng=2,
                                    bc_x3=abi.DK_X3_REFLECT)
    w.mesh, w.blocks, w.name = mesh, blocks, "fp32-reflect-hydro"
    check_against_oracle(w, 10)


def test_fp64_fluxes_on_larger_and_uneven_blocks_match_the_oracle():
    """16^3 blocks, and 24 x 20 x 8 blocks (cells per block differing by axis)."""
    w = workloads.with_hydro(workloads.c1(n=32, nb=2, ppc=1))
    w.setup.dust_noise, w.setup.gas_noise, w.dt = 0.5, 0.2, 2e-3
    w.name = "fp64-16-blocks-hydro"
    fp64_run(w, 6, check_every=3)
    w = crossing()
    mesh, blocks = abi.uniform_mesh((24, 20, 8), (2, 2, 2), ((0.0, 1.2), (0.0, 1.0), (0.0, 0.4)),
                                    ng=2)
    w.mesh, w.blocks, w.name = mesh, blocks, "fp64-uneven-blocks-hydro"
    fp64_run(w, 4, check_every=2)


def test_fp32_imex_with_fluxes_within_1e5_of_the_oracle():
    check_against_oracle(crossing(), 20)


def test_fp32_pc2_with_fluxes_within_1e5_of_the_oracle():
    pc2_check(crossing(abi.DK_PC2, abi.DK_LOCAL), 10)


def test_stage_graphs_with_fluxes_match_eager_stages():
    # FP32: each stage captured and replayed, checked bit for bit against eager stages
    graph_replays(crossing())
    # FP64: the same through the plain form's graphs, to the FP64 parity bar (its deposits are
    # atomic adds, whose order differs from run to run)
    import torch
    from athenak_amd.lib import DustContext
    w = crossing()
    ctx = []
    for own in (False, True):
        g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16,
                        stream=torch.cuda.Stream() if own else None)
        g.setup(w.setup)
        ctx.append(g)
    ctx[1].set_graphs(True)
    for step in range(4):
        for g in ctx:
            g.step(w.dt, w.src)
            g.check_device_error()
        a, b = (g.u0.cpu().numpy() for g in ctx)
        for v in range(4):
            assert np.abs(a[:, v] - b[:, v]).max() <= 1e-11 * np.abs(a[:, v]).max(), (step, v)
        ra, rb = (g.rdata[:, :g.npart].cpu().numpy() for g in ctx)
        vel = [abi.IPVX, abi.IPVY, abi.IPVZ]
        assert np.abs(ra[vel] - rb[vel]).max() <= 1e-11 * np.abs(ra[vel]).max(), step
        pos = [abi.IPX, abi.IPY, abi.IPZ]
        assert np.abs(ra[pos] - rb[pos]).max() <= 1e-12, step
    assert ctx[1].graphs_captured() > 0
    say("fp64 graph replays with fluxes within 1e-11 of eager stages")


def test_refused_options():
    from athenak_amd.lib import DustContext, DustError
    w = workloads.c2(n=16, nbx=2, ppc=1)
    src = hydro_src(1.0)
    for form in (abi.DK_FORM_FP64, abi.DK_FORM_FP32):
        g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=form)
        g.setup(w.setup)
        with pytest.raises(DustError, match="option not provided"):
            g.step(w.dt, src)
        g.close()
    w = crossing()
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16)
    g.setup(w.setup)
    with pytest.raises(DustError, match="invalid argument"):
        g.step(w.dt, hydro_src(0.0))
    g.close()


@pytest.mark.parametrize("form", [abi.DK_FORM_FP64, abi.DK_FORM_FP32])
@pytest.mark.parametrize("coupling", [abi.DK_IMEX, abi.DK_PC2])
def test_dusty_wave_at_256_cells_follows_the_dispersion_relation(form, coupling):
    """The eigenmode of the damped dusty sound wave (eps 1, tau_s 0.1, k = 2 pi, cs 1) at 256
    cells per wavelength to t = 0.5: the damping exp(Im(omega) t) = 0.78 and the phase within
    the second-order errors the oracle shows at 32 and 64 cells (divided by 16 and 4)."""
    from athenak_amd.lib import DustContext
    nx, T = 256, 0.5
    o, w, du, dv = dusty_wave_setup(nx, coupling)
    n = o.npart
    g = DustContext(o.mesh, o.blocks, o.phys, n + 16, form=form)
    g.load(o.rdata()[:, :n].copy(), o.idata()[:, :n].copy(), o.u0().copy(), o.u1().copy())
    src = hydro_src(1.0)
    nst = int(np.ceil(T / (0.3 / nx)))
    for _ in range(nst):
        g.step(T / nst, src, check=False)
    g.check_device_error()
    if form == abi.DK_FORM_FP32:
        g.export()
    else:
        g.stream.synchronize()
    o.rdata()[:, :n] = g.rdata[:, :n].cpu().numpy()
    o.u0()[:] = g.u0.cpu().numpy()
    cg, cd = dusty_wave_modes(o, nx)
    eg, ed = du * np.exp(-1j * w * T), dv * np.exp(-1j * w * T)
    errs = (abs(abs(cg) / abs(eg) - 1.0), abs(abs(cd) / abs(ed) - 1.0),
            abs(np.angle(cg / eg)), abs(np.angle(cd / ed)))
    say("dusty wave form %d coupling %d: |gas| %.2e |dust| %.2e phase gas %.2e dust %.2e"
        % ((form, coupling) + errs))
    assert errs[0] < 1e-3 and errs[1] < 1e-3
    assert errs[2] < 6e-4 and errs[3] < 6e-4
    g.close()
