"""The science outputs through the C-ABI (dk_write_pbin, dk_particle_history, dk_write_phst), read
back with athenak_amd.science_io: the pbin columns equal the particles the library exports (the
active rows, in order), its constants follow the reference's rules, the block table is the
mesh; the history row's sums equal sequential sums over chunks of 4096 rows, its rho_max the
TSC deposit from the B-spline closed form, its stiffness records the cycles' own records."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from athenak_amd import science_io as sio
from helpers import bspline2

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-outputs] " + msg, flush=True)
    sys.stdout.flush()


def ctx(w, form):
    from athenak_amd.lib import DustContext
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=form)
    g.setup(w.setup)
    return g


def exported(g):
    g.export()
    n = g.npart
    return g.rdata[:, :n].cpu().numpy(), g.idata[:, :n].cpu().numpy()


@pytest.mark.parametrize("form", [abi.DK_FORM_FP64, abi.DK_FORM_FP32])
def test_pbin_holds_the_exported_particles(tmp_path, form):
    w = workloads.c5(n=32, nbx=2, ppc=4)
    g = ctx(w, form)
    for _ in range(3):
        g.step(w.dt, w.src)
    info = g.output_info(time=3 * w.dt, dt=w.dt, cycle=3, file_number=1, taus=(0.01, 1.0),
                         input_dump="<job>\nbasename = c5\n")
    path = str(tmp_path / "c5.00001.pbin")
    say("form %d: writing %s" % (form, path))
    g.write_pbin(path, info)
    r, i = exported(g)
    d = sio.read_pbin(path)
    act = r[abi.IPM] != 0.0
    P = d["particles"]
    assert d["preamble"]["n_particles"] == act.sum()
    tag = (i[abi.PTAGH].astype(np.int64).astype(np.uint64) << np.uint64(32)) | \
        i[abi.PTAG].astype(np.uint32).astype(np.uint64)
    np.testing.assert_array_equal(P["id"], tag[act])
    for name, col in (("x1", abi.IPX), ("x2", abi.IPY), ("x3", abi.IPZ), ("v1", abi.IPVX),
                      ("v2", abi.IPVY), ("v3", abi.IPVZ), ("stopping_time", abi.IPTS),
                      ("mass", abi.IPM)):
        np.testing.assert_array_equal(P[name], r[col][act], err_msg=name)
    np.testing.assert_array_equal(P["species"], i[abi.PSP][act])
    B = d["blocks"]
    np.testing.assert_array_equal(B["gid"], np.arange(w.mesh.nmb))
    np.testing.assert_array_equal(B["x1min"], [w.blocks[m].x1min for m in range(w.mesh.nmb)])
    np.testing.assert_array_equal(B["lx3"], [w.blocks[m].lx3 for m in range(w.mesh.nmb)])
    h = d["header"]
    assert h["representation"]["representation"] == "tsc_owner"
    assert int(h["mesh"]["nx1"]) == 32 and int(h["species"]["nspecies"]) == 2
    assert float(h["snapshot"]["time"]) == 3 * w.dt and h["snapshot"]["cycle"] == "3"
    assert int(h["snapshot"]["block_topology_hash"], 16) == sio.topology_hash(
        B["level"], B["lx1"], B["lx2"], B["lx3"])
    # two species keep species and t_s as columns; their particles share one mass
    equal = bool((r[abi.IPM][act] == r[abi.IPM][act][0]).all())
    assert h["particles"]["constant_columns"] == ("mass" if equal else "none")
    assert d["input"] == "<job>\nbasename = c5\n"
    g.close()


def test_pbin_constants_single_precision_and_parked_rows(tmp_path):
    # one species, species_fixed, equal masses: species, stopping time and mass are constants
    # (FP64 form: the FP32 form's store holds t_s in FP32, which then differs from tau_s)
    w = workloads.c1(n=16, nb=2, ppc=2)
    g = ctx(w, abi.DK_FORM_FP64)
    g.step(w.dt, w.src)
    path = str(tmp_path / "c1.pbin")
    g.write_pbin(path, g.output_info(time=w.dt, dt=w.dt, taus=(0.1,), single=True))
    r, _ = exported(g)
    d = sio.read_pbin(path)
    assert d["header"]["particles"]["constant_columns"] == "species stopping_time mass"
    assert float(d["particles"]["stopping_time"][0]) == 0.1
    assert float(d["particles"]["mass"][0]) == r[abi.IPM][0]
    assert d["particles"]["x1"].dtype == np.float32
    np.testing.assert_array_equal(d["particles"]["v2"], r[abi.IPVY].astype(np.float32))
    g.close()
    # an x3 outflow box: the particles parked at the faces (zero mass) are left out
    w = workloads.c3(n=16, nbx=2)
    w.setup.dust_noise, w.dt = 0.4, 0.02     # (a step moves the fastest ~0.05: half a block)
    g = ctx(w, abi.DK_FORM_FP64)
    for _ in range(10):
        g.step(w.dt, w.src)
    r, _ = exported(g)
    nzero = int((r[abi.IPM] == 0.0).sum())
    path = str(tmp_path / "c3.pbin")
    g.write_pbin(path, g.output_info(time=0.0, dt=w.dt, taus=tuple(w.setup.taus[:7]),
                                     stopping_time_mode=1))
    d = sio.read_pbin(path)
    say("c3 small: %d parked rows" % nzero)
    assert int(d["header"]["particles"]["zero_mass_rows_excluded"]) == nzero
    assert d["preamble"]["n_particles"] == r.shape[1] - nzero
    assert d["header"]["frame"]["velocity_frame"] == "shear_relative"
    g.close()


def tsc_density(r, mesh, blk):
    """The TSC mass density of one periodic block from the quadratic B-spline closed form."""
    n = mesh.nx1
    L = blk.x1max - blk.x1min
    dx = L / n
    rho = np.zeros((n, n, n))
    c = (np.arange(n) + 0.5) * dx
    for p in range(r.shape[1]):
        m = r[abi.IPM, p]
        if m == 0.0:
            continue
        w = []
        for col, lo in ((abi.IPX, blk.x1min), (abi.IPY, blk.x2min), (abi.IPZ, blk.x3min)):
            u = (r[col, p] - lo - c) / dx
            u = (u + n / 2) % n - n / 2            # periodic images
            w.append(bspline2(u))
        rho += m / dx ** 3 * np.einsum("k,j,i->kji", w[2], w[1], w[0])
    return rho


@pytest.mark.parametrize("form", [abi.DK_FORM_FP64, abi.DK_FORM_FP32])
def test_history_row(tmp_path, form):
    w = workloads.c1(n=16, nb=1, ppc=2)
    w.setup.dust_noise = 0.3
    g = ctx(w, form)
    for _ in range(2):
        g.step(w.dt, w.src)
    row = g.particle_history(2 * w.dt)
    r, _ = exported(g)
    n = r.shape[1]
    m = r[abi.IPM]
    a = np.zeros(7)
    for c0 in range(0, n, 4096):
        s = slice(c0, min(n, c0 + 4096))
        mm, v1, v2, v3, z = m[s], r[abi.IPVX, s], r[abi.IPVY, s], r[abi.IPVZ, s], r[abi.IPZ, s]
        k = mm != 0.0
        terms = [np.ones(k.sum()), mm[k], (mm * v1)[k], (mm * v2)[k], (mm * v3)[k],
                 (0.5 * mm * (v1 * v1 + v2 * v2 + v3 * v3))[k], (mm * z * z)[k]]
        a += [np.cumsum(t)[-1] if t.size else 0.0 for t in terms]   # sequential sums
    np.testing.assert_array_equal(row[0:6], a[0:6])
    assert row[6] == np.sqrt(a[6] / a[1])
    rho = tsc_density(r, w.mesh, w.blocks[0])
    top = np.sort(rho.ravel())[-2:]
    assert abs(row[7] - top[1]) <= 1e-12 * top[1]
    if top[1] - top[0] > 1e-9 * top[1]:
        k, j, i = np.unravel_index(np.argmax(rho), rho.shape)
        dx = 1.0 / w.mesh.nx1
        np.testing.assert_allclose(row[8:11], [(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx],
                                   rtol=0, atol=1e-15)
        assert row[11] == 0.0
    # imex runs record no pc2 stiffness
    assert row[12] == -1.0 and row[13] == -1.0
    path = str(tmp_path / "c1.phst")
    info = g.output_info(time=2 * w.dt, dt=w.dt)
    g.write_phst(path, info, row, header=True)
    g.step(w.dt, w.src)
    row2 = g.particle_history(3 * w.dt)
    info.time = 3 * w.dt
    g.write_phst(path, info, row2)
    h = sio.read_phst(path)
    np.testing.assert_array_equal(h["time"], [2 * w.dt, 3 * w.dt])
    np.testing.assert_array_equal(h["rho_max"], [row[7], row2[7]])
    np.testing.assert_array_equal(h["mom2"], [row[3], row2[3]])
    assert list(h)[:16] == list(sio.PHST_COLUMNS)
    g.close()


def test_history_stiffness_interval_records():
    w = workloads.c4p(n=32, nbx=2, ppc=4)
    w.phys.stiffness = 1
    g = ctx(w, abi.DK_FORM_FP64)
    chi, rate = [], []
    for _ in range(3):
        g.step(w.dt, w.src)
        s = g.stiffness()
        chi.append(s[0])
        rate.append(s[2])
    row = g.particle_history(3 * w.dt, stiff_thresh=0.5)
    assert row[12] == max(chi)
    assert row[13] == 0.5 / max(rate)
    # the records restart after a row: none until the next cycle
    row = g.particle_history(3 * w.dt, stiff_thresh=0.5)
    assert row[12] == -1.0 and row[13] == -1.0
    g.step(w.dt, w.src)
    s = g.stiffness()
    row = g.particle_history(4 * w.dt, stiff_thresh=0.0)
    assert row[12] == s[0] and row[13] == -1.0
    g.close()


@pytest.mark.parametrize("form", [abi.DK_FORM_FP64, abi.DK_FORM_FP32])
def test_history_rho_max_of_clouds_on_a_uniform_mesh(form):
    # level-0 clouds: the hat's exact cell integrals are the TSC weights
    w = workloads.c4u(n=16, nbx=1, ppc=2)
    g = ctx(w, form)
    g.step(w.dt, w.src)
    row = g.particle_history(w.dt)
    r, _ = exported(g)
    rho = tsc_density(r, w.mesh, w.blocks[0])
    say("clouds form %d: rho_max %.12e (B-spline closed form %.12e)" % (form, row[7], rho.max()))
    assert abs(row[7] - rho.max()) <= 1e-12 * rho.max()
    g.close()
