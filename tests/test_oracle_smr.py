"""Pins the oracle's refinable clouds on a statically refined mesh (the uneven-block case) against
what the mathematics fixes, no GPU.

- the cloud of level L is the tent of half-width one level-L cell (the derivative of the hat
  CDF, src/dust/dust.hpp:169-182): its weight over a leaf cell of any level is the integral of
  the tent over that cell, and its first moment the integral of (s - x) times the tent.  Both
  are computed here by Gauss-Legendre quadrature on the pieces between the tent's kinks and
  the cell faces (exact for the linear pieces), not from the hat-CDF differences the oracle
  takes (src/dust/dust_leaf3d.hpp:383-517);
- the linear gather returns a linear field's value exactly for clouds that straddle
  coarse-fine faces (the weights add to one, the moments to zero and the central differences of
  every level are exact), so a wrong cell centre, level or candidate block fails it;
- the routing puts every particle in the leaf whose box holds it (checked against an
  independent search over all blocks) as particles drift across the level faces;
- pc2 with clouds on the refined mesh conserves total momentum to round-off while they cross
  levels, and with a uniform gas and no back-reaction the clouds move as point particles;
- the uneven block counts: the host draw in Python equals the oracle's, the mean is ppc.
"""

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle

# root 2x2x4 blocks of 8^3 over [0,1)x[0,1)x[0,2) (root cell 1/16), level 1 over the middle half
# in x3, level 2 inside it: 8 root, 56 level-1 and 64 level-2 leaves
REGIONS = (((0.0, 2.0), (0.0, 2.0), (1.0, 3.0)), ((0.5, 1.5), (0.5, 1.5), (1.5, 2.5)))
BOX = ((0.0, 1.0), (0.0, 1.0), (0.0, 2.0))


def smr(nx=8, back_reaction=1, cap=200000):
    mesh, blocks = abi.smr_mesh((nx, nx, nx), (2, 2, 4), BOX, REGIONS)
    phys = workloads._phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL, clouds=1,
                           back_reaction=back_reaction)
    return mesh, blocks, phys, Oracle(mesh, blocks, phys, cap)


def tent_integrals(x, h, a, b):
    """(integral of the tent T((s - x)/h)/h over [a, b], integral of (s - x) times it): two-point
    Gauss-Legendre on every piece between the kinks x - h, x, x + h and the faces."""
    pts = sorted({a, b} | {p for p in (x - h, x, x + h) if a < p < b})
    gl = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    w0 = w1 = 0.0
    for lo, hi in zip(pts[:-1], pts[1:]):
        s = 0.5 * (lo + hi) + 0.5 * (hi - lo) * gl
        f = np.maximum(0.0, 1.0 - np.abs(s - x) / h) / h
        w0 += 0.5 * (hi - lo) * f.sum()
        w1 += 0.5 * (hi - lo) * ((s - x) * f).sum()
    return w0, w1


def leaf_cell_box(blocks, mesh, b, cell):
    n1 = mesh.nx1 + 2 * mesh.ng
    n2 = mesh.nx2 + 2 * mesh.ng
    i = cell % n1 - mesh.ng
    j = (cell // n1) % n2 - mesh.ng
    k = cell // (n1 * n2) - mesh.ng
    B = blocks[b]
    out = []
    for lo, hi, n, idx in ((B.x1min, B.x1max, mesh.nx1, i), (B.x2min, B.x2max, mesh.nx2, j),
                           (B.x3min, B.x3max, mesh.nx3, k)):
        d = (hi - lo) / n
        out.append((lo + idx * d, lo + (idx + 1) * d))
    return out


def sample_points(rng, n):
    """Positions near the level faces in x3 (z = 0.5 and 1.5 between levels 0 and 1; z = 0.75
    and 1.25 between 1 and 2) and the level-2 faces in x1, x2 (0.25, 0.75), inside the mesh."""
    faces_z = (0.5, 0.75, 1.25, 1.5)
    out = []
    for t in range(n):
        z = faces_z[t % 4] + rng.uniform(-0.06, 0.06)
        x = (0.25 if t % 3 == 0 else 0.75) + rng.uniform(-0.06, 0.06)
        y = rng.uniform(0.1, 0.9)
        out.append((x, y, z))
    return out


def test_leaf_weights_and_moments_are_the_tent_integrals():
    mesh, blocks, _, o = smr()
    rng = np.random.default_rng(5)
    ncells_multi = 0
    for x in sample_points(rng, 60):
        b0 = o.leaf_at(x)
        assert b0 >= 0
        lev0 = blocks[b0].level
        for lev in {lev0, min(lev0 + 1, 2)}:
            h = 1.0 / 16 / (1 << lev)
            blk, cell, w, m1 = o.leaf_cells(x, lev)
            assert abs(w.sum() - 1.0) < 1e-13
            np.testing.assert_allclose(m1.sum(axis=0), 0.0, atol=1e-15)
            if len({blocks[b].level for b in blk}) > 1:
                ncells_multi += 1
            for q in range(len(w)):
                box = leaf_cell_box(blocks, mesh, blk[q], cell[q])
                wd = [tent_integrals(x[d], h, *box[d]) for d in range(3)]
                want_w = wd[0][0] * wd[1][0] * wd[2][0]
                want_m = (wd[0][1] * wd[1][0] * wd[2][0], wd[0][0] * wd[1][1] * wd[2][0],
                          wd[0][0] * wd[1][0] * wd[2][1])
                assert abs(w[q] - want_w) < 1e-14
                np.testing.assert_allclose(m1[q], want_m, rtol=0, atol=1e-16)
    # the sample straddles level faces
    assert ncells_multi > 20


def cell_centres(mesh, blocks):
    """(nmb, 3, n3, n2, n1) cell centres of every block, ghosts included (unwrapped)."""
    n = (mesh.nx3 + 2 * mesh.ng, mesh.nx2 + 2 * mesh.ng, mesh.nx1 + 2 * mesh.ng)
    out = np.zeros((mesh.nmb, 3) + n)
    for m in range(mesh.nmb):
        B = blocks[m]
        for d, (lo, hi, nx) in enumerate(((B.x1min, B.x1max, mesh.nx1),
                                          (B.x2min, B.x2max, mesh.nx2),
                                          (B.x3min, B.x3max, mesh.nx3))):
            c = lo + (np.arange(nx + 2 * mesh.ng) - mesh.ng + 0.5) * (hi - lo) / nx
            shape = [1, 1, 1]
            shape[2 - d] = -1
            out[m, d] = np.broadcast_to(c.reshape(shape), n)
    return out


def test_linear_gather_is_exact_across_levels():
    mesh, blocks, _, o = smr()
    xc = cell_centres(mesh, blocks)
    A = np.array([0.3, -0.2, 0.1])
    G = np.array([[1.0, 2.0, -0.5], [0.5, -1.0, 0.25], [-2.0, 0.75, 1.5]])
    vel = A[None, :, None, None, None] + np.einsum("cd,mdkji->mckji", G, xc)
    rng = np.random.default_rng(6)
    for x in sample_points(rng, 40):
        lev = blocks[o.leaf_at(x)].level
        got = o.leaf_gather(vel, x, lev)
        want = A + G @ np.array(x)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)


def test_block_counts_and_levels_of_the_uneven_setup():
    mesh, blocks, _, o = smr()
    su = workloads._setup(seed=3, ppc=2, taus=(0.05,), eps=(1.0,), block_lognormal=1,
                          block_sigma=1.5)
    counts = workloads.block_counts(mesh, su)
    o.setup(su)
    assert o.npart == sum(counts)
    gid = o.idata()[abi.PGID, :o.npart]
    np.testing.assert_array_equal(np.bincount(gid, minlength=mesh.nmb), counts)
    # uneven: lognormal of sigma 1.5
    assert max(counts) > 4 * min(max(c, 1) for c in counts)
    # every cloud carries its block's level; masses scale with the level's cell volume
    lev = o.idata()[abi.PLEV, :o.npart]
    np.testing.assert_array_equal(lev, [blocks[g].level for g in gid])
    mass = o.rdata()[abi.IPM, :o.npart]
    np.testing.assert_allclose(mass * 8.0 ** lev, mass[lev == 0][0], rtol=1e-14)
    # the mean count per cell is ppc (20 % for 128 lognormal draws of sigma 1.5)
    assert abs(np.mean(counts) / (2 * 8 ** 3) - 1.0) < 0.5


def drifting(o, su_kw=None):
    su = workloads._setup(seed=4, ppc=1, taus=(0.05,), eps=(1.0,), gas_noise=1e-2,
                          dust_noise=1e-2, gas_vz=0.6, gas_vx=0.3, block_lognormal=1,
                          block_sigma=0.8, **(su_kw or {}))
    o.setup(su)
    return su


def in_own_block(o, blocks):
    r, i = o.rdata(), o.idata()
    n = o.npart
    g = i[abi.PGID, :n]
    lo = np.array([[blocks[b].x1min, blocks[b].x2min, blocks[b].x3min] for b in g])
    hi = np.array([[blocks[b].x1max, blocks[b].x2max, blocks[b].x3max] for b in g])
    x = np.stack([r[abi.IPX, :n], r[abi.IPY, :n], r[abi.IPZ, :n]], axis=1)
    return np.all((x >= lo) & (x < hi), axis=1)


def test_pc2_on_the_refined_mesh_conserves_momentum_and_routes_across_levels():
    mesh, blocks, _, o = smr()
    drifting(o)
    levels0 = np.array([blocks[b].level for b in o.idata()[abi.PGID, :o.npart]])
    d0 = o.diag()
    for _ in range(12):
        o.step(0.01)
    d1 = o.diag()
    # every particle lies in its block after the last stage-2 routing (the advance moved it)
    o.task("update_particle_gids", abi.rk2_stages(o.time, 0.01)[1])
    assert in_own_block(o, blocks).all()
    levels1 = np.array([blocks[b].level for b in o.idata()[abi.PGID, :o.npart]])
    assert (levels1 != levels0).sum() > 100   # crossed level faces
    p0 = d0[0:3] + d0[3:6]
    p1 = d1[0:3] + d1[3:6]
    assert np.abs(p1 - p0).max() / d0[8] < 1e-13


def test_clouds_without_back_reaction_in_a_uniform_gas_move_as_points():
    """A uniform gas velocity is gathered exactly by every cloud, so without back-reaction the
    pc2 update of each particle is the point-particle update with that gas velocity; checked
    against the implicit-midpoint formula evaluated here."""
    mesh, blocks, _, o = smr(back_reaction=0)
    su = workloads._setup(seed=4, ppc=1, taus=(0.05,), eps=(1.0,), gas_vx=0.2, gas_vy=-0.1,
                          gas_vz=0.6, dust_noise=0.3, block_lognormal=1, block_sigma=0.8)
    o.setup(su)
    n = o.npart
    r = o.rdata()
    v0 = np.stack([r[abi.IPVX, :n], r[abi.IPVY, :n], r[abi.IPVZ, :n]]).copy()
    x0 = np.stack([r[abi.IPX, :n], r[abi.IPY, :n], r[abi.IPZ, :n]]).copy()
    dt, ts = 0.01, 0.05
    o.step(dt)
    u = np.array([0.2, -0.1, 0.6])[:, None]
    lam = 0.5 * dt / ts
    vh = (v0 + lam * u) / (1.0 + lam)
    r = o.rdata()
    v1 = np.stack([r[abi.IPVX, :n], r[abi.IPVY, :n], r[abi.IPVZ, :n]])
    np.testing.assert_allclose(v1, 2.0 * vh - v0, rtol=0, atol=1e-14)
    x1 = np.stack([r[abi.IPX, :n], r[abi.IPY, :n], r[abi.IPZ, :n]])
    want = x0 + dt * vh
    L = np.array([1.0, 1.0, 2.0])[:, None]
    dx = (x1 - want) - L * np.round((x1 - want) / L)
    assert np.abs(dx).max() < 1e-13


def test_refined_meshes_need_clouds_and_a_full_tiling():
    mesh, blocks = abi.smr_mesh((8, 8, 8), (2, 2, 4), BOX, REGIONS)
    with pytest.raises(ValueError):   # point particles: no cross-level point deposit
        Oracle(mesh, blocks, workloads._phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL), 16)
    blocks[3].lx1 = blocks[2].lx1     # two leaves on one slot, one slot empty
    with pytest.raises(ValueError):
        Oracle(mesh, blocks, workloads._phys(coupling=abi.DK_PC2, solver=abi.DK_LOCAL,
                                             clouds=1), 16)
