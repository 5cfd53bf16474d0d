"""Shared helpers of the tests: independent closed forms (not the oracle's code) and
particle-state utilities."""

import numpy as np

from athenak_amd import abi


def bspline2(u):
    """The quadratic B-spline (the TSC assignment function) from its textbook piecewise
    form: 3/4 - u^2 for |u| <= 1/2, (3/2 - |u|)^2 / 2 for 1/2 < |u| <= 3/2, else 0."""
    a = np.abs(u)
    return np.where(a <= 0.5, 0.75 - a * a, np.where(a <= 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def bspline1(u):
    """The hat (CIC) assignment function."""
    return np.maximum(0.0, 1.0 - np.abs(u))


def dense_gather_matrix(xp, n, L=1.0, order=2):
    """Dense periodic gather matrix G (npart x n^3) of a 1-block periodic cube of n^3 cells over
    [0, L)^3, built from the B-spline closed form at the cell-centre offsets."""
    f = bspline2 if order == 2 else bspline1
    dx = L / n
    centres = (np.arange(n) + 0.5) * dx
    G = np.zeros((xp.shape[0], n ** 3))
    for p in range(xp.shape[0]):
        w = []
        for d in range(3):
            off = (xp[p, d] - centres) / dx
            off = off - n * np.round(off / n)     # periodic images
            w.append(f(off))
        G[p] = np.einsum("k,j,i->kji", w[2], w[1], w[0]).ravel()
    return G


def active(field, ng):
    """Active cells of a (nmb, nvar, n3, n2, n1) field."""
    return field[:, :, ng:-ng, ng:-ng, ng:-ng]


def lattice_positions(n, ppc_side, L=1.0, x0=0.0):
    """ppc_side^3 particles per cell on a cubic lattice at the sub-cell centres."""
    k = (np.arange(n * ppc_side) + 0.5) * (L / (n * ppc_side)) + x0
    z, y, x = np.meshgrid(k, k, k, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def by_tag(rdata, idata, npart):
    """Particle columns reordered by PTAG (the order of the set-up)."""
    order = np.argsort(idata[abi.PTAG, :npart], kind="stable")
    return rdata[:, :npart][:, order], idata[:, :npart][:, order]
