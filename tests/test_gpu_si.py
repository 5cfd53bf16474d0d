"""The streaming instability of an unstratified shearing box (Youdin & Johansen 2007, linA:
t_s = 0.1, eps = 3, K_x = K_z = 30) through the whole coupled stage: the isothermal PLM + HLLE
gas update with the shearing-box and radial-forcing source terms, imex2+/dc1 drag with
back-reaction, particles on a lattice displaced by the eigenmode.  The mode must grow at the
rate of the linear system (tests/si_linear.py, 0.4168 Omega for eta v_K = c_s/2) up to the
gas update's numerical dissipation, which falls with resolution: measured 14 % below at 32
cells per wavelength and 2 % below at 64.  The box is one wavelength in x1 and x3, 4 cells in
x2 (axisymmetric: periodic x1 is the r-z box)."""

import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
import si_linear

pytestmark = pytest.mark.gpu

ETAVK, TS, EPS, K = 0.5, 0.1, 3.0, 30.0


def say(msg):
    print("[gpu-si] " + msg, flush=True)
    sys.stdout.flush()


def lina_growth(n, form, T=4.0, sub=4, ny=4, ng=2):
    """(measured growth rate, linear theory) of the mode at n cells per wavelength."""
    import torch
    from athenak_amd.lib import DustContext
    lam, X = si_linear.growing_mode(K, ETAVK, TS, EPS)
    kx = kz = K / ETAVK
    lx = 2.0 * np.pi / kx
    dx = lx / n
    mesh, blocks = abi.uniform_mesh((n, ny, n), (1, 1, 1),
                                    ((0.0, lx), (0.0, ny * dx), (0.0, lx)), ng=ng)
    phys = workloads._phys(sbox=True, omega0=si_linear.OMEGA, qshear=si_linear.Q)
    ax, (ugx, ugy), (vx, vy) = si_linear.equilibrium(ETAVK, TS, EPS)
    X = X * (1e-5 / abs(X[1]))                   # gas |du_x| = 1e-5 c_s
    # particles: sub x 1 x sub per cell on a lattice displaced by xi, drho_d/rho_d = -div xi
    xs = (np.arange(n * sub) + 0.5) * (lx / (n * sub))
    ys = (np.arange(ny) + 0.5) * dx
    Xp, Yp, Zp = np.meshgrid(xs, ys, xs, indexing="ij")
    x0, y0, z0 = Xp.ravel(), Yp.ravel(), Zp.ravel()
    ph = np.exp(1j * (kx * x0 + kz * z0))
    xi = 1j * X[4] / (kx * kx + kz * kz)
    npart = x0.size
    r = np.zeros((abi.DK_NRDATA, npart))
    iv = np.zeros((abi.DK_NIDATA, npart), np.int32)
    r[abi.IPX] = (x0 + (xi * kx * ph).real) % lx
    r[abi.IPY] = y0
    r[abi.IPZ] = (z0 + (xi * kz * ph).real) % lx
    r[abi.IPVX] = vx + (X[5] * ph).real
    r[abi.IPVY] = vy + (X[6] * ph).real
    r[abi.IPVZ] = (X[7] * ph).real
    for c in range(6):
        r[abi.IPX1 + c] = r[abi.IPX + c]
    r[abi.IPTS] = TS
    r[abi.IPM] = EPS * dx ** 3 / (sub * sub)
    iv[abi.PTAG] = np.arange(npart)
    # the gas at the cell centres, ghosts included (periodic)
    cc = (np.arange(-ng, n + ng) + 0.5) * dx
    Xc, Zc = np.meshgrid(cc, cc, indexing="ij")
    pg = np.exp(1j * (kx * Xc + kz * Zc))
    rho = 1.0 + (X[0] * pg).real
    u = [ugx + (X[1] * pg).real, ugy + (X[2] * pg).real, (X[3] * pg).real]
    u0 = np.zeros((1, 4, n + 2 * ng, ny + 2 * ng, n + 2 * ng))
    u0[0, 0] = rho.T[:, None, :]
    for d in range(3):
        u0[0, 1 + d] = (rho * u[d]).T[:, None, :]
    src = abi.dk_gas_src(shearing_box=1, omega0=si_linear.OMEGA, qshear=si_linear.Q,
                         const_accel=ax, const_accel_dir=1, hydro=1, iso_cs=si_linear.CS,
                         dfloor=workloads.FLT_MIN)
    g = DustContext(mesh, blocks, phys, npart + 16, form=form, stream=torch.cuda.Stream())
    g.load(r, iv, u0)
    g.set_graphs(True)
    nst = int(np.ceil(T / (0.3 * dx / (si_linear.CS + 2.0 * ETAVK))))
    dt = T / nst
    every = max(1, nst // 16)
    act = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    cz = (np.arange(n) + 0.5) * dx
    mode = np.exp(-1j * (kx * cz[None, None, :] + kz * cz[:, None, None]))   # [k, j, i]
    times, amp = [], []
    for s in range(nst + 1):
        if s % every == 0 or s == nst:
            g.stream.synchronize()
            ua = g.u0.cpu().numpy()[0][act]
            times.append(s * dt)
            amp.append(abs(((ua[3] / ua[0]) * mode).mean()))
        if s < nst:
            g.step(dt, src, check=False)
    g.check_device_error()
    g.close()
    t, a = np.array(times), np.log(np.array(amp))
    fit = t >= 1.0
    rate = np.polyfit(t[fit], a[fit], 1)[0]
    # (the growth is a clean exponential: the residual of the fit is small)
    resid = np.abs(a[fit] - np.polyval(np.polyfit(t[fit], a[fit], 1), t[fit])).max()
    say("n %d form %d: growth %.4f (linear theory %.4f), fit residual %.1e, %d steps"
        % (n, form, rate, lam.real, resid, nst))
    assert resid < 0.02
    return rate, lam.real


@pytest.mark.parametrize("form", [abi.DK_FORM_FP64, abi.DK_FORM_FP32])
def test_lina_grows_at_32_cells(form):
    rate, theory = lina_growth(32, form)
    assert 0.80 * theory < rate < theory


def test_lina_converges_at_64_cells():
    rate, theory = lina_growth(64, abi.DK_FORM_FP32)
    assert 0.96 * theory < rate < theory
