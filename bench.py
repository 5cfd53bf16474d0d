#!/usr/bin/env python3
"""Throughput of the dust drag coupling on one MI355X.

A step is one imex2+ cycle (two working stages of push, deposit, dc1 solve, gather-kick and
back-reaction apply, plus the harness's gas source-term update) over every particle of the
workload, with all inputs resident in HBM.  The default workload is the stratified shearing box
with a continuous size distribution: 256^3 cells in 64 blocks of 64^3, 7 log-uniform Stokes bins
over [1e-3, 1] with one particle per bin per cell (117,440,512 particles), x3 outflow faces,
imex2+/dc1.  `--workload c2`: the unstratified shearing box at ppc 4 (67,108,864 particles).
`--workload c4p`: a step is one rk2 cycle of the pc2 coupling (predictor, advance) on 128^3
cells in 512 blocks of 16^3 with clumps (8,388,608 point particles, stiff cells damped).
`--workload c1h`: 256^3 periodic in 64 blocks, ppc 1, imex2+/dc1, with the explicit gas update's
isothermal PLM + HLLE fluxes (the JSON line adds the gas update's own roofline).

Prints one JSON line: value = Mparticle-steps/s of the measured form, the roofline of the
dominant kernel (algorithmic bytes per launch / measured average launch time, HIP events on the
library's stream), the whole-step roofline (algorithmic bytes per particle-step), and the CPU
baseline (the oracle's OpenMP restatement of the reference kernels on the host cores, on a
bounded sample of the same workload).
"""

import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of the minimal-pass schedule per particle-step (ppc = particles per cell):
# imex2+/dc1 point TSC, FP64: 648 + 1176/ppc; FP32 form (positions and u0 in FP64): 456 + 732/ppc;
# pc2 point TSC (two particle passes per cycle), FP64: 280 + 384/ppc; FP32 form: 204 + 292/ppc
# pc2 with refinable clouds (the packed 12-component gather field twice per cycle): FP64
# 280 + 760/ppc, FP32 form 204 + 480/ppc
def balg_step(form, ppc, pc2=False, clouds=False):
    if pc2 and clouds:
        return (280.0 + 760.0 / ppc) if form == "fp64" else (204.0 + 480.0 / ppc)
    if pc2:
        return (280.0 + 384.0 / ppc) if form == "fp64" else (204.0 + 292.0 / ppc)
    if form == "fp64":
        return 648.0 + 1176.0 / ppc
    return 456.0 + 732.0 / ppc


# Algorithmic bytes of one launch of a kernel: (bytes per particle, bytes per active cell).
# Each particle column or mesh value the kernel must read or write counts once.
KERNEL_BYTES = {
    # FP32 form (positions 8 B, other particle values 4 B; averaged over the two stages).
    # The key pass is the push: stage 1 reads x, v and writes x, v, x1, v1; stage 2 reads x, v,
    # x1, v1, R and writes x, v.  P1 is the deposit: reads x, v, ts, m; Q, P and the three centre
    # masses written per cell (moving the store into the new order is sort traffic, not counted).
    # P2 reads x, ts, m; u* read, result written.  P3 reads x, v, ts, m and writes v (and R at
    # stage 1); u* read, D (FP64) written.
    "key": ((36 + 72 + 84 + 36) / 2.0, 0),
    "p1_gather_deposit": (44, 7 * 4),
    "p2_operator": (32, 6 * 4),
    "p3_gather_kick": ((68 + 56) / 2.0, 3 * 4 + 3 * 8),
    # FP64 form
    "push_imex": (15 * 8 + 6 * 8, 0),          # x, v, x1, v1, R read; x, v written
    "scatter_imex": (8 * 8 + 4, 4 * 8),        # x, v, ts, m, gid read; Q, P written
    "applya_gtsg": (5 * 8 + 4, 6 * 8),         # x, ts, m, gid; field read, result written
    "gather_kick": (8 * 8 + 4 + 6 * 8, 6 * 8),  # x, v, ts, m, gid; v, R written; u*, dmom
    "consistent_dust": (5 * 8 + 4 + 12 * 8, 4 * 8),
    "center_mass": (5 * 8 + 4, 8),
    "part_update": (3 * 8 + 4, 0),
}

# pc2 (rk2) kernels, per form.  FP32: the key pass reads x, v and writes x, x1, v1 (stage 1: the
# routing of the advanced positions with x1 <- x, v1 <- v; stage 2: the half-step placement with
# x1 <- xh, v1 <- v0); the predictor reads x, v, ts, m (u = M/rho read, Q, g written per cell);
# the advance reads x, v, ts, m and writes x, v, R (u* read, D (FP64) written per cell).
# FP64: the predictor reads x, v, ts, m, gid (u1 read, qdep written); the advance reads x1, v1,
# ts, m, gid and writes x, v, x1, R (u* read, dmom written).
KERNEL_BYTES_PC2 = {
    "fp32": {"key": (36 + 60, 0), "pc2_predictor": (44, 3 * 4 + 4 * 4),
             "pc2_advance": (44 + 60, 3 * 4 + 3 * 8)},
    "fp64": {"pc2_predictor": (8 * 8 + 4, 4 * 8 + 4 * 8),
             "pc2_advance": (8 * 8 + 4 + 12 * 8, 3 * 8 + 3 * 8),
             # refinable clouds: the same particle columns plus the level; per cell the packed
             # gather field (12 values) read and the deposit written
             "cloud_predictor": (8 * 8 + 4, 12 * 8 + 4 * 8),
             "cloud_advance": (8 * 8 + 4 + 12 * 8, 12 * 8 + 3 * 8)},
}


# The explicit gas update with isothermal fluxes (dk_gas_src.hydro), both forms, per active
# cell: the divergence kernel reads u0 (4 values; its 13-cell stencil is served from the caches)
# and writes divf (4); the RK combination reads u0, u1 and divf and writes u0 (4 each); in the
# FP32 form it also applies stage 2's drag history (dmom, 3 values, read: 24 B averaged over
# the two stages)
GAS_UPDATE_BYTES = {"hydro_divf": 8 * 8, "gas_hook": 16 * 8}


def kernel_bytes(label, form, w):
    if w.src is not None and w.src.hydro and label in GAS_UPDATE_BYTES:
        extra = 12 if (label == "gas_hook" and form == "fp32" and w.phys.coupling == 0) else 0
        return (0, GAS_UPDATE_BYTES[label] + extra)
    if w.phys.coupling == 1 and label in KERNEL_BYTES_PC2[form]:
        return KERNEL_BYTES_PC2[form][label]
    return KERNEL_BYTES.get(label, (0, 0))


def gas_update_roofline(r, w, form):
    """The achieved bandwidth of the gas update's kernels (hydro workloads only)."""
    if w.src is None or not w.src.hydro:
        return None
    out = {}
    for label in GAS_UPDATE_BYTES:
        if label not in r["prof"]:
            continue
        tot, cnt = r["prof"][label]
        alg = kernel_bytes(label, form, w)[1] * w.ncells
        gbs = alg / (tot / cnt * 1e-3) / 1e9
        out[label] = {"avg_launch_ms": round(tot / cnt, 4), "alg_bytes_per_launch": alg,
                      "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--form", default="both", choices=["fp64", "fp32", "both"])
    ap.add_argument("--n", type=int, default=None, help="override the mesh size (tests)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-ledger", action="store_true",
                    help="skip the drag-exchange conservation run of workloads with sources")
    ap.add_argument("--no-tune", action="store_true",
                    help="FP32 form: keep the default launch parameters (no autotuner steps)")
    ap.add_argument("--tune-config", default=None, metavar="XCD,KT",
                    help="FP32 form: fix the launch parameters the autotuner would choose (XCD "
                         "dealing 0/1, key-pass workgroup size), e.g. for a profiled run")
    ap.add_argument("--prof-steps", type=int, default=5,
                    help="eager steps timed kernel by kernel after the timed steps (>= 1)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run plain device copies of this many GiB (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters) and exit")
    return ap.parse_args()


def make_workload(name, n=None):
    from athenak_amd import workloads
    f = workloads.WORKLOADS[name]
    return f() if n is None else f(n=n)


def describe(w):
    m = w.mesh
    pc2 = w.phys.coupling == 1
    solver = {0: "local", 1: "applya", 2: "dc1", 3: "dc2", 4: "pcg", 5: "adaptive"}[w.phys.solver]
    scheme = "rk2/pc2" if pc2 else "imex2+/" + solver
    if w.phys.clouds:
        scheme += " refinable clouds"
    levels = sorted({w.blocks[b].level for b in range(m.nmb)})
    mesh = "%dx%dx%d cells in %d blocks of %dx%dx%d" % (
        m.nx1 * m.nmbx1, m.nx2 * m.nmbx2, m.nx3 * m.nmbx3, m.nmb, m.nx1, m.nx2, m.nx3)
    if levels[-1] > 0:
        mesh = ("statically refined, root %dx%dx%d, levels 0-%d, %d blocks of %dx%dx%d "
                "(%d cells)" % (m.nx1 * m.nmbx1, m.nx2 * m.nmbx2, m.nx3 * m.nmbx3, levels[-1],
                                m.nmb, m.nx1, m.nx2, m.nx3, w.ncells))
    per = ("lognormal particle counts per block (sigma %g), mean ppc %d" %
           (w.setup.block_sigma, w.ppc) if w.setup.block_lognormal else "ppc %d" % w.ppc)
    return {"workload": "%s: %s, %s, %d particles, %s, dt %g" % (
                            w.name, mesh, per, w.npart, scheme, w.dt),
            "particles": w.npart, "cells": w.ncells, "ppc": w.ppc, "blocks": m.nmb,
            "shear_periodic_x1": bool(m.bc_x1), "coupling": "pc2" if pc2 else "imex",
            "solver": "local" if pc2 else solver}


def gpu_run(w, form, steps, warmup, tune_steps=0, prof_steps=5, tune=True, tune_config=None):
    """Runs one form: (FP32) the online autotuner's steps, `warmup` steps, `steps` timed steps
    (FP32: every stage one HIP graph launch), then `prof_steps` eager steps timed kernel by kernel
    with HIP events on the library's stream (events cannot enter the graph).  The FP64 form takes
    `tune_steps` extra warm-up steps so that both forms end at the same step."""
    import torch
    from athenak_amd import abi
    from athenak_amd.lib import DustContext
    fid = abi.DK_FORM_FP64 if form == "fp64" else abi.DK_FORM_FP32
    # a stream of the library's own (default priority): graph capture needs a stream other than
    # the null stream, which is torch's current stream here
    ctx = DustContext(w.mesh, w.blocks, w.phys, w.npart, form=fid, stream=torch.cuda.Stream())
    ctx.setup(w.setup)
    if form == "fp64":
        # the reference sorts point particles by cell at cycle 0 and every 50 cycles
        # (SortParticles, src/dust/dust_cloud.cpp:615-628)
        ctx.sort()
    torch.cuda.synchronize()
    d0 = ctx.diag()
    tuned = None
    if form == "fp32" and tune_config is not None:
        xcd, kt = (int(v) for v in tune_config.split(","))
        ctx.set_tune_config(xcd, kt)
        tuned = {"state": "fixed", "xcd_dealing": xcd, "key_threads": kt, "steps": 0}
    elif form == "fp32" and tune:
        # the autotuner times its candidates on these steps (no choice changes a result)
        ctx.autotune(True)
        tune_steps = 0
        while ctx.tune_config()[0] == 1 and tune_steps < 16:
            ctx.step(w.dt, w.src)
            tune_steps += 1
        st, xcd, kt = ctx.tune_config()
        tuned = {"state": {0: "fixed", 1: "tuning", 2: "tuned"}[st], "xcd_dealing": xcd,
                 "key_threads": kt, "steps": tune_steps}
    else:
        for _ in range(tune_steps):
            ctx.step(w.dt, w.src)
    # at least one warm-up step: with graphs its last one captures and instantiates every stage
    # graph, so that no capture falls into the timed steps (both forms take the same steps)
    n_warm = max(warmup, 1)
    for _ in range(n_warm - 1):
        ctx.step(w.dt, w.src)
    if w.mesh.bc_x3 == abi.DK_X3_OUTFLOW:
        # the reference removes the particles parked at an outflow face before every cycle
        # (RemoveOutflowParticles, before_timeintegrator); here once, outside the timed steps
        ctx.remove_outflow_particles()
    ctx.check_device_error()
    n_timed = ctx.npart
    stream = ctx.stream
    graphs = form == "fp32"
    ctx.set_graphs(graphs)
    ctx.step(w.dt, w.src)
    ctx.check_device_error()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        ctx.step(w.dt, w.src, check=False)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    ctx.set_graphs(False)
    ctx.check_device_error()
    # per-kernel HIP events over eager steps
    ctx.prof(True)
    ctx.prof_reset()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(prof_steps):
        ctx.step(w.dt, w.src, check=False)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_prof = e0.elapsed_time(e1)
    prof = ctx.prof_read()
    ctx.prof(False)
    ctx.check_device_error()
    d1 = ctx.diag()
    # (relative to the larger of the two momentum scales: a run may start at rest)
    drift = float(abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() /
                  max(d0[8], d1[8], 1e-300))
    ctx.export()
    n = ctx.npart
    order = torch.argsort(ctx.idata[abi.PTAG, :n])
    vel = ctx.rdata[[abi.IPVX, abi.IPVY, abi.IPVZ], :n][:, order].clone()
    tags = ctx.idata[abi.PTAG, :n][order].clone()
    ctx.close()
    del ctx
    torch.cuda.empty_cache()
    return {"ms": ms, "wall_s": wall, "prof": prof, "drift": drift, "vel": vel, "npart": n,
            "tags": tags, "n_timed": n_timed,
            "ms_prof": ms_prof, "prof_steps": prof_steps, "tune_steps": tune_steps,
            "tuned": tuned, "graphs": graphs}


def has_sources(w):
    return bool(w.phys.shearing_box or w.phys.stratified or
                (w.src is not None and (w.src.shearing_box or w.src.stratified or
                                        w.src.const_accel != 0.0)))


def exchange_drift(name, form, steps=3):
    """The drag exchange alone at the workload's full size: the same mesh (shear-periodic x1 and
    the x3 faces included), particles, stopping times, masses, gas and dt, with the momentum
    sources switched off (no shearing-box Coriolis and tidal terms, no vertical gravity, no gas
    source terms), so that total momentum must be conserved; its drift over `steps` steps
    relative to sum m|v| + sum V|M|.  None for a workload without sources (its run reports its own
    drift)."""
    import torch
    from athenak_amd import abi
    from athenak_amd.lib import DustContext
    w = make_workload(name)
    if not has_sources(w):
        return None
    w.phys.shearing_box, w.phys.stratified = 0, 0
    if w.setup.dust_noise == 0.0:
        w.setup.dust_noise = 1e-3     # (a dust velocity to exchange)
    fid = abi.DK_FORM_FP64 if form == "fp64" else abi.DK_FORM_FP32
    ctx = DustContext(w.mesh, w.blocks, w.phys, w.npart, form=fid, stream=torch.cuda.Stream())
    ctx.setup(w.setup)
    d0 = ctx.diag()
    for _ in range(steps):
        ctx.step(w.dt, None)
    ctx.check_device_error()
    d1 = ctx.diag()
    ctx.close()
    del ctx
    torch.cuda.empty_cache()
    return {"value": float(abs((d1[0:3] + d1[3:6]) - (d0[0:3] + d0[3:6])).max() /
                           max(d0[8], d1[8], 1e-300)),
            "steps": steps,
            "how": "the same mesh, particles and gas at full size with the momentum sources "
                   "off (drag exchange alone)"}


def drift_or_none(w, drift):
    """The total-momentum drift where total momentum is conserved; None where the workload has
    momentum sources (shearing-box Coriolis and tidal terms, vertical gravity, the radial
    forcing), whose change is physics, not a conservation error."""
    return None if has_sources(w) else drift


def cpu_baseline(name, budget_s):
    """The oracle (OpenMP, host cores) on the same physics and block size on a smaller mesh:
    one block of the workload's block size (the whole workload for c1)."""
    from athenak_amd import workloads
    from oracle.pyoracle import Oracle
    full = make_workload(name)
    nx = full.mesh.nx1
    if name == "c1":
        w = full
    elif nx >= 32:
        w = workloads.WORKLOADS[name](n=nx, nbx=1)
    else:
        # small blocks (c4, c4p: 16^3): the 64^3 root level in blocks of the same size
        w = workloads.WORKLOADS[name](n=64, nbx=64 // nx)
    cores = min(16, os.cpu_count() or 1)
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16, threads=cores)
    o.setup(w.setup)
    o.step(w.dt, w.src)          # warm-up
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(w.dt, w.src)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 20:
            break
    rate = w.npart * steps / el / 1e6
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(rate, 4), "unit": "Mparticle-steps/s", "cores": cores, "kind": "port",
            "cpu_model": model,
            "sample": "CPU FP64 oracle (reference kernels restated, OpenMP atomics) on %s physics "
                      "with %d blocks of %d^3 (%d particles), %d steps after 1 warm-up, %.1f s"
                      % (name, w.mesh.nmb, w.mesh.nx1, w.npart, steps, el)}


# measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a plain copy of
# known size) of the workload's FP32 run, committed under profiles/ as hbm_traffic_<name>_fp32.md.
# A timer label covers one launch of every kernel of a group (a tile kernel is two launches: its
# first items and its extra parts); labels with two groups (P3 with and without the stage-2 rate
# store) take their mean.
PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
TRAFFIC_KERNEL = {"p1_gather_deposit": [["k_p1<"]], "p2_operator": [["k_p2<"]],
                  "p3_gather_kick": [["k_p3<"], ["k_p3n<"]], "key": [["k_key<"]]}


def measured_traffic(label, w, form):
    if form != "fp32" or label not in TRAFFIC_KERNEL:
        return None
    try:
        lines = open(os.path.join(PROFILES, "hbm_traffic_%s_fp32.md" % w.name)).read().splitlines()
    except OSError:
        return None
    per_kernel = {}
    for ln in lines:
        parts = ln.split()
        if len(parts) >= 3 and parts[0] == "traffic":
            per_kernel[" ".join(parts[1:-1])] = float(parts[-1])
    vals = []
    for group in TRAFFIC_KERNEL[label]:
        hit = [v for name, v in per_kernel.items() if any(name.startswith(p) for p in group)]
        if hit:
            vals.append(sum(hit))
    return sum(vals) / len(vals) if vals else None


def kernel_roofline(r, w, form="fp32"):
    # the dominant kernel among those whose algorithmic bytes are fixed by the workload (the
    # refined mesh's cross-level walk covers a share of the particles that the run decides)
    known = {k: v for k, v in r["prof"].items() if any(kernel_bytes(k, form, w))}
    label, (tot, cnt) = max((known or r["prof"]).items(), key=lambda kv: kv[1][0])
    per_p, per_c = kernel_bytes(label, form, w)
    alg = per_p * w.npart + per_c * w.ncells
    avg_s = tot / cnt * 1e-3
    achieved = alg / avg_s / 1e9 if alg else None
    return {"bound": "hbm", "kernel": label,
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": measured_traffic(label, w, form),
            "avg_launch_ms": round(tot / cnt, 4), "alg_bytes_per_launch": alg,
            "share_of_step": round(tot / r["ms_prof"], 3)}


def rate(r, w, steps):
    """(ms per step, Mparticle-steps/s) over the particles the timed steps carried (after an
    outflow removal, fewer than the set-up's)."""
    ms_step = r["ms"] / steps
    return ms_step, r["n_timed"] / (ms_step * 1e-3) / 1e6


def copy_calibration(gib):
    """Three launches of a plain elementwise copy of `gib` GiB (fp32) on the device: a kernel
    whose HBM traffic is known (gib GiB read + gib GiB written)."""
    import torch
    n = int(gib * 2**30 / 4)
    a = torch.rand(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": 3}), flush=True)


def main():
    args = parse()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X, so --gpus %d is "
              "not run (the path would shard by MeshBlock; see DESIGN.md)" % args.gpus)
        sys.exit(2)
    w = make_workload(args.workload, args.n)
    forms = ["fp32", "fp64"] if args.form == "both" else [args.form]
    res = {}
    tune_steps = 0
    for f in forms:
        print("[bench] %s form: %s, %d warm-up + %d timed steps" % (f, w.name, args.warmup,
                                                                     args.steps),
              file=sys.stderr, flush=True)
        res[f] = gpu_run(w, f, args.steps, args.warmup, tune_steps=tune_steps,
                         prof_steps=max(1, args.prof_steps), tune=not args.no_tune,
                         tune_config=args.tune_config)
        tune_steps = res[f]["tune_steps"]
    main_form = forms[0]
    r = res[main_form]
    ms_step, mps = rate(r, w, args.steps)
    pc2 = w.phys.coupling == 1
    step_alg = balg_step(main_form, w.ppc, pc2, bool(w.phys.clouds))
    out = {
        "metric": "Mparticle-steps/s & %HBM peak, FP32+batched vs FP64 GPU at 1e-5 rel err",
        "value": round(mps, 3), "unit": "Mparticle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f32 (f64 positions, gas momentum, sums)",
        "data": "synthetic (seeded hash draws, generated on the device)",
        "config": dict(describe(w), form=main_form),
        "roofline": kernel_roofline(r, w, main_form),
        "step_roofline": {"alg_bytes_per_particle_step": step_alg,
                          "achieved": round(mps * 1e6 * step_alg / 1e9, 1), "unit": "GB/s",
                          "frac": round(mps * 1e6 * step_alg / 1e9 / HBM_PEAK_GBS, 4)},
        "momentum_drift": drift_or_none(w, r["drift"]),
        "gas_update_roofline": gas_update_roofline(r, w, main_form),
        "particles_timed": r["n_timed"],
        "graphs": r["graphs"],
        "autotune": r["tuned"],
        "profiled_steps": {"steps": r["prof_steps"], "graphs": False,
                           "ms_per_step": round(r["ms_prof"] / r["prof_steps"], 3)},
        "kernels_ms_per_step": {k: round(v[0] / r["prof_steps"], 3)
                                for k, v in sorted(r["prof"].items(), key=lambda kv: -kv[1][0])},
    }
    if "fp64" in res and "fp32" in res:
        ms64, mps64 = rate(res["fp64"], w, args.steps)
        v64, v32 = res["fp64"]["vel"], res["fp32"]["vel"]
        t64, t32 = res["fp64"]["tags"], res["fp32"]["tags"]
        if not (len(t64) == len(t32) and bool((t64 == t32).all())):
            # (the particles both forms still hold, by tag)
            import torch
            common = torch.from_numpy(np.intersect1d(t64.cpu().numpy(), t32.cpu().numpy()))
            common = common.to(t64.device)
            v64 = v64[:, torch.searchsorted(t64, common)]
            v32 = v32[:, torch.searchsorted(t32, common)]
        # per particle: the largest component difference over the largest velocity
        dp = (v32 - v64).abs().amax(dim=0) / v64.abs().max()
        verr = float(dp.max())
        nbad = int((dp > 1e-5).sum())
        out["fp64_form"] = {"value": round(mps64, 3), "ms_per_step": round(ms64, 3),
                            "step_roofline_frac": round(
                                mps64 * 1e6 * balg_step("fp64", w.ppc, pc2, bool(w.phys.clouds)) /
                                1e9 / HBM_PEAK_GBS, 4),
                            "roofline": kernel_roofline(res["fp64"], w, "fp64"),
                            "gas_update_roofline": gas_update_roofline(res["fp64"], w, "fp64"),
                            "momentum_drift": drift_or_none(w, res["fp64"]["drift"])}
        out["speedup_vs_fp64"] = round(mps / mps64, 3)
        out["max_rel_velocity_err_vs_fp64"] = verr
        # particles beyond 1e-5: the consistent split takes each particle's centre cell (NGP),
        # so a particle whose FP32 and FP64 positions (1e-12..1e-9 apart) straddle a cell face
        # gets a different correction; over many steps a few of millions do
        out["particles_beyond_1e-5_vs_fp64"] = {"count": nbad, "of": int(dp.numel())}
    if out["momentum_drift"] is None and not args.no_ledger:
        print("[bench] drag-exchange conservation run (%s, sources off)" % main_form,
              file=sys.stderr, flush=True)
        ex = exchange_drift(args.workload, main_form)
        if ex is not None:
            out["momentum_drift"] = ex["value"]
            out["momentum_drift_how"] = "%s, %d steps" % (ex["how"], ex["steps"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
