"""Builds libdustk.so (gfx950) in-tree with hipcc.  Used by __graft_entry__.build()."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdustk.so")
SOURCES = ["dk_fp64.hip", "dk_fp32.hip", "dk_setup.hip", "dk_prof.hip", "dk_mesh.hip",
           "dk_clouds.hip", "dk_output.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no fused multiply-adds, so that the FP64 form rounds as the oracle does
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
PER_FILE = {"dk_fp64.hip": ["-ffp-contract=off"], "dk_setup.hip": ["-ffp-contract=off"],
            "dk_fp32.hip": [], "dk_prof.hip": [], "dk_mesh.hip": ["-ffp-contract=off"],
            "dk_clouds.hip": ["-ffp-contract=off"], "dk_output.hip": ["-ffp-contract=off"]}


def build(verbose=False, jobs=3):
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC] + COMMON + PER_FILE.get(src, []) + ["-c", os.path.join(CSRC, src),
                                                          "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, text in failed:
            sys.stderr.write("hipcc failed on %s:\n%s\n" % (src, text))
        raise RuntimeError("libdustk build failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("libdustk link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_capi_driver(verbose=False):
    """The C++ caller of the C-ABI that the tests run (tests/capi/drive_step), linked against
    the library built above."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "capi", "drive_step.cpp")
    out = os.path.join(root, "tests", "capi", "drive_step")
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(root, "include"), src,
           "-L" + HERE, "-ldustk", "-Wl,-rpath,$ORIGIN/../../athenak_amd", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("drive_step build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
