"""The C-ABI driven from C++ without Python (tests/capi/drive_step.cpp, built by build()): the
program creates a context, sets up the seeded problem and runs imex2+ steps with the stage records
an AthenaK task body passes.  Its diagnostic sums must equal those of the same run driven through
the Python binding (bit for bit in the deterministic FP32 form, to round-off in the FP64 form with
its atomic deposits), with and without stage graphs, and agree with the oracle."""

import os
import subprocess
import sys

import numpy as np
import pytest

from athenak_amd import abi, workloads
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "capi", "drive_step")
STEPS = 5


def say(msg):
    print("[gpu-capi] " + msg, flush=True)
    sys.stdout.flush()


def drive(form, graphs):
    if not os.path.exists(EXE):
        raise RuntimeError("tests/capi/drive_step is not built (__graft_entry__.build())")
    say("drive_step form %d graphs %d" % (form, graphs))
    out = subprocess.run([EXE, str(form), str(STEPS), str(graphs)], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, timeout=120)
    text = out.stdout.decode()
    assert out.returncode == 0, text
    vals = {}
    for line in text.splitlines():
        f = line.split()
        if f and f[0] in ("npart", "diag", "vsum"):
            vals[f[0]] = [float(x) for x in f[1:] if x != "graphs"]
    return vals


def python_run(form):
    from athenak_amd.lib import DustContext
    w = workloads.c1(n=16, nb=2, ppc=2)
    g = DustContext(w.mesh, w.blocks, w.phys, w.npart + 16, form=form)
    g.setup(w.setup)
    for _ in range(STEPS):
        g.step(w.dt)
    g.check_device_error()
    d = g.diag()
    g.close()
    return d, w


def test_cpp_caller_matches_the_python_binding_and_the_oracle():
    w = workloads.c1(n=16, nb=2, ppc=2)
    o = Oracle(w.mesh, w.blocks, w.phys, w.npart + 16)
    o.setup(w.setup)
    for _ in range(STEPS):
        o.step(w.dt)
    do = o.diag()
    for form in (abi.DK_FORM_FP32, abi.DK_FORM_FP64):
        dp, _ = python_run(form)
        runs = [drive(form, graphs) for graphs in (0, 1)]
        for graphs, r in enumerate(runs):
            assert r["npart"][0] == w.npart
            if graphs:
                assert r["npart"][1] > 0     # the stages ran as graphs
            dc = np.array(r["diag"])
            if form == abi.DK_FORM_FP32:
                assert np.array_equal(dc, dp), (dc, dp)
            else:
                np.testing.assert_allclose(dc, dp, rtol=1e-12, atol=1e-18)
            # momentum and masses against the oracle (FP32 bulk: 1e-5 of the scale)
            tol = 1e-5 if form == abi.DK_FORM_FP32 else 1e-11
            assert np.abs(dc[:6] - do[:6]).max() <= tol * do[8], (form, dc, do)
            say("form %d graphs %d: diag matches (python %s)" % (form, graphs, "exact"
                if form == abi.DK_FORM_FP32 else "1e-12"))
        if form == abi.DK_FORM_FP32:
            assert runs[0]["diag"] == runs[1]["diag"] and runs[0]["vsum"] == runs[1]["vsum"]
