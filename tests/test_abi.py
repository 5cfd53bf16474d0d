"""The C-ABI library loads and exports every entry point the headers declare (no GPU call)."""

import ctypes
import os
import re

import pytest

from athenak_amd import abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("dustk.h", "dustk_setup.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(dk_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_task_entry_points():
    names = declared_symbols()
    for t in lib.TASKS:
        assert "dk_" + t in names
    assert {"dk_create", "dk_destroy", "dk_stage_run", "dk_diag_sums", "dk_setup_problem"} <= names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libdustk.so not built")
    L = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_struct_sizes_match_the_header_layout(tmp_path):
    """The ctypes mirror has the sizes the C compiler gives the header's structs."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "dustk.h"\n#include "dustk_setup.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(dk_block),'
                   ' sizeof(dk_stage), sizeof(dk_mesh), sizeof(dk_physics), sizeof(dk_gas_src),'
                   ' sizeof(dk_setup), sizeof(dk_output_info)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [ctypes.sizeof(t) for t in (abi.dk_block, abi.dk_stage, abi.dk_mesh, abi.dk_physics,
                                       abi.dk_gas_src, abi.dk_setup, abi.dk_output_info)]
    assert got == want


def test_strerror_without_device():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libdustk.so not built")
    L = lib.lib()
    assert L.dk_strerror(4).decode().startswith("particle outside")


def test_clouds_refused_where_not_built_without_device():
    """dk_create checks its options before any HIP call: clouds need pc2 and a mesh periodic in
    x1 and x3 (in either form)."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libdustk.so not built")
    from athenak_amd import workloads
    L = lib.lib()
    out = ctypes.c_void_p()
    nx = 8
    cases = []
    mesh, blocks = abi.uniform_mesh((nx,) * 3, (1, 1, 1), ((0.0, 1.0),) * 3, ng=2,
                                    bc_x1=abi.DK_X1_SHEAR)
    cases.append((mesh, blocks, abi.DK_PC2, abi.DK_FORM_FP64))
    mesh, blocks = abi.uniform_mesh((nx,) * 3, (1, 1, 1), ((0.0, 1.0),) * 3, ng=2,
                                    bc_x3=abi.DK_X3_OUTFLOW)
    cases.append((mesh, blocks, abi.DK_PC2, abi.DK_FORM_FP64))
    mesh, blocks = abi.uniform_mesh((nx,) * 3, (1, 1, 1), ((0.0, 1.0),) * 3, ng=2)
    cases.append((mesh, blocks, abi.DK_IMEX, abi.DK_FORM_FP64))
    cases.append((mesh, blocks, abi.DK_IMEX, abi.DK_FORM_FP32))
    mesh, blocks = abi.uniform_mesh((nx,) * 3, (1, 1, 1), ((0.0, 1.0),) * 3, ng=2,
                                    bc_x1=abi.DK_X1_SHEAR)
    cases.append((mesh, blocks, abi.DK_PC2, abi.DK_FORM_FP32))
    for mesh, blocks, coupling, form in cases:
        phys = workloads._phys(coupling=coupling, solver=abi.DK_LOCAL, clouds=1)
        e = L.dk_create(ctypes.byref(mesh), blocks, ctypes.byref(phys), form, None,
                        ctypes.byref(out))
        assert e == 7, (coupling, form, e)   # DK_EUNSUPPORTED


def test_gamma_switch_follows_the_reference_rule():
    # GammaSwitch (src/dust/dust_newdt.cpp:160-173): 1/2 only when dt exceeds the largest
    # stopping time (strictly), 1 + 1/sqrt(2) otherwise, nothing for pc2; the imex2+ records
    # built with either coefficient satisfy the tableau's identities
    assert abi.gamma_switch(0.2, 0.1) == 0.5
    assert abi.gamma_switch(0.1, 0.1) == abi.IMEX2P_GAMMA
    assert abi.gamma_switch(0.05, 0.1) == abi.IMEX2P_GAMMA
    assert abi.gamma_switch(0.2, 0.1, abi.DK_PC2) is None
    for gam in (0.5, abi.IMEX2P_GAMMA):
        st = abi.imex2p_stages(0.0, 1.0, gam)[1]
        assert abs(st.gam0[1] + st.gam1[1] - 1.0) < 1e-15
        assert st.a_impl == gam and st.beta[0] == gam
        assert abs(st.a_twid22 - (1.0 - 2.0 * gam * gam) / (2.0 * gam)) < 1e-15
