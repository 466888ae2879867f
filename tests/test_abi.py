"""The C ABI: the built HIP library loads and exports every function include/abd_hip.h declares,
and the ctypes struct mirrors have the C layout (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from colb_amd import _abi, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "abd_hip.h"), os.path.join(ROOT, "include", "abd_mlp.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(abd_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_bound_functions():
    assert declared_functions() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (abd_\w+)", out))
    missing = set(declared_functions()) - exported
    assert not missing, missing
    _lib.load()  # resolves every signature


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "abd_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(AbdParams), sizeof(AbdSceneDesc),"
        " sizeof(AbdStepStats), offsetof(AbdParams, ground_normal), offsetof(AbdParams, max_candidates),"
        " offsetof(AbdStepStats, residual_first), offsetof(AbdParams, solver),"
        " offsetof(AbdParams, direct_max_dofs)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(_abi.AbdParams), C.sizeof(_abi.AbdSceneDesc), C.sizeof(_abi.AbdStepStats),
            _abi.AbdParams.ground_normal.offset, _abi.AbdParams.max_candidates.offset,
            _abi.AbdStepStats.residual_first.offset, _abi.AbdParams.solver.offset,
            _abi.AbdParams.direct_max_dofs.offset]
    assert got == want


def test_params_reach_the_c_struct():
    from colb_amd.params import north_star
    c = north_star(solver=2, direct_max_dofs=1536, precision=1, debug=4).to_c()
    assert (c.solver, c.direct_max_dofs, c.precision, c.reserved0) == (2, 1536, 1, 4)
    d = north_star().to_c()
    assert (d.solver, d.direct_max_dofs) == (0, 6144)
