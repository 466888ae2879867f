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
        " offsetof(AbdParams, direct_max_dofs)); printf(\"%zu %zu %zu %zu\\n\", offsetof(AbdParams, k_max),"
        " offsetof(AbdParams, k_raise_gap), offsetof(AbdStepStats, k_raises), offsetof(AbdStepStats, predictor_alpha));"
        " return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(_abi.AbdParams), C.sizeof(_abi.AbdSceneDesc), C.sizeof(_abi.AbdStepStats),
            _abi.AbdParams.ground_normal.offset, _abi.AbdParams.max_candidates.offset,
            _abi.AbdStepStats.residual_first.offset, _abi.AbdParams.solver.offset,
            _abi.AbdParams.direct_max_dofs.offset, _abi.AbdParams.k_max.offset, _abi.AbdParams.k_raise_gap.offset,
            _abi.AbdStepStats.k_raises.offset, _abi.AbdStepStats.predictor_alpha.offset]
    assert got == want


def test_params_reach_the_c_struct():
    from colb_amd.params import north_star
    c = north_star(solver=2, direct_max_dofs=1536, precision=1, debug=4).to_c()
    assert (c.solver, c.direct_max_dofs, c.precision, c.reserved0) == (2, 1536, 1, 4)
    d = north_star().to_c()
    assert (d.solver, d.direct_max_dofs) == (0, 6144)


INVALID = [("dt", 0.0), ("dt", -1e-3), ("dhat", 0.0), ("kappa", -1.0), ("k_contact", -1.0), ("newton_tol", -1e-5),
           ("max_newton", 0), ("cg_max_iter", 0), ("cg_tol", 0.0), ("cg_tol", 1.5), ("ccd_gap_ratio", 0.0),
           ("ccd_gap_ratio", 1.0), ("ccd_max_iter", 0), ("ls_floor", 0.0), ("ls_floor", 1.0), ("pinned_diag", 0.0),
           ("dt", float("nan")), ("gravity", (0.0, float("inf"), 0.0)), ("predictor", 4), ("hess_proj", 2),
           ("solver", 3), ("precision", 2), ("direct_max_dofs", 1 << 20),
           ("k_raise_gap", 1.0), ("pcg_x0", 2), ("stall_alpha", -1e-3), ("stall_alpha", 1.0)]


@pytest.mark.parametrize("field,value", INVALID)
def test_create_rejects_invalid_parameters(field, value):
    """abd_create validates the parameters before it touches the device (no GPU call is made)."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    from colb_amd import scene as S
    from colb_amd.params import north_star
    sc = S.cubes_on_ground(n=1)
    kw = {field: value}
    if field == "k_raise_gap":
        kw["k_max"] = 1e30  # the gap is checked when the stiffness adapts
    p = north_star(**kw).to_c()
    desc = _abi.make_scene_desc(sc)
    L = _lib.load()
    h = C.c_void_p()
    rc = L.abd_create(C.byref(desc), C.byref(p), C.byref(h))
    msg = L.abd_last_error(h).decode()
    L.abd_destroy(h)
    assert rc == _abi.ABD_ERR_ARG
    assert field.split("_")[0] in msg or field in msg, msg


def test_profiler_npz_round_trip(tmp_path):
    """Profiller (utils/data_collector.py:4-55): the npz keeps `convergence` and `alphas` per frame."""
    import numpy as np
    from colb_amd.profiler import Profiller
    p = Profiller(max_iters=16, frames=5)
    rng = np.random.default_rng(0)
    rows = [rng.uniform(size=n) for n in (3, 16, 7)]
    for f, r in enumerate(rows):
        p.record(f, r, 0.5 * r)
    p.record(9, np.ones(4), np.ones(4))  # beyond the frames: ignored
    path = tmp_path / "profiler.npz"
    p.save(str(path))
    with np.load(path, allow_pickle=False) as z:
        assert sorted(z.files) == ["alphas", "convergence"]
        assert z["convergence"].shape == (5, 16)
        for f, r in enumerate(rows):
            np.testing.assert_array_equal(z["convergence"][f, :len(r)], r)
            np.testing.assert_array_equal(z["alphas"][f, :len(r)], 0.5 * r)
            assert np.all(z["convergence"][f, len(r):] == 0)
    q = Profiller(max_iters=16, frames=5)
    q.load(str(path))
    np.testing.assert_array_equal(q.convergence, p.convergence)
    np.testing.assert_array_equal(q.alphas, p.alphas)


@pytest.mark.parametrize("m,precision,n", [(5, 0, 1), (12, 2, 1), (3, 0, 0), (16, 1, 4)])
def test_psd_projection_check_rejects_invalid_arguments(m, precision, n):
    """abd_psd_project_test validates its arguments before it touches the device (no GPU call is
    made): orders other than 3, 6, 9, 12, a precision other than 0 or 1, no matrices."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("HIP library not built")
    import ctypes as C
    import numpy as np
    A = np.zeros(max(n, 1) * m * m)
    out = np.zeros_like(A)
    P = C.POINTER(C.c_double)
    rc = _lib.load().abd_psd_project_test(m, precision, 0, n, A.ctypes.data_as(P), out.ctypes.data_as(P))
    assert rc == -2  # ABD_ERR_ARG
