"""ctypes binding of colb_amd/lib/libabd_hip.so (the C ABI of include/abd_hip.h and
include/abd_mlp.h).

There is no fallback: if the library is missing the import of the solver fails loudly with the
build command to run.
"""
import ctypes as C
import os

from ._abi import AbdParams, AbdSceneDesc, AbdStepStats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libabd_hip.so")

# exported symbol -> (restype, argtypes); must match include/abd_hip.h and include/abd_mlp.h
_D, _I, _V, _F = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p, C.POINTER(C.c_float)
SIGNATURES = {
    "abd_create": (C.c_int, [C.POINTER(AbdSceneDesc), C.POINTER(AbdParams), C.POINTER(_V)]),
    "abd_destroy": (None, [_V]),
    "abd_last_error": (C.c_char_p, [_V]),
    "abd_get_params": (C.c_int, [_V, C.POINTER(AbdParams)]),
    "abd_set_params": (C.c_int, [_V, C.POINTER(AbdParams)]),
    "abd_set_state": (C.c_int, [_V, _D, _D, _D]),
    "abd_get_state": (C.c_int, [_V, _D, _D, _D]),
    "abd_step": (C.c_int, [_V, C.c_int, C.POINTER(AbdStepStats)]),
    "abd_get_iter_log": (C.c_int, [_V, _D, _D, _I]),
    "abd_get_stiffness": (C.c_int, [_V, _D]),
    "abd_set_stiffness": (C.c_int, [_V, _D]),
    "abd_detect": (C.c_int, [_V, _I, _D, C.c_int, _I]),
    "abd_assemble": (C.c_int, [_V, _D, _D, _I, _D, C.c_int, _I, _D]),
    "abd_solve": (C.c_int, [_V, _D, _D, _I]),
    "abd_energy": (C.c_int, [_V, _D, _D]),
    "abd_ccd": (C.c_int, [_V, _D, _D]),
    "abd_line_search": (C.c_int, [_V, _D]),
    "abd_line_search_along": (C.c_int, [_V, _D, _D]),
    "abd_forward": (C.c_int, [_V]),
    "abd_get_vertices": (C.c_int, [_V, _D]),
    "abd_get_contact_points": (C.c_int, [_V, _D, _D, C.c_int, _I]),
    "abd_get_kernel_times": (C.c_int, [_V, C.c_char_p, C.c_int, _D, _I, _D, C.c_int, _I]),
    "abd_enable_timing": (C.c_int, [_V, C.c_int]),
    "abd_psd_project_test": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _D, _D]),
    "abd_autotune": (C.c_int, [_V, C.c_int, _I, _I, _D]),
    "abd_set_pcg_layout": (C.c_int, [_V, C.c_int, C.c_int]),
    "abd_set_warm_start": (C.c_int, [_V, _V, _F, _F, _F, _F]),
    "abd_set_preconditioner": (C.c_int, [_V, _V, _F, _F, _F, _F]),
    "abd_set_predictor": (C.c_int, [_V, _V, _F, _F, _F, _F]),
    "abd_log_steps": (C.c_int, [_V, C.c_int]),
    "abd_get_step_samples": (C.c_int, [_V, _F, _F, C.c_int, _I]),
    "abd_log_samples": (C.c_int, [_V, C.c_int]),
    "abd_get_samples": (C.c_int, [_V, _F, _F, C.c_int, _I]),
    # include/abd_mlp.h
    "abd_mlp_create": (C.c_int, [_I, C.c_int, C.POINTER(_V)]),
    "abd_mlp_destroy": (None, [_V]),
    "abd_mlp_last_error": (C.c_char_p, [_V]),
    "abd_mlp_num_params": (C.c_int, [_V]),
    "abd_mlp_set_params": (C.c_int, [_V, _F]),
    "abd_mlp_get_params": (C.c_int, [_V, _F]),
    "abd_mlp_set_loss_weights": (C.c_int, [_V, _F]),
    "abd_mlp_forward": (C.c_int, [_V, _F, C.c_int, _F]),
    "abd_mlp_forward_static": (C.c_int, [_V, _F, C.c_int, _F]),
    "abd_mlp_gradient": (C.c_int, [_V, _F, _F, C.c_int, _F, _F]),
    "abd_mlp_train": (C.c_int, [_V, _F, _F, C.c_int, _I, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                C.c_float, _F]),
    "abd_mlp_gemm_test": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, _F, _F, _F]),
}

_lib = None


class AbdError(RuntimeError):
    pass


def load():
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m colb_amd.build` "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, handle=None, what=""):
    if rc != 0:
        msg = load().abd_last_error(handle).decode() if handle else ""
        raise AbdError(f"{what} failed with status {rc}: {msg}")
