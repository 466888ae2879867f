"""ctypes mirrors of the structs declared in include/abd_hip.h.

Field order and types must match the C header exactly; tests/test_abi.py checks the sizes and
offsets against the compiled library.
"""
import ctypes as C

ABD_OK = 0
ABD_ERR_HIP = -1
ABD_ERR_ARG = -2
ABD_ERR_CAPACITY = -3
ABD_ERR_NAN = -4
ABD_ERR_STATE = -5

ABD_PT, ABD_EE, ABD_PG = 0, 1, 2
ABD_D_PP, ABD_D_PE, ABD_D_PT, ABD_D_EE, ABD_D_PG = 0, 1, 2, 3, 4


class AbdParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("gravity", C.c_double * 3),
        ("dhat", C.c_double),
        ("kappa", C.c_double),
        ("k_contact", C.c_double),
        ("newton_tol", C.c_double),
        ("max_newton", C.c_int),
        ("cg_max_iter", C.c_int),
        ("cg_tol", C.c_double),
        ("ground_enabled", C.c_int),
        ("ccd_enabled", C.c_int),
        ("ground_normal", C.c_double * 3),
        ("ground_offset", C.c_double),
        ("ccd_gap_ratio", C.c_double),
        ("ccd_max_iter", C.c_int),
        ("precision", C.c_int),
        ("ls_floor", C.c_double),
        ("pinned_diag", C.c_double),
        ("max_candidates", C.c_int),
        ("reserved0", C.c_int),
        ("solver", C.c_int),
        ("direct_max_dofs", C.c_int),
        ("predictor", C.c_int),
        ("hess_proj", C.c_int),
        ("k_max", C.c_double),
        ("k_raise_gap", C.c_double),
        ("cand_margin", C.c_double),
        ("pcg_x0", C.c_int),
        ("reserved1", C.c_int),
        ("stall_alpha", C.c_double),
    ]


class AbdSceneDesc(C.Structure):
    _fields_ = [
        ("n_bodies", C.c_int),
        ("n_verts", C.c_int),
        ("n_edges", C.c_int),
        ("n_tris", C.c_int),
        ("n_scenes", C.c_int),
        ("rest_x", C.POINTER(C.c_double)),
        ("vert_body", C.POINTER(C.c_int)),
        ("edges", C.POINTER(C.c_int)),
        ("tris", C.POINTER(C.c_int)),
        ("body_scene", C.POINTER(C.c_int)),
        ("body_pinned", C.POINTER(C.c_int)),
        ("mass4", C.POINTER(C.c_double)),
    ]


class AbdStepStats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int),
        ("converged", C.c_int),
        ("cg_iters", C.c_int),
        ("ls_floor_hits", C.c_int),
        ("n_active", C.c_int),
        ("n_candidates", C.c_int),
        ("residual_first", C.c_double),
        ("residual_last", C.c_double),
        ("alpha_min", C.c_double),
        ("energy", C.c_double),
        ("ccd_alpha_min", C.c_double),
        ("k_contact", C.c_double),
        ("k_raises", C.c_int),
        ("pcg_fallbacks", C.c_int),
        ("predictor_alpha", C.c_double),
        ("traversals", C.c_int),
        ("detections", C.c_int),
        ("slow_iterations", C.c_int),
        ("graph_launches", C.c_int),
        ("graph_captures", C.c_int),
        ("stalled", C.c_int),
        ("ls_trials", C.c_int),
        ("ls_round_trips", C.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    """int* of a C-contiguous int32 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype.name == "int32" and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_scene_desc(scene):
    """Build an AbdSceneDesc over the arrays of a colb_amd.scene.Scene (kept alive by scene)."""
    d = AbdSceneDesc()
    d.n_bodies = scene.n_bodies
    d.n_verts = scene.n_verts
    d.n_edges = scene.n_edges
    d.n_tris = scene.n_tris
    d.n_scenes = scene.n_scenes
    d.rest_x = dptr(scene.rest_x)
    d.vert_body = iptr(scene.vert_body)
    d.edges = iptr(scene.edges)
    d.tris = iptr(scene.tris)
    d.body_scene = iptr(scene.body_scene)
    d.body_pinned = iptr(scene.body_pinned)
    d.mass4 = dptr(scene.mass4)
    return d
