"""ctypes wrapper of the CPU FP64 oracle (oracle/abd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  `build()` compiles the oracle with g++ into
oracle/_build/libabd_oracle.so.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from colb_amd._abi import AbdParams, AbdSceneDesc, AbdStepStats, dptr, iptr, make_scene_desc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libabd_oracle.so")
_lib = None


def build(force=True):
    """Compile the oracle (g++, OpenMP) into oracle/_build/."""
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    # make rebuilds when the source or the ABI header is newer than the library
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", HERE], check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(force=False)
        L = C.CDLL(LIB_PATH)
        D, I, V = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
        sig = {
            "orc_barrier": (None, [C.c_double, C.c_double, D]),
            "orc_ortho": (C.c_double, [D, C.c_double, D, D]),
            "orc_pair_dist2": (C.c_double, [C.c_int, D, D, D, I]),
            "orc_seg_params": (None, [D, D, D]),
            "orc_psd_project": (None, [C.c_int, D]),
            "orc_accd": (C.c_double, [C.c_int, D, D, C.c_double, C.c_double, C.c_int]),
            "orc_chol_inverse12": (C.c_int, [D, D]),
            "orc_create": (V, [C.POINTER(AbdSceneDesc), C.POINTER(AbdParams)]),
            "orc_destroy": (None, [V]),
            "orc_set_threads": (None, [V, C.c_int]),
            "orc_set_params": (None, [V, C.POINTER(AbdParams)]),
            "orc_set_state": (None, [V, D, D, D]),
            "orc_get_state": (None, [V, D, D, D]),
            "orc_detect": (C.c_int, [V, D, I, D, C.c_int]),
            "orc_candidates": (C.c_int, [V, D, D, C.c_double, I, C.c_int]),
            "orc_energy": (None, [V, D, D]),
            "orc_assemble": (C.c_int, [V, D, D, D, I, D, C.c_int]),
            "orc_solve": (None, [V, D, I, D]),
            "orc_ccd": (None, [V, D, D]),
            "orc_step": (None, [V, C.c_int, C.POINTER(AbdStepStats)]),
            "orc_get_iter_log": (None, [V, D, D, I]),
            "orc_get_vertices": (None, [V, D]),
            "orc_get_stiffness": (None, [V, D]),
            "orc_set_stiffness": (None, [V, D]),
            "orc_take_error": (C.c_int, [V]),
            "orc_line_search": (None, [V, D, D]),
            "orc_features": (None, [V, C.POINTER(C.c_float)]),
            "orc_set_prediction": (None, [V, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------------------------------------
# pointwise functions
# ----------------------------------------------------------------------------------------------
def barrier(s, shat):
    out = np.zeros(3)
    lib().orc_barrier(float(s), float(shat), dptr(out))
    return out


def ortho(q12, kappa):
    q = f64(q12)
    g, H = np.zeros(9), np.zeros(81)
    e = lib().orc_ortho(dptr(q), float(kappa), dptr(g), dptr(H))
    return e, g, H.reshape(9, 9)


def pair_dist2(kind, x12, derivs=True):
    x = f64(x12).reshape(12)
    g, H = (np.zeros(12), np.zeros(144)) if derivs else (None, None)
    t = C.c_int(0)
    d2 = lib().orc_pair_dist2(int(kind), dptr(x), dptr(g), dptr(H), C.byref(t))
    if derivs:
        return d2, g, H.reshape(12, 12), t.value
    return d2, t.value


def seg_params(x12):
    s, t = C.c_double(), C.c_double()
    lib().orc_seg_params(dptr(f64(x12).reshape(12)), C.byref(s), C.byref(t))
    return s.value, t.value


def psd_project(H):
    H = f64(H).copy()
    n = H.shape[0]
    lib().orc_psd_project(n, dptr(H.reshape(-1)))
    return H


def accd(kind, x0, x1, s=0.1, tmax=1.0, max_iter=10000):
    return lib().orc_accd(int(kind), dptr(f64(x0).reshape(12)), dptr(f64(x1).reshape(12)), s, tmax, max_iter)


def chol_inverse12(A):
    out = np.zeros(144)
    rc = lib().orc_chol_inverse12(dptr(f64(A).reshape(144)), dptr(out))
    return rc, out.reshape(12, 12)


# ----------------------------------------------------------------------------------------------
# scene-level oracle
# ----------------------------------------------------------------------------------------------
class Oracle:
    """FP64 restatement of NewtonAbd (abd.py:583-757) over a colb_amd.scene.Scene."""

    def __init__(self, scene, params, threads=1):
        self.scene = scene
        self.params = params
        self._desc = make_scene_desc(scene)
        self._p = params.to_c()
        self.h = lib().orc_create(C.byref(self._desc), C.byref(self._p))
        lib().orc_set_threads(self.h, int(threads))
        self.nb, self.ns = scene.n_bodies, scene.n_scenes
        self.set_state(scene.q0, scene.q0, scene.qdot0)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_threads(self, n):
        lib().orc_set_threads(self.h, int(n))

    def set_params(self, params):
        self.params = params
        self._p = params.to_c()
        lib().orc_set_params(self.h, C.byref(self._p))

    def set_state(self, q=None, q_now=None, qdot=None):
        a = [None if x is None else f64(x).reshape(-1) for x in (q, q_now, qdot)]
        lib().orc_set_state(self.h, *[dptr(x) for x in a])

    def stiffness(self):
        k = np.zeros(self.nb)
        lib().orc_get_stiffness(self.h, dptr(k))
        return k

    def set_stiffness(self, k):
        lib().orc_set_stiffness(self.h, dptr(f64(k).reshape(-1)))

    def features(self):
        """The 49 body features of the learned models at the current iterate, float32 [nb, 49]."""
        f = np.zeros((self.nb, 49), dtype=np.float32)
        lib().orc_features(self.h, f.ctypes.data_as(C.POINTER(C.c_float)))
        return f

    def set_prediction(self, dq):
        """The step predictor's dq [nb, 12] for the following steps with predictor 3 (None clears)."""
        lib().orc_set_prediction(self.h, None if dq is None else dptr(f64(dq).reshape(-1)))

    def take_error(self):
        """Error bits raised since the last call (64: a vertex on or below the ground plane, 128: a
        step with predictor 3 and no prediction set, 256: a diagonal block without a Cholesky
        factor in features())."""
        return int(lib().orc_take_error(self.h))

    def get_state(self):
        q, qn, qd = (np.zeros(12 * self.nb) for _ in range(3))
        lib().orc_get_state(self.h, dptr(q), dptr(qn), dptr(qd))
        return q.reshape(-1, 12), qn.reshape(-1, 12), qd.reshape(-1, 12)

    def detect(self, q=None, cap=1 << 20):
        qq = None if q is None else f64(q).reshape(-1)
        pairs = np.zeros(3 * cap, dtype=np.int32)
        d2 = np.zeros(cap)
        n = lib().orc_detect(self.h, dptr(qq), iptr(pairs), dptr(d2), cap)
        assert n <= cap
        return pairs[:3 * n].reshape(-1, 3), d2[:n]

    def candidates(self, q0, q1=None, infl=None, cap=1 << 22):
        infl = self.params.dhat if infl is None else infl
        pairs = np.zeros(3 * cap, dtype=np.int32)
        n = lib().orc_candidates(self.h, dptr(f64(q0).reshape(-1)),
                                 None if q1 is None else dptr(f64(q1).reshape(-1)), infl, iptr(pairs), cap)
        assert n <= cap
        return pairs[:3 * n].reshape(-1, 3)

    def energy(self, q=None):
        E = np.zeros(self.ns)
        lib().orc_energy(self.h, None if q is None else dptr(f64(q).reshape(-1)), dptr(E))
        return E

    def assemble(self, q=None, cap=1 << 16):
        qq = None if q is None else f64(q).reshape(-1)
        g = np.zeros(12 * self.nb)
        D = np.zeros(144 * self.nb)
        pij = np.zeros(2 * cap, dtype=np.int32)
        off = np.zeros(144 * cap)
        n = lib().orc_assemble(self.h, dptr(qq), dptr(g), dptr(D), iptr(pij), dptr(off), cap)
        assert n <= cap
        return g.reshape(-1, 12), D.reshape(-1, 12, 12), pij[:2 * n].reshape(-1, 2), off[:144 * n].reshape(-1, 12, 12)

    def dense_hessian(self, q=None):
        g, D, pij, off = self.assemble(q)
        n = 12 * self.nb
        H = np.zeros((n, n))
        for b in range(self.nb):
            H[12 * b:12 * b + 12, 12 * b:12 * b + 12] = D[b]
        for (i, j), B in zip(pij, off):
            H[12 * i:12 * i + 12, 12 * j:12 * j + 12] = B
            H[12 * j:12 * j + 12, 12 * i:12 * i + 12] = B.T
        return g.reshape(-1), H

    def solve(self):
        du = np.zeros(12 * self.nb)
        it = np.zeros(self.ns, dtype=np.int32)
        res = np.zeros(self.ns)
        lib().orc_solve(self.h, dptr(du), iptr(it), dptr(res))
        return du.reshape(-1, 12), it, res

    def ccd(self, du):
        a = np.zeros(self.ns)
        lib().orc_ccd(self.h, dptr(f64(du).reshape(-1)), dptr(a))
        return a

    def line_search(self, du):
        """Line search along du (q -= alpha du applied); alpha per scene."""
        a = np.zeros(self.ns)
        lib().orc_line_search(self.h, dptr(f64(du).reshape(-1)), dptr(a))
        return a

    def step(self, nsteps=1):
        st = (AbdStepStats * self.ns)()
        lib().orc_step(self.h, int(nsteps), st)
        return [s.as_dict() for s in st]

    def iter_log(self):
        m = self.params.max_newton
        r, a = np.zeros(self.ns * m), np.zeros(self.ns * m)
        cg = np.zeros(self.ns * m, dtype=np.int32)
        lib().orc_get_iter_log(self.h, dptr(r), dptr(a), iptr(cg))
        return r.reshape(self.ns, m), a.reshape(self.ns, m), cg.reshape(self.ns, m)

    def vertices(self):
        x = np.zeros(3 * self.scene.n_verts)
        lib().orc_get_vertices(self.h, dptr(x))
        return x.reshape(-1, 3)
