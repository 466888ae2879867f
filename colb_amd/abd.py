"""`NewtonAbd`: drop-in for the reference solver class (abd.py:583-757) on MI355X.

The class keeps the reference's constructor, methods and attributes
(``NewtonAbd(h, config_file)``, ``step``, ``reset``, ``compute_V``, ``get_contact_points``, the
template hooks ``detect_collision``, ``compute_du``, ``line_search``, ``compute_g``, ``add_du``,
``forward_states``, and ``soup.x_transformed``, ``E``, ``F``, ``V``, ``frame``, ``profiler``,
``n_bodies``, ``n_nodes``, ``n_contacts``, ``dt``), while every kernel of the step runs in
colb_amd/lib/libabd_hip.so through the C ABI of include/abd_hip.h.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._abi import AbdStepStats, dptr, iptr, make_scene_desc
from .params import Params, north_star, reference_compat
from .profiler import Profiller
from .scene import Scene, load_json_scene


class _Positions(np.ndarray):
    """World vertex positions (n_nodes, 3) float64.  `.numpy()` returns the plain array, as the Warp
    array the reference hands out does (viewer.py:6 reads ``soup.x_transformed.numpy()``)."""

    def numpy(self):
        return np.asarray(self)


class _Soup:
    """Read-only view mirroring geometry.Soup (geometry.py:22-28)."""

    def __init__(self, solver):
        self._s = solver
        sc = solver.scene
        self.xcs = sc.rest_x
        self.triangles = sc.tris.reshape(-1)
        self.edges = sc.edges.reshape(-1)
        self.body = sc.vert_body

    @property
    def x_transformed(self):
        return self._s.compute_V(ret=True).copy().view(_Positions)


class NewtonAbd:
    """ABD Newton solver with IPC barriers (abd.py:583).

    Parameters
    ----------
    h : float
        time step (links.py:16).
    config_file : str or Scene
        a reference JSON scene (utils/scene.py:65) -- solved with ``reference_compat`` params
        unless ``params`` is given -- or a :class:`colb_amd.scene.Scene`.
    params : Params, optional
    """

    def __init__(self, h, config_file, params: Params = None, max_frames=50):
        if isinstance(config_file, Scene):
            scene = config_file
            params = params or north_star()
        else:
            scene = load_json_scene(config_file)
            params = params or reference_compat()
        self.scene = scene
        self.params = params.replace(dt=float(h))
        self.dt = float(h)
        self.n_bodies = scene.n_bodies
        self.n_nodes = scene.n_verts
        self.n_scenes = scene.n_scenes
        self.V0 = scene.rest_x.copy()
        self.V = self.V0.copy()
        self.E = scene.edges.copy()
        self.F = scene.tris.copy()
        self.frame = 0
        self.n_contacts = 0
        self.profiler = Profiller(max_iters=self.params.max_newton, frames=max_frames)
        self.last_stats = None
        self._du = None
        self.pcg_layout = (5, False)  # PCG bodies per tile, one launch per iteration
        L = _lib.load()
        self._desc = make_scene_desc(scene)
        self._cp = self.params.to_c()
        h_ = C.c_void_p()
        rc = L.abd_create(C.byref(self._desc), C.byref(self._cp), C.byref(h_))
        self._h = h_
        _lib.check(rc, self._h, "abd_create")
        self.soup = _Soup(self)
        self.reset()

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.load().abd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: Params):
        self.params = params
        self.dt = float(params.dt)
        self._cp = params.to_c()
        _lib.check(_lib.load().abd_set_params(self._h, C.byref(self._cp)), self._h, "abd_set_params")

    # -- state -------------------------------------------------------------------------------
    def reset(self):
        """AbdComplex.reset (abd_simple.py:64-69): initial poses and velocities of the scene."""
        self.set_state(self.scene.q0, self.scene.q0, self.scene.qdot0)
        self.frame = 0
        self.compute_V()

    def set_state(self, q, q_now=None, qdot=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (q, q_now, qdot)]
        _lib.check(_lib.load().abd_set_state(self._h, *[dptr(x) for x in a]), self._h, "abd_set_state")

    def get_state(self):
        n = 12 * self.n_bodies
        q, qn, qd = np.zeros(n), np.zeros(n), np.zeros(n)
        _lib.check(_lib.load().abd_get_state(self._h, dptr(q), dptr(qn), dptr(qd)), self._h, "abd_get_state")
        return q.reshape(-1, 12), qn.reshape(-1, 12), qd.reshape(-1, 12)

    @property
    def history(self):
        """State as arrays: q (reference BDFAffine.nxt: c and rows of q), q_now, qdot."""
        q, qn, qd = self.get_state()
        return {"q": q, "q_now": qn, "qdot": qd}

    @property
    def inertia(self):
        return self.scene.mass4

    # -- the step ----------------------------------------------------------------------------
    def step(self, nsteps=1):
        """NewtonAbd.step (abd.py:674-702), device resident; records the Profiller arrays."""
        st = (AbdStepStats * self.n_scenes)()
        for _ in range(nsteps):
            _lib.check(_lib.load().abd_step(self._h, 1, st), self._h, "abd_step")
            res, alpha, _ = self.iter_log()
            self.profiler.record(self.frame, res[0], alpha[0])
            self.frame += 1
        self.last_stats = [s.as_dict() for s in st]
        self.n_contacts = int(sum(s["n_active"] for s in self.last_stats))
        return self.last_stats

    def iter_log(self):
        m = self.params.max_newton
        r, a = np.zeros(self.n_scenes * m), np.zeros(self.n_scenes * m)
        cg = np.zeros(self.n_scenes * m, dtype=np.int32)
        _lib.check(_lib.load().abd_get_iter_log(self._h, dptr(r), dptr(a), iptr(cg)), self._h, "abd_get_iter_log")
        return r.reshape(-1, m), a.reshape(-1, m), cg.reshape(-1, m)

    # -- template hooks (abd.py / contact.py / gauss_newton.py) --------------------------------
    def detect_collision(self, cap=1 << 20):
        """contact.py:270-279: active pairs (kind, i, j) and their d^2 at the current iterate."""
        pairs = np.zeros(3 * cap, dtype=np.int32)
        d2 = np.zeros(cap)
        n = C.c_int(0)
        rc = _lib.load().abd_detect(self._h, iptr(pairs), dptr(d2), cap, C.byref(n))
        _lib.check(rc, self._h, "abd_detect")
        self.n_contacts = n.value
        return pairs[:3 * n.value].reshape(-1, 3), d2[:n.value]

    def assemble(self, cap=1 << 16):
        """inertia_grad_hess + contact_hessian_ee + bsr_from_triplets (abd.py:682-688, 634)."""
        nb = self.n_bodies
        g, D = np.zeros(12 * nb), np.zeros(144 * nb)
        pij = np.zeros(2 * cap, dtype=np.int32)
        off = np.zeros(144 * cap)
        npairs = C.c_int(0)
        E = np.zeros(self.n_scenes)
        rc = _lib.load().abd_assemble(self._h, dptr(g), dptr(D), iptr(pij), dptr(off), cap, C.byref(npairs), dptr(E))
        _lib.check(rc, self._h, "abd_assemble")
        n = npairs.value
        return g.reshape(-1, 12), D.reshape(-1, 12, 12), pij[:2 * n].reshape(-1, 2), off[:144 * n].reshape(-1, 12, 12), E

    def compute_du(self, iter=0):
        """compute_du (abd.py:631-658): assemble + block-Jacobi PCG; returns ||grad||_2."""
        nb = self.n_bodies
        g = np.zeros(12 * nb)
        rc = _lib.load().abd_assemble(self._h, dptr(g), None, None, None, 0, None, None)
        _lib.check(rc, self._h, "abd_assemble")
        du = np.zeros(12 * nb)
        res = np.zeros(self.n_scenes)
        cg = np.zeros(self.n_scenes, dtype=np.int32)
        _lib.check(_lib.load().abd_solve(self._h, dptr(du), dptr(res), iptr(cg)), self._h, "abd_solve")
        self._du = du.reshape(-1, 12)
        self.last_cg_iters = cg
        return float(res[0]) if self.n_scenes == 1 else res

    def line_search(self):
        """line_search_iterative (gauss_newton.py:98-115) with the CCD filter; applies the step."""
        a = np.zeros(self.n_scenes)
        _lib.check(_lib.load().abd_line_search(self._h, dptr(a)), self._h, "abd_line_search")
        return float(a[0]) if self.n_scenes == 1 else a

    def compute_g(self, update_contact=True, q=None):
        """compute_g (abd.py:725-735): incremental potential at the current iterate (or q)."""
        E = np.zeros(self.n_scenes)
        qq = None if q is None else np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        _lib.check(_lib.load().abd_energy(self._h, dptr(qq), dptr(E)), self._h, "abd_energy")
        return float(E[0]) if self.n_scenes == 1 else E

    def line_search_along(self, du):
        """The CCD-filtered line search along a given direction du (applies q -= alpha du)."""
        a = np.zeros(self.n_scenes)
        _lib.check(_lib.load().abd_line_search_along(self._h, dptr(np.ascontiguousarray(du, dtype=np.float64).reshape(-1)),
                                                     dptr(a)), self._h, "abd_line_search_along")
        return a

    def stiffness(self):
        """Current contact stiffness of every body (adaptive when params.k_max > k_contact)."""
        k = np.zeros(self.n_bodies)
        _lib.check(_lib.load().abd_get_stiffness(self._h, dptr(k)), self._h, "abd_get_stiffness")
        return k

    def set_stiffness(self, k):
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1)
        if k.size != self.n_bodies:
            raise ValueError("one stiffness per body")
        _lib.check(_lib.load().abd_set_stiffness(self._h, dptr(k)), self._h, "abd_set_stiffness")

    def ccd(self, du):
        a = np.zeros(self.n_scenes)
        _lib.check(_lib.load().abd_ccd(self._h, dptr(np.ascontiguousarray(du, dtype=np.float64).reshape(-1)), dptr(a)),
                   self._h, "abd_ccd")
        return a

    def add_du(self, alpha):
        """add_du (abd.py:744-746): q -= alpha * du of the last compute_du."""
        if self._du is None:
            raise _lib.AbdError("add_du before compute_du")
        q, qn, qd = self.get_state()
        self.set_state(q - alpha * self._du, None, None)

    def forward_states(self):
        """forward_states (abd.py:670-672)."""
        _lib.check(_lib.load().abd_forward(self._h), self._h, "abd_forward")
        self.frame += 1

    # -- readback for viewers (viewer.py:44-47) --------------------------------------------------
    def compute_V(self, ret=True):
        """compute_V (abd_simple.py:71-76): world vertex positions (n_nodes, 3)."""
        x = np.zeros(3 * self.n_nodes)
        _lib.check(_lib.load().abd_get_vertices(self._h, dptr(x)), self._h, "abd_get_vertices")
        self.V = x.reshape(-1, 3)
        return self.V if ret else None

    def get_contact_points(self, cap=1 << 20):
        """get_contact_points (abd.py:748-757): midpoints of closest points of active pairs and
        their penetration-into-dhat magnitude |d - dhat|."""
        mid, dist = np.zeros(3 * cap), np.zeros(cap)
        n = C.c_int(0)
        rc = _lib.load().abd_get_contact_points(self._h, dptr(mid), dptr(dist), cap, C.byref(n))
        _lib.check(rc, self._h, "abd_get_contact_points")
        k = n.value
        d = dist[:k]
        return mid[:3 * k].reshape(-1, 3), np.abs(d - self.params.dhat)

    # -- per-kernel timing (HIP events) -------------------------------------------------------------
    def autotune(self, repeats=3):
        """Pick the PCG tile size and launch form by timing solves at the current state (no
        counterpart in the reference).  Returns (bodies_per_tile, per_iteration_launch, ms[5, 2])."""
        bpt, multi = C.c_int(0), C.c_int(0)
        ms = np.zeros(10)
        _lib.check(_lib.load().abd_autotune(self._h, int(repeats), C.byref(bpt), C.byref(multi), dptr(ms)), self._h,
                   "abd_autotune")
        self.pcg_layout = (bpt.value, bool(multi.value))
        return bpt.value, multi.value, ms.reshape(5, 2)

    def set_pcg_layout(self, bodies_per_tile, per_iteration_launch=False):
        """Fix the PCG tile size and launch form (e.g. a result of autotune())."""
        _lib.check(_lib.load().abd_set_pcg_layout(self._h, int(bodies_per_tile), int(bool(per_iteration_launch))),
                   self._h, "abd_set_pcg_layout")
        self.pcg_layout = (int(bodies_per_tile), bool(per_iteration_launch))

    def save_checkpoint(self, path):
        """Write state, frame, parameters and PCG layout to an .npz (colb_amd.checkpoint); a
        solver on the same scene that loads it continues the run bit for bit."""
        from . import checkpoint
        checkpoint.save(self, path)

    def load_checkpoint(self, path):
        """Restore what save_checkpoint wrote (the scene must be the same)."""
        from . import checkpoint
        return checkpoint.load(self, path)

    def log_samples(self, on=True):
        """Start (and clear) or stop logging (features, du) training samples of every solve."""
        _lib.check(_lib.load().abd_log_samples(self._h, int(on)), self._h, "abd_log_samples")

    def samples(self):
        """The logged (features[n, N_FEATURES], du[n, 12]) samples, float32 (colb_amd.learned.N_FEATURES)."""
        from .learned import N_FEATURES
        n = C.c_int(0)
        _lib.check(_lib.load().abd_get_samples(self._h, None, None, 0, C.byref(n)), self._h, "abd_get_samples")
        f = np.zeros((n.value, N_FEATURES), dtype=np.float32)
        d = np.zeros((n.value, 12), dtype=np.float32)
        fp = f.ctypes.data_as(C.POINTER(C.c_float))
        dp = d.ctypes.data_as(C.POINTER(C.c_float))
        _lib.check(_lib.load().abd_get_samples(self._h, fp, dp, n.value, C.byref(n)), self._h, "abd_get_samples")
        return f, d

    def log_steps(self, on=True):
        """Start (and clear) or stop logging step-predictor rows (features at the step's start, the
        step q - q_now) of every unpinned body at every step."""
        _lib.check(_lib.load().abd_log_steps(self._h, int(on)), self._h, "abd_log_steps")

    def step_samples(self):
        """The logged (features[n, N_FEATURES], dq[n, 12]) step rows, float32."""
        from .learned import N_FEATURES
        n = C.c_int(0)
        L = _lib.load()
        _lib.check(L.abd_get_step_samples(self._h, None, None, 0, C.byref(n)), self._h, "abd_get_step_samples")
        f = np.zeros((n.value, N_FEATURES), dtype=np.float32)
        d = np.zeros((n.value, 12), dtype=np.float32)
        fp = f.ctypes.data_as(C.POINTER(C.c_float))
        dp = d.ctypes.data_as(C.POINTER(C.c_float))
        _lib.check(L.abd_get_step_samples(self._h, fp, dp, n.value, C.byref(n)), self._h, "abd_get_step_samples")
        return f, d

    def set_predictor(self, sp):
        """Use a trained colb_amd.learned.StepPredictor for the first iterate of every step when
        params.predictor is 3 (None: no model)."""
        L = _lib.load()
        if sp is None:
            _lib.check(L.abd_set_predictor(self._h, None, None, None, None, None), self._h, "abd_set_predictor")
            self._predictor = None
            return
        self._pred_arrays = [np.ascontiguousarray(a, dtype=np.float32) for a in
                             (sp.feat_mean, sp.feat_std, sp.out_mean, sp.out_std)]
        a = [x.ctypes.data_as(C.POINTER(C.c_float)) for x in self._pred_arrays]
        _lib.check(L.abd_set_predictor(self._h, sp.mlp._h, *a), self._h, "abd_set_predictor")
        self._predictor = sp  # keeps the model alive while the solver uses it

    def set_warm_start(self, ws):
        """Use a trained colb_amd.learned.WarmStart to seed every PCG solve (None: plain PCG)."""
        L = _lib.load()
        if ws is None:
            _lib.check(L.abd_set_warm_start(self._h, None, None, None, None, None), self._h, "abd_set_warm_start")
            self._warm = None
            return
        self._warm_arrays = [np.ascontiguousarray(a, dtype=np.float32) for a in
                             (ws.feat_mean, ws.feat_std, ws.out_mean, ws.out_std)]
        a = [x.ctypes.data_as(C.POINTER(C.c_float)) for x in self._warm_arrays]
        _lib.check(L.abd_set_warm_start(self._h, ws.mlp._h, *a), self._h, "abd_set_warm_start")
        self._warm = ws  # keeps the model alive while the solver uses it

    def set_preconditioner(self, lp):
        """Use a trained colb_amd.learned.LearnedPreconditioner in every PCG solve in place of the
        block-Jacobi inverses (None: block Jacobi)."""
        L = _lib.load()
        if lp is None:
            _lib.check(L.abd_set_preconditioner(self._h, None, None, None, None, None), self._h,
                       "abd_set_preconditioner")
            self._precond = None
            return
        self._precond_arrays = [np.ascontiguousarray(a, dtype=np.float32) for a in
                                (lp.feat_mean, lp.feat_std, lp.out_mean, lp.out_std)]
        a = [x.ctypes.data_as(C.POINTER(C.c_float)) for x in self._precond_arrays]
        _lib.check(L.abd_set_preconditioner(self._h, lp.mlp._h, *a), self._h, "abd_set_preconditioner")
        self._precond = lp  # keeps the model alive while the solver uses it

    def enable_timing(self, on=True):
        """HIP-event timing of the step's stages: False/0 off, True/1 every stage, 2 the PCG solve
        launches only (no host synchronisation; light enough for a timed run)."""
        _lib.check(_lib.load().abd_enable_timing(self._h, int(on)), self._h, "abd_enable_timing")

    def kernel_times(self, cap=64):
        names = C.create_string_buffer(4096)
        ms, by = np.zeros(cap), np.zeros(cap)
        calls = np.zeros(cap, dtype=np.int32)
        n = C.c_int(0)
        _lib.load().abd_get_kernel_times(self._h, names, 4096, dptr(ms), iptr(calls), dptr(by), cap, C.byref(n))
        out = {}
        for k, name in enumerate(names.value.decode().split("\n")[:n.value]):
            out[name] = {"ms": float(ms[k]), "calls": int(calls[k]), "bytes": float(by[k])}
        return out
