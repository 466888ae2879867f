"""Learned warm start, learned block preconditioner and learned step predictor of the ABD Newton
step (no counterpart in the reference).

A per-body MLP maps body features to the body's Newton step du (the 12 DOFs of the PCG
solution), which then seeds the PCG as its initial guess.  The model runs on the GPU through
the C ABI of include/abd_mlp.h: BF16 weights read by the matrix cores with FP32 accumulation,
FP32 master weights trained by Adam, forward and backward written by hand (abd_mlp.hip).

Features per body (49, standardised with the training set's mean and standard deviation, then
clipped to +-8 so that a state far outside the training set cannot blow the output up):
    q - q~ (12)          the DOFs relative to the inertial target q~ = q_now + dt qdot + dt^2 g
    qdot (12)            the DOF velocities
    n_contacts (1)       active contacts incident to the body
    g_contact (12)       the summed barrier gradient on the body
    z (12)               the block-Jacobi step D_b^-1 grad_b (D_b the body's 12x12 diagonal block)
Target per body (12): du - z, standardised the same way; the warm start is x0 = z + prediction.

The step predictor (StepPredictor, AbdParams.predictor = 3) reads the same 49 features at the
step's start q_now and predicts the step's correction over the extrapolation, dq - dt qdot; the
first Newton iterate moves toward q_now + dt qdot + prediction (a CCD-safe, energy-decreasing
fraction of the way).
"""
import ctypes as C

import numpy as np

from . import _lib

N_FEATURES = 49
FEAT_Z = 37
FEAT_CLIP = 8.0
N_OUT = 12
DEFAULT_DIMS = (N_FEATURES, 128, 128, 128, N_OUT)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def init_params(dims, seed=0):
    """He-normal weights, zero biases (float32, flat layout of include/abd_mlp.h)."""
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(len(dims) - 1):
        parts.append((rng.normal(size=(dims[l], dims[l + 1])) * np.sqrt(2.0 / dims[l])).astype(np.float32).reshape(-1))
        parts.append(np.zeros(dims[l + 1], dtype=np.float32))
    return np.concatenate(parts)


class Mlp:
    """The GPU MLP (abd_mlp.hip) behind include/abd_mlp.h."""

    def __init__(self, dims=DEFAULT_DIMS, seed=0, params=None):
        self.dims = tuple(int(d) for d in dims)
        self._h = C.c_void_p()
        d = np.asarray(self.dims, dtype=np.int32)
        L = _lib.load()
        rc = L.abd_mlp_create(_iptr(d), len(self.dims) - 1, C.byref(self._h))
        self._check(rc, "abd_mlp_create")
        self.n_params = L.abd_mlp_num_params(self._h)
        self.set_params(init_params(self.dims, seed) if params is None else params)
        self.set_loss_weights(np.ones(self.dims[-1], dtype=np.float32))

    def _check(self, rc, what):
        if rc != 0:
            msg = _lib.load().abd_mlp_last_error(self._h).decode() if self._h else ""
            raise _lib.AbdError(f"{what} failed with status {rc}: {msg}")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().abd_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.n_params
        self._check(_lib.load().abd_mlp_set_params(self._h, _fptr(flat)), "abd_mlp_set_params")

    def get_params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(_lib.load().abd_mlp_get_params(self._h, _fptr(out)), "abd_mlp_get_params")
        return out

    def set_loss_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(_lib.load().abd_mlp_set_loss_weights(self._h, _fptr(w)), "abd_mlp_set_loss_weights")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.dims[-1]), dtype=np.float32)
        self._check(_lib.load().abd_mlp_forward(self._h, _fptr(X), X.shape[0], _fptr(Y)), "abd_mlp_forward")
        return Y

    def forward_static(self, X):
        """forward() through abd_mlp_forward_static: no device allocation (static scratch per GPU)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.dims[-1]), dtype=np.float32)
        self._check(_lib.load().abd_mlp_forward_static(self._h, _fptr(X), X.shape[0], _fptr(Y)),
                    "abd_mlp_forward_static")
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = C.c_float(0.0)
        g = np.zeros(self.n_params, dtype=np.float32)
        self._check(_lib.load().abd_mlp_gradient(self._h, _fptr(X), _fptr(T), X.shape[0], C.byref(loss), _fptr(g)),
                    "abd_mlp_gradient")
        return float(loss.value), g

    def train(self, X, T, steps, batch=256, lr=1e-3, seed=0, beta1=0.9, beta2=0.999, eps=1e-8):
        """Adam over `steps` minibatches drawn (seeded) from the rows of X, T; returns the batch losses."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        n = X.shape[0]
        rng = np.random.default_rng(seed)
        losses = []
        done = 0
        while done < steps:
            k = min(4096, steps - done)
            perm = np.concatenate([rng.permutation(n) for _ in range((k * batch) // n + 1)])[:k * batch]
            perm = np.ascontiguousarray(perm, dtype=np.int32)
            out = np.zeros(k, dtype=np.float32)
            self._check(_lib.load().abd_mlp_train(self._h, _fptr(X), _fptr(T), n, _iptr(perm), k, batch, lr, beta1,
                                                  beta2, eps, _fptr(out)), "abd_mlp_train")
            losses.append(out)
            done += k
        return np.concatenate(losses)

    def gemm_test(self, A, B):
        A = np.ascontiguousarray(A, dtype=np.float32)
        B = np.ascontiguousarray(B, dtype=np.float32)
        Cm = np.zeros((A.shape[0], B.shape[1]), dtype=np.float32)
        self._check(_lib.load().abd_mlp_gemm_test(self._h, A.shape[0], B.shape[1], A.shape[1], _fptr(A), _fptr(B),
                                                  _fptr(Cm)), "abd_mlp_gemm_test")
        return Cm


class WarmStart:
    """The trained warm-start model with its feature/target standardisation."""

    def __init__(self, dims=DEFAULT_DIMS, seed=0):
        self.dims = tuple(dims)
        self.mlp = Mlp(self.dims, seed=seed)
        self.feat_mean = np.zeros(N_FEATURES, dtype=np.float32)
        self.feat_std = np.ones(N_FEATURES, dtype=np.float32)
        self.out_mean = np.zeros(N_OUT, dtype=np.float32)
        self.out_std = np.ones(N_OUT, dtype=np.float32)

    def _inputs(self, feat):
        X = (np.asarray(feat, dtype=np.float64) - self.feat_mean) / self.feat_std
        return np.clip(X, -FEAT_CLIP, FEAT_CLIP).astype(np.float32)

    def fit(self, feat, du, steps=3000, batch=512, lr=1e-3, seed=0):
        """Standardise, then train with Adam on the GPU; returns the batch losses."""
        feat = np.asarray(feat, dtype=np.float64)
        t = np.asarray(du, dtype=np.float64) - feat[:, FEAT_Z:FEAT_Z + N_OUT]
        self.feat_mean = feat.mean(0).astype(np.float32)
        self.feat_std = np.maximum(feat.std(0), 1e-12).astype(np.float32)
        self.out_mean = t.mean(0).astype(np.float32)
        self.out_std = np.maximum(t.std(0), 1e-12).astype(np.float32)
        X = self._inputs(feat)
        T = ((t - self.out_mean) / self.out_std).astype(np.float32)
        return self.mlp.train(X, T, steps=steps, batch=min(batch, X.shape[0]), lr=lr, seed=seed)

    def predict(self, feat):
        """The warm start x0 = z + model output for each row of raw features."""
        feat = np.asarray(feat, dtype=np.float64)
        return feat[:, FEAT_Z:FEAT_Z + N_OUT] + self.mlp.forward(self._inputs(feat)) * self.out_std + self.out_mean

    def save(self, path):
        np.savez(path, dims=np.asarray(self.dims), params=self.mlp.get_params(), feat_mean=self.feat_mean,
                 feat_std=self.feat_std, out_mean=self.out_mean, out_std=self.out_std)

    @classmethod
    def load(cls, path):
        z = np.load(path)
        ws = cls(tuple(int(d) for d in z["dims"]))
        ws.mlp.set_params(z["params"])
        for k in ("feat_mean", "feat_std", "out_mean", "out_std"):
            setattr(ws, k, z[k].astype(np.float32))
        return ws


class StepPredictor(WarmStart):
    """The trained step-predictor model: features at the step's start -> dq - dt qdot."""

    def fit(self, feat, dq, dt, steps=3000, batch=512, lr=1e-3, seed=0):
        """Standardise, then train with Adam on the GPU; returns the batch losses.  feat, dq: the
        rows the solver logs with log_steps (features at the step's start, the step q - q_now)."""
        feat = np.asarray(feat, dtype=np.float64)
        t = np.asarray(dq, dtype=np.float64) - dt * feat[:, 12:24]
        self.feat_mean = feat.mean(0).astype(np.float32)
        self.feat_std = np.maximum(feat.std(0), 1e-12).astype(np.float32)
        self.out_mean = t.mean(0).astype(np.float32)
        self.out_std = np.maximum(t.std(0), 1e-12).astype(np.float32)
        X = self._inputs(feat)
        T = ((t - self.out_mean) / self.out_std).astype(np.float32)
        return self.mlp.train(X, T, steps=steps, batch=min(batch, X.shape[0]), lr=lr, seed=seed)

    def predict(self, feat, dt):
        """The predicted step dq = dt qdot + model output for each row of raw features."""
        feat = np.asarray(feat, dtype=np.float64)
        return dt * feat[:, 12:24] + self.mlp.forward(self._inputs(feat)) * self.out_std + self.out_mean


N_PRECOND = 78
PRECOND_DIMS = (N_FEATURES, 128, 128, 128, N_PRECOND)
_TRIL = np.tril_indices(12, -1)


def precond_target(D, Z):
    """Target outputs of the learned preconditioner for one body: with D = L L^T its diagonal block
    and Z the body's diagonal block of H^-1, T = chol(L^T Z L) (T = I when the body is uncoupled);
    returns [log diag T (12), T below the diagonal row by row (66)]."""
    L = np.linalg.cholesky(D)
    T = np.linalg.cholesky(L.T @ Z @ L)
    return np.concatenate([np.log(np.diag(T)), T[_TRIL]])


def precond_matrix(D, out):
    """P = L^-T T T^T L^-1 from the 78 (de-standardised) outputs, as the device builds it."""
    L = np.linalg.cholesky(D)
    T = np.diag(np.exp(out[:12]))
    T[_TRIL] = out[12:]
    X = np.linalg.solve(L.T, T)
    return X @ X.T


def precond_samples(solver, steps):
    """(features, targets) for the learned preconditioner from `steps` steps of a small scene: at
    the start of every step the assembled Hessian is inverted densely (12 nb x 12 nb) and each
    unpinned body's diagonal block of H^-1 gives its target.  Advances the solver by `steps`."""
    F, T = [], []
    pinned = np.asarray(solver.scene.body_pinned, dtype=bool)
    for _ in range(steps):
        g, D, pij, off, _e = solver.assemble()
        nb = D.shape[0]
        H = np.zeros((12 * nb, 12 * nb))
        for b in range(nb):
            H[12 * b:12 * b + 12, 12 * b:12 * b + 12] = D[b]
        for (i, j), B in zip(pij, off):
            H[12 * i:12 * i + 12, 12 * j:12 * j + 12] = B
            H[12 * j:12 * j + 12, 12 * i:12 * i + 12] = B.T
        Hi = np.linalg.inv(H)
        solver.log_samples(True)
        solver.compute_du()
        f, _ = solver.samples()
        solver.log_samples(False)
        rows = [b for b in range(nb) if not pinned[b]]
        if len(rows) == len(f):
            for k, b in enumerate(rows):
                Z = Hi[12 * b:12 * b + 12, 12 * b:12 * b + 12]
                Zs = 0.5 * (Z + Z.T)
                F.append(f[k])
                T.append(precond_target(D[b], Zs))
        solver.step()
    return np.asarray(F, dtype=np.float32), np.asarray(T, dtype=np.float32)


class LearnedPreconditioner:
    """The trained block-preconditioner model with its feature/target standardisation."""

    def __init__(self, dims=PRECOND_DIMS, seed=0):
        self.dims = tuple(dims)
        self.mlp = Mlp(self.dims, seed=seed)
        self.feat_mean = np.zeros(N_FEATURES, dtype=np.float32)
        self.feat_std = np.ones(N_FEATURES, dtype=np.float32)
        self.out_mean = np.zeros(N_PRECOND, dtype=np.float32)
        self.out_std = np.ones(N_PRECOND, dtype=np.float32)

    def _inputs(self, feat):
        X = (np.asarray(feat, dtype=np.float64) - self.feat_mean) / self.feat_std
        return np.clip(X, -FEAT_CLIP, FEAT_CLIP).astype(np.float32)

    def fit(self, feat, target, steps=3000, batch=512, lr=1e-3, seed=0):
        """Standardise, then train with Adam on the GPU; returns the batch losses."""
        feat = np.asarray(feat, dtype=np.float64)
        t = np.asarray(target, dtype=np.float64)
        self.feat_mean = feat.mean(0).astype(np.float32)
        self.feat_std = np.maximum(feat.std(0), 1e-12).astype(np.float32)
        self.out_mean = t.mean(0).astype(np.float32)
        self.out_std = np.maximum(t.std(0), 1e-12).astype(np.float32)
        X = self._inputs(feat)
        T = ((t - self.out_mean) / self.out_std).astype(np.float32)
        return self.mlp.train(X, T, steps=steps, batch=min(batch, X.shape[0]), lr=lr, seed=seed)

    def predict(self, feat):
        """De-standardised outputs (78 per row)."""
        return self.mlp.forward(self._inputs(feat)) * self.out_std + self.out_mean
