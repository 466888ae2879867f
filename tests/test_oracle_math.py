"""Pins the oracle's pointwise functions with independent checks: closed forms, central finite
differences, brute-force minimisation, numpy.linalg.eigh and dense sampling.  A plausible
mistake in a formula (dropped term, wrong sign, transposed operand, wrong slot) fails here."""
import zlib

import numpy as np
import pytest
from scipy.optimize import minimize

PT, EE = 0, 1
D_PP, D_PE, D_PT, D_EE = 0, 1, 2, 3


def fd_grad(f, x, h=1e-6):
    g = np.zeros_like(x)
    for k in range(len(x)):
        e = np.zeros_like(x)
        e[k] = h
        g[k] = (f(x + e) - f(x - e)) / (2 * h)
    return g


# ---------------------------------------------------------------------------------------------
# barrier.py:47-66
# ---------------------------------------------------------------------------------------------
def test_barrier_closed_form_and_derivatives(oracle_mod):
    shat = 0.095 ** 2
    for s in np.linspace(0.0, 1.2 * shat, 25):
        B, B1, B2 = oracle_mod.barrier(s, shat)
        if s < shat:
            assert B == pytest.approx(-(s - shat) ** 3, rel=1e-14, abs=1e-30)
            h = 1e-9 * shat
            if s > h and s + h < shat:
                assert B1 == pytest.approx((oracle_mod.barrier(s + h, shat)[0] - oracle_mod.barrier(s - h, shat)[0]) / (2 * h), rel=1e-6)
                assert B2 == pytest.approx((oracle_mod.barrier(s + h, shat)[1] - oracle_mod.barrier(s - h, shat)[1]) / (2 * h), rel=1e-6)
            assert B1 <= 0 and B2 >= 0
        else:
            assert B == 0 and B1 == 0 and B2 == 0


# ---------------------------------------------------------------------------------------------
# ortho.py:9-41
# ---------------------------------------------------------------------------------------------
def test_ortho_zero_at_rotation(oracle_mod):
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    q = np.concatenate([rng.normal(size=3), Q.T.reshape(-1)])
    e, g, H = oracle_mod.ortho(q, 1e7)
    assert abs(e) < 1e-15 * 1e7 and np.abs(g).max() < 1e-7


def test_ortho_closed_form_and_fd(oracle_mod):
    rng = np.random.default_rng(1)
    kappa = 1e5
    for _ in range(5):
        A = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        q = np.concatenate([rng.normal(size=3), A.reshape(-1)])  # rows a_i = A[i]
        e, g, H = oracle_mod.ortho(q, kappa)
        G = A @ A.T
        assert e == pytest.approx(kappa * np.sum((G - np.eye(3)) ** 2), rel=1e-12)
        f = lambda a9: oracle_mod.ortho(np.concatenate([q[:3], a9]), kappa)[0]
        np.testing.assert_allclose(g, fd_grad(f, q[3:], 1e-6), rtol=1e-6, atol=1e-6 * kappa)
        gf = lambda a9: oracle_mod.ortho(np.concatenate([q[:3], a9]), kappa)[1]
        Hfd = np.stack([fd_grad(lambda x, k=k: gf(x)[k], q[3:], 1e-6) for k in range(9)])
        np.testing.assert_allclose(H, Hfd, rtol=1e-5, atol=1e-5 * kappa)
        np.testing.assert_allclose(H, H.T, rtol=1e-12, atol=1e-9)


# ---------------------------------------------------------------------------------------------
# distances: classification, value (brute force), gradient and Hessian (finite differences)
# ---------------------------------------------------------------------------------------------
def brute_pt(x):
    """min over the triangle via the square -> triangle map (u, v) = (s (1 - t), s t), so every
    iterate is feasible; multi-start L-BFGS-B."""
    p, a, b, c = x.reshape(4, 3)
    def f(st):
        s, t = st
        y = a + s * (1 - t) * (b - a) + s * t * (c - a)
        return np.sum((p - y) ** 2)
    best = None
    for s0 in np.linspace(0, 1, 7):
        for t0 in np.linspace(0, 1, 7):
            r = minimize(f, [s0, t0], method="L-BFGS-B", bounds=[(0, 1), (0, 1)], options={"ftol": 1e-18, "gtol": 1e-14})
            if best is None or r.fun < best:
                best = r.fun
    return best


def brute_ee(x):
    a0, a1, b0, b1 = x.reshape(4, 3)
    f = lambda st: np.sum((a0 + st[0] * (a1 - a0) - b0 - st[1] * (b1 - b0)) ** 2)
    best = None
    for s0 in np.linspace(0, 1, 6):
        for t0 in np.linspace(0, 1, 6):
            r = minimize(f, [s0, t0], method="L-BFGS-B", bounds=[(0, 1), (0, 1)], options={"ftol": 1e-18, "gtol": 1e-14})
            if best is None or r.fun < best:
                best = r.fun
    return best


CASES = {
    # name: (kind, points, expected distance type)
    "pt_face": (PT, [[0.2, 0.3, 0.5], [0, 0, 0], [1, 0, 0], [0, 1, 0]], D_PT),
    "pt_edge": (PT, [[0.5, -0.3, 0.4], [0, 0, 0], [1, 0, 0], [0, 1, 0]], D_PE),
    "pt_edge_bc": (PT, [[0.7, 0.7, -0.3], [0, 0, 0], [1, 0, 0], [0, 1, 0]], D_PE),
    "pt_vertex": (PT, [[-0.3, -0.2, 0.4], [0, 0, 0], [1, 0, 0], [0, 1, 0]], D_PP),
    "ee_interior": (EE, [[0, 0, 0], [1, 0, 0], [0.4, -0.5, 0.3], [0.5, 0.6, 0.35]], D_EE),
    "ee_point_edge": (EE, [[0, 0, 0], [1, 0, 0], [1.3, -0.5, 0.2], [1.4, 0.5, 0.25]], D_PE),
    "ee_edge_point": (EE, [[1.3, -0.5, 0.2], [1.4, 0.5, 0.25], [0, 0, 0], [1, 0, 0]], D_PE),
    "ee_point_point": (EE, [[0, 0, 0], [1, 0, 0], [1.3, 0.2, 0.2], [2, 1, 0.5]], D_PP),
}


@pytest.mark.parametrize("name", list(CASES))
def test_pair_distance_value_type_and_derivatives(oracle_mod, name):
    kind, pts, want = CASES[name]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    x0 = np.array(pts, dtype=float).reshape(-1)
    for trial in range(3):
        x = x0 + 0.01 * rng.normal(size=12) * (trial > 0)
        d2, g, H, t = oracle_mod.pair_dist2(kind, x)
        assert t == want, (name, t)
        ref = brute_pt(x) if kind == PT else brute_ee(x)
        assert d2 == pytest.approx(ref, rel=1e-7, abs=1e-12)
        f = lambda y: oracle_mod.pair_dist2(kind, y, derivs=False)[0]
        np.testing.assert_allclose(g, fd_grad(f, x, 1e-6), rtol=1e-6, atol=1e-8)
        gk = lambda y: oracle_mod.pair_dist2(kind, y)[1]
        Hfd = np.stack([fd_grad(lambda y, k=k: gk(y)[k], x, 1e-6) for k in range(12)])
        np.testing.assert_allclose(H, Hfd, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(H, H.T, atol=1e-10)
        # translation invariance: gradient blocks sum to zero
        np.testing.assert_allclose(g.reshape(4, 3).sum(0), 0, atol=1e-10)


def test_pair_distance_random_against_brute_force(oracle_mod):
    rng = np.random.default_rng(7)
    seen = set()
    for _ in range(60):
        kind = int(rng.integers(0, 2))
        x = rng.normal(size=12)
        d2, t = oracle_mod.pair_dist2(kind, x, derivs=False)
        seen.add((kind, t))
        ref = brute_pt(x) if kind == PT else brute_ee(x)
        assert d2 == pytest.approx(ref, rel=1e-7, abs=1e-12)
    assert {(PT, D_PT), (PT, D_PE), (EE, D_EE)} <= seen


def test_parallel_edges_fall_back_to_endpoint(oracle_mod):
    x = np.array([0, 0, 0, 1, 0, 0, 0.2, 0.1, 0, 0.9, 0.1, 0], dtype=float)
    d2, t = oracle_mod.pair_dist2(EE, x, derivs=False)
    assert t in (D_PE, D_PP)
    assert d2 == pytest.approx(0.01, rel=1e-12)


def test_segment_params_match_brute_force(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(30):
        x = rng.normal(size=12)
        s, t = oracle_mod.seg_params(x)
        a0, a1, b0, b1 = x.reshape(4, 3)
        d2 = np.sum((a0 + s * (a1 - a0) - b0 - t * (b1 - b0)) ** 2)
        assert d2 == pytest.approx(brute_ee(x), rel=1e-7, abs=1e-12)


# ---------------------------------------------------------------------------------------------
# PSD projection and the 12x12 Cholesky inverse
# ---------------------------------------------------------------------------------------------
def eig_projection(A, absolute=False):
    """(A + |A|) / 2 (or |A|) of a symmetric matrix from its eigendecomposition."""
    w, V = np.linalg.eigh(A)
    return (V * (np.abs(w) if absolute else np.maximum(w, 0))) @ V.T


def psd_cases(m, seed=0, n=24):
    """Symmetric m x m test matrices for a PSD projection: eigenvalues spread over 12 decades, half
    of them zero (a barrier Hessian's null space), definite and indefinite, at scales 1e-6 .. 1e12."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        Q, _ = np.linalg.qr(rng.normal(size=(m, m)))
        w = rng.normal(size=m) * 10.0 ** rng.uniform(-12, 0, size=m)
        if k % 3 == 0:
            w[: m // 2] = 0.0
        if k % 4 == 1:
            w = np.abs(w)
        if k % 4 == 2:
            w = -np.abs(w)
        A = (Q * w) @ Q.T * 10.0 ** rng.uniform(-6, 12)
        out.append(0.5 * (A + A.T))
    return np.ascontiguousarray(np.array(out))


def hip_psd_project(lib, A, precision=0, absolute=False):
    """The assembly's PSD projection (abd_psd_project_test) on matrices A[n, m, m]."""
    import ctypes as C
    A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros_like(A)
    P = C.POINTER(C.c_double)
    rc = lib.abd_psd_project_test(A.shape[1], precision, int(absolute), A.shape[0], A.ctypes.data_as(P),
                                  out.ctypes.data_as(P))
    assert rc == 0
    return out


@pytest.mark.parametrize("n", [3, 6, 9, 12])
def test_psd_projection_matches_eigh(oracle_mod, n):
    rng = np.random.default_rng(n)
    for scale in (1.0, 1e-6, 1e9):
        A = rng.normal(size=(n, n))
        A = scale * (A + A.T)
        P = oracle_mod.psd_project(A)
        w, V = np.linalg.eigh(A)
        ref = (V * np.maximum(w, 0)) @ V.T
        np.testing.assert_allclose(P, ref, rtol=1e-10, atol=1e-12 * scale * n)
        assert np.linalg.eigvalsh(P).min() > -1e-10 * scale


def test_psd_projection_keeps_psd_matrix(oracle_mod):
    rng = np.random.default_rng(11)
    B = rng.normal(size=(12, 5))
    A = B @ B.T
    np.testing.assert_allclose(oracle_mod.psd_project(A), A, rtol=1e-11, atol=1e-11)


def test_chol_inverse12(oracle_mod):
    rng = np.random.default_rng(5)
    B = rng.normal(size=(12, 12))
    A = B @ B.T + 12 * np.eye(12)
    rc, Ai = oracle_mod.chol_inverse12(A)
    assert rc == 0
    np.testing.assert_allclose(Ai, np.linalg.inv(A), rtol=1e-10, atol=1e-12)
    rc, _ = oracle_mod.chol_inverse12(-A)
    assert rc != 0


# ---------------------------------------------------------------------------------------------
# additive CCD
# ---------------------------------------------------------------------------------------------
def true_toi(oracle_mod, kind, x0, x1, n=4000):
    ts = np.linspace(0, 1, n + 1)
    d = np.array([oracle_mod.pair_dist2(kind, x0 + t * (x1 - x0), derivs=False)[0] for t in ts])
    hit = np.nonzero(d < 1e-14)[0]
    return ts[hit[0]] if len(hit) else None, np.sqrt(d)


def test_accd_head_on_closed_form(oracle_mod):
    # point falling straight onto a static triangle: TOI = d0 / v
    x0 = np.array([0.2, 0.2, 0.5, 0, 0, 0, 1, 0, 0, 0, 1, 0], dtype=float)
    x1 = x0.copy()
    x1[2] = -0.5
    s = 0.1
    t = oracle_mod.accd(PT, x0, x1, s=s)
    assert t == pytest.approx((1 - s) * 0.5, rel=1e-12)


def test_accd_is_conservative(oracle_mod):
    rng = np.random.default_rng(9)
    hits = 0
    for _ in range(40):
        kind = int(rng.integers(0, 2))
        x0 = rng.normal(size=12)
        x1 = x0 + 1.5 * rng.normal(size=12)
        t = oracle_mod.accd(kind, x0, x1, s=0.1, tmax=1.0, max_iter=100000)
        toi, d = true_toi(oracle_mod, kind, x0, x1)
        d0 = d[0]
        dt = np.sqrt(oracle_mod.pair_dist2(kind, x0 + t * (x1 - x0), derivs=False)[0])
        # the gap at the returned time keeps >= s * d0, and no crossing happens before it
        assert dt >= 0.1 * d0 * (1 - 1e-9)
        ts = np.linspace(0, 1, len(d))
        assert d[ts <= t].min() > 0.0
        if toi is not None:
            hits += 1
            assert t <= toi
    assert hits >= 0
