"""Pins the oracle's point-ground contact (the static plane n.x = h of the north-star scenes): the
energy against a numpy restatement, the gradient against central differences of the energy, the
Hessian against differences of the gradient where the projection leaves it alone (d > dhat/sqrt 5)
and against a numpy restatement with the clamp where it does not, and the infeasible side of the
plane (infinite energy, the reported error bit, the CCD bound of a vertex on or below the plane).
"""
import numpy as np
import pytest

from colb_amd import scene as S
from colb_amd.params import north_star

DHAT = 0.05


def cube_on_ground(gap, tilt_deg=(3.0, 0.0, 2.0), density=1e3, seed=0, n=1):
    """n unit cubes side by side, slightly rotated, lowest vertex `gap` above the plane y = 0."""
    V, F = S.cube_mesh()
    E = S.edges_of_tris(F)
    M = S.mass4_from_mesh(V, F, density)
    b = S._Builder()
    rng = np.random.default_rng(seed)
    for i in range(n):
        R = S.rot_euler_xyz(np.array(tilt_deg) + rng.uniform(-1, 1, 3))
        b.add(V, F, E, M, S.q_from_pose([1.5 * i, 1.0, 0.0], R))
    sc = b.build("cube_on_ground")
    y = sc.world_vertices()[:, 1]
    sc.q0[:, 1] -= y.min() - gap
    return sc


def ground_params(**kw):
    base = dict(dhat=DHAT, k_contact=1e12, kappa=1e5, ground_enabled=True, cg_tol=1e-12, cg_max_iter=500)
    base.update(kw)
    return north_star(**base)


def numpy_energy(sc, p, q, qdot, orc):
    """Incremental potential restated from the pointwise definitions (barrier.py:47-66, ortho.py:9-16,
    abd.py:53-75) with the ground pairs found here, not by the oracle."""
    qt = sc.q0 + p.dt * qdot
    qt[:, :3] += p.dt ** 2 * np.array(p.gravity)
    e = 0.0
    for b in range(sc.n_bodies):
        dq = (q[b] - qt[b]).reshape(4, 3)
        e += 0.5 * np.sum(sc.mass4[b] * (dq @ dq.T))
        A = q[b, 3:].reshape(3, 3)
        e += p.dt ** 2 * p.kappa * np.sum((A @ A.T - np.eye(3)) ** 2)
    X = sc.world_vertices(q)
    n = np.array(p.ground_normal) / np.linalg.norm(p.ground_normal)
    d = X @ n - p.ground_offset
    shat = p.dhat ** 2
    g = d[d < p.dhat]
    assert np.all(g > 0)
    e += np.sum(p.dt ** 2 * p.k_contact * -(g * g - shat) ** 3)
    # body-body pairs from the oracle's detection (their own restatement is test_oracle_scene's)
    pairs, d2 = orc.detect(q)
    bb = pairs[:, 0] != 2
    e += np.sum(p.dt ** 2 * p.k_contact * -(d2[bb] - shat) ** 3)
    return e, int(np.sum(d < p.dhat))


@pytest.mark.parametrize("gap", [0.3 * DHAT, 0.7 * DHAT])
def test_ground_energy_matches_numpy_restatement(oracle_mod, gap):
    sc = cube_on_ground(gap, n=2)
    p = ground_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(3)
    q = sc.q0 + 1e-4 * rng.normal(size=sc.q0.shape)
    qdot = 0.1 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, qdot)
    E = orc.energy()[0]
    Eref, nground = numpy_energy(sc, p, q, qdot, orc)
    assert nground > 0
    assert np.sum(orc.detect()[0][:, 0] == 2) == nground
    assert E == pytest.approx(Eref, rel=1e-12)


@pytest.mark.parametrize("gap", [0.3 * DHAT, 0.7 * DHAT])
def test_ground_gradient_matches_energy_differences(oracle_mod, gap):
    sc = cube_on_ground(gap)
    p = ground_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(4)
    q = sc.q0 + 1e-4 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, 0.1 * rng.normal(size=sc.q0.shape))
    g = orc.assemble()[0].reshape(-1)
    assert np.sum(orc.detect()[0][:, 0] == 2) > 0
    qf = q.reshape(-1)
    h = 1e-7
    gfd = np.array([(orc.energy(qf + h * e)[0] - orc.energy(qf - h * e)[0]) / (2 * h) for e in np.eye(len(qf))])
    np.testing.assert_allclose(g, gfd, rtol=2e-5, atol=1e-5 * np.abs(gfd).max())


def test_ground_hessian_matches_gradient_differences(oracle_mod):
    """Every ground contact at d > dhat/sqrt(5): the point-plane Hessian kd (2 B' + 4 B'' d^2) n n^T is
    positive and the projection is the identity, so the assembled Hessian is the derivative of the
    assembled gradient (the orthogonality term is PSD at a rotation)."""
    sc = cube_on_ground(0.7 * DHAT)
    p = ground_params()
    orc = oracle_mod.Oracle(sc, p)
    q = sc.q0.copy()
    orc.set_state(q, sc.q0, np.zeros_like(q))
    d = sc.world_vertices(q)[:, 1]
    act = d[d < DHAT]
    assert len(act) > 0 and np.all(act > DHAT / np.sqrt(5.0))
    _, H = orc.dense_hessian()
    qf = q.reshape(-1)
    h = 1e-7
    cols = []
    for e in np.eye(len(qf)):
        orc.set_state(qf + h * e)
        gp = orc.assemble()[0].reshape(-1)
        orc.set_state(qf - h * e)
        gm = orc.assemble()[0].reshape(-1)
        cols.append((gp - gm) / (2 * h))
    Hfd = np.array(cols).T
    np.testing.assert_allclose(H, Hfd, rtol=1e-5, atol=1e-6 * np.abs(H).max())


def test_ground_hessian_matches_numpy_restatement(oracle_mod):
    """Ground contacts on both sides of dhat/sqrt(5) (some eigenvalues clamped): the assembled
    Hessian equals inertia + projected ortho + sum over ground vertices of J^T P(H_x) J, with H_x the
    closed form above and P the eigh clamp of the 3x3."""
    sc = cube_on_ground(0.2 * DHAT, tilt_deg=(1.2, 0.0, 0.8))
    p = ground_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(5)
    q = sc.q0 + 1e-4 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, np.zeros_like(q))
    _, H = orc.dense_hessian()
    psd = lambda A: (lambda w, V: (V * np.maximum(w, 0)) @ V.T)(*np.linalg.eigh(A))
    n12 = 12 * sc.n_bodies
    Href = np.zeros((n12, n12))
    for b in range(sc.n_bodies):
        _, _, Ho = oracle_mod.ortho(q[b], p.kappa)
        blk = np.kron(sc.mass4[b], np.eye(3))
        blk[3:, 3:] += p.dt ** 2 * psd(Ho)
        Href[12 * b:12 * b + 12, 12 * b:12 * b + 12] = blk
    X = sc.world_vertices(q)
    nrm = np.array([0.0, 1.0, 0.0])
    kd = p.k_contact * p.dt ** 2
    shat = p.dhat ** 2
    clamped = unclamped = 0
    for v in range(sc.n_verts):
        d = X[v] @ nrm
        if d >= p.dhat:
            continue
        s = d * d
        B1, B2 = -3 * (s - shat) ** 2, -6 * (s - shat)
        # second derivative of B(d^2) along n: 2 B' + 4 B'' d^2, times n n^T
        Hx = kd * (B1 * 2 * np.outer(nrm, nrm) + B2 * np.outer(2 * d * nrm, 2 * d * nrm))
        lam = kd * (2 * B1 + 4 * B2 * s)
        clamped += lam < 0
        unclamped += lam > 0
        b = sc.vert_body[v]
        J = np.zeros((3, n12))
        J[:, 12 * b:12 * b + 12] = np.kron(np.concatenate([[1.0], sc.rest_x[v]])[None, :], np.eye(3))
        Href += J.T @ psd(Hx) @ J
    assert clamped > 0 and unclamped > 0
    np.testing.assert_allclose(H, Href, rtol=1e-9, atol=1e-9 * np.abs(Href).max())


def test_vertex_on_or_below_the_plane_is_infeasible(oracle_mod):
    sc = cube_on_ground(0.5 * DHAT)
    p = ground_params()
    orc = oracle_mod.Oracle(sc, p)
    assert np.isfinite(orc.energy()[0])
    for depth in (0.0, 1e-18, 1e-3):
        q = sc.q0.copy()
        q[:, 1] -= 0.5 * DHAT + depth  # lowest vertex at -depth
        assert orc.energy(q)[0] == np.inf, depth
        orc.take_error()
        orc.assemble(q)
        assert orc.take_error() == 64, depth
    orc.assemble(sc.q0)
    assert orc.take_error() == 0


@pytest.mark.parametrize("d0,down,expect", [(0.5 * DHAT, 1.0, "gap"), (0.0, 1.0, 0.0), (-1e-16, 1.0, 0.0),
                                            (0.0, -1.0, 1.0), (0.5 * DHAT, 0.0, 1.0)])
def test_ground_ccd_bound(oracle_mod, d0, down, expect):
    """A step du that lowers the cube by `down` (q - alpha du), the plane placed so that the lowest
    vertex is d0 above it: the bound keeps 0.1 of the gap, is 0 for a vertex on or below the plane
    moving down, and 1 for motion away from the plane."""
    sc = cube_on_ground(0.5 * DHAT, tilt_deg=(0.0, 0.0, 0.0))
    y = sc.world_vertices()[:, 1]
    p = ground_params(ground_offset=float(y.min() - d0))
    d = y - p.ground_offset  # the oracle's arithmetic for n = (0, 1, 0)
    assert d.min() == pytest.approx(d0, abs=1e-17)
    orc = oracle_mod.Oracle(sc, p)
    du = np.zeros_like(sc.q0)
    du[:, 1] = down
    a = orc.ccd(du)[0]
    want = (1 - p.ccd_gap_ratio) * d.min() / down if expect == "gap" else expect
    assert a == pytest.approx(want, rel=1e-12, abs=0.0)
