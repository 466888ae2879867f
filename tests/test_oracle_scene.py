"""Pins the oracle's scene-level functions: energy and assembled gradient (central differences of
the energy), the assembled projected Hessian (an independent numpy restatement from the
pointwise functions, eigh projection and the affine Jacobian kron), PCG (dense solve), the
Newton step (free-fall closed form, energy decrease, pinned bodies fixed), mass matrices
(closed forms) and the reference-scene loader (counts from the reference's own assets)."""
import os

import numpy as np
import pytest

from colb_amd import scene as S
from colb_amd.params import north_star, reference_compat

HERE = os.path.dirname(os.path.abspath(__file__))


def two_cube_scene(gap=0.02, ground=False):
    """Cube B rests slightly rotated just above cube A (gap below dhat): PT + EE contacts."""
    V, F = S.cube_mesh()
    E = S.edges_of_tris(F)
    M = S.mass4_from_mesh(V, F, 1e3)
    b = S._Builder()
    b.add(V, F, E, M, S.q_from_pose([0, 0.5, 0], np.eye(3)))
    th = 0.3
    Rm = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    b.add(V, F, E, M, S.q_from_pose([0.1, 1.5 + gap, 0.05], Rm))
    return b.build("two_cubes")


def fd_params():
    return north_star(dhat=0.05, k_contact=1e12, kappa=1e5, ground_enabled=False, cg_tol=1e-12,
                      cg_max_iter=500)


def test_mass_matrix_closed_forms():
    V, F = S.cube_mesh((0.4, 0.6, 1.0))
    M = S.mass4_from_mesh(V, F, 1e3)
    m = 1e3 * 0.4 * 0.6 * 1.0
    assert M[0, 0] == pytest.approx(m, rel=1e-12)
    np.testing.assert_allclose(M[0, 1:], 0, atol=1e-12)
    np.testing.assert_allclose(np.diag(M[1:, 1:]), m / 12 * np.array([0.16, 0.36, 1.0]), rtol=1e-12)
    # shifted mesh: first moments = m * shift, parallel-axis second moments
    sh = np.array([0.3, -0.2, 0.5])
    M2 = S.mass4_from_mesh(V + sh, F, 1e3)
    np.testing.assert_allclose(M2[0, 1:], m * sh, rtol=1e-12)
    np.testing.assert_allclose(M2[1:, 1:], M[1:, 1:] + m * np.outer(sh, sh), rtol=1e-12, atol=1e-9)
    # fine torus approaches the smooth volume 2 pi^2 R r^2
    V, F = S.torus_mesh(1.0, 0.25, 192, 96)
    vol = S.mesh_volume_moments(V, F)[0]
    assert vol == pytest.approx(2 * np.pi ** 2 * 0.25 ** 2, rel=2e-3)


def test_benchmark_scene_sizes():
    c1 = S.cubes_on_ground()
    assert (c1.n_bodies, c1.n_verts, c1.n_tris, c1.n_edges) == (8, 64, 96, 144)
    net = S.chain_net(n=4)
    assert net.n_bodies == 16 and net.n_verts == 16 * 288 and net.n_tris == 16 * 576 and net.n_edges == 16 * 864
    assert net.body_pinned.sum() == 4


def test_reference_chain_scene_loader():
    sc = S.load_json_scene(os.path.join(HERE, "golden", "ref_assets", "chains.json"))
    # sizes of the reference chain net (144 links, 22-vertex polylines)
    assert sc.n_bodies == 144 and sc.n_verts == 3168 and sc.n_edges == 3168 and sc.n_tris == 0
    assert sc.body_pinned.sum() == 24
    # v = V R^T * 2 (utils/scene.py:78-85): polyline radius 0.325*2 around the link centre
    x = sc.world_vertices()
    assert np.abs(x[:22] - sc.q0[0, :3]).max() == pytest.approx(0.65, rel=1e-3)


def test_energy_matches_numpy_restatement(oracle_mod):
    sc = two_cube_scene()
    p = fd_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(0)
    q = sc.q0 + 1e-3 * rng.normal(size=sc.q0.shape)
    qdot = 0.1 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, qdot)
    E = orc.energy()[0]
    # numpy: inertia + ortho + barrier over the active set (from the oracle's detection)
    qt = sc.q0 + p.dt * qdot
    qt[:, :3] += p.dt ** 2 * np.array(p.gravity)
    e_in = 0.0
    for b in range(sc.n_bodies):
        dq = (q[b] - qt[b]).reshape(4, 3)
        e_in += 0.5 * np.sum(sc.mass4[b] * (dq @ dq.T))
        A = q[b, 3:].reshape(3, 3)
        e_in += p.dt ** 2 * p.kappa * np.sum((A @ A.T - np.eye(3)) ** 2)
    pairs, d2 = orc.detect()
    assert len(pairs) > 0
    shat = p.dhat ** 2
    e_c = np.sum(p.dt ** 2 * p.k_contact * -(d2 - shat) ** 3)
    assert E == pytest.approx(e_in + e_c, rel=1e-12)


def test_gradient_matches_energy_differences(oracle_mod):
    sc = two_cube_scene()
    p = fd_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(1)
    q = sc.q0 + 1e-3 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, 0.1 * rng.normal(size=sc.q0.shape))
    g, D, pij, off = orc.assemble()
    assert len(orc.detect()[0]) > 0
    qf = q.reshape(-1)
    h = 1e-7
    gfd = np.zeros_like(qf)
    for k in range(len(qf)):
        e = np.zeros_like(qf)
        e[k] = h
        gfd[k] = (orc.energy(qf + e)[0] - orc.energy(qf - e)[0]) / (2 * h)
    np.testing.assert_allclose(g.reshape(-1), gfd, rtol=2e-5, atol=1e-5 * np.abs(gfd).max())


def test_hessian_matches_numpy_restatement(oracle_mod):
    sc = two_cube_scene()
    p = fd_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(2)
    q = sc.q0 + 1e-3 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, np.zeros_like(q))
    gvec, H = orc.dense_hessian()
    n = 12 * sc.n_bodies
    Href = np.zeros((n, n))
    psd = lambda A: (lambda w, V: (V * np.maximum(w, 0)) @ V.T)(*np.linalg.eigh(A))
    for b in range(sc.n_bodies):
        _, _, Ho = oracle_mod.ortho(q[b], p.kappa)
        blk = np.kron(sc.mass4[b], np.eye(3))
        blk[3:, 3:] += p.dt ** 2 * psd(Ho)
        Href[12 * b:12 * b + 12, 12 * b:12 * b + 12] = blk
    X = sc.world_vertices(q)
    pairs, _ = orc.detect()
    kd = p.k_contact * p.dt ** 2
    shat = p.dhat ** 2
    for kind, i, j in pairs:
        vids = [i, *sc.tris[j]] if kind == 0 else [*sc.edges[i], *sc.edges[j]]
        d2, gx, Hx, _ = oracle_mod.pair_dist2(kind, X[vids].reshape(-1))
        B, B1, B2 = oracle_mod.barrier(d2, shat)
        Hc = psd(kd * (B1 * Hx + B2 * np.outer(gx, gx)))
        # affine Jacobian: x_k = [1, xr_k] (x) I3 applied to the owning body's 12 DOF
        J = np.zeros((12, n))
        for k, v in enumerate(vids):
            b = sc.vert_body[v]
            jv = np.concatenate([[1.0], sc.rest_x[v]])
            J[3 * k:3 * k + 3, 12 * b:12 * b + 12] = np.kron(jv[None, :], np.eye(3))
        Href += J.T @ Hc @ J
    np.testing.assert_allclose(H, Href, rtol=1e-9, atol=1e-9 * np.abs(Href).max())


def test_pcg_matches_dense_solve(oracle_mod):
    sc = two_cube_scene()
    p = fd_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(3)
    q = sc.q0 + 1e-3 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, np.zeros_like(q))
    g, H = orc.dense_hessian()
    du, iters, res = orc.solve()
    np.testing.assert_allclose(du.reshape(-1), np.linalg.solve(H, g), rtol=1e-8, atol=1e-10 * np.abs(du).max())
    assert res[0] == pytest.approx(np.linalg.norm(g), rel=1e-14)
    assert 0 < iters[0] <= 24 * 2


def test_free_fall_closed_form(oracle_mod):
    V, F = S.cube_mesh()
    b = S._Builder()
    v0 = np.array([1.0, 2.0, -0.5])
    qd = np.zeros(12)
    qd[:3] = v0
    b.add(V, F, S.edges_of_tris(F), S.mass4_from_mesh(V, F, 1e3), S.q_from_pose([0, 5, 0], np.eye(3)), qd)
    sc = b.build("free_fall")
    p = north_star(newton_tol=1e-9, cg_tol=1e-14)
    orc = oracle_mod.Oracle(sc, p)
    c = np.array([0.0, 5.0, 0.0])
    v = v0.copy()
    g = np.array(p.gravity)
    for k in range(5):
        st = orc.step()
        c = c + p.dt * v + p.dt ** 2 * g
        v = v + p.dt * g
        q, qn, qdot = orc.get_state()
        np.testing.assert_allclose(q[0, :3], c, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(qdot[0, :3], v, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(q[0, 3:], np.eye(3).reshape(-1), atol=1e-12)
        assert st[0]["converged"] == 1


def test_step_decreases_energy_and_keeps_pinned_fixed(oracle_mod):
    sc = S.cubes_on_ground()
    sc.body_pinned[0] = 1
    p = north_star(ground_enabled=True, k_contact=1e20)
    orc = oracle_mod.Oracle(sc, p)
    q0 = sc.q0.copy()
    for k in range(3):
        st = orc.step()
        res, alpha, cg = orc.iter_log()
        q, _, _ = orc.get_state()
        np.testing.assert_array_equal(q[0], q0[0])
        assert st[0]["newton_iters"] >= 1
    # the cubes fall: every dynamic body moved down
    assert np.all(q[1:, 1] < q0[1:, 1])
    # no vertex crosses the ground (CCD-filtered line search)
    assert orc.vertices()[:, 1].min() > 0


def test_multi_scene_equals_separate_scenes(oracle_mod):
    a = S.cubes_on_ground(seed=1)
    b = S.cubes_on_ground(seed=5)
    both = S.concat_scenes([a, b])
    p = north_star(ground_enabled=True, k_contact=1e20, newton_tol=1e-7)
    o_ab = oracle_mod.Oracle(both, p)
    o_a, o_b = oracle_mod.Oracle(a, p), oracle_mod.Oracle(b, p)
    for k in range(3):
        o_ab.step(), o_a.step(), o_b.step()
    q_ab = o_ab.get_state()[0]
    np.testing.assert_allclose(q_ab[:8], o_a.get_state()[0], rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(q_ab[8:], o_b.get_state()[0], rtol=1e-12, atol=1e-13)


def test_contact_order_invariance(oracle_mod):
    """Permuting the bodies (and so the contact order) changes results only by rounding."""
    sc = S.cubes_on_ground()
    p = north_star(ground_enabled=True, k_contact=1e20, newton_tol=1e-8, cg_tol=1e-10)
    perm = np.array([3, 1, 7, 0, 5, 2, 6, 4])
    sc2 = S.concat_scenes([sc._single(0)])
    sc2 = S._extract_bodies(sc, perm)
    o1, o2 = oracle_mod.Oracle(sc, p), oracle_mod.Oracle(sc2, p)
    for k in range(3):
        o1.step(), o2.step()
    np.testing.assert_allclose(o2.get_state()[0], o1.get_state()[0][perm], rtol=1e-7, atol=1e-9)
