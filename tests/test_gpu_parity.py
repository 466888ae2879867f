"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle on the same seeded inputs.

Tolerances: contact pair sets are compared exactly; d^2, energies, gradients and Hessian blocks
within 1e-9 relative to the largest magnitude (FP64 with a different summation order and FMA
contraction); PCG solutions (solved to cg_tol 1e-12) within 1e-7; trajectories within 1e-6
relative when both sides converge Newton to 1e-9.
"""
import sys

import numpy as np
import pytest

from colb_amd import scene as S
from colb_amd.params import north_star, reference_compat

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def O():
    from oracle import oracle as O
    O.build(force=False)
    return O


def solver(sc, p):
    from colb_amd.abd import NewtonAbd
    return NewtonAbd(p.dt, sc, params=p)


def two_cubes(gap=0.0005):
    V, F = S.cube_mesh()
    E = S.edges_of_tris(F)
    M = S.mass4_from_mesh(V, F, 1e3)
    b = S._Builder()
    b.add(V, F, E, M, S.q_from_pose([0, 0.5, 0], np.eye(3)))
    th = 0.3
    Rm = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    b.add(V, F, E, M, S.q_from_pose([0.1, 1.5 + gap, 0.05], Rm))
    return b.build("two_cubes")


def contact_params(**kw):
    base = dict(dhat=0.05, k_contact=1e12, kappa=1e5, ground_enabled=False, cg_tol=1e-12, cg_max_iter=500)
    base.update(kw)
    return north_star(**base)


def perturbed(sc, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    return sc.q0 + amp * rng.normal(size=sc.q0.shape), 0.1 * rng.normal(size=sc.q0.shape)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = {
    "two_cubes": lambda: (two_cubes(), contact_params()),
    "cubes_ground": lambda: (S.cubes_on_ground(lift=0.02), north_star(ground_enabled=True, dhat=0.05, k_contact=1e14,
                                                                        cg_tol=1e-12)),
    "chain3": lambda: (S.chain_net(n=3), north_star(dhat=2e-3, k_contact=1e19, cg_tol=1e-12)),
    "ref_chain": lambda: (S.load_json_scene(_ref_chain()), reference_compat(cg_tol=1e-12)),
}


def _ref_chain():
    import os
    return os.path.join(os.path.dirname(__file__), "golden", "ref_assets", "chains.json")


@pytest.mark.parametrize("case", list(CASES))
def test_detect_energy_assemble_match_oracle(O, case):
    say("case", case)
    sc, p = CASES[case]()
    q, qd = perturbed(sc, 1, amp=1e-4)
    orc = O.Oracle(sc, p)
    orc.set_state(q, sc.q0, qd)
    gpu = solver(sc, p)
    gpu.set_state(q, sc.q0, qd)
    # contact pair sets: exact
    po, d2o = orc.detect()
    pg, d2g = gpu.detect_collision()
    so = {tuple(r): d for r, d in zip(po.tolist(), d2o)}
    sg = {tuple(r): d for r, d in zip(pg.tolist(), d2g)}
    say("  active pairs", len(so), len(sg))
    assert set(so) == set(sg)
    for k in so:
        assert sg[k] == pytest.approx(so[k], rel=1e-10, abs=1e-18)
    # energy
    Eo, Eg = orc.energy(), gpu.compute_g()
    assert Eg == pytest.approx(Eo[0], rel=1e-11)
    # gradient and Hessian blocks
    go, Do, pijo, offo = orc.assemble()
    gg, Dg, pijg, offg, Eg2 = gpu.assemble()
    assert Eg2[0] == pytest.approx(Eo[0], rel=1e-11)
    assert rel(gg, go) < 1e-9
    assert rel(Dg, Do) < 1e-9
    assert sorted(map(tuple, pijo.tolist())) == sorted(map(tuple, pijg.tolist()))
    og = {tuple(k): B for k, B in zip(pijg.tolist(), offg)}
    for k, B in zip(pijo.tolist(), offo):
        assert rel(og[tuple(k)], B) < 1e-9 or np.abs(og[tuple(k)] - B).max() < 1e-9 * np.abs(Do).max()


@pytest.mark.parametrize("case", ["two_cubes", "cubes_ground", "chain3"])
def test_solve_and_ccd_match_oracle(O, case):
    say("case", case)
    sc, p = CASES[case]()
    q, qd = perturbed(sc, 2, amp=1e-4)
    orc = O.Oracle(sc, p)
    orc.set_state(q, sc.q0, qd)
    gpu = solver(sc, p)
    gpu.set_state(q, sc.q0, qd)
    duo, ito, reso = orc.solve()
    resg = gpu.compute_du()
    dug = gpu._du
    say("  residual", reso[0], resg, "cg", ito[0], gpu.last_cg_iters[0])
    assert resg == pytest.approx(reso[0], rel=1e-10)
    assert rel(dug, duo) < 1e-7
    ao = orc.ccd(duo)
    ag = gpu.ccd(duo)
    say("  ccd", ao, ag)
    assert ag[0] == pytest.approx(ao[0], rel=1e-9)


@pytest.mark.parametrize("case,nsteps", [("cubes_ground", 25), ("two_cubes", 10), ("chain3", 5)])
def test_trajectory_matches_oracle(O, case, nsteps):
    say("case", case)
    sc, p = CASES[case]()
    p = p.replace(newton_tol=1e-9, cg_tol=1e-12, cg_max_iter=2000)
    orc = O.Oracle(sc, p)
    gpu = solver(sc, p)
    for k in range(nsteps):
        so = orc.step()[0]
        sg = gpu.step()[0]
        qo = orc.get_state()[0]
        qg = gpu.get_state()[0]
        err = rel(qg, qo)
        say(f"  step {k} newton {so['newton_iters']}/{sg['newton_iters']} active {so['n_active']}/{sg['n_active']} "
            f"err {err:.2e}")
        assert err < 1e-6
        assert sg["n_active"] == so["n_active"]


def test_multi_scene_batch_equals_single_scenes():
    a = S.cubes_on_ground(seed=1, lift=0.02)
    b = S.cubes_on_ground(seed=5, lift=0.02)
    p = north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, newton_tol=1e-9, cg_tol=1e-12)
    ab = solver(S.concat_scenes([a, b]), p)
    ga, gb = solver(a, p), solver(b, p)
    for k in range(8):
        ab.step(), ga.step(), gb.step()
    q = ab.get_state()[0]
    assert rel(q[:8], ga.get_state()[0]) < 1e-12
    assert rel(q[8:], gb.get_state()[0]) < 1e-12


def test_pinned_bodies_stay_fixed_and_no_penetration():
    sc = S.chain_net(n=4)
    p = north_star(dhat=1e-3, k_contact=1e20)
    g = solver(sc, p)
    q0 = sc.q0.copy()
    for k in range(3):
        st = g.step()
        say("  step", k, st[0])
    q = g.get_state()[0]
    pin = sc.body_pinned.astype(bool)
    np.testing.assert_array_equal(q[pin], q0[pin])
    assert np.all(q[~pin, 1] < q0[~pin, 1])
    pairs, d2 = g.detect_collision()
    assert len(d2) == 0 or d2.min() > 0


def test_deterministic_repeat():
    sc = S.cubes_on_ground(lift=0.02)
    p = north_star(ground_enabled=True, dhat=0.05, k_contact=1e14)
    a, b = solver(sc, p), solver(sc, p)
    for k in range(10):
        a.step(), b.step()
    np.testing.assert_array_equal(a.get_state()[0], b.get_state()[0])


def test_checkpoint_resumes_bit_for_bit(tmp_path):
    """Save after 5 of 10 steps; a fresh solver on the same scene that loads the file ends the
    last 5 steps on the same bits as the uninterrupted run (non-default PCG layout, in contact)."""
    sc = S.cubes_on_ground(lift=0.02)
    p = north_star(ground_enabled=True, dhat=0.05, k_contact=1e14)
    a = solver(sc, p)
    a.set_pcg_layout(3)
    for k in range(5):
        a.step()
    f = tmp_path / "ck.npz"
    a.save_checkpoint(f)
    for k in range(5):
        a.step()
    say("  uninterrupted", a.frame, a.last_stats[0])
    b = solver(sc, north_star())
    b.load_checkpoint(f)
    assert b.frame == 5 and b.pcg_layout == (3, False) and b.params == p
    for k in range(5):
        b.step()
    say("  resumed", b.frame, b.last_stats[0])
    assert b.frame == a.frame
    for x, y in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(x, y)


def lowered(sc, gap=5e-4):
    """Shift every body down so the lowest vertex sits `gap` above the ground plane y = 0."""
    q = sc.q0.copy()
    qb = q[sc.vert_body]
    xr = sc.rest_x
    y = qb[:, 1] + xr[:, 0] * qb[:, 4] + xr[:, 1] * qb[:, 7] + xr[:, 2] * qb[:, 10]
    sc.q0[:, 1] -= y.min() - gap
    return sc


@pytest.mark.parametrize("case", ["hull_batch", "box_pile"])
def test_batched_and_box_scenes_match_oracle(O, case):
    """Batched hull piles of uneven body counts (one context, several scenes) and a box pile, on the
    ground: trajectories after 6 steps."""
    if case == "hull_batch":
        sc = lowered(S.batched_hull_piles(n_scenes=4, lo=4, hi=12))
        p = north_star(dt=4e-3, ground_enabled=True, k_contact=1e14, newton_tol=1e-9, cg_tol=1e-12, cg_max_iter=2000)
    else:
        sc = lowered(S.box_pile(n=27))
        p = north_star(dt=8e-3, ground_enabled=True, k_contact=1e14, newton_tol=1e-9, cg_tol=1e-12, cg_max_iter=2000)
    say(f"{case}: {sc.n_bodies} bodies in {sc.n_scenes} scenes")
    g, o = solver(sc, p), O.Oracle(sc, p)
    for k in range(6):
        sg, so = g.step(), o.step()
        err = rel(g.get_state()[0], o.get_state()[0])
        say(f"  step {k}: rel {err:.2e} newton {[s['newton_iters'] for s in sg]} active {[s['n_active'] for s in sg]}")
        assert err < 1e-6
        assert [s["n_active"] for s in sg] == [s["n_active"] for s in so]


def test_pcg_launch_forms_and_autotune_agree():
    """One cooperative launch per solve, one launch per iteration, the multi-tile cooperative launch
    (its per-stage arithmetic is the per-iteration launches': the same bits), and every autotuned
    tile size give the same solution (to the PCG tolerance)."""
    sc, p = CASES["cubes_ground"]()
    q, qd = perturbed(sc, 5, amp=1e-4)
    g = solver(sc, p)
    g.set_state(q, sc.q0, qd)
    g.compute_du()
    du_a = g._du.copy()
    g2 = solver(sc, p.replace(debug=8))
    g2.set_state(q, sc.q0, qd)
    g2.compute_du()
    assert rel(g2._du, du_a) < 1e-8
    g3 = solver(sc, p.replace(debug=32))
    g3.set_state(q, sc.q0, qd)
    g3.compute_du()
    np.testing.assert_array_equal(g3._du, g2._du)
    np.testing.assert_array_equal(g3.last_cg_iters, g2.last_cg_iters)
    bpt, multi, ms = g.autotune(repeats=2)
    say(f"autotune: {bpt} bodies per tile, per-iteration launch {multi}, ms {ms.tolist()}")
    assert 1 <= bpt <= 5 and multi in (0, 1) and ms[bpt - 1, multi] > 0
    g.set_state(q, sc.q0, qd)
    g.compute_du()
    assert rel(g._du, du_a) < 1e-8


def test_learned_warm_start_trains_and_keeps_the_solution():
    """The warm start is trained on the solver's own logged Newton iterates (cubes on the ground,
    eight seeds), predicts the steps of two held-out scenes better than a zero guess (reported
    beside the block-Jacobi step z it corrects), and leaves the solution and the trajectory
    unchanged (PCG stops relative to ||grad|| either way); prints the A/B counts."""
    from colb_amd.learned import WarmStart, FEAT_Z
    p = north_star(dt=8e-3, ground_enabled=True, dhat=1e-3, kappa=1e7, k_contact=1e20, cg_tol=1e-10)

    def logged(seeds, steps):
        F, D = [], []
        for seed in seeds:
            g = solver(S.cubes_on_ground(seed=seed), p)
            g.log_samples(True)
            g.step(steps)
            f, d = g.samples()
            F.append(f)
            D.append(d)
        return np.concatenate(F), np.concatenate(D)

    F, D = logged(range(11, 19), 60)
    ws = WarmStart(seed=0)
    losses = ws.fit(F, D, steps=2000, batch=512, lr=1e-3)
    assert losses[-100:].mean() < 0.5 * losses[:20].mean()
    f, d = logged((19, 20), 60)
    err = np.linalg.norm(ws.predict(f) - d) / np.linalg.norm(d)
    err_z = np.linalg.norm(f[:, FEAT_Z:FEAT_Z + 12] - d) / np.linalg.norm(d)
    say(f"held-out relative error of the predicted step: {err:.3f} (block-Jacobi step alone {err_z:.3f}, "
        f"zero guess 1.0); training rows {len(F)}")
    assert err < 0.5
    runs = {}
    for name, w in (("plain", None), ("warm", ws)):
        g = solver(S.cubes_on_ground(seed=1), p)
        g.set_warm_start(w)
        cg = nw = 0
        for k in range(40):
            st = g.step()[0]
            cg += st["cg_iters"]
            nw += st["newton_iters"]
        runs[name] = (cg, nw, g.get_state()[0].copy())
        g.set_warm_start(None)
    say(f"A/B over 40 steps: plain {runs['plain'][:2]} (cg, newton), warm start {runs['warm'][:2]}")
    assert rel(runs["warm"][2], runs["plain"][2]) < 1e-6
    assert abs(runs["warm"][1] - runs["plain"][1]) <= 0.05 * runs["plain"][1] + 1


def test_mixed_precision_tracks_fp64():
    """FP32 PCG (AbdParams.precision = 1) against the all-FP64 path on a box pile dropped on the
    ground.  One FP32 solve is inexact (the contact blocks have condition numbers far beyond FP32's
    reach: error up to ~1e-4 of the step, against ~1e-13 for FP64 here), but Newton's FP64 gradient
    corrects it: after 20 steps, every Newton solve converged, the positions and the energy (at
    the start of step 21) agree within 1e-6, the gradient within 1e-3 (contact stiffness)."""
    sc = lowered(S.box_pile(n=64, density=1e1))
    p = north_star(dt=8e-3, ground_enabled=True, k_contact=1e20)
    gx, g64, g32 = solver(sc, p.replace(cg_tol=1e-13, cg_max_iter=5000)), solver(sc, p), solver(sc, p.replace(precision=1))
    q, qd = perturbed(sc, 3, amp=1e-4)
    for g in (gx, g64, g32):
        g.set_state(q, sc.q0, qd)
        g.compute_du()
    e64, e32 = rel(g64._du, gx._du), rel(g32._du, gx._du)
    say(f"one solve at the default PCG tolerance: du error FP64 {e64:.2e}, mixed {e32:.2e} (cg {g64.last_cg_iters}, "
        f"{g32.last_cg_iters})")
    assert e32 < 1e-3
    g64, g32 = solver(sc, p), solver(sc, p.replace(precision=1))
    for k in range(20):
        s64, s32 = g64.step()[0], g32.step()[0]
        assert s64["converged"] and s32["converged"]
    ex = rel(g32.get_state()[0], g64.get_state()[0])
    for g in (g64, g32):
        g.set_state(*g.get_state())  # energy and gradient at the start of the next step (q~ of step 21)
    E64, E32 = g64.compute_g(), g32.compute_g()
    gr64, gr32 = g64.assemble()[0], g32.assemble()[0]
    eE, eg = abs(E32 - E64) / abs(E64), rel(gr32, gr64)
    say(f"after 20 steps: x rel {ex:.2e}, E rel {eE:.2e}, grad rel {eg:.2e}")
    # the gradient magnifies a position difference by the barrier stiffness of the contacts
    assert ex < 1e-6 and eE < 1e-6 and eg < 1e-3


def _dense_H(D, pij, off):
    nb = D.shape[0]
    H = np.zeros((12 * nb, 12 * nb))
    for b in range(nb):
        H[12 * b:12 * b + 12, 12 * b:12 * b + 12] = D[b]
    for (i, j), B in zip(pij, off):
        H[12 * i:12 * i + 12, 12 * j:12 * j + 12] = B
        H[12 * j:12 * j + 12, 12 * i:12 * i + 12] = B.T
    return H


def _block_pcg(H, b, blocks, tol):
    """Textbook PCG with a block-diagonal preconditioner (reference for the device solve)."""
    def prec(v):
        return np.concatenate([P @ v[12 * k:12 * k + 12] for k, P in enumerate(blocks)])
    x, r = np.zeros(len(b)), b.copy()
    z = prec(r)
    pv, rz, bn, it = z.copy(), r @ z, np.linalg.norm(b), 0
    while np.linalg.norm(r) > tol * bn and it < 5000:
        Ap = H @ pv
        a = rz / (pv @ Ap)
        x += a * pv
        r -= a * Ap
        it += 1
        z = prec(r)
        rz2 = r @ z
        pv = z + (rz2 / rz) * pv
        rz = rz2
    return x, it


@pytest.mark.parametrize("c_scale", [0.0, 0.3])
def test_learned_preconditioner_matches_block_pcg(c_scale):
    """A preconditioner model with constant output c gives P_b = L^-T T T^T L^-1 (D_b = L L^T) on every
    unpinned body: the device PCG agrees with a textbook block PCG using precond_matrix(D_b, c) in
    solution (1e-8) and iteration count (+-2); c = 0 is block Jacobi (same count as plain PCG)."""
    from colb_amd.learned import LearnedPreconditioner, precond_matrix, N_PRECOND
    sc, p = CASES["chain3"]()
    p = p.replace(cg_tol=1e-12, cg_max_iter=2000)
    q, qd = perturbed(sc, 1, amp=1e-4)
    g = solver(sc, p)
    g.set_state(q, sc.q0, qd)
    gg, D, pij, off, _ = g.assemble()
    g.compute_du()
    plain_it = int(g.last_cg_iters[0])
    c = c_scale * np.random.default_rng(5).normal(size=N_PRECOND)
    lp = LearnedPreconditioner()
    lp.mlp.set_params(np.zeros(lp.mlp.n_params, dtype=np.float32))  # output 0, de-standardised to c
    lp.out_mean = c.astype(np.float32)
    lp.out_std = np.ones(N_PRECOND, dtype=np.float32)
    g.set_preconditioner(lp)
    g.compute_du()
    pinned = sc.body_pinned.astype(bool)
    blocks = [np.eye(12) / p.pinned_diag if pinned[b] else precond_matrix(D[b], c) for b in range(D.shape[0])]
    x, it = _block_pcg(_dense_H(D, pij, off), gg.reshape(-1), blocks, 1e-12)
    say(f"c_scale {c_scale}: device {int(g.last_cg_iters[0])} iterations, textbook {it}, plain {plain_it}")
    assert rel(g._du.reshape(-1), x) < 1e-8
    assert abs(int(g.last_cg_iters[0]) - it) <= 2
    if c_scale == 0.0:
        assert int(g.last_cg_iters[0]) == plain_it


def test_learned_preconditioner_trains_and_keeps_the_solution():
    """The preconditioner is trained on (features, chol(L^T Z L)) pairs, Z the body's diagonal block of
    the exact H^-1 of the solver's own assembled systems (cubes on the ground, eight seeds); on a
    held-out seed it leaves the trajectory unchanged (1e-6) and the Newton iterations within 5 % (a
    step whose residual ends near the Newton tolerance may take one iteration more or less with
    either PCG's rounding), and prints the PCG iteration counts of both preconditioners."""
    from colb_amd.learned import LearnedPreconditioner, precond_samples
    p = north_star(dt=8e-3, ground_enabled=True, dhat=1e-3, kappa=1e7, k_contact=1e20, cg_tol=1e-10)
    F, Tt = [], []
    for seed in range(11, 19):
        f, t = precond_samples(solver(S.cubes_on_ground(seed=seed), p), 40)
        F.append(f)
        Tt.append(t)
    F, Tt = np.concatenate(F), np.concatenate(Tt)
    lp = LearnedPreconditioner(seed=0)
    losses = lp.fit(F, Tt, steps=2000, batch=512, lr=1e-3)
    assert losses[-100:].mean() < 0.5 * losses[:20].mean()
    runs = {}
    for name, pre in (("block Jacobi", None), ("learned", lp)):
        g = solver(S.cubes_on_ground(seed=19), p)
        g.set_preconditioner(pre)
        cg = nw = 0
        for k in range(40):
            st = g.step()[0]
            cg += st["cg_iters"]
            nw += st["newton_iters"]
        runs[name] = (cg, nw, g.get_state()[0].copy())
    say(f"A/B over 40 steps (held-out seed): block Jacobi {runs['block Jacobi'][:2]} (cg, newton), "
        f"learned {runs['learned'][:2]}; training rows {len(F)}")
    assert rel(runs["learned"][2], runs["block Jacobi"][2]) < 1e-6
    assert abs(runs["learned"][1] - runs["block Jacobi"][1]) <= 0.05 * runs["block Jacobi"][1]


# --- direct solver (AbdParams.solver 1 and 2) -------------------------------------------------------
def _batch_of_sizes():
    """Three scenes of 8, 9 and 1 bodies: padded orders 96, 128 and 32 in one set of launches (the
    chain net lifted clear of the ground plane)."""
    net = S.chain_net(n=3)
    net.q0[:, 1] += 0.02 - net.world_vertices()[:, 1].min()
    return S.concat_scenes([S.cubes_on_ground(n=2, lift=0.02), net, S.cubes_on_ground(n=1, lift=0.02)])


DIRECT_CASES = dict(CASES, batch=lambda: (_batch_of_sizes(), north_star(ground_enabled=True, dhat=0.05,
                                                                        k_contact=1e14, cg_tol=1e-12)))


@pytest.mark.parametrize("case", ["two_cubes", "cubes_ground", "chain3", "ref_chain", "batch"])
def test_direct_solver_matches_dense_solve(O, case):
    """solver = 1: the device's dense Cholesky per scene against numpy's LU solve of the same
    assembled system (both FP64: 1e-9 relative, residual 1e-12) and, for one scene, against the
    oracle's PCG solved to 1e-12 (1e-7, the PCG tolerance the other solve tests use)."""
    say("case", case)
    sc, p = DIRECT_CASES[case]()
    q, qd = perturbed(sc, 2, amp=1e-4)
    g = solver(sc, p.replace(solver=1))
    g.set_state(q, sc.q0, qd)
    grad, D, pij, off, E = g.assemble()
    H = _dense_H(D, pij, off)
    b = grad.reshape(-1)
    res = np.atleast_1d(g.compute_du())
    du = g._du.reshape(-1)
    x = np.linalg.solve(H, b)
    resid = np.linalg.norm(H @ du - b) / np.linalg.norm(b)
    say("  n", len(b), "vs numpy", rel(du, x), "residual", resid, "cg", g.last_cg_iters)
    assert rel(du, x) < 1e-9
    assert resid < 1e-12
    assert np.all(g.last_cg_iters == 0)
    b0 = np.searchsorted(sc.body_scene, np.arange(sc.n_scenes + 1))
    for s in range(sc.n_scenes):
        assert res[s] == pytest.approx(np.linalg.norm(b[12 * b0[s]:12 * b0[s + 1]]), rel=1e-12)
    if sc.n_scenes == 1:
        orc = O.Oracle(sc, p)
        orc.set_state(q, sc.q0, qd)
        duo, ito, reso = orc.solve()
        assert rel(du, np.reshape(duo, -1)) < 1e-7


@pytest.mark.parametrize("case,nsteps", [("cubes_ground", 25), ("chain3", 5)])
def test_direct_trajectory_matches_oracle(O, case, nsteps):
    say("case", case)
    sc, p = CASES[case]()
    p = p.replace(newton_tol=1e-9, cg_tol=1e-12, cg_max_iter=2000)
    orc = O.Oracle(sc, p)
    gpu = solver(sc, p.replace(solver=1))
    for k in range(nsteps):
        so = orc.step()[0]
        sg = gpu.step()[0]
        err = rel(gpu.get_state()[0], orc.get_state()[0])
        say(f"  step {k} newton {so['newton_iters']}/{sg['newton_iters']} err {err:.2e}")
        assert err < 1e-6
        assert sg["cg_iters"] == 0


def test_direct_fallback_replaces_capped_pcg():
    """solver = 2 with a PCG cap of 3: each solve that ends at the cap is redone by the direct solve,
    so the run follows solver = 1 (1e-6, the trajectory tolerance: the solves the PCG finishes within
    3 iterations are converged to 1e-12), while the same cap without the fallback leaves inexact
    Newton steps that drift at least 100 times further."""
    sc, p = CASES["cubes_ground"]()
    p = p.replace(newton_tol=1e-9)
    a = solver(sc, p.replace(solver=1))
    b = solver(sc, p.replace(solver=2, cg_max_iter=3))
    d = solver(sc, p.replace(solver=0, cg_max_iter=3))
    for k in range(5):
        sa, sb, sd = a.step()[0], b.step()[0], d.step()[0]
        eb = rel(b.get_state()[0], a.get_state()[0])
        ed = rel(d.get_state()[0], a.get_state()[0])
        say(f"  step {k} newton {sa['newton_iters']}/{sb['newton_iters']}/{sd['newton_iters']} "
            f"cg {sa['cg_iters']}/{sb['cg_iters']}/{sd['cg_iters']} err fallback {eb:.2e} capped {ed:.2e}")
        assert sa["cg_iters"] == 0 and 0 < sb["cg_iters"] <= 3 * sb["newton_iters"]
        assert eb < 1e-6
    assert ed > 100 * eb


def test_direct_solver_rejects_what_it_cannot_hold():
    from colb_amd._lib import AbdError
    sc, p = CASES["two_cubes"]()
    g = solver(sc, p.replace(solver=1, direct_max_dofs=12))
    with pytest.raises(AbdError, match="direct_max_dofs"):
        g.step()
    with pytest.raises(AbdError, match="solver must be"):
        solver(sc, p.replace(solver=3))
    with pytest.raises(AbdError, match="direct_max_dofs above"):
        solver(sc, p.replace(solver=1, direct_max_dofs=1 << 20))


# --- line search: the accepted step against a brute-force search on the oracle's energy -----------
@pytest.mark.parametrize("case,amp", [("two_cubes", 3e-3), ("cubes_ground", 3e-3), ("chain3", 1e-3),
                                      ("ref_chain", 3e-2)])
def test_line_search_alpha_matches_brute_force(O, case, amp):
    """line_search_iterative (gauss_newton.py:98-115) with the CCD bound as the first trial: from
    the device's own du and CCD bound alpha0, halve alpha until the oracle's energy (pinned by
    finite differences and closed forms) falls below E(q), or until 0.5 alpha < ls_floor alpha0.
    The device must accept exactly that alpha (both are alpha0 / 2^k) and apply q - alpha du.
    Decisions closer than 1e-9 relative between E(q - alpha du) and E(q) are ties the two
    summation orders may split: comparison stops at the first such iteration (a tie at the first
    skips the case)."""
    say("case", case)
    sc, p = CASES[case]()
    q0, qd = perturbed(sc, 3, amp=amp)
    orc = O.Oracle(sc, p)
    orc.set_state(q0, sc.q0, qd)
    g = solver(sc, p)
    g.set_state(q0, sc.q0, qd)
    halvings = 0
    for it in range(4):
        q = g.get_state()[0].copy()
        g.compute_du()
        du = g._du.copy()
        a0 = float(g.ccd(du)[0]) if p.ccd_enabled else 1.0
        E0 = orc.energy(q)[0]
        a, margins = a0, []
        while True:
            E1 = orc.energy(q - a * du)[0]
            margins.append(abs(E1 - E0) / max(abs(E0), 1e-300))
            if E1 < E0 or 0.5 * a < p.ls_floor * a0:
                break
            a *= 0.5
        if min(margins) < 1e-9:
            if it == 0:
                pytest.skip(f"energy tie at the first iteration (margin {min(margins):.1e})")
            say(f"  iteration {it}: energy tie (margin {min(margins):.1e}), stop comparing")
            break
        alpha = g.line_search()
        say(f"  iteration {it}: alpha0 {a0:.6g} device {alpha:.6g} brute force {a:.6g} trials {len(margins)}")
        assert alpha == a
        halvings += len(margins) - 1
        assert rel(g.get_state()[0], q - alpha * du) < 1e-14
    say("  halvings", halvings)

