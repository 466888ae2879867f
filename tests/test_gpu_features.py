"""GPU parity of the solver options and host-forced failure paths, through the C ABI, against the
CPU FP64 oracle on the same inputs.

Tolerances as in test_gpu_parity: d^2, energies, gradients and Hessian blocks within 1e-9 relative,
line-search steps equal (halvings of a CCD bound that agrees to 1e-9), trajectories within 1e-6
relative while both sides converge Newton.
"""
import os
import sys

import numpy as np
import pytest

from colb_amd import scene as S
from colb_amd.params import north_star

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_ground import DHAT, cube_on_ground, ground_params  # noqa: E402
from test_oracle_scene import fd_params, two_cube_scene  # noqa: E402
from test_oracle_math import eig_projection, hip_psd_project, psd_cases  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def O():
    from oracle import oracle as O
    O.build(force=False)
    return O


def solver(sc, p):
    from colb_amd.abd import NewtonAbd
    return NewtonAbd(p.dt, sc, params=p)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ------------------------------------------------------------------------------------------------
# the assembly's PSD projection (matrix-core sign iteration) against eigendecompositions
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("m", [3, 6, 9, 12])
def test_psd_projection_matches_eigh_on_the_matrix_cores(m):
    """FP64 within 1e-13 of ||A||_F, FP32 (the mixed assembly) within 2e-6; the result exactly
    symmetric.  A wrong operand or output layout of the matrix instruction fails here."""
    from colb_amd import _lib
    A = psd_cases(m, seed=100 + m)
    for prec, tol in ((0, 1e-13), (1, 2e-6)):
        for absolute in (False, True):
            out = hip_psd_project(_lib.load(), A, prec, absolute)
            err = max(np.abs(out[i] - eig_projection(A[i], absolute)).max() / np.linalg.norm(A[i])
                      for i in range(len(A)))
            say(f"  m {m} precision {prec} absolute {absolute} err {err:.2e}")
            assert err < tol
            np.testing.assert_array_equal(out, out.transpose(0, 2, 1))


# ------------------------------------------------------------------------------------------------
# ground contact, both sides of the projection clamp
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("gap,hess_proj", [(0.2 * DHAT, 0), (0.7 * DHAT, 0), (0.2 * DHAT, 1)])
def test_ground_assembly_matches_oracle(O, gap, hess_proj):
    say("gap", gap, "hess_proj", hess_proj)
    sc = cube_on_ground(gap, tilt_deg=(1.2, 0.0, 0.8), n=2)
    p = ground_params(hess_proj=hess_proj)
    rng = np.random.default_rng(7)
    q = sc.q0 + 1e-4 * rng.normal(size=sc.q0.shape)
    qd = 0.1 * rng.normal(size=sc.q0.shape)
    orc, gpu = O.Oracle(sc, p), solver(sc, p)
    orc.set_state(q, sc.q0, qd)
    gpu.set_state(q, sc.q0, qd)
    po, d2o = orc.detect()
    pg, d2g = gpu.detect_collision()
    assert sorted(map(tuple, po.tolist())) == sorted(map(tuple, pg.tolist()))
    assert np.sum(po[:, 0] == 2) > 0
    go, Do, pijo, offo = orc.assemble()
    gg, Dg, pijg, offg, Eg = gpu.assemble()
    assert Eg[0] == pytest.approx(orc.energy()[0], rel=1e-11)
    assert rel(gg, go) < 1e-9
    assert rel(Dg, Do) < 1e-9


def test_penetrating_state_is_reported():
    from colb_amd._lib import AbdError
    sc = cube_on_ground(0.5 * DHAT)
    p = ground_params()
    gpu = solver(sc, p)
    q = sc.q0.copy()
    q[:, 1] -= 0.5 * DHAT + 1e-3
    gpu.set_state(q, q, np.zeros_like(q))
    assert gpu.compute_g() == np.inf
    with pytest.raises(AbdError, match="ground plane"):
        gpu.assemble()
    gpu.set_state(sc.q0, sc.q0, np.zeros_like(q))
    gpu.assemble()  # the error word was read and cleared


@pytest.mark.parametrize("case,d0", [("gap", 0.5 * DHAT), ("touch", 0.0), ("below", -1e-16)])
def test_ground_ccd_matches_oracle(O, case, d0):
    sc = cube_on_ground(0.5 * DHAT, tilt_deg=(0.0, 0.0, 0.0))
    y = sc.world_vertices()[:, 1]
    p = ground_params(ground_offset=float(y.min() - d0))
    du = np.zeros_like(sc.q0)
    du[:, 1] = 1.0
    a_o = O.Oracle(sc, p).ccd(du)[0]
    a_g = solver(sc, p).ccd(du)[0]
    say(case, a_o, a_g)
    assert a_g == pytest.approx(a_o, rel=1e-12, abs=0.0)
    if d0 <= 0:
        assert a_g == 0.0


# ------------------------------------------------------------------------------------------------
# line search along a given direction, including the floor branch
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_line_search_along_matches_oracle(O, sign):
    """sign -1: an ascent direction, which the reference's search ends at the floor alpha0/64."""
    sc = two_cube_scene()
    p = fd_params()
    rng = np.random.default_rng(6)
    q = sc.q0 + 1e-3 * rng.normal(size=sc.q0.shape)
    qd = 0.1 * rng.normal(size=sc.q0.shape)
    orc, gpu = O.Oracle(sc, p), solver(sc, p)
    orc.set_state(q, sc.q0, qd)
    gpu.set_state(q, sc.q0, qd)
    du = sign * orc.solve()[0]
    a0 = orc.ccd(du)[0]
    ao = orc.line_search(du)[0]
    ag = gpu.line_search_along(du)[0]
    say("alpha0", a0, "oracle", ao, "gpu", ag)
    if sign < 0:
        assert ao == a0 / 64
    assert ag == pytest.approx(ao, rel=1e-9)
    assert rel(gpu.get_state()[0], orc.get_state()[0]) < 1e-12


def test_line_search_through_the_plane_takes_no_step(O):
    sc = cube_on_ground(0.5 * DHAT)
    p = ground_params(ccd_enabled=False)
    gpu = solver(sc, p)
    du = np.zeros_like(sc.q0)
    du[:, 1] = 10.0
    assert gpu.line_search_along(du)[0] == 0.0
    np.testing.assert_array_equal(gpu.get_state()[0], sc.q0)
    assert O.Oracle(sc, p).line_search(du)[0] == 0.0


# ------------------------------------------------------------------------------------------------
# predictor, adaptive stiffness, eigenvalue mirroring: trajectories against the oracle
# ------------------------------------------------------------------------------------------------
TRAJ = {
    "predictor1": lambda: (S.cubes_on_ground(lift=0.02),
                           north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, predictor=1), 20),
    "predictor2": lambda: (S.cubes_on_ground(lift=0.02),
                           north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, predictor=2), 20),
    "adaptive": lambda: (cube_on_ground(0.8 * DHAT, tilt_deg=(0.0, 0.0, 0.0), density=1e4),
                         ground_params(k_contact=1e8, k_max=1e14, k_raise_gap=0.45), 25),
    "mirror": lambda: (S.chain_net(n=3), north_star(dhat=2e-3, k_contact=1e19, hess_proj=1), 5),
    "pcg_x0": lambda: (S.cubes_on_ground(lift=0.02),
                       north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, pcg_x0=1), 20),
}


@pytest.mark.parametrize("case", list(TRAJ))
def test_option_trajectories_match_oracle(O, case):
    sc, p, nsteps = TRAJ[case]()
    p = p.replace(newton_tol=1e-9, cg_tol=1e-12, cg_max_iter=2000)
    orc, gpu = O.Oracle(sc, p), solver(sc, p)
    raises = 0
    for k in range(nsteps):
        so, sg = orc.step()[0], gpu.step()[0]
        err = rel(gpu.get_state()[0], orc.get_state()[0])
        raises += sg["k_raises"]
        say(f"  {case} step {k} newton {so['newton_iters']}/{sg['newton_iters']} active {so['n_active']}/"
            f"{sg['n_active']} pred {so['predictor_alpha']:.3g}/{sg['predictor_alpha']:.3g} "
            f"k {so['k_contact']:.3g}/{sg['k_contact']:.3g} err {err:.2e}")
        assert err < 1e-6
        assert sg["n_active"] == so["n_active"]
        assert sg["k_raises"] == so["k_raises"]
        assert sg["predictor_alpha"] == pytest.approx(so["predictor_alpha"], rel=1e-9, abs=0.0)
    np.testing.assert_array_equal(gpu.stiffness(), orc.stiffness())
    if case == "adaptive":
        assert raises > 0


def test_stall_trajectory_matches_oracle(O):
    """stall_alpha stops a scene after a CCD-pinned step: the GPU stops on the same iterations as
    the oracle (pinned steps have alpha ~1e-11, far from the 1e-6 threshold) and ends in the same
    state, on one scene and batched with a 3 x 3 net.  The first stall comes at step 15; the run
    stops five steps later: a pinned contact (d ~1e-16) turns rounding differences of 1e-10 into
    2e-7 at the stall, and some steps further on into a different CCD bound, a trait of the
    scene that the plain method shares."""
    p = north_star(dhat=1e-3, k_contact=1e20, stall_alpha=1e-6)
    for ns in (1, 2):
        sc = S.chain_net(n=4) if ns == 1 else S.concat_scenes([S.chain_net(n=4), S.chain_net(n=3)])
        orc, gpu = O.Oracle(sc, p), solver(sc, p)
        stalls = 0
        for k in range(20):
            so, sg = orc.step(), gpu.step()
            err = rel(gpu.get_state()[0], orc.get_state()[0])
            say(f"  scenes {ns} step {k} newton {[x['newton_iters'] for x in so]}/{[x['newton_iters'] for x in sg]} "
                f"stalled {[x['stalled'] for x in so]}/{[x['stalled'] for x in sg]} err {err:.2e}")
            for a, b in zip(so, sg):
                assert (b["newton_iters"], b["stalled"], b["converged"]) == (a["newton_iters"], a["stalled"], a["converged"])
                stalls += b["stalled"]
            assert err < 1e-6
        assert stalls > 0


def test_stiffness_round_trip_and_reset():
    sc = cube_on_ground(0.5 * DHAT, n=2)
    p = ground_params()
    gpu = solver(sc, p)
    np.testing.assert_array_equal(gpu.stiffness(), p.k_contact)
    k = np.array([2e12, 5e12])
    gpu.set_stiffness(k)
    np.testing.assert_array_equal(gpu.stiffness(), k)
    gpu.set_params(p)
    np.testing.assert_array_equal(gpu.stiffness(), p.k_contact)


# ------------------------------------------------------------------------------------------------
# host-forced failure paths
# ------------------------------------------------------------------------------------------------
def test_forced_pcg_fallback_gives_the_same_step():
    """Debug bit 4 raises the persistent PCG's failure flag before its launch: every workgroup leaves
    at its first barrier that does not fill at once, the host counts the fallback and reruns the
    solve with one launch per iteration.  The result equals a run that uses that form directly."""
    sc = S.chain_net(n=3)
    p = north_star(dhat=2e-3, k_contact=1e19, cg_tol=1e-12, newton_tol=1e-9)
    ref = solver(sc, p)
    ref.set_pcg_layout(5, True)
    forced = solver(sc, p.replace(debug=16))
    normal = solver(sc, p)
    for k in range(4):
        sr, sf, sn = ref.step()[0], forced.step()[0], normal.step()[0]
        say(k, sr["newton_iters"], sf["newton_iters"], sf["pcg_fallbacks"], sn["pcg_fallbacks"])
        assert sf["pcg_fallbacks"] == sf["newton_iters"]
        assert sn["pcg_fallbacks"] == 0
    np.testing.assert_array_equal(forced.get_state()[0], ref.get_state()[0])
    assert rel(normal.get_state()[0], ref.get_state()[0]) < 1e-6


def test_non_finite_state_is_rejected():
    from colb_amd._lib import AbdError
    sc = two_cube_scene()
    gpu = solver(sc, fd_params())
    q = sc.q0.copy()
    q[1, 4] = np.nan
    with pytest.raises(AbdError, match="non-finite"):
        gpu.set_state(q)
    qd = np.zeros_like(q)
    qd[0, 0] = np.inf
    with pytest.raises(AbdError, match="non-finite"):
        gpu.set_state(None, None, qd)
    gpu.step()  # the context is still usable


def test_newton_cap_is_reported():
    sc = two_cube_scene(gap=0.01)
    p = fd_params().replace(max_newton=1, newton_tol=1e-12)
    st = solver(sc, p).step()[0]
    assert st["newton_iters"] == 1 and st["converged"] == 0 and st["residual_last"] > 1e-12


def test_capacity_growth_gives_the_same_result():
    """The smallest initial capacities force every list to grow during the first detections."""
    sc = S.chain_net(n=3)
    p = north_star(dhat=0.2, k_contact=1e12, cg_tol=1e-12, newton_tol=1e-9)
    big, small = solver(sc, p), solver(sc, p.replace(max_candidates=1))
    for k in range(3):
        sb, ss = big.step()[0], small.step()[0]
        say(k, sb["n_candidates"], sb["n_active"])
    assert sb["n_candidates"] > 1024  # beyond the smallest candidate capacity
    np.testing.assert_array_equal(small.get_state()[0], big.get_state()[0])


def test_corrupted_learned_output_falls_back_to_plain_pcg():
    """A warm-start model whose weights are NaN predicts NaN: the residual guard rejects it and the
    solve equals plain PCG."""
    from colb_amd.learned import WarmStart
    sc = S.cubes_on_ground(lift=0.02)
    p = north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, cg_tol=1e-12, newton_tol=1e-9)
    plain, warm = solver(sc, p), solver(sc, p)
    ws = WarmStart(seed=0)
    flat = np.full(ws.mlp.n_params, np.nan, dtype=np.float32)
    ws.mlp.set_params(flat)
    ws.feat_mean = np.zeros(49, np.float32)
    ws.feat_std = np.ones(49, np.float32)
    ws.out_mean = np.zeros(12, np.float32)
    ws.out_std = np.ones(12, np.float32)
    warm.set_warm_start(ws)
    for k in range(5):
        sp, sw = plain.step()[0], warm.step()[0]
        assert sw["newton_iters"] == sp["newton_iters"]
    # same solve; only the norm of the right-hand side is summed in another order
    assert rel(warm.get_state()[0], plain.get_state()[0]) < 1e-10


# ------------------------------------------------------------------------------------------------
# readback for viewers and the Profiller
# ------------------------------------------------------------------------------------------------
def test_readback_matches_oracle(O, tmp_path):
    """compute_V, soup.x_transformed.numpy() (viewer.py:6, 46) and get_contact_points (abd.py:748-757)
    against the oracle's vertices and a numpy restatement of the closest points."""
    sc = two_cube_scene(gap=0.01)
    p = fd_params()
    orc, gpu = O.Oracle(sc, p), solver(sc, p)
    orc.step()
    gpu.step()
    Vg = gpu.compute_V()
    assert rel(Vg, orc.vertices()) < 1e-9
    xt = gpu.soup.x_transformed.numpy()
    assert isinstance(xt, np.ndarray) and type(xt) is np.ndarray
    np.testing.assert_array_equal(xt, Vg)
    pts, mag = gpu.get_contact_points()
    pairs, d2 = gpu.detect_collision()
    assert len(pts) == len(pairs) > 0
    np.testing.assert_allclose(mag, np.abs(np.sqrt(d2) - p.dhat), rtol=1e-9)
    X = Vg
    from scipy.optimize import minimize
    for (kind, i, j), m, dd in zip(pairs, pts, d2):
        # closest points by a bounded minimisation over the primitives' parameters, from a grid start
        if kind == 0:
            P, tri = X[i], X[sc.tris[j]]
            f = lambda w: np.sum((tri[0] + w[0] * (tri[1] - tri[0]) + w[1] * (tri[2] - tri[0]) - P) ** 2)
            cons = [{"type": "ineq", "fun": lambda w: 1.0 - w[0] - w[1]}]
            grid = [(u, v) for u in np.linspace(0, 1, 21) for v in np.linspace(0, 1, 21) if u + v <= 1]
            w0 = min(grid, key=f)
            w = minimize(f, w0, method="SLSQP", bounds=[(0, 1), (0, 1)], constraints=cons, options={"ftol": 1e-16}).x
            c0, c1 = P, tri[0] + w[0] * (tri[1] - tri[0]) + w[1] * (tri[2] - tri[0])
        else:
            a, b = X[sc.edges[i]], X[sc.edges[j]]
            f = lambda st: np.sum((a[0] + st[0] * (a[1] - a[0]) - b[0] - st[1] * (b[1] - b[0])) ** 2)
            grid = [(u, v) for u in np.linspace(0, 1, 21) for v in np.linspace(0, 1, 21)]
            st = minimize(f, min(grid, key=f), method="SLSQP", bounds=[(0, 1), (0, 1)], options={"ftol": 1e-16}).x
            c0, c1 = a[0] + st[0] * (a[1] - a[0]), b[0] + st[1] * (b[1] - b[0])
        assert np.linalg.norm(c0 - c1) == pytest.approx(np.sqrt(dd), rel=1e-6)
        assert np.linalg.norm(m - 0.5 * (c0 + c1)) < 1e-6 * max(1.0, np.abs(X).max())


def test_iter_log_and_profiler_match_oracle(O, tmp_path):
    sc = two_cube_scene(gap=0.01)
    # the first frame takes the oracle 15 Newton iterations (steps at the line-search floor while the
    # residual is near 1e-7): a cap of 32 keeps rounding-level differences from ending it unconverged
    p = fd_params().replace(newton_tol=1e-8, max_newton=32)
    orc, gpu = O.Oracle(sc, p), solver(sc, p)
    for f in range(6):
        so, sg = orc.step()[0], gpu.step()[0]
        ro, ao, co = orc.iter_log()
        rg, ag, cg = gpu.iter_log()
        n, no = sg["newton_iters"], so["newton_iters"]
        assert sg["converged"] and so["converged"]
        # steps from iterates within 1e3 tol of convergence change the energy by less than the 1e-11
        # to which the two energies agree: accepting or reaching the floor there is a rounding tie,
        # after which the two iteration paths (and counts) may differ.  Up to that point both logs
        # hold the same iterations: residuals within 1e-6 of the step's largest, equal steps.
        far = int(np.argmax(ro[0, :no] <= 1e3 * p.newton_tol))
        assert far > 0 and int(np.argmax(rg[0, :n] <= 1e3 * p.newton_tol)) == far
        np.testing.assert_allclose(rg[0, :far + 1], ro[0, :far + 1], rtol=0, atol=1e-6 * ro[0, :no].max())
        np.testing.assert_allclose(ag[0, :far], ao[0, :far], rtol=1e-6)
        assert rg[0, n - 1] < p.newton_tol and ro[0, no - 1] < p.newton_tol
        assert np.all(rg[0, n:] == 0) and np.all(ag[0, n:] == 0)
        np.testing.assert_array_equal(gpu.profiler.convergence[f, :n], rg[0, :n])
    path = tmp_path / "profiler.npz"
    gpu.profiler.save(str(path))
    with np.load(path, allow_pickle=False) as z:
        np.testing.assert_array_equal(z["convergence"], gpu.profiler.convergence)
        np.testing.assert_array_equal(z["alphas"], gpu.profiler.alphas)


# ------------------------------------------------------------------------------------------------
# candidate cache: the same candidates as a traversal at every detection, so the same bits
# ------------------------------------------------------------------------------------------------
CACHE_CASES = {
    "cubes_ground": lambda: (S.cubes_on_ground(lift=0.02), north_star(ground_enabled=True, dhat=0.05, k_contact=1e14),
                             15),
    "chain3": lambda: (S.chain_net(n=3), north_star(dhat=2e-3, k_contact=1e19), 8),
    "box27": lambda: (cube_on_ground(0.5 * DHAT, n=3), ground_params(), 15),
}


@pytest.mark.parametrize("case", list(CACHE_CASES))
@pytest.mark.parametrize("margin", [None, 1e-4])
def test_candidate_cache_gives_the_same_bits(case, margin):
    """margin None: 4 dhat (few traversals); 1e-4: small enough that iterates leave the cached region
    and the cache is rebuilt within steps."""
    sc, p, nsteps = CACHE_CASES[case]()
    full, cached = solver(sc, p.replace(cand_margin=0.0)), solver(sc, p.replace(cand_margin=margin))
    trav = det = 0
    for k in range(nsteps):
        sf, sc_ = full.step()[0], cached.step()[0]
        assert sf["traversals"] == sf["detections"]
        trav += sc_["traversals"]
        det += sc_["detections"]
        assert sc_["n_candidates"] == sf["n_candidates"] and sc_["n_active"] == sf["n_active"]
        np.testing.assert_array_equal(cached.get_state()[0], full.get_state()[0])
    say(f"  {case} margin {margin}: {trav} traversals for {det} detections")
    assert 0 < trav < det


# ------------------------------------------------------------------------------------------------
# learned step predictor: features and the predicted first iterate against the oracle
# ------------------------------------------------------------------------------------------------
def _untrained_predictor(seed=1):
    from colb_amd.learned import StepPredictor
    sp = StepPredictor(dims=(49, 64, 64, 12), seed=seed)
    rng = np.random.default_rng(seed)
    sp.feat_mean = rng.normal(size=49).astype(np.float32)
    sp.feat_std = np.exp(rng.normal(size=49)).astype(np.float32)
    sp.out_std = np.full(12, 1e-4, dtype=np.float32)
    return sp


LEARNED_CASES = {
    "two_cubes": lambda: (two_cube_scene(gap=0.01), fd_params()),
    "chain3": lambda: (S.chain_net(n=3), north_star(dhat=2e-3, k_contact=1e19)),
}


@pytest.mark.parametrize("case", list(LEARNED_CASES))
def test_learned_predictor_matches_oracle(O, case):
    """Each step: the GPU's logged features at the step's start equal the oracle's (FP32 rounding of
    FP64 values: 1e-5 of the largest magnitude of the feature's group: q - q~, qdot, contact count,
    contact gradient, block-Jacobi step; a component that is zero up to rounding differs with the
    order of the FP64 sums), and the step equals the oracle's given
    the prediction of the CPU model restatement on the oracle's features (predictor step equal).
    The scenes are ones whose Newton iterates do not magnify the model's rounding (a BF16 output
    differs between the two models by its last bits)."""
    from oracle import mlp_oracle
    sc, p = LEARNED_CASES[case]()
    p = p.replace(predictor=3, newton_tol=1e-9, cg_tol=1e-12, cg_max_iter=2000)
    sp = _untrained_predictor()
    gpu, orc = solver(sc, p), O.Oracle(sc, p)
    gpu.set_predictor(sp)
    gpu.log_steps(True)
    contacts = 0
    for k in range(8):
        fo = orc.features()
        contacts += int(fo[:, 24].sum())
        qd = orc.get_state()[2].reshape(-1, 12)
        y = mlp_oracle.forward(sp.mlp.get_params(), sp.dims, sp._inputs(fo))[1][-1]
        orc.set_prediction(p.dt * qd + (y * sp.out_std + sp.out_mean).astype(np.float64))
        so, sg = orc.step()[0], gpu.step()[0]
        free = sc.body_pinned == 0  # the log holds the unpinned bodies
        fo = fo[free]
        fg = gpu.step_samples()[0][-len(fo):]
        scale = np.zeros(49)
        for lo, hi in ((0, 12), (12, 24), (24, 25), (25, 37), (37, 49)):
            scale[lo:hi] = np.abs(fo[:, lo:hi]).max()
        cerr = (np.abs(fg - fo) / np.maximum(scale, 1e-30)).max(0)
        ferr = cerr.max()
        err = rel(gpu.get_state()[0], orc.get_state()[0])
        say(f"  {case} step {k}: features {ferr:.2e} (column {int(np.argmax(cerr))}) pred {so['predictor_alpha']:.3g}/{sg['predictor_alpha']:.3g} "
            f"newton {so['newton_iters']}/{sg['newton_iters']} err {err:.2e}")
        assert ferr < 1e-5
        assert sg["predictor_alpha"] == so["predictor_alpha"]
        assert err < 1e-6
    assert orc.take_error() == 0 and contacts > 0


def test_learned_predictor_without_a_model_is_an_error():
    from colb_amd._lib import AbdError
    sc, p, _ = CACHE_CASES["cubes_ground"]()
    gpu = solver(sc, p.replace(predictor=3))
    with pytest.raises(AbdError, match="abd_set_predictor"):
        gpu.step()


def test_learned_predictor_trains_and_keeps_the_solution():
    """Trained on the logged steps of runs from other initial velocities, the predictor leaves the
    trajectory unchanged (1e-6) over steps that all converge (from the state after the landing
    step, which reaches the Newton cap) and is reported as taken (predictor step > 0)."""
    from colb_amd.learned import StepPredictor
    sc = S.cubes_on_ground(lift=0.02)
    p = north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, newton_tol=1e-6, cg_tol=1e-12, cg_max_iter=2000)
    rng = np.random.default_rng(4)
    g = solver(sc, p)
    g.log_steps(True)
    for r in range(3):
        qd = sc.qdot0.copy()
        qd[:, :3] += rng.normal(scale=0.3, size=qd[:, :3].shape)
        g.set_state(sc.q0, sc.q0, qd)
        for _ in range(20):
            g.step()
    F, D = g.step_samples()
    assert len(F) == 3 * 20 * int((sc.body_pinned == 0).sum())
    sp = StepPredictor(seed=0)
    losses = sp.fit(F, D, p.dt, steps=1500, batch=256, lr=1e-3)
    assert losses[-100:].mean() < 0.5 * losses[:20].mean()
    g.log_steps(False)
    g.set_state(sc.q0, sc.q0, sc.qdot0)
    for _ in range(2):
        g.step()
    start = g.get_state()
    runs = {}
    for name, pred in (("plain", 0), ("learned", 3)):
        s = solver(sc, p.replace(predictor=pred))
        s.set_state(*start)
        if pred == 3:
            s.set_predictor(sp)
        it, taken, conv = 0, 0, True
        for _ in range(20):
            st = s.step()[0]
            it += st["newton_iters"]
            taken += st["predictor_alpha"] > 0
            conv &= bool(st["converged"])
        runs[name] = (it, taken, conv, s.get_state()[0])
    say(f"  newton iterations over 20 steps: plain {runs['plain'][0]}, learned {runs['learned'][0]} "
        f"(prediction taken in {runs['learned'][1]} steps)")
    assert runs["plain"][2] and runs["learned"][2]
    assert rel(runs["learned"][3], runs["plain"][3]) < 1e-6
    assert runs["learned"][1] > 0


# ------------------------------------------------------------------------------------------------
# the Newton iteration replayed from a captured HIP graph
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("case", ["chain3", "cubes_ground", "warm_chain3"])
def test_graph_replay_matches_eager_launches(case):
    """Every step of a run whose Newton iterations are replayed from a captured HIP graph equals,
    bit for bit, the same run with eager launches (debug bit 6).  Replays alternate with the eager
    work of each step (its static broad phase, the first sighting of a key, the redone parts of
    slow iterations, the forward update), and each step is checked."""
    if case == "chain3":
        sc, p = S.chain_net(n=3), north_star(dhat=2e-3, k_contact=1e19, cg_tol=1e-12)
    elif case == "cubes_ground":
        sc, p = S.cubes_on_ground(lift=0.02), north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, cg_tol=1e-12)
    else:
        sc, p = S.chain_net(n=3), north_star(dhat=2e-3, k_contact=1e19, cg_tol=1e-12, pcg_x0=1)
    say("case", case)
    g, e = solver(sc, p), solver(sc, p.replace(debug=64))
    g.enable_timing(2)  # the PCG solves time themselves on the device, inside the replays too
    launches = captures = iters = 0
    for k in range(12):
        sg, se = g.step()[0], e.step()[0]
        launches += sg["graph_launches"]
        captures += sg["graph_captures"]
        iters += sg["newton_iters"]
        assert se["graph_launches"] == 0 and se["graph_captures"] == 0
        assert np.array_equal(g.get_state()[0], e.get_state()[0]), f"step {k}"
        assert sg["newton_iters"] == se["newton_iters"] and sg["cg_iters"] == se["cg_iters"]
        assert sg["residual_last"] == se["residual_last"]
    t = g.kernel_times()["pcg_solve"]
    say(f"  {launches} graph replays, {captures} captures over 12 steps; {t['calls']} timed solves, "
        f"{t['ms'] / max(t['calls'], 1) * 1e3:.1f} us each")
    assert launches > 0 and captures >= 1
    assert t["calls"] >= iters and t["ms"] > 0 and t["bytes"] > 0
    g.close()
    e.close()


@pytest.mark.parametrize("case", ["cubes_8", "cube_batch"])
def test_line_search_trials_launched_together_give_the_same_bits(case):
    """Line-search trials launched several at a time (the k-th at 2^-k of the step, decided in
    order as one trial per launch decides them) give every step bit for bit as one trial per host
    round trip (debug bit 128), with fewer round trips where the searches halve several times.  The
    8-cube pile at d^ 1e-3, kappa 1e7 halves about four times per Newton iteration once it lands;
    the batch mixes scenes whose searches end after different trials of one launch."""
    if case == "cubes_8":
        sc = S.cubes_on_ground()
    else:
        sc = S.concat_scenes([S.cubes_on_ground(seed=1), S.cubes_on_ground(seed=2, lift=0.02),
                              S.cubes_on_ground(seed=3, n=1, lift=0.01)])
    p = north_star(dt=8e-3, ground_enabled=True, dhat=1e-3, kappa=1e7, k_contact=1e20)
    say("case", case)
    a, b = solver(sc, p), solver(sc, p.replace(debug=128))
    trips = {"together": 0, "single": 0}
    trials = {"together": 0, "single": 0}
    for k in range(40):
        sa, sb = a.step(), b.step()
        assert np.array_equal(a.get_state()[0], b.get_state()[0]), f"step {k}"
        for xa, xb in zip(sa, sb):
            for key in ("newton_iters", "cg_iters", "residual_last", "energy", "alpha_min", "ls_floor_hits"):
                assert xa[key] == xb[key], (k, key)
        trips["together"] += sa[0]["ls_round_trips"]
        trips["single"] += sb[0]["ls_round_trips"]
        trials["together"] += sa[0]["ls_trials"]
        trials["single"] += sb[0]["ls_trials"]
    say(f"  line-search round trips {trips}, trials {trials} over 40 steps ({len(sa)} scenes)")
    assert trips["together"] <= trips["single"]
    if case == "cubes_8":
        assert trips["together"] < trips["single"]
    a.close()
    b.close()


@pytest.mark.parametrize("shape", ["stretched", "compressed", "near_rotation", "flat"])
def test_orthogonality_projection_matches_oracle(O, shape):
    """The body block's orthogonality Hessian projected on the device from its analytic eigensystem
    (the negative scaling, twist and flip modes, from the singular values of the 3 x 3 block), or
    by the matrix-sign iteration for a near-singular block ("flat"), equals the oracle's projection
    through a Jacobi eigendecomposition.  kappa 1e9 makes the orthogonality term dominate the
    block."""
    sv = {"stretched": (1.3, 1.0, 0.8), "compressed": (0.5, 0.7, 0.9), "near_rotation": (1 - 1e-6, 1.0, 1 + 2e-6),
          "flat": (1.0, 0.9, 1e-7)}[shape]
    sc = S.cubes_on_ground(n=1, lift=0.5)
    p = north_star(ground_enabled=False, dhat=1e-3, kappa=1e9)
    rng = np.random.default_rng(5)
    A = S.random_rotation(rng) @ np.diag(sv) @ S.random_rotation(rng)
    q = sc.q0.copy()
    q[0, 3:] = A.ravel()
    qd = np.zeros_like(q)
    orc = O.Oracle(sc, p)
    orc.set_state(q, sc.q0, qd)
    gpu = solver(sc, p)
    gpu.set_state(q, sc.q0, qd)
    go, Do, _, _ = orc.assemble()
    gg, Dg, _, _, _ = gpu.assemble()
    say("case", shape, "block difference", rel(Dg, Do), "gradient", rel(gg, go))
    assert rel(Dg, Do) < 1e-12
    assert rel(gg, go) < 1e-12
    Dg = np.asarray(Dg).reshape(12, 12)
    assert np.array_equal(Dg, Dg.T)
    assert np.linalg.eigvalsh(Dg).min() > -1e-12 * np.abs(Dg).max()
