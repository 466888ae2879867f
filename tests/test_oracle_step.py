"""Pins the oracle's line search (gauss_newton.py:98-115), the Newton predictor and the adaptive
contact stiffness:

* line search: the accepted step is the first of alpha0 / 2^j (alpha0 the CCD bound) that lowers the
  energy, found here by brute force on the oracle's energy (itself pinned by test_oracle_scene); a
  direction along which no trial lowers the energy ends at the floor alpha0 / 64 (0.5 alpha <
  1e-2 alpha0), and at 0 when the floor trial's energy is infinite (through the ground plane);
* predictor: on a free-falling body the predicted first iterate changes nothing but the iteration
  count (the incremental potential has one minimiser near q~), and it starts closer to it;
* stall test (an option with no counterpart in the reference): a scene whose accepted step falls
  below stall_alpha stops iterating, reported as stalled, and the state it ends the step in is the
  state of the reference's loop cut at that iteration (max_newton = the iterations taken);
* adaptive stiffness: a heavy cube on the ground whose contacts cannot carry its weight at the
  initial stiffness (the cubic barrier's force 6 k d (dhat^2 - d^2)^2 peaks at 1.717 k dhat^5) ends
  with a stiffness whose capacity holds the weight, with every contact outside k_raise_gap * dhat,
  and never above k_max.
"""
import numpy as np
import pytest

from colb_amd import scene as S
from colb_amd.params import north_star

from test_oracle_ground import DHAT, cube_on_ground, ground_params
from test_oracle_scene import fd_params, two_cube_scene


def brute_force_alpha(orc, q, du, alpha0, floor=1e-2):
    """The halving search of gauss_newton.py:98-115 from alpha0, by direct energy evaluations."""
    E0 = orc.energy(q)[0]
    a = alpha0
    while True:
        E1 = orc.energy(q - a * du)[0]
        if E1 < E0:
            return a, False
        if a * 0.5 < floor * alpha0:
            return (a if np.isfinite(E1) else 0.0), True
        a *= 0.5


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_line_search_matches_brute_force(oracle_mod, sign):
    """sign 1: the Newton direction (a decrease at some halving); -1: an ascent direction (the floor)."""
    sc = two_cube_scene()
    p = fd_params()
    orc = oracle_mod.Oracle(sc, p)
    rng = np.random.default_rng(6)
    q = sc.q0 + 1e-3 * rng.normal(size=sc.q0.shape)
    orc.set_state(q, sc.q0, 0.1 * rng.normal(size=sc.q0.shape))
    du = sign * orc.solve()[0]
    alpha0 = orc.ccd(du)[0]
    want, at_floor = brute_force_alpha(orc, q, du, alpha0)
    assert at_floor == (sign < 0)
    if sign < 0:
        assert want == alpha0 / 64
    a = orc.line_search(du)[0]
    assert a == want
    np.testing.assert_array_equal(orc.get_state()[0], q - a * du)


def test_line_search_floor_through_the_plane_takes_no_step(oracle_mod):
    """Without CCD every trial of a step through the plane has infinite energy: no step is taken."""
    sc = cube_on_ground(0.5 * DHAT)
    p = ground_params(ccd_enabled=False)
    orc = oracle_mod.Oracle(sc, p)
    du = np.zeros_like(sc.q0)
    du[:, 1] = 10.0  # q - alpha du lowers the cube by 10 alpha: below the plane at every trial
    assert brute_force_alpha(orc, sc.q0, du, 1.0) == (0.0, True)
    assert orc.line_search(du)[0] == 0.0
    np.testing.assert_array_equal(orc.get_state()[0], sc.q0)


def free_cube():
    V, F = S.cube_mesh()
    E = S.edges_of_tris(F)
    b = S._Builder()
    b.add(V, F, E, S.mass4_from_mesh(V, F, 1e3), S.q_from_pose([0.0, 5.0, 0.0], S.rot_euler_xyz([10, 20, 30])))
    return b.build("free_cube")


@pytest.mark.parametrize("predictor", [1, 2])
def test_predictor_keeps_the_free_fall_solution(oracle_mod, predictor):
    sc = free_cube()
    base = north_star(newton_tol=1e-10, cg_tol=1e-12)
    runs = {}
    for pr in (0, predictor):
        orc = oracle_mod.Oracle(sc, base.replace(predictor=pr))
        orc.set_state(sc.q0, sc.q0, np.tile([0.3, 1.0, -0.2, 0.1, 0.0, 0.0, 0.0, 0.05, 0.0, 0.0, 0.0, -0.1], (1, 1)))
        its, pa, r0 = [], [], []
        for _ in range(5):
            st = orc.step()[0]
            its.append(st["newton_iters"])
            pa.append(st["predictor_alpha"])
            r0.append(st["residual_first"])
        runs[pr] = (orc.get_state()[0], its, pa, r0)
    q0, it0, pa0, r00 = runs[0]
    q1, it1, pa1, r01 = runs[predictor]
    np.testing.assert_allclose(q1, q0, rtol=0, atol=1e-12 * np.abs(q0).max())
    assert all(a == 0.0 for a in pa0)
    assert all(a == 1.0 for a in pa1)
    assert sum(it1) <= sum(it0)
    assert all(b < a for a, b in zip(r00, r01))


def test_adaptive_stiffness_holds_a_heavy_cube(oracle_mod):
    g, dens = 9.8, 1e4  # 10 t cube
    sc = cube_on_ground(0.8 * DHAT, tilt_deg=(0.0, 0.0, 0.0), density=dens)
    W = dens * g  # weight of the unit cube, on 4 bottom corners
    cap = lambda k: 6 * k * (DHAT / np.sqrt(5)) * (DHAT ** 2 * 0.8) ** 2  # peak barrier force
    k0, kmax, gap = 1e8, 1e14, 0.45
    assert cap(k0) < W / 4 < cap(kmax)
    p = ground_params(k_contact=k0, k_max=kmax, k_raise_gap=gap, newton_tol=1e-8)
    orc = oracle_mod.Oracle(sc, p)
    raises = 0
    for _ in range(40):
        st = orc.step()[0]
        raises += st["k_raises"]
    k = orc.stiffness()
    assert raises > 0 and np.all(k <= kmax) and np.all(k >= k0)
    assert st["k_contact"] == k.max()
    assert cap(k[0]) > W / 4
    d = sc.world_vertices(orc.get_state()[0])[:, 1]
    assert d.min() > gap * DHAT
    # fixed stiffness (k_max <= k_contact): the contacts collapse instead
    orc2 = oracle_mod.Oracle(sc, p.replace(k_max=0.0))
    for _ in range(40):
        st2 = orc2.step()[0]
    assert st2["k_raises"] == 0 and np.all(orc2.stiffness() == k0)
    assert sc.world_vertices(orc2.get_state()[0])[:, 1].min() < gap * DHAT


def test_set_params_resets_the_stiffness(oracle_mod):
    sc = cube_on_ground(0.5 * DHAT)
    p = ground_params()
    orc = oracle_mod.Oracle(sc, p)
    orc.set_stiffness(np.full(sc.n_bodies, 3.0 * p.k_contact))
    np.testing.assert_array_equal(orc.stiffness(), 3.0 * p.k_contact)
    orc.set_params(p)
    np.testing.assert_array_equal(orc.stiffness(), p.k_contact)


def test_stall_equals_the_loop_cut_at_that_iteration(oracle_mod):
    """A 4 x 4 chain net at d^ 1e-3, k 1e20 has CCD-pinned steps (alpha ~1e-11): with stall_alpha
    1e-6 a step that takes one ends there.  Rerunning that step from the same state with the
    reference's loop and max_newton = its iteration count gives the same state bit for bit."""
    sc = S.chain_net(n=4)
    p = north_star(dhat=1e-3, k_contact=1e20)
    orc = oracle_mod.Oracle(sc, p.replace(stall_alpha=1e-6))
    stalls = 0
    for _ in range(18):  # the first stall comes at step 15
        q, qn, qd = orc.get_state()
        st = orc.step()[0]
        if not st["stalled"]:
            assert st["converged"] or st["newton_iters"] == p.max_newton
            continue
        stalls += 1
        assert not st["converged"] and st["newton_iters"] < p.max_newton
        assert st["alpha_min"] < 1e-6 and st["residual_last"] >= p.newton_tol
        cut = oracle_mod.Oracle(sc, p.replace(max_newton=st["newton_iters"]))
        cut.set_state(q, qn, qd)
        sc_ = cut.step()[0]
        assert sc_["newton_iters"] == st["newton_iters"] and not sc_["stalled"]
        for a, b in zip(cut.get_state(), orc.get_state()):
            np.testing.assert_array_equal(a, b)
    assert stalls > 0
