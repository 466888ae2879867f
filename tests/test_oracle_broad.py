"""Pins the oracle's candidate rule (broad phase over a linear path q0 -> q1) against brute force.

The rule keeps a pair when its swept boxes overlap and its distance at q0 minus the additive-CCD
motion bound is below dhat.  Brute force here evaluates every cross-body PT / EE pair's distance
(orc_pair_dist2, itself pinned against brute-force minimisation) at many points of the path:
  * every pair that comes within dhat at a sampled point is a candidate (nothing active is lost);
  * with q1 = q0 the candidates are exactly the pairs closer than dhat (the active set);
  * the motion bound is valid: no sampled distance of a pair lies below d(q0) - bound;
  * the rule prunes: far fewer candidates at dhat than at a wide inflation.
"""
import numpy as np
import pytest

from colb_amd import scene as S
from colb_amd.params import north_star

PT, EE = 0, 1


def two_cubes(gap):
    V, F = S.cube_mesh()
    E = S.edges_of_tris(F)
    M = S.mass4_from_mesh(V, F, 1e3)
    b = S._Builder()
    b.add(V, F, E, M, S.q_from_pose([0, 0.5, 0], np.eye(3)))
    th = 0.3
    Rm = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    b.add(V, F, E, M, S.q_from_pose([0.1, 1.5 + gap, 0.05], Rm))
    return b.build("two_cubes")


def positions(sc, q):
    q = np.asarray(q).reshape(-1, 12)
    qb = q[sc.vert_body]
    xr = sc.rest_x
    return qb[:, :3] + xr[:, 0:1] * qb[:, 3:6] + xr[:, 1:2] * qb[:, 6:9] + xr[:, 2:3] * qb[:, 9:12]


def cross_pairs(sc):
    vb = sc.vert_body
    out = []
    for v in range(sc.n_verts):
        for t in range(sc.n_tris):
            if vb[sc.tris[t, 0]] != vb[v]:
                out.append((PT, v, t))
    for e in range(sc.n_edges):
        for f in range(e + 1, sc.n_edges):
            if vb[sc.edges[e, 0]] != vb[sc.edges[f, 0]]:
                out.append((EE, e, f))
    return out


def pair_x(sc, X, kind, i, j):
    if kind == PT:
        return np.concatenate([X[i], X[sc.tris[j, 0]], X[sc.tris[j, 1]], X[sc.tris[j, 2]]])
    return np.concatenate([X[sc.edges[i, 0]], X[sc.edges[i, 1]], X[sc.edges[j, 0]], X[sc.edges[j, 1]]])


def path_distances(O, sc, q0, q1, pairs, ns=61):
    al = np.linspace(0.0, 1.0, ns)
    D = np.zeros((len(pairs), ns))
    for a_i, a in enumerate(al):
        X = positions(sc, (1 - a) * q0 + a * q1)
        for k, (kind, i, j) in enumerate(pairs):
            D[k, a_i] = np.sqrt(O.pair_dist2(kind, pair_x(sc, X, kind, i, j), derivs=False)[0])
    return D


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_swept_candidates_cover_every_pair_reaching_dhat(oracle_mod, seed):
    O = oracle_mod
    sc = two_cubes(gap=0.03)
    p = north_star(dhat=0.05)
    orc = O.Oracle(sc, p)
    rng = np.random.default_rng(seed)
    q0 = sc.q0.copy()
    q1 = q0.copy()
    # the upper cube moves down through the gap with a random twist, the lower one jitters
    q1[1, :3] += np.array([0.0, -0.06, 0.0]) + rng.normal(0, 0.01, 3)
    q1[1, 3:] += rng.normal(0, 0.02, 9)
    q1[0, :] += rng.normal(0, 0.003, 12)
    cand = {tuple(c) for c in orc.candidates(q0, q1).tolist()}
    pairs = cross_pairs(sc)
    D = path_distances(O, sc, q0, q1, pairs)
    reach = {pairs[k] for k in np.nonzero(D.min(axis=1) < p.dhat)[0]}
    assert len(reach) > 0
    assert reach <= cand
    # motion bound: no sampled distance below d0 - bound (bound recomputed from the path ends)
    X0, X1 = positions(sc, q0), positions(sc, q1)
    for k, (kind, i, j) in enumerate(pairs):
        a, b = pair_x(sc, X0, kind, i, j).reshape(4, 3), pair_x(sc, X1, kind, i, j).reshape(4, 3)
        d = (b - a) - (b - a).mean(axis=0)
        n = np.linalg.norm(d, axis=1)
        lp = n[0] + n[1:].max() if kind == PT else n[:2].max() + n[2:].max()
        assert D[k].min() >= D[k, 0] - lp - 1e-12
    # candidates are never pairs whose lower bound stays above dhat
    for (kind, i, j) in cand:
        k = pairs.index((kind, i, j))
        a, b = pair_x(sc, X0, kind, i, j).reshape(4, 3), pair_x(sc, X1, kind, i, j).reshape(4, 3)
        d = (b - a) - (b - a).mean(axis=0)
        n = np.linalg.norm(d, axis=1)
        lp = n[0] + n[1:].max() if kind == PT else n[:2].max() + n[2:].max()
        assert D[k, 0] - lp < p.dhat * (1 + 1e-6)


def test_static_candidates_are_the_active_set(oracle_mod):
    O = oracle_mod
    sc = two_cubes(gap=0.02)
    p = north_star(dhat=0.05)
    orc = O.Oracle(sc, p)
    q0 = sc.q0
    cand = {tuple(c) for c in orc.candidates(q0).tolist()}
    pairs = cross_pairs(sc)
    D = path_distances(O, sc, q0, q0, pairs, ns=1)[:, 0]
    active = {pairs[k] for k in np.nonzero(D < p.dhat)[0]}
    assert len(active) > 0
    assert cand == active


def test_rule_prunes_swept_box_overlaps(oracle_mod):
    """Interlocked chain links moving: with a wide inflation most nearby primitive pairs are
    candidates, with dhat only the few that can come that close."""
    O = oracle_mod
    sc = S.chain_net(n=3, pin_corners=False)
    p = north_star(dhat=1e-3)
    orc = O.Oracle(sc, p)
    rng = np.random.default_rng(7)
    q0 = sc.q0.copy()
    q1 = q0.copy()
    q1[:, :3] += rng.normal(0, 0.02, (sc.n_bodies, 3))
    q1[:, 3:] += rng.normal(0, 0.005, (sc.n_bodies, 9))
    big = orc.candidates(q0, q1, infl=1.0)   # swept boxes inflated far beyond any contact
    small = orc.candidates(q0, q1)
    assert len(small) * 20 < len(big)
