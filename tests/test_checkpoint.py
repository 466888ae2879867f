"""Checkpoint file format (colb_amd.checkpoint) on the CPU: round trip, scene check, shapes.

The solver here is a stand-in holding the attributes and methods that checkpoint.save/load use;
the bit-for-bit resume of the real solver is a GPU test (test_gpu_parity.py).
"""
import numpy as np
import pytest

from colb_amd import checkpoint
from colb_amd import scene as S
from colb_amd.params import north_star


class _Stub:
    def __init__(self, sc, params):
        self.scene, self.params, self.n_bodies = sc, params, sc.n_bodies
        self.frame, self.pcg_layout = 0, (5, False)
        self.state = (sc.q0.copy(), sc.q0.copy(), sc.qdot0.copy())
        self.k = np.full(sc.n_bodies, params.k_contact)
        self.calls = []

    def get_state(self):
        return self.state

    def set_state(self, q, qn, qd):
        self.state = (np.array(q), np.array(qn), np.array(qd))
        self.calls.append("set_state")

    def set_params(self, p):
        self.params = p
        self.calls.append("set_params")

    def set_pcg_layout(self, bpt, multi):
        self.pcg_layout = (bpt, bool(multi))
        self.calls.append("set_pcg_layout")

    def stiffness(self):
        return self.k

    def set_stiffness(self, k):
        self.k = np.array(k)
        self.calls.append("set_stiffness")

    def compute_V(self):
        self.calls.append("compute_V")


def test_round_trip(tmp_path):
    sc = S.cubes_on_ground(n=2)
    p = north_star(ground_enabled=True, dhat=0.05, k_contact=1e14, gravity=(0.0, -9.8, 0.5))
    a = _Stub(sc, p)
    rng = np.random.default_rng(0)
    a.state = tuple(rng.normal(size=(sc.n_bodies, 12)) for _ in range(3))
    a.frame, a.pcg_layout = 17, (3, True)
    a.k = rng.uniform(1e14, 1e16, size=sc.n_bodies)
    f = tmp_path / "ck.npz"
    checkpoint.save(a, f)
    b = _Stub(sc, north_star())
    d = checkpoint.load(b, f)
    for x, y in zip(a.state, b.state):
        np.testing.assert_array_equal(x, y)
    assert b.frame == 17 and b.pcg_layout == (3, True)
    assert b.params == p and isinstance(b.params.gravity, tuple)
    np.testing.assert_array_equal(a.k, b.k)
    assert b.calls == ["set_params", "set_pcg_layout", "set_state", "set_stiffness", "compute_V"]
    assert d["scene_sha256"] == checkpoint.scene_fingerprint(sc)


def test_rejects_other_scene_and_other_files(tmp_path):
    a = _Stub(S.cubes_on_ground(n=2), north_star())
    f = tmp_path / "ck.npz"
    checkpoint.save(a, f)
    other = S.cubes_on_ground(n=2)
    other.rest_x *= 1.01  # same sizes, other rest geometry
    with pytest.raises(ValueError, match="another scene"):
        checkpoint.load(_Stub(other, north_star()), f)
    g = tmp_path / "other.npz"
    np.savez(g, q=np.zeros((8, 12)))
    with pytest.raises(ValueError, match="not a"):
        checkpoint.read(g)


def test_fingerprint_sees_every_field():
    sc = S.cubes_on_ground(n=2)
    base = checkpoint.scene_fingerprint(sc)
    assert checkpoint.scene_fingerprint(S.cubes_on_ground(n=2)) == base
    for name in ("rest_x", "mass4"):
        s2 = S.cubes_on_ground(n=2)
        getattr(s2, name).reshape(-1)[5] += 1e-12
        assert checkpoint.scene_fingerprint(s2) != base, name
    s2 = S.cubes_on_ground(n=2)
    s2.body_pinned[0] = 1
    assert checkpoint.scene_fingerprint(s2) != base
    s2 = S.cubes_on_ground(n=2)
    s2.tris[0] = s2.tris[0, ::-1]
    assert checkpoint.scene_fingerprint(s2) != base
