"""The per-body MLP of the learned warm start.

CPU: the oracle's backward (oracle/mlp_oracle.py) against central finite differences of its
own FP64 forward; BF16 rounding against its definition.
GPU (through include/abd_mlp.h): the matrix-core product against numpy on BF16-rounded inputs
(pins the operand and result lane layout of the matrix instruction), forward and gradient against
the oracle in BF16 mode, and Adam training that lowers the loss of a fitted function.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from colb_amd.learned import init_params

DIMS = (37, 48, 32, 12)


def test_bf16_rounding_matches_definition():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -3.1415926, 65504.0, 1e-30], dtype=np.float32)
    r = MO.to_bf16(x)
    # BF16 keeps 8 significant bits: the rounded values are representable and within half an ulp
    for v, w in zip(x, r):
        assert np.float32(w).view(np.uint32) & 0xFFFF == 0
        ulp = 2.0 ** (np.floor(np.log2(abs(v))) - 7)
        assert abs(w - v) <= 0.5 * ulp
    assert r[1] == 1.0  # tie 1 + 2^-8 rounds to even (1.0)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    flat = init_params(DIMS, seed=1).astype(np.float64)
    X = rng.normal(size=(9, DIMS[0]))
    T = rng.normal(size=(9, DIMS[-1]))
    w = rng.uniform(0.5, 1.5, DIMS[-1])
    loss, g = MO.loss_and_grad(flat, DIMS, X, T, w, bf16=False)
    idx = rng.choice(flat.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(flat)
        e[i] = h
        lp, _ = MO.loss_and_grad(flat + e, DIMS, X, T, w, bf16=False)
        lm, _ = MO.loss_and_grad(flat - e, DIMS, X, T, w, bf16=False)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-8)


@pytest.mark.gpu
def test_matrix_core_product_layout():
    from colb_amd.learned import Mlp
    m = Mlp(DIMS)
    rng = np.random.default_rng(2)
    for (M, N, K) in [(16, 16, 16), (48, 32, 64), (128, 96, 48)]:
        A = rng.normal(size=(M, K)).astype(np.float32)
        B = rng.normal(size=(K, N)).astype(np.float32)
        Cg = m.gemm_test(A, B)
        Cr = MO.to_bf16(A).astype(np.float64) @ MO.to_bf16(B).astype(np.float64)
        assert np.abs(Cg - Cr).max() <= 1e-4 * np.abs(Cr).max()


@pytest.mark.gpu
def test_forward_and_gradient_match_oracle():
    from colb_amd.learned import Mlp
    rng = np.random.default_rng(3)
    flat = init_params(DIMS, seed=4)
    m = Mlp(DIMS, params=flat)
    w = rng.uniform(0.5, 1.5, DIMS[-1]).astype(np.float32)
    m.set_loss_weights(w)
    X = rng.normal(size=(70, DIMS[0])).astype(np.float32)
    T = rng.normal(size=(70, DIMS[-1])).astype(np.float32)
    Yg = m.forward(X)
    Yo = MO.forward(flat, DIMS, X, bf16=True)[1][-1]
    assert np.abs(Yg - Yo).max() <= 2e-3 * np.abs(Yo).max()
    lg, gg = m.gradient(X, T)
    lo, go = MO.loss_and_grad(flat, DIMS, X, T, w, bf16=True)
    assert lg == pytest.approx(lo, rel=2e-3)
    assert np.abs(gg - go).max() <= 2e-2 * np.abs(go).max()
    assert np.dot(gg, go) / (np.linalg.norm(gg) * np.linalg.norm(go)) > 0.999


@pytest.mark.gpu
def test_training_lowers_the_loss():
    from colb_amd.learned import Mlp
    rng = np.random.default_rng(5)
    X = rng.normal(size=(2000, DIMS[0])).astype(np.float32)
    Wt = rng.normal(size=(DIMS[0], DIMS[-1])).astype(np.float32) / np.sqrt(DIMS[0])
    T = np.tanh(X @ Wt).astype(np.float32)
    m = Mlp(DIMS, seed=6)
    l0, _ = m.gradient(X, T)
    losses = m.train(X, T, steps=600, batch=128, lr=2e-3, seed=7)
    l1, _ = m.gradient(X, T)
    assert np.all(np.isfinite(losses))
    assert l1 < 0.2 * l0


def test_preconditioner_factor_round_trip():
    """The learned preconditioner's parametrisation: precond_matrix(D, precond_target(D, Z)) == Z for
    an SPD block D and an SPD target Z (the exact block of H^-1 of a coupled system), and all-zero
    outputs give D^-1 (block Jacobi)."""
    from colb_amd.learned import precond_matrix, precond_target
    rng = np.random.default_rng(3)
    A = rng.normal(size=(24, 24))
    H = A @ A.T + 24 * np.eye(24)
    D = H[:12, :12]
    Z = np.linalg.inv(H)[:12, :12]
    Z = 0.5 * (Z + Z.T)
    out = precond_target(D, Z)
    assert out.shape == (78,)
    assert np.allclose(precond_matrix(D, out), Z, rtol=1e-10, atol=1e-12 * np.abs(Z).max())
    assert np.allclose(precond_matrix(D, np.zeros(78)), np.linalg.inv(D), rtol=1e-10, atol=1e-14)
