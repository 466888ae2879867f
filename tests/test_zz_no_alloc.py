"""abd_mlp_forward_static (include/abd_mlp.h), the allocation-free forward: the same rows as
abd_mlp_forward, bit for bit, on calls in a row (the second smaller than the first, then a model of
other widths sharing the static scratch), and a batch beyond the scratch is rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_static_forward_matches_forward():
    from colb_amd.learned import Mlp
    m = Mlp((49, 128, 128, 128, 12), seed=3)
    rng = np.random.default_rng(0)
    for n in (3000, 517):
        X = rng.normal(size=(n, 49)).astype(np.float32)
        Ys, Y = m.forward_static(X), m.forward(X)
        print("rows", n, "max |Y|", float(np.abs(Y).max()), flush=True)
        np.testing.assert_array_equal(Ys, Y)
    m2 = Mlp((49, 64, 78), seed=4)
    X = rng.normal(size=(100, 49)).astype(np.float32)
    np.testing.assert_array_equal(m2.forward_static(X), m2.forward(X))
    np.testing.assert_array_equal(m.forward_static(X), m.forward(X))


def test_static_forward_rejects_too_many_rows():
    from colb_amd._lib import AbdError
    from colb_amd.learned import Mlp
    m = Mlp((49, 128, 128, 128, 12), seed=3)
    with pytest.raises(AbdError, match="at most"):
        m.forward_static(np.zeros((100000, 49), np.float32))
