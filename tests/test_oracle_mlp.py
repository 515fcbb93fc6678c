"""CPU checks of the order-model oracle (oracle/mlp_oracle.py) and the host
side of the learned treecode (lpmcxx_amd/learned.py).

The oracle's float64 backward pass is pinned by central finite differences;
its BF16 rounding by torch's float32 -> bfloat16 conversion (round to
nearest even); the label rule by hand-made error tables.
"""
import numpy as np
import pytest
import torch

from lpmcxx_amd import learned as LN
from oracle import mlp_oracle as MO


def _params(seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.3, LN.NPARAMS)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=5000) * 10.0 ** rng.integers(-20, 20, 5000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 65504.0]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(MO.bf16(x), ref)
    assert MO.bf16(np.float32(1.0 + 2 ** -8)) == 1.0          # tie to even
    assert MO.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def test_float64_backward_matches_finite_differences():
    p = _params(1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(37, 16))
    x[:, 10:] = 0.0
    y = rng.integers(0, 13, 37)
    _, g = MO.loss_and_grad(p, x, y, bf=False)
    idx = np.concatenate([rng.choice(1024, 12, replace=False),                 # W1
                          1024 + rng.choice(64, 4, replace=False),             # b1
                          1088 + rng.choice(4096, 12, replace=False),          # W2
                          5184 + rng.choice(64, 4, replace=False),             # b2
                          5248 + rng.choice(1024, 12, replace=False),          # W3
                          6272 + np.arange(13)])                               # b3 (used classes)
    num = MO.numeric_grad(p, x, y, idx)
    assert np.allclose(g[idx], num, rtol=1e-6, atol=1e-9)
    assert np.all(g[6272 + 13:] == 0.0)                                        # unused classes


def test_bf16_graph_close_to_float64():
    p = _params(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(200, 16))
    x[:, 10:] = 0.0
    y = rng.integers(0, 13, 200)
    l64, g64 = MO.loss_and_grad(p, x, y, bf=False)
    l16, g16 = MO.loss_and_grad(p, x, y, bf=True)
    assert abs(l16 - l64) < 0.05 * abs(l64)
    assert np.linalg.norm(g16 - g64) < 0.1 * np.linalg.norm(g64)


def test_pack_unpack_layout():
    p = np.arange(LN.NPARAMS, dtype=np.float32)
    parts = LN.unpack(p)
    assert parts["W1"].shape == (16, 64) and parts["b3"].shape == (16,)
    assert parts["W2"][0, 0] == 16 * 64 + 64
    assert np.array_equal(LN.pack(parts), p)
    assert np.array_equal(MO.pack(MO.unpack(p)), p)


def test_order_labels_rule():
    f = np.zeros((5, 16), np.float32)
    e = np.full((5, 12), 1.0, np.float32)
    e[0, 3:] = 1e-9            # order 4 suffices
    e[1, :] = 1e-9             # order 1
    e[2, 11] = 1e-9            # only order 12 -> beyond max_order 10 -> descend
    e[3, 7:] = 1e-9            # order 8 at a leaf of 64 particles: 120 terms x 9 > 64 x 11 -> descend
    f[3, 9] = 1.0
    f[3, 6] = 6.0
    e[4, 2:] = 1e-9            # order 3 at the same leaf: 10 x 9 < 704 -> accept
    f[4, 9] = 1.0
    f[4, 6] = 6.0
    y = LN.order_labels(LN.LabelSet(f, e, np.zeros((5, 2), np.int32)), 1e-8, 10)
    assert list(y) == [4, 1, 0, 0, 3]


def test_normalization():
    rng = np.random.default_rng(0)
    f = np.zeros((1000, 16), np.float32)
    f[:, :10] = rng.normal(3.0, 2.0, (1000, 10))
    f[:, 9] = 1.0               # constant input: unit-ish scale, no division by zero
    mean, inv = LN.normalization(f)
    assert np.allclose(mean[:9], 3.0, atol=0.2) and np.allclose(1 / inv[:9], 2.0, atol=0.2)
    assert np.isfinite(inv).all() and inv[9] == pytest.approx(1e3) and np.all(inv[10:] == 1.0)
