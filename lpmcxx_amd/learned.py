"""Learned Taylor order and acceptance for the batch treecode.

Python access to the order-model entry points of ``include/lpm_gpu.h``
(``lpm_model_*``, ``lpm_tree_labels``, ``lpm_tree_eval_learned``) and the
training pipeline built on them:

1. ``labels``: the FP64 treecode scores every order p = 1..12 on a sample of
   (target batch, node) pairs against the exact sum over the node;
2. ``order_labels``: class = lowest order meeting the per-pair error bound
   ``tau`` (0 = descend), with leaves kept in the near field when that is
   cheaper than the series;
3. ``OrderModel.train``: Adam on the GPU through the forward/backward MFMA
   kernels;
4. ``evaluate``: the batch treecode with the model choosing p and acceptance.

Nothing here has a counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import LPM_F32C, LPM_F64, Context, _as_host, _dev, _numel, _ptr

NPARAMS = 6288
NIN = 16
NCLASSES = 13
NORDERS = 12
NFEATURES = 10
FEATURE_NAMES = ("log2 r_n/s", "log2 r_b/s", "(r_n+r_b)/s", "log2 M0/(u_ref s)", "log2 M1/(M0 r_n)",
                 "log2 M2/(M0 r_n^2)", "log2 npart", "nvalid/64", "log2 s", "leaf")
# parameter layout (row = input): W1[16][64] b1[64] W2[64][64] b2[64] W3[64][16] b3[16]
LAYOUT = (("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 16)), ("b3", (16,)))


def unpack(params: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = params[o:o + n].reshape(shape)
        o += n
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], np.float32).ravel() for name, _ in LAYOUT]).astype(np.float32)


def nterms(p) -> np.ndarray:
    p = np.asarray(p, np.int64)
    return p * (p + 1) * (p + 2) // 6


def _rows(x, y):
    """Model rows as float32 [n][16] and labels as int32 [n]: ValueError on
    other shapes (the kernels read 16 inputs per row and one label per row)."""
    x = np.ascontiguousarray(x, np.float32)
    if x.ndim != 2 or x.shape[1] != NIN:
        raise ValueError(f"model inputs must be [n][{NIN}], got {x.shape}")
    if y is not None:
        y = np.ascontiguousarray(y, np.int32).reshape(-1)
        if y.size != x.shape[0]:
            raise ValueError(f"{y.size} labels for {x.shape[0]} rows")
    return x, y


class OrderModel:
    """The order model of a context (BF16 MFMA kernels, FP32 master weights)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def init(self, seed: int = 1) -> None:
        self.ctx.check(self.ctx.lib.lpm_model_init(self.ctx.h, seed))

    def set_params(self, params, mean=None, inv_std=None) -> None:
        (p,) = _as_host({"params": params}, NPARAMS, np.float32)
        m, s = _as_host({"mean": mean, "inv_std": inv_std}, NIN, np.float32)
        self.ctx.check(self.ctx.lib.lpm_model_set_params(self.ctx.h, _ptr(p), _ptr(m), _ptr(s)))

    def get_params(self):
        p = np.empty(NPARAMS, np.float32)
        m, s = np.empty(NIN, np.float32), np.empty(NIN, np.float32)
        self.ctx.check(self.ctx.lib.lpm_model_get_params(self.ctx.h, _ptr(p), _ptr(m), _ptr(s)))
        return p, m, s

    def set_range(self, lo=None, hi=None) -> None:
        """Clamp inputs to [lo, hi] (None: unbounded); training sets the range
        of its rows."""
        if (lo is None) != (hi is None):
            raise ValueError("give both lo and hi, or neither")
        a, b = _as_host({"lo": lo, "hi": hi}, NIN, np.float32)
        self.ctx.check(self.ctx.lib.lpm_model_set_range(self.ctx.h, _ptr(a), _ptr(b)))

    def get_range(self):
        lo, hi = np.empty(NIN, np.float32), np.empty(NIN, np.float32)
        self.ctx.check(self.ctx.lib.lpm_model_get_range(self.ctx.h, _ptr(lo), _ptr(hi)))
        return lo, hi

    def gradient(self, x, y):
        x, y = _rows(x, y)
        g = np.empty(NPARAMS, np.float32)
        loss = C.c_float()
        self.ctx.check(self.ctx.lib.lpm_model_gradient(self.ctx.h, _ptr(x), _ptr(y), len(y), _ptr(g),
                                                       C.byref(loss)))
        return g, loss.value

    def train(self, x, y, epochs: int = 10, batch: int = 4096, lr: float = 2e-3, seed: int = 1) -> float:
        x, y = _rows(x, y)
        loss = C.c_float()
        self.ctx.check(self.ctx.lib.lpm_model_train(self.ctx.h, _ptr(x), _ptr(y), len(y), epochs, batch,
                                                    C.c_float(lr), seed, C.byref(loss)))
        return loss.value

    def forward(self, x):
        x, _ = _rows(x, None)
        n = x.shape[0]
        logits = np.empty((n, NIN), np.float32)
        cls = np.empty(n, np.int32)
        self.ctx.check(self.ctx.lib.lpm_model_forward(self.ctx.h, _ptr(x), n, _ptr(logits), _ptr(cls)))
        return logits, cls


@dataclass
class LabelSet:
    features: np.ndarray  # [n][16] raw model inputs
    errors: np.ndarray    # [n][12] |du| / u_ref of the order-1..12 far field (max over the batch)
    ids: np.ndarray       # [n][2] (batch, node)


def labels(ctx: Context, targets, theta_max: float, u_ref: float, delta: float = 0.0, batch_stride: int = 1,
           keep_rate: float = 1.0, frontier_levels: int = 2, seed: int = 1, max_samples: int = 1 << 20) -> LabelSet:
    """Training data from the FP64 treecode (lpm_tree_labels); the current tree
    needs moments of order >= 12 with moment_sign +1.  frontier_levels limits
    the candidates to the first levels at and below the first node of each
    branch that passes theta_max: the pairs the learned traversal meets."""
    tx, ty, tz = _as_host({"tx": targets[0], "ty": targets[1], "tz": targets[2]})
    f = np.empty((max_samples, NIN), np.float32)
    e = np.empty((max_samples, NORDERS), np.float32)
    ids = np.empty((max_samples, 2), np.int32)
    n = C.c_int64()
    ctx.check(ctx.lib.lpm_tree_labels(ctx.h, theta_max, delta, u_ref, batch_stride, keep_rate, frontier_levels,
                                      seed, max_samples,
                                      _ptr(tx), _ptr(ty), _ptr(tz), tx.size, _ptr(f), _ptr(e), _ptr(ids),
                                      C.byref(n)))
    k = n.value
    return LabelSet(f[:k].copy(), e[:k].copy(), ids[:k].copy())


# relative cost of one far-field series term and of one near-field pair
# (vector instructions per lane: ~9 for a term, ~11 for a pair)
FAR_TERM_COST = 9.0
NEAR_PAIR_COST = 11.0
# stereographic entries (LPM_OPT_STEREO): ~25 instructions per entry and ~4
# per complex term; ~13 per near pair in that kernel
STEREO_ENTRY_COST = 25.0
STEREO_TERM_COST = 4.0
STEREO_NEAR_PAIR_COST = 13.0


def order_labels(ls: LabelSet, tau: float, max_order: int, stereo_extra: int | None = None) -> np.ndarray:
    """Class of every sample: the lowest order p <= max_order whose error is at
    most tau, or 0 (descend) when there is none or, at a leaf, when the
    series would cost more than summing the leaf directly.  With stereo_extra
    (the LPM_OPT_STEREO_EXTRA_ORDER of the evaluation) a far leaf is priced as
    a stereographic entry of order p + stereo_extra."""
    ok = ls.errors[:, :max_order] <= tau
    has = ok.any(axis=1)
    p = np.where(has, ok.argmax(axis=1) + 1, 0)
    leaf = ls.features[:, 9] > 0.5
    npart = np.exp2(ls.features[:, 6])
    if stereo_extra is None:
        dearer = leaf & (FAR_TERM_COST * nterms(p) >= NEAR_PAIR_COST * npart)
    else:
        far = STEREO_ENTRY_COST + STEREO_TERM_COST * (p + stereo_extra + 1)
        dearer = leaf & (far >= STEREO_NEAR_PAIR_COST * npart)
    return np.where(dearer, 0, p).astype(np.int32)


def normalization(features: np.ndarray):
    """Per-input mean and inverse standard deviation (padded inputs: 0 / 1)."""
    mean = np.zeros(NIN, np.float32)
    inv = np.ones(NIN, np.float32)
    x = features[:, :NFEATURES].astype(np.float64)
    mean[:NFEATURES] = x.mean(axis=0)
    sd = x.std(axis=0)
    inv[:NFEATURES] = 1.0 / np.maximum(sd, 1e-3)
    return mean, inv


def evaluate(ctx: Context, targets, theta_max: float, margin: int, max_order: int, u_ref: float,
             delta: float = 0.0, prec: int = LPM_F32C):
    tx, ty, tz = _as_host({"tx": targets[0], "ty": targets[1], "tz": targets[2]})
    n = tx.size
    ux, uy, uz = np.zeros(n), np.zeros(n), np.zeros(n)
    ctx.check(ctx.lib.lpm_tree_eval_learned(ctx.h, prec, theta_max, margin, max_order, u_ref, delta, _ptr(tx),
                                            _ptr(ty), _ptr(tz), n, _ptr(ux), _ptr(uy), _ptr(uz)))
    return ux, uy, uz


def evaluate_dev(ctx: Context, tx, ty, tz, out, theta_max: float, margin: int, max_order: int, u_ref: float,
                 delta: float = 0.0, prec: int = LPM_F32C) -> None:
    _dev(ctx, {"tx": tx, "ty": ty, "tz": tz, "ux": out[0], "uy": out[1], "uz": out[2]}, _numel(tx, "tx"))
    ctx.after_torch()
    ctx.check(ctx.lib.lpm_tree_eval_learned_dev(ctx.h, prec, theta_max, margin, max_order, u_ref, delta, _ptr(tx),
                                                _ptr(ty), _ptr(tz), tx.numel(), _ptr(out[0]), _ptr(out[1]),
                                                _ptr(out[2])))


def constant_model(ctx: Context, cls: int) -> None:
    """A model that answers `cls` for every input (zero weights, one large
    output bias): the learned traversal then reduces to the plain one."""
    parts = {name: np.zeros(shape, np.float32) for name, shape in LAYOUT}
    parts["b3"][cls] = 10.0
    OrderModel(ctx).set_params(pack(parts))


def u_ref_of(gamma) -> float:
    """Velocity scale sum |gamma| / (4 pi) used by the features and labels."""
    return float(np.abs(np.asarray(gamma, np.float64)).sum() / (4.0 * np.pi))


@dataclass
class FitResult:
    n_samples: int
    class_counts: np.ndarray
    final_loss: float
    accuracy: float
    mean: np.ndarray
    inv_std: np.ndarray
    lo: np.ndarray        # input range recorded by training
    hi: np.ndarray


def fit(ctx: Context, ls: LabelSet, tau: float, max_order: int, *, epochs: int = 20, batch: int = 4096,
        lr: float = 2e-3, seed: int = 1, stereo_extra: int | None = None) -> FitResult:
    """Train the context's order model on a label set for the bound tau
    (stereo_extra: leaves priced for the stereographic far field)."""
    y = order_labels(ls, tau, max_order, stereo_extra)
    mean, inv = normalization(ls.features)
    m = OrderModel(ctx)
    m.init(seed)
    p, _, _ = m.get_params()
    m.set_params(p, mean, inv)
    loss = m.train(ls.features, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    _, cls = m.forward(ls.features)
    lo, hi = m.get_range()
    return FitResult(len(y), np.bincount(y, minlength=NCLASSES), loss, float((cls == y).mean()), mean, inv, lo, hi)
