"""Argument checks of the Python wrappers (CPU).

Every wrapper validates lengths, dtypes and devices before it calls into the
native library: the kernels take the count from one argument, so a short,
float32 or host tensor would otherwise be read or written out of bounds.  A
context stub without a native handle stands in for lpm_create here; reaching
the library at all fails the test with an AttributeError instead of the
expected ValueError.
"""
import numpy as np
import pytest
import torch

import lpmcxx_amd as L
from lpmcxx_amd import learned as LN
from lpmcxx_amd import tree as T


@pytest.fixture
def stub():
    c = object.__new__(L.Context)
    c.device, c.lib, c.h = 0, None, None
    return c


def _pts(n):
    return tuple(np.zeros(n) for _ in range(3))


def test_direct_sum_host_lengths(stub):
    with pytest.raises(ValueError, match="ty"):
        stub.direct_sum((np.zeros(4), np.zeros(3), np.zeros(4)), _pts(5), np.zeros(5))
    with pytest.raises(ValueError, match="gamma"):
        stub.direct_sum(_pts(4), _pts(5), np.zeros(4))
    with pytest.raises(ValueError, match="skip"):
        stub.direct_sum(_pts(4), _pts(5), np.zeros(5), skip=np.zeros(3, np.int64))


def test_direct_sum_dev_tensors(stub):
    t = [torch.zeros(8, dtype=torch.float64) for _ in range(7)]
    with pytest.raises(ValueError, match="cuda"):  # host tensors
        stub.direct_sum_dev(*t[:3], *t[3:6], t[6], None)
    t32 = [torch.zeros(8, dtype=torch.float32) for _ in range(7)]
    with pytest.raises(ValueError, match="dtype"):
        stub.direct_sum_dev(*t32[:3], *t32[3:6], t32[6], None)
    with pytest.raises(TypeError):
        stub.direct_sum_dev(*(np.zeros(8),) * 7, None)


def test_dev_check_length_and_device():
    if not torch.cuda.is_available():
        stub = object.__new__(L.Context)
        stub.device = 0
        a = torch.zeros(4, dtype=torch.float64)
        with pytest.raises(ValueError, match="cuda:0"):
            L._dev(stub, {"x": a}, 4)
    else:
        stub = object.__new__(L.Context)
        stub.device = 0
        a = torch.zeros(4, dtype=torch.float64, device="cuda")
        with pytest.raises(ValueError, match="entries"):
            L._dev(stub, {"x": a}, 5)
        with pytest.raises(ValueError, match="contiguous"):
            L._dev(stub, {"x": torch.zeros(8, dtype=torch.float64, device="cuda")[::2]}, 4)
    with pytest.raises(ValueError, match="required"):
        L._dev(stub, {"x": None}, 4)


def test_tree_wrappers(stub):
    with pytest.raises(ValueError, match="build"):
        T.moments(stub, np.zeros(4), 4, 1)
    stub.tree_sources = 10
    with pytest.raises(ValueError, match="gamma has 4"):
        T.moments(stub, np.zeros(4), 4, 1)
    with pytest.raises(ValueError, match="gamma has 4"):
        T.moments_psi(stub, np.zeros(4), 4)
    with pytest.raises(ValueError, match="sy"):
        T.build(stub, (np.zeros(4), np.zeros(5), np.zeros(4)), 1.0, 8)
    with pytest.raises(ValueError, match="box_max"):
        T.build_box(stub, _pts(4), np.zeros(3), np.zeros(2), 0, 8)
    with pytest.raises(ValueError, match="tz"):
        T.evaluate(stub, (np.zeros(4), np.zeros(4), np.zeros(3)), 0.5)
    with pytest.raises(ValueError, match="cuda"):
        T.evaluate_dev(stub, *(torch.zeros(4, dtype=torch.float64) for _ in range(3)),
                       [torch.zeros(4, dtype=torch.float64) for _ in range(3)], 0.5)
    with pytest.raises(ValueError, match="cuda"):
        T.moments_dev(stub, torch.zeros(10, dtype=torch.float64), 4, 1)


def test_rk4_and_bve_wrappers(stub):
    cfg = L.SolverConfig(method=L.LPM_METHOD_DIRECT)  # a ctypes struct: no library call
    with pytest.raises(ValueError, match="area"):
        stub.rk4_steps(cfg, 0.01, 0.0, 1, *_pts(6), np.zeros(6), np.zeros(5))
    with pytest.raises(ValueError, match="tracer y"):
        stub.bve_steps(cfg, 0.01, 0.0, 1, _pts(6), np.zeros(6), np.zeros(6),
                       tracers=(np.zeros(3), np.zeros(2), np.zeros(3)))
    with pytest.raises(ValueError, match="n_src"):
        stub.bve_step_dev(cfg, 0.01, 0.0, 9, *(torch.zeros(6, dtype=torch.float64) for _ in range(5)))


def test_learned_wrappers(stub):
    m = LN.OrderModel(stub)
    with pytest.raises(ValueError, match="params"):
        m.set_params(np.zeros(LN.NPARAMS - 1))
    with pytest.raises(ValueError, match="mean"):
        m.set_params(np.zeros(LN.NPARAMS), np.zeros(3), np.ones(LN.NIN))
    with pytest.raises(ValueError, match="both"):
        m.set_range(np.zeros(LN.NIN), None)
    with pytest.raises(ValueError, match="inputs"):
        m.gradient(np.zeros((4, 10)), np.zeros(4))
    with pytest.raises(ValueError, match="labels"):
        m.train(np.zeros((4, LN.NIN)), np.zeros(3))
    with pytest.raises(ValueError, match="ty"):
        LN.evaluate(stub, (np.zeros(4), np.zeros(3), np.zeros(4)), 0.9, 0, 6, 1.0)
