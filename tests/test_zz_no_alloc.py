"""GPU: rf_submit_queue_static, the batched entry point that allocates
nothing (its queue buffers are a fixed device array), against
rf_submit_queue on the same workloads through the C ABI: two calls in a row,
the second smaller, and the refusal of a workload beyond its fixed size."""
import ctypes as C

import numpy as np
import pytest

from rtlflow_amd.designs import registry
from rtlflow_amd.model import RFError, runtime

SEED = registry.SEED_CFG[2]
RF_QUEUE_STATIC_MAX = 65536


def _static(m, seed, budgets, order=None):
    rt = runtime()
    b = np.ascontiguousarray(budgets, np.uint32)
    out = np.zeros(len(b), np.uint64)
    optr = None
    if order is not None:
        o = np.ascontiguousarray(order, np.uint32)
        optr = o.ctypes.data
    rc = rt.rf_submit_queue_static(m.h, seed, b.ctypes.data, len(b), optr, out.ctypes.data)
    if rc:
        raise RFError(f"rf error {rc}: {rt.rf_last_error().decode()}")
    return out


@pytest.mark.gpu
def test_static_queue_matches_allocating_queue(gpu):
    rng = np.random.default_rng(11)
    for name, lanes in (("net_small", 128), ("circ", 64)):
        cd = registry.get(name)
        with cd.model(lanes) as m:
            b1 = rng.integers(0, 50, size=700).astype(np.uint32)
            b2 = rng.integers(0, 30, size=90).astype(np.uint32)
            want1 = m.submit_queue(SEED, b1)
            want2 = m.submit_queue(SEED, b2)
            got1 = _static(m, SEED, b1)
            got2 = _static(m, SEED, b2)
            assert (got1 == want1).all() and (got2 == want2).all(), name
            perm = rng.permutation(len(b2)).astype(np.uint32)
            assert (_static(m, SEED, b2, perm) == want2).all()
        if name == "net_small":
            import oracle
            ref = oracle.netlist_run(registry.table_bytes(name), SEED, nstim=len(b2), budgets=b2)
            assert (got2 == ref).all()


@pytest.mark.gpu
def test_static_queue_refuses_oversize_and_bad_order(gpu):
    cd = registry.get("net_tiny")
    with cd.model(64) as m:
        with pytest.raises(RFError, match="exceed its fixed queue"):
            _static(m, SEED, np.ones(RF_QUEUE_STATIC_MAX + 1, np.uint32))
        with pytest.raises(RFError, match="permutation"):
            _static(m, SEED, np.ones(4, np.uint32), order=[0, 1, 1, 2])
        # the model still works after the refusals
        b = np.array([3, 0, 5, 9], np.uint32)
        assert (_static(m, SEED, b) == m.submit_queue(SEED, b)).all()
