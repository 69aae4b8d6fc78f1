"""GPU parity for the systolic array + 4,096-bit chain (synthetic design 4)."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry
from rtlflow_amd.model import MODE_FUSED, MODE_GRAPH

SEED = registry.SEED_CFG[4]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_FUSED])
def test_systolic_protocol(gpu, mode):
    n, cycles = 512, 200
    ref = oracle.systolic_run(SEED, n, cycles)
    with registry.get("systolic").model(n) as m:
        m.begin(SEED)
        m.run_cycles(cycles, mode)
        got = m.hashes()
    assert (got == ref).all(), f"{int((got != ref).sum())} of {n} differ"


@pytest.mark.gpu
def test_systolic_queue(gpu):
    budgets = np.array([0, 1, 2, 50, 199, 7, 64, 128] * 40, np.uint32)
    want = np.array([oracle.systolic_run(SEED, 1, int(b), stim0=i)[0] for i, b in enumerate(budgets)],
                    dtype=np.uint64)
    with registry.get("systolic").model(128) as m:
        got = m.submit_queue(SEED, budgets)
    assert (got == want).all()
