"""GPU parity for the pipelined ALU (synthetic design 1) against its C golden model."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry
from rtlflow_amd.model import MODE_FUSED, MODE_GRAPH

SEED = registry.SEED_CFG[1]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_FUSED])
def test_alu_protocol(gpu, mode):
    n, cycles = 1024, 300
    ref = oracle.alu_run(SEED, n, cycles)
    with registry.get("alu").model(n) as m:
        m.begin(SEED)
        m.run_cycles(cycles, mode)
        got = m.hashes()
    assert (got == ref).all(), f"{int((got != ref).sum())} of {n} differ"


@pytest.mark.gpu
def test_alu_queue_uneven(gpu):
    rng = np.random.default_rng(2)
    budgets = rng.integers(0, 400, 3000).astype(np.uint32)
    want = np.array([oracle.alu_run(SEED, 1, int(b), stim0=i)[0] for i, b in enumerate(budgets)],
                    dtype=np.uint64)
    with registry.get("alu").model(512) as m:
        got = m.submit_queue(SEED, budgets)
    assert (got == want).all()
