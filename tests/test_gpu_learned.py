"""The learned partition (granularity chosen by the MLP) changes the mtask
graph, not the results: graph-form hashes of net2 equal the heuristic
partition's and the oracle's."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry
from rtlflow_amd.model import MODE_GRAPH


@pytest.mark.gpu
def test_learned_partition_is_bit_exact(gpu):
    lanes, cycles, seed = 2048, 20, registry.SEED_CFG[2]
    base, learned = registry.get("net2"), registry.get("net2_learned")
    assert learned.manifest["mtasks"] != base.manifest["mtasks"]
    out = []
    for cd in (base, learned):
        with cd.model(lanes) as m:
            m.begin(seed)
            m.run_cycles(cycles, MODE_GRAPH)
            out.append(m.hashes())
    assert (out[0] == out[1]).all()
    ref = oracle.netlist_run(registry.table_bytes("net2"), seed, 32, cycles)
    assert (out[1][:32] == ref).all()
