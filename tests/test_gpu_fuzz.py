"""GPU parity on random operator-coverage designs against the IR interpreter
(every operator, widths 1-130 bits, signed casts, dynamic selects, memories,
conditional / partial non-blocking writes)."""
import numpy as np
import pytest

import ir_interp
from rtlflow_amd.designs import registry
from rtlflow_amd.designs.fuzz import verilog
from rtlflow_amd.verilog import elaborate


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_fuzz_design(gpu, seed):
    n, cycles, stim_seed = 24, 12, 0xF00D + seed
    ref = np.array(ir_interp.protocol(elaborate(verilog(seed, name=f"fuzz{seed}")), stim_seed, n, cycles),
                   dtype=np.uint64)
    cd = registry.get(f"fuzz{seed}")
    for mode in (0, 1):
        with cd.model(n) as m:
            m.begin(stim_seed)
            m.run_cycles(cycles, mode)
            got = m.hashes()
        assert (got == ref).all(), f"mode {mode}: {int((got != ref).sum())} of {n} differ"
    with cd.model(7) as m:
        got = m.submit_queue(stim_seed, [cycles] * n)
    assert (got == ref).all()
