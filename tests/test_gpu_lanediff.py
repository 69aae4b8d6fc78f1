"""Lane-k state diff on the GPU: rf_lane_state of one lane after the
clock-cycle protocol (plain graph and fused forms), decoded through the
design manifest, equals the IR interpreter's state for the same stimulus,
variable by variable and memory by memory."""
import pytest

from lanediff_util import compare, interp_lane_state
from rtlflow_amd.designs import registry
from rtlflow_amd.designs.fuzz import verilog
from rtlflow_amd.statediff import decode_state
from rtlflow_amd.verilog import elaborate


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_lane_state(gpu, seed, mode):
    cd = registry.get(f"fuzz{seed}")
    lanes, lane, cycles, stim_seed = 40, 37, 10, 0xD1FF + seed
    with cd.model(lanes) as m:
        m.begin(stim_seed)
        m.run_cycles(cycles, mode)
        got = decode_state(cd.manifest, m.lane_state(lane))
    want = interp_lane_state(elaborate(verilog(seed, name=f"fuzz{seed}")), stim_seed, lane, cycles)
    names = [v["name"] for v in cd.manifest["vars"] if v["name"] in want] + \
        [x["name"] for x in cd.manifest.get("mems", []) if x["name"] in want]
    assert len(names) >= 5
    bad = compare(got, want, names)
    assert not bad, bad[:5]
