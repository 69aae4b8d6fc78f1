"""Lane-k state diff (the role of the reference's checkN asserter,
src/V3EmitC.cpp:4380-4536): every stored variable and memory of one lane of
the one-lane host build, decoded through the design manifest, equals the IR
interpreter's state for the same stimulus."""
import pytest

from lanediff_util import compare, interp_lane_state
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.designs.fuzz import verilog
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.verilog import elaborate


def state_names(man, want):
    return [v["name"] for v in man["vars"] if v["name"] in want] + \
        [m["name"] for m in man.get("mems", []) if m["name"] in want]


@pytest.mark.parametrize("seed", [1, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_host_lane_state_matches_interpreter(seed, mode):
    src = verilog(seed, name=f"fuzz{seed}")
    comp = Compiled(elaborate(src), Options(threads=3))
    hd = HostDesign(comp)
    lane, nlanes, cycles, stim_seed = 5, 8, 9, 0xBEEF
    got = hd.lane_state(lane, nlanes, stim_seed, cycles, mode=mode)
    want = interp_lane_state(elaborate(src), stim_seed, lane, cycles)
    names = state_names(comp.manifest(), want)
    assert len(names) >= 5
    assert not compare(got, want, names)


def test_diff_detects_one_cycle_of_drift():
    src = verilog(2, name="fuzz2")
    comp = Compiled(elaborate(src), Options(threads=3))
    got = HostDesign(comp).lane_state(1, 4, 0xBEEF, 6)
    want = interp_lane_state(elaborate(src), 0xBEEF, 1, 7)
    assert compare(got, want, state_names(comp.manifest(), want))


@pytest.mark.parametrize("mode", [0, 1])
def test_netlist_lane_state_all_stored_values(mode):
    """Every stored value, including combinational values kept for transfer
    between mtasks, holds its settled value after a run (a 2,000-gate
    netlist: the fused form packs 1-bit values into shared register slots)."""
    from rtlflow_amd.designs.netlist import generate
    src, _ = generate(n_gates=2000, n_ff=200, n_pi=64, n_po=64, depth=20, seed=99, name="nls")
    comp = Compiled(elaborate(src), Options(threads=4))
    got = HostDesign(comp).lane_state(3, 8, 0x5EED, 11, mode=mode)
    want = interp_lane_state(elaborate(src), 0x5EED, 3, 11)
    names = state_names(comp.manifest(), want)
    assert len(names) > 300
    assert not compare(got, want, names)
