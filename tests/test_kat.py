"""Known answers from the reference's own self-checking tests: the vectors
of test_regress/t/t_math_shift.v, t_math_div.v, t_math_signed.v and
t_math_red.v (tests/golden/kat_math.json, written by tests/golden/make_kat.py)
stepped through by this project's known-answer designs
(rtlflow_amd/designs/kat.py). A design counts mismatches against the
reference's expected values, so every correct simulator hashes the closed
form err = 0, idx = min(c + 1, N). CPU: the IR interpreter, the one-lane host
build in every schedule form, the bit-plane program's NumPy model; and a
vector with a planted wrong expected value is caught."""
import pytest

import ir_interp
import oracle
from bitplane_sim import Program
from rtlflow_amd import bitplane as bp
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.designs import kat
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.verilog import elaborate

FAMILIES = list(kat.FAMILIES)


def test_fixture_holds_the_reference_checks():
    v = kat.vectors()
    assert len(v["shift"]) == 129 and len(v["div"]) == 28 and len(v["signed"]) == 7 and len(v["red"]) == 12
    # spot checks against the reference file lines the fixture names
    d = {x["src"]: x for x in v["div"]}
    assert d["test_regress/t/t_math_div.v:54"]["expect"] == 0x00000403AD81C0DA
    s = [x for x in v["shift"] if x["op"] == "rights" and x["width"] == 96 and x["amount"] == 28]
    assert s and s[0]["expect"] == 0xFFFFFFFF784BF8F127340891
    assert all(x["src"].startswith("test_regress/t/t_math_") for k in v for x in v[k])


@pytest.mark.parametrize("family", FAMILIES)
def test_kat_interpreter_and_host_build(family):
    n = len(kat.vectors()[family])
    cycles = n + 3
    want = kat.expected_hashes(oracle.fnv, n, cycles, 3)
    d = elaborate(kat.verilog(family))
    assert ir_interp.protocol(d, 0x4B41, 3, cycles) == want
    comp = Compiled(elaborate(kat.verilog(family)), Options(threads=2))
    hd = HostDesign(comp)
    for mode in (0, 1, 2, 3):
        assert list(hd.protocol(3, 0x4B41, cycles, mode=mode, order_seed=2)) == want, mode
    assert list(hd.queue(2, 0x4B41, [cycles] * 3)) == want
    if comp.bitplane_info["supported"]:
        got = Program(bp.encode(bp.compile_program(comp))).run(0x4B41, 3, cycles)
        assert list(got) == want


@pytest.mark.parametrize("family", ["shift", "div"])
def test_kat_catches_a_planted_wrong_answer(family):
    """One expected value changed by one bit: the design reports one
    mismatch from that vector on, so the hash leaves the closed form."""
    vecs = [dict(x) for x in kat.vectors()[family]]
    k = len(vecs) // 2
    vecs[k]["expect"] ^= 1
    src = kat.FAMILIES[family](vecs, name=f"kat_{family}_bad")
    n = len(vecs)
    got = ir_interp.protocol(elaborate(src), 1, 1, n + 1)
    assert got != kat.expected_hashes(oracle.fnv, n, n + 1, 1)
    h = 0xCBF29CE484222325
    for c in range(n + 1):
        err = 1 if c >= k else 0
        h = oracle.fnv(err.to_bytes(4, "little") + bytes([min(c + 1, n)]), h)
    assert got == [h]


@pytest.mark.gpu
@pytest.mark.parametrize("family", FAMILIES)
def test_kat_every_gpu_form(gpu, family):
    from rtlflow_amd.designs import registry
    from rtlflow_amd.model import MODE_BITPLANE, MODE_FUSED, MODE_GRAPH
    n = len(kat.vectors()[family])
    cycles = n + 3
    lanes = 130
    want = kat.expected_hashes(oracle.fnv, n, cycles, lanes)[0]
    cd = registry.get(f"kat_{family}")
    modes = [MODE_GRAPH, MODE_FUSED]
    if cd.manifest.get("bitplane", {}).get("supported"):
        modes.append(MODE_BITPLANE)
    for mode in modes:
        with cd.model(lanes) as m:
            m.begin(0x4B41)
            m.run_cycles(cycles, mode)
            got = m.hashes()
        assert (got == want).all(), (family, mode)
    with cd.model(64) as m:
        got = m.submit_queue(0x4B41, [cycles] * 100)
    assert (got == want).all()
