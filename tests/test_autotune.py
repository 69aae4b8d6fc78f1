"""Autotuner bookkeeping (no GPU): candidate spaces and the committed choices."""
import json

from rtlflow_amd import autotune


def test_compile_candidates_cover_block_and_memory_layout():
    c = autotune.compile_candidates("rv32i")
    assert {o["block"] for o in c} == {64, 128, 256}
    assert {o["lane_major_min_depth"] for o in c} == set(autotune.MEM_LAYOUT)
    assert {o["threads"] for o in c} == {2, 4, 8, 16}
    assert autotune.compile_candidates("net2") == [] or all("block" in o for o in autotune.compile_candidates("net2"))


def test_runtime_candidates():
    wl = autotune.WORKLOADS["rv32i"]
    assert [f for f, _ in autotune.runtime_candidates("rv32i", {}, wl)] == ["batched"] * 6
    assert {k["lane_pad"] for _, k in autotune.runtime_candidates("rv32i", {}, wl)} == {0, 64}
    man = {"bitplane": {"supported": True}}
    forms = [f for f, _ in autotune.runtime_candidates("net2", man, autotune.WORKLOADS["net2"])]
    assert forms.count("bitplane") == 3 and "graph" in forms and "fused" in forms


def test_committed_choices_are_the_fastest_exact_rows():
    for name in autotune.WORKLOADS:
        p = autotune.TUNED / f"{name}.json"
        if not p.exists():
            continue
        d = json.loads(p.read_text())
        exact = [r for r in d["table"] if r["bit_exact"]]
        assert d["best"] == max(exact, key=lambda r: r["rate"])
        assert all(r["bit_exact"] for r in d["table"])
