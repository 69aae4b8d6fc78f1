"""The heuristic partition judges the learned one: kept only when faster and
bit-exact against both the heuristic's hashes and the checker's; otherwise
the heuristic is the fallback (no GPU: the decision and the lookup)."""
import json

import numpy as np

from rtlflow_amd import judge


def test_learned_kept_only_when_faster_and_exact():
    h = np.arange(10, dtype=np.uint64)
    v = judge.decide(1.0, 1.2, h, h.copy(), expected=h[:4], sample=np.arange(4))
    assert v["kept"] and v["choice"] == "learned" and abs(v["speedup"] - 1.2) < 1e-12
    # slower: the heuristic stays
    v = judge.decide(1.0, 0.93, h, h.copy(), expected=h[:4], sample=np.arange(4))
    assert not v["kept"] and v["choice"] == "heuristic" and v["bit_exact"]
    # faster but one hash differs from the heuristic's
    bad = h.copy()
    bad[7] ^= 1
    v = judge.decide(1.0, 2.0, h, bad)
    assert not v["kept"] and not v["bit_exact"]
    # faster and equal to the heuristic, but both differ from the checker
    v = judge.decide(1.0, 2.0, h, h.copy(), expected=h[:4] + 1, sample=np.arange(4))
    assert not v["kept"]


def test_scheduled_follows_the_committed_verdicts(tmp_path, monkeypatch):
    monkeypatch.setattr(judge, "JUDGE", tmp_path / "judge.json")
    judge.record("net5s", "graph", {"kept": True})
    judge.record("net5s", "batched", {"kept": False})
    d = json.loads((tmp_path / "judge.json").read_text())
    assert d["net5s"]["graph"]["kept"] and not d["net5s"]["batched"]["kept"]


def test_committed_verdicts_are_consistent():
    for name, forms in judge.verdicts().items():
        for form, v in forms.items():
            assert v["kept"] == (v["bit_exact"] and v["learned"] > v["heuristic"]
                                 and not v.get("identical", False)), (name, form)


def test_scheduled_takes_the_fastest_kept_variant(tmp_path, monkeypatch):
    from rtlflow_amd.designs import registry
    monkeypatch.setattr(judge, "JUDGE", tmp_path / "judge.json")
    monkeypatch.setattr(registry, "load_index", lambda: {"d": 1, "d_learned": 1, "d_lmerge": 1})
    monkeypatch.setattr(registry, "get", lambda name: name)
    judge.record("d", "graph", {"kept": True, "speedup": 1.1})
    judge.record("d", "graph" + judge.MERGE, {"kept": True, "speedup": 1.3})
    judge.record("d", "batched", {"kept": True, "speedup": 1.2})
    judge.record("d", "batched" + judge.MERGE, {"kept": False, "speedup": 1.5})
    assert judge.scheduled("d", "graph") == "d_lmerge"
    assert judge.scheduled("d", "batched") == "d_learned"
    assert judge.scheduled("d", "other") == "d"
