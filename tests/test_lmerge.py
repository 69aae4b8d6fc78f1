"""Learned merge scoring in the contraction (costmodel.LearnedMerge): the
incremental union features equal the features recomputed from the vertex
sets, an additive merge model reproduces the instruction-count partition,
and a learned-merge partition simulates bit-exactly (host build against the
C oracle)."""
import json
import random
from pathlib import Path

import numpy as np
import pytest

import oracle
from rtlflow_amd import costmodel
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.costmodel import FeatureExtractor, LearnedMerge
from rtlflow_amd.designs import registry
from rtlflow_amd.designs.netlist import generate
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.partition import partition
from rtlflow_amd.verilog import elaborate

MODEL = json.loads((Path(costmodel.__file__).parent / "costtables" / "model.json").read_text())


def _comp(name, threads=4):
    spec = registry.ENTRIES[name].make()
    return Compiled(elaborate(spec["verilog"], spec["top"]), Options(threads=threads, bitplane=False))


def _features_of(lm, comp, vids):
    """The six features of a vertex set from the definitions (sets, not the
    merge model's bit sets), with the merge model's stored-variable rule."""
    sch = comp.sch
    stored = LearnedMerge._stored_like(comp)
    fx = FeatureExtractor(comp)
    inside = set(vids)
    reads, writes = set(), set()
    instr = br = wide = 0
    for v in vids:
        vx = sch.vertices[v]
        reads |= vx.reads | vx.need_old
        writes |= vx.writes
        instr += sum(costmodel.process_cost(p) for p in vx.procs)
        b, w = fx.counts[v]
        br += b
        wide += w
    fan_in = len([x for x in reads if x not in writes])
    fan_out = len([x for x in writes if fx.readers.get(x, set()) - inside])
    state = 0
    for x in reads | writes:
        if x in stored and not x.is_mem:
            state += x.slots * costmodel.CLS_BYTES[costmodel.class_of(x)]
        elif x.is_mem:
            state += costmodel.CLS_BYTES[costmodel.class_of(x)] * x.words
    return [float(instr), float(fan_in), float(fan_out), float(state), float(br), float(wide)]


def _union_features(lm, a, b):
    x = lm._x_union(a, b)
    r, w = lm.R[a] | lm.R[b], lm.W[a] | lm.W[b]
    ca, cb = lm.cnt[a], lm.cnt[b]
    state = lm.state[a] + lm.state[b] - lm._bytes((lm.R[a] | lm.W[a]) & (lm.R[b] | lm.W[b]))
    return [float(ca[0] + cb[0]), float(bin(r & ~w).count("1")), float(bin(x).count("1")),
            float(state), float(ca[1] + cb[1]), float(ca[2] + cb[2])]


@pytest.mark.parametrize("name", ["rv32i", "alu", "net_tiny"])
def test_incremental_union_features_match_definition(name):
    comp = _comp(name)
    lm = LearnedMerge(MODEL, comp)
    rng = random.Random(3)
    nv = len(comp.sch.vertices)
    # merge random pairs, checking every union before it is taken
    alive = list(range(nv))
    for _ in range(min(60, nv - 1)):
        a, b = rng.sample(alive, 2)
        want = _features_of(lm, comp, lm.members[a] + lm.members[b])
        assert _union_features(lm, a, b) == want
        lm.merge(a, b)
        alive.remove(b)
        assert sorted(lm.members[a]) == sorted(v for v in range(nv) if lm.owner[v] == a)


def test_forward_cpu_matches_learned_cost_definition():
    """exp of the de-normalised network output on normalised log1p
    features; the empty kernel costs the launch only."""
    x = [[10.0, 3.0, 2.0, 40.0, 1.0, 0.0], [0.0] * 6]
    us = costmodel.forward_cpu(MODEL, x)
    assert us.shape == (2,) and (us > 0).all()
    lm = LearnedMerge(MODEL, _comp("alu"))
    assert abs(lm._predict(x[0]) - us[0]) < 1e-9 * us[0]
    assert abs(lm.base - us[1]) < 1e-9 * us[1]


class _Additive:
    """A merge model with the instruction-count heuristic's additive cost."""

    def cost(self, ta, tb):
        return ta.cost + tb.cost

    def merge(self, keep, gone):
        pass


def test_additive_merge_model_reproduces_heuristic_partition():
    comp = _comp("net_tiny", threads=3)
    edges = [(vx.id, s) for vx in comp.sch.vertices for s in vx.succs]
    base = partition(comp.vcost, edges, 3, order=comp.sch.topo)
    hooked = partition(comp.vcost, edges, 3, order=comp.sch.topo, merged=_Additive())
    assert [sorted(t.vids) for t in base] == [sorted(t.vids) for t in hooked]


def test_learned_merge_partition_simulates_bit_exact():
    src, table = generate(n_gates=300, n_ff=30, n_pi=16, n_po=16, depth=8, seed=7)
    comp = Compiled(elaborate(src), Options(threads=3, merge_fn=costmodel.learned_merge(MODEL)))
    seen = sorted(v for t in comp.mtasks for v in t.vids)
    assert seen == list(range(len(comp.sch.vertices)))
    seed = 0x52544C4600000002
    ref = oracle.netlist_run(table.to_bytes(), seed, nstim=8, cycles=25)
    hd = HostDesign(comp)
    for mode in (0, 1, 2):
        assert (hd.protocol(8, seed, 25, mode=mode, order_seed=5) == ref).all()


def test_learned_merge_coarsens_rv32i():
    """rv32i's vertices share most of their inputs (pipeline registers,
    register file): costed as unions, merges that an additive count rejects
    fit under the critical-path limit, so the learned partition has fewer
    mtasks and a lower predicted eval time."""
    spec = registry.ENTRIES["rv32i"].make()
    d = elaborate(spec["verilog"], spec["top"])
    heur = Compiled(d, Options(threads=2, bitplane=False))
    d = elaborate(spec["verilog"], spec["top"])
    lm = Compiled(d, Options(threads=2, bitplane=False, merge_fn=costmodel.learned_merge(MODEL)))
    assert len(lm.mtasks) < len(heur.mtasks)
    pred = []
    for c in (heur, lm):
        fx = FeatureExtractor(c)
        pred.append(costmodel.forward_cpu(MODEL, [fx.mtask(k) for k in range(len(c.mtasks))]).sum())
    assert pred[1] < pred[0]


def test_forward_cpu_matches_pinned_fp64_reference():
    """The compiler's host forward reads the parameters in the device
    kernels' layout: with identity normalisation it is exp of the FP64
    reference network (pinned by finite differences) on log1p features."""
    import mlp_ref
    rng = np.random.default_rng(4)
    p = rng.normal(0, 0.3, 64 * 6 + 64 + 64 * 64 + 64 + 64 + 1)
    model = {"params": p.tolist(), "mu": [0.0] * 6, "sd": [1.0] * 6, "ym": 0.0, "ys": 1.0}
    raw = rng.integers(0, 500, (9, 6)).astype(float)
    got = costmodel.forward_cpu(model, raw)
    want = np.exp(np.asarray(mlp_ref.forward(p, np.log1p(raw), 6)).reshape(-1))
    assert np.allclose(got, want, rtol=1e-12)
