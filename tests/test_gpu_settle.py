"""GPU: the settle loop (change detection on combinational cycles, the
OR-reduced change flag, the on-device loop of the fused form and the
host-driven loop of eval()) against a closed form, and the reporting of a
loop that never settles."""
import numpy as np
import pytest

import oracle
from rtlflow_amd.designs import registry
from rtlflow_amd.model import MODE_FUSED, MODE_GRAPH

SEED = 0x5E77


def _closed(nstim, cycles):
    out = np.zeros(nstim, np.uint64)
    for s in range(nstim):
        h = 0xcbf29ce484222325
        for c in range(cycles):
            a = oracle.stim(SEED, s, c, 0) & 0xff
            h = oracle.fnv(bytes([a & 0x0f, a]), h)
        out[s] = h
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_FUSED])
def test_settle_loop_protocol(gpu, mode):
    cd = registry.get("circ")
    n, cycles = 130, 25
    want = _closed(n, cycles)
    with cd.model(n) as m:
        m.begin(SEED)
        m.run_cycles(cycles, mode)
        got = m.hashes()
    assert (got == want).all()


@pytest.mark.gpu
def test_settle_loop_eval_api_and_queue(gpu):
    """eval() runs the reference's host-driven settle loop (graph replays
    until no lane requests a change)."""
    cd = registry.get("circ")
    n, cycles = 64, 6
    with cd.model(n) as m:
        h = np.full(n, 0xcbf29ce484222325, dtype=np.uint64)
        for c in range(cycles):
            a = np.array([oracle.stim(SEED, s, c, 0) & 0xff for s in range(n)], np.uint8)
            m.set_port("a", a)
            m.set_port("clk", np.zeros(n, np.uint8))
            m.eval()
            m.set_port("clk", np.ones(n, np.uint8))
            m.eval()
            o, ox = m.get_port("o"), m.get_port("ox")
            assert (o == (a & 0x0f)).all() and (ox == a).all()
            for s in range(n):
                h[s] = oracle.fnv(bytes([int(o[s]), int(ox[s])]), int(h[s]))
    assert (h == _closed(n, cycles)).all()
    budgets = np.array([25, 0, 3, 17] * 10, np.uint32)
    with cd.model(16) as m:
        got = m.submit_queue(SEED, budgets)
    for i, b in enumerate(budgets):
        h = 0xcbf29ce484222325
        for c in range(int(b)):
            a = oracle.stim(SEED, i, c, 0) & 0xff
            h = oracle.fnv(bytes([a & 0x0f, a]), h)
        assert int(got[i]) == h, i


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_FUSED])
def test_unsettled_loop_reports(gpu, mode):
    """A loop that never settles: the run reports non-convergence at the
    next call that waits for the device, in every form."""
    cd = registry.get("osc")
    with cd.model(64, converge_limit=20) as m:
        m.begin(SEED)
        m.run_cycles(3, mode)
        with pytest.raises(RuntimeError, match="converge"):
            m.hashes()
    with cd.model(64, converge_limit=20) as m:
        m.set_port("clk", np.zeros(64, np.uint8))
        with pytest.raises(RuntimeError, match="converge"):
            m.eval()


def _gated_want(seed, n, cycles, budgets=None):
    from rtlflow_amd.designs import settle
    out = np.zeros(n, np.uint64)
    for s in range(n):
        h = 0xcbf29ce484222325
        cyc = cycles if budgets is None else int(budgets[s])
        for b in settle.gated_closed_form(oracle.stim, seed, s, cyc):
            h = oracle.fnv(b, h)
        out[s] = h
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_FUSED])
def test_gated_clock(gpu, mode):
    """A gated clock (clk & en): the counter it clocks advances exactly in
    the cycles whose enable is set, in the graph and fused forms and the
    batched queue."""
    cd = registry.get("gated")
    seed, n, cycles = 0x6A7E, 100, 40
    with cd.model(n) as m:
        m.begin(seed)
        m.run_cycles(cycles, mode)
        got = m.hashes()
    assert (got == _gated_want(seed, n, cycles)).all()
    budgets = np.array([40, 0, 5, 17, 33] * 6, np.uint32)
    with cd.model(8) as m:
        got = m.submit_queue(seed, budgets)
    assert (got == _gated_want(seed, len(budgets), 0, budgets)).all()


def _smear_want(seed, n, cycles, budgets=None):
    from rtlflow_amd.designs import settle
    out = np.zeros(n, np.uint64)
    for s in range(n):
        h = 0xcbf29ce484222325
        cyc = cycles if budgets is None else int(budgets[s])
        for b in settle.smear_closed_form(oracle.stim, seed, s, cyc):
            h = oracle.fnv(b, h)
        out[s] = h
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_FUSED])
def test_deep_settle_protocol(gpu, mode):
    """smear needs up to 17 settle passes per eval: the graph form's two
    captured passes are finished by its settle-tail kernel, with the default
    configuration, and every form matches the closed form."""
    cd = registry.get("smear")
    seed, n, cycles = 0x5EA1, 200, 40
    with cd.model(n) as m:
        m.begin(seed)
        m.run_cycles(cycles, mode)
        got = m.hashes()
    assert (got == _smear_want(seed, n, cycles)).all()


@pytest.mark.gpu
def test_deep_settle_eval_api_and_queue(gpu):
    from rtlflow_amd.designs import settle
    cd = registry.get("smear")
    seed, n, cycles = 0x5EA1, 64, 8
    r = np.zeros(n, np.uint32)
    with cd.model(n) as m:
        for c in range(cycles):
            a = np.array([oracle.stim(seed, s, c, 0) & 15 for s in range(n)], np.uint8)
            m.set_port("a", a)
            m.set_port("clk", np.zeros(n, np.uint8))
            m.eval()
            m.set_port("clk", np.ones(n, np.uint8))
            m.eval()
            x = (0xFFFF << a.astype(np.uint32)) & 0xFFFF
            r = (r + x) & 0xFFFF
            assert (m.get_port("ox") == x).all() and (m.get_port("o") == r).all()
    budgets = np.array([40, 0, 3, 17, 29] * 8, np.uint32)
    with cd.model(16) as m:
        got = m.submit_queue(seed, budgets)
    assert (got == _smear_want(seed, len(budgets), 0, budgets)).all()
    # a converge limit below the passes the design needs is reported by the
    # graph form's tail as by the other forms
    with cd.model(64, converge_limit=3) as m:
        m.begin(seed)
        m.run_cycles(4, MODE_GRAPH)
        with pytest.raises(RuntimeError, match="converge"):
            m.hashes()
