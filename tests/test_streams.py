"""User input streams and output capture on the CPU: the one-lane host build
of the generated kernels (identical bodies to the device build) against the
IR interpreter given the same [cycle][word][stimulus] input words, in every
schedule form and the batched queue."""
import numpy as np
import pytest

import ir_interp
from rtlflow_amd.compiler import Compiled, Options
from rtlflow_amd.designs.mix import VERILOG as MIX
from rtlflow_amd.designs.netlist import generate
from rtlflow_amd.hostsim import HostDesign
from rtlflow_amd.verilog import elaborate


def _words(rng, ncyc, words, nstim):
    return rng.integers(0, 2**64, size=(ncyc, words, nstim), dtype=np.uint64, endpoint=False)


@pytest.fixture(scope="module")
def mix():
    d = elaborate(MIX)
    comp = Compiled(d, Options(threads=2))
    return d, comp, HostDesign(comp)


def test_stream_table_layout(mix):
    _, comp, _ = mix
    st = comp.stream_table()
    assert st["in"] == [["a", 0], ["w", 0], ["w", 1], ["b", 0]]
    assert st["out"] == [["r", 0], ["r", 1], ["s", 0], ["z", 0]]
    assert comp.manifest()["streams"]["in_words"] == 4


@pytest.mark.parametrize("mode", [0, 1, 3])
def test_stream_and_capture_match_interpreter(mix, mode):
    d, comp, hd = mix
    rng = np.random.default_rng(11 + mode)
    nstim, ncyc, stim0 = 5, 23, 7
    x = _words(rng, ncyc, 4, nstim)
    got = np.zeros((ncyc, 4, nstim), np.uint64)
    want = np.zeros_like(got)
    hd.set_streams(inputs=x, capture=got, in_stim0=stim0, cap_stim0=stim0)
    try:
        h = hd.protocol(nstim, 0, ncyc, mode=mode, stim0=stim0)
    finally:
        hd.set_streams()
    ref = ir_interp.protocol(d, 0, nstim, ncyc, stim0=stim0, inputs=x, capture=want)
    assert [int(v) for v in h] == ref
    assert (got == want).all()
    # the capture holds the outputs the hash folded: r's top word is 36 bits
    assert (got[:, 1, :] >> np.uint64(36) == 0).all()


def test_stream_of_generator_words_gives_generator_hashes(mix):
    """Words equal to the generator's r(seed, s, c, (p << 3) | j), in stream
    order, reproduce the generator run: pins the word order and keys."""
    d, comp, hd = mix
    seed, nstim, ncyc = 0x52544C4600000001, 4, 12
    keys = [(0 << 3) | 0, (1 << 3) | 0, (1 << 3) | 1, (2 << 3) | 0]
    x = np.zeros((ncyc, 4, nstim), np.uint64)
    for c in range(ncyc):
        for n, k in enumerate(keys):
            for s in range(nstim):
                x[c, n, s] = ir_interp._stim(seed, s, c, k)
    base = hd.protocol(nstim, seed, ncyc, mode=1)
    hd.set_streams(inputs=x)
    try:
        streamed = hd.protocol(nstim, 0, ncyc, mode=1)
    finally:
        hd.set_streams()
    assert (base == streamed).all()


def test_partial_capture_window(mix):
    _, comp, hd = mix
    rng = np.random.default_rng(3)
    nstim, ncyc = 6, 10
    x = _words(rng, ncyc, 4, nstim)
    full = np.zeros((ncyc, 4, nstim), np.uint64)
    part = np.zeros((4, 4, 3), np.uint64)
    hd.set_streams(inputs=x, capture=full)
    try:
        hd.protocol(nstim, 0, ncyc, mode=1)
        hd.set_streams(inputs=x, capture=part, cap_stim0=2, cap_cycle0=3)
        hd.protocol(nstim, 0, ncyc, mode=0)
    finally:
        hd.set_streams()
    assert (part == full[3:7, :, 2:5]).all()


def test_stream_too_short_is_an_error(mix):
    _, comp, hd = mix
    x = np.zeros((5, 4, 3), np.uint64)
    hd.set_streams(inputs=x)
    try:
        with pytest.raises(RuntimeError):
            hd.protocol(3, 0, 6, mode=1)      # cycle 5 is not in the stream
        with pytest.raises(RuntimeError):
            hd.protocol(4, 0, 5, mode=0)      # stimulus 3 is not in the stream
        hd.protocol(3, 0, 5, mode=0)          # covered exactly
    finally:
        hd.set_streams()


def test_queue_with_stream_matches_interpreter(mix):
    d, comp, hd = mix
    rng = np.random.default_rng(5)
    budgets = [9, 0, 4, 9, 1, 7]
    nstim, ncyc = len(budgets), max(budgets)
    x = _words(rng, ncyc, 4, nstim)
    got = np.zeros((ncyc, 4, nstim), np.uint64)
    want = np.zeros_like(got)
    hd.set_streams(inputs=x, capture=got)
    try:
        h = hd.queue(2, 0, budgets)
    finally:
        hd.set_streams()
    ref = ir_interp.protocol(d, 0, nstim, ncyc, budgets=budgets, inputs=x, capture=want)
    assert [int(v) for v in h] == ref
    assert (got == want).all()


def test_netlist_with_rams_stream():
    """A random netlist with synchronous RAMs and a 70-bit input, calls form
    and inline form."""
    src, _ = generate(n_gates=300, n_ff=20, n_pi=70, n_po=20, depth=7, seed=41, n_rams=2,
                      ram_depth=8, ram_width=5)
    d = elaborate(src)
    rng = np.random.default_rng(9)
    nstim, ncyc = 4, 15
    for opts in (Options(threads=3), Options(threads=3, inline_limit=1)):
        comp = Compiled(d, opts)
        st = comp.stream_table()
        hd = HostDesign(comp)
        x = _words(rng, ncyc, st["in_words"], nstim)
        got = np.zeros((ncyc, st["out_words"], nstim), np.uint64)
        want = np.zeros_like(got)
        hd.set_streams(inputs=x, capture=got)
        try:
            h = hd.protocol(nstim, 0, ncyc, mode=1)
        finally:
            hd.set_streams()
        ref = ir_interp.protocol(d, 0, nstim, ncyc, inputs=x, capture=want)
        assert [int(v) for v in h] == ref
        assert (got == want).all()


def test_bench_generator_words_match_the_generator():
    """The benchmark builds its stream of generator words with int64 torch
    arithmetic; on the CPU it must give r(seed, s, c, key) exactly."""
    import torch
    import bench
    man = {"ports": [{"dir": 0, "stim_index": 0, "width": 1}, {"dir": 2, "stim_index": 0xffffffff, "width": 1},
                     {"dir": 0, "stim_index": 1, "width": 70}]}
    seed = 0x52544C4600000002
    for c in (0, 3, 1999):
        w = bench.generator_words_dev(torch, man, seed, 9, c, device="cpu").numpy().view(np.uint64)
        for n, k in enumerate([0, 8, 9]):
            for s in range(9):
                assert int(w[n, s]) == ir_interp._stim(seed, s, c, k)
