"""Interpreter-side lane state for the lane-k differential check."""
import numpy as np

import ir_interp


def interp_lane_state(design, seed: int, s: int, cycles: int) -> dict:
    """Stored state of stimulus ``s`` after ``cycles`` protocol cycles,
    computed by the IR interpreter: {name: int | array[depth, words]}."""
    data_inputs = [v for v in design.vars if v.kind == "input" and v not in design.clocks]
    lane = ir_interp.Lane(design)
    for c in range(cycles):
        for p, var in enumerate(data_inputs):
            val = 0
            for j in range((var.width + 63) // 64):
                val |= ir_interp._stim(seed, s, c, (p << 3) | j) << (64 * j)
            lane.v[var] = val & ir_interp._mask(var.width)
        for ph in (0, 1):
            for clk in design.clocks:
                lane.v[clk] = ph
            lane.eval()
    out = {v.name: x for v, x in lane.v.items()}
    for m, vals in lane.m.items():
        words = (m.width + 31) // 32 if m.width > 64 else 1
        if m.width > 64:
            arr = np.array([[(x >> (32 * i)) & 0xffffffff for i in range(words)] for x in vals],
                           dtype=np.uint64)
        else:
            arr = np.array([[x] for x in vals], dtype=np.uint64)
        out[m.name] = arr
    return out


def compare(got: dict, want: dict, names) -> list:
    bad = []
    for n in names:
        g, w = got[n], want[n]
        if isinstance(g, np.ndarray):
            if not np.array_equal(np.asarray(g, np.uint64), np.asarray(w, np.uint64)):
                bad.append((n, "memory"))
        elif int(g) != int(w):
            bad.append((n, hex(int(g)), hex(int(w))))
    return bad
