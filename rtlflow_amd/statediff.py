"""Lane state decoding and diffing (the role of the reference's emitted
``checkN`` differential asserter, src/V3EmitC.cpp:4380-4536, which compares
lane 0 of the GPU model with a CPU Verilator model variable by variable).

``rf_lane_state`` (GPU runtime) and the host build return one lane's raw
slots in class order C, S, I, Q followed by the lane-major memory region;
``decode_state`` maps them back to variable names with the design manifest,
and ``diff_states`` lists the variables that differ -- including wide values
and memories, which the reference's asserter skips (:4443-4445).
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

CLS_BYTES = [1, 2, 4, 8]
DT = [np.uint8, np.uint16, np.uint32, np.uint64]


def state_bytes(man: dict) -> int:
    s = man["slots"]
    return s["C"] + 2 * s["S"] + 4 * s["I"] + 8 * s["Q"] + man["lane_major_bytes"]


def decode_state(man: dict, raw: bytes) -> Dict[str, object]:
    s = man["slots"]
    off = [0, s["C"], s["C"] + 2 * s["S"], s["C"] + 2 * s["S"] + 4 * s["I"]]
    n = [s["C"], s["S"], s["I"], s["Q"]]
    arrays = [np.frombuffer(raw, DT[c], n[c], off[c]) for c in range(4)]
    lm0 = off[3] + 8 * s["Q"]
    out: Dict[str, object] = {}
    for v in man["vars"]:
        a = arrays[v["cls"]]
        words = [int(x) for x in a[v["loc"]:v["loc"] + v["words"]]]
        if v["words"] == 1:
            out[v["name"]] = words[0]
        else:
            out[v["name"]] = sum(w << (32 * i) for i, w in enumerate(words))
    for m in man.get("mems", []):
        c = m["cls"]
        cnt = m["depth"] * m["words"]
        if m["lane_major"]:
            vals = np.frombuffer(raw, DT[c], cnt, lm0 + m["loc"])
        else:
            vals = arrays[c][m["loc"]:m["loc"] + cnt]
        out[m["name"]] = np.array(vals).reshape(m["depth"], m["words"])
    return out


def diff_states(a: Dict[str, object], b: Dict[str, object]) -> List[Tuple[str, object, object]]:
    out = []
    for k in sorted(set(a) | set(b)):
        x, y = a.get(k), b.get(k)
        if isinstance(x, np.ndarray) or isinstance(y, np.ndarray):
            if x is None or y is None or not np.array_equal(x, y):
                idx = None if x is None or y is None else np.argwhere(np.asarray(x) != np.asarray(y))[:4]
                out.append((k, "memory", idx))
        elif x != y:
            out.append((k, x, y))
    return out
