"""Struct-of-arrays state layout for the generated kernels.

The reference assigns every non-local variable a per-class slot ("memLoc") and
stores scalars at ``X[lane + THREADS*memLoc]`` but wide values and arrays
lane-contiguous (``X + lane*words + THREADS*memLoc``), which makes their
accesses strided (src/V3EmitC.cpp:1209-1233, allocator :4538-5045).

Here every slot is lane-minor: word ``w`` of variable ``v`` for lane ``l``
lives at ``class_array[(loc(v) + w) * npad + l]`` with ``npad`` the lane count
padded to a multiple of the 64-lane wavefront, so any access by a wavefront to
a scalar or to one word of a wide value is a single coalesced transaction.

Memories (unpacked arrays) are either element-major in the class arrays
(``[(loc + e*words + w) * npad + l]``: coalesced when lanes index the same
element) or lane-major in a separate byte region
(``mem + off*npad + l*depth*esize + e*esize``: each lane's memory is
contiguous, which keeps per-lane locality for divergent addresses such as
instruction fetch). The choice is per memory.

Class ordering inside each array follows the reference's rule of top-level
variables first, then by name (src/V3EmitC.cpp:4618-4628), simplified to
declaration order within roles: ports, state, then everything else.
"""
from __future__ import annotations

from typing import Dict, List, Optional

from .ir import CLASS_BYTES, Var

CLASSES = ("C", "S", "I", "Q")
WAVE = 64


def class_of(v: Var) -> str:
    c = v.cls
    return "I" if c == "W" else c


class Layout:
    def __init__(self):
        self.loc: Dict[Var, int] = {}
        self.slots: Dict[str, int] = {c: 0 for c in CLASSES}
        self.lane_major: Dict[Var, int] = {}     # memory -> byte offset (per padded lane count)
        self.lane_major_bytes = 0                # bytes per lane of the lane-major region

    def add(self, v: Var, lane_major: bool = False):
        if v in self.loc or v in self.lane_major:
            return
        if v.is_mem and lane_major:
            esize = CLASS_BYTES[v.cls] if v.cls != "W" else 4 * v.words
            self.lane_major[v] = self.lane_major_bytes
            self.lane_major_bytes += v.depth * esize
            # keep 8-byte alignment of the next region per lane
            self.lane_major_bytes = (self.lane_major_bytes + 7) & ~7
            return
        c = class_of(v)
        self.loc[v] = self.slots[c]
        self.slots[c] += v.slots

    def has(self, v: Var) -> bool:
        return v in self.loc or v in self.lane_major

    def bytes_per_lane(self) -> int:
        return sum(self.slots[c] * {"C": 1, "S": 2, "I": 4, "Q": 8}[c] for c in CLASSES) + \
            self.lane_major_bytes


def pad_lanes(n: int, block: int = 256) -> int:
    q = max(block, WAVE)
    return ((n + q - 1) // q) * q
