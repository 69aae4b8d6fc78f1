"""Value change dump of one lane (IEEE 1364-2005 clause 18 four-state VCD,
written two-state).

The reference traces a whole model through ``verilated_vcd_c`` when built
with ``--trace`` (src/V3EmitC.cpp:3927-3961); on the GPU one lane of a batch
is traced instead: ``Model.trace`` (``rf_trace_lane``) copies the lane's
state after every clock cycle on the device, and ``write_vcd`` turns those
records into a VCD of every stored variable (ports, registers, values that
cross mtasks) with the design manifest's names and widths. Time advances 10
units per cycle; ``clk`` is drawn low at the middle of each cycle so the
waveform shows the edges the protocol applies.
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Sequence

from .statediff import decode_state


def _ident(k: int) -> str:
    """Short identifier codes from the printable range 33..126."""
    out = ""
    k += 1
    while k:
        k -= 1
        out += chr(33 + k % 94)
        k //= 94
    return out


def _val(v: int, width: int, ident: str) -> str:
    if width == 1:
        return f"{v & 1}{ident}"
    return f"b{v:b} {ident}"


def write_vcd(manifest: dict, frames: Sequence[bytes], lane: int = 0, stim: int = 0,
              timescale: str = "1ns") -> str:
    """The VCD text of per-cycle lane states ``frames`` (rf_lane_state
    records, one per cycle)."""
    top = manifest["name"]
    vars_ = [v for v in manifest["vars"] if not v["name"].startswith("__V")]
    idents = {v["name"]: _ident(k) for k, v in enumerate(vars_)}
    clocks = [p["name"] for p in manifest["ports"] if p["dir"] == 2]
    L: List[str] = [
        "$date rtlflow_amd $end",
        f"$version rtlflow_amd lane {lane} (stimulus {stim}) $end",
        f"$timescale {timescale} $end",
        f"$scope module {top} $end",
    ]
    for v in vars_:
        name = v["name"]
        ref = name if v["width"] == 1 else f"{name} [{v['width'] - 1}:0]"
        L.append(f"$var wire {v['width']} {idents[name]} {ref} $end")
    L += ["$upscope $end", "$enddefinitions $end"]
    prev: Dict[str, int] = {}
    for c, raw in enumerate(frames):
        vals = decode_state(manifest, raw)
        if c:
            # the falling clock in the middle of the cycle
            L.append(f"#{10 * c + 5}")
            for clk in clocks:
                if clk in idents:
                    L.append(_val(0, 1, idents[clk]))
                    prev[clk] = 0
        L.append(f"#{10 * (c + 1)}")
        if c == 0:
            L.append("$dumpvars")
        for v in vars_:
            x = int(vals[v["name"]]) & ((1 << v["width"]) - 1)
            if c == 0 or prev.get(v["name"]) != x:
                L.append(_val(x, v["width"], idents[v["name"]]))
                prev[v["name"]] = x
        if c == 0:
            L.append("$end")
    return "\n".join(L) + "\n"


def read_vcd(text: str) -> Dict[str, List[tuple]]:
    """Minimal reader (tests): variable name -> [(time, value)] changes."""
    names: Dict[str, str] = {}
    out: Dict[str, List[tuple]] = {}
    t = 0
    for line in text.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "$var":
            names[w[3]] = w[4]
            out[w[4]] = []
        elif w[0].startswith("#"):
            t = int(w[0][1:])
        elif w[0].startswith("b") and len(w) == 2:
            out[names[w[1]]].append((t, int(w[0][1:], 2)))
        elif w[0][0] in "01" and w[0][1:] in names:
            out[names[w[0][1:]]].append((t, int(w[0][0])))
    return out


def values_at_cycles(changes: Dict[str, List[tuple]], name: str, ncycles: int) -> List[int]:
    """A variable's value at the end of each cycle (times 10, 20, ...)."""
    ch = changes[name]
    out = []
    k = 0
    cur = 0
    for c in range(ncycles):
        t = 10 * (c + 1)
        while k < len(ch) and ch[k][0] <= t:
            cur = ch[k][1]
            k += 1
        out.append(cur)
    return out
