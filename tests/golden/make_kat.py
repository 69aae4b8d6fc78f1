"""Extract known-answer vectors from the reference's self-checking Verilog
tests (test_regress/t/t_math_*.v) into tests/golden/kat_math.json.

Run in a container that holds the reference checkout (its path is the first
argument, default /root/reference). Only data leaves the reference: operand
values, shift amounts and the expected results its ``$stop`` checks compare
against, each with the file and line it came from. The designs that check
them are this project's own (rtlflow_amd/designs/kat.py).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def _num(tok: str) -> int:
    """A Verilog number literal (sized or unsized, hex/decimal/binary)."""
    tok = tok.replace("_", "").replace(" ", "")
    m = re.fullmatch(r"(\d*)'s?([hHdDbB])([0-9a-fA-F]+)", tok)
    if m:
        base = {"h": 16, "d": 10, "b": 2}[m.group(2).lower()]
        return int(m.group(3), base)
    return int(tok)


def _blocks(text: str):
    """(cycle, first line, body lines) of each ``if (cyc==N) begin`` block."""
    lines = text.splitlines()
    out = []
    k = 0
    while k < len(lines):
        m = re.search(r"if \(cyc\s*==\s*(\d+)\) begin", lines[k])
        if m:
            depth, body, start = 0, [], k + 1
            for j in range(k, len(lines)):
                depth += lines[j].count("begin") - lines[j].count("end")
                if j > k:
                    body.append((j + 1, lines[j]))
                if depth <= 0 and j > k:
                    break
            out.append((int(m.group(1)), start, body))
        k += 1
    return out


def shift_vectors(ref: Path):
    """t_math_shift.v: shifts of three constants (32, 64 and 96 bits) by
    the amounts the testbench sets; the checks of cycle N see the amounts
    set in cycle N - 1."""
    f = ref / "test_regress/t/t_math_shift.v"
    text = f.read_text()
    consts = {}
    for w, name in ((32, "i"), (64, "q"), (96, "w")):
        m = re.search(rf"\b{name}right\s*=\s*(\d+'h[0-9a-fA-F_]+)\s*>>\s*iamt", text)
        consts[name] = (w, _num(m.group(1)))
    amts = {}
    vecs = []
    for cyc, line, body in _blocks(text):
        for ln, t in body:
            for var in ("iamt", "qamt", "wamt"):
                m = re.search(rf"\b{var}\s*<=\s*([^;]+);", t)
                if m:
                    v = m.group(1).strip()
                    cat = re.fullmatch(r"\{(.+)\}", v)
                    if cat:
                        val = 0
                        for part in cat.group(1).split(","):
                            pm = re.fullmatch(r"(\d+)'h([0-9a-fA-F_]+)", part.strip())
                            val = (val << int(pm.group(1))) | int(pm.group(2).replace("_", ""), 16)
                    else:
                        val = _num(v)
                    amts.setdefault(cyc, {})[var] = val
        prev = amts.get(cyc - 1)
        if not prev:
            continue
        for ln, t in body:
            m = re.search(r"if \((\w+)\s*!=\s*([^)]+)\) \$stop;", t)
            if not m or m.group(1)[1:] not in ("left", "right", "rights"):
                continue
            name, lit = m.group(1), m.group(2).strip()
            kind = name[0]
            w, val = consts[kind]
            if lit == "'0":
                exp = 0
            elif lit == "'1":
                exp = (1 << w) - 1
            else:
                exp = _num(lit)
            for var in ("iamt", "qamt", "wamt"):
                vecs.append({"src": f"test_regress/t/t_math_shift.v:{ln}", "op": name[1:],
                             "width": w, "value": val, "amount": prev[var],
                             "amount_width": {"iamt": 32, "qamt": 64, "wamt": 96}[var],
                             "expect": exp})
    return vecs


def div_vectors(ref: Path):
    """t_math_div.v: 61-bit unsigned and signed quotient and remainder of
    a[60:0] by divisor (and by 4), checked one cycle after the operands are
    set; bit 60 of each operand is forced by the testbench."""
    f = ref / "test_regress/t/t_math_div.v"
    text = f.read_text()
    ops = {}
    vecs = []
    a = d = None
    for cyc, line, body in _blocks(text):
        for ln, t in body:
            m = re.search(r"\ba <= (256'h[0-9a-fA-F_]+);", t)
            if m:
                a = _num(m.group(1)) & ((1 << 61) - 1)
            m = re.search(r"\bdivisor <= (\d+'[hd][0-9a-fA-F_]+);", t)
            if m:
                d = _num(m.group(1)) & ((1 << 61) - 1)
            m = re.search(r"a\[60\] <= 1'b([01]); divisor\[60\] <= 1'b([01]);", t)
            if m:
                a = (a & ~(1 << 60)) | (int(m.group(1)) << 60)
                d = (d & ~(1 << 60)) | (int(m.group(2)) << 60)
        ops[cyc] = (a, d)
        prev = ops.get(cyc - 1)
        if not prev or prev[0] is None:
            continue
        for ln, t in body:
            m = re.search(r"if \((qq|rq|qqs|rqs|qq4|rq4)\s*!==?\s*61'h([0-9a-fA-F_]+)\) \$stop;", t)
            if m:
                vecs.append({"src": f"test_regress/t/t_math_div.v:{ln}", "op": m.group(1), "width": 61,
                             "a": prev[0], "b": prev[1], "expect": int(m.group(2).replace("_", ""), 16)})
            # division by zero gives 0 in Verilator's two-state semantics
            m = re.search(r"if \((qq|rq) !== \{61\{1'b0\}\}\) \$stop;", t)
            if m:
                vecs.append({"src": f"test_regress/t/t_math_div.v:{ln}", "op": m.group(1), "width": 61,
                             "a": prev[0], "b": prev[1], "expect": 0})
    return vecs


def signed_vectors(ref: Path):
    """t_math_signed.v: a (16-bit signed) and b (5-bit signed) set by the
    case of cycle N, the concatenation {sr, srs, sl, sls, b_s, b_us}
    checked in cycle N + 1."""
    f = ref / "test_regress/t/t_math_signed.v"
    lines = f.read_text().splitlines()
    sets = {}
    vecs = []
    for k, t in enumerate(lines):
        m = re.match(r"\s*(\d+): begin", t)
        if m:
            cyc = int(m.group(1))
            for t2 in lines[k + 1:k + 4]:
                m2 = re.search(r"a <= (16\s*'sh\s*[0-9a-fA-F]+); b <= (5'sh[0-9a-fA-F]+);", t2)
                if m2:
                    sets[cyc] = (_num(m2.group(1)), _num(m2.group(2)))
        m = re.match(r"\s*(\d+): if \(\{sr,srs,sl,sls,b_s,b_us\}!==96'sh([0-9a-fA-F_]+)\) \$stop;", t)
        if m:
            cyc = int(m.group(1))
            if cyc - 1 in sets:
                a, b = sets[cyc - 1]
                vecs.append({"src": f"test_regress/t/t_math_signed.v:{k + 1}", "op": "signed_shifts",
                             "a": a, "b": b, "expect": int(m.group(2).replace("_", ""), 16)})
    return vecs


def red_vectors(ref: Path):
    """t_math_red.v: reductions of a 68-bit value, checked one cycle after
    it is set."""
    f = ref / "test_regress/t/t_math_red.v"
    text = f.read_text()
    vecs = []
    r = None
    for cyc, line, body in _blocks(text):
        checks = {}
        for ln, t in body:
            m = re.search(r"`checkh\((\w+), '([01])\);", t)
            if m:
                checks[m.group(1)] = (int(m.group(2)), ln)
            m = re.search(r"`checkh\(r, 68'h([0-9a-fA-F_]+)\);", t)
            if m:
                r = int(m.group(1).replace("_", ""), 16)
        for name, (exp, ln) in checks.items():
            vecs.append({"src": f"test_regress/t/t_math_red.v:{ln}", "op": name, "width": 68,
                         "value": r, "expect": exp})
    return vecs


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = {"shift": shift_vectors(ref), "div": div_vectors(ref), "signed": signed_vectors(ref),
           "red": red_vectors(ref)}
    (HERE / "kat_math.json").write_text(json.dumps(out, indent=1) + "\n")
    for k, v in out.items():
        print(k, len(v))


if __name__ == "__main__":
    main()
