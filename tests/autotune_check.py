"""Run the autotuner on the GPU with the CPU oracle as the checker of every
candidate's hashes (``python tests/autotune_check.py NAME...``; default:
every tuned design). Writes rtlflow_amd/tuned/<name>.json."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import oracle  # noqa: E402
from rtlflow_amd import autotune  # noqa: E402
from rtlflow_amd.designs import registry  # noqa: E402


def oracle_hashes(name, wl, idx):
    """The oracle's hashes of stimuli ``idx`` of a tuning workload."""
    seed, progs, budgets = autotune._inputs(name, wl)
    idx = np.asarray(idx)
    out = np.zeros(len(idx), np.uint64)
    # contiguous runs of stimulus ids, one oracle call each
    k = 0
    while k < len(idx):
        j = k
        while j + 1 < len(idx) and idx[j + 1] == idx[j] + 1:
            j += 1
        s0, cnt = int(idx[k]), j - k + 1
        if name in ("net2", "net5s"):
            r = oracle.netlist_run(registry.table_bytes(name), seed, cnt, cycles=wl["cycles"], stim0=s0)
        elif name == "alu":
            r = oracle.alu_run(seed, cnt, wl["cycles"], stim0=s0)
        elif name == "systolic":
            r = oracle.systolic_run(seed, cnt, wl["cycles"], stim0=s0)
        elif name == "rv32i":
            r = oracle.rv32i_run(progs[s0:s0 + cnt], seed, budgets[s0:s0 + cnt], stim0=s0)
        else:
            raise KeyError(f"no oracle for {name}")
        out[k:j + 1] = r
        k = j + 1
    return out


def main(argv):
    names = argv or list(autotune.WORKLOADS)
    for n in names:
        r = autotune.tune(n, oracle_hashes, log=lambda *a: print(*a, flush=True))
        b = r["best"]
        print(f"best {n}: {b['library']} {b['form']} {b['knobs']} {b['rate']:.4g} stimulus-cycles/s",
              flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
