"""Judge the learned partitions on the GPU and commit the verdicts
(``python tests/judge_schedules.py``): for each design and form, the learned
partition is kept only when it is faster than the heuristic one and both are
bit-exact against the CPU oracle on a sample. Writes
rtlflow_amd/costtables/judge.json (read by rtlflow_amd.judge.scheduled)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import oracle  # noqa: E402
from rtlflow_amd import judge  # noqa: E402
from rtlflow_amd.designs import registry, rv32i  # noqa: E402


def main():
    sample = np.arange(16)
    jobs = []
    for name, lanes, cycles in (("net2", 65536, 64), ("net5s", 4096, 100)):
        seed = registry.SEED_CFG[2 if name == "net2" else 5]
        tb = registry.table_bytes(name)
        jobs.append((name, "graph", lambda cd, l=lanes, c=cycles, s=seed: judge.measure_graph(cd, l, c, s),
                     lambda smp, t=tb, s=seed, c=cycles: oracle.netlist_run(t, s, len(smp), cycles=c)))
    seed = registry.SEED_CFG[5]
    budgets = rv32i.budgets(seed, 2048, median=150, lo=20, hi=600)
    tb = registry.table_bytes("net5s")
    jobs.append(("net5s", "batched", lambda cd: judge.measure_batched(cd, 1024, budgets, seed),
                 lambda smp: oracle.netlist_run(tb, seed, len(smp), budgets=budgets[:len(smp)])))
    seed = registry.SEED_CFG[4]
    jobs.append(("systolic", "graph", lambda cd: judge.measure_graph(cd, 8192, 100, seed),
                 lambda smp: oracle.systolic_run(seed, len(smp), 100)))
    seed = registry.SEED_CFG[3]
    progs = rv32i.programs(seed, np.arange(4096))
    rb = rv32i.budgets(seed, 4096, median=1000, lo=200, hi=5000)

    def rv_measure(cd):
        import time
        with cd.model(1024) as m:
            m.set_mem_images("imem", progs)
            m.submit_queue(seed, rb[:64])
            m.sync()
            t0 = time.perf_counter()
            h = m.submit_queue(seed, rb)
            dt = time.perf_counter() - t0
        return float(rb.astype(np.int64).sum()) / dt, h

    jobs.append(("rv32i", "batched", rv_measure,
                 lambda smp: oracle.rv32i_run(progs[:len(smp)], seed, rb[:len(smp)])))
    for name, form, measure, checker in jobs:
        if f"{name}_learned" not in registry.load_index():
            continue
        v = judge.judge(registry.get(name), registry.get(f"{name}_learned"), measure, checker, sample)
        judge.record(name, form, v)
        print(f"{name} {form}: learned/heuristic {v['speedup']:.3f} bit-exact {v['bit_exact']} "
              f"-> {v['choice']}", flush=True)


if __name__ == "__main__":
    main()
