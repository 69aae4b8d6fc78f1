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
    # configuration 5 at full size, the gate-net engine's per-mtask graph form
    if "learned" in registry.gatenet_meta("net5")["schedules"]:
        from rtlflow_amd.gatenet import GateNet
        seed5 = registry.SEED_CFG[5]
        b5 = np.minimum(rv32i.budgets(seed5, 4096, median=2000, lo=200, hi=20000), 8)
        tb5 = registry.gatenet_table("net5")

        class _Sched:
            def __init__(self, tag):
                self.tag = tag
                self.manifest = {"mtasks": registry.gatenet_meta("net5")["schedules"][tag]["mtasks"],
                                 "partition": tag}

        def gn_measure(s):
            import time
            with GateNet(registry.gatenet_program("net5", s.tag), 4096) as g:
                g.run(GateNet.GRAPH, seed5, np.minimum(b5, 1))
                t0 = time.perf_counter()
                h, _ = g.run(GateNet.GRAPH, seed5, b5)
                dt = time.perf_counter() - t0
            return float(b5.astype(np.int64).sum()) / dt, h

        v = judge.judge(_Sched("heuristic"), _Sched("learned"), gn_measure,
                        lambda smp: oracle.netlist_run(tb5, seed5, len(smp), budgets=b5[:len(smp)]), sample)
        judge.record("net5", "graph", v)
        print(f"net5 graph: learned/heuristic {v['speedup']:.3f} bit-exact {v['bit_exact']} -> {v['choice']}",
              flush=True)
    for name, form, measure, checker in jobs:
        for key, suffix in judge.VARIANTS:
            if f"{name}{suffix}" not in registry.load_index():
                continue
            v = judge.judge(registry.get(name), registry.get(f"{name}{suffix}"), measure, checker, sample)
            judge.record(name, form + key, v)
            print(f"{name} {form + key}: {suffix[1:]}/heuristic {v['speedup']:.3f} bit-exact "
                  f"{v['bit_exact']} -> {v['choice']}", flush=True)


if __name__ == "__main__":
    main()
