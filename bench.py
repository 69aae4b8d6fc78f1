"""Throughput benchmark: stimulus-cycles/s on one MI355X.

Workload: the seeded random gate-level netlist (20,000 gates, 2,000 flip-flops,
64 primary inputs, 64 primary outputs, depth 40), 65,536 stimuli x 2,000
cycles, 64-bit uniform inputs per cycle generated on the device, per-cycle
FNV-1a-64 output hash per stimulus. A "step" is ``--cycles-per-step`` clock
cycles of all stimuli; ``--steps`` steps are timed (default 20 x 100 = the full
2,000 cycles). Three GPU forms are measured:

* bitplane: the design lowered to single-bit logic, 64 stimuli per 64-bit
  word, every node value in LDS, one kernel per chunk of cycles;
* fused: one per-lane kernel runs whole cycles (one stimulus per thread);
* graph: one hipGraph of the per-mtask kernels replayed per cycle (the plain
  form, one stimulus per thread).

``value`` is the fastest form whose hashes match the CPU oracle on the
checked sample; every form is reported beside it. Every timed run starts
from a fresh model. The CPU oracle (16 threads, one stimulus per thread,
byte-per-gate evaluation) is timed on a bounded sample as
the reported CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

HBM_PEAK_GBS = 8000.0
VALU_PEAK = 256 * 4 * 16 * 2.4e9     # int32 lane-ops/s: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(cd, lanes, mode, steps, cps, warmup, seed, chunk):
    if warmup:
        with cd.model(lanes, fused_chunk=chunk) as m:
            m.begin(seed)
            for _ in range(warmup):
                m.run_cycles(cps, mode)
            m.sync()
    with cd.model(lanes, fused_chunk=chunk) as m:
        m.begin(seed)
        m.sync()
        m.reset_stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.run_cycles(cps, mode)
        m.sync()
        dt = time.perf_counter() - t0
        st = m.stats()
        h = m.hashes()
    return dt, st, h


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cycles-per-step", type=int, default=100)
    ap.add_argument("--lanes", type=int, default=65536)
    ap.add_argument("--design", default="net2")
    ap.add_argument("--forms", default="bitplane,fused,graph")
    ap.add_argument("--chunk", type=int, default=512, help="cycles per fused / bit-plane launch")
    ap.add_argument("--check", type=int, default=64, help="stimuli checked against the oracle")
    ap.add_argument("--cpu-sample", type=int, default=64, help="stimuli in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="only run a plain device-to-device copy of MB megabytes three times "
                         "(the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "stimuli shard trivially across GPUs but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    from rtlflow_amd.designs import registry
    from rtlflow_amd.model import MODE_BITPLANE, MODE_FUSED, MODE_GRAPH
    import oracle

    seed = registry.SEED_CFG[2]
    cd = registry.get(args.design)
    man = cd.manifest
    lanes, cps, steps = args.lanes, args.cycles_per_step, args.steps
    cycles = cps * steps
    stim_cycles = lanes * cycles
    forms = {}
    bp_ok = bool(man.get("bitplane", {}).get("supported"))
    for form in args.forms.split(","):
        if form == "bitplane" and not bp_ok:
            log(f"[bench] {args.design}: no bit-plane program ({man.get('bitplane')}); skipped")
            continue
        mode = {"fused": MODE_FUSED, "graph": MODE_GRAPH, "bitplane": MODE_BITPLANE}[form]
        log(f"[bench] {args.design} form={form} lanes={lanes} cycles={cycles}")
        dt, st, h = run_form(cd, lanes, mode, steps, cps, args.warmup, seed, args.chunk)
        nchk = min(args.check, lanes)
        ref = oracle.netlist_run(registry.table_bytes(args.design), seed, nstim=nchk, cycles=cycles,
                                 threads=16)
        ok = bool((h[:nchk] == ref).all())
        log(f"[bench] form={form} {dt:.3f}s {stim_cycles / dt:.4g} stim-cycles/s oracle_match={ok}")
        forms[form] = {"seconds": dt, "value": stim_cycles / dt, "bit_exact_sample": ok,
                       "checked_stimuli": nchk, "stats": st}

    # the reported form: the fastest one whose sample is bit-exact
    exact = {k: v for k, v in forms.items() if v["bit_exact_sample"]}
    main_form = max(exact or forms, key=lambda k: forms[k]["value"])
    f = forms[main_form]
    st = f["stats"]
    # algorithmic bytes per stimulus-cycle: 2 * S_seq + T_mem + B_in + B_out in
    # the form's own state layout (inputs are generated and outputs hashed on
    # the device, so B_in = B_out = 0 in HBM)
    if main_form == "bitplane":
        state_bits = man["bitplane"]["state_bits"]
        b_alg = 2 * ((state_bits + 7) // 8)
        per_launch_ms = st["bitplane_ms"] / max(1, st["bitplane_launches"])
        cyc_per_launch = min(cps, args.chunk)
        bytes_per_launch = b_alg * lanes * cyc_per_launch
        kernel = "rf_bp_kernel"
    elif main_form == "fused":
        b_alg = 2 * man["bytes_seq"]
        per_launch_ms = st["fused_ms"] / max(1, st["fused_launches"])
        bytes_per_launch = b_alg * lanes * min(cps, args.chunk)
        kernel = "rf_k_fused"
    else:
        b_alg = 2 * man["bytes_seq"]
        per_launch_ms = st["graph_ms"] / max(1, st["graph_launches"])
        bytes_per_launch = b_alg * lanes
        kernel = "cycle graph"
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    ops = man["ops_per_cycle"]
    valu = stim_cycles * ops / forms[main_form]["seconds"]

    cpu = None
    if not args.no_cpu:
        ns = args.cpu_sample
        cpu_cycles = cycles
        t0 = time.perf_counter()
        oracle.netlist_run(registry.table_bytes(args.design), seed, nstim=ns, cycles=cpu_cycles,
                           threads=16)
        cdt = time.perf_counter() - t0
        cpu = {"value": ns * cpu_cycles / cdt, "unit": "stimulus-cycles/s", "cores": 16,
               "kind": "port",
               "sample": f"{ns} stimuli x {cpu_cycles} cycles of the same netlist and protocol "
                         f"({cdt:.1f} s), C oracle, one stimulus per thread"}

    line = {
        "metric": "stimulus-cycles/sec, bit-exact vs Verilator CPU; batched+learned vs plain GPU",
        "value": f["value"],
        "unit": "stimulus-cycles/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": f["seconds"] / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 bit planes" if main_form == "bitplane" else "u8/u32 integer",
        "data": "synthetic (seeded netlist, on-device SplitMix64 stimulus)",
        "config": {"workload": f"random netlist 20k gates / 2k flops / 64 PI / 64 PO / depth 40, "
                               f"{lanes} stimuli x {cycles} cycles",
                   "design": args.design, "stimuli": lanes, "cycles": cycles,
                   "cycles_per_step": cps, "form": main_form, "mtasks": man["mtasks"]},
        "forms": {k: {kk: vv for kk, vv in v.items() if kk != "stats"} for k, v in forms.items()},
        "bit_exact_sample": all(v["bit_exact_sample"] for v in forms.values()),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel,
                     "alg_bytes_per_stimulus_cycle": b_alg, "avg_launch_ms": per_launch_ms},
        "valu": {"scalar_equivalent_ops": valu, "peak_ops": VALU_PEAK, "frac": valu / VALU_PEAK,
                 "ops_per_stimulus_cycle": ops},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
