#!/usr/bin/env python3
"""Vacancy-hop AKMC benchmark (one GPU).

Workload (plain form): FP64 AKMC on a BCC box of 256^3 unit cells (33.5M
sites, one uint8 per site), Fe-1.34Cu-1.30Mn-0.70Ni at% placed from the seed,
6,711 vacancies (2e-4), T = 573 K, attempt frequency 6e12 /s, the 8-sector
sublattice scheme with a 1e-6 s threshold per sector cycle.  A "step" is one
akmc_run of --events-per-step events; the initial occupancy (host) and the
upload are outside the timed region, so the lattice is resident in HBM when
timing starts.  value = events performed in the timed steps / seconds.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--events-per-step E]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "misa-akmc-gpu_amd"))

METRIC = "vacancy-hop events/s, plain FP64 vs batched FP32-rate; p50/p99 job latency"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(box=(256, 256, 256), alloy=[("Fe", 9666), ("Cu", 134), ("Mn", 130), ("Ni", 70)], va_count=6711,
                seeds=dict(create_types=100565, create_vacancy=16650, event_selection=563264, time_inc=594689),
                temperature=573.0, attempt_freq=6e12, sector_T=1e-6, physics_time=1.0)


def algorithmic_bytes(d_events, d_eval, d_fold, d_shift):
    """Per-event algorithmic traffic of the plain path:
    B = n_eval*(55 + 8*8) + 8*K + 2 + 16 per event, with
    55    uint8 sites of one vacancy's rate stencil (gathered per evaluation),
    8*8   its eight FP64 rates written,
    K     FP64 running-sum words touched: 9 per folded vacancy (8 rates read,
          1 sum written), 2 per shifted vacancy (1 read, 1 written),
    2     the swap's two site bytes, 16 the event's two random numbers."""
    K = 9 * d_fold + 2 * d_shift
    return d_eval * (55 + 64) + 8 * K + (2 + 16) * d_events


def cpu_baseline(events, seconds_budget):
    """The oracle (the reference CPU algorithm restated: hash containers,
    full-sector sweeps) on the same workload, one thread, first `events`
    events (initial occupancy excluded from the timing)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    ref = refcpu.RefCpu(WORKLOAD)
    t0 = time.perf_counter()
    done = 0
    while done < events and time.perf_counter() - t0 < seconds_budget:
        ref.run(10)
        done = ref.state()["events"]
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "events/s", "cores": 1, "kind": "port",
            "sample": "first %d events of the same 256^3 FP64 workload, oracle/refcpu.cpp "
                      "(reference CPU algorithm: hash occupancy, full-sector sweep per event), "
                      "1 thread, initial occupancy excluded, %.1f s" % (done, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events-per-step", type=int, default=1_000_000)
    ap.add_argument("--cpu-events", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the streaming copy kernel (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (one lattice per GPU; replica and job "
              "sharding across GPUs is described in DESIGN.md); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        return 2

    import akmc
    print("bench: creating the 256^3 lattice", file=sys.stderr, flush=True)
    eng = akmc.Engine(WORKLOAD, device=0)
    E = args.events_per_step
    for _ in range(args.warmup):
        eng.run(E)
    akmc.sync(0)
    s0 = eng.stats()
    t0 = time.perf_counter()
    ksec = 0.0
    launches = 0
    for i in range(args.steps):
        st = eng.run(E)
        ksec += st["kernel_seconds"]
        launches += st["launches"]
        print("bench: step %d/%d events %d" % (i + 1, args.steps, st["events"]), file=sys.stderr, flush=True)
    akmc.sync(0)
    dt = time.perf_counter() - t0
    s1 = eng.stats()
    d_events = s1["events"] - s0["events"]
    value = d_events / dt
    nbytes = algorithmic_bytes(d_events, s1["n_eval"] - s0["n_eval"], s1["n_fold"] - s0["n_fold"],
                               s1["n_shift"] - s0["n_shift"])
    per_launch = nbytes / max(1, launches)
    avg_launch_s = ksec / max(1, launches)
    achieved = per_launch / avg_launch_s / 1e9
    if args.copy_calibration:
        akmc.copy_bench(1 << 30, iters=5)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_events, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain FP64 AKMC: BCC 256^3 cells (33,554,432 sites), Fe-1.34Cu-1.30Mn-0.70Ni at%, "
                               "6711 vacancies, T=573 K, nu=6e12/s, 8-sector sublattice, T_sector=1e-6 s",
                   "events_per_step": E, "events_timed": d_events, "launches": launches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_event": nbytes / max(1, d_events),
                     "kernel_us_per_event": ksec / max(1, d_events) * 1e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
