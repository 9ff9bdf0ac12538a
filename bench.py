#!/usr/bin/env python3
"""Vacancy-hop AKMC benchmark (one GPU).

Workload (plain form): FP64 AKMC on a BCC box of 256^3 unit cells (33.5M
sites, one uint8 per site), Fe-1.34Cu-1.30Mn-0.70Ni at% placed from the seed,
6,711 vacancies (2e-4), T = 573 K, attempt frequency 6e12 /s, the 8-sector
sublattice scheme with a 1e-6 s threshold per sector cycle.  A "step" is one
akmc_run of --events-per-step events; the initial occupancy (host) and the
upload are outside the timed region, so the lattice is resident in HBM when
timing starts.  value = events performed in the timed steps / seconds.

Beside the plain line, the same JSON line carries three more legs of the same
metric (disable with --no-batched / --no-served / --plain-engines 0) and the
dumbbell-rate step (--no-itl):
  batched  the replica engine on 1,024 independent replicas (edge U{16..96},
           random Fe-Cu-Ni-Mn compositions, T U[500,700] K, 2e-4 vacancies),
           FP32 rates with FP64 sums, autotuned, hipGraph replays; a "step"
           advances every replica by --batch-events events.
  served   the serving loop on a Poisson stream of --jobs jobs (edge U{32..320},
           vacancies U[1e-4,5e-4], targets log-uniform in [1e5,1e7] events scaled
           by --serve-scale; a tenth of them also carry dumbbells at 1e-5),
           admitted by a 250 GB HBM budget; reports throughput and p50/p99
           completion latency, overall and for the dumbbell jobs.
  plain_engines  four plain FP64 engines at once (the 256^3 workload, seeds
           differing), one host thread and stream each: aggregate events/s of
           bit-exact trajectories.
  itl      the abvi dumbbell model's rate step (akmc_itl_rates) on a 256^3
           typed lattice with 1e-5 dumbbells: rates per second of kernel time.

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
# HBM traffic measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on the
# same workloads (FETCH_SIZE scaled by 2.0: it reads 524,306 KiB for the
# 1 GiB read of the library's copy kernel; WRITE_SIZE reads it exactly):
# profiles/plain_fp64_256cube_hbm_counters.md, profiles/replica_fp32_1024_hbm_counters.md.
PLAIN_TRAFFIC_BYTES_PER_EVENT = 75.4
REPLICA_TRAFFIC_BYTES_PER_EVENT = 60.1

WORKLOAD = dict(box=(256, 256, 256), alloy=[("Fe", 9666), ("Cu", 134), ("Mn", 130), ("Ni", 70)], va_count=6711,
                seeds=dict(create_types=100565, create_vacancy=16650, event_selection=563264, time_inc=594689),
                temperature=573.0, attempt_freq=6e12, sector_T=1e-6, physics_time=1.0)


def host_cpu():
    """The host CPU model and its logical CPU count, for the CPU baselines."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return "%s, %d logical CPUs" % (model, os.cpu_count() or 0)


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
    # Pinned to one core for the timed loop (as taskset -c would): the oracle
    # runs in this thread, so its affinity is set and restored afterwards.
    allowed = os.sched_getaffinity(0)
    core = min(allowed)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        done = 0
        while done < events and time.perf_counter() - t0 < seconds_budget:
            ref.run(10)
            done = ref.state()["events"]
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, allowed)
    return {"value": done / dt, "unit": "events/s", "cores": 1, "kind": "port", "host": host_cpu(),
            "pinned_cpu": core,
            "sample": "first %d events of the same 256^3 FP64 workload, oracle/refcpu.cpp "
                      "(reference CPU algorithm: hash occupancy, full-sector sweep per event), "
                      "1 thread pinned to CPU %d, initial occupancy excluded, %.1f s" % (done, core, dt)}


def _oracle_replica(cfg, seconds):
    """Events/s of the oracle on one replica (a worker process: no GPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refcpu
    ref = refcpu.RefCpu(cfg)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        ref.run(5)
    return ref.state()["events"], time.perf_counter() - t0


def cpu_batched_baseline(cfgs, seconds):
    """The oracle on independent replicas, one per host core (one process
    each, started fresh), for `seconds`: aggregate events/s."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    cores = max(1, min(16, os.cpu_count() or 1))
    with ProcessPoolExecutor(max_workers=cores, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(_oracle_replica, cfgs[:cores], [seconds] * cores))
    value = sum(e / t for e, t in res)
    return {"value": value, "unit": "events/s", "cores": cores, "kind": "port", "host": host_cpu(),
            "sample": "oracle/refcpu.cpp on the first %d replicas of the same batch, one process per core, "
                      "%.0f s each, %d events in total" % (cores, seconds, sum(e for e, _ in res))}


def batched_leg(args, plain_value):
    """Replica engine throughput (events of all replicas per second)."""
    import akmc
    from akmc import workloads
    cfgs = workloads.replicas(args.replicas)
    t0 = time.perf_counter()
    tuning = {}
    if args.batch_vcap:
        tuning = dict(lds_vcap=args.batch_vcap, waves_per_group=args.batch_waves, order=args.batch_order,
                      interleave=args.batch_interleave)
    b = akmc.Batch(cfgs, device=0, precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_DEVICE,
                   slice_events=args.batch_slice, **tuning)
    t_create = time.perf_counter() - t0
    print("bench: %d replicas created in %.1f s" % (len(cfgs), t_create), file=sys.stderr, flush=True)
    tun = b.tune(trial_seconds=args.tune_seconds) if not (args.no_tune or tuning) else b.tuning()
    print("bench: tuning %s" % tun, file=sys.stderr, flush=True)
    # The interleave-width sweep: the best trial of the tuner at each width
    # (replica waves per SIMD), the other knobs free.
    sweep = {}
    for t in (b.tune_trials() if not (args.no_tune or tuning) else []):
        k = str(t["interleave"])
        if k not in sweep or t["events_per_second"] > sweep[k]["events_per_second"]:
            sweep[k] = {"events_per_second": t["events_per_second"], "lds_vcap": t["lds_vcap"],
                        "waves_per_group": t["waves_per_group"], "order": t["order"]}
    if sweep:
        print("bench: interleave sweep %s" % {k: "%.3g" % v["events_per_second"] for k, v in sweep.items()},
              file=sys.stderr, flush=True)
    for _ in range(args.batch_warmup):
        b.run(args.batch_events)
    akmc.sync(0)
    before = [b.replica_stats(i) for i in range(b.n)]
    # Throughput: each step advances every replica by --batch-events events
    # through replays of the captured hipGraph (a step ends with its slowest
    # replica).
    ev, ksec, launches = 0, 0.0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = b.run(args.batch_events)
        ev += st["events"]
        ksec += st["kernel_seconds"]
        launches += st["launches"]
    akmc.sync(0)
    dt = time.perf_counter() - t0
    after = [b.replica_stats(i) for i in range(b.n)]
    d_eval = sum(a["n_eval"] - x["n_eval"] for a, x in zip(after, before))
    # Algorithmic HBM bytes: 55 stencil bytes per evaluation, the 2-byte swap
    # and the 4-byte vacancy key per event (the sector state stays in LDS).
    nbytes = d_eval * 55 + ev * 6
    achieved = nbytes / max(1, launches) / (ksec / max(1, launches)) / 1e9
    # Beside it: deadline launches, in which every replica runs until a
    # deadline so that no replica waits for another (a launch of its own,
    # outside the graph).
    deadline = None
    if not args.no_deadline:
        dev, dks = 0, 0.0
        t1 = time.perf_counter()
        for _ in range(args.deadline_steps):
            st = b.run_for(args.batch_seconds)
            dev += st["events"]
            dks += st["kernel_seconds"]
        akmc.sync(0)
        ddt = time.perf_counter() - t1
        deadline = {"events_per_s": dev / ddt, "launches": args.deadline_steps, "events": dev,
                    "seconds": ddt, "step": "one launch, every replica runs until a %.0f ms deadline"
                                            % (args.batch_seconds * 1e3)}
    nvac = sum(b.replica_stats(i)["nvac"] for i in range(b.n))
    sites = sum(b.num_sites(i) for i in range(b.n))
    b.close()
    rate_error = None
    if not args.no_error_stats:
        # FP32 rate error against FP64 on a subset (every rate of every evaluation).
        sub = akmc.Batch(cfgs[:64], device=0, precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_DEVICE,
                         error_stats=1)
        sub.run(2000)
        errs = [sub.replica_stats(i) for i in range(sub.n)]
        sub.close()
        rate_error = {"max_rel": max(e["err_max"] for e in errs),
                      "mean_rel": sum(e["err_mean"] * e["err_n"] for e in errs) / max(1, sum(e["err_n"] for e in errs)),
                      "rates_compared": sum(e["err_n"] for e in errs),
                      "sample": "64 replicas x 2000 events, every rate of every evaluation"}
    value = ev / dt
    return {"value": value, "unit": "events/s", "replicas": len(cfgs), "sites": sites, "vacancies": nvac,
            "step": "every replica advances by %d events through hipGraph replays" % args.batch_events,
            "events_timed": ev, "ms_per_step": dt / args.steps * 1e3,
            "deadline": deadline,
            "kernel_seconds": ksec, "launches": launches, "dtype": "f32 rates, f64 sums/search/dt",
            "tuning": tun, "interleave_sweep": sweep or None,
            "vs_plain": value / plain_value if plain_value else None,
            "evaluations_per_event": d_eval / max(1, ev),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": REPLICA_TRAFFIC_BYTES_PER_EVENT * ev / max(1, launches),
                         "algorithmic_bytes_per_event": nbytes / max(1, ev)},
            "rate_error": rate_error,
            "cpu_baseline": None if args.no_cpu_baseline else cpu_batched_baseline(cfgs, args.cpu_batched_seconds)}


def served_leg(args, batched_value):
    """Serving loop on a Poisson job stream: throughput and completion latency.
    A fraction --serve-itl-frac of the jobs also carry self-interstitial
    dumbbells at 1e-5 of their sites (akmc_server_submit_itl)."""
    import random
    import akmc
    from akmc import workloads
    lo, hi = 1e5 * args.serve_scale, 1e7 * args.serve_scale
    jobs = workloads.jobs_with_itl(args.jobs, ev_lo=lo, ev_hi=hi, itl_frac=args.serve_itl_frac)
    mean_events = sum(t for _, t, _ in jobs) / len(jobs)

    def submit(srv, cfg, target, n_itl):
        return srv.submit_itl(cfg, n_itl, target) if n_itl else srv.submit(cfg, target)

    capacity = None
    if not args.serve_rate:
        # Capacity: a burst of jobs of the same law, all submitted at once.
        # Its event rate while the server is saturated (until half of the
        # jobs are done, at most --probe-seconds; the tail after that is the
        # longest jobs' serial event chains) divided by the mean target gives
        # the jobs/s the server sustains, and the offered rate is 70% of it.
        probe = workloads.jobs_with_itl(args.probe_jobs, base=0x5EED0003, ev_lo=lo, ev_hi=hi,
                                        itl_frac=args.serve_itl_frac)
        srv = akmc.Server(device=0, hbm_budget=int(250e9))
        t0 = srv.clock()
        for cfg, target, n_itl in probe:
            submit(srv, cfg, target, n_itl)
        while True:
            time.sleep(0.05)
            st = srv.stats()
            if st["done"] + st["failed"] >= len(probe) // 2 or srv.clock() - t0 >= args.probe_seconds:
                break
        probe_events = st["events"]
        probe_span = srv.clock() - t0
        srv.close()
        probe_mean = sum(t for _, t, _ in probe) / len(probe)
        capacity = probe_events / probe_span / probe_mean
        print("bench: serving capacity %.0f jobs/s (%.3g events/s over %.1f s)" %
              (capacity, probe_events / probe_span, probe_span), file=sys.stderr, flush=True)
    rate = args.serve_rate or 0.7 * capacity
    srv = akmc.Server(device=0, hbm_budget=int(250e9))
    rng = random.Random(12345)
    t_start = srv.clock()
    t_next = t_start
    ids = []
    t_note = t_start

    def note(what):
        st = srv.stats()
        print("bench: served %s: %d submitted, %d done, %d live, %.3g events" %
              (what, len(ids), st["done"], st["live"], st["events"]), file=sys.stderr, flush=True)

    for cfg, target, n_itl in jobs:
        t_next += rng.expovariate(rate)
        while True:
            now = srv.clock()
            if now >= t_next:
                break
            time.sleep(min(0.002, t_next - now))
        ids.append((submit(srv, cfg, target, n_itl), target, n_itl))
        if srv.clock() - t_note > 30:
            t_note = srv.clock()
            note("%.0f s" % (t_note - t_start))
    while True:  # a progress line every 30 s until the last job is done
        try:
            srv.wait_idle(30)
            break
        except akmc.AkmcError as e:
            if "timeout" not in str(e):
                raise
            note("%.0f s" % (srv.clock() - t_start))
    st = srv.stats()
    infos = [srv.info(j) for j, _, _ in ids]
    srv.close()

    def quantiles(sel):
        lat = sorted(i["t_done"] - i["t_submit"] for i in sel)
        if not lat:
            return None, None, None
        q = lambda p: lat[min(len(lat) - 1, int(p * len(lat)))]
        return q(0.50), q(0.99), lat[-1]

    p50, p99, pmax = quantiles(infos)
    dumb = [i for i, (_, _, n) in zip(infos, ids) if n]
    vac = [i for i, (_, _, n) in zip(infos, ids) if not n]
    d50, d99, dmax = quantiles(dumb)
    v50, v99, vmax = quantiles(vac)
    span = max(i["t_done"] for i in infos) - t_start
    ev = sum(i["events"] for i in infos)
    vspan = max(i["t_done"] for i in vac) - t_start if vac else None
    return {"value": ev / span, "unit": "events/s", "jobs": len(ids), "offered_jobs_per_s": rate,
            "capacity_jobs_per_s": capacity,
            "mean_target_events": mean_events, "target_scale": args.serve_scale,
            "p50_latency_s": p50, "p99_latency_s": p99, "max_latency_s": pmax,
            "dumbbell_jobs": len(dumb), "dumbbell_events": sum(i["events"] for i in dumb),
            "dumbbell_p50_latency_s": d50, "dumbbell_p99_latency_s": d99, "dumbbell_max_latency_s": dmax,
            "vacancy_jobs_events_per_s": sum(i["events"] for i in vac) / vspan if vac else None,
            "vacancy_jobs_p50_latency_s": v50, "vacancy_jobs_p99_latency_s": v99,
            "makespan_s": span, "gpu_busy_frac": st["busy_seconds"] / span, "launches": st["launches"],
            "hbm_peak_bytes": st["bytes_peak"], "failed": st["failed"],
            "data": "synthetic jobs (edge U{32..320}, vacancies U[1e-4,5e-4], T U[500,700] K, targets "
                    "log-uniform in [%g, %g] events); %d of them also carry dumbbells at 1e-5 of the sites, "
                    "served by the vacancy+dumbbell engine" % (lo, hi, len(dumb))}


CONFIG0 = dict(box=(40, 40, 40), alloy=[("Fe", 9666), ("Cu", 134), ("Mn", 130), ("Ni", 70)], va_count=26,
               seeds=dict(create_types=100565, create_vacancy=16650, event_selection=563264, time_inc=594689),
               temperature=573.0, attempt_freq=6e12, sector_T=3e-7, physics_time=1e3)
CONFIG0_EVENTS = 100000


def config0_line(args):
    """The reference's CPU-runnable case (40^3, 26 vacancies, sector T 3e-7 s,
    1e5 events from creation) on the plain engine, and the oracle over the
    same 1e5 events as the CPU baseline, with the final lattices compared."""
    import hashlib
    import akmc
    eng = akmc.Engine(CONFIG0, device=0)
    per = CONFIG0_EVENTS // args.steps
    akmc.sync(0)
    t0 = time.perf_counter()
    ksec = 0.0
    for _ in range(args.steps):
        st = eng.run(per)
        ksec += st["kernel_seconds"]
    akmc.sync(0)
    dt = time.perf_counter() - t0
    events = st["events"]
    gpu_hash = hashlib.sha256(eng.download().tobytes()).hexdigest()
    eng.close()
    out = {"metric": METRIC, "value": events / dt, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
           "warmup": 0, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "plain FP64 AKMC, the reference's CPU case: BCC 40^3 cells (128,000 sites), "
                                  "Fe-1.34Cu-1.30Mn-0.70Ni at%, 26 vacancies, T=573 K, T_sector=3e-7 s, "
                                  "1e5 events from creation", "events_timed": events},
           "kernel_us_per_event": ksec / max(1, events) * 1e6}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import refcpu
        ref = refcpu.RefCpu(CONFIG0)
        t0 = time.perf_counter()
        ref.run(events)
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": events / cdt, "unit": "events/s", "cores": 1, "kind": "port",
                               "host": host_cpu(), "sample": "the same %d events, oracle/refcpu.cpp, 1 thread, %.1f s" % (events, cdt)}
        out["matches_oracle"] = hashlib.sha256(ref.sites().tobytes()).hexdigest() == gpu_hash
    return out


def plain_engines_leg(args):
    """Several plain FP64 engines on the 256^3 workload (seeds differ), one
    per host thread, each on its own stream: the aggregate events/s of
    bit-exact trajectories run at once."""
    import akmc
    from concurrent.futures import ThreadPoolExecutor
    n = args.plain_engines
    cfgs = []
    for k in range(n):
        c = dict(WORKLOAD)
        c["seeds"] = {key: v + k for key, v in WORKLOAD["seeds"].items()}
        cfgs.append(c)
    with ThreadPoolExecutor(n) as ex:
        engs = list(ex.map(lambda c: akmc.Engine(c, device=0), cfgs))
    E = args.plain_engines_events
    with ThreadPoolExecutor(n) as ex:
        list(ex.map(lambda e: e.run(E // 10), engs))  # warmup
    akmc.sync(0)
    before = [e.stats()["events"] for e in engs]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(n) as ex:
        sts = list(ex.map(lambda e: e.run(E), engs))
    akmc.sync(0)
    dt = time.perf_counter() - t0
    for e in engs:
        e.close()
    # Events each engine reports performing in the timed region.
    done = [s["events"] - b for s, b in zip(sts, before)]
    assert all(s["stop"] == akmc.AKMC_STOP_EVENTS for s in sts) and all(d == E for d in done), (done, sts)
    events = sum(done)
    return {"value": events / dt, "unit": "events/s", "engines": n, "events_per_engine": E, "seconds": dt,
            "kernel_us_per_event_per_engine": sum(s["kernel_seconds"] for s in sts) / events * 1e6,
            "dtype": "f64", "note": "plain engines (the path tested bit-exact against the oracle), "
                                    "one host thread and stream each; events from each engine's own counter"}


def itl_leg(args):
    """Dumbbell (self-interstitial) rates on one serving job's lattice: the
    rate step of the abvi dumbbell model (akmc_itl_rates), device time of its
    kernels, and the oracle on the same input as the CPU baseline."""
    import akmc
    from akmc import workloads
    L = args.itl_edge
    box = (L, L, L)
    types, sites, orient = workloads.typed_lattice(box, 0x5EED0004, args.itl_fraction)
    akmc.itl_rates(box, types, sites, orient, 6e12, 573.0)  # warmup
    ksec, calls, nrates = 0.0, 0, 0
    t0 = time.perf_counter()
    for _ in range(args.itl_calls):
        tm = {}
        av, r = akmc.itl_rates(box, types, sites, orient, 6e12, 573.0, timing=tm)
        ksec += tm["kernel_seconds"]
        calls += 1
        nrates += int(sum(bin(int(a)).count("1") for a in av))
    wall = time.perf_counter() - t0
    leg = {"value": nrates / ksec, "unit": "dumbbell rates/s", "dtype": "f64",
           "config": {"workload": "abvi dumbbell rates: BCC %d^3 cells (%d sites), Fe-1.34Cu-0.70Ni-1.30Mn at%%, "
                                  "%d dumbbells (fraction %g), T=573 K" % (L, types.size, sites.size, args.itl_fraction),
                      "calls": calls, "rates_per_call": nrates // max(1, calls)},
           "kernel_ms_per_call": ksec / calls * 1e3, "wall_ms_per_call": wall / calls * 1e3,
           "note": "kernel time of the two kernels; a call also checks and uploads the whole typed lattice"}
    if not args.no_cpu_baseline:
        from oracle import refcpu
        o = refcpu.Itl()
        t0 = time.perf_counter()
        av_o, r_o = o.rates(box, types, sites, orient, 6e12, 573.0)
        dt = time.perf_counter() - t0
        leg["cpu_baseline"] = {"value": nrates // max(1, calls) / dt, "unit": "dumbbell rates/s", "cores": 1,
                               "kind": "port", "host": host_cpu(),
                               "sample": "oracle/refcpu.cpp (refitl) on the same lattice, one call, "
                                                         "%.2f s" % dt}
        leg["matches_oracle"] = bool((av_o == av).all() and (r_o.view("u8") == r.view("u8")).all())
    return leg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events-per-step", type=int, default=250_000)
    ap.add_argument("--cpu-events", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-batched-seconds", type=float, default=8.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the streaming copy kernel (profiler counter calibration)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--replicas", type=int, default=1024)
    ap.add_argument("--batch-events", type=int, default=20000, help="events per replica of the fixed-event run")
    ap.add_argument("--batch-seconds", type=float, default=0.1, help="deadline of one batched step")
    ap.add_argument("--batch-slice", type=int, default=4096)
    ap.add_argument("--tune-seconds", type=float, default=0.01)
    ap.add_argument("--no-tune", action="store_true", help="batched leg: keep the default tuning")
    ap.add_argument("--batch-vcap", type=int, default=0, help="batched leg: fixed tuning (LDS capacity) instead of the tuner")
    ap.add_argument("--batch-waves", type=int, default=1)
    ap.add_argument("--batch-order", type=int, default=0)
    ap.add_argument("--batch-interleave", type=int, default=1, help="batched leg, fixed tuning: replica waves per SIMD")
    ap.add_argument("--batch-warmup", type=int, default=1)
    ap.add_argument("--no-deadline", action="store_true", help="batched leg: skip the deadline launches")
    ap.add_argument("--deadline-steps", type=int, default=3)
    ap.add_argument("--no-error-stats", action="store_true", help="batched leg: skip the FP32 error sample")
    ap.add_argument("--config0", action="store_true",
                    help="instead: the plain engine and the CPU baseline on the reference's 40^3 case, 1e5 events")
    ap.add_argument("--no-itl", action="store_true", help="skip the dumbbell-rate leg")
    ap.add_argument("--plain-engines", type=int, default=4, help="plain engines run at once (0: skip that leg)")
    ap.add_argument("--plain-engines-events", type=int, default=500_000)
    ap.add_argument("--itl-edge", type=int, default=256)
    ap.add_argument("--itl-fraction", type=float, default=1e-5)
    ap.add_argument("--itl-calls", type=int, default=5)
    ap.add_argument("--jobs", type=int, default=4000)
    ap.add_argument("--serve-scale", type=float, default=0.01,
                    help="scale of the job targets (1 = the full 1e5..1e7 events)")
    ap.add_argument("--serve-rate", type=float, default=0.0, help="offered jobs/s (0: 70%% of the probed capacity)")
    ap.add_argument("--serve-itl-frac", type=float, default=0.1,
                    help="fraction of served jobs that also carry dumbbells at 1e-5 of the sites")
    ap.add_argument("--probe-jobs", type=int, default=400, help="jobs in the capacity probe burst")
    ap.add_argument("--probe-seconds", type=float, default=20.0, help="longest capacity probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only (one lattice, batch or server per GPU; sharding replicas, "
              "jobs or sub-boxes across GPUs is not built); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        return 2

    # Streams of this process map onto HIP's hardware queues (4 by default,
    # some taken by the runtime's own streams); the plain-engines leg runs one
    # stream per engine, so ask for at least 8 before HIP starts.
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"
    if args.config0:
        print(json.dumps(config0_line(args)))
        return 0
    import akmc
    out = {"metric": METRIC, "value": None, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    value = None
    if not args.no_plain:
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
        per_step = []
        prev, tp = s0, t0
        for i in range(args.steps):
            st = eng.run(E)
            ksec += st["kernel_seconds"]
            launches += st["launches"]
            # Per-step figures (the host reads counters the run already
            # returned; no extra device work): they show how the cost per
            # event moves as the lattice evolves.
            tn = time.perf_counter()
            de = max(1, st["events"] - prev["events"])
            rec = {"events": st["events"], "seconds": tn - tp, "us_per_event": (tn - tp) / de * 1e6,
                   "kernel_us_per_event": st["kernel_seconds"] / de * 1e6,
                   "eval_per_event": (st["n_eval"] - prev["n_eval"]) / de,
                   "fold_per_event": (st["n_fold"] - prev["n_fold"]) / de,
                   "shift_per_event": (st["n_shift"] - prev["n_shift"]) / de}
            per_step.append(rec)
            prev, tp = st, tn
            print("bench: step %d/%d events %d  %.3f s  %.2f us/event  eval %.2f  fold %.0f  shift %.0f per event"
                  % (i + 1, args.steps, st["events"], rec["seconds"], rec["us_per_event"], rec["eval_per_event"],
                     rec["fold_per_event"], rec["shift_per_event"]), file=sys.stderr, flush=True)
        akmc.sync(0)
        dt = time.perf_counter() - t0
        s1 = eng.stats()
        eng.close()
        d_events = s1["events"] - s0["events"]
        value = d_events / dt
        nbytes = algorithmic_bytes(d_events, s1["n_eval"] - s0["n_eval"], s1["n_fold"] - s0["n_fold"],
                                   s1["n_shift"] - s0["n_shift"])
        per_launch = nbytes / max(1, launches)
        avg_launch_s = ksec / max(1, launches)
        achieved = per_launch / avg_launch_s / 1e9
        out.update({
            "value": value, "ms_per_step": dt / args.steps * 1e3,
            "config": {"workload": "plain FP64 AKMC: BCC 256^3 cells (33,554,432 sites), Fe-1.34Cu-1.30Mn-0.70Ni "
                                   "at%, 6711 vacancies, T=573 K, nu=6e12/s, 8-sector sublattice, T_sector=1e-6 s",
                       "events_per_step": E, "events_timed": d_events, "launches": launches},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": PLAIN_TRAFFIC_BYTES_PER_EVENT * d_events / max(1, launches),
                         "algorithmic_bytes_per_event": nbytes / max(1, d_events),
                         "kernel_us_per_event": ksec / max(1, d_events) * 1e6,
                         # HBM bytes the counters measure per event (the algorithmic bytes are mostly
                         # sector state kept in LDS): the fraction of peak the device traffic reaches.
                         "hbm_counter_bytes_per_event": PLAIN_TRAFFIC_BYTES_PER_EVENT,
                         "hbm_counter_frac": PLAIN_TRAFFIC_BYTES_PER_EVENT * d_events / max(ksec, 1e-12)
                                             / 1e9 / HBM_PEAK_GBS,
                         "note": "one workgroup (an event wavefront and two helpers) runs the serial "
                                 "event chain: bound by dependent latency (LDS, L2/HBM gathers, FP64 "
                                 "folds), not by HBM bandwidth"},
            "per_step": per_step})
    if args.copy_calibration:
        akmc.copy_bench(1 << 30, iters=5)
    batched = None
    if not args.no_batched:
        batched = batched_leg(args, value)
        out["batched"] = batched
        print("bench: batched %.3g events/s" % batched["value"], file=sys.stderr, flush=True)
    if not args.no_served:
        out["served"] = served_leg(args, batched["value"] if batched else 5e7)
        print("bench: served %.3g events/s" % out["served"]["value"], file=sys.stderr, flush=True)
    if args.plain_engines > 0 and not args.no_plain:
        out["plain_engines"] = plain_engines_leg(args)
        print("bench: %d plain engines %.3g events/s" % (args.plain_engines, out["plain_engines"]["value"]),
              file=sys.stderr, flush=True)
    if not args.no_itl:
        out["itl"] = itl_leg(args)
        print("bench: dumbbell rates %.3g /s" % out["itl"]["value"], file=sys.stderr, flush=True)
    out["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(args.cpu_events, args.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
