"""Precision study of the FP32 rate form (512^3 workload, 16 seeds).

    python3 misa-akmc-gpu_amd/akmc/studies.py precision --L 512 --seeds 16 --events 10000000 --out result.json
    python3 misa-akmc-gpu_amd/akmc/studies.py plain --L 512 --seeds 16 --events 10000000 --threads 8 --out result.json
    (one precision per run: --only fp32 / --only fp64, then
     studies.py combine fp32.json fp64.json --out result.json)

`precision`: the same seeds (initial lattices from creation::createRandom)
advanced by the replica engine with FP32 rates and with FP64 rates (the
reference's exact rate evaluation), both batches at the same time, then
Cu-cluster statistics of every final lattice, the per-rate error of the FP32
form (every rate of every evaluation against FP64, on a separate sample run
of --error-events events per seed), and whether the two final lattices of a
seed are identical.  `plain`: the plain FP64 engine (the reference's trajectory) and
the replica engine's FP32 form on the same seeds, compared statistically
(their random streams differ).
"""
import argparse
import hashlib
import json
import math
import os
import sys
import time

if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np

import akmc
from akmc import workloads


def _stats(sites, box, hist_cap):
    c = akmc.cu_clusters(sites, box, hist_cap)
    sizes = np.arange(hist_cap)
    n = max(1, c["clusters"])
    return {"clusters": c["clusters"], "cu": c["cu"], "isolated": c["isolated"],
            "mean_size": float((c["hist"] * sizes).sum() / n), "max_size": int(sizes[c["hist"] > 0].max()),
            "hist": c["hist"].tolist()}


def _summary(per):
    out = {}
    for k in ("clusters", "isolated", "mean_size", "max_size"):
        v = np.array([p[k] for p in per], dtype=np.float64)
        out[k] = {"mean": float(v.mean()), "std": float(v.std(ddof=1)) if len(v) > 1 else 0.0}
    out["hist_total"] = np.sum([p["hist"] for p in per], axis=0).tolist()
    return out


def _welch(a, b, key):
    x = np.array([p[key] for p in a], dtype=np.float64)
    y = np.array([p[key] for p in b], dtype=np.float64)
    se = math.sqrt(x.var(ddof=1) / len(x) + y.var(ddof=1) / len(y)) if len(x) > 1 else 0.0
    return float((x.mean() - y.mean()) / se) if se > 0 else 0.0


def _advance(b, name, events, chunk, log, deadline=None):
    """Advances every replica by `events` in chunks; with a deadline
    (time.perf_counter() value) it stops before a chunk that would end past
    it.  Returns (seconds, events per replica done)."""
    done = 0
    t0 = time.perf_counter()
    last = 0.0
    while done < events:
        if deadline is not None and time.perf_counter() + last > deadline:
            log("%s: deadline reached at %d events per seed" % (name, done))
            break
        n = min(chunk, events - done)
        t1 = time.perf_counter()
        b.run(n)
        last = time.perf_counter() - t1
        done += n
        log("%s: %d / %d events per seed, %.1f s" % (name, done, events, time.perf_counter() - t0))
    return time.perf_counter() - t0, done


def precision(L, seeds, events, hist_cap, chunk, log, error_events=1_000_000, which=("fp32", "fp64"),
              deadline=None):
    """FP32 and FP64 rates on the same seeds, the two batches advanced at the
    same time (one host thread and stream each).  The FP32 rate error (every
    rate of every evaluation against the FP64 formula, which doubles the cost
    of an evaluation) is measured on a separate run of error_events events per
    seed from the same initial lattices.  `which` runs one precision only (a
    run too long for one sitting); combine() then compares two such results,
    the final lattices through their SHA-256.  `deadline` (seconds from the
    start) stops the runs early, at a chunk boundary, when time is short."""
    t_start = time.perf_counter()
    t_end = t_start + deadline if deadline else None
    cfgs = [workloads.precision_512(k, L) for k in range(seeds)]
    box = cfgs[0]["box"]
    res = {"workload": "BCC %d^3, Fe-1.34Cu-0.70Ni-1.30Mn at%%, %d vacancies, T=563 K, %d seeds, %d events each"
                       % (L, cfgs[0]["va_count"], seeds, events)}
    if error_events > 0:
        t0 = time.perf_counter()
        b = akmc.Batch(cfgs, precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_REFERENCE, error_stats=1)
        _advance(b, "fp32 error sample", error_events, chunk, log)
        rs = [b.replica_stats(i) for i in range(seeds)]
        res["rate_error"] = {"max_rel": max(r["err_max"] for r in rs),
                             "mean_rel": sum(r["err_mean"] * r["err_n"] for r in rs) / max(1, sum(r["err_n"] for r in rs)),
                             "rates_compared": sum(r["err_n"] for r in rs),
                             "rates_below_flt_min": sum(r["err_tiny"] for r in rs),
                             "sample": "%d seeds x %d events, every rate of every evaluation" % (seeds, error_events),
                             "seconds": time.perf_counter() - t0}
        b.close()
    batches = {}
    for name, prec in (("fp32", akmc.AKMC_PREC_FP32), ("fp64", akmc.AKMC_PREC_FP64)):
        if name not in which:
            continue
        t0 = time.perf_counter()
        # Direct launches: the two batches are driven from two host threads.
        batches[name] = akmc.Batch(cfgs, precision=prec, init=akmc.AKMC_INIT_REFERENCE, use_graph=0)
        log("%s: %d lattices created in %.1f s" % (name, seeds, time.perf_counter() - t0))
    if "fp32" in batches:
        # (the FP64 batch starts from the same lattices: creation does not depend on the precision)
        res["initial"] = _summary([_stats(batches["fp32"].download(i), box, hist_cap) for i in range(seeds)])
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(2) as ex:
        futs = {name: ex.submit(_advance, b, name, events, chunk, log, t_end) for name, b in batches.items()}
        secs = {name: f.result()[0] for name, f in futs.items()}
    finals = {}
    for name, b in batches.items():
        rs = [b.replica_stats(i) for i in range(seeds)]
        finals[name] = [b.download(i) for i in range(seeds)]
        per = [_stats(s, box, hist_cap) for s in finals[name]]
        ev_done = sum(int(r["events"]) for r in rs)
        res[name] = {"seconds": secs[name], "events_per_s": ev_done / secs[name], "summary": _summary(per),
                     "per_seed": per, "sim_time": [float(np.sum(r["evolution"])) / 8 for r in rs],
                     "events": [int(r["events"]) for r in rs],
                     "final_sha256": [hashlib.sha256(f.tobytes()).hexdigest() for f in finals[name]]}
        b.close()
    if len(batches) == 2:
        res.update(_compare(res))
    return res


def _compare(res):
    a, b = res["fp32"], res["fp64"]
    return {"identical_final_lattices": int(sum(x == y for x, y in zip(a["final_sha256"], b["final_sha256"]))),
            "welch_t": {k: _welch(a["per_seed"], b["per_seed"], k) for k in ("clusters", "isolated", "mean_size")}}


def combine(paths):
    """Merges precision results of one precision each (same workload)."""
    res = {}
    for p in paths:
        with open(p) as f:
            r = json.load(f)
        for k, v in r.items():
            if k in ("fp32", "fp64", "rate_error") or k not in res:
                res[k] = v
    res.update(_compare(res))
    return res


def plain(L, seeds, events, hist_cap, log, threads=1):
    cfgs = []
    for k in range(seeds):
        c = workloads.precision_512(k, L)
        c["physics_time"] = 1e3
        cfgs.append(c)
    box = cfgs[0]["box"]
    res = {"workload": "BCC %d^3, %d vacancies, T=563 K, %d seeds, %d events each" % (L, cfgs[0]["va_count"], seeds,
                                                                                       events)}
    t0 = time.perf_counter()

    def one(k):
        # One engine per host thread: each runs on its own stream.
        e = akmc.Engine(cfgs[k])
        done = 0
        while done < events:
            n = min(1_000_000, events - done)
            e.run(n)
            done += n
            log("plain: seed %d, %d / %d events, %.1f s" % (k, done, events, time.perf_counter() - t0))
        s = e.download()
        e.close()
        return _stats(s, box, hist_cap)

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max(1, threads)) as ex:
        per_plain = list(ex.map(one, range(len(cfgs))))
    res["plain_fp64"] = {"seconds": time.perf_counter() - t0, "summary": _summary(per_plain), "per_seed": per_plain}
    t0 = time.perf_counter()
    b = akmc.Batch(cfgs, precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_REFERENCE)
    done = 0
    while done < events:
        n = min(1_000_000, events - done)
        b.run(n)
        done += n
        log("replica fp32: %d / %d events per seed, %.1f s" % (done, events, time.perf_counter() - t0))
    per_b = [_stats(b.download(i), box, hist_cap) for i in range(seeds)]
    b.close()
    log("replica fp32 done, %.1f s" % (time.perf_counter() - t0))
    res["replica_fp32"] = {"seconds": time.perf_counter() - t0, "summary": _summary(per_b), "per_seed": per_b}
    res["welch_t"] = {k: _welch(per_b, per_plain, k) for k in ("clusters", "isolated", "mean_size")}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["precision", "plain", "combine"])
    ap.add_argument("inputs", nargs="*", help="combine: precision results to merge")
    ap.add_argument("--only", choices=["fp32", "fp64"], default="", help="precision study: one precision only")
    ap.add_argument("--deadline", type=float, default=0.0,
                    help="precision study: stop advancing (at a chunk boundary) this many seconds after the start")
    ap.add_argument("--L", type=int, default=512)
    ap.add_argument("--seeds", type=int, default=16)
    ap.add_argument("--events", type=int, default=10_000_000)
    ap.add_argument("--chunk", type=int, default=1_000_000)
    ap.add_argument("--error-events", type=int, default=1_000_000,
                    help="precision study: events per seed of the FP32 rate-error sample (0: none)")
    ap.add_argument("--hist-cap", type=int, default=64)
    ap.add_argument("--out", default="")
    ap.add_argument("--threads", type=int, default=1, help="plain study: engines run at once, one per host thread")
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.study == "precision":
        res = precision(a.L, a.seeds, a.events, a.hist_cap, a.chunk, log, a.error_events,
                        (a.only,) if a.only else ("fp32", "fp64"), a.deadline or None)
    elif a.study == "combine":
        res = combine(a.inputs)
    else:
        res = plain(a.L, a.seeds, a.events, a.hist_cap, log, a.threads)
    text = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    print(json.dumps({k: v for k, v in res.items() if k not in ("fp32", "fp64", "plain_fp64", "replica_fp32")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
