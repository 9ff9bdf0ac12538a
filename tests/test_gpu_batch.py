"""Replica engine (batched form) on the GPU, through the C ABI.

Parity bars:
* FP64 rates: bit-exact with the oracle's rates for every vacancy of every
  sector (same lattice, same side effect on stale tags);
* FP32 rates: relative error <= 1e-6 against the oracle's FP64 rates (the
  form's budget; the table bound is 1.5e-7);
* trajectories: the engine's own sums and xoshiro streams are not the
  reference's, so a replica is pinned by invariants, by selection statistics
  (hop directions against the rate proportions, residence times against
  1/total) and by reproducibility: the same replica gives the same lattice
  whatever batch, slicing and tuning it runs under, and FP32 and FP64 rates
  give the same trajectory from the same seed.
"""
import sys

import numpy as np
import pytest

import akmc
import refcpu
from conftest import make_cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _counts(s):
    return [int((s & 1).sum()), int(((s >> 1) & 1).sum())]


RATE_CASES = [
    dict(n=16, va=6),
    dict(n=12, va=60, T=650.0, alloy=[("Fe", 6000), ("Mn", 2000), ("Ni", 1500), ("Cu", 500)]),
    dict(box=(17, 13, 11), va=30, T=520.0),
    dict(n=32, va=700),   # several 64-vacancy blocks per sector
    dict(n=48, va=40000),  # > 4096 vacancies per sector: the superblock level
]


@pytest.mark.parametrize("case", RATE_CASES)
def test_fp64_sector_rates_bit_exact(case):
    say("test_fp64_sector_rates_bit_exact", case)
    cfg = make_cfg(**case)
    b = akmc.Batch([cfg], precision=akmc.AKMC_PREC_FP64, init=akmc.AKMC_INIT_REFERENCE)
    ref = refcpu.RefCpu(cfg)
    assert np.array_equal(b.download(0), ref.sites())
    for s in range(8):
        gs, gr, gt = b.sector_rates(0, s)
        rs, rr, rt = ref.sector_rates(s)
        o = np.argsort(gs)
        assert np.array_equal(gs[o], rs), s           # the same vacancies (sweep order vs list order)
        assert np.array_equal(_bits(gr[o]), _bits(rr)), s
        assert gt == pytest.approx(rt, rel=1e-13), s
    assert np.array_equal(b.download(0), ref.sites())  # same stale-tag side effect
    b.close()


@pytest.mark.parametrize("case", RATE_CASES)
def test_fp32_sector_rates_within_budget(case):
    say("test_fp32_sector_rates_within_budget", case)
    cfg = make_cfg(**case)
    b = akmc.Batch([cfg], precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_REFERENCE)
    ref = refcpu.RefCpu(cfg)
    for s in range(8):
        gs, gr, gt = b.sector_rates(0, s)
        rs, rr, rt = ref.sector_rates(s)
        o = np.argsort(gs)
        assert np.array_equal(gs[o], rs), s
        g = gr[o]
        assert np.array_equal(g == 0, rr == 0), s     # the same available directions
        nz = rr > 0
        if nz.any():
            rel = np.abs(g[nz] - rr[nz]) / rr[nz]
            assert rel.max() <= 1e-6, (s, rel.max())
            assert rel.max() <= 2.0 ** -22, (s, rel.max())
        assert gt == pytest.approx(rt, rel=1e-6), s
    assert np.array_equal(b.download(0), ref.sites())
    b.close()


def _mixed_configs(k, seed=5, lo=8, hi=20):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        L = int(rng.integers(lo, hi + 1))
        cu, ni, mn = rng.uniform(10, 150), rng.uniform(20, 350), rng.uniform(30, 150)
        va = max(1, int(round(2e-4 * 2 * L ** 3 * 10)))
        out.append(make_cfg(n=L, va=va, T=float(rng.uniform(500, 700)),
                            alloy=[("Fe", int(10000 - cu - ni - mn)), ("Cu", int(cu)), ("Ni", int(ni)), ("Mn", int(mn))],
                            seeds=dict(create_types=1000 + i, create_vacancy=2000 + i, time_inc=3000 + i),
                            physics=1e3))
    return out


def test_invariants_and_event_counts():
    say("test_invariants_and_event_counts")
    cfgs = _mixed_configs(24)
    b = akmc.Batch(cfgs, precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_DEVICE)
    s0 = [b.download(i) for i in range(b.n)]
    st = b.run(3000)
    assert st["failed"] == 0
    for i in range(b.n):
        r = b.replica_stats(i)
        assert r["events"] == 3000, (i, r)
        s1 = b.download(i)
        assert _counts(s1) == _counts(s0[i]), i     # vacancies and Cu conserved
        assert ((s0[i] & 1) != (s1 & 1)).any(), i   # the vacancies moved
        assert r["nvac"] == int((s1 & 1).sum())
    assert st["events"] == 3000 * b.n
    b.close()


def test_superblock_replica_invariants():
    """A replica with more than 4096 vacancies per sector (three scan levels),
    FP32 and FP64, both kinds of launch."""
    say("test_superblock_replica_invariants")
    cfg = make_cfg(n=48, va=40000, T=600.0, physics=1e3)
    for prec in (akmc.AKMC_PREC_FP32, akmc.AKMC_PREC_FP64):
        b = akmc.Batch([cfg, make_cfg(n=12, va=3)], precision=prec, init=akmc.AKMC_INIT_REFERENCE)
        s0 = b.download(0)
        st = b.run(3000)
        assert st["failed"] == 0
        s1 = b.download(0)
        assert _counts(s1) == _counts(s0)
        assert b.replica_stats(0)["events"] == 3000
        b.close()


def _final(cfgs, idx, events, runs=1, **opt):
    b = akmc.Batch(cfgs, init=akmc.AKMC_INIT_DEVICE, **opt)
    for _ in range(runs):
        b.run(events // runs)
    out = [(b.download(i), b.replica_stats(i)) for i in idx]
    b.close()
    return out


def test_results_do_not_depend_on_batch_slicing_or_tuning():
    say("test_results_do_not_depend_on_batch_slicing_or_tuning")
    cfgs = _mixed_configs(12, seed=9, lo=8, hi=24)
    n = 4000
    base = _final(cfgs, range(12), n)
    variants = [
        dict(slice_events=97, runs=4),                   # many launches, resumes mid-sector
        dict(lds_vcap=-1, waves_per_group=1),            # global kind only
        dict(lds_vcap=64, waves_per_group=4, order=0),
        dict(use_graph=0, slice_events=333),
        dict(interleave=2),                               # two replica waves per SIMD
        dict(interleave=4, lds_vcap=-1),
    ]
    for v in variants:
        runs = v.pop("runs", 1)
        got = _final(cfgs, range(12), n, runs=runs, **v)
        for i, ((s_a, r_a), (s_b, r_b)) in enumerate(zip(base, got)):
            assert np.array_equal(s_a, s_b), (v, i)
            assert r_a["events"] == r_b["events"]
            assert np.array_equal(_bits(r_a["evolution"]), _bits(r_b["evolution"])), (v, i)
            assert r_a["sector_time"] == r_b["sector_time"], (v, i)
    # A replica alone and inside a larger batch.
    alone = _final([cfgs[3]], [0], n)[0]
    mixed = _final(_mixed_configs(40, seed=11) + [cfgs[3]], [40], n)[0]
    assert np.array_equal(alone[0], base[3][0]) and np.array_equal(mixed[0], base[3][0])


def test_fp32_and_fp64_rates_give_the_same_trajectory():
    say("test_fp32_and_fp64_rates_give_the_same_trajectory")
    cfgs = _mixed_configs(8, seed=21)
    a = _final(cfgs, range(8), 2000, precision=akmc.AKMC_PREC_FP32)
    b = _final(cfgs, range(8), 2000, precision=akmc.AKMC_PREC_FP64)
    same = sum(np.array_equal(x[0], y[0]) for x, y in zip(a, b))
    # A selection can differ only when u * total falls within ~1e-7 of a
    # cumulative boundary; over 16,000 events that is rare.
    assert same >= 7, same
    for (sa, ra), (sb, rb) in zip(a, b):
        if np.array_equal(sa, sb):
            assert np.allclose(ra["evolution"], rb["evolution"], rtol=1e-5, atol=0)


def test_error_stats_recorded():
    say("test_error_stats_recorded")
    b = akmc.Batch(_mixed_configs(6, seed=31), precision=akmc.AKMC_PREC_FP32, error_stats=1)
    b.run(2000)
    for i in range(b.n):
        r = b.replica_stats(i)
        assert r["err_n"] > 100
        assert 0 < r["err_max"] <= 1e-6, r["err_max"]
        assert r["err_mean"] < r["err_max"]
    b.close()


def test_selection_statistics():
    """One vacancy in identical 8^3 lattices, 4096 seeds, one event each: the
    hop directions follow the rate proportions (chi-square) and the residence
    time has mean 1/total."""
    say("test_selection_statistics")
    base = make_cfg(n=8, va=1, T=600.0, sector_T=1.0, physics=1e3,
                    alloy=[("Fe", 6000), ("Cu", 1500), ("Ni", 1500), ("Mn", 1000)])
    nrep = 4096
    cfgs = []
    for i in range(nrep):
        c = dict(base)
        c["seeds"] = dict(base["seeds"], time_inc=77 + i)
        cfgs.append(c)
    b = akmc.Batch(cfgs, precision=akmc.AKMC_PREC_FP64, init=akmc.AKMC_INIT_REFERENCE)
    ref = refcpu.RefCpu(base)
    s0 = ref.sites()
    v = int(np.nonzero(s0 & 1)[0][0])
    X, Y = 16, 8
    x, y, z = v % X, (v // X) % Y, v // (X * Y)
    targets = []
    for d in range(8):
        dy0 = 0 if x % 2 else -1
        targets.append(ref.site_index(x + (1 if d & 4 else -1), y + dy0 + ((d >> 1) & 1), z + dy0 + (d & 1)))
    rates = np.array([ref.rate(v, d) if not (s0[targets[d]] & 1) else 0.0 for d in range(8)])
    b.run(1)
    hits = np.zeros(8)
    times = []
    for i in range(nrep):
        s1 = b.download(i)
        w = int(np.nonzero(s1 & 1)[0][0])
        hits[targets.index(w)] += 1
        times.append(b.replica_stats(i)["sector_time"])
    p = rates / rates.sum()
    exp_counts = p * nrep
    m = exp_counts > 5
    chi2 = (((hits - exp_counts) ** 2)[m] / exp_counts[m]).sum()
    assert chi2 < 30.0, (chi2, hits, exp_counts)   # 7 degrees of freedom: p < 1e-4 beyond 30
    assert hits[~m].sum() <= 20
    t = np.array(times)
    mean, sd = t.mean(), t.std() / np.sqrt(nrep)
    assert abs(mean - 1.0 / rates.sum()) < 5 * sd, (mean, 1.0 / rates.sum(), sd)
    b.close()


def test_autotuner_picks_a_valid_setting():
    say("test_autotuner_picks_a_valid_setting")
    cfgs = _mixed_configs(32, seed=41)
    b = akmc.Batch(cfgs, init=akmc.AKMC_INIT_DEVICE)
    t = b.tune(trial_seconds=0.001)
    assert t["trials"] >= 6 and t["events_per_second"] > 0
    assert t["interleave"] in (1, 2, 4)
    assert t["waves_per_group"] == (4 * t["interleave"] if t["interleave"] > 1 else t["waves_per_group"])
    assert t["waves_per_group"] in (1, 2, 4, 8, 16)
    trials = b.tune_trials()
    assert len(trials) == t["trials"]
    assert {x["interleave"] for x in trials} == {1, 2, 4}          # the interleave widths are swept
    assert max(x["events_per_second"] for x in trials) == t["events_per_second"]
    done = [b.replica_stats(i)["events"] for i in range(32)]
    assert len(set(done)) == 1 and done[0] > 0    # every replica brought to the same count
    b.run(500)
    ref = _final(cfgs, range(32), done[0] + 500)
    for i in range(32):
        assert np.array_equal(b.download(i), ref[i][0]), i   # tuning trials advanced the same trajectory
    b.close()


def test_server_alone_vs_mix_and_budget():
    say("test_server_alone_vs_mix_and_budget")
    jobs = _mixed_configs(6, seed=51, lo=10, hi=24)
    targets = [1500, 3000, 800, 2500, 1200, 2000]
    # Alone: one server per job.
    alone = []
    for c, t in zip(jobs, targets):
        srv = akmc.Server(keep_results=1)
        j = srv.submit(c, t)
        srv.wait_idle(120)
        info = srv.info(j)
        assert info["status"] == akmc.JOB_DONE and info["events"] == t
        alone.append(srv.result(j, 2 * c["box"][0] * c["box"][1] * c["box"][2]))
        srv.close()
    # In a mix: other jobs before, between and after, under a budget that
    # admits only part of the queue at once.
    others = _mixed_configs(20, seed=61, lo=16, hi=40)
    per_job = max(2 * c["box"][0] ** 3 for c in others) * 3
    srv = akmc.Server(keep_results=1, hbm_budget=int(per_job * 4), slice_seconds=5e-4)
    ids = []
    k = 0
    for i, c in enumerate(others):
        srv.submit(c, 4000)
        if i % 3 == 0 and k < len(jobs):
            ids.append(srv.submit(jobs[k], targets[k]))
            k += 1
    while k < len(jobs):
        ids.append(srv.submit(jobs[k], targets[k]))
        k += 1
    srv.wait_idle(300)
    st = srv.stats()
    assert st["done"] == len(others) + len(jobs) and st["failed"] == 0
    assert st["bytes_peak"] <= int(per_job * 4)
    for j, c, a in zip(ids, jobs, alone):
        got = srv.result(j, 2 * c["box"][0] * c["box"][1] * c["box"][2])
        assert np.array_equal(got, a)
    with pytest.raises(akmc.AkmcError) as e:
        srv.submit(make_cfg(n=200, va=10), 10)   # 16 MB of sites alone exceeds the budget
    assert e.value.code == -6
    srv.close()


def test_brick_layout_round_trips_and_device_creation():
    # Replica sites live in 8 x 4 x 4 bricks on the device; the host sees the
    # logical order.  An odd box (padding on every axis): the vacancy list of a
    # device-created lattice sits exactly on the V-tagged sites of its
    # download, events conserve the vacancies, and upload . download is the
    # identity for an evolved lattice and for arbitrary tag bytes.
    say("test_brick_layout_round_trips_and_device_creation")
    cfg = make_cfg(box=(17, 13, 11), va=40, T=600.0)
    b = akmc.Batch([cfg, cfg], precision=akmc.AKMC_PREC_FP64, init=akmc.AKMC_INIT_DEVICE)
    s0 = b.download(0)
    assert s0.size == 2 * 17 * 13 * 11
    X, Y = 2 * 17, 13
    keys = set()
    for sec in range(8):
        gs, _, _ = b.sector_rates(0, sec)
        keys.update(int(k) for k in gs)
    assert keys == set(int(i) for i in np.flatnonzero(s0 & 1))
    b.run(3000)
    s1 = b.download(0)
    assert int((s1 & 1).sum()) == int((s0 & 1).sum())
    b.upload(1, s1)
    assert np.array_equal(b.download(1), s1)
    rng = np.random.default_rng(5)
    junk = rng.integers(0, 16, size=s1.size, dtype=np.uint8) & np.uint8(0x0E)
    junk[np.flatnonzero(s1 & 1)] |= 1
    b.upload(1, junk)
    assert np.array_equal(b.download(1), junk)
    b.close()
