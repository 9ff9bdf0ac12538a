"""Plain FP64 path on the GPU, through the C ABI, against the oracle.

The bar is bit-exact: every rate, every event's (from, to), the final site
array (including stale species tags), the 8 sector clocks and the partial
sector time.  Full-size checks (the 256^3 benchmark box) compare a prefix of
the trajectory with the oracle and the rest through invariants.
"""
import glob
import hashlib
import json
import os
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


def compare_run(cfg, n, chunk=None, pieces=1, select_mode=None, log=True):
    eng = akmc.Engine(cfg)
    if chunk:
        eng.set_rng_chunk(chunk)
    if select_mode is not None:
        eng.set_select_mode(select_mode)
    if log:
        eng.set_event_log(n)
    ref = refcpu.RefCpu(cfg)
    assert np.array_equal(eng.download(), ref.sites())
    done = 0
    for k in range(pieces):
        m = n // pieces if k < pieces - 1 else n - done
        st = eng.run(m)
        done = st["events"]
    rc, lf, lt = ref.run(n, log=True)
    rs = ref.state()
    assert st["events"] == rs["events"]
    m = rs["events"]
    if log:
        gf, gt = eng.event_log(n)
        assert np.array_equal(gf[:m], lf[:m]) and np.array_equal(gt[:m], lt[:m]), \
            "first divergence at %s" % np.nonzero((gf[:m] != lf[:m]) | (gt[:m] != lt[:m]))[0][:3]
    assert np.array_equal(eng.download(), ref.sites())
    assert np.array_equal(_bits(st["evolution"]), _bits(rs["evolution"]))
    assert st["sector_time"] == rs["sector_time"] and st["step"] == rs["step"]
    assert st["sector_pos"] == rs["sect"] and st["in_sector"] == rs["in_sector"]
    eng.close()
    return st


def test_device_exp_matches_libc():
    say("test_device_exp_matches_libc")
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-50, 0, 4_000_000), rng.uniform(-745.2, 709.8, 500_000)])
    assert np.array_equal(_bits(akmc.exp_device(x)), _bits(refcpu.exp_array(x)))


RATE_CASES = [
    dict(n=16, va=40),
    dict(n=24, va=300, alloy=[("Fe", 5000), ("Mn", 2500), ("Ni", 2500)]),
    dict(box=(17, 12, 9), va=30, T=650.0),
    dict(n=32, va=800, alloy=[("Fe", 8000), ("Cu", 1000), ("Mn", 500), ("Ni", 500)]),
]


@pytest.mark.parametrize("case", RATE_CASES)
def test_sector_rates_bit_exact(case):
    say("test_sector_rates_bit_exact", case)
    cfg = make_cfg(**case)
    eng = akmc.Engine(cfg)
    ref = refcpu.RefCpu(cfg)
    for s in range(8):
        gs, gr, gt = eng.sector_rates(s)
        rs, rr, rt = ref.sector_rates(s)
        assert np.array_equal(gs, rs), s
        assert np.array_equal(_bits(gr), _bits(rr)), s
        assert _bits(np.array([gt]))[0] == _bits(np.array([rt]))[0], s
    # the evaluation side effect (stale tags dropped) leaves identical lattices
    assert np.array_equal(eng.download(), ref.sites())
    eng.close()


TRAJ_CASES = [
    ("cfg0_40cube", make_cfg(n=40, va=26, T=573.0), 20000),
    ("dense_24", make_cfg(n=24, va=200), 20000),
    ("mnni_rich", make_cfg(n=16, va=60, T=600.0, alloy=[("Fe", 6000), ("Mn", 2000), ("Ni", 1500), ("Cu", 500)],
                           seeds=dict(create_types=7, create_vacancy=11, time_inc=13)), 20000),
    ("odd_box", make_cfg(box=(17, 13, 11), va=20, T=520.0, sector_T=5e-7,
                         seeds=dict(create_types=21, create_vacancy=22, time_inc=23)), 10000),
    ("hot_long_T", make_cfg(n=20, va=30, T=700.0, sector_T=2e-6), 20000),
    ("cluster_heavy", make_cfg(n=16, va=400, T=560.0), 10000),
    # 1024 < vacancies per sector <= capacity: a sector list that is not a
    # power of two above 1024 (the sort must not touch entries past the count).
    ("crowded_sector", make_cfg(n=32, va=9400, T=573.0), 300),
]


@pytest.mark.parametrize("name,cfg,n", TRAJ_CASES, ids=[c[0] for c in TRAJ_CASES])
def test_trajectory_bit_exact(name, cfg, n):
    say("test_trajectory_bit_exact", name, n)
    compare_run(cfg, n)


def test_default_kernel_bit_exact():
    """The event log, the self-check, the selection modes and the parity hook
    run in a second instance of the kernel; with none of them set the
    default instance runs: final lattices, sector clocks and counters
    bit-exact vs the oracle on the trajectory cases."""
    for name, cfg, n in TRAJ_CASES:
        say("test_default_kernel_bit_exact", name, n)
        compare_run(cfg, n, log=False)


@pytest.mark.parametrize("mode", [1, 2])
def test_select_modes_bit_exact(mode):
    """While the helper wavefront repairs the running sums, the event loop
    selects from their bounded estimate; on real trajectories the bound
    almost always decides, so its exact fallback (wait for the repair, copy
    the repaired sums, select exactly) is forced here: always (mode 1) or on
    alternate events (mode 2, mixing both paths and the copies each posts).
    Trajectories, clocks and lattices bit-exact vs the oracle."""
    for name, cfg, n in (TRAJ_CASES[1], TRAJ_CASES[3], TRAJ_CASES[5], TRAJ_CASES[6]):
        say("test_select_modes_bit_exact", mode, name, n)
        compare_run(cfg, n, select_mode=mode)


def test_resume_and_small_chunks():
    say("test_resume_and_small_chunks")
    cfg = make_cfg(n=24, va=120, T=590.0)
    compare_run(cfg, 12000, pieces=7)       # event-cap stops mid-sector, then resumes
    compare_run(cfg, 12000, chunk=97)      # a relaunch (full re-evaluation) every 97 events
    compare_run(cfg, 3000, chunk=1)        # a relaunch at every event


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "traj_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_trajectories(path):
    say("test_golden_trajectories", os.path.basename(path))
    z = np.load(path, allow_pickle=False)
    cfg = json.loads(str(z["cfg"]))
    n = int(z["events"])
    eng = akmc.Engine(cfg)
    eng.set_event_log(n)
    st = eng.run(n)
    gf, gt = eng.event_log(n)
    assert np.array_equal(gf, z["from_"]) and np.array_equal(gt, z["to"])
    assert hashlib.sha256(eng.download().tobytes()).hexdigest() == str(z["sites_sha256"])
    assert np.array_equal(_bits(st["evolution"]), _bits(z["evolution"]))
    assert st["sector_time"] == float(z["sector_time"])


SUM_CASES = [c[1] for c in TRAJ_CASES] + [
    make_cfg(n=24, va=200, T=520.0),
    make_cfg(n=24, va=400, T=700.0, alloy=[("Fe", 7000), ("Cu", 1500), ("Mn", 1000), ("Ni", 500)]),
    make_cfg(n=20, va=900, T=573.0),
]


def test_running_sums_self_check():
    """The incrementally kept running sums (block folds, exact tail shifts
    made by the helper wavefront into its own array and copied back, serial
    folds at binade crossings and ties) against a serial refold of the
    sector after every event, on the trajectory cases, the golden cases and
    the 256^3 benchmark box."""
    say("test_running_sums_self_check")
    cfgs = list(SUM_CASES)
    for path in GOLDEN:
        cfgs.append(json.loads(str(np.load(path, allow_pickle=False)["cfg"])))
    for cfg in cfgs:
        eng = akmc.Engine(cfg)
        eng.set_check_sums(True)
        st = eng.run(6000)
        assert st["sum_errors"] == 0, (cfg, st["sum_errors"], st["events"])
        eng.close()
    cfg = make_cfg(n=256, va=6711, T=573.0, sector_T=1e-6, physics=1.0)
    eng = akmc.Engine(cfg)
    eng.set_check_sums(True)
    st = eng.run(20000)
    assert st["events"] == 20000 and st["sum_errors"] == 0
    eng.close()


def test_run_steps_stops_at_step_boundaries():
    """akmc_run_steps ends the time loop after whole steps: the state equals
    the oracle's at the end of a time loop of that many steps (physics time
    (k - 0.5) T), bit for bit, and further runs continue the same trajectory."""
    say("test_run_steps_stops_at_step_boundaries")
    T = 3e-7
    cfg = make_cfg(n=16, va=40, T=600.0, sector_T=T, physics=1.0)
    eng = akmc.Engine(cfg)
    done = 0
    for k in (3, 4, 1):
        st = eng.run_steps(k)
        done += k
        assert st["stop"] == akmc.AKMC_STOP_STEPS and st["step"] == done and st["in_sector"] == 0
        ref = refcpu.RefCpu(make_cfg(n=16, va=40, T=600.0, sector_T=T, physics=(done - 0.5) * T))
        ref.run(10 ** 9)
        rs = ref.state()
        assert rs["step"] == done and st["events"] == rs["events"]
        assert np.array_equal(eng.download(), ref.sites())
        assert np.array_equal(_bits(st["evolution"]), _bits(rs["evolution"]))
    eng.close()


def test_step_hooks_dump_and_log(tmp_path):
    """The reference's step hook around a run: isolated-Cu log lines after
    steps 0, 100, 200 and dumps every 100 steps, each dump equal to the
    lattice at that step (readout)."""
    say("test_step_hooks_dump_and_log")
    cfg = make_cfg(n=12, va=8, T=650.0, sector_T=3e-8, physics=1.0)
    eng = akmc.Engine(cfg)
    logs = []
    out = {"dump_interval": 100, "dump_file_path": str(tmp_path / "kmc.{}.out")}
    st = akmc.run_with_hooks(eng, (12, 12, 12), 2, 201, out, log=logs.append)
    assert st["step"] == 201
    assert [l.split()[1] for l in logs] == ["0", "100", "200"]
    assert sorted(os.listdir(tmp_path)) == ["kmc.0.out", "kmc.100.out", "kmc.200.out"]
    sites, box, ghost = akmc.read_dump(tmp_path / "kmc.200.out")
    cur = eng.download()
    rb = np.zeros_like(cur)
    for bit in (8, 4, 2, 1):
        rb[(cur & bit) != 0] = bit
    assert box == (12, 12, 12) and ghost == 2 and np.array_equal(sites, rb)
    assert int(logs[-1].split()[-1].rstrip(".")) == akmc.cu_clusters(cur, box)["isolated"]
    eng.close()


def test_checkpoint_restart_is_exact(tmp_path):
    """Save after an event-cap stop inside a sector, load into a new engine
    and continue: the same sites, sector clocks and time as one straight run,
    and the same as the oracle."""
    say("test_checkpoint_restart_is_exact")
    cfg = make_cfg(n=20, va=60, T=620.0)
    a = akmc.Engine(cfg)
    st = a.run(3333)
    assert st["in_sector"] == 1
    path = tmp_path / "eng.ckpt"
    a.save(path)
    a.run(5000)
    b = akmc.Engine.load(path)
    assert b.stats()["events"] == 3333
    b.run(5000)
    sa, sb = a.stats(), b.stats()
    assert np.array_equal(a.download(), b.download())
    assert np.array_equal(_bits(sa["evolution"]), _bits(sb["evolution"]))
    assert sa["sector_time"] == sb["sector_time"] and sa["step"] == sb["step"] and sa["events"] == sb["events"]
    ref = refcpu.RefCpu(cfg)
    ref.run(8333)
    assert np.array_equal(b.download(), ref.sites())
    assert np.array_equal(_bits(sb["evolution"]), _bits(ref.state()["evolution"]))
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"not a checkpoint")
    with pytest.raises(akmc.AkmcError):
        akmc.Engine.load(bad)
    a.close()
    b.close()


def test_time_loop_end_and_empty_box():
    say("test_time_loop_end_and_empty_box")
    cfg = make_cfg(n=12, va=0, physics=3e-5, sector_T=3e-7)
    eng = akmc.Engine(cfg)
    st = eng.run(10)
    assert st["stop"] == akmc.AKMC_STOP_TIME and st["events"] == 0 and st["step"] == 100
    assert np.allclose(st["evolution"], 3e-5, rtol=1e-12)
    ref = refcpu.RefCpu(cfg)
    ref.run(10)
    assert np.array_equal(_bits(st["evolution"]), _bits(ref.state()["evolution"]))
    # a short physics time with vacancies: the loop ends before the cap
    cfg = make_cfg(n=12, va=5, physics=2e-6, sector_T=3e-7)
    eng = akmc.Engine(cfg)
    st = eng.run(10 ** 7)
    ref = refcpu.RefCpu(cfg)
    ref.run(10 ** 7)
    rs = ref.state()
    assert st["stop"] == akmc.AKMC_STOP_TIME and st["events"] == rs["events"]
    assert np.array_equal(_bits(st["evolution"]), _bits(rs["evolution"]))
    assert np.array_equal(eng.download(), ref.sites())


def test_sectors_beyond_lds_capacity():
    """Sectors of more vacancies than the LDS form holds (akmc_plain_vcap)
    run with the sector state in device memory: rates and trajectory
    bit-exact vs the oracle."""
    say("test_sectors_beyond_lds_capacity")
    cfg = make_cfg(n=32, va=16000, T=600.0)  # ~2,000 vacancies per sector
    eng = akmc.Engine(cfg)
    ref = refcpu.RefCpu(cfg)
    for s in (0, 5):
        gs, gr, gt = eng.sector_rates(s, cap=4096)
        rs, rr, rt = ref.sector_rates(s, cap=4096)
        assert len(gs) > akmc.lib().akmc_plain_vcap()
        assert np.array_equal(gs, rs) and np.array_equal(_bits(gr), _bits(rr)) and gt == rt
    eng.close()
    st = compare_run(cfg, 400)
    assert st["events"] == 400
    eng = akmc.Engine(cfg)
    eng.set_check_sums(True)
    st = eng.run(1500)
    assert st["sum_errors"] == 0
    eng.close()
    # More than 4,096 vacancies per sector: three search rounds over the sums.
    cfg = make_cfg(n=48, va=40000, T=620.0)
    eng = akmc.Engine(cfg)
    ref = refcpu.RefCpu(cfg)
    gs, gr, gt = eng.sector_rates(3, cap=8192)
    rs, rr, rt = ref.sector_rates(3, cap=8192)
    assert len(gs) > 4096
    assert np.array_equal(gs, rs) and np.array_equal(_bits(gr), _bits(rr)) and gt == rt
    eng.close()
    compare_run(cfg, 150)


def test_full_size_prefix_and_invariants():
    """The 256^3 benchmark box: the first 150 events bit-exact against the
    oracle, then 300k more events through invariants."""
    say("test_full_size_prefix_and_invariants")
    cfg = make_cfg(n=256, va=6711, T=573.0, sector_T=1e-6, physics=1.0)
    eng = akmc.Engine(cfg)
    eng.set_event_log(150)
    s0 = eng.download()
    ref = refcpu.RefCpu(cfg)
    assert np.array_equal(s0, ref.sites())
    st = eng.run(150)
    rc, lf, lt = ref.run(150, log=True)
    gf, gt = eng.event_log(150)
    assert np.array_equal(gf, lf) and np.array_equal(gt, lt)
    assert np.array_equal(eng.download(), ref.sites())
    assert st["sector_time"] == ref.state()["sector_time"]
    del ref
    st = eng.run(300_000)
    s1 = eng.download()
    assert st["events"] == 300_150
    assert (s1 & 1).sum() == (s0 & 1).sum() == 6711
    assert ((s1 >> 1) & 1).sum() == ((s0 >> 1) & 1).sum()
    assert s1.max() < 16
    eng.close()


def test_config0_full_length_hashes():
    """Config 0 at full length: 1e5 events of the 40^3 box, the occupancy
    hash after every 1000 events, the sector clocks and the partial sector
    time, against the oracle's fixture (tests/golden/make_golden.py cfg0)."""
    say("test_config0_full_length_hashes")
    path = os.path.join(os.path.dirname(__file__), "golden", "hashes_cfg0_40cube.npz")
    z = np.load(path, allow_pickle=False)
    cfg = json.loads(str(z["cfg"]))
    every = int(z["every"])
    eng = akmc.Engine(cfg)
    for k, want in enumerate(z["sha256"]):
        st = eng.run(every)
        assert st["events"] == (k + 1) * every
        got = np.frombuffer(hashlib.sha256(eng.download().tobytes()).digest(), dtype=np.uint8)
        assert np.array_equal(got, want), "occupancy differs after %d events" % ((k + 1) * every)
    assert np.array_equal(_bits(st["evolution"]), _bits(z["evolution"]))
    assert st["sector_time"] == float(z["sector_time"]) and st["step"] == int(z["step"])
    eng.close()


def test_config1_golden_hashes():
    """Config 1 (256^3 cells, 6,711 vacancies, 573 K, 1e-6 s sector
    threshold) for the oracle fixture's 1e5 events: SHA-256 of the (from, to)
    site pairs of every 1000 events, of the whole occupancy every 10,000
    events, and the sector clocks at the end (tests/golden/make_golden.py
    cfg1)."""
    say("test_config1_golden_hashes")
    path = os.path.join(os.path.dirname(__file__), "golden", "hashes_cfg1_256cube.npz")
    z = np.load(path, allow_pickle=False)
    cfg = json.loads(str(z["cfg"]))
    every, sites_every, total = int(z["every"]), int(z["sites_every"]), int(z["events"])
    assert total >= 100000 and total % sites_every == 0
    eng = akmc.Engine(cfg)
    eng.set_event_log(total)
    for k, want in enumerate(z["sites_sha256"]):
        st = eng.run(sites_every)
        assert st["events"] == (k + 1) * sites_every
        got = np.frombuffer(hashlib.sha256(eng.download().tobytes()).digest(), dtype=np.uint8)
        assert np.array_equal(got, want), "occupancy differs after %d events" % ((k + 1) * sites_every)
    lf, lt = eng.event_log(total)
    assert lf.size == total
    for k, want in enumerate(z["log_sha256"]):
        pairs = np.stack([lf[k * every:(k + 1) * every], lt[k * every:(k + 1) * every]], axis=1).astype(np.int64)
        got = np.frombuffer(hashlib.sha256(pairs.tobytes()).digest(), dtype=np.uint8)
        assert np.array_equal(got, want), "events %d..%d differ" % (k * every, (k + 1) * every)
    assert np.array_equal(_bits(st["evolution"]), _bits(z["evolution"]))
    assert st["sector_time"] == float(z["sector_time"]) and st["step"] == int(z["step"])
    assert st["sector_pos"] == int(z["sect"])
    eng.close()


def test_pipe_created_lattice_bit_exact(tmp_path):
    # Pipe creation: the random creation's atoms and vacancies from a file;
    # the engine and the oracle start from the same sites and must agree.
    cfg = make_cfg(n=16, va=0)
    rng = np.random.default_rng(9)
    f = tmp_path / "vac.txt"
    f.write_text("".join("%d %d %d\n" % (rng.integers(32), rng.integers(16), rng.integers(16)) for _ in range(14)))
    sites = akmc.pipe_sites(cfg, f)
    nv = int((sites & akmc.BIT_V).sum())
    say("pipe case: 16^3, %d vacancies" % nv)
    assert 10 <= nv <= 14
    eng = akmc.Engine(cfg, init_random=False)
    eng.upload(sites)
    ref = refcpu.RefCpu(cfg, create=False)
    ref.set_sites(sites)
    st = eng.run(4000)
    ref.run(4000)
    rs = ref.state()
    assert st["events"] == rs["events"] == 4000
    assert np.array_equal(eng.download(), ref.sites())
    assert np.array_equal(_bits(st["evolution"]), _bits(rs["evolution"]))
    eng.close()


def test_run_with_int64_max_event_cap():
    # INT64_MAX as max_events means "no cap": the run ends with the time loop.
    say("test_run_with_int64_max_event_cap")
    cfg = make_cfg(n=16, va=12, T=600.0, sector_T=3e-7, physics=5 * 3e-7)
    eng = akmc.Engine(cfg)
    st = eng.run((1 << 63) - 1)
    ref = refcpu.RefCpu(cfg)
    ref.run(10 ** 9)
    rs = ref.state()
    assert st["stop"] == akmc.AKMC_STOP_TIME and st["events"] == rs["events"] > 0
    assert np.array_equal(eng.download(), ref.sites())
    eng.close()


def test_engines_in_host_threads_are_independent():
    # One engine per host thread, each on its own stream: the same results as
    # the engines run one after another, and as the oracle.
    from concurrent.futures import ThreadPoolExecutor
    say("test_engines_in_host_threads_are_independent")
    cfgs = [make_cfg(n=16, va=10 + k, T=560.0 + 10 * k, seeds=dict(create_types=40 + k, create_vacancy=50 + k,
                                                                     time_inc=60 + k)) for k in range(6)]
    # Two crowded lattices (~2,000 vacancies per sector) whose sector state
    # lives in device memory: their engines allocate it during the run, while
    # the other threads' kernels are in flight.
    cfgs += [make_cfg(n=32, va=16000, T=600.0 + 10 * k, seeds=dict(create_types=70 + k, create_vacancy=80 + k,
                                                                     time_inc=90 + k)) for k in range(2)]
    events = [3000] * 6 + [300] * 2

    def run(arg):
        c, n = arg
        e = akmc.Engine(c)
        st = e.run(n)
        s = e.download()
        e.close()
        return st["events"], s

    with ThreadPoolExecutor(len(cfgs)) as ex:
        par = list(ex.map(run, zip(cfgs, events)))
    for c, want, (n, s) in zip(cfgs, events, par):
        ref = refcpu.RefCpu(c)
        ref.run(want)
        assert n == ref.state()["events"] == want
        assert np.array_equal(s, ref.sites())
