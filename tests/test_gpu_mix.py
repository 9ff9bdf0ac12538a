"""Vacancies and dumbbells on the GPU (akmc_mix_*) against the oracle's
restatement of the abvi model's event loop (refmix), bit for bit: every rate
slot of every sector, every event's (from, to, kind), the final typed
lattice, the dumbbell list with its orientations, the sector clocks."""
import numpy as np
import pytest

import akmc
import refcpu
from akmc import workloads
from conftest import make_cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _lattice(box, seed, f_itl, f_v, solutes=(0.03, 0.02, 0.02)):
    t, s, o = workloads.typed_lattice(box, seed, f_itl, solutes=solutes, f_v=f_v)
    return t, s, o


def _pair(box, t, s, o, **kw):
    """The same lattice on the GPU (host MT streams) and in the oracle."""
    p = akmc.mix_params(box=box, **kw)
    m = akmc.Mix([(p, t, s, o)], device=0)
    r = refcpu.RefMix(box, t, s, o, nu=p.attempt_freq, T=p.temperature, sector_T=p.sector_time,
                      physics_time=p.physics_time, ef110=p.ef110, gen_rate=p.defect_gen_rate,
                      seed_time_inc=p.seed_time_inc, seed_rec=p.seed_recombine)
    return m, r


RATE_CASES = [((8, 8, 8), 1, 0.02, 0.01), ((10, 9, 12), 2, 0.005, 0.003), ((16, 16, 16), 3, 0.002, 0.01),
              ((9, 13, 8), 4, 0.03, 0.0)]


@pytest.mark.parametrize("box,seed,f_itl,f_v", RATE_CASES)
def test_sector_rates_bit_exact(box, seed, f_itl, f_v):
    say("mix sector rates", box, seed, f_itl, f_v)
    t, s, o = _lattice(box, seed, f_itl, f_v)
    m, r = _pair(box, t, s, o, temperature=600.0)
    for sec in range(8):
        gs, gk, gr, gt = m.sector_rates(0, sec)
        rs, rk, rr, rt = r.sector_rates(sec)
        assert np.array_equal(gs, rs) and np.array_equal(gk, rk), sec
        assert np.array_equal(_bits(gr), _bits(rr)), (sec, np.argwhere(_bits(gr) != _bits(rr))[:4])
        assert _bits([gt])[0] == _bits([rt])[0], sec
    m.close()


def _compare(box, t, s, o, n, pieces=1, **kw):
    m, r = _pair(box, t, s, o, **kw)
    m.set_event_log(n)
    done = 0
    for k in range(pieces):
        step = n // pieces if k < pieces - 1 else n - done
        m.run(step)
        done += step
    rc, (rf, rt, rk) = r.run(n, log=True)
    gs, rs = m.stats(0), r.state()
    gf, gt, gk = m.event_log(0, n)
    assert gs["events"] == rs["events"], (gs["events"], rs["events"])
    k = rs["events"]
    bad = np.nonzero((gf[:k] != rf[:k]) | (gt[:k] != rt[:k]) | (gk[:k] != rk[:k]))[0]
    assert bad.size == 0, "first divergence at event %d" % bad[0]
    tg, sg, og = m.download(0)
    assert np.array_equal(tg, r.types())
    sr, orr = r.itl()
    assert np.array_equal(sg, sr) and np.array_equal(og, orr)
    assert np.array_equal(_bits(gs["evolution"]), _bits(rs["evolution"]))
    assert gs["sector_time"] == rs["sector_time"] and gs["step"] == rs["step"]
    assert gs["n_recombined"] == rs["recombined"] and gs["nvac"] == rs["n_vac"] and gs["nitl"] == rs["n_itl"]
    m.close()
    return gs, rk


TRAJ_CASES = [
    ("dense_small", (8, 8, 8), 11, 0.01, 0.01, dict(temperature=600.0, sector_T=2e-9), 3000),
    ("moderate", (12, 10, 9), 12, 0.004, 0.004, dict(temperature=573.0, sector_T=5e-9), 3000),
    ("dumbbells_only", (9, 9, 10), 13, 0.02, 0.0, dict(temperature=560.0, sector_T=1e-9), 2000),
    ("vacancies_only", (10, 10, 10), 14, 0.0, 0.01, dict(temperature=650.0, sector_T=3e-7), 3000),
    ("generation", (8, 9, 8), 15, 0.0, 0.002, dict(temperature=600.0, sector_T=1e-9, defect_gen_rate=5e9), 1500),
    ("long_list", (16, 12, 16), 16, 0.006, 0.002, dict(temperature=620.0, sector_T=2e-9), 1500),
]


@pytest.mark.parametrize("name,box,seed,f_itl,f_v,kw,n", TRAJ_CASES, ids=[c[0] for c in TRAJ_CASES])
def test_trajectory_bit_exact(name, box, seed, f_itl, f_v, kw, n):
    say("mix trajectory", name, n)
    t, s, o = _lattice(box, seed, f_itl, f_v)
    if f_itl == 0.0:
        t[s] = 1
        s, o = s[:0], o[:0]
    kw = dict(kw)
    kw["sector_time"] = kw.pop("sector_T")
    st, kinds = _compare(box, t, s, o, n, pieces=3, **kw)
    if name == "generation":
        assert (kinds == 2).sum() > 0 and st["n_generated"] > 0
    elif name != "vacancies_only":
        assert (kinds == 1).sum() > 0


def _sparse_lattice(box, seed, n_vac, n_itl):
    """Fe with 1 % Cu and 1 % Ni, n_vac vacancies and n_itl Fe dumbbells at
    random sites: once the vacancies have recombined, the dumbbells sit in
    configurations whose rates differ by ten orders of magnitude, so sector
    clocks run far ahead and long stretches of time steps pass without an
    event."""
    n = 2 * box[0] * box[1] * box[2]
    rng = np.random.default_rng(seed)
    t = np.ones(n, np.uint16)
    u = rng.random(n)
    t[u < 0.02] = 4
    t[u < 0.01] = 2
    idx = rng.choice(n, n_vac + n_itl, replace=False)
    t[idx[:n_vac]] = 0
    s = np.sort(idx[n_vac:])
    for q in s:
        a = int(t[q])
        t[q] = (a << 8) | 1 if a <= 1 else (1 << 8) | a
    o = rng.choice([-6, -5, -4, -3, -2, -1, 1, 2, 3, 4, 5, 6], s.size).astype(np.int8)
    return t, s, o


@pytest.mark.parametrize("physics_time,n", [(1e3, 3000), (2e-4, 5000)], ids=["events", "time_end"])
def test_idle_steps_taken_in_one_go_match(physics_time, n):
    """Time steps with no possible event are passed in one go on the device;
    the oracle walks every sector of every step.  Same events, clocks and
    step counts bit for bit, also where the time loop ends among them."""
    say("mix idle steps", physics_time, n)
    box = (12, 12, 12)
    t, s, o = _sparse_lattice(box, 1, 7, 20)
    st, kinds = _compare(box, t, s, o, n, pieces=3, temperature=650.0, sector_time=2e-9, physics_time=physics_time)
    assert st["n_skipped_steps"] > 10000
    if physics_time < 1:
        assert st["events"] < n and st["stop"] == akmc.AKMC_STOP_TIME


def test_recombination_happens_and_matches():
    say("mix recombination")
    box = (8, 8, 8)
    t, s, o = _lattice(box, 21, 0.015, 0.015)
    st, kinds = _compare(box, t, s, o, 2500, temperature=640.0, sector_time=2e-9)
    assert st["n_recombined"] > 0


def test_lattices_in_one_launch_equal_one_at_a_time():
    say("mix batch independence")
    cases = [((8, 8, 8), 31, 0.01, 0.01), ((10, 9, 12), 32, 0.004, 0.004), ((9, 9, 9), 33, 0.0, 0.01),
             ((12, 8, 10), 34, 0.02, 0.0)]
    lats = []
    for box, seed, f_itl, f_v in cases:
        t, s, o = _lattice(box, seed, f_itl, f_v)
        if f_itl == 0.0:
            t[s] = 1
            s, o = s[:0], o[:0]
        lats.append((akmc.mix_params(box=box, temperature=600.0, sector_time=2e-9, seed_time_inc=seed), t, s, o))
    both = akmc.Mix(lats)
    both.set_event_log(1500)
    both.run(1500)
    for i, lat in enumerate(lats):
        one = akmc.Mix([lat])
        one.set_event_log(1500)
        one.run(1500)
        for a, b in zip(one.event_log(0, 1500), both.event_log(i, 1500)):
            assert np.array_equal(a, b), i
        for a, b in zip(one.download(0), both.download(i)):
            assert np.array_equal(a, b), i
        one.close()
    both.close()


DUMB = {0x0101, 0x0102, 0x0104, 0x0108, 0x0202, 0x0204, 0x0208, 0x0404, 0x0408, 0x0808}


def _atoms(t):
    c = {1: 0, 2: 0, 4: 0, 8: 0}
    for v, cnt in zip(*np.unique(t, return_counts=True)):
        v = int(v)
        if v in DUMB:
            c[v >> 8] += int(cnt)
            c[v & 0xFF] += int(cnt)
        elif v:
            c[v] += int(cnt)
    return c


def test_device_streams_conserve_and_reproduce():
    """xoshiro streams on the device (the serving form): atoms conserved,
    lists consistent with the lattice, the same result twice."""
    say("mix device streams")
    box = (16, 16, 16)
    t, s, o = _lattice(box, 41, 0.002, 0.004)
    p = akmc.mix_params(box=box, temperature=600.0, sector_time=1e-8, device_rng=1)
    res = []
    for _ in range(2):
        m = akmc.Mix([(p, t, s, o)])
        m.run(20000)
        res.append((m.download(0), m.stats(0)))
        m.close()
    (t1, s1, o1), st = res[0]
    assert st["events"] == 20000 and st["n_itl_events"] > 0 and st["n_recombined"] > 0
    assert all(np.array_equal(a, b) for a, b in zip(res[0][0], res[1][0]))
    assert _atoms(t1) == _atoms(t)
    assert set(np.nonzero(np.isin(t1, list(DUMB)))[0]) == set(s1.tolist()) and np.all(o1 != 0)
    assert st["nvac"] == int((t1 == 0).sum())
