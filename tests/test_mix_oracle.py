"""The vacancy + dumbbell event loop of the oracle (refmix), pinned on the CPU.

* orientation::trans against the reference's unit test
  (tests/unit-abvi/orientation_test.cpp:52-62) and the table's own structure;
* with no dumbbells and no stale tags, refmix is the vacancy-only oracle
  (refcpu, the reference CPU build restated) event for event;
* dumbbell rate slots equal the rate step's oracle (refitl, pinned by the
  reference's Itl / orientation unit tests) on the same lattice;
* closed forms: a lone FeFe dumbbell in pure iron walks with the constant
  total 4 nu exp(-0.32 / kT); a vacancy that reaches a dumbbell recombines
  into two single atoms (high end to the vacancy site);
* invariants over long mixed runs: atoms conserved, defect lists consistent
  with the lattice, orientations valid, vacancies and dumbbells removed in
  pairs only.
"""
import math

import numpy as np
import pytest

import refcpu
from akmc import workloads
from conftest import make_cfg

KB = 8.617330350e-5
DUMB = {0x0101, 0x0102, 0x0104, 0x0108, 0x0202, 0x0204, 0x0208, 0x0404, 0x0408, 0x0808}


def test_trans_matches_reference_unit_test():
    # orientation_test.cpp:52-62
    assert refcpu.mix_trans(1, 0, True, True) == -3   # s_110.trans(0, true, true) == r_011
    assert refcpu.mix_trans(1, 1, True, True) == -4   # == r_01_1
    assert refcpu.mix_trans(3, 3, True, False) == -2  # s_011.trans(3, true, false) == r_1_10


def test_trans_table_structure():
    """Every direction an orientation can jump along has a successor; flipping
    either flag negates it (the table's rows are {{a, b}, {-a, -b}}), and
    the successor is never the unknown orientation."""
    for o in range(-6, 7):
        if o == 0:
            continue
        dirs = refcpu.lib().ref_ori_avail(o, 0)
        for tag in range(8):
            vals = [refcpu.mix_trans(o, tag, h, r) for h in (True, False) for r in (True, False)]
            if (dirs >> tag) & 1:
                assert 0 not in vals, (o, tag)
                assert vals[0] == -vals[2] and vals[1] == -vals[3], (o, tag, vals)
                assert abs(vals[0]) != abs(o) and abs(vals[1]) != abs(o)
            else:
                assert vals == [0, 0, 0, 0]


def _typed_from_flags(sites):
    """Typed lattice of a flag-byte lattice without stale tags."""
    t = np.full(sites.size, 1, np.uint16)
    t[(sites & 8) != 0] = 4
    t[(sites & 4) != 0] = 8
    t[(sites & 2) != 0] = 2
    t[(sites & 1) != 0] = 0
    return t


@pytest.mark.parametrize("box,va,T", [((8, 8, 8), 20, 600.0), ((9, 11, 8), 40, 650.0), ((12, 12, 12), 30, 573.0)])
def test_no_dumbbells_equals_the_vacancy_oracle(box, va, T):
    cfg = make_cfg(box=box, va=va, T=T, sector_T=3e-7, physics=1.0,
                   seeds=dict(create_types=3, create_vacancy=4, time_inc=5))
    a = refcpu.RefCpu(cfg)
    s = a.sites()
    s = np.where(s & 1, 1, s).astype(np.uint8)  # drop stale tags under vacancies
    a.set_sites(s)
    m = refcpu.RefMix(box, _typed_from_flags(s), [], [], T=T, sector_T=3e-7, physics_time=1.0, seed_time_inc=5)
    rc, fa, ta = a.run(3000, log=True)
    rm, (fm, tm, km) = m.run(3000, log=True)
    sa, sm = a.state(), m.state()
    assert sa["events"] == sm["events"] == 3000
    assert np.array_equal(fa, fm) and np.array_equal(ta, tm) and np.all(km == 0)
    assert np.array_equal(sa["evolution"], sm["evolution"]) and sa["sector_time"] == sm["sector_time"]
    assert np.array_equal(_typed_from_flags(a.sites()), m.types())


def test_dumbbell_slots_equal_the_rate_oracle():
    box = (8, 8, 8)
    t, sites, orient = workloads.typed_lattice(box, 7, 0.02, solutes=(0.05, 0.05, 0.05), f_v=0.01)
    m = refcpu.RefMix(box, t, sites, orient, T=563.0)
    av, r = refcpu.Itl().rates(box, t, sites, orient, 6e12, 563.0)
    seen = 0
    for sec in range(8):
        ss, kk, rr, tot = m.sector_rates(sec)
        for site, k, row in zip(ss, kk, rr):
            if k == 1:
                j = int(np.searchsorted(sites, site))
                assert sites[j] == site
                assert np.array_equal(row.view(np.uint64), r[j].view(np.uint64))
                seen += 1
        # The total folds every slot in sweep order.
        acc = 0.0
        for row in rr:
            for v in row:
                acc += v
        assert acc + 0.0 == tot
    assert seen == len(sites)


def test_lone_fefe_dumbbell_walks_with_a_constant_total():
    box = (8, 8, 8)
    n = 2 * 8 * 8 * 8
    t = np.full(n, 1, np.uint16)
    s0 = 3 + 16 * (4 + 8 * 4)
    t[s0] = 0x0101
    T = 600.0
    m = refcpu.RefMix(box, t, [s0], [1], T=T, sector_T=1e-8, physics_time=1.0)
    rc, (f, to, k) = m.run(400, log=True)
    assert np.all(k == 1) and len(f) == 400
    # Every event is a 1nn step of the one dumbbell, which stays FeFe.
    assert np.array_equal(f[1:], to[:-1]) and f[0] == s0
    st = m.state()
    assert st["n_itl"] == 1
    ts = m.types()
    assert int((ts == 0x0101).sum()) == 1 and int((ts == 1).sum()) == n - 1
    sites, ori = m.itl()
    assert sites[0] == to[-1] and ori[0] != 0
    # Each sector holding the dumbbell has total 8 slots x rate / 2 = 4 rate.
    rate = 6e12 * math.exp(-0.32 / (KB * T))
    sec = [s for s in range(8) if len(m.sector_rates(s)[0])][0]
    ss, kk, rr, tot = m.sector_rates(sec)
    assert np.allclose(rr[0], rate / 2, rtol=0, atol=0) and tot == sum([rate / 2] * 8)


def test_vacancy_meets_dumbbell_and_recombines():
    """A FeCu dumbbell with a vacancy two 1nn steps away in pure iron: the
    first event that brings them within 2nn recombines them: the vacancy site
    takes the high end (Fe) and the dumbbell site keeps the low end (Cu)."""
    box = (8, 8, 8)
    X = 16
    n = 2 * 8 * 8 * 8
    seen = 0
    for seed in range(20):
        t = np.full(n, 1, np.uint16)
        d = 6 + X * (4 + 8 * 4)
        v = 6 + 4 + X * (4 + 8 * 4)  # (x + 4, y, z): 2 cells away in x, not within 2nn
        t[d] = 0x0102
        t[v] = 0
        m = refcpu.RefMix(box, t, [d], [3], T=700.0, sector_T=1e-9, physics_time=1.0, seed_time_inc=seed)
        for _ in range(200):
            rc, _ = m.run(1)
            if m.state()["recombined"]:
                break
        st = m.state()
        if not st["recombined"]:
            continue
        seen += 1
        ts = m.types()
        assert st["n_itl"] == 0 and st["n_vac"] == 0
        assert int((ts == 2).sum()) == 1 and int((ts == 1).sum()) == n - 1  # Cu conserved, no defects left
    assert seen >= 5


def _atoms(t):
    c = {1: 0, 2: 0, 4: 0, 8: 0}
    for v, cnt in zip(*np.unique(t, return_counts=True)):
        v = int(v)
        if v == 0:
            continue
        if v in DUMB:
            c[v >> 8] += cnt
            c[v & 0xFF] += cnt
        else:
            c[v] += cnt
    return c


@pytest.mark.parametrize("seed,f_itl", [(11, 0.004), (12, 0.01)])
def test_mixed_run_invariants(seed, f_itl):
    box = (10, 9, 8)
    t, sites, orient = workloads.typed_lattice(box, seed, f_itl, solutes=(0.03, 0.02, 0.02), f_v=0.01)
    m = refcpu.RefMix(box, t, sites, orient, T=600.0, sector_T=2e-9, physics_time=1.0, seed_time_inc=seed)
    a0 = _atoms(t)
    v0, i0 = int((t == 0).sum()), len(sites)
    rc, (f, to, k) = m.run(3000, log=True)
    assert rc == 1 and len(f) == 3000
    assert set(np.unique(k)) <= {0, 1} and (k == 1).sum() > 20 and (k == 0).sum() > 10
    ts = m.types()
    st = m.state()
    s_itl, o_itl = m.itl()
    assert _atoms(ts) == a0
    assert st["n_vac"] == int((ts == 0).sum()) and st["n_itl"] == len(s_itl)
    assert set(np.nonzero(np.isin(ts, list(DUMB)))[0]) == set(s_itl.tolist())
    assert np.all(o_itl != 0) and np.all(np.abs(o_itl) <= 6)
    assert v0 - st["n_vac"] == i0 - st["n_itl"] == st["recombined"]


def test_random_lattice_places_dumbbells_as_the_reference_driver():
    """akmc_mix_random_lattice = creation::createRandom's lattice, then
    placeDumbbells restated in the oracle, on the same mt19937 draws."""
    import akmc
    cfg = make_cfg(box=(12, 10, 9), va=30, seeds=dict(create_types=8, create_vacancy=9, time_inc=10))
    t, s, o = akmc.mix_random_lattice(cfg, itl=(("Fe", 40), ("Cu", 7), ("Mn", 3)), seed=77)
    flags = refcpu.RefCpu(cfg).sites()
    base = np.where(flags & 1, 0, np.where(flags & 2, 2, np.where(flags & 4, 8, np.where(flags & 8, 4, 1))))
    t2, s2, o2 = refcpu.mix_place(base, ["Fe", "Cu", "Mn"], [40, 7, 3], 77)
    order = np.argsort(s2, kind="stable")
    assert np.array_equal(t, t2) and np.array_equal(s, s2[order]) and np.array_equal(o, o2[order])
    assert 40 < len(s) <= 50 and np.all(np.diff(s) > 0)


def test_defect_generation_invariants():
    box = (8, 9, 8)
    t, sites, orient = workloads.typed_lattice(box, 5, 0.0, solutes=(0.02, 0.02, 0.02), f_v=0.0)
    t[sites] = 1  # no dumbbells, no vacancies: every event is a generation until defects appear
    m = refcpu.RefMix(box, t, [], [], T=600.0, sector_T=1e-9, physics_time=1.0, gen_rate=5e9)
    a0 = _atoms(t)
    rc, (f, to, k) = m.run(400, log=True)
    st = m.state()
    ts = m.types()
    assert (k == 2).sum() > 0 and (k == 1).sum() > 0
    assert _atoms(ts) == a0
    gen = int(((k == 2) & (to >= 0)).sum())
    assert st["n_vac"] == gen - st["recombined"] and st["n_itl"] == gen - st["recombined"]
