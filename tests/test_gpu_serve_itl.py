"""Jobs with self-interstitial dumbbells in the serving loop
(akmc_server_submit_itl / akmc_server_result_typed): the job's initial
lattice is the replica engine's lattice for the same composition and seeds
with the dumbbells added at atom sites; a job's result is the same alone and
in a stream with vacancy jobs and other dumbbell jobs under an HBM budget;
atoms of each species are conserved."""
import numpy as np
import pytest

import akmc
from conftest import make_cfg

pytestmark = pytest.mark.gpu

DUMB = {0x0101, 0x0102, 0x0104, 0x0108, 0x0202, 0x0204, 0x0208, 0x0404, 0x0408, 0x0808}


def say(*a):
    print(*a, flush=True)


def _typed_of_flags(f):
    """Replica-engine site flags (bit0 V, bit1 Cu, bit2 Mn, bit3 Ni) as
    LatticeTypes values (V 0, Fe 1, Cu 2, Ni 4, Mn 8)."""
    t = np.ones(f.size, np.uint16)
    t[(f & akmc.BIT_NI) != 0] = 4
    t[(f & akmc.BIT_MN) != 0] = 8
    t[(f & akmc.BIT_CU) != 0] = 2
    t[(f & akmc.BIT_V) != 0] = 0
    return t


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


def _cfg(L, i, T=600.0, va_frac=2e-3, sector_T=2e-9):
    rng = np.random.default_rng(700 + i)
    cu, ni, mn = rng.uniform(10, 150), rng.uniform(20, 350), rng.uniform(30, 150)
    return make_cfg(n=L, va=max(1, int(round(va_frac * 2 * L ** 3))), T=T,
                    alloy=[("Fe", int(10000 - cu - ni - mn)), ("Cu", int(cu)), ("Ni", int(ni)), ("Mn", int(mn))],
                    seeds=dict(create_types=100 + i, create_vacancy=200 + i, time_inc=300 + i, event_selection=400 + i),
                    physics=1e3, sector_T=sector_T)


def _nsites(c):
    return 2 * c["box"][0] * c["box"][1] * c["box"][2]


def _served_alone(c, n_itl, target, atom=1):
    srv = akmc.Server(keep_results=1)
    j = srv.submit_itl(c, n_itl, target, itl_atom=atom)
    srv.wait_idle(120)
    info = srv.info(j)
    out = srv.result_typed(j, _nsites(c))
    srv.close()
    return info, out


@pytest.mark.parametrize("atom", [1, 2, 8])
def test_initial_lattice_follows_the_replica_composition(atom):
    say("served dumbbell job: initial lattice, atom", atom)
    c = _cfg(14, 1)
    n_itl = 25
    info, t0 = _served_alone(c, n_itl, 0, atom)
    assert info["status"] == akmc.JOB_DONE and info["events"] == 0
    b = akmc.Batch([c], init=akmc.AKMC_INIT_DEVICE)
    base = _typed_of_flags(b.download(0))
    b.close()
    d = np.nonzero(np.isin(t0, list(DUMB)))[0]
    assert d.size == n_itl
    rest = np.ones(t0.size, bool)
    rest[d] = False
    assert np.array_equal(t0[rest], base[rest])
    a = base[d]
    assert np.all(a != 0)
    want = np.where(a <= atom, (a.astype(np.uint32) << 8) | atom, (atom << 8) | a.astype(np.uint32))
    assert np.array_equal(t0[d], want.astype(np.uint16))


def test_served_alone_equals_served_in_a_mixed_stream():
    say("served dumbbell jobs: alone vs mixed stream")
    jobs = [(_cfg(12, 11), 20, 3000, 1), (_cfg(16, 12), 40, 6000, 2), (_cfg(10, 13), 8, 1500, 1),
            (_cfg(14, 14, T=650.0), 30, 4000, 4)]
    alone = []
    for c, n_itl, target, atom in jobs:
        say("  alone", c["box"], n_itl, target, atom)
        info, t = _served_alone(c, n_itl, target, atom)
        assert info["status"] == akmc.JOB_DONE and info["events"] == target, info
        _, t0 = _served_alone(c, n_itl, 0, atom)
        assert _atoms(t) == _atoms(t0)
        assert int((t == 0).sum()) <= int((t0 == 0).sum())   # recombination only removes vacancies
        alone.append(t)
    vac = [_cfg(int(L), 20 + k, va_frac=4e-4, sector_T=3e-7) for k, L in enumerate([12, 20, 16, 24, 10, 18, 22, 14])]
    per_job = max(2 * c["box"][0] ** 3 for c in vac) * 3
    srv = akmc.Server(keep_results=1, hbm_budget=int(per_job * 4), slice_events=700)
    ids = []
    k = 0
    for i, c in enumerate(vac):
        srv.submit(c, 5000)
        if i % 2 == 0 and k < len(jobs):
            c2, n_itl, target, atom = jobs[k]
            ids.append(srv.submit_itl(c2, n_itl, target, itl_atom=atom))
            srv.submit_itl(_cfg(12, 40 + i), 15, 2500)
            k += 1
    say("  mixed stream of", len(vac) + 2 * len(jobs), "jobs")
    srv.wait_idle(300)
    st = srv.stats()
    assert st["failed"] == 0 and st["done"] == len(vac) + 2 * len(jobs)
    assert st["bytes_peak"] <= int(per_job * 4)
    for j, (c, *_), a in zip(ids, jobs, alone):
        assert srv.info(j)["status"] == akmc.JOB_DONE
        assert np.array_equal(srv.result_typed(j, _nsites(c)), a)
    srv.close()


def test_bad_dumbbell_jobs_are_refused():
    say("served dumbbell job: bad arguments")
    srv = akmc.Server(keep_results=1, hbm_budget=1 << 22)
    with pytest.raises(akmc.AkmcError):
        srv.submit_itl(_cfg(12, 1), 10, 100, itl_atom=3)
    with pytest.raises(akmc.AkmcError):
        srv.submit_itl(_cfg(12, 1), -1, 100)
    with pytest.raises(akmc.AkmcError):
        srv.submit_itl(_cfg(12, 1), 2 * 12 ** 3, 100)     # more dumbbells than the box holds: refused at once
    j0 = srv.submit_itl(_cfg(10, 3), 5, 50)              # and the server still serves
    srv.wait_idle(60)
    assert srv.info(j0)["status"] == akmc.JOB_DONE
    with pytest.raises(akmc.AkmcError) as e:
        srv.submit_itl(make_cfg(n=200, va=10), 10, 10)
    assert e.value.code == -6
    j = srv.submit(_cfg(10, 2), 100)
    srv.wait_idle(60)
    with pytest.raises(akmc.AkmcError):
        srv.result_typed(j, _nsites(_cfg(10, 2)))     # a vacancy job keeps no typed result
    srv.close()
