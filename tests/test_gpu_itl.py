"""Dumbbell (self-interstitial) rates on the GPU (akmc_itl_rates) against the
oracle's restatement of the reference's model, bit for bit."""
import math

import numpy as np
import pytest

import akmc
from akmc import workloads
from oracle import refcpu

pytestmark = pytest.mark.gpu

def typed_lattice(box, seed, f_itl, solutes=(0.01, 0.02, 0.01), f_v=2e-4):
    return workloads.typed_lattice(box, seed, f_itl, solutes=solutes, f_v=f_v)


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def _check(box, t, sites, orient, nu=6e12, T=573.0):
    print("itl case: box", box, "dumbbells", len(sites), flush=True)
    av_g, r_g = akmc.itl_rates(box, t, sites, orient, nu, T)
    av_o, r_o = refcpu.Itl().rates(box, t, sites, orient, nu, T)
    assert np.array_equal(av_g, av_o)
    assert _same(r_g, r_o), np.argwhere(r_g != r_o)[:5]
    return av_g, r_g


def test_fefe_in_pure_iron_closed_form():
    box = (5, 4, 3)
    t = np.full(2 * 5 * 4 * 3, 1, dtype=np.uint16)
    s = 2 + 10 * (1 + 4 * 1)
    t[s] = 0x0101
    for o in (1, -2, 3, -4, 5, -6):
        av, r = _check(box, t, [s], [o])
        assert list(r[0]) == [6e12 * math.exp(-0.32 / (8.617330350e-5 * 573.0)) / 2] * 8


@pytest.mark.parametrize("box,seed,f_itl", [((3, 3, 3), 1, 0.08), ((4, 5, 6), 2, 0.05), ((8, 8, 8), 3, 0.03),
                                            ((6, 4, 9), 4, 0.10)])
def test_random_lattices_bit_exact(box, seed, f_itl):
    t, sites, orient = typed_lattice(box, seed, f_itl, solutes=(0.05, 0.05, 0.05), f_v=0.02)
    av, r = _check(box, t, sites, orient, T=500.0 + 10 * seed)
    assert (r > 0).sum() > 0


def test_job_sized_lattice():
    # A serving job's box (L = 64, 524,288 sites) with 1e-4 dumbbells.
    box = (64, 64, 64)
    t, sites, orient = typed_lattice(box, 11, 1e-4)
    _check(box, t, sites, orient, T=650.0)


def test_many_dumbbells():
    box = (16, 16, 16)
    t, sites, orient = typed_lattice(box, 12, 0.02)
    assert len(sites) > 150
    _check(box, t, sites, orient)


def test_dense_dumbbells_many_chunks():
    # ~1,400 dumbbells (5 % of the sites): 22 chunks of the prefix fold and
    # long neighbour lists.
    box = (24, 24, 24)
    t, sites, orient = typed_lattice(box, 13, 0.05, solutes=(0.05, 0.05, 0.05), f_v=0.01)
    assert len(sites) > 1300
    _check(box, t, sites, orient, T=540.0)


def test_runs_beside_an_engine():
    # The call uses a stream of its own: an engine on the same device keeps
    # its trajectory when dumbbell rates are computed between its runs.
    from conftest import make_cfg
    cfg = make_cfg(n=16, va=12)
    a = akmc.Engine(cfg)
    a.run(3000)
    box = (8, 8, 8)
    t, sites, orient = typed_lattice(box, 21, 0.05)
    av1, r1 = akmc.itl_rates(box, t, sites, orient, 6e12, 573.0)
    a.run(3000)
    b = akmc.Engine(cfg)
    b.run(6000)
    assert np.array_equal(a.download(), b.download())
    av2, r2 = akmc.itl_rates(box, t, sites, orient, 6e12, 573.0)
    assert np.array_equal(av1, av2) and _same(r1, r2)
    a.close()
    b.close()


def test_golden_fixture():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "itl_dense_6x4x9.npz"))
    box = tuple(int(x) for x in g["box"])
    print("itl case: golden fixture", box, len(g["sites"]), flush=True)
    av, r = akmc.itl_rates(box, g["types"], g["sites"], g["orient"], float(g["nu"]), float(g["T"]))
    assert np.array_equal(av, g["avail"])
    assert _same(r, g["rates"])
