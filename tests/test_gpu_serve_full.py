"""Serving jobs across config 4's whole edge range (32 to 320 cells,
vacancies 1e-4 to 5e-4 of the sites): every job reaches its target with its
vacancies and Cu atoms conserved, and the final lattice of the 320^3 job has
the oracle's sector rates (within the FP32-rate budget of 1e-6) when the
replica engine that served it evaluates them."""
import numpy as np
import pytest

import akmc
import refcpu
from conftest import make_cfg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _counts(s):
    return [int((s & 1).sum()), int(((s >> 1) & 1).sum())]


def _job(L, va_frac, i):
    return make_cfg(n=L, va=max(1, int(round(va_frac * 2 * L ** 3))), T=500.0 + 40 * i,
                    alloy=[("Fe", 9700 - 30 * i), ("Cu", 60 + 20 * i), ("Ni", 120), ("Mn", 120 + 10 * i)],
                    seeds=dict(create_types=900 + i, create_vacancy=910 + i, time_inc=920 + i),
                    physics=1e3)


JOBS = [(32, 1e-4), (96, 5e-4), (160, 3e-4), (240, 2e-4), (320, 5e-4)]


def test_served_jobs_across_the_edge_range():
    cfgs = [_job(L, f, i) for i, (L, f) in enumerate(JOBS)]
    target = 20000
    say("served jobs, edges", [c["box"][0] for c in cfgs])
    srv = akmc.Server(keep_results=1)
    ids = [srv.submit(c, target) for c in cfgs]
    srv.wait_idle(300)
    st = srv.stats()
    assert st["done"] == len(cfgs) and st["failed"] == 0
    last = None
    for j, c in zip(ids, cfgs):
        info = srv.info(j)
        assert info["status"] == akmc.JOB_DONE and info["events"] == target, info
        n = 2 * c["box"][0] ** 3
        b = akmc.Batch([c], init=akmc.AKMC_INIT_DEVICE)   # the job's initial lattice (same law)
        s0 = b.download(0)
        b.close()
        s1 = srv.result(j, n)
        assert _counts(s1) == _counts(s0), c["box"]
        assert ((s0 & 1) != (s1 & 1)).any(), c["box"]
        say("  edge", c["box"][0], "vacancies", _counts(s1)[0], "conserved")
        last = (c, s1)
    srv.close()
    c, s1 = last
    say("  320^3 job: sector rates of its final lattice vs the oracle")
    b = akmc.Batch([c], precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_DEVICE)
    b.upload(0, s1)
    ref = refcpu.RefCpu(c, create=False, list_sweep=True)
    ref.set_sites(s1)
    for s in (0, 5):
        gs, gr, gt = b.sector_rates(0, s, cap=1 << 16)
        rs, rr, rt = ref.sector_rates(s, cap=1 << 16)
        o = np.argsort(gs)
        assert np.array_equal(gs[o], rs), s
        g = gr[o]
        assert np.array_equal(g == 0, rr == 0), s
        nz = rr > 0
        rel = np.abs(g[nz] - rr[nz]) / rr[nz]
        say("    sector", s, "rates", int(nz.sum()), "max rel error %.3g" % rel.max())
        assert rel.max() <= 1e-6, (s, rel.max())
        assert gt == pytest.approx(rt, rel=1e-6), s
    b.close()
