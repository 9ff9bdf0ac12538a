"""Config 2's lattice on the GPU (512^3 cells, Fe-1.34Cu-0.70Ni-1.30Mn,
2e-4 vacancies, 563 K) against the oracle's fixture
(tests/golden/make_golden.py cfg2): the FP64 engine's sector rates and first
200 events bit for bit (its sectors hold ~6,700 vacancies, so it runs the
device-memory sector form), and the FP32-rate replica engine's rates of the
same sector within the 1e-6 relative budget."""
import hashlib
import json
import os

import numpy as np
import pytest

import akmc

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "cfg2_512cube.npz")


def say(*a):
    print(*a, flush=True)


def _sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _fixture():
    z = np.load(FIXTURE, allow_pickle=False)
    return z, json.loads(str(z["cfg"]))


def test_plain_fp64_sector_rates_and_first_events():
    z, cfg = _fixture()
    say("config 2: FP64 engine, box", cfg["box"], "vacancies", cfg["va_count"])
    eng = akmc.Engine(cfg)
    gs, gr, gt = eng.sector_rates(int(z["sector"]), cap=1 << 15)
    assert gs.size == int(z["n_rows"])
    assert np.array_equal(gs[:len(z["head_sites"])], z["head_sites"])
    assert np.array_equal(_bits(gr[:len(z["head_rates"])]), _bits(z["head_rates"]))
    assert np.array_equal(_sha(gs), z["sites_sha256_rows"]) and np.array_equal(_sha(gr), z["rates_sha256"])
    assert _bits([gt])[0] == _bits([float(z["total"])])[0]
    n = int(z["events"])
    eng.set_event_log(n)
    st = eng.run(n)
    assert st["events"] == n
    gf, gt2 = eng.event_log(n)
    log = z["log"]
    bad = np.nonzero((gf != log[:, 0]) | (gt2 != log[:, 1]))[0]
    assert bad.size == 0, "first divergence at event %d" % bad[0]
    assert np.array_equal(_sha(eng.download()), z["final_sha256"])
    assert np.array_equal(_bits(st["evolution"]), _bits(z["evolution"]))
    assert st["sector_time"] == float(z["sector_time"]) and st["step"] == int(z["step"])
    assert st["sector_pos"] == int(z["sect"])
    eng.close()


def test_fp32_replica_rates_within_budget():
    z, cfg = _fixture()
    say("config 2: FP32-rate replica engine")
    b = akmc.Batch([cfg], precision=akmc.AKMC_PREC_FP32, init=akmc.AKMC_INIT_REFERENCE)
    gs, gr, gt = b.sector_rates(0, int(z["sector"]), cap=1 << 15)
    o = np.argsort(gs)
    gs, gr = gs[o], gr[o]
    assert np.array_equal(_sha(gs), z["sites_sha256_rows"])
    rr = z["head_rates"]
    g = gr[:rr.shape[0]]
    assert np.array_equal(g == 0, rr == 0)
    nz = rr > 0
    rel = np.abs(g[nz] - rr[nz]) / rr[nz]
    say("  max relative error over %d rates: %.3g" % (nz.sum(), rel.max()))
    assert rel.max() <= 1e-6 and rel.max() <= 2.0 ** -22
    assert gt == pytest.approx(float(z["total"]), rel=1e-6)
    b.close()
