"""The oracle reproduces its committed golden trajectories (regression pin of
oracle/refcpu.cpp; the same fixtures are the GPU path's targets)."""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import refcpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "traj_*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    return json.loads(str(z["cfg"])), z


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_matches_golden(path):
    cfg, z = load(path)
    n = int(z["events"])
    r = refcpu.RefCpu(cfg)
    rc, lf, lt = r.run(n, log=True)
    st = r.state()
    assert st["events"] == n
    assert np.array_equal(lf[:n], z["from_"]) and np.array_equal(lt[:n], z["to"])
    assert hashlib.sha256(r.sites().tobytes()).hexdigest() == str(z["sites_sha256"])
    assert np.array_equal(st["evolution"], z["evolution"])
    assert st["sector_time"] == float(z["sector_time"])


def test_oracle_invariants():
    from conftest import make_cfg
    cfg = make_cfg(n=16, va=60, alloy=[("Fe", 6000), ("Mn", 2000), ("Ni", 1500), ("Cu", 500)])
    r = refcpu.RefCpu(cfg)
    s0 = r.sites()
    rc, lf, lt = r.run(2000, log=True)
    s1 = r.sites()
    assert (s1 & 1).sum() == (s0 & 1).sum()                 # vacancies conserved
    assert ((s1 >> 1) & 1).sum() == ((s0 >> 1) & 1).sum()   # Cu conserved (never left stale)
    X, Y, Z = 32, 16, 16
    fx, fy, fz = lf % X, (lf // X) % Y, lf // (X * Y)
    tx, ty, tz = lt % X, (lt // X) % Y, lt // (X * Y)
    dx = (tx - fx + X // 2) % X - X // 2
    dy = (ty - fy + Y // 2) % Y - Y // 2
    dz = (tz - fz + Z // 2) % Z - Z // 2
    assert np.all(np.abs(dx) == 1) and np.all(np.abs(dy) <= 1) and np.all(np.abs(dz) <= 1)  # 1nn hops


def test_oracle_matches_cfg0_hashes_prefix():
    """The first 20,000 events of the config-0 fixture (1e5 events, hashed
    every 1000) reproduce on the oracle."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "hashes_cfg0_40cube.npz"), allow_pickle=False)
    r = refcpu.RefCpu(json.loads(str(z["cfg"])))
    every = int(z["every"])
    for k in range(20):
        r.run(every)
        got = np.frombuffer(hashlib.sha256(r.sites().tobytes()).digest(), dtype=np.uint8)
        assert np.array_equal(got, z["sha256"][k]), k
