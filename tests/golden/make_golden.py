"""Writes the golden trajectory fixtures tests/golden/traj_*.npz from the
oracle (oracle/refcpu.cpp).  Each fixture holds the configuration, the first
events' (from, to) site indices, a SHA-256 of the final site array, the
sector clocks and the partial sector time.  Run from the repository root:
    python3 tests/golden/make_golden.py        # traj_*.npz
    python3 tests/golden/make_golden.py cfg0   # hashes_cfg0_40cube.npz (1e5 events, a few minutes)
    python3 tests/golden/make_golden.py itl    # itl_dense_6x4x9.npz (dumbbell rates)
    python3 tests/golden/make_golden.py cfg1 [resume_dir]   # hashes_cfg1_256cube.npz (1e5 events of 256^3, ~1 h)
    python3 tests/golden/make_golden.py cfg2   # cfg2_512cube.npz (512^3: sector rates, 200 events)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))
import refcpu  # noqa: E402
from conftest import make_cfg  # noqa: E402

CASES = {
    "cfg0_40cube": (make_cfg(n=40, va=26, T=573.0, sector_T=3e-7), 3000),
    "mnni_rich_16": (make_cfg(n=16, va=60, T=600.0, alloy=[("Fe", 6000), ("Mn", 2000), ("Ni", 1500), ("Cu", 500)],
                              seeds=dict(create_types=7, create_vacancy=11, time_inc=13)), 3000),
    "odd_box": (make_cfg(box=(17, 13, 11), va=20, T=520.0, sector_T=5e-7,
                         seeds=dict(create_types=21, create_vacancy=22, time_inc=23)), 2000),
}


# Config 0 of the benchmark suite at full length: 1e5 events of the 40^3
# box, with the occupancy hashed every 1000 events.
CFG0 = make_cfg(n=40, va=26, T=573.0, sector_T=3e-7, physics=1e3)
CFG0_EVENTS, CFG0_EVERY = 100000, 1000


def cfg0_hashes():
    r = refcpu.RefCpu(CFG0)
    digests = []
    for k in range(CFG0_EVENTS // CFG0_EVERY):
        r.run(CFG0_EVERY)
        assert r.state()["events"] == (k + 1) * CFG0_EVERY
        digests.append(np.frombuffer(hashlib.sha256(r.sites().tobytes()).digest(), dtype=np.uint8))
    st = r.state()
    np.savez_compressed(os.path.join(HERE, "hashes_cfg0_40cube.npz"), cfg=json.dumps(CFG0),
                        events=np.int64(st["events"]), every=np.int64(CFG0_EVERY), sha256=np.stack(digests),
                        evolution=st["evolution"], sector_time=np.float64(st["sector_time"]),
                        step=np.int64(st["step"]))
    print("cfg0 hashes", st["events"], "events")


# Config 1 of the benchmark suite (256^3 cells, 6,711 vacancies, T = 573 K,
# sector threshold 1e-6 s): the first 1e5 events.  Every 1000 events the
# SHA-256 of that block's (from, to) site indices (int64 pairs) is stored,
# every 10,000 events the SHA-256 of the whole site array, and the sector
# clocks at the end.  The oracle visits the sector's vacancies through its
# list (the same sweep order as the site sweep, tests/test_ghost_oracle.py).
CFG1 = make_cfg(n=256, va=6711, T=573.0, sector_T=1e-6, physics=1.0)
CFG1_EVENTS, CFG1_EVERY, CFG1_SITES_EVERY = 100000, 1000, 10000


def cfg1_hashes(resume_dir=None):
    """resume_dir (optional): where the oracle state and the digests so far
    are kept after every 10,000 events, so that an interrupted run resumes."""
    r = refcpu.RefCpu(CFG1, list_sweep=True)
    log_digests, site_digests = [], []
    out = os.path.join(HERE, "hashes_cfg1_256cube.npz")
    k0 = 0
    if resume_dir:
        os.makedirs(resume_dir, exist_ok=True)
        st_path, dg_path = os.path.join(resume_dir, "state.txt"), os.path.join(resume_dir, "digests.npz")
        if os.path.exists(dg_path) and os.path.exists(st_path):
            z = np.load(dg_path, allow_pickle=False)
            log_digests = list(z["log_sha256"])
            site_digests = list(z["sites_sha256"])
            r.load(st_path)
            k0 = len(log_digests)
            assert r.state()["events"] == k0 * CFG1_EVERY
            print("cfg1 resumes at", k0 * CFG1_EVERY, "events", flush=True)
    for k in range(k0, CFG1_EVENTS // CFG1_EVERY):
        rc, lf, lt = r.run(CFG1_EVERY, log=True)
        assert rc == 1 and r.state()["events"] == (k + 1) * CFG1_EVERY, rc
        pairs = np.stack([lf, lt], axis=1).astype(np.int64)
        log_digests.append(np.frombuffer(hashlib.sha256(pairs.tobytes()).digest(), dtype=np.uint8))
        if (k + 1) * CFG1_EVERY % CFG1_SITES_EVERY == 0:
            site_digests.append(np.frombuffer(hashlib.sha256(r.sites().tobytes()).digest(), dtype=np.uint8))
            st = r.state()
            np.savez_compressed(out, cfg=json.dumps(CFG1), events=np.int64(st["events"]),
                                every=np.int64(CFG1_EVERY), sites_every=np.int64(CFG1_SITES_EVERY),
                                log_sha256=np.stack(log_digests), sites_sha256=np.stack(site_digests),
                                evolution=st["evolution"], sector_time=np.float64(st["sector_time"]),
                                step=np.int64(st["step"]), sect=np.int64(st["sect"]))
            print("cfg1", st["events"], "events", flush=True)
            if resume_dir:
                r.save(st_path + ".tmp")
                np.savez(dg_path + ".tmp.npz", log_sha256=np.stack(log_digests), sites_sha256=np.stack(site_digests))
                os.replace(st_path + ".tmp", st_path)
                os.replace(dg_path + ".tmp.npz", dg_path)


# Config 2's lattice (512^3 cells, Fe-1.34Cu-0.70Ni-1.30Mn, 2e-4 vacancies,
# 563 K): sector 0's rates (SHA-256 of all rows, the first 512 rows in full)
# and the first 200 events' site pairs, occupancy and clocks.
CFG2_EVENTS, CFG2_ROWS = 200, 512


def cfg2_fixture():
    sys.path.insert(0, os.path.join(HERE, "..", "..", "misa-akmc-gpu_amd"))
    from akmc import workloads
    cfg = workloads.precision_512(0)
    r = refcpu.RefCpu(cfg, list_sweep=True)
    print("cfg2 created", flush=True)
    sites, rates, total = r.sector_rates(0, cap=1 << 15)
    print("cfg2 sector 0:", sites.size, "vacancies", flush=True)
    rc, lf, lt = r.run(CFG2_EVENTS, log=True)
    assert rc == 1 and r.state()["events"] == CFG2_EVENTS
    st = r.state()
    pairs = np.stack([lf, lt], axis=1).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "cfg2_512cube.npz"), cfg=json.dumps(cfg), sector=np.int64(0),
                        n_rows=np.int64(sites.size), rates_sha256=np.frombuffer(
                            hashlib.sha256(np.ascontiguousarray(rates).tobytes()).digest(), dtype=np.uint8),
                        sites_sha256_rows=np.frombuffer(hashlib.sha256(sites.tobytes()).digest(), dtype=np.uint8),
                        head_sites=sites[:CFG2_ROWS], head_rates=rates[:CFG2_ROWS], total=np.float64(total),
                        events=np.int64(CFG2_EVENTS), log=pairs,
                        final_sha256=np.frombuffer(hashlib.sha256(r.sites().tobytes()).digest(), dtype=np.uint8),
                        evolution=st["evolution"], sector_time=np.float64(st["sector_time"]),
                        step=np.int64(st["step"]), sect=np.int64(st["sect"]))
    print("cfg2 written", flush=True)


def itl_rates():
    """Dumbbell rates of a dense typed lattice (6 x 4 x 9 cells, 10 % dumbbell
    sites of every kind and orientation, Cu/Ni/Mn atoms and vacancies) from
    the oracle's restatement of the reference's dumbbell model."""
    rng = np.random.default_rng(2024)
    box = (6, 4, 9)
    n = 2 * box[0] * box[1] * box[2]
    u = rng.random(n)
    t = np.full(n, 1, dtype=np.uint16)
    t[u < 0.17] = 0
    t[u < 0.15] = 8
    t[u < 0.10] = 4
    t[u < 0.05] = 2
    sites = np.sort(rng.choice(n, size=n // 10, replace=False)).astype(np.int64)
    kinds = np.array([0x0101, 0x0102, 0x0104, 0x0108, 0x0202, 0x0204, 0x0208, 0x0404, 0x0408, 0x0808], np.uint16)
    t[sites] = kinds[np.arange(sites.size) % kinds.size]
    orient = np.array([-6, -5, -4, -3, -2, -1, 1, 2, 3, 4, 5, 6], np.int8)[np.arange(sites.size) % 12]
    nu, T = 6e12, 563.0
    av, r = refcpu.Itl().rates(box, t, sites, orient, nu, T)
    np.savez_compressed(os.path.join(HERE, "itl_dense_6x4x9.npz"), box=np.array(box, np.int64), types=t,
                        sites=sites, orient=orient, nu=np.float64(nu), T=np.float64(T), avail=av, rates=r)
    print("itl", sites.size, "dumbbells", int((r > 0).sum()), "non-zero rate slots")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "cfg0":
        cfg0_hashes()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "cfg1":
        cfg1_hashes(sys.argv[2] if len(sys.argv) > 2 else None)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "cfg2":
        cfg2_fixture()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "itl":
        itl_rates()
        return
    for name, (cfg, n) in CASES.items():
        r = refcpu.RefCpu(cfg)
        rc, lf, lt = r.run(n, log=True)
        st = r.state()
        sites = r.sites()
        np.savez_compressed(os.path.join(HERE, "traj_%s.npz" % name),
                            cfg=json.dumps(cfg), events=np.int64(st["events"]),
                            from_=lf[:st["events"]].astype(np.int32), to=lt[:st["events"]].astype(np.int32),
                            sites_sha256=hashlib.sha256(sites.tobytes()).hexdigest(),
                            evolution=st["evolution"], sector_time=np.float64(st["sector_time"]),
                            step=np.int64(st["step"]))
        print(name, st["events"], "events")


if __name__ == "__main__":
    main()
