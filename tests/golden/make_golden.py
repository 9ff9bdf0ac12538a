"""Writes the golden trajectory fixtures tests/golden/traj_*.npz from the
oracle (oracle/refcpu.cpp).  Each fixture holds the configuration, the first
events' (from, to) site indices, a SHA-256 of the final site array, the
sector clocks and the partial sector time.  Run from the repository root:
    python3 tests/golden/make_golden.py        # traj_*.npz
    python3 tests/golden/make_golden.py cfg0   # hashes_cfg0_40cube.npz (1e5 events, a few minutes)
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


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "cfg0":
        cfg0_hashes()
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
