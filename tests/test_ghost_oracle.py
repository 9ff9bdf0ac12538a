"""One rank's ghost images are the periodic wrap: pinned on the CPU.

The oracle (and the engine) read a neighbour outside the box at its periodic
image directly.  The reference's single-rank CPU build instead keeps ghost
copies: GhostInitPacker fills them from the images
(ghost_init_packer.cpp:107-179, 266-403), a hop into a ghost is recorded
(ABVIModel::perform, kmc.cpp:385-392) and applied to the owner after the
sector (SimSyncPacker, sim_sync_packer.cpp:149-210), and changed sites
refresh their copies before the next sector (GhostSyncPacker,
ghost_sync_packer.cpp:270-448; sublattice.inl:161-163).  The oracle's
ghost_images form restates that bookkeeping; these tests run both forms from
the same lattices and require identical trajectories, clocks and sites
(stale tags included), with many hops across the box faces.
"""
import numpy as np
import pytest

import refcpu
from conftest import make_cfg

ALLOY_RICH = [("Fe", 6500), ("Cu", 1200), ("Mn", 1200), ("Ni", 1100)]


def _cfg(box, va, seed):
    return make_cfg(box=box, va=va, T=700.0, sector_T=2e-7, physics=1.0, alloy=ALLOY_RICH,
                    seeds=dict(create_types=seed, create_vacancy=seed + 1, time_inc=seed + 2))


def _crossings(box, f, t):
    """Hops whose (from, to) lie on opposite faces of the periodic box."""
    X, Y, Z = 2 * box[0], box[1], box[2]
    fx, fy, fz = f % X, (f // X) % Y, f // (X * Y)
    tx, ty, tz = t % X, (t // X) % Y, t // (X * Y)
    return int(np.sum((np.abs(fx - tx) > 1) | (np.abs(fy - ty) > 1) | (np.abs(fz - tz) > 1)))


def _compare(cfg, n, sites=None):
    a = refcpu.RefCpu(cfg, create=sites is None)
    b = refcpu.RefCpu(cfg, create=sites is None, ghost_images=True)
    if sites is not None:
        a.set_sites(sites)
        b.set_sites(sites)
    s0 = a.sites()
    assert np.array_equal(s0, b.sites())
    ra, fa, ta = a.run(n, log=True)
    rb, fb, tb = b.run(n, log=True)
    sa, sb = a.state(), b.state()
    assert ra == rb and sa["events"] == sb["events"] > 0
    k = sa["events"]
    assert np.array_equal(fa[:k], fb[:k]) and np.array_equal(ta[:k], tb[:k]), "trajectories differ"
    assert np.array_equal(sa["evolution"], sb["evolution"]) and sa["sector_time"] == sb["sector_time"]
    assert sa["step"] == sb["step"] and sa["sect"] == sb["sect"]
    assert np.array_equal(a.sites(), b.sites()), "sites differ"
    return _crossings(cfg["box"], fa[:k], ta[:k]), s0, a


@pytest.mark.parametrize("box,va", [((8, 8, 8), 24), ((9, 9, 9), 24), ((8, 11, 9), 30), ((12, 10, 13), 40),
                                    ((16, 16, 16), 60)])
def test_ghost_images_equal_wrap(box, va):
    cross, s0, _ = _compare(_cfg(box, va, 31), 10000)
    assert cross >= 100, cross
    # Stale tags (Ni/Mn flags kept under a vacancy, creation.cpp:149-158)
    # are present, so their handling at the faces is exercised too.
    assert np.count_nonzero((s0 & 1) & ((s0 & 12) != 0)) > 0


def test_vacancies_on_faces_edges_and_corners():
    box = (9, 8, 10)
    cfg = _cfg(box, 0, 41)
    cfg["va_count"] = 0
    sites = refcpu.RefCpu(cfg).sites()
    X, Y, Z = 2 * box[0], box[1], box[2]
    pts = [(x, y, z) for x in (0, 1, X - 2, X - 1) for y in (0, Y - 1) for z in (0, Z - 1)]
    pts += [(X // 2, 0, Z // 2), (X // 2 + 1, Y - 1, 3), (0, Y // 2, Z // 2), (X - 1, 3, 0), (7, Y // 2 - 1, Z // 2)]
    for (x, y, z) in pts:
        sites[x + X * (y + Y * z)] = (sites[x + X * (y + Y * z)] & 12) | 1  # keep Ni/Mn as stale tags
    cross, _, _ = _compare(cfg, 8000, sites)
    assert cross >= 200, cross


def test_ghost_copies_current_at_sector_ends():
    """Run a whole time loop: every ghost site then holds the readout of its image."""
    box = (8, 9, 10)
    cfg = _cfg(box, 30, 51)
    cfg["physics"] = cfg["physics_time"] = 3e-6
    g = refcpu.RefCpu(cfg, ghost_images=True)
    rc, _, _ = g.run(10 ** 7)
    assert rc == 0 and g.state()["events"] > 1000
    ext = g.ext_sites()
    G = 2
    zz, yy, xx = np.meshgrid(np.arange(ext.shape[0]), np.arange(ext.shape[1]), np.arange(ext.shape[2]),
                             indexing="ij")
    iz = (zz - G) % box[2] + G
    iy = (yy - G) % box[1] + G
    ix = (xx - 2 * G) % (2 * box[0]) + 2 * G
    img = ext[iz, iy, ix]

    def readout_bit(f):
        return np.where(f & 1, 1, np.where(f & 2, 2, np.where(f & 4, 4, np.where(f & 8, 8, 0))))

    ghost = (iz != zz) | (iy != yy) | (ix != xx)
    assert np.array_equal(ext[ghost], readout_bit(img[ghost]))


def test_ghost_form_catches_a_missing_sync():
    """The comparison has teeth: with the sector-end push and refresh switched
    off in the ghost form (a mistake planted in the oracle, on the CPU), hops
    across the faces are lost and the trajectories part."""
    cfg = _cfg((8, 8, 8), 24, 61)
    cfg["physics_time"] = 2e-5  # vacancies that leave through a face are lost: keep the empty loop short
    a = refcpu.RefCpu(cfg)
    b = refcpu.RefCpu(cfg, ghost_images=True)
    b.L.ref_set_ghost_sync(b.h, 0)
    ra, fa, ta = a.run(3000, log=True)
    rb, fb, tb = b.run(3000, log=True)
    k = min(a.state()["events"], b.state()["events"])
    assert not (np.array_equal(fa[:k], fb[:k]) and np.array_equal(ta[:k], tb[:k]))


@pytest.mark.parametrize("box,va", [((16, 16, 16), 40), ((17, 13, 11), 60), ((12, 12, 12), 200)])
def test_list_sweep_equals_the_site_sweep(box, va):
    """The oracle's vacancy-list sweep (used to write long fixtures) visits the
    sector's vacancies in the reference's sweep order: identical runs."""
    cfg = make_cfg(box=box, va=va, T=600.0, alloy=ALLOY_RICH)
    a = refcpu.RefCpu(cfg)
    b = refcpu.RefCpu(cfg, list_sweep=True)
    ra, fa, ta = a.run(3000, log=True)
    rb, fb, tb = b.run(3000, log=True)
    assert np.array_equal(fa, fb) and np.array_equal(ta, tb) and np.array_equal(a.sites(), b.sites())
    assert np.array_equal(a.state()["evolution"], b.state()["evolution"])
    for s in (0, 5):
        sa, pa, qa = a.sector_rates(s)
        sb, pb, qb = b.sector_rates(s)
        assert np.array_equal(sa, sb) and np.array_equal(pa, pb) and qa == qb
