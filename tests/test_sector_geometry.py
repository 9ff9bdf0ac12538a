"""Sector geometry pinned by the reference's own per-sector ghost bounds.

The reference's CPU build takes its sector regions from libcomm (unvendored);
only comments describe them (src/simulation.cpp:91-110: low half n/2 cells,
local_split_coord = n/2 + ghost).  Its GPU code, however, spells out, for
every sector id, the ghost sites a hop out of that sector can land in:
add_ghost (gpu/gpu_simulate.hip:2093-3086) tests seven boxes per sector id,
each bound written in local_sub_box_lattice_region (low = ghost, high =
ghost + n), local_split_coord and lattice_size_ghost.  ADD_GHOST below
restates those bounds as data; a few of the reference's lines name another
axis's field for a low bound (e.g. `z_high = ...x_low` in case 1, :2227),
whose value is the same ghost width, and the table records that value.

Those seven boxes must be exactly the sector's region grown by the ghost
width on every side, minus the sub-box, which pins every bound of the
region: a low side at the sub-box's low face, a high side at the split, and
the reverse for a high side.  The oracle's region() and the engine's
sector assignment (akmc_site_sectors, the host copy of the kernels'
key_sector) are checked against it for even and odd boxes.
"""
import itertools

import numpy as np
import pytest

import akmc
import refcpu
from conftest import make_cfg

# add_ghost, per sector id: (exchange_ghost index, x, y, z) with each axis
# (low base, low k, high base, high k) meaning [base + k*ghost, ...) in cells,
# base "lo" = local_sub_box_lattice_region.*_low, "hi" = .*_high,
# "split" = local_split_coord[d].  (x is doubled in the reference's
# coordinates; the test doubles it.)
ADD_GHOST = {
    0: [(0, ("lo", 0, "split", 1), ("lo", 0, "split", 1), ("lo", -1, "lo", 0)),
        (1, ("lo", 0, "split", 1), ("lo", -1, "lo", 0), ("lo", -1, "lo", 0)),
        (2, ("lo", -1, "lo", 0), ("lo", 0, "split", 1), ("lo", -1, "lo", 0)),
        (3, ("lo", -1, "lo", 0), ("lo", -1, "lo", 0), ("lo", -1, "lo", 0)),
        (4, ("lo", 0, "split", 1), ("lo", -1, "lo", 0), ("lo", 0, "split", 1)),
        (5, ("lo", -1, "lo", 0), ("lo", -1, "lo", 0), ("lo", 0, "split", 1)),
        (6, ("lo", -1, "lo", 0), ("lo", 0, "split", 1), ("lo", 0, "split", 1))],
    1: [(0, ("split", -1, "hi", 0), ("lo", 0, "split", 1), ("lo", -1, "lo", 0)),
        (1, ("split", -1, "hi", 0), ("lo", -1, "lo", 0), ("lo", -1, "lo", 0)),
        (2, ("hi", 0, "hi", 1), ("lo", 0, "split", 1), ("lo", -1, "lo", 0)),
        (3, ("hi", 0, "hi", 1), ("lo", -1, "lo", 0), ("lo", -1, "lo", 0)),
        (4, ("split", -1, "hi", 0), ("lo", -1, "lo", 0), ("lo", 0, "split", 1)),
        (5, ("hi", 0, "hi", 1), ("lo", -1, "lo", 0), ("lo", 0, "split", 1)),
        (6, ("hi", 0, "hi", 1), ("lo", 0, "split", 1), ("lo", 0, "split", 1))],
    2: [(0, ("lo", 0, "split", 1), ("split", -1, "hi", 0), ("lo", -1, "lo", 0)),
        (1, ("lo", 0, "split", 1), ("hi", 0, "hi", 1), ("lo", -1, "lo", 0)),
        (2, ("lo", -1, "lo", 0), ("split", -1, "hi", 0), ("lo", -1, "lo", 0)),
        (3, ("lo", -1, "lo", 0), ("hi", 0, "hi", 1), ("lo", -1, "lo", 0)),
        (4, ("lo", 0, "split", 1), ("hi", 0, "hi", 1), ("lo", 0, "split", 1)),
        (5, ("lo", -1, "lo", 0), ("hi", 0, "hi", 1), ("lo", 0, "split", 1)),
        (6, ("lo", -1, "lo", 0), ("split", -1, "hi", 0), ("lo", 0, "split", 1))],
    3: [(0, ("split", -1, "hi", 0), ("split", -1, "hi", 0), ("lo", -1, "lo", 0)),
        (1, ("split", -1, "hi", 0), ("hi", 0, "hi", 1), ("lo", -1, "lo", 0)),
        (2, ("hi", 0, "hi", 1), ("split", -1, "hi", 0), ("lo", -1, "lo", 0)),
        (3, ("hi", 0, "hi", 1), ("hi", 0, "hi", 1), ("lo", -1, "lo", 0)),
        (4, ("split", -1, "hi", 0), ("hi", 0, "hi", 1), ("lo", 0, "split", 1)),
        (5, ("hi", 0, "hi", 1), ("hi", 0, "hi", 1), ("lo", 0, "split", 1)),
        (6, ("hi", 0, "hi", 1), ("split", -1, "hi", 0), ("lo", 0, "split", 1))],
    4: [(0, ("lo", 0, "split", 1), ("lo", 0, "split", 1), ("hi", 0, "hi", 1)),
        (1, ("lo", 0, "split", 1), ("lo", -1, "lo", 0), ("hi", 0, "hi", 1)),
        (2, ("lo", -1, "lo", 0), ("lo", 0, "split", 1), ("hi", 0, "hi", 1)),
        (3, ("lo", -1, "lo", 0), ("lo", -1, "lo", 0), ("hi", 0, "hi", 1)),
        (4, ("lo", 0, "split", 1), ("lo", -1, "lo", 0), ("split", -1, "hi", 0)),
        (5, ("lo", -1, "lo", 0), ("lo", -1, "lo", 0), ("split", -1, "hi", 0)),
        (6, ("lo", -1, "lo", 0), ("lo", 0, "split", 1), ("split", -1, "hi", 0))],
    5: [(0, ("split", -1, "hi", 0), ("lo", 0, "split", 1), ("hi", 0, "hi", 1)),
        (1, ("split", -1, "hi", 0), ("lo", -1, "lo", 0), ("hi", 0, "hi", 1)),
        (2, ("hi", 0, "hi", 1), ("lo", 0, "split", 1), ("hi", 0, "hi", 1)),
        (3, ("hi", 0, "hi", 1), ("lo", -1, "lo", 0), ("hi", 0, "hi", 1)),
        (4, ("split", -1, "hi", 0), ("lo", -1, "lo", 0), ("split", -1, "hi", 0)),
        (5, ("hi", 0, "hi", 1), ("lo", -1, "lo", 0), ("split", -1, "hi", 0)),
        (6, ("hi", 0, "hi", 1), ("lo", 0, "split", 1), ("split", -1, "hi", 0))],
    6: [(0, ("lo", 0, "split", 1), ("split", -1, "hi", 0), ("hi", 0, "hi", 1)),
        (1, ("lo", 0, "split", 1), ("hi", 0, "hi", 1), ("hi", 0, "hi", 1)),
        (2, ("lo", -1, "lo", 0), ("split", -1, "hi", 0), ("hi", 0, "hi", 1)),
        (3, ("lo", -1, "lo", 0), ("hi", 0, "hi", 1), ("hi", 0, "hi", 1)),
        (4, ("lo", 0, "split", 1), ("hi", 0, "hi", 1), ("split", -1, "hi", 0)),
        (5, ("lo", -1, "lo", 0), ("hi", 0, "hi", 1), ("split", -1, "hi", 0)),
        (6, ("lo", -1, "lo", 0), ("split", -1, "hi", 0), ("split", -1, "hi", 0))],
    7: [(0, ("split", -1, "hi", 0), ("split", -1, "hi", 0), ("hi", 0, "hi", 1)),
        (1, ("split", -1, "hi", 0), ("hi", 0, "hi", 1), ("hi", 0, "hi", 1)),
        (2, ("hi", 0, "hi", 1), ("split", -1, "hi", 0), ("hi", 0, "hi", 1)),
        (3, ("hi", 0, "hi", 1), ("hi", 0, "hi", 1), ("hi", 0, "hi", 1)),
        (4, ("split", -1, "hi", 0), ("hi", 0, "hi", 1), ("split", -1, "hi", 0)),
        (5, ("hi", 0, "hi", 1), ("hi", 0, "hi", 1), ("split", -1, "hi", 0)),
        (6, ("hi", 0, "hi", 1), ("split", -1, "hi", 0), ("split", -1, "hi", 0))],
}

BOXES = [(8, 8, 8), (9, 9, 9), (8, 11, 9), (12, 10, 13), (17, 13, 11), (10, 15, 16)]
G = 2  # ghost width: ceil(cutoff 5.6 / a0 2.855)


def _bases(n):
    """Per-axis values of lo, hi and split for a box of n cells (one rank)."""
    return {"lo": G, "hi": G + n, "split": G + n // 2}  # simulation.cpp:107-110


def _piece_mask(shape_cells, box, piece):
    """Boolean mask over the ghost-extended cell grid (z, y, x) of one add_ghost box."""
    _, *axes = piece
    lo, hi = [], []
    for d, (lb, lk, hb, hk) in enumerate(axes):
        b = _bases(box[d])
        lo.append(b[lb] + lk * G)
        hi.append(b[hb] + hk * G)
    m = np.zeros(shape_cells, dtype=bool)
    m[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = True
    return m


def _ghost_shell(shape_cells, box, lo, hi):
    """Region [lo, hi) grown by G on every side, minus the sub-box."""
    grown = np.zeros(shape_cells, dtype=bool)
    grown[lo[2] - G:hi[2] + G, lo[1] - G:hi[1] + G, lo[0] - G:hi[0] + G] = True
    sub = np.zeros(shape_cells, dtype=bool)
    sub[G:G + box[2], G:G + box[1], G:G + box[0]] = True
    return grown & ~sub


def test_table_covers_every_sector():
    assert sorted(ADD_GHOST) == list(range(8))
    for s, pieces in ADD_GHOST.items():
        assert [p[0] for p in pieces] == list(range(7)), s


@pytest.mark.parametrize("box", BOXES)
def test_oracle_regions_match_add_ghost(box):
    r = refcpu.RefCpu(make_cfg(box=box, va=1), create=False)
    shape = (box[2] + 2 * G, box[1] + 2 * G, box[0] + 2 * G)
    cover = np.zeros(shape, dtype=np.int32)
    for s in range(8):
        lo, hi = r.region(s)
        # The seven add_ghost boxes of sector s are disjoint and are exactly
        # the sector's ghost shell.
        union = np.zeros(shape, dtype=np.int32)
        for piece in ADD_GHOST[s]:
            union += _piece_mask(shape, box, piece)
        assert union.max() == 1, (box, s)
        assert np.array_equal(union.astype(bool), _ghost_shell(shape, box, lo, hi)), (box, s)
        cover[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += 1
    # The eight regions partition the sub-box.
    assert np.all(cover[G:G + box[2], G:G + box[1], G:G + box[0]] == 1)
    assert cover.sum() == box[0] * box[1] * box[2]


def _table_region(box, s):
    """Sector s's cell bounds (real coordinates) read off the add_ghost table:
    an axis whose ghost boxes lie below the sub-box is a low half [0, n//2),
    above it a high half [n//2, n)."""
    lo, hi = [], []
    for d in range(3):
        below = any(p[1 + d][0] == "lo" and p[1 + d][1] == -1 for p in ADD_GHOST[s])
        above = any(p[1 + d][2] == "hi" and p[1 + d][3] == 1 for p in ADD_GHOST[s])
        assert below != above
        lo.append(0 if below else box[d] // 2)
        hi.append(box[d] // 2 if below else box[d])
    return lo, hi


@pytest.mark.parametrize("box", BOXES)
def test_engine_sector_of_every_site(box):
    X, Y, Z = 2 * box[0], box[1], box[2]
    idx = np.arange(X * Y * Z, dtype=np.int64)
    sec = akmc.site_sectors(box, idx)
    x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
    for s in range(8):
        lo, hi = _table_region(box, s)
        inside = ((x >= 2 * lo[0]) & (x < 2 * hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) & (z < hi[2]))
        assert np.array_equal(sec == s, inside), (box, s)


def test_a_wrong_split_fails_the_table():
    """The check has teeth: an oracle region with the split moved by one cell,
    or with the x and y bits swapped, does not match the add_ghost boxes."""
    box = (9, 10, 11)
    shape = (box[2] + 2 * G, box[1] + 2 * G, box[0] + 2 * G)
    good = 0
    for s, shift, swap in itertools.product(range(8), (0, 1), (False, True)):
        bits = [(s >> d) & 1 for d in range(3)]
        if swap:
            bits[0], bits[1] = bits[1], bits[0]
        lo, hi = [], []
        for d in range(3):
            half = box[d] // 2 + (shift if d == 2 else 0)
            lo.append(G + (half if bits[d] else 0))
            hi.append(G + (box[d] if bits[d] else half))
        union = np.zeros(shape, dtype=bool)
        for piece in ADD_GHOST[s]:
            union |= _piece_mask(shape, box, piece)
        ok = np.array_equal(union, _ghost_shell(shape, box, lo, hi))
        assert ok == (shift == 0 and (not swap or bits[0] == bits[1])), (s, shift, swap)
        good += ok
    assert good == 8 + 4


@pytest.mark.gpu
@pytest.mark.parametrize("box", [(8, 8, 8), (17, 13, 11), (9, 10, 16)])
def test_device_sectors_match_add_ghost(box):
    """The kernels' own sector lists (the parity hooks of both event loops)
    hold exactly the vacancies inside the table's region of each sector."""
    X, Y, Z = 2 * box[0], box[1], box[2]
    cfg = make_cfg(box=box, va=X * Y * Z // 20, seeds=dict(create_types=5, create_vacancy=6, time_inc=7))
    eng = akmc.Engine(cfg, device=0)
    sites = eng.download()
    vac = np.nonzero(sites & 1)[0]
    x, y, z = vac % X, (vac // X) % Y, vac // (X * Y)
    bat = akmc.Batch([cfg], device=0, precision=akmc.AKMC_PREC_FP64, init=akmc.AKMC_INIT_REFERENCE)
    assert np.array_equal(bat.download(0), sites)
    for s in range(8):
        lo, hi = _table_region(box, s)
        inside = (x >= 2 * lo[0]) & (x < 2 * hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) & (z < hi[2])
        want = np.sort(vac[inside])
        got, _, _ = eng.sector_rates(s, cap=1 << 14)
        assert np.array_equal(np.sort(got), want), (box, s)
        got_b, _, _ = bat.sector_rates(0, s, cap=1 << 14)
        assert np.array_equal(np.sort(got_b), want), (box, s)
    bat.close()
    eng.close()
