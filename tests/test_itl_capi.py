"""akmc_itl_rates on the host side: argument checks run before any device is
touched, so they are tested without a GPU."""
import numpy as np
import pytest

import akmc


def _lattice(box=(4, 4, 4)):
    return np.full(2 * box[0] * box[1] * box[2], akmc.LAT_FE, dtype=np.uint16)


def test_rejects_bad_input():
    box = (4, 4, 4)
    t = _lattice(box)
    t[5] = akmc.dumbbell(akmc.LAT_FE, akmc.LAT_CU)
    t[9] = akmc.dumbbell(akmc.LAT_CU, akmc.LAT_CU)
    bad = [
        dict(sites=[9, 5], orient=[1, 1]),                  # not ascending
        dict(sites=[5, 6], orient=[1, 1]),                  # site 6 holds an atom
        dict(sites=[5], orient=[0]),                        # unknown orientation
        dict(sites=[5], orient=[7]),
        dict(sites=[600], orient=[1]),                      # outside the box
    ]
    for b in bad:
        with pytest.raises(akmc.AkmcError):
            akmc.itl_rates(box, t, b["sites"], b["orient"], 6e12, 573.0)
    with pytest.raises(akmc.AkmcError):
        akmc.itl_rates((2, 4, 4), np.full(64, 1, dtype=np.uint16), [], [], 6e12, 573.0)   # box < 3
    u = t.copy()
    u[7] = 0x0201                                           # high end above low end
    with pytest.raises(akmc.AkmcError):
        akmc.itl_rates(box, u, [5], [1], 6e12, 573.0)
    u = t.copy()
    u[7] = 3                                                # not a lattice type
    with pytest.raises(akmc.AkmcError):
        akmc.itl_rates(box, u, [5], [1], 6e12, 573.0)
    with pytest.raises(akmc.AkmcError):
        akmc.itl_rates(box, t, [5], [1], 6e12, 0.0)        # temperature


def test_no_dumbbells_needs_no_device():
    av, r = akmc.itl_rates((4, 4, 4), _lattice(), [], [], 6e12, 573.0)
    assert av.shape == (0,) and r.shape == (0, 8)


def test_dumbbell_encoding():
    assert akmc.dumbbell(akmc.LAT_CU, akmc.LAT_FE) == 0x0102
    assert akmc.dumbbell(akmc.LAT_MN, akmc.LAT_NI) == 0x0408
