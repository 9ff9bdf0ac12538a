"""Pins of the oracle (oracle/refcpu.cpp) that do not retype its formulas.

* Closed forms of the reference's own unit tests
  (tests/unit-abvi/rate/bonds/bonds_counter_test.cpp:17-100,
  tests/unit-abvi/rate/vac_rates_solver_test.cpp:20-106), on a periodic box,
  for both the 6-bond (0x3F) 2nn mask those tests assume and the production
  0x2F mask (the closed forms hold for either, as the uniform Fe background
  cancels).
* The 0x2F mask itself (src/lattice/period_lattice_list.cpp:15-18): the
  (x, y+1, z) 2nn bond is not counted.
* Detailed balance: the reverse hop of a configuration's swap has exactly
  -dE (the same two addends, commuted), for random configurations.
* The reference's randomAtomsType vectors (tests/unit/lattice/
  lattice_types_test.cpp:8-55), and the reference's own lattice_types.cpp /
  pair_bond.hpp compiled from source when available (oracle/_ref).
"""
import ctypes
import math

import numpy as np
import pytest

import refcpu
from conftest import make_cfg

KB = 8.617330350e-5


def _box(n=8, mask=0x2F, T=700.0):
    cfg = make_cfg(n=n, va=0, T=T)
    cfg["mask_2nn"] = mask
    r = refcpu.RefCpu(cfg, create=False)
    r.set_sites(np.zeros(r.n_sites, dtype=np.uint8))  # all Fe
    return r


def _set(r, xyz, name):
    s = r.sites()
    bits = {"V": 1, "Cu": 2, "Mn": 4, "Ni": 8, "Fe": 0}
    s[r.site_index(*xyz)] = bits[name]
    r.set_sites(s)


B1, B2 = refcpu.bond1, refcpu.bond2


@pytest.mark.parametrize("mask,n2", [(0x3F, 6), (0x2F, 5)])
def test_count_all_fe(mask, n2):
    r = _box(mask=mask)
    e = r.count(r.site_index(4, 2, 2), "Fe")
    assert e == pytest.approx(B1("Fe", "Fe") * 8 + B2("Fe", "Fe") * n2, abs=1e-12)


@pytest.mark.parametrize("mask,n2", [(0x3F, 6), (0x2F, 5)])
def test_count_with_vacancy(mask, n2):
    r = _box(mask=mask)
    _set(r, (3, 1, 2), "Ni")  # 1nn of the source
    _set(r, (4, 2, 2), "V")
    src, tar = r.site_index(4, 2, 2), r.site_index(5, 2, 2)
    assert r.count(src, "V") == pytest.approx(B1("V", "Fe") * 7 + B1("V", "Ni") + B2("V", "Fe") * n2, abs=1e-12)
    assert r.count(tar, "Fe") == pytest.approx(B1("V", "Fe") + B1("Fe", "Fe") * 7 + B2("Fe", "Fe") * n2, abs=1e-12)


@pytest.mark.parametrize("mask,n2", [(0x3F, 6), (0x2F, 5)])
def test_count_with_vacancy_exchanged(mask, n2):
    r = _box(mask=mask)
    _set(r, (3, 1, 2), "Ni")
    _set(r, (5, 2, 2), "V")
    src, tar = r.site_index(4, 2, 2), r.site_index(5, 2, 2)
    assert r.count(src, "Fe") == pytest.approx(
        B1("Fe", "Fe") * 6 + B1("V", "Fe") + B1("Fe", "Ni") + B2("Fe", "Fe") * n2, abs=1e-12)
    assert r.count(tar, "V") == pytest.approx(B1("V", "Fe") * 8 + B2("V", "Fe") * n2, abs=1e-12)


@pytest.mark.parametrize("mask", [0x3F, 0x2F])
def test_delta_e_null(mask):
    T = 700.0
    r = _box(mask=mask, T=T)
    _set(r, (4, 2, 2), "V")
    src = r.site_index(4, 2, 2)
    assert r.delta_e(src, 7) == pytest.approx(0.0, abs=1e-10)  # (5,2,2) is 1nn direction 7 of an even site
    expected = 6e12 * math.exp(-0.62 / (KB * T))
    assert r.rate(src, 7) / expected == pytest.approx(1.0, abs=1e-10)


@pytest.mark.parametrize("mask", [0x3F, 0x2F])
def test_delta_e_both_neighbour(mask):
    r = _box(mask=mask)
    _set(r, (4, 2, 3), "Ni")  # 2nn of the source, 1nn of the target
    _set(r, (4, 2, 2), "V")
    de = r.delta_e(r.site_index(4, 2, 2), 7)
    expected = (B1("Fe", "Fe") + B1("V", "Ni") + B2("Fe", "Ni") + B2("V", "Fe")
                - B1("V", "Fe") - B1("Fe", "Ni") - B2("V", "Ni") - B2("Fe", "Fe"))
    assert de == pytest.approx(expected, abs=1e-10)


@pytest.mark.parametrize("mask", [0x3F, 0x2F])
def test_delta_e_src_1nn_neighbour(mask):
    r = _box(mask=mask)
    _set(r, (3, 1, 2), "Ni")
    _set(r, (4, 2, 2), "V")
    de = r.delta_e(r.site_index(4, 2, 2), 7)
    expected = (B1("V", "Fe") + B1("Fe", "Ni")) - (B1("V", "Ni") + B1("Fe", "Fe"))
    assert de == pytest.approx(expected, abs=1e-10)


def test_masked_2nn_direction_is_not_counted():
    for mask, counted in [(0x3F, True), (0x2F, False)]:
        r = _box(mask=mask)
        _set(r, (4, 3, 2), "Ni")  # (x, y+1, z): 2nn index 4
        e = r.count(r.site_index(4, 2, 2), "V")
        base = B1("V", "Fe") * 8 + B2("V", "Fe") * (5 if mask == 0x2F else 5)
        extra = B2("V", "Ni") if counted else 0.0
        assert e == pytest.approx(base + extra, abs=1e-12)
    # ...while the (x, y-1, z) one is counted under the production mask.
    r = _box(mask=0x2F)
    _set(r, (4, 1, 2), "Ni")
    e = r.count(r.site_index(4, 2, 2), "V")
    assert e == pytest.approx(B1("V", "Fe") * 8 + B2("V", "Fe") * 4 + B2("V", "Ni"), abs=1e-12)


def test_detailed_balance_bitwise():
    rng = np.random.default_rng(7)
    r = _box(n=8)
    checked = 0
    for trial in range(40):
        s = rng.choice(np.array([0, 0, 0, 0, 2, 4, 8, 1], dtype=np.uint8), size=r.n_sites)
        r.set_sites(s)
        vac = np.nonzero(s & 1)[0]
        for v in vac[:4]:
            x, y, z = v % 16, (v // 16) % 8, v // 128
            for b in range(8):
                dy0 = 0 if x % 2 else -1
                t = r.site_index(x + (1 if b & 4 else -1), y + dy0 + ((b >> 1) & 1), z + dy0 + (b & 1))
                if s[t] & 1:
                    continue
                de_f = r.delta_e(v, b)
                sw = s.copy()
                sw[v], sw[t] = s[t], 1
                r.set_sites(sw)
                de_r = r.delta_e(t, 7 - b)
                r.set_sites(s)
                assert de_r == -de_f, (trial, v, b, de_f, de_r)
                checked += 1
    assert checked > 100


REF_TYPES = [0, 1, 2, 8]  # V, Fe, Cu, Mn as in the reference test


def _oracle_random_atoms_type(types, ratio, hit):
    cfg = make_cfg(n=8, va=0)
    name = {0: "V", 1: "Fe", 2: "Cu", 4: "Ni", 8: "Mn"}
    cfg["alloy"] = [(name[t], q) for t, q in zip(types, ratio)]
    return refcpu.RefCpu(cfg, create=False).random_atoms_type(hit)


def test_random_atoms_type_reference_vectors():
    ratio = [4, 4, 8, 8]
    want = {**{i: 0 for i in range(1, 5)}, **{i: 1 for i in range(5, 9)},
            **{i: 2 for i in range(9, 17)}, **{i: 8 for i in range(17, 25)}, 25: 1}
    for hit, t in want.items():
        assert _oracle_random_atoms_type(REF_TYPES, ratio, hit) == t
    for hit in range(1, 17):
        assert _oracle_random_atoms_type(REF_TYPES, [0, 0, 8, 8], hit) == (2 if hit <= 8 else 8)


def test_against_reference_source():
    L = refcpu.refsrc()
    if L is None:
        pytest.skip("oracle/_ref not built (the reference sources are not mounted here)")
    types = (ctypes.c_int32 * 4)(*REF_TYPES)
    for ratio in ([4, 4, 8, 8], [0, 0, 8, 8], [9666, 134, 130, 70]):
        rr = (ctypes.c_int64 * 4)(*ratio)
        for hit in list(range(1, 30)) + [sum(ratio), sum(ratio) + 1]:
            assert L.refsrc_random_atoms_type(types, rr, 4, hit) == _oracle_random_atoms_type(REF_TYPES, ratio, hit)
    lib = refcpu.lib()
    for a in (0, 1, 2, 4, 8):
        for b in (0, 1, 2, 4, 8):
            assert L.refsrc_make_bond(a, b) == lib.ref_make_bond(a, b)
