"""Brute-force pin of the oracle's rate on tiny periodic boxes (CPU only).

The rates are recomputed here from BCC geometry alone: positions in units of
a0/2 (corner sites at even coordinates, body sites at odd ones), first
neighbours at (+-1, +-1, +-1), second neighbours at +-2 along one axis with
the (0, +2, 0) bond left out (the reference's 0x2F mask,
period_lattice_list.cpp:15-18), hop direction b = (b&4, b&2, b&1 -> +1, else
-1), dE = E_after - E_before from the bond tables, Ea = e0 + dE/2,
k = nu exp(-Ea / (kB T)).  Nothing is shared with the oracle's index
arithmetic, so a wrong offset, order or sign there fails these tests; the
comparison is to 1e-12 because the summation order differs.
"""
import math

import numpy as np
import pytest

import refcpu
from conftest import make_cfg

KB = 8.617330350e-5
NAMES = ["V", "Fe", "Cu", "Ni", "Mn"]
E0 = {"Fe": 0.62, "Cu": 0.54, "Ni": 0.68, "Mn": 1.03}


def species(f):
    return "V" if f & 1 else "Cu" if f & 2 else "Mn" if f & 4 else "Ni" if f & 8 else "Fe"


class Geo:
    def __init__(self, sites, box):
        self.box = box
        self.L = (2 * box[0], 2 * box[1], 2 * box[2])
        self.sp = {}
        X = 2 * box[0]
        for i, f in enumerate(sites):
            x, y, z = i % X, (i // X) % box[1], i // (X * box[1])
            p = (2 * (x // 2) + (x & 1), 2 * y + (x & 1), 2 * z + (x & 1))
            self.sp[p] = species(int(f))
        self.index = {}
        for i in range(len(sites)):
            x, y, z = i % X, (i // X) % box[1], i // (X * box[1])
            self.index[(2 * (x // 2) + (x & 1), 2 * y + (x & 1), 2 * z + (x & 1))] = i

    def wrap(self, p):
        return (p[0] % self.L[0], p[1] % self.L[1], p[2] % self.L[2])

    def nn1(self, p):
        return [self.wrap((p[0] + dx, p[1] + dy, p[2] + dz)) for dx in (-1, 1) for dy in (-1, 1) for dz in (-1, 1)]

    def nn2(self, p):
        offs = [(-2, 0, 0), (2, 0, 0), (0, -2, 0), (0, 0, -2), (0, 0, 2)]   # (0, +2, 0) not counted
        return [self.wrap((p[0] + a, p[1] + b, p[2] + c)) for a, b, c in offs]

    def energy(self, p, typ, override):
        e = 0.0
        for q in self.nn1(p):
            e += refcpu.bond1(typ, override.get(q, self.sp[q]))
        for q in self.nn2(p):
            e += refcpu.bond2(typ, override.get(q, self.sp[q]))
        return e

    def delta_e(self, v, b):
        t = self.wrap((v[0] + (1 if b & 4 else -1), v[1] + (1 if b & 2 else -1), v[2] + (1 if b & 1 else -1)))
        X = self.sp[t]
        before = self.energy(v, "V", {}) + self.energy(t, X, {})
        after = self.energy(v, X, {t: "V"}) + self.energy(t, "V", {v: X})
        return after - before, X, t


@pytest.mark.parametrize("box,va,seed", [((4, 4, 4), 3, 1), ((5, 4, 6), 5, 2), ((6, 6, 6), 6, 3), ((4, 5, 4), 2, 4)])
def test_rates_from_geometry(box, va, seed):
    rng = np.random.default_rng(seed)
    cfg = make_cfg(box=box, va=0, T=600.0)
    r = refcpu.RefCpu(cfg)
    n = 2 * box[0] * box[1] * box[2]
    s = rng.choice(np.array([0, 0, 0, 2, 4, 8], dtype=np.uint8), size=n)
    vac = rng.choice(n, size=va, replace=False)
    s[vac] |= 1
    r.set_sites(s)
    g = Geo(s, box)
    checked = 0
    for v in vac:
        pv = [p for p, i in g.index.items() if i == v][0]
        for b in range(8):
            de, X, t = g.delta_e(pv, b)
            if X == "V":
                continue
            assert r.delta_e(int(v), b) == pytest.approx(de, abs=1e-12), (v, b)
            k = 6e12 * math.exp(-(E0[X] + de / 2) / (KB * 600.0))
            assert r.rate(int(v), b) == pytest.approx(k, rel=1e-12), (v, b)
            # the target of direction b is the site the oracle hops to
            checked += 1
    assert checked >= 8 * va - 2 * va
