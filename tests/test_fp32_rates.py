"""The FP32 rate form on the host: the table exp and the integer-valued bond
sums, against the oracle's FP64 rates (CPU only).

The replica engine computes an FP32 rate as tab_hi[(m - m_min) >> 8] *
tab_lo[(m - m_min) & 255] where m = (2*e0(X) + dE) / q is an integer
because every bond energy is a multiple of q.  These tests pin both halves:
the table over its whole range against exp in FP64, and m against the
oracle's FP64 dE on random lattices; the device kernel is then compared to
the oracle directly (tests/test_gpu_batch.py).
"""
import numpy as np
import pytest

import akmc
import refcpu
from conftest import make_cfg

KB = 8.617330350E-5
BOUND = 1e-6  # the FP32 form's relative rate error budget against FP64


@pytest.mark.parametrize("T", [500.0, 563.0, 573.0, 620.0, 700.0])
def test_table_exhaustive(T):
    tab, nhi, m_min, q, e1q, e2q, e0q2 = akmc.fp32_rate_table(T)
    kT = KB * T
    m = np.arange(m_min, m_min + 256 * nhi)
    with np.errstate(over="ignore"):
        r32 = (tab[256 + ((m - m_min) >> 8)] * tab[(m - m_min) & 255]).astype(np.float64)
    r64 = 6e12 * np.exp(-m * q / (2 * kT))
    ok = np.isfinite(r32) & (r64 >= np.finfo(np.float32).tiny)
    # The physical window (0 < Ea < 3 eV) lies well inside the finite range.
    ea = m * q / 2
    assert ok[(ea > 0) & (ea < 3)].all()
    rel = np.abs(r32[ok] - r64[ok]) / r64[ok]
    assert rel.max() < 2.0 ** -22, rel.max()   # three FP32 roundings
    assert rel.max() < BOUND


def test_energy_quantum_tables():
    tab, nhi, m_min, q, e1q, e2q, e0q2 = akmc.fp32_rate_table(573.0)
    assert q == 1e-3
    e1 = np.array([refcpu.bond1(a, b) for a in "V Fe Cu Ni Mn".split() for b in "V Fe Cu Ni Mn".split()])
    e2 = np.array([refcpu.bond2(a, b) for a in "V Fe Cu Ni Mn".split() for b in "V Fe Cu Ni Mn".split()])
    assert np.array_equal(e1q, np.round(e1 / q)) and np.allclose(e1q * q, e1, atol=1e-12)
    assert np.array_equal(e2q, np.round(e2 / q)) and np.allclose(e2q * q, e2, atol=1e-12)
    assert list(e0q2) == [0, 1240, 1080, 1360, 2060]   # 2*e0/q for V, Fe, Cu, Ni, Mn


@pytest.mark.parametrize("T,alloy", [(573.0, None), (650.0, [("Fe", 6000), ("Mn", 2000), ("Ni", 1500), ("Cu", 500)])])
def test_quantized_rates_match_oracle(T, alloy):
    """m from the oracle's FP64 dE is an integer, and the table rate of m is
    within 2e-7 of the oracle's FP64 rate, for every direction of every
    vacancy of random lattices."""
    tab, nhi, m_min, q, e1q, e2q, e0q2 = akmc.fp32_rate_table(T)
    cfg = make_cfg(n=12, va=40, T=T, alloy=alloy)
    r = refcpu.RefCpu(cfg)
    s = r.sites()
    e0 = {"Fe": 0.62, "Cu": 0.54, "Ni": 0.68, "Mn": 1.03}
    worst, n = 0.0, 0
    X, Y = 24, 12
    for v in np.nonzero(s & 1)[0]:
        x, y, z = v % X, (v // X) % Y, v // (X * Y)
        for b in range(8):
            dy0 = 0 if x % 2 else -1
            t = r.site_index(x + (1 if b & 4 else -1), y + dy0 + ((b >> 1) & 1), z + dy0 + (b & 1))
            f = s[t]
            if f & 1:
                continue
            sp = "Cu" if f & 2 else "Mn" if f & 4 else "Ni" if f & 8 else "Fe"
            de = r.delta_e(v, b)
            mf = (2 * e0[sp] + de) / q
            m = int(round(mf))
            assert abs(mf - m) < 1e-9
            r32 = float(np.float32(tab[256 + ((m - m_min) >> 8)]) * np.float32(tab[(m - m_min) & 255]))
            r64 = r.rate(v, b)
            worst = max(worst, abs(r32 - r64) / r64)
            n += 1
    assert n > 200
    assert worst < 2.0 ** -22, worst
