"""The two bit-exactness building blocks of the FP64 path, pinned on the CPU.

* exp: the table-driven exp of the rate kernel (akmc_exp.h) against the host C
  library's exp, which the reference's arrhenius calls
  (models/abvi/rate/rates_solver.cpp:25).
* random streams: the per-event u and -log(r) against the oracle, and the
  oracle's streams against an independent pure-Python MT19937 with the C++
  standard's known answer (the 10000th output of a default-constructed
  mt19937 is 4123659995, [rand.predef]) and libstdc++'s 2-word
  generate_canonical (r = fl(w1 + w2 * 2^32) / 2^64, clamped below 1).
"""
import math

import numpy as np

import akmc
import refcpu


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_exp_matches_libc_on_rate_arguments():
    rng = np.random.default_rng(1234)
    # Rate arguments -Ea/(kB T): Ea in ~[0.2, 2] eV, T in [500, 700] K.
    x = rng.uniform(-50.0, 0.0, 2_000_000)
    assert np.array_equal(_bits(akmc.exp_host(x)), _bits(refcpu.exp_array(x)))


def test_exp_matches_libc_full_range_and_specials():
    rng = np.random.default_rng(99)
    x = np.concatenate([
        rng.uniform(-745.2, 709.8, 500_000),      # whole finite range incl. scaled special cases
        rng.uniform(-1e-6, 1e-6, 50_000),         # near zero
        np.array([0.0, -0.0, 1e-300, -1e-300, 2.0 ** -54, -(2.0 ** -54), 709.78, 709.79, -708.4, -745.13,
                  -745.14, -1000.0, 1000.0, 512.0, -512.0, 1.0, -1.0]),
    ])
    a = akmc.exp_host(x)
    b = refcpu.exp_array(x)
    assert np.array_equal(_bits(a), _bits(b))
    inf = akmc.exp_host(np.array([np.inf, -np.inf]))
    assert inf[0] == np.inf and inf[1] == 0.0
    assert np.isnan(akmc.exp_host(np.array([np.nan]))[0])


def _mt19937(seed, n):
    """Plain MT19937 (Matsumoto & Nishimura 1998) -> n 32-bit outputs."""
    mt = [0] * 624
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    out = []
    idx = 624
    while len(out) < n:
        if idx >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        out.append(y)
    return out


def test_mt19937_known_answer():
    assert _mt19937(5489, 10000)[-1] == 4123659995


def test_oracle_streams_against_independent_mt():
    n = 3000
    seed = 594689
    u_ref, l_ref = refcpu.streams(seed, n)
    w_sel = _mt19937(5489, n)
    w_time = _mt19937(seed, 2 * n)
    c = 1.0 / 4294967295.0
    u = np.array([w * c for w in w_sel])
    ll = []
    for i in range(n):
        s = float(w_time[2 * i]) + float(w_time[2 * i + 1]) * 4294967296.0
        r = s / 18446744073709551616.0
        if r >= 1.0:
            r = math.nextafter(1.0, 0.0)
        ll.append(-math.log(r))
    assert np.array_equal(_bits(u), _bits(u_ref))
    assert np.array_equal(_bits(np.array(ll)), _bits(l_ref))


def test_engine_streams_match_oracle():
    u, l = akmc.rng_streams(594689, 0, 200_000)
    ru, rl = refcpu.streams(594689, 200_000)
    assert np.array_equal(_bits(u), _bits(ru)) and np.array_equal(_bits(l), _bits(rl))
    u2, l2 = akmc.rng_streams(594689, 123_456, 1000)
    assert np.array_equal(_bits(u2), _bits(ru[123_456:124_456]))
    assert np.array_equal(_bits(l2), _bits(rl[123_456:124_456]))


def test_selection_draw_can_reach_one():
    # u = 4294967295 * (1.0 / 4294967295): the reference's r::random() can
    # return exactly 1.0, and then no event is above u * sum (the engine
    # reports AKMC_ERR_NO_EVENT where the reference asserts).
    assert 4294967295 * (1.0 / 4294967295.0) == 1.0
