// Table-driven double-precision exp, bit-compatible with the host C library's
// exp on x86-64 glibc >= 2.28 (FMA code path).
//
// The vacancy-hop rate is k = nu * exp(-Ea / (kB*T)) (reference:
// models/abvi/rate/rates_solver.cpp:25, which calls the C library exp).  For
// the device trajectory to be bit-identical to the CPU build, the device must
// return exactly what the host library returns.  This header restates the
// published algorithm of ARM optimized-routines' exp (the one glibc adopted in
// 2.28): x = k*ln2/128 + r, exp(x) = 2^(k/128) * exp(r), with a degree-5
// polynomial for exp(r) - 1 and a 128-entry 2^(k/128) table generated by
// gen_exp_table.py.  The fused multiply-adds are written out explicitly so the
// result does not depend on compiler contraction; the placement matches the
// glibc x86-64 FMA multiarch build and is checked against the host library on
// every test run.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define AKMC_HD __host__ __device__
#else
#define AKMC_HD
#endif

namespace akmc {

AKMC_HD inline uint64_t as_u64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint64_t)__double_as_longlong(x);
#else
  uint64_t u;
  memcpy(&u, &x, sizeof u);
  return u;
#endif
}

AKMC_HD inline double as_f64(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)u);
#else
  double x;
  memcpy(&x, &u, sizeof x);
  return x;
#endif
}

AKMC_HD inline double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Published constants of the algorithm (N = 128, polynomial order 5).
#define AKMC_EXP_INVLN2N 0x1.71547652b82fep7
#define AKMC_EXP_SHIFT 0x1.8p52
#define AKMC_EXP_NEGLN2HIN (-0x1.62e42fefa0000p-8)
#define AKMC_EXP_NEGLN2LON (-0x1.cf79abc9e3b3ap-47)
#define AKMC_EXP_C2 0x1.ffffffffffdbdp-2
#define AKMC_EXP_C3 0x1.555555555543cp-3
#define AKMC_EXP_C4 0x1.55555cf172b91p-5
#define AKMC_EXP_C5 0x1.1111167a4d017p-7

AKMC_HD inline uint32_t exp_top12(double x) { return (uint32_t)(as_u64(x) >> 52); }

// Results whose scale exponent over- or underflows (|x| > ~708).
AKMC_HD inline double exp_specialcase(double tmp, uint64_t sbits, uint64_t ki) {
  if ((ki & 0x80000000u) == 0) {
    sbits -= 1009ull << 52;
    double scale = as_f64(sbits);
    return 0x1p1009 * fma_(scale, tmp, scale);
  }
  sbits += 1022ull << 52;
  double scale = as_f64(sbits);
  double y = scale + scale * tmp;
  if (y < 1.0) {
    // Round y to the target precision before scaling into the subnormal
    // range to avoid double rounding.
    double lo = scale - y + scale * tmp;
    double hi = 1.0 + y;
    lo = 1.0 - hi + y + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

// tab: 256 words, akmc_exp_tab_init from akmc_exp_table.h (host) or its
// device copy.
AKMC_HD inline double exp_glibc(double x, const uint64_t* tab) {
  uint32_t abstop = exp_top12(x) & 0x7ff;
  if (abstop - exp_top12(0x1p-54) >= exp_top12(512.0) - exp_top12(0x1p-54)) {
    if ((int32_t)(abstop - exp_top12(0x1p-54)) < 0) return 1.0 + x;  // tiny |x|
    if (abstop >= exp_top12(1024.0)) {
      if (as_u64(x) == as_u64(-__builtin_inf())) return 0.0;
      if (abstop >= exp_top12(__builtin_inf())) return 1.0 + x;
      return (as_u64(x) >> 63) ? 0.0 : __builtin_inf();
    }
    abstop = 0;  // large |x|: handled by exp_specialcase
  }
  double kd = fma_(AKMC_EXP_INVLN2N, x, AKMC_EXP_SHIFT);
  uint64_t ki = as_u64(kd);
  kd -= AKMC_EXP_SHIFT;
  double r = fma_(kd, AKMC_EXP_NEGLN2LON, fma_(kd, AKMC_EXP_NEGLN2HIN, x));
  uint64_t idx = 2 * (ki % 128);
  uint64_t top = ki << (52 - 7);
  double tail = as_f64(tab[idx]);
  uint64_t sbits = tab[idx + 1] + top;
  double r2 = r * r;
  double tmp = fma_(r2 * r2, fma_(r, AKMC_EXP_C5, AKMC_EXP_C4),
                    fma_(r2, fma_(r, AKMC_EXP_C3, AKMC_EXP_C2), tail + r));
  if (abstop == 0) return exp_specialcase(tmp, sbits, ki);
  double scale = as_f64(sbits);
  return fma_(scale, tmp, scale);
}

}  // namespace akmc
