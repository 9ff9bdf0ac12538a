// Wavefront-level sector engine: the vacancy-hop event loop of one lattice's
// current sector, run by one 64-lane wavefront on LDS-resident state (in the
// plain LDS kernel with two helper wavefronts beside it, akmc_plain.hip).
//
// The event chain of one lattice is serial (every hop changes the next
// rates), so the engine is written for latency, not throughput: no workgroup
// barriers, cross-lane work through ballots, DPP and readlane, loads issued
// together and consumed late.  The plain path runs one engine on the whole
// GPU; the batched path runs one engine per replica, many per CU.
//
// Exactness of the running sums.  The reference folds the sector's rates in
// sweep order, sum_rates += rate (models/abvi/kmc.cpp:103), and select()
// re-accumulates the same sequence until it exceeds u*sum (:266-270).  The
// engine keeps P[v], the running sum after vacancy v's 8 slots, bit-identical
// to that serial fold:
//  * Inside one binade [2^e, 2^(e+1)) every representable sum is a multiple
//    of u = 2^(e-52), and fl(P + a) = P + u*rn(a/u) unless a/u ends in exactly
//    1/2 (a tie, whose rounding depends on P).  So while P stays in one
//    binade and no slot ties, the fold is an exact integer prefix sum of
//    rn(a_j/u), done across the wave (fold_range); the first slot that
//    leaves the binade or ties is folded serially and the scan resumes.
//  * After a hop only a short block of sorted positions changes.  Behind it,
//    rn(a_j/u) does not depend on P, so the tail's running sums all shift by
//    the same exact D = P_new - P_old as long as old and new sums stay in the
//    same binade and no slot ties (tail_update).  A per-vacancy mask of the
//    rates' lowest set bit decides ties without touching the rates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_device.h"

namespace akmc {

// ---------------------------------------------------------------- wave ops
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }
__device__ __forceinline__ int first_set(uint64_t m) { return m ? (int)__ffsll((unsigned long long)m) - 1 : 64; }
__device__ __forceinline__ int last_set(uint64_t m) { return m ? 63 - (int)__clzll((unsigned long long)m) : -1; }

// Value of v in lane `lane` (a wave-uniform index).
__device__ __forceinline__ double bcast_d(double v, int lane) {
  const uint64_t b = as_u64(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return as_f64(((uint64_t)hi << 32) | lo);
}

// A value every lane holds identically, moved to scalar registers so that
// the compiler treats it (and branches on it) as wave-uniform.  Only for
// values computed the same way in all active lanes.
__device__ __forceinline__ int uni_i(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uni_d(double x) {
  const uint64_t b = as_u64(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return as_f64(((uint64_t)hi << 32) | lo);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v) {
  const uint64_t b = as_u64(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xF, false);
  return as_f64(((uint64_t)hi << 32) | lo);
}

// OR of a 64-bit value over each group of eight lanes (8k .. 8k + 7), in
// every lane of the group: quad_perm [1,0,3,2], quad_perm [2,3,0,1], then
// row_half_mirror (lane i <-> 7 - i within each half row).
__device__ __forceinline__ uint64_t or8_u64(uint64_t v) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0xB1, 0xF, 0xF, false);
  hi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0xB1, 0xF, 0xF, false);
  lo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x4E, 0xF, 0xF, false);
  hi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x4E, 0xF, 0xF, false);
  lo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x141, 0xF, 0xF, false);
  hi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, 0x141, 0xF, 0xF, false);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive scan across the wave (Hillis-Steele inside each 16-lane row with
// row_shr 1/2/4/8, then row_bcast 15/31 across rows).  Exact for integers
// whose totals stay below 2^53, which is all it is used for.
__device__ __forceinline__ double wave_incl_scan(double v) {
  v += dpp_d<0x111, 0xF>(v);  // row_shr:1
  v += dpp_d<0x112, 0xF>(v);  // row_shr:2
  v += dpp_d<0x114, 0xF>(v);  // row_shr:4
  v += dpp_d<0x118, 0xF>(v);  // row_shr:8
  v += dpp_d<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_d<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

__device__ __forceinline__ int exponent_of(double p) {  // unbiased exponent of a positive normal
  return (int)((as_u64(p) >> 52) & 0x7FF) - 1023;
}

// Lowest set bit exponent of a positive finite double (its "ulp floor").
__device__ __forceinline__ int lsb_exponent(double a) {
  const uint64_t b = as_u64(a);
  const int ex = (int)((b >> 52) & 0x7FF);
  const uint64_t m = b & 0xFFFFFFFFFFFFFull;
  if (ex == 0) return -1074 + (int)__builtin_ctzll(m);
  return ex - 1075 + (int)__builtin_ctzll(m | (1ull << 52));
}
// Tie-mask bit index of a rate: lsb + 60 for lsb in [-60, 3], else -1.
__device__ __forceinline__ int tie_index(double a) {
  if (!(a > 0.0)) return -1;
  const int L = lsb_exponent(a);
  return ((L >= -60) & (L <= 3)) ? L + 60 : -1;
}
// True if folding a vacancy with this tie mask inside binade e cannot hit a
// rounding tie (no rate is an odd multiple of 2^(e-53)).
__device__ __forceinline__ bool tie_free(uint64_t mask, int e) {
  const int t = e - 53 + 60;
  const bool in_window = (t >= 0) & (t <= 63);
  return in_window & !((mask >> (t & 63)) & 1);
}

// Shift risk of a vacancy whose running sum went from So to Eo over its 8
// slots.  RISK_HARD: the sums do not lie in one normal binade e.  RISK_TIE:
// some rate is an odd multiple of half the binade's ulp u, so a slot rounds
// a tie; a shift by an even multiple of u keeps the parity of every partial
// sum and with it the tie's rounding, an odd one does not.  A later exact
// shift inside the binade reproduces the serial fold unless a flag applies.
enum : uint8_t { RISK_TIE = 1, RISK_HARD = 2 };
__device__ __forceinline__ uint8_t fold_risk(double So, double Eo, uint64_t tm) {
  const int e = exponent_of(So);
  const bool hard = !((So >= 0x1p-969) & (exponent_of(Eo) == e));
  return (uint8_t)(hard ? RISK_HARD : (tie_free(tm, e) ? 0 : RISK_TIE));
}

// ---------------------------------------------------------------- state
constexpr int EVB = 8;  // vacancies evaluated per gather/compute pass

template <int VCAP, typename RateT>
struct SectorLDS {
  using rate_t = RateT;
  uint32_t key[VCAP];      // sorted lattice keys of the sector's vacancies
  int32_t slot[VCAP];      // index into the lattice's global vacancy list
  RateT rate[VCAP * 8];
  double pvac[VCAP];       // running sum after each vacancy
  uint64_t tmask[VCAP];    // tie mask of each vacancy's rates
  uint8_t risk[VCAP];      // RISK_* flags of shifting this vacancy's sums (fold_risk)
  uint8_t own[VCAP];       // flag byte of the vacancy site
  uint8_t nb1[VCAP * 8];   // flag bytes of its 8 1nn sites
  uint8_t sten[EVB][STENCIL_MAX];
  uint8_t cflag[EVB];      // gathered flag byte of each batch vacancy's own site
  int32_t list[64];
  int32_t nv;
  int32_t pad;
};

// The same sector state with the per-vacancy arrays in device memory (any
// capacity: sectors too crowded for LDS) and the per-pass staging arrays in
// LDS.  The helpers below take either form.
struct SectorRef {
  using rate_t = double;
  uint32_t* key;
  int32_t* slot;
  double* rate;
  double* pvac;
  uint64_t* tmask;
  uint8_t* risk;
  uint8_t* own;
  uint8_t* nb1;
  uint8_t (*sten)[STENCIL_MAX];
  uint8_t* cflag;
  int32_t* list;
};

// Tables every engine of a workgroup reads (staged once into LDS).
struct SharedTables {
  StencilTables st;
  double e1[25], e2[25], e0[5];
  float e1f[25], e2f[25], e0f[5];
  uint64_t exp_tab[256];
};

// One lattice as the engine sees it.
struct LatticeView {
  uint8_t* sites;
  uint32_t* vac_key;
  int32_t nvac;
  BoxDims d;
  double nu;
  double kT;
};

struct Patch {
  uint32_t k0, k1;  // keys whose global bytes may be stale (the last hop)
  uint8_t f0, f1;
  int32_t moved;    // sorted index of the vacancy that just hopped (own = f1), or -1
  uint64_t i0, i1;  // site indices of k0 and k1
};

// ---------------------------------------------------------------- sort
// Sorts key[0..n) ascending (sweep order) and slot[] with it, by a bitonic
// network in its all-ascending form: each stage first compares i with its
// mirror i ^ (k - 1), then with i ^ d for d = k/4 .. 1.  Every comparator
// puts the smaller key at the lower index, so positions >= n act as
// +infinity and are never read or written: no padding beyond n.
__device__ __forceinline__ void sort_by_key(uint32_t* key, int32_t* slot, int n) {
  const int l = lane_id();
  int n2 = 1;
  while (n2 < n) n2 <<= 1;
  for (int k = 2; k <= n2; k <<= 1) {
    for (int d = k - 1; d > 0; d = (d == k - 1) ? (k >> 2) : (d >> 1)) {
      for (int i = l; i < n; i += 64) {
        const int p = i ^ d;
        if (p > i && p < n) {
          const uint32_t a = key[i], b = key[p];
          if (a > b) {
            key[i] = b; key[p] = a;
            const int32_t t = slot[i]; slot[i] = slot[p]; slot[p] = t;
          }
        }
      }
      wave_sync();
    }
  }
}

// ---------------------------------------------------------------- searches
// First index i in [0, n) with pred(a[i]), n if none, for a sorted so that
// pred is false and then true.  Each round splits the remaining range into
// 64 chunks and keeps the first chunk whose last element satisfies pred (all
// elements before it fail): two rounds up to 4,096 elements, three up to
// 262,144.
template <typename T, class Pred>
__device__ __forceinline__ int first_true(const T* a, int n, Pred pred) {
  if (n <= 0) return 0;
  const int l = lane_id();
  int lo = 0, len = n;
  while (len > 64) {
    const int S = (len + 63) >> 6;
    const int ci = lo + min(len - 1, l * S + S - 1);
    const int f = first_set(ballot((l * S < len) & pred(a[ci])));
    if (f == 64) return n;  // only possible in the first round
    lo += f * S;
    len = min(S, len - f * S);
  }
  const int g = first_set(ballot((l < len) & pred(a[lo + min(l, len - 1)])));
  return g == 64 ? n : lo + g;
}

// First index i in [0, n) with key[i] >= target, n if none (key sorted).
__device__ __forceinline__ int lower_bound_key(const uint32_t* key, int n, uint32_t target) {
  return first_true(key, n, [=](uint32_t k) { return k >= target; });
}

// First index v in [0, n) with pvac[v] > target, n if none.
__device__ __forceinline__ int first_above(const double* pvac, int n, double target) {
  return first_true(pvac, n, [=](double p) { return p > target; });
}

// ---------------------------------------------------------------- fold
// Serial fold of the cnt <= 8 vacancies v .. v + cnt - 1 from P, in the
// reference order.  Every lane reads the same rates (broadcast LDS reads, all
// issued before the adds) and folds them in registers; lane j writes vacancy
// v + j's running sum (into out) and tie risk.  Returns the sum after the last.
template <class SecT>
__device__ __forceinline__ double fold_serial(SecT& S, int v, int cnt, double P, double* out) {
  const int l = lane_id();
  double pw = 0.0, pw0 = 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    if (j < cnt) {
      double r[8];
#pragma unroll
      for (int b = 0; b < 8; b++) r[b] = (double)S.rate[8 * (v + j) + b];
      const double p0 = P;
#pragma unroll
      for (int b = 0; b < 8; b++) P += r[b];
      if (l == j) { pw = P; pw0 = p0; }
    }
  }
  if (l < cnt) {
    out[v + l] = pw;
    S.risk[v + l] = fold_risk(pw0, pw, S.tmask[v + l]);
  }
  return uni_d(P);
}

// Exact running sums for vacancies [v0, v1) starting from P (the sum before
// v0).  Writes them into out (pvac unless a caller keeps the old sums) with
// their tie risks, and returns the sum after v1 - 1.
template <class SecT>
__device__ __forceinline__ double fold_range(SecT& S, int v0, int v1, double P, double* out) {
  const int l = lane_id();
  int v = uni_i(v0);
  v1 = uni_i(v1);
  P = uni_d(P);
  while (v < v1) {
    if (v1 - v <= 6) {
      // A few vacancies: serially (cheaper than a scan pass).
      P = fold_serial(S, v, v1 - v, P, out);
      break;
    }
    if (!(P >= 0x1p-969)) {
      // Zero or tiny running sum: one vacancy serially.
      double a[8];
#pragma unroll
      for (int b = 0; b < 8; b++) a[b] = (double)S.rate[8 * v + b];
      const uint64_t tm = S.tmask[v];
      double p = P;
#pragma unroll
      for (int b = 0; b < 8; b++) p += a[b];
      p = uni_d(p);
      if (l == 0) {
        out[v] = p;
        S.risk[v] = fold_risk(P, p, tm);
      }
      P = p;
      v++;
      continue;
    }
    const int e = exponent_of(P);
    const double inv_u = __builtin_ldexp(1.0, 52 - e);
    const double u = __builtin_ldexp(1.0, e - 52);
    const double limit = __builtin_ldexp(1.0, e + 1);
    const int vi = v + l;
    const bool active = vi < v1;
    const int vc = active ? vi : v;
    double a[8];
#pragma unroll
    for (int b = 0; b < 8; b++) a[b] = (double)S.rate[8 * vc + b];
#pragma unroll
    for (int b = 0; b < 8; b++) a[b] = active ? a[b] : 0.0;
    double q[8];
    bool bad = false;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      // A rate outside the binade forces the serial path; so does a tie.
      const bool inb = a[b] < limit;
      const double f = inb ? a[b] * inv_u : 0.0;
      const double r = __builtin_rint(f);
      const double fr = f - r;
      bad = bad | !inb | (fr == 0.5) | (fr == -0.5);
      q[b] = inb ? r : 0x1p53;
    }
    // Integer sums below 2^53 are exact in any order: a depth-3 tree.
    const double local = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
    const double incl = wave_incl_scan(local);
    // The running sum grows monotonically, so the lane's last slot decides
    // whether all of its slots stayed inside the binade.
    bad = bad | !(P + incl * u < limit);
    const int fb = first_set(ballot(active & bad));
    if (active & (l < fb)) {
      // All of this vacancy's sums lie in P's binade e: only ties matter.
      out[vi] = P + incl * u;
      S.risk[vi] = tie_free(S.tmask[vi], e) ? (uint8_t)0 : (uint8_t)RISK_TIE;
    }
    if (fb == 64) {
      const int last = min(63, v1 - 1 - v);
      P = P + bcast_d(incl, last) * u;
      v += last + 1;
    } else {
      // The sum before this vacancy is the inclusive sum of the lane before
      // it, which stayed in the binade (below 2^52, so exact); incl - local
      // here would round once the bad lane's own total passes 2^53.
      const double excl = fb == 0 ? 0.0 : bcast_d(incl, fb - 1);
      double p = P + excl * u;
      const int vb = v + fb;
      const uint64_t tm = S.tmask[vb];
      double ab[8];
#pragma unroll
      for (int b = 0; b < 8; b++) ab[b] = (double)S.rate[8 * vb + b];
      const double p0 = p;
#pragma unroll
      for (int b = 0; b < 8; b++) p += ab[b];
      p = uni_d(p);
      if (l == 0) {
        out[vb] = p;
        S.risk[vb] = fold_risk(p0, p, tm);
      }
      P = p;
      v = vb + 1;
    }
  }
  wave_sync();
  return P;
}

// Shifts the running sums of vacancies [v0, n) whose rates did not change
// after the sum before them moved from Po to Pn.  pvac[v] holds the old sums
// for v >= v0 on entry; the new ones go to out[v0, hi), and from hi on they
// equal the old ones.  Returns the new total.  OOP: out is a separate array,
// not pvac; then a region round stores all its shifted sums before it looks
// for the region's end (a store past the end is overwritten later in the
// repair, which writes every position of [v0, hi) in increasing order, or
// lies at or past hi), so the stores overlap the search.
//
// Region by region: with the old sum before v in binade e (S_old) and the
// new one (P_new) in the same binade, D = P_new - S_old is a multiple of
// that binade's ulp, and every following vacancy whose old sums lie in
// [2^e, 2^(e+1) - max(D, 0)) and which carries no tie risk (risk[], kept by
// the folds) shifts by exactly D.  One pass over the region loads, checks
// (two compares per vacancy: the sums are sorted, so the first vacancy
// whose end passes the bound ends the region) and writes 256 vacancies per
// round.  The vacancies where it ends are folded serially, four per LDS
// round trip, until old and new sums share a binade again, and the next
// region starts with the new exact difference.
template <bool OOP, class SecT>
__device__ __forceinline__ double tail_update(SecT& S, int v0, int n, double Pn, double Po, double* out, int& hi) {
  const int l = lane_id();
  int v = uni_i(v0);
  n = uni_i(n);
  double Sold = uni_d(Po);   // old running sum before v
  double Pprev = uni_d(Pn);  // new running sum before v
  bool capped = false;
  hi = v;
  while (v < n) {
    const double D = Pprev - Sold;
    if (D == 0.0) {
      wave_sync();
      hi = v;
      return uni_d(S.pvac[n - 1]);  // the rest is unchanged
    }
    const int e = exponent_of(Sold);
    const bool start_ok = (Sold >= 0x1p-969) & (exponent_of(Pprev) == e) & (Sold + D == Pprev) & (Pprev - D == Sold);
    if (!start_ok && capped) {
      hi = n;
      return fold_range(S, v, n, Pprev, out);
    }
    int stop = v;
    if (start_ok) {
      // D < 2^e here, so top - D is exact, and D / u is an integer below 2^52.
      const double top = __builtin_ldexp(1.0, e + 1);
      const double thr = D > 0.0 ? top - D : top;
      const double half = D * __builtin_ldexp(1.0, 51 - e);  // (D / u) / 2
      const uint8_t rmask = half != __builtin_rint(half) ? (uint8_t)(RISK_TIE | RISK_HARD) : (uint8_t)RISK_HARD;
      stop = n;
      // Lane l takes the four consecutive vacancies c + 4l .. c + 4l + 3.
      for (int c = v; c < n; c += 256) {
        const int i0 = c + 4 * l;
        double x[4];
        uint8_t rk[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int ic = min(i0 + k, n - 1);
          x[k] = S.pvac[ic];
          rk[k] = S.risk[ic];
        }
        if (OOP) {
#pragma unroll
          for (int k = 0; k < 4; k++) out[min(i0 + k, n - 1)] = x[k] + D;
        }
        int kf = 4;  // first of this lane's four that ends the region
#pragma unroll
        for (int k = 3; k >= 0; k--)
          if ((i0 + k < n) & !((x[k] < thr) & ((rk[k] & rmask) == 0))) kf = k;
        const uint64_t m = ballot(kf < 4);
        const int fl = first_set(m);
        const int f = fl == 64 ? n : c + 4 * fl + __builtin_amdgcn_readlane(kf, fl);
        if (!OOP) {
#pragma unroll
          for (int k = 0; k < 4; k++)
            if (i0 + k < f) out[i0 + k] = x[k] + D;
        }
        if (f < n) {
          stop = f;
          break;
        }
      }
      wave_sync();
      if (stop >= n) {
        hi = n;
        return uni_d(out[n - 1]);
      }
    }
    // Exact fold from `stop` until a folded vacancy's old and new sums share
    // a binade again (normally one to three vacancies next to a power of
    // two, or the risky vacancy alone; a large change of the block moves
    // more), at least 64 unless that is reached.  The first vacancy goes
    // serially: every lane reads its eight rates, old sum and tie mask
    // (broadcast LDS reads) and folds in registers; lane 0 writes its new sum
    // and risk.  If the zone goes on, its end is estimated as the first
    // vacancy whose old sum and old sum + D share a binade, and everything up
    // to it is folded in one fold_range; the check after it decides.
    double p = uni_d(stop == v ? Pprev : out[stop - 1]);
    double old_w = 0.0;
    int w = stop;
    bool same = false;
    {
      double r[8];
#pragma unroll
      for (int b = 0; b < 8; b++) r[b] = (double)S.rate[8 * w + b];
      const double ow = S.pvac[w];
      const uint64_t tw = S.tmask[w];
      const double p0 = p;
#pragma unroll
      for (int b = 0; b < 8; b++) p += r[b];
      old_w = ow;
      same = uni_i((exponent_of(old_w) == exponent_of(p)) & (old_w >= 0x1p-969)) != 0;
      if (l == 0) {
        out[w] = p;
        S.risk[w] = fold_risk(p0, p, tw);
      }
      w++;
    }
    const double Dz = Pprev - Sold;
    while (!same && w < n && w - stop < 64) {
      const int i = w + l;
      const int ic = min(i, n - 1);
      const double xo = S.pvac[ic];  // old sums (fold_range below writes [w, w + cnt) only)
      const bool es = (xo >= 0x1p-969) & (exponent_of(xo) == exponent_of(xo + Dz));
      const int j = first_set(ballot((i < n) & (es | (i == n - 1))));
      const int cnt = j == 64 ? 64 : j + 1;
      old_w = bcast_d(xo, cnt - 1);
      p = fold_range(S, w, w + cnt, p, out);
      w += cnt;
      same = (exponent_of(old_w) == exponent_of(p)) & (old_w >= 0x1p-969);
    }
    wave_sync();
    if (w >= n) {
      hi = n;
      return p;
    }
    capped = !same;
    Sold = old_w;
    Pprev = p;
    v = w;
  }
  wave_sync();
  return Pprev;
}

// ---------------------------------------------------------------- rates
// Stencil indices of one lane's direction b = l & 7, packed four per word:
// the 13 counted sites around the vacancy (8 1nn, then 5 2nn) and the 13
// around its target t_b, for both parities, plus t_b itself.  Held in
// registers so the species loads of a rate issue together.
struct LaneStencil {
  uint32_t v[2][4];
  uint32_t t[2][4];
  uint32_t x;  // byte par: stencil index of t_b
  int32_t lin[2];  // by parity: linear site-index offset of stencil site l (no wrap)
  int32_t off[2];  // by parity: (dx, dy, dz) of stencil site l, one signed byte each
};

// Linear offsets of this lane's stencil site (site l, clamped to the
// stencil) for a box: dx + X*(dy + Y*dz), valid when no coordinate wraps.
__device__ __forceinline__ void lane_linear(LaneStencil& R, const StencilTables& st, const BoxDims& d) {
  const int lk = min(lane_id(), st.n_sites - 1);
  for (int par = 0; par < 2; par++) {
    const int dx = st.off[par][lk][0], dy = st.off[par][lk][1], dz = st.off[par][lk][2];
    R.lin[par] = dx + d.X * (dy + d.Y * dz);
    R.off[par] = (int32_t)(((uint32_t)(uint8_t)dx) | ((uint32_t)(uint8_t)dy << 8) | ((uint32_t)(uint8_t)dz << 16));
  }
}

// Site index of this lane's stencil site around the vacancy at key w (a
// wave-uniform key): the vacancy's index plus the linear offset when the
// 7 x 5 x 5 stencil box stays inside the lattice, else wrapped per
// coordinate as key_add does.
__device__ __forceinline__ uint64_t stencil_site(const BoxDims& d, const StencilTables& st, const LaneStencil& LS,
                                                 uint32_t w) {
  const int x = (int)key_x(w), y = (int)key_y(w), z = (int)key_z(w);
  const int par = x & 1;
  const bool inside = (x >= 3) & (x < d.X - 3) & (y >= 2) & (y < d.Y - 2) & (z >= 2) & (z < d.Z - 2);
  if (inside) return (uint64_t)((int64_t)key_index(d, w) + (par ? LS.lin[1] : LS.lin[0]));
  const int32_t o = par ? LS.off[1] : LS.off[0];
  return key_index(d, key_add(d, w, (int8_t)(o & 255), (int8_t)((o >> 8) & 255), (int8_t)((o >> 16) & 255)));
}
__device__ __forceinline__ LaneStencil lane_stencil(const StencilTables& st) {
  const int b = lane_id() & 7;
  LaneStencil R;
  for (int par = 0; par < 2; par++) {
    uint8_t bv[16], bt[16];
    for (int k = 0; k < N1; k++) { bv[k] = st.w1[par][k]; bt[k] = st.t1[par][b][k]; }
    for (int k = 0; k < N2; k++) { bv[N1 + k] = st.w2[par][k]; bt[N1 + k] = st.t2[par][b][k]; }
    for (int k = N1 + N2; k < 16; k++) { bv[k] = 0; bt[k] = 0; }
    for (int q = 0; q < 4; q++) {
      R.v[par][q] = (uint32_t)bv[4 * q] | ((uint32_t)bv[4 * q + 1] << 8) | ((uint32_t)bv[4 * q + 2] << 16) |
                    ((uint32_t)bv[4 * q + 3] << 24);
      R.t[par][q] = (uint32_t)bt[4 * q] | ((uint32_t)bt[4 * q + 1] << 8) | ((uint32_t)bt[4 * q + 2] << 16) |
                    ((uint32_t)bt[4 * q + 3] << 24);
    }
  }
  R.x = (uint32_t)st.w1[0][b] | ((uint32_t)st.w1[1][b] << 8);
  return R;
}

// rate_fp64 (akmc_device.h) for the stencil species sp[] of one vacancy,
// direction b = l & 7, with the same FP64 operations in the same order; the
// loads are staged (all species, then all bond terms) so that they overlap.
__device__ __forceinline__ double rate_fp64_staged(const uint8_t* sp, int par, const LaneStencil& LS,
                                                   const double* e1, const double* e2, const double* e0,
                                                   double nu, double kT, const uint64_t* exp_tab, int& Xout) {
  const int b = lane_id() & 7;
  uint32_t vw[4], tw[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    vw[q] = par ? LS.v[1][q] : LS.v[0][q];
    tw[q] = par ? LS.t[1][q] : LS.t[0][q];
  }
  const int X = sp[par ? (LS.x >> 8) & 255u : LS.x & 255u];
  Xout = X;
  int sv[N1 + N2], st[N1 + N2];
#pragma unroll
  for (int k = 0; k < N1 + N2; k++) {
    sv[k] = sp[(vw[k >> 2] >> (8 * (k & 3))) & 255u];
    st[k] = sp[(tw[k >> 2] >> (8 * (k & 3))) & 255u];
  }
  if (X == SP_V) return 0.0;
  const double* v1 = e1 + SP_V * 5;
  const double* v2 = e2 + SP_V * 5;
  const double* x1 = e1 + X * 5;
  const double* x2 = e2 + X * 5;
  double tsb[N1 + N2], tdb[N1 + N2], tsa[N1 + N2], tda[N1 + N2];
#pragma unroll
  for (int k = 0; k < N1; k++) {
    tsb[k] = v1[sv[k]];
    tdb[k] = x1[st[k]];
    tsa[k] = x1[k == b ? (int)SP_V : sv[k]];
    tda[k] = v1[k == 7 - b ? X : st[k]];
  }
#pragma unroll
  for (int k = N1; k < N1 + N2; k++) {
    tsb[k] = v2[sv[k]];
    tdb[k] = x2[st[k]];
    tsa[k] = x2[sv[k]];
    tda[k] = v2[st[k]];
  }
  double src_b = 0.0, des_b = 0.0, src_a = 0.0, des_a = 0.0;
#pragma unroll
  for (int k = 0; k < N1 + N2; k++) {
    src_b += tsb[k];
    des_b += tdb[k];
    src_a += tsa[k];
    des_a += tda[k];
  }
  const double e_before = src_b + des_b;
  const double e_after = src_a + des_a;
  const double ea = e0[X] + (e_after - e_before) / 2;
  return nu * exp_glibc(-ea / kT, exp_tab);
}

// Evaluates the rates of the vacancies idx[0..n) (sorted indices; idx ==
// nullptr means base + i), as the reference's calcRates evaluates a vacancy:
// 8 direction rates, plus the side effect of deltaE on the vacancy's own
// species tags (a tag equal to an available neighbour's species is dropped,
// vacancy_rates_solver.cpp:81-163).  Global bytes at the patch keys are
// replaced by the patch values (the hop that was just written may not be
// visible yet).  pf/pf_vi: this lane's stencil byte (site index = lane) of
// vacancy pf_vi, gathered earlier by the caller (pf_vi < 0: none).
template <class SecT>
__device__ void eval_list(SecT& S, const SharedTables& T, const LaneStencil& LS, const LatticeView& L,
                          const int32_t* idx, int base, int n, const Patch& pt, bool own_from_lds, uint8_t pf,
                          int pf_vi) {
  const int l = lane_id();
  const int ns = T.st.n_sites;
  for (int bs = 0; bs < n; bs += EVB) {
    const int bc = min(EVB, n - bs);
    // Gather: load i is (vacancy i of the batch, stencil site = lane).
    uint32_t f[EVB];
    // Site indices: the vacancy's index plus this lane's linear stencil
    // offset, unless the stencil wraps around the box (then per coordinate).
    uint64_t qi[EVB];
#pragma unroll
    for (int i = 0; i < EVB; i++) {
      qi[i] = 0;
      if (i < bc) {
        const int vv = uni_i(idx ? idx[bs + i] : base + bs + i);
        const uint32_t w = (uint32_t)uni_i((int)S.key[vv]);
        qi[i] = stencil_site(L.d, T.st, LS, w);
      }
    }
    // Wave-uniform conditions only, so the loads issue back to back.
#pragma unroll
    for (int i = 0; i < EVB; i++) {
      f[i] = pf;
      if (i < bc) {
        const int vv = uni_i(idx ? idx[bs + i] : base + bs + i);
        if (vv != pf_vi) f[i] = L.sites[qi[i]];
      }
    }
#pragma unroll
    for (int i = 0; i < EVB; i++) {
      if (qi[i] == pt.i0) f[i] = pt.f0;
      if (qi[i] == pt.i1) f[i] = pt.f1;
    }
#pragma unroll
    for (int i = 0; i < EVB; i++) {
      if (i < bc && l < ns) {
        const int vv = idx ? idx[bs + i] : base + bs + i;
        S.sten[i][l] = (uint8_t)readout((uint8_t)f[i]);
        if (l >= 1 && l <= 8) S.nb1[vv * 8 + (l - 1)] = (uint8_t)f[i];
        if (l == 0) S.cflag[i] = (uint8_t)f[i];
      }
    }
    wave_sync();
    // Compute: lane = (a, b).
    const int a = l >> 3, b = l & 7;
    const bool act = a < bc;
    const int vi = act ? (idx ? idx[bs + a] : base + bs + a) : 0;
    double r = 0.0;
    int X = SP_V;
    if (act) {
      const int par = (int)(key_x(S.key[vi]) & 1);
      r = rate_fp64_staged(S.sten[a], par, LS, T.e1, T.e2, T.e0, L.nu, L.kT, T.exp_tab, X);
      S.rate[vi * 8 + b] = (typename SecT::rate_t)r;
    }
    // Tie mask of each vacancy: OR of its eight rates' bits, across the eight
    // lanes of the vacancy.
    const int ti = act ? tie_index((double)(typename SecT::rate_t)r) : -1;
    const uint64_t tmv = or8_u64(ti >= 0 ? 1ull << ti : 0ull);
    // Side effect: species present among the available targets of each vacancy.
    const uint64_t mcu = ballot(act & (X == SP_CU));
    const uint64_t mmn = ballot(act & (X == SP_MN));
    const uint64_t mni = ballot(act & (X == SP_NI));
    wave_sync();
    if (act && b == 0) {
      const uint32_t sh = 8u * (uint32_t)a;
      const uint8_t clr = (uint8_t)((((mcu >> sh) & 0xFF) ? BIT_CU : 0) | (((mmn >> sh) & 0xFF) ? BIT_MN : 0) |
                                    (((mni >> sh) & 0xFF) ? BIT_NI : 0));
      // The vacancy's own flags: after a hop, the LDS copy (kept current by
      // this engine) or the patch value for the vacancy that just arrived; at
      // a rebuild, the gathered byte.
      const uint8_t f0 = vi == pt.moved ? pt.f1 : (own_from_lds ? S.own[vi] : S.cflag[a]);
      const uint8_t nf = (uint8_t)(f0 & ~clr);
      if (nf != f0) L.sites[key_index(L.d, S.key[vi])] = nf;
      S.own[vi] = nf;
      S.tmask[vi] = tmv;
    }
    wave_sync();
  }
}

// eval_list for one vacancy: the one that just hopped (key `to`, own flags
// pt.f1), which is the only vacancy re-evaluated after most hops.  Its
// stencil bytes were gathered by the caller right after the hop (pf, lane =
// stencil site, site index pq = stencil_site(to)), so there are no index
// loads and no per-slot loops; the operations on the rates and the side
// effect are eval_list's.  Split in two: the rates into registers (nothing
// depends on the vacancy's sorted index yet) and the stores at its index vi.
struct HoppedEval {
  double r;            // lane b < 8: rate of direction b
  uint64_t tmv;        // tie mask of the eight rates
  uint8_t clr;         // species tags to drop from the vacancy's own byte
  uint8_t f;           // lane 1..8: the 1nn flag bytes
};
template <class SecT>
__device__ __forceinline__ HoppedEval eval_hopped_rates(SecT& S, const SharedTables& T, const LaneStencil& LS,
                                                        const LatticeView& L, uint32_t to, const Patch& pt,
                                                        uint8_t pf, uint64_t pq) {
  const int l = lane_id();
  uint32_t f = pf;
  if (pq == pt.i0) f = pt.f0;
  if (pq == pt.i1) f = pt.f1;
  if (l < T.st.n_sites) S.sten[0][l] = (uint8_t)readout((uint8_t)f);
  wave_sync();
  const bool act = l < 8;
  double r = 0.0;
  int X = SP_V;
  if (act) r = rate_fp64_staged(S.sten[0], (int)(key_x(to) & 1), LS, T.e1, T.e2, T.e0, L.nu, L.kT, T.exp_tab, X);
  const int ti = act ? tie_index((double)(typename SecT::rate_t)r) : -1;
  HoppedEval h;
  h.r = r;
  h.tmv = or8_u64(ti >= 0 ? 1ull << ti : 0ull);
  const uint64_t mcu = ballot(act & (X == SP_CU));
  const uint64_t mmn = ballot(act & (X == SP_MN));
  const uint64_t mni = ballot(act & (X == SP_NI));
  h.clr = (uint8_t)((mcu ? BIT_CU : 0) | (mmn ? BIT_MN : 0) | (mni ? BIT_NI : 0));
  h.f = (uint8_t)f;
  return h;
}
template <class SecT>
__device__ __forceinline__ void eval_hopped_store(SecT& S, const LatticeView& L, int vi, uint32_t to, const Patch& pt,
                                                  const HoppedEval& h) {
  const int l = lane_id();
  if (l < 8) S.rate[vi * 8 + l] = (typename SecT::rate_t)h.r;
  if (l >= 1 && l <= 8) S.nb1[vi * 8 + (l - 1)] = h.f;
  if (l == 0) {
    const uint8_t f0 = pt.f1;
    const uint8_t nf = (uint8_t)(f0 & ~h.clr);
    if (nf != f0) L.sites[key_index(L.d, to)] = nf;
    S.own[vi] = nf;
    S.tmask[vi] = h.tmv;
  }
  wave_sync();
}
template <class SecT>
__device__ __forceinline__ void eval_hopped(SecT& S, const SharedTables& T, const LaneStencil& LS,
                                            const LatticeView& L, int vi, uint32_t to, const Patch& pt, uint8_t pf,
                                            uint64_t pq) {
  const HoppedEval h = eval_hopped_rates(S, T, LS, L, to, pt, pf, pq);
  eval_hopped_store(S, L, vi, to, pt, h);
}

}  // namespace akmc
