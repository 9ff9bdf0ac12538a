// One-wavefront sector engine: the vacancy-hop event loop of one lattice's
// current sector, run by a single 64-lane wavefront on LDS-resident state.
//
// The event chain of one lattice is serial (every hop changes the next
// rates), so the engine is written for latency, not throughput: one wave, no
// workgroup barriers, cross-lane work through ballots and shuffles, loads
// issued together and consumed late.  The plain path runs one engine on the
// whole GPU; the batched path runs one engine per replica, many per CU.
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
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ int shfl_i(int v, int src) { return __shfl(v, src, 64); }

__device__ __forceinline__ double wave_incl_scan(double v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_up(v, o, 64);
    if (l >= o) v += y;
  }
  return v;
}

__device__ __forceinline__ int exponent_of(double p) {  // biased-free exponent of a positive normal
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
// Tie-mask bit of a rate: bit (lsb + 60) for lsb in [-60, 3].
__device__ __forceinline__ uint64_t tie_bit(double a) {
  if (!(a > 0.0)) return 0;
  const int L = lsb_exponent(a);
  return (L >= -60 && L <= 3) ? (1ull << (L + 60)) : 0;
}
// True if folding a vacancy with this tie mask inside binade e cannot hit a
// rounding tie (no rate is an odd multiple of 2^(e-53)).
__device__ __forceinline__ bool tie_free(uint64_t mask, int e) {
  const int t = e - 53 + 60;
  if (t < 0 || t > 63) return false;
  return !((mask >> t) & 1);
}

// ---------------------------------------------------------------- state
constexpr int EVB = 8;  // vacancies evaluated per gather/compute pass

template <int VCAP, typename RateT>
struct SectorLDS {
  uint32_t key[VCAP];      // sorted lattice keys of the sector's vacancies
  int32_t slot[VCAP];      // index into the lattice's global vacancy list
  RateT rate[VCAP * 8];
  double pvac[VCAP];       // running sum after each vacancy
  uint64_t tmask[VCAP];    // tie mask of each vacancy's rates
  uint8_t own[VCAP];       // flag byte of the vacancy site
  uint8_t nb1[VCAP * 8];   // flag bytes of its 8 1nn sites
  uint8_t sten[EVB][STENCIL_MAX];
  uint8_t cflag[EVB];      // gathered flag byte of each batch vacancy's own site
  int32_t list[64];
  int32_t nv;
  int32_t pad;
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
};

// ---------------------------------------------------------------- searches
// First index i in [0, n) with key[i] >= target, n if none (key sorted).
template <int VCAP>
__device__ int lower_bound_key(const uint32_t* key, int n, uint32_t target) {
  if (n <= 0) return 0;
  const int l = lane_id();
  const int S = (n + 63) >> 6;
  const int ci = min(n - 1, l * S + S - 1);
  const int f = first_set(ballot(l * S < n && key[ci] >= target));
  if (f == 64) return n;
  const int fi = f * S + l;
  const int g = first_set(ballot(l < S && fi < n && key[min(fi, n - 1)] >= target));
  return g == 64 ? min(n, (f + 1) * S) : f * S + g;
}

// First index v in [0, n) with pvac[v] > target, n if none.
__device__ __forceinline__ int first_above(const double* pvac, int n, double target) {
  if (n <= 0) return 0;
  const int l = lane_id();
  const int S = (n + 63) >> 6;
  const int ci = min(n - 1, l * S + S - 1);
  const int f = first_set(ballot(l * S < n && pvac[ci] > target));
  if (f == 64) return n;
  const int fi = f * S + l;
  const int g = first_set(ballot(l < S && fi < n && pvac[min(fi, n - 1)] > target));
  return g == 64 ? min(n, (f + 1) * S) : f * S + g;
}

// ---------------------------------------------------------------- fold
// Exact running sums for vacancies [v0, v1) starting from P (the sum before
// v0).  Writes pvac and returns the sum after v1 - 1.
template <int VCAP, typename RateT>
__device__ double fold_range(SectorLDS<VCAP, RateT>& S, int v0, int v1, double P) {
  const int l = lane_id();
  int v = v0;
  while (v < v1) {
    if (!(P >= 0x1p-969)) {
      // Zero (or tiny) running sum: fold one vacancy serially.
      double p = P;
#pragma unroll
      for (int b = 0; b < 8; b++) p += (double)S.rate[8 * v + b];
      if (l == 0) S.pvac[v] = p;
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
    for (int b = 0; b < 8; b++) a[b] = active ? (double)S.rate[8 * vc + b] : 0.0;
    double q[8];
    double local = 0.0;
    int bad = 8;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      double r = 0.0;
      if (!(a[b] < limit)) {
        r = 0x1p53;
        bad = min(bad, b);
      } else {
        const double f = a[b] * inv_u;
        r = __builtin_rint(f);
        const double fr = f - r;
        if (fr == 0.5 || fr == -0.5) bad = min(bad, b);
      }
      q[b] = r;
      local += r;
    }
    const double incl = wave_incl_scan(local);
    double R = incl - local;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      R += q[b];
      if (b < bad && !(P + R * u < limit)) bad = b;
    }
    const int fb = first_set(ballot(active && bad < 8));
    if (active && l < fb) S.pvac[vi] = P + incl * u;
    if (fb == 64) {
      const int last = min(63, v1 - 1 - v);
      P = P + shfl_d(incl, last) * u;
      v += last + 1;
    } else {
      const double excl = shfl_d(incl - local, fb);
      double p = P + excl * u;  // exact: every slot before this vacancy stayed in the binade
      const int vb = v + fb;
#pragma unroll
      for (int b = 0; b < 8; b++) p += (double)S.rate[8 * vb + b];
      if (l == 0) S.pvac[vb] = p;
      P = p;
      v = vb + 1;
    }
  }
  wave_sync();
  return P;
}

// Shifts the running sums of vacancies [v0, n) whose rates did not change
// after the sums before them moved from Po to Pn (pvac[v0 - 1] new = Pn, old
// = Po).  pvac[v] still holds old values on entry for v >= v0.  Returns the
// new total (pvac[n - 1]).
template <int VCAP, typename RateT>
__device__ double tail_update(SectorLDS<VCAP, RateT>& S, int v0, int n, double Pn, double Po) {
  const int l = lane_id();
  double prev_old = Po, prev_new = Pn;
  double D = Pn - Po;
  int v = v0;
  if (!(Po + D == Pn && Pn - D == Po)) {
    // The shift itself is not exact: fold the tail exactly.
    return n > v0 ? fold_range<VCAP, RateT>(S, v0, n, Pn) : Pn;
  }
  while (v < n && D != 0.0) {
    const int vi = v + l;
    const bool active = vi < n;
    const double Eo = active ? S.pvac[vi] : 0.0;
    const uint64_t tm = active ? S.tmask[vi] : 0;
    double So = __shfl_up(Eo, 1, 64);
    if (l == 0) So = prev_old;
    const double Sn = So + D, En = Eo + D;
    const int e = exponent_of(So);
    const bool ok = So >= 0x1p-969 && exponent_of(Eo) == e && exponent_of(Sn) == e && exponent_of(En) == e &&
                    tie_free(tm, e);
    const int fb = first_set(ballot(active && !ok));
    if (active && l < fb) S.pvac[vi] = En;
    if (fb == 64) {
      const int last = min(63, n - 1 - v);
      prev_old = shfl_d(Eo, last);
      prev_new = shfl_d(En, last);
      v += last + 1;
    } else {
      // Exact fold of the offending vacancy from its new start.
      const int vb = v + fb;
      double p = fb == 0 ? prev_new : shfl_d(En, fb - 1);
#pragma unroll
      for (int b = 0; b < 8; b++) p += (double)S.rate[8 * vb + b];
      const double Eob = shfl_d(Eo, fb);
      if (l == 0) S.pvac[vb] = p;
      prev_old = Eob;
      prev_new = p;
      D = p - Eob;
      v = vb + 1;
      if (!(Eob + D == p && p - D == Eob)) {
        wave_sync();
        return v < n ? fold_range<VCAP, RateT>(S, v, n, p) : p;
      }
    }
  }
  wave_sync();
  if (v >= n) return prev_new;
  // D == 0: the remaining sums are unchanged.
  return S.pvac[n - 1];
}

// ---------------------------------------------------------------- rates
// Evaluates the rates of the vacancies idx[0..n) (sorted indices; idx ==
// nullptr means base + i), as the
// reference's calcRates evaluates a vacancy: 8 direction rates, plus the
// side effect of deltaE on the vacancy's own species tags (a tag equal to an
// available neighbour's species is dropped, vacancy_rates_solver.cpp:81-163).
// Global bytes at the patch keys are replaced by the patch values (the hop
// that was just written may not be visible yet).
template <int VCAP, typename RateT, bool FP32>
__device__ void eval_list(SectorLDS<VCAP, RateT>& S, const SharedTables& T, const LatticeView& L, const int32_t* idx,
                          int base, int n, const Patch& pt, bool own_from_lds) {
  const int l = lane_id();
  const int ns = T.st.n_sites;
  for (int bs = 0; bs < n; bs += EVB) {
    const int bc = min(EVB, n - bs);
    // Gather: pairs (a, k) = (vacancy in batch, stencil site), 7 per lane.
    constexpr int PER = (EVB * STENCIL_MAX + 63) / 64;
    uint8_t f[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int pidx = l + 64 * i;
      const int a = pidx / STENCIL_MAX, k = pidx % STENCIL_MAX;
      f[i] = 0;
      if (a < bc && k < ns) {
        const uint32_t w = S.key[idx ? idx[bs + a] : base + bs + a];
        const int par = (int)(key_x(w) & 1);
        const uint32_t q = key_add(L.d, w, T.st.off[par][k][0], T.st.off[par][k][1], T.st.off[par][k][2]);
        f[i] = L.sites[key_index(L.d, q)];
        if (q == pt.k0) f[i] = pt.f0;
        if (q == pt.k1) f[i] = pt.f1;
      }
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int pidx = l + 64 * i;
      const int a = pidx / STENCIL_MAX, k = pidx % STENCIL_MAX;
      if (a < bc && k < ns) {
        const int vv = idx ? idx[bs + a] : base + bs + a;
        S.sten[a][k] = (uint8_t)readout(f[i]);
        if (k >= 1 && k <= 8) S.nb1[vv * 8 + (k - 1)] = f[i];
        if (k == 0) S.cflag[a] = f[i];
      }
    }
    wave_sync();
    // Compute: lane = (a, b).
    const int a = l >> 3, b = l & 7;
    const bool act = a < bc;
    const int vi = act ? (idx ? idx[bs + a] : base + bs + a) : 0;
    double r = 0.0;
    uint32_t clr = 0;
    if (act) {
      const int par = (int)(key_x(S.key[vi]) & 1);
      if constexpr (FP32) {
        r = 0.0;  // replaced by the FP32 path of the batched engine
      } else {
        r = rate_fp64(S.sten[a], par, b, T.st, T.e1, T.e2, T.e0, L.nu, L.kT, T.exp_tab);
      }
      clr = species_bit(S.sten[a][T.st.w1[par][b]]);
      S.rate[vi * 8 + b] = (RateT)r;
    }
    uint64_t tb = tie_bit((double)(RateT)r);
    clr |= __shfl_xor(clr, 1, 64);
    clr |= __shfl_xor(clr, 2, 64);
    clr |= __shfl_xor(clr, 4, 64);
    tb |= __shfl_xor(tb, 1, 64);
    tb |= __shfl_xor(tb, 2, 64);
    tb |= __shfl_xor(tb, 4, 64);
    if (act && b == 0) {
      // The vacancy's own flags: after a hop, the LDS copy (kept current by
      // this engine) or the patch value for the vacancy that just arrived; at
      // a rebuild, the gathered byte.
      const uint8_t f0 = vi == pt.moved ? pt.f1 : (own_from_lds ? S.own[vi] : S.cflag[a]);
      const uint8_t nf = (uint8_t)(f0 & ~clr);
      if (nf != f0) L.sites[key_index(L.d, S.key[vi])] = nf;
      S.own[vi] = nf;
      S.tmask[vi] = tb;
    }
    wave_sync();
  }
}

}  // namespace akmc
