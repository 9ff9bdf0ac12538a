// Replica engine: a persistent kernel in which every wavefront advances one
// replica (an independent lattice) at a time through the vacancy-hop loop.
//
// Per replica and event (the reference's order, sublattice.inl:61-179):
//   sector start : collect the sector's vacancies (in vacancy-list order),
//                  evaluate all their rates, build block sums;
//   event        : total = sum of block sums; u * total located by three
//                  wave scans (superblocks, blocks, vacancies) and an 8-rate
//                  walk; the hop is performed (ABVIModel::perform,
//                  models/abvi/kmc.cpp:358-396); dt = -log(r) / total;
//   update       : the vacancies whose 55-site stencil holds `from` or `to`
//                  are re-evaluated and their blocks re-summed.
// Blocks whose z range misses the hop are skipped without reading their
// vacancies, so an update touches O(affected) entries.
//
// Two kinds of launch: LDS-resident (the replica's sector arrays are copied
// into LDS for the work item and back at its end; replicas with at most
// `vcap` vacancies) and global (the arrays are used in place in HBM).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "akmc_batch.h"
#include "akmc_exp_table.h"
#include "akmc_wave.h"

namespace akmc {

__constant__ StencilTables r_sten;
__constant__ uint64_t r_exp_tab[256] = AKMC_EXP_TAB_INIT;
__constant__ double r_e1[25] = AKMC_E1_TABLE;
__constant__ double r_e2[25] = AKMC_E2_TABLE;
__constant__ double r_e0[5] = AKMC_E0_TABLE;

// Sector id at ring position p: 0,7,2,5,3,4,1,6 (SectorMeta, sublattice.h:21-30).
__device__ __forceinline__ int ring_sector(int p) { return (int)((0x61435270u >> (4 * p)) & 0xFu); }

struct RTabs {
  StencilTables st;
  double e1[25], e2[25], e0[5];
  float e1q[25], e2q[25], e0q2[5];
  uint64_t exp_tab[256];
};

struct RStage {
  float tab[256 + BATCH_NHI_MAX];  // FP32 exp table of the replica: lo[256], hi[nhi]
  double evo[8];
  int32_t cnt[8];                  // vacancies per sector
  uint8_t sten[BATCH_EVB][64];
  int32_t list[64];
  uint8_t cflag[BATCH_EVB];
  uint8_t pad[8];
};
constexpr int RSTAGE_BYTES = (int)((sizeof(RStage) + 15) & ~(size_t)15);

template <typename RateT>
struct Sec {
  RateT* rate;
  double* vt;
  double* bsum;
  double* ssum;
  uint32_t* key;
  int32_t* slot;
  uint32_t* bz;  // z range of each block's vacancies: zmin | zmax << 16
  uint8_t* own;
  uint8_t* nb1;
};

template <typename RateT>
__device__ __forceinline__ Sec<RateT> sec_at(uint8_t* base, int64_t cap) {
  const WsLayout w = ws_layout(cap, (int)sizeof(RateT));
  return Sec<RateT>{reinterpret_cast<RateT*>(base + w.rate), reinterpret_cast<double*>(base + w.vt),
                    reinterpret_cast<double*>(base + w.bsum), reinterpret_cast<double*>(base + w.ssum),
                    reinterpret_cast<uint32_t*>(base + w.key), reinterpret_cast<int32_t*>(base + w.slot),
                    reinterpret_cast<uint32_t*>(base + w.bz), base + w.own, base + w.nb1};
}

struct RView {
  uint8_t* sites;
  uint32_t* vac_key;
  int32_t nvac;
  BoxDims d;
  double nu, kT;
  int32_t nhi, m_min;
  int32_t nbx, nby;
};

// Device address of the site with key k (brick layout, akmc_batch.h).
__device__ __forceinline__ uint64_t site_at(const RView& R, uint32_t k) {
  return (uint64_t)brick_index(R.nbx, R.nby, key_x(k), key_y(k), key_z(k));
}

struct ErrAcc {
  double emax, esum;
  int64_t n, tiny;
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
// xoshiro256** (Blackman & Vigna 2018).
__device__ __forceinline__ uint64_t xo_next(uint64_t* s) {
  const uint64_t r = rotl64(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return r;
}

__device__ __forceinline__ double shfl_d(double v, int src) {
  const uint64_t b = as_u64(v);
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)b);
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(b >> 32));
  return as_f64(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, shfl_d(v, lane_id() ^ o));
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += shfl_d(v, lane_id() ^ o);
  return v;
}

// FP32 rate of direction b.  The bond energies are integer multiples of the
// quantum q, held as integer-valued floats, so the four bond counts and dE
// (at most 52 terms of magnitude below 2^12) are exact in FP32.  The barrier
// index m = (2*e0(X) + dE) / q is an integer and the rate
// nu * exp(-m * q / (2 kT)) is the product of two FP32 table entries,
// tab_hi[m >> 8] * tab_lo[m & 255] (m counted from m_min).
__device__ __forceinline__ float rate_fp32q(const uint8_t* sp, int par, int b, const StencilTables& S,
                                            const float* e1q, const float* e2q, const float* e0q2, const float* tab,
                                            int nhi, int m_min, bool& bad) {
  const int X = sp[S.w1[par][b]];
  if (X == SP_V) return 0.0f;
  float src_b = 0.f, des_b = 0.f, src_a = 0.f, des_a = 0.f;
#pragma unroll
  for (int k = 0; k < N1; k++) src_b += e1q[SP_V * 5 + sp[S.w1[par][k]]];
#pragma unroll
  for (int k = 0; k < N2; k++) src_b += e2q[SP_V * 5 + sp[S.w2[par][k]]];
#pragma unroll
  for (int k = 0; k < N1; k++) des_b += e1q[X * 5 + sp[S.t1[par][b][k]]];
#pragma unroll
  for (int k = 0; k < N2; k++) des_b += e2q[X * 5 + sp[S.t2[par][b][k]]];
#pragma unroll
  for (int k = 0; k < N1; k++) src_a += e1q[X * 5 + (k == b ? (int)SP_V : (int)sp[S.w1[par][k]])];
#pragma unroll
  for (int k = 0; k < N2; k++) src_a += e2q[X * 5 + sp[S.w2[par][k]]];
#pragma unroll
  for (int k = 0; k < N1; k++) des_a += e1q[SP_V * 5 + (k == 7 - b ? X : (int)sp[S.t1[par][b][k]])];
#pragma unroll
  for (int k = 0; k < N2; k++) des_a += e2q[SP_V * 5 + sp[S.t2[par][b][k]]];
  const float dE = (src_a + des_a) - (src_b + des_b);
  const int m = (int)(e0q2[X] + dE) - m_min;
  const bool ok = (m >= 0) & ((m >> 8) < nhi);
  const int mc = ok ? m : 0;
  const float r = tab[256 + (mc >> 8)] * tab[mc & 255];
  // Outside the table, or a barrier so negative that the rate overflows FP32.
  bad = bad | !ok | !(r <= 3.40282347e38f);
  return r;
}

// Evaluates vacancies idx[0..n) (list indices; idx == nullptr: base + i):
// 8 rates, the per-vacancy total, the 1nn flag bytes, and the side effect on
// the vacancy's own stale species tags (vacancy_rates_solver.cpp:81-163).
template <typename RateT, bool FP32>
__device__ void eval_rep(const Sec<RateT>& S, RStage& G, const RTabs& T, const RView& R, const int32_t* idx, int base,
                         int n, const Patch& pt, bool own_from_sec, uint8_t pf, int pf_vi, int flags, ErrAcc& E,
                         bool& bad) {
  const int l = lane_id();
  const int ns = T.st.n_sites;
  for (int bs = 0; bs < n; bs += BATCH_EVB) {
    const int bc = min(BATCH_EVB, n - bs);
    uint32_t f[BATCH_EVB];
    uint32_t qk[BATCH_EVB];
    const int lk = min(l, ns - 1);
#pragma unroll
    for (int i = 0; i < BATCH_EVB; i++) {
      const int vv = idx ? idx[bs + min(i, bc - 1)] : base + bs + min(i, bc - 1);
      const uint32_t w = S.key[vv];
      const int par = (int)(key_x(w) & 1);
      qk[i] = key_add(R.d, w, T.st.off[par][lk][0], T.st.off[par][lk][1], T.st.off[par][lk][2]);
    }
#pragma unroll
    for (int i = 0; i < BATCH_EVB; i++) {
      const int vv = idx ? idx[bs + min(i, bc - 1)] : base + bs + min(i, bc - 1);
      f[i] = pf;
      if (i < bc && vv != pf_vi) f[i] = R.sites[site_at(R, qk[i])];
    }
#pragma unroll
    for (int i = 0; i < BATCH_EVB; i++) {
      if (qk[i] == pt.k0) f[i] = pt.f0;
      if (qk[i] == pt.k1) f[i] = pt.f1;
    }
#pragma unroll
    for (int i = 0; i < BATCH_EVB; i++) {
      if (i < bc && l < ns) {
        const int vv = idx ? idx[bs + i] : base + bs + i;
        G.sten[i][l] = (uint8_t)readout((uint8_t)f[i]);
        if (l >= 1 && l <= 8) S.nb1[vv * 8 + (l - 1)] = (uint8_t)f[i];
        if (l == 0) G.cflag[i] = (uint8_t)f[i];
      }
    }
    wave_sync();
    const int a = l >> 3, b = l & 7;
    const bool act = a < bc;
    const int vi = act ? (idx ? idx[bs + a] : base + bs + a) : 0;
    double r = 0.0;
    int X = SP_V;
    if (act) {
      const int par = (int)(key_x(S.key[vi]) & 1);
      X = G.sten[a][T.st.w1[par][b]];
      if constexpr (FP32) {
        const float rf = rate_fp32q(G.sten[a], par, b, T.st, T.e1q, T.e2q, T.e0q2, G.tab, R.nhi, R.m_min, bad);
        r = (double)rf;
        if (flags & AKMC_BF_ERROR_STATS) {
          const double r64 = rate_fp64(G.sten[a], par, b, T.st, T.e1, T.e2, T.e0, R.nu, R.kT, T.exp_tab);
          if (r64 >= 1.1754943508222875e-38) {
            const double rel = fabs(r - r64) / r64;
            E.emax = fmax(E.emax, rel);
            E.esum += rel;
            E.n += 1;
          } else if (r64 > 0.0) {
            E.tiny += 1;
          }
        }
        S.rate[vi * 8 + b] = (RateT)rf;
      } else {
        r = rate_fp64(G.sten[a], par, b, T.st, T.e1, T.e2, T.e0, R.nu, R.kT, T.exp_tab);
        S.rate[vi * 8 + b] = (RateT)r;
      }
    }
    // Per-vacancy total, a fixed-order FP64 tree over the 8 rates.
    double rr[8];
#pragma unroll
    for (int k = 0; k < 8; k++) rr[k] = shfl_d(r, (l & ~7) + k);
    const double vt = ((rr[0] + rr[1]) + (rr[2] + rr[3])) + ((rr[4] + rr[5]) + (rr[6] + rr[7]));
    const uint64_t mcu = ballot(act & (X == SP_CU));
    const uint64_t mmn = ballot(act & (X == SP_MN));
    const uint64_t mni = ballot(act & (X == SP_NI));
    if (act && b == 0) {
      const uint32_t sh = 8u * (uint32_t)a;
      const uint8_t clr = (uint8_t)((((mcu >> sh) & 0xFF) ? BIT_CU : 0) | (((mmn >> sh) & 0xFF) ? BIT_MN : 0) |
                                    (((mni >> sh) & 0xFF) ? BIT_NI : 0));
      const uint8_t f0 = vi == pt.moved ? pt.f1 : (own_from_sec ? S.own[vi] : G.cflag[a]);
      const uint8_t nf = (uint8_t)(f0 & ~clr);
      if (nf != f0) R.sites[site_at(R, S.key[vi])] = nf;
      S.own[vi] = nf;
      S.vt[vi] = vt;
    }
    wave_sync();
  }
}

// Stencil indices that lane (b = lane >> 3, p = lane & 7) reads in the
// single-vacancy FP32 evaluation, for both parities: bytes w1[p], t1[b][p],
// w2[min(p,4)], t2[b][min(p,4)] in idx[par], and w1[b] in xidx (byte par).
struct LaneIdx {
  uint32_t idx[2];
  uint32_t xidx;
};
__device__ __forceinline__ LaneIdx lane_indices(const StencilTables& st) {
  const int l = lane_id(), b = l >> 3, p = l & 7, q = min(p, 4);
  LaneIdx L;
  for (int par = 0; par < 2; par++)
    L.idx[par] = (uint32_t)st.w1[par][p] | ((uint32_t)st.t1[par][b][p] << 8) | ((uint32_t)st.w2[par][q] << 16) |
                 ((uint32_t)st.t2[par][b][q] << 24);
  L.xidx = (uint32_t)st.w1[0][b] | ((uint32_t)st.w1[1][b] << 8);
  return L;
}

__device__ __forceinline__ int bperm_i(int src_lane, int v) { return __builtin_amdgcn_ds_bpermute(src_lane << 2, v); }
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// FP32 evaluation of one vacancy (list index vi) whose stencil bytes are
// already in registers: f = flag byte of stencil site `lane` (lanes < the
// stencil size).  All 64 lanes work: lane (b, p) takes the p-th 1nn and 2nn
// terms of the four bond counts of direction b, and the eight lanes of b sum
// their parts.  The terms are integer multiples of q held exactly in FP32,
// so the sum is the same in any order -- the same dE as the term-by-term
// evaluation (rate_fp32q).  Writes the 8 rates, the total, the 1nn flag
// bytes and the stale-tag side effect like eval_rep.
template <typename RateT>
__device__ void eval_one(const Sec<RateT>& S, RStage& G, const RTabs& T, const RView& R, const LaneIdx& LI, int vi,
                         uint32_t f, const Patch& pt, bool own_from_sec, int flags, ErrAcc& E, bool& bad) {
  const int l = lane_id(), b = l >> 3, p = l & 7;
  const uint32_t w = S.key[vi];
  const int par = (int)(key_x(w) & 1);
  const int spl = readout((uint8_t)f);
  const uint32_t P = LI.idx[par];
  const int s_w1 = bperm_i((int)(P & 255u), spl);
  const int s_t1 = bperm_i((int)((P >> 8) & 255u), spl);
  const int s_w2 = bperm_i((int)((P >> 16) & 255u), spl);
  const int s_t2 = bperm_i((int)(P >> 24), spl);
  const int X = bperm_i((int)((LI.xidx >> (8 * par)) & 255u), spl);
  const int Xc = X == SP_V ? SP_FE : X;  // a vacancy target has rate 0; any row keeps the lookups in range
  // src_a - src_b + des_a - des_b, 1nn term k = p.
  const int a_src = p == b ? (int)SP_V : s_w1;
  const int a_des = p == 7 - b ? Xc : s_t1;
  float d = (T.e1q[Xc * 5 + a_src] + T.e1q[SP_V * 5 + a_des]) - (T.e1q[SP_V * 5 + s_w1] + T.e1q[Xc * 5 + s_t1]);
  // 2nn term k = p (five of them).
  const float d2 = (T.e2q[Xc * 5 + s_w2] + T.e2q[SP_V * 5 + s_t2]) - (T.e2q[SP_V * 5 + s_w2] + T.e2q[Xc * 5 + s_t2]);
  d += p < 5 ? d2 : 0.0f;
  d += dpp_f<0xB1>(d);   // quad_perm [1,0,3,2]
  d += dpp_f<0x4E>(d);   // quad_perm [2,3,0,1]
  d += dpp_f<0x141>(d);  // row_half_mirror: the other quad of the eight lanes
  const int m = (int)(T.e0q2[Xc] + d) - R.m_min;
  const bool ok = (m >= 0) & ((m >> 8) < R.nhi);
  const int mc = ok ? m : 0;
  const float rf32 = G.tab[256 + (mc >> 8)] * G.tab[mc & 255];
  const bool vac = X == SP_V;
  bad = bad | (!vac & (!ok | !(rf32 <= 3.40282347e38f)));
  const float rf = vac ? 0.0f : rf32;
  double r = (double)rf;
  if (flags & AKMC_BF_ERROR_STATS) {
    if (l < T.st.n_sites) G.sten[0][l] = (uint8_t)spl;
    wave_sync();
    if (p == 0 && !vac) {
      const double r64 = rate_fp64(G.sten[0], par, b, T.st, T.e1, T.e2, T.e0, R.nu, R.kT, T.exp_tab);
      if (r64 >= 1.1754943508222875e-38) {
        const double rel = fabs(r - r64) / r64;
        E.emax = fmax(E.emax, rel);
        E.esum += rel;
        E.n += 1;
      } else if (r64 > 0.0) {
        E.tiny += 1;
      }
    }
    wave_sync();
  }
  if (p == 0) S.rate[vi * 8 + b] = (RateT)rf;
  if (l >= 1 && l <= 8) S.nb1[vi * 8 + (l - 1)] = (uint8_t)f;
  double rr[8];
#pragma unroll
  for (int k = 0; k < 8; k++) rr[k] = bcast_d(r, 8 * k);
  const double vt = ((rr[0] + rr[1]) + (rr[2] + rr[3])) + ((rr[4] + rr[5]) + (rr[6] + rr[7]));
  const uint64_t sel = 0x0101010101010101ull;  // lane p == 0 of each direction
  const bool cu = (ballot((p == 0) & (X == SP_CU)) & sel) != 0;
  const bool mn = (ballot((p == 0) & (X == SP_MN)) & sel) != 0;
  const bool ni = (ballot((p == 0) & (X == SP_NI)) & sel) != 0;
  const uint32_t f_own = (uint32_t)__builtin_amdgcn_readlane((int)f, 0);  // stencil site 0: the vacancy
  if (l == 0) {
    const uint8_t clr = (uint8_t)((cu ? BIT_CU : 0) | (mn ? BIT_MN : 0) | (ni ? BIT_NI : 0));
    const uint8_t f0 = vi == pt.moved ? pt.f1 : (own_from_sec ? S.own[vi] : (uint8_t)f_own);
    const uint8_t nf = (uint8_t)(f0 & ~clr);
    if (nf != f0) R.sites[site_at(R, w)] = nf;
    S.own[vi] = nf;
    S.vt[vi] = vt;
  }
  wave_sync();
}

// Block j's sum (and, at a rebuild, its z range) from the per-vacancy totals.
template <typename RateT>
__device__ void block_sum(const Sec<RateT>& S, int j, int nv, bool with_z) {
  const int l = lane_id();
  const int i = 64 * j + l;
  const bool in = i < nv;
  const double x = in ? S.vt[i] : 0.0;
  const double tot = bcast_d(wave_incl_scan(x), 63);
  if (with_z) {
    const bool live = in && S.slot[i] >= 0;
    const int z = live ? (int)key_z(S.key[i]) : 0;
    const int zmin = wave_min_i(live ? z : 0xFFFF), zmax = wave_max_i(live ? z : 0);
    if (l == 0) S.bz[j] = (uint32_t)zmin | ((uint32_t)zmax << 16);
  }
  if (l == 0) S.bsum[j] = tot;
}

template <typename RateT>
__device__ void super_sum(const Sec<RateT>& S, int k, int nb) {
  const int l = lane_id();
  const int j = 64 * k + l;
  const double x = j < nb ? S.bsum[j] : 0.0;
  const double tot = bcast_d(wave_incl_scan(x), 63);
  if (l == 0) S.ssum[k] = tot;
}

// One scan level: the first lane (< n) with a positive entry whose inclusive
// sum exceeds `tl`, or, when rounding leaves none, the last lane with a
// positive entry.  Only positive entries are ever picked (a zero entry is an
// empty block or a vacancy that left the sector).  Returns the lane (-1 if
// every entry is zero) and adds the exclusive sum before it to `base`.
__device__ __forceinline__ int scan_pick(double x, int n, double tl, double& base) {
  const int l = lane_id();
  const double inc = wave_incl_scan(x);
  int k = first_set(ballot((l < n) & (x > 0.0) & (inc > tl)));
  if (k == 64) k = last_set(ballot((l < n) & (x > 0.0)));
  if (k >= 0) base += k > 0 ? bcast_d(inc, k - 1) : 0.0;
  return k;
}

// Collects the vacancies of sector s and sorts them by key (z, y, x), so
// that each 64-vacancy block covers a few z planes: after a hop only the
// blocks whose z range reaches the hop are scanned for stencil hits.
template <typename RateT>
__device__ int collect(const Sec<RateT>& S, const RView& R, int s) {
  const int l = lane_id();
  int n = 0;
  for (int b = 0; b < R.nvac; b += 64) {
    const int j = b + l;
    const uint32_t k = R.vac_key[min(j, R.nvac - 1)];
    const bool in = (j < R.nvac) & (key_sector(R.d, k) == s);
    const uint64_t m = ballot(in);
    const int r = n + __popcll(m & ((1ull << l) - 1));
    if (in) { S.key[r] = k; S.slot[r] = j; }
    n += __popcll(m);
  }
  wave_sync();
  sort_by_key(S.key, S.slot, n);
  return n;
}

template <typename RateT, bool FP32>
__device__ int rebuild(const Sec<RateT>& S, RStage& G, const RTabs& T, const RView& R, int s, int flags, ErrAcc& E,
                       bool& bad) {
  const int nv = collect(S, R, s);
  const Patch nopatch{0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, -1};
  eval_rep<RateT, FP32>(S, G, T, R, nullptr, 0, nv, nopatch, false, 0, -1, flags, E, bad);
  const int nb = (nv + 63) >> 6;
  for (int j = 0; j < nb; j++) block_sum(S, j, nv, true);
  wave_sync();
  const int nsup = (nb + 63) >> 6;
  if (nsup > 1)
    for (int k = 0; k < nsup; k++) super_sum(S, k, nb);
  wave_sync();
  return nv;
}

// Copies the sector arrays of nv vacancies between two layouts.
__device__ void copy_words(uint32_t* dst, const uint32_t* src, int64_t words) {
  for (int64_t i = lane_id(); i < words; i += 64) dst[i] = src[i];
}
template <typename RateT>
__device__ void copy_sector(const Sec<RateT>& D, const Sec<RateT>& S, int nv) {
  const int nb = (nv + 63) >> 6, nsup = (nb + 63) >> 6;
  copy_words((uint32_t*)D.rate, (const uint32_t*)S.rate, (int64_t)nv * 8 * sizeof(RateT) / 4);
  copy_words((uint32_t*)D.vt, (const uint32_t*)S.vt, 2 * nv);
  copy_words((uint32_t*)D.bsum, (const uint32_t*)S.bsum, 2 * nb);
  copy_words((uint32_t*)D.ssum, (const uint32_t*)S.ssum, 2 * nsup);
  copy_words(D.key, S.key, nv);
  copy_words((uint32_t*)D.slot, (const uint32_t*)S.slot, nv);
  copy_words(D.bz, S.bz, nb);
  copy_words((uint32_t*)D.own, (const uint32_t*)S.own, (nv + 3) / 4);
  copy_words((uint32_t*)D.nb1, (const uint32_t*)S.nb1, 2 * nv);
  wave_sync();
}

template <typename RateT, bool FP32, bool LDS>
__device__ void run_item(const BatchArgs& a, const RTabs& T, RStage& G, uint8_t* lds_sec, int r) {
  const int l = lane_id();
  const ReplicaDesc D = a.desc[r];
  ReplicaState st = a.state[r];
  const bool rates_only = (a.flags & AKMC_BF_RATES_ONLY) != 0;
  if (!rates_only && st.status != AKMC_REP_RUNNING) return;
  const uint64_t t_item = (uint64_t)wall_clock64();
  const RView R{D.sites, D.vac_key, D.nvac, D.d, D.nu, D.kT, D.nhi, D.m_min, D.nbx, D.nby};
  if (FP32)
    for (int i = l; i < 256 + D.nhi; i += 64) G.tab[i] = D.exp_tab[i];
  if (l < 8) {
    G.evo[l] = a.state[r].evo[l];
    G.cnt[l] = a.state[r].sec_count[l];
  }
  wave_sync();
  const Sec<RateT> Sg = sec_at<RateT>(D.ws, D.cap);
  const Sec<RateT> S = LDS ? sec_at<RateT>(lds_sec, a.vcap) : Sg;
  ErrAcc E{0.0, 0.0, 0, 0};
  bool bad = false;

  if (rates_only) {
    const int nv = rebuild<RateT, FP32>(S, G, T, R, a.rates_sector, a.flags, E, bad);
    if (LDS) copy_sector(Sg, S, nv);
    if (l == 0) a.state[r].nv = bad ? -1 : nv;
    return;
  }
  if (LDS && st.in_sector && st.nv >= 0) copy_sector(S, Sg, st.nv);
  const uint64_t t_start = a.queue[1];
  // Work budget of this item: slice_events events, and at most 64 times as
  // many sector visits, so that a launch ends even when a replica performs
  // no events.
  int64_t budget = a.slice_events;
  int64_t visits = 64 * a.slice_events;
  const double inv_T = 1.0 / D.sector_T;
  const LaneIdx LI = lane_indices(T.st);
  int zero_run = 0;  // consecutive sector visits with a zero total and no event
  int32_t status = AKMC_REP_RUNNING;
  const Patch nopatch{0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, -1};
  (void)nopatch;
  for (;;) {
    if (!st.in_sector) {
      if (st.step >= D.time_steps) { status = AKMC_REP_TIME_DONE; break; }
      if (st.sect == 0) {
        // Time steps in which every sector holding a vacancy is already past
        // its threshold have no event: those sectors end at once and the
        // empty ones only advance their clocks to (step + 1) T.  Skip them:
        // jump to the first step at which an occupied sector can run.  The
        // jump depends on the state alone, so it is taken wherever a run
        // reaches this ring start.
        int64_t K = INT64_MAX;
        for (int q = 0; q < 8; q++) {
          if (G.cnt[q] == 0) continue;
          const double e = G.evo[q];
          int64_t j = max(st.step, (int64_t)ceil(e * inv_T) - 1);
          while (!(static_cast<double>(j + 1) * D.sector_T - e > 0.0)) j++;
          while (j > st.step && static_cast<double>(j) * D.sector_T - e > 0.0) j--;
          K = min(K, j);
        }
        if (K != INT64_MAX && K > st.step) {
          K = min(K, D.time_steps);
          const double tk = static_cast<double>(K) * D.sector_T;
          if (l < 8 && G.cnt[l] == 0) G.evo[l] = fmax(G.evo[l], tk);
          wave_sync();
          st.n_skip += K - st.step;
          st.step = K;
          if (st.step >= D.time_steps) { status = AKMC_REP_TIME_DONE; break; }
        }
      }
      const int s0 = ring_sector(st.sect);
      // step_threshold_time (sublattice.inl:72)
      st.thr = static_cast<double>(st.step + 1) * D.sector_T - G.evo[s0];
      st.t = 0.0;
      st.in_sector = 1;
      st.nv = -1;  // the sector's rates are evaluated at its first event
    }
    const int s = ring_sector(st.sect);
    if (!(st.t < st.thr)) {
      if (l == 0) G.evo[s] += st.t;  // sublattice.inl:140
      wave_sync();
      // A sector that ended before its rates were evaluated (its threshold
      // was already passed) says nothing about its total: restart the count
      // of empty sectors.
      if (st.nv < 0) zero_run = 0;
      st.in_sector = 0;
      if (++st.sect == 8) { st.sect = 0; st.step++; }
      visits--;
      continue;
    }
    if (st.events >= st.target) { status = AKMC_REP_TARGET_DONE; break; }
    if (budget <= 0 || visits <= 0) break;
    if (st.nv < 0 && G.cnt[s] == 0) {
      // An empty sector: its total is zero without evaluating anything.
      st.nv = 0;
      st.t = st.thr;
      st.n_jump++;
      if (++zero_run >= 8) {
        if (l < 8) G.evo[l] = (double)D.time_steps * D.sector_T;
        wave_sync();
        st.step = D.time_steps;
        st.sect = 0;
        st.in_sector = 0;
        status = AKMC_REP_TIME_DONE;
        break;
      }
      continue;
    }
    if (st.nv < 0) {
      // calcRates of the sector's first event (kmc.cpp:68-115).
      st.nv = rebuild<RateT, FP32>(S, G, T, R, s, a.flags, E, bad);
      st.n_rebuild++;
      st.n_eval += st.nv;
      if (bad) { status = AKMC_REP_ERR_RANGE; break; }
    }
    if (a.budget_ticks && (st.events & 7) == 0 && (uint64_t)wall_clock64() - t_start > a.budget_ticks) break;

    // ---- total and selection (three scan levels).
    const int nv = st.nv;
    const int nb = (nv + 63) >> 6, nsup = (nb + 63) >> 6;
    double sum, target, base = 0.0;
    int j0 = 0, nbk = nb;
    double xt = 0.0, inc_top = 0.0;
    if (nb == 1) {
      sum = S.bsum[0];  // one block: its sum is the total
    } else {
      xt = nsup > 1 ? (l < nsup ? S.ssum[l] : 0.0) : (l < nb ? S.bsum[l] : 0.0);
      inc_top = wave_incl_scan(xt);
      sum = bcast_d(inc_top, 63);
    }
    // sum_rates == 0 or below epsilon: jump to the threshold (sublattice.inl:108-112).
    if (sum == 0.0 || fabs(sum) < 2.220446049250313e-16) {
      st.t = st.thr;
      st.n_jump++;
      if (++zero_run >= 8) {
        // Eight consecutive sectors (the whole ring) evaluated with a zero
        // total and no event between them: the lattice can no longer
        // change, so every later sector is empty too.  Jump to the end of
        // the time loop.
        if (l < 8) G.evo[l] = (double)D.time_steps * D.sector_T;
        wave_sync();
        st.step = D.time_steps;
        st.sect = 0;
        st.in_sector = 0;
        status = AKMC_REP_TIME_DONE;
        break;
      }
      continue;
    }
    const double u = (double)(xo_next(st.rng) >> 11) * 0x1p-53;  // [0, 1)
    target = u * sum;
    int k;
    if (nsup > 1) {
      k = first_set(ballot((l < nsup) & (xt > 0.0) & (inc_top > target)));
      if (k == 64) k = last_set(ballot((l < nsup) & (xt > 0.0)));
      if (k > 0) base += bcast_d(inc_top, k - 1);
      j0 = 64 * max(k, 0);
      nbk = min(64, nb - j0);
      k = scan_pick(l < nbk ? S.bsum[j0 + l] : 0.0, nbk, target - base, base);
    } else if (nb > 1) {
      k = first_set(ballot((l < nb) & (xt > 0.0) & (inc_top > target)));
      if (k == 64) k = last_set(ballot((l < nb) & (xt > 0.0)));
      if (k > 0) base += bcast_d(inc_top, k - 1);
    } else {
      k = 0;
    }
    const int blk = j0 + max(k, 0);
    const int i0 = 64 * blk;
    const int nvk = min(64, nv - i0);
    const int kv = scan_pick(l < nvk ? S.vt[i0 + l] : 0.0, nvk, target - base, base);
    if (kv < 0) { status = AKMC_REP_ERR_NO_EVENT; break; }
    const int v = i0 + kv;
    // Direction: an inclusive scan of the vacancy's 8 rates in lanes 0..7.
    const int bsel = scan_pick(l < 8 ? (double)S.rate[v * 8 + min(l, 7)] : 0.0, 8, target - base, base);
    if (bsel < 0) { status = AKMC_REP_ERR_NO_EVENT; break; }
    const double rr = (double)((xo_next(st.rng) >> 11) + 1) * 0x1p-53;  // (0, 1]

    // ---- perform (ABVIModel::perform, kmc.cpp:358-396).
    const uint32_t from = S.key[v];
    const int par = (int)(key_x(from) & 1);
    int ox, oy, oz;
    nn1_offset(par, bsel, &ox, &oy, &oz);
    const uint32_t to = key_add(R.d, from, ox, oy, oz);
    uint8_t pf = 0;
    uint32_t pf_key = 0xFFFFFFFFu;
    if (l < T.st.n_sites) {
      const int pp = par ^ 1;
      pf_key = key_add(R.d, to, T.st.off[pp][l][0], T.st.off[pp][l][1], T.st.off[pp][l][2]);
      pf = R.sites[site_at(R, pf_key)];
    }
    const uint8_t ff = S.own[v];
    const uint8_t gt = S.nb1[v * 8 + bsel];
    const uint8_t xb = (gt & BIT_CU) ? BIT_CU : (gt & BIT_MN) ? BIT_MN : (gt & BIT_NI) ? BIT_NI : 0;
    const uint8_t nf_from = (uint8_t)((ff & ~BIT_V) | xb);
    const uint8_t nf_to = (uint8_t)((gt & ~xb) | BIT_V);
    const int32_t vslot = S.slot[v];
    const bool stays = key_sector(R.d, to) == s;
    if (l == 0) {
      R.sites[site_at(R, from)] = nf_from;
      R.sites[site_at(R, to)] = nf_to;
      R.vac_key[vslot] = to;
      if (stays) {
        S.key[v] = to;
        const uint32_t z = key_z(to), bzv = S.bz[blk];
        S.bz[blk] = min(bzv & 0xFFFFu, z) | (max(bzv >> 16, z) << 16);
      } else {
        S.slot[v] = -1;
        S.vt[v] = 0.0;
        S.own[v] = 0;
        G.cnt[s] -= 1;
        G.cnt[key_sector(R.d, to)] += 1;
      }
    }
    // ---- residence time (sublattice.inl:118-120).
    st.t += -log(rr) / sum;
    st.events++;
    budget--;
    zero_run = 0;
    if (!(st.t < st.thr)) continue;  // the sector ends before any further evaluation

    // ---- re-evaluate the vacancies whose stencil holds `from` or `to`.
    wave_sync();
    const int zf = (int)key_z(from);
    const int zlo = min(zf, zf + oz) - 2, zhi = max(zf, zf + oz) + 2;
    const Patch pt{from, to, nf_from, nf_to, stays ? v : -1};
    uint64_t dsup = 0;
    for (int c0 = 0; c0 < nb; c0 += 64) {
      const int j = c0 + l;
      const uint32_t bzj = S.bz[min(j, nb - 1)];
      const bool hb = (j < nb) & ((((int)(bzj >> 16) >= zlo) & ((int)(bzj & 0xFFFFu) <= zhi)) | (j == blk));
      uint64_t mb = ballot(hb);
      while (mb) {
        const int jb = c0 + first_set(mb);
        mb &= mb - 1;
        const int i = 64 * jb + l;
        const int ic = min(i, nv - 1);
        const uint32_t w = S.key[ic];
        const bool live = (i < nv) & (S.slot[ic] >= 0);
        const bool hit = (i == v) ? stays : (live & (in_stencil(R.d, T.st, w, from) | in_stencil(R.d, T.st, w, to)));
        const uint64_t m = ballot(hit);
        if (FP32 && __popcll(m) == 1) {
          // One vacancy: the all-lane evaluation, from the prefetched stencil
          // when it is the one that hopped.
          const int ih = 64 * jb + first_set(m);
          uint32_t f = pf, fk = pf_key;
          if (ih != v) {
            const uint32_t wh = S.key[ih];
            const int ph = (int)(key_x(wh) & 1);
            const int lk = min(l, T.st.n_sites - 1);
            fk = key_add(R.d, wh, T.st.off[ph][lk][0], T.st.off[ph][lk][1], T.st.off[ph][lk][2]);
            f = R.sites[site_at(R, fk)];
          }
          if (fk == from) f = nf_from;
          if (fk == to) f = nf_to;
          eval_one<RateT>(S, G, T, R, LI, ih, f, pt, true, a.flags, E, bad);
          st.n_eval += 1;
        } else if (m) {
          const int rk = __popcll(m & ((1ull << l) - 1));
          if (hit) G.list[rk] = i;
          wave_sync();
          eval_rep<RateT, FP32>(S, G, T, R, G.list, 0, __popcll(m), pt, true, pf, stays ? v : -1, a.flags, E, bad);
          st.n_eval += __popcll(m);
        }
        if (m || jb == blk) {
          block_sum(S, jb, nv, false);
          dsup |= 1ull << ((jb >> 6) & 63);
        }
      }
    }
    wave_sync();
    if (nsup > 1) {
      while (dsup) {
        const int ks = first_set(dsup);
        dsup &= dsup - 1;
        super_sum(S, ks, nb);
      }
      wave_sync();
    }
    if (bad) { status = AKMC_REP_ERR_RANGE; break; }
  }
  if (LDS && st.in_sector && st.nv >= 0) copy_sector(Sg, S, st.nv);
  // Error statistics of the FP32 rates, reduced over the wave.
  if (FP32 && (a.flags & AKMC_BF_ERROR_STATS)) {
    const double emax = wave_max_d(E.emax);
    const double esum = wave_sum_d(E.esum);
    const double en = wave_sum_d((double)E.n), et = wave_sum_d((double)E.tiny);
    st.err_max = fmax(st.err_max, emax);
    st.err_sum += esum;
    st.err_n += (int64_t)en;
    st.err_tiny += (int64_t)et;
  }
  st.status = status;
  st.busy_ticks += (int64_t)((uint64_t)wall_clock64() - t_item);
  for (int i = 0; i < 8; i++) {
    st.evo[i] = G.evo[i];
    st.sec_count[i] = G.cnt[i];
  }
  wave_sync();
  if (l == 0) a.state[r] = st;
}

// IL (interleave width) is the number of replica waves that share one SIMD:
// the kernel is compiled for at most 512 / IL registers per lane, so that IL
// waves fit on a SIMD, and launched in workgroups of 4 * IL waves (one
// workgroup per CU, IL waves on each of its four SIMDs; IL = 1 also takes
// workgroups of 1 or 2 waves).  While one wave waits on a gather or a
// dependent chain, the SIMD issues another replica's instructions.
template <typename RateT, bool FP32, bool LDS, int IL>
__global__ __launch_bounds__(256 * IL, IL) void replica_kernel(BatchArgs a) {
  __shared__ RTabs T;
  extern __shared__ __align__(16) uint8_t dyn[];
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&r_sten);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&T.st);
    for (int i = threadIdx.x; i < (int)(sizeof(StencilTables) / 4); i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < 25; i += blockDim.x) {
      T.e1[i] = r_e1[i]; T.e2[i] = r_e2[i]; T.e1q[i] = a.e1q[i]; T.e2q[i] = a.e2q[i];
    }
    for (int i = threadIdx.x; i < 5; i += blockDim.x) { T.e0[i] = r_e0[i]; T.e0q2[i] = a.e0q2[i]; }
    for (int i = threadIdx.x; i < 256; i += blockDim.x) T.exp_tab[i] = r_exp_tab[i];
  }
  __syncthreads();
  // Work items are dealt to the waves of the grid round-robin.  The item
  // index is computed from uniform values only (never from a value that one
  // lane produced), so every lane of a wave runs the same item sequence.
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int waves = (int)(blockDim.x >> 6);
  uint8_t* region = dyn + (size_t)wave * a.region_bytes;
  RStage& G = *reinterpret_cast<RStage*>(region);
  uint8_t* lds_sec = region + RSTAGE_BYTES;
  const int stride = (int)gridDim.x * waves;
  for (int item = (int)blockIdx.x * waves + wave; item < a.n_items; item += stride)
    run_item<RateT, FP32, LDS>(a, T, G, lds_sec, __builtin_amdgcn_readfirstlane(a.items[item]));
}

// Records the launch start tick (queue[1]) for the time budget.
__global__ void reset_kernel(uint64_t* q) {
  if (threadIdx.x == 0) {
    q[0] = 0;
    q[1] = (uint64_t)wall_clock64();
  }
}

// Initial occupancy of a replica from a counter-based generator: site i takes
// the type whose cumulative ratio bracket holds splitmix64(seed + i) mod total.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// The loop runs over the brick layout (every byte, padding included, is
// written once); the draw of a site uses its logical index.
__global__ void init_sites_kernel(uint8_t* sites, BoxDims d, uint64_t seed, const uint32_t* cum, const uint8_t* flag,
                                  int32_t nt) {
  const uint32_t total = cum[nt - 1];
  const BrickDims b = brick_dims(d);
  const int64_t nb = brick_bytes(b);
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < nb; a += (int64_t)gridDim.x * blockDim.x) {
    const int64_t brick = a >> 7;
    const int w = (int)(a & 127);
    const int64_t bxy = (int64_t)b.nbx * b.nby;
    const int x = (int)(brick % b.nbx) * 8 + (w & 7);
    const int y = (int)((brick / b.nbx) % b.nby) * 4 + ((w >> 3) & 3);
    const int z = (int)(brick / bxy) * 4 + (w >> 5);
    uint8_t f = 0;
    if (x < d.X && y < d.Y && z < d.Z) {
      const int64_t i = (int64_t)x + (int64_t)d.X * ((int64_t)y + (int64_t)d.Y * z);
      const uint64_t h = splitmix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(i + 1)));
      const uint32_t r = (uint32_t)(((h >> 32) * (uint64_t)total) >> 32);  // [0, total)
      for (int t = nt - 1; t >= 0; t--)
        if (r < cum[t]) f = flag[t];
    }
    sites[a] = f;
  }
}
// Vacancy placement as the reference's createRandom does it: the site gains
// the V tag and loses a Cu tag; Mn/Ni tags stay (creation.cpp:131-164).
// idx holds logical site indices.
__global__ void place_vacancies_kernel(uint8_t* sites, BoxDims d, const int64_t* idx, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int64_t li = idx[i];
    const BrickDims b = brick_dims(d);
    const uint32_t x = (uint32_t)(li % d.X), y = (uint32_t)((li / d.X) % d.Y), z = (uint32_t)(li / ((int64_t)d.X * d.Y));
    const int64_t s = brick_index(b.nbx, b.nby, x, y, z);
    sites[s] = (uint8_t)((sites[s] & ~BIT_CU) | BIT_V);
  }
}

template <typename RateT, bool FP32, bool LDS>
static int launch_t(const BatchArgs* a, int il, int waves, int grid, hipStream_t stream) {
  const size_t lds = (size_t)waves * a->region_bytes;
  if (il == 4)
    hipLaunchKernelGGL((replica_kernel<RateT, FP32, LDS, 4>), dim3(grid), dim3(64 * waves), lds, stream, *a);
  else if (il == 2)
    hipLaunchKernelGGL((replica_kernel<RateT, FP32, LDS, 2>), dim3(grid), dim3(64 * waves), lds, stream, *a);
  else
    hipLaunchKernelGGL((replica_kernel<RateT, FP32, LDS, 1>), dim3(grid), dim3(64 * waves), lds, stream, *a);
  return (int)hipGetLastError();
}

template <typename RateT, bool FP32, bool LDS>
static int occupancy_t(int il, int waves, int region_bytes, int* blocks) {
  const size_t lds = (size_t)waves * region_bytes;
  if (il == 4)
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, replica_kernel<RateT, FP32, LDS, 4>, 64 * waves, lds);
  if (il == 2)
    return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, replica_kernel<RateT, FP32, LDS, 2>, 64 * waves, lds);
  return (int)hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, replica_kernel<RateT, FP32, LDS, 1>, 64 * waves, lds);
}

}  // namespace akmc

using akmc::BatchArgs;

extern "C" int akmc_batch_launch(const BatchArgs* a, int fp32, int lds, int il, int waves, int grid,
                                 hipStream_t stream) {
  if (il != 1 && il != 2 && il != 4) return (int)hipErrorInvalidValue;
  if (fp32) {
    return lds ? akmc::launch_t<float, true, true>(a, il, waves, grid, stream)
               : akmc::launch_t<float, true, false>(a, il, waves, grid, stream);
  }
  return lds ? akmc::launch_t<double, false, true>(a, il, waves, grid, stream)
             : akmc::launch_t<double, false, false>(a, il, waves, grid, stream);
}

extern "C" int akmc_batch_occupancy(int fp32, int lds, int il, int waves, int region_bytes, int* blocks) {
  if (il != 1 && il != 2 && il != 4) return (int)hipErrorInvalidValue;
  if (fp32) {
    return lds ? akmc::occupancy_t<float, true, true>(il, waves, region_bytes, blocks)
               : akmc::occupancy_t<float, true, false>(il, waves, region_bytes, blocks);
  }
  return lds ? akmc::occupancy_t<double, false, true>(il, waves, region_bytes, blocks)
             : akmc::occupancy_t<double, false, false>(il, waves, region_bytes, blocks);
}

extern "C" int akmc_batch_reset_launch(uint64_t* queue, hipStream_t stream) {
  hipLaunchKernelGGL(akmc::reset_kernel, dim3(1), dim3(64), 0, stream, queue);
  return (int)hipGetLastError();
}

extern "C" int akmc_batch_set_stencil(const akmc::StencilTables* t) {
  return (int)hipMemcpyToSymbol(HIP_SYMBOL(akmc::r_sten), t, sizeof(*t));
}

extern "C" int akmc_batch_region_bytes(int fp32, int vcap) {
  const int64_t sec = vcap > 0 ? akmc::ws_layout(vcap, fp32 ? 4 : 8).total : 0;
  return (int)(akmc::RSTAGE_BYTES + sec);
}

extern "C" int akmc_batch_init_sites_launch(uint8_t* sites, akmc::BoxDims d, uint64_t seed, const uint32_t* cum,
                                            const uint8_t* flag, int32_t nt, hipStream_t stream) {
  const int64_t n = akmc::brick_bytes(akmc::brick_dims(d));
  if (n <= 0) return 0;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(akmc::init_sites_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, sites, d, seed, cum, flag,
                     nt);
  return (int)hipGetLastError();
}

extern "C" int akmc_batch_place_vacancies_launch(uint8_t* sites, akmc::BoxDims d, const int64_t* idx, int64_t n,
                                                 hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(akmc::place_vacancies_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, sites, d,
                     idx, n);
  return (int)hipGetLastError();
}
