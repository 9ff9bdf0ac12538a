// The abvi model's event loop with vacancies and self-interstitial dumbbells
// on a typed lattice, one wavefront per lattice (a launch carries any number
// of lattices, one workgroup each).
//
// The reference has the model's pieces but no wired loop (the dumbbell branch
// of ABVIModel::select is commented out, models/abvi/kmc.cpp:251-268; perform
// asserts on anything but a vacancy hop, :393-395).  This loop wires them as
// the oracle (oracle/refcpu.cpp, refmix) states, bit for bit in FP64:
//   * per sector (ring 0,7,2,5,3,4,1,6, SubLattice::startTimeLoop,
//     src/algorithms/sl/sublattice.inl:61-143) the sector's vacancies and
//     dumbbells are kept as records sorted in sweep order, each with 8 rate
//     slots: a vacancy's 8 direction rates (VacRatesSolver on the typed
//     lattice), a dumbbell's 8 slots (Itl::updateRates with ItlRatesSolver,
//     models/abvi/defect/itl_list.cpp:31-50, rate/itl_rates_solver.cpp:31-80);
//     the total is the left fold of every slot in that order plus
//     defect_gen_rate (ABVIModel::defectGenRate, kmc.cpp:207);
//   * select: the first slot whose running sum exceeds u * total
//     (kmc.cpp:242-309 with the dumbbell branch :251-268); past the last slot,
//     defect generation (the default event, :245);
//   * perform: a vacancy hop (kmc.cpp:358-396 on the typed lattice), a
//     dumbbell transition (orientation::tranAtom and orientation::trans,
//     orientation.cc:238-252), or a generated Frenkel pair; then
//     recombination around the moved defect (rec::RecList, recombine.cpp).
// Exact running sums use the same binade-integer folds as the vacancy engine
// (akmc_wave.h).  After an event only the records within reach of a changed
// site are re-evaluated, and every dumbbell of the sector when a term of the
// whole-list dumbbell energy Edumb changed (each dumbbell rate folds the whole
// list, bond::Edumb, energy_dumbbell.cpp:7-137).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_mix.h"
#include "akmc_plain.h"
#include "akmc_wave.h"

namespace akmc {
namespace mix {

using itl::c_e1;
using itl::c_e2;
using itl::c_exp_tab;
using itl::c_high;
using itl::c_low;
using itl::Dims;
using itl::edumb_term;
using itl::is_atom;
using itl::is_dumbbell;
using itl::T_CU;
using itl::T_FE;
using itl::T_MN;
using itl::T_NI;
using itl::T_V;
using itl::View;

// orientation::TransHashTable (orientation.cc:23-140): [|o| - 1][1nn tag]
// [!new_higher][!rotate]; 0 where the orientation cannot jump.
static __constant__ int8_t c_trans[6][8][2][2] = {
    {{{-3, -5}, {3, 5}}, {{-4, -6}, {4, 6}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}},
     {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{4, 6}, {-4, -6}}, {{3, 5}, {-3, -5}}},
    {{{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{4, -5}, {-4, 5}}, {{3, -6}, {-3, 6}},
     {{-3, 6}, {3, -6}}, {{-4, 5}, {4, -5}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}},
    {{{-1, -5}, {1, 5}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{-6, -2}, {6, 2}},
     {{6, 2}, {-6, -2}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{1, 5}, {-1, -5}}},
    {{{0, 0}, {0, 0}}, {{-1, -6}, {1, 6}}, {{-2, -5}, {2, 5}}, {{0, 0}, {0, 0}},
     {{0, 0}, {0, 0}}, {{2, 5}, {-2, -5}}, {{1, 6}, {-1, -6}}, {{0, 0}, {0, 0}}},
    {{{-3, -1}, {3, 1}}, {{0, 0}, {0, 0}}, {{-4, 2}, {4, -2}}, {{0, 0}, {0, 0}},
     {{0, 0}, {0, 0}}, {{4, -2}, {-4, 2}}, {{0, 0}, {0, 0}}, {{3, 1}, {-3, -1}}},
    {{{0, 0}, {0, 0}}, {{-4, -1}, {4, 1}}, {{0, 0}, {0, 0}}, {{-3, 2}, {3, -2}},
     {{3, -2}, {-3, 2}}, {{0, 0}, {0, 0}}, {{4, 1}, {-4, -1}}, {{0, 0}, {0, 0}}},
};

// Third neighbours (doubled x, <110> type): the partner shell of a generated
// Frenkel pair, the nearest outside the 1nn/2nn recombination range.
static __constant__ int8_t c_third[12][3] = {{-2, -1, 0}, {-2, 1, 0}, {-2, 0, -1}, {-2, 0, 1}, {2, -1, 0}, {2, 1, 0},
                                             {2, 0, -1},  {2, 0, 1},  {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}};

__device__ __forceinline__ int ring_sector(int pos) { return (int)((0x61435270u >> (4 * pos)) & 7u); }

__device__ __forceinline__ int hi_end(int t) { return t >> 8; }
__device__ __forceinline__ int lo_end(int t) { return t & 0xFF; }
// LatticeTypes::combineToInter and diff (src/lattice/lattice_types.cpp:21-48).
__device__ __forceinline__ int combine(int a, int b) { return a <= b ? (a << 8) | b : (b << 8) | a; }
__device__ __forceinline__ int diff_type(int t, int g) {
  return lo_end(t) == g ? hi_end(t) : hi_end(t) == g ? lo_end(t) : t;
}
// VacRatesSolver::e0 (vacancy_rates_solver.cpp:14-27) and ItlRatesSolver::e0
// (itl_rates_solver.cpp:16-29) by LatticeTypes value.
__device__ __forceinline__ double e0_vac(int X) {
  return X == T_FE ? 0.62 : X == T_CU ? 0.54 : X == T_NI ? 0.68 : X == T_MN ? 1.03 : 0.0;
}
__device__ __forceinline__ double e0_itl(int g) {
  return g == T_FE ? 0.32 : g == T_CU ? 0.36 : g == T_NI ? 0.45 : g == T_MN ? 0.31 : 0.0;
}
__device__ __forceinline__ uint8_t dirs_of(int o) { return (uint8_t)(c_high[o + 6] | c_low[o + 6]); }
// Itl::get1nnIdByRatesIndex (itl_list.cpp:76-91).
__device__ __forceinline__ int first_nn_by_slot(int slot, uint8_t dirs) {
  const int want = slot >> 1;
  int i = 0;
  for (int b = 0; b < N1; b++)
    if ((dirs >> b) & 1) {
      if (i == want) return b;
      i++;
    }
  return 0;
}
// orientation::random (orientation.h:38-42) from one 32-bit draw.
__device__ __forceinline__ int orientation_from_draw(uint32_t r) {
  const uint32_t v = r % 12u + 1u;
  return v % 2 == 0 ? (int)(v / 2) : -(int)((v + 1) / 2);
}

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
__device__ __forceinline__ uint64_t xo_next(uint64_t* s) {  // xoshiro256**
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

__device__ __forceinline__ void key_xyz(uint32_t k, int& x, int& y, int& z) {
  x = (int)key_x(k);
  y = (int)key_y(k);
  z = (int)key_z(k);
}
// Periodic box test: q within |dx| <= rx (doubled x), |dy| <= ry, |dz| <= rz of w.
__device__ __forceinline__ bool near_box(const BoxDims& d, uint32_t w, uint32_t q, int rx, int ry, int rz) {
  const int dx = pdiff((int)key_x(q), (int)key_x(w), d.X);
  const int dy = pdiff((int)key_y(q), (int)key_y(w), d.Y);
  const int dz = pdiff((int)key_z(q), (int)key_z(w), d.Z);
  return (dx >= -rx) & (dx <= rx) & (dy >= -ry) & (dy <= ry) & (dz >= -rz) & (dz <= rz);
}
// A record's rates read sites within 1nn + 2nn of itself and of its 1nn
// targets: |dx| <= 3, |dy| <= 2, |dz| <= 2.
__device__ __forceinline__ bool in_reach(const BoxDims& d, uint32_t w, uint32_t q) { return near_box(d, w, q, 3, 2, 2); }
// A dumbbell's Edumb term and availability mask read its 1nn and 2nn sites.
__device__ __forceinline__ bool term_reads(const BoxDims& d, uint32_t w, uint32_t q) { return near_box(d, w, q, 2, 1, 1); }

// Sector records (device memory) in the form the exact folds take.
struct MixSector {
  double* rate;
  double* pvac;
  uint64_t* tmask;
  uint8_t* risk;
};

// First index i in [0, n) with key[i] >= k (keys sorted), one lane's search.
__device__ __forceinline__ int lb_lane(const uint32_t* key, int n, uint32_t k) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// LDS working space of the dumbbell evaluation passes (16 records a pass).
constexpr int kItlPass = 16;
constexpr int kMaxCand = 48;  // dumbbells within reach of one record
constexpr int kMaxAff = 16;   // terms one jump changes
struct ItlWork {
  uint32_t wkey[kItlPass];
  int32_t own[kItlPass];
  int32_t ncand[kItlPass];
  int32_t cand[kItlPass][kMaxCand];
  uint32_t ckey[kItlPass][kMaxCand];
  int32_t aq[kMaxAff][64];  // per lane: list index of each changed term, ascending
  double at[kMaxAff][64];   // and the term on the lattice after the jump
  int32_t vlist[64], ilist[64];
  int32_t ibuf[128];        // dumbbell records gathered for full passes
  double tbuf[64];          // one chunk of Edumb terms, read by every lane
};

struct Ctx {
  const MixDesc* D;
  Dims td;
  BoxDims d;
  ItlWork* W;
};

__device__ __forceinline__ int64_t site_of_key(const Ctx& c, uint32_t k) { return (int64_t)key_index(c.d, k); }

// Availability mask (Itl::availTranDirs, itl_list.cpp:9-23) and Edumb term of
// dumbbell j on the lattice as it is.
__device__ void refresh_term(const Ctx& c, int j) {
  const MixDesc& D = *c.D;
  const uint32_t k = D.ikey[j];
  int x, y, z;
  key_xyz(k, x, y, z);
  const int o = D.io[j];
  const View L{D.types, -1, -1, 0, 0};
  uint8_t atoms = 0;
  for (int b = 0; b < N1; b++)
    if (is_atom(L.at(itl::nn1(c.td, x, y, z, b)))) atoms |= (uint8_t)(1u << b);
  const uint8_t av = (uint8_t)(dirs_of(o) & atoms);
  D.iav[j] = av;
  D.iterm[j] = edumb_term(c.td, L, x, y, z, L.at(site_of_key(c, k)), o, av);
}

// Left fold of the terms from index j0 (prefix[j] = fold of term[0 .. j]),
// 64 terms per load, folded in order through lane reads.
__device__ void prefix_from(const Ctx& c, int j0, int n) {
  const MixDesc& D = *c.D;
  ItlWork& W = *c.W;
  const int l = lane_id();
  j0 = uni_i(j0);
  double acc = uni_d(j0 > 0 ? D.ipre[j0 - 1] : 0.0);
  double tn = j0 + l < n ? D.iterm[j0 + l] : 0.0;  // the next chunk loads while this one folds
  for (int base = j0; base < n; base += 64) {
    const int j = base + l;
    W.tbuf[l] = tn;
    __threadfence_block();
    wave_sync();
    if (base + 64 < n) tn = j + 64 < n ? D.iterm[j + 64] : 0.0;
    const int cnt = min(64, n - base);
    double mine = 0.0;
#pragma unroll 8
    for (int i = 0; i < cnt; i++) {
      acc += W.tbuf[i];
      mine = l == i ? acc : mine;
    }
    if (j < n) D.ipre[j] = mine;
    __threadfence_block();
    wave_sync();
  }
}

// Masks and terms of every dumbbell whose term reads one of the sites cs[0 ..
// nc); returns the first list index whose term may have changed (n if none).
__device__ int refresh_near(const Ctx& c, const uint32_t* cs, int nc, int n) {
  const MixDesc& D = *c.D;
  const int l = lane_id();
  int jmin = n;
  // Keys load four chunks ahead of the tests.
  uint32_t k0 = l < n ? D.ikey[l] : 0u, k1 = 64 + l < n ? D.ikey[64 + l] : 0u;
  uint32_t k2 = 128 + l < n ? D.ikey[128 + l] : 0u, k3 = 192 + l < n ? D.ikey[192 + l] : 0u;
  for (int base = 0; base < n; base += 64) {
    const int j = base + l;
    const uint32_t k = k0;
    k0 = k1;
    k1 = k2;
    k2 = k3;
    k3 = j + 256 < n ? D.ikey[j + 256] : 0u;
    bool hit = false;
    if (j < n) {
      for (int i = 0; i < nc; i++) hit = hit | term_reads(c.d, k, cs[i]);
    }
    if (hit) refresh_term(c, j);
    const uint64_t m = ballot(hit);
    if (m && jmin == n) jmin = base + first_set(m);
  }
  __threadfence_block();
  wave_sync();
  return uni_i(jmin);
}

// Rate of the vacancy at key w hopping in direction b (VacRatesSolver::deltaE,
// vacancy_rates_solver.cpp:29-172, and RatesSolver::rate, rates_solver.cpp:
// 13-26, on the typed lattice: dumbbell neighbours carry no bond,
// bonds_counter.cpp:35-81; the direction is available iff the target is a
// single atom, vac_hash.cpp:7-15).
__device__ double vac_rate(const Ctx& c, uint32_t w, int b) {
  const MixDesc& D = *c.D;
  int x, y, z;
  key_xyz(w, x, y, z);
  int ox, oy, oz;
  nn1_offset(x & 1, b, &ox, &oy, &oz);
  const int tx = wrapc(x + ox, c.d.X), ty = wrapc(y + oy, c.d.Y), tz = wrapc(z + oz, c.d.Z);
  const int64_t vs = itl::site_of(c.td, x, y, z), ts = itl::site_of(c.td, tx, ty, tz);
  const int X = D.types[ts];
  if (!is_atom(X)) return 0.0;
  const View L0{D.types, -1, -1, 0, 0};
  const double e_before = itl::count(c.td, L0, x, y, z, T_V) + itl::count(c.td, L0, tx, ty, tz, X);
  const View L1{D.types, ts, vs, T_V, X};
  const double e_after = itl::count(c.td, L1, x, y, z, X) + itl::count(c.td, L1, tx, ty, tz, T_V);
  const double ea = e0_vac(X) + (e_after - e_before) / 2;
  return D.nu * exp_glibc(-ea / D.kT, c_exp_tab);
}

// A dumbbell jump's rate from its energies: e_before = bonds + Edumb before,
// e_after = bonds after + Edumb after, summed in this order.
__device__ __forceinline__ double itl_finish(const Ctx& c, int ghost, double e_before, double bonds_after, double sa) {
  const double e_after = bonds_after + sa;
  const double ea = e0_itl(ghost) + (e_after - e_before) / 2;
  return c.D->nu * exp_glibc(-ea / c.D->kT, c_exp_tab);
}

// Rate of dumbbell j (at key w) jumping towards 1nn tag b (ItlRatesSolver::
// deltaE, itl_rates_solver.cpp:31-80): both sites' bonds and the whole list's
// Edumb before and after the jumping atom moves; Edumb after is the prefix
// fold up to the first term the jump can change, then the fold continues term
// by term with those terms recomputed on the changed lattice.
__device__ double itl_rate(const Ctx& c, uint32_t w, int j, int b, int n) {
  const MixDesc& D = *c.D;
  int x, y, z;
  key_xyz(w, x, y, z);
  const int o = D.io[j];
  const int64_t src = itl::site_of(c.td, x, y, z);
  const int type = D.types[src];
  const int ghost = ((c_high[o + 6] >> b) & 1) ? hi_end(type) : lo_end(type);
  int ox, oy, oz;
  nn1_offset(x & 1, b, &ox, &oy, &oz);
  const int tx = wrapc(x + ox, c.d.X), ty = wrapc(y + oy, c.d.Y), tz = wrapc(z + oz, c.d.Z);
  const int64_t dst = itl::site_of(c.td, tx, ty, tz);
  const int tdst = D.types[dst];
  const int src_after = diff_type(type, ghost);
  const int dst_after = combine(tdst, ghost);
  const View L0{D.types, -1, -1, 0, 0};
  const View L1{D.types, src, dst, src_after, dst_after};
  const double e_src_b = itl::count(c.td, L0, x, y, z, ghost);
  const double e_des_b = itl::count(c.td, L0, tx, ty, tz, tdst);
  const double e_src_a = itl::count(c.td, L1, x, y, z, src_after);
  const double e_des_a = itl::count(c.td, L1, tx, ty, tz, ghost);
  const double sb = n > 0 ? D.ipre[n - 1] : 0.0;
  const uint32_t wd = make_key((uint32_t)tx, (uint32_t)ty, (uint32_t)tz);
  int q = 0;
  while (q < n && !term_reads(c.d, D.ikey[q], w) && !term_reads(c.d, D.ikey[q], wd)) q++;
  double sa = q > 0 ? D.ipre[q - 1] : 0.0;
  for (; q < n; q++) {
    const uint32_t kq = D.ikey[q];
    if (term_reads(c.d, kq, w) || term_reads(c.d, kq, wd)) {
      int qx, qy, qz;
      key_xyz(kq, qx, qy, qz);
      sa += edumb_term(c.td, L1, qx, qy, qz, L1.at(site_of_key(c, kq)), D.io[q], D.iav[q]);
    } else {
      sa += D.iterm[q];
    }
  }
  return itl_finish(c, ghost, e_src_b + e_des_b + sb, e_src_a + e_des_a, sa);
}

// The jump of the dumbbell at key w (list entry j, orientation o) towards 1nn
// tag b: the sites, types and bond energies before and after.
struct ItlJump {
  uint32_t w, wd;
  int64_t src, dst;
  int ghost, src_after, dst_after;
  double e_before_bonds, e_after_bonds;
};
__device__ ItlJump itl_jump(const Ctx& c, uint32_t w, int o, int b) {
  const MixDesc& D = *c.D;
  ItlJump J;
  int x, y, z;
  key_xyz(w, x, y, z);
  J.w = w;
  J.src = itl::site_of(c.td, x, y, z);
  const int type = D.types[J.src];
  J.ghost = ((c_high[o + 6] >> b) & 1) ? hi_end(type) : lo_end(type);
  int ox, oy, oz;
  nn1_offset(x & 1, b, &ox, &oy, &oz);
  const int tx = wrapc(x + ox, c.d.X), ty = wrapc(y + oy, c.d.Y), tz = wrapc(z + oz, c.d.Z);
  J.dst = itl::site_of(c.td, tx, ty, tz);
  J.wd = make_key((uint32_t)tx, (uint32_t)ty, (uint32_t)tz);
  const int tdst = D.types[J.dst];
  J.src_after = diff_type(type, J.ghost);
  J.dst_after = combine(tdst, J.ghost);
  const View L0{D.types, -1, -1, 0, 0};
  const View L1{D.types, J.src, J.dst, J.src_after, J.dst_after};
  J.e_before_bonds = itl::count(c.td, L0, x, y, z, J.ghost) + itl::count(c.td, L0, tx, ty, tz, tdst);
  J.e_after_bonds = itl::count(c.td, L1, x, y, z, J.src_after) + itl::count(c.td, L1, tx, ty, tz, J.ghost);
  return J;
}

__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return uni_i(v);
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return uni_i(v);
}

// Rates of the dumbbell records rr[0 .. m), m <= 16: lane (a, i) takes the
// i-th direction of record a's orientation, and slots 2i and 2i + 1 hold half
// its rate when that direction is available (Itl::ratesIndex, itl_list.cpp:
// 67-74).  Edumb after the jump is the same left fold as itl_rate's, taken by
// all lanes in step: one pass over the list from the lowest changed index,
// each lane adding the stored term or, at its own changed indices, the term
// recomputed on the lattice after its jump.  The changed terms (dumbbells
// whose term reads the jump's source or target) are found among the dumbbells
// within reach of the record, which one scan of the list collects for the 16
// records at once; a lane with more candidates or changed terms than the
// working space holds takes itl_rate's own serial fold.
__device__ void itl_pass(const Ctx& c, MixSector& S, const int32_t* rr, int m, int nitl) {
  const MixDesc& D = *c.D;
  ItlWork& W = *c.W;
  const int l = lane_id();
  if (l < kItlPass) {
    W.wkey[l] = l < m ? D.rkey[rr[l]] : 0u;
    W.ncand[l] = 0;
    W.own[l] = -1;
  }
  __threadfence_block();
  wave_sync();
  // The list is sorted by key (z major): only entries within two planes of
  // the records' z range can be within reach.
  const int zlo = uni_i((int)key_z(W.wkey[0]) - 2), zhi = uni_i((int)key_z(W.wkey[m - 1]) + 2);
  int q0 = 0, q1 = nitl;
  if (zlo >= 0 && zhi < c.d.Z) {
    q0 = uni_i(lower_bound_key(D.ikey, nitl, make_key(0u, 0u, (uint32_t)zlo)));
    q1 = zhi + 1 < c.d.Z ? uni_i(lower_bound_key(D.ikey, nitl, make_key(0u, 0u, (uint32_t)(zhi + 1)))) : nitl;
  }
  for (int base = q0; base < q1; base += 64) {
    const int q = base + l;
    const uint32_t kq = q < q1 ? D.ikey[q] : 0u;
    for (int a = 0; a < m; a++) {
      const uint32_t wr = W.wkey[a];
      const bool hit = q < q1 && in_reach(c.d, wr, kq);
      const uint64_t hm = ballot(hit);
      if (!hm) continue;
      const int n0 = W.ncand[a];
      if (hit) {
        const int pos = n0 + __popcll(hm & ((1ull << l) - 1));
        if (pos < kMaxCand) {
          W.cand[a][pos] = q;
          W.ckey[a][pos] = kq;
        }
        if (kq == wr) W.own[a] = q;
      }
      __threadfence_block();
      wave_sync();
      if (l == 0) W.ncand[a] = n0 + __popcll(hm);
      __threadfence_block();
      wave_sync();
    }
  }
  const int a = l >> 2, i = l & 3;
  const bool rec = a < m;
  const int r = rec ? rr[a] : 0;
  const int j = rec ? W.own[a] : -1;
  const int nca = rec ? W.ncand[a] : 0;
  int o = 0;
  uint8_t dirs = 0, av = 0;
  if (j >= 0) {
    o = D.io[j];
    dirs = dirs_of(o);
    av = D.iav[j];
  }
  int dsel = 0;
  for (int b = 0, seen = 0; b < N1; b++)
    if ((dirs >> b) & 1) {
      if (seen == i) dsel = b;
      seen++;
    }
  const bool act = j >= 0 && ((av >> dsel) & 1);
  bool slow = act && nca > kMaxCand;
  ItlJump J{};
  if (act) J = itl_jump(c, W.wkey[a], o, dsel);
  const View L1{D.types, J.src, J.dst, J.src_after, J.dst_after};
  int na = 0, qf = nitl;
  const int cmax = wave_max_i(act && !slow ? nca : 0);
  for (int t = 0; t < cmax; t++) {
    if (act && !slow && t < nca) {
      const int q = W.cand[a][t];
      const uint32_t kq = W.ckey[a][t];
      if (term_reads(c.d, kq, J.w) || term_reads(c.d, kq, J.wd)) {
        if (na < kMaxAff) {
          int qx, qy, qz;
          key_xyz(kq, qx, qy, qz);
          W.aq[na][l] = q;
          W.at[na][l] = edumb_term(c.td, L1, qx, qy, qz, L1.at(site_of_key(c, kq)), D.io[q], D.iav[q]);
          if (na == 0) qf = q;
          na++;
        } else {
          slow = true;
        }
      }
    }
  }
  const bool fast = act && !slow && na > 0;
  if (!fast) qf = nitl;
  const int qa = fast ? W.aq[na - 1][l] : -1;  // the lane's last changed term
  double sa = fast && qf > 0 ? D.ipre[qf - 1] : 0.0;
  int k = 0, nq = fast ? W.aq[0][l] : nitl;
  // Up to the last changed term of any lane: each lane adds from its first
  // changed term on, its own changed terms in place of the stored ones.
  // After it: every lane adds the stored terms (lanes without a fast fold
  // add too; their sums are not used).  The next chunk loads while this one
  // folds.
  const int qmin = wave_min_i(qf), qend = wave_max_i(qa + 1);
  int base = qend > 0 ? qmin & ~63 : nitl;  // no fast lane: nothing to fold
  double tn = base + l < nitl ? D.iterm[base + l] : 0.0;
  for (; base < nitl; base += 64) {
    // The chunk's terms go to LDS, where every lane reads them (one
    // broadcast read per term); the next chunk loads meanwhile.
    W.tbuf[l] = tn;
    __threadfence_block();
    wave_sync();
    if (base + 64 < nitl) tn = base + 64 + l < nitl ? D.iterm[base + 64 + l] : 0.0;
    const int cnt = min(64, nitl - base);
    int ii = 0;
    for (; ii < cnt && base + ii < qend; ii++) {
      const int q = base + ii;
      const double tq = W.tbuf[ii];
      if (q >= qf) {
        double v = tq;
        if (q == nq) {
          v = W.at[k][l];
          k++;
          nq = k < na ? W.aq[k][l] : nitl;
        }
        sa += v;
      }
    }
#pragma unroll 8
    for (; ii < cnt; ii++) sa += W.tbuf[ii];
    __threadfence_block();
    wave_sync();
  }
  double rate = 0.0;
  if (fast) {
    const double sb = nitl > 0 ? D.ipre[nitl - 1] : 0.0;
    rate = itl_finish(c, J.ghost, J.e_before_bonds + sb, J.e_after_bonds, sa);
  } else if (act) {
    rate = itl_rate(c, W.wkey[a], j, dsel, nitl);
  }
  const double v = act ? rate / 2 : 0.0;
  if (rec) {
    S.rate[8 * r + 2 * i] = v;
    S.rate[8 * r + 2 * i + 1] = v;
  }
  const int ti = rec ? tie_index(v) : -1;
  uint64_t tm = ti >= 0 ? 1ull << ti : 0ull;
  uint32_t lo = (uint32_t)tm, hi = (uint32_t)(tm >> 32);
  lo |= __shfl_xor(lo, 1);
  hi |= __shfl_xor(hi, 1);
  lo |= __shfl_xor(lo, 2);
  hi |= __shfl_xor(hi, 2);
  if (rec && i == 0) S.tmask[r] = ((uint64_t)hi << 32) | lo;
  __threadfence_block();
  wave_sync();
}

// Evaluates the records idx[0 .. cnt), cnt <= 64 (a wave-uniform list in
// LDS): the vacancies eight per pass, lane (a, b) taking direction b of
// record a; the dumbbells sixteen per pass (itl_pass).
__device__ void eval_records(const Ctx& c, MixSector& S, const int32_t* idx, int cnt, int nitl, int64_t& n_eval) {
  const MixDesc& D = *c.D;
  ItlWork& W = *c.W;
  const int l = lane_id();
  const int r0 = l < cnt ? idx[l] : 0;
  const bool isv = l < cnt && D.rkind[r0] == K_VAC;
  const uint64_t mv = ballot(isv), mi = ballot(l < cnt && !isv);
  if (isv) W.vlist[__popcll(mv & ((1ull << l) - 1))] = r0;
  if (l < cnt && !isv) W.ilist[__popcll(mi & ((1ull << l) - 1))] = r0;
  __threadfence_block();
  wave_sync();
  const int nv = __popcll(mv), ni = __popcll(mi);
  for (int p = 0; p < nv; p += 8) {
    const int a = l >> 3, b = l & 7;
    const bool act = p + a < nv;
    const int r = act ? W.vlist[p + a] : 0;
    const double v = act ? vac_rate(c, D.rkey[r], b) : 0.0;
    if (act) S.rate[8 * r + b] = v;
    const int ti = act ? tie_index(v) : -1;
    const uint64_t tm = or8_u64(ti >= 0 ? 1ull << ti : 0ull);
    if (act && b == 0) S.tmask[r] = tm;
  }
  if (ni > 0 && nitl > 0)
    for (int p = 0; p < ni; p += kItlPass) itl_pass(c, S, W.ilist + p, min(kItlPass, ni - p), nitl);
  n_eval += cnt;
  __threadfence_block();
  wave_sync();
}

// ---------------------------------------------------------------- list edits
// Moves entry `from` of the sorted record arrays to position `to` (the
// entries between shift by one), or deletes it (to < 0).  Rates and tie masks
// move with their records; running sums are refolded by the caller.
__device__ void rec_shift(const MixDesc& D, int nr, int from, int to) {
  const int l = lane_id();
  int lo, hi, delta;
  if (to < 0) { lo = from + 1; hi = nr - 1; delta = -1; }
  else if (to >= from) { lo = from + 1; hi = to; delta = -1; }
  else { lo = to; hi = from - 1; delta = 1; }
  const int cnt = hi - lo + 1;
  if (cnt <= 0) return;
  const int rounds = (cnt + 63) >> 6;
  for (int rr = 0; rr < rounds; rr++) {
    const int q = delta < 0 ? rr : rounds - 1 - rr;
    const int v = lo + q * 64 + l;
    const bool act = v <= hi;
    const int vc = min(v, hi);
    const uint32_t k = D.rkey[vc];
    const uint8_t kd = D.rkind[vc];
    const int32_t sl = D.rslot[vc];
    const uint64_t tm = D.tmask[vc];
    double rt[8];
#pragma unroll
    for (int b = 0; b < 8; b++) rt[b] = D.rate[8 * vc + b];
    __threadfence_block();
    wave_sync();
    if (act) {
      const int w = v + delta;
      D.rkey[w] = k;
      D.rkind[w] = kd;
      D.rslot[w] = sl;
      D.tmask[w] = tm;
#pragma unroll
      for (int b = 0; b < 8; b++) D.rate[8 * w + b] = rt[b];
    }
    __threadfence_block();
    wave_sync();
  }
}

// The same for the dumbbell list (key, orientation; masks, terms and
// prefixes are refreshed by the caller).
__device__ void itl_shift(const MixDesc& D, int n, int from, int to) {
  const int l = lane_id();
  int lo, hi, delta;
  if (to < 0) { lo = from + 1; hi = n - 1; delta = -1; }
  else if (to >= from) { lo = from + 1; hi = to; delta = -1; }
  else { lo = to; hi = from - 1; delta = 1; }
  const int cnt = hi - lo + 1;
  if (cnt <= 0) return;
  const int rounds = (cnt + 63) >> 6;
  for (int rr = 0; rr < rounds; rr++) {
    const int q = delta < 0 ? rr : rounds - 1 - rr;
    const int v = lo + q * 64 + l;
    const bool act = v <= hi;
    const int vc = min(v, hi);
    const uint32_t k = D.ikey[vc];
    const int8_t o = D.io[vc];
    const uint8_t av = D.iav[vc];
    const double tm = D.iterm[vc];
    __threadfence_block();
    wave_sync();
    if (act) {
      const int w = v + delta;
      D.ikey[w] = k;
      D.io[w] = o;
      D.iav[w] = av;
      D.iterm[w] = tm;
    }
    __threadfence_block();
    wave_sync();
  }
}

// Inserts a record (sorted by key); returns its position.
__device__ int rec_insert(const MixDesc& D, int& nr, uint32_t key, int kind, int slot) {
  const int p = uni_i(lower_bound_key(D.rkey, nr, key));
  // Grow by one at the end, then move that entry to p.
  if (lane_id() == 0) {
    D.rkey[nr] = key;
    D.rkind[nr] = (uint8_t)kind;
    D.rslot[nr] = slot;
  }
  __threadfence_block();
  wave_sync();
  nr++;
  rec_shift(D, nr, nr - 1, p);
  if (lane_id() == 0) {
    D.rkey[p] = key;
    D.rkind[p] = (uint8_t)kind;
    D.rslot[p] = slot;
  }
  __threadfence_block();
  wave_sync();
  return p;
}

__device__ int itl_insert(const MixDesc& D, int& n, uint32_t key, int o) {
  const int p = uni_i(lower_bound_key(D.ikey, n, key));
  n++;
  itl_shift(D, n, n - 1, p);
  if (lane_id() == 0) {
    D.ikey[p] = key;
    D.io[p] = (int8_t)o;
  }
  __threadfence_block();
  wave_sync();
  return p;
}

// Index of vacancy key k in the unordered list (-1 if absent).
__device__ int find_vac(const MixDesc& D, int nvac, uint32_t k) {
  const int l = lane_id();
  for (int base = 0; base < nvac; base += 64) {
    const int i = base + l;
    const uint64_t m = ballot(i < nvac && D.vkey[i] == k);
    if (m) return base + first_set(m);
  }
  return -1;
}

// Removes vacancy key k from the unordered list (swap with the last) and
// keeps the record slots of the current sector in step.
__device__ void vac_remove(const MixDesc& D, int& nvac, int nr, uint32_t k) {
  int i = -1;
  const int r = uni_i(lower_bound_key(D.rkey, nr, k));
  if (r < nr && D.rkey[r] == k && D.rkind[r] == K_VAC) i = D.rslot[r];
  if (i < 0) i = find_vac(D, nvac, k);
  i = uni_i(i);
  if (i < 0) return;
  const uint32_t last = D.vkey[nvac - 1];
  if (lane_id() == 0) D.vkey[i] = last;
  __threadfence_block();
  wave_sync();
  nvac--;
  if (i < nvac) {
    const int rl = uni_i(lower_bound_key(D.rkey, nr, last));
    if (rl < nr && D.rkey[rl] == last && D.rkind[rl] == K_VAC && lane_id() == 0) D.rslot[rl] = i;
    __threadfence_block();
    wave_sync();
  }
}

// Collects the sector's vacancies and dumbbells as records sorted by key.
// Returns the count, or -1 beyond the capacity.
__device__ int collect(const Ctx& c, int s, int nvac, int nitl) {
  const MixDesc& D = *c.D;
  const int l = lane_id();
  int n = 0;
  for (int pass = 0; pass < 2; pass++) {
    const int tot = pass == 0 ? nvac : nitl;
    const uint32_t* src = pass == 0 ? D.vkey : D.ikey;
    for (int base = 0; base < tot; base += 64) {
      const int j = base + l;
      const uint32_t k = j < tot ? src[j] : 0u;
      const bool in = j < tot && key_sector(c.d, k) == s;
      const uint64_t m = ballot(in);
      const int r = n + __popcll(m & ((1ull << l) - 1));
      if (in && r < D.rcap) {
        D.rkey[r] = k;
        D.rkind[r] = (uint8_t)(pass == 0 ? K_VAC : K_ITL);
        D.rslot[r] = pass == 0 ? j : -1;
      }
      n += __popcll(m);
    }
  }
  if (n > D.rcap) return -1;
  __threadfence_block();
  wave_sync();
  // All-ascending bitonic network over [0, n) (positions >= n never touched).
  int n2 = 1;
  while (n2 < n) n2 <<= 1;
  for (int k = 2; k <= n2; k <<= 1) {
    for (int dd = k - 1; dd > 0; dd = (dd == k - 1) ? (k >> 2) : (dd >> 1)) {
      for (int i = l; i < n; i += 64) {
        const int p = i ^ dd;
        if (p > i && p < n) {
          const uint32_t a = D.rkey[i], b = D.rkey[p];
          if (a > b) {
            D.rkey[i] = b;
            D.rkey[p] = a;
            const uint8_t ka = D.rkind[i];
            D.rkind[i] = D.rkind[p];
            D.rkind[p] = ka;
            const int32_t sa = D.rslot[i];
            D.rslot[i] = D.rslot[p];
            D.rslot[p] = sa;
          }
        }
      }
      __threadfence_block();
      wave_sync();
    }
  }
  return n;
}

// Evaluates every record in [lo, hi) that is within reach of a changed site
// (or every dumbbell when `all_itl`, or every record when `every`), through
// the LDS index list.  Vacancies are evaluated chunk by chunk; dumbbells are
// gathered until 64 are pending, so that their passes run full.  Returns the
// lowest evaluated position (hi if none).
__device__ int eval_near(const Ctx& c, MixSector& S, int lo, int hi, const uint32_t* cs, int nc, bool all_itl,
                         bool every, int nitl, int32_t* list, int64_t& n_eval) {
  const MixDesc& D = *c.D;
  ItlWork& W = *c.W;
  const int l = lane_id();
  int first = hi, ni = 0;
  // Keys and kinds load four chunks ahead of the tests.
  uint32_t k0 = lo + l < hi ? D.rkey[lo + l] : 0u, k1 = lo + 64 + l < hi ? D.rkey[lo + 64 + l] : 0u;
  uint32_t k2 = lo + 128 + l < hi ? D.rkey[lo + 128 + l] : 0u, k3 = lo + 192 + l < hi ? D.rkey[lo + 192 + l] : 0u;
  uint8_t d0 = lo + l < hi ? D.rkind[lo + l] : 0, d1 = lo + 64 + l < hi ? D.rkind[lo + 64 + l] : 0;
  uint8_t d2 = lo + 128 + l < hi ? D.rkind[lo + 128 + l] : 0, d3 = lo + 192 + l < hi ? D.rkind[lo + 192 + l] : 0;
  for (int base = lo; base < hi; base += 64) {
    const int r = base + l;
    const uint32_t k = k0;
    const uint8_t kd = d0;
    k0 = k1;
    k1 = k2;
    k2 = k3;
    d0 = d1;
    d1 = d2;
    d2 = d3;
    k3 = r + 256 < hi ? D.rkey[r + 256] : 0u;
    d3 = r + 256 < hi ? D.rkind[r + 256] : 0;
    bool hit = false, dumb = false;
    if (r < hi) {
      dumb = kd == K_ITL;
      hit = every | (all_itl & dumb);
      for (int i = 0; i < nc; i++) hit = hit | in_reach(c.d, k, cs[i]);
    }
    const uint64_t m = ballot(hit);
    if (!m) continue;
    if (first == hi) first = base + first_set(m);
    const bool hv = hit & !dumb, hd = hit & dumb;
    const uint64_t mv = ballot(hv), md = ballot(hd);
    if (hv) list[__popcll(mv & ((1ull << l) - 1))] = r;
    if (hd) W.ibuf[ni + __popcll(md & ((1ull << l) - 1))] = r;
    __threadfence_block();
    wave_sync();
    ni += __popcll(md);
    if (mv) eval_records(c, S, list, __popcll(mv), nitl, n_eval);
    if (ni >= 64) {
      eval_records(c, S, W.ibuf, 64, nitl, n_eval);
      const int rest = ni - 64;
      const int32_t v = l < rest ? W.ibuf[64 + l] : 0;
      __threadfence_block();
      wave_sync();
      if (l < rest) W.ibuf[l] = v;
      __threadfence_block();
      wave_sync();
      ni = rest;
    }
  }
  if (ni > 0) eval_records(c, S, W.ibuf, ni, nitl, n_eval);
  return uni_i(first);
}

// Candidate of rec::RecList::stageCombination (recombine.cpp:84-140): the
// vacancy at key v takes the high end of the dumbbell at key dk (list index
// j), which keeps its low end; returns the energy change including -ef110.
__device__ double rec_delta(const Ctx& c, uint32_t v, uint32_t dk, int nitl) {
  const MixDesc& D = *c.D;
  int vx, vy, vz, dx, dy, dz;
  key_xyz(v, vx, vy, vz);
  key_xyz(dk, dx, dy, dz);
  const int64_t vs = site_of_key(c, v), ds = site_of_key(c, dk);
  const int td = D.types[ds], hi = hi_end(td), lo = lo_end(td);
  const View L0{D.types, -1, -1, 0, 0};
  const View L1{D.types, vs, ds, hi, lo};
  const double sb = nitl > 0 ? D.ipre[nitl - 1] : 0.0;
  int q = 0;
  while (q < nitl && !term_reads(c.d, D.ikey[q], v) && !term_reads(c.d, D.ikey[q], dk)) q++;
  double sa = q > 0 ? D.ipre[q - 1] : 0.0;
  for (; q < nitl; q++) {
    const uint32_t kq = D.ikey[q];
    if (term_reads(c.d, kq, v) || term_reads(c.d, kq, dk)) {
      int qx, qy, qz;
      key_xyz(kq, qx, qy, qz);
      sa += edumb_term(c.td, L1, qx, qy, qz, L1.at(site_of_key(c, kq)), D.io[q], D.iav[q]);
    } else {
      sa += D.iterm[q];
    }
  }
  const double e_before = itl::count(c.td, L0, vx, vy, vz, T_V) + itl::count(c.td, L0, dx, dy, dz, hi) + sb;
  const double e_after = itl::count(c.td, L1, vx, vy, vz, hi) + itl::count(c.td, L1, dx, dy, dz, lo) + sa;
  return (-1 * D.ef110) + e_after - e_before;
}

// Sectors that hold a vacancy or a dumbbell: bit s for sector s.
__device__ uint32_t occupied_sectors(const Ctx& c, int nvac, int nitl) {
  const MixDesc& D = *c.D;
  const int l = lane_id();
  uint32_t m = 0;
  for (int j = l; j < nvac; j += 64) m |= 1u << key_sector(c.d, D.vkey[j]);
  for (int j = l; j < nitl; j += 64) m |= 1u << key_sector(c.d, D.ikey[j]);
  uint32_t out = 0;
  for (int b = 0; b < 8; b++)
    if (ballot((m >> b) & 1u)) out |= 1u << b;
  return out;
}

// The first step k >= k0 at whose start a sector with clock e has time left:
// thr = fl((k + 1) T) - e > 0, i.e. fl((k + 1) T) > e, monotone in k.
// Returns kmax when that step lies at or beyond kmax (a NaN clock never runs
// again), and k0 when the search does not settle (no skipping then).
__device__ int64_t first_step_above(double e, double T, int64_t k0, int64_t kmax) {
  if (static_cast<double>(k0 + 1) * T > e) return k0;
  if (!(e < static_cast<double>(kmax) * T)) return kmax;
  int64_t k = (int64_t)(e / T) - 2;  // (k + 1) T <= e - T: not yet
  if (k < k0) k = k0;
  for (int i = 0; i < 64; i++, k++)
    if (static_cast<double>(k + 1) * T > e) return k < kmax ? k : kmax;
  return k0;
}

struct Draws {
  const MixDesc* D;
  MixState* st;
  __device__ bool have() const { return !D->rec || st->rec_pos < D->rec_end; }
  __device__ uint32_t next() {
    uint32_t r;
    if (D->rec) r = D->rec[st->rec_pos - D->rec_base];
    else r = (uint32_t)(xo_next(st->rng) >> 32);
    st->rec_pos++;
    return r;
  }
};

__global__ __launch_bounds__(64) void mix_kernel(const MixDesc* descs) {
  const MixDesc& D = descs[blockIdx.x];
  const uint64_t tick0 = __builtin_amdgcn_s_memrealtime();
  __shared__ double evo[8];
  __shared__ int32_t list[64];
  __shared__ uint32_t cs[8];
  const int l = lane_id();
  __shared__ ItlWork iw;
  Ctx c{&D, Dims{D.d.X, D.d.Y, D.d.Z}, D.d, &iw};
  MixSector S{D.rate, D.pvac, D.tmask, D.risk};
  MixState st = *D.st;
  if (l < 8) evo[l] = st.evo[l];
  wave_sync();
  int nvac = st.nvac, nitl = st.nitl, nr = 0;
  bool rebuild = true;
  double sum = 0.0;
  // A launch that stopped inside a sector left its records current: resume
  // on them (rebuilding would give the same rates and sums).
  if (D.rates_sector < 0 && st.in_sector && st.rec_valid) {
    rebuild = false;
    nr = st.nr;
    sum = st.rsum;
  }
  int32_t status = AKMC_RUN_OK;
  int64_t visits = 0;
  Draws dr{&D, &st};
  if (D.rates_sector >= 0) {
    st.in_sector = 1;
    st.t = 0.0;
    st.thr = 1.0;
    st.events = 0;
    for (int i = 0; i < 8; i++)
      if (ring_sector(i) == D.rates_sector) st.sect = i;
  }
  for (;;) {
    if (!st.in_sector) {
      if (st.step >= D.time_steps) { status = AKMC_RUN_TIME_DONE; break; }
      if (++visits > (int64_t)1 << 16) { status = AKMC_RUN_YIELD; break; }
      // Whole steps in which no event can happen: every occupied sector's
      // clock is ahead (thr <= 0: the visit changes nothing) and every empty
      // sector (zero total rate) only catches its clock up.  An empty sector
      // entering step j with clock e < a_j = fl((j + 1) T) ends it at
      // fl(e + fl(a_j - e)) = a_j exactly: e >= a_{j-1} >= a_j / 2 from the
      // step before (j >= 1), so the difference is exact (Sterbenz).  Such
      // steps are taken in one go, up to the first step at which an occupied
      // sector has time left.  Not with defect generation (no sector has a
      // zero total then), nor where the loop would stop at the next entry.
      if (st.sect == 0 && st.step >= 1 && !(D.gen_rate > 0) && D.rates_sector < 0 && st.events < D.max_events &&
          !(D.u && st.events >= D.rng_end) && !(D.rec && st.rec_pos + 4 > D.rec_end)) {
        const uint32_t occ = occupied_sectors(c, nvac, nitl);
        int64_t k = D.time_steps;
        for (int q = 0; q < 8; q++)
          if ((occ >> q) & 1u) k = min(k, first_step_above(evo[q], D.sector_T, st.step, D.time_steps));
        if (k > st.step) {
          const double a = static_cast<double>(k) * D.sector_T;  // a_{k-1}
          // The last visit of step k - 1 leaves its thr and t behind (the
          // loop's state where the time loop ends).
          const int s7 = ring_sector(7);
          double e7 = evo[s7];
          wave_sync();
          if (l < 8 && !((occ >> l) & 1u) && evo[l] < a) evo[l] = a;
          wave_sync();
          if (k - 1 > st.step) {
            const double ap = static_cast<double>(k - 1) * D.sector_T;  // a_{k-2}
            if (!((occ >> s7) & 1u) && e7 < ap) e7 = ap;
          }
          st.thr = a - e7;
          st.t = (!((occ >> s7) & 1u) && st.thr > 0.0) ? st.thr : 0.0;
          st.n_skipped += k - st.step;
          st.step = k;
          continue;
        }
      }
      const int s0 = ring_sector(st.sect);
      st.thr = static_cast<double>(st.step + 1) * D.sector_T - evo[s0];  // sublattice.inl:72
      st.t = 0.0;
      st.in_sector = 1;
      rebuild = true;
    }
    const int s = ring_sector(st.sect);
    if (!(st.t < st.thr)) {
      if (l == 0) evo[s] += st.t;
      wave_sync();
      st.in_sector = 0;
      if (++st.sect == 8) { st.sect = 0; st.step++; }
      continue;
    }
    if (st.events >= D.max_events) { status = AKMC_RUN_EVENT_CAP; break; }
    if (D.u && st.events >= D.rng_end) { status = AKMC_RUN_NEED_RNG; break; }
    // Worst case per event: a candidate draw and a generation's three.
    if (D.rec && st.rec_pos + 4 > D.rec_end) { status = MIX_NEED_REC; break; }
    if (rebuild) {
      __threadfence_block();
      nr = uni_i(collect(c, s, nvac, nitl));
      if (nr < 0) { status = AKMC_RUN_ERR_CAPACITY; break; }
      // Edumb terms and prefixes serve the sector's rates and the edits that
      // follow an event; a sector with no record and no generation has a zero
      // total and performs none.
      if (nr > 0 || D.gen_rate > 0 || D.rates_sector >= 0) {
        for (int base = 0; base < nitl; base += 64)
          if (base + l < nitl) refresh_term(c, base + l);
        __threadfence_block();
        wave_sync();
        prefix_from(c, 0, nitl);
      }
      eval_near(c, S, 0, nr, cs, 0, false, true, nitl, list, st.n_eval);
      sum = nr > 0 ? fold_range(S, 0, nr, 0.0, S.pvac) : 0.0;
      rebuild = false;
      if (D.rates_sector >= 0) {
        for (int i = l; i < nr && i < D.out_cap; i += 64) {
          D.out_keys[i] = D.rkey[i];
          D.out_kinds[i] = D.rkind[i];
          for (int b = 0; b < 8; b++) D.out_rates[8 * i + b] = D.rate[8 * i + b];
        }
        if (l == 0) { *D.out_total = sum + D.gen_rate; *D.out_n = nr; }
        status = AKMC_RUN_OK;
        break;
      }
    }
    // ABVIModel::calcRates ends with sum_rates += defectGenRate() (kmc.cpp:113).
    const double total = uni_d(sum + D.gen_rate);
    if (total == 0.0 || fabs(total) < 2.220446049250313e-16) {  // sublattice.inl:108-112
      st.t = st.thr;
      continue;
    }
    double u, lg;
    if (D.u) {
      u = D.u[st.events - D.rng_base];
      lg = D.l[st.events - D.rng_base];
    } else {
      u = (double)(uint32_t)(xo_next(st.rng) >> 32) * (1.0 / 4294967295.0);
      const double r = ((double)(xo_next(st.rng) >> 11) + 1.0) * 0x1p-53;
      lg = -log(r);
    }
    u = uni_d(u);
    lg = uni_d(lg);
    const double target = u * total;
    const int vsel = nr > 0 ? uni_i(first_above(D.pvac, nr, target)) : 0;
    int kind = K_GEN, slot = 0;
    if (vsel < nr) {
      kind = D.rkind[vsel];
      double p = vsel == 0 ? 0.0 : D.pvac[vsel - 1];
      slot = -1;
      for (int b = 0; b < 8; b++) {
        p += D.rate[8 * vsel + b];
        if (slot < 0 && p > target) slot = b;
      }
      if (slot < 0) { status = AKMC_RUN_ERR_NO_EVENT; break; }
    } else if (!(D.gen_rate > 0)) {
      status = AKMC_RUN_ERR_NO_EVENT;
      break;
    }
    kind = uni_i(kind);
    slot = uni_i(slot);

    // ---- perform
    int nc = 0;       // changed sites (cs)
    int edit_lo = nr; // lowest record position an edit touched
    int jlo = nitl;   // lowest dumbbell list position an edit touched
    uint32_t from = 0xFFFFFFFFu, to = 0xFFFFFFFFu;
    if (kind == K_VAC) {
      from = (uint32_t)uni_i((int)D.rkey[vsel]);
      int ox, oy, oz;
      nn1_offset((int)(key_x(from) & 1), slot, &ox, &oy, &oz);
      to = key_add(c.d, from, ox, oy, oz);
      const int64_t fs = site_of_key(c, from), ts = site_of_key(c, to);
      const int X = D.types[ts];
      const int32_t vs = D.rslot[vsel];
      if (l == 0) {
        D.types[fs] = (uint16_t)X;
        D.types[ts] = (uint16_t)T_V;
        D.vkey[vs] = to;
      }
      __threadfence_block();
      wave_sync();
      // The record follows the vacancy (or leaves the sector).
      if (key_sector(c.d, to) == s) {
        const uint32_t t2 = to;
        int p = uni_i(lower_bound_key(D.rkey, nr, t2));
        if (p > vsel) p--;
        rec_shift(D, nr, vsel, p);
        if (l == 0) { D.rkey[p] = to; D.rkind[p] = K_VAC; D.rslot[p] = vs; }
        __threadfence_block();
        wave_sync();
        edit_lo = min(vsel, p);
      } else {
        rec_shift(D, nr, vsel, -1);
        nr--;
        edit_lo = vsel;
      }
    } else if (kind == K_ITL) {
      from = (uint32_t)uni_i((int)D.rkey[vsel]);
      const int j = uni_i(lower_bound_key(D.ikey, nitl, from));
      const int o = D.io[j];
      const uint8_t dirs = dirs_of(o);
      const int tag = first_nn_by_slot(slot, dirs);
      int ox, oy, oz;
      nn1_offset((int)(key_x(from) & 1), tag, &ox, &oy, &oz);
      to = key_add(c.d, from, ox, oy, oz);
      const int64_t fs = site_of_key(c, from), ts = site_of_key(c, to);
      const int type = D.types[fs];
      const int g = ((c_high[o + 6] >> tag) & 1) ? hi_end(type) : lo_end(type);
      const int nd = combine(D.types[ts], g);
      const bool new_higher = hi_end(nd) == g;
      const int ha = o < 0 ? -o - 1 : o - 1;  // orientation::tranHash
      const int o2 = c_trans[ha][tag][new_higher ? 0 : 1][(slot & 1) ? 0 : 1];
      if (l == 0) {
        D.types[fs] = (uint16_t)diff_type(type, g);
        D.types[ts] = (uint16_t)nd;
      }
      __threadfence_block();
      wave_sync();
      // The list entry moves to the target (ItlList::mp keeps site order).
      int p = uni_i(lower_bound_key(D.ikey, nitl, to));
      if (p > j) p--;
      itl_shift(D, nitl, j, p);
      if (l == 0) { D.ikey[p] = to; D.io[p] = (int8_t)o2; }
      __threadfence_block();
      wave_sync();
      jlo = min(j, p);
      if (key_sector(c.d, to) == s) {
        int q = uni_i(lower_bound_key(D.rkey, nr, to));
        if (q > vsel) q--;
        rec_shift(D, nr, vsel, q);
        if (l == 0) { D.rkey[q] = to; D.rkind[q] = K_ITL; D.rslot[q] = -1; }
        __threadfence_block();
        wave_sync();
        edit_lo = min(vsel, q);
      } else {
        rec_shift(D, nr, vsel, -1);
        nr--;
        edit_lo = vsel;
      }
      st.n_itl_events++;
    } else {
      // Defect generation: a sector site and a third neighbour from the draws.
      const int wx = (s & 1) ? c.d.X - c.d.hx : c.d.hx, wy = (s & 2) ? c.d.Y - c.d.hy : c.d.hy;
      const int wz = (s & 4) ? c.d.Z - c.d.hz : c.d.hz;
      const int x0 = (s & 1) ? c.d.hx : 0, y0 = (s & 2) ? c.d.hy : 0, z0 = (s & 4) ? c.d.hz : 0;
      const uint64_t ncell = (uint64_t)wx * wy * wz;
      const uint64_t i = (uint64_t)dr.next() % ncell;
      const int b = (int)(dr.next() % 12u);
      from = make_key((uint32_t)(x0 + i % wx), (uint32_t)(y0 + (i / wx) % wy), (uint32_t)(z0 + i / ((uint64_t)wx * wy)));
      const uint32_t cdk = key_add(c.d, from, c_third[b][0], c_third[b][1], c_third[b][2]);
      const int64_t as = site_of_key(c, from), csite = site_of_key(c, cdk);
      const int ta = D.types[as], tc = D.types[csite];
      if (is_atom(ta) && is_atom(tc) && nvac < D.vcap && nitl < D.icap && nr + 2 <= D.rcap) {
        const int o = orientation_from_draw(dr.next());
        if (l == 0) {
          D.types[as] = (uint16_t)T_V;
          D.types[csite] = (uint16_t)combine(tc, ta);
          D.vkey[nvac] = from;
        }
        __threadfence_block();
        wave_sync();
        nvac++;
        const int pv = rec_insert(D, nr, from, K_VAC, nvac - 1);
        edit_lo = min(edit_lo, pv);
        const int pj = itl_insert(D, nitl, cdk, o);
        jlo = min(jlo, pj);
        if (key_sector(c.d, cdk) == s) {
          const int pi = rec_insert(D, nr, cdk, K_ITL, -1);
          edit_lo = min(edit_lo, pi);
        }
        to = from;  // recombination looks around the new vacancy
        if (l == 0) cs[nc] = cdk;
        nc++;
        st.n_generated++;
      }
    }
    if (l == 0) {
      if (from != 0xFFFFFFFFu) cs[nc] = from;
    }
    if (from != 0xFFFFFFFFu) nc++;
    if (to != 0xFFFFFFFFu && to != from) {
      if (l == 0) cs[nc] = to;
      nc++;
    }
    wave_sync();
    // Masks, terms and Edumb prefixes after the transition.
    {
      const int jm = refresh_near(c, cs, nc, nitl);
      jlo = min(jlo, jm);
      if (jlo < nitl) prefix_from(c, jlo, nitl);
    }
    // ---- recombination around the moved defect (rec::RecList::create,
    // recombine.cpp:35-82): 1nn then masked 2nn partners.
    if (to != 0xFFFFFFFFu) {
      const int64_t tsite = site_of_key(c, to);
      const int tt = D.types[tsite];
      const bool is_vac = tt == T_V;
      if (is_vac || is_dumbbell(tt)) {
        uint32_t pk[14];
        int np = 0;
        for (int q = 0; q < N1 + N2ALL; q++) {
          int ox, oy, oz;
          if (q < N1) {
            nn1_offset((int)(key_x(to) & 1), q, &ox, &oy, &oz);
          } else {
            if (!((itl::STATUS_2NN >> (q - N1)) & 1)) continue;
            nn2_offset(q - N1, &ox, &oy, &oz);
          }
          const uint32_t nk = key_add(c.d, to, ox, oy, oz);
          const int tn = D.types[site_of_key(c, nk)];
          if (is_vac ? is_dumbbell(tn) : tn == T_V) pk[np++] = nk;
        }
        if (np > 0) {
          // Lane q evaluates candidate q (all candidates at once).
          uint32_t mine = pk[0];
          for (int q = 1; q < np; q++) mine = l == q ? pk[q] : mine;
          double de = 0.0;
          if (l < np) de = rec_delta(c, is_vac ? to : mine, is_vac ? mine : to, nitl);
          // Stable sort by energy, then the candidates whose difference from
          // the lowest truncates to 0 as an int (abs() of a double in
          // pickMinimum resolves to the int overload, recombine.cpp:25).
          double dv[14];
          int ord[14];
          for (int q = 0; q < np; q++) {
            dv[q] = bcast_d(de, q);
            ord[q] = q;
          }
          for (int q = 1; q < np; q++) {
            const int oq = ord[q];
            int w = q;
            while (w > 0 && dv[ord[w - 1]] > dv[oq]) {
              ord[w] = ord[w - 1];
              w--;
            }
            ord[w] = oq;
          }
          const double dmin = dv[ord[0]];
          int ncand = 0;
          for (int q = 0; q < np; q++)
            if (trunc(dv[ord[q]] - dmin) == 0.0) ncand++;
          int pick = 0;
          if (ncand > 1) pick = (int)(dr.next() % (uint32_t)ncand);
          const int qsel = ord[pick];
          const uint32_t v = is_vac ? to : pk[qsel], dk = is_vac ? pk[qsel] : to;
          const int64_t vs = site_of_key(c, v), ds = site_of_key(c, dk);
          const int td = D.types[ds];
          if (l == 0) {
            D.types[vs] = (uint16_t)hi_end(td);
            D.types[ds] = (uint16_t)lo_end(td);
          }
          __threadfence_block();
          wave_sync();
          // Rec::recombine: both leave their lists.
          vac_remove(D, nvac, nr, v);
          {
            const int r = uni_i(lower_bound_key(D.rkey, nr, v));
            if (r < nr && D.rkey[r] == v) {
              rec_shift(D, nr, r, -1);
              nr--;
              edit_lo = min(edit_lo, r);
            }
          }
          {
            const int j = uni_i(lower_bound_key(D.ikey, nitl, dk));
            itl_shift(D, nitl, j, -1);
            nitl--;
            jlo = min(jlo, j);
            const int r = uni_i(lower_bound_key(D.rkey, nr, dk));
            if (r < nr && D.rkey[r] == dk) {
              rec_shift(D, nr, r, -1);
              nr--;
              edit_lo = min(edit_lo, r);
            }
          }
          if (l == 0) {
            cs[nc] = v;
            cs[nc + 1] = dk;
          }
          nc += 2;
          wave_sync();
          const int jm = refresh_near(c, cs + nc - 2, 2, nitl);
          jlo = min(jlo, jm);
          if (jlo < nitl) prefix_from(c, jlo, nitl);
          st.n_recombined++;
        }
      }
    }
    if (l == 0) {
      const int64_t li = st.events - D.log_base;
      if (D.log && li >= 0 && li < D.log_cap) {
        D.log[3 * li] = from;
        D.log[3 * li + 1] = to;
        D.log[3 * li + 2] = (uint32_t)kind;
      }
    }
    // ---- residence time (sublattice.inl:118-120)
    st.t += lg / total;
    st.events++;
    if (!(st.t < st.thr)) {
      rebuild = true;
      continue;
    }
    // (At an event or stream limit the records are still brought up to date,
    // so that the next launch resumes on them.)
    // ---- the next event's rates: records within reach of a changed site,
    // and every dumbbell when a term of Edumb changed.
    const bool all_itl = jlo < nitl || nitl != st.nitl;
    const int e_lo = eval_near(c, S, 0, nr, cs, nc, all_itl, false, nitl, list, st.n_eval);
    const int c_lo = min(e_lo, edit_lo);
    // Records before c_lo kept their sums; when nothing at or after it is
    // left (the last record deleted), the total is the last kept sum.
    if (nr == 0) sum = 0.0;
    else if (c_lo < nr) sum = fold_range(S, c_lo, nr, c_lo > 0 ? D.pvac[c_lo - 1] : 0.0, S.pvac);
    else sum = uni_d(D.pvac[nr - 1]);
    st.nitl = nitl;  // Edumb terms are current again
  }
  if (l == 0 && D.rates_sector < 0) {
    st.busy_ticks += (int64_t)(__builtin_amdgcn_s_memrealtime() - tick0);
    st.nr = nr;
    st.rsum = sum;
    st.rec_valid = (!rebuild && st.in_sector) ? 1 : 0;
    st.status = status;
    st.nvac = nvac;
    st.nitl = nitl;
    for (int i = 0; i < 8; i++) st.evo[i] = evo[i];
    *D.st = st;
  }
  if (l == 0 && D.rates_sector >= 0 && status != AKMC_RUN_OK) *D.out_n = -100 - status;
  if (l == 0 && D.rates_sector >= 0) D.st->rec_valid = 0;  // the hook used the record arrays
}

}  // namespace mix
}  // namespace akmc

extern "C" int akmc_mix_launch(const akmc::mix::MixDesc* d_descs, int32_t n, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(akmc::mix::mix_kernel, dim3((unsigned)n), dim3(64), 0, stream, d_descs);
  return (int)hipGetLastError();
}

namespace akmc {
namespace mix {

// Initial lattice of a generated job: site i takes the type whose cumulative
// ratio bracket holds splitmix64(seed ^ K (i + 1)) mod total (the replica
// engine's composition draw, akmc_batch.hip init_sites_kernel).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__global__ void fill_kernel(uint16_t* types, int64_t n, uint64_t seed, const uint32_t* cum, const uint16_t* tval,
                            int32_t nt) {
  const uint32_t total = cum[nt - 1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h = splitmix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(i + 1)));
    const uint32_t x = (uint32_t)(((h >> 32) * (uint64_t)total) >> 32);  // [0, total)
    uint16_t v = (uint16_t)T_FE;
    for (int t = nt - 1; t >= 0; t--)
      if (x < cum[t]) v = tval[t];
    types[i] = v;
  }
}
// Vacancies at vidx, dumbbells (the site's atom and `atom`) at iidx.
__global__ void defects_kernel(uint16_t* types, const int64_t* vidx, int64_t nv, const int64_t* iidx, int64_t ni,
                               int32_t atom) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) types[vidx[i]] = (uint16_t)T_V;
  if (i < ni) {
    const int64_t s = iidx[i];
    types[s] = (uint16_t)combine(types[s], atom);
  }
}

}  // namespace mix
}  // namespace akmc

extern "C" int akmc_mix_fill_launch(uint16_t* types, int64_t n, uint64_t seed, const uint32_t* cum,
                                    const uint16_t* tval, int32_t nt, hipStream_t stream) {
  if (n <= 0) return 0;
  const int64_t blocks = n / 256 + 1 < 8192 ? n / 256 + 1 : 8192;
  hipLaunchKernelGGL(akmc::mix::fill_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, types, n, seed, cum, tval,
                     nt);
  return (int)hipGetLastError();
}

extern "C" int akmc_mix_defects_launch(uint16_t* types, const int64_t* vidx, int64_t nv, const int64_t* iidx,
                                       int64_t ni, int32_t atom, hipStream_t stream) {
  const int64_t m = nv > ni ? nv : ni;
  if (m <= 0) return 0;
  hipLaunchKernelGGL(akmc::mix::defects_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, types, vidx,
                     nv, iidx, ni, atom);
  return (int)hipGetLastError();
}
