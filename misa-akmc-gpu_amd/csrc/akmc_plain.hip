// Plain FP64 vacancy-hop event loop for one lattice on one GPU.
//
// One wavefront runs the reference's sublattice time loop
// (src/algorithms/sl/sublattice.inl:61-179) on the device: for each step the
// 8 sectors in ring order, and inside a sector the event loop
//   calcRates -> (sum == 0 ? jump to the threshold : select, perform, dt).
// The chain is serial by construction, so the loop is written for latency
// (akmc_wave.h): the sector's vacancy list, rates and running sums stay in
// LDS; after a hop only the vacancies whose 55-site stencil holds one of the
// two changed sites are re-evaluated, and the running sums are repaired
// exactly (block fold + tail shift) instead of re-folded.  In the LDS kernel
// two more wavefronts take the tail shift and the list edit off the event
// wave, which meanwhile selects the next event from sums known to within a
// proven bound (and falls back to waiting for the exact ones when that
// bound cannot decide), evaluates the hopped vacancy and advances the clock
// only with the exact total.
//
// Evaluation points match the reference.  The reference re-evaluates every
// vacancy of the sector at the start of every event (ABVIModel::calcRates,
// models/abvi/kmc.cpp:68-115), and that evaluation has one side effect
// (VacRatesSolver::deltaE moves the target's species tag to the vacancy and
// back, which drops a stale copy held by the vacancy,
// vacancy_rates_solver.cpp:81-163).  Here every vacancy of the sector is
// evaluated at the sector's first event and at every launch start (always an
// event start); afterwards, at the start of each further event of the sector,
// exactly the vacancies whose stencil changed are evaluated.  The others
// would reproduce their stored rates and an already-applied, idempotent side
// effect.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_exp_table.h"
#include "akmc_plain.h"
#include "akmc_wave.h"

namespace akmc {

__constant__ StencilTables c_sten;
__constant__ uint64_t c_exp_tab[256] = AKMC_EXP_TAB_INIT;
__constant__ double c_e1[25] = AKMC_E1_TABLE;
__constant__ double c_e2[25] = AKMC_E2_TABLE;
__constant__ double c_e0[5] = AKMC_E0_TABLE;

static constexpr int PLAIN_VCAP = 1472;
// SectorMeta ring 0,7,2,5,3,4,1,6 (sublattice.h:21-30), one nibble per
// position: arithmetic instead of a table load in the event loop.
__device__ __forceinline__ int ring_sector(int pos) { return (int)((0x61435270u >> (4 * pos)) & 7u); }

using PlainSector = SectorLDS<PLAIN_VCAP, double>;

__device__ void stage_tables(SharedTables& T) {
  const int l = lane_id();
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&c_sten);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&T.st);
  for (int i = l; i < (int)(sizeof(StencilTables) / 4); i += 64) dst[i] = src[i];
  for (int i = l; i < 25; i += 64) { T.e1[i] = c_e1[i]; T.e2[i] = c_e2[i]; }
  for (int i = l; i < 5; i += 64) T.e0[i] = c_e0[i];
  for (int i = l; i < 256; i += 64) T.exp_tab[i] = c_exp_tab[i];
  wave_sync();
}

// Collects the vacancies of sector s from the global list and sorts them by
// key (sweep order).  Returns the count, or -1 if it exceeds the capacity.
template <class SecT>
__device__ int collect_sector(SecT& S, const LatticeView& L, int s, int cap) {
  const int l = lane_id();
  int n = 0;
  for (int b = 0; b < L.nvac; b += 64) {
    const int j = b + l;
    const uint32_t k = j < L.nvac ? L.vac_key[j] : 0u;
    const bool in = j < L.nvac && key_sector(L.d, k) == s;
    const uint64_t m = ballot(in);
    const int r = n + __popcll(m & ((1ull << l) - 1));
    if (in && r < cap) { S.key[r] = k; S.slot[r] = j; }
    n += __popcll(m);
  }
  if (n > cap) return -1;
  wave_sync();
  sort_by_key(S.key, S.slot, n);
  return n;
}

// Moves entry `from_i` to position `to_i` (shifting the entries between), or
// deletes it when to_i < 0.  pvac is not moved (it is repaired afterwards).
template <class SecT>
__device__ void move_entry(SecT& S, int nv, int from_i, int to_i) {
  const int l = lane_id();
  int lo, hi, delta;
  if (to_i < 0) { lo = from_i + 1; hi = nv - 1; delta = -1; }
  else if (to_i >= from_i) { lo = from_i + 1; hi = to_i; delta = -1; }
  else { lo = to_i; hi = from_i - 1; delta = 1; }
  const int cnt = hi - lo + 1;
  if (cnt <= 0) return;
  const int rounds = (cnt + 63) >> 6;
  for (int rr = 0; rr < rounds; rr++) {
    // Moving down: process from the low end; moving up: from the high end.
    const int r = delta < 0 ? rr : rounds - 1 - rr;
    const int v = lo + r * 64 + l;
    const bool act = v <= hi;
    const int vc = min(v, hi);
    const uint32_t k = S.key[vc];
    const int32_t sl = S.slot[vc];
    const uint8_t o = S.own[vc];
    const uint64_t tm = S.tmask[vc];
    double rt[8];
    uint8_t nb[8];
#pragma unroll
    for (int b = 0; b < 8; b++) { rt[b] = S.rate[8 * vc + b]; nb[b] = S.nb1[8 * vc + b]; }
    wave_sync();
    if (act) {
      const int w = v + delta;
      S.key[w] = k; S.slot[w] = sl; S.own[w] = o; S.tmask[w] = tm;
#pragma unroll
      for (int b = 0; b < 8; b++) { S.rate[8 * w + b] = rt[b]; S.nb1[8 * w + b] = nb[b]; }
    }
    wave_sync();
  }
}

// The list edit after a hop of entry vsel to key `to`: the rank of `to`
// among the sector's other keys (the hop moves at most one z-plane, so a
// window around the old position first), the entries between moved by one,
// and the key written at its new position; an entry that left the sector is
// deleted.  Returns the new position (-1 if it left) and updates nv.
template <class SecT>
__device__ __forceinline__ int rank_and_move(SecT& S, int& nv, int vsel, uint32_t to, bool stays, int32_t vslot) {
  const int l = lane_id();
  int knew = -1;
  if (stays) {
    const int w0 = max(0, min(vsel - 32, nv - 64));
    const int wv = w0 + l;
    const uint32_t kw = S.key[min(wv, nv - 1)];
    const uint64_t below = ballot((wv < nv) & (wv != vsel) & (kw < to));
    // Every key before the window is below `to` and every key after it above.
    const bool inside = (w0 == 0 || S.key[w0 - 1] < to) && (w0 + 64 >= nv || S.key[w0 + 64] > to);
    if (inside) {
      knew = w0 + __popcll(below) - (vsel < w0 ? 1 : 0);
    } else {
      const int lb = lower_bound_key(S.key, nv, to);
      knew = lb > vsel ? lb - 1 : lb;
    }
  }
  move_entry(S, nv, vsel, stays ? knew : -1);
  if (stays) {
    if (l == 0) { S.key[knew] = to; S.slot[knew] = vslot; }
  } else {
    nv -= 1;
  }
  wave_sync();
  return knew;
}

// Candidates of re-evaluation: sorted positions whose keys lie within 2
// z-planes of the hop (the stencil's reach; `to` is one plane from `from`),
// unwrapped since a sector never wraps around the box.  A 64-entry window
// around position `pos` usually holds them all.
template <class SecT>
__device__ __forceinline__ void cand_range(const SecT& S, int nv, int pos, int zlo, int zhi, int Z, int& ilo,
                                           int& ihi) {
  const int c0 = max(0, min(pos - 32, nv - 64));
  const bool covers = (c0 == 0 || (int)key_z(S.key[c0 - 1]) < zlo) &&
                      (c0 + 64 >= nv || (int)key_z(S.key[c0 + 64]) > zhi);
  if (covers) {
    ilo = c0;
    ihi = min(nv, c0 + 64);
  } else {
    ilo = zlo <= 0 ? 0 : lower_bound_key(S.key, nv, make_key(0, 0, (uint32_t)zlo));
    ihi = zhi + 1 >= Z ? nv : lower_bound_key(S.key, nv, make_key(0, 0, (uint32_t)(zhi + 1)));
  }
}

// True if no position of [ilo, ihi) holds a vacancy whose stencil contains
// `from` or `to`, apart from the one that hopped (position knew, when it
// stayed in the sector): then it is the only one to re-evaluate.
template <class SecT>
__device__ __forceinline__ bool only_hopped(const SecT& S, const BoxDims& d, const StencilTables& st, int ilo,
                                            int ihi, int knew, uint32_t from, uint32_t to) {
  const int l = lane_id();
  bool only = true;
  for (int b0 = ilo; b0 < ihi; b0 += 64) {
    const int v = b0 + l;
    const uint32_t w = S.key[min(v, ihi - 1)];
    const bool hit = (v < ihi) & (v != knew) & (in_stencil(d, st, w, from) | in_stencil(d, st, w, to));
    only = only & (ballot(hit) == 0);
  }
  return only;
}

// The read-only half of the list edit after a hop of entry vsel from key
// `from` to key `to`, done on the list as it is before the edit (nv
// entries): the new position knew of the hopped vacancy (-1 if it left the
// sector) and the candidate range [ilo, ihi) and only_hopped flag as they
// are after it.  Positions before min(vsel, knew) and after max(vsel, knew)
// keep their entries; the hopped key lies inside the candidates' z-planes,
// so a range that covers them before the edit covers them after it, one
// shorter when the entry was deleted.
template <class SecT>
__device__ __forceinline__ void analyze_hop(const SecT& S, const BoxDims& d, const StencilTables& st, int nv,
                                            int vsel, uint32_t from, uint32_t to, bool stays, int zlo, int zhi,
                                            int& knew, int& ilo, int& ihi, bool& only) {
  const int l = lane_id();
  knew = -1;
  if (stays) {
    const int w0 = max(0, min(vsel - 32, nv - 64));
    const int wv = w0 + l;
    const uint32_t kw = S.key[min(wv, nv - 1)];
    const uint64_t below = ballot((wv < nv) & (wv != vsel) & (kw < to));
    // Every key before the window is below `to` and every key after it above.
    const bool inside = (w0 == 0 || S.key[w0 - 1] < to) && (w0 + 64 >= nv || S.key[w0 + 64] > to);
    if (inside) {
      knew = w0 + __popcll(below) - (vsel < w0 ? 1 : 0);
    } else {
      const int lb = lower_bound_key(S.key, nv, to);
      knew = lb > vsel ? lb - 1 : lb;
    }
  }
  cand_range(S, nv, vsel, zlo, zhi, d.Z, ilo, ihi);
  only = only_hopped(S, d, st, ilo, ihi, vsel, from, to);
  if (!stays) ihi -= 1;
}

// ---------------------------------------------------------------- helper waves
// The LDS kernel runs two helper wavefronts beside the event loop, each with
// its own command ring in LDS.
//  * The sum helper repairs the running-sum tail.  After the block fold the
//    event wave hands it the tail shift (TAIL) and goes on to the next event:
//    it selects from the exact prefix and the old tail sums plus the block's
//    change, each within a bound it checks, and collects the exact total
//    before the clock advances.
//  * The list helper does the list edit while the event wave evaluates the
//    hopped vacancy's rates into registers: first its read-only half
//    (analyze_hop: rank of the new key, candidate range, whether any vacancy
//    besides the hopped one needs re-evaluation), which may overlap the tail
//    repair, then, once the repair (which reads the rates and tie masks that
//    move) is done, the move of the entries.
// Counters are written by lane 0 after an LDS release fence and read by
// every lane before an LDS acquire fence (the handed-over state is all in
// LDS; a fence over global memory would also wait for the hop's stores to
// the lattice).  A helper runs its commands in order.  Every wait is
// bounded: one that runs out ends the launch with AKMC_RUN_ERR_HELPER
// instead of hanging.  Handing over only the tail's region shifts, with the
// event wave waiting for them, was measured and not kept (posting and
// waiting cost more than the stores).
enum : int32_t { CMD_ANALYZE = 1, CMD_TAIL = 2, CMD_EXIT = 3, CMD_COPY = 4 };
constexpr int MB_RING = 16;
constexpr int64_t MB_SPIN_LIMIT = (int64_t)1 << 26;  // x4: the bound of a wait, seconds
// A command: arguments in three 16-byte words, results in two, so that each
// side writes and reads them with a few wide LDS accesses.
//   ANALYZE  a0 = op, vsel, nv, stays; a1 = from, to, zlo, zhi;
//            a2 = tail repair to wait for before the move (-1: none), vslot
//            r0 = knew, ilo, ihi, only; r1 = status (nonzero: that wait ran out)
//   TAIL     a0 = op, v0, n; a1 = Pn, Po (the bits of two doubles)
//            r0 = total (bits), hi
//   COPY     a0 = op, vsel, stays
// A tail repair writes the new sums into a separate array (pnew) and leaves
// the old ones in pvac, where the event wave still reads them while it
// selects the next event.  The COPY posted once that selection is done
// moves pnew[v0, hi) into pvac; the list edit moves no running sums, and the
// positions it shifts are refolded anyway (a vacancy that left the sector
// shifts every position from vsel on, which the fold then covers, so those
// are not copied).
struct __align__(16) MbCmd {
  int4 a0;
  int4 a1;
  int4 a2;
  int4 r0;
  int4 r1;
};
__device__ __forceinline__ int4 pack_dd(double x, double y) {
  const uint64_t a = as_u64(x), b = as_u64(y);
  return make_int4((int)(uint32_t)a, (int)(uint32_t)(a >> 32), (int)(uint32_t)b, (int)(uint32_t)(b >> 32));
}
__device__ __forceinline__ double unpack_d(int lo, int hi) {
  return as_f64(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
struct Mailbox {
  int32_t posted;  // commands posted by the event wave
  int32_t done;    // commands completed by the helper wave
  int32_t pad[2];
  MbCmd q[MB_RING];
};

__device__ __forceinline__ void lds_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local"); }
__device__ __forceinline__ int mb_load(int32_t* p) {
  const int v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  lds_acquire();
  return uni_i(v);
}
__device__ __forceinline__ void mb_store(int32_t* p, int32_t v) {
  // The wave's earlier LDS writes (every lane's) complete before the count.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  if (lane_id() == 0) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Event wave: the slot of the next command (seq = commands posted so far),
// once the ring has room; nullptr if it stayed full past the spin limit.
// Lane 0 fills the slot, then mb_commit posts it.
__device__ __forceinline__ MbCmd* mb_reserve(Mailbox& mb, int seq) {
  int64_t spins = 0;
  while (seq - mb_load(&mb.done) >= MB_RING) {
    if (++spins > MB_SPIN_LIMIT) return nullptr;
    __builtin_amdgcn_s_sleep(1);
  }
  return &mb.q[seq & (MB_RING - 1)];
}
__device__ __forceinline__ int mb_commit(Mailbox& mb, int& seq) {
  const int k = seq;
  mb_store(&mb.posted, k + 1);
  seq = k + 1;
  return k;
}
// Waits until commands 0 .. k are done, polling without a sleep (a hand-over
// on the event chain costs the poll's LDS round trip, not a sleep quantum):
// ~2^28 polls, seconds, before it reports failure.
__device__ __forceinline__ bool mb_wait(Mailbox& mb, int k) {
  int64_t spins = 0;
  while (mb_load(&mb.done) <= k) {
    if (++spins > 4 * MB_SPIN_LIMIT) return false;
  }
  return true;
}

// Copies the new running sums [lo, hi) from pnew into pvac.
template <class SecT>
__device__ __forceinline__ void copy_sums(SecT& S, const double* pnew, int lo, int hi) {
  const int l = lane_id();
  for (int c = lo; c < hi; c += 256) {
    double x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = pnew[min(c + 64 * k + l, hi - 1)];
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (c + 64 * k + l < hi) S.pvac[c + 64 * k + l] = x[k];
  }
  wave_sync();
}

// tmb: the sum helper's mailbox (for the list helper's wait on a tail repair).
template <class SecT>
__device__ __forceinline__ void helper_loop(SecT& S, const SharedTables& T, const BoxDims d, Mailbox& mb,
                                            double* pnew, Mailbox* tmb) {
  const int l = lane_id();
  int pend_lo = 0, pend_hi = 0;  // new sums of the last tail repair, not yet in pvac
  for (int k = 0;; k++) {
    int64_t spins = 0;
    bool got = true;
    while (mb_load(&mb.posted) <= k) {
      // The event wave always posts CMD_EXIT before it ends; this bound only
      // keeps a broken protocol from running forever.
      if (++spins > 4 * MB_SPIN_LIMIT) { got = false; break; }
      __builtin_amdgcn_s_sleep(1);
    }
    if (!got) return;
    MbCmd& c = mb.q[k & (MB_RING - 1)];
    const int4 a0 = c.a0, a1 = c.a1, a2 = c.a2;
    const int op = uni_i(a0.x);
    if (op == CMD_EXIT) return;
    if (op == CMD_ANALYZE) {
      int knew, ilo, ihi;
      bool only;
      const int nv = uni_i(a0.z), vsel = uni_i(a0.y);
      const bool stays = uni_i(a0.w) != 0;
      const uint32_t to = (uint32_t)uni_i(a1.y);
      analyze_hop(S, d, T.st, nv, vsel, (uint32_t)uni_i(a1.x), to, stays, uni_i(a1.z), uni_i(a1.w), knew, ilo,
                  ihi, only);
      const int wait_k = uni_i(a2.x);
      int status = 0;
      // The entries between move by one (as move_entry does).  Up to 64 of
      // them are loaded before the wait for the tail repair, which only
      // reads the rates and tie masks, and stored after it.
      const int to_i = stays ? knew : -1;
      const int lo = to_i < 0 || to_i >= vsel ? vsel + 1 : to_i;
      const int hi = to_i < 0 ? nv - 1 : to_i >= vsel ? to_i : vsel - 1;
      const int delta = to_i < 0 || to_i >= vsel ? -1 : 1;
      if (hi - lo + 1 <= 64) {
        const int v = lo + l;
        const bool act = v <= hi;
        const int vc = hi >= lo ? min(v, hi) : 0;  // [lo, hi] lies in [0, nv)
        const uint32_t kk = S.key[vc];
        const int32_t sl = S.slot[vc];
        const uint8_t o = S.own[vc];
        const uint64_t tm = S.tmask[vc];
        double rt[8];
        uint8_t nb[8];
#pragma unroll
        for (int b = 0; b < 8; b++) { rt[b] = S.rate[8 * vc + b]; nb[b] = S.nb1[8 * vc + b]; }
        if (wait_k >= 0 && !mb_wait(*tmb, wait_k)) {
          status = 1;
        } else {
          wave_sync();
          if (act) {
            const int w = v + delta;
            S.key[w] = kk; S.slot[w] = sl; S.own[w] = o; S.tmask[w] = tm;
#pragma unroll
            for (int b = 0; b < 8; b++) { S.rate[8 * w + b] = rt[b]; S.nb1[8 * w + b] = nb[b]; }
          }
          wave_sync();
        }
      } else if (wait_k >= 0 && !mb_wait(*tmb, wait_k)) {
        status = 1;
      } else {
        move_entry(S, nv, vsel, to_i);
      }
      if (status == 0) {
        if (stays && l == 0) {
          S.key[knew] = to;
          S.slot[knew] = uni_i(a2.y);
        }
        wave_sync();
      }
      if (l == 0) {
        c.r0 = make_int4(knew, ilo, ihi, only ? 1 : 0);
        c.r1 = make_int4(status, 0, 0, 0);
      }
    } else if (op == CMD_TAIL) {
      // The exact running-sum repair behind the changed block.
      const double Pn = uni_d(unpack_d(a1.x, a1.y)), Po = uni_d(unpack_d(a1.z, a1.w));
      int hi = 0;
      const double tot = tail_update<true>(S, uni_i(a0.y), uni_i(a0.z), Pn, Po, pnew, hi);
      pend_lo = uni_i(a0.y);
      pend_hi = hi;
      const uint64_t tb = as_u64(tot);
      if (l == 0) c.r0 = make_int4((int)(uint32_t)tb, (int)(uint32_t)(tb >> 32), hi, 0);
    } else if (op == CMD_COPY) {
      copy_sums(S, pnew, pend_lo, uni_i(a0.z) ? pend_hi : min(pend_hi, uni_i(a0.y)));
      pend_lo = pend_hi = 0;
    }
    mb_store(&mb.done, k + 1);
  }
}

// Self-check: the running sums and the total against a serial refold of the
// sector's rates in sweep order (every lane folds the same values, so the
// result is wave-uniform).
template <class SecT>
__device__ bool sums_agree(const SecT& S, int nv, double sum) {
  wave_sync();
  double p = 0.0;
  bool ok = true;
  for (int v = 0; v < nv; v++) {
#pragma unroll
    for (int b = 0; b < 8; b++) p += S.rate[8 * v + b];
    ok = ok & (p == S.pvac[v]);
  }
  return ok & (p == (nv > 0 ? sum : 0.0));
}

// First index i in [0, n) whose estimated running sum exceeds t, n if none:
// pvac[i] itself below pre_n, the old sum pvac[i] plus D from there on (a
// tail whose repair is still running).  The search assumes sorted sums; the
// caller checks the margins around whatever it returns.
__device__ __forceinline__ int first_above_est(const double* pvac, int n, int pre_n, double D, double t) {
  if (n <= 0) return 0;
  const int l = lane_id();
  auto pred = [=](int i) {
    const double p = pvac[i];
    return (i < pre_n ? p : p + D) > t;
  };
  int lo = 0, len = n;
  while (len > 64) {
    const int S = (len + 63) >> 6;
    const int ci = lo + min(len - 1, l * S + S - 1);
    const int f = first_set(ballot((l * S < len) & pred(ci)));
    if (f == 64) return n;
    lo += f * S;
    len = min(S, len - f * S);
  }
  const int g = first_set(ballot((l < len) & pred(lo + min(l, len - 1))));
  return g == 64 ? n : lo + g;
}

// Event wave: waits for the tail repair posted as command k; its total and
// the end of the range it wrote into pnew.
__device__ __forceinline__ bool tail_collect(Mailbox& mb, int k, double& tot, int& hi) {
  if (!mb_wait(mb, k)) return false;
  const int4 r0 = mb.q[k & (MB_RING - 1)].r0;
  tot = uni_d(unpack_d(r0.x, r0.y));
  hi = uni_i(r0.z);
  return true;
}

// The event loop over a sector state S of capacity `cap` (LDS or device
// memory); T and evo[8] are LDS.
// HELPER: two helper wavefronts run helper_loop on the same LDS state, mb
// and mbc are the mailboxes of the sum helper and the list helper and pnew
// the tail-repair array (the LDS kernel); otherwise all three are unused.
// AIDS: the test aids and hooks of the arguments (check_sums, select_mode,
// the event log, the parity hook rates_sector) are read; the instance
// without them is the one launched when all are off.
template <class SecT, bool HELPER, bool AIDS>
__device__ __forceinline__ void plain_body(const PlainArgs& a, SecT& S, int cap, SharedTables& T, double* evo,
                                           Mailbox* mb, Mailbox* mbc, double* pnew) {
  const int l = lane_id();
  stage_tables(T);
  LaneStencil LS = lane_stencil(T.st);
  lane_linear(LS, T.st, a.dims);
  const LatticeView L{a.sites, a.vac_key, a.nvac, a.dims, a.nu, a.kT};
  const Patch nopatch{0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, -1, ~0ull, ~0ull};
  PlainState st = *a.state;  // every lane keeps an identical copy
  const int rsec = AIDS ? a.rates_sector : -1;  // the parity hook's sector (-1: none)
  if (l < 8) evo[l] = a.state->evo[l];
  wave_sync();
  bool rebuild = true;
  double sum = 0.0;
  int nv = 0;
  int s = ring_sector(st.sect);
  int32_t status = AKMC_RUN_OK;
  int64_t visits = 0;  // sector visits this launch (bounded so a launch always returns)
  int mb_seq = 0;      // commands posted to the sum helper
  int mc_seq = 0;      // commands posted to the list helper
  // A running-sum tail handed to the helper wave and not yet collected:
  // its command index (-1: none), the total before the event and its
  // estimate after it, the shift d_est = fl(Pn - Po) of the old tail sums,
  // and the exact prefix (positions < pre_n, sum pre_sum).
  int tail_k = -1;
  double t_old = 0.0, t_apx = 0.0, d_est = 0.0, pre_sum = 0.0;
  int pre_n = 0;
  int copy_k = -1;  // a COPY posted and not yet waited for
  // Random numbers of the next event, loaded one event ahead.
  double u_next = 0.0, l_next = 0.0;
  if (rsec < 0 && a.u_stream && a.l_stream && st.events >= a.rng_base && st.events < a.rng_end) {
    u_next = a.u_stream[st.events - a.rng_base];
    l_next = a.l_stream[st.events - a.rng_base];
  }

  if (rsec >= 0) {
    // Parity hook: evaluate one sector as that sector's first event would.
    st.in_sector = 1;
    st.t = 0.0;
    st.thr = 1.0;
    st.events = 0;
    for (int i = 0; i < 8; i++)
      if (ring_sector(i) == rsec) st.sect = i;
  }
  // After an event that went the whole way through the loop body the checks
  // below all pass (its residence step has tested the clock against the
  // threshold and the event counts against their limits): skip them.
  bool cont = false;
  for (;;) {
    if (!cont) {
      if (!st.in_sector) {
        if (st.step >= a.time_steps) { status = AKMC_RUN_TIME_DONE; break; }
        if (++visits > (int64_t)1 << 16) { status = AKMC_RUN_YIELD; break; }
        s = ring_sector(st.sect);
        // step_threshold_time (sublattice.inl:72)
        st.thr = static_cast<double>(st.step + 1) * a.sector_T - evo[s];
        st.t = 0.0;
        st.in_sector = 1;
        rebuild = true;
      }
      s = ring_sector(st.sect);
      if (!(st.t < st.thr)) {
        if (l == 0) evo[s] += st.t;  // sublattice.inl:140
        wave_sync();
        st.in_sector = 0;
        if (++st.sect == 8) { st.sect = 0; st.step++; }
        continue;
      }
      if (st.events >= a.max_events) { status = AKMC_RUN_EVENT_CAP; break; }
      if (st.events >= a.rng_end) { status = AKMC_RUN_NEED_RNG; break; }

      if (rebuild) {
        // The whole sector is re-read from HBM: make this wave's earlier
        // writes visible first.  (Within a sector the engine never reads back
        // a byte it wrote in the same event; those are patched from registers.)
        __threadfence_block();
        nv = collect_sector(S, L, s, cap);
        if (nv < 0) { status = AKMC_RUN_ERR_CAPACITY; break; }
        eval_list(S, T, LS, L, nullptr, 0, nv, nopatch, false, 0, -1);
        st.n_eval += nv;
        st.n_fold += nv;
        sum = fold_range(S, 0, nv, 0.0, S.pvac);
        if (AIDS && a.check_sums && !sums_agree(S, nv, sum)) st.sum_errors++;
        rebuild = false;
        if (rsec >= 0) {
          for (int i = l; i < nv && i < a.out_cap; i += 64) {
            a.out_keys[i] = S.key[i];
            for (int b = 0; b < 8; b++) a.out_rates[8 * i + b] = S.rate[8 * i + b];
          }
          if (l == 0) { *a.out_total = sum; *a.out_n = nv; }
          status = AKMC_RUN_OK;
          break;
        }
      }
    }
    cont = false;
    // A tail repair still with the helper: its total exceeds the exact
    // prefix sum, so the zero test below is decided unless that is tiny.
    if (HELPER && tail_k >= 0 && !(pre_sum >= 1e-10)) {
      int hi = 0;
      if (!tail_collect(*mb, tail_k, sum, hi)) { status = AKMC_RUN_ERR_HELPER; break; }
      copy_sums(S, pnew, pre_n, hi);
      tail_k = -1;
    }
    // sum_rates == 0 or below epsilon: jump to the threshold (sublattice.inl:108-112).
    if (tail_k < 0 && (sum == 0.0 || fabs(sum) < 2.220446049250313e-16)) {
      st.t = st.thr;
      continue;
    }

    // ---- select: first (vacancy, dir) whose running sum exceeds u * sum
    // (ModelAdapter::selectAndPerform, model_adapter.h:95-101; ABVIModel::select, kmc.cpp:242-309).
    const double lg = uni_d(l_next);
    int vsel = -1, bsel = -1;
    if (HELPER && tail_k >= 0) {
      // While the helper repairs the tail, every running sum is known to
      // within eps of its exact value: exact below pre_n, and from there on
      // old + d_est, off by at most one ulp of the running sum (<= ulp(tm))
      // per add of the tail (old and new folds round differently) plus the
      // rounding of d_est and of the sum; the slot's own adds and the two
      // products u * total add a few more.  If u * estimate lies farther
      // than eps from the running sums on either side of the slot it
      // selects, the exact target selects the same slot.
      const double u = uni_d(u_next);
      const double tm = 2.0 * fmax(t_old, t_apx);
      const bool try_est = !AIDS || a.select_mode == 0 || (a.select_mode == 2 && (st.events & 1));
      if (try_est && tm > 0x1p-969) {
        const double ulp_t = __builtin_ldexp(1.0, exponent_of(tm) - 52);
        const double eps = (16.0 * (double)(nv - pre_n) + 32.0) * ulp_t;
        const double tt = u * t_apx;
        const int v = first_above_est(S.pvac, nv, pre_n, d_est, tt);
        if (v < nv) {
          double r8[8];
#pragma unroll
          for (int b = 0; b < 8; b++) r8[b] = S.rate[v * 8 + b];
          const double pv = v == 0 ? 0.0 : S.pvac[v - 1];
          const double pb = v == 0 ? 0.0 : (v - 1 < pre_n ? pv : pv + d_est);
          double p = pb, prev = pb;
          int b_apx = -1;
          bool ok = true;
#pragma unroll
          for (int b = 0; b < 8; b++) {
            p += r8[b];
            if (b_apx < 0 && p > tt) {
              b_apx = b;
              ok = (p > tt + eps) & (prev <= tt - eps);
            }
            prev = p;
          }
          if (uni_i(ok && b_apx >= 0 ? 1 : 0)) {
            vsel = uni_i(v);
            bsel = uni_i(b_apx);
          }
        }
      }
      if (vsel < 0) {
        int hi = 0;
        if (!tail_collect(*mb, tail_k, sum, hi)) { status = AKMC_RUN_ERR_HELPER; break; }
        copy_sums(S, pnew, pre_n, hi);
        tail_k = -1;
      }
    }
    if (vsel < 0) {
      const double target = uni_d(u_next) * sum;
      vsel = first_above(S.pvac, nv, target);
      if (vsel >= nv) { status = AKMC_RUN_ERR_NO_EVENT; break; }
      {
        double r8[8];
#pragma unroll
        for (int b = 0; b < 8; b++) r8[b] = S.rate[vsel * 8 + b];
        double p = vsel == 0 ? 0.0 : S.pvac[vsel - 1];
#pragma unroll
        for (int b = 0; b < 8; b++) {
          p += r8[b];
          if (bsel < 0 && p > target) bsel = b;
        }
      }
      bsel = uni_i(bsel);
      if (bsel < 0) { status = AKMC_RUN_ERR_NO_EVENT; break; }
    }

    // ---- perform (ABVIModel::perform, kmc.cpp:358-396): the vacancy moves to
    // the target; the target's first species tag of Cu, Mn, Ni moves back.
    const uint32_t from = (uint32_t)uni_i((int)S.key[vsel]);
    const int par = (int)(key_x(from) & 1);
    int ox, oy, oz;
    nn1_offset(par, bsel, &ox, &oy, &oz);
    const uint32_t to = key_add(L.d, from, ox, oy, oz);
    // Gather the stencil around `to` now (lane = stencil site); the bytes are
    // used when the hopped vacancy is re-evaluated below, after the global
    // latency has been hidden behind the list update.
    uint8_t pf = 0;
    const uint64_t pq = stencil_site(L.d, T.st, LS, (uint32_t)uni_i((int)to));
    if (l < T.st.n_sites) pf = L.sites[pq];
    const uint8_t ff = S.own[vsel];
    const uint8_t gt = S.nb1[vsel * 8 + bsel];
    const uint8_t xb = (gt & BIT_CU) ? BIT_CU : (gt & BIT_MN) ? BIT_MN : (gt & BIT_NI) ? BIT_NI : 0;
    const uint8_t nf_from = (uint8_t)((ff & ~BIT_V) | xb);
    const uint8_t nf_to = (uint8_t)((gt & ~xb) | BIT_V);
    const int32_t vslot = S.slot[vsel];
    if (l == 0) {
      L.sites[key_index(L.d, from)] = nf_from;
      L.sites[key_index(L.d, to)] = nf_to;
      L.vac_key[vslot] = to;
      const int64_t li = st.events - a.log_base;
      if (AIDS && a.log && li >= 0 && li < a.log_cap) { a.log[2 * li] = from; a.log[2 * li + 1] = to; }
    }
    st.events++;
    if (st.events < a.rng_end) {
      u_next = a.u_stream[st.events - a.rng_base];
      l_next = a.l_stream[st.events - a.rng_base];
    }
    const bool last = st.events >= a.max_events || st.events >= a.rng_end;
    // ---- residence time (sublattice.inl:118-120): dt = -log(r) / sum, with
    // the exact total.  While the sum helper still repairs the tail, the
    // clock advances after the list analysis posted below, with the
    // repair's total.
    if (!HELPER || tail_k < 0 || last) {
      if (HELPER && tail_k >= 0) {
        int hi = 0;
        if (!tail_collect(*mb, tail_k, sum, hi)) { status = AKMC_RUN_ERR_HELPER; break; }
        copy_sums(S, pnew, pre_n, hi);
        tail_k = -1;
      }
      st.t += lg / sum;
      if (!(st.t < st.thr) || last) {
        rebuild = true;
        continue;
      }
    }

    // ---- the next event's calcRates, incrementally.
    const bool stays = key_sector(L.d, to) == s;
    const int zf = (int)key_z(from);
    const int zlo = min(zf, zf + oz) - 2, zhi = max(zf, zf + oz) + 2;
    int knew = -1, ilo = 0, ihi = 0;
    bool only = false;
    HoppedEval hop{};
    if (HELPER) {
      // The list helper analyses the edit while this wave evaluates the
      // hopped vacancy's rates (they do not depend on its new position) and
      // the sum helper may still be repairing the tail.  Every earlier
      // analysis is done, so the slot is free.
      MbCmd* c = &mbc->q[mc_seq & (MB_RING - 1)];
      if (l == 0) {
        c->a0 = make_int4(CMD_ANALYZE, vsel, nv, stays ? 1 : 0);
        c->a1 = make_int4((int)from, (int)to, zlo, zhi);
        c->a2 = make_int4(tail_k, vslot, 0, 0);
      }
      const int ka = mb_commit(*mbc, mc_seq);
      if (tail_k >= 0) {
        // This event's selection is done with the old tail sums: the sum
        // helper moves the new ones into pvac after its repair.
        MbCmd* cc = &mb->q[mb_seq & (MB_RING - 1)];
        if (l == 0) cc->a0 = make_int4(CMD_COPY, vsel, stays ? 1 : 0, 0);
        copy_k = mb_commit(*mb, mb_seq);
      }
      if (stays) {
        const Patch pt0{from, to, nf_from, nf_to, -1, key_index(L.d, from), key_index(L.d, to)};
        hop = eval_hopped_rates(S, T, LS, L, to, pt0, pf, pq);
      }
      if (!mb_wait(*mbc, ka)) { status = AKMC_RUN_ERR_HELPER; break; }
      const int4 r0 = mbc->q[ka & (MB_RING - 1)].r0;
      knew = uni_i(r0.x);
      ilo = uni_i(r0.y);
      ihi = uni_i(r0.z);
      only = uni_i(r0.w) != 0;
      if (uni_i(mbc->q[ka & (MB_RING - 1)].r1.x) != 0) { status = AKMC_RUN_ERR_HELPER; break; }
      if (!stays) nv -= 1;
      if (tail_k >= 0) {
        // The list helper waited for the repair before its move (and its
        // answer was read after an acquire), so the total is there.
        sum = uni_d(unpack_d(mb->q[tail_k & (MB_RING - 1)].r0.x, mb->q[tail_k & (MB_RING - 1)].r0.y));
        tail_k = -1;
        st.t += lg / sum;
        if (!(st.t < st.thr)) {
          // The sector ends: it is collected afresh at its next visit, once
          // the sum helper has stopped writing pvac.
          if (copy_k >= 0 && !mb_wait(*mb, copy_k)) { status = AKMC_RUN_ERR_HELPER; break; }
          copy_k = -1;
          rebuild = true;
          continue;
        }
      }
    } else {
      knew = rank_and_move(S, nv, vsel, to, stays, vslot);
      cand_range(S, nv, stays ? knew : vsel, zlo, zhi, L.d.Z, ilo, ihi);
    }
    int amin = 0x7FFFFFFF, amax = -1;
    const Patch pt{from, to, nf_from, nf_to, stays ? knew : -1, key_index(L.d, from), key_index(L.d, to)};
    if (HELPER && only) {
      // Only the vacancy that hopped changed its stencil (or none stayed).
      if (stays) {
        eval_hopped_store(S, L, knew, to, pt, hop);
        amin = amax = knew;
        st.n_eval += 1;
      }
    } else {
      for (int b0 = ilo; b0 < ihi; b0 += 64) {
        const int v = b0 + l;
        bool hit = false;
        {
          const uint32_t w = S.key[min(v, ihi - 1)];
          hit = (v < ihi) & (in_stencil(L.d, T.st, w, from) | in_stencil(L.d, T.st, w, to));
        }
        const uint64_t m = ballot(hit);
        if (stays && knew >= b0 && knew < b0 + 64 && m == 1ull << (knew - b0)) {
          // Only the vacancy that hopped: from the stencil gathered at the hop.
          amin = min(amin, knew);
          amax = max(amax, knew);
          eval_hopped(S, T, LS, L, knew, to, pt, pf, pq);
          st.n_eval += 1;
        } else if (m) {
          const int r = __popcll(m & ((1ull << l) - 1));
          if (hit) S.list[r] = v;
          amin = min(amin, b0 + first_set(m));
          amax = max(amax, b0 + 63 - (int)__clzll(m));
          wave_sync();
          eval_list(S, T, LS, L, S.list, 0, __popcll(m), pt, true, pf, stays ? knew : -1);
          st.n_eval += __popcll(m);
        }
      }
    }
    // ---- running sums.
    if (HELPER && copy_k >= 0) {
      if (!mb_wait(*mb, copy_k)) { status = AKMC_RUN_ERR_HELPER; break; }
      copy_k = -1;
    }
    if (nv == 0) {
      sum = 0.0;
    } else if (!stays) {
      const int c_lo = min(min(vsel, amin), nv);
      st.n_fold += nv - c_lo;
      sum = c_lo < nv ? fold_range(S, c_lo, nv, c_lo > 0 ? S.pvac[c_lo - 1] : 0.0, S.pvac)
                      : S.pvac[nv - 1];
    } else {
      const int c_lo = min(min(vsel, knew), amin);
      const int c_hi = max(max(vsel, knew), amax);
      const double Po = S.pvac[c_hi];
      const double Pn = fold_range(S, c_lo, c_hi + 1, c_lo > 0 ? S.pvac[c_lo - 1] : 0.0, S.pvac);
      st.n_fold += c_hi + 1 - c_lo;
      st.n_shift += nv - (c_hi + 1);
      if (HELPER && c_hi + 1 < nv) {
        // The helper repairs the tail while the next event is selected
        // (from the estimated sums when their bound decides) and performed.
        MbCmd* c = &mb->q[mb_seq & (MB_RING - 1)];
        if (l == 0) {
          c->a0 = make_int4(CMD_TAIL, c_hi + 1, nv, 0);
          c->a1 = pack_dd(Pn, Po);
        }
        tail_k = mb_commit(*mb, mb_seq);
        t_old = sum;
        d_est = uni_d(Pn - Po);
        t_apx = uni_d(sum + d_est);
        pre_n = c_hi + 1;
        pre_sum = Pn;
        if (AIDS && a.check_sums) {
          // The self-check compares the repaired sums: wait for them.
          int hi = 0;
          if (!tail_collect(*mb, tail_k, sum, hi)) { status = AKMC_RUN_ERR_HELPER; break; }
          copy_sums(S, pnew, pre_n, hi);
          tail_k = -1;
        }
      } else {
        int hi = 0;
        sum = c_hi + 1 < nv ? tail_update<false>(S, c_hi + 1, nv, Pn, Po, S.pvac, hi) : Pn;
      }
    }
    if (AIDS && a.check_sums && !sums_agree(S, nv, sum)) st.sum_errors++;
    cont = true;
  }
  if (HELPER) {
    MbCmd* c = mb_reserve(*mb, mb_seq);
    if (c) {
      if (l == 0) c->a0 = make_int4(CMD_EXIT, 0, 0, 0);
      mb_commit(*mb, mb_seq);
    }
    c = mb_reserve(*mbc, mc_seq);
    if (c) {
      if (l == 0) c->a0 = make_int4(CMD_EXIT, 0, 0, 0);
      mb_commit(*mbc, mc_seq);
    }
  }
  if (l == 0 && rsec < 0) {
    st.status = status;
    for (int i = 0; i < 8; i++) st.evo[i] = evo[i];
    *a.state = st;
  }
  if (l == 0 && rsec >= 0 && status != AKMC_RUN_OK) *a.out_n = status;
}

// Three wavefronts: the event loop (wave 0), the sum helper (wave 1) and the
// list helper (wave 2).
template <bool AIDS>
__device__ __forceinline__ void plain_lds(const PlainArgs& a) {
  __shared__ PlainSector S;
  __shared__ SharedTables T;
  __shared__ double evo[8];  // sector clocks (indexed by the running sector id)
  __shared__ Mailbox mb, mbc;
  __shared__ double pnew[PLAIN_VCAP];
  if (threadIdx.x == 0) {
    mb.posted = 0;
    mb.done = 0;
    mbc.posted = 0;
    mbc.done = 0;
  }
  __syncthreads();
  if (threadIdx.x < 64)
    plain_body<PlainSector, true, AIDS>(a, S, PLAIN_VCAP, T, evo, &mb, &mbc, pnew);
  else if (threadIdx.x < 128)
    helper_loop(S, T, a.dims, mb, pnew, nullptr);
  else
    helper_loop(S, T, a.dims, mbc, pnew, &mb);
}
__global__ __launch_bounds__(192) void plain_kernel(PlainArgs a) { plain_lds<false>(a); }
// The same with the test aids and hooks (akmc_set_check_sums,
// akmc_set_select_mode, the event log, the parity hook).
__global__ __launch_bounds__(192) void plain_kernel_aids(PlainArgs a) { plain_lds<true>(a); }

// The same loop with the sector arrays in device memory, for sectors of more
// than PLAIN_VCAP vacancies.
__global__ __launch_bounds__(64) void plain_kernel_hbm(PlainArgs a) {
  __shared__ SharedTables T;
  __shared__ double evo[8];
  __shared__ uint8_t sten[EVB][STENCIL_MAX];
  __shared__ uint8_t cflag[EVB];
  __shared__ int32_t list[64];
  int64_t off[9];
  plain_sector_layout(a.sector_cap, off);
  uint8_t* w = a.sector_ws;
  SectorRef S{reinterpret_cast<uint32_t*>(w + off[0]), reinterpret_cast<int32_t*>(w + off[1]),
              reinterpret_cast<double*>(w + off[2]),   reinterpret_cast<double*>(w + off[3]),
              reinterpret_cast<uint64_t*>(w + off[4]), w + off[5], w + off[6], w + off[7], sten, cflag, list};
  plain_body<SectorRef, false, true>(a, S, a.sector_cap, T, evo, nullptr, nullptr, nullptr);
}

}  // namespace akmc

extern "C" int akmc_plain_launch(const akmc::PlainArgs* args, hipStream_t stream) {
  if (args->sector_ws)
    hipLaunchKernelGGL(akmc::plain_kernel_hbm, dim3(1), dim3(64), 0, stream, *args);
  else if (args->check_sums || args->select_mode || args->log || args->rates_sector >= 0)
    hipLaunchKernelGGL(akmc::plain_kernel_aids, dim3(1), dim3(192), 0, stream, *args);
  else
    hipLaunchKernelGGL(akmc::plain_kernel, dim3(1), dim3(192), 0, stream, *args);
  return (int)hipGetLastError();
}

extern "C" int akmc_plain_set_stencil(const akmc::StencilTables* t) {
  return (int)hipMemcpyToSymbol(HIP_SYMBOL(akmc::c_sten), t, sizeof(*t));
}

extern "C" int akmc_plain_vcap() { return akmc::PLAIN_VCAP; }

namespace akmc {
// The device exp of the rate kernel over an array (verification hook).
__global__ void exp_kernel(const double* x, double* y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = exp_glibc(x[i], c_exp_tab);
}
}  // namespace akmc

namespace akmc {
// Plain streaming copy (HBM traffic calibration for the profiler counters).
__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace akmc

extern "C" int akmc_copy_launch(const void* src, void* dst, int64_t bytes, hipStream_t stream) {
  const int64_t n = bytes / 16;
  hipLaunchKernelGGL(akmc::copy_kernel, dim3(2048), dim3(256), 0, stream, (const uint4*)src, (uint4*)dst, n);
  return (int)hipGetLastError();
}

extern "C" int akmc_exp_launch(const double* d_x, double* d_y, int64_t n, hipStream_t stream) {
  if (n <= 0) return 0;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(akmc::exp_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_x, d_y, n);
  return (int)hipGetLastError();
}
