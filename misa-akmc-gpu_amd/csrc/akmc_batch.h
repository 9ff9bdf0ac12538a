// Host/device interface of the replica engine: many independent lattices
// ("replicas"), each advanced by one wavefront of a persistent kernel.
//
// The replica engine is the batched and served form of the vacancy-hop loop.
// Its physics is the reference's (same stencil, bond counting, side effect on
// stale species tags, perform rule, sublattice time loop), but its sums are
// not the reference's serial fold: each replica keeps per-vacancy totals and
// 64-wide block sums in FP64, selects through wave scans, and draws its
// random numbers from its own xoshiro256** stream.  Rates are FP64 (the
// reference's exp, bit for bit) or FP32 (exact integer-valued bond sums and a
// two-level FP32 table exp, relative error below 2e-7).
//
// A replica's trajectory depends only on its own parameters, initial lattice
// and seed: the sector state persists in the replica's workspace between
// launches, so the slicing of a run into launches, the other replicas of the
// batch and the kernel's tuning do not change it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/akmc.h"
#include "akmc_device.h"

// Replica status (ReplicaState::status): AKMC_REP_RUNNING, _TIME_DONE and
// _TARGET_DONE of include/akmc.h, or one of these errors.
enum {
  AKMC_REP_ERR_NO_EVENT = -10, // a positive total with no selectable event
  AKMC_REP_ERR_RANGE = -12     // an FP32 rate left the exp table's range
};

// Launch flags.
enum {
  AKMC_BF_ERROR_STATS = 1,  // FP32 rates: also evaluate the FP64 rate and record the error
  AKMC_BF_RATES_ONLY = 2    // parity hook: rebuild sector `rates_sector` and stop
};

namespace akmc {

constexpr int BATCH_NHI_MAX = 512;  // entries of the FP32 exp table's high level
constexpr int BATCH_EVB = 8;        // vacancies evaluated per gather pass

// Sector workspace of one replica: arrays for `cap` vacancies, at offsets
// that the host (global workspace) and the kernel (LDS copy) both use.
struct WsLayout {
  int64_t rate, vt, bsum, ssum, key, slot, bz, own, nb1, total;
};
AKMC_HD inline int64_t ws_align(int64_t x) { return (x + 15) & ~(int64_t)15; }
AKMC_HD inline WsLayout ws_layout(int64_t cap, int rate_bytes) {
  const int64_t nb = (cap + 63) / 64, ns = (nb + 63) / 64;
  WsLayout w;
  int64_t o = 0;
  w.rate = o; o = ws_align(o + 8 * cap * rate_bytes);
  w.vt = o; o = ws_align(o + 8 * cap);
  w.bsum = o; o = ws_align(o + 8 * nb);
  w.ssum = o; o = ws_align(o + 8 * ns);
  w.key = o; o = ws_align(o + 4 * cap);
  w.slot = o; o = ws_align(o + 4 * cap);
  w.bz = o; o = ws_align(o + 4 * nb);
  w.own = o; o = ws_align(o + cap);
  w.nb1 = o; o = ws_align(o + 8 * cap);
  w.total = o;
  return w;
}

// Brick layout of a replica's sites in device memory: bricks of 8 (x,
// doubled) x 4 (y) x 4 (z) sites, 128 bytes = one cache line each, stored
// x-fastest, then y, then z, the box padded to whole bricks.  A vacancy's
// 55-site stencil (x +-3, y +-2, z +-2) then spans about 7 lines instead of
// the ~26 rows of the plain x-fastest order.  Inside a brick: byte
// (z & 3) * 32 + (y & 3) * 8 + (x & 7).  The host sees logical order
// (x + X (y + Y z)) only: uploads and downloads convert.
struct BrickDims {
  int32_t nbx, nby, nbz;
};
AKMC_HD inline BrickDims brick_dims(const BoxDims& d) { return BrickDims{(d.X + 7) >> 3, (d.Y + 3) >> 2, (d.Z + 3) >> 2}; }
AKMC_HD inline int64_t brick_bytes(const BrickDims& b) { return (int64_t)b.nbx * b.nby * b.nbz * 128; }
AKMC_HD inline int64_t brick_index(int32_t nbx, int32_t nby, uint32_t x, uint32_t y, uint32_t z) {
  return ((((int64_t)(z >> 2) * nby + (y >> 2)) * nbx + (x >> 3)) << 7) | (int64_t)(((z & 3u) << 5) | ((y & 3u) << 3) | (x & 7u));
}

// Per-replica constants (read-only during a launch).
struct ReplicaDesc {
  uint8_t* sites;       // one flag byte per site, x + X*(y + Y*z), x doubled
  uint32_t* vac_key;    // key of every vacancy (its order is the sweep order)
  uint8_t* ws;          // sector workspace (ws_layout(cap, rate bytes))
  const float* exp_tab; // FP32 form: tab_lo[256] then tab_hi[nhi]
  BoxDims d;
  int32_t nvac;
  int32_t cap;
  int32_t nhi;          // entries of tab_hi
  int32_t m_min;        // barrier index of tab_hi[0] * tab_lo[0]
  int32_t nbx, nby;     // bricks along x and y (the site layout)
  int32_t pad;
  double nu;
  double kT;
  double sector_T;
  int64_t time_steps;   // ceil(physics_time / sector_T)
};

// Per-replica mutable state.
struct ReplicaState {
  double evo[8];        // sector clocks by sector id
  double thr;           // threshold of the sector in progress
  double t;             // partial time of the sector in progress
  int64_t step;
  int64_t events;
  int64_t target;       // stop after this many events
  uint64_t rng[4];      // xoshiro256** state
  int32_t sect;         // ring position 0..7
  int32_t in_sector;
  int32_t nv;           // entries of the sector list (dead ones included)
  int32_t status;       // AKMC_REP_*
  int64_t n_eval;       // vacancy evaluations
  int64_t n_rebuild;    // sector rebuilds
  int64_t n_jump;       // sector visits with a zero total rate
  // FP32 rate error against the FP64 rate (AKMC_BF_ERROR_STATS).
  double err_max;
  double err_sum;
  int64_t err_n;
  int64_t err_tiny;     // rates below FLT_MIN (FP64 value), not in the error
  int32_t sec_count[8]; // vacancies in each sector (kept by the kernel)
  int64_t n_skip;       // time steps skipped because no sector could have an event
  int64_t busy_ticks;   // s_memrealtime ticks (100 MHz) spent on this replica's work items
};

struct BatchArgs {
  const ReplicaDesc* desc;
  ReplicaState* state;
  const int32_t* items;     // replica index of each work item
  int32_t n_items;
  int32_t flags;            // AKMC_BF_*
  uint64_t* queue;          // [1]: launch start tick (set by the reset kernel)
  int64_t slice_events;     // events per replica per launch
  uint64_t budget_ticks;    // s_memrealtime ticks (100 MHz) after the launch start
                            // at which waves stop at the next event (0: none)
  int32_t region_bytes;     // LDS bytes per wave (LDS-resident kind)
  int32_t vcap;             // LDS capacity in vacancies (LDS-resident kind)
  int32_t rates_sector;     // AKMC_BF_RATES_ONLY: sector id to evaluate
  int32_t pad;
  // Bond tables in units of the energy quantum q (integer valued) and 2*e0/q.
  float e1q[25];
  float e2q[25];
  float e0q2[5];
  float pad2[3];
};

}  // namespace akmc

// il: interleave width (replica waves per SIMD: 1, 2 or 4; workgroups of
// 4 * il waves when il > 1).
extern "C" int akmc_batch_launch(const akmc::BatchArgs* a, int fp32, int lds, int il, int waves_per_group, int grid,
                                 hipStream_t stream);
extern "C" int akmc_batch_reset_launch(uint64_t* queue, hipStream_t stream);
extern "C" int akmc_batch_set_stencil(const akmc::StencilTables* t);
// Both take the replica's box and write its brick layout; site i of the
// logical order draws from counter i whatever the layout.
extern "C" int akmc_batch_init_sites_launch(uint8_t* sites, akmc::BoxDims d, uint64_t seed, const uint32_t* cum,
                                            const uint8_t* flag, int32_t n_types, hipStream_t stream);
extern "C" int akmc_batch_place_vacancies_launch(uint8_t* sites, akmc::BoxDims d, const int64_t* idx, int64_t n,
                                                 hipStream_t stream);
extern "C" int akmc_batch_region_bytes(int fp32, int vcap);
extern "C" int akmc_batch_occupancy(int fp32, int lds, int il, int waves_per_group, int region_bytes, int* blocks_per_cu);
