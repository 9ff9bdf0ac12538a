// Host/device interface of the plain FP64 event-loop kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_device.h"

// Why a launch of the event loop returned.
enum {
  AKMC_RUN_OK = 0,
  AKMC_RUN_TIME_DONE = 1,      // all ceil(physics_time / T) steps finished
  AKMC_RUN_EVENT_CAP = 2,      // max_events events performed (mid-sector stop)
  AKMC_RUN_NEED_RNG = 3,       // the uploaded random-number chunk is used up
  AKMC_RUN_YIELD = 4,          // sector-visit budget of one launch used up; relaunch
  AKMC_RUN_ERR_NO_EVENT = -10, // u * sum >= sum: no event above the target
  AKMC_RUN_ERR_CAPACITY = -11, // a sector holds more vacancies than the kernel's capacity
  AKMC_RUN_ERR_HELPER = -12    // a helper wavefront did not answer within its bound
};

namespace akmc {

// Sublattice time-loop state (SubLattice::startTimeLoop locals and the
// per-sector evolution_time of SectorMeta, reference sublattice.inl:25-143).
struct PlainState {
  double evo[8];      // evolution time by sector id
  double thr;         // threshold of the sector in progress
  double t;           // partial sector time of the sector in progress
  int64_t step;       // time step
  int64_t events;     // events performed so far
  int32_t sect;       // ring position 0..7 (ring 0,7,2,5,3,4,1,6)
  int32_t in_sector;  // 1 while a sector is in progress
  int32_t status;     // AKMC_RUN_* of the last launch
  int32_t sum_errors; // events whose running sums a serial refold disagreed with (check_sums)
  // Work counters (algorithmic traffic bookkeeping).
  int64_t n_eval;     // vacancy evaluations (55-site stencil gathers)
  int64_t n_fold;     // vacancies whose running sum was folded from their rates
  int64_t n_shift;    // vacancies whose running sum was shifted
};

struct PlainArgs {
  uint8_t* sites;        // one flag byte per site, x + X*(y + Y*z)
  uint32_t* vac_key;     // every vacancy's key
  int32_t nvac;
  BoxDims dims;
  PlainState* state;
  const double* u_stream;  // u_i = mt19937() * (1 / 4294967295) for event i
  const double* l_stream;  // l_i = -log(r_i), r_i the time stream draw of event i
  int64_t rng_base;        // event index of u_stream[0]
  int64_t rng_end;         // first event index not covered by the streams
  int64_t max_events;      // absolute event cap
  int64_t time_steps;      // ceil(physics_time / T)
  double sector_T;
  double nu;
  double kT;               // kB * T
  uint32_t* log;           // optional (from, to) keys per event
  int64_t log_base;
  int64_t log_cap;
  // Parity hook: when rates_sector >= 0 the kernel only evaluates that
  // sector (calcRates with its side effect) and writes the sorted keys,
  // rates and total, then returns.
  int32_t rates_sector;
  // Self-check: after every update of the running sums, every lane refolds
  // the sector serially in sweep order and compares (counted in sum_errors).
  int32_t check_sums;
  // Selection while the tail repair runs (LDS kernel): 0 from the estimated
  // sums when their bound decides, 1 never (always wait for the exact sums),
  // 2 only on events of odd index.  A test aid: every mode selects the same
  // events.
  int32_t select_mode;
  uint32_t* out_keys;
  double* out_rates;
  double* out_total;
  int32_t* out_n;
  int32_t out_cap;
  // Sector state in device memory (plain_kernel_hbm) for sectors holding more
  // vacancies than the LDS form: plain_sector_layout(sector_cap) bytes.
  int32_t sector_cap;
  uint8_t* sector_ws;
};

// Byte offsets of the device-memory sector arrays for a capacity of `cap`
// vacancies: key, slot, rate (8 per vacancy), pvac, tmask, risk, own, nb1
// (8 per vacancy); off[8] is the total.  Every array starts on 256 bytes.
__host__ __device__ inline void plain_sector_layout(int64_t cap, int64_t off[9]) {
  const int64_t sz[8] = {4 * cap, 4 * cap, 64 * cap, 8 * cap, 8 * cap, cap, cap, 8 * cap};
  int64_t o = 0;
  for (int i = 0; i < 8; i++) {
    off[i] = o;
    o += (sz[i] + 255) / 256 * 256;
  }
  off[8] = o;
}

}  // namespace akmc

extern "C" int akmc_plain_launch(const akmc::PlainArgs* args, hipStream_t stream);
extern "C" int akmc_plain_set_stencil(const akmc::StencilTables* t);
extern "C" int akmc_plain_vcap();
extern "C" int akmc_exp_launch(const double* d_x, double* d_y, int64_t n, hipStream_t stream);
extern "C" int akmc_copy_launch(const void* src, void* dst, int64_t bytes, hipStream_t stream);
