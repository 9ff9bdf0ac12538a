// Host/device interface of the vacancy + dumbbell event loop (akmc_mix.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_device.h"
#include "akmc_typed.h"

namespace akmc {
namespace mix {

// Why a launch returned for a lattice (AKMC_RUN_* of akmc_plain.h, plus:)
enum { MIX_NEED_REC = 5 };  // the stand-in draws for recombination / generation are used up

enum : uint8_t { K_VAC = 0, K_ITL = 1, K_GEN = 2 };

// Sector-loop state of one lattice (SubLattice::startTimeLoop locals and the
// sector clocks), kept in device memory between launches.
struct MixState {
  double evo[8];
  double thr, t;
  int64_t step, events;
  int32_t sect, in_sector, status, nvac;
  int32_t nitl, nr;         // dumbbells; records of the current sector
  double rsum;              // the records' total (valid with rec_valid)
  int32_t rec_valid, pad;   // 1: the sector's records, rates and sums are current (a launch resumes on them)
  int64_t rec_pos;         // stand-in draws consumed
  int64_t n_recombined;    // vacancy-dumbbell pairs recombined
  int64_t n_generated;     // Frenkel pairs generated
  int64_t n_eval;          // record evaluations
  int64_t n_itl_events;    // dumbbell transitions
  int64_t n_skipped;       // time steps taken in one go (no event could happen in them)
  int64_t busy_ticks;      // device time of this lattice's wavefront in its launches (100 MHz ticks)
  uint64_t rng[4];         // xoshiro256** state (device streams)
};

// One lattice as the kernel sees it (all device pointers).
struct MixDesc {
  uint16_t* types;          // LatticeTypes per site, x + X (y + Y z)
  uint32_t* vkey;           // vacancy keys, unordered (state.nvac)
  int32_t vcap, icap, rcap, pad0;
  uint32_t* ikey;           // dumbbell keys ascending (ItlList::mp order, state.nitl)
  int8_t* io;               // orientation of each
  uint8_t* iav;             // availability mask of each (from the current lattice)
  double* iterm;            // Edumb term of each (current lattice, current masks)
  double* ipre;             // left-fold prefix of the terms
  // Sector records, sorted by key: the sector's vacancies and dumbbells.
  uint32_t* rkey;
  uint8_t* rkind;
  int32_t* rslot;           // vacancy: index into vkey
  double* rate;             // 8 slots per record
  double* pvac;             // running sum after each record
  uint64_t* tmask;
  uint8_t* risk;
  MixState* st;
  const double* u;          // host streams (u_i, -log r_i) for event i, or null: device xoshiro
  const double* l;
  int64_t rng_base, rng_end;
  const uint32_t* rec;      // host stand-in draws (mt19937), index rec_pos - rec_base; null: device xoshiro
  int64_t rec_base, rec_end;
  uint32_t* log;            // 3 words per event: from key, to key (~0: none), kind
  int64_t log_base, log_cap;
  BoxDims d;
  int64_t max_events, time_steps;
  double sector_T, nu, kT, ef110, gen_rate;
  // Parity hook (rates_sector >= 0): evaluate that sector and write it out.
  int32_t rates_sector, out_cap;
  uint32_t* out_keys;
  uint8_t* out_kinds;
  double* out_rates;
  double* out_total;
  int32_t* out_n;
};

}  // namespace mix
}  // namespace akmc

extern "C" int akmc_mix_launch(const akmc::mix::MixDesc* d_descs, int32_t n, hipStream_t stream);
extern "C" int akmc_mix_fill_launch(uint16_t* types, int64_t n, uint64_t seed, const uint32_t* cum,
                                    const uint16_t* tval, int32_t nt, hipStream_t stream);
extern "C" int akmc_mix_defects_launch(uint16_t* types, const int64_t* vidx, int64_t nv, const int64_t* iidx,
                                       int64_t ni, int32_t atom, hipStream_t stream);
