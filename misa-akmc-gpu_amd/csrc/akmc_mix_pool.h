// Host-side pool of vacancy + dumbbell lattices (akmc_mix_host.cpp): slots
// holding device-resident lattices, launched together, one wavefront each.
// Used by the C ABI (akmc_mix_*) and by the serving loop for jobs that carry
// dumbbells.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/akmc.h"
#include "akmc_mix.h"

namespace akmc_mixpool {

struct Pool;

int pool_create(int device, int32_t slots, Pool** out);
void pool_destroy(Pool* p);

// A lattice from host arrays (LatticeTypes per site, dumbbells ascending).
int pool_add(Pool* p, int32_t slot, const akmc_mix_params& prm, const uint16_t* types, int64_t n_itl,
             const int64_t* itl_sites, const int8_t* itl_orient);
// A lattice drawn on the device: the replica engine's composition law
// (akmc_batch's AKMC_INIT_DEVICE draw from comp's alloy and seeds, va_count
// vacancies), then n_itl dumbbells at distinct atom sites, each the site's
// atom and one of type itl_atom, with orientations drawn on the host.  Runs
// on `st` without waiting (the host data stays in the slot until
// pool_setup_done).  Device streams (xoshiro) drive its events.
int pool_add_generated(Pool* p, int32_t slot, const akmc_params& comp, int64_t n_itl, int32_t itl_atom, double ef110,
                       hipStream_t st);
void pool_setup_done(Pool* p, int32_t slot, hipStream_t st);
int pool_remove(Pool* p, int32_t slot, hipStream_t st);
// Device bytes a generated lattice of these parameters holds.  A generated
// lattice's device buffers are stream-ordered: allocated on the stream of
// pool_add_generated, freed on the stream given to pool_setup_done (set-up
// buffers) and pool_remove (after the work queued there).
int64_t pool_bytes(const akmc_params& comp, int64_t n_itl);

// Launches the given slots on `st`, each up to events limits[i] (absolute),
// and enqueues the copy of their states back (pinned, asynchronous); read
// pool_state once the stream has passed that copy.
int pool_launch(Pool* p, const std::vector<int32_t>& slots, const std::vector<int64_t>& limits, hipStream_t st);
const akmc::mix::MixState& pool_state(Pool* p, int32_t slot);
int pool_download(Pool* p, int32_t slot, uint16_t* types, int64_t n_sites);
int64_t pool_sites(Pool* p, int32_t slot);

}  // namespace akmc_mixpool
