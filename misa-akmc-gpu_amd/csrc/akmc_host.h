// Host-side helpers shared by the single-lattice engine (akmc_capi.cpp) and
// the replica engine (akmc_batch_host.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/akmc.h"
#include "akmc_model.h"

namespace akmc_host {

// Records the message returned by akmc_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);
extern thread_local std::string g_err;

// The 55-site rate stencil tables (akmc_model.h).
void build_stencil(akmc::StencilTables* S);

// The stencil tables, built once per process (thread-safe).
const akmc::StencilTables& stencil_tables();

// Copies the stencil tables into the __constant__ tables of both event-loop
// modules on the current device, once per device per process (thread-safe).
// Kernels of every engine on the device read these shared tables.
int upload_stencil(int device);

}  // namespace akmc_host

#define HIP_TRY(expr)                                                                                    \
  do {                                                                                                   \
    hipError_t _e = (expr);                                                                              \
    if (_e != hipSuccess) return akmc_host::fail(AKMC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)
