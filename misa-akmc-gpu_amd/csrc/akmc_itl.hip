// Self-interstitial (dumbbell) transition rates of the abvi model on a typed
// lattice, and the C entry point akmc_itl_rates.
//
// Lattice: one reference LatticeTypes value per real site (uint16: V 0, Fe 1,
// Cu 2, Ni 4, Mn 8, dumbbell XY = X << 8 | Y with X <= Y;
// src/lattice/lattice_types.h:20-36), index x + 2nx (y + ny z) with x
// doubled, periodic.  The dumbbells are listed in ascending site order, the
// order of the reference's ItlList::mp (a std::map keyed by lattice id,
// models/abvi/defect/itl_list.h), each with an orientation in [-6, 6] \ {0}
// (models/abvi/orientation.h:12-30).
//
// Per dumbbell, as Itl::beforeRatesUpdate then Itl::updateRates with
// ItlRatesSolver::rate as the callback (models/abvi/defect/itl_list.cpp:9-50,
// models/abvi/rate/rates_solver.cpp:13-26, itl_rates_solver.cpp:16-80):
//   avail  = orientation directions whose 1nn holds a single atom
//   for each available direction b: the jumping atom g (the high or low end
//   of XY by the orientation), dE = e_after - e_before with
//     e = count(src) + count(dst) + Edumb(all dumbbells)
//   before and after g moves (src keeps the other atom, dst becomes a
//   dumbbell), k = nu exp(-(e0(g) + dE / 2) / kB T), stored halved in the
//   direction's two rate slots.
// Edumb (models/abvi/rate/bonds/energy_dumbbell.cpp:7-206) is a left fold over
// the whole list, so each rate re-folds the list: the kernel keeps every
// dumbbell's own term from the unchanged lattice and recomputes only the terms
// of dumbbells within reach of src or dst; the fold is exact to the bit.
//
// All stored availability masks are computed before any rate (the reference
// reads each dumbbell's stored mask inside FeX_comp, energy_dumbbell.cpp:144).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/akmc.h"
#include "akmc_exp.h"
#include "akmc_exp_table.h"
#include "akmc_host.h"
#include "akmc_model.h"

using akmc_host::fail;

#include "akmc_typed.h"

namespace akmc {
namespace itl {

struct Args {
  Dims d;
  const uint16_t* types;
  const int64_t* sites;
  const int8_t* orient;
  int64_t n;
  double nu, kT;
  int4* info;        // per dumbbell: x, y, z, type | avail << 16 | (o + 6) << 24
  double* term;      // per dumbbell: its Edumb term on the unchanged lattice
  double* prefix;    // prefix[j]: left fold of term[0 .. j]
  int32_t* nbr;      // per dumbbell: list indices of the dumbbells near it, ascending
  int32_t* nbr_cnt;
  uint8_t* avail;    // out
  double* rates;     // out, 8 per dumbbell
  const int32_t* status;  // input check of the no-allocation entry point (non-zero: skip), or null
};

// Reach of the neighbour lists: every dumbbell whose Edumb term can read the
// dumbbell's site or one of its 1nn lies within |dx| <= 2 + 1 (doubled x),
// |dy| <= 1 + 1, |dz| <= 1 + 1 of it.
constexpr int NBR_RX = 3, NBR_RY = 2, NBR_RZ = 2;
constexpr int NBR_CAP = (2 * NBR_RX + 1) * (2 * NBR_RY + 1) * (2 * NBR_RZ + 1);

// The i-th smallest of the distinct values (c + d) mod n, d in [-R, R]
// (there are min(2R + 1, n) of them).
__device__ __forceinline__ int ring_value(int c, int R, int n, int i) {
  if (n <= 2 * R + 1) return i;
  const int lo = c - R, hi = c + R;
  if (lo < 0) return i <= hi ? i : n + lo + (i - hi - 1);
  if (hi >= n) {
    const int w = hi - n + 1;
    return i < w ? i : lo + (i - w);
  }
  return lo + i;
}

// List index of site s (sites ascending), -1 if it is not listed.
__device__ int64_t find_site(const int64_t* sites, int64_t n, int64_t s) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (sites[mid] < s) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && sites[lo] == s) ? lo : -1;
}

// Per dumbbell, one wavefront each: availability mask and Edumb term (lane
// 0), and the ascending list of dumbbells within reach: lane l takes the
// l-th candidate site in ascending index order (z, then y, then x over the
// ascending coordinate values), and a ballot keeps that order.
constexpr int PREP_THREADS = 64;
__global__ __launch_bounds__(PREP_THREADS) void itl_prep_kernel(Args a) {
  if (a.status && *a.status) return;
  const int64_t k = blockIdx.x;
  const int l = threadIdx.x;
  const int64_t s = a.sites[k];
  const int x = (int)(s % a.d.X), y = (int)((s / a.d.X) % a.d.Y), z = (int)(s / ((int64_t)a.d.X * a.d.Y));
  if (l == 0) {
    const int type = a.types[s];
    const int o = a.orient[k];
    const View L{a.types, -1, -1, 0, 0};
    uint8_t atoms = 0;
    for (int b = 0; b < N1; b++)
      if (is_atom(L.at(nn1(a.d, x, y, z, b)))) atoms |= (uint8_t)(1u << b);
    const uint8_t av = (uint8_t)((c_high[o + 6] | c_low[o + 6]) & atoms);
    a.avail[k] = av;
    a.term[k] = edumb_term(a.d, L, x, y, z, type, o, av);
    a.info[k] = make_int4(x, y, z, type | ((int)av << 16) | ((o + 6) << 24));
  }
  const int mx = min(2 * NBR_RX + 1, a.d.X), my = min(2 * NBR_RY + 1, a.d.Y), mz = min(2 * NBR_RZ + 1, a.d.Z);
  const int total = mx * my * mz;
  int32_t* out = a.nbr + k * NBR_CAP;
  int cnt = 0;
  for (int c0 = 0; c0 < total; c0 += PREP_THREADS) {
    const int ci = c0 + l;
    int64_t j = -1;
    if (ci < total) {
      const int ix = ci % mx, iy = (ci / mx) % my, iz = ci / (mx * my);
      const int64_t c = ring_value(x, NBR_RX, a.d.X, ix) +
                        (int64_t)a.d.X * (ring_value(y, NBR_RY, a.d.Y, iy) +
                                          (int64_t)a.d.Y * ring_value(z, NBR_RZ, a.d.Z, iz));
      if (is_dumbbell(a.types[c])) j = find_site(a.sites, a.n, c);
    }
    const uint64_t m = __ballot(j >= 0);
    if (j >= 0) out[cnt + __popcll(m & ((1ull << l) - 1))] = (int32_t)j;
    cnt += __popcll(m);
  }
  if (l == 0) a.nbr_cnt[k] = cnt;
}

// Lane i's double, read by every lane (i uniform).
__device__ __forceinline__ double read_lane(double v, int i) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, i);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), i);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Prefix folds of the terms: Edumb of the unchanged lattice and every partial
// sum of it, in list order.  One wavefront: 64 terms per load, folded in
// order through lane reads (every lane holds the same running sum).
__global__ __launch_bounds__(64) void itl_prefix_kernel(Args a) {
  if (a.status && *a.status) return;
  const int l = threadIdx.x;
  double acc = 0;
  double t_next = l < a.n ? a.term[l] : 0.0;
  for (int64_t base = 0; base < a.n; base += 64) {
    const int64_t j = base + l;
    const double t = t_next;  // the next chunk loads while this one folds
    t_next = j + 64 < a.n ? a.term[j + 64] : 0.0;
    const int cnt = (int)min<int64_t>(64, a.n - base);
    double mine = 0;
    if (cnt == 64) {
#pragma unroll
      for (int i = 0; i < 64; i++) {
        acc += read_lane(t, i);
        mine = l == i ? acc : mine;
      }
    } else {
      for (int i = 0; i < cnt; i++) {
        acc += read_lane(t, i);
        mine = l == i ? acc : mine;
      }
    }
    if (j < a.n) a.prefix[j] = mine;
  }
}

// sa folded with term[j .. end), eight loads ahead of the adds.
__device__ __forceinline__ double fold_terms(double sa, const double* __restrict__ tm, int64_t j, int64_t end) {
  for (; j + 8 <= end; j += 8) {
    double t[8];
#pragma unroll
    for (int i = 0; i < 8; i++) t[i] = tm[j + i];
#pragma unroll
    for (int i = 0; i < 8; i++) sa += t[i];
  }
  for (; j < end; j++) sa += tm[j];
  return sa;
}

// True if (x, y, z) is within the reach of one Edumb term of (cx, cy, cz):
// |dx| <= 2 (doubled x), |dy| <= 1, |dz| <= 1, periodically.
__device__ __forceinline__ bool near(const Dims& d, int x, int y, int z, int cx, int cy, int cz) {
  int dx = x - cx, dy = y - cy, dz = z - cz;
  dx = dx > d.X / 2 ? dx - d.X : dx < -(d.X / 2) ? dx + d.X : dx;
  dy = dy > d.Y / 2 ? dy - d.Y : dy < -(d.Y / 2) ? dy + d.Y : dy;
  dz = dz > d.Z / 2 ? dz - d.Z : dz < -(d.Z / 2) ? dz + d.Z : dz;
  return dx >= -2 && dx <= 2 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
}

// Rates: eight lanes per dumbbell, lane b for direction b.  Edumb after the
// jump is the prefix fold up to the first term the jump changes, then the
// fold continues term by term, recomputing the changed ones (found in the
// dumbbell's neighbour list); Edumb before is the full prefix fold.
constexpr int RATE_THREADS = 256;
__global__ __launch_bounds__(RATE_THREADS) void itl_rate_kernel(Args a) {
  if (a.status && *a.status) return;
  const double* tm = a.term;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = g >> 3;
  const int b = (int)(g & 7);
  const bool live = k < a.n;
  const int4 me = live ? a.info[k] : make_int4(0, 0, 0, 0);
  const int type = me.w & 0xFFFF;
  const uint8_t av = (uint8_t)((me.w >> 16) & 0xFF);
  const int o = (me.w >> 24) - 6;
  double rate = 0.0;
  if (live && ((av >> b) & 1)) {
    const int64_t src = site_of(a.d, me.x, me.y, me.z);
    const int lo = (me.x & 1) ? 0 : -1;
    const int tx = me.x + ((b & 4) ? 1 : -1), ty = me.y + lo + ((b >> 1) & 1), tz = me.z + lo + (b & 1);
    const int64_t dst = site_of(a.d, tx, ty, tz);
    const int dx = wrap(tx, a.d.X), dy = wrap(ty, a.d.Y), dz = wrap(tz, a.d.Z);
    // orientation::tranAtom (orientation.cc:238-253): X of XY in the high
    // end's directions, Y in the low end's.
    const int ghost = ((c_high[o + 6] >> b) & 1) ? (type >> 8) : (type & 0xFF);
    const int tdst = a.types[dst];
    // LatticeTypes::diff and combineToInter (lattice_types.cpp:21-48).
    const int src_after = (type & 0xFF) == ghost ? (type >> 8) : (type >> 8) == ghost ? (type & 0xFF) : type;
    const int dst_after = tdst <= ghost ? (tdst << 8) | ghost : (ghost << 8) | tdst;
    const View L0{a.types, -1, -1, 0, 0};
    const View L1{a.types, src, dst, src_after, dst_after};
    const double e_src_b = count(a.d, L0, me.x, me.y, me.z, ghost);
    const double e_des_b = count(a.d, L0, dx, dy, dz, tdst);
    const double e_src_a = count(a.d, L1, me.x, me.y, me.z, src_after);
    const double e_des_a = count(a.d, L1, dx, dy, dz, ghost);
    const double sb = a.prefix[a.n - 1];
    const int m = a.nbr_cnt[k];
    const int32_t* nb = a.nbr + k * NBR_CAP;
    double sa = sb;
    int64_t j = -1;  // next list index to fold, once the first changed term is found
    for (int q = 0; q < m; q++) {
      const int32_t ja = nb[q];
      const int4 qj = a.info[ja];
      if (!near(a.d, qj.x, qj.y, qj.z, me.x, me.y, me.z) && !near(a.d, qj.x, qj.y, qj.z, dx, dy, dz)) continue;
      if (j < 0) {
        sa = ja > 0 ? a.prefix[ja - 1] : 0.0;
        j = ja;
      }
      sa = fold_terms(sa, tm, j, ja);
      const int64_t sj = site_of(a.d, qj.x, qj.y, qj.z);
      sa += edumb_term(a.d, L1, qj.x, qj.y, qj.z, L1.at(sj), (qj.w >> 24) - 6, (uint8_t)((qj.w >> 16) & 0xFF));
      j = ja + 1;
    }
    if (j >= 0) sa = fold_terms(sa, tm, j, a.n);
    const double e_before = e_src_b + e_des_b + sb;
    const double e_after = e_src_a + e_des_a + sa;
    const double e0 = ghost == T_FE ? 0.32 : ghost == T_CU ? 0.36 : ghost == T_NI ? 0.45 : ghost == T_MN ? 0.31 : 0.0;
    const double ea = e0 + (e_after - e_before) / 2;
    rate = a.nu * exp_glibc(-ea / a.kT, c_exp_tab);
  }
  // Slot b holds half the rate of the (b / 2)-th orientation direction when
  // that direction is available (Itl::ratesIndex, itl_list.cpp:67-74).
  const uint8_t dirs = live ? (uint8_t)(c_high[o + 6] | c_low[o + 6]) : 0;
  int dsel = 0;
  for (int i = 0, seen = 0; i < N1; i++)
    if ((dirs >> i) & 1) {
      if (seen == (b >> 1)) dsel = i;
      seen++;
    }
  const double rd = __shfl(rate, dsel, 8);
  if (live) a.rates[8 * k + b] = ((av >> dsel) & 1) ? rd / 2 : 0.0;
}

}  // namespace itl
}  // namespace akmc

namespace {

template <class T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)); }
};

using akmc::itl::valid_dumbbell;
using akmc::itl::valid_single;

}  // namespace

extern "C" int akmc_itl_rates(const int64_t box[3], const uint16_t* types, int64_t n_itl, const int64_t* sites,
                              const int8_t* orient, double attempt_freq, double temperature, int device,
                              uint8_t* avail, double* rates, double* kernel_seconds) {
  if (kernel_seconds) *kernel_seconds = 0.0;
  if (!box || !types || n_itl < 0 || (n_itl > 0 && (!sites || !orient || !avail || !rates)))
    return fail(AKMC_ERR_INVALID, "null argument");
  if (n_itl >= (int64_t)1 << 31) return fail(AKMC_ERR_INVALID, "at most 2^31 - 1 dumbbells");
  for (int i = 0; i < 3; i++)
    if (box[i] < 3 || box[i] > (1 << 20)) return fail(AKMC_ERR_INVALID, "box dimension %d is %lld (3 .. 2^20)", i, (long long)box[i]);
  if (!(temperature > 0) || !(attempt_freq >= 0) || !std::isfinite(attempt_freq))
    return fail(AKMC_ERR_INVALID, "bad temperature or attempt frequency");
  const int64_t X = 2 * box[0], nsites = X * box[1] * box[2];
  for (int64_t s = 0; s < nsites; s++)
    if (!valid_single(types[s]) && !valid_dumbbell(types[s]))
      return fail(AKMC_ERR_INVALID, "site %lld holds 0x%x, not a lattice type", (long long)s, (unsigned)types[s]);
  for (int64_t k = 0; k < n_itl; k++) {
    if (sites[k] < 0 || sites[k] >= nsites) return fail(AKMC_ERR_INVALID, "dumbbell %lld: site out of the box", (long long)k);
    if (k > 0 && sites[k] <= sites[k - 1])
      return fail(AKMC_ERR_INVALID, "dumbbell sites must be strictly ascending (at %lld)", (long long)k);
    if (!valid_dumbbell(types[sites[k]]))
      return fail(AKMC_ERR_INVALID, "dumbbell %lld: site %lld holds 0x%x, not a dumbbell", (long long)k,
                  (long long)sites[k], (unsigned)types[sites[k]]);
    if (orient[k] == 0 || orient[k] < -6 || orient[k] > 6)
      return fail(AKMC_ERR_INVALID, "dumbbell %lld: orientation %d", (long long)k, (int)orient[k]);
  }
  if (n_itl == 0) return AKMC_OK;
  HIP_TRY(hipSetDevice(device));
  DevBuf<uint16_t> d_types;
  DevBuf<int64_t> d_sites;
  DevBuf<int8_t> d_orient;
  DevBuf<int4> d_info;
  DevBuf<double> d_term, d_prefix, d_rates;
  DevBuf<int32_t> d_nbr, d_nbr_cnt;
  DevBuf<uint8_t> d_avail;
  HIP_TRY(d_types.alloc(nsites));
  HIP_TRY(d_sites.alloc(n_itl));
  HIP_TRY(d_orient.alloc(n_itl));
  HIP_TRY(d_info.alloc(n_itl));
  HIP_TRY(d_term.alloc(n_itl));
  HIP_TRY(d_prefix.alloc(n_itl));
  HIP_TRY(d_nbr.alloc((size_t)n_itl * akmc::itl::NBR_CAP));
  HIP_TRY(d_nbr_cnt.alloc(n_itl));
  HIP_TRY(d_avail.alloc(n_itl));
  HIP_TRY(d_rates.alloc(8 * n_itl));
  // A stream of its own: the call waits for its own work only, not for an
  // engine or server running on the same device.
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};
  HIP_TRY(hipMemcpyAsync(d_types.p, types, nsites * sizeof(uint16_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d_sites.p, sites, n_itl * sizeof(int64_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d_orient.p, orient, n_itl * sizeof(int8_t), hipMemcpyHostToDevice, st));
  akmc::itl::Args a;
  a.d = akmc::itl::Dims{(int32_t)X, (int32_t)box[1], (int32_t)box[2]};
  a.types = d_types.p;
  a.sites = d_sites.p;
  a.orient = d_orient.p;
  a.n = n_itl;
  a.nu = attempt_freq;
  a.kT = AKMC_KB * temperature;  // RatesSolver::arrhenius: kB * T first
  a.info = d_info.p;
  a.term = d_term.p;
  a.prefix = d_prefix.p;
  a.nbr = d_nbr.p;
  a.nbr_cnt = d_nbr_cnt.p;
  a.avail = d_avail.p;
  a.rates = d_rates.p;
  a.status = nullptr;  // inputs were checked on the host above
  hipEvent_t ev[2];
  HIP_TRY(hipEventCreate(&ev[0]));
  if (hipEventCreate(&ev[1]) != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return fail(AKMC_ERR_HIP, "hipEventCreate failed");
  }
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() {
      (void)hipEventDestroy(e[0]);
      (void)hipEventDestroy(e[1]);
    }
  } evg{ev};
  HIP_TRY(hipEventRecord(ev[0], st));
  hipLaunchKernelGGL(akmc::itl::itl_prep_kernel, dim3((unsigned)n_itl), dim3(akmc::itl::PREP_THREADS), 0, st, a);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(akmc::itl::itl_prefix_kernel, dim3(1), dim3(64), 0, st, a);
  HIP_TRY(hipGetLastError());
  constexpr int rt = akmc::itl::RATE_THREADS;
  hipLaunchKernelGGL(akmc::itl::itl_rate_kernel, dim3((unsigned)((8 * n_itl + rt - 1) / rt)), dim3(rt), 0, st, a);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(ev[1], st));
  HIP_TRY(hipMemcpyAsync(avail, d_avail.p, n_itl, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(rates, d_rates.p, 8 * n_itl * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
  if (kernel_seconds) *kernel_seconds = ms * 1e-3;
  return AKMC_OK;
}

// ---------------------------------------------------------------- no allocation
// akmc_itl_rates_noalloc: the same kernels with caller-supplied device buffers
// and the working arrays (per dumbbell: info, term, prefix, neighbour list and
// count) in one static device array of a fixed maximum size.  Each device
// holds its own copy of the array, at its own address, from the library's
// load until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device.
namespace akmc {
namespace itl {

constexpr int64_t NOALLOC_MAX = AKMC_ITL_NOALLOC_MAX;
constexpr size_t SCR_STATUS = 0;                                  // int32 input-check flag
constexpr size_t SCR_INFO = 256;                                  // int4 x MAX
constexpr size_t SCR_TERM = SCR_INFO + 16 * NOALLOC_MAX;          // double x MAX
constexpr size_t SCR_PREFIX = SCR_TERM + 8 * NOALLOC_MAX;         // double x MAX
constexpr size_t SCR_NBR = SCR_PREFIX + 8 * NOALLOC_MAX;          // int32 x MAX x NBR_CAP
constexpr size_t SCR_NBR_CNT = SCR_NBR + 4 * NOALLOC_MAX * NBR_CAP;  // int32 x MAX
constexpr size_t SCR_BYTES = SCR_NBR_CNT + 4 * NOALLOC_MAX;

static __device__ __align__(256) unsigned char scratch[SCR_BYTES];

// The flag starts at zero on every call (a kernel, not a memset).
__global__ void itl_clear_kernel(int32_t* status) { *status = 0; }

// Device-side input check: the conditions akmc_itl_rates checks on the host.
// Any failure sets the flag, and the rate kernels then return at once, so an
// invalid input reads nothing out of bounds and writes no output.
__global__ __launch_bounds__(256) void itl_check_kernel(const uint16_t* types, int64_t nsites, const int64_t* sites,
                                                        const int8_t* orient, int64_t n, int32_t* status) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsites; i += stride) {
    const int t = types[i];
    bad = bad || !(valid_single(t) || valid_dumbbell(t));
  }
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const int64_t sk = sites[k];
    const int o = orient[k];
    if (sk < 0 || sk >= nsites) bad = true;
    else bad = bad || !valid_dumbbell(types[sk]) || (k > 0 && !(sites[k - 1] < sk));
    bad = bad || o == 0 || o < -6 || o > 6;
  }
  if (bad) *status = 1;  // every writer stores the same value
}

}  // namespace itl
}  // namespace akmc

namespace {
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;  // per device, from hipGetSymbolAddress

// The scratch array's device address on `device` (current), looked up once.
int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[(size_t)device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(akmc::itl::scratch));
    if (e != hipSuccess || !p)
      return fail(AKMC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[(size_t)device] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[(size_t)device];
  return AKMC_OK;
}
}  // namespace

extern "C" int akmc_itl_rates_noalloc(const int64_t box[3], const uint16_t* d_types, int64_t n_itl,
                                      const int64_t* d_sites, const int8_t* d_orient, double attempt_freq,
                                      double temperature, int device, void* stream, uint8_t* d_avail,
                                      double* d_rates) {
  using namespace akmc::itl;
  if (!box || !d_types || n_itl < 0 || (n_itl > 0 && (!d_sites || !d_orient || !d_avail || !d_rates)))
    return fail(AKMC_ERR_INVALID, "null argument");
  if (n_itl > NOALLOC_MAX)
    return fail(AKMC_ERR_CAPACITY, "%lld dumbbells: the no-allocation entry point takes at most %lld",
                (long long)n_itl, (long long)NOALLOC_MAX);
  for (int i = 0; i < 3; i++)
    if (box[i] < 3 || box[i] > (1 << 20)) return fail(AKMC_ERR_INVALID, "box dimension %d is %lld (3 .. 2^20)", i, (long long)box[i]);
  if (!(temperature > 0) || !(attempt_freq >= 0) || !std::isfinite(attempt_freq))
    return fail(AKMC_ERR_INVALID, "bad temperature or attempt frequency");
  if (device < 0) return fail(AKMC_ERR_INVALID, "bad device");
  if (n_itl == 0) return AKMC_OK;
  const int64_t X = 2 * box[0], nsites = X * box[1] * box[2];
  HIP_TRY(hipSetDevice(device));
  unsigned char* base = nullptr;
  {
    const int rc = scratch_address(device, &base);
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  int32_t* status = reinterpret_cast<int32_t*>(base + SCR_STATUS);
  Args a;
  a.d = Dims{(int32_t)X, (int32_t)box[1], (int32_t)box[2]};
  a.types = d_types;
  a.sites = d_sites;
  a.orient = d_orient;
  a.n = n_itl;
  a.nu = attempt_freq;
  a.kT = AKMC_KB * temperature;  // RatesSolver::arrhenius: kB * T first
  a.info = reinterpret_cast<int4*>(base + SCR_INFO);
  a.term = reinterpret_cast<double*>(base + SCR_TERM);
  a.prefix = reinterpret_cast<double*>(base + SCR_PREFIX);
  a.nbr = reinterpret_cast<int32_t*>(base + SCR_NBR);
  a.nbr_cnt = reinterpret_cast<int32_t*>(base + SCR_NBR_CNT);
  a.avail = d_avail;
  a.rates = d_rates;
  a.status = status;
  hipLaunchKernelGGL(itl_clear_kernel, dim3(1), dim3(1), 0, st, status);
  HIP_TRY(hipGetLastError());
  const int64_t work = std::max(nsites, n_itl);
  const unsigned cblocks = (unsigned)std::min<int64_t>(4096, (work + 255) / 256);
  hipLaunchKernelGGL(itl_check_kernel, dim3(cblocks), dim3(256), 0, st, d_types, nsites, d_sites, d_orient, n_itl,
                     status);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(itl_prep_kernel, dim3((unsigned)n_itl), dim3(PREP_THREADS), 0, st, a);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(itl_prefix_kernel, dim3(1), dim3(64), 0, st, a);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(itl_rate_kernel, dim3((unsigned)((8 * n_itl + RATE_THREADS - 1) / RATE_THREADS)),
                     dim3(RATE_THREADS), 0, st, a);
  HIP_TRY(hipGetLastError());
  int32_t flag = 0;
  HIP_TRY(hipMemcpyAsync(&flag, status, sizeof(flag), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (flag)
    return fail(AKMC_ERR_INVALID, "invalid input: a site type, a dumbbell site (range, order, type) or an orientation");
  return AKMC_OK;
}
