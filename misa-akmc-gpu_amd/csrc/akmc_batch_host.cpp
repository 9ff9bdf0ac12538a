// libakmc host side of the replica engine (akmc_batch_*) and of the serving
// loop (akmc_server_*): replica set-up, HBM layout, launch configuration,
// hipGraph capture, the autotuner and the request loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/akmc.h"
#include "akmc_batch.h"
#include "akmc_host.h"
#include "akmc_mix_pool.h"
#include "akmc_model.h"
#include "akmc_plain.h"

using akmc_host::fail;

namespace {

int64_t a256(int64_t x) { return (x + 255) & ~(int64_t)255; }

uint64_t splitmix_next(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------- FP32 form
// The bond energies (bonds_counter.cpp:13-33) and e0 (vacancy_rates_solver.cpp:
// 14-27) are decimal numbers with three digits; on the grid q = 10^-k eV
// (smallest k that represents every one of them) they are small integers,
// so the FP32 bond sums are exact.
struct QTables {
  bool ok = false;
  double q = 0;
  float e1q[25], e2q[25], e0q2[5];
  int32_t dmax = 0;  // bound on |dE| / q: four counts of 8 + 5 terms
};

const QTables& qtables() {
  static const QTables Q = [] {
    QTables t;
    const double e1[25] = AKMC_E1_TABLE, e2[25] = AKMC_E2_TABLE, e0[5] = AKMC_E0_TABLE;
    for (int k = 0; k <= 6 && !t.ok; k++) {
      const double q = std::pow(10.0, -k);
      auto fits = [&](double v) {
        const double m = v / q;
        return std::fabs(m - std::nearbyint(m)) < 1e-6 && std::fabs(m) < 4096.0;
      };
      bool good = true;
      for (int i = 0; i < 25; i++) good = good && fits(e1[i]) && fits(e2[i]);
      for (int i = 0; i < 5; i++) good = good && fits(2.0 * e0[i]);
      if (!good) continue;
      t.ok = true;
      t.q = q;
      int A = 0, B = 0;
      for (int i = 0; i < 25; i++) {
        t.e1q[i] = (float)std::nearbyint(e1[i] / q);
        t.e2q[i] = (float)std::nearbyint(e2[i] / q);
        A = std::max(A, (int)std::fabs(t.e1q[i]));
        B = std::max(B, (int)std::fabs(t.e2q[i]));
      }
      for (int i = 0; i < 5; i++) t.e0q2[i] = (float)std::nearbyint(2.0 * e0[i] / q);
      t.dmax = 4 * (8 * A + 5 * B);
    }
    return t;
  }();
  return Q;
}

// Rate table of one temperature: rate(m) = nu * exp(-m q / (2 kT)) for the
// barrier index m = 2 Ea / q, as tab_hi[(m - m_min) >> 8] * tab_lo[(m - m_min) & 255].
int make_exp_table(double kT, double nu, std::vector<float>* tab, int32_t* nhi, int32_t* m_min) {
  const QTables& Q = qtables();
  if (!Q.ok) return fail(AKMC_ERR_CONFIG, "FP32 rates need bond energies on a decimal grid");
  int lo = INT_MAX, hi = INT_MIN;
  for (int X = 1; X < 5; X++) {
    lo = std::min(lo, (int)Q.e0q2[X] - Q.dmax);
    hi = std::max(hi, (int)Q.e0q2[X] + Q.dmax);
  }
  const int span = hi - lo + 1;
  *nhi = (span + 255) / 256;
  *m_min = lo;
  if (*nhi > akmc::BATCH_NHI_MAX) return fail(AKMC_ERR_CONFIG, "FP32 rate table too large (%d)", *nhi);
  const double c = Q.q / (2.0 * kT);
  tab->assign(256 + *nhi, 0.0f);
  for (int i = 0; i < 256; i++) (*tab)[i] = (float)std::exp(-(double)i * c);
  for (int j = 0; j < *nhi; j++) (*tab)[256 + j] = (float)(nu * std::exp(-(double)(lo + 256 * j) * c));
  return AKMC_OK;
}

// ---------------------------------------------------------------- replicas
struct HostRep {
  akmc_params p{};
  akmc::BoxDims d{};
  int64_t nsites = 0;
  int64_t sbytes = 0;  // device bytes of the sites in the brick layout (>= nsites)
  int32_t cap = 1;
  int32_t nvac = 0;
  double kT = 0;
  int64_t time_steps = 0;
  int32_t nhi = 0, m_min = 0;
  int32_t sec_count[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // vacancies per sector
  std::vector<float> tab;
  int64_t bytes = 0;  // device bytes of the replica's block
  uint8_t* mem = nullptr;  // base of this replica's device block
  uint8_t* sites = nullptr;
  uint32_t* vac = nullptr;
  uint8_t* ws = nullptr;
  float* dtab = nullptr;
  int64_t* didx = nullptr;   // vacancy site draws (set-up only)
  uint32_t* dcum = nullptr;  // composition table (set-up only)
  uint8_t* dflag = nullptr;
  // Host data of an asynchronous set-up, alive until its stream is synchronized.
  std::vector<int64_t> h_idx;
  std::vector<uint32_t> h_keys;
  uint32_t h_cum[AKMC_MAX_ALLOY] = {};
  uint8_t h_flag[AKMC_MAX_ALLOY] = {};
  void release_setup() {
    std::vector<int64_t>().swap(h_idx);
    std::vector<uint32_t>().swap(h_keys);
  }
};

int setup_rep(HostRep* r, const akmc_params& p, bool fp32) {
  for (int d = 0; d < 3; d++)
    if (p.box[d] < 8) return fail(AKMC_ERR_INVALID, "box edge %lld < 8 cells is not supported", (long long)p.box[d]);
  if (2 * p.box[0] > 2047 || p.box[1] > 2047 || p.box[2] > 1023)
    return fail(AKMC_ERR_INVALID, "box exceeds 1023 x 2047 x 1023 cells");
  if (p.n_alloy <= 0 || p.n_alloy > AKMC_MAX_ALLOY) return fail(AKMC_ERR_INVALID, "bad alloy list");
  if (p.va_count < 0 || p.va_count > 262144)
    return fail(AKMC_ERR_CAPACITY, "va_count %lld outside 0..262144", (long long)p.va_count);
  if (!(p.sector_time > 0) || !(p.temperature > 0) || !(p.attempt_freq > 0))
    return fail(AKMC_ERR_INVALID, "sector_time, temperature and attempt_freq must be positive");
  r->p = p;
  r->d = akmc::box_dims(p.box);
  r->nsites = (int64_t)r->d.X * r->d.Y * r->d.Z;
  r->sbytes = akmc::brick_bytes(akmc::brick_dims(r->d));
  r->cap = (int32_t)std::max<int64_t>(1, p.va_count);
  r->kT = AKMC_KB * p.temperature;
  r->time_steps = (int64_t)std::ceil(p.physics_time / p.sector_time);
  if (fp32) {
    const int rc = make_exp_table(r->kT, p.attempt_freq, &r->tab, &r->nhi, &r->m_min);
    if (rc) return rc;
  }
  r->bytes = a256(r->sbytes) + a256(4 * (int64_t)r->cap) + a256(akmc::ws_layout(r->cap, fp32 ? 4 : 8).total) +
             a256(4 * (int64_t)(256 + r->nhi)) + a256(8 * (int64_t)r->cap) + a256(4 * AKMC_MAX_ALLOY) +
             a256(AKMC_MAX_ALLOY);
  return AKMC_OK;
}

void carve(HostRep* r, uint8_t* base, bool fp32) {
  r->mem = base;
  r->sites = base;
  base += a256(r->sbytes);
  r->vac = reinterpret_cast<uint32_t*>(base);
  base += a256(4 * (int64_t)r->cap);
  r->ws = base;
  base += a256(akmc::ws_layout(r->cap, fp32 ? 4 : 8).total);
  r->dtab = reinterpret_cast<float*>(base);
  base += a256(4 * (int64_t)(256 + r->nhi));
  r->didx = reinterpret_cast<int64_t*>(base);
  base += a256(8 * (int64_t)r->cap);
  r->dcum = reinterpret_cast<uint32_t*>(base);
  base += a256(4 * AKMC_MAX_ALLOY);
  r->dflag = base;
}

akmc::ReplicaDesc make_desc(const HostRep& r) {
  akmc::ReplicaDesc D{};
  D.sites = r.sites;
  D.vac_key = r.vac;
  D.ws = r.ws;
  D.exp_tab = r.dtab;
  D.d = r.d;
  D.nvac = r.nvac;
  D.cap = r.cap;
  D.nhi = r.nhi;
  D.m_min = r.m_min;
  const akmc::BrickDims bd = akmc::brick_dims(r.d);
  D.nbx = bd.nbx;
  D.nby = bd.nby;
  D.nu = r.p.attempt_freq;
  D.kT = r.kT;
  D.sector_T = r.p.sector_time;
  D.time_steps = r.time_steps;
  return D;
}

akmc::ReplicaState initial_state(const HostRep& r) {
  akmc::ReplicaState st{};
  st.target = INT64_MAX;
  uint64_t x = (uint64_t)r.p.seed_time_inc;
  for (int k = 0; k < 4; k++) st.rng[k] = splitmix_next(x);
  st.status = AKMC_REP_RUNNING;
  for (int q = 0; q < 8; q++) st.sec_count[q] = r.sec_count[q];
  return st;
}

uint32_t key_of(const akmc::BoxDims& d, int64_t i) { return akmc::key_of_index(d, i); }
int sector_of(const akmc::BoxDims& d, uint32_t k) { return akmc::key_sector(d, k); }

void count_sectors(HostRep* r, const std::vector<uint32_t>& keys) {
  for (int q = 0; q < 8; q++) r->sec_count[q] = 0;
  for (uint32_t k : keys) r->sec_count[sector_of(r->d, k)]++;
}

std::vector<uint32_t> vacancy_keys(const akmc::BoxDims& d, const uint8_t* sites, int64_t n) {
  std::vector<uint32_t> keys;
  for (int64_t i = 0; i < n; i++)
    if (sites[i] & akmc::BIT_V) keys.push_back(key_of(d, i));
  return keys;
}

// Device initial occupancy (AKMC_INIT_DEVICE), then the vacancy list,
// enqueued on `st` without waiting: the host data it copies stays in the
// replica (h_idx, h_keys, h_cum, h_flag) until the caller synchronizes `st`
// and calls release_setup().  Also uploads the FP32 rate table.
int init_device_async(HostRep* r, hipStream_t st) {
  int nt = 0;
  uint64_t total = 0;
  for (int i = 0; i < r->p.n_alloy; i++) {
    const int t = r->p.alloy_type[i];
    const int64_t q = r->p.alloy_ratio[i];
    if (q < 0) return fail(AKMC_ERR_INVALID, "negative alloy ratio");
    if (q == 0) continue;
    uint8_t f;
    switch (t) {
      case AKMC_TYPE_FE: f = 0; break;
      case AKMC_TYPE_CU: f = akmc::BIT_CU; break;
      case AKMC_TYPE_MN: f = akmc::BIT_MN; break;
      case AKMC_TYPE_NI: f = akmc::BIT_NI; break;
      default: return fail(AKMC_ERR_INVALID, "alloy type 0x%x is not modelled by the vacancy-hop path", t);
    }
    total += (uint64_t)q;
    if (total > 0xFFFFFFFFull) return fail(AKMC_ERR_INVALID, "alloy ratios too large");
    r->h_cum[nt] = (uint32_t)total;
    r->h_flag[nt] = f;
    nt++;
  }
  if (nt == 0) return fail(AKMC_ERR_INVALID, "empty alloy");
  HIP_TRY(hipMemcpyAsync(r->dcum, r->h_cum, sizeof(uint32_t) * nt, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(r->dflag, r->h_flag, nt, hipMemcpyHostToDevice, st));
  if (akmc_batch_init_sites_launch(r->sites, r->d, (uint64_t)r->p.seed_create_types, r->dcum, r->dflag, nt, st))
    return fail(AKMC_ERR_HIP, "init_sites launch failed");
  // Vacancy sites: va_count uniform draws (repeats collapse, as in the
  // reference's hash insertions), in site order.
  std::vector<int64_t>& idx = r->h_idx;
  idx.resize((size_t)r->p.va_count);
  uint64_t x = (uint64_t)r->p.seed_create_vacancy ^ 0x5A5A5A5A5A5A5A5Aull;
  for (auto& v : idx) v = (int64_t)(((unsigned __int128)splitmix_next(x) * (uint64_t)r->nsites) >> 64);
  std::sort(idx.begin(), idx.end());
  idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
  std::vector<uint32_t>& keys = r->h_keys;
  keys.resize(idx.size());
  for (size_t i = 0; i < idx.size(); i++) keys[i] = key_of(r->d, idx[i]);
  if (!idx.empty()) {
    HIP_TRY(hipMemcpyAsync(r->didx, idx.data(), idx.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    if (akmc_batch_place_vacancies_launch(r->sites, r->d, r->didx, (int64_t)idx.size(), st))
      return fail(AKMC_ERR_HIP, "place_vacancies launch failed");
    HIP_TRY(hipMemcpyAsync(r->vac, keys.data(), keys.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  }
  if (!r->tab.empty())
    HIP_TRY(hipMemcpyAsync(r->dtab, r->tab.data(), r->tab.size() * sizeof(float), hipMemcpyHostToDevice, st));
  r->nvac = (int32_t)keys.size();
  count_sectors(r, keys);
  return AKMC_OK;
}

// A replica's sites between the logical order x + X (y + Y z) and the
// device's brick layout (akmc_batch.h), eight x-consecutive bytes at a time.
void to_bricks(const akmc::BoxDims& d, const uint8_t* logical, uint8_t* bricks) {
  const akmc::BrickDims b = akmc::brick_dims(d);
  memset(bricks, 0, (size_t)akmc::brick_bytes(b));
  for (int z = 0; z < d.Z; z++)
    for (int y = 0; y < d.Y; y++) {
      const uint8_t* row = logical + (int64_t)d.X * (y + (int64_t)d.Y * z);
      for (int x = 0; x < d.X; x += 8)
        memcpy(bricks + akmc::brick_index(b.nbx, b.nby, x, y, z), row + x, (size_t)std::min(8, d.X - x));
    }
}
void from_bricks(const akmc::BoxDims& d, const uint8_t* bricks, uint8_t* logical) {
  const akmc::BrickDims b = akmc::brick_dims(d);
  for (int z = 0; z < d.Z; z++)
    for (int y = 0; y < d.Y; y++) {
      uint8_t* row = logical + (int64_t)d.X * (y + (int64_t)d.Y * z);
      for (int x = 0; x < d.X; x += 8)
        memcpy(row + x, bricks + akmc::brick_index(b.nbx, b.nby, x, y, z), (size_t)std::min(8, d.X - x));
    }
}

// Copies a replica's sites to the host in logical order (synchronous).
int download_sites(const HostRep& r, uint8_t* out) {
  std::vector<uint8_t> tmp((size_t)r.sbytes);
  HIP_TRY(hipMemcpy(tmp.data(), r.sites, (size_t)r.sbytes, hipMemcpyDeviceToHost));
  from_bricks(r.d, tmp.data(), out);
  return AKMC_OK;
}

int upload_rep(HostRep* r, const uint8_t* sites, hipStream_t st) {
  const std::vector<uint32_t> keys = vacancy_keys(r->d, sites, r->nsites);
  if ((int64_t)keys.size() > r->cap)
    return fail(AKMC_ERR_CAPACITY, "lattice holds %zu vacancies, replica capacity %d", keys.size(), r->cap);
  std::vector<uint8_t> bricks((size_t)r->sbytes);
  to_bricks(r->d, sites, bricks.data());
  HIP_TRY(hipMemcpyAsync(r->sites, bricks.data(), (size_t)r->sbytes, hipMemcpyHostToDevice, st));
  if (!keys.empty())
    HIP_TRY(hipMemcpyAsync(r->vac, keys.data(), keys.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  r->nvac = (int32_t)keys.size();
  count_sectors(r, keys);
  HIP_TRY(hipStreamSynchronize(st));
  return AKMC_OK;
}

// ---------------------------------------------------------------- launches
struct Tuning {
  int32_t vcap = 256;   // LDS class capacity (-1: no LDS class)
  int32_t il = 1;       // interleave width: replica waves per SIMD (1, 2, 4)
  int32_t waves = 2;    // waves per workgroup (4 * il when il > 1)
  int32_t order = 1;
  int64_t slice = 1024;
};

// Launch configuration and work lists of a set of replica slots.
struct Runner {
  int device = 0;
  bool fp32 = true;
  int flags = 0;
  uint64_t budget_ticks = 0;
  hipStream_t s0 = nullptr, s1 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, e0 = nullptr, e1 = nullptr;
  akmc::ReplicaDesc* d_desc = nullptr;
  akmc::ReplicaState* d_state = nullptr;
  int32_t slots = 0;
  int32_t* d_items = nullptr;
  uint64_t* d_queue = nullptr;
  int32_t n_lds = 0, n_glob = 0;
  int grid_lds = 0, grid_glob = 0;
  int region_lds = 0, region_glob = 0;
  int n_cu = 0;
  Tuning tun;
  bool use_graph = false;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;

  int init(int dev, bool f32, int32_t nslots) {
    device = dev;
    fp32 = f32;
    slots = nslots;
    HIP_TRY(hipSetDevice(dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    n_cu = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&s0, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipMalloc(&d_desc, sizeof(akmc::ReplicaDesc) * std::max(1, nslots)));
    HIP_TRY(hipMalloc(&d_state, sizeof(akmc::ReplicaState) * std::max(1, nslots)));
    HIP_TRY(hipMalloc(&d_items, sizeof(int32_t) * std::max(1, nslots)));
    HIP_TRY(hipMalloc(&d_queue, 4 * sizeof(uint64_t)));
    HIP_TRY(hipMemset(d_queue, 0, 4 * sizeof(uint64_t)));
    return akmc_host::upload_stencil(device);
  }

  void destroy() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (d_desc) (void)hipFree(d_desc);
    if (d_state) (void)hipFree(d_state);
    if (d_items) (void)hipFree(d_items);
    if (d_queue) (void)hipFree(d_queue);
    for (hipEvent_t e : {fork, join, e0, e1})
      if (e) (void)hipEventDestroy(e);
    if (s0) (void)hipStreamDestroy(s0);
    if (s1) (void)hipStreamDestroy(s1);
  }

  // Splits the live slots into the LDS-resident and global kinds and orders
  // each work list.
  int set_items(const std::vector<int32_t>& live, const std::vector<int32_t>& cap_of_slot) {
    std::vector<int32_t> a, b;
    for (int32_t s : live) (tun.vcap > 0 && cap_of_slot[s] <= tun.vcap ? a : b).push_back(s);
    if (tun.order == 1) {
      auto big_first = [&](int32_t x, int32_t y) {
        return cap_of_slot[x] != cap_of_slot[y] ? cap_of_slot[x] > cap_of_slot[y] : x < y;
      };
      std::stable_sort(a.begin(), a.end(), big_first);
      std::stable_sort(b.begin(), b.end(), big_first);
    }
    n_lds = (int32_t)a.size();
    n_glob = (int32_t)b.size();
    std::vector<int32_t> all(a);
    all.insert(all.end(), b.begin(), b.end());
    if (!all.empty())
      HIP_TRY(hipMemcpy(d_items, all.data(), all.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    region_lds = tun.vcap > 0 ? akmc_batch_region_bytes(fp32, tun.vcap) : 0;
    region_glob = akmc_batch_region_bytes(fp32, 0);
    grid_lds = grid_glob = 0;
    if (n_lds) {
      int per_cu = 0;
      if (akmc_batch_occupancy(fp32, 1, tun.il, tun.waves, region_lds, &per_cu) || per_cu <= 0)
        return fail(AKMC_ERR_CAPACITY, "LDS class does not fit (vcap %d, %d waves)", tun.vcap, tun.waves);
      grid_lds = std::min(per_cu * n_cu, (n_lds + tun.waves - 1) / tun.waves);
    }
    if (n_glob) {
      int per_cu = 0;
      if (akmc_batch_occupancy(fp32, 0, tun.il, tun.waves, region_glob, &per_cu) || per_cu <= 0)
        return fail(AKMC_ERR_CAPACITY, "global class does not fit");
      grid_glob = std::min(per_cu * n_cu, (n_glob + tun.waves - 1) / tun.waves);
    }
    graph_valid = false;
    return AKMC_OK;
  }

  akmc::BatchArgs args(bool lds) const {
    const auto& Q = qtables();
    akmc::BatchArgs a{};
    a.desc = d_desc;
    a.state = d_state;
    a.items = lds ? d_items : d_items + n_lds;
    a.n_items = lds ? n_lds : n_glob;
    a.flags = flags;
    a.queue = lds ? d_queue : d_queue + 2;
    a.slice_events = tun.slice;
    a.budget_ticks = budget_ticks;
    a.region_bytes = lds ? region_lds : region_glob;
    a.vcap = lds ? tun.vcap : 0;
    a.rates_sector = -1;
    for (int i = 0; i < 25; i++) { a.e1q[i] = Q.e1q[i]; a.e2q[i] = Q.e2q[i]; }
    for (int i = 0; i < 5; i++) a.e0q2[i] = Q.e0q2[i];
    return a;
  }

  // One slice: both kinds, the global kind on a forked stream.
  int enqueue() {
    const bool both = n_lds && n_glob;
    if (both) {
      HIP_TRY(hipEventRecord(fork, s0));
      HIP_TRY(hipStreamWaitEvent(s1, fork, 0));
    }
    if (n_glob) {
      hipStream_t s = both ? s1 : s0;
      const akmc::BatchArgs a = args(false);
      if (const int e = akmc_batch_reset_launch(d_queue + 2, s))
        return fail(AKMC_ERR_HIP, "reset launch failed: %s", hipGetErrorString((hipError_t)e));
      if (const int e = akmc_batch_launch(&a, fp32, 0, tun.il, tun.waves, grid_glob, s))
        return fail(AKMC_ERR_HIP, "launch failed: %s", hipGetErrorString((hipError_t)e));
    }
    if (n_lds) {
      const akmc::BatchArgs a = args(true);
      if (const int e = akmc_batch_reset_launch(d_queue, s0))
        return fail(AKMC_ERR_HIP, "reset launch failed: %s", hipGetErrorString((hipError_t)e));
      if (const int e = akmc_batch_launch(&a, fp32, 1, tun.il, tun.waves, grid_lds, s0))
        return fail(AKMC_ERR_HIP, "launch failed: %s", hipGetErrorString((hipError_t)e));
    }
    if (both) {
      HIP_TRY(hipEventRecord(join, s1));
      HIP_TRY(hipStreamWaitEvent(s0, join, 0));
    }
    return AKMC_OK;
  }

  int launch() {
    if (!n_lds && !n_glob) return AKMC_OK;
    if (!use_graph) return enqueue();
    if (!graph_valid) {
      if (exec) {
        HIP_TRY(hipGraphExecDestroy(exec));
        exec = nullptr;
      }
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue();
      const hipError_t ce = hipStreamEndCapture(s0, &g);
      if (rc) return rc;
      if (ce != hipSuccess) return fail(AKMC_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
      const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(AKMC_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
      graph_valid = true;
    }
    HIP_TRY(hipGraphLaunch(exec, s0));
    return AKMC_OK;
  }
};

void fill_rep_stats(const HostRep& r, const akmc::ReplicaState& st, akmc_replica_stats* s) {
  memset(s, 0, sizeof(*s));
  s->events = st.events;
  s->step = st.step;
  s->status = st.status;
  s->sector_pos = st.sect;
  for (int i = 0; i < 8; i++) s->evolution[i] = st.evo[i];
  s->sector_time = st.t;
  s->nvac = r.nvac;
  s->n_eval = st.n_eval;
  s->n_rebuild = st.n_rebuild;
  s->n_jump = st.n_jump;
  s->err_max = st.err_max;
  s->err_mean = st.err_n ? st.err_sum / (double)st.err_n : 0.0;
  s->err_n = st.err_n;
  s->err_tiny = st.err_tiny;
  s->n_skip = st.n_skip;
  s->busy_seconds = st.busy_ticks * 1e-8;  // 100 MHz
}

}  // namespace

// ================================================================ batch
struct akmc_batch {
  akmc_batch_options opt{};
  Runner R;
  std::vector<HostRep> reps;
  std::vector<int32_t> cap_of_slot;
  std::vector<int32_t> live;
  uint8_t* pool = nullptr;
  std::vector<akmc::ReplicaState> h_state;
  akmc_batch_tuning last_tune{};
  std::vector<akmc_batch_tuning> trials;  // every setting of the last akmc_batch_tune, with its rate
};

namespace {

int read_states(akmc_batch* b) {
  HIP_TRY(hipStreamSynchronize(b->R.s0));
  HIP_TRY(hipMemcpy(b->h_state.data(), b->R.d_state, b->h_state.size() * sizeof(akmc::ReplicaState),
                    hipMemcpyDeviceToHost));
  return AKMC_OK;
}
int write_states(akmc_batch* b) {
  HIP_TRY(hipMemcpy(b->R.d_state, b->h_state.data(), b->h_state.size() * sizeof(akmc::ReplicaState),
                    hipMemcpyHostToDevice));
  return AKMC_OK;
}

// Runs `launches` slices; returns the events of all replicas and device seconds.
int run_slices(akmc_batch* b, int64_t launches, int64_t* events, double* ksec) {
  int64_t before = 0, after = 0;
  for (const auto& s : b->h_state) before += s.events;
  HIP_TRY(hipEventRecord(b->R.e0, b->R.s0));
  for (int64_t k = 0; k < launches; k++) {
    const int rc = b->R.launch();
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(b->R.e1, b->R.s0));
  HIP_TRY(hipEventSynchronize(b->R.e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->R.e0, b->R.e1));
  const int rc = read_states(b);
  if (rc) return rc;
  for (const auto& s : b->h_state) after += s.events;
  *events = after - before;
  *ksec = ms * 1e-3;
  return AKMC_OK;
}

int set_targets(akmc_batch* b, int64_t events) {
  for (auto& s : b->h_state)
    if (s.status == AKMC_REP_RUNNING) s.target = events > INT64_MAX - s.events ? INT64_MAX : s.events + events;
  return write_states(b);
}

// Advances every running replica by exactly `events` events (fewer only when
// its time loop ends or it fails): launches ceil(events / slice) slices, then
// more while a replica is short (a slice's budget also counts sector visits).
int advance(akmc_batch* b, int64_t events, int64_t* total, double* ksec, int64_t* launches) {
  int rc = read_states(b);
  if (rc) return rc;
  rc = set_targets(b, events);
  if (rc) return rc;
  *total = 0;
  *ksec = 0;
  *launches = 0;
  int64_t need = events / b->R.tun.slice + (events % b->R.tun.slice != 0 ? 1 : 0);
  while (need > 0) {
    int64_t ev = 0;
    double ks = 0;
    rc = run_slices(b, need, &ev, &ks);
    if (rc) return rc;
    *total += ev;
    *ksec += ks;
    *launches += need;
    need = 0;
    for (const auto& s : b->h_state)
      if (s.status == AKMC_REP_RUNNING && s.events < s.target) need = 1;
  }
  // Replicas that reached this target continue in the next call.
  for (auto& s : b->h_state)
    if (s.status == AKMC_REP_TARGET_DONE) s.status = AKMC_REP_RUNNING;
  return write_states(b);
}

}  // namespace

extern "C" {

int akmc_batch_options_default(akmc_batch_options* o) {
  if (!o) return fail(AKMC_ERR_INVALID, "null options");
  memset(o, 0, sizeof(*o));
  o->precision = AKMC_PREC_FP32;
  o->init = AKMC_INIT_DEVICE;
  o->use_graph = 1;
  o->lds_vcap = 0;
  o->waves_per_group = 0;
  o->order = -1;
  o->slice_events = 0;
  return AKMC_OK;
}

int akmc_batch_create(const akmc_params* reps, int32_t n, const akmc_batch_options* o, int device,
                      akmc_batch** out) {
  if (!reps || !out || n <= 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  *out = nullptr;
  akmc_batch_options opt;
  if (o) opt = *o; else akmc_batch_options_default(&opt);
  if (opt.precision != AKMC_PREC_FP32 && opt.precision != AKMC_PREC_FP64)
    return fail(AKMC_ERR_INVALID, "bad precision");
  const bool fp32 = opt.precision == AKMC_PREC_FP32;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(AKMC_ERR_INVALID, "device %d out of range", device);
  HIP_TRY(hipSetDevice(device));
  std::unique_ptr<akmc_batch> b(new akmc_batch());
  b->opt = opt;
  b->reps.resize(n);
  int64_t total = 0;
  for (int i = 0; i < n; i++) {
    const int rc = setup_rep(&b->reps[i], reps[i], fp32);
    if (rc) return fail(rc, "replica %d: %s", i, akmc_host::g_err.c_str());
    total += b->reps[i].bytes;
  }
  int rc = b->R.init(device, fp32, n);
  if (rc) return rc;
  HIP_TRY(hipMalloc(&b->pool, (size_t)total));
  int64_t off = 0;
  for (auto& r : b->reps) {
    carve(&r, b->pool + off, fp32);
    off += r.bytes;
  }
  // Initial lattices.
  if (opt.init == AKMC_INIT_REFERENCE) {
    // creation::createRandom per replica on the host, in parallel.
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<uint8_t>> host(n);
    std::vector<int> rcs(n, 0);
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; t++)
      th.emplace_back([&] {
        for (int i = next++; i < n; i = next++) {
          host[i].resize(b->reps[i].nsites);
          rcs[i] = akmc_random_sites(&b->reps[i].p, host[i].data(), b->reps[i].nsites);
        }
      });
    for (auto& t : th) t.join();
    for (int i = 0; i < n; i++) {
      if (rcs[i]) return fail(rcs[i], "replica %d: creation failed", i);
      rc = upload_rep(&b->reps[i], host[i].data(), b->R.s0);
      if (rc) return rc;
      host[i].clear();
      host[i].shrink_to_fit();
    }
  } else if (opt.init == AKMC_INIT_DEVICE) {
    for (int i = 0; i < n; i++) {
      rc = init_device_async(&b->reps[i], b->R.s0);
      if (rc) return fail(rc, "replica %d: %s", i, akmc_host::g_err.c_str());
    }
    HIP_TRY(hipStreamSynchronize(b->R.s0));
    for (auto& r : b->reps) r.release_setup();
  } else {
    return fail(AKMC_ERR_INVALID, "bad init mode");
  }
  std::vector<akmc::ReplicaDesc> desc(n);
  b->h_state.resize(n);
  b->cap_of_slot.resize(n);
  for (int i = 0; i < n; i++) {
    if (fp32)
      HIP_TRY(hipMemcpy(b->reps[i].dtab, b->reps[i].tab.data(), b->reps[i].tab.size() * sizeof(float),
                        hipMemcpyHostToDevice));
    desc[i] = make_desc(b->reps[i]);
    b->h_state[i] = initial_state(b->reps[i]);
    b->cap_of_slot[i] = b->reps[i].cap;
    b->live.push_back(i);
  }
  HIP_TRY(hipMemcpy(b->R.d_desc, desc.data(), n * sizeof(akmc::ReplicaDesc), hipMemcpyHostToDevice));
  rc = write_states(b.get());
  if (rc) return rc;
  b->R.use_graph = opt.use_graph != 0;
  b->R.flags = opt.error_stats ? AKMC_BF_ERROR_STATS : 0;
  b->R.tun.vcap = opt.lds_vcap > 0 ? opt.lds_vcap : (opt.lds_vcap < 0 ? -1 : 256);
  b->R.tun.il = opt.interleave > 0 ? opt.interleave : 1;
  b->R.tun.waves = b->R.tun.il > 1 ? 4 * b->R.tun.il : (opt.waves_per_group > 0 ? opt.waves_per_group : 2);
  b->R.tun.order = opt.order >= 0 ? opt.order : 1;
  b->R.tun.slice = opt.slice_events > 0 ? opt.slice_events : 1024;
  if (b->R.tun.il != 1 && b->R.tun.il != 2 && b->R.tun.il != 4)
    return fail(AKMC_ERR_INVALID, "interleave must be 1, 2 or 4");
  if (b->R.tun.il == 1 && b->R.tun.waves != 1 && b->R.tun.waves != 2 && b->R.tun.waves != 4)
    return fail(AKMC_ERR_INVALID, "waves_per_group must be 1, 2 or 4");
  rc = b->R.set_items(b->live, b->cap_of_slot);
  if (rc) return rc;
  HIP_TRY(hipDeviceSynchronize());
  *out = b.release();
  return AKMC_OK;
}

int akmc_batch_run(akmc_batch* b, int64_t events, akmc_batch_stats* stats) {
  if (!b || events < 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(b->R.device));
  const auto t0 = std::chrono::steady_clock::now();
  int64_t total = 0, launches = 0;
  double ksec = 0;
  const int rc = advance(b, events, &total, &ksec, &launches);
  if (rc) return rc;
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->events = total;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats->kernel_seconds = ksec;
    stats->launches = launches;
    for (const auto& s : b->h_state) {
      if (s.status == AKMC_REP_RUNNING) stats->running++;
      if (s.status < 0) stats->failed++;
    }
  }
  return AKMC_OK;
}

int akmc_batch_run_for(akmc_batch* b, double seconds, akmc_batch_stats* stats) {
  if (!b || !(seconds > 0)) return fail(AKMC_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(b->R.device));
  const auto t0 = std::chrono::steady_clock::now();
  int rc = read_states(b);
  if (rc) return rc;
  for (auto& s : b->h_state)
    if (s.status == AKMC_REP_RUNNING) s.target = INT64_MAX;
  rc = write_states(b);
  if (rc) return rc;
  // One direct launch with a deadline instead of an event budget.
  const Tuning keep = b->R.tun;
  const bool graph = b->R.use_graph;
  b->R.tun.slice = (int64_t)1 << 40;
  b->R.budget_ticks = (uint64_t)(seconds * 100e6);  // s_memrealtime runs at 100 MHz
  b->R.use_graph = false;
  int64_t ev = 0;
  double ks = 0;
  rc = run_slices(b, 1, &ev, &ks);
  b->R.tun = keep;
  b->R.budget_ticks = 0;
  b->R.use_graph = graph;
  if (rc) return rc;
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->events = ev;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats->kernel_seconds = ks;
    stats->launches = 1;
    for (const auto& s : b->h_state) {
      if (s.status == AKMC_REP_RUNNING) stats->running++;
      if (s.status < 0) stats->failed++;
    }
  }
  return AKMC_OK;
}

int akmc_batch_get_tuning(akmc_batch* b, akmc_batch_tuning* t) {
  if (!b || !t) return fail(AKMC_ERR_INVALID, "bad arguments");
  memset(t, 0, sizeof(*t));
  t->lds_vcap = b->R.tun.vcap;
  t->waves_per_group = b->R.tun.waves;
  t->interleave = b->R.tun.il;
  t->order = b->R.tun.order;
  t->slice_events = b->R.tun.slice;
  t->events_per_second = b->last_tune.events_per_second;
  t->trials = b->last_tune.trials;
  return AKMC_OK;
}

int akmc_batch_tune(akmc_batch* b, double trial_seconds, akmc_batch_tuning* chosen) {
  if (!b || !(trial_seconds > 0)) return fail(AKMC_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(b->R.device));
  const Tuning keep = b->R.tun;
  int32_t maxcap = 1;
  for (int32_t c : b->cap_of_slot) maxcap = std::max(maxcap, c);
  std::vector<int32_t> vcaps = {-1};
  for (int v : {32, 64, 128, 256, 512, 1024}) {
    vcaps.push_back(v);
    if (v >= maxcap) break;  // larger capacities only add LDS
  }
  Tuning best = keep;
  double best_rate = -1;
  int trials = 0;
  b->trials.clear();
  // Interleave width il > 1: workgroups of 4 * il waves, il per SIMD.
  const std::pair<int32_t, int32_t> shapes[] = {{1, 1}, {1, 2}, {1, 4}, {2, 8}, {4, 16}};
  for (int32_t vc : vcaps)
    for (const auto& sh : shapes)
      for (int32_t ord : {0, 1}) {
        Tuning t = keep;
        t.vcap = vc;
        t.il = sh.first;
        t.waves = sh.second;
        t.order = ord;
        const int w = sh.second;
        if (vc > 0 && (int64_t)w * akmc_batch_region_bytes(b->R.fp32, vc) > 160 * 1024) continue;
        b->R.tun = t;
        if (b->R.set_items(b->live, b->cap_of_slot)) continue;  // does not fit: skip
        akmc_batch_stats st;
        const int rc = akmc_batch_run_for(b, trial_seconds, &st);
        if (rc) return rc;
        trials++;
        const double rate = st.kernel_seconds > 0 ? st.events / st.kernel_seconds : 0;
        akmc_batch_tuning rec{};
        rec.lds_vcap = t.vcap;
        rec.waves_per_group = t.waves;
        rec.interleave = t.il;
        rec.order = t.order;
        rec.slice_events = t.slice;
        rec.events_per_second = rate;
        rec.trials = 1;
        b->trials.push_back(rec);
        if (rate > best_rate) {
          best_rate = rate;
          best = t;
        }
      }
  b->R.tun = best;
  int rc = b->R.set_items(b->live, b->cap_of_slot);
  if (rc) return rc;
  // The deadline trials leave the replicas at different event counts: bring
  // every running replica up to the largest count.
  rc = read_states(b);
  if (rc) return rc;
  int64_t top = 0;
  for (const auto& s : b->h_state)
    if (s.status == AKMC_REP_RUNNING) top = std::max(top, s.events);
  for (auto& s : b->h_state)
    if (s.status == AKMC_REP_RUNNING) s.target = top;
  rc = write_states(b);
  if (rc) return rc;
  for (;;) {
    bool need = false;
    for (const auto& s : b->h_state)
      if (s.status == AKMC_REP_RUNNING && s.events < s.target) need = true;
    if (!need) break;
    int64_t ev = 0;
    double ks = 0;
    rc = run_slices(b, 1, &ev, &ks);
    if (rc) return rc;
  }
  for (auto& s : b->h_state)
    if (s.status == AKMC_REP_TARGET_DONE) s.status = AKMC_REP_RUNNING;
  rc = write_states(b);
  if (rc) return rc;
  b->last_tune.events_per_second = best_rate;
  b->last_tune.trials = trials;
  return chosen ? akmc_batch_get_tuning(b, chosen) : AKMC_OK;
}

int akmc_batch_tune_trials(akmc_batch* b, akmc_batch_tuning* out, int32_t cap, int32_t* n) {
  if (!b || !n || cap < 0 || (cap > 0 && !out)) return fail(AKMC_ERR_INVALID, "bad arguments");
  *n = (int32_t)b->trials.size();
  for (int32_t i = 0; i < std::min(cap, *n); i++) out[i] = b->trials[i];
  return AKMC_OK;
}

int32_t akmc_batch_size(const akmc_batch* b) { return b ? (int32_t)b->reps.size() : 0; }

int64_t akmc_batch_num_sites(const akmc_batch* b, int32_t i) {
  if (!b || i < 0 || i >= (int32_t)b->reps.size()) return -1;
  return b->reps[i].nsites;
}

int akmc_batch_replica_stats(akmc_batch* b, int32_t i, akmc_replica_stats* s) {
  if (!b || !s || i < 0 || i >= (int32_t)b->reps.size()) return fail(AKMC_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(b->R.device));
  HIP_TRY(hipStreamSynchronize(b->R.s0));
  akmc::ReplicaState st;
  HIP_TRY(hipMemcpy(&st, b->R.d_state + i, sizeof(st), hipMemcpyDeviceToHost));
  fill_rep_stats(b->reps[i], st, s);
  return AKMC_OK;
}

int akmc_batch_download(akmc_batch* b, int32_t i, uint8_t* sites, int64_t n) {
  if (!b || !sites || i < 0 || i >= (int32_t)b->reps.size()) return fail(AKMC_ERR_INVALID, "bad arguments");
  if (n != b->reps[i].nsites) return fail(AKMC_ERR_INVALID, "site count %lld != %lld", (long long)n,
                                          (long long)b->reps[i].nsites);
  HIP_TRY(hipSetDevice(b->R.device));
  HIP_TRY(hipStreamSynchronize(b->R.s0));
  return download_sites(b->reps[i], sites);
}

int akmc_batch_upload(akmc_batch* b, int32_t i, const uint8_t* sites, int64_t n) {
  if (!b || !sites || i < 0 || i >= (int32_t)b->reps.size()) return fail(AKMC_ERR_INVALID, "bad arguments");
  if (n != b->reps[i].nsites) return fail(AKMC_ERR_INVALID, "site count mismatch");
  HIP_TRY(hipSetDevice(b->R.device));
  int rc = read_states(b);
  if (rc) return rc;
  rc = upload_rep(&b->reps[i], sites, b->R.s0);
  if (rc) return rc;
  const akmc::ReplicaDesc D = make_desc(b->reps[i]);
  HIP_TRY(hipMemcpy(b->R.d_desc + i, &D, sizeof(D), hipMemcpyHostToDevice));
  b->h_state[i].in_sector = 0;
  for (int q = 0; q < 8; q++) b->h_state[i].sec_count[q] = b->reps[i].sec_count[q];
  return write_states(b);
}

int akmc_batch_sector_rates(akmc_batch* b, int32_t i, int sector, int64_t cap, int64_t* sites, double* rates,
                            double* total, int64_t* n) {
  if (!b || i < 0 || i >= (int32_t)b->reps.size() || sector < 0 || sector > 7 || !n)
    return fail(AKMC_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(b->R.device));
  HIP_TRY(hipStreamSynchronize(b->R.s0));
  const HostRep& r = b->reps[i];
  const bool fp32 = b->R.fp32;
  const akmc::WsLayout L = akmc::ws_layout(r.cap, fp32 ? 4 : 8);
  uint8_t* ws = nullptr;
  akmc::ReplicaDesc* dd = nullptr;
  akmc::ReplicaState* ds = nullptr;
  int32_t* di = nullptr;
  uint64_t* dq = nullptr;
  int rc = AKMC_OK;
  auto cleanup = [&] {
    if (ws) (void)hipFree(ws);
    if (dd) (void)hipFree(dd);
    if (ds) (void)hipFree(ds);
    if (di) (void)hipFree(di);
    if (dq) (void)hipFree(dq);
  };
  do {
    if (hipMalloc(&ws, L.total) || hipMalloc(&dd, sizeof(*dd)) || hipMalloc(&ds, sizeof(*ds)) ||
        hipMalloc(&di, sizeof(*di)) || hipMalloc(&dq, 4 * sizeof(uint64_t))) {
      rc = fail(AKMC_ERR_HIP, "scratch allocation failed");
      break;
    }
    akmc::ReplicaDesc D = make_desc(r);
    D.ws = ws;
    akmc::ReplicaState st = b->h_state[i];
    const int32_t zero = 0;
    if (hipMemcpy(dd, &D, sizeof(D), hipMemcpyHostToDevice) || hipMemcpy(ds, &st, sizeof(st), hipMemcpyHostToDevice) ||
        hipMemcpy(di, &zero, sizeof(zero), hipMemcpyHostToDevice) || hipMemset(dq, 0, 4 * sizeof(uint64_t))) {
      rc = fail(AKMC_ERR_HIP, "scratch upload failed");
      break;
    }
    akmc::BatchArgs a = b->R.args(false);
    a.desc = dd;
    a.state = ds;
    a.items = di;
    a.n_items = 1;
    a.queue = dq;
    a.flags = AKMC_BF_RATES_ONLY;
    a.rates_sector = sector;
    if (akmc_batch_launch(&a, fp32, 0, 1, 1, 1, b->R.s0)) {
      rc = fail(AKMC_ERR_HIP, "rates launch failed");
      break;
    }
    if (hipStreamSynchronize(b->R.s0) || hipMemcpy(&st, ds, sizeof(st), hipMemcpyDeviceToHost)) {
      rc = fail(AKMC_ERR_HIP, "rates launch did not complete");
      break;
    }
    if (st.nv < 0) {
      rc = fail(AKMC_ERR_CONFIG, "an FP32 rate left the table range");
      break;
    }
    *n = st.nv;
    const int64_t m = std::min<int64_t>(cap, st.nv);
    std::vector<uint8_t> host(L.total);
    if (hipMemcpy(host.data(), ws, L.total, hipMemcpyDeviceToHost)) {
      rc = fail(AKMC_ERR_HIP, "workspace download failed");
      break;
    }
    const uint32_t* keys = reinterpret_cast<const uint32_t*>(host.data() + L.key);
    const double* vt = reinterpret_cast<const double*>(host.data() + L.vt);
    double sum = 0;
    for (int64_t v = 0; v < st.nv; v++) sum += vt[v];
    if (total) *total = sum;
    for (int64_t v = 0; v < m; v++) {
      const uint32_t k = keys[v];
      if (sites)
        sites[v] = (int64_t)(k & 0x7FF) + (int64_t)r.d.X * ((int64_t)((k >> 11) & 0x7FF) + (int64_t)r.d.Y * (k >> 22));
      if (rates)
        for (int bb = 0; bb < 8; bb++)
          rates[8 * v + bb] = fp32 ? (double)reinterpret_cast<const float*>(host.data() + L.rate)[8 * v + bb]
                                   : reinterpret_cast<const double*>(host.data() + L.rate)[8 * v + bb];
    }
  } while (0);
  cleanup();
  return rc;
}

int akmc_fp32_rate_table(double temperature, double nu, float* tab, int32_t cap, int32_t* nhi, int32_t* m_min,
                         double* q, float* e1q, float* e2q, float* e0q2) {
  if (!tab || !nhi || !m_min || !(temperature > 0) || !(nu > 0)) return fail(AKMC_ERR_INVALID, "bad arguments");
  std::vector<float> t;
  const int rc = make_exp_table(AKMC_KB * temperature, nu, &t, nhi, m_min);
  if (rc) return rc;
  if ((int64_t)t.size() > cap) return fail(AKMC_ERR_INVALID, "table needs %zu floats", t.size());
  memcpy(tab, t.data(), t.size() * sizeof(float));
  const QTables& Q = qtables();
  if (q) *q = Q.q;
  if (e1q) memcpy(e1q, Q.e1q, sizeof(Q.e1q));
  if (e2q) memcpy(e2q, Q.e2q, sizeof(Q.e2q));
  if (e0q2) memcpy(e0q2, Q.e0q2, sizeof(Q.e0q2));
  return AKMC_OK;
}

void akmc_batch_destroy(akmc_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->R.device);
  (void)hipDeviceSynchronize();
  b->R.destroy();
  if (b->pool) (void)hipFree(b->pool);
  delete b;
}

}  // extern "C"

// ================================================================ server
namespace {

using Clock = std::chrono::steady_clock;

// Device memory for jobs: large chunks from hipMalloc, carved first-fit in
// 256-byte granules and coalesced on release, so that admitting or retiring
// a job calls neither hipMalloc nor hipFree once the chunks exist.
struct Arena {
  struct Chunk {
    uint8_t* base = nullptr;
    int64_t size = 0;
    std::map<int64_t, int64_t> free;  // offset -> bytes
  };
  std::vector<Chunk> chunks;
  int64_t chunk_bytes = (int64_t)4 << 30;

  int alloc(int64_t n, uint8_t** out) {
    n = a256(n);
    for (auto& c : chunks)
      for (auto it = c.free.begin(); it != c.free.end(); ++it)
        if (it->second >= n) {
          const int64_t off = it->first, sz = it->second;
          c.free.erase(it);
          if (sz > n) c.free[off + n] = sz - n;
          *out = c.base + off;
          return AKMC_OK;
        }
    Chunk c;
    c.size = std::max(chunk_bytes, n);
    HIP_TRY(hipMalloc(&c.base, (size_t)c.size));
    if (c.size > n) c.free[n] = c.size - n;
    *out = c.base;
    chunks.push_back(std::move(c));
    return AKMC_OK;
  }

  void release(uint8_t* p, int64_t n) {
    n = a256(n);
    for (auto& c : chunks) {
      if (p < c.base || p >= c.base + c.size) continue;
      auto it = c.free.emplace(p - c.base, n).first;
      auto nx = std::next(it);
      if (nx != c.free.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        c.free.erase(nx);
      }
      if (it != c.free.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
          pv->second += it->second;
          c.free.erase(it);
        }
      }
      return;
    }
  }

  void destroy() {
    for (auto& c : chunks) (void)hipFree(c.base);
    chunks.clear();
  }
};

struct Job {
  int64_t id = 0;
  akmc_params p{};
  int64_t target = 0;
  HostRep rep;
  int32_t status = AKMC_JOB_PENDING;
  int32_t rep_status = AKMC_REP_RUNNING;
  int32_t slot = -1;
  int64_t events = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  std::vector<uint8_t> result;
  akmc::ReplicaDesc h_desc{};   // sources of the admission copies
  akmc::ReplicaState h_state{};
  int64_t bytes = 0;            // HBM bytes while live
  // A job with self-interstitial dumbbells runs on the vacancy + dumbbell
  // engine (akmc_mix_pool.h): its own slot space.
  bool mix = false;
  int64_t n_itl = 0;
  int32_t itl_atom = 1;
  double ef110 = 3.53;
  std::vector<uint16_t> result_types;
  double ev_per_tick = 0;       // events per 100 MHz tick of its wavefront, from its last launch
  int64_t ticks = 0;            // its wavefront's device time so far
};

}  // namespace

struct akmc_server {
  akmc_server_options opt{};
  Runner R;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job*> pending;
  std::map<int64_t, std::unique_ptr<Job>> jobs;
  std::vector<Job*> slot_job;
  std::vector<int32_t> free_slots;
  std::vector<int32_t> cap_of_slot;
  std::vector<akmc::ReplicaState> h_state;
  Arena arena;
  int64_t next_id = 1;
  int64_t bytes_live = 0;
  akmc_server_stats st{};
  std::atomic<bool> stop{false};
  std::thread worker;
  hipStream_t init_stream = nullptr;  // job set-up, overlapping the running slice
  // Jobs with dumbbells: the vacancy + dumbbell pool, its stream and slots.
  akmc_mixpool::Pool* mix = nullptr;
  hipStream_t mix_stream = nullptr;
  std::vector<Job*> mix_job;
  std::vector<int32_t> mix_free;
  // The pool's launch runs beside the replica slices and is collected once
  // it has ended (mix_done), so a slow lattice never holds the others back.
  hipEvent_t mix_done = nullptr;
  bool mix_inflight = false;
  std::vector<int32_t> mix_launched;
  double t_mix_launch = 0;
  bool items_dirty = false;
  int error = AKMC_OK;
  std::string error_msg;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Admits pending jobs in arrival order while the budget and slots allow (at
// most 256 per step).  Their set-up is enqueued on the set-up stream and
// waited for once.
int admit(akmc_server* s, bool* changed) {
  std::vector<Job*> batch;
  while (batch.size() < 256) {
    Job* j = nullptr;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->pending.empty()) break;
      Job* h = s->pending.front();
      if ((h->mix ? s->mix_free : s->free_slots).empty()) break;
      if (s->bytes_live + h->bytes > s->opt.hbm_budget) break;
      s->pending.pop_front();
      j = h;
      std::vector<int32_t>& fl = j->mix ? s->mix_free : s->free_slots;
      j->slot = fl.back();
      fl.pop_back();
      s->bytes_live += j->bytes;
      s->st.bytes_peak = std::max(s->st.bytes_peak, s->bytes_live);
      s->st.live++;  // counted from here, so that the server never looks idle while it sets the job up
    }
    if (j->mix) {
      const int rc = akmc_mixpool::pool_add_generated(s->mix, j->slot, j->p, j->n_itl, j->itl_atom, j->ef110,
                                                      s->init_stream);
      if (rc) return rc;
      batch.push_back(j);
      continue;
    }
    uint8_t* mem = nullptr;
    int rc = s->arena.alloc(j->rep.bytes, &mem);
    if (rc) return rc;
    carve(&j->rep, mem, s->R.fp32);
    rc = init_device_async(&j->rep, s->init_stream);
    if (rc) return rc;
    j->h_desc = make_desc(j->rep);
    j->h_state = initial_state(j->rep);
    j->h_state.target = j->target;
    HIP_TRY(hipMemcpyAsync(s->R.d_desc + j->slot, &j->h_desc, sizeof(j->h_desc), hipMemcpyHostToDevice,
                           s->init_stream));
    HIP_TRY(hipMemcpyAsync(s->R.d_state + j->slot, &j->h_state, sizeof(j->h_state), hipMemcpyHostToDevice,
                           s->init_stream));
    batch.push_back(j);
  }
  if (batch.empty()) return AKMC_OK;
  HIP_TRY(hipStreamSynchronize(s->init_stream));
  const double t = s->now();
  std::lock_guard<std::mutex> lk(s->mu);
  for (Job* j : batch) {
    if (j->mix) {
      akmc_mixpool::pool_setup_done(s->mix, j->slot, s->init_stream);
      s->mix_job[j->slot] = j;
      j->status = AKMC_JOB_RUNNING;
      j->t_admit = t;
      s->st.admitted++;
      continue;
    }
    j->rep.release_setup();
    s->cap_of_slot[j->slot] = j->rep.cap;
    s->slot_job[j->slot] = j;
    j->status = AKMC_JOB_RUNNING;
    j->t_admit = t;
    s->st.admitted++;
  }
  s->st.bytes_live = s->bytes_live;
  *changed = true;
  return AKMC_OK;
}

int retire(akmc_server* s, Job* j, int32_t rep_status) {
  if (j->mix) {
    if (s->opt.keep_results) {
      j->result_types.resize((size_t)akmc_mixpool::pool_sites(s->mix, j->slot));
      const int rc = akmc_mixpool::pool_download(s->mix, j->slot, j->result_types.data(),
                                                 (int64_t)j->result_types.size());
      if (rc) return rc;
    }
    const int rc = akmc_mixpool::pool_remove(s->mix, j->slot, s->mix_stream);
    if (rc) return rc;
  } else {
    if (s->opt.keep_results) {
      j->result.resize(j->rep.nsites);
      const int rc = download_sites(j->rep, j->result.data());
      if (rc) return rc;
    }
    s->arena.release(j->rep.mem, j->rep.bytes);
    j->rep.mem = nullptr;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  j->rep_status = rep_status;
  j->status = rep_status < 0 ? AKMC_JOB_FAILED : AKMC_JOB_DONE;
  j->t_done = s->now();
  if (j->mix) {
    s->mix_job[j->slot] = nullptr;
    s->mix_free.push_back(j->slot);
  } else {
    s->slot_job[j->slot] = nullptr;
    s->free_slots.push_back(j->slot);
  }
  j->slot = -1;
  s->bytes_live -= j->bytes;
  s->st.bytes_live = s->bytes_live;
  s->st.live--;
  if (rep_status < 0) s->st.failed++; else s->st.done++;
  s->cv.notify_all();
  return AKMC_OK;
}

// Launches the pool over the live dumbbell jobs.  Each lattice runs up to
// the events that its measured rate fits into a few replica slices (at least
// 16, at most slice_events, never past its target), so that one launch takes
// about that long whatever the lattices' sizes.
int mix_launch(akmc_server* s) {
  std::vector<int32_t> mlive;
  std::vector<int64_t> mlim;
  const double ticks = std::max(1e5, 5e8 * s->opt.slice_seconds);  // 100 MHz ticks
  const int64_t cap = std::max<int64_t>(16, s->opt.slice_events);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (int32_t k = 0; k < (int32_t)s->mix_job.size(); k++)
      if (Job* j = s->mix_job[k]) {
        const int64_t want = j->ev_per_tick > 0 ? (int64_t)(j->ev_per_tick * ticks) : 256;
        mlive.push_back(k);
        mlim.push_back(std::min(j->target, j->events + std::min(cap, std::max<int64_t>(16, want))));
      }
  }
  if (mlive.empty()) return AKMC_OK;
  int rc = akmc_mixpool::pool_launch(s->mix, mlive, mlim, s->mix_stream);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(s->mix_done, s->mix_stream));
  s->mix_inflight = true;
  s->mix_launched = mlive;
  s->t_mix_launch = s->now();
  return AKMC_OK;
}

// Takes the results of the ended pool launch: progress, rates, retirement.
int mix_collect(akmc_server* s) {
  s->mix_inflight = false;
  int64_t mev = 0;
  for (int32_t k : s->mix_launched) {
    Job* j = s->mix_job[k];
    const akmc::mix::MixState& st = akmc_mixpool::pool_state(s->mix, k);
    const int64_t dt = st.busy_ticks - j->ticks;
    if (dt > 0 && st.events > j->events) j->ev_per_tick = (double)(st.events - j->events) / (double)dt;
    j->ticks = st.busy_ticks;
    mev += st.events - j->events;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      j->events = st.events;
    }
    int32_t rs = AKMC_REP_RUNNING;
    if (st.status == AKMC_RUN_TIME_DONE) rs = AKMC_REP_TIME_DONE;
    else if (st.events >= j->target) rs = AKMC_REP_TARGET_DONE;
    else if (st.status < 0) rs = st.status;
    if (rs != AKMC_REP_RUNNING) {
      const int rc = retire(s, j, rs);
      if (rc) return rc;
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.launches++;
  s->st.events += mev;
  return AKMC_OK;
}

// One step of the request loop: launch a slice over the live jobs, admit
// pending jobs while it runs (their set-up overlaps the device work), then
// collect the slice and retire the jobs that finished.  The dumbbell jobs'
// pool launch runs on its own stream beside the slices: collected when it
// has ended, then launched again.
int serve_step(akmc_server* s) {
  std::vector<int32_t> live;
  int32_t hi = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (int32_t k = 0; k < (int32_t)s->slot_job.size(); k++)
      if (s->slot_job[k]) { live.push_back(k); hi = k + 1; }
  }
  int rc;
  if (s->mix_inflight) {
    const hipError_t q = hipEventQuery(s->mix_done);
    if (q == hipSuccess) {
      rc = mix_collect(s);
      if (rc) return rc;
    } else if (q != hipErrorNotReady) {
      return fail(AKMC_ERR_HIP, "dumbbell jobs' launch: %s", hipGetErrorString(q));
    }
  }
  if (!s->mix_inflight) {
    rc = mix_launch(s);
    if (rc) return rc;
  }
  if (s->items_dirty) {
    rc = s->R.set_items(live, s->cap_of_slot);
    if (rc) return rc;
    s->items_dirty = false;
  }
  const double t_a = s->now();
  const bool launched = !live.empty();
  if (launched) {
    rc = s->R.launch();
    if (rc) return rc;
  }
  bool changed = false;
  rc = admit(s, &changed);
  if (rc) return rc;
  if (changed) s->items_dirty = true;
  if (!launched) {
    // Only dumbbell jobs run: wait for their launch instead of spinning.
    if (s->mix_inflight) {
      HIP_TRY(hipEventSynchronize(s->mix_done));
      const double t_m = s->t_mix_launch;
      rc = mix_collect(s);
      if (rc) return rc;
      std::lock_guard<std::mutex> lk(s->mu);
      s->st.busy_seconds += s->now() - std::max(t_m, t_a);
    }
    return AKMC_OK;
  }
  HIP_TRY(hipStreamSynchronize(s->R.s0));
  const double t_b = s->now();
  HIP_TRY(hipMemcpy(s->h_state.data(), s->R.d_state, hi * sizeof(akmc::ReplicaState), hipMemcpyDeviceToHost));
  int64_t ev = 0;
  for (int32_t k : live) {
    Job* j = s->slot_job[k];
    const akmc::ReplicaState& st = s->h_state[k];
    ev += st.events - j->events;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      j->events = st.events;
    }
    if (st.status != AKMC_REP_RUNNING) {
      rc = retire(s, j, st.status);
      if (rc) return rc;
      s->items_dirty = true;
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->st.launches++;
  s->st.events += ev;
  s->st.busy_seconds += t_b - t_a;
  return AKMC_OK;
}

void serve_loop(akmc_server* s) {
  (void)hipSetDevice(s->R.device);
  while (!s->stop.load()) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->pending.empty() && s->st.live == 0) {
        s->cv.wait_for(lk, std::chrono::milliseconds(20));
        continue;
      }
    }
    const int rc = serve_step(s);
    if (rc) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->error = rc;
      s->error_msg = akmc_host::g_err;
      s->cv.notify_all();
      return;
    }
  }
}

}  // namespace

extern "C" {

int akmc_server_options_default(akmc_server_options* o) {
  if (!o) return fail(AKMC_ERR_INVALID, "null options");
  memset(o, 0, sizeof(*o));
  o->hbm_budget = (int64_t)250e9;
  o->max_live = 4096;
  o->precision = AKMC_PREC_FP32;
  o->lds_vcap = 256;
  o->waves_per_group = 2;
  o->slice_events = 4096;
  o->slice_seconds = 2e-3;
  o->keep_results = 0;
  return AKMC_OK;
}

int akmc_server_create(int device, const akmc_server_options* o, akmc_server** out) {
  if (!out) return fail(AKMC_ERR_INVALID, "null argument");
  *out = nullptr;
  akmc_server_options opt;
  if (o) opt = *o; else akmc_server_options_default(&opt);
  if (opt.max_live <= 0 || opt.hbm_budget <= 0) return fail(AKMC_ERR_INVALID, "bad server options");
  if (opt.precision != AKMC_PREC_FP32 && opt.precision != AKMC_PREC_FP64)
    return fail(AKMC_ERR_INVALID, "bad precision");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(AKMC_ERR_INVALID, "device %d out of range", device);
  HIP_TRY(hipSetDevice(device));
  std::unique_ptr<akmc_server> s(new akmc_server());
  s->opt = opt;
  s->t0 = Clock::now();
  int rc = s->R.init(device, opt.precision == AKMC_PREC_FP32, opt.max_live);
  if (rc) return rc;
  s->R.use_graph = false;  // the live set changes between launches
  HIP_TRY(hipStreamCreateWithFlags(&s->init_stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&s->mix_stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&s->mix_done, hipEventDisableTiming));
  rc = akmc_mixpool::pool_create(device, opt.max_live, &s->mix);
  if (rc) return rc;
  s->mix_job.assign(opt.max_live, nullptr);
  for (int32_t k = opt.max_live - 1; k >= 0; k--) s->mix_free.push_back(k);
  s->R.tun.vcap = opt.lds_vcap > 0 ? opt.lds_vcap : (opt.lds_vcap < 0 ? -1 : 256);
  s->R.tun.il = 1;
  s->R.tun.waves = opt.waves_per_group > 0 ? opt.waves_per_group : 2;
  s->R.tun.order = 1;
  s->R.tun.slice = opt.slice_events > 0 ? opt.slice_events : 4096;
  s->R.budget_ticks = opt.slice_seconds > 0 ? (uint64_t)(opt.slice_seconds * 100e6) : 0;  // 100 MHz clock
  s->slot_job.assign(opt.max_live, nullptr);
  s->cap_of_slot.assign(opt.max_live, 1);
  s->h_state.resize(opt.max_live);
  for (int32_t k = opt.max_live - 1; k >= 0; k--) s->free_slots.push_back(k);
  // Unused slots hold a stopped state.
  std::vector<akmc::ReplicaState> idle(opt.max_live);
  for (auto& x : idle) x.status = AKMC_REP_TARGET_DONE;
  HIP_TRY(hipMemcpy(s->R.d_state, idle.data(), idle.size() * sizeof(akmc::ReplicaState), hipMemcpyHostToDevice));
  akmc_server* raw = s.release();
  raw->worker = std::thread(serve_loop, raw);
  *out = raw;
  return AKMC_OK;
}

int akmc_server_submit(akmc_server* s, const akmc_params* p, int64_t target_events, int64_t* job) {
  if (!s || !p || target_events < 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  std::unique_ptr<Job> j(new Job());
  j->p = *p;
  j->target = target_events;
  int rc = setup_rep(&j->rep, *p, s->R.fp32);
  if (rc) return rc;
  j->bytes = j->rep.bytes;
  if (j->bytes > s->opt.hbm_budget)
    return fail(AKMC_ERR_BUDGET, "job needs %lld bytes, budget %lld", (long long)j->bytes,
                (long long)s->opt.hbm_budget);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
  j->id = s->next_id++;
  j->t_submit = s->now();
  if (job) *job = j->id;
  s->pending.push_back(j.get());
  s->jobs[j->id] = std::move(j);
  s->st.submitted++;
  s->cv.notify_all();
  return AKMC_OK;
}

int akmc_server_submit_itl(akmc_server* s, const akmc_params* p, int64_t n_itl, int32_t itl_atom, double ef110,
                            int64_t target_events, int64_t* job) {
  if (!s || !p || target_events < 0 || n_itl < 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  if (!(itl_atom == AKMC_TYPE_FE || itl_atom == AKMC_TYPE_CU || itl_atom == AKMC_TYPE_NI || itl_atom == AKMC_TYPE_MN))
    return fail(AKMC_ERR_INVALID, "dumbbell atom 0x%x", itl_atom);
  for (int d = 0; d < 3; d++)
    if (p->box[d] < 8) return fail(AKMC_ERR_INVALID, "box edge %lld < 8 cells", (long long)p->box[d]);
  if (2 * p->box[0] > 2048 || p->box[1] > 2047 || p->box[2] > 1023)
    return fail(AKMC_ERR_INVALID, "box exceeds 1024 x 2047 x 1023 cells");
  // Everything admission checks, checked here, so that a bad job is refused
  // at once instead of stopping the server when its turn comes.
  const int64_t nsites = 2 * p->box[0] * p->box[1] * p->box[2];
  if (p->va_count < 0 || p->va_count + n_itl > nsites / 2)
    return fail(AKMC_ERR_INVALID, "%lld vacancies and %lld dumbbells do not fit the box", (long long)p->va_count,
                (long long)n_itl);
  if (!(p->sector_time > 0) || !(p->temperature > 0) || !(p->attempt_freq > 0))
    return fail(AKMC_ERR_INVALID, "bad time, temperature or frequency");
  if (p->n_alloy < 0 || p->n_alloy > AKMC_MAX_ALLOY || !std::isfinite(ef110))
    return fail(AKMC_ERR_INVALID, "bad alloy count or dumbbell energy");
  {
    uint64_t total = 0;
    for (int i = 0; i < p->n_alloy; i++) {
      const int t = p->alloy_type[i];
      if (p->alloy_ratio[i] < 0 || !(t == AKMC_TYPE_FE || t == AKMC_TYPE_CU || t == AKMC_TYPE_NI || t == AKMC_TYPE_MN))
        return fail(AKMC_ERR_INVALID, "bad alloy entry %d", i);
      total += (uint64_t)p->alloy_ratio[i];
    }
    if (total == 0 || total > 0xFFFFFFFFull) return fail(AKMC_ERR_INVALID, "alloy ratios sum to %llu", (unsigned long long)total);
  }
  std::unique_ptr<Job> j(new Job());
  j->p = *p;
  j->target = target_events;
  j->mix = true;
  j->n_itl = n_itl;
  j->itl_atom = itl_atom;
  j->ef110 = ef110;
  j->bytes = akmc_mixpool::pool_bytes(*p, n_itl);
  if (j->bytes > s->opt.hbm_budget)
    return fail(AKMC_ERR_BUDGET, "job needs %lld bytes, budget %lld", (long long)j->bytes,
                (long long)s->opt.hbm_budget);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
  j->id = s->next_id++;
  j->t_submit = s->now();
  if (job) *job = j->id;
  s->pending.push_back(j.get());
  s->jobs[j->id] = std::move(j);
  s->st.submitted++;
  s->cv.notify_all();
  return AKMC_OK;
}

int akmc_server_result_typed(akmc_server* s, int64_t id, uint16_t* types, int64_t n) {
  if (!s || !types) return fail(AKMC_ERR_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(AKMC_ERR_INVALID, "no job %lld", (long long)id);
  const Job& j = *it->second;
  if (j.status != AKMC_JOB_DONE) return fail(AKMC_ERR_INVALID, "job %lld has not finished", (long long)id);
  if ((int64_t)j.result_types.size() != n) return fail(AKMC_ERR_INVALID, "no kept typed result of %lld sites", (long long)n);
  memcpy(types, j.result_types.data(), (size_t)n * 2);
  return AKMC_OK;
}

int akmc_server_job_info(akmc_server* s, int64_t id, akmc_job_info* info) {
  if (!s || !info) return fail(AKMC_ERR_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(AKMC_ERR_INVALID, "no job %lld", (long long)id);
  const Job& j = *it->second;
  memset(info, 0, sizeof(*info));
  info->status = j.status;
  info->rep_status = j.rep_status;
  info->events = j.events;
  info->bytes = j.bytes;
  info->t_submit = j.t_submit;
  info->t_admit = j.t_admit;
  info->t_done = j.t_done;
  return AKMC_OK;
}

int akmc_server_result(akmc_server* s, int64_t id, uint8_t* sites, int64_t n) {
  if (!s || !sites) return fail(AKMC_ERR_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(id);
  if (it == s->jobs.end()) return fail(AKMC_ERR_INVALID, "no job %lld", (long long)id);
  const Job& j = *it->second;
  if (j.status != AKMC_JOB_DONE) return fail(AKMC_ERR_INVALID, "job %lld has not finished", (long long)id);
  if ((int64_t)j.result.size() != n) return fail(AKMC_ERR_INVALID, "no kept result of %lld sites", (long long)n);
  memcpy(sites, j.result.data(), n);
  return AKMC_OK;
}

int akmc_server_wait_idle(akmc_server* s, double timeout_s) {
  if (!s) return fail(AKMC_ERR_INVALID, "null server");
  const auto until = Clock::now() + std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    if (s->error) return fail(s->error, "server stopped: %s", s->error_msg.c_str());
    if (s->pending.empty() && s->st.live == 0) return AKMC_OK;
    if (Clock::now() >= until) return fail(AKMC_ERR_INVALID, "timeout");
    s->cv.wait_for(lk, std::chrono::milliseconds(5));
  }
}

int akmc_server_get_stats(akmc_server* s, akmc_server_stats* st) {
  if (!s || !st) return fail(AKMC_ERR_INVALID, "bad arguments");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->st;
  return s->error ? fail(s->error, "server stopped: %s", s->error_msg.c_str()) : AKMC_OK;
}

double akmc_server_clock(akmc_server* s) { return s ? s->now() : 0.0; }

void akmc_server_destroy(akmc_server* s) {
  if (!s) return;
  s->stop.store(true);
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->R.device);
  (void)hipDeviceSynchronize();
  s->arena.destroy();
  akmc_mixpool::pool_destroy(s->mix);
  if (s->mix_done) (void)hipEventDestroy(s->mix_done);
  if (s->mix_stream) (void)hipStreamDestroy(s->mix_stream);
  if (s->init_stream) (void)hipStreamDestroy(s->init_stream);
  s->R.destroy();
  delete s;
}

}  // extern "C"
