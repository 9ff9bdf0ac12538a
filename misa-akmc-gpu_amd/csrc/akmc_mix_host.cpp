// Host side of the vacancy + dumbbell event loop: the C ABI (akmc_mix_*),
// device buffers per lattice, the random streams and the launch loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <unordered_set>
#include <vector>

#include "../../include/akmc.h"
#include "akmc_host.h"
#include "akmc_mix.h"
#include "akmc_mix_pool.h"
#include "akmc_plain.h"

using akmc_host::fail;

namespace {

constexpr int64_t kChunk = 1 << 16;  // events of host streams per upload
constexpr int64_t kRecWindow = 4096;  // stand-in draws per upload

bool valid_single(int t) { return t == 0 || t == 1 || t == 2 || t == 4 || t == 8; }
bool valid_dumbbell(int t) {
  const int hi = t >> 8, lo = t & 0xFF;
  return hi != 0 && valid_single(hi) && valid_single(lo) && hi <= lo && lo != 0 && (t >> 16) == 0;
}

struct MixLat {
  akmc_mix_params p{};
  hipStream_t mem_stream = nullptr;  // set: device buffers are stream-ordered (hipMallocAsync / hipFreeAsync)
  akmc::BoxDims d{};
  int64_t nsites = 0;
  int32_t vcap = 0, icap = 0, rcap = 0;
  // device
  uint16_t* types = nullptr;
  uint32_t* vkey = nullptr;
  uint32_t* ikey = nullptr;
  int8_t* io = nullptr;
  uint8_t* iav = nullptr;
  double* iterm = nullptr;
  double* ipre = nullptr;
  uint32_t* rkey = nullptr;
  uint8_t* rkind = nullptr;
  int32_t* rslot = nullptr;
  double* rate = nullptr;
  double* pvac = nullptr;
  uint64_t* tmask = nullptr;
  uint8_t* risk = nullptr;
  double* d_u = nullptr;
  double* d_l = nullptr;
  uint32_t* d_rec = nullptr;
  uint32_t* d_log = nullptr;
  int64_t log_cap = 0;
  // host streams (reference MT streams)
  std::mt19937 sel, time_rng, rec;
  std::uniform_real_distribution<double> dis{0, 1};
  int64_t generated = 0;  // events whose (u, l) were drawn
  int64_t rng_base = 0, rng_end = 0;
  std::vector<uint32_t> rec_draws;
  int64_t rec_base = 0, rec_end = 0;
  std::vector<double> h_u, h_l;
  int64_t target = 0;
  bool done = false;
  int32_t last_status = 0;
  // Device set-up of a generated lattice (freed by pool_setup_done).
  uint32_t* d_cum = nullptr;
  uint16_t* d_tval = nullptr;
  int64_t* d_idx = nullptr;
  std::vector<int64_t> h_idx;
  uint32_t h_cum[AKMC_MAX_ALLOY] = {};
  uint16_t h_tval[AKMC_MAX_ALLOY] = {};
  std::vector<uint32_t> h_vkey, h_ikey;
  std::vector<int8_t> h_io;
  akmc::mix::MixState h_init{};  // source of the initial state's upload
  int64_t bytes = 0;
};

// Frees a lattice's device buffers; stream-ordered ones on `fs` (their own
// allocation stream if null), after the work queued there.
void free_lat(MixLat& L, hipStream_t fs = nullptr) {
  void* ptrs[] = {L.types, L.vkey, L.ikey, L.io, L.iav, L.iterm, L.ipre, L.rkey, L.rkind, L.rslot, L.rate,
                  L.pvac, L.tmask, L.risk, L.d_u, L.d_l, L.d_rec, L.d_log, L.d_cum, L.d_tval, L.d_idx};
  for (void* q : ptrs)
    if (q) (void)(L.mem_stream ? hipFreeAsync(q, fs ? fs : L.mem_stream) : hipFree(q));
}

// Device bytes of a lattice: types, lists and sector records.
int64_t lat_bytes(int64_t nsites, int64_t vcap, int64_t icap) {
  return 2 * nsites + 4 * vcap + 22 * icap + 90 * (vcap + icap);
}

template <class T>
hipError_t dalloc(T** p, int64_t n, hipStream_t st = nullptr) {
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
  return st ? hipMallocAsync((void**)p, bytes, st) : hipMalloc((void**)p, bytes);
}

}  // namespace

struct akmc_mix {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<std::unique_ptr<MixLat>> lat;
  akmc::mix::MixState* d_state = nullptr;  // one per lattice
  akmc::mix::MixDesc* d_desc = nullptr;
  std::vector<akmc::mix::MixState> h_state;
  std::vector<akmc::mix::MixDesc> h_desc;
  double seconds = 0, kernel_seconds = 0;
  int64_t launches = 0;
};

namespace {

void release(akmc_mix* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (auto& L : m->lat) free_lat(*L);
  if (m->d_state) (void)hipFree(m->d_state);
  if (m->d_desc) (void)hipFree(m->d_desc);
  if (m->ev0) (void)hipEventDestroy(m->ev0);
  if (m->ev1) (void)hipEventDestroy(m->ev1);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

akmc::mix::MixDesc desc_of(MixLat& L, akmc::mix::MixState* d_st) {
  akmc::mix::MixDesc D{};
  D.types = L.types;
  D.vkey = L.vkey;
  D.vcap = L.vcap;
  D.icap = L.icap;
  D.rcap = L.rcap;
  D.ikey = L.ikey;
  D.io = L.io;
  D.iav = L.iav;
  D.iterm = L.iterm;
  D.ipre = L.ipre;
  D.rkey = L.rkey;
  D.rkind = L.rkind;
  D.rslot = L.rslot;
  D.rate = L.rate;
  D.pvac = L.pvac;
  D.tmask = L.tmask;
  D.risk = L.risk;
  D.st = d_st;
  const bool host = L.p.device_rng == 0;
  D.u = host ? L.d_u : nullptr;
  D.l = host ? L.d_l : nullptr;
  D.rng_base = L.rng_base;
  D.rng_end = L.rng_end;
  D.rec = host ? L.d_rec : nullptr;
  D.rec_base = L.rec_base;
  D.rec_end = L.rec_end;
  D.log = L.d_log;
  D.log_base = 0;
  D.log_cap = L.log_cap;
  D.d = L.d;
  D.max_events = L.target;
  D.time_steps = (int64_t)std::ceil(L.p.physics_time / L.p.sector_time);
  D.sector_T = L.p.sector_time;
  D.nu = L.p.attempt_freq;
  D.kT = AKMC_KB * L.p.temperature;  // RatesSolver::arrhenius: kB * T first
  D.ef110 = L.p.ef110;
  D.gen_rate = L.p.defect_gen_rate;
  D.rates_sector = -1;
  return D;
}

akmc::mix::MixDesc make_desc(akmc_mix* m, int i) { return desc_of(*m->lat[(size_t)i], m->d_state + i); }

// Device buffers of a lattice (lists of the given capacities, host streams
// when the lattice draws from the host).
int alloc_lat(MixLat& L) {
  HIP_TRY(dalloc(&L.types, L.nsites, L.mem_stream));
  HIP_TRY(dalloc(&L.vkey, L.vcap, L.mem_stream));
  HIP_TRY(dalloc(&L.ikey, L.icap, L.mem_stream));
  HIP_TRY(dalloc(&L.io, L.icap, L.mem_stream));
  HIP_TRY(dalloc(&L.iav, L.icap, L.mem_stream));
  HIP_TRY(dalloc(&L.iterm, L.icap, L.mem_stream));
  HIP_TRY(dalloc(&L.ipre, L.icap, L.mem_stream));
  HIP_TRY(dalloc(&L.rkey, L.rcap, L.mem_stream));
  HIP_TRY(dalloc(&L.rkind, L.rcap, L.mem_stream));
  HIP_TRY(dalloc(&L.rslot, L.rcap, L.mem_stream));
  HIP_TRY(dalloc(&L.rate, 8 * (int64_t)L.rcap, L.mem_stream));
  HIP_TRY(dalloc(&L.pvac, L.rcap, L.mem_stream));
  HIP_TRY(dalloc(&L.tmask, L.rcap, L.mem_stream));
  HIP_TRY(dalloc(&L.risk, L.rcap, L.mem_stream));
  if (L.p.device_rng == 0) {
    HIP_TRY(dalloc(&L.d_u, kChunk, L.mem_stream));
    HIP_TRY(dalloc(&L.d_l, kChunk, L.mem_stream));
    HIP_TRY(dalloc(&L.d_rec, kRecWindow, L.mem_stream));
    L.h_u.resize(kChunk);
    L.h_l.resize(kChunk);
    // r::rng is default-constructed (rand_generators.cpp:8); SubLattice
    // seeds its time stream with the (double -> int64) config seed
    // (simulation.h:70-76, sublattice.cpp:17).
    L.time_rng.seed((std::mt19937::result_type)(int64_t)(double)L.p.seed_time_inc);
    L.rec.seed((std::mt19937::result_type)L.p.seed_recombine);
  }
  L.bytes = lat_bytes(L.nsites, L.vcap, L.icap);
  return AKMC_OK;
}

// Initial state: defect counts and the device streams' xoshiro256** state
// (splitmix64 of the two seeds).
akmc::mix::MixState initial_state(const MixLat& L, int64_t nv, int64_t ni) {
  akmc::mix::MixState st{};
  st.nvac = (int32_t)nv;
  st.nitl = (int32_t)ni;
  uint64_t x = (uint64_t)L.p.seed_time_inc * 0x9E3779B97F4A7C15ull ^ (uint64_t)L.p.seed_recombine;
  for (int k = 0; k < 4; k++) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    st.rng[k] = z ^ (z >> 31);
  }
  return st;
}

// Host streams of lattice L cover events [ev, ev + n) and the stand-in draws
// [rec_pos, rec_pos + kRecWindow).
int fill_streams(akmc_mix* m, MixLat& L, int64_t ev, int64_t rec_pos) {
  if (L.p.device_rng) return AKMC_OK;
  if (ev >= L.rng_end || ev < L.rng_base) {
    if (ev != L.generated) return fail(AKMC_ERR_INVALID, "random stream out of step (%lld vs %lld)", (long long)ev,
                                       (long long)L.generated);
    const int64_t n = std::min<int64_t>(kChunk, std::max<int64_t>(1, L.target - ev));
    for (int64_t k = 0; k < n; k++) {
      // r::random() (random.cpp:35-41) and the time stream's -log(r)
      // (sublattice.inl:118-120).
      L.h_u[(size_t)k] = (double)L.sel() * (1.0 / 4294967295.0);
      L.h_l[(size_t)k] = -std::log(L.dis(L.time_rng));
    }
    L.generated += n;
    HIP_TRY(hipMemcpyAsync(L.d_u, L.h_u.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemcpyAsync(L.d_l, L.h_l.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, m->stream));
    L.rng_base = ev;
    L.rng_end = ev + n;
  }
  if (rec_pos + 4 > L.rec_end || rec_pos < L.rec_base) {
    while ((int64_t)L.rec_draws.size() < rec_pos + kRecWindow) L.rec_draws.push_back((uint32_t)L.rec());
    HIP_TRY(hipMemcpyAsync(L.d_rec, L.rec_draws.data() + rec_pos, (size_t)kRecWindow * sizeof(uint32_t),
                           hipMemcpyHostToDevice, m->stream));
    L.rec_base = rec_pos;
    L.rec_end = rec_pos + kRecWindow;
  }
  return AKMC_OK;
}

}  // namespace

extern "C" {

int akmc_mix_params_default(akmc_mix_params* p) {
  if (!p) return fail(AKMC_ERR_INVALID, "null argument");
  memset(p, 0, sizeof(*p));
  p->box[0] = p->box[1] = p->box[2] = 16;
  p->temperature = 573.0;
  p->attempt_freq = 6e12;
  p->sector_time = 3e-7;
  p->physics_time = 1.0;
  p->ef110 = 3.53;  // the reference's serial driver (profile_serial/main.cpp:45)
  p->defect_gen_rate = 0.0;
  p->seed_time_inc = 594689;
  p->seed_recombine = 563264;
  return AKMC_OK;
}

int akmc_mix_create(int32_t n, const akmc_mix_params* ps, const uint16_t* const* types, const int64_t* n_itl,
                    const int64_t* const* itl_sites, const int8_t* const* itl_orient, int device, akmc_mix** out) {
  if (!out || !ps || !types || !n_itl || n <= 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(AKMC_ERR_INVALID, "device %d out of range (%d devices)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  auto* m = new akmc_mix();
  m->device = device;
  std::unique_ptr<akmc_mix, void (*)(akmc_mix*)> guard(m, release);
  HIP_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&m->ev0));
  HIP_TRY(hipEventCreate(&m->ev1));
  HIP_TRY(hipMalloc(&m->d_state, sizeof(akmc::mix::MixState) * (size_t)n));
  HIP_TRY(hipMalloc(&m->d_desc, sizeof(akmc::mix::MixDesc) * (size_t)n));
  m->h_state.resize((size_t)n);
  for (int i = 0; i < n; i++) {
    const akmc_mix_params& p = ps[i];
    for (int d = 0; d < 3; d++)
      if (p.box[d] < 8) return fail(AKMC_ERR_INVALID, "lattice %d: box edge %lld < 8 cells", i, (long long)p.box[d]);
    if (2 * p.box[0] > 2048 || p.box[1] > 2047 || p.box[2] > 1023)
      return fail(AKMC_ERR_INVALID, "lattice %d: box exceeds 1024 x 2047 x 1023 cells", i);
    if (!(p.sector_time > 0) || !(p.temperature > 0) || !(p.attempt_freq > 0) || !(p.defect_gen_rate >= 0))
      return fail(AKMC_ERR_INVALID, "lattice %d: bad time, temperature, frequency or generation rate", i);
    if (!types[i] || n_itl[i] < 0 || (n_itl[i] > 0 && (!itl_sites || !itl_sites[i] || !itl_orient || !itl_orient[i])))
      return fail(AKMC_ERR_INVALID, "lattice %d: null lattice or dumbbell list", i);
    auto L = std::make_unique<MixLat>();
    L->p = p;
    L->d = akmc::box_dims(p.box);
    L->nsites = 2 * p.box[0] * p.box[1] * p.box[2];
    const uint16_t* t = types[i];
    int64_t nv = 0, ni = 0;
    for (int64_t s = 0; s < L->nsites; s++) {
      if (!valid_single(t[s]) && !valid_dumbbell(t[s]))
        return fail(AKMC_ERR_INVALID, "lattice %d: site %lld holds 0x%x, not a lattice type", i, (long long)s, t[s]);
      nv += t[s] == 0;
      ni += t[s] > 8;
    }
    if (ni != n_itl[i]) return fail(AKMC_ERR_INVALID, "lattice %d: %lld dumbbell sites, %lld listed", i, (long long)ni,
                                    (long long)n_itl[i]);
    std::vector<uint32_t> ik((size_t)ni);
    std::vector<int8_t> io((size_t)ni);
    for (int64_t k = 0; k < ni; k++) {
      const int64_t s = itl_sites[i][k];
      if (s < 0 || s >= L->nsites || t[s] <= 8) return fail(AKMC_ERR_INVALID, "lattice %d: dumbbell %lld not at a dumbbell site", i, (long long)k);
      if (k > 0 && s <= itl_sites[i][k - 1]) return fail(AKMC_ERR_INVALID, "lattice %d: dumbbell sites must ascend", i);
      const int o = itl_orient[i][k];
      if (o == 0 || o < -6 || o > 6) return fail(AKMC_ERR_INVALID, "lattice %d: orientation %d", i, o);
      ik[(size_t)k] = akmc::key_of_index(L->d, s);
      io[(size_t)k] = (int8_t)o;
    }
    const int64_t extra = p.defect_gen_rate > 0 ? 4096 : 0;
    L->vcap = (int32_t)std::max<int64_t>(1, p.vac_capacity > 0 ? p.vac_capacity : nv + extra);
    L->icap = (int32_t)std::max<int64_t>(1, p.itl_capacity > 0 ? p.itl_capacity : ni + extra);
    if (L->vcap < nv || L->icap < ni) return fail(AKMC_ERR_INVALID, "lattice %d: capacities below the defect counts", i);
    L->rcap = L->vcap + L->icap;
    std::vector<uint32_t> vk;
    for (int64_t s = 0; s < L->nsites; s++)
      if (t[s] == 0) vk.push_back(akmc::key_of_index(L->d, s));
    MixLat* Lp = L.get();
    m->lat.push_back(std::move(L));  // freed with the handle from here on
    {
      const int rc = alloc_lat(*Lp);
      if (rc) return rc;
    }
    HIP_TRY(hipMemcpy(Lp->types, t, (size_t)Lp->nsites * 2, hipMemcpyHostToDevice));
    if (nv) HIP_TRY(hipMemcpy(Lp->vkey, vk.data(), vk.size() * 4, hipMemcpyHostToDevice));
    if (ni) {
      HIP_TRY(hipMemcpy(Lp->ikey, ik.data(), ik.size() * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(Lp->io, io.data(), io.size(), hipMemcpyHostToDevice));
    }
    m->h_state[(size_t)i] = initial_state(*Lp, nv, ni);
  }
  HIP_TRY(hipMemcpy(m->d_state, m->h_state.data(), m->h_state.size() * sizeof(m->h_state[0]), hipMemcpyHostToDevice));
  *out = guard.release();
  return AKMC_OK;
}

int akmc_mix_run(akmc_mix* m, int64_t max_events, int64_t* total_events) {
  if (!m || max_events < 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(m->device));
  const auto t0 = std::chrono::steady_clock::now();
  int64_t before = 0;
  for (size_t i = 0; i < m->lat.size(); i++) {
    MixLat& L = *m->lat[i];
    const int64_t ev = m->h_state[i].events;
    before += ev;
    L.target = max_events > INT64_MAX - ev ? INT64_MAX : ev + max_events;
    L.done = false;
  }
  m->kernel_seconds = 0;
  m->launches = 0;
  for (;;) {
    std::vector<int> which;
    for (size_t i = 0; i < m->lat.size(); i++) {
      MixLat& L = *m->lat[i];
      if (L.done) continue;
      const int rc = fill_streams(m, L, m->h_state[i].events, m->h_state[i].rec_pos);
      if (rc) return rc;
      which.push_back((int)i);
    }
    if (which.empty()) break;
    // Descriptors index lattices in `which` order; every state lives at its
    // own lattice's slot, so map the launch onto the lattices themselves.
    {
      std::vector<akmc::mix::MixDesc> ds;
      for (int i : which) ds.push_back(make_desc(m, i));
      HIP_TRY(hipMemcpyAsync(m->d_desc, ds.data(), ds.size() * sizeof(ds[0]), hipMemcpyHostToDevice, m->stream));
      HIP_TRY(hipEventRecord(m->ev0, m->stream));
      const int rc = akmc_mix_launch(m->d_desc, (int32_t)ds.size(), m->stream);
      if (rc) return fail(AKMC_ERR_HIP, "mix kernel launch failed: %s", hipGetErrorString((hipError_t)rc));
      HIP_TRY(hipEventRecord(m->ev1, m->stream));
      HIP_TRY(hipMemcpyAsync(m->h_state.data(), m->d_state, m->h_state.size() * sizeof(m->h_state[0]),
                             hipMemcpyDeviceToHost, m->stream));
      HIP_TRY(hipStreamSynchronize(m->stream));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
      m->kernel_seconds += ms * 1e-3;
      m->launches++;
    }
    for (int i : which) {
      MixLat& L = *m->lat[(size_t)i];
      const akmc::mix::MixState& st = m->h_state[(size_t)i];
      L.last_status = st.status;
      switch (st.status) {
        case AKMC_RUN_TIME_DONE:
        case AKMC_RUN_EVENT_CAP:
          L.done = true;
          break;
        case AKMC_RUN_NEED_RNG:
        case akmc::mix::MIX_NEED_REC:
        case AKMC_RUN_YIELD:
        case AKMC_RUN_OK:
          break;
        case AKMC_RUN_ERR_CAPACITY:
          return fail(AKMC_ERR_CAPACITY, "lattice %d: a list outgrew its capacity", i);
        case AKMC_RUN_ERR_NO_EVENT:
          return fail(AKMC_ERR_NO_EVENT, "lattice %d: u * sum >= sum and no generation rate", i);
        default:
          return fail(AKMC_ERR_HIP, "lattice %d: unexpected status %d", i, st.status);
      }
      if (st.events >= L.target) L.done = true;
    }
  }
  int64_t after = 0;
  for (auto& s : m->h_state) after += s.events;
  if (total_events) *total_events = after - before;
  m->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return AKMC_OK;
}

int akmc_mix_stats_of(akmc_mix* m, int32_t i, akmc_mix_stats* s) {
  if (!m || !s || i < 0 || i >= (int32_t)m->lat.size()) return fail(AKMC_ERR_INVALID, "bad arguments");
  const akmc::mix::MixState& st = m->h_state[(size_t)i];
  memset(s, 0, sizeof(*s));
  s->events = st.events;
  s->step = st.step;
  s->sector_pos = st.sect;
  s->in_sector = st.in_sector;
  for (int k = 0; k < 8; k++) s->evolution[k] = st.evo[k];
  s->sector_time = st.t;
  const int32_t ls = m->lat[(size_t)i]->last_status;
  s->stop = ls == AKMC_RUN_TIME_DONE ? AKMC_STOP_TIME : AKMC_STOP_EVENTS;
  s->nvac = st.nvac;
  s->nitl = st.nitl;
  s->n_recombined = st.n_recombined;
  s->n_generated = st.n_generated;
  s->n_itl_events = st.n_itl_events;
  s->n_eval = st.n_eval;
  s->seconds = m->seconds;
  s->kernel_seconds = m->kernel_seconds;
  s->launches = m->launches;
  s->n_skipped_steps = st.n_skipped;
  return AKMC_OK;
}

int akmc_mix_download(akmc_mix* m, int32_t i, uint16_t* types, int64_t n_sites, int64_t cap, int64_t* itl_sites,
                      int8_t* itl_orient, int64_t* n_itl) {
  if (!m || i < 0 || i >= (int32_t)m->lat.size()) return fail(AKMC_ERR_INVALID, "bad arguments");
  MixLat& L = *m->lat[(size_t)i];
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (types) {
    if (n_sites != L.nsites) return fail(AKMC_ERR_INVALID, "site count %lld != %lld", (long long)n_sites, (long long)L.nsites);
    HIP_TRY(hipMemcpy(types, L.types, (size_t)L.nsites * 2, hipMemcpyDeviceToHost));
  }
  const int32_t ni = m->h_state[(size_t)i].nitl;
  if (n_itl) *n_itl = ni;
  if (itl_sites && itl_orient && cap > 0) {
    std::vector<uint32_t> k((size_t)ni);
    std::vector<int8_t> o((size_t)ni);
    if (ni) {
      HIP_TRY(hipMemcpy(k.data(), L.ikey, k.size() * 4, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(o.data(), L.io, o.size(), hipMemcpyDeviceToHost));
    }
    for (int32_t q = 0; q < ni && q < cap; q++) {
      const uint32_t key = k[(size_t)q];
      itl_sites[q] = (int64_t)(key & 0x7FF) + (int64_t)L.d.X * ((int64_t)((key >> 11) & 0x7FF) + (int64_t)L.d.Y * (key >> 22));
      itl_orient[q] = o[(size_t)q];
    }
  }
  return AKMC_OK;
}

int akmc_mix_sector_rates(akmc_mix* m, int32_t i, int sector, int64_t cap, int64_t* sites, int32_t* kinds,
                          double* rates, double* total, int64_t* n) {
  if (!m || i < 0 || i >= (int32_t)m->lat.size() || sector < 0 || sector > 7 || !total || !n || cap < 0 ||
      (cap > 0 && (!sites || !kinds || !rates)))
    return fail(AKMC_ERR_INVALID, "bad arguments");
  MixLat& L = *m->lat[(size_t)i];
  HIP_TRY(hipSetDevice(m->device));
  const int64_t oc = std::max<int64_t>(1, std::min<int64_t>(cap, L.rcap));
  uint32_t* ok = nullptr;
  uint8_t* okd = nullptr;
  double *orr = nullptr, *ot = nullptr;
  int32_t* on = nullptr;
  struct Frees {
    std::vector<void*> p;
    ~Frees() {
      for (void* q : p) (void)hipFree(q);
    }
  } fr;
  HIP_TRY(dalloc(&ok, oc));
  fr.p.push_back(ok);
  HIP_TRY(dalloc(&okd, oc));
  fr.p.push_back(okd);
  HIP_TRY(dalloc(&orr, 8 * oc));
  fr.p.push_back(orr);
  HIP_TRY(dalloc(&ot, 1));
  fr.p.push_back(ot);
  HIP_TRY(dalloc(&on, 1));
  fr.p.push_back(on);
  akmc::mix::MixDesc D = make_desc(m, i);
  D.rates_sector = sector;
  D.out_cap = (int32_t)oc;
  D.out_keys = ok;
  D.out_kinds = okd;
  D.out_rates = orr;
  D.out_total = ot;
  D.out_n = on;
  D.u = nullptr;  // no event is drawn
  D.rec = nullptr;
  D.max_events = INT64_MAX;
  HIP_TRY(hipMemcpyAsync(m->d_desc, &D, sizeof(D), hipMemcpyHostToDevice, m->stream));
  const int rc = akmc_mix_launch(m->d_desc, 1, m->stream);
  if (rc) return fail(AKMC_ERR_HIP, "mix kernel launch failed");
  int32_t nn = 0;
  double tt = 0;
  HIP_TRY(hipMemcpyAsync(&nn, on, 4, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipMemcpyAsync(&tt, ot, 8, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  if (nn < 0) return fail(AKMC_ERR_CAPACITY, "sector rates failed (status %d)", nn);
  const int64_t w = std::min<int64_t>(nn, cap);
  std::vector<uint32_t> k((size_t)std::max<int64_t>(w, 1));
  std::vector<uint8_t> kd((size_t)std::max<int64_t>(w, 1));
  if (w > 0) {
    HIP_TRY(hipMemcpy(k.data(), ok, (size_t)w * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(kd.data(), okd, (size_t)w, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rates, orr, (size_t)w * 64, hipMemcpyDeviceToHost));
  }
  for (int64_t q = 0; q < w; q++) {
    const uint32_t key = k[(size_t)q];
    sites[q] = (int64_t)(key & 0x7FF) + (int64_t)L.d.X * ((int64_t)((key >> 11) & 0x7FF) + (int64_t)L.d.Y * (key >> 22));
    kinds[q] = kd[(size_t)q];
  }
  *total = tt;
  *n = nn;
  return AKMC_OK;
}

int akmc_mix_set_event_log(akmc_mix* m, int64_t capacity) {
  if (!m || capacity < 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(m->device));
  for (auto& L : m->lat) {
    if (L->d_log) {
      HIP_TRY(hipFree(L->d_log));
      L->d_log = nullptr;
    }
    L->log_cap = capacity;
    if (capacity > 0) HIP_TRY(dalloc(&L->d_log, 3 * capacity));
  }
  return AKMC_OK;
}

int akmc_mix_get_event_log(akmc_mix* m, int32_t i, int64_t* from, int64_t* to, int32_t* kind, int64_t cap, int64_t* n) {
  if (!m || i < 0 || i >= (int32_t)m->lat.size() || !n || cap < 0 || (cap > 0 && (!from || !to || !kind)))
    return fail(AKMC_ERR_INVALID, "bad arguments");
  MixLat& L = *m->lat[(size_t)i];
  const int64_t have = std::min<int64_t>(m->h_state[(size_t)i].events, L.log_cap);
  *n = have;
  const int64_t w = std::min(have, cap);
  if (w <= 0) return AKMC_OK;
  std::vector<uint32_t> buf((size_t)(3 * w));
  HIP_TRY(hipSetDevice(m->device));
  HIP_TRY(hipMemcpy(buf.data(), L.d_log, buf.size() * 4, hipMemcpyDeviceToHost));
  auto site = [&](uint32_t key) -> int64_t {
    if (key == 0xFFFFFFFFu) return -1;
    return (int64_t)(key & 0x7FF) + (int64_t)L.d.X * ((int64_t)((key >> 11) & 0x7FF) + (int64_t)L.d.Y * (key >> 22));
  };
  for (int64_t q = 0; q < w; q++) {
    from[q] = site(buf[(size_t)(3 * q)]);
    to[q] = site(buf[(size_t)(3 * q + 1)]);
    kind[q] = (int32_t)buf[(size_t)(3 * q + 2)];
  }
  return AKMC_OK;
}

void akmc_mix_destroy(akmc_mix* m) { release(m); }

int akmc_mix_random_lattice(const akmc_params* p, const int32_t* itl_types, const int64_t* itl_counts, int32_t n_kinds,
                            int64_t seed, uint16_t* types, int64_t n_sites, int64_t cap, int64_t* itl_sites,
                            int8_t* itl_orient, int64_t* n_itl) {
  if (!p || !types || !n_itl || n_kinds < 0 || (n_kinds > 0 && (!itl_types || !itl_counts)))
    return fail(AKMC_ERR_INVALID, "bad arguments");
  const int64_t N = 2 * p->box[0] * p->box[1] * p->box[2];
  if (n_sites != N) return fail(AKMC_ERR_INVALID, "site count %lld != %lld", (long long)n_sites, (long long)N);
  std::vector<uint8_t> f((size_t)N);
  const int rc = akmc_random_sites(p, f.data(), N);
  if (rc) return rc;
  // The typed lattice of creation::createRandom's occupancy (a vacancy's
  // leftover Ni/Mn tag has no place in a typed lattice).
  for (int64_t s = 0; s < N; s++) {
    const uint8_t b = f[(size_t)s];
    types[s] = (b & 1) ? 0 : (b & 2) ? 2 : (b & 4) ? 8 : (b & 8) ? 4 : 1;
  }
  // placeDumbbells (profile_serial/main.cpp:137-161): for each kind, draw a
  // site r::rand32(0, N) = mt() % N; a single atom there becomes the dumbbell
  // of itself and the new atom, with orientation::random(); other sites are
  // skipped.
  std::mt19937 rng((std::mt19937::result_type)seed);
  std::vector<std::pair<int64_t, int8_t>> placed;
  for (int32_t k = 0; k < n_kinds; k++) {
    const int add = itl_types[k];
    if (!(add == 1 || add == 2 || add == 4 || add == 8)) return fail(AKMC_ERR_INVALID, "dumbbell atom 0x%x", add);
    for (int64_t c = 0; c < itl_counts[k]; c++) {
      const int64_t id = (int64_t)((uint64_t)rng() % (uint64_t)N);
      const int t = types[id];
      if (t == 1 || t == 2 || t == 4 || t == 8) {
        types[id] = (uint16_t)(t <= add ? (t << 8) | add : (add << 8) | t);
        const uint32_t r = (uint32_t)rng();
        const uint32_t v = r % 12u + 1u;  // orientation::random (orientation.h:38-42)
        placed.emplace_back(id, (int8_t)(v % 2 == 0 ? (int)(v / 2) : -(int)((v + 1) / 2)));
      }
    }
  }
  std::sort(placed.begin(), placed.end());
  *n_itl = (int64_t)placed.size();
  if (itl_sites && itl_orient)
    for (size_t q = 0; q < placed.size() && (int64_t)q < cap; q++) {
      itl_sites[q] = placed[q].first;
      itl_orient[q] = placed[q].second;
    }
  return AKMC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- pool
namespace akmc_mixpool {

struct Pool {
  int device = 0;
  int32_t nslots = 0;
  akmc::mix::MixState* d_state = nullptr;
  akmc::mix::MixDesc* d_desc = nullptr;
  // Pinned host copies, so that a launch's copies stay asynchronous.
  akmc::mix::MixState* h_state = nullptr;
  akmc::mix::MixDesc* h_desc = nullptr;
  std::vector<std::unique_ptr<MixLat>> slot;
  int32_t hi = 0;  // states copied back: slots [0, hi)
};

int pool_create(int device, int32_t slots, Pool** out) {
  if (!out || slots <= 0) return fail(AKMC_ERR_INVALID, "bad pool arguments");
  std::unique_ptr<Pool, void (*)(Pool*)> p(new Pool(), pool_destroy);  // frees what was made on an error
  p->device = device;
  p->nslots = slots;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipMalloc(&p->d_state, sizeof(akmc::mix::MixState) * (size_t)slots));
  HIP_TRY(hipMalloc(&p->d_desc, sizeof(akmc::mix::MixDesc) * (size_t)slots));
  HIP_TRY(hipHostMalloc((void**)&p->h_state, sizeof(akmc::mix::MixState) * (size_t)slots, hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&p->h_desc, sizeof(akmc::mix::MixDesc) * (size_t)slots, hipHostMallocDefault));
  memset(p->h_state, 0, sizeof(akmc::mix::MixState) * (size_t)slots);
  p->slot.resize((size_t)slots);
  *out = p.release();
  return AKMC_OK;
}

void pool_destroy(Pool* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  for (auto& L : p->slot)
    if (L) free_lat(*L);
  if (p->d_state) (void)hipFree(p->d_state);
  if (p->d_desc) (void)hipFree(p->d_desc);
  if (p->h_state) (void)hipHostFree(p->h_state);
  if (p->h_desc) (void)hipHostFree(p->h_desc);
  delete p;
}

int64_t pool_bytes(const akmc_params& comp, int64_t n_itl) {
  const int64_t n = 2 * comp.box[0] * comp.box[1] * comp.box[2];
  return lat_bytes(n, std::max<int64_t>(1, comp.va_count), std::max<int64_t>(1, n_itl));
}

int pool_add_generated(Pool* p, int32_t slot, const akmc_params& comp, int64_t n_itl, int32_t itl_atom, double ef110,
                       hipStream_t st) {
  if (!p || slot < 0 || slot >= p->nslots || p->slot[(size_t)slot]) return fail(AKMC_ERR_INVALID, "bad pool slot");
  for (int d = 0; d < 3; d++)
    if (comp.box[d] < 8) return fail(AKMC_ERR_INVALID, "box edge %lld < 8 cells", (long long)comp.box[d]);
  if (2 * comp.box[0] > 2048 || comp.box[1] > 2047 || comp.box[2] > 1023)
    return fail(AKMC_ERR_INVALID, "box exceeds 1024 x 2047 x 1023 cells");
  if (!(itl_atom == 1 || itl_atom == 2 || itl_atom == 4 || itl_atom == 8) || n_itl < 0 || comp.va_count < 0)
    return fail(AKMC_ERR_INVALID, "bad dumbbell atom or counts");
  auto L = std::make_unique<MixLat>();
  akmc_mix_params& mp = L->p;
  akmc_mix_params_default(&mp);
  for (int d = 0; d < 3; d++) mp.box[d] = comp.box[d];
  mp.temperature = comp.temperature;
  mp.attempt_freq = comp.attempt_freq;
  mp.sector_time = comp.sector_time;
  mp.physics_time = comp.physics_time;
  mp.ef110 = ef110;
  mp.seed_time_inc = comp.seed_time_inc;
  mp.seed_recombine = comp.seed_event_selection;
  mp.device_rng = 1;
  if (!(mp.sector_time > 0) || !(mp.temperature > 0) || !(mp.attempt_freq > 0))
    return fail(AKMC_ERR_INVALID, "bad time, temperature or frequency");
  L->d = akmc::box_dims(mp.box);
  L->nsites = 2 * mp.box[0] * mp.box[1] * mp.box[2];
  if (comp.va_count + n_itl > L->nsites / 2) return fail(AKMC_ERR_INVALID, "too many defects for the box");
  // Composition brackets of the alloy, as the replica engine draws them.
  int nt = 0;
  uint64_t total = 0;
  for (int i = 0; i < comp.n_alloy; i++) {
    const int t = comp.alloy_type[i];
    const int64_t q = comp.alloy_ratio[i];
    if (q < 0 || !(t == 1 || t == 2 || t == 4 || t == 8)) return fail(AKMC_ERR_INVALID, "bad alloy entry");
    if (q == 0) continue;
    total += (uint64_t)q;
    if (total > 0xFFFFFFFFull) return fail(AKMC_ERR_INVALID, "alloy ratios too large");
    L->h_cum[nt] = (uint32_t)total;
    L->h_tval[nt] = (uint16_t)t;
    nt++;
  }
  if (nt == 0) return fail(AKMC_ERR_INVALID, "empty alloy");
  // Vacancy sites as the replica engine draws them (repeats collapse), then
  // distinct dumbbell sites off them, and orientation::random draws.
  std::vector<int64_t> vidx((size_t)comp.va_count);
  uint64_t x = (uint64_t)comp.seed_create_vacancy ^ 0x5A5A5A5A5A5A5A5Aull;
  auto next = [](uint64_t& z) {
    uint64_t r = (z += 0x9E3779B97F4A7C15ull);
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
    return r ^ (r >> 31);
  };
  for (auto& v : vidx) v = (int64_t)(((unsigned __int128)next(x) * (uint64_t)L->nsites) >> 64);
  std::sort(vidx.begin(), vidx.end());
  vidx.erase(std::unique(vidx.begin(), vidx.end()), vidx.end());
  std::unordered_set<int64_t> taken(vidx.begin(), vidx.end());
  std::vector<std::pair<int64_t, int8_t>> itl;
  uint64_t y = (uint64_t)comp.seed_create_vacancy ^ 0xA5A5A5A5A5A5A5A5ull;
  while ((int64_t)itl.size() < n_itl) {
    const int64_t s = (int64_t)(((unsigned __int128)next(y) * (uint64_t)L->nsites) >> 64);
    if (!taken.insert(s).second) continue;
    const uint32_t v = (uint32_t)(next(y) >> 32) % 12u + 1u;  // orientation::random (orientation.h:38-42)
    itl.emplace_back(s, (int8_t)(v % 2 == 0 ? (int)(v / 2) : -(int)((v + 1) / 2)));
  }
  std::sort(itl.begin(), itl.end());
  L->vcap = (int32_t)std::max<int64_t>(1, (int64_t)vidx.size());
  L->icap = (int32_t)std::max<int64_t>(1, n_itl);
  L->rcap = L->vcap + L->icap;
  L->mem_stream = st;
  MixLat* Lp = L.get();
  p->slot[(size_t)slot] = std::move(L);
  int rc = alloc_lat(*Lp);
  if (rc) return rc;
  Lp->h_idx = vidx;
  for (auto& q : itl) Lp->h_idx.push_back(q.first);
  for (int64_t s : vidx) Lp->h_vkey.push_back(akmc::key_of_index(Lp->d, s));
  for (auto& q : itl) {
    Lp->h_ikey.push_back(akmc::key_of_index(Lp->d, q.first));
    Lp->h_io.push_back(q.second);
  }
  HIP_TRY(dalloc(&Lp->d_cum, nt, st));
  HIP_TRY(dalloc(&Lp->d_tval, nt, st));
  HIP_TRY(dalloc(&Lp->d_idx, (int64_t)Lp->h_idx.size(), st));
  HIP_TRY(hipMemcpyAsync(Lp->d_cum, Lp->h_cum, sizeof(uint32_t) * nt, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(Lp->d_tval, Lp->h_tval, sizeof(uint16_t) * nt, hipMemcpyHostToDevice, st));
  if (!Lp->h_idx.empty())
    HIP_TRY(hipMemcpyAsync(Lp->d_idx, Lp->h_idx.data(), Lp->h_idx.size() * 8, hipMemcpyHostToDevice, st));
  if (akmc_mix_fill_launch(Lp->types, Lp->nsites, (uint64_t)comp.seed_create_types, Lp->d_cum, Lp->d_tval, nt, st))
    return fail(AKMC_ERR_HIP, "fill launch failed");
  if (akmc_mix_defects_launch(Lp->types, Lp->d_idx, (int64_t)vidx.size(), Lp->d_idx + vidx.size(), (int64_t)itl.size(),
                              itl_atom, st))
    return fail(AKMC_ERR_HIP, "defect placement launch failed");
  if (!Lp->h_vkey.empty())
    HIP_TRY(hipMemcpyAsync(Lp->vkey, Lp->h_vkey.data(), Lp->h_vkey.size() * 4, hipMemcpyHostToDevice, st));
  if (!Lp->h_ikey.empty()) {
    HIP_TRY(hipMemcpyAsync(Lp->ikey, Lp->h_ikey.data(), Lp->h_ikey.size() * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(Lp->io, Lp->h_io.data(), Lp->h_io.size(), hipMemcpyHostToDevice, st));
  }
  // The upload's source stays in the lattice record: the pool's own state
  // array is overwritten by launches that may run meanwhile.
  Lp->h_init = initial_state(*Lp, (int64_t)vidx.size(), (int64_t)itl.size());
  HIP_TRY(hipMemcpyAsync(p->d_state + slot, &Lp->h_init, sizeof(akmc::mix::MixState), hipMemcpyHostToDevice, st));
  return AKMC_OK;
}

void pool_setup_done(Pool* p, int32_t slot, hipStream_t st) {
  MixLat* L = p->slot[(size_t)slot].get();
  if (!L) return;
  void* q[] = {L->d_cum, L->d_tval, L->d_idx};
  for (void* v : q)
    if (v) (void)(L->mem_stream ? hipFreeAsync(v, st) : hipFree(v));
  L->d_cum = nullptr;
  L->d_tval = nullptr;
  L->d_idx = nullptr;
  std::vector<int64_t>().swap(L->h_idx);
  std::vector<uint32_t>().swap(L->h_vkey);
  std::vector<uint32_t>().swap(L->h_ikey);
  std::vector<int8_t>().swap(L->h_io);
}

int pool_remove(Pool* p, int32_t slot, hipStream_t st) {
  if (!p || slot < 0 || slot >= p->nslots || !p->slot[(size_t)slot]) return fail(AKMC_ERR_INVALID, "bad pool slot");
  free_lat(*p->slot[(size_t)slot], st);
  p->slot[(size_t)slot].reset();
  return AKMC_OK;
}

int pool_launch(Pool* p, const std::vector<int32_t>& slots, const std::vector<int64_t>& limits, hipStream_t st) {
  if (slots.empty()) return AKMC_OK;
  if ((int64_t)slots.size() > p->nslots) return fail(AKMC_ERR_INVALID, "more lattices than pool slots");
  int32_t hi = 0;
  for (size_t k = 0; k < slots.size(); k++) {
    MixLat& L = *p->slot[(size_t)slots[k]];
    L.target = limits[k];
    p->h_desc[k] = desc_of(L, p->d_state + slots[k]);
    hi = std::max(hi, slots[k] + 1);
  }
  HIP_TRY(hipMemcpyAsync(p->d_desc, p->h_desc, slots.size() * sizeof(akmc::mix::MixDesc), hipMemcpyHostToDevice, st));
  const int rc = akmc_mix_launch(p->d_desc, (int32_t)slots.size(), st);
  if (rc) return fail(AKMC_ERR_HIP, "mix kernel launch failed");
  HIP_TRY(hipMemcpyAsync(p->h_state, p->d_state, (size_t)hi * sizeof(akmc::mix::MixState),
                         hipMemcpyDeviceToHost, st));
  return AKMC_OK;
}

const akmc::mix::MixState& pool_state(Pool* p, int32_t slot) { return p->h_state[(size_t)slot]; }

int64_t pool_sites(Pool* p, int32_t slot) { return p->slot[(size_t)slot] ? p->slot[(size_t)slot]->nsites : 0; }

int pool_download(Pool* p, int32_t slot, uint16_t* types, int64_t n) {
  MixLat* L = p->slot[(size_t)slot].get();
  if (!L || !types || n != L->nsites) return fail(AKMC_ERR_INVALID, "bad download");
  HIP_TRY(hipMemcpy(types, L->types, (size_t)n * 2, hipMemcpyDeviceToHost));
  return AKMC_OK;
}

}  // namespace akmc_mixpool
