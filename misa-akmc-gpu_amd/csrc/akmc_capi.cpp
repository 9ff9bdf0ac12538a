// libakmc host side: the C ABI of include/akmc.h.
//
// The engine keeps one lattice resident in HBM and drives the device event
// loop (akmc_plain.hip).  The two random streams the reference consumes once
// per event -- the selection draw u = r::random() (src/utils/random/
// random.cpp:35-41, a default-constructed mt19937) and the residence-time
// draw r of SubLattice's uniform_real_distribution (sublattice.inl:118-119) --
// do not depend on the simulation state, so the host produces them ahead in
// chunks (u_i and l_i = -log(r_i), with this C library's log) and streams
// them to the device while the previous chunk is being consumed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/akmc.h"
#include "akmc_exp.h"
#include "akmc_exp_table.h"
#include "akmc_batch.h"
#include "akmc_host.h"
#include "akmc_model.h"
#include "akmc_plain.h"
#include "akmc_yaml.h"

namespace akmc_host {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Builds the 55-site rate stencil tables (see akmc_model.h).
void build_stencil(akmc::StencilTables* S) {
  memset(S, 0, sizeof(*S));
  for (int par = 0; par < 2; par++) {
    std::vector<std::array<int, 3>> sites;
    auto add = [&](int dx, int dy, int dz) -> uint8_t {
      for (size_t i = 0; i < sites.size(); i++)
        if (sites[i][0] == dx && sites[i][1] == dy && sites[i][2] == dz) return (uint8_t)i;
      sites.push_back({dx, dy, dz});
      return (uint8_t)(sites.size() - 1);
    };
    add(0, 0, 0);
    for (int b = 0; b < 8; b++) {
      int dx, dy, dz;
      akmc::nn1_offset(par, b, &dx, &dy, &dz);
      S->w1[par][b] = add(dx, dy, dz);
    }
    int kk = 0;
    for (int k = 0; k < 6; k++) {
      if (!((akmc::MASK_2NN >> k) & 1)) continue;
      int dx, dy, dz;
      akmc::nn2_offset(k, &dx, &dy, &dz);
      S->w2[par][kk++] = add(dx, dy, dz);
    }
    for (int b = 0; b < 8; b++) {
      int tx, ty, tz;
      akmc::nn1_offset(par, b, &tx, &ty, &tz);
      for (int k = 0; k < 8; k++) {
        int dx, dy, dz;
        akmc::nn1_offset(par ^ 1, k, &dx, &dy, &dz);
        S->t1[par][b][k] = add(tx + dx, ty + dy, tz + dz);
      }
      int k2 = 0;
      for (int k = 0; k < 6; k++) {
        if (!((akmc::MASK_2NN >> k) & 1)) continue;
        int dx, dy, dz;
        akmc::nn2_offset(k, &dx, &dy, &dz);
        S->t2[par][b][k2++] = add(tx + dx, ty + dy, tz + dz);
      }
    }
    S->n_sites = (int32_t)sites.size();
    for (size_t i = 0; i < sites.size(); i++) {
      for (int c = 0; c < 3; c++) S->off[par][i][c] = (int8_t)sites[i][c];
      S->reach[par][sites[i][2] + 2][sites[i][1] + 2] |= (uint8_t)(1u << (sites[i][0] + 3));
    }
  }
}

}  // namespace akmc_host

namespace akmc_host {

const akmc::StencilTables& stencil_tables() {
  static akmc::StencilTables tables;
  static std::once_flag once;
  std::call_once(once, [] { build_stencil(&tables); });
  return tables;
}

int upload_stencil(int device) {
  static std::mutex m;
  static std::vector<bool> done;
  const akmc::StencilTables& t = stencil_tables();
  if (t.n_sites > akmc::STENCIL_MAX) return fail(AKMC_ERR_INVALID, "stencil too large");
  std::lock_guard<std::mutex> g(m);
  if (device < 0) return fail(AKMC_ERR_INVALID, "bad device");
  if ((size_t)device >= done.size()) done.resize((size_t)device + 1, false);
  if (done[(size_t)device]) return AKMC_OK;
  if (akmc_plain_set_stencil(&t) != 0 || akmc_batch_set_stencil(&t) != 0)
    return fail(AKMC_ERR_HIP, "copying stencil tables failed");
  done[(size_t)device] = true;
  return AKMC_OK;
}

}  // namespace akmc_host

namespace {

using akmc_host::build_stencil;
using akmc_host::fail;
using akmc_host::g_err;

constexpr int64_t kChunkMax = int64_t(1) << 20;  // events per random-number chunk (default and maximum)

int lattype_from_name(const std::string& s) {
  static const std::map<std::string, int> m = {
      {"V", 0},        {"Fe", 0x1},     {"Cu", 0x2},     {"Ni", 0x4},     {"Mn", 0x8},
      {"FeFe", 0x101}, {"FeCu", 0x102}, {"FeNi", 0x104}, {"FeMn", 0x108}, {"CuCu", 0x202},
      {"CuNi", 0x204}, {"CuMn", 0x208}, {"NiNi", 0x404}, {"NiMn", 0x408}, {"MnMn", 0x808}};
  auto it = m.find(s);
  return it == m.end() ? -1 : it->second;
}

}  // namespace

struct akmc_engine {
  akmc_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  akmc::BoxDims dims{};
  int64_t nsites = 0;
  int64_t time_steps = 0;
  double kT = 0;
  uint8_t* d_sites = nullptr;
  uint32_t* d_vac = nullptr;
  int32_t nvac = 0;
  int32_t vac_capacity = 0;
  akmc::PlainState* d_state = nullptr;
  akmc::PlainState h_state{};
  // random streams
  std::mt19937 sel_rng;   // r::rng, default seed 5489 (rand_generators.cpp:8)
  std::mt19937 time_rng;  // SubLattice::time_inc_rng
  std::uniform_real_distribution<double> time_dis{0.0, 1.0};
  double* h_u[2] = {nullptr, nullptr};
  double* h_l[2] = {nullptr, nullptr};
  double* d_u[2] = {nullptr, nullptr};
  double* d_l[2] = {nullptr, nullptr};
  int64_t chunk_base[2] = {-1, -1};  // first event of each buffered chunk
  int64_t generated = 0;             // events whose random numbers were generated
  int64_t chunk = kChunkMax;         // events per chunk (one launch consumes at most one chunk)
  int cur = 0;
  // event log
  uint32_t* d_log = nullptr;
  int64_t log_cap = 0;
  int64_t log_base = 0;
  bool lattice_ready = false;
  int32_t check_sums = 0;
  int32_t select_mode = 0;
  int64_t step_limit = INT64_MAX;  // akmc_run_steps: the time loop ends at this step
  uint8_t* d_sector = nullptr;     // device-memory sector state (sectors beyond the LDS capacity)
  int32_t sector_cap = 0;
  akmc_stats last{};
};

namespace {

int generate_chunk(akmc_engine* e, int buf) {
  const int64_t base = e->generated;
  double* u = e->h_u[buf];
  double* l = e->h_l[buf];
  for (int64_t i = 0; i < e->chunk; i++) {
    u[i] = e->sel_rng() * (1.0 / std::mt19937::max());
    const double r = e->time_dis(e->time_rng);
    l[i] = -std::log(r);
  }
  e->chunk_base[buf] = base;
  e->generated = base + e->chunk;
  HIP_TRY(hipMemcpyAsync(e->d_u[buf], u, e->chunk * sizeof(double), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipMemcpyAsync(e->d_l[buf], l, e->chunk * sizeof(double), hipMemcpyHostToDevice, e->stream));
  return AKMC_OK;
}

int upload_vacancies(akmc_engine* e, const uint8_t* sites) {
  std::vector<uint32_t> keys;
  const int64_t X = e->dims.X, Y = e->dims.Y;
  for (int64_t i = 0; i < e->nsites; i++) {
    if (sites[i] & akmc::BIT_V) {
      const uint32_t x = (uint32_t)(i % X), y = (uint32_t)((i / X) % Y), z = (uint32_t)(i / (X * Y));
      keys.push_back(x | (y << 11) | (z << 22));
    }
  }
  if (e->d_vac && (int32_t)keys.size() > e->vac_capacity) {
    HIP_TRY(hipFree(e->d_vac));
    e->d_vac = nullptr;
  }
  if (!e->d_vac) {
    e->vac_capacity = std::max<int32_t>((int32_t)keys.size(), 1);
    HIP_TRY(hipMalloc(&e->d_vac, (size_t)e->vac_capacity * sizeof(uint32_t)));
  }
  e->nvac = (int32_t)keys.size();
  if (e->d_sector && e->nvac > e->sector_cap) {
    HIP_TRY(hipFree(e->d_sector));
    e->d_sector = nullptr;
    e->sector_cap = 0;
  }
  if (!keys.empty())
    HIP_TRY(hipMemcpy(e->d_vac, keys.data(), keys.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  return AKMC_OK;
}

// Switches the engine to the device-memory sector state (plain_kernel_hbm),
// sized for every vacancy in one sector.
int enable_hbm_sectors(akmc_engine* e) {
  if (e->d_sector) return AKMC_OK;
  const int64_t cap = std::max<int64_t>(e->nvac, 1);
  int64_t off[9];
  akmc::plain_sector_layout(cap, off);
  HIP_TRY(hipMalloc(&e->d_sector, (size_t)off[8]));
  e->sector_cap = (int32_t)cap;
  return AKMC_OK;
}

void fill_stats(akmc_engine* e, akmc_stats* s) {
  const akmc::PlainState& st = e->h_state;
  s->events = st.events;
  s->step = st.step;
  s->sector_pos = st.sect;
  s->in_sector = st.in_sector;
  for (int i = 0; i < 8; i++) s->evolution[i] = st.evo[i];
  s->sector_threshold = st.thr;
  s->sector_time = st.t;
  s->n_eval = st.n_eval;
  s->n_fold = st.n_fold;
  s->n_shift = st.n_shift;
  s->sum_errors = st.sum_errors;
}

akmc::PlainArgs base_args(akmc_engine* e) {
  akmc::PlainArgs a{};
  a.sites = e->d_sites;
  a.vac_key = e->d_vac;
  a.nvac = e->nvac;
  a.dims = e->dims;
  a.state = e->d_state;
  a.time_steps = std::min(e->time_steps, e->step_limit);
  a.sector_T = e->p.sector_time;
  a.nu = e->p.attempt_freq;
  a.kT = e->kT;
  a.rates_sector = -1;
  a.check_sums = e->check_sums;
  a.select_mode = e->select_mode;
  a.sector_ws = e->d_sector;
  a.sector_cap = e->sector_cap;
  return a;
}

}  // namespace

extern "C" {

const char* akmc_last_error(void) { return g_err.c_str(); }

int akmc_params_default(akmc_params* p) {
  if (!p) return fail(AKMC_ERR_INVALID, "null params");
  memset(p, 0, sizeof(*p));
  p->box[0] = p->box[1] = p->box[2] = 80;
  p->lattice_const = 2.85532;
  p->cutoff_radius = 5.6;
  const int types[4] = {AKMC_TYPE_FE, AKMC_TYPE_CU, AKMC_TYPE_MN, AKMC_TYPE_NI};
  const int64_t ratio[4] = {9598, 134, 134, 134};
  p->n_alloy = 4;
  for (int i = 0; i < 4; i++) { p->alloy_type[i] = types[i]; p->alloy_ratio[i] = ratio[i]; }
  p->va_count = 1;
  p->seed_create_types = 100565;
  p->seed_create_vacancy = 16650;
  p->seed_event_selection = 563264;
  p->seed_time_inc = 594689;
  p->temperature = 673;
  p->physics_time = 8e4;
  p->attempt_freq = 6e12;
  p->sector_time = 3.0e-7;
  p->precision = AKMC_PREC_FP64;
  return AKMC_OK;
}

int akmc_params_from_yaml(const char* path, akmc_params* p) {
  if (!path || !p) return fail(AKMC_ERR_INVALID, "null argument");
  akmc_params_default(p);
  akmc::yaml::Node root;
  std::string err;
  if (!akmc::yaml::parse_file(path, &root, &err)) return fail(AKMC_ERR_CONFIG, "%s: %s", path, err.c_str());
  // parseBox (profiles/config/config_parsing.cpp:114-142)
  const akmc::yaml::Node* box = root.get("box");
  if (!box) return fail(AKMC_ERR_CONFIG, "box must be set in config");
  const akmc::yaml::Node* size = box->get("size");
  if (!size || !size->is_seq() || size->seq.size() != 3) return fail(AKMC_ERR_CONFIG, "box size must be 3.");
  for (int i = 0; i < 3; i++) p->box[i] = size->seq[i].as_int(0);
  if (!box->get("lattice_const")) return fail(AKMC_ERR_CONFIG, "lattice const must be set in config.");
  p->lattice_const = box->get("lattice_const")->as_double(0.0);
  if (!box->get("cutoff_radius")) return fail(AKMC_ERR_CONFIG, "cutoff radius must be set in config.");
  p->cutoff_radius = box->get("cutoff_radius")->as_double(0.0);
  // parseSim (:144-152)
  const akmc::yaml::Node* sim = root.get("simulation");
  if (!sim) return fail(AKMC_ERR_CONFIG, "simulation must be set in config");
  p->temperature = sim->get_double("temperature", 0.0);
  p->physics_time = sim->get_double("physics_time", 0.0);
  p->attempt_freq = sim->get_double("attempt_freq", 0.0);
  p->sector_time = sim->get_double("sector_time", 3.0e-7);  // extension; the reference hard-codes 3e-7
  // parseCreate (:150-206): random creation, else pipe creation, whose alloy
  // is still read from create.random.alloy (:175-177); a pipe places no
  // random vacancies (its file is read by akmc_pipe_sites).
  const akmc::yaml::Node* create = root.get("create");
  const akmc::yaml::Node* rnd = create ? create->get("random") : nullptr;
  const akmc::yaml::Node* pipe = create ? create->get("pipe") : nullptr;
  const bool random_on = rnd && rnd->get_bool("enable", false);
  const bool pipe_on = !random_on && pipe && pipe->get_bool("enable", false);
  if (!random_on && !pipe_on)
    return fail(AKMC_ERR_CONFIG, "no create option in config (create.random or create.pipe)");
  if (!rnd) return fail(AKMC_ERR_CONFIG, "alloy must be a map in config.");
  p->va_count = random_on ? rnd->get_int("va_count", 0) : 0;
  const akmc::yaml::Node* alloy = rnd->get("alloy");
  if (!alloy || !alloy->is_map()) return fail(AKMC_ERR_CONFIG, "alloy must be a map in config.");
  p->n_alloy = 0;
  for (const auto& kv : alloy->map) {
    const int t = lattype_from_name(kv.first);
    if (t < 0) return fail(AKMC_ERR_CONFIG, "unknown alloy key %s", kv.first.c_str());
    if (p->n_alloy >= AKMC_MAX_ALLOY) return fail(AKMC_ERR_CONFIG, "too many alloy keys");
    p->alloy_type[p->n_alloy] = t;
    p->alloy_ratio[p->n_alloy] = kv.second.as_int(0);
    p->n_alloy++;
  }
  // parseSeeds (:208-214)
  const akmc::yaml::Node* seeds = root.get("seeds");
  if (!seeds) return fail(AKMC_ERR_CONFIG, "seeds must be set in config");
  p->seed_create_types = seeds->get_int("create_types", 0);
  p->seed_create_vacancy = seeds->get_int("create_vacancy", 0);
  p->seed_event_selection = seeds->get_int("event_selection", 0);
  p->seed_time_inc = seeds->get_int("time_inc", 0);
  return AKMC_OK;
}

int akmc_create(const akmc_params* p, int device, akmc_engine** out) {
  if (!p || !out) return fail(AKMC_ERR_INVALID, "null argument");
  *out = nullptr;
  for (int d = 0; d < 3; d++)
    if (p->box[d] < 8) return fail(AKMC_ERR_INVALID, "box edge %lld < 8 cells is not supported", (long long)p->box[d]);
  if (2 * p->box[0] > 2048 || p->box[1] > 2047 || p->box[2] > 1023)
    return fail(AKMC_ERR_INVALID, "box exceeds 1024 x 2047 x 1023 cells");
  if (p->precision != AKMC_PREC_FP64) return fail(AKMC_ERR_INVALID, "the single-lattice engine runs FP64 rates");
  if (!(p->sector_time > 0)) return fail(AKMC_ERR_INVALID, "sector_time must be positive");
  if (p->n_alloy <= 0 || p->n_alloy > AKMC_MAX_ALLOY) return fail(AKMC_ERR_INVALID, "bad alloy list");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(AKMC_ERR_INVALID, "device %d out of range (%d devices)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  std::unique_ptr<akmc_engine> e(new akmc_engine());
  e->p = *p;
  e->device = device;
  e->dims = akmc::box_dims(p->box);
  e->nsites = (int64_t)e->dims.X * e->dims.Y * e->dims.Z;
  // SubLattice::startTimeLoop: time_steps = ceil(time_limit / T) (sublattice.inl:26)
  e->time_steps = (int64_t)ceil(p->physics_time / p->sector_time);
  // RatesSolver::arrhenius evaluates BoltzmannConstant * T first (rates_solver.cpp:25)
  e->kT = AKMC_KB * p->temperature;
  // seeds: SubLattice seeds time_inc_rng with the (double -> int64) config seed
  // (simulation.h:70-76, sublattice.cpp:17); mt19937::seed takes it mod 2^32.
  e->time_rng.seed((std::mt19937::result_type)(int64_t)(double)p->seed_time_inc);
  {
    const int rc = akmc_host::upload_stencil(device);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&e->ev0));
  HIP_TRY(hipEventCreate(&e->ev1));
  HIP_TRY(hipMalloc(&e->d_sites, (size_t)e->nsites));
  HIP_TRY(hipMemset(e->d_sites, 0, (size_t)e->nsites));
  HIP_TRY(hipMalloc(&e->d_state, sizeof(akmc::PlainState)));
  for (int b = 0; b < 2; b++) {
    HIP_TRY(hipHostMalloc(&e->h_u[b], kChunkMax * sizeof(double), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&e->h_l[b], kChunkMax * sizeof(double), hipHostMallocDefault));
    HIP_TRY(hipMalloc(&e->d_u[b], kChunkMax * sizeof(double)));
    HIP_TRY(hipMalloc(&e->d_l[b], kChunkMax * sizeof(double)));
  }
  memset(&e->h_state, 0, sizeof(e->h_state));
  HIP_TRY(hipMemcpy(e->d_state, &e->h_state, sizeof(e->h_state), hipMemcpyHostToDevice));
  *out = e.release();
  return AKMC_OK;
}

void akmc_destroy(akmc_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  (void)hipFree(e->d_sites);
  (void)hipFree(e->d_vac);
  (void)hipFree(e->d_state);
  (void)hipFree(e->d_log);
  (void)hipFree(e->d_sector);
  for (int b = 0; b < 2; b++) {
    (void)hipHostFree(e->h_u[b]);
    (void)hipHostFree(e->h_l[b]);
    (void)hipFree(e->d_u[b]);
    (void)hipFree(e->d_l[b]);
  }
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int64_t akmc_num_sites(const akmc_engine* e) { return e ? e->nsites : 0; }

int akmc_site_sectors(const int64_t box[3], const int64_t* sites, int64_t n, int8_t* sectors) {
  if (!box || (n > 0 && (!sites || !sectors)) || n < 0) return fail(AKMC_ERR_INVALID, "null argument");
  for (int d = 0; d < 3; d++)
    if (box[d] < 1) return fail(AKMC_ERR_INVALID, "box edge %lld < 1", (long long)box[d]);
  if (2 * box[0] > 2048 || box[1] > 2047 || box[2] > 1023)
    return fail(AKMC_ERR_INVALID, "box exceeds 1024 x 2047 x 1023 cells");
  const akmc::BoxDims d = akmc::box_dims(box);
  const int64_t ns = (int64_t)d.X * d.Y * d.Z;
  for (int64_t i = 0; i < n; i++) {
    if (sites[i] < 0 || sites[i] >= ns) return fail(AKMC_ERR_INVALID, "site %lld outside the box", (long long)sites[i]);
    sectors[i] = (int8_t)akmc::key_sector(d, akmc::key_of_index(d, sites[i]));
  }
  return AKMC_OK;
}

int akmc_random_sites(const akmc_params* pp, uint8_t* out, int64_t n) {
  if (!pp || !out) return fail(AKMC_ERR_INVALID, "null argument");
  const akmc_params& p = *pp;
  const int64_t nsites = 2 * p.box[0] * p.box[1] * p.box[2];
  if (n != nsites) return fail(AKMC_ERR_INVALID, "site count %lld != %lld", (long long)n, (long long)nsites);
  if (!(p.lattice_const > 0) || !(p.cutoff_radius > 0)) return fail(AKMC_ERR_INVALID, "bad lattice_const/cutoff");
  // Ghost-extended local grid (LatListMeta, lattice_list_meta.cpp:9-16):
  // ghost width g = ceil(cutoff / a0) (src/simulation.cpp:54-58).
  const int64_t g = (int64_t)ceil(p.cutoff_radius / p.lattice_const);
  const int64_t box_x = 2 * p.box[0], box_y = p.box[1], box_z = p.box[2];
  const int64_t gx = 2 * g, gy = g, gz = g;
  const int64_t size_x = box_x + 2 * gx, size_y = box_y + 2 * gy, size_z = box_z + 2 * gz;
  uint8_t* sites = out;
  memset(sites, 0, (size_t)n);
  // creation::createAtomsRandom (src/creation.cpp:18-94): one
  // uniform_int_distribution<>(1, sum of ratios) draw per ghost-extended site
  // in z, y, x order, mapped through LatticeTypes::randomAtomsType
  // (lattice_types.cpp:9-19).  Ghost sites become periodic images later, so
  // only real sites keep their draw.
  int64_t ratio_total = 0;
  for (int i = 0; i < p.n_alloy; i++) ratio_total += p.alloy_ratio[i];
  if (ratio_total <= 0) return fail(AKMC_ERR_INVALID, "alloy ratios sum to zero");
  int64_t seed_types = p.seed_create_types, seed_vac = p.seed_create_vacancy;
  std::random_device rd;  // seeds 0 = random device (profiles/seed_relocate.h:24-42)
  if (seed_types == 0) seed_types = rd();
  if (seed_vac == 0) seed_vac = rd();
  std::uniform_int_distribution<> types_dis(1, (int)ratio_total);
  std::mt19937 types_rng((std::mt19937::result_type)seed_types);
  uint8_t bit_of_index[AKMC_MAX_ALLOY];
  for (int i = 0; i < p.n_alloy; i++) {
    const int t = p.alloy_type[i];
    bit_of_index[i] = t == AKMC_TYPE_CU ? akmc::BIT_CU : t == AKMC_TYPE_MN ? akmc::BIT_MN
                      : t == AKMC_TYPE_NI ? akmc::BIT_NI : 0;  // Fe and dumbbell keys store nothing
  }
  for (int64_t z = 0; z < size_z; z++) {
    const bool zin = z >= gz && z < gz + box_z;
    for (int64_t y = 0; y < size_y; y++) {
      const bool yin = zin && y >= gy && y < gy + box_y;
      uint8_t* row = yin ? &sites[(size_t)(box_x * ((y - gy) + box_y * (z - gz)))] : nullptr;
      for (int64_t x = 0; x < size_x; x++) {
        const int64_t hit = types_dis(types_rng);
        int64_t rank_local = 0;
        int idx = -1;
        for (int i = 0; i < p.n_alloy; i++) {
          rank_local += p.alloy_ratio[i];
          if (rank_local >= hit) { idx = i; break; }
        }
        if (row && x >= gx && x < gx + box_x && idx >= 0) row[x - gx] = bit_of_index[idx];
      }
    }
  }
  // createRandom (:128-160): va_count vacancies at uniform_int_distribution
  // <long long>(0, box_x*box_y*box_z - 1) draws; only a Cu tag is removed when
  // a vacancy is placed (Ni/Mn tags stay).
  std::uniform_int_distribution<long long> va_dis(0, box_x * box_y * box_z - 1);
  std::mt19937 va_rng((std::mt19937::result_type)seed_vac);
  for (int64_t va_i = 0; va_i < p.va_count;) {
    const int64_t id = va_dis(va_rng);
    if (!(sites[(size_t)id] & akmc::BIT_V)) {
      sites[(size_t)id] = (uint8_t)((sites[(size_t)id] & ~akmc::BIT_CU) | akmc::BIT_V);
      va_i++;
    }
  }
  return AKMC_OK;
}

int akmc_init_random(akmc_engine* e) {
  if (!e) return fail(AKMC_ERR_INVALID, "null engine");
  std::vector<uint8_t> sites((size_t)e->nsites, 0);
  int rc = akmc_random_sites(&e->p, sites.data(), e->nsites);
  if (rc) return rc;
  return akmc_upload(e, sites.data(), e->nsites);
}

int akmc_upload(akmc_engine* e, const uint8_t* sites, int64_t n) {
  if (!e || !sites) return fail(AKMC_ERR_INVALID, "null argument");
  if (n != e->nsites) return fail(AKMC_ERR_INVALID, "site count %lld != %lld", (long long)n, (long long)e->nsites);
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipMemcpy(e->d_sites, sites, (size_t)n, hipMemcpyHostToDevice));
  int rc = upload_vacancies(e, sites);
  if (rc) return rc;
  e->lattice_ready = true;
  return AKMC_OK;
}

int akmc_download(akmc_engine* e, uint8_t* sites, int64_t n) {
  if (!e || !sites) return fail(AKMC_ERR_INVALID, "null argument");
  if (n != e->nsites) return fail(AKMC_ERR_INVALID, "site count %lld != %lld", (long long)n, (long long)e->nsites);
  HIP_TRY(hipSetDevice(e->device));
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipMemcpy(sites, e->d_sites, (size_t)n, hipMemcpyDeviceToHost));
  return AKMC_OK;
}

int akmc_set_event_log(akmc_engine* e, int64_t capacity) {
  if (!e || capacity < 0) return fail(AKMC_ERR_INVALID, "bad argument");
  HIP_TRY(hipSetDevice(e->device));
  if (e->d_log) { HIP_TRY(hipFree(e->d_log)); e->d_log = nullptr; }
  e->log_cap = capacity;
  e->log_base = e->h_state.events;
  if (capacity > 0) HIP_TRY(hipMalloc(&e->d_log, (size_t)capacity * 2 * sizeof(uint32_t)));
  return AKMC_OK;
}

int akmc_get_event_log(akmc_engine* e, int64_t* from, int64_t* to, int64_t cap, int64_t* n) {
  if (!e || !n) return fail(AKMC_ERR_INVALID, "null argument");
  HIP_TRY(hipSetDevice(e->device));
  const int64_t avail = std::min(e->h_state.events - e->log_base, e->log_cap);
  *n = avail;
  const int64_t m = std::min(avail, cap);
  if (m <= 0) return AKMC_OK;
  std::vector<uint32_t> buf((size_t)m * 2);
  HIP_TRY(hipMemcpy(buf.data(), e->d_log, buf.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
  const int64_t X = e->dims.X, Y = e->dims.Y;
  for (int64_t i = 0; i < m; i++) {
    const uint32_t kf = buf[2 * i], kt = buf[2 * i + 1];
    from[i] = (int64_t)(kf & 0x7FF) + X * ((int64_t)((kf >> 11) & 0x7FF) + Y * (int64_t)(kf >> 22));
    to[i] = (int64_t)(kt & 0x7FF) + X * ((int64_t)((kt >> 11) & 0x7FF) + Y * (int64_t)(kt >> 22));
  }
  return AKMC_OK;
}

int akmc_run(akmc_engine* e, int64_t max_events, akmc_stats* stats) {
  if (!e) return fail(AKMC_ERR_INVALID, "null engine");
  if (!e->lattice_ready) return fail(AKMC_ERR_INVALID, "no lattice: call akmc_init_random or akmc_upload first");
  if (max_events < 0) return fail(AKMC_ERR_INVALID, "max_events < 0");
  HIP_TRY(hipSetDevice(e->device));
  const auto t0 = std::chrono::steady_clock::now();
  // Saturating: any max_events up to INT64_MAX means "no cap" beyond it.
  const int64_t cap = max_events > INT64_MAX - e->h_state.events ? INT64_MAX : e->h_state.events + max_events;
  double ksec = 0.0;
  int64_t launches = 0;
  int stop = 0;
  for (;;) {
    const int64_t ev = e->h_state.events;
    // Make sure the chunk holding event `ev` is on the device.
    int buf = -1;
    for (int b = 0; b < 2; b++)
      if (e->chunk_base[b] >= 0 && ev >= e->chunk_base[b] && ev < e->chunk_base[b] + e->chunk) buf = b;
    if (buf < 0) {
      buf = e->chunk_base[0] < 0 ? 0 : (e->chunk_base[0] < e->chunk_base[1] ? 0 : 1);
      int rc = generate_chunk(e, buf);
      if (rc) return rc;
    }
    akmc::PlainArgs a = base_args(e);
    a.u_stream = e->d_u[buf];
    a.l_stream = e->d_l[buf];
    a.rng_base = e->chunk_base[buf];
    a.rng_end = e->chunk_base[buf] + e->chunk;
    a.max_events = cap;
    a.log = e->d_log;
    a.log_base = e->log_base;
    a.log_cap = e->log_cap;
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    if (akmc_plain_launch(&a, e->stream) != 0) return fail(AKMC_ERR_HIP, "event-loop launch failed");
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    launches++;
    // Overlap: produce the next chunk while this one is consumed.
    const int other = buf ^ 1;
    const bool need_next = e->chunk_base[other] < e->chunk_base[buf] || e->chunk_base[other] < 0;
    if (need_next && e->generated == e->chunk_base[buf] + e->chunk && cap > e->chunk_base[buf] + e->chunk) {
      int rc = generate_chunk(e, other);
      if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    ksec += ms * 1e-3;
    HIP_TRY(hipMemcpy(&e->h_state, e->d_state, sizeof(e->h_state), hipMemcpyDeviceToHost));
    const int st = e->h_state.status;
    if (st == AKMC_RUN_NEED_RNG || st == AKMC_RUN_YIELD) continue;
    if (st == AKMC_RUN_TIME_DONE) { stop = AKMC_STOP_TIME; break; }
    if (st == AKMC_RUN_EVENT_CAP) { stop = AKMC_STOP_EVENTS; break; }
    if (st == AKMC_RUN_ERR_NO_EVENT)
      return fail(AKMC_ERR_NO_EVENT, "no event selected at event %lld (u * sum >= sum; the reference asserts here)",
                  (long long)e->h_state.events);
    if (st == AKMC_RUN_ERR_CAPACITY && !e->d_sector) {
      // A sector outgrew the LDS form: continue with the device-memory one
      // (the launch stopped before the sector's first event).
      int rc = enable_hbm_sectors(e);
      if (rc) return rc;
      continue;
    }
    if (st == AKMC_RUN_ERR_CAPACITY)
      return fail(AKMC_ERR_CAPACITY, "a sector holds more than %d vacancies", e->sector_cap);
    if (st == AKMC_RUN_ERR_HELPER)
      return fail(AKMC_ERR_HIP, "event loop: a helper wavefront did not answer (event %lld)",
                  (long long)e->h_state.events);
    return fail(AKMC_ERR_HIP, "event loop returned status %d", st);
  }
  const auto t1 = std::chrono::steady_clock::now();
  e->last = akmc_stats{};
  fill_stats(e, &e->last);
  e->last.stop = stop;
  e->last.seconds = std::chrono::duration<double>(t1 - t0).count();
  e->last.kernel_seconds = ksec;
  e->last.launches = launches;
  if (stats) *stats = e->last;
  return AKMC_OK;
}

int akmc_set_rng_chunk(akmc_engine* e, int64_t events) {
  if (!e) return fail(AKMC_ERR_INVALID, "null engine");
  if (events < 1 || events > kChunkMax)
    return fail(AKMC_ERR_INVALID, "chunk must be in [1, %lld]", (long long)kChunkMax);
  if (e->generated != 0) return fail(AKMC_ERR_INVALID, "the chunk size can only be set before the first run");
  e->chunk = events;
  return AKMC_OK;
}

// Checkpoint file: magic, the sizes of the two records, akmc_params,
// PlainState, the site count and one byte per site.  The vacancy list is
// rebuilt from the sites (each sector's list is collected and sorted by key,
// so the list order does not matter) and the two random streams are replayed
// up to the next event's draws.
static const char kCkptMagic[8] = {'A', 'K', 'M', 'C', 'C', 'K', 'P', '1'};

int akmc_checkpoint_save(akmc_engine* e, const char* path) {
  if (!e || !path) return fail(AKMC_ERR_INVALID, "null argument");
  if (!e->lattice_ready) return fail(AKMC_ERR_INVALID, "no lattice to save");
  std::vector<uint8_t> sites((size_t)e->nsites);
  int rc = akmc_download(e, sites.data(), e->nsites);
  if (rc) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(AKMC_ERR_INVALID, "cannot open %s", path);
  const uint32_t sz[2] = {(uint32_t)sizeof(akmc_params), (uint32_t)sizeof(akmc::PlainState)};
  bool ok = fwrite(kCkptMagic, 1, 8, f) == 8 && fwrite(sz, sizeof(sz), 1, f) == 1 &&
            fwrite(&e->p, sizeof(e->p), 1, f) == 1 && fwrite(&e->h_state, sizeof(e->h_state), 1, f) == 1 &&
            fwrite(&e->nsites, sizeof(e->nsites), 1, f) == 1 &&
            fwrite(sites.data(), 1, sites.size(), f) == sites.size();
  ok = (fclose(f) == 0) && ok;
  return ok ? AKMC_OK : fail(AKMC_ERR_INVALID, "write failed: %s", path);
}

int akmc_checkpoint_load(const char* path, int device, akmc_engine** out) {
  if (!path || !out) return fail(AKMC_ERR_INVALID, "null argument");
  *out = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(AKMC_ERR_INVALID, "cannot open %s", path);
  char magic[8];
  uint32_t sz[2];
  akmc_params p;
  akmc::PlainState st;
  int64_t n = 0;
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kCkptMagic, 8) == 0 && fread(sz, sizeof(sz), 1, f) == 1 &&
            sz[0] == sizeof(akmc_params) && sz[1] == sizeof(akmc::PlainState) && fread(&p, sizeof(p), 1, f) == 1 &&
            fread(&st, sizeof(st), 1, f) == 1 && fread(&n, sizeof(n), 1, f) == 1 && n > 0;
  // The site count must be the one the stored box implies, and the file must
  // hold exactly that many bytes after the header, before anything is
  // allocated from it.
  if (ok) {
    for (int d = 0; d < 3; d++) ok = ok && p.box[d] >= 1 && p.box[d] <= 2048;
    ok = ok && n == 2 * p.box[0] * p.box[1] * p.box[2];
  }
  if (ok) {
    const long here = ftell(f);
    ok = here >= 0 && fseek(f, 0, SEEK_END) == 0 && ftell(f) - here == (long)n && fseek(f, here, SEEK_SET) == 0;
  }
  std::vector<uint8_t> sites;
  if (ok) {
    try {
      sites.resize((size_t)n);
    } catch (const std::exception&) {
      fclose(f);
      return fail(AKMC_ERR_INVALID, "%s: cannot hold %lld sites", path, (long long)n);
    }
    ok = fread(sites.data(), 1, sites.size(), f) == sites.size();
  }
  fclose(f);
  if (!ok) return fail(AKMC_ERR_INVALID, "%s is not a checkpoint of this build", path);
  akmc_engine* e = nullptr;
  int rc = akmc_create(&p, device, &e);
  if (rc) return rc;
  if (n != e->nsites || st.events < 0) {
    akmc_destroy(e);
    return fail(AKMC_ERR_INVALID, "%s: inconsistent checkpoint", path);
  }
  rc = akmc_upload(e, sites.data(), n);
  if (rc) { akmc_destroy(e); return rc; }
  st.status = AKMC_RUN_OK;
  e->h_state = st;
  if (hipMemcpy(e->d_state, &e->h_state, sizeof(e->h_state), hipMemcpyHostToDevice) != hipSuccess) {
    akmc_destroy(e);
    return fail(AKMC_ERR_HIP, "state upload failed");
  }
  // Replay the streams: one draw of each per event performed so far.
  for (int64_t i = 0; i < st.events; i++) {
    (void)e->sel_rng();
    (void)e->time_dis(e->time_rng);
  }
  e->generated = st.events;
  e->chunk_base[0] = e->chunk_base[1] = -1;
  fill_stats(e, &e->last);
  *out = e;
  return AKMC_OK;
}

int akmc_run_steps(akmc_engine* e, int64_t steps, int64_t max_events, akmc_stats* stats) {
  if (!e) return fail(AKMC_ERR_INVALID, "null engine");
  if (steps < 0) return fail(AKMC_ERR_INVALID, "steps < 0");
  const int64_t step0 = e->h_state.step;
  e->step_limit = steps > INT64_MAX - step0 ? INT64_MAX : step0 + steps;
  const int rc = akmc_run(e, max_events, nullptr);
  e->step_limit = INT64_MAX;
  if (rc != AKMC_OK) return rc;
  if (e->last.stop == AKMC_STOP_TIME && e->h_state.step < e->time_steps) e->last.stop = AKMC_STOP_STEPS;
  if (stats) *stats = e->last;
  return AKMC_OK;
}

int akmc_set_check_sums(akmc_engine* e, int on) {
  if (!e) return fail(AKMC_ERR_INVALID, "null engine");
  e->check_sums = on ? 1 : 0;
  return AKMC_OK;
}

int akmc_set_select_mode(akmc_engine* e, int mode) {
  if (!e) return fail(AKMC_ERR_INVALID, "null engine");
  if (mode < 0 || mode > 2) return fail(AKMC_ERR_INVALID, "select mode must be 0, 1 or 2");
  e->select_mode = mode;
  return AKMC_OK;
}

int akmc_rng_streams(int64_t seed_time_inc, int64_t first, int64_t n, double* u, double* l) {
  if (first < 0 || n < 0 || ((!u || !l) && n > 0)) return fail(AKMC_ERR_INVALID, "bad argument");
  std::mt19937 sel;
  std::mt19937 tr((std::mt19937::result_type)(int64_t)(double)seed_time_inc);
  std::uniform_real_distribution<double> dis(0.0, 1.0);
  for (int64_t i = 0; i < first; i++) {
    (void)sel();
    (void)dis(tr);
  }
  for (int64_t i = 0; i < n; i++) {
    u[i] = sel() * (1.0 / std::mt19937::max());
    l[i] = -std::log(dis(tr));
  }
  return AKMC_OK;
}

int akmc_sync(int device) {
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipDeviceSynchronize());
  return AKMC_OK;
}

int akmc_copy_bench(int device, int64_t bytes, int iters, double* seconds) {
  if (bytes < 16 || iters < 1) return fail(AKMC_ERR_INVALID, "bad argument");
  HIP_TRY(hipSetDevice(device));
  void *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, (size_t)bytes));
  HIP_TRY(hipMalloc(&b, (size_t)bytes));
  HIP_TRY(hipMemset(a, 1, (size_t)bytes));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = AKMC_OK;
  HIP_TRY(hipEventRecord(e0, nullptr));
  for (int i = 0; i < iters && rc == AKMC_OK; i++)
    if (akmc_copy_launch(a, b, bytes, nullptr) != 0) rc = fail(AKMC_ERR_HIP, "copy launch failed");
  HIP_TRY(hipEventRecord(e1, nullptr));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  if (seconds) *seconds = ms * 1e-3 / iters;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

int akmc_exp_host(const double* x, double* y, int64_t n) {
  if ((!x || !y) && n > 0) return fail(AKMC_ERR_INVALID, "null argument");
  for (int64_t i = 0; i < n; i++) y[i] = akmc::exp_glibc(x[i], akmc_exp_tab_init);
  return AKMC_OK;
}

int akmc_exp_device(int device, const double* x, double* y, int64_t n) {
  if ((!x || !y) && n > 0) return fail(AKMC_ERR_INVALID, "null argument");
  if (n <= 0) return AKMC_OK;
  HIP_TRY(hipSetDevice(device));
  double *dx = nullptr, *dy = nullptr;
  HIP_TRY(hipMalloc(&dx, (size_t)n * sizeof(double)));
  HIP_TRY(hipMalloc(&dy, (size_t)n * sizeof(double)));
  int rc = AKMC_OK;
  if (hipMemcpy(dx, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      akmc_exp_launch(dx, dy, n, nullptr) != 0 || hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(y, dy, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(AKMC_ERR_HIP, "device exp failed");
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

int akmc_get_stats(akmc_engine* e, akmc_stats* stats) {
  if (!e || !stats) return fail(AKMC_ERR_INVALID, "null argument");
  *stats = e->last;
  fill_stats(e, stats);
  return AKMC_OK;
}

int akmc_sector_rates(akmc_engine* e, int sector, int64_t cap, int64_t* sites, double* rates, double* total,
                      int64_t* n) {
  if (!e || !n || !total || sector < 0 || sector > 7 || cap < 0) return fail(AKMC_ERR_INVALID, "bad argument");
  if (!e->lattice_ready) return fail(AKMC_ERR_INVALID, "no lattice");
  HIP_TRY(hipSetDevice(e->device));
  const int32_t ocap = (int32_t)std::min<int64_t>(cap, std::max<int64_t>(akmc_plain_vcap(), e->nvac));
  uint32_t* d_keys = nullptr;
  double* d_rates = nullptr;
  double* d_total = nullptr;
  int32_t* d_n = nullptr;
  akmc::PlainState* d_tmp = nullptr;
  HIP_TRY(hipMalloc(&d_keys, std::max<int32_t>(ocap, 1) * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&d_rates, std::max<int32_t>(ocap, 1) * 8 * sizeof(double)));
  HIP_TRY(hipMalloc(&d_total, sizeof(double)));
  HIP_TRY(hipMalloc(&d_n, sizeof(int32_t)));
  HIP_TRY(hipMalloc(&d_tmp, sizeof(akmc::PlainState)));
  HIP_TRY(hipMemcpy(d_tmp, &e->h_state, sizeof(e->h_state), hipMemcpyHostToDevice));
  int rc = AKMC_OK;
  int32_t hn = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
  akmc::PlainArgs a = base_args(e);
  a.state = d_tmp;
  a.rates_sector = sector;
  a.max_events = 1;
  a.rng_base = 0;
  a.rng_end = 1;
  a.time_steps = INT64_MAX;
  a.out_keys = d_keys;
  a.out_rates = d_rates;
  a.out_total = d_total;
  a.out_n = d_n;
  a.out_cap = ocap;
  if (akmc_plain_launch(&a, e->stream) != 0) rc = fail(AKMC_ERR_HIP, "rates launch failed");
  hipError_t he = hipStreamSynchronize(e->stream);
  if (rc == AKMC_OK && he != hipSuccess) rc = fail(AKMC_ERR_HIP, "rates kernel: %s", hipGetErrorString(he));
  if (rc == AKMC_OK) {
    (void)hipMemcpy(&hn, d_n, sizeof(hn), hipMemcpyDeviceToHost);
    (void)hipMemcpy(total, d_total, sizeof(double), hipMemcpyDeviceToHost);
    if (hn == AKMC_RUN_ERR_CAPACITY && !e->d_sector && attempt == 0) {
      rc = enable_hbm_sectors(e);
      if (rc == AKMC_OK) continue;
    }
    if (rc == AKMC_OK && hn < 0)
      rc = fail(hn == AKMC_RUN_ERR_CAPACITY ? AKMC_ERR_CAPACITY : AKMC_ERR_HIP, "rates kernel status %d", hn);
  }
  break;
  }
  if (rc == AKMC_OK) {
    *n = hn;
    const int32_t m = std::min<int32_t>(hn, ocap);
    std::vector<uint32_t> keys((size_t)std::max(m, 1));
    if (m > 0) {
      (void)hipMemcpy(keys.data(), d_keys, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost);
      if (rates) (void)hipMemcpy(rates, d_rates, (size_t)m * 8 * sizeof(double), hipMemcpyDeviceToHost);
    }
    const int64_t X = e->dims.X, Y = e->dims.Y;
    for (int32_t i = 0; i < m && sites; i++) {
      const uint32_t k = keys[i];
      sites[i] = (int64_t)(k & 0x7FF) + X * ((int64_t)((k >> 11) & 0x7FF) + Y * (int64_t)(k >> 22));
    }
  }
  (void)hipFree(d_keys);
  (void)hipFree(d_rates);
  (void)hipFree(d_total);
  (void)hipFree(d_n);
  (void)hipFree(d_tmp);
  return rc;
}

}  // extern "C"
