/* akmc_sim.hpp -- the reference's host-side C++ shapes for the vacancy-hop
 * path, header-only (C++17) over the C ABI of akmc.h.
 *
 *   parseConfig           ConfigParsing::newInstance + parse
 *                         (profiles/config/config_parsing.cpp:54-230)
 *   EventHooks            src/hook/event_hooks.hpp:8-20
 *   ABVIModel             models/abvi/kmc.h: the vacancy model of one lattice,
 *                         in place of gpu_prepare / initialize_gpu (the
 *                         constructor) and gpu_final (the destructor) of
 *                         gpu/gpu_simulate.h:31-66
 *   simulation::simulate  src/simulation.h:61-85, the sublattice time loop
 *                         with the hooks after every step
 *
 * Errors are thrown as akmc_sim::Error (the reference exits the process).
 * The sector loop itself runs on the device: calcRates, select and perform
 * are not called per event from the host as in the reference's
 * ModelAdapter; calcRates(sector) evaluates a sector as its first event would,
 * as a parity hook. */
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "akmc.h"

namespace akmc_sim {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
  int code;
};

inline void check(int rc) {
  if (rc != AKMC_OK) throw Error(rc, akmc_last_error());
}

/* conf::ConfigValues (profiles/config/config_values.h:15-70), the parts this
 * engine reads. */
struct ConfigValues {
  akmc_params params{};
  akmc_output output{};
  bool create_pipe = false;        /* create.pipe is the create option */
  std::string pipe_input_box;      /* create.pipe.input_box */
};

inline ConfigValues parseConfig(const std::string& path) {
  ConfigValues c;
  check(akmc_params_from_yaml(path.c_str(), &c.params));
  check(akmc_output_from_yaml(path.c_str(), &c.output));
  std::vector<char> buf(4096);
  if (akmc_pipe_path_from_yaml(path.c_str(), buf.data(), (int64_t)buf.size()) == AKMC_OK) {
    c.create_pipe = true;
    c.pipe_input_box = buf.data();
  }
  return c;
}

/* dev_Vacancy (gpu/DeviceVacRatesSolver.h:7-14): one vacancy's 8 direction
 * rates; id is the real-site index x + 2 nx (y + ny z) (the reference keeps
 * its ghost-extended local id). */
struct dev_Vacancy {
  int64_t id;
  double rates[8];
  int valid;
};

/* Called after every finished time step (step counted from 0) and once when
 * the time loop has ended. */
class EventHooks {
 public:
  virtual ~EventHooks() = default;
  virtual void onStepFinished(unsigned long step) = 0;
  virtual void onAllDone() = 0;
};

/* One lattice of the abvi vacancy model, resident on one GPU. */
class ABVIModel {
 public:
  ABVIModel(const akmc_params& p, int device) : p_(p) { check(akmc_create(&p_, device, &e_)); }
  ~ABVIModel() {
    if (e_) akmc_destroy(e_);
  }
  ABVIModel(const ABVIModel&) = delete;
  ABVIModel& operator=(const ABVIModel&) = delete;

  /* creation::createRandom (src/creation.cpp:96-164). */
  void createRandom() { check(akmc_init_random(e_)); }

  /* creation::createFromPile (src/creation.cpp:184-230). */
  void createFromPile(const std::string& pipe_file) {
    std::vector<uint8_t> s((size_t)numSites());
    check(akmc_pipe_sites(&p_, pipe_file.c_str(), s.data(), (int64_t)s.size(), nullptr));
    check(akmc_upload(e_, s.data(), (int64_t)s.size()));
  }

  /* ABVIModel::calcRates for sector `sector` (models/abvi/kmc.cpp:68-115):
   * the total rate of the sector as its first event computes it, and, when
   * asked, its vacancies (real-site index) with their 8 rates each. */
  double calcRates(int sector, std::vector<int64_t>* sites = nullptr, std::vector<double>* rates = nullptr) {
    int64_t cap = 1 << 16, n = 0;
    double total = 0.0;
    for (;;) {
      std::vector<int64_t> s((size_t)cap);
      std::vector<double> r((size_t)cap * 8);
      check(akmc_sector_rates(e_, sector, cap, s.data(), r.data(), &total, &n));
      if (n <= cap) {
        s.resize((size_t)n);
        r.resize((size_t)n * 8);
        if (sites) *sites = std::move(s);
        if (rates) *rates = std::move(r);
        return total;
      }
      cap = n;
    }
  }

  /* ABVIModel::calcRatesGPU's result for one sector (models/abvi/kmc.cpp:
   * 118-179, first_calculate_GPU / calculate_GPU of gpu/gpu_simulate.h:33-35)
   * as the reference's device records:
   * every vacancy of the sector with its 8 rates; *total the sector total. */
  std::vector<dev_Vacancy> calcRatesGPU(int sector, double* total = nullptr) {
    std::vector<int64_t> s;
    std::vector<double> r;
    const double t = calcRates(sector, &s, &r);
    if (total) *total = t;
    std::vector<dev_Vacancy> v(s.size());
    for (size_t i = 0; i < s.size(); i++) {
      v[i].id = s[i];
      for (int b = 0; b < 8; b++) v[i].rates[b] = r[8 * i + b];
      v[i].valid = 1;
    }
    return v;
  }

  /* ModelAdapter::defectSize: the number of vacancies. */
  unsigned long defectSize() const {
    unsigned long c = 0;
    for (uint8_t f : lattice()) c += (f & 1u) ? 1 : 0;
    return c;
  }

  /* The occupancy: one flag byte per real site (bit 0 V, 1 Cu, 2 Mn, 3 Ni). */
  std::vector<uint8_t> lattice() const {
    std::vector<uint8_t> s((size_t)numSites());
    check(akmc_download(e_, s.data(), (int64_t)s.size()));
    return s;
  }

  akmc_stats stats() const {
    akmc_stats st;
    check(akmc_get_stats(e_, &st));
    return st;
  }

  int64_t numSites() const { return akmc_num_sites(e_); }
  const akmc_params& params() const { return p_; }
  akmc_engine* handle() { return e_; }

 private:
  akmc_params p_;
  akmc_engine* e_ = nullptr;
};

/* Itl (models/abvi/defect/itl_list.h): a dumbbell of the interstitial list,
 * with its orientation (orientation.h:12-30, -6..6 without 0), the directions
 * it can jump to and its 8 rate slots (two per orientation direction). */
struct Itl {
  int8_t orient = 0;
  uint8_t avail_trans_dir = 0;
  double rates[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

/* Itl::beforeRatesUpdate + Itl::updateRates with ItlRatesSolver::rate as the
 * callback (itl_list.cpp:25-50, rate/itl_rates_solver.cpp:16-80) for every
 * dumbbell of `list` (site index -> Itl, ascending sites as ItlList::mp).
 * types: one LatticeTypes value per real site of the periodic box. */
template <class Map>
inline void updateItlRates(const int64_t box[3], const std::vector<uint16_t>& types, Map& list, double attempt_freq,
                           double temperature, int device) {
  std::vector<int64_t> sites;
  std::vector<int8_t> ori;
  sites.reserve(list.size());
  ori.reserve(list.size());
  for (const auto& kv : list) {
    sites.push_back((int64_t)kv.first);
    ori.push_back(kv.second.orient);
  }
  std::vector<uint8_t> avail(sites.size());
  std::vector<double> rates(8 * sites.size());
  check(akmc_itl_rates(box, types.data(), (int64_t)sites.size(), sites.data(), ori.data(), attempt_freq, temperature,
                       device, avail.data(), rates.data(), nullptr));
  size_t k = 0;
  for (auto& kv : list) {
    kv.second.avail_trans_dir = avail[k];
    for (int b = 0; b < 8; b++) kv.second.rates[b] = rates[8 * k + b];
    k++;
  }
}

namespace simulation {

/* simulation::simulate (src/simulation.h:61-85): runs the sublattice time
 * loop of `time_limit` seconds (ceil(time_limit / T) steps,
 * sublattice.inl:26) and calls p_event_hooks->onStepFinished(step) after
 * every step s with s % hook_steps == 0 (every step for 1, as the reference
 * does; its own hooks act on multiples of their intervals only), then
 * onAllDone().  The engine stops at those step boundaries only.  The model
 * was created with its seed and time limit, so seed_time_inc and time_limit
 * must be the ones of its parameters. */
inline akmc_stats simulate(ABVIModel* p_model, EventHooks* p_event_hooks, const double seed_time_inc,
                           const double time_limit, const long hook_steps = 1) {
  const akmc_params& p = p_model->params();
  if ((int64_t)seed_time_inc != p.seed_time_inc || time_limit != p.physics_time || hook_steps < 1)
    throw Error(AKMC_ERR_INVALID, "simulate: seed, time limit or hook_steps differ from the model's");
  const int64_t total = (int64_t)std::ceil(time_limit / p.sector_time);
  akmc_stats st = p_model->stats();
  while (st.step < total) {
    const int64_t s0 = st.step;  // steps finished so far; s0 is the next to finish
    int64_t h = (s0 + hook_steps - 1) / hook_steps * hook_steps;
    if (h > total - 1) h = total - 1;
    check(akmc_run_steps(p_model->handle(), h + 1 - s0, INT64_MAX, &st));
    if (st.step != h + 1) throw Error(AKMC_ERR_INVALID, "simulate: the time loop stopped early");
    if (p_event_hooks && h % hook_steps == 0) p_event_hooks->onStepFinished((unsigned long)h);
  }
  if (p_event_hooks) p_event_hooks->onAllDone();
  return st;
}

}  // namespace simulation
}  // namespace akmc_sim
