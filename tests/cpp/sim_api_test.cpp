// Drives the C++ host interface of include/akmc_sim.hpp as the reference's
// PKMC::onCreate / onSimulation would (profiles/pkmc.cpp:148-231); the
// Python test tests/test_cpp_api.py builds it with g++ and checks its output
// against the oracle.
//   sim_api_test config <config.yaml>
//   sim_api_test simulate <config.yaml> <hook_steps> <sites_out>
//   sim_api_test itl
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "akmc_sim.hpp"

namespace {

class RecordingHooks : public akmc_sim::EventHooks {
 public:
  std::vector<unsigned long> steps;
  int all_done = 0;
  void onStepFinished(unsigned long step) override { steps.push_back(step); }
  void onAllDone() override { all_done++; }
};

int config(const char* path) {
  const akmc_sim::ConfigValues c = akmc_sim::parseConfig(path);
  const akmc_params& p = c.params;
  printf("box=%lld,%lld,%lld\n", (long long)p.box[0], (long long)p.box[1], (long long)p.box[2]);
  printf("n_alloy=%d\nva_count=%lld\n", p.n_alloy, (long long)p.va_count);
  printf("temperature=%.17g\nphysics_time=%.17g\nsector_time=%.17g\n", p.temperature, p.physics_time,
         p.sector_time);
  printf("seed_time_inc=%lld\n", (long long)p.seed_time_inc);
  printf("create_pipe=%d\npipe_input_box=%s\n", c.create_pipe ? 1 : 0, c.pipe_input_box.c_str());
  printf("dump_interval=%lld\n", (long long)c.output.dump_interval);
  return 0;
}

int simulate(const char* path, long hook_steps, const char* out) {
  const akmc_sim::ConfigValues c = akmc_sim::parseConfig(path);
  akmc_sim::ABVIModel model(c.params, 0);
  if (c.create_pipe)
    model.createFromPile(c.pipe_input_box);
  else
    model.createRandom();
  printf("defects=%lu\n", model.defectSize());
  RecordingHooks hooks;
  const akmc_stats st = akmc_sim::simulation::simulate(&model, &hooks, (double)c.params.seed_time_inc,
                                                       c.params.physics_time, hook_steps);
  printf("events=%lld\nstep=%lld\n", (long long)st.events, (long long)st.step);
  printf("hook_steps=");
  for (size_t i = 0; i < hooks.steps.size(); i++) printf(i ? ",%lu" : "%lu", hooks.steps[i]);
  printf("\nall_done=%d\n", hooks.all_done);
  const std::vector<uint8_t> s = model.lattice();
  FILE* f = fopen(out, "wb");
  if (!f || fwrite(s.data(), 1, s.size(), f) != s.size()) return 3;
  fclose(f);
  // A wrong seed is rejected before anything runs.
  try {
    akmc_sim::simulation::simulate(&model, &hooks, (double)c.params.seed_time_inc + 1, c.params.physics_time);
    printf("seed_check=missed\n");
  } catch (const akmc_sim::Error& e) {
    printf("seed_check=ok\n");
  }
  std::vector<int64_t> sites;
  std::vector<double> rates;
  const double total = model.calcRates(0, &sites, &rates);
  printf("sector0_vacancies=%zu\nsector0_total=%.17g\n", sites.size(), total);
  double total2 = 0.0;
  const std::vector<akmc_sim::dev_Vacancy> dv = model.calcRatesGPU(0, &total2);
  bool same = dv.size() == sites.size() && total2 == total;
  for (size_t i = 0; same && i < dv.size(); i++) {
    same = dv[i].id == sites[i] && dv[i].valid == 1;
    for (int b = 0; b < 8; b++) same = same && dv[i].rates[b] == rates[8 * i + b];
  }
  printf("dev_vacancy_check=%s\n", same ? "ok" : "differs");
  return 0;
}

// Dumbbell rates through the C++ layer: FeFe dumbbells in pure Fe, one per
// orientation, far apart.  Each jump leaves the bond sums unchanged (the same
// two sums in the other order) and the dumbbell energy zero, so every rate
// slot is nu exp(-0.32 / kB T) / 2.
int itl() {
  const int64_t box[3] = {24, 6, 6};
  const int64_t X = 2 * box[0];
  std::vector<uint16_t> types((size_t)(X * box[1] * box[2]), 1);
  std::map<int64_t, akmc_sim::Itl> list;
  const int8_t oris[6] = {1, -2, 3, -4, 5, -6};
  for (int i = 0; i < 6; i++) {
    const int64_t s = 8 * i + X * (2 + box[1] * 3);
    types[(size_t)s] = 0x0101;
    list[s].orient = oris[i];
  }
  akmc_sim::updateItlRates(box, types, list, 6e12, 573.0, 0);
  const double want = 6e12 * std::exp(-0.32 / (8.617330350E-5 * 573.0)) / 2;
  int bad = 0;
  for (const auto& kv : list) {
    bad += __builtin_popcount(kv.second.avail_trans_dir) != 4;
    for (int b = 0; b < 8; b++) bad += kv.second.rates[b] != want;
  }
  printf("itl_dumbbells=%zu\nitl_check=%s\n", list.size(), bad ? "differs" : "ok");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc == 3 && std::string(argv[1]) == "config") return config(argv[2]);
    if (argc == 2 && std::string(argv[1]) == "itl") return itl();
    if (argc == 5 && std::string(argv[1]) == "simulate") return simulate(argv[2], atol(argv[3]), argv[4]);
  } catch (const akmc_sim::Error& e) {
    printf("error=%d %s\n", e.code, e.what());
    return 2;
  }
  fprintf(stderr, "usage: sim_api_test config <yaml> | simulate <yaml> <hook_steps> <sites_out>\n");
  return 1;
}
