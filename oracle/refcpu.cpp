// refcpu -- CPU restatement of the reference's vacancy-hop AKMC CPU build.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/, the smoke
// check in __graft_entry__.py and bench.py's cpu_baseline leg may load it, and
// only as the checker or the timed CPU baseline.  The product library never
// links or calls it.
//
// It restates the reference CPU build (-DKMC_HIP_ENABLE_FLAG=OFF, KMC_RAND=MT,
// one MPI rank) with the reference's own data structures: occupancy lives in
// four hash containers (vac/cu/mn/ni), every event re-sweeps the whole sector
// through those hashes, and the C library's exp/log are used.  Each function
// cites the reference lines it follows.  Single-rank ghost sites are periodic
// images of the box (ghost init + per-sector sync of src/pack/*); by default
// that is expressed as an index wrap on the ghost-extended coordinates
// (wrapId).  The ghost_images form keeps real ghost copies and the packers'
// bookkeeping instead (sector_sync); tests/test_ghost_oracle.py shows that the
// two give the same trajectories.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <array>
#include <algorithm>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace refcpu {

// LatticeTypes::lat_type values (src/lattice/lattice_types.h:20-36).
enum LatType : int { V = 0, Fe = 1, Cu = 2, Ni = 4, Mn = 8 };

// PairBond::makeBond key (models/abvi/rate/bonds/pair_bond.hpp:58-65).
static inline int make_bond(int a, int b) {
  if (a > b) { int t = a; a = b; b = t; }
  return (a << 8) | b;
}

// Bond tables (models/abvi/rate/bonds/bonds_counter.cpp:13-33), in eV.
static double bond1(int key) {
  switch (key) {
  case 0x0101: return -0.778; case 0x0000: return 0.315; case 0x0202: return -0.581;
  case 0x0404: return -0.793; case 0x0808: return -0.438;
  case 0x0001: return -0.161; case 0x0102: return -0.609; case 0x0104: return -0.821;
  case 0x0108: return -0.648;
  case 0x0002: return -0.103; case 0x0004: return -0.234; case 0x0008: return -0.151;
  case 0x0204: return -0.692; case 0x0208: return -0.519; case 0x0408: return -0.831;
  }
  fprintf(stderr, "refcpu: bad 1nn bond key %x\n", key);
  abort();
}
static double bond2(int key) {
  switch (key) {
  case 0x0101: return -0.389; case 0x0000: return -0.214; case 0x0202: return -0.389;
  case 0x0404: return -0.389; case 0x0808: return -0.389;
  case 0x0001: return -0.161; case 0x0102: return -0.344; case 0x0104: return -0.399;
  case 0x0108: return -0.364;
  case 0x0002: return -0.206; case 0x0004: return -0.351; case 0x0008: return -0.206;
  case 0x0204: return -0.344; case 0x0208: return -0.249; case 0x0408: return -0.464;
  }
  fprintf(stderr, "refcpu: bad 2nn bond key %x\n", key);
  abort();
}

// VacRatesSolver::e0 (models/abvi/rate/vacancy_rates_solver.cpp:14-27).
static double e0(int ghost_atom) {
  switch (ghost_atom) {
  case Fe: return 0.62;
  case Cu: return 0.54;
  case Ni: return 0.68;
  case Mn: return 1.03;
  default: return 0;
  }
}

// models/abvi/rate/rates_types.h:16
static const double kBoltzmann = 8.617330350E-5;

struct VacancyHash {  // models/abvi/defect/vac_hash.h:11-20
  double nn_rates[8];
  unsigned char avail_trans_dir;
};

struct Params {
  int64_t box[3];           // cells (x not doubled)
  int64_t ghost;            // ghost width in cells = ceil(cutoff / a0)
  std::vector<int> types;   // alloy keys in YAML order (lat_type values)
  std::vector<int64_t> ratio;
  int64_t va_count;
  int64_t seed_create_types, seed_create_vacancy, seed_time_inc;
  double temperature, attempt_freq, sector_T, physics_time;
  uint32_t mask_2nn;        // 0x2F in the periodic lattice list
};

class RefCpu {
 public:
  explicit RefCpu(const Params& p) : P(p) {
    // LatListMeta (src/lattice/lattice_list_meta.cpp:9-16), BCC_DBX = 2.
    box_x = 2 * p.box[0]; box_y = p.box[1]; box_z = p.box[2];
    ghost_x = 2 * p.ghost; ghost_y = p.ghost; ghost_z = p.ghost;
    size_x = box_x + 2 * ghost_x; size_y = box_y + 2 * ghost_y; size_z = box_z + 2 * ghost_z;
  }

  // LatListMeta::getLId (src/lattice/lattice_list_meta.h:103-105).
  int64_t getId(int64_t x, int64_t y, int64_t z) const { return x + size_x * (y + z * size_y); }
  void coord(int64_t id, int64_t* x, int64_t* y, int64_t* z) const {
    *x = id % size_x; id /= size_x; *y = id % size_y; *z = id / size_y;
  }
  // Ghost coordinates resolve to their periodic image inside the box.
  static int64_t wrap1(int64_t c, int64_t g, int64_t n) {
    int64_t r = (c - g) % n;
    if (r < 0) r += n;
    return r + g;
  }
  int64_t wrapId(int64_t x, int64_t y, int64_t z) const {
    if (ghost_images) {
      // Ghost-image form: ids address the ghost-extended grid directly, as
      // LatListMeta::getIdBy1nnOffset does (lattice_list_meta.cpp:22-70).
      if (x < 0 || x >= size_x || y < 0 || y >= size_y || z < 0 || z >= size_z) {
        fprintf(stderr, "refcpu: stencil leaves the ghost-extended grid\n");
        abort();
      }
      return getId(x, y, z);
    }
    return getId(wrap1(x, ghost_x, box_x), wrap1(y, ghost_y, box_y), wrap1(z, ghost_z, box_z));
  }
  // The real site a ghost-extended id is an image of.
  int64_t imageId(int64_t id) const {
    int64_t x, y, z;
    coord(id, &x, &y, &z);
    return getId(wrap1(x, ghost_x, box_x), wrap1(y, ghost_y, box_y), wrap1(z, ghost_z, box_z));
  }
  bool isGhost(int64_t id) const { return imageId(id) != id; }

  // PeriodLatticeList::getnnlattice readout order V > Cu > Mn > Ni > Fe
  // (src/lattice/period_lattice_list.cpp:307-324).
  int readout(int64_t id) const {
    if (vac_hash.find(id) != vac_hash.end()) return V;
    if (cu_hash.find(id) != cu_hash.end()) return Cu;
    if (mn_hash.find(id) != mn_hash.end()) return Mn;
    if (ni_hash.find(id) != ni_hash.end()) return Ni;
    return Fe;
  }

  // 1nn ids of (x,y,z), PeriodLatticeList::get1nn2 (period_lattice_list.cpp:45-100).
  void nn1(int64_t x, int64_t y, int64_t z, int64_t out[8]) const {
    if (x % 2 == 0) {
      out[0] = wrapId(x - 1, y - 1, z - 1); out[1] = wrapId(x - 1, y - 1, z);
      out[2] = wrapId(x - 1, y, z - 1);     out[3] = wrapId(x - 1, y, z);
      out[4] = wrapId(x + 1, y - 1, z - 1); out[5] = wrapId(x + 1, y - 1, z);
      out[6] = wrapId(x + 1, y, z - 1);     out[7] = wrapId(x + 1, y, z);
    } else {
      out[0] = wrapId(x - 1, y, z);         out[1] = wrapId(x - 1, y, z + 1);
      out[2] = wrapId(x - 1, y + 1, z);     out[3] = wrapId(x - 1, y + 1, z + 1);
      out[4] = wrapId(x + 1, y, z);         out[5] = wrapId(x + 1, y, z + 1);
      out[6] = wrapId(x + 1, y + 1, z);     out[7] = wrapId(x + 1, y + 1, z + 1);
    }
  }
  // 2nn ids, PeriodLatticeList::get2nn2 (period_lattice_list.cpp:234-256).
  void nn2(int64_t x, int64_t y, int64_t z, int64_t out[6]) const {
    out[0] = wrapId(x - 2, y, z); out[1] = wrapId(x, y - 1, z); out[2] = wrapId(x, y, z - 1);
    out[3] = wrapId(x, y, z + 1); out[4] = wrapId(x, y + 1, z); out[5] = wrapId(x + 2, y, z);
  }

  // BondsCounter::count (models/abvi/rate/bonds/bonds_counter.cpp:35-81) and
  // count_exchange (:83-129, get1nn3/get2nn3 substitute `atom` at source_id).
  double count_impl(int64_t now_id, int src_atom_type, int64_t source_id, int atom, bool exchange) const {
    int64_t x, y, z;
    coord(now_id, &x, &y, &z);
    int64_t n1[8], n2[6];
    nn1(x, y, z, n1);
    double energy = 0.0;
    for (int b = 0; b < 8; b++) {  // get1nnStatus = 0xFF
      int t = (exchange && n1[b] == source_id) ? atom : readout(n1[b]);
      energy += bond1(make_bond(t, src_atom_type));
    }
    nn2(x, y, z, n2);
    for (int b = 0; b < 6; b++) {
      if ((P.mask_2nn >> b) & 1) {
        int t = (exchange && n2[b] == source_id) ? atom : readout(n2[b]);
        energy += bond2(make_bond(t, src_atom_type));
      }
    }
    return energy;
  }
  double count(int64_t id, int type) const { return count_impl(id, type, -1, 0, false); }
  double count_exchange(int64_t now_id, int64_t source_id, int atom, int src_atom_type) const {
    return count_impl(now_id, src_atom_type, source_id, atom, true);
  }

  std::unordered_set<int64_t>* species_hash(int t) {
    switch (t) { case Cu: return &cu_hash; case Mn: return &mn_hash; case Ni: return &ni_hash; }
    return nullptr;
  }

  // VacRatesSolver::deltaE (models/abvi/rate/vacancy_rates_solver.cpp:29-172),
  // including the temporary hash exchange and exchange back.
  double deltaE(int64_t source_id, int64_t target_id, int target_type, int ghost_atom) {
    double e_before = 0;
    {
      double e_src = count(source_id, V);
      double e_des = count(target_id, target_type);
      e_before = e_src + e_des;
    }
    vac_hash.emplace(target_id, VacancyHash{});  // source keeps its V entry
    std::unordered_set<int64_t>* h = species_hash(ghost_atom);
    if (h) {
      auto it = h->find(target_id);
      if (it == h->end()) { fprintf(stderr, "refcpu: deltaE missing species\n"); abort(); }
      h->erase(it);
      h->emplace(source_id);
    }
    double e_after = 0;
    {
      double e_src = count(source_id, ghost_atom);
      double e_des = count_exchange(target_id, source_id, ghost_atom, V);
      e_after = e_src + e_des;
    }
    vac_hash.erase(vac_hash.find(target_id));
    if (h) {
      auto it = h->find(source_id);
      if (it == h->end()) { fprintf(stderr, "refcpu: deltaE missing species (back)\n"); abort(); }
      h->erase(it);
      h->emplace(target_id);
    }
    return e_after - e_before;
  }

  // RatesSolver::rate and arrhenius (models/abvi/rate/rates_solver.cpp:13-26).
  double rate(int64_t source_id, int64_t target_id, int ghost_atom) {
    const double active_energy = e0(ghost_atom) + deltaE(source_id, target_id, ghost_atom, ghost_atom) / 2;
    return P.attempt_freq * exp(-active_energy / (kBoltzmann * P.temperature));
  }

  // Sector region: local_sector_region[s] of the colored domain (low half
  // floor(n/2) cells, bit0 = x high, bit1 = y high, bit2 = z high, shifted by
  // the ghost width; described at src/simulation.cpp:91-110).
  void region(int s, int64_t lo[3], int64_t hi[3]) const {
    const int64_t n[3] = {P.box[0], P.box[1], P.box[2]};
    for (int d = 0; d < 3; d++) {
      const int64_t half = n[d] / 2;
      const bool high = (s >> d) & 1;
      lo[d] = P.ghost + (high ? half : 0);
      hi[d] = P.ghost + (high ? n[d] : half);
    }
  }

  // The sector's vacancies in sweep order, z, then y, then x in [2 x_lo,
  // 2 x_hi) (kmc.cpp:72-75): f(id, x, y, z, record) returns false to stop.
  // The reference visits every site of the sector through the hash
  // (list_sweep false); list_sweep visits only the vacancies, sorted by id,
  // which is the same sequence (ids ascend in z, y, x) at a fraction of the
  // cost, for writing long fixtures.
  template <class F>
  void sweep_vacancies(int s, F f) {
    int64_t lo[3], hi[3];
    region(s, lo, hi);
    if (!list_sweep) {
      for (int64_t z = lo[2]; z < hi[2]; z++)
        for (int64_t y = lo[1]; y < hi[1]; y++)
          for (int64_t x = 2 * lo[0]; x < 2 * hi[0]; x++) {
            const int64_t id = getId(x, y, z);
            auto it = vac_hash.find(id);
            if (it == vac_hash.end()) continue;
            if (!f(id, x, y, z, it->second)) return;
          }
      return;
    }
    std::vector<int64_t> ids;
    for (auto& kv : vac_hash) {
      int64_t x, y, z;
      coord(kv.first, &x, &y, &z);
      if (z >= lo[2] && z < hi[2] && y >= lo[1] && y < hi[1] && x >= 2 * lo[0] && x < 2 * hi[0])
        ids.push_back(kv.first);
    }
    std::sort(ids.begin(), ids.end());
    for (int64_t id : ids) {
      int64_t x, y, z;
      coord(id, &x, &y, &z);
      if (!f(id, x, y, z, vac_hash.at(id))) return;
    }
  }

  // ABVIModel::calcRates (models/abvi/kmc.cpp:68-115) with
  // VacancyHash::beforeRatesUpdate2/updateRates2 (vac_hash.cpp:7-46).
  double calcRates(int s) {
    double sum_rates = 0;
    sweep_vacancies(s, [&](int64_t id, int64_t x, int64_t y, int64_t z, VacancyHash& vh) {
      int64_t n1[8];
      int t1[8];
      nn1(x, y, z, n1);
      for (int b = 0; b < 8; b++) t1[b] = readout(n1[b]);
      for (int b = 0; b < 8; b++) vh.nn_rates[b] = 0;
      unsigned char avail = 0;
      for (int b = 0; b < 8; b++)
        if (t1[b] != V) avail |= (unsigned char)(1u << b);
      vh.avail_trans_dir = avail;
      for (int b = 0; b < 8; b++) {
        if ((avail >> b) & 1) {
          const double r = rate(id, n1[b], t1[b]);
          sum_rates += r;
          vh.nn_rates[b] = r;
        }
      }
      return true;
    });
    sum_rates += 0.0;  // defectGenRate(): env::global_env.defect_gen_rate == 0
    return sum_rates;
  }

  // ABVIModel::select (models/abvi/kmc.cpp:242-309): first (vacancy, dir) in
  // sweep order whose running sum is > excepted_rate.
  bool select(int s, double excepted_rate, int64_t* from, int64_t* to, int* dir) {
    double acc = 0.0;
    bool found = false;
    sweep_vacancies(s, [&](int64_t id, int64_t x, int64_t y, int64_t z, VacancyHash& vh) {
      for (int b = 0; b < 8; b++) {
        acc += vh.nn_rates[b];
        if (acc > excepted_rate) {
          int64_t n1[8];
          nn1(x, y, z, n1);  // LatListMeta::getIdBy1nnOffset
          *from = id; *to = n1[b]; *dir = b;
          found = true;
          return false;
        }
      }
      return true;
    });
    return found;  // not found: the reference returns event::DefectGen, which perform() asserts on
  }

  // ABVIModel::perform (models/abvi/kmc.cpp:358-396), vacancy branch.
  void perform(int64_t from, int64_t to) {
    auto it = vac_hash.find(from);
    if (it == vac_hash.end()) { fprintf(stderr, "refcpu: perform without vacancy\n"); abort(); }
    vac_hash.erase(it);
    vac_hash.emplace(to, VacancyHash{});
    auto ic = cu_hash.find(to), im = mn_hash.find(to), in = ni_hash.find(to);
    if (ic != cu_hash.end()) { cu_hash.erase(ic); cu_hash.emplace(from); }
    else if (im != mn_hash.end()) { mn_hash.erase(im); mn_hash.emplace(from); }
    else if (in != ni_hash.end()) { ni_hash.erase(in); ni_hash.emplace(from); }
    if (ghost_images) {
      // addExchange_ghost / addExchange_surface (kmc.cpp:385-392): a hop into
      // a ghost site is sent to its owner after the sector; changed real
      // sites refresh their ghost copies.
      if (isGhost(to)) exchange_ghost.push_back(to);
      else changed.insert(to);
      changed.insert(from);
    }
  }

  // ---- Ghost-image form of one rank (src/pack/*).  Ghost sites hold copies
  // of their periodic images; the copies a hop changed are pushed to the
  // owner after the sector, and every ghost copy of a changed real site is
  // refreshed before the next sector.

  // Sets the flags of `id` to `type` the way the packers' receive handlers do
  // (ghost_init_packer.cpp:281-357, sim_sync_packer.cpp:169-208,
  // ghost_sync_packer.cpp:285-318): erase the first of V, Cu, Mn, Ni that is
  // present, then insert `type` (nothing for Fe).
  void receive_type(int64_t id, int type) {
    auto iv = vac_hash.find(id);
    if (iv != vac_hash.end()) vac_hash.erase(iv);
    else if (cu_hash.erase(id)) {}
    else if (mn_hash.erase(id)) {}
    else ni_hash.erase(id);
    if (type == V) vac_hash.emplace(id, VacancyHash{});
    else if (auto* h = species_hash(type)) h->emplace(id);
  }
  // Every ghost copy of real site `id` receives its readout.
  void refresh_copies(int64_t id) {
    int64_t x, y, z;
    coord(id, &x, &y, &z);
    const int t = readout(id);
    const int64_t c[3] = {x, y, z}, n[3] = {box_x, box_y, box_z}, s[3] = {size_x, size_y, size_z};
    int64_t opts[3][3];
    int cnt[3];
    for (int d = 0; d < 3; d++) {
      cnt[d] = 0;
      for (int k = -1; k <= 1; k++) {
        const int64_t v = c[d] + k * n[d];
        if (v >= 0 && v < s[d]) opts[d][cnt[d]++] = v;
      }
    }
    for (int i = 0; i < cnt[0]; i++)
      for (int j = 0; j < cnt[1]; j++)
        for (int k = 0; k < cnt[2]; k++) {
          const int64_t g = getId(opts[0][i], opts[1][j], opts[2][k]);
          if (g != id) receive_type(g, t);
        }
  }
  // GhostInitPacker (ghost_init_packer.cpp:107-179, 266-403): every ghost
  // site receives the readout of its image.
  void ghost_init() {
    for (int64_t id = 0; id < size_x * size_y * size_z; id++)
      if (!isGhost(id)) refresh_copies(id);
  }
  // syncSimRegions then syncNextSectorGhostRegions (sublattice.inl:161-163).
  void sector_sync() {
    if (!ghost_sync) return;
    for (int64_t g : exchange_ghost) {  // SimSyncPacker::onReceive2: the ghost's type is V
      const int64_t img = imageId(g);
      receive_type(img, readout(g));
      changed.insert(img);
    }
    exchange_ghost.clear();
    for (int64_t id : changed) refresh_copies(id);
    changed.clear();
  }

  // LatticeTypes::randomAtomsType (src/lattice/lattice_types.cpp:9-19).
  int randomAtomsType(int64_t hit) const {
    int64_t rank_local = 0;
    for (size_t i = 0; i < P.types.size(); i++) {
      rank_local += P.ratio[i];
      if (rank_local >= hit) return P.types[i];
    }
    return Fe;
  }

  // creation::createAtomsRandom + createRandom (src/creation.cpp:18-164).  One
  // type draw per ghost-extended site in z,y,x order; ghost sites are later
  // overwritten by their periodic images, so only real sites are stored.
  void createRandom() {
    int64_t ratio_total = 0;
    for (size_t i = 0; i < P.types.size(); i++) ratio_total += P.ratio[i];
    std::uniform_int_distribution<> types_dis(1, (int)ratio_total);
    std::mt19937 types_rng((std::mt19937::result_type)P.seed_create_types);
    for (int64_t z = 0; z < size_z; z++)
      for (int64_t y = 0; y < size_y; y++)
        for (int64_t x = 0; x < size_x; x++) {
          const int64_t hit = types_dis(types_rng);
          const int t = randomAtomsType(hit);
          const bool real = x >= ghost_x && x < ghost_x + box_x && y >= ghost_y && y < ghost_y + box_y &&
                            z >= ghost_z && z < ghost_z + box_z;
          if (!real && !ghost_images) continue;  // wrap form: ghost draws are overwritten by images
          const int64_t id = getId(x, y, z);
          if (t == Cu) cu_hash.emplace(id);
          else if (t == Ni) ni_hash.emplace(id);
          else if (t == Mn) mn_hash.emplace(id);
        }
    const int64_t max_lattice_size = box_x * box_y * box_z;
    std::uniform_int_distribution<long long> va_dis(0, max_lattice_size - 1);
    std::mt19937 va_rng((std::mt19937::result_type)P.seed_create_vacancy);
    for (int64_t va_i = 0; va_i < P.va_count;) {
      int64_t local_id = va_dis(va_rng);
      const int64_t x = local_id % box_x;
      local_id = local_id / box_x;
      const int64_t y = local_id % box_y;
      const int64_t z = local_id / box_y;
      const int64_t id = getId(x + ghost_x, y + ghost_y, z + ghost_z);
      if (vac_hash.find(id) == vac_hash.end()) {
        auto ic = cu_hash.find(id);
        if (ic != cu_hash.end()) cu_hash.erase(ic);
        vac_hash.emplace(id, VacancyHash{});
        va_i++;
      }
    }
    if (ghost_images) ghost_init();
  }

  // SubLattice::startTimeLoop (src/algorithms/sl/sublattice.inl:25-190),
  // sector ring order 0,7,2,5,3,4,1,6 (sublattice.h:21-30), time stream from
  // SubLattice's constructor (sublattice.cpp:10-18).  Runs until the time
  // loop ends or max_events events have been performed (mid-sector stop).
  int run(int64_t max_events) {
    static const int ring[8] = {0, 7, 2, 5, 3, 4, 1, 6};
    const int64_t time_steps = (int64_t)ceil(P.physics_time / P.sector_T);
    for (; step < time_steps; step++) {
      for (; sect < 8; sect++) {
        const int s = ring[sect];
        if (!in_sector) {
          thr = static_cast<double>(step + 1) * P.sector_T - evolution[s];
          sector_time = 0.0;
          in_sector = true;
        }
        while (sector_time < thr) {
          if (events >= max_events) return 1;
          const double total_rates = calcRates(s);
          if (total_rates == 0.0 || std::abs(total_rates) < std::numeric_limits<double>::epsilon()) {
            sector_time = thr;
          } else {
            // ModelAdapter::selectAndPerform (src/models/model_adapter.h:95-101),
            // r::random() = rand32() * (1.0 / mt19937::max()) (random.cpp:35-41).
            const double u = sel_rng() * (1.0 / std::mt19937::max());
            int64_t from, to;
            int dir;
            if (!select(s, u * total_rates, &from, &to, &dir)) return -2;
            perform(from, to);
            if (log_from) { log_from->push_back(from); log_to->push_back(to); }
            const double rnd = time_inc_dis(time_inc_rng);
            const double delta_t = -std::log(rnd) / total_rates;
            sector_time += delta_t;
            events++;
          }
        }
        evolution[s] += sector_time;
        in_sector = false;
        if (ghost_images) sector_sync();
      }
      sect = 0;
    }
    return 0;
  }

  void seed_time() { time_inc_rng.seed((std::mt19937::result_type)P.seed_time_inc); }

  Params P;
  int64_t box_x, box_y, box_z, ghost_x, ghost_y, ghost_z, size_x, size_y, size_z;
  std::unordered_map<int64_t, VacancyHash> vac_hash;
  std::unordered_set<int64_t> cu_hash, mn_hash, ni_hash;
  std::mt19937 sel_rng;  // r::rng: default-constructed, never seeded (rand_generators.cpp:8)
  std::mt19937 time_inc_rng;
  std::uniform_real_distribution<double> time_inc_dis{0, 1};
  double evolution[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t step = 0;
  int sect = 0;
  bool in_sector = false;
  double thr = 0, sector_time = 0;
  int64_t events = 0;
  std::vector<int64_t>* log_from = nullptr;
  std::vector<int64_t>* log_to = nullptr;
  bool list_sweep = false;               // sweep_vacancies: the vacancy list instead of every site
  // Ghost-image form (set before createRandom): see sector_sync().
  bool ghost_images = false;
  bool ghost_sync = true;                // false only in a test that plants a missing sync
  std::vector<int64_t> exchange_ghost;   // ghost sites hopped into during this sector
  std::unordered_set<int64_t> changed;   // real sites changed since their copies were refreshed
};

}  // namespace refcpu

// ---------------------------------------------------------------- C ABI
// Flag byte written by ref_get_sites(): bit0 V, bit1 Cu, bit2 Mn, bit3 Ni
// (membership in vac/cu/mn/ni hashes); one byte per real site in
// x + 2N*(y + N*z) order with x doubled.
extern "C" {

struct ref_params {
  int64_t box[3];
  int64_t ghost;
  int32_t n_types;
  int32_t types[32];
  int64_t ratio[32];
  int64_t va_count;
  int64_t seed_create_types, seed_create_vacancy, seed_time_inc;
  double temperature, attempt_freq, sector_T, physics_time;
  uint32_t mask_2nn;
};

void* ref_create(const ref_params* p) {
  refcpu::Params P;
  for (int d = 0; d < 3; d++) P.box[d] = p->box[d];
  P.ghost = p->ghost;
  for (int i = 0; i < p->n_types; i++) { P.types.push_back(p->types[i]); P.ratio.push_back(p->ratio[i]); }
  P.va_count = p->va_count;
  P.seed_create_types = p->seed_create_types;
  P.seed_create_vacancy = p->seed_create_vacancy;
  P.seed_time_inc = p->seed_time_inc;
  P.temperature = p->temperature;
  P.attempt_freq = p->attempt_freq;
  P.sector_T = p->sector_T;
  P.physics_time = p->physics_time;
  P.mask_2nn = p->mask_2nn;
  auto* r = new refcpu::RefCpu(P);
  r->seed_time();
  return r;
}

void ref_destroy(void* h) { delete (refcpu::RefCpu*)h; }
void ref_create_random(void* h) { ((refcpu::RefCpu*)h)->createRandom(); }

int64_t ref_n_sites(void* h) {
  auto* r = (refcpu::RefCpu*)h;
  return r->box_x * r->box_y * r->box_z;
}

static int64_t real_to_local(refcpu::RefCpu* r, int64_t i) {
  const int64_t x = i % r->box_x, y = (i / r->box_x) % r->box_y, z = i / (r->box_x * r->box_y);
  return r->getId(x + r->ghost_x, y + r->ghost_y, z + r->ghost_z);
}
static int64_t local_to_real(refcpu::RefCpu* r, int64_t id) {
  int64_t x, y, z;
  r->coord(r->imageId(id), &x, &y, &z);  // a ghost id names its image
  return (x - r->ghost_x) + r->box_x * ((y - r->ghost_y) + r->box_y * (z - r->ghost_z));
}

void ref_get_sites(void* h, uint8_t* out) {
  auto* r = (refcpu::RefCpu*)h;
  const int64_t n = ref_n_sites(h);
  memset(out, 0, (size_t)n);
  for (auto& kv : r->vac_hash) if (!r->isGhost(kv.first)) out[local_to_real(r, kv.first)] |= 1;
  for (auto id : r->cu_hash) if (!r->isGhost(id)) out[local_to_real(r, id)] |= 2;
  for (auto id : r->mn_hash) if (!r->isGhost(id)) out[local_to_real(r, id)] |= 4;
  for (auto id : r->ni_hash) if (!r->isGhost(id)) out[local_to_real(r, id)] |= 8;
  // Ghost-image form stopped inside a sector: the owner's sites as the
  // sector-end push will leave them (erase the first flag present, set V).
  for (int64_t g : r->exchange_ghost) {
    uint8_t& f = out[local_to_real(r, g)];
    const uint8_t first = (f & 1) ? 1 : (f & 2) ? 2 : (f & 4) ? 4 : (f & 8) ? 8 : 0;
    f = (uint8_t)((f & ~first) | 1);
  }
}

// Vacancy-list sweep on or off (the same sweep order; see sweep_vacancies).
void ref_set_list_sweep(void* h, int on) { ((refcpu::RefCpu*)h)->list_sweep = on != 0; }

// Ghost-image form on or off (before ref_create_random / ref_set_sites).
void ref_set_ghost_images(void* h, int on) { ((refcpu::RefCpu*)h)->ghost_images = on != 0; }
// Sector-end push and refresh of the ghost form (on by default; a test
// switches it off to show that the comparison detects a missing sync).
void ref_set_ghost_sync(void* h, int on) { ((refcpu::RefCpu*)h)->ghost_sync = on != 0; }

// Flags of every ghost-extended site, x + size_x*(y + size_y*z) (size_x =
// 2(N + 2g)): real sites and, in the ghost-image form, the ghost copies.
int64_t ref_get_ext_sites(void* h, uint8_t* out, int64_t cap) {
  auto* r = (refcpu::RefCpu*)h;
  const int64_t n = r->size_x * r->size_y * r->size_z;
  if (cap < n) return -n;
  memset(out, 0, (size_t)n);
  for (auto& kv : r->vac_hash) out[kv.first] |= 1;
  for (auto id : r->cu_hash) out[id] |= 2;
  for (auto id : r->mn_hash) out[id] |= 4;
  for (auto id : r->ni_hash) out[id] |= 8;
  return n;
}

// Saves / restores the whole oracle state (sites with every tag, both random
// engines and the time-stream distribution, the sector-loop position and
// clocks), so that a long fixture run can resume.  Wrap form only.
void ref_set_sites(void* h, const uint8_t* in);

int ref_save(void* h, const char* path) {
  auto* r = (refcpu::RefCpu*)h;
  if (r->ghost_images) return -1;
  std::ofstream f(path);
  if (!f) return -1;
  f.precision(17);
  f << r->sel_rng << "\n" << r->time_inc_rng << "\n" << r->time_inc_dis << "\n";
  f << r->step << " " << r->sect << " " << (r->in_sector ? 1 : 0) << " " << r->events << "\n";
  f << std::hexfloat << r->thr << " " << r->sector_time;
  for (int i = 0; i < 8; i++) f << " " << r->evolution[i];
  f << std::defaultfloat << "\n";
  const int64_t n = ref_n_sites(h);
  std::vector<uint8_t> sites((size_t)n);
  ref_get_sites(h, sites.data());
  for (int64_t i = 0; i < n; i++)
    if (sites[(size_t)i]) f << i << " " << (int)sites[(size_t)i] << "\n";
  f << -1 << "\n";
  return f.good() ? 0 : -1;
}

int ref_load(void* h, const char* path) {
  auto* r = (refcpu::RefCpu*)h;
  std::ifstream f(path);
  if (!f) return -1;
  int ins = 0;
  f >> r->sel_rng >> r->time_inc_rng >> r->time_inc_dis;
  f >> r->step >> r->sect >> ins >> r->events;
  std::string t;
  f >> t; r->thr = std::strtod(t.c_str(), nullptr);
  f >> t; r->sector_time = std::strtod(t.c_str(), nullptr);
  for (int i = 0; i < 8; i++) { f >> t; r->evolution[i] = std::strtod(t.c_str(), nullptr); }
  r->in_sector = ins != 0;
  const int64_t n = ref_n_sites(h);
  std::vector<uint8_t> sites((size_t)n, 0);
  for (;;) {
    int64_t i = -1;
    int v = 0;
    if (!(f >> i) || i < 0) break;
    f >> v;
    if (i >= n) return -1;
    sites[(size_t)i] = (uint8_t)v;
  }
  if (!f.good() && !f.eof()) return -1;
  ref_set_sites(h, sites.data());
  return 0;
}

// local_sector_region[s] of the oracle (ghost-extended cell coordinates, x
// not doubled).
void ref_region(void* h, int s, int64_t lo[3], int64_t hi[3]) { ((refcpu::RefCpu*)h)->region(s, lo, hi); }

void ref_set_sites(void* h, const uint8_t* in) {
  auto* r = (refcpu::RefCpu*)h;
  r->vac_hash.clear(); r->cu_hash.clear(); r->mn_hash.clear(); r->ni_hash.clear();
  const int64_t n = ref_n_sites(h);
  for (int64_t i = 0; i < n; i++) {
    if (!in[i]) continue;
    const int64_t id = real_to_local(r, i);
    if (in[i] & 1) r->vac_hash.emplace(id, refcpu::VacancyHash{});
    if (in[i] & 2) r->cu_hash.emplace(id);
    if (in[i] & 4) r->mn_hash.emplace(id);
    if (in[i] & 8) r->ni_hash.emplace(id);
  }
  if (r->ghost_images) r->ghost_init();
}

// Runs up to max_events; logs (from, to) real-site indices when log arrays
// are given (each of capacity max_events).  Returns 0 when the time loop
// ended, 1 when the event cap stopped it, -2 when no event was found
// (u * sum >= sum, the reference's DefectGen assert).
int ref_run(void* h, int64_t max_events, int64_t* log_from, int64_t* log_to) {
  auto* r = (refcpu::RefCpu*)h;
  std::vector<int64_t> lf, lt;
  if (log_from) { r->log_from = &lf; r->log_to = &lt; }
  const int rc = r->run(r->events + max_events);
  if (log_from) {
    for (size_t i = 0; i < lf.size(); i++) { log_from[i] = local_to_real(r, lf[i]); log_to[i] = local_to_real(r, lt[i]); }
  }
  r->log_from = r->log_to = nullptr;
  return rc;
}

// Clocks: evolution[8] by sector id, then step, ring position, in-sector flag,
// current threshold and partial sector time, and the event count.
void ref_state(void* h, double evolution[8], int64_t* step, int32_t* sect, int32_t* in_sector, double* thr,
               double* sector_time, int64_t* events) {
  auto* r = (refcpu::RefCpu*)h;
  for (int i = 0; i < 8; i++) evolution[i] = r->evolution[i];
  *step = r->step; *sect = r->sect; *in_sector = r->in_sector ? 1 : 0;
  *thr = r->thr; *sector_time = r->sector_time; *events = r->events;
}

// Rates of every vacancy of sector s as computed by calcRates (which applies
// its side effects); out holds 8 doubles per vacancy in sweep order, sites
// the vacancy real-site indices.  Returns the vacancy count (capacity cap)
// and writes the sector total.
int64_t ref_sector_rates(void* h, int s, int64_t cap, int64_t* sites, double* out, double* total) {
  auto* r = (refcpu::RefCpu*)h;
  *total = r->calcRates(s);
  int64_t lo[3], hi[3];
  r->region(s, lo, hi);
  int64_t n = 0;
  for (int64_t z = lo[2]; z < hi[2]; z++)
    for (int64_t y = lo[1]; y < hi[1]; y++)
      for (int64_t x = 2 * lo[0]; x < 2 * hi[0]; x++) {
        const int64_t id = r->getId(x, y, z);
        auto it = r->vac_hash.find(id);
        if (it == r->vac_hash.end()) continue;
        if (n < cap) {
          sites[n] = local_to_real(r, id);
          for (int b = 0; b < 8; b++) out[8 * n + b] = it->second.nn_rates[b];
        }
        n++;
      }
  return n;
}

// deltaE of one (source vacancy, direction) pair through the hash path.
double ref_delta_e(void* h, int64_t real_src, int dir) {
  auto* r = (refcpu::RefCpu*)h;
  const int64_t src = real_to_local(r, real_src);
  int64_t x, y, z, n1[8];
  r->coord(src, &x, &y, &z);
  r->nn1(x, y, z, n1);
  const int t = r->readout(n1[dir]);
  return r->deltaE(src, n1[dir], t, t);
}

// BondsCounter::count of a real site holding `type` (lat_type value).
double ref_count(void* h, int64_t real_site, int type) {
  auto* r = (refcpu::RefCpu*)h;
  return r->count(real_to_local(r, real_site), type);
}

// BondsCounter::count_exchange: neighbours of now_site, with source_site read
// as `atom`, bonded to src_atom_type.
double ref_count_exchange(void* h, int64_t now_site, int64_t source_site, int atom, int src_atom_type) {
  auto* r = (refcpu::RefCpu*)h;
  return r->count_exchange(real_to_local(r, now_site), real_to_local(r, source_site), atom, src_atom_type);
}

int ref_random_atoms_type(void* h, int64_t hit) { return ((refcpu::RefCpu*)h)->randomAtomsType(hit); }
int ref_make_bond(int a, int b) { return refcpu::make_bond(a, b); }
double ref_bond1(int key) { return refcpu::bond1(key); }
double ref_bond2(int key) { return refcpu::bond2(key); }

// The C library's exp over an array (what RatesSolver::arrhenius calls).
void ref_exp_array(const double* x, double* y, int64_t n) {
  for (int64_t i = 0; i < n; i++) y[i] = exp(x[i]);
}

// The two random numbers the time loop consumes per event, in order:
// u = r::random() on a default-constructed mt19937 (random.cpp:35-41) and
// l = -log(r) with r from uniform_real_distribution<double>(0,1) on
// mt19937(seed_time_inc) (sublattice.cpp:12-17, sublattice.inl:118-120).
void ref_streams(int64_t seed_time_inc, int64_t n, double* u, double* l) {
  std::mt19937 sel;
  std::mt19937 tr((std::mt19937::result_type)seed_time_inc);
  std::uniform_real_distribution<double> dis(0, 1);
  for (int64_t i = 0; i < n; i++) {
    u[i] = sel() * (1.0 / std::mt19937::max());
    l[i] = -std::log(dis(tr));
  }
}

double ref_rate(void* h, int64_t real_src, int dir) {
  auto* r = (refcpu::RefCpu*)h;
  const int64_t src = real_to_local(r, real_src);
  int64_t x, y, z, n1[8];
  r->coord(src, &x, &y, &z);
  r->nn1(x, y, z, n1);
  const int t = r->readout(n1[dir]);
  return r->rate(src, n1[dir], t);
}

}  // extern "C"

// ============================================================ dumbbells (SIA)
// The abvi self-interstitial (dumbbell) model, restated on a typed periodic
// lattice: one LatticeTypes value per real site (the reference's array
// lattice, src/lattice/lattices_list.h:345), dumbbells listed in ascending
// site order.  ItlList::mp is a std::map keyed by lattice id
// (models/abvi/defect/itl_list.h), and ascending local id on the
// ghost-extended grid is ascending real index x + 2N(y + N z).  The
// reference's hash-set lattice no longer allocates that array
// (lattices_list.cpp:13), so this part of the reference cannot run as
// shipped; its semantics are those of the sources cited below.
namespace refitl {

using refcpu::bond1;
using refcpu::bond2;
using refcpu::make_bond;
using refcpu::Fe;
using refcpu::Cu;
using refcpu::Ni;
using refcpu::Mn;

// LatticeTypes predicates and combinations (src/lattice/lattice_types.h:48-95,
// lattice_types.cpp:21-48).
static inline bool is_dumbbell(int t) { return t > 0x0008; }
static inline bool is_atom(int t) { return t <= 0x0008 && t != 0; }
static inline int high_end(int t) { return t >> 8; }
static inline int low_end(int t) { return t & 0xFF; }
static inline int combine_to_inter(int a, int b) {
  if (a > b) { const int t = a; a = b; b = t; }
  return (a << 8) | b;
}
static inline int diff(int t, int b) {
  if (low_end(t) == b) return high_end(t);
  if (high_end(t) == b) return low_end(t);
  return t;
}

// orientation (models/abvi/orientation.h:12-30; orientation.cc:7-21 jump
// table, :143-236 direction masks).  o in [-6, 6], 0 = unknown.
static const char kJump[13][8] = {
    {'_', 'Y', '_', 'Y', 'X', '_', 'X', '_'},  // -6
    {'Y', '_', 'Y', '_', '_', 'X', '_', 'X'},  // -5
    {'_', 'Y', 'X', '_', '_', 'Y', 'X', '_'},  // -4
    {'Y', '_', '_', 'X', 'Y', '_', '_', 'X'},  // -3
    {'_', '_', 'Y', 'Y', 'X', 'X', '_', '_'},  // -2
    {'Y', 'Y', '_', '_', '_', '_', 'X', 'X'},  // -1
    {'_', '_', '_', '_', '_', '_', '_', '_'},  // 0
    {'X', 'X', '_', '_', '_', '_', 'Y', 'Y'},  // 1
    {'_', '_', 'X', 'X', 'Y', 'Y', '_', '_'},  // 2
    {'X', '_', '_', 'Y', 'X', '_', '_', 'Y'},  // 3
    {'_', 'X', 'Y', '_', '_', 'X', 'Y', '_'},  // 4
    {'X', '_', 'X', '_', '_', 'Y', '_', 'Y'},  // 5
    {'_', 'X', '_', 'X', 'Y', '_', 'Y', '_'},  // 6
};
static uint8_t avail_all(int o) {
  switch (o) {
  case 1: case -1: return 0xC3;
  case 2: case -2: return 0x3C;
  case 3: case -3: return 0x99;
  case 4: case -4: return 0x66;
  case 5: case -5: return 0xA5;
  case 6: case -6: return 0x5A;
  default: return 0;
  }
}
static uint8_t avail_high(int o) {
  switch (o) {
  case 1: return 0x03; case -1: return 0xC0;
  case 2: return 0x0C; case -2: return 0x30;
  case 3: return 0x11; case -3: return 0x88;
  case 4: return 0x22; case -4: return 0x44;
  case 5: return 0x05; case -5: return 0xA0;
  case 6: return 0x0A; case -6: return 0x50;
  default: return 0;
  }
}
static uint8_t avail_low(int o) {
  switch (o) {
  case 1: return 0xC0; case -1: return 0x03;
  case 2: return 0x30; case -2: return 0x0C;
  case 3: return 0x88; case -3: return 0x11;
  case 4: return 0x44; case -4: return 0x22;
  case 5: return 0xA0; case -5: return 0x05;
  case 6: return 0x50; case -6: return 0x0A;
  default: return 0;
  }
}
// orientation::tranAtom (orientation.cc:238-253): the atom of dumbbell type t
// that jumps towards 1nn tag b; t itself when neither may.
static int tran_atom(int o, int t, int b) {
  switch (kJump[o + 6][b]) {
  case 'X': return high_end(t);
  case 'Y': return low_end(t);
  default: return t;
  }
}
// orientation::random (orientation.h:38-42) from one 32-bit draw r:
// rand32(0, 12) = r % 12, then 1..12 -> +r/2 for even, -(r+1)/2 for odd.
static int orientation_from_draw(uint32_t r) {
  const uint32_t v = r % 12u + 1u;
  return v % 2 == 0 ? (int)(v / 2) : -(int)((v + 1) / 2);
}

// Itl (models/abvi/defect/itl_list.cpp).
// availTranDirs (:9-23): orientation directions whose 1nn is a single atom.
static uint8_t itl_avail(int o, uint8_t status_1nn, const int nb[8]) {
  uint8_t atoms = 0;
  for (int b = 0; b < 8; b++)
    if (((status_1nn >> b) & 1) && is_atom(nb[b])) atoms |= (uint8_t)(1 << b);
  return (uint8_t)(avail_all(o) & atoms);
}
// ratesIndex (:67-74): two slots per orientation direction, in direction order.
static int rates_index(int b, uint8_t dirs, bool up) {
  int below = 0;
  for (int k = 0; k < b; k++) below += (dirs >> k) & 1;
  return up ? 2 * below + 1 : 2 * below;
}
// get1nnIdByRatesIndex (:76-91).
static int first_nn_by_rates_index(int idx, uint8_t dirs) {
  idx /= 2;
  int i = 0;
  for (int b = 0; b < 8; b++)
    if ((dirs >> b) & 1) {
      if (i == idx) return b;
      i++;
    }
  return 0;
}
// getRatesStatus (:52-65).
static uint8_t rates_status(int o, uint8_t avail) {
  uint8_t st = 0;
  const uint8_t dirs = avail_all(o);
  int i = 0;
  for (int b = 0; b < 8; b++)
    if ((dirs >> b) & 1) {
      if ((avail >> b) & 1) st |= (uint8_t)((1 << i) | (1 << (i + 1)));
      i += 2;
    }
  return st;
}

// The typed periodic lattice.
struct Typed {
  int64_t X, Y, Z;     // 2 nx, ny, nz
  uint16_t* t;         // one LatticeTypes value per real site
  int64_t wrap_site(int64_t s, int dx, int dy, int dz) const {
    int64_t x = s % X, y = (s / X) % Y, z = s / (X * Y);
    x = ((x + dx) % X + X) % X;
    y = ((y + dy) % Y + Y) % Y;
    z = ((z + dz) % Z + Z) % Z;
    return x + X * (y + Y * z);
  }
  // get1nn / get2nn order (src/lattice/period_lattice_list.cpp:20-44,223-232).
  int64_t nn1(int64_t s, int b) const {
    const int par = (int)((s % X) & 1);
    const int lo = par ? 0 : -1;
    return wrap_site(s, (b & 4) ? 1 : -1, lo + ((b >> 1) & 1), lo + (b & 1));
  }
  int64_t nn2(int64_t s, int k) const {
    static const int o[6][3] = {{-2, 0, 0}, {0, -1, 0}, {0, 0, -1}, {0, 0, 1}, {0, 1, 0}, {2, 0, 0}};
    return wrap_site(s, o[k][0], o[k][1], o[k][2]);
  }
};
static const uint8_t kStatus1 = 0xFF, kStatus2 = 0x2F;  // period_lattice_list.cpp:10-18

// BondsCounter::count (models/abvi/rate/bonds/bonds_counter.cpp:35-81) on the
// typed lattice: dumbbell neighbours carry no pair bond.
static double count(const Typed& L, int64_t s, int src) {
  double e = 0.0;
  for (int b = 0; b < 8; b++)
    if ((kStatus1 >> b) & 1) {
      const int t = L.t[L.nn1(s, b)];
      if (is_atom(t) || t == 0) e += bond1(make_bond(t, src));
    }
  for (int b = 0; b < 6; b++)
    if ((kStatus2 >> b) & 1) {
      const int t = L.t[L.nn2(s, b)];
      if (is_atom(t) || t == 0) e += bond2(make_bond(t, src));
    }
  return e;
}

struct Itl {
  int64_t site;
  int o;
  uint8_t avail;
  double rates[8];
};

// bond::FeX_comp (models/abvi/rate/bonds/energy_dumbbell.cpp:139-206).
static double fex_comp(const Typed& L, int type, const Itl& itl, const int n1[8], const int n2[6]) {
  double ecomp = 0, eb = 0;
  const uint8_t comp = itl.avail & avail_low(itl.o);
  const int x = low_end(type);
  for (int b = 0; b < 8; b++) {
    if (!((comp >> b) & 1)) continue;
    const int t = n1[b];
    switch (x) {
    case Cu:
      if (is_dumbbell(t)) eb += 0.70;
      else if (t == Cu) ecomp += -0.18;
      else if (t == Ni) ecomp += -0.35;
      else if (t == Mn) ecomp += -0.46;
      break;
    case Ni:
      if (is_dumbbell(t)) eb += 0.70;
      else if (t == Cu) ecomp += -0.21;
      else if (t == Ni) ecomp += -0.32;
      else if (t == Mn) ecomp += -0.30;
      break;
    case Mn:
      if (is_dumbbell(t)) eb += 0.70;
      else if (t == Cu) ecomp += 0.28;
      else if (t == Ni) ecomp += 0.29;
      else if (t == Mn) ecomp += 0.28;
      break;
    default:
      break;
    }
  }
  (void)L;
  for (int b = 0; b < 8; b++)
    if (b < 6 && ((kStatus2 >> b) & 1) && is_dumbbell(n2[b])) eb += 0.48;
  return ecomp + eb / 2;
}

// bond::Edumb (energy_dumbbell.cpp:7-137): the dumbbell energy of the whole
// list, a left fold in list order.
static double edumb(const Typed& L, const std::vector<Itl>& list) {
  double edumb = 0;
  for (const Itl& itl : list) {
    const int type = L.t[itl.site];
    int n1[8], n2[6];
    for (int b = 0; b < 8; b++) n1[b] = L.t[L.nn1(itl.site, b)];
    for (int b = 0; b < 6; b++) n2[b] = L.t[L.nn2(itl.site, b)];
    auto contribute = [&]() {
      double c = 0;
      for (int b = 0; b < 8; b++) {
        if (((kStatus1 >> b) & 1) && is_dumbbell(n1[b])) c += 0.70;
        if (b < 6 && ((kStatus2 >> b) & 1) && is_dumbbell(n2[b])) c += 0.48;
      }
      return c;
    };
    double emix = 0, etens = 0, ecomp = 0, eb = 0;
    switch (type) {
    case 0x0102: emix = -0.46; ecomp = fex_comp(L, type, itl, n1, n2); break;
    case 0x0202: emix = -0.36; eb += contribute(); break;
    case 0x0104: emix = -0.19; ecomp = fex_comp(L, type, itl, n1, n2); break;
    case 0x0204: emix = -0.5; eb += contribute(); break;
    case 0x0404: emix = -0.3; eb += contribute(); break;
    case 0x0108: emix = 0.39; ecomp = fex_comp(L, type, itl, n1, n2); break;
    case 0x0208: emix = -0.11; eb += contribute(); break;
    case 0x0408: emix = 0.08; eb += contribute(); break;
    case 0x0808: emix = 0.41; eb += contribute(); break;
    case 0x0101: {
      etens = 0;
      const uint8_t dirs = avail_all(itl.o);
      for (int b = 0; b < 8; b++) {
        if (!((kStatus1 >> b) & 1)) continue;
        const int t = n1[b];
        if (!((dirs >> b) & 1)) {
          if (is_dumbbell(t)) eb += 0.70;
          else if (t == Cu) etens += 0.13;
          else if (t == Ni) etens += 0.02;
          else if (t == Mn) etens += -0.36;
        } else {
          if (is_dumbbell(t)) eb += 0.70;
          else if (t == Cu) etens += -0.13;
          else if (t == Ni) etens += 0.15;
          else if (t == Mn) etens += 0.02;
        }
      }
      // The reference tests the 1nn list under the 2nn status bits here
      // (energy_dumbbell.cpp:118-122); kept as written.
      for (int b = 0; b < 8; b++)
        if (((kStatus2 >> b) & 1) && is_dumbbell(n1[b])) eb += 0.48;
      break;
    }
    default: emix = 0; etens = 0; ecomp = 0; eb = 0; break;
    }
    edumb += emix + etens + ecomp + eb / 2;
  }
  return edumb;
}

// ItlRatesSolver::e0 (models/abvi/rate/itl_rates_solver.cpp:16-29).
static double e0_itl(int ghost) {
  switch (ghost) {
  case Fe: return 0.32;
  case Cu: return 0.36;
  case Ni: return 0.45;
  case Mn: return 0.31;
  default: return 0;
  }
}

// ItlRatesSolver::deltaE (itl_rates_solver.cpp:31-80): both sites' bond
// energies and the whole list's dumbbell energy, before and after the
// jumping atom leaves the dumbbell at src and joins the atom at dst.
static double delta_e(Typed& L, const std::vector<Itl>& list, int64_t src, int64_t dst, int ghost) {
  double e_before = 0;
  {
    const double e_src = count(L, src, ghost);
    const double e_des = count(L, dst, L.t[dst]);
    const double e_d = edumb(L, list);
    e_before = e_src + e_des + e_d;
  }
  const uint16_t cs = L.t[src], cd = L.t[dst];
  L.t[src] = (uint16_t)diff(cs, ghost);
  L.t[dst] = (uint16_t)combine_to_inter(cd, ghost);
  double e_after = 0;
  {
    const double e_src = count(L, src, L.t[src]);
    const double e_des = count(L, dst, ghost);
    const double e_d = edumb(L, list);
    e_after = e_src + e_des + e_d;
  }
  L.t[src] = cs;
  L.t[dst] = cd;
  return e_after - e_before;
}

// Itl::beforeRatesUpdate for every dumbbell, then Itl::updateRates for every
// dumbbell with RatesSolver::rate (rates_solver.cpp:13-26) as the callback
// (itl_list.cpp:25-50): each available direction's rate, halved, in the
// direction's two slots.
static void rates(Typed& L, std::vector<Itl>& list, double nu, double T) {
  for (Itl& itl : list) {
    int nb[8];
    for (int b = 0; b < 8; b++) nb[b] = L.t[L.nn1(itl.site, b)];
    for (double& r : itl.rates) r = 0;
    itl.avail = itl_avail(itl.o, kStatus1, nb);
  }
  for (Itl& itl : list) {
    const uint8_t dirs = avail_all(itl.o);
    for (int b = 0; b < 8; b++) {
      if (!((itl.avail >> b) & 1)) continue;
      const int64_t dst = L.nn1(itl.site, b);
      const int ghost = tran_atom(itl.o, L.t[itl.site], b);
      const double ea = e0_itl(ghost) + delta_e(L, list, itl.site, dst, ghost) / 2;
      const double rate = nu * exp(-ea / (refcpu::kBoltzmann * T));
      itl.rates[rates_index(b, dirs, false)] = rate / 2;
      itl.rates[rates_index(b, dirs, true)] = rate / 2;
    }
  }
}

}  // namespace refitl

// ---------------------------------------------------------------- refmix
// The abvi model's event loop with vacancies and self-interstitial dumbbells
// on a typed lattice.  The reference has the pieces but no wired loop: the
// dumbbell branch of ABVIModel::select is commented out
// (models/abvi/kmc.cpp:251-268), perform() handles vacancy hops only and
// asserts otherwise (:358-396), and its comments (:310-355) describe the
// dumbbell transition, recombination and defect generation in words.  This
// restatement wires them as follows (each choice is stated where it is made):
//  * calcRates over a sector sweeps z, y, x like ABVIModel::calcRates
//    (kmc.cpp:68-115); a vacancy contributes its 8 direction rates
//    (VacRatesSolver on the typed lattice: dumbbell neighbours carry no bond,
//    bonds_counter.cpp:35-81; a direction is available iff its 1nn is a single
//    atom, vac_hash.cpp:7-15), a dumbbell its 8 rate slots (Itl::updateRates,
//    itl_list.cpp:31-50, ItlRatesSolver).  The total folds every slot in that
//    order (zero slots add nothing), then adds defect_gen_rate
//    (ABVIModel::defectGenRate, kmc.cpp:207).  Every dumbbell's availability
//    mask is recomputed from the lattice before any rate (the masks
//    bond::Edumb reads, energy_dumbbell.cpp:144).
//  * select follows the commented branch (kmc.cpp:251-268): the first slot
//    whose running sum exceeds u * total; a dumbbell slot i is the
//    transition towards 1nn tag Itl::get1nnIdByRatesIndex(i, dirs)
//    (itl_list.cpp:76-91) with rotate = i odd; past the last slot the
//    default event is defect generation (kmc.cpp:245).
//  * a dumbbell transition (kmc.cpp:324-333 in words): the jumping atom g =
//    orientation::tranAtom (orientation.cc:238-253) leaves the dumbbell, whose
//    other atom stays (LatticeTypes::diff), and forms a dumbbell with the atom
//    at the target (combineToInter); the new orientation is
//    orientation::trans(tag, g at the high end of the new dumbbell, rotate)
//    (orientation.cc:245-252); the list entry moves to the target.
//  * after every transition, recombination around the moved defect
//    (rec::RecList::create / pickMinimum / Rec::recombine, recombine.cpp:
//    11-140): every vacancy-dumbbell pair it forms at 1nn or (masked) 2nn
//    is a candidate.  stageCombination compares the high end with itself
//    (recombine.cpp:91), so only "high atom to the vacancy, low atom stays"
//    is staged; pickMinimum keeps the candidates within 1e-6 of the lowest
//    energy through abs(), which for a double argument resolves to the int
//    overload with these includes under g++ 11 (a |difference| below 1 eV
//    truncates to 0), and draws one with std::rand(); here a stable sort and
//    an mt19937 seeded from seed_event_selection (parsed but unused by the
//    reference) stand in.  Edumb inside a candidate uses the masks of the
//    lattice before the candidate change, as a rate does.
//  * defect generation (kmc.cpp:334-353 in words: a random lattice, a
//    partner next to it, one becomes a vacancy and the other a dumbbell with
//    a random orientation): a site of the sector and one of its 12 third
//    neighbours (<110> type, the nearest shell outside the 1nn/2nn
//    recombination range) are drawn from the same stand-in stream; if both
//    hold single atoms the first becomes a vacancy and its atom joins the
//    second as a dumbbell of orientation::random() (orientation.h:38-42);
//    otherwise nothing changes.  Recombination follows around the new
//    vacancy.
namespace refmix {

using refitl::Itl;
using refitl::Typed;

// orientation::TransHashTable (orientation.cc:23-140), indexed
// [|o| - 1][1nn tag][!new_higher][!rotate]; 0 = unknown.
static const int8_t kTrans[6][8][2][2] = {
    {{{-3, -5}, {3, 5}}, {{-4, -6}, {4, 6}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}},
     {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{4, 6}, {-4, -6}}, {{3, 5}, {-3, -5}}},
    {{{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{4, -5}, {-4, 5}}, {{3, -6}, {-3, 6}},
     {{-3, 6}, {3, -6}}, {{-4, 5}, {4, -5}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}},
    {{{-1, -5}, {1, 5}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{-6, -2}, {6, 2}},
     {{6, 2}, {-6, -2}}, {{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{1, 5}, {-1, -5}}},
    {{{0, 0}, {0, 0}}, {{-1, -6}, {1, 6}}, {{-2, -5}, {2, 5}}, {{0, 0}, {0, 0}},
     {{0, 0}, {0, 0}}, {{2, 5}, {-2, -5}}, {{1, 6}, {-1, -6}}, {{0, 0}, {0, 0}}},
    {{{-3, -1}, {3, 1}}, {{0, 0}, {0, 0}}, {{-4, 2}, {4, -2}}, {{0, 0}, {0, 0}},
     {{0, 0}, {0, 0}}, {{4, -2}, {-4, 2}}, {{0, 0}, {0, 0}}, {{3, 1}, {-3, -1}}},
    {{{0, 0}, {0, 0}}, {{-4, -1}, {4, 1}}, {{0, 0}, {0, 0}}, {{-3, 2}, {3, -2}},
     {{3, -2}, {-3, 2}}, {{0, 0}, {0, 0}}, {{4, 1}, {-4, -1}}, {{0, 0}, {0, 0}}},
};
// Third neighbours (doubled x), in the order generation draws them.
static const int kThird[12][3] = {{-2, -1, 0}, {-2, 1, 0}, {-2, 0, -1}, {-2, 0, 1}, {2, -1, 0}, {2, 1, 0},
                                  {2, 0, -1},  {2, 0, 1},  {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}};

static int trans(int o, int tag, bool new_higher, bool rotate) {
  if (o == 0) return 0;
  const int h = o < 0 ? -o - 1 : o - 1;  // orientation::tranHash
  return kTrans[h][tag][new_higher ? 0 : 1][rotate ? 0 : 1];
}

enum Kind : int { K_VAC = 0, K_ITL = 1, K_GEN = 2, K_NONE = -1 };

struct MixParams {
  int64_t box[3];
  double nu, T, sector_T, physics_time, ef110, gen_rate;
  int64_t seed_time_inc, seed_rec;
};

struct Rec {
  int64_t v, d;
  int tv, td;
  double de;
};

class RefMix {
 public:
  RefMix(const MixParams& p, const uint16_t* types, int64_t n_itl, const int64_t* sites, const int8_t* ori)
      : P(p), t(types, types + 2 * p.box[0] * p.box[1] * p.box[2]) {
    L = Typed{2 * p.box[0], p.box[1], p.box[2], t.data()};
    for (int64_t k = 0; k < n_itl; k++) itl[sites[k]] = Itl{sites[k], ori[k], 0, {0}};
    for (int64_t s = 0; s < (int64_t)t.size(); s++)
      if (t[s] == refcpu::V) vac[s] = {};
    // SubLattice seeds its time stream with the (double -> int64) config seed
    // (simulation.h:70-76, sublattice.cpp:17).
    time_inc_rng.seed((std::mt19937::result_type)(int64_t)(double)p.seed_time_inc);
    rec_rng.seed((std::mt19937::result_type)p.seed_rec);
  }

  // Real-site sector region (cells; x not doubled): low half floor(n/2).
  void region(int s, int64_t lo[3], int64_t hi[3]) const {
    for (int d = 0; d < 3; d++) {
      const int64_t half = P.box[d] / 2;
      const bool high = (s >> d) & 1;
      lo[d] = high ? half : 0;
      hi[d] = high ? P.box[d] : half;
    }
  }

  void refresh_masks() {
    for (auto& kv : itl) {
      int nb[8];
      for (int b = 0; b < 8; b++) nb[b] = t[L.nn1(kv.first, b)];
      kv.second.avail = refitl::itl_avail(kv.second.o, refitl::kStatus1, nb);
    }
  }
  std::vector<Itl> list() const {
    std::vector<Itl> v;
    v.reserve(itl.size());
    for (auto& kv : itl) v.push_back(kv.second);
    return v;
  }

  // VacRatesSolver::deltaE / RatesSolver::rate on the typed lattice.
  double vac_rate(int64_t v, int b) {
    const int64_t tb = L.nn1(v, b);
    const int X = t[tb];
    if (!refitl::is_atom(X)) return 0.0;
    const double e_before = refitl::count(L, v, refcpu::V) + refitl::count(L, tb, X);
    t[tb] = refcpu::V;
    t[v] = (uint16_t)X;
    const double e_after = refitl::count(L, v, X) + refitl::count(L, tb, refcpu::V);
    t[tb] = (uint16_t)X;
    t[v] = refcpu::V;
    const double ea = refcpu::e0(X) + (e_after - e_before) / 2;
    return P.nu * exp(-ea / (refcpu::kBoltzmann * P.T));
  }
  // Itl::updateRates of one dumbbell (masks already current).
  void itl_rates(Itl& d, const std::vector<Itl>& lst) {
    for (double& r : d.rates) r = 0;
    const uint8_t dirs = refitl::avail_all(d.o);
    for (int b = 0; b < 8; b++) {
      if (!((d.avail >> b) & 1)) continue;
      const int64_t dst = L.nn1(d.site, b);
      const int ghost = refitl::tran_atom(d.o, t[d.site], b);
      const double ea = refitl::e0_itl(ghost) + refitl::delta_e(L, lst, d.site, dst, ghost) / 2;
      const double rate = P.nu * exp(-ea / (refcpu::kBoltzmann * P.T));
      d.rates[refitl::rates_index(b, dirs, false)] = rate / 2;
      d.rates[refitl::rates_index(b, dirs, true)] = rate / 2;
    }
  }

  template <class F>
  void sweep(int s, F f) {
    int64_t lo[3], hi[3];
    region(s, lo, hi);
    const int64_t X = L.X, Y = L.Y;
    for (int64_t z = lo[2]; z < hi[2]; z++)
      for (int64_t y = lo[1]; y < hi[1]; y++)
        for (int64_t x = 2 * lo[0]; x < 2 * hi[0]; x++)
          if (!f(x + X * (y + Y * z))) return;
  }

  double calcRates(int s) {
    refresh_masks();
    const std::vector<Itl> lst = list();
    double sum = 0;
    sweep(s, [&](int64_t site) {
      auto iv = vac.find(site);
      if (iv != vac.end()) {
        for (int b = 0; b < 8; b++) iv->second[b] = vac_rate(site, b);
        for (int b = 0; b < 8; b++) sum += iv->second[b];
      } else {
        auto ii = itl.find(site);
        if (ii != itl.end()) {
          itl_rates(ii->second, lst);
          for (int b = 0; b < 8; b++) sum += ii->second.rates[b];
        }
      }
      return true;
    });
    sum += P.gen_rate;
    return sum;
  }

  // First slot of the sweep whose running sum exceeds target.
  int select(int s, double target, int64_t* from, int* slot) {
    double acc = 0;
    int kind = K_GEN;
    sweep(s, [&](int64_t site) {
      const double* r = nullptr;
      int k = K_NONE;
      auto iv = vac.find(site);
      if (iv != vac.end()) { r = iv->second.data(); k = K_VAC; }
      else {
        auto ii = itl.find(site);
        if (ii != itl.end()) { r = ii->second.rates; k = K_ITL; }
      }
      if (!r) return true;
      for (int b = 0; b < 8; b++) {
        acc += r[b];
        if (acc > target) { *from = site; *slot = b; kind = k; return false; }
      }
      return true;
    });
    return kind;
  }

  // rec::RecList::create + pickMinimum + Rec::recombine around site c.
  void recombine(int64_t c) {
    const bool is_vac = t[c] == refcpu::V;
    if (!is_vac && !refitl::is_dumbbell(t[c])) return;
    std::vector<std::pair<int64_t, int64_t>> pairs;
    for (int b = 0; b < 8; b++) {
      const int64_t n = L.nn1(c, b);
      if (is_vac ? refitl::is_dumbbell(t[n]) : t[n] == refcpu::V) pairs.push_back(is_vac ? std::make_pair(c, n) : std::make_pair(n, c));
    }
    for (int k = 0; k < 6; k++) {
      if (!((refitl::kStatus2 >> k) & 1)) continue;
      const int64_t n = L.nn2(c, k);
      if (is_vac ? refitl::is_dumbbell(t[n]) : t[n] == refcpu::V) pairs.push_back(is_vac ? std::make_pair(c, n) : std::make_pair(n, c));
    }
    if (pairs.empty()) return;
    refresh_masks();
    const std::vector<Itl> lst = list();
    std::vector<Rec> recs;
    for (auto& pr : pairs) {
      const int64_t v = pr.first, d = pr.second;
      const int td = t[d], hi = refitl::high_end(td), lo = refitl::low_end(td);
      const double e_before = refitl::count(L, v, refcpu::V) + refitl::count(L, d, hi) + refitl::edumb(L, lst);
      t[v] = (uint16_t)hi;
      t[d] = (uint16_t)lo;
      const double e_after = refitl::count(L, v, hi) + refitl::count(L, d, lo) + refitl::edumb(L, lst);
      t[v] = refcpu::V;
      t[d] = (uint16_t)td;
      recs.push_back(Rec{v, d, hi, lo, (-1 * P.ef110) + e_after - e_before});
    }
    std::stable_sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.de < b.de; });
    std::vector<Rec> cand;
    for (const Rec& r : recs)
      if (std::trunc(r.de - recs[0].de) == 0.0) cand.push_back(r);  // int abs(): |d| < 1
    const Rec& r = cand.size() == 1 ? cand[0] : cand[rec_rng() % cand.size()];
    t[r.v] = (uint16_t)r.tv;
    t[r.d] = (uint16_t)r.td;
    vac.erase(r.v);
    itl.erase(r.d);
    n_recombined++;
  }

  // Applies the selected event; for a generation `from` receives the drawn
  // site.  *to_out: the moved defect's new site (-1: nothing moved).
  void perform(int kind, int s, int64_t& from, int slot, int64_t* to_out) {
    int64_t to = -1;
    if (kind == K_VAC) {
      to = L.nn1(from, slot);
      t[from] = t[to];
      t[to] = refcpu::V;
      vac.erase(from);
      vac[to] = {};
    } else if (kind == K_ITL) {
      const Itl d = itl.at(from);
      const uint8_t dirs = refitl::avail_all(d.o);
      const int tag = refitl::first_nn_by_rates_index(slot, dirs);
      to = L.nn1(from, tag);
      const int g = refitl::tran_atom(d.o, t[from], tag);
      t[from] = (uint16_t)refitl::diff(t[from], g);
      t[to] = (uint16_t)refitl::combine_to_inter(t[to], g);
      const bool new_higher = refitl::high_end(t[to]) == g;
      itl.erase(from);
      itl[to] = Itl{to, trans(d.o, tag, new_higher, (slot % 2) != 0), 0, {0}};
    } else {  // K_GEN
      std::vector<int64_t> sec;
      sweep(s, [&](int64_t site) { sec.push_back(site); return true; });
      const int64_t a = sec[rec_rng() % sec.size()];
      const int b = (int)(rec_rng() % 12u);
      const int64_t c = L.wrap_site(a, kThird[b][0], kThird[b][1], kThird[b][2]);
      if (refitl::is_atom(t[a]) && refitl::is_atom(t[c])) {
        const int ta = t[a];
        t[a] = refcpu::V;
        t[c] = (uint16_t)refitl::combine_to_inter(t[c], ta);
        vac[a] = {};
        itl[c] = Itl{c, refitl::orientation_from_draw(rec_rng()), 0, {0}};
        to = a;
      }
      from = a;
    }
    *to_out = to;
    if (to >= 0) recombine(to);
  }

  // SubLattice::startTimeLoop (sublattice.inl:61-143), as RefCpu::run.
  int run(int64_t max_events, std::vector<int64_t>* lf, std::vector<int64_t>* lt, std::vector<int>* lk) {
    static const int ring[8] = {0, 7, 2, 5, 3, 4, 1, 6};
    const int64_t time_steps = (int64_t)ceil(P.physics_time / P.sector_T);
    for (; step < time_steps; step++) {
      for (; sect < 8; sect++) {
        const int s = ring[sect];
        if (!in_sector) {
          thr = static_cast<double>(step + 1) * P.sector_T - evolution[s];
          sector_time = 0.0;
          in_sector = true;
        }
        while (sector_time < thr) {
          if (events >= max_events) return 1;
          const double total = calcRates(s);
          if (total == 0.0 || std::abs(total) < std::numeric_limits<double>::epsilon()) {
            sector_time = thr;
          } else {
            const double u = sel_rng() * (1.0 / std::mt19937::max());
            int64_t from = -1, to = -1;
            int slot = -1;
            const int kind = select(s, u * total, &from, &slot);
            if (kind == K_GEN && !(P.gen_rate > 0)) return -2;  // DefectGen with no generation rate
            perform(kind, s, from, slot, &to);
            if (lf) { lf->push_back(from); lt->push_back(to); lk->push_back(kind); }
            const double rnd = time_inc_dis(time_inc_rng);
            sector_time += -std::log(rnd) / total;
            events++;
          }
        }
        evolution[s] += sector_time;
        in_sector = false;
      }
      sect = 0;
    }
    return 0;
  }

  MixParams P;
  std::vector<uint16_t> t;
  Typed L;
  std::map<int64_t, Itl> itl;
  std::map<int64_t, std::array<double, 8>> vac;
  std::mt19937 sel_rng;       // r::rng: default-seeded (rand_generators.cpp:8)
  std::mt19937 time_inc_rng;  // SubLattice's time stream
  std::mt19937 rec_rng;       // stand-in for std::rand() (recombination) and generation draws
  std::uniform_real_distribution<double> time_inc_dis{0, 1};
  double evolution[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int64_t step = 0, events = 0, n_recombined = 0;
  int sect = 0;
  bool in_sector = false;
  double thr = 0, sector_time = 0;
};

}  // namespace refmix

extern "C" {

// ---- refmix (see the namespace comment)
void* refmix_create(const int64_t box[3], double nu, double T, double sector_T, double physics_time, double ef110,
                    double gen_rate, int64_t seed_time_inc, int64_t seed_rec, const uint16_t* types, int64_t n_itl,
                    const int64_t* sites, const int8_t* ori) {
  refmix::MixParams p{{box[0], box[1], box[2]}, nu, T, sector_T, physics_time, ef110, gen_rate, seed_time_inc,
                      seed_rec};
  return new refmix::RefMix(p, types, n_itl, sites, ori);
}
void refmix_destroy(void* h) { delete (refmix::RefMix*)h; }
// Runs up to max_events more events; logs (from, to, kind) per event when the
// arrays are given (kind 0 vacancy hop, 1 dumbbell transition, 2 defect
// generation; to = -1 when a generation changed nothing).
int refmix_run(void* h, int64_t max_events, int64_t* lf, int64_t* lt, int32_t* lk, int64_t* n_logged) {
  auto* r = (refmix::RefMix*)h;
  std::vector<int64_t> f, t;
  std::vector<int> k;
  const bool log = lf != nullptr;
  const int rc = r->run(r->events + max_events, log ? &f : nullptr, log ? &t : nullptr, log ? &k : nullptr);
  if (log) {
    for (size_t i = 0; i < f.size(); i++) { lf[i] = f[i]; lt[i] = t[i]; lk[i] = k[i]; }
    *n_logged = (int64_t)f.size();
  }
  return rc;
}
void refmix_state(void* h, double evolution[8], int64_t* ints, double* dbls) {
  auto* r = (refmix::RefMix*)h;
  for (int i = 0; i < 8; i++) evolution[i] = r->evolution[i];
  ints[0] = r->step; ints[1] = r->sect; ints[2] = r->in_sector ? 1 : 0; ints[3] = r->events;
  ints[4] = (int64_t)r->itl.size(); ints[5] = (int64_t)r->vac.size(); ints[6] = r->n_recombined;
  dbls[0] = r->thr; dbls[1] = r->sector_time;
}
void refmix_types(void* h, uint16_t* out) {
  auto* r = (refmix::RefMix*)h;
  memcpy(out, r->t.data(), r->t.size() * sizeof(uint16_t));
}
// Dumbbell list in site order: sites and orientations (cap entries).
int64_t refmix_itl(void* h, int64_t cap, int64_t* sites, int8_t* ori) {
  auto* r = (refmix::RefMix*)h;
  int64_t n = 0;
  for (auto& kv : r->itl) {
    if (n < cap) { sites[n] = kv.first; ori[n] = (int8_t)kv.second.o; }
    n++;
  }
  return n;
}
// calcRates of sector s: the defects in sweep order (site, kind) and their 8
// slots; returns the count and writes the total (with defect_gen_rate).
int64_t refmix_sector_rates(void* h, int s, int64_t cap, int64_t* sites, int32_t* kinds, double* rates,
                            double* total) {
  auto* r = (refmix::RefMix*)h;
  *total = r->calcRates(s);
  int64_t n = 0;
  r->sweep(s, [&](int64_t site) {
    const double* q = nullptr;
    int k = -1;
    auto iv = r->vac.find(site);
    if (iv != r->vac.end()) { q = iv->second.data(); k = 0; }
    else {
      auto ii = r->itl.find(site);
      if (ii != r->itl.end()) { q = ii->second.rates; k = 1; }
    }
    if (q) {
      if (n < cap) {
        sites[n] = site;
        kinds[n] = k;
        for (int b = 0; b < 8; b++) rates[8 * n + b] = q[b];
      }
      n++;
    }
    return true;
  });
  return n;
}
int refmix_trans(int o, int tag, int new_higher, int rotate) { return refmix::trans(o, tag, new_higher != 0, rotate != 0); }

// placeDumbbells of the reference's serial driver (profile_serial/main.cpp:
// 137-161) on a typed lattice of n sites: for each kind, counts[k] draws
// rand32(0, n) = mt() % n (random.cpp:27-32); a single atom there becomes
// combineToInter(atom, types[k]) with orientation::random(), drawn next.
// Writes the placed sites and orientations in placement order; returns the
// count.
int64_t refmix_place(uint16_t* t, int64_t n, const int32_t* kinds, const int64_t* counts, int32_t nk, int64_t seed,
                     int64_t* sites, int8_t* ori) {
  std::mt19937 rng((std::mt19937::result_type)seed);
  int64_t m = 0;
  for (int32_t k = 0; k < nk; k++)
    for (int64_t c = 0; c < counts[k]; c++) {
      const uint32_t id = (uint32_t)(rng() % (uint32_t)n);
      if (refitl::is_atom(t[id])) {
        t[id] = (uint16_t)refitl::combine_to_inter(t[id], kinds[k]);
        sites[m] = id;
        ori[m] = (int8_t)refitl::orientation_from_draw((uint32_t)rng());
        m++;
      }
    }
  return m;
}

}  // extern "C"

extern "C" {

// Orientation tables: which 0 = all, 1 = high end (X of XY), 2 = low end.
int ref_ori_avail(int o, int which) {
  return which == 1 ? refitl::avail_high(o) : which == 2 ? refitl::avail_low(o) : refitl::avail_all(o);
}
int ref_ori_jump(int o, int b) { return refitl::kJump[o + 6][b]; }
int ref_ori_tran_atom(int o, int type, int b) { return refitl::tran_atom(o, type, b); }
int ref_ori_from_draw(uint32_t r) { return refitl::orientation_from_draw(r); }
int ref_itl_avail(int o, int status, const int* nb) { return refitl::itl_avail(o, (uint8_t)status, nb); }
int ref_itl_rates_index(int b, int dirs, int up) { return refitl::rates_index(b, (uint8_t)dirs, up != 0); }
int ref_itl_first_nn(int idx, int dirs) { return refitl::first_nn_by_rates_index(idx, (uint8_t)dirs); }
int ref_itl_rates_status(int o, int avail) { return refitl::rates_status(o, (uint8_t)avail); }
// Itl::updateRates with a given value per direction in place of the rate
// callback (the unit tests' test_rates_updater, itl_rates_test.cpp).
void ref_itl_update_rates(int o, int avail, const double* cb, double* rates) {
  const uint8_t dirs = refitl::avail_all(o);
  for (int b = 0; b < 8; b++)
    if ((avail >> b) & 1) {
      rates[refitl::rates_index(b, dirs, false)] = cb[b] / 2;
      rates[refitl::rates_index(b, dirs, true)] = cb[b] / 2;
    }
}
int ref_combine_to_inter(int a, int b) { return refitl::combine_to_inter(a, b); }
int ref_type_diff(int t, int b) { return refitl::diff(t, b); }

// Dumbbell energy of a typed periodic lattice (box in cells; sites ascending).
double ref_edumb(const int64_t* box, uint16_t* types, int64_t n, const int64_t* sites, const int8_t* ori,
                 const uint8_t* avail) {
  refitl::Typed L{2 * box[0], box[1], box[2], types};
  std::vector<refitl::Itl> list((size_t)n);
  for (int64_t k = 0; k < n; k++) list[k] = refitl::Itl{sites[k], ori[k], avail[k], {0}};
  return refitl::edumb(L, list);
}
double ref_count_typed(const int64_t* box, uint16_t* types, int64_t site, int src) {
  refitl::Typed L{2 * box[0], box[1], box[2], types};
  return refitl::count(L, site, src);
}
// All dumbbells' availability masks and rate slots.
void ref_itl_rates(const int64_t* box, uint16_t* types, int64_t n, const int64_t* sites, const int8_t* ori,
                   double nu, double T, uint8_t* avail, double* rates) {
  refitl::Typed L{2 * box[0], box[1], box[2], types};
  std::vector<refitl::Itl> list((size_t)n);
  for (int64_t k = 0; k < n; k++) list[k] = refitl::Itl{sites[k], ori[k], 0, {0}};
  refitl::rates(L, list, nu, T);
  for (int64_t k = 0; k < n; k++) {
    avail[k] = list[k].avail;
    for (int b = 0; b < 8; b++) rates[8 * k + b] = list[k].rates[b];
  }
}

}  // extern "C"
