// C entry points over the reference's own, standalone-compilable source
// (src/lattice/lattice_types.cpp and the header-only
// models/abvi/rate/bonds/pair_bond.hpp), compiled from where they lie by
// oracle/Makefile (target `ref`).  Test infrastructure only: the tests use
// it to pin the oracle's restatement of randomAtomsType and makeBond.
#include <cstdint>
#include <vector>

#include "abvi/rate/bonds/pair_bond.hpp"
#include "lattice/lattice_types.h"

extern "C" int refsrc_random_atoms_type(const int32_t* types, const int64_t* ratio, unsigned len, int64_t hit) {
  std::vector<LatticeTypes::lat_type> t(len);
  for (unsigned i = 0; i < len; i++) t[i] = static_cast<LatticeTypes::lat_type>(types[i]);
  return (int)LatticeTypes::randomAtomsType(t.data(), ratio, len, hit);
}

extern "C" int refsrc_make_bond(int a, int b) {
  return (int)bonds::PairBond::makeBond(static_cast<LatticeTypes::lat_type>(a), static_cast<LatticeTypes::lat_type>(b));
}

// LatticeTypes' dumbbell helpers (lattice_types.h:48-95, lattice_types.cpp:21-48):
// they pin the oracle's restatement used by the dumbbell model.
extern "C" int refsrc_combine_to_inter(int a, int b) {
  return (int)LatticeTypes::combineToInter(static_cast<LatticeTypes::lat_type>(a), static_cast<LatticeTypes::lat_type>(b));
}

extern "C" int refsrc_diff(int t, int b) {
  const LatticeTypes T{static_cast<LatticeTypes::lat_type>(t)};
  return (int)T.diff(LatticeTypes{static_cast<LatticeTypes::lat_type>(b)});
}

// bit 0: isDumbbell, bit 1: isAtom, bit 2: isVacancy; then getHighEnd << 8 | getLowEnd << 16.
extern "C" int refsrc_type_info(int t) {
  const LatticeTypes T{static_cast<LatticeTypes::lat_type>(t)};
  return (T.isDumbbell() ? 1 : 0) | (T.isAtom() ? 2 : 0) | (T.isVacancy() ? 4 : 0) | ((int)T.getHighEnd() << 8) |
         ((int)T.getLowEnd() << 16);
}
