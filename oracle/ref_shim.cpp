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
