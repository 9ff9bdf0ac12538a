// The abvi model on a typed lattice: device helpers shared by the dumbbell
// rate step (akmc_itl.hip) and the vacancy + dumbbell event loop
// (akmc_mix.hip).
//
// Lattice: one reference LatticeTypes value per real site (uint16: V 0, Fe 1,
// Cu 2, Ni 4, Mn 8, dumbbell XY = X << 8 | Y with X <= Y;
// src/lattice/lattice_types.h:20-36), index x + 2nx (y + ny z) with x
// doubled, periodic.  Each translation unit that includes this header gets
// its own copy of the constant tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_exp.h"
#include "akmc_exp_table.h"
#include "akmc_model.h"

namespace akmc {
namespace itl {

static __constant__ uint64_t c_exp_tab[256] = AKMC_EXP_TAB_INIT;
static __constant__ double c_e1[25] = AKMC_E1_TABLE;
static __constant__ double c_e2[25] = AKMC_E2_TABLE;
// Direction masks by orientation o + 6: the high end's (X of XY) and the low
// end's (orientation.cc:180-236).
static __constant__ uint8_t c_high[13] = {0x50, 0xA0, 0x44, 0x88, 0x30, 0xC0, 0, 0x03, 0x0C, 0x11, 0x22, 0x05, 0x0A};
static __constant__ uint8_t c_low[13] = {0x0A, 0x05, 0x22, 0x11, 0x0C, 0x03, 0, 0xC0, 0x30, 0x88, 0x44, 0xA0, 0x50};

enum : int { T_V = 0, T_FE = 1, T_CU = 2, T_NI = 4, T_MN = 8 };
constexpr uint8_t STATUS_2NN = 0x2F;  // period_lattice_list.cpp:15-18

__device__ __forceinline__ bool is_dumbbell(int t) { return t > 0x0008; }
__device__ __forceinline__ bool is_atom_or_v(int t) { return t <= 0x0008; }
__device__ __forceinline__ bool is_atom(int t) { return t <= 0x0008 && t != 0; }
__device__ __forceinline__ int species(int t) {  // single type -> bond-table index
  return t == T_V ? SP_V : t == T_FE ? SP_FE : t == T_CU ? SP_CU : t == T_NI ? SP_NI : SP_MN;
}

struct Dims {
  int32_t X, Y, Z;
};

// LatticeTypes values the model takes (lattice_types.h:20-36): V or a single
// atom, or a dumbbell XY = X << 8 | Y of two atoms with X <= Y.
__host__ __device__ inline bool valid_single(int t) { return t == 0 || t == 1 || t == 2 || t == 4 || t == 8; }
__host__ __device__ inline bool valid_dumbbell(int t) {
  const int hi = t >> 8, lo = t & 0xFF;
  return hi != 0 && valid_single(hi) && valid_single(lo) && hi <= lo && (t >> 16) == 0;
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : v >= n ? v - n : v; }
__device__ __forceinline__ int64_t site_of(const Dims& d, int x, int y, int z) {
  return (int64_t)wrap(x, d.X) + (int64_t)d.X * ((int64_t)wrap(y, d.Y) + (int64_t)d.Y * wrap(z, d.Z));
}
// 1nn b of (x, y, z) in the reference's order (period_lattice_list.cpp:20-44).
__device__ __forceinline__ int64_t nn1(const Dims& d, int x, int y, int z, int b) {
  const int lo = (x & 1) ? 0 : -1;
  return site_of(d, x + ((b & 4) ? 1 : -1), y + lo + ((b >> 1) & 1), z + lo + (b & 1));
}
// 2nn k (period_lattice_list.cpp:223-232): (-2,0,0) (0,-1,0) (0,0,-1) (0,0,1) (0,1,0) (2,0,0).
__device__ __forceinline__ int64_t nn2(const Dims& d, int x, int y, int z, int k) {
  const int dx = k == 0 ? -2 : k == 5 ? 2 : 0;
  const int dy = k == 1 ? -1 : k == 4 ? 1 : 0;
  const int dz = k == 2 ? -1 : k == 3 ? 1 : 0;
  return site_of(d, x + dx, y + dy, z + dz);
}

// The lattice with up to two sites replaced (the state after a jump).
struct View {
  const uint16_t* t;
  int64_t s0, s1;
  int v0, v1;
  __device__ __forceinline__ int at(int64_t s) const { return s == s0 ? v0 : s == s1 ? v1 : (int)t[s]; }
};

// BondsCounter::count (bonds_counter.cpp:35-81): pair bonds of an atom of type
// src at (x, y, z) with its 1nn and counted 2nn; dumbbells carry no bond.
__device__ double count(const Dims& d, const View& L, int x, int y, int z, int src) {
  const int ps = species(src);
  double e = 0.0;
#pragma unroll
  for (int b = 0; b < N1; b++) {
    const int t = L.at(nn1(d, x, y, z, b));
    if (is_atom_or_v(t)) e += c_e1[species(t) * 5 + ps];
  }
#pragma unroll
  for (int k = 0; k < N2ALL; k++) {
    if (!((STATUS_2NN >> k) & 1)) continue;
    const int t = L.at(nn2(d, x, y, z, k));
    if (is_atom_or_v(t)) e += c_e2[species(t) * 5 + ps];
  }
  return e;
}

// One dumbbell's term of bond::Edumb, emix + etens + ecomp + eb / 2
// (energy_dumbbell.cpp:29-134, FeX_comp :139-206), on the lattice L.
__device__ double edumb_term(const Dims& d, const View& L, int x, int y, int z, int type, int o, uint8_t avail) {
  int n1[N1], n2[N2ALL];
#pragma unroll
  for (int b = 0; b < N1; b++) n1[b] = L.at(nn1(d, x, y, z, b));
#pragma unroll
  for (int k = 0; k < N2ALL; k++) n2[k] = ((STATUS_2NN >> k) & 1) ? L.at(nn2(d, x, y, z, k)) : 0;
  double emix = 0, etens = 0, ecomp = 0, eb = 0;
  const int lo = type & 0xFF;
  switch (type) {
    case 0x0102: emix = -0.46; break;
    case 0x0104: emix = -0.19; break;
    case 0x0108: emix = 0.39; break;
    case 0x0202: emix = -0.36; break;
    case 0x0204: emix = -0.5; break;
    case 0x0208: emix = -0.11; break;
    case 0x0404: emix = -0.3; break;
    case 0x0408: emix = 0.08; break;
    case 0x0808: emix = 0.41; break;
    default: break;
  }
  if (type == 0x0102 || type == 0x0104 || type == 0x0108) {
    // FeX: tension of the solute X towards its own jump targets.
    double c = 0, ebx = 0;
    const uint8_t comp = avail & c_low[o + 6];
    for (int b = 0; b < N1; b++) {
      if (!((comp >> b) & 1)) continue;
      const int t = n1[b];
      if (is_dumbbell(t)) {
        ebx += 0.70;
      } else if (lo == T_CU) {
        if (t == T_CU) c += -0.18;
        else if (t == T_NI) c += -0.35;
        else if (t == T_MN) c += -0.46;
      } else if (lo == T_NI) {
        if (t == T_CU) c += -0.21;
        else if (t == T_NI) c += -0.32;
        else if (t == T_MN) c += -0.30;
      } else {
        if (t == T_CU) c += 0.28;
        else if (t == T_NI) c += 0.29;
        else if (t == T_MN) c += 0.28;
      }
    }
    for (int k = 0; k < N2ALL; k++)
      if (((STATUS_2NN >> k) & 1) && is_dumbbell(n2[k])) ebx += 0.48;
    ecomp = c + ebx / 2;
  } else if (type == 0x0101) {
    const uint8_t dirs = c_high[o + 6] | c_low[o + 6];
    for (int b = 0; b < N1; b++) {
      const int t = n1[b];
      const bool jump = (dirs >> b) & 1;
      if (is_dumbbell(t)) eb += 0.70;
      else if (t == T_CU) etens += jump ? -0.13 : 0.13;
      else if (t == T_NI) etens += jump ? 0.15 : 0.02;
      else if (t == T_MN) etens += jump ? 0.02 : -0.36;
    }
    // The 2nn status bits select entries of the 1nn list here, as written in
    // energy_dumbbell.cpp:118-122.
    for (int k = 0; k < N2ALL; k++)
      if (((STATUS_2NN >> k) & 1) && is_dumbbell(n1[k])) eb += 0.48;
  } else if (type > 0x0108) {
    // Dumbbells of two solutes: 0.70 per 1nn and 0.48 per 2nn dumbbell,
    // interleaved by neighbour index.
    double c = 0;
    for (int b = 0; b < N1; b++) {
      if (is_dumbbell(n1[b])) c += 0.70;
      if (b < N2ALL && ((STATUS_2NN >> b) & 1) && is_dumbbell(n2[b])) c += 0.48;
    }
    eb += c;
  }
  return emix + etens + ecomp + eb / 2;
}


}  // namespace itl
}  // namespace akmc
