// abvi vacancy pair-bond model: parameters, site encoding and the BCC stencil.
//
// Site encoding (one uint8 per lattice site): four membership bits that mirror
// the reference's four occupancy hash containers (src/lattice/lattices_list.h:
// 329-333): bit0 V, bit1 Cu, bit2 Mn, bit3 Ni; no bit = Fe.  A site reads as
// the first present of V, Cu, Mn, Ni, else Fe
// (src/lattice/period_lattice_list.cpp:307-324).  Several bits can be set at
// once (stale Ni/Mn tags under a vacancy, reference src/creation.cpp:149-158);
// they are kept exactly because they change later readouts.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define AKMC_HD __host__ __device__
#else
#define AKMC_HD
#endif

namespace akmc {

enum SiteBits : uint8_t { BIT_V = 1, BIT_CU = 2, BIT_MN = 4, BIT_NI = 8 };

// Species index used by the bond tables.
enum Species : int { SP_V = 0, SP_FE = 1, SP_CU = 2, SP_NI = 3, SP_MN = 4 };

// Readout of a flag byte as a species index.
AKMC_HD inline int readout(uint8_t f) {
  return (f & BIT_V) ? SP_V : (f & BIT_CU) ? SP_CU : (f & BIT_MN) ? SP_MN : (f & BIT_NI) ? SP_NI : SP_FE;
}

// Flag bit carried by a species (0 for V and Fe).
AKMC_HD inline uint8_t species_bit(int sp) {
  return sp == SP_CU ? BIT_CU : sp == SP_MN ? BIT_MN : sp == SP_NI ? BIT_NI : 0;
}

// Pair-bond energies (eV), Vincent et al. 2006 table 7, as compiled into the
// reference (models/abvi/rate/bonds/bonds_counter.cpp:13-33).  Symmetric
// 5x5 tables indexed by Species.
#define AKMC_E1_TABLE                                           \
  {                                                             \
    /*      V       Fe      Cu      Ni      Mn  */              \
    /*V */ 0.315, -0.161, -0.103, -0.234, -0.151,              \
    /*Fe*/ -0.161, -0.778, -0.609, -0.821, -0.648,             \
    /*Cu*/ -0.103, -0.609, -0.581, -0.692, -0.519,             \
    /*Ni*/ -0.234, -0.821, -0.692, -0.793, -0.831,             \
    /*Mn*/ -0.151, -0.648, -0.519, -0.831, -0.438              \
  }
#define AKMC_E2_TABLE                                           \
  {                                                             \
    /*      V       Fe      Cu      Ni      Mn  */              \
    /*V */ -0.214, -0.161, -0.206, -0.351, -0.206,             \
    /*Fe*/ -0.161, -0.389, -0.344, -0.399, -0.364,             \
    /*Cu*/ -0.206, -0.344, -0.389, -0.344, -0.249,             \
    /*Ni*/ -0.351, -0.399, -0.344, -0.389, -0.464,             \
    /*Mn*/ -0.206, -0.364, -0.249, -0.464, -0.389              \
  }
// Migration barrier prefactor e0 by exchanged atom
// (models/abvi/rate/vacancy_rates_solver.cpp:14-27); V entry unused.
#define AKMC_E0_TABLE \
  { 0.0, 0.62, 0.54, 0.68, 1.03 }
// Boltzmann constant in eV/K (models/abvi/rate/rates_types.h:16).
#define AKMC_KB 8.617330350E-5

// Rate-stencil geometry for a vacancy at an even-x (corner) or odd-x (body)
// site, in doubled-x coordinates.  1nn order (lattice_list_meta.cpp:22-70):
// even x: (-1,-1,-1) (-1,-1,0) (-1,0,-1) (-1,0,0) (+1,-1,-1) (+1,-1,0)
// (+1,0,-1) (+1,0,0); odd x: the same with y,z offsets 0/+1 instead of -1/0.
// Direction b and 7-b are opposite.  2nn order (period_lattice_list.cpp:
// 234-256): (-2,0,0) (0,-1,0) (0,0,-1) (0,0,+1) (0,+1,0) (+2,0,0), of which
// the periodic lattice list uses mask 0x2F (:15-18): the (0,+1,0) bond is
// not counted.
enum { N1 = 8, N2ALL = 6, N2 = 5, STENCIL_MAX = 64 };
static const uint32_t MASK_2NN = 0x2F;

AKMC_HD inline void nn1_offset(int parity, int b, int* dx, int* dy, int* dz) {
  *dx = (b & 4) ? 1 : -1;
  const int lo = parity ? 0 : -1;  // even: {-1,0}; odd: {0,+1}
  *dy = lo + ((b >> 1) & 1);
  *dz = lo + (b & 1);
}
AKMC_HD inline void nn2_offset(int k, int* dx, int* dy, int* dz) {
  *dx = 0; *dy = 0; *dz = 0;
  switch (k) {
    case 0: *dx = -2; break;
    case 1: *dy = -1; break;
    case 2: *dz = -1; break;
    case 3: *dz = 1; break;
    case 4: *dy = 1; break;
    default: *dx = 2; break;
  }
}

// Unique sites of a vacancy's rate stencil: the vacancy, its 8 1nn and
// counted 2nn, and for every 1nn target t_b the 1nn and counted 2nn of t_b.
// Index 0 is the vacancy.  Built on the host by build_stencil().
struct StencilTables {
  int8_t off[2][STENCIL_MAX][3];
  int32_t n_sites;
  uint8_t w1[2][N1];           // stencil index of 1nn_b(vacancy)
  uint8_t w2[2][N2];           // counted 2nn of the vacancy, in order
  uint8_t t1[2][N1][N1];       // stencil index of 1nn_k(t_b)
  uint8_t t2[2][N1][N2];       // counted 2nn of t_b
  uint8_t reach[2][5][5];      // bit (dx+3) set if (dx,dy,dz) is in the stencil, [dz+2][dy+2]
};

}  // namespace akmc
