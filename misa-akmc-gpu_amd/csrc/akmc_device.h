// Device-side lattice geometry and the FP64 rate, shared by the plain (one
// lattice) and batched (many replicas) event loops.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "akmc_exp.h"
#include "akmc_model.h"

namespace akmc {

// Lattice key: x (doubled, 11 bits) | y << 11 (11 bits) | z << 22 (10 bits).
// Sorting keys sorts sites in the reference's z, y, x sweep order
// (models/abvi/kmc.cpp:72-75).
__host__ __device__ __forceinline__ uint32_t key_x(uint32_t k) { return k & 0x7FFu; }
__host__ __device__ __forceinline__ uint32_t key_y(uint32_t k) { return (k >> 11) & 0x7FFu; }
__host__ __device__ __forceinline__ uint32_t key_z(uint32_t k) { return k >> 22; }
__host__ __device__ __forceinline__ uint32_t make_key(uint32_t x, uint32_t y, uint32_t z) {
  return x | (y << 11) | (z << 22);
}

struct BoxDims {
  int32_t X;  // 2 * cells in x (doubled)
  int32_t Y;
  int32_t Z;
  int32_t hx, hy, hz;  // low-half sizes: 2*floor(nx/2), floor(ny/2), floor(nz/2)
};

// Dimensions of a box of box[0] x box[1] x box[2] cells.  The low half of
// each axis has floor(n/2) cells: local_split_coord = sector_lattice_size[0]
// + ghost with sector_lattice_size[0][d] = n/2 (src/simulation.cpp:91-110);
// the sector bounds that add_ghost spells out per sector id
// (gpu/gpu_simulate.hip:2093-3086) split every axis there.
__host__ __device__ inline BoxDims box_dims(const int64_t box[3]) {
  BoxDims d;
  d.X = (int32_t)(2 * box[0]);
  d.Y = (int32_t)box[1];
  d.Z = (int32_t)box[2];
  d.hx = (int32_t)(2 * (box[0] / 2));
  d.hy = (int32_t)(box[1] / 2);
  d.hz = (int32_t)(box[2] / 2);
  return d;
}

// Key of site index i = x + X (y + Y z).
__host__ __device__ __forceinline__ uint32_t key_of_index(const BoxDims& d, int64_t i) {
  return make_key((uint32_t)(i % d.X), (uint32_t)((i / d.X) % d.Y), (uint32_t)(i / ((int64_t)d.X * d.Y)));
}

__device__ __forceinline__ uint64_t key_index(const BoxDims& d, uint32_t k) {
  return (uint64_t)key_x(k) + (uint64_t)d.X * ((uint64_t)key_y(k) + (uint64_t)d.Y * key_z(k));
}

__device__ __forceinline__ int wrapc(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

__device__ __forceinline__ uint32_t key_add(const BoxDims& d, uint32_t k, int dx, int dy, int dz) {
  return make_key((uint32_t)wrapc((int)key_x(k) + dx, d.X), (uint32_t)wrapc((int)key_y(k) + dy, d.Y),
                  (uint32_t)wrapc((int)key_z(k) + dz, d.Z));
}

// Sector of a site: bit0 x high, bit1 y high, bit2 z high (the low half has
// floor(n/2) cells; the per-sector ghost bounds of add_ghost,
// gpu/gpu_simulate.hip:2103-3086, place sector s on the high side of x, y, z
// exactly where bits 0, 1, 2 of s are set).
__host__ __device__ __forceinline__ int key_sector(const BoxDims& d, uint32_t k) {
  return ((int)key_x(k) >= d.hx ? 1 : 0) | ((int)key_y(k) >= d.hy ? 2 : 0) | ((int)key_z(k) >= d.hz ? 4 : 0);
}

// Periodic difference a - b folded into [-n/2, n/2).
__device__ __forceinline__ int pdiff(int a, int b, int n) {
  int d = a - b;
  if (d >= n / 2) d -= n;
  if (d < -(n / 2)) d += n;
  return d;
}

// True if site q lies in the rate stencil of the vacancy at w.
__device__ __forceinline__ bool in_stencil(const BoxDims& d, const StencilTables& S, uint32_t w, uint32_t q) {
  const int dx = pdiff((int)key_x(q), (int)key_x(w), d.X);
  const int dy = pdiff((int)key_y(q), (int)key_y(w), d.Y);
  const int dz = pdiff((int)key_z(q), (int)key_z(w), d.Z);
  // Branch-free: every lane evaluates the same instructions.
  const bool in_box = (dx >= -3) & (dx <= 3) & (dy >= -2) & (dy <= 2) & (dz >= -2) & (dz <= 2);
  const int iz = min(max(dz + 2, 0), 4), iy = min(max(dy + 2, 0), 4), ix = min(max(dx + 3, 0), 7);
  return in_box & (((S.reach[key_x(w) & 1][iz][iy] >> ix) & 1) != 0);
}

// ---------------------------------------------------------------- rates
// Rate of vacancy (stencil species sp[]) hopping in direction b, in FP64,
// following the reference evaluation order exactly:
//   VacRatesSolver::deltaE (models/abvi/rate/vacancy_rates_solver.cpp:29-172)
//   e_before = count(v, V) + count(t, X)
//   e_after  = count(v, X with t read as V) + count_exchange(t, V with v read as X)
//   BondsCounter::count (bonds_counter.cpp:35-81): 8 1nn terms then the
//   counted 2nn terms, a left fold from 0.0.
//   RatesSolver::rate (rates_solver.cpp:13-26): Ea = e0(X) + dE/2,
//   k = nu * exp(-Ea / (kB*T)).
// Returns 0 for a direction whose target is a vacancy (VacancyHash::
// availTranDirs2, models/abvi/defect/vac_hash.cpp:7-16).
__device__ __forceinline__ double rate_fp64(const uint8_t* sp, int par, int b, const StencilTables& S,
                                            const double* e1, const double* e2, const double* e0, double nu,
                                            double kT, const uint64_t* exp_tab) {
  const int X = sp[S.w1[par][b]];
  if (X == SP_V) return 0.0;
  // e_src before: bonds of V at v.
  double src_b = 0.0;
#pragma unroll
  for (int k = 0; k < N1; k++) src_b += e1[SP_V * 5 + sp[S.w1[par][k]]];
#pragma unroll
  for (int k = 0; k < N2; k++) src_b += e2[SP_V * 5 + sp[S.w2[par][k]]];
  // e_des before: bonds of X at t.
  double des_b = 0.0;
#pragma unroll
  for (int k = 0; k < N1; k++) des_b += e1[X * 5 + sp[S.t1[par][b][k]]];
#pragma unroll
  for (int k = 0; k < N2; k++) des_b += e2[X * 5 + sp[S.t2[par][b][k]]];
  // e_src after: bonds of X at v, with t now a vacancy.
  double src_a = 0.0;
#pragma unroll
  for (int k = 0; k < N1; k++) src_a += e1[X * 5 + (k == b ? (int)SP_V : (int)sp[S.w1[par][k]])];
#pragma unroll
  for (int k = 0; k < N2; k++) src_a += e2[X * 5 + sp[S.w2[par][k]]];
  // e_des after: bonds of V at t, with v now holding X (v is 1nn 7-b of t).
  double des_a = 0.0;
#pragma unroll
  for (int k = 0; k < N1; k++) des_a += e1[SP_V * 5 + (k == 7 - b ? X : (int)sp[S.t1[par][b][k]])];
#pragma unroll
  for (int k = 0; k < N2; k++) des_a += e2[SP_V * 5 + sp[S.t2[par][b][k]]];
  const double e_before = src_b + des_b;
  const double e_after = src_a + des_a;
  const double ea = e0[X] + (e_after - e_before) / 2;
  return nu * exp_glibc(-ea / kT, exp_tab);
}

}  // namespace akmc
