// Solute-cluster statistics of a lattice (host side): Cu clusters by 1nn
// connectivity, and the reference's isolated-Cu count
// (profiles/m_event_hook.cpp:19-55: a site in the Cu set none of whose 1nn
// sites reads as Cu).
#include <stdint.h>

#include <cstring>
#include <vector>

#include "../../include/akmc.h"
#include "akmc_host.h"
#include "akmc_model.h"

using akmc_host::fail;

namespace {

int32_t find_root(std::vector<int32_t>& parent, int32_t a) {
  while (parent[a] != a) {
    parent[a] = parent[parent[a]];
    a = parent[a];
  }
  return a;
}

}  // namespace

extern "C" int akmc_cu_clusters(const uint8_t* sites, const int64_t box[3], int64_t hist_cap, int64_t* hist,
                                int64_t* n_clusters, int64_t* n_cu, int64_t* n_isolated) {
  if (!sites || !box || box[0] < 2 || box[1] < 2 || box[2] < 2 || (hist_cap > 0 && !hist))
    return fail(AKMC_ERR_INVALID, "bad arguments");
  const int64_t X = 2 * box[0], Y = box[1], Z = box[2], n = X * Y * Z;
  auto is_cu = [&](int64_t i) { return akmc::readout(sites[i]) == akmc::SP_CU; };
  auto nb = [&](int64_t x, int64_t y, int64_t z, int b) {
    int dx, dy, dz;
    akmc::nn1_offset((int)(x & 1), b, &dx, &dy, &dz);
    const int64_t xx = (x + dx + X) % X, yy = (y + dy + Y) % Y, zz = (z + dz + Z) % Z;
    return xx + X * (yy + Y * zz);
  };
  // Dense ids of the Cu atoms.
  std::vector<int32_t> id(n, -1);
  int32_t m = 0;
  for (int64_t i = 0; i < n; i++)
    if (is_cu(i)) id[i] = m++;
  std::vector<int32_t> parent(m);
  for (int32_t a = 0; a < m; a++) parent[a] = a;
  int64_t isolated = 0;
  for (int64_t z = 0; z < Z; z++)
    for (int64_t y = 0; y < Y; y++)
      for (int64_t x = 0; x < X; x++) {
        const int64_t i = x + X * (y + Y * z);
        const bool member = (sites[i] & akmc::BIT_CU) != 0;
        if (!member && id[i] < 0) continue;
        bool any = false;
        for (int b = 0; b < 8; b++) {
          const int64_t j = nb(x, y, z, b);
          if (id[j] < 0) continue;
          any = true;
          if (id[i] >= 0) {
            const int32_t ra = find_root(parent, id[i]), rb = find_root(parent, id[j]);
            if (ra != rb) parent[ra < rb ? rb : ra] = ra < rb ? ra : rb;
          }
        }
        if (member && !any) isolated++;
      }
  std::vector<int32_t> size(m, 0);
  for (int32_t a = 0; a < m; a++) size[find_root(parent, a)]++;
  int64_t clusters = 0;
  if (hist_cap > 0) memset(hist, 0, sizeof(int64_t) * hist_cap);
  for (int32_t a = 0; a < m; a++) {
    if (size[a] == 0) continue;
    clusters++;
    if (hist_cap > 0) hist[size[a] < hist_cap ? size[a] : hist_cap - 1]++;
  }
  if (n_clusters) *n_clusters = clusters;
  if (n_cu) *n_cu = m;
  if (n_isolated) *n_isolated = isolated;
  return AKMC_OK;
}
