// Outputs of the reference's step hook (profiles/m_event_hook.cpp:17-70):
// lattice dumps in LatticeDump's text format, reading such a dump back as the
// reference's XYZ reader does, and the `output` section of config.yaml.
//
// Dump format (profiles/io/lattice_dump.cpp:14-63): the site count
// (2 nx * ny * nz), an empty line, then one line per site in z, y, x order
// over the real (non-ghost) sites of the ghost-extended grid:
//   corner site (even doubled x):  Type \t x/2 \t y \t z
//   body site (odd doubled x):     Type \t x/2 + 0.5 \t y + 0.5 \t z + 0.5
// with coordinates of that grid (real coordinate + ghost width g in every
// axis), integers for corner sites, and doubles as an ostream prints them by
// default (six significant digits).  Type is the readout of the site's
// membership set (V, Cu, Mn, Ni, else Fe; period_lattice_list.cpp:307-324).
#include <stdint.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/akmc.h"
#include "akmc_host.h"
#include "akmc_model.h"
#include "akmc_yaml.h"

using akmc_host::fail;

namespace {

const char* type_name(uint8_t f) {
  switch (akmc::readout(f)) {
    case akmc::SP_V: return "V";
    case akmc::SP_CU: return "Cu";
    case akmc::SP_MN: return "Mn";
    case akmc::SP_NI: return "Ni";
    default: return "Fe";
  }
}

// Flag byte of a dumped type name (lat::LatTypes, lattice_types_string.cpp);
// -1 for the dumbbell names, which this engine does not model.
int flags_of(const std::string& t) {
  if (t == "V") return akmc::BIT_V;
  if (t == "Fe") return 0;
  if (t == "Cu") return akmc::BIT_CU;
  if (t == "Mn") return akmc::BIT_MN;
  if (t == "Ni") return akmc::BIT_NI;
  return -1;
}

void copy_str(char* dst, size_t cap, const std::string& s) {
  const size_t n = std::min(cap - 1, s.size());
  memcpy(dst, s.data(), n);
  dst[n] = 0;
}

}  // namespace

extern "C" {

int akmc_output_from_yaml(const char* path, akmc_output* o) {
  if (!path || !o) return fail(AKMC_ERR_INVALID, "null argument");
  memset(o, 0, sizeof(*o));
  akmc::yaml::Node root;
  std::string err;
  if (!akmc::yaml::parse_file(path, &root, &err)) return fail(AKMC_ERR_CONFIG, "%s: %s", path, err.c_str());
  // ConfigParsing::parseOutput (profiles/config/config_parsing.cpp:216-229); the
  // section and its keys are optional.
  const akmc::yaml::Node* out = root.get("output");
  if (!out) return AKMC_OK;
  if (const akmc::yaml::Node* d = out->get("dump")) {
    o->dump_interval = d->get_int("interval", 0);
    if (const akmc::yaml::Node* f = d->get("file_path")) copy_str(o->dump_file_path, sizeof(o->dump_file_path), f->scalar);
  }
  if (const akmc::yaml::Node* l = out->get("logs")) {
    o->logs_interval = l->get_int("interval", 0);
    if (const akmc::yaml::Node* f = l->get("logs_file")) copy_str(o->logs_file, sizeof(o->logs_file), f->scalar);
  }
  return AKMC_OK;
}

int akmc_dump_path(const char* pattern, int64_t step, char* out, int64_t cap) {
  if (!pattern || !out || cap <= 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  // fmt::format(dump_file_path, step) (m_event_hook.cpp:62): every "{}" is the
  // step; other replacement fields are not supported.
  std::string p(pattern), r;
  for (size_t i = 0; i < p.size(); i++) {
    if (p[i] == '{') {
      if (i + 1 < p.size() && p[i + 1] == '}') {
        r += std::to_string((long long)step);
        i++;
        continue;
      }
      if (i + 1 < p.size() && p[i + 1] == '{') { r += '{'; i++; continue; }
      return fail(AKMC_ERR_INVALID, "unsupported replacement field in \"%s\"", pattern);
    }
    if (p[i] == '}') {
      if (i + 1 < p.size() && p[i + 1] == '}') { r += '}'; i++; continue; }
      return fail(AKMC_ERR_INVALID, "unmatched '}' in \"%s\"", pattern);
    }
    r += p[i];
  }
  if ((int64_t)r.size() >= cap) return fail(AKMC_ERR_INVALID, "path longer than %lld bytes", (long long)cap - 1);
  memcpy(out, r.c_str(), r.size() + 1);
  return AKMC_OK;
}

int akmc_write_dump(const uint8_t* sites, const int64_t box[3], int64_t ghost, const char* path) {
  if (!sites || !box || !path || box[0] < 1 || box[1] < 1 || box[2] < 1 || ghost < 0)
    return fail(AKMC_ERR_INVALID, "bad arguments");
  FILE* f = fopen(path, "w");
  if (!f) return fail(AKMC_ERR_INVALID, "invalid file path %s", path);
  const int64_t X = 2 * box[0], Y = box[1], Z = box[2];
  std::vector<char> buf;
  buf.reserve(1 << 20);
  char line[96];
  int k = snprintf(line, sizeof(line), "%lld\n\n", (long long)(X * Y * Z));
  buf.insert(buf.end(), line, line + k);
  for (int64_t z = 0; z < Z; z++)
    for (int64_t y = 0; y < Y; y++)
      for (int64_t x = 0; x < X; x++) {
        const uint8_t s = sites[x + X * (y + Y * z)];
        // Coordinates of the ghost-extended grid: doubled x + 2g, y + g, z + g.
        const int64_t gx = x + 2 * ghost, gy = y + ghost, gz = z + ghost;
        if ((gx & 1) == 0)
          k = snprintf(line, sizeof(line), "%s\t%lld\t%lld\t%lld\n", type_name(s), (long long)(gx / 2),
                       (long long)gy, (long long)gz);
        else
          k = snprintf(line, sizeof(line), "%s\t%g\t%g\t%g\n", type_name(s), (double)(gx / 2) + 0.5,
                       (double)gy + 0.5, (double)gz + 0.5);
        buf.insert(buf.end(), line, line + k);
        if (buf.size() > (1u << 20)) {
          if (fwrite(buf.data(), 1, buf.size(), f) != buf.size()) { fclose(f); return fail(AKMC_ERR_INVALID, "write failed: %s", path); }
          buf.clear();
        }
      }
  const bool ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  if (fclose(f) != 0 || !ok) return fail(AKMC_ERR_INVALID, "write failed: %s", path);
  return AKMC_OK;
}

int akmc_read_dump(const char* path, int64_t box[3], int64_t* ghost, uint8_t* sites, int64_t cap) {
  if (!path || !box) return fail(AKMC_ERR_INVALID, "bad arguments");
  std::ifstream in(path);
  if (!in.good()) return fail(AKMC_ERR_INVALID, "bad argument of file path %s", path);
  // XYZReader::XYZReader (tools/xyz_reader.cpp:15-80): the count, the rest of
  // that line, then one site per non-empty line; particle(line)
  // (tools/particle.h): the type up to a tab, then x, y, z as doubles,
  // stored as (int)(2x), (int)y, (int)z.
  unsigned long count = 0;
  in >> count;
  std::string line;
  std::getline(in, line);
  struct P { int64_t x, y, z; int f; };
  std::vector<P> ps;
  int64_t xmin = INT64_MAX, ymin = INT64_MAX, zmin = INT64_MAX, xmax = INT64_MIN, ymax = INT64_MIN, zmax = INT64_MIN;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    std::istringstream ls(line);
    std::string t;
    std::getline(ls, t, '\t');
    double xf = 0, yf = 0, zf = 0;
    ls >> xf >> yf >> zf;
    const int fl = flags_of(t);
    if (fl < 0) return fail(AKMC_ERR_INVALID, "lattice type \"%s\" is not supported (vacancy and atom types only)", t.c_str());
    const double lim = 1e9;  // far beyond any box the engines take; keeps the integer conversions defined
    if (!(std::fabs(xf) < lim && std::fabs(yf) < lim && std::fabs(zf) < lim))
      return fail(AKMC_ERR_INVALID, "coordinates out of range in line \"%s\"", line.c_str());
    const P p{(int64_t)(2 * xf), (int64_t)yf, (int64_t)zf, fl};
    xmin = std::min(xmin, p.x); ymin = std::min(ymin, p.y); zmin = std::min(zmin, p.z);
    xmax = std::max(xmax, p.x); ymax = std::max(ymax, p.y); zmax = std::max(zmax, p.z);
    ps.push_back(p);
  }
  if (ps.size() != count || ps.empty()) return fail(AKMC_ERR_INVALID, "bad lattice count");
  // Whole unit cells: doubled x must start at a corner site (even) and span
  // corner/body pairs (an even count), as LatticeDump writes it.
  if ((xmin & 1) != 0 || ((xmax - xmin + 1) & 1) != 0)
    return fail(AKMC_ERR_INVALID, "doubled x runs from %lld to %lld: not whole unit cells", (long long)xmin,
                (long long)xmax);
  box[0] = (xmax - xmin + 1) / 2;
  box[1] = ymax - ymin + 1;
  box[2] = zmax - zmin + 1;
  if (ghost) *ghost = xmin / 2;
  const int64_t X = 2 * box[0], Y = box[1], Z = box[2], n = X * Y * Z;
  if ((int64_t)count != n) return fail(AKMC_ERR_INVALID, "the sites do not fill a %lld x %lld x %lld box",
                                       (long long)box[0], (long long)box[1], (long long)box[2]);
  if (!sites) return AKMC_OK;  // size query
  if (cap < n) return fail(AKMC_ERR_INVALID, "buffer holds %lld sites, the dump %lld", (long long)cap, (long long)n);
  std::vector<uint8_t> seen(n, 0);
  for (const P& p : ps) {
    const int64_t i = (p.x - xmin) + X * ((p.y - ymin) + Y * (p.z - zmin));
    if (i < 0 || i >= n) return fail(AKMC_ERR_INVALID, "site (%lld, %lld, %lld) lies outside the box", (long long)p.x,
                                     (long long)p.y, (long long)p.z);
    if (seen[i]) return fail(AKMC_ERR_INVALID, "site (%lld, %lld, %lld) appears twice", (long long)p.x,
                             (long long)p.y, (long long)p.z);
    seen[i] = 1;
    sites[i] = (uint8_t)p.f;
  }
  return AKMC_OK;
}

int akmc_pipe_path_from_yaml(const char* path, char* out, int64_t cap) {
  if (!path || !out || cap <= 0) return fail(AKMC_ERR_INVALID, "bad arguments");
  akmc::yaml::Node root;
  std::string err;
  if (!akmc::yaml::parse_file(path, &root, &err)) return fail(AKMC_ERR_CONFIG, "%s: %s", path, err.c_str());
  // ConfigParsing::parseCreate (profiles/config/config_parsing.cpp:150-188):
  // create.random wins when enabled; otherwise create.pipe.input_box.
  const akmc::yaml::Node* create = root.get("create");
  const akmc::yaml::Node* rnd = create ? create->get("random") : nullptr;
  const akmc::yaml::Node* pipe = create ? create->get("pipe") : nullptr;
  if ((rnd && rnd->get_bool("enable", false)) || !pipe || !pipe->get_bool("enable", false))
    return fail(AKMC_ERR_CONFIG, "create.pipe is not the create option of %s", path);
  const akmc::yaml::Node* f = pipe->get("input_box");
  const std::string v = f ? f->scalar : std::string();
  if ((int64_t)v.size() >= cap) return fail(AKMC_ERR_INVALID, "path longer than %lld bytes", (long long)cap - 1);
  memcpy(out, v.c_str(), v.size() + 1);
  return AKMC_OK;
}

int akmc_pipe_sites(const akmc_params* p, const char* pipe_file, uint8_t* sites, int64_t n, int64_t* n_vacancies) {
  if (!p || !pipe_file || !sites) return fail(AKMC_ERR_INVALID, "null argument");
  // creation::createFromPile (src/creation.cpp:184-230): the atoms of
  // createAtomsRandom (the random creation's draw without its vacancies) ...
  akmc_params q = *p;
  q.va_count = 0;
  const int rc = akmc_random_sites(&q, sites, n);
  if (rc != AKMC_OK) return rc;
  // ... then one vacancy per "x y z" triple of the file (x doubled, real
  // coordinates), read until the first that does not parse; triples outside
  // the box are skipped (the local sub-box test of :216, one rank).  The
  // reference writes the site's type as V, so the atom is replaced.
  std::ifstream in(pipe_file);
  if (!in.good()) return fail(AKMC_ERR_INVALID, "cannot open pipe file %s", pipe_file);
  const int64_t X = 2 * p->box[0], Y = p->box[1], Z = p->box[2];
  long long x, y, z;
  while (in >> x >> y >> z) {
    if (x < 0 || x >= X || y < 0 || y >= Y || z < 0 || z >= Z) continue;
    uint8_t& f = sites[x + X * (y + Y * z)];
    if (!(f & akmc::BIT_V)) f = akmc::BIT_V;
  }
  if (n_vacancies) {
    int64_t c = 0;
    for (int64_t i = 0; i < n; i++) c += (sites[i] & akmc::BIT_V) ? 1 : 0;
    *n_vacancies = c;
  }
  return AKMC_OK;
}

}  // extern "C"
