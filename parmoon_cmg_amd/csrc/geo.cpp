// Reader of ParMooN GEO files (the text format of TDomain::ReadGeo,
// ReadGeo.C:46-200) for the meshes this library refines: structured
// m x m x m macro meshes of hexahedra.
//
// The file holds two title lines; NEL NVT NVpF NVpEL NBF; a header line and
// NVT vertex lines (x y z when the vertex's KNPR entry is 0, boundary
// parameters otherwise); "KVERT" and NEL lines of NVpEL 1-based vertex
// numbers; "KNPR" and NVT integers; then the boundary faces, which the
// lattice form does not need (every outer face is Dirichlet).
//
// Accepted: NVpEL 8, NEL = m^3, NVT = (m+1)^3, every KNPR 0, vertices
// numbered along the lattice (x fastest) and cell (i, j, k) listed at
// i + m (j + m k) with its vertices in the hexahedron order of the reference
// (bottom face counter-clockwise from the lattice corner (i, j, k), then the
// top face), i.e. the identity local frame of the lattice hierarchy.  The
// vertex coordinates are free: general hexahedra, mapped trilinearly.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cd3d.h"

namespace {

struct GeoError {
  std::string msg;
};

void fail(const std::string &m) { throw GeoError{m}; }

int read_geo(const char *path, std::vector<double> &xyz) {
  std::ifstream in(path);
  if (!in) fail(std::string("cannot open ") + path);
  std::string line;
  for (int k = 0; k < 2; k++)
    if (!std::getline(in, line)) fail("missing title lines");
  if (!std::getline(in, line)) fail("missing the NEL NVT NVpF NVpEL NBF line");
  long nel = 0, nvt = 0, nvpf = 0, nvpel = 0, nbf = 0;
  {
    std::istringstream ss(line);
    if (!(ss >> nel >> nvt >> nvpf >> nvpel >> nbf)) fail("bad NEL NVT NVpF NVpEL NBF line");
  }
  if (nvpel != 8) fail("only hexahedral meshes (NVpEL 8) are supported");
  long m = (long)std::llround(std::cbrt((double)nel));
  if (m < 1 || m * m * m != nel || (m + 1) * (m + 1) * (m + 1) != nvt)
    fail("not a structured m^3 macro mesh (NEL = m^3, NVT = (m+1)^3)");
  if (!std::getline(in, line)) fail("missing the DCORVG header");
  xyz.assign(3 * (size_t)nvt, 0.0);
  for (long v = 0; v < nvt; v++)
    if (!(in >> xyz[3 * v] >> xyz[3 * v + 1] >> xyz[3 * v + 2])) fail("bad vertex line " + std::to_string(v + 1));
  std::string word;
  if (!(in >> word) || word != "KVERT") fail("missing KVERT");
  std::vector<long> kv(8 * (size_t)nel);
  for (auto &x : kv)
    if (!(in >> x)) fail("bad KVERT entry");
  if (!(in >> word) || word != "KNPR") fail("missing KNPR");
  for (long v = 0; v < nvt; v++) {
    long knpr = 0;
    if (!(in >> knpr)) fail("bad KNPR entry");
    if (knpr != 0)
      fail("vertex " + std::to_string(v + 1) +
           " is given by boundary parameters (KNPR != 0); explicit coordinates (KNPR 0) are needed");
  }
  static const int hexv[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (long k = 0, c = 0; k < m; k++)
    for (long j = 0; j < m; j++)
      for (long i = 0; i < m; i++, c++)
        for (int v = 0; v < 8; v++) {
          const long want = 1 + (i + hexv[v][0]) + (m + 1) * ((j + hexv[v][1]) + (m + 1) * (k + hexv[v][2]));
          if (kv[8 * c + v] != want)
            fail("cell " + std::to_string(c + 1) +
                 " is not in lattice order (cells x fastest, vertices numbered along the lattice, "
                 "the reference's hexahedron vertex order)");
        }
  for (double x : xyz)
    if (!std::isfinite(x)) fail("non-finite vertex coordinate");
  return (int)m;
}

}  // namespace

extern "C" cd3d_status cd3d_read_geo(const char *path, int *m, double *xyz, size_t cap, char *msg, size_t msg_len) {
  if (msg && msg_len) msg[0] = 0;
  if (!path || !m) return CD3D_ERR_ARG;
  try {
    std::vector<double> v;
    *m = read_geo(path, v);
    if (xyz) {
      if (cap < v.size()) fail("vertex buffer too small: " + std::to_string(v.size()) + " doubles needed");
      std::memcpy(xyz, v.data(), sizeof(double) * v.size());
    }
    return CD3D_OK;
  } catch (const GeoError &e) {
    if (msg && msg_len) snprintf(msg, msg_len, "%s: %s", path, e.msg.c_str());
    return CD3D_ERR_ARG;
  } catch (...) {
    if (msg && msg_len) snprintf(msg, msg_len, "%s: read failed", path);
    return CD3D_ERR_NOMEM;
  }
}
