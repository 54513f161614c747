// Host code of the library under AddressSanitizer and UndefinedBehaviorSanitizer:
// the mesh, DOF, CSR, transfer, colour and slot-map construction of setup.cpp
// (Q1/Q2, one and several macro cells, several depths), the vertex-lattice
// refinement of mapped hierarchies and the GEO reader of geo.cpp on good and
// malformed files.  Built with g++ by tests/test_host_sanitizers.py (no GPU
// code is compiled); every structural invariant checked here would also be
// an out-of-range access in the device kernels that consume these arrays.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "cd3d_internal.hpp"

using namespace cd3d;

static int bad = 0;
#define EXPECT(c)                                                       \
  do {                                                                  \
    if (!(c)) {                                                         \
      printf("FAILED %s:%d %s\n", __FILE__, __LINE__, #c);              \
      bad++;                                                            \
    }                                                                   \
  } while (0)

static void check_level(const HostLevel &L, int order) {
  const int nb = (order + 1) * (order + 1) * (order + 1);
  EXPECT((int)L.rowptr.size() == L.ndof + 1 && L.rowptr[0] == 0 && L.rowptr[L.ndof] == L.nnz);
  for (int r = 0; r < L.ndof; r++) {
    EXPECT(L.rowptr[r] < L.rowptr[r + 1]);
    for (int k = L.rowptr[r]; k < L.rowptr[r + 1]; k++) {
      EXPECT(L.kcol[k] >= 0 && L.kcol[k] < L.ndof);
      if (k > L.rowptr[r]) EXPECT(L.kcol[k] > L.kcol[k - 1]);
    }
    EXPECT(L.kcol[L.diag[r]] == r);
  }
  EXPECT((int)L.gn.size() == L.ncell * nb);
  for (int g : L.gn) EXPECT(g >= 0 && g < L.ndof);
  EXPECT((int)L.colour_start.size() == 9 && L.colour_start[8] == L.ncell);
  std::vector<int> seen(L.ncell, 0);
  for (int c : L.colour_cells) {
    EXPECT(c >= 0 && c < L.ncell);
    if (c >= 0 && c < L.ncell) seen[c]++;
  }
  for (int s : seen) EXPECT(s == 1);
  std::vector<int> slots;
  build_slots(L, order, slots);
  EXPECT((int64_t)slots.size() == (int64_t)L.ncell * nb * nb);
  for (int s : slots) EXPECT(s >= -1 && s < L.nnz);
}

// the wavefront schedules of the row-order SOR: each active row once; every
// neighbour visited before it in the sequential sweep sits on a lower level
// and every one visited after it on a higher level (so a level's rows read
// exactly the values the sequential sweep reads)
static void check_gs(const HostLevel &L) {
  for (int b = 0; b < 2; b++) {
    std::vector<int> rows, start;
    build_gs_schedule(L, b == 1, rows, start);
    EXPECT((int)rows.size() == L.nact && !start.empty() && start.back() == L.nact);
    std::vector<int> lev(L.nact, -1);
    for (int k = 0; k + 1 < (int)start.size(); k++)
      for (int r = start[k]; r < start[k + 1]; r++) {
        EXPECT(rows[r] >= 0 && rows[r] < L.nact);
        if (rows[r] >= 0 && rows[r] < L.nact) {
          EXPECT(lev[rows[r]] == -1);
          lev[rows[r]] = k;
        }
      }
    for (int i = 0; i < L.nact; i++)
      for (int j = L.rowptr[i]; j < L.rowptr[i + 1]; j++) {
        const int c = L.kcol[j];
        if (c >= L.nact || c == i) continue;
        const bool before = b == 0 ? c < i : c > i;
        EXPECT(before ? lev[c] < lev[i] : lev[c] > lev[i]);
      }
  }
}

// the gather lists of the two-pass assembly give, entry by entry, the sums of
// the colour-by-colour scatter bit for bit
static void check_gather(const HostLevel &L, int order) {
  const int nb = (order + 1) * (order + 1) * (order + 1);
  std::vector<int> slots, gptr, gidx, rptr, ridx;
  build_slots(L, order, slots);
  build_gather(L, order, slots, gptr, gidx, rptr, ridx);
  std::vector<double> E((size_t)L.ncell * nb * nb), Er((size_t)L.ncell * nb);
  unsigned s = 12345;
  auto rnd = [&] {
    s = s * 1103515245u + 12345u;
    return (double)(s >> 8) / (1 << 24) - 0.5;
  };
  for (auto &e : E) e = rnd();
  for (auto &e : Er) e = rnd();
  std::vector<double> val(L.nnz, 0.0), rhs(L.ndof, 0.0);
  for (int k = 0; k < 8; k++)
    for (int q = L.colour_start[k]; q < L.colour_start[k + 1]; q++) {
      const int c = L.colour_cells[q];
      for (int i = 0; i < nb; i++) {
        const int r = L.gn[(size_t)c * nb + i];
        if (r >= L.nact) continue;
        for (int j = 0; j < nb; j++) val[slots[((size_t)c * nb + i) * nb + j]] += E[((size_t)c * nb + i) * nb + j];
        rhs[r] += Er[(size_t)c * nb + i];
      }
    }
  const int nnz_act = L.rowptr[L.nact];
  EXPECT((int)gptr.size() == nnz_act + 1 && (int)rptr.size() == L.nact + 1);
  for (int e = 0; e < nnz_act; e++) {
    double t = 0.0;
    for (int k = gptr[e]; k < gptr[e + 1]; k++) {
      EXPECT(gidx[k] >= 0 && (size_t)gidx[k] < E.size());
      t += E[gidx[k]];
    }
    EXPECT(t == val[e]);
  }
  for (int r = 0; r < L.nact; r++) {
    double t = 0.0;
    for (int k = rptr[r]; k < rptr[r + 1]; k++) t += Er[ridx[k]];
    EXPECT(t == rhs[r]);
  }
}

static int geo_status(const std::string &path, const std::string &text) {
  std::ofstream(path) << text;
  int m = 0;
  char msg[256];
  std::vector<double> xyz(3 * 1000);
  return cd3d_read_geo(path.c_str(), &m, xyz.data(), xyz.size(), msg, sizeof msg);
}

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  for (int order = 1; order <= 2; order++)
    for (int m : {1, 2, 3})
      for (int depth = 0; depth <= (m == 1 ? 3 : 2); depth++) {
        HostLevel C, F;
        build_host_level(C, order, m, depth);
        build_host_level(F, order, m, depth + 1);
        build_transfer(F, C, order);
        build_colours(C, order);
        build_colours(F, order);
        build_mc_colours(F, order);
        check_level(C, order);
        check_level(F, order);
        if (depth <= 1) check_gather(F, order);
        check_gs(C);
        check_gs(F);
        for (int c : F.p_col) EXPECT(c >= 0 && c < C.ndof);
        for (int c : F.r_col) EXPECT(c >= 0 && c < F.ndof);
        EXPECT((int)F.mc_rows.size() == F.nact);
      }
  // vertex lattices: refinement keeps the corners and halves the spacing
  std::vector<double> v(3 * 27), w, u;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        v[3 * (i + 3 * (j + 3 * k))] = i * 0.5 + 0.05 * (i == 1 && j == 1 && k == 1);
        v[3 * (i + 3 * (j + 3 * k)) + 1] = j * 0.5;
        v[3 * (i + 3 * (j + 3 * k)) + 2] = k * 0.5;
      }
  refine_vertices(v, 2, w);
  refine_vertices(w, 4, u);
  EXPECT(w.size() == 3 * 125 && u.size() == 3 * 729);
  EXPECT(u[0] == 0.0 && u[3 * 728] == 1.0 && u[3 * 728 + 1] == 1.0 && u[3 * 728 + 2] == 1.0);
  // the GEO reader on a valid 1-cell file and on malformed ones
  const std::string good =
      "t\nt\n 1 8 4 8 6 NEL NVT NVpF NVpEL NBF\nDCORVG\n0 0 0\n1 0 0\n0 1 0\n1 1 0\n0 0 1\n1 0 1\n0 1 1\n1 1 1\n"
      "KVERT\n1 2 4 3 5 6 8 7\nKNPR\n0 0 0 0 0 0 0 0\n";
  EXPECT(geo_status(dir + "/good.GEO", good) == CD3D_OK);
  EXPECT(geo_status(dir + "/trunc.GEO", good.substr(0, 60)) == CD3D_ERR_ARG);
  EXPECT(geo_status(dir + "/empty.GEO", "") == CD3D_ERR_ARG);
  EXPECT(geo_status(dir + "/huge.GEO", "t\nt\n 1000000000 8 4 8 6\n") == CD3D_ERR_ARG);
  EXPECT(geo_status(dir + "/neg.GEO", "t\nt\n -8 -27 4 8 6\n") == CD3D_ERR_ARG);
  std::string badvert = good;
  badvert.replace(badvert.find("1 2 4 3"), 7, "1 2 9 3");
  EXPECT(geo_status(dir + "/badvert.GEO", badvert) == CD3D_ERR_ARG);
  int m = 0;
  EXPECT(cd3d_read_geo((dir + "/none.GEO").c_str(), &m, nullptr, 0, nullptr, 0) == CD3D_ERR_ARG);
  printf(bad ? "FAILED %d\n" : "OK\n", bad);
  return bad ? 1 : 0;
}
