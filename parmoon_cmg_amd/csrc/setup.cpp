// Host-side construction of the multilevel problem: mesh, DOF numbering, CSR
// pattern, grid-transfer operators, element slot maps and colours.
//
// The mesh is the uniform refinement of m^3 macro bricks in depth-first child
// order with the rotated child frames of the regular hexahedron refinement
// (RefHexaRegDesc.C:29-33).  Everything is generated on the integer node
// lattice of each level, so shared DOFs, boundary DOFs and matrix columns are
// found by lattice arithmetic instead of the reference's joint/mapper graph.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "cd3d_internal.hpp"

namespace cd3d {

namespace {

// child frames of the regular refinement: child j starts at parent vertex j;
// its local axes (v1-v0, v3-v0, v4-v0) in the parent's local axes
// (RefHexaRegDesc.C:29-33, DatChildVertex)
const int kChildAxes[8][3][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},   {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}},
    {{-1, 0, 0}, {0, -1, 0}, {0, 0, 1}}, {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 1, 0}, {1, 0, 0}, {0, 0, -1}},  {{-1, 0, 0}, {0, 1, 0}, {0, 0, -1}},
    {{0, -1, 0}, {-1, 0, 0}, {0, 0, -1}}, {{1, 0, 0}, {0, -1, 0}, {0, 0, -1}}};
// hexahedron vertex j in local coordinates (bottom 0-3 counter-clockwise, 4-7 above)
const int kHexV[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                         {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct Frame {
  int O[3];
  int E[3][3];
};

void emit(HostLevel &L, int &cnt, const Frame &f, int S, int depth) {
  if (depth == 0) {
    for (int d = 0; d < 3; d++) L.org[3 * cnt + d] = f.O[d];
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++) L.ax[9 * cnt + 3 * a + d] = (int8_t)f.E[a][d];
    cnt++;
    return;
  }
  for (int j = 0; j < 8; j++) {
    Frame c;
    for (int d = 0; d < 3; d++)
      c.O[d] = f.O[d] + S * (kHexV[j][0] * f.E[0][d] + kHexV[j][1] * f.E[1][d] + kHexV[j][2] * f.E[2][d]);
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++)
        c.E[a][d] = kChildAxes[j][a][0] * f.E[0][d] + kChildAxes[j][a][1] * f.E[1][d] +
                    kChildAxes[j][a][2] * f.E[2][d];
    emit(L, cnt, c, S / 2, depth - 1);
  }
}

template <class F>
void parallel_for(int n, F f) {
  int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([=]() {
      int b = (int)((int64_t)n * t / nt), e = (int)((int64_t)n * (t + 1) / nt);
      for (int i = b; i < e; i++) f(i);
    });
  for (auto &x : th) x.join();
}

inline int cell_node(const HostLevel &L, int c, int i, int j, int k) {
  const int *O = &L.org[3 * c];
  const int8_t *E = &L.ax[9 * c];
  int X = O[0] + i * E[0] + j * E[3] + k * E[6];
  int Y = O[1] + i * E[1] + j * E[4] + k * E[7];
  int Z = O[2] + i * E[2] + j * E[5] + k * E[8];
  return X + L.nn * (Y + L.nn * Z);
}

}  // namespace

int quad_points_1d(int order) { return order + 2; }  // Gauss3 (Q1), Gauss4 (Q2): FEDatabase3D.C:1009-1016

// New vertices of the regular hexahedron refinement (Refinement.C with
// RefHexaRegDesc.C): an edge's at 0.5 v0 + 0.5 v1, a face's at the mean of
// its 4 vertices, a cell's at the mean of its 8; the sums run over the
// corners in lattice order.  The vertex lattice of an n^3 structured mesh
// becomes that of its 2n^3 refinement.
void refine_vertices(const std::vector<double> &src, int n, std::vector<double> &dst) {
  const int64_t ns = n + 1, nd = 2 * (int64_t)n + 1;
  dst.assign(3 * nd * nd * nd, 0.0);
  for (int64_t K = 0; K < nd; K++)
    for (int64_t J = 0; J < nd; J++)
      for (int64_t I = 0; I < nd; I++) {
        const int odd[3] = {(int)(I & 1), (int)(J & 1), (int)(K & 1)};
        const int64_t base[3] = {I >> 1, J >> 1, K >> 1};
        const int nodd = odd[0] + odd[1] + odd[2];
        const double wt = nodd == 0 ? 1.0 : nodd == 1 ? 0.5 : nodd == 2 ? 0.25 : 0.125;
        double *o = &dst[3 * (I + nd * (J + nd * K))];
        for (int t = 0; t < 8; t++) {
          const int e[3] = {t & 1, (t >> 1) & 1, (t >> 2) & 1};
          if ((e[0] && !odd[0]) || (e[1] && !odd[1]) || (e[2] && !odd[2])) continue;
          const double *x = &src[3 * ((base[0] + e[0]) + ns * ((base[1] + e[1]) + ns * (base[2] + e[2])))];
          for (int d = 0; d < 3; d++) o[d] += wt * x[d];
        }
      }
}

void build_host_level(HostLevel &L, int order, int m, int depth) {
  const int p = order, nb = (p + 1) * (p + 1) * (p + 1);
  L.n = m << depth;
  L.ncell = L.n * L.n * L.n;
  L.nn = p * L.n + 1;
  L.org.assign(3 * (size_t)L.ncell, 0);
  L.ax.assign(9 * (size_t)L.ncell, 0);
  int cnt = 0, S = p << depth;
  for (int k = 0; k < m; k++)
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) {
        Frame f{{i * S, j * S, k * S}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
        emit(L, cnt, f, S, depth);
      }

  // DOF numbering: first touch in (cell, local DOF) order, inner DOFs first,
  // then Dirichlet DOFs (boundary nodes), each class counted separately --
  // the result of the union-find marks and final pass of
  // TFESpace3D::ConstructSpace (FESpace3D.C:459-2035, FE3DMapper.C:134-230).
  const int nn = L.nn;
  const int64_t nnodes = (int64_t)nn * nn * nn;
  std::vector<int> node_dof(nnodes, -1);
  L.gn.resize((size_t)nb * L.ncell);
  int n_in = 0, n_dir = 0;
  for (int c = 0; c < L.ncell; c++)
    for (int b = 0; b < nb; b++) {
      int i = b % (p + 1), j = (b / (p + 1)) % (p + 1), k = b / ((p + 1) * (p + 1));
      int node = cell_node(L, c, i, j, k);
      if (node_dof[node] == -1) {
        int X = node % nn, Y = (node / nn) % nn, Z = node / (nn * nn);
        bool bd = X == 0 || Y == 0 || Z == 0 || X == nn - 1 || Y == nn - 1 || Z == nn - 1;
        node_dof[node] = bd ? -2 - (n_dir++) : n_in++;
      }
      L.gn[(size_t)c * nb + b] = node;
    }
  L.nact = n_in;
  L.ndof = n_in + n_dir;
  L.dof_node.assign(L.ndof, 0);
  for (int64_t i = 0; i < nnodes; i++) {
    int v = node_dof[i];
    if (v == -1) continue;
    int d = v >= 0 ? v : n_in + (-2 - v);
    node_dof[i] = d;
    L.dof_node[d] = (int)i;
  }
  for (auto &g : L.gn) g = node_dof[g];

  // CSR pattern (TSquareStructure3D + Sort): an active row holds every DOF of
  // the cells around its node, sorted; a Dirichlet row holds its diagonal.
  // The cells around lattice node X along one axis span [X-p, X+p] when X is
  // on a cell face plane, otherwise the single cell containing X.
  std::vector<int> len(L.ndof);
  auto range = [&](int X, int &lo, int &hi) {
    if (X % p == 0) {
      lo = std::max(0, X - p);
      hi = std::min(nn - 1, X + p);
    } else {
      lo = p * (X / p);
      hi = lo + p;
    }
  };
  parallel_for(L.ndof, [&](int r) {
    if (r >= L.nact) { len[r] = 1; return; }
    int node = L.dof_node[r], X = node % nn, Y = (node / nn) % nn, Z = node / (nn * nn), a[3], b[3];
    range(X, a[0], b[0]);
    range(Y, a[1], b[1]);
    range(Z, a[2], b[2]);
    len[r] = (b[0] - a[0] + 1) * (b[1] - a[1] + 1) * (b[2] - a[2] + 1);
  });
  L.rowptr.assign(L.ndof + 1, 0);
  for (int r = 0; r < L.ndof; r++) L.rowptr[r + 1] = L.rowptr[r] + len[r];
  L.nnz = L.rowptr[L.ndof];
  L.kcol.assign(L.nnz, 0);
  L.diag.assign(L.ndof, 0);
  parallel_for(L.ndof, [&](int r) {
    int *out = &L.kcol[L.rowptr[r]];
    if (r >= L.nact) {
      out[0] = r;
      L.diag[r] = L.rowptr[r];
      return;
    }
    int node = L.dof_node[r], X = node % nn, Y = (node / nn) % nn, Z = node / (nn * nn), a[3], b[3];
    range(X, a[0], b[0]);
    range(Y, a[1], b[1]);
    range(Z, a[2], b[2]);
    int t = 0;
    for (int z = a[2]; z <= b[2]; z++)
      for (int y = a[1]; y <= b[1]; y++)
        for (int x = a[0]; x <= b[0]; x++) out[t++] = node_dof[x + nn * (y + (int64_t)nn * z)];
    std::sort(out, out + t);
    L.diag[r] = (int)(std::lower_bound(out, out + t, r) - &L.kcol[0]);
  });
  build_colours(L, order);
  build_row_blocks(L.rowptr, kStreamNnz, kStreamRows, L.blk);
  build_lattice(L, order);
}

// Lattice-stencil form: classes, the CSR position of every stencil value and
// the interior node -> DOF map.  Stencil entry order within a class: z, y, x
// offsets ascending (x fastest).
void build_lattice(HostLevel &L, int order) {
  const int p = order, nn = L.nn;
  std::vector<int> node_dof((size_t)nn * nn * nn, -1);
  for (int d = 0; d < L.ndof; d++) node_dof[L.dof_node[d]] = d;
  L.lat.clear();
  const int ncls = p == 1 ? 1 : 8;
  int64_t off = 0, first = 0;
  for (int c = 0; c < ncls; c++) {
    LatClass C;
    C.step = p == 1 ? 1 : 2;
    int sten = 1;
    for (int d = 0; d < 3; d++) {
      C.px[d] = p == 1 ? 0 : (c >> d) & 1;
      // interior coordinates 1..nn-2 with the class parity
      C.lo[d] = p == 1 ? 1 : (C.px[d] ? 1 : 2);
      C.cnt[d] = p == 1 ? nn - 2 : (C.px[d] ? (nn - 1) / 2 : (nn - 3) / 2);
      sten *= (p == 1 || C.px[d] == 0) ? 2 * p + 1 : 3;
    }
    C.nodes = (int64_t)C.cnt[0] * C.cnt[1] * C.cnt[2];
    C.nsten = sten;
    C.off = off;
    C.first = first;
    off += C.nodes * sten;
    first += C.nodes;
    L.lat.push_back(C);
  }
  L.lat_nval = off;
  L.lat_perm.assign(off, -1);
  L.lat_node2dof.assign(first, -1);
  L.lat_sid.assign(off, 0);
  L.lat_first.assign(8 * 128, 0);
  std::atomic<int> missing{0};
  for (int c = 0; c < (int)L.lat.size(); c++) {
    const LatClass &C = L.lat[c];
    parallel_for((int)C.nodes, [&](int k) {
      const int kx = k % C.cnt[0], ky = (k / C.cnt[0]) % C.cnt[1], kz = k / (C.cnt[0] * C.cnt[1]);
      const int X = C.lo[0] + C.step * kx, Y = C.lo[1] + C.step * ky, Z = C.lo[2] + C.step * kz;
      const int row = node_dof[X + nn * (Y + (int64_t)nn * Z)];
      L.lat_node2dof[C.first + k] = row;
      int rng[3][2];
      const int co[3] = {X, Y, Z};
      for (int d = 0; d < 3; d++) {
        if (co[d] % p == 0) { rng[d][0] = -p; rng[d][1] = p; }
        else { rng[d][0] = -(co[d] % p); rng[d][1] = p - co[d] % p; }
      }
      const int *b = &L.kcol[L.rowptr[row]], *e = &L.kcol[L.rowptr[row + 1]];
      int s = 0;
      for (int dz = rng[2][0]; dz <= rng[2][1]; dz++)
        for (int dy = rng[1][0]; dy <= rng[1][1]; dy++)
          for (int dx = rng[0][0]; dx <= rng[0][1]; dx++, s++) {
            const int cdof = node_dof[(X + dx) + nn * ((Y + dy) + (int64_t)nn * (Z + dz))];
            const int *q = std::lower_bound(b, e, cdof);
            if (q == e || *q != cdof) {
              missing++;
              continue;
            }
            const int pos = (int)(q - &L.kcol[0]);
            L.lat_perm[C.off + (int64_t)s * C.nodes + k] = pos;
            L.lat_sid[pos] = (uint16_t)(c * 128 + s);
            if (k == 0) L.lat_first[c * 128 + s] = pos;
          }
    });
  }
  if (missing) throw std::logic_error("lattice stencil entry missing from the CSR pattern");
}

// Consecutive rows grouped so that each group holds at most max_nnz nonzeros
// and max_rows rows (a longer row forms a group of its own).
void build_row_blocks(const std::vector<int> &rowptr, int max_nnz, int max_rows, std::vector<int> &blk) {
  const int n = (int)rowptr.size() - 1;
  blk.clear();
  blk.push_back(0);
  int r = 0;
  while (r < n) {
    int e = r + 1;
    while (e < n && e - r < max_rows && rowptr[e + 1] - rowptr[r] <= max_nnz) e++;
    blk.push_back(e);
    r = e;
  }
}

// Wavefront schedule of the reference's sequential SOR sweep over the active
// rows (TMGLevel3D::SOR / SSOR, MGLevel3D.C:254-441): forward, row i reads
// the new values of its lower-numbered neighbours and the old values of the
// higher-numbered ones; backward (the second half of SSOR), the other way
// round.  level(i) = 1 + the largest level of i's earlier-visited active
// neighbours, so that with the symmetric pattern of the level matrices every
// row of a level depends only on rows of lower levels and the rows of one
// level can update at once with the values the sequential sweep would read.
// rows: the active rows by level, ascending within a level; start: nlev + 1
// offsets.
void build_gs_schedule(const HostLevel &L, bool backward, std::vector<int> &rows, std::vector<int> &start) {
  const int na = L.nact;
  std::vector<int> lev(na, 0);
  int nlev = 0;
  for (int t = 0; t < na; t++) {
    const int i = backward ? na - 1 - t : t;
    int m = -1;
    for (int j = L.rowptr[i]; j < L.rowptr[i + 1]; j++) {
      const int c = L.kcol[j];
      if (c < na && (backward ? c > i : c < i)) m = std::max(m, lev[c]);
    }
    lev[i] = m + 1;
    nlev = std::max(nlev, m + 2);
  }
  start.assign(nlev + 1, 0);
  for (int i = 0; i < na; i++) start[lev[i] + 1]++;
  for (int k = 0; k < nlev; k++) start[k + 1] += start[k];
  rows.assign(na, 0);
  std::vector<int> pos(start.begin(), start.end() - 1);
  for (int i = 0; i < na; i++) rows[pos[lev[i]]++] = i;
}

// Grid transfer between coarse level and the next finer one: P(i,k) is the
// value of coarse basis function k at fine node i -- what Prolongate
// computes cell by cell with the prolongation matrices and the multiplicity
// average (MGComponents3D.C:182-463, FEDatabase3D.C:1630-1790); the weights
// are dyadic (Q1: 1, 1/2; Q2: 1, 3/4, 3/8, -1/8 per axis) so the products are
// exact.  Restriction is R = P^T (DefectRestriction, MGComponents3D.C:706-989).
void build_transfer(HostLevel &F, const HostLevel &C, int order) {
  const int p = order, nnf = F.nn, nnc = C.nn, nc = C.n;
  std::vector<int> cnode_dof((size_t)nnc * nnc * nnc, -1);
  for (int d = 0; d < C.ndof; d++) cnode_dof[C.dof_node[d]] = d;
  struct W { int idx[3]; double w[3]; int n; };
  auto w1d = [&](int X, W &o) {
    int c = std::min(X / (2 * p), nc - 1), t = X - 2 * p * c;
    double xi = (double)t / p - 1.0;
    double v[3];
    if (p == 1) { v[0] = 0.5 - 0.5 * xi; v[1] = 0.5 + 0.5 * xi; }
    else { v[0] = 0.5 * xi * (xi - 1); v[1] = 1 - xi * xi; v[2] = 0.5 * xi * (xi + 1); }
    o.n = 0;
    for (int k = 0; k <= p; k++)
      if (v[k] != 0.0) { o.idx[o.n] = c * p + k; o.w[o.n] = v[k]; o.n++; }
  };
  std::vector<int> len(F.ndof);
  std::vector<std::vector<std::pair<int, double>>> rows(F.ndof);
  parallel_for(F.ndof, [&](int i) {
    int node = F.dof_node[i], X = node % nnf, Y = (node / nnf) % nnf, Z = node / (nnf * nnf);
    W a, b, c;
    w1d(X, a); w1d(Y, b); w1d(Z, c);
    auto &r = rows[i];
    for (int kz = 0; kz < c.n; kz++)
      for (int ky = 0; ky < b.n; ky++)
        for (int kx = 0; kx < a.n; kx++) {
          int cn = a.idx[kx] + nnc * (b.idx[ky] + nnc * c.idx[kz]);
          r.push_back({cnode_dof[cn], a.w[kx] * b.w[ky] * c.w[kz]});
        }
    std::sort(r.begin(), r.end());
  });
  F.p_rowptr.assign(F.ndof + 1, 0);
  for (int i = 0; i < F.ndof; i++) F.p_rowptr[i + 1] = F.p_rowptr[i] + (int)rows[i].size();
  F.p_col.resize(F.p_rowptr[F.ndof]);
  F.p_val.resize(F.p_rowptr[F.ndof]);
  for (int i = 0; i < F.ndof; i++)
    for (size_t t = 0; t < rows[i].size(); t++) {
      F.p_col[F.p_rowptr[i] + t] = rows[i][t].first;
      F.p_val[F.p_rowptr[i] + t] = rows[i][t].second;
    }
  // transpose
  F.r_rowptr.assign(C.ndof + 1, 0);
  for (int c : F.p_col) F.r_rowptr[c + 1]++;
  for (int k = 0; k < C.ndof; k++) F.r_rowptr[k + 1] += F.r_rowptr[k];
  F.r_col.resize(F.p_col.size());
  F.r_val.resize(F.p_col.size());
  std::vector<int> pos(F.r_rowptr.begin(), F.r_rowptr.end() - 1);
  for (int i = 0; i < F.ndof; i++)
    for (int t = F.p_rowptr[i]; t < F.p_rowptr[i + 1]; t++) {
      int k = F.p_col[t], q = pos[k]++;
      F.r_col[q] = i;
      F.r_val[q] = F.p_val[t];
    }
}

// CSR position of every local matrix entry (row DOF i, column DOF j of the
// cell), -1 on Dirichlet rows -- the column search of AddLocalSqMatToGlobal
// (AssembleMat3D.C:765-787) done once at setup.
void build_slots(const HostLevel &L, int order, std::vector<int> &slots) {
  const int nb = (order + 1) * (order + 1) * (order + 1);
  slots.assign((size_t)L.ncell * nb * nb, -1);
  parallel_for(L.ncell, [&](int c) {
    const int *g = &L.gn[(size_t)c * nb];
    for (int i = 0; i < nb; i++) {
      int r = g[i];
      if (r >= L.nact) continue;
      const int *b = &L.kcol[L.rowptr[r]], *e = &L.kcol[L.rowptr[r + 1]];
      for (int j = 0; j < nb; j++)
        slots[((size_t)c * nb + i) * nb + j] = (int)(std::lower_bound(b, e, g[j]) - &L.kcol[0]);
    }
  });
}

// Gather maps of the two-pass cell-by-cell assembly.  The element kernel
// writes row i of cell c's element matrix to E[(c nb + i) nb ..] and its rhs
// entry to Er[c nb + i]; CSR entry e of an active row then sums
// E[gidx[gptr[e] .. gptr[e+1])] and active row r sums Er[ridx[rptr[r] ..
// rptr[r+1])].  The lists run over the cells by colour, so every entry adds
// its contributions in the order of the colour-by-colour scatter.
void build_gather(const HostLevel &L, int order, const std::vector<int> &slots, std::vector<int> &gptr,
                  std::vector<int> &gidx, std::vector<int> &rptr, std::vector<int> &ridx) {
  const int nb = (order + 1) * (order + 1) * (order + 1);
  const int64_t nnz_act = L.rowptr[L.nact];
  gptr.assign(nnz_act + 1, 0);
  rptr.assign(L.nact + 1, 0);
  // cells of one colour share no DOF, so within a colour every row and entry
  // is touched by one cell at most: the colour's cells run in parallel
  auto each_colour = [&](auto &&cell) {
    for (int k = 0; k < 8; k++) {
      const int b = L.colour_start[k];
      parallel_for(L.colour_start[k + 1] - b, [&](int q) { cell(L.colour_cells[b + q]); });
    }
  };
  each_colour([&](int c) {
    for (int i = 0; i < nb; i++) {
      const int r = L.gn[(size_t)c * nb + i];
      if (r >= L.nact) continue;
      rptr[r + 1]++;
      for (int j = 0; j < nb; j++) gptr[slots[((size_t)c * nb + i) * nb + j] + 1]++;
    }
  });
  for (int64_t e = 0; e < nnz_act; e++) gptr[e + 1] += gptr[e];
  for (int r = 0; r < L.nact; r++) rptr[r + 1] += rptr[r];
  gidx.resize(gptr[nnz_act]);
  ridx.resize(rptr[L.nact]);
  std::vector<int> gp(gptr.begin(), gptr.end() - 1), rp(rptr.begin(), rptr.end() - 1);
  each_colour([&](int c) {
    for (int i = 0; i < nb; i++) {
      const int r = L.gn[(size_t)c * nb + i];
      if (r >= L.nact) continue;
      const int row = c * nb + i;
      ridx[rp[r]++] = row;
      for (int j = 0; j < nb; j++) gidx[gp[slots[(size_t)row * nb + j]]++] = row * nb + j;
    }
  });
}

// 8 colours by the parity of the cell's integer position: two cells of one
// colour never share a node, so one colour can be scattered without atomics.
void build_colours(HostLevel &L, int order) {
  const int p = order;
  std::vector<int> col(L.ncell);
  std::vector<int> cnt(8, 0);
  for (int c = 0; c < L.ncell; c++) {
    int par = 0;
    for (int d = 0; d < 3; d++) {
      int s = L.ax[9 * c + d] + L.ax[9 * c + 3 + d] + L.ax[9 * c + 6 + d];  // sign of the axis along d
      int lo = L.org[3 * c + d] + (s < 0 ? -p : 0);
      par |= ((lo / p) & 1) << d;
    }
    col[c] = par;
    cnt[par]++;
  }
  L.colour_start.assign(9, 0);
  for (int k = 0; k < 8; k++) L.colour_start[k + 1] = L.colour_start[k] + cnt[k];
  L.colour_cells.resize(L.ncell);
  std::vector<int> pos(L.colour_start.begin(), L.colour_start.end() - 1);
  for (int c = 0; c < L.ncell; c++) L.colour_cells[pos[col[c]]++] = c;
}

// Row lists of the multicolour SSOR smoother: every active row by the colour
// of its lattice node, X%k + k (Y%k + k Z%k) with k = order + 1; rows of one
// colour never share a cell, so their matrix rows never couple.
void build_mc_colours(HostLevel &L, int order) {
  const int k = order + 1, ncol = k * k * k, nn = L.nn;
  std::vector<int> colour(L.nact);
  L.mc_start.assign(ncol + 1, 0);
  for (int i = 0; i < L.nact; i++) {
    const int node = L.dof_node[i], X = node % nn, Y = (node / nn) % nn, Z = node / (nn * nn);
    colour[i] = X % k + k * (Y % k + k * (Z % k));
    L.mc_start[colour[i] + 1]++;
  }
  for (int c = 0; c < ncol; c++) L.mc_start[c + 1] += L.mc_start[c];
  L.mc_rows.resize(L.nact);
  std::vector<int> pos(L.mc_start.begin(), L.mc_start.end() - 1);
  for (int i = 0; i < L.nact; i++) L.mc_rows[pos[colour[i]]++] = i;
}

// Reference basis at the quadrature points, per point q:
//   [w_q, xi, eta, zeta, then per basis b: phi, dphi/dxi, dphi/deta, dphi/dzeta]
// Gauss rules in the point order of QuadFormulaHexa.C:62-120; tensor-product
// Lagrange basis, x fastest (BF_C_H_Q1_3D.h, BF_C_H_Q2_3D.h).
void basis_tables(int order, std::vector<double> &tab) {
  static const double g3x[3] = {-0.7745966692414833770358530, 0, 0.7745966692414833770358530};
  static const double g3w[3] = {0.5555555555555555555555558, 0.8888888888888888888888888,
                                0.5555555555555555555555558};
  static const double g4x[4] = {0.33998104358485626480266576, 0.86113631159405257522394649,
                                -0.86113631159405257522394649, -0.33998104358485626480266576};
  static const double g4w[4] = {0.652145154862546142626936051, 0.347854845137453857373063949,
                                0.347854845137453857373063949, 0.652145154862546142626936051};
  const int p = order, n1 = p + 2, nb = (p + 1) * (p + 1) * (p + 1), stride = 4 + 4 * nb;
  const double *gx = p == 1 ? g3x : g4x, *gw = p == 1 ? g3w : g4w;
  tab.assign((size_t)n1 * n1 * n1 * stride, 0.0);
  auto lag = [&](double x, double *v, double *d) {
    if (p == 1) { v[0] = 0.5 - 0.5 * x; v[1] = 0.5 + 0.5 * x; d[0] = -0.5; d[1] = 0.5; }
    else {
      v[0] = 0.5 * x * (x - 1); v[1] = 1 - x * x; v[2] = 0.5 * x * (x + 1);
      d[0] = x - 0.5; d[1] = -2 * x; d[2] = x + 0.5;
    }
  };
  int q = 0;
  for (int qi = 0; qi < n1; qi++)
    for (int qj = 0; qj < n1; qj++)
      for (int qk = 0; qk < n1; qk++, q++) {
        double xi = gx[qk], eta = gx[qj], zeta = gx[qi];
        double *t = &tab[(size_t)q * stride];
        t[0] = gw[qi] * gw[qj] * gw[qk];
        t[1] = xi; t[2] = eta; t[3] = zeta;
        double vx[3], dx[3], vy[3], dy[3], vz[3], dz[3];
        lag(xi, vx, dx); lag(eta, vy, dy); lag(zeta, vz, dz);
        int b = 0;
        for (int k = 0; k <= p; k++)
          for (int j = 0; j <= p; j++)
            for (int i = 0; i <= p; i++, b++) {
              t[4 + 4 * b + 0] = vx[i] * vy[j] * vz[k];
              t[4 + 4 * b + 1] = dx[i] * vy[j] * vz[k];
              t[4 + 4 * b + 2] = vx[i] * dy[j] * vz[k];
              t[4 + 4 * b + 3] = vx[i] * vy[j] * dz[k];
            }
      }
}

}  // namespace cd3d
