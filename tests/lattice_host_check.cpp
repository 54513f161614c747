// Host check of the lattice form's index arithmetic (the __host__ __device__
// node functions of lattice.hip) against the CSR matrices and transfer
// operators that setup.cpp builds: class stencils against CSR rows, compact
// against per-node stencils, restriction and prolongation against R and P, stencil ids against the lattice map, and the folding of one
// element matrix into class stencils against cell-by-cell accumulation.
// Built by tests/test_lattice_host.py with hipcc (host code only runs).
#include "lattice.hip"
#include <cstdio>
#include <cstring>
#include <random>
using namespace cd3d;

static double rnd(std::mt19937 &g) { return std::uniform_real_distribution<double>(-1, 1)(g); }

template <int P>
int check(int m, int k0) {
  HostLevel C, F;
  build_host_level(C, P, m, k0);
  build_host_level(F, P, m, k0 + 1);
  build_transfer(F, C, P);
  std::mt19937 g(7);
  int bad = 0;
  // ---- stencil: defect with f = 0 gives -A x on interior nodes
  const int nn = F.nn;
  const size_t nn3 = (size_t)nn * nn * nn;
  std::vector<double> val(F.nnz);
  for (auto &v : val) v = rnd(g);
  std::vector<float> lval(F.lat_nval);
  for (int64_t i = 0; i < F.lat_nval; i++) {
    if (F.lat_perm[i] < 0 || F.lat_perm[i] >= F.nnz) { printf("perm out of range\n"); return 1; }
    lval[i] = (float)val[F.lat_perm[i]];
  }
  std::vector<float> x(nn3, 0.f), f(nn3, 0.f), out(nn3, 0.f);
  std::vector<double> xd(F.ndof, 0.0);
  for (int d = 0; d < F.nact; d++) { xd[d] = rnd(g); x[F.dof_node[d]] = (float)xd[d]; }
  DevLevel D;
  D.nn = nn;
  D.nclass = (int)F.lat.size();
  for (int c = 0; c < D.nclass; c++) D.cls[c] = F.lat[c];
  LatArgs a = lat_args(D);
  // node2dof consistency
  for (int c = 0; c < D.nclass; c++) {
    const LatClass &L = F.lat[c];
    for (int64_t k = 0; k < L.nodes; k++) {
      int kx = k % L.cnt[0], ky = (k / L.cnt[0]) % L.cnt[1], kz = k / ((int64_t)L.cnt[0] * L.cnt[1]);
      int X = L.lo[0] + L.step * kx, Y = L.lo[1] + L.step * ky, Z = L.lo[2] + L.step * kz;
      int node = X + nn * (Y + nn * Z);
      int dof = F.lat_node2dof[L.first + k];
      if (dof < 0 || dof >= F.nact || F.dof_node[dof] != node) bad++;
    }
    for (int64_t k = 0; k < L.nodes; k++) lat_class_node<P, kLDefect>(a, c, k, lval.data(), nullptr, nullptr, x.data(), f.data(), out.data(), 0.f);
  }
  // stencil ids and first-node positions agree with the lattice map
  for (int c = 0; c < D.nclass; c++) {
    const LatClass &L = F.lat[c];
    for (int s2 = 0; s2 < L.nsten; s2++) {
      if (F.lat_first[c * 128 + s2] != F.lat_perm[L.off + (int64_t)s2 * L.nodes]) bad++;
      for (int64_t k = 0; k < L.nodes; k++)
        if (F.lat_sid[F.lat_perm[L.off + (int64_t)s2 * L.nodes + k]] != c * 128 + s2) bad++;
    }
  }
  if ((int64_t)F.lat_sid.size() != F.rowptr[F.nact]) { printf("sid size\n"); bad++; }
  int64_t nint = 0;
  for (auto &L : F.lat) nint += L.nodes;
  if (nint != F.nact) { printf("interior count %lld != nact %d\n", (long long)nint, F.nact); bad++; }
  double emax = 0;
  for (int d = 0; d < F.nact; d++) {
    double y = 0;
    for (int k = F.rowptr[d]; k < F.rowptr[d + 1]; k++) y += val[k] * xd[F.kcol[k]];
    emax = std::max(emax, fabs(-y - out[F.dof_node[d]]));
  }
  // boundary entries untouched
  for (int d = F.nact; d < F.ndof; d++) if (out[F.dof_node[d]] != 0.f) bad++;
  printf("P%d m%d k0=%d nn=%d stencil max err %.3g  (nval %lld nnz %lld)\n", P, m, k0, nn, emax,
         (long long)F.lat_nval, (long long)F.nnz);
  if (emax > 1e-4) bad++;
  // ---- compact stencil: values that depend only on (class, offset)
  {
    auto cls_of = [&](int X, int Y, int Z) { return P == 1 ? 0 : (X & 1) + 2 * (Y & 1) + 4 * (Z & 1); };
    std::vector<double> uval(F.nnz, 0.0);
    for (int d = 0; d < F.nact; d++) {
      int nd = F.dof_node[d], X = nd % nn, Y = (nd / nn) % nn, Z = nd / (nn * nn);
      for (int k = F.rowptr[d]; k < F.rowptr[d + 1]; k++) {
        int nc = F.dof_node[F.kcol[k]], dx = nc % nn - X, dy = (nc / nn) % nn - Y, dz = nc / (nn * nn) - Z;
        uval[k] = 1.0 + cls_of(X, Y, Z) * 0.37 + dx * 0.11 + dy * 0.013 + dz * 0.0017 + (dx == 0 && dy == 0 && dz == 0) * 5;
      }
    }
    std::vector<float> cst(8 * kCst, 0.f);
    for (int c = 0; c < D.nclass; c++) {
      const LatClass &L = F.lat[c];
      for (int s2 = 0; s2 < L.nsten; s2++) cst[c * kCst + s2] = (float)uval[F.lat_perm[L.off + (int64_t)s2 * L.nodes]];
      a.uni[c] = 1;
    }
    std::fill(out.begin(), out.end(), 0.f);
    for (int c = 0; c < D.nclass; c++)
      for (int64_t k = 0; k < F.lat[c].nodes; k++)
        lat_class_node<P, kLDefect>(a, c, k, nullptr, nullptr, cst.data(), x.data(), f.data(), out.data(), 0.f);
    double e2 = 0;
    for (int d = 0; d < F.nact; d++) {
      double y = 0;
      for (int k = F.rowptr[d]; k < F.rowptr[d + 1]; k++) y += uval[k] * xd[F.kcol[k]];
      e2 = std::max(e2, fabs(-y - out[F.dof_node[d]]) / (1 + fabs(y)));
    }
    printf("   compact stencil max rel err %.3g\n", e2);
    if (e2 > 1e-5) bad++;
    for (int c = 0; c < D.nclass; c++) a.uni[c] = 0;
  }
  // ---- class stencils folded from one element matrix == cell-by-cell assembly
  {
    constexpr int P1 = P + 1, NB = P1 * P1 * P1;
    std::vector<double> E(NB * NB + NB);
    for (auto &v : E) v = rnd(g);
    std::vector<double> aval(F.nnz, 0.0), arhs(F.ndof, 0.0);
    std::vector<int> nd2dof((size_t)nn * nn * nn, -1);
    for (int d = 0; d < F.ndof; d++) nd2dof[F.dof_node[d]] = d;
    const int ncs = (nn - 1) / P;
    for (int cz = 0; cz < ncs; cz++) for (int cy = 0; cy < ncs; cy++) for (int cx = 0; cx < ncs; cx++)
      for (int li = 0; li < NB; li++) {
        int ix = li % P1, iy = (li / P1) % P1, iz = li / (P1 * P1);
        int r = nd2dof[(P * cx + ix) + nn * ((P * cy + iy) + nn * (P * cz + iz))];
        if (r >= F.nact) continue;
        arhs[r] += E[NB * NB + li];
        for (int lj = 0; lj < NB; lj++) {
          int jx = lj % P1, jy = (lj / P1) % P1, jz = lj / (P1 * P1);
          int cc = nd2dof[(P * cx + jx) + nn * ((P * cy + jy) + nn * (P * cz + jz))];
          const int *b = &F.kcol[F.rowptr[r]], *e = &F.kcol[F.rowptr[r + 1]];
          const int *q = std::lower_bound(b, e, cc);
          if (q == e || *q != cc) { printf("missing column\n"); bad++; continue; }
          aval[q - &F.kcol[0]] += E[li * NB + lj];
        }
      }
    std::vector<double> st(8 * kCst + 8, 0.0), fval(F.nnz, -99.0), frhs(F.ndof, -99.0);
    for (int c = 0; c < D.nclass; c++) {
      const LatClass &L = F.lat[c];
      int nl[3], l[3][2], lo[3], len[3];
      for (int d = 0; d < 3; d++) class_axis<P>(L.px[d], &nl[d], l[d], &lo[d], &len[d]);
      if (len[0] * len[1] * len[2] != L.nsten) { printf("nsten mismatch\n"); bad++; }
      for (int s2 = 0; s2 < L.nsten; s2++) {
        const int dd[3] = {lo[0] + s2 % len[0], lo[1] + (s2 / len[0]) % len[1], lo[2] + s2 / (len[0] * len[1])};
        st[c * kCst + s2] = stencil_entry<P>(E.data(), nl, l, dd);
      }
      double r = 0;
      for (int az = 0; az < nl[2]; az++) for (int ay = 0; ay < nl[1]; ay++) for (int ax = 0; ax < nl[0]; ax++)
        r += E[NB * NB + l[0][ax] + P1 * (l[1][ay] + P1 * l[2][az])];
      st[8 * kCst + c] = r;
      for (int s2 = 0; s2 < L.nsten; s2++)
        for (int64_t k = 0; k < L.nodes; k++) fval[F.lat_perm[L.off + s2 * L.nodes + k]] = st[c * kCst + s2];
      for (int64_t k = 0; k < L.nodes; k++) frhs[F.lat_node2dof[L.first + k]] = st[8 * kCst + c];
    }
    double e3 = 0;
    for (int d = 0; d < F.nact; d++) {
      for (int k = F.rowptr[d]; k < F.rowptr[d + 1]; k++) e3 = std::max(e3, fabs(fval[k] - aval[k]));
      e3 = std::max(e3, fabs(frhs[d] - arhs[d]));
    }
    printf("   folded class stencils max err %.3g\n", e3);
    if (e3 > 1e-12) bad++;
  }
  // ---- affine class stencils == cell-by-cell accumulation with coefficients
  //      at the cell centres: A_cell = E0 + beta1(cell) Dx + beta2(cell) Dy
  {
    constexpr int P1 = P + 1, NB = P1 * P1 * P1;
    std::vector<double> E0(NB * NB), Dx(NB * NB), Dy(NB * NB);
    for (auto *v : {&E0, &Dx, &Dy})
      for (auto &e : *v) e = rnd(g);
    const double dnx = 0.37, dny = 0.29, r0 = 0.41, r1 = 0.53;
    std::vector<double> aval(F.nnz, 0.0);
    std::vector<int> nd2dof((size_t)nn * nn * nn, -1);
    for (int d = 0; d < F.ndof; d++) nd2dof[F.dof_node[d]] = d;
    const int ncs = (nn - 1) / P;
    for (int cz = 0; cz < ncs; cz++) for (int cy = 0; cy < ncs; cy++) for (int cx = 0; cx < ncs; cx++) {
      const double xc = dnx * (P * cx + 0.5 * P), yc = dny * (P * cy + 0.5 * P);
      const double b1 = -(yc - r1), b2 = xc - r0;
      for (int li = 0; li < NB; li++) {
        int ix = li % P1, iy = (li / P1) % P1, iz = li / (P1 * P1);
        int r = nd2dof[(P * cx + ix) + nn * ((P * cy + iy) + nn * (P * cz + iz))];
        if (r >= F.nact) continue;
        for (int lj = 0; lj < NB; lj++) {
          int jx = lj % P1, jy = (lj / P1) % P1, jz = lj / (P1 * P1);
          int cc = nd2dof[(P * cx + jx) + nn * ((P * cy + jy) + nn * (P * cz + jz))];
          const int *b = &F.kcol[F.rowptr[r]], *e = &F.kcol[F.rowptr[r + 1]];
          const int *q = std::lower_bound(b, e, cc);
          const int ent = li * NB + lj;
          aval[q - &F.kcol[0]] += E0[ent] + b1 * Dx[ent] + b2 * Dy[ent];
        }
      }
    }
    std::vector<double> st(3 * 8 * kCst + 8, 0.0);
    for (int c = 0; c < D.nclass; c++) {
      const LatClass &L = F.lat[c];
      int nl[3], l[3][2], lo[3], len[3];
      for (int d = 0; d < 3; d++) class_axis<P>(L.px[d], &nl[d], l[d], &lo[d], &len[d]);
      for (int s2 = 0; s2 < L.nsten; s2++) {
        const int dd[3] = {lo[0] + s2 % len[0], lo[1] + (s2 / len[0]) % len[1], lo[2] + s2 / (len[0] * len[1])};
        affine_stencil_entry<P>(E0.data(), Dx.data(), Dy.data(), nl, l, dd, dnx, dny, r0, r1, &st[c * kCst + s2],
                                &st[8 * kCst + 8 + c * kCst + s2], &st[2 * 8 * kCst + 8 + c * kCst + s2]);
      }
    }
    double e6 = 0, scale = 0;
    for (int d = 0; d < F.nact; d++) {
      const int node = F.dof_node[d];
      for (int k = F.rowptr[d]; k < F.rowptr[d + 1]; k++) {
        const double v = affine_entry(st.data(), F.lat_sid[k], node % nn, (node / nn) % nn);
        e6 = std::max(e6, fabs(v - aval[k]));
        scale = std::max(scale, fabs(aval[k]));
      }
    }
    printf("   affine class stencils max err %.3g (scale %.3g)\n", e6, scale);
    if (e6 > 1e-12 * scale) bad++;
  }
  // ---- restriction: R = P^T, Dirichlet coarse rows zero
  const int nc = C.nn;
  std::vector<float> fl(nn3, 0.f), cl((size_t)nc * nc * nc, 0.f);
  std::vector<double> fd(F.ndof, 0.0);
  for (int d = 0; d < F.nact; d++) { fd[d] = rnd(g); fl[F.dof_node[d]] = (float)fd[d]; }
  {
    DevLevel DC;
    DC.nn = C.nn;
    DC.nclass = (int)C.lat.size();
    for (int c = 0; c < DC.nclass; c++) DC.cls[c] = C.lat[c];
    LatArgs ca = lat_args(DC);
    // uniform class stencils: the restriction also writes the zero-start
    // sweep omega d_c f, which must be the bits of the kLJacobiZero sweep
    std::vector<float> cst(8 * kCst, 0.f), zc((size_t)nc * nc * nc, 0.f), zref(zc.size(), 0.f);
    for (int c = 0; c < 8; c++) cst[c * kCst + kCst - 1] = (float)(0.5 + 0.5 * std::fabs(rnd(g)));
    const float om = 0.7f;
    const ZeroSweep zs{zc.data(), cst.data(), om};
    for (int c = 0; c < DC.nclass; c++)
      for (int64_t k = 0; k < C.lat[c].nodes; k++) restrict_class_node<P>(ca, c, k, nn, fl.data(), cl.data(), zs);
    LatArgs cu = ca;
    for (int c = 0; c < DC.nclass; c++) cu.uni[c] = 1;
    for (int c = 0; c < DC.nclass; c++)
      for (int64_t k = 0; k < C.lat[c].nodes; k++)
        lat_class_node<P, kLJacobiZero>(cu, c, k, nullptr, nullptr, cst.data(), nullptr, cl.data(), zref.data(), om);
    int64_t zd = 0;
    for (size_t i = 0; i < zc.size(); i++) zd += memcmp(&zc[i], &zref[i], sizeof(float)) != 0;
    printf("   restriction's zero-start sweep vs the sweep: %lld differ\n", (long long)zd);
    if (zd) bad++;
    if (P == 2) {  // by coarse vertex/midpoint pairs: the same bits
      std::vector<float> cp((size_t)nc * nc * nc, 0.f), zp(zc.size(), 0.f);
      const ZeroSweep zsp{zp.data(), cst.data(), om};
      for (int pc = 0; pc < 4; pc++)
        for (int64_t k = 0; k < ca.c[2 * pc + 1].nodes; k++)
          restrict_pair_node(ca, pc, k, nn, fl.data(), cp.data(), zsp);
      int64_t rd = 0;
      for (size_t i = 0; i < cp.size(); i++)
        rd += (memcmp(&cp[i], &cl[i], sizeof(float)) != 0) + (memcmp(&zp[i], &zref[i], sizeof(float)) != 0);
      printf("   restriction by pairs (and its zero-start sweep): %lld differ\n", (long long)rd);
      if (rd) bad++;
    }
  }
  emax = 0;
  for (int d = 0; d < C.ndof; d++) {
    double y = 0;
    for (int k = F.r_rowptr[d]; k < F.r_rowptr[d + 1]; k++) y += F.r_val[k] * fd[F.r_col[k]];
    if (d >= C.nact) y = 0;
    emax = std::max(emax, fabs(y - cl[C.dof_node[d]]));
  }
  printf("   restriction max err %.3g\n", emax);
  if (emax > 1e-4) bad++;
  // ---- prolongation: fine interior += P c (coarse Dirichlet zero)
  std::vector<double> cd(C.ndof, 0.0);
  std::fill(cl.begin(), cl.end(), 0.f);
  for (int d = 0; d < C.nact; d++) { cd[d] = rnd(g); cl[C.dof_node[d]] = (float)cd[d]; }
  std::vector<float> xl(nn3, 0.f);
  const long long ncc = (nc - 1) / P;
  for (long long t = 0; t < ncc * ncc * ncc; t++) prolong_cell<P>(t, (int)ncc, nn, nc, cl.data(), xl.data(), 1.f);
  emax = 0;
  for (int d = 0; d < F.ndof; d++) {
    double y = 0;
    for (int k = F.p_rowptr[d]; k < F.p_rowptr[d + 1]; k++) y += F.p_val[k] * cd[F.p_col[k]];
    if (d >= F.nact) y = 0;
    emax = std::max(emax, fabs(y - xl[F.dof_node[d]]));
  }
  printf("   prolongation max err %.3g\n", emax);
  if (emax > 1e-4) bad++;
  // the row form (one thread per cell, fine Z, fine Y) writes the same bits,
  // onto an arbitrary fine vector, and nothing else
  {
    std::vector<float> x1(nn3), x2;
    for (auto &v : x1) v = (float)rnd(g);
    x2 = x1;
    for (long long t = 0; t < ncc * ncc * ncc; t++) prolong_cell<P>(t, (int)ncc, nn, nc, cl.data(), x1.data(), 0.7f);
    for (long long t = 0; t < ncc * ncc * ncc * 4 * P * P; t++)
      prolong_row<P>(t, (int)ncc, nn, nc, cl.data(), x2.data(), 0.7f);
    int64_t diff = 0;
    for (size_t i = 0; i < nn3; i++) diff += memcmp(&x1[i], &x2[i], sizeof(float)) != 0;
    printf("   prolongation rows vs cells: %lld differ\n", (long long)diff);
    if (diff) bad++;
  }
  return bad;
}

// Q2: the vertex/midpoint pair sweep writes exactly what the per-class sweep
// writes (bitwise), in every mode, with compact and with per-node stencils;
// x, f and the output's untouched entries hold arbitrary values everywhere.
int check_pairs(int m, int k0) {
  HostLevel F;
  build_host_level(F, 2, m, k0);
  const int nn = F.nn;
  const size_t nn3 = (size_t)nn * nn * nn;
  std::mt19937 g(11);
  std::vector<float> x(nn3), f(nn3), lval(F.lat_nval), inv(F.nact), cst(8 * kCst);
  for (auto &v : x) v = (float)rnd(g);
  for (auto &v : f) v = (float)rnd(g);
  for (auto &v : lval) v = (float)rnd(g);
  for (auto &v : inv) v = (float)rnd(g);
  for (auto &v : cst) v = (float)rnd(g);
  DevLevel D;
  D.nn = nn;
  D.nclass = (int)F.lat.size();
  for (int c = 0; c < D.nclass; c++) D.cls[c] = F.lat[c];
  int bad = 0;
  for (int uni = 0; uni < 2; uni++) {
    LatArgs a = lat_args(D);
    for (int c = 0; c < 8; c++) a.uni[c] = uni;
    int64_t npairs = 0;
    for (int pc = 0; pc < 4; pc++) npairs += a.c[2 * pc + 1].nodes;
    if (a.pstart[4] != (int)((a.c[1].nodes + kB - 1) / kB + (a.c[3].nodes + kB - 1) / kB +
                             (a.c[5].nodes + kB - 1) / kB + (a.c[7].nodes + kB - 1) / kB))
      bad++;
    for (int mode = 0; mode < 3; mode++) {
      std::vector<float> o1(nn3, 7.f), o2(nn3, 7.f);
      for (int c = 0; c < 8; c++)
        for (int64_t k = 0; k < a.c[c].nodes; k++) {
          if (mode == kLJacobi) lat_class_node<2, kLJacobi>(a, c, k, lval.data(), inv.data(), cst.data(), x.data(), f.data(), o1.data(), 0.7f);
          if (mode == kLJacobiZero) lat_class_node<2, kLJacobiZero>(a, c, k, lval.data(), inv.data(), cst.data(), x.data(), f.data(), o1.data(), 0.7f);
          if (mode == kLDefect) lat_class_node<2, kLDefect>(a, c, k, lval.data(), inv.data(), cst.data(), x.data(), f.data(), o1.data(), 0.7f);
        }
      for (int pc = 0; pc < 4; pc++)
        for (int64_t k = 0; k < a.c[2 * pc + 1].nodes; k++) {
          if (mode == kLJacobi) (uni ? lat_pair_node<kLJacobi, true> : lat_pair_node<kLJacobi, false>)(a, pc, k, lval.data(), inv.data(), cst.data() + 2 * pc * kCst, cst.data() + (2 * pc + 1) * kCst, x.data(), f.data(), o2.data(), 0.7f);
          if (mode == kLJacobiZero) (uni ? lat_pair_node<kLJacobiZero, true> : lat_pair_node<kLJacobiZero, false>)(a, pc, k, lval.data(), inv.data(), cst.data() + 2 * pc * kCst, cst.data() + (2 * pc + 1) * kCst, x.data(), f.data(), o2.data(), 0.7f);
          if (mode == kLDefect) (uni ? lat_pair_node<kLDefect, true> : lat_pair_node<kLDefect, false>)(a, pc, k, lval.data(), inv.data(), cst.data() + 2 * pc * kCst, cst.data() + (2 * pc + 1) * kCst, x.data(), f.data(), o2.data(), 0.7f);
        }
      int64_t diff = 0, written = 0;
      for (size_t i = 0; i < nn3; i++) {
        diff += memcmp(&o1[i], &o2[i], sizeof(float)) != 0;
        written += o1[i] != 7.f;
      }
      printf("   pairs m%d k0=%d uni %d mode %d: %lld pairs, %lld nodes written, %lld differ\n", m, k0, uni, mode,
             (long long)npairs, (long long)written, (long long)diff);
      if (diff || written < F.nact - 2) bad++;
      // groups of 2 and 4 pairs per thread (compact stencils) against single pairs
      for (int G = 2; uni && G <= 4; G += 2) {
        if (a.c[1].cnt[0] % G) continue;
        std::vector<float> o3(nn3, 7.f);
        for (int pc = 0; pc < 4; pc++)
          for (int64_t k = 0; k < a.c[2 * pc + 1].nodes / G; k++) {
            const float *c0 = cst.data() + 2 * pc * kCst, *c1 = cst.data() + (2 * pc + 1) * kCst;
#define CD3D_GRP(GG, M) lat_pair_group_node<GG, M>(a, pc, k, c0, c1, x.data(), f.data(), o3.data(), 0.7f)
            if (G == 2 && mode == kLJacobi) CD3D_GRP(2, kLJacobi);
            if (G == 2 && mode == kLJacobiZero) CD3D_GRP(2, kLJacobiZero);
            if (G == 2 && mode == kLDefect) CD3D_GRP(2, kLDefect);
            if (G == 4 && mode == kLJacobi) CD3D_GRP(4, kLJacobi);
            if (G == 4 && mode == kLJacobiZero) CD3D_GRP(4, kLJacobiZero);
            if (G == 4 && mode == kLDefect) CD3D_GRP(4, kLDefect);
#undef CD3D_GRP
          }
        int64_t gd = 0;
        for (size_t i = 0; i < nn3; i++) gd += memcmp(&o2[i], &o3[i], sizeof(float)) != 0;
        printf("   groups of %d m%d k0=%d mode %d: %lld differ\n", G, m, k0, mode, (long long)gd);
        if (gd) bad++;
      }
      // quads of lines (G pairs on each of 4 neighbouring lines per thread) against single pairs
      for (int G = 1; uni && G <= 4; G *= 2) {
        const long long n = a.c[1].cnt[0];
        if (n % G) continue;
        std::vector<float> o4(nn3, 7.f);
        for (long long k = 0; k < n / G * n * n; k++) {
#define CD3D_QD(GG, M) lat_quad_node<GG, M>(a, k, cst.data(), x.data(), f.data(), o4.data(), 0.7f)
          if (G == 1 && mode == kLJacobi) CD3D_QD(1, kLJacobi);
          if (G == 1 && mode == kLJacobiZero) CD3D_QD(1, kLJacobiZero);
          if (G == 1 && mode == kLDefect) CD3D_QD(1, kLDefect);
          if (G == 2 && mode == kLJacobi) CD3D_QD(2, kLJacobi);
          if (G == 2 && mode == kLJacobiZero) CD3D_QD(2, kLJacobiZero);
          if (G == 2 && mode == kLDefect) CD3D_QD(2, kLDefect);
          if (G == 4 && mode == kLJacobi) CD3D_QD(4, kLJacobi);
          if (G == 4 && mode == kLJacobiZero) CD3D_QD(4, kLJacobiZero);
          if (G == 4 && mode == kLDefect) CD3D_QD(4, kLDefect);
#undef CD3D_QD
        }
        int64_t qd = 0;
        for (size_t i = 0; i < nn3; i++) qd += memcmp(&o2[i], &o4[i], sizeof(float)) != 0;
        printf("   quads of %d m%d k0=%d mode %d: %lld differ\n", G, m, k0, mode, (long long)qd);
        if (qd) bad++;
      }
    }
  }
  // affine stencils evaluated node by node against per-node values filled
  // with the same formula (affine_value32 on the FP32-rounded stencils, as
  // lat_fill_affine_k), every class and mode
  {
    std::vector<double> st(kLst64), stm(8 * kCst);
    for (auto &v : st) v = rnd(g);
    for (auto &v : stm) v = rnd(g);
    std::vector<float> st32(st.begin(), st.end()), stm32(stm.begin(), stm.end());
    LatArgs a = lat_args(D);
    for (int c = 0; c < 8; c++) a.uni[c] = 0;
    int centre[8];
    for (int c = 0; c < 8; c++) {
      int lo[3], len[3];
      for (int d = 0; d < 3; d++) {
        const bool mid = a.c[c].px[d] == 1;
        lo[d] = mid ? -1 : -2;
        len[d] = mid ? 3 : 5;
      }
      centre[c] = -lo[0] + len[0] * (-lo[1] + len[1] * -lo[2]);
    }
    const float alpha = 0.37f, alpha_m = 1.0f;
    std::vector<float> av(F.lat_nval), ai(F.nact);
    for (int c = 0; c < 8; c++) {
      const LatDev &C = a.c[c];
      const float *s0 = st32.data() + c * kCst, *sx = st32.data() + 8 * kCst + 8 + c * kCst,
                  *sy = st32.data() + 2 * 8 * kCst + 8 + c * kCst, *sm = stm32.data() + c * kCst;
      for (int s2 = 0; s2 < C.nsten; s2++)
        for (int64_t k = 0; k < C.nodes; k++) {
          const int X = C.lo[0] + C.step * (int)(k % C.cnt[0]), Y = C.lo[1] + C.step * (int)((k / C.cnt[0]) % C.cnt[1]);
          const float v = affine_value32(s0, sx, sy, sm, alpha, alpha_m, s2, X, Y);
          av[C.off + (int64_t)s2 * C.nodes + k] = v;
          if (s2 == centre[c]) ai[C.first + k] = 1.0f / v;
        }
    }
    for (int mode = 0; mode < 3; mode++) {
      std::vector<float> o1(nn3, 7.f), o2(nn3, 7.f);
      for (int c = 0; c < 8; c++) {
        const float *s0 = st32.data() + c * kCst, *sx = st32.data() + 8 * kCst + 8 + c * kCst,
                    *sy = st32.data() + 2 * 8 * kCst + 8 + c * kCst, *sm = stm32.data() + c * kCst;
        for (int64_t k = 0; k < a.c[c].nodes; k++) {
#define CD3D_BOTH(M)                                                                                       \
  lat_class_node<2, M>(a, c, k, av.data(), ai.data(), nullptr, x.data(), f.data(), o1.data(), 0.7f);       \
  lat_class_node_aff<2, M>(a, c, k, s0, sx, sy, sm, alpha, alpha_m, centre[c], x.data(), f.data(), o2.data(), 0.7f)
          if (mode == kLJacobi) { CD3D_BOTH(kLJacobi); }
          if (mode == kLJacobiZero) { CD3D_BOTH(kLJacobiZero); }
          if (mode == kLDefect) { CD3D_BOTH(kLDefect); }
#undef CD3D_BOTH
        }
      }
      int64_t diff = 0;
      for (size_t i = 0; i < nn3; i++) diff += memcmp(&o1[i], &o2[i], sizeof(float)) != 0;
      printf("   affine m%d k0=%d mode %d: %lld differ\n", m, k0, mode, (long long)diff);
      if (diff) bad++;
    }
  }
  // FP64 operator: pairs against the per-class nodes, both modes
  {
    std::vector<double> xd(nn3), fd(nn3), st(8 * kCst);
    for (auto &v : xd) v = rnd(g);
    for (auto &v : fd) v = rnd(g);
    for (auto &v : st) v = rnd(g);
    const LatArgs a = lat_args(D);
    for (int mode = 0; mode < 2; mode++) {
      std::vector<double> o1(nn3, 7.0), o2(nn3, 7.0);
      for (int c = 0; c < 8; c++)
        for (int64_t k = 0; k < a.c[c].nodes; k++) {
          const double *sc = st.data() + c * kCst;
#define CD3D_L64(PX, PY, PZ)                                                                          \
  (mode ? lat64_node<2, PX, PY, PZ, kL64Defect> : lat64_node<2, PX, PY, PZ, kL64Spmv>)(a, a.c[c], k, sc, \
                                                                                      xd.data(), fd.data(), o1.data())
          switch (c) {
            case 0: CD3D_L64(0, 0, 0); break;
            case 1: CD3D_L64(1, 0, 0); break;
            case 2: CD3D_L64(0, 1, 0); break;
            case 3: CD3D_L64(1, 1, 0); break;
            case 4: CD3D_L64(0, 0, 1); break;
            case 5: CD3D_L64(1, 0, 1); break;
            case 6: CD3D_L64(0, 1, 1); break;
            default: CD3D_L64(1, 1, 1); break;
          }
#undef CD3D_L64
        }
      for (int pc = 0; pc < 4; pc++)
        for (int64_t k = 0; k < a.c[2 * pc + 1].nodes; k++)
          (mode ? lat64_pair_node<kL64Defect> : lat64_pair_node<kL64Spmv>)(
              a, pc, k, st.data() + 2 * pc * kCst, st.data() + (2 * pc + 1) * kCst, xd.data(), fd.data(), o2.data());
      int64_t diff = 0;
      for (size_t i = 0; i < nn3; i++) diff += memcmp(&o1[i], &o2[i], sizeof(double)) != 0;
      printf("   f64 pairs m%d k0=%d mode %d: %lld differ\n", m, k0, mode, (long long)diff);
      if (diff) bad++;
      for (int G = 2; G <= 4; G += 2) {
        if (a.c[1].cnt[0] % G) continue;
        std::vector<double> o3(nn3, 7.0);
        for (int pc = 0; pc < 4; pc++)
          for (int64_t k = 0; k < a.c[2 * pc + 1].nodes / G; k++) {
            const double *c0 = st.data() + 2 * pc * kCst, *c1 = st.data() + (2 * pc + 1) * kCst;
            if (G == 2) (mode ? lat64_pair_group_node<2, kL64Defect> : lat64_pair_group_node<2, kL64Spmv>)(a, pc, k, c0, c1, xd.data(), fd.data(), o3.data());
            else (mode ? lat64_pair_group_node<4, kL64Defect> : lat64_pair_group_node<4, kL64Spmv>)(a, pc, k, c0, c1, xd.data(), fd.data(), o3.data());
          }
        int64_t gd = 0;
        for (size_t i = 0; i < nn3; i++) gd += memcmp(&o2[i], &o3[i], sizeof(double)) != 0;
        printf("   f64 groups of %d m%d k0=%d mode %d: %lld differ\n", G, m, k0, mode, (long long)gd);
        if (gd) bad++;
      }
    }
  }
  return bad;
}

int main() {
  int bad = 0;
  {  // the XCD-contiguous block order of the CSR kernels is a permutation
    int nb = 0;
    for (int n = 1; n <= 5000; n++) {
      std::vector<char> seen(n, 0);
      for (int b = 0; b < n; b++) {
        const int l = xcd_block(b, n);
        if (l < 0 || l >= n || seen[l]++) nb++;
      }
    }
    printf("   xcd_block: %d collisions\n", nb);
    if (nb) bad++;
  }
  bad += check_pairs(1, 0);
  bad += check_pairs(1, 2);
  bad += check_pairs(3, 1);
  bad += check_pairs(1, 5);  // 32^3 cells: many tiles per XCD
  bad += check<1>(1, 1);
  bad += check<1>(2, 1);
  bad += check<2>(1, 1);
  bad += check<2>(1, 2);
  bad += check<2>(3, 0);
  bad += check<2>(2, 1);
  printf(bad ? "FAIL %d\n" : "OK\n", bad);
  return bad != 0;
}
