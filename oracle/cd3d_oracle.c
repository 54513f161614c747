/* CPU restatement of ParMooN's CD3D/TCD3D hot path -- TEST INFRASTRUCTURE ONLY.
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark use
 * this file; the product never links or calls it.
 *
 * Each routine cites the reference routine it restates.  What pins each one
 * (reference fixtures in tests/golden/ (npz files), closed forms, invariants) is
 * listed in tests/test_oracle_golden.py and tests/test_oracle_math.py.
 */
#include "cd3d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXNB 27
#define MAXQ 64

typedef struct {
  int n;             /* cells per side */
  int ncell;
  double *vx;        /* mapped hierarchy: (n+1)^3 x 3 vertex coordinates, x fastest; NULL: the box */
  int *org;          /* ncell*3: cell vertex 0, in node-lattice units (spacing 1/(p n)) */
  int *ax;           /* ncell*9: ax[9c+3a+d] = component d of local axis a (signed unit) */
  int nn;            /* nodes per side p n + 1 */
  int ndof, nact;
  int *gn;           /* ncell*nb global DOF numbers (FESpace3D GlobalNumbers) */
  int *dof_node;     /* DOF -> lattice node index */
  int *rowptr, *kcol;
  long nnz;
  double *A;         /* matrix used by the multigrid level (A, or the TCD system M) */
  double *Aconv;     /* TCD: conv-diff matrix */
  double *Ksupg;     /* TCD SUPG: time-derivative SUPG matrix K */
  double *qp;        /* coefficients per quadrature point (ncell*nq*6: eps, b1, b2, b3, c, f) or NULL */
  double *rhs, *sol; /* level vectors (TMGLevel3D Rhs / X) */
  double *aux0, *aux1;
  double *mult;      /* number of cells sharing each DOF (Prolongate aux) */
} level_t;

struct orc_hier {
  int order, nb, L, m, k0;
  int mapped;  /* vertices given per macro vertex (general hexahedra, trilinear maps) */
  double box[3];
  level_t *lv;
  double *lu;  /* coarse dense LU (exact coarse solve) */
  int *piv;
  int lu_valid;
};

/* ------------------------------------------------------------------ tables */

/* Child frames of the regular hexahedron refinement (RefHexaRegDesc.C:29-33,
 * DatChildVertex): child j has its vertex 0 at parent vertex j; its local
 * axes (v1-v0, v3-v0, v4-v0) expressed in the parent's local axes. */
static const int CHILD_AX[8][3][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},   {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}},
    {{-1, 0, 0}, {0, -1, 0}, {0, 0, 1}}, {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 1, 0}, {1, 0, 0}, {0, 0, -1}},  {{-1, 0, 0}, {0, 1, 0}, {0, 0, -1}},
    {{0, -1, 0}, {-1, 0, 0}, {0, 0, -1}}, {{1, 0, 0}, {0, -1, 0}, {0, 0, -1}}};
/* hexahedron vertex j in local coordinates (Hexahedron shape: 0-3 bottom
 * counter-clockwise, 4-7 above them) */
static const int HEX_V[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* 1D Gauss rules in the point order of QuadFormulaHexa.C:62-120 */
static const double G3X[3] = {-0.7745966692414833770358530, 0, 0.7745966692414833770358530};
static const double G3W[3] = {0.5555555555555555555555558, 0.8888888888888888888888888,
                              0.5555555555555555555555558};
static const double G4X[4] = {0.33998104358485626480266576, 0.86113631159405257522394649,
                              -0.86113631159405257522394649, -0.33998104358485626480266576};
static const double G4W[4] = {0.652145154862546142626936051, 0.347854845137453857373063949,
                              0.347854845137453857373063949, 0.652145154862546142626936051};

/* 1D Lagrange basis on [-1,1] with nodes {-1,1} (p=1) or {-1,0,1} (p=2);
 * BF_C_H_Q1_3D.h / BF_C_H_Q2_3D.h are their tensor products, x fastest. */
static void lag1d(int p, double x, double *v, double *d) {
  if (p == 1) {
    v[0] = 0.5 - 0.5 * x; v[1] = 0.5 + 0.5 * x;
    d[0] = -0.5; d[1] = 0.5;
  } else {
    v[0] = 0.5 * x * (x - 1); v[1] = 1 - x * x; v[2] = 0.5 * x * (x + 1);
    d[0] = x - 0.5; d[1] = -2 * x; d[2] = x + 0.5;
  }
}

static void basis3d(int p, double xi, double eta, double zeta, double *phi, double *dxi, double *deta,
                    double *dzeta) {
  double vx[3], dx[3], vy[3], dy[3], vz[3], dz[3];
  lag1d(p, xi, vx, dx); lag1d(p, eta, vy, dy); lag1d(p, zeta, vz, dz);
  int q = p + 1, b = 0;
  for (int k = 0; k < q; k++)
    for (int j = 0; j < q; j++)
      for (int i = 0; i < q; i++, b++) {
        phi[b] = vx[i] * vy[j] * vz[k];
        if (dxi) {
          dxi[b] = dx[i] * vy[j] * vz[k];
          deta[b] = vx[i] * dy[j] * vz[k];
          dzeta[b] = vx[i] * vy[j] * dz[k];
        }
      }
}

/* ------------------------------------------------------------------ mesh */

/* Uniform refinement in depth-first child order (Domain.C:1037 RegRefineAll,
 * It_Finest leaf order): emit leaves with origin and axes. */
static void refine_emit(level_t *L, int *cnt, const int O[3], const int E[3][3], int S, int depth, int p) {
  if (depth == 0) {
    int c = (*cnt)++;
    for (int d = 0; d < 3; d++) L->org[3 * c + d] = O[d];
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++) L->ax[9 * c + 3 * a + d] = E[a][d];
    (void)S; (void)p;
    return;
  }
  for (int j = 0; j < 8; j++) {
    int Oc[3], Ec[3][3];
    for (int d = 0; d < 3; d++)
      Oc[d] = O[d] + S * (HEX_V[j][0] * E[0][d] + HEX_V[j][1] * E[1][d] + HEX_V[j][2] * E[2][d]);
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 3; d++)
        Ec[a][d] = CHILD_AX[j][a][0] * E[0][d] + CHILD_AX[j][a][1] * E[1][d] + CHILD_AX[j][a][2] * E[2][d];
    refine_emit(L, cnt, Oc, Ec, S / 2, depth - 1, p);
  }
}

/* level mesh: m^3 macro cells (x fastest, standard frame), each refined
 * `depth` times.  m = 1 is data/Wuerfel.GEO (vertex 0 at the origin, axes
 * x, y, z -- ReadGeo.C:1790-1812). */
static void build_mesh(level_t *L, int m, int depth, int p) {
  int n = m << depth;
  L->n = n;
  L->ncell = n * n * n;
  L->nn = p * n + 1;
  L->org = malloc(sizeof(int) * 3 * L->ncell);
  L->ax = malloc(sizeof(int) * 9 * L->ncell);
  int cnt = 0, S = p << depth;
  int E[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int k = 0; k < m; k++)
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) {
        int O[3] = {i * S, j * S, k * S};
        refine_emit(L, &cnt, O, E, S, depth, p);
      }
}

static void cell_vertices(const orc_hier *h, const level_t *L, int c, double V[8][3]);

static inline int cell_node(const level_t *L, int c, int p, int i, int j, int k) {
  const int *O = L->org + 3 * c, *E = L->ax + 9 * c;
  int X = O[0] + i * E[0] + j * E[3] + k * E[6];
  int Y = O[1] + i * E[1] + j * E[4] + k * E[7];
  int Z = O[2] + i * E[2] + j * E[5] + k * E[8];
  return X + L->nn * (Y + L->nn * Z);
}

/* DOF numbering (FESpace3D.C:459-2035, FE3DMapper.C:42-230).  The mappers
 * link every shared DOF to the lowest (cell, local DOF) position of its class
 * (MapDOF keeps MIN(w0,w1) as root); the final pass (FESpace3D.C:1794-1870)
 * numbers each root when it is first met, inner and Dirichlet classes with
 * separate counters, Dirichlet DOFs after all others.  A DOF is Dirichlet when
 * it lies on a boundary face (MapBound with DirichletCounter). */
static void build_dofs(level_t *L, int p) {
  int nb = (p + 1) * (p + 1) * (p + 1), nn = L->nn, nnodes = nn * nn * nn;
  int *node_dof = malloc(sizeof(int) * nnodes);
  for (int i = 0; i < nnodes; i++) node_dof[i] = -1;
  L->gn = malloc(sizeof(int) * nb * L->ncell);
  int n_in = 0, n_dir = 0;
  /* first pass: class-local first-touch numbers (Dirichlet ones negative) */
  for (int c = 0; c < L->ncell; c++)
    for (int b = 0; b < nb; b++) {
      int i = b % (p + 1), j = (b / (p + 1)) % (p + 1), k = b / ((p + 1) * (p + 1));
      int node = cell_node(L, c, p, i, j, k);
      if (node_dof[node] == -1) {
        int X = node % nn, Y = (node / nn) % nn, Z = node / (nn * nn);
        int bd = X == 0 || Y == 0 || Z == 0 || X == nn - 1 || Y == nn - 1 || Z == nn - 1;
        node_dof[node] = bd ? -2 - (n_dir++) : n_in++;
      }
      L->gn[c * nb + b] = node;
    }
  L->nact = n_in;
  L->ndof = n_in + n_dir;
  L->dof_node = malloc(sizeof(int) * L->ndof);
  for (int i = 0; i < nnodes; i++) {
    int v = node_dof[i];
    if (v == -1) continue;
    int d = v >= 0 ? v : n_in + (-2 - v);
    node_dof[i] = d;
    L->dof_node[d] = i;
  }
  for (int e = 0; e < nb * L->ncell; e++) L->gn[e] = node_dof[L->gn[e]];
  free(node_dof);
  L->mult = calloc(L->ndof, sizeof(double));
  for (int e = 0; e < nb * L->ncell; e++) L->mult[L->gn[e]] += 1;
}

static int cmp_int(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }

/* CSR pattern (SquareStructure3D.C:32-330, SquareStructure.C:65-84 Sort):
 * active row = sorted set of all DOFs sharing a cell with the row DOF,
 * Dirichlet columns included; Dirichlet row = its diagonal only. */
static void build_structure(level_t *L, int p) {
  int nb = (p + 1) * (p + 1) * (p + 1), cap = (2 * p + 1) * (2 * p + 1) * (2 * p + 1);
  int *cols = malloc(sizeof(int) * (size_t)cap * L->nact);
  int *len = calloc(L->nact, sizeof(int));
  for (int c = 0; c < L->ncell; c++) {
    const int *g = L->gn + c * nb;
    for (int a = 0; a < nb; a++) {
      int r = g[a];
      if (r >= L->nact) continue;
      int *row = cols + (size_t)cap * r;
      for (int b = 0; b < nb; b++) {
        int col = g[b], found = 0;
        for (int t = 0; t < len[r]; t++)
          if (row[t] == col) { found = 1; break; }
        if (!found) row[len[r]++] = col;
      }
    }
  }
  L->rowptr = malloc(sizeof(int) * (L->ndof + 1));
  L->rowptr[0] = 0;
  for (int r = 0; r < L->ndof; r++) L->rowptr[r + 1] = L->rowptr[r] + (r < L->nact ? len[r] : 1);
  L->nnz = L->rowptr[L->ndof];
  L->kcol = malloc(sizeof(int) * L->nnz);
  for (int r = 0; r < L->ndof; r++) {
    if (r < L->nact) {
      qsort(cols + (size_t)cap * r, len[r], sizeof(int), cmp_int);
      memcpy(L->kcol + L->rowptr[r], cols + (size_t)cap * r, sizeof(int) * len[r]);
    } else {
      L->kcol[L->rowptr[r]] = r;
    }
  }
  free(cols);
  free(len);
}

/* One regular refinement of the vertex lattice of an n^3 structured mesh
 * (Refinement.C, RefHexaRegDesc.C: the new vertex of an edge at
 * 0.5 v0 + 0.5 v1, of a face at the mean of its 4 vertices, inside a cell at
 * the mean of its 8 vertices, DatOldFaceNewVertexPos / DatNewVertexPos).
 * Sums run over the corners in lattice order (x fastest). */
static double *refine_vertices(const double *src, int n) {
  int ns = n + 1, nd = 2 * n + 1;
  double *dst = malloc(sizeof(double) * 3 * nd * nd * nd);
  for (int K = 0; K < nd; K++)
    for (int J = 0; J < nd; J++)
      for (int I = 0; I < nd; I++) {
        int odd[3] = {I & 1, J & 1, K & 1}, base[3] = {I >> 1, J >> 1, K >> 1};
        int nodd = odd[0] + odd[1] + odd[2];
        double wt = nodd == 0 ? 1.0 : nodd == 1 ? 0.5 : nodd == 2 ? 0.25 : 0.125;
        double *o = dst + 3 * (I + nd * (J + nd * K));
        o[0] = o[1] = o[2] = 0;
        for (int t = 0; t < 8; t++) {
          int e[3] = {t & 1, (t >> 1) & 1, (t >> 2) & 1}, skip = 0, v[3];
          for (int d = 0; d < 3; d++) {
            if (e[d] && !odd[d]) skip = 1;
            v[d] = base[d] + e[d];
          }
          if (skip) continue;
          const double *x = src + 3 * (v[0] + ns * (v[1] + ns * v[2]));
          for (int d = 0; d < 3; d++) o[d] += wt * x[d];
        }
      }
  return dst;
}

static orc_hier *create_(int order, int m, int k0, int levels, const double *box, const double *macro_xyz) {
  orc_hier *h = calloc(1, sizeof(orc_hier));
  for (int d = 0; d < 3; d++) h->box[d] = box ? box[d] : 1.0;
  h->order = order;
  h->nb = (order + 1) * (order + 1) * (order + 1);
  h->L = levels;
  h->m = m;
  h->k0 = k0;
  h->mapped = macro_xyz != NULL;
  h->lv = calloc(levels, sizeof(level_t));
  double *vx = NULL;
  int nv = m;
  if (macro_xyz) {
    vx = malloc(sizeof(double) * 3 * (m + 1) * (m + 1) * (m + 1));
    memcpy(vx, macro_xyz, sizeof(double) * 3 * (m + 1) * (m + 1) * (m + 1));
    for (int r = 0; r < k0; r++, nv *= 2) {
      double *t = refine_vertices(vx, nv);
      free(vx);
      vx = t;
    }
  }
  for (int l = 0; l < levels; l++) {
    level_t *L = h->lv + l;
    if (macro_xyz) {
      if (l > 0) {
        double *t = refine_vertices(vx, nv);
        vx = t;
        nv *= 2;
      }
      L->vx = vx;
    }
    build_mesh(L, m, k0 + l, order);
    build_dofs(L, order);
    build_structure(L, order);
    L->A = calloc(L->nnz, sizeof(double));
    L->rhs = calloc(L->ndof, sizeof(double));
    L->sol = calloc(L->ndof, sizeof(double));
    L->aux0 = calloc(L->ndof, sizeof(double));
    L->aux1 = calloc(L->ndof, sizeof(double));
  }
  if (macro_xyz && levels == 0) free(vx);
  return h;
}

orc_hier *orc_create(int order, int m, int k0, int levels, const double *box) {
  return create_(order, m, k0, levels, box, NULL);
}

orc_hier *orc_create_mapped(int order, int m, int k0, int levels, const double *macro_xyz) {
  return create_(order, m, k0, levels, NULL, macro_xyz);
}

void orc_destroy(orc_hier *h) {
  if (!h) return;
  for (int l = 0; l < h->L; l++) {
    level_t *L = h->lv + l;
    free(L->vx); free(L->org); free(L->ax); free(L->gn); free(L->dof_node); free(L->rowptr); free(L->kcol);
    free(L->A); free(L->Aconv); free(L->Ksupg); free(L->qp); free(L->rhs); free(L->sol); free(L->aux0); free(L->aux1); free(L->mult);
  }
  free(h->lv); free(h->lu); free(h->piv);
  free(h);
}

int orc_level_info(const orc_hier *h, int l, int *n_active, long *nnz, int *n_cells) {
  const level_t *L = h->lv + l;
  if (n_active) *n_active = L->nact;
  if (nnz) *nnz = L->nnz;
  if (n_cells) *n_cells = L->ncell;
  return L->ndof;
}

void orc_get_cells(const orc_hier *h, int l, double *xyz) {
  const level_t *L = h->lv + l;
  int S = h->order;
  if (L->vx) {
    for (int c = 0; c < L->ncell; c++) cell_vertices(h, L, c, (double(*)[3])(xyz + 24 * (size_t)c));
    return;
  }
  for (int c = 0; c < L->ncell; c++) {
    const int *O = L->org + 3 * c, *E = L->ax + 9 * c;
    for (int v = 0; v < 8; v++)
      for (int d = 0; d < 3; d++)
        xyz[24 * c + 3 * v + d] = (O[d] + S * (HEX_V[v][0] * E[d] + HEX_V[v][1] * E[3 + d] + HEX_V[v][2] * E[6 + d])) *
                                  (h->box[d] / (h->order * L->n));
  }
}

void orc_get_gn(const orc_hier *h, int l, int *gn) {
  const level_t *L = h->lv + l;
  memcpy(gn, L->gn, sizeof(int) * h->nb * L->ncell);
}

void orc_get_csr(const orc_hier *h, int l, int *rowptr, int *kcol, double *entries) {
  const level_t *L = h->lv + l;
  if (rowptr) memcpy(rowptr, L->rowptr, sizeof(int) * (L->ndof + 1));
  if (kcol) memcpy(kcol, L->kcol, sizeof(int) * L->nnz);
  if (entries) memcpy(entries, L->A, sizeof(double) * L->nnz);
}

void orc_get_dof_nodes(const orc_hier *h, int l, int *node) {
  memcpy(node, h->lv[l].dof_node, sizeof(int) * h->lv[l].ndof);
}

void orc_get_rhs(const orc_hier *h, int l, double *rhs) {
  memcpy(rhs, h->lv[l].rhs, sizeof(double) * h->lv[l].ndof);
}

void orc_set_entries(orc_hier *h, int l, const double *entries) {
  memcpy(h->lv[l].A, entries, sizeof(double) * h->lv[l].nnz);
  if (l == 0) h->lu_valid = 0;
}

/* ------------------------------------------------------------------ coefficients */

static void coeff_at(const orc_coeffs *co, double x, double y, double z, double *c7) {
  double b1, b2, b3;
  if (co->conv_kind == 1) {
    b1 = -(y - co->rot_c[1]); b2 = x - co->rot_c[0]; b3 = 0.1 * sin(2 * M_PI * co->t);
  } else {
    b1 = co->b[0]; b2 = co->b[1]; b3 = co->b[2];
  }
  c7[0] = co->eps; c7[1] = b1; c7[2] = b2; c7[3] = b3; c7[4] = co->c;
  if (co->f_kind == 1) {
    double sx = sin(M_PI * x), sy = sin(M_PI * y), sz = sin(M_PI * z);
    double cx = cos(M_PI * x), cy = cos(M_PI * y), cz = cos(M_PI * z), u = sx * sy * sz;
    c7[5] = co->eps * 3 * M_PI * M_PI * u + b1 * M_PI * cx * sy * sz + b2 * M_PI * sx * cy * sz +
            b3 * M_PI * sx * sy * cz + co->c * u;
  } else {
    c7[5] = co->f_val;
  }
  c7[6] = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
}

/* Mesh_size_in_convection_direction_without_storing, hexahedron branch
 * (ConvDiff.C:411-469): 2|b| / sum_i |b . grad phi_i(barycentre)| with the
 * trilinear vertex basis.  On an axis-aligned brick with sides (hx,hy,hz)
 * every vertex gradient at the barycentre is (+-1/hx, +-1/hy, +-1/hz)/4, so
 * the sum runs over the 8 sign patterns of |+-b1/hx +- b2/hy +- b3/hz|/4. */
double orc_hconv_brick(const double *h, const double *b) {
  double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]), den = 0;
  for (int s = 0; s < 8; s++) {
    double v = ((s & 1) ? -b[0] : b[0]) / h[0] + ((s & 2) ? -b[1] : b[1]) / h[1] + ((s & 4) ? -b[2] : b[2]) / h[2];
    den += fabs(v) / 4;
  }
  if (den < 1e-10) return sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]); /* hK = cell diameter */
  return 2 * nb / den;
}

/* Mesh_size_in_convection_direction_without_storing, hexahedron branch
 * (ConvDiff.C:411-469) for any hexahedron: the trilinear function of
 * (x, y, z) -- coefficients of 1, x, y, z, xy, xz, yz, xyz -- that is 1 at
 * vertex i and 0 at the others, by LU with partial pivoting of the 8 x 8
 * vertex matrix for 8 right-hand sides at once (SolveMultipleSystems,
 * LinAlg.C:466-578), and its gradient at the barycentre. */
static double hconv_hexa(const double V[8][3], const double *bv) {
  double a[64], r[64], s[3] = {0, 0, 0};
  for (int i = 0; i < 8; i++)
    for (int d = 0; d < 3; d++) s[d] += V[i][d];
  for (int d = 0; d < 3; d++) s[d] /= 8;
  for (int i = 0; i < 8; i++) {
    double x = V[i][0], y = V[i][1], z = V[i][2];
    double row[8] = {1, x, y, z, x * y, x * z, y * z, x * y * z};
    memcpy(a + 8 * i, row, sizeof row);
  }
  memset(r, 0, sizeof r);
  for (int i = 0; i < 8; i++) r[9 * i] = 1; /* right-hand side j is r[8 j ..] */
  for (int k = 0; k < 7; k++) {
    double best = 0;
    int pr = k;
    for (int i = k; i < 8; i++)
      if (fabs(a[8 * i + k]) > best) { best = fabs(a[8 * i + k]); pr = i; }
    if (pr > k) {
      for (int j = 0; j < 8; j++) { double t = a[8 * k + j]; a[8 * k + j] = a[8 * pr + j]; a[8 * pr + j] = t; }
      for (int j = 0; j < 8; j++) { double t = r[8 * j + k]; r[8 * j + k] = r[8 * j + pr]; r[8 * j + pr] = t; }
    }
    double piv = a[8 * k + k];
    for (int i = k + 1; i < 8; i++) a[8 * i + k] /= piv;
    for (int j = k + 1; j < 8; j++) {
      double t = a[8 * k + j];
      for (int i = k + 1; i < 8; i++) a[8 * i + j] -= a[8 * i + k] * t;
    }
  }
  for (int k = 0; k < 8; k++)
    for (int j = 0; j < 8; j++) {
      double t = r[8 * j + k];
      for (int i = k + 1; i < 8; i++) r[8 * j + i] -= a[8 * i + k] * t;
    }
  for (int k = 7; k >= 0; k--)
    for (int j = 0; j < 8; j++) {
      r[8 * j + k] /= a[8 * k + k];
      double t = r[8 * j + k];
      for (int i = 0; i < k; i++) r[8 * j + i] -= a[8 * i + k] * t;
    }
  double nb = sqrt(bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2]), den = 0;
  for (int i = 0; i < 8; i++) {
    const double *q = r + 8 * i;
    double v = bv[0] * (q[1] + q[4] * s[1] + q[5] * s[2] + q[7] * s[1] * s[2]);
    v += bv[1] * (q[2] + q[4] * s[0] + q[6] * s[2] + q[7] * s[0] * s[2]);
    v += bv[2] * (q[3] + q[5] * s[0] + q[6] * s[1] + q[7] * s[0] * s[1]);
    den += fabs(v);
  }
  if (den < 1e-10) { /* hK: the longest space diagonal (THexahedron::GetDiameter, Hexahedron.C:78-104) */
    static const int DG[4][2] = {{0, 6}, {1, 7}, {2, 4}, {3, 5}};
    double dm = 0;
    for (int k = 0; k < 4; k++) {
      double l2 = 0;
      for (int d = 0; d < 3; d++) l2 += (V[DG[k][0]][d] - V[DG[k][1]][d]) * (V[DG[k][0]][d] - V[DG[k][1]][d]);
      if (l2 > dm) dm = l2;
    }
    return sqrt(dm);
  }
  return 2 * nb / den;
}

/* ------------------------------------------------------------------ cell geometry */

/* The map of a cell from the reference cube.  The reference maps every Q1/Q2
 * hexahedron with THexaTrilinear (the elements' RefTrans, FEDatabase3D.C:632-634).
 * On the box hierarchy the cells are bricks whose frames are signed axis
 * permutations, and the trilinear map reduces to the scaled frame; on a
 * mapped hierarchy the map is the trilinear interpolation of the cell's 8
 * vertices (THexaTrilinear::SetCell, HexaTrilinear.C:884-948): x(r) =
 * sum_k xc_k m_k(r) over m = 1, xi, eta, zeta, xi eta, xi zeta, eta zeta,
 * xi eta zeta, xc_k = 1/8 sum_v m_k(r_v) x_v. */
typedef struct {
  int mapped;
  double V[8][3];  /* vertices in the cell's local order */
  double xc[3][8];
  const int *O, *E;
  double hd[3], dn[3], detj;
} cellgeo_t;

typedef struct {
  double X[3];     /* the point in physical coordinates */
  double C[3][3];  /* mapped: cofactors, physical gradient = C grad_ref / det */
  double rdet;     /* mapped: 1 / det */
  double absdet;   /* |det J| */
} geopt_t;

/* vertices of cell c in its local order (mapped hierarchies) */
static void cell_vertices(const orc_hier *h, const level_t *L, int c, double V[8][3]) {
  const int *O = L->org + 3 * c, *E = L->ax + 9 * c;
  int nv = L->n + 1, p = h->order;
  for (int v = 0; v < 8; v++) {
    int I[3];
    for (int d = 0; d < 3; d++) I[d] = O[d] / p + HEX_V[v][0] * E[d] + HEX_V[v][1] * E[3 + d] + HEX_V[v][2] * E[6 + d];
    const double *x = L->vx + 3 * (I[0] + nv * (I[1] + nv * I[2]));
    for (int d = 0; d < 3; d++) V[v][d] = x[d];
  }
}

static void geo_cell(const orc_hier *h, const level_t *L, int c, cellgeo_t *G) {
  G->mapped = L->vx != NULL;
  G->O = L->org + 3 * c;
  G->E = L->ax + 9 * c;
  for (int d = 0; d < 3; d++) { G->hd[d] = h->box[d] / L->n; G->dn[d] = h->box[d] / (h->order * L->n); }
  G->detj = (G->hd[0] / 2) * (G->hd[1] / 2) * (G->hd[2] / 2);
  if (!G->mapped) return;
  cell_vertices(h, L, c, G->V);
  /* vertices in tensor order (x fastest): 0 1 3 2 4 5 7 6 of the local order */
  static const int TV[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 8; k++) {
      double acc = 0;
      for (int t = 0; t < 8; t++) {
        double sx = (t & 1) ? 1 : -1, sy = (t & 2) ? 1 : -1, sz = (t & 4) ? 1 : -1;
        double mk[8] = {1, sx, sy, sz, sx * sy, sx * sz, sy * sz, sx * sy * sz};
        acc += mk[k] * G->V[TV[t]][d];
      }
      G->xc[d][k] = acc * 0.125;
    }
}

/* THexaTrilinear::GetOrigFromRef / GetOrigValues (HexaTrilinear.C): the point,
 * the Jacobian's cofactors and |det J| at the reference point (xi, eta, zeta) */
static void geo_point(const cellgeo_t *G, double xi, double eta, double zeta, geopt_t *P) {
  if (!G->mapped) {
    double ref[3] = {xi, eta, zeta};
    const int *O = G->O, *E = G->E;
    for (int d = 0; d < 3; d++)
      P->X[d] = O[d] * G->dn[d] + (ref[0] + 1) / 2 * G->hd[d] * E[d] + (ref[1] + 1) / 2 * G->hd[d] * E[3 + d] +
                (ref[2] + 1) / 2 * G->hd[d] * E[6 + d];
    P->absdet = G->detj;
    return;
  }
  double J[3][3];
  for (int d = 0; d < 3; d++) {
    const double *c = G->xc[d];
    P->X[d] = c[0] + c[1] * xi + c[2] * eta + c[3] * zeta + c[4] * xi * eta + c[5] * xi * zeta + c[6] * eta * zeta +
              c[7] * xi * eta * zeta;
    J[d][0] = c[1] + c[4] * eta + c[5] * zeta + c[7] * eta * zeta;
    J[d][1] = c[2] + c[4] * xi + c[6] * zeta + c[7] * xi * zeta;
    J[d][2] = c[3] + c[5] * xi + c[6] * eta + c[7] * xi * eta;
  }
  /* rows x, y, z of J: (dx1 dx2 dx3), (dy1 dy2 dy3), (dz1 dz2 dz3) */
  double dx1 = J[0][0], dx2 = J[0][1], dx3 = J[0][2], dy1 = J[1][0], dy2 = J[1][1], dy3 = J[1][2];
  double dz1 = J[2][0], dz2 = J[2][1], dz3 = J[2][2];
  double det = dx1 * dy2 * dz3 + dx2 * dy3 * dz1 + dx3 * dy1 * dz2 - dx3 * dy2 * dz1 - dx2 * dy1 * dz3 - dx1 * dy3 * dz2;
  P->C[0][0] = dy2 * dz3 - dy3 * dz2; P->C[0][1] = dy3 * dz1 - dy1 * dz3; P->C[0][2] = dy1 * dz2 - dy2 * dz1;
  P->C[1][0] = dx3 * dz2 - dx2 * dz3; P->C[1][1] = dx1 * dz3 - dx3 * dz1; P->C[1][2] = dx2 * dz1 - dx1 * dz2;
  P->C[2][0] = dx2 * dy3 - dy2 * dx3; P->C[2][1] = dx3 * dy1 - dx1 * dy3; P->C[2][2] = dx1 * dy2 - dx2 * dy1;
  P->rdet = 1 / det;
  P->absdet = fabs(det);
}

/* physical gradient of a basis function from its reference derivatives */
static void geo_grad(const cellgeo_t *G, const geopt_t *P, double d0, double d1, double d2, double g[3]) {
  if (!G->mapped) {
    const int *E = G->E;
    for (int d = 0; d < 3; d++) g[d] = (2 / G->hd[d]) * (E[d] * d0 + E[3 + d] * d1 + E[6 + d] * d2);
    return;
  }
  for (int d = 0; d < 3; d++) g[d] = (P->C[d][0] * d0 + P->C[d][1] * d1 + P->C[d][2] * d2) * P->rdet;
}

static double geo_hconv(const cellgeo_t *G, const double *b) {
  return G->mapped ? hconv_hexa(G->V, b) : orc_hconv_brick(G->hd, b);
}

/* Compute_SDFEM_delta, SDFEM_TYPE 2 (ConvDiff.C:494-600) with DELTA0 = 1 and
 * INTERNAL_POLYNOMIAL_DEGREE = order (SetParametersCD, ConvDiff.C:160-177). */
double orc_sdfem_delta(double hconv, double eps, const double *b, int order) {
  if (fabs(eps) < 1e-20) eps = 1e-20;
  double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  double nu = 1.0 / order;
  if (!(nb > 0)) return 0;
  double alpha = nu * nb * hconv / (2 * eps);
  return nu * hconv * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
}

/* ------------------------------------------------------------------ assembly */

/* Local matrices of one cell (TFEDatabase3D::GetOrig + THexaAffin,
 * TDiscreteForm3D::GetLocalForms DiscreteForm3D.C:187-270, kernels
 * BilinearAssemble / BilinearAssemble_SD ConvDiff3D.C:61-190,
 * MatrixMRhsAssemble TCD3D.C:11-46, MatrixMRhsAssemble_SUPG TCD3D.C:48-100,
 * MatricesAKRhsAssemble_SUPG TCD3D.C:102-240).
 * what: 0 conv-diff A (SUPG test functions when co->supg), 1 mass M (rhs with
 * the SUPG test functions when co->supg), 2 the SUPG time-derivative matrix
 * K_ij = sum w|J| phi_j tau b.grad phi_i (no rhs).
 * tscale = theta1 * tau > 0: the time-dependent SUPG parameter of
 * MatricesAKRhsAssemble_SUPG -- Compute_SDFEM_delta of the scaled
 * coefficients (tscale eps, tscale b), multiplied by tscale -- for what 0 and
 * 2; tscale 0: the steady parameter of BilinearAssemble_SD. */
static void local_forms(const orc_hier *h, const level_t *L, int c, const orc_coeffs *co, int what,
                        double tscale, double *Aloc, double *bloc) {
  int p = h->order, nb = h->nb, n1 = p + 2; /* Gauss3 for Q1, Gauss4 for Q2 */
  const double *gx = p == 1 ? G3X : G4X, *gw = p == 1 ? G3W : G4W;
  cellgeo_t G;
  geo_cell(h, L, c, &G);
  memset(Aloc, 0, sizeof(double) * nb * nb);
  memset(bloc, 0, sizeof(double) * nb);
  double hconv = -1;
  for (int qi = 0; qi < n1; qi++)
    for (int qj = 0; qj < n1; qj++)
      for (int qk = 0; qk < n1; qk++) {
        double xi = gx[qk], eta = gx[qj], zeta = gx[qi];
        double w = gw[qi] * gw[qj] * gw[qk];
        geopt_t gp;
        geo_point(&G, xi, eta, zeta, &gp);
        double Mult = w * gp.absdet;
        const double *X = gp.X;
        double phi[MAXNB], d0[MAXNB], d1[MAXNB], d2[MAXNB], gx_[MAXNB], gy_[MAXNB], gz_[MAXNB];
        basis3d(p, xi, eta, zeta, phi, d0, d1, d2);
        for (int b = 0; b < nb; b++) {
          double g[3];
          geo_grad(&G, &gp, d0[b], d1[b], d2[b], g);
          gx_[b] = g[0]; gy_[b] = g[1]; gz_[b] = g[2];
        }
        double cf[7];
        if (L->qp) { /* the CoeffFct3D values handed in for this point */
          const double *v = L->qp + ((size_t)c * n1 * n1 * n1 + (qi * n1 + qj) * n1 + qk) * 6;
          for (int k = 0; k < 6; k++) cf[k] = v[k];
          cf[6] = sqrt(cf[1] * cf[1] + cf[2] * cf[2] + cf[3] * cf[3]);
        } else {
          coeff_at(co, X[0], X[1], X[2], cf);
        }
        double delta = 0;
        if (co->supg && what != 1 && tscale > 0) {
          /* TCD3D.C:141-156: tilde tau of (tscale eps, tscale b), then tau *= tscale */
          double sb[3] = {tscale * cf[1], tscale * cf[2], tscale * cf[3]};
          if (hconv < 0) hconv = geo_hconv(&G, sb); /* cached per cell: AssembleMat3D.C:457 */
          delta = orc_sdfem_delta(hconv, tscale * cf[0], sb, p) * tscale;
        } else if (co->supg) {
          if (hconv < 0) hconv = geo_hconv(&G, cf + 1); /* cached per cell: AssembleMat3D.C:457 */
          delta = orc_sdfem_delta(hconv, cf[0], cf + 1, p);
        }
        if (what == 1) {
          for (int i = 0; i < nb; i++) {
            double bgradv = cf[1] * gx_[i] + cf[2] * gy_[i] + cf[3] * gz_[i];
            bloc[i] += Mult * (co->supg ? phi[i] + delta * bgradv : phi[i]) * cf[5];
            for (int j = 0; j < nb; j++) Aloc[i * nb + j] += Mult * phi[j] * phi[i];
          }
          continue;
        }
        if (what == 2) {
          for (int i = 0; i < nb; i++) {
            double bgradv = delta * (cf[1] * gx_[i] + cf[2] * gy_[i] + cf[3] * gz_[i]);
            for (int j = 0; j < nb; j++) Aloc[i * nb + j] += Mult * phi[j] * bgradv;
          }
          continue;
        }
        for (int i = 0; i < nb; i++) {
          double t100 = gx_[i], t010 = gy_[i], t001 = gz_[i], t000 = phi[i];
          double bgradv = cf[1] * t100 + cf[2] * t010 + cf[3] * t001;
          if (co->supg)
            bloc[i] += Mult * (t000 + delta * bgradv) * cf[5];
          else
            bloc[i] += Mult * t000 * cf[5];
          for (int j = 0; j < nb; j++) {
            double a100 = gx_[j], a010 = gy_[j], a001 = gz_[j], a000 = phi[j];
            double val = cf[0] * (t100 * a100 + t010 * a010 + t001 * a001);
            val += (cf[1] * a100 + cf[2] * a010 + cf[3] * a001) * t000;
            val += cf[4] * a000 * t000;
            if (co->supg) val += delta * (cf[1] * a100 + cf[2] * a010 + cf[3] * a001 + cf[4] * a000) * bgradv;
            Aloc[i * nb + j] += Mult * val;
          }
        }
      }
}

/* The quadrature points of level l per cell (cell order) and point (Gauss
 * rule order): the points TFEDatabase3D::GetOrig hands to the CoeffFct3D
 * callback (FEDatabase3D.C:1116-1480, DiscreteForm3D.C:187-270). */
void orc_quad_points(const orc_hier *h, int l, double *xyz) {
  const level_t *L = h->lv + l;
  int p = h->order, n1 = p + 2;
  const double *gx = p == 1 ? G3X : G4X;
  size_t k = 0;
  for (int c = 0; c < L->ncell; c++) {
    cellgeo_t G;
    geo_cell(h, L, c, &G);
    for (int qi = 0; qi < n1; qi++)
      for (int qj = 0; qj < n1; qj++)
        for (int qk = 0; qk < n1; qk++) {
          geopt_t gp;
          geo_point(&G, gx[qk], gx[qj], gx[qi], &gp);
          for (int d = 0; d < 3; d++) xyz[k++] = gp.X[d];
        }
  }
}

/* coefficients per quadrature point for level l (NULL: back to the descriptor) */
void orc_set_coeffs_qp(orc_hier *h, int l, const double *qp) {
  level_t *L = h->lv + l;
  free(L->qp);
  L->qp = NULL;
  if (!qp) return;
  int n1 = h->order + 2;
  size_t n = (size_t)L->ncell * n1 * n1 * n1 * 6;
  L->qp = malloc(sizeof(double) * n);
  memcpy(L->qp, qp, sizeof(double) * n);
}

/* TAssembleMat3D::Assemble3D cell loop (AssembleMat3D.C:386-491) with
 * AddLocalSqMatToGlobal (:721-827: active rows accumulate, Dirichlet rows get
 * 1 on the diagonal) and AddLocalRhsToGlobal (:883-1066: active rows
 * accumulate, Dirichlet rows get g = 0). */
static void assemble_level_t(const orc_hier *h, int l, const orc_coeffs *co, int what, double tscale,
                             double *entries, double *rhs) {
  const level_t *L = h->lv + l;
  int nb = h->nb;
  double Aloc[MAXNB * MAXNB], bloc[MAXNB];
  memset(entries, 0, sizeof(double) * L->nnz);
  memset(rhs, 0, sizeof(double) * L->ndof);
  for (int c = 0; c < L->ncell; c++) {
    local_forms(h, L, c, co, what, tscale, Aloc, bloc);
    const int *g = L->gn + c * nb;
    for (int m = 0; m < nb; m++) {
      int r = g[m];
      if (r < L->nact) {
        for (int k = 0; k < nb; k++) {
          int n = L->rowptr[r];
          while (L->kcol[n] != g[k]) n++;
          entries[n] += Aloc[m * nb + k];
        }
        rhs[r] += bloc[m];
      } else {
        entries[L->rowptr[r]] = 1.0;
        rhs[r] = 0.0;
      }
    }
  }
}

void orc_assemble_level(const orc_hier *h, int l, const orc_coeffs *co, int what, double *entries, double *rhs) {
  assemble_level_t(h, l, co, what, 0.0, entries, rhs);
}

/* TSystemCD3D::Assemble (SystemCD3D.C:314-339): every level, then the
 * Dirichlet part of the rhs is copied into the level solution. */
void orc_assemble(orc_hier *h, const orc_coeffs *co) {
  for (int l = 0; l < h->L; l++) {
    level_t *L = h->lv + l;
    orc_assemble_level(h, l, co, 0, L->A, L->rhs);
    memcpy(L->sol + L->nact, L->rhs + L->nact, sizeof(double) * (L->ndof - L->nact));
  }
  h->lu_valid = 0;
}

/* ------------------------------------------------------------------ BLAS */

/* MatVect (Blas2.C:60-87) */
static void matvec_(const level_t *L, const double *A, const double *x, double *y) {
  for (int i = 0; i < L->ndof; i++) {
    double s = 0;
    for (int j = L->rowptr[i]; j < L->rowptr[i + 1]; j++) s += A[j] * x[L->kcol[j]];
    y[i] = s;
  }
}
void orc_matvec(const orc_hier *h, int l, const double *x, double *y) { matvec_(h->lv + l, h->lv[l].A, x, y); }

/* ScalarDefect (Blas2.C:11-56) */
static double defect_(const level_t *L, const double *x, const double *f, double *d) {
  double res = 0;
  for (int i = 0; i < L->ndof; i++) {
    double s = f[i];
    for (int j = L->rowptr[i]; j < L->rowptr[i + 1]; j++) s -= L->A[j] * x[L->kcol[j]];
    d[i] = s;
    res += s * s;
  }
  return sqrt(res);
}
double orc_defect(const orc_hier *h, int l, const double *x, const double *f, double *d) {
  return defect_(h->lv + l, x, f, d);
}

static double ddot(int n, const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* ------------------------------------------------------------------ smoothers */

/* TMGLevel3D::Jacobi (MGLevel3D.C:443-537) */
static void jacobi(level_t *L, double omega) {
  double *sol = L->sol, *f = L->rhs, *aux = L->aux1;
  memcpy(aux, sol, sizeof(double) * L->ndof);
  memcpy(sol + L->nact, f + L->nact, sizeof(double) * (L->ndof - L->nact));
  for (int i = 0; i < L->nact; i++) {
    double s = f[i], diag = 1;
    for (int j = L->rowptr[i]; j < L->rowptr[i + 1]; j++) {
      int c = L->kcol[j];
      if (c == i) diag = L->A[j];
      else s -= L->A[j] * aux[c];
    }
    double t = aux[i];
    sol[i] = (1 - omega) * t + omega * s / diag;
  }
}

/* TMGLevel3D::SOR (MGLevel3D.C:254-335): Dirichlet values, then one
 * relaxed Gauss-Seidel pass over the active rows in ascending order */
static void sor(level_t *L, double omega) {
  double *sol = L->sol, *f = L->rhs;
  memcpy(sol + L->nact, f + L->nact, sizeof(double) * (L->ndof - L->nact));
  for (int i = 0; i < L->nact; i++) {
    double s = f[i], diag = 1;
    for (int j = L->rowptr[i]; j < L->rowptr[i + 1]; j++) {
      int c = L->kcol[j];
      if (c == i) diag = L->A[j];
      else s -= L->A[j] * sol[c];
    }
    double t = sol[i];
    sol[i] = omega * (s / diag - t) + t;
  }
}

/* TMGLevel3D::SSOR (MGLevel3D.C:337-441) */
static void ssor(level_t *L, double omega) {
  double *sol = L->sol, *f = L->rhs;
  memcpy(sol + L->nact, f + L->nact, sizeof(double) * (L->ndof - L->nact));
  for (int i = 0; i < L->nact; i++) {
    double s = f[i], diag = 1;
    for (int j = L->rowptr[i]; j < L->rowptr[i + 1]; j++) {
      int c = L->kcol[j];
      if (c == i) diag = L->A[j];
      else s -= L->A[j] * sol[c];
    }
    double t = sol[i];
    sol[i] = omega * (s / diag - t) + t;
  }
  for (int i = L->nact - 1; i >= 0; i--) {
    double s = f[i], diag = 1;
    for (int j = L->rowptr[i + 1] - 1; j >= L->rowptr[i]; j--) {
      int c = L->kcol[j];
      if (c == i) diag = L->A[j];
      else s -= L->A[j] * sol[c];
    }
    double t = sol[i];
    sol[i] = omega * (s / diag - t) + t;
  }
}

/* exact coarse solve (TMGLevel3D::SolveExact, MGLevel3D.C:839-880): dense
 * LU with partial pivoting of the whole coarse matrix.  The reference calls
 * LAPACK dgetrf/dgetrs (LinAlg.C:106, third party, not vendored); this is the
 * textbook row-pivoted Doolittle LU (Golub & Van Loan, Alg. 3.4.1). */
static void coarse_factor(orc_hier *h) {
  level_t *L = h->lv;
  int n = L->ndof;
  free(h->lu); free(h->piv);
  h->lu = calloc((size_t)n * n, sizeof(double));
  h->piv = malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++)
    for (int j = L->rowptr[i]; j < L->rowptr[i + 1]; j++) h->lu[(size_t)i * n + L->kcol[j]] = L->A[j];
  double *a = h->lu;
  for (int k = 0; k < n; k++) {
    int pr = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(a[(size_t)i * n + k]) > fabs(a[(size_t)pr * n + k])) pr = i;
    h->piv[k] = pr;
    if (pr != k)
      for (int j = 0; j < n; j++) {
        double t = a[(size_t)k * n + j]; a[(size_t)k * n + j] = a[(size_t)pr * n + j]; a[(size_t)pr * n + j] = t;
      }
    for (int i = k + 1; i < n; i++) {
      double lik = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      a[(size_t)i * n + k] = lik;
      if (lik != 0)
        for (int j = k + 1; j < n; j++) a[(size_t)i * n + j] -= lik * a[(size_t)k * n + j];
    }
  }
  h->lu_valid = 1;
}

static void coarse_solve(orc_hier *h) {
  if (!h->lu_valid) coarse_factor(h);
  level_t *L = h->lv;
  int n = L->ndof;
  double *x = L->sol, *a = h->lu;
  memcpy(x, L->rhs, sizeof(double) * n);
  for (int k = 0; k < n; k++) {
    int pr = h->piv[k];
    if (pr != k) { double t = x[k]; x[k] = x[pr]; x[pr] = t; }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) x[i] -= a[(size_t)i * n + j] * x[j];
  for (int i = n - 1; i >= 0; i--) {
    for (int j = i + 1; j < n; j++) x[i] -= a[(size_t)i * n + j] * x[j];
    x[i] /= a[(size_t)i * n + i];
  }
}

/* ------------------------------------------------------------------ transfers */

/* values of the parent basis at the nodes of child cell c (the rows of
 * TFEDatabase3D::GetProlongationMatrix3D, FEDatabase3D.C:1630-1790): the
 * child's local DOF k sits at a lattice node; its parent-reference
 * coordinates follow from the parent's origin and axes. */
static void child_qq(const orc_hier *h, int l, int c, double *QQ /* nb x nb: [k][lp] */) {
  const level_t *F = h->lv + l, *C = h->lv + l - 1;
  int p = h->order, nb = h->nb, P = c >> 3;
  const int *OP = C->org + 3 * P, *EP = C->ax + 9 * P;
  for (int k = 0; k < nb; k++) {
    int i = k % (p + 1), j = (k / (p + 1)) % (p + 1), kk = k / ((p + 1) * (p + 1));
    int node = cell_node(F, c, p, i, j, kk);
    int X[3] = {node % F->nn, (node / F->nn) % F->nn, node / (F->nn * F->nn)};
    double ref[3];
    for (int a = 0; a < 3; a++) {
      int t = 0;
      for (int d = 0; d < 3; d++) t += EP[3 * a + d] * (X[d] - 2 * OP[d]);
      ref[a] = 2.0 * t / (2 * p) - 1.0;
    }
    basis3d(p, ref[0], ref[1], ref[2], QQ + k * nb, NULL, NULL, NULL);
  }
}

/* Prolongate (MGComponents3D.C:182-463): per fine cell, QQ * parent values,
 * accumulated and divided by the DOF multiplicity. */
void orc_prolongate(const orc_hier *h, int l, const double *coarse, double *fine) {
  const level_t *F = h->lv + l, *C = h->lv + l - 1;
  int nb = h->nb;
  double QQ[MAXNB * MAXNB];
  memset(fine, 0, sizeof(double) * F->ndof);
  for (int c = 0; c < F->ncell; c++) {
    child_qq(h, l, c, QQ);
    const int *gP = C->gn + (c >> 3) * nb, *gc = F->gn + c * nb;
    double Val[MAXNB];
    for (int a = 0; a < nb; a++) Val[a] = coarse[gP[a]];
    for (int k = 0; k < nb; k++) {
      double s = 0;
      for (int a = 0; a < nb; a++) s += QQ[k * nb + a] * Val[a];
      fine[gc[k]] += s;
    }
  }
  for (int i = 0; i < F->ndof; i++)
    if (F->mult[i]) fine[i] /= F->mult[i];
}

/* DefectRestriction (MGComponents3D.C:706-989): fine defect divided by the
 * multiplicity, then QQ^T per fine cell, accumulated into the parent DOFs. */
void orc_restrict(const orc_hier *h, int l, const double *fine, double *coarse) {
  const level_t *F = h->lv + l, *C = h->lv + l - 1;
  int nb = h->nb;
  double QQ[MAXNB * MAXNB];
  double *fs = malloc(sizeof(double) * F->ndof);
  for (int i = 0; i < F->ndof; i++) fs[i] = fine[i] / F->mult[i];
  memset(coarse, 0, sizeof(double) * C->ndof);
  for (int c = 0; c < F->ncell; c++) {
    child_qq(h, l, c, QQ);
    const int *gP = C->gn + (c >> 3) * nb, *gc = F->gn + c * nb;
    double Val[MAXNB];
    for (int a = 0; a < nb; a++) Val[a] = fs[gc[a]];
    for (int k = 0; k < nb; k++) {
      double s = 0;
      for (int a = 0; a < nb; a++) s += QQ[a * nb + k] * Val[a];
      coarse[gP[k]] += s;
    }
  }
  free(fs);
}

/* ------------------------------------------------------------------ multigrid */

/* Multicolour SSOR (smoother 13, no reference counterpart: a parallel
 * smoother of this build for convection-dominated problems, where damped
 * Jacobi stalls).  Interior nodes are coloured by their lattice coordinates
 * modulo 2 (Q1) or 3 (Q2): c = X%k + k (Y%k + k Z%k), so two nodes of one
 * colour never share a cell and their rows never couple.  One sweep: x_D = f_D,
 * then relaxed Gauss-Seidel colour by colour, c = 0 .. k^3 - 1 (every row of a
 * colour updated from the current x), then back, c = k^3 - 1 .. 0.
 * Pinned by tests/test_oracle_math.py against a numpy statement of the same
 * definition. */
static int node_colour(const level_t *L, int dof, int p) {
  int k = p == 1 ? 2 : 3, node = L->dof_node[dof], nn = L->nn;
  int X = node % nn, Y = (node / nn) % nn, Z = node / (nn * nn);
  return X % k + k * (Y % k + k * (Z % k));
}
static void colour_pass(level_t *L, double omega, int p, int colour) {
  double *sol = L->sol, *f = L->rhs;
  for (int i = 0; i < L->nact; i++) {
    if (node_colour(L, i, p) != colour) continue;
    double s = f[i], diag = 1;
    for (int j = L->rowptr[i]; j < L->rowptr[i + 1]; j++) {
      int c = L->kcol[j];
      if (c == i) diag = L->A[j];
      else s -= L->A[j] * sol[c];
    }
    sol[i] = (1 - omega) * sol[i] + omega * s / diag;
  }
}
static void mcssor(level_t *L, double omega, int p) {
  int k = p == 1 ? 2 : 3, ncol = k * k * k;
  memcpy(L->sol + L->nact, L->rhs + L->nact, sizeof(double) * (L->ndof - L->nact));
  for (int c = 0; c < ncol; c++) colour_pass(L, omega, p, c);
  for (int c = ncol - 1; c >= 0; c--) colour_pass(L, omega, p, c);
}
/* one SOR (kind 2) or SSOR (kind 3) sweep of level l on x with rhs f */
void orc_gs_sweep(orc_hier *h, int l, int kind, double omega, const double *f, double *x) {
  level_t *L = h->lv + l;
  double *s0 = L->sol, *r0 = L->rhs;
  L->sol = x;
  L->rhs = (double *)f;
  if (kind == 2) sor(L, omega);
  else ssor(L, omega);
  L->sol = s0;
  L->rhs = r0;
}

void orc_mcssor_sweep(orc_hier *h, int l, double omega, const double *f, double *x) {
  level_t *L = h->lv + l;
  double *s0 = L->sol, *r0 = L->rhs;
  L->sol = x;
  L->rhs = (double *)f;
  mcssor(L, omega, h->order);
  L->sol = s0;
  L->rhs = r0;
}

static void smooth(orc_hier *h, level_t *L, const orc_solver *s, int n) {
  for (int k = 0; k < n; k++) {
    if (s->smoother == 3) ssor(L, s->omega);
    else if (s->smoother == 2) sor(L, s->omega);
    else if (s->smoother == 13) mcssor(L, s->omega, h->order);
    else jacobi(L, s->omega);
  }
}

/* TMultiGrid3D::Cycle (MultiGrid3D.C:423-669), no step-length control; the
 * Defect calls whose result only feeds verbose output are omitted (they
 * change no state that is read afterwards).  rec[i] is mg_recursions[i]
 * (SetRecursion, MultiGrid3D.C:690-703): 1 on the level above the coarsest,
 * SC_MG_CYCLE_SCALAR (V 1, W 2) elsewhere; the F-cycle starts with 2 and
 * drops to 1 after a level's first coarse-grid correction. */
static void cycle(orc_hier *h, int i, const orc_solver *s, int *rec) {
  level_t *L = h->lv + i;
  if (i == 0) {
    /* TMultiGrid3D::Smooth, CoarseSmooth (MultiGrid3D.C:197-260) */
    double res = defect_(L, L->sol, L->rhs, L->aux0);
    double reduction = s->coarse_red * res;
    int it = 0;
    while (res > reduction && it < s->coarse_maxit) {
      if (s->coarse_smoother == 17) coarse_solve(h);
      else if (s->coarse_smoother == 1) jacobi(L, s->omega);
      else if (s->coarse_smoother == 2) sor(L, s->omega);
      else ssor(L, s->omega);
      res = defect_(L, L->sol, L->rhs, L->aux0);
      it++;
    }
    return;
  }
  level_t *Cl = h->lv + i - 1;
  smooth(h, L, s, s->pre);
  defect_(L, L->sol, L->rhs, L->aux0);
  orc_restrict(h, i, L->aux0, Cl->rhs);
  memset(Cl->rhs + Cl->nact, 0, sizeof(double) * (Cl->ndof - Cl->nact)); /* CorrectDefect */
  memset(Cl->sol, 0, sizeof(double) * Cl->ndof);                         /* Reset */
  for (int r = 0; r < rec[i]; r++) cycle(h, i - 1, s, rec);
  if (s->cycle < 1) rec[i] = 1;
  orc_prolongate(h, i, Cl->sol, L->aux0);
  memset(L->aux0 + L->nact, 0, sizeof(double) * (L->ndof - L->nact)); /* CorrectNodes */
  for (int k = 0; k < L->ndof; k++) L->sol[k] += s->gmg_damp * L->aux0[k]; /* Update */
  smooth(h, L, s, s->post);
}

/* TMultiGridScaIte::Iterate (MultiGridScaIte.C:63-104), zero start */
void orc_vcycle(orc_hier *h, const orc_solver *s, const double *r, double *z) {
  level_t *L = h->lv + h->L - 1;
  memset(L->sol, 0, sizeof(double) * L->ndof);
  memcpy(L->rhs, r, sizeof(double) * L->ndof);
  memcpy(L->sol + L->nact, L->rhs + L->nact, sizeof(double) * (L->ndof - L->nact)); /* SetDirichletNodes */
  int rec[32];
  rec[1] = 1;
  for (int k = 2; k < h->L; k++) rec[k] = s->cycle > 0 ? s->cycle : 2;
  cycle(h, h->L - 1, s, rec);
  memcpy(z, L->sol, sizeof(double) * L->ndof);
}

/* ------------------------------------------------------------------ FGMRES */

/* GeneratePlaneRotation / ApplyPlaneRotation (FgmresIte.C:168-195) */
static void gen_rot(double dx, double dy, double *cs, double *sn) {
  if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
  else if (fabs(dy) > fabs(dx)) { double t = dx / dy; *sn = 1.0 / sqrt(1.0 + t * t); *cs = t * *sn; }
  else { double t = dy / dx; *cs = 1.0 / sqrt(1.0 + t * t); *sn = t * *cs; }
}
static void app_rot(double *dx, double *dy, double cs, double sn) {
  double t = cs * *dx + sn * *dy;
  *dy = -sn * *dx + cs * *dy;
  *dx = t;
}
/* UpdateGmresIterate (FgmresIte.C:198-216) */
static void update_iter(double *x, int n, int k, double *H, int ldh, double *s, double **zv) {
  for (int i = k; i >= 0; i--) {
    s[i] /= H[i * ldh + i];
    for (int j = i - 1; j >= 0; j--) s[j] -= H[j * ldh + i] * s[i];
  }
  for (int j = 0; j <= k; j++)
    for (int i = 0; i < n; i++) x[i] += zv[j][i] * s[j];
}

/* TFgmresIte::Iterate, flexible branch (FgmresIte.C:219-595): right
 * preconditioned by one V-cycle, modified Gram-Schmidt, Givens rotations;
 * stops when |s_{i+1}| < res_norm_min or < red_factor * (residual at the
 * start of the current restart cycle). */
int orc_fgmres(orc_hier *h, const orc_solver *sv, const double *rhs, double *sol, double *hist) {
  level_t *L = h->lv + h->L - 1;
  int N = L->ndof, R = sv->restart;
  if (R > sv->maxit) R = sv->maxit;
  double **v = malloc(sizeof(double *) * (R + 1)), **zv = malloc(sizeof(double *) * (R + 1));
  for (int i = 0; i <= R; i++) { v[i] = calloc(N, sizeof(double)); zv[i] = calloc(N, sizeof(double)); }
  double *H = calloc((size_t)(R + 1) * R, sizeof(double)), *s = calloc(R + 1, sizeof(double));
  double *cs = calloc(R + 1, sizeof(double)), *sn = calloc(R + 1, sizeof(double));
  double *w = malloc(sizeof(double) * N);
  double res, res0, beta;
  int i = 0, j, nh = 0, ret = -1;
  res = defect_(L, sol, rhs, v[0]);
  res = sqrt(ddot(N, v[0], v[0]));
  res0 = beta = res;
  double res_norm_min = sv->res_norm_min, red_factor = sv->red_factor;
  if (sv->rel_tol > 0) { res_norm_min = sv->rel_tol * beta; red_factor = 0.0; }
  if (beta <= res_norm_min) { ret = 0; goto done; }
  j = 1;
  while (j <= sv->maxit) {
    for (int k = 0; k < N; k++) v[0][k] *= 1.0 / beta;
    memset(s, 0, sizeof(double) * (R + 1));
    s[0] = beta;
    for (i = 0; i < R && j <= sv->maxit; i++, j++) {
      orc_vcycle(h, sv, v[i], zv[i]);
      matvec_(L, L->A, zv[i], w);
      for (int k = 0; k <= i; k++) {
        H[k * R + i] = ddot(N, w, v[k]);
        for (int t = 0; t < N; t++) w[t] += -H[k * R + i] * v[k][t];
      }
      H[(i + 1) * R + i] = sqrt(ddot(N, w, w));
      memcpy(v[i + 1], w, sizeof(double) * N);
      if (H[(i + 1) * R + i] != 0)
        for (int t = 0; t < N; t++) v[i + 1][t] *= 1 / H[(i + 1) * R + i];
      for (int k = 0; k < i; k++) app_rot(&H[k * R + i], &H[(k + 1) * R + i], cs[k], sn[k]);
      gen_rot(H[i * R + i], H[(i + 1) * R + i], &cs[i], &sn[i]);
      app_rot(&H[i * R + i], &H[(i + 1) * R + i], cs[i], sn[i]);
      app_rot(&s[i], &s[i + 1], cs[i], sn[i]);
      res = fabs(s[i + 1]);
      if (hist) hist[nh] = res;
      nh++;
      if (res < res_norm_min || res < red_factor * res0) {
        memset(w, 0, sizeof(double) * N);
        update_iter(w, N, i, H, R, s, zv);
        for (int t = 0; t < N; t++) sol[t] += w[t];
        ret = j;
        goto done;
      }
      if (res > sv->div_factor * res0) { ret = -2; goto done; }
    }
    memset(v[0], 0, sizeof(double) * N);
    update_iter(v[0], N, R - 1, H, R, s, zv);
    for (int t = 0; t < N; t++) sol[t] += v[0][t];
    defect_(L, sol, rhs, v[0]);
    res = sqrt(ddot(N, v[0], v[0]));
    res0 = beta = res;
  }
  ret = sv->maxit;
done:
  for (int k = 0; k <= R; k++) { free(v[k]); free(zv[k]); }
  free(v); free(zv); free(H); free(s); free(cs); free(sn); free(w);
  return ret;
}

/* ------------------------------------------------------------------ TCD3D */

static void interp_mapped(const orc_hier *h, int l, int fn, const orc_coeffs *co, double *u);

void orc_interp_pulse(const orc_hier *h, int l, const orc_coeffs *co, double *u) {
  const level_t *L = h->lv + l;
  if (L->vx) {
    interp_mapped(h, l, 2, co, u);
    return;
  }
  const double *pc = co->pulse_c;
  for (int d = 0; d < L->ndof; d++) {
    int node = L->dof_node[d];
    double x = (node % L->nn) * (h->box[0] / (h->order * L->n));
    double y = ((node / L->nn) % L->nn) * (h->box[1] / (h->order * L->n));
    double z = (node / (L->nn * L->nn)) * (h->box[2] / (h->order * L->n));
    double r2 = (x - pc[0]) * (x - pc[0]) + (y - pc[1]) * (y - pc[1]) + (z - pc[2]) * (z - pc[2]);
    u[d] = exp(-r2 / (2 * 0.05 * 0.05));
  }
}

/* ------------------------------------------------------------------ diagnostics */

/* Known functions: 0 zero; 1 the sine of Examples/CD_3D/Laplace.h:12-19
 * (value and gradient); 2 the Gaussian pulse (centre co->pulse_c, sigma 0.05);
 * 3 the linear function x + 2y + 3z. */
static void known_fn(int fn, const orc_coeffs *co, double x, double y, double z, double *v) {
  v[0] = v[1] = v[2] = v[3] = 0;
  if (fn == 1) {
    v[0] = sin(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    v[1] = M_PI * cos(M_PI * x) * sin(M_PI * y) * sin(M_PI * z);
    v[2] = M_PI * sin(M_PI * x) * cos(M_PI * y) * sin(M_PI * z);
    v[3] = M_PI * sin(M_PI * x) * sin(M_PI * y) * cos(M_PI * z);
  } else if (fn == 2) {
    const double *pc = co->pulse_c, s2 = 0.05 * 0.05;
    double dx = x - pc[0], dy = y - pc[1], dz = z - pc[2];
    v[0] = exp(-(dx * dx + dy * dy + dz * dz) / (2 * s2));
    v[1] = -dx / s2 * v[0];
    v[2] = -dy / s2 * v[0];
    v[3] = -dz / s2 * v[0];
  } else if (fn == 3) { /* x + 2y + 3z: in every Q1/Q2 space on trilinear cells */
    v[0] = x + 2 * y + 3 * z;
    v[1] = 1; v[2] = 2; v[3] = 3;
  }
}

/* TFEFunction3D::Interpolate (FEFunction3D.C:1446-1560) for Lagrange Q1/Q2:
 * the nodal functionals are point values at the nodes. */
/* on a mapped hierarchy: cell by cell, the nodes' reference points mapped by
 * the cell's trilinear map (GetOrigFromRef); a DOF shared by several cells
 * keeps the value of the last of them, as in the reference's loop */
static void interp_mapped(const orc_hier *h, int l, int fn, const orc_coeffs *co, double *u) {
  const level_t *L = h->lv + l;
  int p = h->order, nb = h->nb, q = p + 1;
  memset(u, 0, sizeof(double) * L->ndof);
  for (int c = 0; c < L->ncell; c++) {
    cellgeo_t G;
    geo_cell(h, L, c, &G);
    for (int b = 0; b < nb; b++) {
      int i = b % q, j = (b / q) % q, k = b / (q * q);
      geopt_t gp;
      geo_point(&G, -1.0 + 2.0 * i / p, -1.0 + 2.0 * j / p, -1.0 + 2.0 * k / p, &gp);
      double v[4];
      known_fn(fn, co, gp.X[0], gp.X[1], gp.X[2], v);
      u[L->gn[c * nb + b]] = v[0];
    }
  }
}

void orc_interpolate(const orc_hier *h, int l, int fn, const orc_coeffs *co, double *u) {
  const level_t *L = h->lv + l;
  if (L->vx) {
    interp_mapped(h, l, fn, co, u);
    return;
  }
  double dn = 1.0 / (h->order * L->n);
  for (int d = 0; d < L->ndof; d++) {
    int node = L->dof_node[d];
    double v[4];
    known_fn(fn, co, (node % L->nn) * (h->box[0] * dn), ((node / L->nn) % L->nn) * (h->box[1] * dn),
             (node / (L->nn * L->nn)) * (h->box[2] * dn), v);
    u[d] = v[0];
  }
}

/* TFEFunction3D::GetErrors with L2H1Errors (FEFunction3D.C:57-240,
 * MainUtilities.C:1623-1653): per cell, at the quadrature points of the
 * assembly, e = u_h - u and grad e; err[0] = sqrt(sum w|J| e^2), err[1] =
 * sqrt(sum w|J| |grad e|^2), err[2] = max |e| over the points of the LAST
 * cell (the reference resets that maximum for every cell, :211-219), err[3]
 * = max |e| over all points. */
void orc_errors(const orc_hier *h, int l, const double *u, int fn, const orc_coeffs *co, double *err) {
  const level_t *L = h->lv + l;
  int p = h->order, nb = h->nb, n1 = p + 2;
  const double *gx = p == 1 ? G3X : G4X, *gw = p == 1 ? G3W : G4W;
  double e0 = 0, e1 = 0, emax = 0, elast = 0;
  for (int c = 0; c < L->ncell; c++) {
    const int *g = L->gn + c * nb;
    cellgeo_t G;
    geo_cell(h, L, c, &G);
    double cellmax = 0;
    for (int qi = 0; qi < n1; qi++)
      for (int qj = 0; qj < n1; qj++)
        for (int qk = 0; qk < n1; qk++) {
          double xi = gx[qk], eta = gx[qj], zeta = gx[qi];
          geopt_t gp;
          geo_point(&G, xi, eta, zeta, &gp);
          double w = gw[qi] * gw[qj] * gw[qk] * gp.absdet;
          const double *X = gp.X;
          double phi[MAXNB], d0[MAXNB], d1[MAXNB], d2[MAXNB];
          basis3d(p, xi, eta, zeta, phi, d0, d1, d2);
          double uh = 0, gr[3] = {0, 0, 0};
          for (int b = 0; b < nb; b++) {
            double ub = u[g[b]], gb[3];
            uh += ub * phi[b];
            geo_grad(&G, &gp, d0[b], d1[b], d2[b], gb);
            for (int d = 0; d < 3; d++) gr[d] += ub * gb[d];
          }
          double v[4];
          known_fn(fn, co, X[0], X[1], X[2], v);
          double t = uh - v[0];
          e0 += w * t * t;
          for (int d = 0; d < 3; d++) e1 += w * (gr[d] - v[1 + d]) * (gr[d] - v[1 + d]);
          if (fabs(t) > cellmax) cellmax = fabs(t);
        }
    if (cellmax > emax) emax = cellmax;
    elast = cellmax;
  }
  err[0] = sqrt(e0);
  err[1] = sqrt(e1);
  err[2] = elast;
  err[3] = emax;
}

/* MatAdd, active rows only (Blas2.C:132-150) */
static void matadd(level_t *L, double *C, const double *D, double alpha) {
  long n = L->rowptr[L->nact];
  for (long k = 0; k < n; k++) C[k] += alpha * D[k];
}

/* Crank-Nicolson loop of 3DPrograms/TCD3D_ParMooN.C:539-637 over
 * TSystemTCD3D (SystemTCD3D.C:176-376): AssembleMRhs once, then per step
 * AssembleARhs at t_new, AssembleSystMat, Solve, RestoreMassMat.
 * SUPG (co->supg): M and the SUPG rhs of MatrixMRhsAssemble_SUPG at t0; per
 * step A, K and rhs of MatricesAKRhsAssemble_SUPG (TCD3D.C:48,102), and K
 * beside M in the system and rhs matrices, added and removed where the SUPG
 * branch of a reference time loop does it (Main_Users/Sashi/3DPrograms/
 * TCD_3D.C:1184-1188,1314-1317): system M + K + tau th1 A, rhs matrix
 * M + K - tau th2 A. */
int orc_tcd_run(orc_hier *h, const orc_coeffs *co0, const orc_solver *sv, double t0, double tau, int nsteps,
                double *u, double *states) {
  int Lf = h->L - 1, tot = 0;
  orc_coeffs co = *co0;
  double th1 = 0.5, th2 = 0.5, th3 = 0.5, th4 = 0.5;
  double **M = malloc(sizeof(double *) * h->L), **rhsl = malloc(sizeof(double *) * h->L);
  for (int l = 0; l < h->L; l++) {
    level_t *L = h->lv + l;
    if (!L->Aconv) L->Aconv = calloc(L->nnz, sizeof(double));
    if (co.supg && !L->Ksupg) L->Ksupg = calloc(L->nnz, sizeof(double));
    M[l] = L->A;
    rhsl[l] = L->rhs;
  }
  level_t *F = h->lv + Lf;
  int N = F->ndof, NA = F->nact;
  co.t = t0;
  for (int l = 0; l < h->L; l++) orc_assemble_level(h, l, &co, 1, M[l], rhsl[l]);
  memcpy(u + NA, F->rhs + NA, sizeof(double) * (N - NA));
  double *oldrhs = malloc(sizeof(double) * N), *oldsol = malloc(sizeof(double) * N);
  double *B = malloc(sizeof(double) * N), *def = malloc(sizeof(double) * N), *x = malloc(sizeof(double) * N);
  memcpy(oldrhs, F->rhs, sizeof(double) * N);
  double t = t0;
  for (int m = 0; m < nsteps; m++) {
    memcpy(oldsol, u, sizeof(double) * N);
    t += tau;
    co.t = t;
    for (int l = 0; l < h->L; l++) {
      level_t *L = h->lv + l;
      assemble_level_t(h, l, &co, 0, th1 * tau, L->Aconv, rhsl[l]);
      if (co.supg) {
        double *scratch = calloc(L->ndof, sizeof(double));
        assemble_level_t(h, l, &co, 2, th1 * tau, L->Ksupg, scratch);
        free(scratch);
      }
    }
    memcpy(u + NA, F->rhs + NA, sizeof(double) * (N - NA));
    /* AssembleSystMat (SystemTCD3D.C:228-294) */
    memset(B, 0, sizeof(double) * N);
    for (int k = 0; k < NA; k++) B[k] += tau * th3 * oldrhs[k];
    for (int k = 0; k < NA; k++) B[k] += tau * th4 * F->rhs[k];
    matadd(F, M[Lf], F->Aconv, -tau * th2);
    double gamma = -tau * th2;
    if (co.supg)
      for (int l = 0; l < h->L; l++) matadd(h->lv + l, M[l], h->lv[l].Ksupg, 1.0);
    memset(def, 0, sizeof(double) * N);
    for (int i = 0; i < NA; i++) { /* MatVectActive (Blas2.C:102-129) */
      double s = 0;
      for (int j = F->rowptr[i]; j < F->rowptr[i + 1]; j++) s += M[Lf][j] * oldsol[F->kcol[j]];
      def[i] = s;
    }
    for (int k = 0; k < NA; k++) B[k] += 1 * def[k];
    memcpy(B + NA, F->rhs + NA, sizeof(double) * (N - NA));
    memcpy(u + NA, F->rhs + NA, sizeof(double) * (N - NA));
    for (int l = 0; l < h->L; l++) {
      level_t *L = h->lv + l;
      matadd(L, M[l], L->Aconv, l == Lf ? -gamma + tau * th1 : tau * th1);
    }
    gamma = tau * th1;
    memcpy(oldrhs, F->rhs, sizeof(double) * N);
    h->lu_valid = 0;
    /* Solve (SystemTCD3D.C:317-376) */
    memcpy(x, u, sizeof(double) * N);
    int its = orc_fgmres(h, sv, B, x, NULL);
    tot += its;
    memcpy(u, x, sizeof(double) * N);
    /* RestoreMassMat (SystemTCD3D.C:296-315) */
    for (int l = 0; l < h->L; l++) {
      matadd(h->lv + l, M[l], h->lv[l].Aconv, -gamma);
      if (co.supg) matadd(h->lv + l, M[l], h->lv[l].Ksupg, -1.0);
    }
    if (states) memcpy(states + (size_t)m * N, u, sizeof(double) * N);
  }
  free(oldrhs); free(oldsol); free(B); free(def); free(x); free(M); free(rhsl);
  return tot;
}
