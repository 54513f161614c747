// Driver for the reference build in oracle/_ref (checker only; never shipped
// as product code).  It wires the reference's own public call surface --
// TFESpace3D, TSquareStructure3D::Sort, TSquareMatrix3D, TDiscreteForm3D,
// TAssembleMat3D, TMGLevel3D, TMultiGrid3D, TMultiGridScaIte, TFgmresIte
// (src/System/SystemCD3D.C:36-203,227-311 wires the same objects for
// GALERKIN only; this driver also builds the SDFEM form from
// BilinearAssemble_SD, src/FE/ConvDiff3D.C:124) -- and dumps what the
// reference computes, for the committed golden fixtures.
//
// usage: cd3d_ref <file.dat> <outdir|-> [mode]
//   mode "cd"  (default): assemble all levels, V-cycle probe, FGMRES solve
//   mode "tcd": Crank-Nicolson steps of the TCD3D system (TSystemTCD3D)
//   mode "tcdsupg": Crank-Nicolson steps of the SUPG-stabilised TCD3D system:
//     the mass matrix and rhs of MatrixMRhsAssemble_SUPG (TCD3D.C:48) at t0,
//     then per step A, K and rhs of MatricesAKRhsAssemble_SUPG (TCD3D.C:102)
//     at t_new and the steps of TSystemTCD3D::AssembleSystMat / Solve /
//     RestoreMassMat (SystemTCD3D.C:228-376) with the SUPG matrix K added to
//     M wherever the SUPG branch of a reference time loop adds it
//     (Main_Users/Sashi/3DPrograms/TCD_3D.C:1184-1188,1314-1317): system
//     M + K + tau theta1 A, rhs tau theta3 f_old + tau theta4 f_new +
//     (M + K - tau theta2 A) u_old; GMG-preconditioned FGMRES as in "cd",
//     stopping at SC_LIN_RED_FACTOR_SCALAR times each step's initial residual
//
// Coefficients come from the .dat P-parameters:
//   P0 eps, P1..P3 convection b (P7 1: rotating field
//   b = (-(y - P13), x - P12, 0.1 sin(2 pi t)) of the TCD3D config),
//   P4 reaction c, P5 source kind (0: f = P6, 1: manufactured
//   u = sin(pi x) sin(pi y) sin(pi z)), P8 number of CN steps,
//   P9..P11 centre of the Gaussian pulse initial condition (sigma 0.05).
//   P7 2: coefficients varying in space, evaluated by this CoeffFct3D at
//   every quadrature point: eps = P0 (1 + xyz), b = (P1 sin(pi y),
//   P2 cos(pi x), P3 z), c = P4 x, f = P6 (1 + x).
// The mesh is what the reference builds from data/Wuerfel.GEO: ReadGeo.C:1753-1755
// scales parametrised vertices by (1, 0.346, 0.346), so the domain is the box
// [0,1] x [0,0.346] x [0,0.346].
// Boundary: Dirichlet u = 0 on every face.

#include <Domain.h>
#include <Database.h>
#include <FEDatabase3D.h>
#include <FESpace3D.h>
#include <SquareStructure3D.h>
#include <SquareMatrix3D.h>
#include <DiscreteForm3D.h>
#include <AssembleMat3D.h>
#include <AuxParam3D.h>
#include <MultiGrid3D.h>
#include <MGLevel3D.h>
#include <MultiGridScaIte.h>
#include <FgmresIte.h>
// ConvDiff3D.h defines data objects; declare the two kernels used here (ConvDiff3D.C:61,124)
void BilinearAssemble(double, double *, double, double **, int *, double ***, double **);
void BilinearAssemble_SD(double, double *, double, double **, int *, double ***, double **);
// TCD3D SUPG kernels (TCD3D.C:48,102)
void MatrixMRhsAssemble_SUPG(double, double *, double *, double, double **, int *, double ***, double **);
void MatricesAKRhsAssemble_SUPG(double, double *, double *, double, double **, int *, double ***, double **);
#include <SystemTCD3D.h>
#include <LinAlg.h>
#include <MainUtilities.h>
#include <MooNMD_Io.h>
#include <It_Finest.h>
#include <TimeDiscRout.h>
#include <FEFunction3D.h>
#include <Output3D.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <iostream>
#include <iomanip>

void SetPolynomialDegree();

static std::string g_out;


static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static void dump(const std::string &name, const void *p, size_t bytes) {
  if (g_out == "-") return;
  FILE *f = fopen((g_out + "/" + name).c_str(), "wb");
  if (!f) { perror(name.c_str()); exit(2); }
  fwrite(p, 1, bytes, f);
  fclose(f);
}

// ---------------------------------------------------------------- callbacks
static void conv_field(double x, double y, double z, double t, double &b1, double &b2, double &b3) {
  if (TDatabase::ParamDB->P7 == 2) {  // variable field of the per-point coefficient cases
    b1 = TDatabase::ParamDB->P1 * sin(Pi * y);
    b2 = TDatabase::ParamDB->P2 * cos(Pi * x);
    b3 = TDatabase::ParamDB->P3 * z;
  } else if (TDatabase::ParamDB->P7 > 0) {
    b1 = -(y - TDatabase::ParamDB->P13);
    b2 = x - TDatabase::ParamDB->P12;
    b3 = 0.1 * sin(2 * Pi * t);
  } else {
    b1 = TDatabase::ParamDB->P1;
    b2 = TDatabase::ParamDB->P2;
    b3 = TDatabase::ParamDB->P3;
  }
}

void DrvCoeffs(int n_points, double *X, double *Y, double *Z, double **parameters, double **coeffs) {
  double eps = TDatabase::ParamDB->P0, c = TDatabase::ParamDB->P4;
  double t = TDatabase::TimeDB->CURRENTTIME;
  for (int i = 0; i < n_points; i++) {
    double *co = coeffs[i], b1, b2, b3;
    conv_field(X[i], Y[i], Z[i], t, b1, b2, b3);
    co[0] = eps;
    co[1] = b1; co[2] = b2; co[3] = b3;
    co[4] = c;
    if ((int)TDatabase::ParamDB->P7 == 2) {  // eps, c and f vary too
      co[0] = eps * (1 + X[i] * Y[i] * Z[i]);
      co[4] = c * X[i];
      co[5] = TDatabase::ParamDB->P6 * (1 + X[i]);
      co[6] = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
      continue;
    }
    if ((int)TDatabase::ParamDB->P5 == 1) {
      double sx = sin(Pi * X[i]), sy = sin(Pi * Y[i]), sz = sin(Pi * Z[i]);
      double cx = cos(Pi * X[i]), cy = cos(Pi * Y[i]), cz = cos(Pi * Z[i]);
      double u = sx * sy * sz;
      co[5] = eps * 3 * Pi * Pi * u + b1 * Pi * cx * sy * sz + b2 * Pi * sx * cy * sz +
              b3 * Pi * sx * sy * cz + c * u;
    } else {
      co[5] = TDatabase::ParamDB->P6;
    }
    co[6] = sqrt(b1 * b1 + b2 * b2 + b3 * b3);
  }
}

// the exact solution of Examples/CD_3D/Laplace.h (value, gradient, Laplacian)
void DrvExact(double x, double y, double z, double *values) {
  values[0] = sin(Pi * x) * sin(Pi * y) * sin(Pi * z);
  values[1] = Pi * cos(Pi * x) * sin(Pi * y) * sin(Pi * z);
  values[2] = Pi * sin(Pi * x) * cos(Pi * y) * sin(Pi * z);
  values[3] = Pi * sin(Pi * x) * sin(Pi * y) * cos(Pi * z);
  values[4] = -3 * Pi * Pi * sin(Pi * x) * sin(Pi * y) * sin(Pi * z);
}

void DrvBoundCondition(int, double, double, double, BoundCond &cond) { cond = DIRICHLET; }
void DrvBoundValue(int, double, double, double, double &value) { value = 0; }
void DrvInitialCondition(double x, double y, double z, double *values) {
  double cx = TDatabase::ParamDB->P9, cy = TDatabase::ParamDB->P10, cz = TDatabase::ParamDB->P11;
  double s = 0.05, r2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
  values[0] = exp(-r2 / (2 * s * s));
}

// ---------------------------------------------------------------- helpers
static double norm2(int n, const double *v) {
  double s = 0;
  for (int i = 0; i < n; i++) s += v[i] * v[i];
  return sqrt(s);
}

static void dump_level(int lev, TFESpace3D *sp, TSquareMatrix3D *A, double *rhs) {
  char b[64];
  TCollection *coll = sp->GetCollection();
  int nc = coll->GetN_Cells();
  std::vector<double> xyz(nc * 24);
  for (int c = 0; c < nc; c++) {
    TBaseCell *cell = coll->GetCell(c);
    for (int v = 0; v < 8; v++)
      cell->GetVertex(v)->GetCoords(xyz[c * 24 + 3 * v], xyz[c * 24 + 3 * v + 1], xyz[c * 24 + 3 * v + 2]);
  }
  int nb = sp->GetBeginIndex()[1];
  snprintf(b, sizeof b, "L%d_cells.f64", lev); dump(b, xyz.data(), xyz.size() * 8);
  snprintf(b, sizeof b, "L%d_gn.i32", lev); dump(b, sp->GetGlobalNumbers(), (size_t)nc * nb * 4);
  int N = A->GetN_Rows(), nnz = A->GetN_Entries();
  snprintf(b, sizeof b, "L%d_rowptr.i32", lev); dump(b, A->GetRowPtr(), (size_t)(N + 1) * 4);
  snprintf(b, sizeof b, "L%d_kcol.i32", lev); dump(b, A->GetKCol(), (size_t)nnz * 4);
  snprintf(b, sizeof b, "L%d_entries.f64", lev); dump(b, A->GetEntries(), (size_t)nnz * 8);
  snprintf(b, sizeof b, "L%d_rhs.f64", lev); dump(b, rhs, (size_t)N * 8);
  printf("LEVEL %d cells %d nbf %d ndof %d active %d hanging_bound %d dirichlet %d nnz %d\n", lev, nc, nb,
         sp->GetN_DegreesOfFreedom(), sp->GetN_ActiveDegrees(), sp->GetHangingBound(), sp->GetN_Dirichlet(), nnz);
}

// ---------------------------------------------------------------- main
int main(int argc, char **argv) {
  if (argc < 3) { fprintf(stderr, "usage: %s file.dat outdir|- [cd|tcd|tcdsupg]\n", argv[0]); return 1; }
  g_out = argv[2];
  std::string mode = argc > 3 ? argv[3] : "cd";
  TDatabase *Database = new TDatabase();
  TFEDatabase3D *FEDatabase = new TFEDatabase3D();
  TDomain *Domain = new TDomain(argv[1]);
  OpenFiles();
  std::cout << std::setprecision(17);
  OutFile << std::setprecision(17);
  if (TDatabase::ParamDB->DISCTYPE == SDFEM) SetPolynomialDegree();

  Domain->Init(TDatabase::ParamDB->BNDFILE, TDatabase::ParamDB->GEOFILE);
  for (int i = 0; i < TDatabase::ParamDB->UNIFORM_STEPS; i++) Domain->RegRefineAll();

  int L = TDatabase::ParamDB->LEVELS, ORDER = TDatabase::ParamDB->ANSATZ_ORDER;
  char Name[] = "name", Desc[] = "desc";
  std::vector<TFESpace3D *> sp(L);
  std::vector<TSquareMatrix3D *> A(L), M(L);
  std::vector<double *> sol(L), rhs(L), oldrhs(L);
  double t0 = now_s();
  for (int i = 0; i < L; i++) {
    if (i) Domain->RegRefineAll();
    TCollection *coll = Domain->GetCollection(It_Finest, 0);
    sp[i] = new TFESpace3D(coll, Name, Desc, DrvBoundCondition, ORDER);
    int N = sp[i]->GetN_DegreesOfFreedom();
    sol[i] = new double[N]();
    rhs[i] = new double[N]();
  }
  double t_space = now_s() - t0;

  if (mode == "tcd") {
    // TSystemTCD3D path (src/System/SystemTCD3D.C), Crank-Nicolson, following
    // the time loop of 3DPrograms/TCD3D_ParMooN.C:539-637.
    TDatabase::TimeDB->TIME_DISC = 2;
    int Nf = sp[L - 1]->GetN_DegreesOfFreedom();
    TFEFunction3D *u0 = new TFEFunction3D(sp[L - 1], Name, Desc, sol[L - 1], Nf);
    u0->Interpolate(DrvInitialCondition);
    printf("TCD interpolated |u0| %.17e\n", norm2(Nf, sol[L - 1]));
    TSystemTCD3D *S = new TSystemTCD3D(L, sp.data(), sol.data(), rhs.data(), TDatabase::ParamDB->DISCTYPE,
                                       TDatabase::ParamDB->SOLVER_TYPE);
    TDatabase::TimeDB->CURRENTTIME = TDatabase::TimeDB->STARTTIME;
    S->Init(DrvCoeffs, DrvBoundCondition, DrvBoundValue, NULL);
    printf("TCD after Init |u0| %.17e\n", norm2(Nf, sol[L - 1]));
    S->AssembleMRhs();
    printf("TCD after AssembleMRhs |u0| %.17e\n", norm2(Nf, sol[L - 1]));
    std::vector<double> old_rhs(rhs[L - 1], rhs[L - 1] + Nf), old_sol(Nf);
    dump("tcd_u0.f64", sol[L - 1], (size_t)Nf * 8);
    int nsteps = (int)TDatabase::ParamDB->P8;
    double ts = now_s();
    for (int m = 0; m < nsteps; m++) {
      SetTimeDiscParameters(1);
      memcpy(old_sol.data(), sol[L - 1], Nf * 8);
      double tau = TDatabase::TimeDB->CURRENTTIMESTEPLENGTH;
      TDatabase::TimeDB->CURRENTTIME += tau;
      S->AssembleARhs();
      S->AssembleSystMat(old_rhs.data(), old_sol.data(), rhs[L - 1], sol[L - 1]);
      memcpy(old_rhs.data(), rhs[L - 1], Nf * 8);
      S->Solve(sol[L - 1]);
      S->RestoreMassMat();
      char b[64];
      snprintf(b, sizeof b, "tcd_u%d.f64", m + 1);
      dump(b, sol[L - 1], (size_t)Nf * 8);
      printf("TCD step %d t %.17e |u| %.17e\n", m + 1, TDatabase::TimeDB->CURRENTTIME, norm2(Nf, sol[L - 1]));
    }
    printf("TIME tcd_steps %.6f\n", now_s() - ts);
    return 0;
  }

  if (mode == "tcdsupg") {
    TDatabase::TimeDB->TIME_DISC = 2;
    int Nf = sp[L - 1]->GetN_DegreesOfFreedom(), NAf = sp[L - 1]->GetActiveBound();
    TFEFunction3D *u0 = new TFEFunction3D(sp[L - 1], Name, Desc, sol[L - 1], Nf);
    u0->Interpolate(DrvInitialCondition);
    char cm[] = "M-SUPG", ca[] = "AK-SUPG", fm[] = "form";
    static MultiIndex3D Der[4] = {D100, D010, D001, D000};
    static int SpNum[4] = {0, 0, 0, 0}, RowSp[3] = {0, 0, 0}, ColSp[3] = {0, 0, 0}, RhsSp[1] = {0};
    TDiscreteForm3D *formM = new TDiscreteForm3D(cm, fm, 4, Der, SpNum, 1, 1, RowSp, ColSp, RhsSp,
                                                 MatrixMRhsAssemble_SUPG, DrvCoeffs, NULL);
    TDiscreteForm3D *formAK = new TDiscreteForm3D(ca, fm, 4, Der, SpNum, 3, 1, RowSp, ColSp, RhsSp,
                                                  MatricesAKRhsAssemble_SUPG, DrvCoeffs, NULL);
    BoundCondFunct3D *BC[3] = {DrvBoundCondition, DrvBoundCondition, DrvBoundCondition};
    BoundValueFunct3D *BV[3] = {DrvBoundValue, DrvBoundValue, DrvBoundValue};
    std::vector<TSquareMatrix3D *> Am(L), Km(L), Sm(L);
    std::vector<TAssembleMat3D *> asmM(L), asmAK(L);
    for (int i = 0; i < L; i++) {
      TSquareStructure3D *st = new TSquareStructure3D(sp[i]);
      st->Sort();
      M[i] = new TSquareMatrix3D(st);
      Am[i] = new TSquareMatrix3D(st);
      Km[i] = new TSquareMatrix3D(st);
      Sm[i] = new TSquareMatrix3D(st);
      TFESpace3D *fesp[1] = {sp[i]};
      TSquareMatrix3D *SQM1[1] = {M[i]}, *SQM3[3] = {Am[i], Km[i], Sm[i]};
      double *R[1] = {rhs[i]};
      TAuxParam3D *aux = new TAuxParam3D(1, 0, 0, 0, fesp, NULL, NULL, NULL, NULL, 0, NULL);
      asmM[i] = new TAssembleMat3D(1, fesp, 1, SQM1, 0, NULL, 1, R, fesp, formM, BC, BV, aux);
      asmM[i]->Init();
      asmAK[i] = new TAssembleMat3D(1, fesp, 3, SQM3, 0, NULL, 1, R, fesp, formAK, BC, BV, aux);
      asmAK[i]->Init();
    }
    TDatabase::TimeDB->CURRENTTIME = TDatabase::TimeDB->STARTTIME;
    for (int i = 0; i < L; i++) {  // TSystemTCD3D::AssembleMRhs (SystemTCD3D.C:176-202)
      int N = sp[i]->GetN_DegreesOfFreedom(), NA = sp[i]->GetActiveBound();
      asmM[i]->Reset();
      asmM[i]->Assemble3D();
      memcpy(sol[i] + NA, rhs[i] + NA, (N - NA) * 8);
    }
    dump("tcd_u0.f64", sol[L - 1], (size_t)Nf * 8);
    double Parameters[2] = {TDatabase::ParamDB->SC_SMOOTH_DAMP_FACTOR_SCALAR,
                            TDatabase::ParamDB->SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR};
    TMultiGrid3D *MG = new TMultiGrid3D(1, 2, Parameters);
    for (int i = 0; i < L; i++) MG->AddLevel(new TMGLevel3D(i, M[i], rhs[i], sol[i], 2, NULL));
    TSquareMatrix *SQM[1] = {M[L - 1]};
    TMultiGridScaIte *prec = new TMultiGridScaIte(MatVect_Scalar, Defect_Scalar, NULL, 0, Nf, MG, 1);
    TFgmresIte *fg = new TFgmresIte(MatVect_Scalar, Defect_Scalar, prec, 0, Nf, 1);
    std::vector<double> oldrhs(rhs[L - 1], rhs[L - 1] + Nf), oldsol(Nf), B(Nf), def(Nf), x(Nf), b(Nf);
    int nsteps = (int)TDatabase::ParamDB->P8;
    const double rel = TDatabase::ParamDB->SC_LIN_RED_FACTOR_SCALAR;
    for (int m = 0; m < nsteps; m++) {
      SetTimeDiscParameters(1);
      memcpy(oldsol.data(), sol[L - 1], Nf * 8);
      double tau = TDatabase::TimeDB->CURRENTTIMESTEPLENGTH;
      double th1 = TDatabase::TimeDB->THETA1, th2 = TDatabase::TimeDB->THETA2;
      double th3 = TDatabase::TimeDB->THETA3, th4 = TDatabase::TimeDB->THETA4;
      TDatabase::TimeDB->CURRENTTIME += tau;
      for (int i = 0; i < L; i++) {  // AssembleARhs with the SUPG form (A, K and rhs)
        int N = sp[i]->GetN_DegreesOfFreedom(), NA = sp[i]->GetActiveBound();
        asmAK[i]->Reset();
        asmAK[i]->Assemble3D();
        memcpy(sol[i] + NA, rhs[i] + NA, (N - NA) * 8);
      }
      // AssembleSystMat (SystemTCD3D.C:228-294) with K
      std::fill(B.begin(), B.end(), 0.0);
      Daxpy(NAf, tau * th3, oldrhs.data(), B.data());
      Daxpy(NAf, tau * th4, rhs[L - 1], B.data());
      MatAdd(M[L - 1], Am[L - 1], -tau * th2);
      double gamma = -tau * th2;
      for (int i = 0; i < L; i++) MatAdd(M[i], Km[i], 1.0);
      std::fill(def.begin(), def.end(), 0.0);
      MatVectActive(M[L - 1], oldsol.data(), def.data());
      Daxpy(NAf, 1.0, def.data(), B.data());
      memcpy(B.data() + NAf, rhs[L - 1] + NAf, (Nf - NAf) * 8);
      memcpy(sol[L - 1] + NAf, rhs[L - 1] + NAf, (Nf - NAf) * 8);
      for (int i = 0; i < L; i++) MatAdd(M[i], Am[i], i == L - 1 ? -gamma + tau * th1 : tau * th1);
      gamma = tau * th1;
      memcpy(oldrhs.data(), rhs[L - 1], Nf * 8);
      // Solve (SystemTCD3D.C:317-376): the solver works on copies; the stop
      // is relative to this step's initial residual (the reference's
      // SC_LIN_RED_FACTOR_SCALAR test restarts from each restart cycle)
      memcpy(x.data(), sol[L - 1], Nf * 8);
      memcpy(b.data(), B.data(), Nf * 8);
      double r0;
      ScalarDefect(M[L - 1], x.data(), b.data(), def.data(), r0);
      TDatabase::ParamDB->SC_LIN_RES_NORM_MIN_SCALAR = rel * r0;
      TDatabase::ParamDB->SC_LIN_RED_FACTOR_SCALAR = 0.0;
      int its = fg->Iterate(SQM, NULL, x.data(), b.data());
      memcpy(sol[L - 1], x.data(), Nf * 8);
      // RestoreMassMat (SystemTCD3D.C:296-315) and the SUPG matrix
      for (int i = 0; i < L; i++) {
        MatAdd(M[i], Am[i], -gamma);
        MatAdd(M[i], Km[i], -1.0);
      }
      char bn[64];
      snprintf(bn, sizeof bn, "tcd_u%d.f64", m + 1);
      dump(bn, sol[L - 1], (size_t)Nf * 8);
      printf("TCD step %d t %.17e its %d |u| %.17e\n", m + 1, TDatabase::TimeDB->CURRENTTIME, its,
             norm2(Nf, sol[L - 1]));
    }
    return 0;
  }

  // ---- CD3D: structures, matrices, discrete form (Galerkin or SDFEM)
  t0 = now_s();
  char cd[] = "Conv-Diff", gal[] = "form";
  static MultiIndex3D Der[4] = {D100, D010, D001, D000};
  static int SpNum[4] = {0, 0, 0, 0}, RowSp[1] = {0}, ColSp[1] = {0}, RhsSp[1] = {0};
  AssembleFct3D *kern = TDatabase::ParamDB->DISCTYPE == SDFEM ? BilinearAssemble_SD : BilinearAssemble;
  TDiscreteForm3D *form = new TDiscreteForm3D(cd, gal, 4, Der, SpNum, 1, 1, RowSp, ColSp, RhsSp, kern, DrvCoeffs, NULL);
  BoundCondFunct3D *BC[1] = {DrvBoundCondition};
  BoundValueFunct3D *BV[1] = {DrvBoundValue};
  std::vector<TAssembleMat3D *> asmb(L);
  for (int i = 0; i < L; i++) {
    TSquareStructure3D *st = new TSquareStructure3D(sp[i]);
    st->Sort();
    A[i] = new TSquareMatrix3D(st);
    TFESpace3D *fesp[1] = {sp[i]};
    TSquareMatrix3D *SQ[1] = {A[i]};
    double *R[1] = {rhs[i]};
    TAuxParam3D *aux = new TAuxParam3D(1, 0, 0, 0, fesp, NULL, NULL, NULL, NULL, 0, NULL);
    asmb[i] = new TAssembleMat3D(1, fesp, 1, SQ, 0, NULL, 1, R, fesp, form, BC, BV, aux);
    asmb[i]->Init();
  }
  double t_struct = now_s() - t0;
  t0 = now_s();
  for (int i = 0; i < L; i++) {
    int N = sp[i]->GetN_DegreesOfFreedom(), NA = sp[i]->GetActiveBound();
    asmb[i]->Reset();
    asmb[i]->Assemble3D();
    memcpy(sol[i] + NA, rhs[i] + NA, (N - NA) * 8);  // SystemCD3D.C:330
  }
  double t_asm = now_s() - t0;
  for (int i = 0; i < L; i++) dump_level(i, sp[i], A[i], rhs[i]);

  // ---- multigrid hierarchy + FGMRES (SystemCD3D.C:151-203)
  double Parameters[2] = {TDatabase::ParamDB->SC_SMOOTH_DAMP_FACTOR_SCALAR,
                          TDatabase::ParamDB->SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR};
  TMultiGrid3D *MG = new TMultiGrid3D(1, 2, Parameters);
  for (int i = 0; i < L; i++) MG->AddLevel(new TMGLevel3D(i, A[i], rhs[i], sol[i], 2, NULL));
  int N = sp[L - 1]->GetN_DegreesOfFreedom();
  TSquareMatrix *SQM[1] = {A[L - 1]};
  TMultiGridScaIte *prec = new TMultiGridScaIte(MatVect_Scalar, Defect_Scalar, NULL, 0, N, MG, 1);

  // the solver works on its own copies (SystemCD3D.C:351-355)
  std::vector<double> x(sol[L - 1], sol[L - 1] + N), b(rhs[L - 1], rhs[L - 1] + N), r(N), z(N);
  double r0;
  ScalarDefect(A[L - 1], x.data(), b.data(), r.data(), r0);

  // V-cycle probe: z = V(r / |r|)
  std::vector<double> rn(r);
  for (int i = 0; i < N; i++) rn[i] /= r0;
  std::vector<double> save_sol(sol[L - 1], sol[L - 1] + N), save_rhs(rhs[L - 1], rhs[L - 1] + N);
  prec->Iterate(SQM, NULL, z.data(), rn.data());
  dump("vcycle_in.f64", rn.data(), (size_t)N * 8);
  dump("vcycle_out.f64", z.data(), (size_t)N * 8);
  memcpy(sol[L - 1], save_sol.data(), N * 8);
  memcpy(rhs[L - 1], save_rhs.data(), N * 8);

  double rel = TDatabase::ParamDB->SC_LIN_RED_FACTOR_SCALAR;
  TDatabase::ParamDB->SC_LIN_RES_NORM_MIN_SCALAR = rel * r0;
  TDatabase::ParamDB->SC_LIN_RED_FACTOR_SCALAR = 0.0;
  TFgmresIte *fg = new TFgmresIte(MatVect_Scalar, Defect_Scalar, prec, 0, N, 1);
  double ts = now_s();
  int its = fg->Iterate(SQM, NULL, x.data(), b.data());
  double t_solve = now_s() - ts;
  double rf;
  ScalarDefect(A[L - 1], x.data(), b.data(), r.data(), rf);
  dump("sol.f64", x.data(), (size_t)N * 8);
  printf("SOLVE its %d r0 %.17e rfinal %.17e rel %.17e\n", its, r0, rf, rf / r0);

  // ---- diagnostics and output of the solution (CD3D_ParMooN.C:496-520):
  // L2/H1 errors against the sine of Laplace.h (TFEFunction3D::GetErrors with
  // L2H1Errors), the nodal interpolant of the sine and its errors
  // (Interpolate), the VTK file (TOutput3D::WriteVtk) and the binary matrix
  // (TSquareMatrix::Write) of the finest level
  {
    char nm[] = "u", ds[] = "solution";
    static MultiIndex3D AllDer[4] = {D000, D100, D010, D001};
    TFESpace3D *fespL[1] = {sp[L - 1]};
    TAuxParam3D *auxE = new TAuxParam3D(1, 0, 0, 0, fespL, NULL, NULL, NULL, NULL, 0, NULL);
    TFEFunction3D *fsol = new TFEFunction3D(sp[L - 1], nm, ds, x.data(), N);
    double err[4] = {0, 0, 0, 0};
    fsol->GetErrors(DrvExact, 4, AllDer, 2, L2H1Errors, DrvCoeffs, auxE, 1, fespL, err);
    dump("errors_sol.f64", err, 3 * 8);
    std::vector<double> ui(N, 0.0);
    TFEFunction3D *fint = new TFEFunction3D(sp[L - 1], nm, ds, ui.data(), N);
    fint->Interpolate(DrvExact);
    dump("interp_sine.f64", ui.data(), (size_t)N * 8);
    double erri[4] = {0, 0, 0, 0};
    fint->GetErrors(DrvExact, 4, AllDer, 2, L2H1Errors, DrvCoeffs, auxE, 1, fespL, erri);
    dump("errors_interp.f64", erri, 3 * 8);
    printf("ERRORS sol %.17e %.17e %.17e interp %.17e %.17e %.17e\n", err[0], err[1], err[2], erri[0], erri[1],
           erri[2]);
    if (g_out != "-") {
      TOutput3D *out = new TOutput3D(1, 1, 0, 0, Domain, sp[L - 1]->GetCollection());
      out->AddFEFunction(fsol);
      out->WriteVtk((g_out + "/sol.vtk").c_str());
      A[L - 1]->Write((g_out + "/matrix.bin").c_str());
    }
  }
  printf("TIME space %.6f struct %.6f assemble %.6f solve %.6f\n", t_space, t_struct, t_asm, t_solve);
  return 0;
}
