// A CD3D / TCD3D main program in the shape of 3DPrograms/CD3D_ParMooN.C and
// TCD3D_ParMooN.C, written against the drop-in classes of parmoon_cd3d.hpp.
//
//   cd3d_shim_main key=value ...
//     mode=cd|cdfct|mg|tcd   cd: TSystemCD3D (cdfct: Init with a CoeffFct3D,
//                      coefficients varying in space); mg: TMultiGrid3D + TMGLevel3D +
//                      TMultiGridScaIte + TFgmresIte wired by hand as in
//                      SystemCD3D.C:138-203; tcd: TSystemTCD3D
//     geo=<file.GEO>   mesh from a GEO file through TDomain::Init, k0 + levels - 1
//                      RegRefineAll and TFESpace3D(coll, ...) per level, as
//                      CD3D_ParMooN.C:137-330 (otherwise the box m, box=)
//     order k0 levels box=x,y,z eps b=x,y,z c f_kind f_val supg conv_kind
//     rot=x,y pulse=x,y,z  (problem)     tau nsteps time_disc (tcd)
//     tcd_stop=step  each TCD step stops at rel times its own initial residual
//     rel=1e-10  stop at |r| <= rel |r0| (SC_LIN_RES_NORM_MIN = rel |r0|)
//     mixed=0|1  out=<prefix>
// Writes <prefix>.sol (finest-level solution or states, f64) and
// <prefix>.txt (iterations, |r0|).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/parmoon_cd3d.hpp"

using namespace parmoon_cd3d;

namespace {

std::map<std::string, std::string> args;

double num(const char *k, double d) { return args.count(k) ? atof(args[k].c_str()) : d; }
void vec(const char *k, double *v, int n) {
  if (!args.count(k)) return;
  const char *s = args[k].c_str();
  for (int i = 0; i < n; i++) {
    v[i] = strtod(s, (char **)&s);
    if (*s == ',') s++;
  }
}

void write_out(const std::string &prefix, const std::vector<double> &x, int its, double r0) {
  FILE *f = fopen((prefix + ".sol").c_str(), "wb");
  if (!f || fwrite(x.data(), sizeof(double), x.size(), f) != x.size()) Fail("cannot write " + prefix + ".sol");
  fclose(f);
  f = fopen((prefix + ".txt").c_str(), "w");
  if (!f) Fail("cannot write " + prefix + ".txt");
  fprintf(f, "its %d\nr0 %.17g\n", its, r0);
  fclose(f);
}

// BoundCondFunct3D of the examples: Dirichlet everywhere
void BoundCondition(int, double, double, double, int &cond) { cond = 0; }

// A coefficient function in the shape of an Examples/CD_3D/*.h BilinearCoeffs
// (coefficients varying in space: eps (1 + xyz), b = (b1 sin(pi y),
// b2 cos(pi x), b3 z), c x, f (1 + x)); its constants come from the command line
double g_eps = 1.0, g_b[3] = {0, 0, 0}, g_c = 0.0, g_f = 1.0;
void VarCoeffs(int n_points, double *X, double *Y, double *Z, double **, double **coeffs) {
  for (int i = 0; i < n_points; i++) {
    double *co = coeffs[i];
    co[0] = g_eps * (1 + X[i] * Y[i] * Z[i]);
    co[1] = g_b[0] * sin(M_PI * Y[i]);
    co[2] = g_b[1] * cos(M_PI * X[i]);
    co[3] = g_b[2] * Z[i];
    co[4] = g_c * X[i];
    co[5] = g_f * (1 + X[i]);
    co[6] = sqrt(co[1] * co[1] + co[2] * co[2] + co[3] * co[3]);
  }
}

}  // namespace

int main(int argc, char **argv) {
  for (int i = 1; i < argc; i++) {
    const char *e = strchr(argv[i], '=');
    if (e) args[std::string(argv[i], e - argv[i])] = e + 1;
  }
  const std::string mode = args.count("mode") ? args["mode"] : "cd";
  const std::string out = args.count("out") ? args["out"] : "cd3d_shim";

  // ParamDB, as read from the .dat file by the reference
  TParamDB &db = *TDatabase::ParamDB;
  db.SC_SOLVER_SCALAR = 16;
  db.SC_PRECONDITIONER_SCALAR = 5;
  db.SC_SMOOTHER_SCALAR = (int)num("smoother", 1);
  db.SC_PRE_SMOOTH_SCALAR = (int)num("pre", 2);
  db.SC_POST_SMOOTH_SCALAR = (int)num("post", 2);
  db.SC_SMOOTH_DAMP_FACTOR_SCALAR = num("omega", 0.8);
  db.SC_COARSE_SMOOTHER_SCALAR = (int)num("coarse", 3);
  db.SC_COARSE_MAXIT_SCALAR = 10;
  db.SC_COARSE_RED_FACTOR_SCALAR = 0.1;
  db.SC_GMRES_RESTART = 20;
  db.SC_LIN_MAXIT_SCALAR = (int)num("maxit", 200);
  db.SC_LIN_RED_FACTOR_SCALAR = 0.0;
  db.SC_MIXED_PRECISION_SCALAR = (int)num("mixed", 0);
  db.TIME_DISC = (int)num("time_disc", 2);

  // domain, refinement and FE spaces of every level
  cd3d_mesh_desc mesh{};
  mesh.order = (int)num("order", 1);
  mesh.m = (int)num("m", 1);
  mesh.k0 = (int)num("k0", 2);
  mesh.levels = (int)num("levels", 3);
  mesh.box[0] = mesh.box[1] = mesh.box[2] = 1.0;
  vec("box", mesh.box, 3);
  const int L = mesh.levels;
  std::unique_ptr<TDomain> Domain;
  std::vector<std::unique_ptr<TFESpace3D>> spaces;
  if (args.count("geo")) {
    // CD3D_ParMooN.C:137-330: Init(BNDFILE, GEOFILE), UNIFORM_STEPS
    // refinements, then one refinement and one space per level
    Domain.reset(new TDomain());
    Domain->Init(nullptr, args["geo"].c_str());
    for (int i = 0; i < mesh.k0; i++) Domain->RegRefineAll();
    for (int i = 0; i < L; i++) {
      if (i) Domain->RegRefineAll();
      TCollection *coll = Domain->GetCollection(It_Finest, 0);
      spaces.emplace_back(new TFESpace3D(coll, "c", "c", BoundCondition, mesh.order));
      printf("level %d: %d dofs\n", i, spaces.back()->GetN_DegreesOfFreedom());
    }
  } else {
    spaces = CreateHierarchy(mesh);
  }
  std::vector<TFESpace3D *> fes;
  std::vector<std::vector<double>> sol(L), rhs(L);
  std::vector<double *> solp, rhsp;
  for (int l = 0; l < L; l++) {
    fes.push_back(spaces[l].get());
    sol[l].assign(spaces[l]->GetN_DegreesOfFreedom(), 0.0);
    rhs[l].assign(spaces[l]->GetN_DegreesOfFreedom(), 0.0);
    solp.push_back(sol[l].data());
    rhsp.push_back(rhs[l].data());
  }
  const int N = spaces[L - 1]->GetN_DegreesOfFreedom(), NA = spaces[L - 1]->GetActiveBound();

  // problem coefficients (the example's BilinearCoeffs)
  cd3d_coeffs co{};
  co.eps = num("eps", 1.0);
  vec("b", co.b, 3);
  co.c = num("c", 0.0);
  co.conv_kind = (int)num("conv_kind", 0);
  co.rot_c[0] = co.rot_c[1] = 0.5;
  vec("rot", co.rot_c, 2);
  co.pulse_c[0] = 0.3;
  co.pulse_c[1] = co.pulse_c[2] = 0.5;
  vec("pulse", co.pulse_c, 3);
  co.f_kind = (int)num("f_kind", 0);
  co.f_val = num("f_val", 1.0);
  co.supg = (int)num("supg", 0);
  const double rel = num("rel", 1e-10);

  const int disctype = co.supg ? SDFEM : GALERKIN;
  if (mode == "tcd") {
    // TCD3D_ParMooN.C:390-634: initial condition, AssembleMRhs, time loop
    TSystemTCD3D sys(L, fes.data(), solp.data(), rhsp.data(), disctype, GMG);
    sys.Init(co);
    Check(cd3d_interp_pulse(spaces[0]->handle(), L - 1, &co, sol[L - 1].data()), "cd3d_interp_pulse",
          spaces[0]->handle());
    db.SC_LIN_RES_NORM_MIN_SCALAR = 1e-30;
    db.SC_LIN_RED_FACTOR_SCALAR = rel;
    if (args.count("tcd_stop") && args["tcd_stop"] == "step") sys.SetStepRelativeTolerance(rel);
    sys.AssembleMRhs(0.0);
    const int nsteps = (int)num("nsteps", 3);
    const double tau = num("tau", 1e-3);
    std::vector<double> states;
    int its = 0;
    for (int m = 0; m < nsteps; m++) {
      its += sys.Step(tau);
      states.insert(states.end(), sol[L - 1].begin(), sol[L - 1].end());
    }
    write_out(out, states, its, 0.0);
    return 0;
  }

  if (mode == "cd" || mode == "cdfct") {
    // CD3D_ParMooN.C:388-441; cdfct: Init with a CoeffFct3D (VarCoeffs)
    TSystemCD3D sys(L, fes.data(), solp.data(), rhsp.data(), disctype, GMG);
    if (mode == "cdfct") {
      g_eps = co.eps;
      for (int d = 0; d < 3; d++) g_b[d] = co.b[d];
      g_c = co.c;
      g_f = co.f_val;
      sys.Init(VarCoeffs, nullptr, nullptr, nullptr);
    } else {
      sys.Init(co);
    }
    sys.Assemble();
    // stopping rule of the fixture runs: |r| <= rel |r0|
    std::vector<double> r(N);
    TSquareMatrix *A[1] = {sys.GetMatrix(L - 1)};
    Defect_Scalar(A, nullptr, sol[L - 1].data(), rhs[L - 1].data(), r.data());
    double r0 = 0;
    for (int k = 0; k < NA; k++) r0 += r[k] * r[k];
    r0 = std::sqrt(r0);
    db.SC_LIN_RES_NORM_MIN_SCALAR = rel * r0;
    sys.Solve(sol[L - 1].data(), rhs[L - 1].data());
    write_out(out, sol[L - 1], sys.LastIterations(), r0);
    return 0;
  }
  if (mode == "mg") {
    // the wiring inside TSystemCD3D (SystemCD3D.C:52-203, 227-311)
    std::vector<std::unique_ptr<TSquareStructure3D>> st;
    std::vector<std::unique_ptr<TSquareMatrix3D>> A;
    std::vector<std::unique_ptr<TMGLevel3D>> lev;
    std::vector<std::unique_ptr<TAssembleMat3D>> am;
    TDiscreteForm3D form(co, disctype);
    double params[2] = {db.SC_SMOOTH_DAMP_FACTOR_SCALAR, db.SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR};
    TMultiGrid3D mg(1, 2, params);
    for (int l = 0; l < L; l++) {
      st.emplace_back(new TSquareStructure3D(fes[l]));
      st.back()->Sort();
      A.emplace_back(new TSquareMatrix3D(st.back().get()));
      TFESpace3D *fe[1] = {fes[l]};
      TSquareMatrix3D *sq[1] = {A.back().get()};
      double *r[1] = {rhsp[l]};
      am.emplace_back(new TAssembleMat3D(1, fe, 1, sq, 0, nullptr, 1, r, fe, &form));
      am.back()->Init();
      lev.emplace_back(new TMGLevel3D(l, sq[0], rhsp[l], solp[l], 2, nullptr));
      mg.AddLevel(lev.back().get());
    }
    for (int l = 0; l < L; l++) {
      am[l]->Reset();
      am[l]->Assemble3D();
      const int n = fes[l]->GetN_DegreesOfFreedom(), na = fes[l]->GetActiveBound();
      memcpy(solp[l] + na, rhsp[l] + na, sizeof(double) * (n - na));
    }
    TMultiGridScaIte prec(MatVect_Scalar, Defect_Scalar, nullptr, 0, N, &mg, 1);
    TFgmresIte fgmres(MatVect_Scalar, Defect_Scalar, &prec, 0, N, 1);
    std::vector<double> r(N), itsol(sol[L - 1]), itrhs(rhs[L - 1]);
    TSquareMatrix *Af[1] = {A.back().get()};
    Defect_Scalar(Af, nullptr, itsol.data(), itrhs.data(), r.data());
    double r0 = 0;
    for (int k = 0; k < NA; k++) r0 += r[k] * r[k];
    r0 = std::sqrt(r0);
    db.SC_LIN_RES_NORM_MIN_SCALAR = rel * r0;
    const int its = fgmres.Iterate(Af, nullptr, itsol.data(), itrhs.data());
    write_out(out, itsol, its, r0);
    return 0;
  }
  Fail("unknown mode " + mode);
}
