// parmoon_cd3d.hpp -- drop-in C++ classes for ParMooN's CD3D / TCD3D solver
// path, forwarding to the C ABI of libcd3d_hip.so (cd3d.h).
//
// The classes keep the reference's names, constructor arguments and call
// order (3DPrograms/CD3D_ParMooN.C:388-441, TCD3D_ParMooN.C:390-634), so a
// program written against TSystemCD3D / TSystemTCD3D, or one that wires
// TMultiGrid3D + TMGLevel3D + TMultiGridScaIte + TFgmresIte by hand as
// SystemCD3D.C:138-203 does, runs on the GPU unchanged apart from two
// points the device cannot take over:
//   * the mesh hierarchy comes from TDomain::Init with a GEO file of a
//     structured m^3 macro mesh with explicit vertex coordinates (general
//     hexahedra; the PRM file is not read: every outer face is Dirichlet),
//     or from a cd3d_mesh_desc (CreateHierarchy) for boxes, ParMooN's
//     parametrised Wuerfel.GEO among them; the FE spaces have the
//     reference's DOF numbering (FESpace3D.C), so GlobalNumbers, RowPtr and
//     KCol are bit-identical;
//   * the CoeffFct3D callback (Constants.h:123) is replaced by a
//     cd3d_coeffs descriptor (TDiscreteForm3D below), since a host callback
//     cannot be evaluated inside the device assembly.
// Solver parameters are read from TDatabase::ParamDB->SC_* as in the
// reference; SC_MIXED_PRECISION_SCALAR (not in the reference) selects the
// FP32 V-cycle inside the FP64 FGMRES.
//
// Errors: the reference prints a message and calls exit(4711)
// (SystemCD3D.C:176, FgmresIte.C:509).  The shim does the same by default;
// SetErrorMode(CD3D_SHIM_THROW) turns errors into std::runtime_error.
#ifndef PARMOON_CD3D_HPP
#define PARMOON_CD3D_HPP

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cd3d.h"

namespace parmoon_cd3d {

// ------------------------------------------------------------------ errors

enum ErrorMode { CD3D_SHIM_EXIT = 0, CD3D_SHIM_THROW = 1 };

inline ErrorMode &error_mode() {
  static ErrorMode m = CD3D_SHIM_EXIT;
  return m;
}
inline void SetErrorMode(ErrorMode m) { error_mode() = m; }

[[noreturn]] inline void Fail(const std::string &msg) {
  if (error_mode() == CD3D_SHIM_THROW) throw std::runtime_error(msg);
  fprintf(stderr, "%s\n", msg.c_str());
  fflush(stderr);
  exit(4711);
}

inline void Check(cd3d_status st, const char *what, const cd3d_problem *h) {
  if (st == CD3D_OK) return;
  Fail(std::string(what) + " failed (status " + std::to_string((int)st) + "): " + (h ? cd3d_last_error(h) : ""));
}

// ------------------------------------------------------------------ constants (Constants.h:188-190,302-304)

constexpr int GALERKIN = 1;
constexpr int SDFEM = 2;
constexpr int SUPG = 2;
constexpr int AMG_SOLVE = 0;
constexpr int GMG = 1;
constexpr int DIRECT = 2;

// ------------------------------------------------------------------ parameter database

// The SC_* entries of TDatabase::ParamDB that the path reads, with the
// reference defaults (Database.C:641-757).
struct TParamDB {
  int SC_SOLVER_SCALAR = 16;          // 16 FGMRES (11 fixed point: not on the device path)
  int SC_PRECONDITIONER_SCALAR = 5;   // 5 multigrid
  int SC_MG_TYPE_SCALAR = 0;          // 0 geometric
  int SC_MG_CYCLE_SCALAR = 1;         // 1 V, 2 W, 0 F
  int SC_SMOOTHER_SCALAR = 3;         // 1 Jacobi, 2 SOR, 3 SSOR (13: multicolour SSOR, an extension)
  int SC_PRE_SMOOTH_SCALAR = 2;
  int SC_POST_SMOOTH_SCALAR = 2;
  double SC_SMOOTH_DAMP_FACTOR_SCALAR = 1.0;
  double SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR = 1.0;
  int SC_COARSE_SMOOTHER_SCALAR = 3;  // 17 exact, 3 SSOR, 2 SOR, 1 Jacobi
  int SC_COARSE_MAXIT_SCALAR = 10;
  double SC_COARSE_RED_FACTOR_SCALAR = 0.1;
  double SC_GMG_DAMP_FACTOR_SCALAR = 1.0;
  int SC_GMRES_RESTART = 10;
  int SC_LIN_MAXIT_SCALAR = 10000;
  double SC_LIN_RED_FACTOR_SCALAR = 0.0;
  double SC_LIN_RES_NORM_MIN_SCALAR = 1e-10;
  double SC_DIV_FACTOR = 1e10;
  int SC_STEP_LENGTH_CONTROL_ALL_SCALAR = 0;
  int SC_STEP_LENGTH_CONTROL_FINE_SCALAR = 0;
  int SC_VERBOSE = 0;
  int SC_MIXED_PRECISION_SCALAR = 0;  // extension: 1 = FP32 V-cycle inside FP64 FGMRES
  int ANSATZ_ORDER = 1;
  int DISCTYPE = GALERKIN;
  // TCD3D (TimeDiscRout.C): TIME_DISC 1 backward Euler, 2 Crank-Nicolson
  int TIME_DISC = 2;
  double TIMESTEPLENGTH = 1e-3;
};

// TDatabase::ParamDB->SC_... as in the reference (Database.h)
struct TDatabase {
  static inline TParamDB DefaultParamDB{};
  static inline TParamDB *ParamDB = &DefaultParamDB;
};

inline cd3d_solver_params SolverParamsFromDB() {
  const TParamDB &db = *TDatabase::ParamDB;
  cd3d_solver_params sp;
  cd3d_default_solver(&sp);
  if (db.SC_SOLVER_SCALAR != 16) Fail("Unknown solver !!! (the device path runs FGMRES, SC_SOLVER_SCALAR 16)");
  if (db.SC_PRECONDITIONER_SCALAR != 5)
    Fail("Unknown preconditioner !!! (the device path runs multigrid, SC_PRECONDITIONER_SCALAR 5)");
  if (db.SC_STEP_LENGTH_CONTROL_ALL_SCALAR || db.SC_STEP_LENGTH_CONTROL_FINE_SCALAR)
    Fail("step length control is not available on the device path");
  if (db.SC_SMOOTHER_SCALAR != 1 && db.SC_SMOOTHER_SCALAR != 2 && db.SC_SMOOTHER_SCALAR != 3 &&
      db.SC_SMOOTHER_SCALAR != 13)
    Fail("smoother not implemented !!! (the device path has SC_SMOOTHER_SCALAR 1 Jacobi, 2 SOR, 3 SSOR, 13)");
  sp.smoother = db.SC_SMOOTHER_SCALAR;
  sp.pre = db.SC_PRE_SMOOTH_SCALAR;
  sp.post = db.SC_POST_SMOOTH_SCALAR;
  sp.omega = db.SC_SMOOTH_DAMP_FACTOR_SCALAR;
  sp.coarse_smoother = db.SC_COARSE_SMOOTHER_SCALAR;
  sp.coarse_maxit = db.SC_COARSE_MAXIT_SCALAR;
  sp.coarse_red = db.SC_COARSE_RED_FACTOR_SCALAR;
  sp.gmg_damp = db.SC_GMG_DAMP_FACTOR_SCALAR;
  sp.restart = db.SC_GMRES_RESTART;
  sp.maxit = db.SC_LIN_MAXIT_SCALAR;
  sp.res_norm_min = db.SC_LIN_RES_NORM_MIN_SCALAR;
  sp.red_factor = db.SC_LIN_RED_FACTOR_SCALAR;
  sp.div_factor = db.SC_DIV_FACTOR;
  sp.rel_tol = 0.0;  // the reference's stopping rule (FgmresIte.C:71-85)
  sp.precision = db.SC_MIXED_PRECISION_SCALAR ? CD3D_MIXED : CD3D_FP64;
  sp.cycle = db.SC_MG_CYCLE_SCALAR;
  return sp;
}

// ------------------------------------------------------------------ mesh hierarchy and FE spaces

// One multilevel problem on the device; shared by the spaces, structures and
// matrices of all its levels.
class THierarchy {
 public:
  explicit THierarchy(const cd3d_mesh_desc &m) : mesh_(m) {
    Check(cd3d_create(&m, &h_), "cd3d_create", nullptr);
  }
  // general hexahedra: the (m+1)^3 macro vertices of a GEO file
  THierarchy(const cd3d_mesh_desc &m, const std::vector<double> &macro_xyz) : mesh_(m) {
    Check(cd3d_create_mapped(&m, macro_xyz.data(), &h_), "cd3d_create_mapped", nullptr);
  }
  ~THierarchy() {
    if (h_) cd3d_destroy(h_);
  }
  THierarchy(const THierarchy &) = delete;
  THierarchy &operator=(const THierarchy &) = delete;
  cd3d_problem *handle() const { return h_; }
  const cd3d_mesh_desc &mesh() const { return mesh_; }
  int levels() const { return mesh_.levels; }
  // assembly bookkeeping for TAssembleMat3D (every level is assembled at once)
  long generation = 0;
  cd3d_coeffs last{};

 private:
  cd3d_mesh_desc mesh_;
  cd3d_problem *h_ = nullptr;
};

typedef void BoundCondFunct3D(int, double, double, double, int &);
enum Iterators { It_EQ, It_LE, It_Finest, It_EQLevel, It_LELevel, It_Between, It_OCAF, It_Mortar1, It_Mortar2 };

class TDomain;

// TCollection (Collection.C): the finest cells of a domain after a number of
// refinements (TDomain::GetCollection(It_Finest, 0))
class TCollection {
 public:
  TCollection(TDomain *d, int refinements) : domain_(d), refinements_(refinements) {}
  TDomain *GetDomain() const { return domain_; }
  int Refinements() const { return refinements_; }
  int GetN_Cells() const;

 private:
  TDomain *domain_;
  int refinements_;
};

// TDomain (Domain.C) over a GEO file: Init reads the macro mesh
// (TDomain::Init, Domain.C:818 / ReadGeo.C:46-200, through cd3d_read_geo),
// RegRefineAll (Domain.C:1037) counts refinements, GetCollection hands out
// the finest cells.  The device hierarchy of an element order spans the
// refinements of the spaces built on the domain; it is (re)built when a space
// on a finer collection appears, so a main program that creates its spaces
// level by level, as CD3D_ParMooN.C:262-330 does, sees every space on one
// hierarchy once the last one exists.
class TDomain {
 public:
  explicit TDomain(const char * /*ParamFile*/ = nullptr) {}
  TDomain(const TDomain &) = delete;
  TDomain &operator=(const TDomain &) = delete;
  void Init(const char * /*PRM*/, const char *GEO) {
    char msg[512] = {0};
    int m = 0;
    cd3d_status st = cd3d_read_geo(GEO, &m, nullptr, 0, msg, sizeof msg);
    if (st == CD3D_OK) {
      xyz_.assign(3 * (size_t)(m + 1) * (m + 1) * (m + 1), 0.0);
      st = cd3d_read_geo(GEO, &m, xyz_.data(), xyz_.size(), msg, sizeof msg);
    }
    if (st != CD3D_OK) Fail(std::string("TDomain::Init: ") + msg);
    m_ = m;
  }
  int RegRefineAll() {
    ++refinements_;
    return 0;
  }
  void GenerateEdgeInfo() {}
  TCollection *GetCollection(Iterators /*it*/, int /*level*/) {
    colls_.emplace_back(new TCollection(this, refinements_));
    return colls_.back().get();
  }
  int MacroCells() const { return m_; }
  // the hierarchy of element order `order` covering refinements
  // [coarsest registered, finest registered]
  void Register(int order, int refinements) {
    Span &s = spans_[order];
    if (s.lo < 0 || refinements < s.lo || refinements > s.hi) {
      s.lo = s.lo < 0 ? refinements : std::min(s.lo, refinements);
      s.hi = std::max(s.hi, refinements);
      s.hier.reset();
    }
  }
  const std::shared_ptr<THierarchy> &Hierarchy(int order) {
    Span &s = spans_[order];
    if (!s.hier) {
      if (m_ < 1) Fail("TDomain: Init with a GEO file first");
      cd3d_mesh_desc md{};
      md.order = order;
      md.m = m_;
      md.k0 = s.lo;
      md.levels = s.hi - s.lo + 1;
      md.box[0] = md.box[1] = md.box[2] = 1.0;
      s.hier = std::make_shared<THierarchy>(md, xyz_);
    }
    return s.hier;
  }
  int Level(int order, int refinements) { return refinements - spans_[order].lo; }

 private:
  struct Span {
    int lo = -1, hi = -1;
    std::shared_ptr<THierarchy> hier;
  };
  int m_ = 0, refinements_ = 0;
  std::vector<double> xyz_;
  std::map<int, Span> spans_;
  std::vector<std::unique_ptr<TCollection>> colls_;
};

inline int TCollection::GetN_Cells() const {
  const long n = (long)domain_->MacroCells() << refinements_;
  return (int)(n * n * n);
}

// TFESpace3D of one level (FESpace3D.C:55): conforming Q1/Q2, Dirichlet on
// every face, DOFs [inner | Dirichlet].
class TFESpace3D {
 public:
  TFESpace3D(std::shared_ptr<THierarchy> hier, int level) : hier_(std::move(hier)), level_(level) { Query(); }
  // the reference's constructor (FESpace3D.h:52-53, FESpace3D.C:55): space
  // of order k on a collection of a TDomain; name, description and the
  // boundary condition are not used (every outer face is Dirichlet)
  TFESpace3D(TCollection *coll, const char * /*name*/, const char * /*description*/,
             BoundCondFunct3D * /*BoundaryCondition*/, int k)
      : dom_(coll->GetDomain()), order_(k), refinements_(coll->Refinements()) {
    if (k != 1 && k != 2) Fail("TFESpace3D: order must be 1 or 2");
    dom_->Register(k, refinements_);
    Query();
  }
  int GetN_DegreesOfFreedom() const { return ndof_; }
  int GetActiveBound() const { return nact_; }
  int GetN_Inner() const { return nact_; }
  int GetN_Dirichlet() const { return ndof_ - nact_; }
  int GetN_Cells() const { return ncells_; }
  int GetN_BaseFunctions() const { return nbf_; }
  // GlobalNumbers[BeginIndex[c] + l] (cell c, local DOF l, lexicographic)
  int *GetGlobalNumbers() {
    if (gn_.empty()) {
      gn_.resize((size_t)ncells_ * nbf_);
      Check(cd3d_get_dofmap(handle(), GetLevel(), gn_.data()), "cd3d_get_dofmap", handle());
    }
    return gn_.data();
  }
  int *GetBeginIndex() {
    if (bi_.empty()) {
      bi_.resize(ncells_ + 1);
      for (int c = 0; c <= ncells_; c++) bi_[c] = c * nbf_;
    }
    return bi_.data();
  }
  int GetLevel() const { return dom_ ? dom_->Level(order_, refinements_) : level_; }
  const std::shared_ptr<THierarchy> &Hierarchy() const { return dom_ ? dom_->Hierarchy(order_) : hier_; }
  cd3d_problem *handle() const { return Hierarchy()->handle(); }

 private:
  void Query() {
    int64_t nnz;
    Check(cd3d_level_info(handle(), GetLevel(), &ndof_, &nact_, &nnz, &ncells_), "cd3d_level_info", handle());
    const int o = Hierarchy()->mesh().order;
    nbf_ = (o + 1) * (o + 1) * (o + 1);
  }
  std::shared_ptr<THierarchy> hier_;
  TDomain *dom_ = nullptr;
  int order_ = 0, refinements_ = 0;
  int level_ = 0, ndof_ = 0, nact_ = 0, ncells_ = 0, nbf_ = 0;
  std::vector<int> gn_, bi_;
};

// The refined box and one FE space per level (TDomain::Init + RegRefineAll
// + TFESpace3D per level, CD3D_ParMooN.C:120-262).  fespaces[l] is owned by
// the returned vector.
inline std::vector<std::unique_ptr<TFESpace3D>> CreateHierarchy(const cd3d_mesh_desc &m) {
  auto hier = std::make_shared<THierarchy>(m);
  std::vector<std::unique_ptr<TFESpace3D>> v;
  for (int l = 0; l < m.levels; l++) v.emplace_back(new TFESpace3D(hier, l));
  return v;
}

// ------------------------------------------------------------------ CSR structure and matrices

// TStructure / TSquareStructure3D(TFESpace3D*) (SquareStructure3D.C:32):
// host copies of RowPtr / KCol, already sorted by column (Sort() is a no-op).
class TSquareStructure3D {
 public:
  explicit TSquareStructure3D(TFESpace3D *space) : space_(space) {
    int64_t nnz;
    int n, na, nc;
    Check(cd3d_level_info(space->handle(), space->GetLevel(), &n, &na, &nnz, &nc), "cd3d_level_info",
          space->handle());
    rowptr_.resize(n + 1);
    kcol_.resize(nnz);
    Check(cd3d_get_csr(space->handle(), space->GetLevel(), rowptr_.data(), kcol_.data(), nullptr), "cd3d_get_csr",
          space->handle());
    active_ = na;
  }
  void Sort() {}
  int GetN_Rows() const { return (int)rowptr_.size() - 1; }
  int GetN_Columns() const { return GetN_Rows(); }
  int GetN_Entries() const { return (int)kcol_.size(); }
  int GetActiveBound() const { return active_; }
  int *GetRowPtr() { return rowptr_.data(); }
  int *GetKCol() { return kcol_.data(); }
  TFESpace3D *GetFESpace() const { return space_; }

 private:
  TFESpace3D *space_;
  std::vector<int> rowptr_, kcol_;
  int active_ = 0;
};

class TMatrix {
 public:
  virtual ~TMatrix() {}
};

// TSquareMatrix / TSquareMatrix3D (SquareMatrix3D.C): the values live on the
// device.  GetEntries() returns a host mirror read back on demand; after
// changing it on the host, SetEntriesToDevice() writes it back.
class TSquareMatrix : public TMatrix {
 public:
  explicit TSquareMatrix(TSquareStructure3D *s) : structure_(s) {}
  int GetN_Rows() const { return structure_->GetN_Rows(); }
  int GetN_Columns() const { return structure_->GetN_Columns(); }
  int GetN_Entries() const { return structure_->GetN_Entries(); }
  int GetActiveBound() const { return structure_->GetActiveBound(); }
  int *GetRowPtr() const { return structure_->GetRowPtr(); }
  int *GetKCol() const { return structure_->GetKCol(); }
  TSquareStructure3D *GetStructure() const { return structure_; }
  cd3d_problem *handle() const { return structure_->GetFESpace()->handle(); }
  int level() const { return structure_->GetFESpace()->GetLevel(); }
  double *GetEntries() {
    entries_.resize(GetN_Entries());
    Check(cd3d_get_csr(handle(), level(), nullptr, nullptr, entries_.data()), "cd3d_get_csr", handle());
    return entries_.data();
  }
  void SetEntriesToDevice() {
    if ((int)entries_.size() != GetN_Entries()) Fail("SetEntriesToDevice before GetEntries");
    Check(cd3d_set_entries(handle(), level(), entries_.data()), "cd3d_set_entries", handle());
  }
  // Matrix.C:48: zero the entries (on the device)
  void Reset() {
    entries_.assign(GetN_Entries(), 0.0);
    SetEntriesToDevice();
  }

 protected:
  TSquareStructure3D *structure_;
  std::vector<double> entries_;
};

class TSquareMatrix3D : public TSquareMatrix {
 public:
  explicit TSquareMatrix3D(TSquareStructure3D *s) : TSquareMatrix(s) {}
};

// MatVecProc / DefectProc (Constants.h:160-161) and their scalar
// instances (Blas2.C:89-99): y = A x, r = b - A x on A[0]'s level
typedef void MatVecProc(TSquareMatrix **, TMatrix **, double *, double *);
typedef void DefectProc(TSquareMatrix **, TMatrix **, double *, double *, double *);

inline void MatVect_Scalar(TSquareMatrix **A, TMatrix **, double *x, double *y) {
  Check(cd3d_matvec(A[0]->handle(), A[0]->level(), x, y), "cd3d_matvec", A[0]->handle());
}

inline void Defect_Scalar(TSquareMatrix **A, TMatrix **B, double *x, double *b, double *r) {
  MatVect_Scalar(A, B, x, r);
  const int n = A[0]->GetN_Rows();
  for (int i = 0; i < n; i++) r[i] = b[i] - r[i];
}

// ------------------------------------------------------------------ discrete forms and assembly

// The reference's callback and kernel types (Constants.h:123,136,158) and
// derivative names (MultiIndex3D).
enum MultiIndex3D { D000, D100, D010, D001, D200, D110, D101, D020, D011, D002 };
typedef void CoeffFct3D(int n_points, double *X, double *Y, double *Z, double **parameters, double **coeffs);
typedef void AssembleFct3D(double, double *, double, double **, int *, double ***, double **);
typedef void ManipulateFct3D(int, double **, double **, void *);
typedef void BoundCondFunct3D(int, double, double, double, int &);
typedef void BoundValueFunct3D(int, double, double, double, double &);

// The local kernels the device runs (ConvDiff3D.C:61,124): a TDiscreteForm3D
// built from these selects Galerkin or SDFEM; the bodies never run, the
// device's element kernels integrate the same forms
inline void BilinearAssemble(double, double *, double, double **, int *, double ***, double **) {}
inline void BilinearAssemble_SD(double, double *, double, double **, int *, double ***, double **) {}

// TDiscreteForm3D: the local form (BilinearAssemble for GALERKIN,
// BilinearAssemble_SD for SDFEM, ConvDiff3D.C:61,124; MatrixMRhsAssemble
// for the TCD3D mass, TCD3D.C:11) with its coefficients as a descriptor, or
// -- through the reference's constructor -- as any CoeffFct3D, which the
// host evaluates at the quadrature points of every cell before each assembly
// (cd3d_quad_points / cd3d_set_coeffs_qp).
class TDiscreteForm3D {
 public:
  enum Kind { STIFFNESS = 0, MASS = 1 };
  TDiscreteForm3D(const cd3d_coeffs &co, int disctype, Kind kind = STIFFNESS) : co_(co), kind_(kind) {
    if (disctype != GALERKIN && disctype != SDFEM) Fail("Unknown DISCTYPE");
    co_.supg = disctype == SDFEM ? 1 : 0;
  }
  // the reference's constructor (DiscreteForm3D.h:81-86): derivatives
  // {D100, D010, D001, D000}, one matrix and one rhs on one space
  TDiscreteForm3D(char *name, char *description, int n_terms, MultiIndex3D *derivatives, int *fespacenumber,
                  int n_matrices, int n_rhs, int *rowspace, int *columnspace, int *rhsspace,
                  AssembleFct3D *assemble, CoeffFct3D *coeffs, ManipulateFct3D *manipulate)
      : kind_(STIFFNESS), fct_(coeffs) {
    (void)name;
    (void)description;
    (void)fespacenumber;
    (void)rowspace;
    (void)columnspace;
    (void)rhsspace;
    if (assemble == &BilinearAssemble) co_.supg = 0;
    else if (assemble == &BilinearAssemble_SD) co_.supg = 1;
    else Fail("TDiscreteForm3D: the device path integrates BilinearAssemble and BilinearAssemble_SD");
    if (n_matrices != 1 || n_rhs != 1 || n_terms != 4 || manipulate != nullptr)
      Fail("TDiscreteForm3D: one matrix, one rhs, derivatives D100 D010 D001 D000 and no manipulate function");
    const MultiIndex3D want[4] = {D100, D010, D001, D000};
    for (int k = 0; k < 4; k++)
      if (derivatives[k] != want[k]) Fail("TDiscreteForm3D: derivatives must be D100 D010 D001 D000");
    if (!coeffs) Fail("TDiscreteForm3D: no coefficient function");
  }
  const cd3d_coeffs &coeffs() const { return co_; }
  Kind kind() const { return kind_; }
  CoeffFct3D *callback() const { return fct_; }
  // the callback at every quadrature point of every level of h (as
  // TFEDatabase3D::GetOrig + GetLocalForms would call it, cell by cell);
  // Gauss3 (27 points) for Q1, Gauss4 (64 points) for Q2
  void HandCoefficientsTo(cd3d_problem *h, int levels, int order) const {
    if (!fct_) return;
    const int nq = (order + 2) * (order + 2) * (order + 2);
    for (int l = 0; l < levels; l++) {
      int ndof, nact, ncells;
      int64_t nnz;
      Check(cd3d_level_info(h, l, &ndof, &nact, &nnz, &ncells), "cd3d_level_info", h);
      std::vector<double> xyz((size_t)ncells * nq * 3), qp((size_t)ncells * nq * 6);
      Check(cd3d_quad_points(h, l, xyz.data()), "cd3d_quad_points", h);
      std::vector<double> X(nq), Y(nq), Z(nq), aux((size_t)nq * 20, 0.0);
      std::vector<double *> cp(nq), pp(nq, nullptr);
      for (int q = 0; q < nq; q++) cp[q] = aux.data() + 20 * q;
      for (int c = 0; c < ncells; c++) {
        for (int q = 0; q < nq; q++) {
          X[q] = xyz[((size_t)c * nq + q) * 3];
          Y[q] = xyz[((size_t)c * nq + q) * 3 + 1];
          Z[q] = xyz[((size_t)c * nq + q) * 3 + 2];
        }
        fct_(nq, X.data(), Y.data(), Z.data(), pp.data(), cp.data());
        for (int q = 0; q < nq; q++)
          for (int k = 0; k < 6; k++) qp[((size_t)c * nq + q) * 6 + k] = cp[q][k];
      }
      Check(cd3d_set_coeffs_qp(h, l, qp.data()), "cd3d_set_coeffs_qp", h);
    }
  }

 private:
  cd3d_coeffs co_{};
  Kind kind_;
  CoeffFct3D *fct_ = nullptr;
};

// TAssembleMat3D (AssembleMat3D.h:43-65) for one square matrix and one rhs
// on one level.  The device assembles every level of the hierarchy in one
// call; the first level that asks for a new assembly triggers it and the
// others pick up their matrices and rhs from it.
class TAssembleMat3D {
 public:
  TAssembleMat3D(int n_fespaces, TFESpace3D **fespaces, int n_sqmatrices, TSquareMatrix3D **sqmatrices,
                 int n_matrices, TMatrix **, int n_rhs, double **rhs, TFESpace3D **, TDiscreteForm3D *form,
                 void * /*BoundCondFunct3D** */ = nullptr, void * /*BoundValueFunct3D** */ = nullptr,
                 void * /*TAuxParam3D* */ = nullptr)
      : space_(fespaces[0]), matrix_(sqmatrices[0]), rhs_(n_rhs > 0 ? rhs[0] : nullptr), form_(form) {
    if (n_fespaces != 1 || n_sqmatrices != 1 || n_matrices != 0 || n_rhs > 1)
      Fail("TAssembleMat3D: the device path assembles one square matrix and one rhs");
    if (form->kind() != TDiscreteForm3D::STIFFNESS) Fail("TAssembleMat3D: mass forms are assembled by TSystemTCD3D");
  }
  void Init() {}
  void Reset() {}
  void DeAllocate() {}
  void Assemble3D() {
    THierarchy &H = *space_->Hierarchy();
    const bool same = memcmp(&H.last, &form_->coeffs(), sizeof(cd3d_coeffs)) == 0 && !form_->callback();
    if (!same || seen_ == H.generation) {
      form_->HandCoefficientsTo(H.handle(), H.levels(), H.mesh().order);
      Check(cd3d_assemble(H.handle(), &form_->coeffs()), "cd3d_assemble", H.handle());
      H.generation++;
      H.last = form_->coeffs();
    }
    seen_ = H.generation;
    if (rhs_) Check(cd3d_get_rhs(H.handle(), space_->GetLevel(), rhs_), "cd3d_get_rhs", H.handle());
  }

 private:
  TFESpace3D *space_;
  TSquareMatrix3D *matrix_;
  double *rhs_;
  TDiscreteForm3D *form_;
  long seen_ = -1;
};

// ------------------------------------------------------------------ multigrid and iterative methods

// TMGLevel3D(level, A, rhs, sol, n_aux, permutation) (MGLevel3D.C:39-83)
class TMGLevel3D {
 public:
  TMGLevel3D(int level, TSquareMatrix3D *A, double *rhs, double *sol, int n_aux, int *permutation)
      : level_(level), A_(A), rhs_(rhs), sol_(sol) {
    (void)n_aux;
    (void)permutation;
  }
  int GetLevel() const { return level_; }
  TSquareMatrix3D *GetMatrix() const { return A_; }
  double *GetRhs() const { return rhs_; }
  double *GetSolution() const { return sol_; }
  int GetN_DOF() const { return A_->GetN_Rows(); }
  int GetActiveBound() const { return A_->GetActiveBound(); }

 private:
  int level_;
  TSquareMatrix3D *A_;
  double *rhs_, *sol_;
};

// TMultiGrid3D(n_problems, n_parameters, parameters) (MultiGrid3D.C:32)
class TMultiGrid3D {
 public:
  TMultiGrid3D(int n_problems, int n_parameters, double *parameters) {
    (void)n_problems;
    for (int i = 0; i < n_parameters; i++) params_.push_back(parameters[i]);
  }
  void AddLevel(TMGLevel3D *L) { levels_.push_back(L); }
  int GetN_Levels() const { return (int)levels_.size(); }
  TMGLevel3D *GetLevel(int i) const { return levels_[i]; }
  // MultiGrid3D.C:671-688: sol_D = rhs_D on level i
  void SetDirichletNodes(int i) {
    TMGLevel3D *L = levels_[i];
    for (int k = L->GetActiveBound(); k < L->GetN_DOF(); k++) L->GetSolution()[k] = L->GetRhs()[k];
  }
  // MultiGrid3D.C:690-703: the W/F recursion follows SC_MG_CYCLE_SCALAR on the device
  void SetRecursion(int) {}
  // one cycle on the finest level (MultiGrid3D.C:423-669): for the linear
  // cycle, sol_new = sol + V(rhs - A sol); res = |rhs - A sol_new|
  void Cycle(int i, double &res) {
    if (i != GetN_Levels() - 1) Fail("TMultiGrid3D::Cycle: the device cycle starts on the finest level");
    TMGLevel3D *L = levels_[i];
    TSquareMatrix *A[1] = {L->GetMatrix()};
    const int n = L->GetN_DOF(), na = L->GetActiveBound();
    cd3d_problem *h = A[0]->handle();
    const cd3d_solver_params sp = SolverParamsFromDB();
    double *x = L->GetSolution(), *f = L->GetRhs();
    bool zero = true;
    for (int k = 0; k < na && zero; k++) zero = x[k] == 0.0;
    std::vector<double> r(n), z(n);
    if (zero) {
      // zero start: the cycle takes the Dirichlet data from the rhs
      Check(cd3d_vcycle(h, &sp, f, z.data()), "cd3d_vcycle", h);
      for (int k = 0; k < n; k++) x[k] = z[k];
    } else {
      Defect_Scalar(A, nullptr, x, f, r.data());
      for (int k = na; k < n; k++) r[k] = 0.0;
      Check(cd3d_vcycle(h, &sp, r.data(), z.data()), "cd3d_vcycle", h);
      for (int k = 0; k < na; k++) x[k] += z[k];
    }
    Defect_Scalar(A, nullptr, x, f, r.data());
    double s = 0;
    for (int k = 0; k < na; k++) s += r[k] * r[k];
    res = std::sqrt(s);
  }

 private:
  std::vector<double> params_;
  std::vector<TMGLevel3D *> levels_;
};

// TItMethod::Iterate (ItMethod.h:79): returns the iteration count
class TItMethod {
 public:
  virtual ~TItMethod() {}
  virtual int Iterate(TSquareMatrix **A, TMatrix **B, double *sol, double *rhs) = 0;
};

// TMultiGridScaIte (MultiGridScaIte.C:37-104): one multigrid cycle as an
// iterative method / preconditioner; returns 0 as the reference does
class TMultiGridScaIte : public TItMethod {
 public:
  TMultiGridScaIte(MatVecProc *, DefectProc *, TItMethod *, int n_aux, int n_dof, TMultiGrid3D *mg, int zero_start)
      : mg_(mg), n_dof_(n_dof), zero_start_(zero_start) {
    (void)n_aux;
  }
  int Iterate(TSquareMatrix **, TMatrix **, double *sol, double *rhs) override {
    TMGLevel3D *L = mg_->GetLevel(mg_->GetN_Levels() - 1);
    if (zero_start_) memset(L->GetSolution(), 0, sizeof(double) * n_dof_);
    else memcpy(L->GetSolution(), sol, sizeof(double) * n_dof_);
    memcpy(L->GetRhs(), rhs, sizeof(double) * n_dof_);
    mg_->SetDirichletNodes(mg_->GetN_Levels() - 1);
    mg_->SetRecursion(mg_->GetN_Levels() - 1);
    double res;
    mg_->Cycle(mg_->GetN_Levels() - 1, res);
    memcpy(sol, L->GetSolution(), sizeof(double) * n_dof_);
    return 0;
  }
  TMultiGrid3D *GetMultiGrid() const { return mg_; }

 private:
  TMultiGrid3D *mg_;
  int n_dof_, zero_start_;
};

// TFgmresIte(MatVec, Defect, Prec, n_aux, N_DOF, scalar) (FgmresIte.C:44):
// flexible GMRES(SC_GMRES_RESTART) right-preconditioned by a multigrid
// TMultiGridScaIte; the whole iteration runs on the device.
class TFgmresIte : public TItMethod {
 public:
  TFgmresIte(MatVecProc *, DefectProc *, TItMethod *prec, int n_aux, int n_dof, int scalar)
      : prec_(dynamic_cast<TMultiGridScaIte *>(prec)), n_dof_(n_dof) {
    (void)n_aux;
    (void)scalar;
    if (!prec_) Fail("Unknown preconditioner !!! (the device FGMRES takes a TMultiGridScaIte)");
  }
  int Iterate(TSquareMatrix **A, TMatrix **, double *sol, double *rhs) override {
    cd3d_problem *h = A[0]->handle();
    if (A[0]->GetN_Rows() != n_dof_) Fail("TFgmresIte::Iterate: matrix size differs from N_DOF");
    const cd3d_solver_params sp = SolverParamsFromDB();
    cd3d_stats st{};
    const cd3d_status rc = cd3d_solve(h, &sp, rhs, sol, 1, &st);
    if (rc == CD3D_DIVERGED) Fail("FGMRES: iteration diverged (FgmresIte.C:509)");
    if (rc != CD3D_OK && rc != CD3D_NOT_CONVERGED) Check(rc, "cd3d_solve", h);
    last_ = st;
    return st.iterations;
  }
  const cd3d_stats &LastStats() const { return last_; }

 private:
  TMultiGridScaIte *prec_;
  int n_dof_;
  cd3d_stats last_{};
};

// ------------------------------------------------------------------ systems

// TSystemCD3D (SystemCD3D.C:36-396) with SOLVER == GMG
class TSystemCD3D {
 public:
  TSystemCD3D(int N_levels, TFESpace3D **fespaces, double **sol, double **rhs, int disctype, int solver)
      : n_levels_(N_levels), spaces_(fespaces, fespaces + N_levels), sol_(sol, sol + N_levels),
        rhs_(rhs, rhs + N_levels), disctype_(disctype) {
    if (solver != GMG) Fail("Unknown Solver (the device path runs GMG)");
    if (N_levels != spaces_[0]->Hierarchy()->levels())
      Fail("TSystemCD3D: N_levels must equal the number of levels of the hierarchy");
    for (int l = 0; l < N_levels; l++) {
      structures_.emplace_back(new TSquareStructure3D(spaces_[l]));
      matrices_.emplace_back(new TSquareMatrix3D(structures_.back().get()));
    }
    const TParamDB &db = *TDatabase::ParamDB;
    params_[0] = db.SC_SMOOTH_DAMP_FACTOR_SCALAR;
    params_[1] = db.SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR;
    mg_.reset(new TMultiGrid3D(1, 2, params_));
    const int n = spaces_.back()->GetN_DegreesOfFreedom();
    prec_.reset(new TMultiGridScaIte(MatVect_Scalar, Defect_Scalar, nullptr, 0, n, mg_.get(), 1));
    it_.reset(new TFgmresIte(MatVect_Scalar, Defect_Scalar, prec_.get(), 0, n, 1));
  }
  // Init(CoeffFct3D*, BoundCond, BoundValue, aux) (SystemCD3D.C:227-311):
  // the reference's signature, any coefficient function (evaluated at the
  // quadrature points on the host); g = 0 on every face
  void Init(CoeffFct3D *BilinearCoeffs, BoundCondFunct3D *BoundCond, BoundValueFunct3D *BoundValue, void *aux) {
    (void)BoundCond;
    (void)BoundValue;
    (void)aux;
    static char nm[] = "Conv-Diff", ds[] = "form";
    static MultiIndex3D der[4] = {D100, D010, D001, D000};
    static int zero[4] = {0, 0, 0, 0};
    if (disctype_ != GALERKIN && disctype_ != SDFEM) Fail("Unknown or not yet implemented DISCTYPE");
    InitForm(new TDiscreteForm3D(nm, ds, 4, der, zero, 1, 1, zero, zero, zero,
                                 disctype_ == SDFEM ? &BilinearAssemble_SD : &BilinearAssemble, BilinearCoeffs,
                                 nullptr));
  }
  // the same with the coefficients as a descriptor (no host callback)
  void Init(const cd3d_coeffs &coeffs) { InitForm(new TDiscreteForm3D(coeffs, disctype_)); }
  void InitForm(TDiscreteForm3D *form) {
    form_.reset(form);
    for (int l = 0; l < n_levels_; l++) {
      TFESpace3D *fe[1] = {spaces_[l]};
      TSquareMatrix3D *A[1] = {matrices_[l].get()};
      double *r[1] = {rhs_[l]};
      assemble_.emplace_back(new TAssembleMat3D(1, fe, 1, A, 0, nullptr, 1, r, fe, form_.get()));
      assemble_.back()->Init();
      levels_.emplace_back(new TMGLevel3D(l, A[0], rhs_[l], sol_[l], 2, nullptr));
      mg_->AddLevel(levels_.back().get());
    }
  }
  // SystemCD3D.C:314-339: assemble every level; the Dirichlet part of the
  // rhs goes into the solution
  void Assemble() {
    if (!form_) Fail("TSystemCD3D::Assemble before Init");
    for (int l = 0; l < n_levels_; l++) {
      assemble_[l]->Reset();
      assemble_[l]->Assemble3D();
      const int n = spaces_[l]->GetN_DegreesOfFreedom(), na = spaces_[l]->GetActiveBound();
      memcpy(sol_[l] + na, rhs_[l] + na, sizeof(double) * (n - na));
    }
  }
  // SystemCD3D.C:342-396, SOLVER GMG
  void Solve(double *sol, double *rhs) {
    TSquareMatrix *A[1] = {matrices_.back().get()};
    iterations_ = it_->Iterate(A, nullptr, sol, rhs);
  }
  TSquareMatrix3D *GetMatrix(int level) const { return matrices_[level].get(); }
  int LastIterations() const { return iterations_; }
  const cd3d_stats &LastStats() const { return it_->LastStats(); }

 private:
  int n_levels_;
  std::vector<TFESpace3D *> spaces_;
  std::vector<double *> sol_, rhs_;
  int disctype_;
  double params_[2] = {0, 0};
  std::vector<std::unique_ptr<TSquareStructure3D>> structures_;
  std::vector<std::unique_ptr<TSquareMatrix3D>> matrices_;
  std::unique_ptr<TDiscreteForm3D> form_;
  std::vector<std::unique_ptr<TAssembleMat3D>> assemble_;
  std::vector<std::unique_ptr<TMGLevel3D>> levels_;
  std::unique_ptr<TMultiGrid3D> mg_;
  std::unique_ptr<TMultiGridScaIte> prec_;
  std::unique_ptr<TFgmresIte> it_;
  int iterations_ = 0;
};

// TSystemTCD3D (SystemTCD3D.C:22-408): theta-scheme with the mass matrix in
// the level matrices; the state lives on the device between steps.
// DISCTYPE GALERKIN as in the reference; SDFEM/SUPG runs the SUPG forms of
// TCD3D.C:48,102 with the time-derivative matrix K beside M (the SUPG case
// that SystemTCD3D.C:42-51,114-131 leaves commented out).
class TSystemTCD3D {
 public:
  TSystemTCD3D(int N_levels, TFESpace3D **fespaces, double **sol, double **rhs, int disctype, int solver)
      : n_levels_(N_levels), spaces_(fespaces, fespaces + N_levels), sol_(sol, sol + N_levels),
        rhs_(rhs, rhs + N_levels), supg_(disctype == SDFEM || disctype == SUPG) {
    if (solver != GMG) Fail("Unknown Solver (the device path runs GMG)");
    if (disctype != GALERKIN && !supg_) Fail("Unknown or not yet implemented DISCTYPE");
    if (N_levels != spaces_[0]->Hierarchy()->levels())
      Fail("TSystemTCD3D: N_levels must equal the number of levels of the hierarchy");
  }
  void Init(const cd3d_coeffs &coeffs) { co_ = coeffs; co_.supg = supg_ ? 1 : 0; }
  // stop every step's FGMRES at rel times that step's initial residual
  // instead of the reference's restart-relative SC_LIN_RED_FACTOR_SCALAR
  // test (which, once a step needs a second restart cycle, demands the
  // reduction again from that cycle's start); 0: the reference's rule
  void SetStepRelativeTolerance(double rel) { step_rel_tol_ = rel; }
  // AssembleMRhs (SystemTCD3D.C:176-202) at the current time with the
  // initial state in sol[N_levels-1]
  void AssembleMRhs(double t0) {
    cd3d_problem *h = handle();
    Check(cd3d_tcd_init(h, &co_, t0, sol_.back()), "cd3d_tcd_init", h);
    Check(cd3d_tcd_get_state(h, sol_.back()), "cd3d_tcd_get_state", h);
  }
  // AssembleARhs + AssembleSystMat + Solve + RestoreMassMat for one step of
  // length tau (TCD3D_ParMooN.C:569-622); theta from TIME_DISC
  int Step(double tau) {
    cd3d_problem *h = handle();
    cd3d_solver_params sp = SolverParamsFromDB();
    if (step_rel_tol_ > 0) sp.rel_tol = step_rel_tol_;
    const int td = TDatabase::ParamDB->TIME_DISC;
    if (td != 1 && td != 2) Fail("TIME_DISC: the device path runs backward Euler (1) or Crank-Nicolson (2)");
    cd3d_stats st{};
    const cd3d_status rc = cd3d_tcd_step(h, &co_, &sp, tau, td == 1 ? 1.0 : 0.5, sol_.back(), &st);
    if (rc != CD3D_OK && rc != CD3D_NOT_CONVERGED) Check(rc, "cd3d_tcd_step", h);
    return st.iterations;
  }

 private:
  cd3d_problem *handle() const { return spaces_[0]->handle(); }
  int n_levels_;
  std::vector<TFESpace3D *> spaces_;
  std::vector<double *> sol_, rhs_;
  bool supg_ = false;
  double step_rel_tol_ = 0.0;
  cd3d_coeffs co_{};
};

}  // namespace parmoon_cd3d

#endif
