"""MI355X implementation of ParMooN's CD3D/TCD3D hot path (Q1/Q2 SUPG assembly
into CSR, CSR SpMV, geometric multigrid V-cycle, FGMRES).

The compute lives in ``libcd3d_hip.so`` (hand-written HIP for gfx950 behind a
C ABI, ``include/cd3d.h``).  This module is a thin ctypes mirror of that ABI
for Python callers; it never falls back to a CPU path: if the library is
missing or no GPU is present, calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CD3D_LIB") or os.path.join(HERE, "libcd3d_hip.so")

CD3D_OK, CD3D_ERR_ARG, CD3D_ERR_HIP, CD3D_ERR_NOMEM = 0, 1, 2, 3
CD3D_NOT_CONVERGED, CD3D_DIVERGED, CD3D_NONFINITE, CD3D_ERR_STATE = 4, 5, 6, 7
STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_NOMEM", 4: "NOT_CONVERGED", 5: "DIVERGED",
                6: "NONFINITE", 7: "ERR_STATE"}
FP64, MIXED = 0, 1

# every entry point declared in include/cd3d.h
EXPORTS = ["cd3d_default_solver", "cd3d_create", "cd3d_create_mapped", "cd3d_read_geo", "cd3d_solve_batch",
           "cd3d_destroy", "cd3d_last_error", "cd3d_device_bytes",
           "cd3d_level_info", "cd3d_get_dofmap", "cd3d_get_cells", "cd3d_get_csr", "cd3d_get_rhs",
           "cd3d_set_entries", "cd3d_quad_points", "cd3d_set_coeffs_qp", "cd3d_assemble", "cd3d_matvec", "cd3d_matvec_dev", "cd3d_matvec_noalloc",
           "cd3d_prolongate",
           "cd3d_restrict", "cd3d_vcycle", "cd3d_solve", "cd3d_get_solution", "cd3d_tcd_run",
           "cd3d_tcd_init", "cd3d_tcd_step", "cd3d_tcd_get_state",
           "cd3d_interp_pulse", "cd3d_interpolate", "cd3d_errors", "cd3d_write_vtk", "cd3d_write_matrix",
           "cd3d_kernel_bench", "cd3d_device_ptr", "cd3d_stream", "cd3d_synchronize",
           # cd3d_server.h
           "cd3d_server_default_params", "cd3d_estimate_bytes", "cd3d_server_create", "cd3d_server_submit",
           "cd3d_server_wait", "cd3d_server_drain", "cd3d_server_get_stats", "cd3d_server_destroy"]


class CD3DError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class MeshDesc(C.Structure):
    _fields_ = [("order", C.c_int), ("m", C.c_int), ("k0", C.c_int), ("levels", C.c_int),
                ("box", C.c_double * 3)]


class Coeffs(C.Structure):
    _fields_ = [("eps", C.c_double), ("b", C.c_double * 3), ("c", C.c_double), ("conv_kind", C.c_int),
                ("rot_c", C.c_double * 2), ("pulse_c", C.c_double * 3), ("f_kind", C.c_int),
                ("f_val", C.c_double), ("supg", C.c_int), ("t", C.c_double)]


class SolverParams(C.Structure):
    _fields_ = [("smoother", C.c_int), ("pre", C.c_int), ("post", C.c_int), ("omega", C.c_double),
                ("coarse_smoother", C.c_int), ("coarse_maxit", C.c_int), ("coarse_red", C.c_double),
                ("gmg_damp", C.c_double), ("restart", C.c_int), ("maxit", C.c_int),
                ("res_norm_min", C.c_double), ("red_factor", C.c_double), ("div_factor", C.c_double),
                ("rel_tol", C.c_double), ("precision", C.c_int), ("cycle", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("r0", C.c_double), ("rfinal", C.c_double),
                ("ms_assemble", C.c_double), ("ms_solve", C.c_double)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_inflight", C.c_int), ("workers", C.c_int),
                ("keep_resident", C.c_int), ("max_bypass", C.c_int), ("simulate", C.c_int),
                ("sim_us_per_dof", C.c_double), ("robust_after", C.c_int), ("max_batch", C.c_int)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("mesh", MeshDesc), ("coeffs", Coeffs), ("solver", SolverParams),
                ("t0", C.c_double), ("tau", C.c_double), ("theta", C.c_double), ("nsteps", C.c_int),
                ("want_solution", C.c_int)]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("iterations", C.c_int), ("rel_residual", C.c_double),
                ("queue_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double),
                ("device_bytes", C.c_size_t), ("reused_hierarchy", C.c_int), ("n_dof", C.c_int),
                ("steps_done", C.c_int), ("escalated", C.c_int), ("batch_size", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("rejected", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("solves_per_s", C.c_double), ("bytes_in_use", C.c_size_t), ("peak_bytes", C.c_size_t),
                ("budget", C.c_size_t), ("inflight", C.c_int), ("queued", C.c_int), ("resident", C.c_int)]


REQ_CD, REQ_TCD = 0, 1

_lib = None


def lib():
    """Load libcd3d_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.cd3d_default_solver.argtypes = [C.POINTER(SolverParams)]
    L.cd3d_default_solver.restype = None
    L.cd3d_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(P)]
    L.cd3d_create_mapped.argtypes = [C.POINTER(MeshDesc), dp, C.POINTER(P)]
    L.cd3d_solve_batch.argtypes = [C.POINTER(P), C.c_int, C.POINTER(SolverParams), C.POINTER(Stats)]
    L.cd3d_read_geo.argtypes = [C.c_char_p, ip, dp, C.c_size_t, C.c_char_p, C.c_size_t]
    L.cd3d_destroy.argtypes = [P]
    L.cd3d_last_error.argtypes = [P]
    L.cd3d_last_error.restype = C.c_char_p
    L.cd3d_device_bytes.argtypes = [P]
    L.cd3d_device_bytes.restype = C.c_size_t
    L.cd3d_level_info.argtypes = [P, I, ip, ip, C.POINTER(I64), ip]
    L.cd3d_get_dofmap.argtypes = [P, I, ip]
    L.cd3d_get_cells.argtypes = [P, I, dp]
    L.cd3d_get_csr.argtypes = [P, I, ip, ip, dp]
    L.cd3d_get_rhs.argtypes = [P, I, dp]
    L.cd3d_set_entries.argtypes = [P, I, dp]
    L.cd3d_assemble.argtypes = [P, C.POINTER(Coeffs)]
    L.cd3d_matvec.argtypes = [P, I, dp, dp]
    L.cd3d_matvec_noalloc.argtypes = [P, I, dp, dp]
    L.cd3d_matvec_dev.argtypes = [P, I, C.c_void_p, C.c_void_p]
    L.cd3d_prolongate.argtypes = [P, I, dp, dp]
    L.cd3d_restrict.argtypes = [P, I, dp, dp]
    L.cd3d_vcycle.argtypes = [P, C.POINTER(SolverParams), dp, dp]
    L.cd3d_solve.argtypes = [P, C.POINTER(SolverParams), dp, dp, I, C.POINTER(Stats)]
    L.cd3d_get_solution.argtypes = [P, dp]
    L.cd3d_tcd_run.argtypes = [P, C.POINTER(Coeffs), C.POINTER(SolverParams), D, D, I, dp, dp, C.POINTER(Stats)]
    L.cd3d_tcd_init.argtypes = [P, C.POINTER(Coeffs), D, dp]
    L.cd3d_tcd_step.argtypes = [P, C.POINTER(Coeffs), C.POINTER(SolverParams), D, D, dp, C.POINTER(Stats)]
    L.cd3d_tcd_get_state.argtypes = [P, dp]
    L.cd3d_interp_pulse.argtypes = [P, I, C.POINTER(Coeffs), dp]
    L.cd3d_device_ptr.argtypes = [P, I, I, C.POINTER(C.c_void_p)]
    L.cd3d_kernel_bench.argtypes = [P, I, I, I, dp, dp]
    L.cd3d_interpolate.argtypes = [P, I, I, C.POINTER(Coeffs), dp]
    L.cd3d_quad_points.argtypes = [P, I, dp]
    L.cd3d_set_coeffs_qp.argtypes = [P, I, dp]
    L.cd3d_errors.argtypes = [P, I, dp, I, C.POINTER(Coeffs), dp]
    L.cd3d_write_vtk.argtypes = [P, I, dp, C.c_char_p, C.c_char_p]
    L.cd3d_write_matrix.argtypes = [P, I, C.c_char_p]
    L.cd3d_stream.argtypes = [P]
    L.cd3d_stream.restype = C.c_void_p
    L.cd3d_synchronize.argtypes = [P]
    L.cd3d_server_default_params.argtypes = [C.POINTER(ServerParams)]
    L.cd3d_server_default_params.restype = None
    L.cd3d_estimate_bytes.argtypes = [C.POINTER(MeshDesc), I, I, I]
    L.cd3d_estimate_bytes.restype = C.c_size_t
    L.cd3d_server_create.argtypes = [C.POINTER(ServerParams), C.POINTER(P)]
    L.cd3d_server_submit.argtypes = [P, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.cd3d_server_wait.argtypes = [P, C.c_uint64, D, C.POINTER(Result), dp, I64]
    L.cd3d_server_drain.argtypes = [P]
    L.cd3d_server_get_stats.argtypes = [P, C.POINTER(ServerStats)]
    L.cd3d_server_destroy.argtypes = [P]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def coeffs(eps=1.0, b=(0, 0, 0), c=0.0, conv_kind=0, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5), f_kind=0,
           f_val=1.0, supg=0, t=0.0):
    return Coeffs(eps, (C.c_double * 3)(*b), c, conv_kind, (C.c_double * 2)(*rot_c), (C.c_double * 3)(*pulse_c),
                  f_kind, f_val, supg, t)


def solver_params(**kw):
    sp = SolverParams()
    lib().cd3d_default_solver(C.byref(sp))
    for k, v in kw.items():
        if not hasattr(sp, k):
            raise KeyError(k)
        setattr(sp, k, v)
    return sp


def solve_batch(problems, sp):
    """Solve assembled problems together (cd3d_solve_batch); returns the
    list of their Stats (each with its own status) -- the solutions are in
    each problem's solution()."""
    n = len(problems)
    hs = (C.c_void_p * n)(*[p.h for p in problems])
    st = (Stats * n)()
    rc = lib().cd3d_solve_batch(hs, n, C.byref(sp), st)
    if rc not in (CD3D_OK, 4, 5, 6):
        raise CD3DError(rc, "cd3d_solve_batch: " + problems[0].L.cd3d_last_error(problems[0].h).decode())
    return list(st)


def read_geo(path):
    """(m, vertices) of a GEO file of a structured m^3 macro mesh (cd3d_read_geo)."""
    L = lib()
    m = C.c_int()
    msg = C.create_string_buffer(512)
    st = L.cd3d_read_geo(str(path).encode(), C.byref(m), None, 0, msg, 512)
    if st != CD3D_OK:
        raise CD3DError(st, msg.value.decode())
    xyz = np.zeros(((m.value + 1) ** 3, 3))
    st = L.cd3d_read_geo(str(path).encode(), C.byref(m), _dp(xyz), xyz.size, msg, 512)
    if st != CD3D_OK:
        raise CD3DError(st, msg.value.decode())
    return m.value, xyz


class Problem:
    """One multilevel CD3D problem resident on the GPU (a cd3d_problem handle).

    Mirrors the reference's TSystemCD3D flow: construct (spaces, structures,
    multigrid levels), assemble(coeffs), solve(params)."""

    def __init__(self, order, m=1, k0=1, levels=2, box=(1.0, 1.0, 1.0), vertices=None):
        """vertices: ((m+1)^3, 3) macro vertex coordinates, x fastest (general
        hexahedra, cd3d_create_mapped); None: the box."""
        self.L = lib()
        self.order, self.levels = order, levels
        self.nb = (order + 1) ** 3
        md = MeshDesc(order, m, k0, levels, (C.c_double * 3)(*box))
        h = C.c_void_p()
        if vertices is not None:
            v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1)
            if v.size != 3 * (m + 1) ** 3:
                raise ValueError(f"vertices: {3 * (m + 1) ** 3} coordinates expected, got {v.size}")
            st = self.L.cd3d_create_mapped(C.byref(md), _dp(v), C.byref(h))
        else:
            st = self.L.cd3d_create(C.byref(md), C.byref(h))
        if st != CD3D_OK:
            raise CD3DError(st, "cd3d_create failed")
        self.h = h

    @classmethod
    def from_geo(cls, path, order, k0=0, levels=2):
        """The hierarchy of a GEO file (structured m^3 macro mesh, explicit
        vertex coordinates; TDomain::ReadGeo)."""
        m, xyz = read_geo(path)
        return cls(order, m, k0, levels, vertices=xyz)

    def close(self):
        if getattr(self, "h", None):
            self.L.cd3d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st, what):
        if st != CD3D_OK:
            raise CD3DError(st, f"{what}: {self.L.cd3d_last_error(self.h).decode()}")

    def info(self, l):
        n, na, nc, nnz = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        self._chk(self.L.cd3d_level_info(self.h, l, C.byref(n), C.byref(na), C.byref(nnz), C.byref(nc)), "info")
        return dict(ndof=n.value, nact=na.value, nnz=nnz.value, ncell=nc.value)

    def device_bytes(self):
        return self.L.cd3d_device_bytes(self.h)

    def gn(self, l):
        i = self.info(l)
        a = np.zeros((i["ncell"], self.nb), dtype=np.int32)
        self._chk(self.L.cd3d_get_dofmap(self.h, l, _ip(a)), "get_dofmap")
        return a

    def cells(self, l):
        i = self.info(l)
        a = np.zeros((i["ncell"], 8, 3))
        self._chk(self.L.cd3d_get_cells(self.h, l, _dp(a)), "get_cells")
        return a

    def csr(self, l):
        i = self.info(l)
        rp = np.zeros(i["ndof"] + 1, dtype=np.int32)
        kc = np.zeros(i["nnz"], dtype=np.int32)
        e = np.zeros(i["nnz"])
        self._chk(self.L.cd3d_get_csr(self.h, l, _ip(rp), _ip(kc), _dp(e)), "get_csr")
        return rp, kc, e

    def rhs(self, l):
        a = np.zeros(self.info(l)["ndof"])
        self._chk(self.L.cd3d_get_rhs(self.h, l, _dp(a)), "get_rhs")
        return a

    def set_entries(self, l, e):
        e = np.ascontiguousarray(e, dtype=np.float64)
        self._chk(self.L.cd3d_set_entries(self.h, l, _dp(e)), "set_entries")

    def assemble(self, co):
        self._chk(self.L.cd3d_assemble(self.h, C.byref(co)), "assemble")

    def matvec(self, l, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._chk(self.L.cd3d_matvec(self.h, l, _dp(x), _dp(y)), "matvec")
        return y

    def matvec_noalloc(self, l, x):
        """y = A x without device allocation (staged through the library's
        per-device array; at most 4 194 304 unknowns)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._chk(self.L.cd3d_matvec_noalloc(self.h, l, _dp(x), _dp(y)), "matvec_noalloc")
        return y

    def prolongate(self, l, coarse):
        coarse = np.ascontiguousarray(coarse, dtype=np.float64)
        fine = np.zeros(self.info(l)["ndof"])
        self._chk(self.L.cd3d_prolongate(self.h, l, _dp(coarse), _dp(fine)), "prolongate")
        return fine

    def restrict(self, l, fine):
        fine = np.ascontiguousarray(fine, dtype=np.float64)
        coarse = np.zeros(self.info(l - 1)["ndof"])
        self._chk(self.L.cd3d_restrict(self.h, l, _dp(fine), _dp(coarse)), "restrict")
        return coarse

    def vcycle(self, sp, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        self._chk(self.L.cd3d_vcycle(self.h, C.byref(sp), _dp(r), _dp(z)), "vcycle")
        return z

    def solve(self, sp, rhs=None, x0=None, check=True):
        """FGMRES + V-cycle; returns (solution, Stats)."""
        n = self.info(self.levels - 1)["ndof"]
        x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        st = Stats()
        r = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
        rc = self.L.cd3d_solve(self.h, C.byref(sp), None if r is None else _dp(r), _dp(x), int(x0 is not None),
                               C.byref(st))
        if check:
            self._chk(rc, "solve")
        return x, st

    def solve_device(self, sp):
        """solve with rhs and solution kept on the device (timing path)"""
        st = Stats()
        rc = self.L.cd3d_solve(self.h, C.byref(sp), None, None, 0, C.byref(st))
        self._chk(rc, "solve")
        return st

    def solution(self):
        x = np.zeros(self.info(self.levels - 1)["ndof"])
        self._chk(self.L.cd3d_get_solution(self.h, _dp(x)), "get_solution")
        return x

    def interp_pulse(self, l, co):
        u = np.zeros(self.info(l)["ndof"])
        self._chk(self.L.cd3d_interp_pulse(self.h, l, C.byref(co), _dp(u)), "interp_pulse")
        return u

    def tcd_run(self, co, sp, t0, tau, nsteps, u0):
        u = np.array(u0, dtype=np.float64)
        states = np.zeros((nsteps, len(u)))
        st = Stats()
        self._chk(self.L.cd3d_tcd_run(self.h, C.byref(co), C.byref(sp), t0, tau, nsteps, _dp(u), _dp(states),
                                      C.byref(st)), "tcd_run")
        return u, states, st

    def tcd_init(self, co, t0, u0):
        """TSystemTCD3D::AssembleMRhs at t0 and the initial state."""
        u = np.ascontiguousarray(u0, dtype=np.float64)
        self._chk(self.L.cd3d_tcd_init(self.h, C.byref(co), t0, _dp(u)), "tcd_init")

    def tcd_step(self, co, sp, tau, theta=0.5, want_state=True):
        """One theta-scheme step; returns (state or None, Stats)."""
        u = np.zeros(self.info(self.levels - 1)["ndof"]) if want_state else None
        st = Stats()
        rc = self.L.cd3d_tcd_step(self.h, C.byref(co), C.byref(sp), tau, theta, None if u is None else _dp(u),
                                  C.byref(st))
        self._chk(rc, "tcd_step")
        return u, st

    def tcd_state(self):
        u = np.zeros(self.info(self.levels - 1)["ndof"])
        self._chk(self.L.cd3d_tcd_get_state(self.h, _dp(u)), "tcd_get_state")
        return u

    KERNELS = {0: "csr_stream<kSpmv> FP64", 1: "lat_sweep_k<kLJacobi> FP32", 2: "lat64_k<kL64Spmv> FP64",
               3: "lat_restrict_k FP32", 4: "lat_prolong_k FP32", 5: "csr_fill_k", 6: "copy_k FP64"}

    def quad_points(self, l):
        """Physical quadrature points of level l: (n_cells, n_q, 3)."""
        nq = (self.order + 2) ** 3
        a = np.zeros((self.info(l)["ncell"], nq, 3))
        self._chk(self.L.cd3d_quad_points(self.h, l, _dp(a)), "quad_points")
        return a

    def set_coeffs_qp(self, l, qp):
        """Per-point coefficients (n_cells, n_q, 6: eps, b1, b2, b3, c, f) of
        level l; None returns to the descriptor."""
        if qp is None:
            self._chk(self.L.cd3d_set_coeffs_qp(self.h, l, None), "set_coeffs_qp")
            return
        qp = np.ascontiguousarray(qp, dtype=np.float64)
        self._chk(self.L.cd3d_set_coeffs_qp(self.h, l, _dp(qp)), "set_coeffs_qp")

    # known functions of interpolate / errors: zero, the sine of Laplace.h, the Gaussian pulse
    FN_ZERO, FN_SINE, FN_PULSE, FN_LINEAR = 0, 1, 2, 3

    def interpolate(self, l, fn, co=None):
        """Nodal interpolation of a known function on level l (on the device)."""
        u = np.zeros(self.info(l)["ndof"])
        self._chk(self.L.cd3d_interpolate(self.h, l, fn, None if co is None else C.byref(co), _dp(u)), "interpolate")
        return u

    def errors(self, l, fn, u=None, co=None):
        """[L2, H1 semi-norm, last cell's max, max] of u - fn (u None: the
        level's current solution on the device)."""
        e = np.zeros(4)
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        self._chk(self.L.cd3d_errors(self.h, l, None if uu is None else _dp(uu), fn,
                                     None if co is None else C.byref(co), _dp(e)), "errors")
        return e

    def write_vtk(self, l, u, path, name="u"):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._chk(self.L.cd3d_write_vtk(self.h, l, _dp(u), name.encode(), str(path).encode()), "write_vtk")

    def write_matrix(self, l, path):
        self._chk(self.L.cd3d_write_matrix(self.h, l, str(path).encode()), "write_matrix")

    def kernel_bench(self, which, level, reps=50):
        """(ms per launch, algorithmic bytes per launch) of one kernel of the path,
        timed with HIP events on the handle's stream."""
        ms, b = C.c_double(), C.c_double()
        self._chk(self.L.cd3d_kernel_bench(self.h, which, level, reps, C.byref(ms), C.byref(b)), "kernel_bench")
        return ms.value, b.value

    def device_ptr(self, l, which):
        p = C.c_void_p()
        self._chk(self.L.cd3d_device_ptr(self.h, l, which, C.byref(p)), "device_ptr")
        return p.value

    def matvec_dev(self, l, x_dev, y_dev):
        self._chk(self.L.cd3d_matvec_dev(self.h, l, C.c_void_p(x_dev), C.c_void_p(y_dev)), "matvec_dev")

    def synchronize(self):
        self._chk(self.L.cd3d_synchronize(self.h), "synchronize")


def mesh_desc(order, m=1, k0=1, levels=2, box=(1.0, 1.0, 1.0)):
    return MeshDesc(order, m, k0, levels, (C.c_double * 3)(*box))


def estimate_bytes(mesh, restart=20, cellwise=False, tcd=False):
    return lib().cd3d_estimate_bytes(C.byref(mesh), restart, int(cellwise), int(tcd))


class Server:
    """Serving loop (cd3d_server.h): requests admitted by an HBM budget and run
    concurrently, each on its own resident hierarchy and stream."""

    def __init__(self, hbm_budget=0, max_inflight=8, workers=0, keep_resident=True, max_bypass=16,
                 simulate=False, sim_us_per_dof=0.01, robust_after=None, max_batch=None):
        self.L = lib()
        prm = ServerParams()
        self.L.cd3d_server_default_params(C.byref(prm))
        prm.hbm_budget, prm.max_inflight, prm.workers = hbm_budget, max_inflight, workers
        prm.keep_resident, prm.max_bypass = int(keep_resident), max_bypass
        prm.simulate, prm.sim_us_per_dof = int(simulate), sim_us_per_dof
        if robust_after is not None:
            prm.robust_after = robust_after
        if max_batch is not None:
            prm.max_batch = max_batch
        h = C.c_void_p()
        st = self.L.cd3d_server_create(C.byref(prm), C.byref(h))
        if st != CD3D_OK:
            raise CD3DError(st, "cd3d_server_create failed")
        self.h = h

    def submit(self, mesh, co, sp, kind=REQ_CD, t0=0.0, tau=1e-3, theta=0.5, nsteps=0, want_solution=False):
        """Returns the request id (raises CD3DError when it is refused)."""
        r = Request(kind, mesh, co, sp, t0, tau, theta, nsteps, int(want_solution))
        rid = C.c_uint64()
        st = self.L.cd3d_server_submit(self.h, C.byref(r), C.byref(rid))
        if st != CD3D_OK:
            raise CD3DError(st, "cd3d_server_submit")
        return rid.value

    def wait(self, rid, timeout_ms=-1.0, n_dof=0, check=True):
        """Returns (Result, solution or None).  A request that ended in an
        error raises CD3DError unless check is False (then its Result, with
        the status, is returned); a wait that times out always raises
        (status CD3D_ERR_STATE) and leaves the request outstanding."""
        res = Result()
        sol = np.zeros(n_dof) if n_dof else None
        st = self.L.cd3d_server_wait(self.h, rid, timeout_ms, C.byref(res), None if sol is None else _dp(sol),
                                     n_dof)
        timed_out = st == CD3D_ERR_STATE and res.id == 0  # no result was written (ids start at 1)
        if timed_out or (check and st not in (CD3D_OK, 4)):
            raise CD3DError(st, f"request {rid}")
        return res, sol

    def drain(self):
        self.L.cd3d_server_drain(self.h)

    def stats(self):
        st = ServerStats()
        self.L.cd3d_server_get_stats(self.h, C.byref(st))
        return st

    def close(self):
        if getattr(self, "h", None):
            self.L.cd3d_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
