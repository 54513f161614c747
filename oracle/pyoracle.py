"""ctypes binding of the C oracle (oracle/cd3d_oracle.c) -- test infrastructure.

Used only by the tests, the smoke check and the CPU-baseline leg of the
benchmark.  The shared library is built by ``build()`` (gcc, no GPU).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_cd3d.so")
SRC = os.path.join(HERE, "cd3d_oracle.c")

# The box the reference reads data/Wuerfel.GEO into (ReadGeo.C:1753-1755).
REF_BOX = (1.0, 0.346, 0.346)
UNIT_BOX = (1.0, 1.0, 1.0)


def build():
    """Compile the oracle shared library (always recompiles)."""
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-o", LIB, SRC, "-lm"], check=True)


class Coeffs(C.Structure):
    _fields_ = [("eps", C.c_double), ("b", C.c_double * 3), ("c", C.c_double),
                ("conv_kind", C.c_int), ("rot_c", C.c_double * 2), ("pulse_c", C.c_double * 3),
                ("f_kind", C.c_int), ("f_val", C.c_double), ("supg", C.c_int), ("t", C.c_double)]


class Solver(C.Structure):
    _fields_ = [("smoother", C.c_int), ("pre", C.c_int), ("post", C.c_int), ("omega", C.c_double),
                ("coarse_smoother", C.c_int), ("coarse_maxit", C.c_int), ("coarse_red", C.c_double),
                ("gmg_damp", C.c_double), ("restart", C.c_int), ("maxit", C.c_int),
                ("res_norm_min", C.c_double), ("red_factor", C.c_double), ("div_factor", C.c_double),
                ("cycle", C.c_int), ("rel_tol", C.c_double)]


def coeffs(eps=1.0, b=(0, 0, 0), c=0.0, conv_kind=0, rot_c=(0.5, 0.5), pulse_c=(0.3, 0.5, 0.5),
           f_kind=0, f_val=1.0, supg=0, t=0.0):
    return Coeffs(eps, (C.c_double * 3)(*b), c, conv_kind, (C.c_double * 2)(*rot_c),
                  (C.c_double * 3)(*pulse_c), f_kind, f_val, supg, t)


def solver(smoother=1, pre=2, post=2, omega=0.8, coarse_smoother=3, coarse_maxit=10, coarse_red=0.1,
           gmg_damp=1.0, restart=20, maxit=200, res_norm_min=1e-30, red_factor=1e-10, div_factor=1e10, cycle=1,
           rel_tol=0.0):
    return Solver(smoother, pre, post, omega, coarse_smoother, coarse_maxit, coarse_red, gmg_damp,
                  restart, maxit, res_norm_min, red_factor, div_factor, cycle, rel_tol)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_create.restype = P
        L.orc_create.argtypes = [I, I, I, I, dp]
        L.orc_create_mapped.restype = P
        L.orc_create_mapped.argtypes = [I, I, I, I, dp]
        L.orc_destroy.argtypes = [P]
        L.orc_level_info.restype = I
        L.orc_level_info.argtypes = [P, I, ip, C.POINTER(C.c_long), ip]
        L.orc_get_cells.argtypes = [P, I, dp]
        L.orc_get_gn.argtypes = [P, I, ip]
        L.orc_get_csr.argtypes = [P, I, ip, ip, dp]
        L.orc_get_rhs.argtypes = [P, I, dp]
        L.orc_get_dof_nodes.argtypes = [P, I, ip]
        L.orc_set_entries.argtypes = [P, I, dp]
        L.orc_assemble.argtypes = [P, C.POINTER(Coeffs)]
        L.orc_assemble_level.argtypes = [P, I, C.POINTER(Coeffs), I, dp, dp]
        L.orc_vcycle.argtypes = [P, C.POINTER(Solver), dp, dp]
        L.orc_fgmres.restype = I
        L.orc_fgmres.argtypes = [P, C.POINTER(Solver), dp, dp, dp]
        L.orc_matvec.argtypes = [P, I, dp, dp]
        L.orc_defect.restype = D
        L.orc_defect.argtypes = [P, I, dp, dp, dp]
        L.orc_prolongate.argtypes = [P, I, dp, dp]
        L.orc_restrict.argtypes = [P, I, dp, dp]
        L.orc_tcd_run.restype = I
        L.orc_tcd_run.argtypes = [P, C.POINTER(Coeffs), C.POINTER(Solver), D, D, I, dp, dp]
        L.orc_interp_pulse.argtypes = [P, I, C.POINTER(Coeffs), dp]
        L.orc_mcssor_sweep.argtypes = [P, I, D, dp, dp]
        L.orc_gs_sweep.argtypes = [P, I, I, D, dp, dp]
        L.orc_interpolate.argtypes = [P, I, I, C.POINTER(Coeffs), dp]
        L.orc_quad_points.argtypes = [P, I, dp]
        L.orc_set_coeffs_qp.argtypes = [P, I, dp]
        L.orc_errors.argtypes = [P, I, dp, I, C.POINTER(Coeffs), dp]
        L.orc_sdfem_delta.restype = D
        L.orc_sdfem_delta.argtypes = [D, D, dp, I]
        L.orc_hconv_brick.restype = D
        L.orc_hconv_brick.argtypes = [dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Hierarchy:
    """Multilevel CD3D problem on the oracle side."""

    def __init__(self, order, m, k0, levels, box=REF_BOX, vertices=None):
        """vertices: ((m+1)^3, 3) macro vertex coordinates (x fastest) of a
        mapped hierarchy (general hexahedra); None: the box."""
        self.L = lib()
        self.order, self.levels = order, levels
        self.nb = (order + 1) ** 3
        if vertices is not None:
            v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1)
            assert v.size == 3 * (m + 1) ** 3
            self.h = self.L.orc_create_mapped(order, m, k0, levels, _dp(v))
        else:
            b = np.array(box, dtype=np.float64)
            self.h = self.L.orc_create(order, m, k0, levels, _dp(b))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def info(self, l):
        na, nc, nnz = C.c_int(), C.c_int(), C.c_long()
        n = self.L.orc_level_info(self.h, l, C.byref(na), C.byref(nnz), C.byref(nc))
        return dict(ndof=n, nact=na.value, nnz=nnz.value, ncell=nc.value)

    def cells(self, l):
        i = self.info(l)
        a = np.zeros((i["ncell"], 8, 3))
        self.L.orc_get_cells(self.h, l, _dp(a))
        return a

    def gn(self, l):
        i = self.info(l)
        a = np.zeros((i["ncell"], self.nb), dtype=np.int32)
        self.L.orc_get_gn(self.h, l, _ip(a))
        return a

    def csr(self, l):
        i = self.info(l)
        rp = np.zeros(i["ndof"] + 1, dtype=np.int32)
        kc = np.zeros(i["nnz"], dtype=np.int32)
        e = np.zeros(i["nnz"])
        self.L.orc_get_csr(self.h, l, _ip(rp), _ip(kc), _dp(e))
        return rp, kc, e

    def rhs(self, l):
        a = np.zeros(self.info(l)["ndof"])
        self.L.orc_get_rhs(self.h, l, _dp(a))
        return a

    def dof_nodes(self, l):
        a = np.zeros(self.info(l)["ndof"], dtype=np.int32)
        self.L.orc_get_dof_nodes(self.h, l, _ip(a))
        return a

    def set_entries(self, l, e):
        e = np.ascontiguousarray(e, dtype=np.float64)
        self.L.orc_set_entries(self.h, l, _dp(e))

    def assemble(self, co):
        self.L.orc_assemble(self.h, C.byref(co))

    def assemble_level(self, l, co, what=0):
        i = self.info(l)
        e, r = np.zeros(i["nnz"]), np.zeros(i["ndof"])
        self.L.orc_assemble_level(self.h, l, C.byref(co), what, _dp(e), _dp(r))
        return e, r

    def vcycle(self, sv, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        self.L.orc_vcycle(self.h, C.byref(sv), _dp(r), _dp(z))
        return z

    def fgmres(self, sv, rhs, x0=None):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros_like(rhs) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(sv.maxit + 1)
        its = self.L.orc_fgmres(self.h, C.byref(sv), _dp(rhs), _dp(x), _dp(hist))
        return x, its, hist[:max(its, 0)]

    def matvec(self, l, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self.L.orc_matvec(self.h, l, _dp(x), _dp(y))
        return y

    def defect(self, l, x, f):
        x = np.ascontiguousarray(x, dtype=np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64)
        d = np.zeros_like(x)
        r = self.L.orc_defect(self.h, l, _dp(x), _dp(f), _dp(d))
        return d, r

    def prolongate(self, l, coarse):
        coarse = np.ascontiguousarray(coarse, dtype=np.float64)
        fine = np.zeros(self.info(l)["ndof"])
        self.L.orc_prolongate(self.h, l, _dp(coarse), _dp(fine))
        return fine

    def restrict(self, l, fine):
        fine = np.ascontiguousarray(fine, dtype=np.float64)
        coarse = np.zeros(self.info(l - 1)["ndof"])
        self.L.orc_restrict(self.h, l, _dp(fine), _dp(coarse))
        return coarse

    def interp_pulse(self, l, co):
        u = np.zeros(self.info(l)["ndof"])
        self.L.orc_interp_pulse(self.h, l, C.byref(co), _dp(u))
        return u

    def quad_points(self, l):
        nq = (self.order + 2) ** 3
        a = np.zeros((self.info(l)["ncell"], nq, 3))
        self.L.orc_quad_points(self.h, l, _dp(a))
        return a

    def set_coeffs_qp(self, l, qp):
        if qp is None:
            self.L.orc_set_coeffs_qp(self.h, l, None)
            return
        qp = np.ascontiguousarray(qp, dtype=np.float64)
        self.L.orc_set_coeffs_qp(self.h, l, _dp(qp))

    def interpolate(self, l, fn, co=None):
        u = np.zeros(self.info(l)["ndof"])
        self.L.orc_interpolate(self.h, l, fn, C.byref(co or coeffs()), _dp(u))
        return u

    def errors(self, l, u, fn, co=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        e = np.zeros(4)
        self.L.orc_errors(self.h, l, _dp(u), fn, C.byref(co or coeffs()), _dp(e))
        return e

    def gs_sweep(self, l, kind, omega, f, x):
        """one SOR (kind 2) or SSOR (kind 3) sweep on level l (returns the new x)"""
        f = np.ascontiguousarray(f, dtype=np.float64)
        x = np.array(x, dtype=np.float64)
        self.L.orc_gs_sweep(self.h, l, kind, omega, _dp(f), _dp(x))
        return x

    def mcssor_sweep(self, l, omega, f, x):
        """one multicolour SSOR sweep on level l (returns the new x)"""
        f = np.ascontiguousarray(f, dtype=np.float64)
        x = np.array(x, dtype=np.float64)
        self.L.orc_mcssor_sweep(self.h, l, omega, _dp(f), _dp(x))
        return x

    def tcd_run(self, co, sv, t0, tau, nsteps, u0):
        u = np.array(u0, dtype=np.float64)
        states = np.zeros((nsteps, len(u)))
        its = self.L.orc_tcd_run(self.h, C.byref(co), C.byref(sv), t0, tau, nsteps, _dp(u), _dp(states))
        return u, states, its


def sdfem_delta(hconv, eps, b, order):
    b = np.array(b, dtype=np.float64)
    return lib().orc_sdfem_delta(hconv, eps, _dp(b), order)


def hconv_brick(h, b):
    h = np.array(h, dtype=np.float64)
    b = np.array(b, dtype=np.float64)
    return lib().orc_hconv_brick(_dp(h), _dp(b))
