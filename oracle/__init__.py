"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

The oracle restates the reference's full-order MLQD iteration in C
(oracle/trt_oracle.c).  It is the checker for the HIP implementation and the
CPU baseline timed beside it; only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it.

Parity pinning: every routine is pinned by tests/test_oracle_*.py against
closed forms, invariants and brute-force solves (the reference ships no golden
vectors for this path).  Parity is unpinned only where the reference depends
on unvendored code: the 1.95 < z <= 2.05 band of sigint (INTERPOLATORS),
GAUSS_COMPLETE's pivoting order (rounding level only) and the SPARSKIT Krylov
iterate path (only the converged solution is meaningful).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_trt.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(os.path.join(HERE, f)) for f in ("trt_oracle.c", "trt_oracle.h")):
        subprocess.check_call(["make", "-s", "-C", HERE, "-B" if force else "all"])
    return LIB


class _Problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nm", C.c_int), ("ng", C.c_int), ("nmat", C.c_int),
        ("dx", C.POINTER(C.c_double)), ("dy", C.POINTER(C.c_double)),
        ("ox", C.POINTER(C.c_double)), ("oy", C.POINTER(C.c_double)), ("w", C.POINTER(C.c_double)),
        ("nu", C.POINTER(C.c_double)),
        ("mat", C.POINTER(C.c_int)), ("kmult", C.POINTER(C.c_double)),
        ("c", C.c_double), ("dt", C.c_double), ("theta", C.c_double), ("comp_unit", C.c_double),
        ("cv", C.c_double), ("kap0", C.c_double), ("tini", C.c_double),
        ("bc_type", C.c_int * 4), ("bcT", C.c_double * 4),
        ("opacity_model", C.c_int),
        ("maxit_rte", C.c_int), ("maxit_mloqd", C.c_int), ("maxit_gloqd", C.c_int), ("conv_type", C.c_int),
        ("conv_ho", C.c_double), ("conv_lo", C.c_double), ("conv_gr1", C.c_double), ("conv_gr2", C.c_double),
        ("chi", C.c_double), ("line_src", C.c_double),
        ("use_line_search", C.c_int), ("use_safety_search", C.c_int), ("kapE_dT_flag", C.c_int),
        ("E_bound_low", C.c_double), ("T_bound_low", C.c_double),
        ("threads", C.c_int),
        ("grey_emission", C.c_int), ("no_grey", C.c_int),
    ]


_libs = {}


def build_mutant(mutant: int, outdir: str) -> str:
    """The oracle compiled with one deliberate mistake (ORC_MUTANT = mutant, see the
    ORC_MUTANT switches in trt_oracle.c) into outdir: the pin tests check that each
    mistake makes them fail.  Never used by anything but tests."""
    path = os.path.join(outdir, f"liboracle_mutant{mutant}.so")
    subprocess.check_call(["gcc", "-O2", "-fopenmp", "-fPIC", "-std=c99", "-shared", f"-DORC_MUTANT={int(mutant)}",
                           "-o", path, os.path.join(HERE, "trt_oracle.c"), "-lm"])
    return path


def lib(path: str | None = None):
    """The oracle library (or, with path, a mutant build of it)."""
    key = path or LIB
    if key not in _libs:
        if path is None:
            build()
        L = C.CDLL(key)
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(_Problem)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_array.restype = dp
        L.orc_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int * 4)]
        L.orc_iarray.restype = C.POINTER(C.c_int)
        L.orc_iarray.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int * 4)]
        for f in ("orc_material_update", "orc_transport_scb", "orc_collapse_intensities", "orc_fg_calc",
                  "orc_cg_calc", "orc_mgqd_in_calc", "orc_gqd_in_calc", "orc_rt_bc_update", "orc_old_mgqd_coefs",
                  "orc_meb_solve", "orc_old_grey_coefs", "orc_mloqd_fv", "orc_collapse_mg_ef", "orc_cbar_calc",
                  "orc_grey_coefs"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = None
        L.orc_egp_fv_newt.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_lin_noise.argtypes = [C.c_double]
        L.orc_set_qd_capture.argtypes = [C.c_int, C.c_void_p, C.c_int]
        L.orc_set_qd_capture.restype = None
        L.orc_set_lin_noise.restype = None
        L.orc_egp_fv_newt.restype = C.c_int
        L.orc_step.argtypes = [C.c_void_p, C.POINTER(C.c_int * 4)]
        L.orc_step.restype = C.c_int
        L.orc_step_ex.argtypes = [C.c_void_p, C.POINTER(C.c_int * 4), C.c_int]
        L.orc_step_ex.restype = C.c_int
        L.orc_time.argtypes = [C.c_void_p]
        L.orc_time.restype = C.c_double
        for f in ("orc_sigint",):
            getattr(L, f).argtypes = [C.c_double]
            getattr(L, f).restype = C.c_double
        L.orc_Bg_planck.argtypes = [C.c_double] * 4
        L.orc_Bg_planck.restype = C.c_double
        L.orc_kapB.argtypes = [C.c_double] * 3
        L.orc_kapB.restype = C.c_double
        L.orc_kap_midpoint.argtypes = [C.c_double] * 3
        L.orc_kap_midpoint.restype = C.c_double
        L.orc_convergence.argtypes = [C.c_int, C.c_double, dp, dp, dp, C.c_int, C.c_int, dp, dp]
        L.orc_convergence.restype = C.c_int
        L.orc_qd_build_csr.argtypes = [C.c_int, C.c_int, C.POINTER(dp), C.POINTER(dp),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int), dp, dp]
        L.orc_qd_build_csr.restype = C.c_int
        L.orc_qd_fv.argtypes = [C.c_int, C.c_int, C.POINTER(dp), C.POINTER(dp), dp, dp, dp]
        L.orc_qd_fv.restype = C.c_int
        _libs[key] = L
    return _libs[key]


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One oracle state (the reference driver's arrays) for a Problem."""

    def __init__(self, prob, threads: int | None = None, lib_path: str | None = None):
        L = lib(lib_path)
        self.prob = prob
        self._keep = []

        def d(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return _dptr(a)

        mat = np.ascontiguousarray(np.asfortranarray(prob.mat).ravel(order="F"), dtype=np.int32)
        self._keep.append(mat)
        s = _Problem()
        s.nx, s.ny, s.nm, s.ng, s.nmat = prob.nx, prob.ny, prob.nm, prob.ng, prob.nmat
        s.dx, s.dy, s.ox, s.oy, s.w, s.nu = d(prob.dx), d(prob.dy), d(prob.ox), d(prob.oy), d(prob.w), d(prob.nu)
        s.mat = mat.ctypes.data_as(C.POINTER(C.c_int))
        s.kmult = d(prob.kmult)
        s.c, s.dt, s.theta, s.comp_unit = prob.c, prob.dt, prob.theta, prob.comp_unit
        s.cv, s.kap0, s.tini = prob.cv, prob.kap0, prob.tini
        for k in range(4):
            s.bc_type[k] = int(prob.bc_type[k])
            s.bcT[k] = float(prob.bcT[k])
        s.opacity_model = prob.opacity_model
        s.maxit_rte, s.maxit_mloqd, s.maxit_gloqd, s.conv_type = (prob.maxit_rte, prob.maxit_mloqd,
                                                                  prob.maxit_gloqd, prob.conv_type)
        s.conv_ho, s.conv_lo, s.conv_gr1, s.conv_gr2 = prob.conv_ho, prob.conv_lo, prob.conv_gr1, prob.conv_gr2
        s.chi, s.line_src = prob.chi, prob.line_src
        s.use_line_search, s.use_safety_search, s.kapE_dT_flag = (prob.use_line_search, prob.use_safety_search,
                                                                  prob.kapE_dT_flag)
        s.E_bound_low, s.T_bound_low = prob.E_bound_low, prob.T_bound_low
        s.threads = threads if threads is not None else prob.threads
        s.grey_emission, s.no_grey = int(getattr(prob, "grey_emission", 0)), int(getattr(prob, "no_grey", 0))
        self._p = s
        self.h = L.orc_create(C.byref(s))
        self.L = L

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def array(self, name: str) -> np.ndarray:
        """A live view (Fortran-ordered, reference shape) of a state array."""
        dims = (C.c_int * 4)()
        ptr = self.L.orc_array(self.h, name.encode(), C.byref(dims))
        if ptr:
            shape = tuple(dims)
            n = int(np.prod(shape))
            flat = np.ctypeslib.as_array(ptr, shape=(n,))
        else:
            ptr = self.L.orc_iarray(self.h, name.encode(), C.byref(dims))
            if not ptr:
                raise KeyError(name)
            shape = tuple(dims)
            n = int(np.prod(shape))
            flat = np.ctypeslib.as_array(ptr, shape=(n,))
        shape = tuple(x for x in shape)
        while len(shape) > 1 and shape[-1] == 1:
            shape = shape[:-1]
        return flat.reshape(shape, order="F")

    def __getitem__(self, name):
        return self.array(name)

    def call(self, routine: str, *args):
        return getattr(self.L, "orc_" + routine)(self.h, *args)

    def step(self):
        it = (C.c_int * 4)()
        self.L.orc_step(self.h, C.byref(it))
        return tuple(it)

    def step_given_qdf(self):
        """The step of the ROM runs: QD factors and C_g from the state, no sweep."""
        it = (C.c_int * 4)()
        self.L.orc_step_ex(self.h, C.byref(it), 1)
        return tuple(it)

    @property
    def time(self):
        return self.L.orc_time(self.h)


def sigint(z):
    return lib().orc_sigint(float(z))


def Bg_planck(T, nl, nh, comp_unit=1e13):
    return lib().orc_Bg_planck(float(T), float(nl), float(nh), float(comp_unit))


def kapB(T, nl, nh):
    return lib().orc_kapB(float(T), float(nl), float(nh))


def kap_midpoint(T, nl, nh):
    return lib().orc_kap_midpoint(float(T), float(nl), float(nh))


def convergence(conv_type, eps, x, x_old, x_old2, its):
    x, x_old, x_old2 = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, x_old, x_old2))
    nrm = C.c_double()
    rho = C.c_double()
    ok = lib().orc_convergence(int(conv_type), float(eps), _dptr(x), _dptr(x_old), _dptr(x_old2), x.size,
                               int(its), C.byref(nrm), C.byref(rho))
    return bool(ok), nrm.value, rho.value


QD_COEF_NAMES = ("EB_L", "EB_B", "EB_C", "EB_R", "EB_T", "MBx_C", "MBx_R", "MBx_B", "MBx_T",
                 "MBy_C", "MBy_T", "MBy_L", "MBy_R", "EB_RHS", "MBx_RHS", "MBy_RHS")
QD_BND_NAMES = ("Cp_L", "Cp_B", "Cp_R", "Cp_T", "BC_L", "BC_B", "BC_R", "BC_T")


def _coef_ptrs(nx, ny, coef, bnd):
    keep = []
    cptr = (C.POINTER(C.c_double) * 16)()
    bptr = (C.POINTER(C.c_double) * 8)()
    for k, nm in enumerate(QD_COEF_NAMES):
        a = np.ascontiguousarray(np.asfortranarray(coef[nm]).ravel(order="F"), dtype=np.float64)
        assert a.size == nx * ny
        keep.append(a)
        cptr[k] = _dptr(a)
    for k, nm in enumerate(QD_BND_NAMES):
        a = np.ascontiguousarray(bnd[nm], dtype=np.float64)
        keep.append(a)
        bptr[k] = _dptr(a)
    return cptr, bptr, keep


def qd_build_csr(nx, ny, coef, bnd):
    """Reference-ordered CSR (0-based) of the reduced QD system, QD_SOLVERS.f08:68-402."""
    nrows = nx + ny + 3 * nx * ny
    nelm = 19 * nx * ny + nx + ny
    cptr, bptr, keep = _coef_ptrs(nx, ny, coef, bnd)
    ia = np.zeros(nrows + 1, dtype=np.int32)
    ja = np.zeros(nelm, dtype=np.int32)
    a = np.zeros(nelm)
    rhs = np.zeros(nrows)
    nnz = lib().orc_qd_build_csr(nx, ny, cptr, bptr, ia.ctypes.data_as(C.POINTER(C.c_int)),
                                 ja.ctypes.data_as(C.POINTER(C.c_int)), _dptr(a), _dptr(rhs))
    return ia, ja[:nnz].copy(), a[:nnz].copy(), rhs


def qd_fv(nx, ny, coef, bnd):
    cptr, bptr, keep = _coef_ptrs(nx, ny, coef, bnd)
    Ea = np.zeros(nx * ny)
    Ev = np.zeros((nx + 1) * ny)
    Eh = np.zeros(nx * (ny + 1))
    err = lib().orc_qd_fv(nx, ny, cptr, bptr, _dptr(Ea), _dptr(Ev), _dptr(Eh))
    if err:
        raise RuntimeError(f"orc_qd_fv failed: {err}")
    return Ea.reshape((nx, ny), order="F"), Ev.reshape((nx + 1, ny), order="F"), Eh.reshape((nx, ny + 1), order="F")


def set_lin_noise(eps: float) -> None:
    """Sensitivity studies: perturb every QD linear solution by relative noise eps (0 = off)."""
    lib().orc_set_lin_noise(float(eps))
