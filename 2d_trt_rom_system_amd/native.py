"""ctypes binding of the C ABI in include/trt_capi.h (libtrt_hip.so).

The product path: every call goes to the HIP library.  There is no CPU
fallback; if the library is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libtrt_hip.so")

_dp = C.POINTER(C.c_double)


class TrtProblem(C.Structure):
    """Mirror of trt_problem (include/trt_capi.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nm", C.c_int), ("ng", C.c_int), ("nmat", C.c_int),
        ("dx", _dp), ("dy", _dp), ("ox", _dp), ("oy", _dp), ("w", _dp), ("nu", _dp),
        ("mat", C.POINTER(C.c_int)), ("kmult", _dp),
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
        ("precision", C.c_int), ("tile_x", C.c_int), ("tile_y", C.c_int),
        ("lin_rtol", C.c_double), ("lin_maxit", C.c_int), ("sweep_mode", C.c_int), ("qd_solver", C.c_int),
        ("grey_emission", C.c_int), ("no_grey", C.c_int),
    ]


EXPORTS = [
    "trt_ctx_create", "trt_ctx_destroy", "trt_last_error", "trt_ctx_bytes", "trt_array_size", "trt_push", "trt_pull",
    "trt_material_update", "trt_transport_scb", "trt_collapse_intensities", "trt_fg_calc", "trt_cg_calc",
    "trt_mgqd_in_calc", "trt_old_mgqd_coefs", "trt_old_grey_coefs", "trt_mloqd_fv", "trt_collapse_mg_ef",
    "trt_cbar_calc", "trt_grey_coefs", "trt_gqd_in_calc", "trt_egp_fv_newt", "trt_convergence", "trt_step",
    "trt_time", "trt_sweep_timed", "trt_sweep_kernel_timed", "trt_calib_copy", "trt_sweep_stats", "trt_qd_stats",
    "trt_sweep_config", "trt_ctx_device_bytes", "trt_tgen_step", "trt_batch_create", "trt_batch_destroy",
    "trt_batch_last_error", "trt_batch_sweep_timed", "trt_batch_tgen_step", "trt_batch_use_graph", "trt_batch_info",
    "trt_batch_create_seg", "trt_batch_segment_length", "trt_autotune_batch", "trt_autotune_segments",
    "trt_server_create", "trt_server_destroy", "trt_server_last_error", "trt_server_submit", "trt_server_run",
    "trt_server_request", "trt_server_request_error", "trt_server_summary", "trt_server_output_size",
    "trt_server_output", "trt_autotune_tile", "trt_autotune_shapes", "trt_autotune_last_error",
    "trt_restart_out", "trt_restart_in", "trt_restart_time_in", "trt_frecords_write", "trt_frecords_read",
    "trt_frecords_last_error", "trt_step_given_qdf", "trt_pod_create", "trt_pod_reconstruct", "trt_pod_destroy",
    "trt_pod_timed", "trt_pod_basis_bytes", "trt_dmd_create", "trt_dmd_reconstruct", "trt_dmd_destroy",
    "trt_convergence_noalloc", "trt_meb_solve", "trt_outfile_create", "trt_timestep_outs", "trt_tgen_outs",
    "trt_outfile_put", "trt_outfile_close", "trt_outfile_last_error", "trt_nc_open", "trt_nc_dim", "trt_nc_get",
    "trt_nc_last_error", "trt_nc_close", "trt_nc_att", "trt_romgrid_create", "trt_romgrid_tmap", "trt_romgrid_map",
    "trt_romgrid_maps", "trt_romgrid_destroy", "trt_rt_bc_update", "trt_batch_step", "trt_server_set_mode",
]

_lib = None


class TrtError(RuntimeError):
    pass


class RequestStats(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("updates", C.c_longlong), ("device_bytes", C.c_size_t), ("steps", C.c_int), ("status", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("ticks", C.c_longlong), ("updates", C.c_longlong), ("sweep_ms", C.c_double),
                ("peak_bytes", C.c_size_t), ("peak_active", C.c_int), ("admit_s", C.c_double),
                ("rebuild_s", C.c_double), ("step_s", C.c_double), ("retire_s", C.c_double),
                ("setup_s", C.c_double)]


def load(path: str = LIBPATH):
    """Load libtrt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TrtError(f"{path} is missing: build it with __graft_entry__.build() "
                       f"(python 2d_trt_rom_system_amd/build.py)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.trt_ctx_create.argtypes = [C.POINTER(vp), C.POINTER(TrtProblem)]
    L.trt_ctx_destroy.argtypes = [vp]
    L.trt_last_error.argtypes = [vp]
    L.trt_last_error.restype = C.c_char_p
    L.trt_ctx_bytes.argtypes = [C.POINTER(TrtProblem)]
    L.trt_ctx_bytes.restype = C.c_size_t
    L.trt_array_size.argtypes = [vp, C.c_char_p]
    L.trt_array_size.restype = C.c_size_t
    L.trt_push.argtypes = [vp, C.c_char_p, _dp, C.c_size_t]
    L.trt_pull.argtypes = [vp, C.c_char_p, _dp, C.c_size_t]
    L.trt_meb_solve.argtypes = [vp, C.c_char_p]
    L.trt_convergence_noalloc.argtypes = [vp, vp, vp, C.c_size_t, C.c_int, C.c_double, C.c_int, vp, vp]
    for f in ("trt_material_update", "trt_transport_scb", "trt_collapse_intensities", "trt_fg_calc", "trt_cg_calc",
              "trt_mgqd_in_calc", "trt_old_mgqd_coefs", "trt_old_grey_coefs", "trt_collapse_mg_ef", "trt_cbar_calc",
              "trt_grey_coefs", "trt_gqd_in_calc", "trt_rt_bc_update"):
        getattr(L, f).argtypes = [vp]
    L.trt_mloqd_fv.argtypes = [vp, C.POINTER(C.c_int)]
    L.trt_egp_fv_newt.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.trt_convergence.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_double, C.c_int, _dp, _dp,
                                  C.POINTER(C.c_int)]
    L.trt_step.argtypes = [vp, C.POINTER(C.c_int * 4)]
    L.trt_step_given_qdf.argtypes = [vp, C.POINTER(C.c_int * 4)]
    L.trt_pod_create.argtypes = [vp, C.POINTER(vp), C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp,
                                 _dp]
    L.trt_pod_reconstruct.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
    L.trt_pod_destroy.argtypes = [vp]
    L.trt_pod_timed.argtypes = [vp, C.c_int, _dp]
    L.trt_pod_basis_bytes.argtypes = [vp]
    L.trt_pod_basis_bytes.restype = C.c_size_t
    L.trt_dmd_create.argtypes = [vp, C.POINTER(vp), C.c_char_p, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp, _dp]
    L.trt_dmd_reconstruct.argtypes = [vp, C.c_double]
    L.trt_dmd_destroy.argtypes = [vp]
    L.trt_time.argtypes = [vp]
    L.trt_time.restype = C.c_double
    L.trt_sweep_timed.argtypes = [vp, C.c_int, _dp]
    L.trt_sweep_stats.argtypes = [vp, C.POINTER(C.c_longlong), _dp]
    L.trt_qd_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    ip = C.POINTER(C.c_int)
    L.trt_sweep_config.argtypes = [vp, ip, ip, ip, ip]
    L.trt_sweep_kernel_timed.argtypes = [vp, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
    L.trt_calib_copy.argtypes = [C.c_size_t, C.c_int, _dp]
    L.trt_ctx_device_bytes.argtypes = [vp]
    L.trt_ctx_device_bytes.restype = C.c_size_t
    L.trt_tgen_step.argtypes = [vp]
    L.trt_batch_create.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_int]
    L.trt_batch_create_seg.argtypes = [C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int]
    L.trt_batch_segment_length.argtypes = [vp]
    L.trt_batch_destroy.argtypes = [vp]
    L.trt_batch_last_error.argtypes = [vp]
    L.trt_batch_last_error.restype = C.c_char_p
    L.trt_batch_sweep_timed.argtypes = [vp, C.c_int, _dp]
    L.trt_batch_tgen_step.argtypes = [vp, _dp]
    L.trt_batch_step.argtypes = [vp, C.POINTER(C.c_int)]
    L.trt_server_set_mode.argtypes = [vp, C.c_int]
    L.trt_batch_use_graph.argtypes = [vp, C.c_int]
    L.trt_batch_info.argtypes = [vp, C.POINTER(C.c_longlong), ip, ip]
    L.trt_server_create.argtypes = [C.POINTER(vp), C.c_size_t, C.c_int]
    L.trt_server_destroy.argtypes = [vp]
    L.trt_server_last_error.argtypes = [vp]
    L.trt_server_last_error.restype = C.c_char_p
    L.trt_server_submit.argtypes = [vp, C.POINTER(TrtProblem), _dp, C.c_int, C.c_double, ip]
    L.trt_server_run.argtypes = [vp]
    L.trt_server_request.argtypes = [vp, C.c_int, C.POINTER(RequestStats)]
    L.trt_server_request_error.argtypes = [vp, C.c_int]
    L.trt_server_request_error.restype = C.c_char_p
    L.trt_server_summary.argtypes = [vp, C.POINTER(ServerStats)]
    L.trt_server_output_size.argtypes = [vp, C.c_int, C.c_char_p]
    L.trt_server_output_size.restype = C.c_size_t
    L.trt_server_output.argtypes = [vp, C.c_int, C.c_char_p, _dp, C.c_size_t]
    L.trt_autotune_tile.argtypes = [C.POINTER(TrtProblem), C.c_int, C.c_int, C.c_int, ip, ip, _dp, _dp]
    L.trt_autotune_shapes.argtypes = [C.c_int, ip, ip, C.c_int]
    L.trt_autotune_batch.argtypes = [C.POINTER(TrtProblem), C.c_int, C.c_int, C.c_int, ip, ip, ip, _dp, _dp]
    L.trt_autotune_segments.argtypes = [ip, C.c_int]
    L.trt_autotune_last_error.restype = C.c_char_p
    L.trt_restart_out.argtypes = [vp, C.c_char_p]
    L.trt_restart_in.argtypes = [vp, C.c_char_p]
    L.trt_restart_time_in.argtypes = [C.c_char_p, _dp]
    L.trt_frecords_write.argtypes = [C.c_char_p, C.c_int, C.POINTER(_dp), C.POINTER(C.c_size_t), C.c_int]
    L.trt_frecords_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(_dp), C.POINTER(C.c_size_t)]
    L.trt_frecords_last_error.restype = C.c_char_p
    L.trt_outfile_create.argtypes = [vp, C.c_char_p, C.POINTER(OutOpts), C.POINTER(vp)]
    L.trt_timestep_outs.argtypes = [vp, C.c_int, C.c_int]
    L.trt_tgen_outs.argtypes = [vp, C.c_int]
    L.trt_outfile_put.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), _dp]
    L.trt_outfile_close.argtypes = [vp]
    L.trt_outfile_last_error.argtypes = [vp]
    L.trt_outfile_last_error.restype = C.c_char_p
    L.trt_nc_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.trt_nc_dim.argtypes = [vp, C.c_char_p, C.POINTER(C.c_size_t)]
    L.trt_nc_get.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), _dp]
    L.trt_nc_last_error.argtypes = [vp]
    L.trt_nc_last_error.restype = C.c_char_p
    L.trt_nc_close.argtypes = [vp]
    L.trt_nc_att.argtypes = [vp, C.c_char_p, C.c_char_p, _dp, C.c_size_t, C.c_char_p, C.c_size_t,
                             C.POINTER(C.c_size_t)]
    L.trt_romgrid_create.argtypes = [vp, C.POINTER(vp), C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                     C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double]
    L.trt_romgrid_tmap.argtypes = [vp, C.c_int, ip, ip, _dp, _dp, _dp]
    L.trt_romgrid_map.argtypes = [vp, C.c_int]
    L.trt_romgrid_maps.argtypes = [vp, C.c_int, ip, ip, _dp, _dp]
    L.trt_romgrid_destroy.argtypes = [vp]
    _lib = L
    return L


def make_problem_struct(prob, precision: int = 0, tile=(0, 0), lin_rtol: float = 0.0, lin_maxit: int = 0,
                        sweep_mode: int = 0, qd_solver: int = 0):
    """Fill a TrtProblem from a problem.Problem; returns (struct, keep-alive list)."""
    keep = []

    def d(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(_dp)

    mat = np.ascontiguousarray(np.asfortranarray(prob.mat).ravel(order="F"), dtype=np.int32)
    keep.append(mat)
    s = TrtProblem()
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
    s.maxit_rte, s.maxit_mloqd, s.maxit_gloqd, s.conv_type = (prob.maxit_rte, prob.maxit_mloqd, prob.maxit_gloqd,
                                                              prob.conv_type)
    s.conv_ho, s.conv_lo, s.conv_gr1, s.conv_gr2 = prob.conv_ho, prob.conv_lo, prob.conv_gr1, prob.conv_gr2
    s.chi, s.line_src = prob.chi, prob.line_src
    s.use_line_search, s.use_safety_search, s.kapE_dT_flag = (prob.use_line_search, prob.use_safety_search,
                                                              prob.kapE_dT_flag)
    s.E_bound_low, s.T_bound_low = prob.E_bound_low, prob.T_bound_low
    s.threads = prob.threads
    s.precision = precision
    s.tile_x, s.tile_y = tile
    s.lin_rtol, s.lin_maxit = lin_rtol, lin_maxit
    s.sweep_mode = sweep_mode
    s.qd_solver = qd_solver
    s.grey_emission = int(getattr(prob, "grey_emission", 0))
    s.no_grey = int(getattr(prob, "no_grey", 0))
    return s, keep


def calib_copy(nbytes: int, reps: int = 10) -> float:
    """Mean ms of a plain device copy kernel moving nbytes (read) + nbytes (write)."""
    L = load()
    ms = C.c_double()
    if L.trt_calib_copy(int(nbytes), int(reps), C.byref(ms)) != 0:
        raise TrtError("trt_calib_copy failed: " + L.trt_last_error(None).decode())
    return ms.value


def device_bytes(prob, precision: int = 0, tile=(0, 0), sweep_mode: int = 0) -> int:
    s, keep = make_problem_struct(prob, precision, tile, sweep_mode=sweep_mode)
    return int(load().trt_ctx_bytes(C.byref(s)))


OUT_FLAGS = ("I_out", "HO_Eg_out", "HO_Fg_out", "HO_E_out", "HO_F_out", "Eg_out", "Fg_out", "MGQD_E_out",
             "MGQD_F_out", "QDfg_out", "E_out", "F_out", "D_out", "old_parms_out", "its_out", "conv_out", "kap_out",
             "Src_out")


class OutOpts(C.Structure):
    """Mirror of trt_outopts (include/trt_capi.h)."""
    _fields_ = ([("kind", C.c_int), ("run_type", C.c_char_p), ("quadrature", C.c_char_p), ("enrgy_strc", C.c_char_p),
                 ("xlen", C.c_double), ("ylen", C.c_double), ("tlen", C.c_double), ("start_time", C.c_double),
                 ("n_t", C.c_int), ("init_out", C.c_int)]
                + [(f, C.c_int) for f in OUT_FLAGS]
                + [("nofill", C.c_int), ("format", C.c_int), ("pod_type", C.c_char_p), ("dmd_type", C.c_char_p), ("direc_diff", C.c_int),
                   ("n_dsets", C.c_int), ("dset_times", _dp), ("dsets", C.POINTER(C.c_char_p)), ("tpts", _dp)])


class Context:
    """One device-resident problem (a trt_ctx)."""

    def __init__(self, prob, precision: int = 0, tile=(0, 0), lin_rtol: float = 0.0, lin_maxit: int = 0,
                 sweep_mode: int = 0, qd_solver: int = 0):
        self.L = load()
        self.prob = prob
        self.precision = precision
        s, keep = make_problem_struct(prob, precision, tile, lin_rtol, lin_maxit, sweep_mode, qd_solver)
        h = C.c_void_p()
        rc = self.L.trt_ctx_create(C.byref(h), C.byref(s))
        if rc != 0:
            raise TrtError(f"trt_ctx_create failed ({rc}): {self.L.trt_last_error(None).decode()}")
        self.h = h
        self._bases = weakref.WeakSet()     # POD/DMD bases created on this context

    def close(self):
        if getattr(self, "h", None):
            # a basis uses the context's stream: close it first (the library also frees
            # any basis still open when the context is destroyed)
            for b in list(getattr(self, "_bases", ())):
                b.close()
            self.L.trt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise TrtError(f"{what} failed ({rc}): {self.L.trt_last_error(self.h).decode()}")

    def size(self, name: str) -> int:
        return int(self.L.trt_array_size(self.h, name.encode()))

    def pull(self, name: str, shape=None) -> np.ndarray:
        n = self.size(name)
        if n == 0:
            raise KeyError(name)
        out = np.empty(n, dtype=np.float64)
        self._check(self.L.trt_pull(self.h, name.encode(), out.ctypes.data_as(_dp), n), f"pull {name}")
        if shape is not None:
            return out.reshape(shape, order="F")
        return out

    def push(self, name: str, arr) -> None:
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel(order="F"))
        self._check(self.L.trt_push(self.h, name.encode(), a.ctypes.data_as(_dp), a.size), f"push {name}")

    def call(self, routine: str, *args):
        rc = getattr(self.L, "trt_" + routine)(self.h, *args)
        self._check(rc, routine)

    def mloqd_fv(self):
        kits = (C.c_int * self.prob.ng)()
        self._check(self.L.trt_mloqd_fv(self.h, kits), "mloqd_fv")
        return list(kits)

    def egp_fv_newt(self, mgqd_it: int) -> int:
        its = C.c_int()
        self._check(self.L.trt_egp_fv_newt(self.h, int(mgqd_it), C.byref(its)), "egp_fv_newt")
        return its.value

    def convergence(self, x, xo, xo2, conv_type, eps, its):
        nrm, rho, conv = C.c_double(), C.c_double(), C.c_int()
        self._check(self.L.trt_convergence(self.h, x.encode(), xo.encode(), xo2.encode(), int(conv_type), float(eps),
                                           int(its), C.byref(nrm), C.byref(rho), C.byref(conv)), "convergence")
        return bool(conv.value), nrm.value, rho.value

    def step(self):
        it = (C.c_int * 4)()
        self._check(self.L.trt_step(self.h, C.byref(it)), "step")
        return tuple(it)

    def step_given_qdf(self, direc_diff: int = 0):
        """The ROM runs' time step: QD factors and C_g pushed by the caller, no sweep.
        direc_diff=1 zeroes the reconstructed off-diagonal face factors first, as the
        reference does for Direc_Diff = 1 (MainSrc/ALGORITHMS.f08:593-596)."""
        if direc_diff == 1:
            for n in ("fg_edgV_xy_next", "fg_edgH_xy_next"):
                self.push(n, np.zeros(self.size(n)))
        it = (C.c_int * 4)()
        self._check(self.L.trt_step_given_qdf(self.h, C.byref(it)), "step_given_qdf")
        return tuple(it)

    @property
    def time(self) -> float:
        return float(self.L.trt_time(self.h))

    def restart_out(self, path: str) -> None:
        """RESTART_OUT (MainSrc/OUTPUTS.f08:1624-1674): the context's state as the
        reference's 25-record restart file."""
        self._check(self.L.trt_restart_out(self.h, os.fsencode(path)), "restart_out")

    def restart_in(self, path: str) -> None:
        """RESTART_IN (MainSrc/INPUTS.f08:1734-1783); the context is unchanged on error."""
        self._check(self.L.trt_restart_in(self.h, os.fsencode(path)), "restart_in")

    def sweep_timed(self, nsweeps: int) -> float:
        ms = C.c_double()
        self._check(self.L.trt_sweep_timed(self.h, int(nsweeps), C.byref(ms)), "sweep_timed")
        return ms.value

    def sweep_kernel_timed(self, nsweeps: int):
        """(wall ms, summed kernel ms, launches) over nsweeps sweeps."""
        wall, kern, nl = C.c_double(), C.c_double(), C.c_int()
        self._check(self.L.trt_sweep_kernel_timed(self.h, int(nsweeps), C.byref(wall), C.byref(kern), C.byref(nl)),
                    "sweep_kernel_timed")
        return wall.value, kern.value, nl.value

    def sweep_config(self):
        """(sweep_mode, tile_x, tile_y, persistent grid)."""
        v = [C.c_int() for _ in range(4)]
        self._check(self.L.trt_sweep_config(self.h, *[C.byref(x) for x in v]), "sweep_config")
        return tuple(x.value for x in v)

    def sweep_stats(self):
        n, ms = C.c_longlong(), C.c_double()
        self.L.trt_sweep_stats(self.h, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def qd_stats(self):
        """(factorized, reused, refinement steps) of the direct QD solves so far."""
        v = [C.c_longlong() for _ in range(3)]
        self._check(self.L.trt_qd_stats(self.h, *(C.byref(x) for x in v)), "qd_stats")
        return tuple(x.value for x in v)

    def tgen_step(self) -> None:
        """One Tgen_QDf step (Temp prescribed, one sweep, QD factors)."""
        self._check(self.L.trt_tgen_step(self.h), "tgen_step")

    @property
    def device_bytes(self) -> int:
        return int(self.L.trt_ctx_device_bytes(self.h))


class Batch:
    """Contexts swept together in one persistent launch (a trt_batch).  The
    contexts must stay open while the batch exists."""

    def __init__(self, ctxs, graph: bool = False, seg: int = 0):
        """seg: column-segment length of the work items (0 = chosen from the parallelism)."""
        self.L = load()
        self.ctxs = list(ctxs)
        arr = (C.c_void_p * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = C.c_void_p()
        rc = self.L.trt_batch_create_seg(C.byref(h), arr, len(self.ctxs), int(seg))
        if rc != 0:
            raise TrtError(f"trt_batch_create failed ({rc}): {self.L.trt_batch_last_error(None).decode()}")
        self.h = h
        if graph:
            self._check(self.L.trt_batch_use_graph(self.h, int(graph)), "batch_use_graph")

    def close(self):
        if getattr(self, "h", None):
            self.L.trt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def segment_length(self) -> int:
        return int(self.L.trt_batch_segment_length(self.h))

    def _check(self, rc, what):
        if rc != 0:
            raise TrtError(f"{what} failed ({rc}): {self.L.trt_batch_last_error(self.h).decode()}")

    def sweep_timed(self, nsweeps: int) -> float:
        ms = C.c_double()
        self._check(self.L.trt_batch_sweep_timed(self.h, int(nsweeps), C.byref(ms)), "batch_sweep_timed")
        return ms.value

    def tgen_step(self) -> float:
        ms = C.c_double()
        self._check(self.L.trt_batch_tgen_step(self.h, C.byref(ms)), "batch_tgen_step")
        return ms.value

    def step(self):
        """One full time step of every member (trt_batch_step): per member (RT, MLOQD,
        grey Newton, sweeps)."""
        its = (C.c_int * (4 * len(self.ctxs)))()
        self._check(self.L.trt_batch_step(self.h, its), "batch_step")
        return [tuple(its[4 * i:4 * i + 4]) for i in range(len(self.ctxs))]

    def info(self):
        """(updates per batched sweep, persistent grid, work items)."""
        u, g, n = C.c_longlong(), C.c_int(), C.c_int()
        self._check(self.L.trt_batch_info(self.h, C.byref(u), C.byref(g), C.byref(n)), "batch_info")
        return u.value, g.value, n.value


class PodBasis:
    """A POD basis on the device reconstructing into one context array (POD_RECONSTRUCT,
    MainSrc/POD_ROUTINES.f08:372-432).  Arrays in the reference's storage order."""

    def __init__(self, ctx: "Context", name: str, C_, S, U, V, rank, nt: int, gsum: int = 0):
        self.ctx, self.L = ctx, ctx.L
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F")) for a in (C_, S, U, V)]
        rk = np.ascontiguousarray(np.asarray(rank, dtype=np.int32).ravel())
        self._keep = arrs + [rk]
        h = C.c_void_p()
        ctx._check(self.L.trt_pod_create(ctx.h, C.byref(h), name.encode(), int(gsum), int(nt),
                                         rk.ctypes.data_as(C.POINTER(C.c_int)), *[a.ctypes.data_as(_dp) for a in arrs]),
                   "pod_create")
        self.h = h
        ctx._bases.add(self)

    def reconstruct(self, t1: int, t2: int | None = None, x: float = 0.0, x1: float = 0.0, x2: float = 0.0):
        """The field at data time t1 (1-based), or interpolated to x between t1 (at x1) and t2 (at x2)."""
        self.ctx._check(self.L.trt_pod_reconstruct(self.h, int(t1), int(t1 if t2 is None else t2), float(x), float(x1),
                                                   float(x2)), "pod_reconstruct")

    def timed(self, reps: int = 10) -> float:
        ms = C.c_double()
        self.ctx._check(self.L.trt_pod_timed(self.h, int(reps), C.byref(ms)), "pod_timed")
        return ms.value

    @property
    def basis_bytes(self) -> int:
        return int(self.L.trt_pod_basis_bytes(self.h))

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:                  # else the context's destroy has freed it
                self.L.trt_pod_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DmdBasis:
    """A DMD basis on the device reconstructing into one context array (DMD_RECONSTRUCT,
    MainSrc/DMD_ROUTINES.f08:28-103).  L, B, W complex; W in PodBasis's U layout."""

    def __init__(self, ctx: "Context", name: str, C_, L_, B, W, rank, gsum: int = 0):
        self.ctx, self.L = ctx, ctx.L
        c = np.ascontiguousarray(np.asarray(C_, dtype=np.float64).ravel())
        cx = [np.ascontiguousarray(np.asarray(a, dtype=np.complex128).ravel()).view(np.float64) for a in (L_, B, W)]
        rk = np.ascontiguousarray(np.asarray(rank, dtype=np.int32).ravel())
        self._keep = [c, rk] + cx
        h = C.c_void_p()
        ctx._check(self.L.trt_dmd_create(ctx.h, C.byref(h), name.encode(), int(gsum),
                                         rk.ctypes.data_as(C.POINTER(C.c_int)), c.ctypes.data_as(_dp),
                                         *[a.ctypes.data_as(_dp) for a in cx]), "dmd_create")
        self.h = h
        ctx._bases.add(self)

    def reconstruct(self, time: float):
        self.ctx._check(self.L.trt_dmd_reconstruct(self.h, float(time)), "dmd_reconstruct")

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.L.trt_dmd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _io_check(rc: int, what: str) -> None:
    if rc != 0:
        raise TrtError(f"{what}: {load().trt_frecords_last_error().decode()}")


def restart_time_in(path: str) -> float:
    """RESTART_TIME_IN (MainSrc/INPUTS.f08:1785-1803): the Time record of a restart file."""
    t = C.c_double()
    _io_check(load().trt_restart_time_in(os.fsencode(path), C.byref(t)), "restart_time_in")
    return t.value


RESTART_RECORDS = ("Time", "Temp", "I_crn", "fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy",
                   "fg_edgH_xy", "RT_Src", "MGQD_Src", "GQD_Src", "E_avg", "Eg_avg", "Eg_edgV", "Eg_edgH", "Fxg_edgV",
                   "Fyg_edgH", "KapE_Bar", "KapE", "KapR", "Cg_L", "Cg_B", "Cg_R", "Cg_T")


def write_records(path: str, arrays, append: bool = False) -> None:
    """Fortran unformatted sequential records of FP64 arrays (Fortran order), by the
    library's writer (host only)."""
    arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F")) for a in arrays]
    n = len(arrs)
    ptrs = (_dp * n)(*[a.ctypes.data_as(_dp) for a in arrs])
    cnt = (C.c_size_t * n)(*[a.size for a in arrs])
    _io_check(load().trt_frecords_write(os.fsencode(path), n, ptrs, cnt, int(bool(append))), "write_records")


def read_records(path: str, sizes):
    """The first len(sizes) records of a Fortran unformatted sequential file, record k
    holding exactly sizes[k] doubles, by the library's reader (host only)."""
    arrs = [np.empty(int(k), dtype=np.float64) for k in sizes]
    n = len(arrs)
    ptrs = (_dp * n)(*[a.ctypes.data_as(_dp) for a in arrs])
    cnt = (C.c_size_t * n)(*[a.size for a in arrs])
    _io_check(load().trt_frecords_read(os.fsencode(path), n, ptrs, cnt), "read_records")
    return arrs


def autotune_shapes(precision: int):
    """Tile shapes the tile engine offers for the precision."""
    L = load()
    n = L.trt_autotune_shapes(int(precision), None, None, 0)
    tx, ty = (C.c_int * n)(), (C.c_int * n)()
    L.trt_autotune_shapes(int(precision), tx, ty, n)
    return list(zip(tx, ty))


def autotune_tile(probs, precision: int = 1, sweeps: int = 3):
    """Fastest tile shape for a batched sweep of probs: (tile, best ms, {shape: ms})."""
    L = load()
    structs, keep = [], []
    for p in probs:
        s, k = make_problem_struct(p, precision)
        structs.append(s)
        keep.append(k)
    arr = (TrtProblem * len(structs))(*structs)
    shapes = autotune_shapes(precision)
    tx, ty, best = C.c_int(), C.c_int(), C.c_double()
    all_ms = (C.c_double * len(shapes))()
    rc = L.trt_autotune_tile(arr, len(structs), int(precision), int(sweeps), C.byref(tx), C.byref(ty), C.byref(best),
                             all_ms)
    if rc != 0:
        raise TrtError(f"trt_autotune_tile failed ({rc}): {L.trt_autotune_last_error().decode()}")
    return (tx.value, ty.value), best.value, dict(zip(shapes, list(all_ms)))


def autotune_segments():
    """Column-segment lengths autotune_batch tries."""
    L = load()
    n = L.trt_autotune_segments(None, 0)
    seg = (C.c_int * n)()
    L.trt_autotune_segments(seg, n)
    return list(seg)


def autotune_batch(probs, precision: int = 1, sweeps: int = 3):
    """Fastest (tile shape, segment length) for a batched sweep of probs:
    (tile, seg, best ms, {(tile, seg): ms})."""
    L = load()
    structs, keep = [], []
    for p in probs:
        s, k = make_problem_struct(p, precision)
        structs.append(s)
        keep.append(k)
    arr = (TrtProblem * len(structs))(*structs)
    shapes, segs = autotune_shapes(precision), autotune_segments()
    tx, ty, seg, best = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    all_ms = (C.c_double * (len(shapes) * len(segs)))()
    rc = L.trt_autotune_batch(arr, len(structs), int(precision), int(sweeps), C.byref(tx), C.byref(ty), C.byref(seg),
                              C.byref(best), all_ms)
    if rc != 0:
        raise TrtError(f"trt_autotune_batch failed ({rc}): {L.trt_autotune_last_error().decode()}")
    table = {(sh, sg): all_ms[i * len(segs) + q] for i, sh in enumerate(shapes) for q, sg in enumerate(segs)}
    return (tx.value, ty.value), seg.value, best.value, table


class OutFile:
    """The reference's NetCDF output file of a context (trt_outfile_*): OUTFILE_INIT +
    OUTFILE_VARDEFS (kind 0) or OUTFILE_INIT_Tgen (kind 1), MainSrc/OUTPUTS.f08.  Keyword
    arguments set trt_outopts fields by name (the OUTPUT_OPTS flags, run_type, tlen, ...)."""

    def __init__(self, ctx: "Context", path: str, **opts):
        self.L = load()
        self.ctx = ctx
        o = OutOpts()
        self._keep = []
        for k, v in opts.items():
            if k in ("run_type", "quadrature", "enrgy_strc", "pod_type", "dmd_type"):
                b = str(v).encode()
                self._keep.append(b)
                setattr(o, k, b)
            elif k in ("dset_times", "tpts"):
                a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
                self._keep.append(a)
                setattr(o, k, a.ctypes.data_as(_dp))
            elif k == "dsets":
                arr = (C.c_char_p * len(v))(*[str(x).encode() for x in v])
                self._keep.append(arr)
                o.dsets = arr
            else:
                setattr(o, k, v)
        h = C.c_void_p()
        rc = self.L.trt_outfile_create(ctx.h, os.fsencode(path), C.byref(o), C.byref(h))
        if rc != 0:
            raise TrtError(f"trt_outfile_create failed ({rc}): {self.L.trt_last_error(None).decode()}")
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            raise TrtError(f"{what} failed ({rc}): {self.L.trt_outfile_last_error(self.h).decode()}")

    def timestep_outs(self, t: int, after_step: bool = True) -> None:
        self._check(self.L.trt_timestep_outs(self.h, int(t), int(bool(after_step))), "timestep_outs")

    def tgen_outs(self, t: int) -> None:
        self._check(self.L.trt_tgen_outs(self.h, int(t)), "tgen_outs")

    def put(self, var: str, start, count, vals) -> None:
        """nf90_put_var: Fortran-order start (1-based) and count."""
        a = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).ravel(order="F"))
        n = len(start)
        st, ct = (C.c_size_t * n)(*start), (C.c_size_t * n)(*count)
        self._check(self.L.trt_outfile_put(self.h, var.encode(), n, st, ct, a.ctypes.data_as(_dp)), f"put {var}")

    def close(self) -> None:
        if getattr(self, "h", None):
            h, self.h = self.h, None
            if self.L.trt_outfile_close(h) != 0:
                raise TrtError(f"trt_outfile_close failed: {self.L.trt_last_error(None).decode()}")

    def __del__(self):
        try:
            self.close()
        except TrtError:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class NcReader:
    """A classic-format NetCDF file read by the library (trt_nc_*): NF_INQ_DIM and
    NF_INQ_VAR_nD of MainSrc/NCDF_IO.f08.  Host only."""

    def __init__(self, path: str):
        self.L = load()
        h = C.c_void_p()
        if self.L.trt_nc_open(os.fsencode(path), C.byref(h)) != 0:
            raise TrtError(f"trt_nc_open failed: {self.L.trt_last_error(None).decode()}")
        self.h = h

    def dim(self, name: str) -> int:
        n = C.c_size_t()
        if self.L.trt_nc_dim(self.h, name.encode(), C.byref(n)) != 0:
            raise TrtError(self.L.trt_nc_last_error(self.h).decode())
        return int(n.value)

    def get(self, var: str, start, count) -> np.ndarray:
        """Fortran-order start (1-based) and count; the values as a Fortran-order array."""
        nd = len(count)
        out = np.empty(int(np.prod(count)) if nd else 1, dtype=np.float64)
        st, ct = (C.c_size_t * max(nd, 1))(*start), (C.c_size_t * max(nd, 1))(*count)
        if self.L.trt_nc_get(self.h, var.encode(), nd, st, ct, out.ctypes.data_as(_dp)) != 0:
            raise TrtError(self.L.trt_nc_last_error(self.h).decode())
        return out.reshape(tuple(count), order="F") if nd else out[0]

    def att(self, name: str, var: str | None = None):
        """An attribute (nf90_get_att): text as str, numbers as an FP64 array."""
        n = C.c_size_t()
        self.L.trt_nc_att(self.h, var.encode() if var else None, name.encode(), None, 0, None, 0, C.byref(n))
        vals = np.empty(max(int(n.value), 1))
        txt = C.create_string_buffer(int(n.value) + 1)
        rc = self.L.trt_nc_att(self.h, var.encode() if var else None, name.encode(), vals.ctypes.data_as(_dp),
                               vals.size, txt, len(txt), C.byref(n))
        if rc < 0:
            raise TrtError(self.L.trt_nc_last_error(self.h).decode())
        return txt.value.decode() if rc == 1 else vals[:int(n.value)]

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.trt_nc_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


class RomGrid:
    """The ROM runs' data grid on a context (trt_romgrid_*): GENERATE_GRIDS_XYT and the
    map of data-grid fields (dat_*) onto the simulation grid's *_next arrays
    (MainSrc/GRID_FUNCTIONS.f08, POD_RECONSTRUCT_fg / POD_RECONSTRUCT_BCg)."""

    LAYOUT_POINTS = ("cells", "V faces", "H faces", "left", "bottom", "right", "top")

    def __init__(self, ctx: "Context", delx_d, dely_d, xlen_d: float, ylen_d: float, dnt: int, delt_d: float,
                 start_time_d: float, nt: int, start_time: float, xlen: float, ylen: float):
        self.L = load()
        self.ctx = ctx
        dx = np.ascontiguousarray(np.asarray(delx_d, dtype=np.float64))
        dy = np.ascontiguousarray(np.asarray(dely_d, dtype=np.float64))
        h = C.c_void_p()
        ctx._check(self.L.trt_romgrid_create(ctx.h, C.byref(h), len(dx), len(dy), dx.ctypes.data_as(_dp),
                                             dy.ctypes.data_as(_dp), float(xlen_d), float(ylen_d), int(dnt),
                                             float(delt_d), float(start_time_d), int(nt), float(start_time),
                                             float(xlen), float(ylen)), "romgrid_create")
        self.h = h
        self.dnx, self.dny = len(dx), len(dy)
        ctx._bases.add(self)

    def tmap(self, t: int):
        """(t1, t2, x, x1, x2) of simulation time t (1-based) for PodBasis.reconstruct."""
        t1, t2 = C.c_int(), C.c_int()
        x, x1, x2 = C.c_double(), C.c_double(), C.c_double()
        self.ctx._check(self.L.trt_romgrid_tmap(self.h, int(t), C.byref(t1), C.byref(t2), C.byref(x), C.byref(x1),
                                                C.byref(x2)), "romgrid_tmap")
        return t1.value, t2.value, x.value, x1.value, x2.value

    def map(self, with_bcg: bool = True) -> None:
        self.ctx._check(self.L.trt_romgrid_map(self.h, int(bool(with_bcg))), "romgrid_map")

    def maps(self, layout: int, nsim: int, ndat: int):
        gm = np.empty(4 * nsim, dtype=np.int32)
        vm = np.empty(4 * nsim, dtype=np.int32)
        sg = np.empty(2 * nsim)
        dg = np.empty(2 * ndat)
        ip_ = C.POINTER(C.c_int)
        self.L.trt_romgrid_maps(self.h, int(layout), gm.ctypes.data_as(ip_), vm.ctypes.data_as(ip_),
                                sg.ctypes.data_as(_dp), dg.ctypes.data_as(_dp))
        return gm, vm, sg, dg

    def close(self) -> None:
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.L.trt_romgrid_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()
