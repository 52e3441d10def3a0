"""ctypes binding of libiroiro_hip.so (the C ABI declared in include/iroiro_hip.h).

The library is the only compute path: if it is missing or fails to load,
every entry point raises; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libiroiro_hip.so")

II_OK, II_ERR_BAD_ARG, II_ERR_NOT_CONVERGED, II_ERR_NAN, II_ERR_DEVICE, II_ERR_STATE = range(6)
II_EVEN, II_ODD, II_FULL = 0, 1, 2
II_PREC_DOUBLE, II_PREC_SINGLE, II_PREC_HALF, II_PREC_BF16 = 0, 1, 2, 3

PRECISIONS = {"double": II_PREC_DOUBLE, "single": II_PREC_SINGLE, "half": II_PREC_HALF,
              "bf16": II_PREC_BF16}

# every symbol include/iroiro_hip.h declares
EXPORTS = [
    "ii_context_create", "ii_context_destroy", "ii_last_error", "ii_device_info",
    "ii_gauge_upload", "ii_gauge_invalidate", "ii_gauge_plaquette", "ii_gauge_download",
    "ii_gauge_unit", "ii_gauge_heatbath", "ii_gauge_apply_apbc",
    "ii_dirac_mult_eo", "ii_dirac_mult_oe", "ii_dirac_mult_eo_dag", "ii_dirac_mult_oe_dag",
    "ii_dirac_mult", "ii_dirac_mult_dag", "ii_dirac_gamma5", "ii_fopr_ddagd_mult",
    "ii_fopr_ddagd_mult_noalloc",
    "ii_dirac_full_mult", "ii_dirac_mult_eo_prec", "ii_solver_cg", "ii_multishift_cg", "ii_eo_invert",
    "ii_field_create", "ii_field_destroy", "ii_field_upload", "ii_field_download",
    "ii_field_point_sources", "ii_field_gaussian", "ii_propagator_solve",
    "ii_multishift_cg_full",
    "ii_multishift_mixed",
    "ii_qprop_multishift",
    "ii_net_init", "ii_net_set_weights", "ii_net_get_weights", "ii_net_layer_apply",
    "ii_net_layer_wgrad", "ii_net_loss_grad", "ii_net_apply", "ii_net_apply_batched",
    "ii_net_train", "ii_net_save",
    "ii_net_load", "ii_net_bench", "ii_bench_net_batched", "ii_net_tune", "ii_fgcr_solve",
    "ii_autotune", "ii_tune_save", "ii_tune_load", "ii_bench_hop", "ii_bench_copy",
    "ii_dirac_md_force", "ii_dirac_full_md_force", "ii_dirac_full_md_force_pm",
    "ii_traceless_antihermite", "ii_action_nf2_md_force", "ii_bench_md_force",
    "ii_rational_approx", "ii_rational_solve", "ii_action_nf_md_force",
    "ii_gauge_read", "ii_gauge_write", "ii_meson_correlators",
]

NET_LAYERS = 4
NET_WIDTHS = (1, 8, 8, 8, 1)


class IIError(RuntimeError):
    """A failure reported by the device library."""

    def __init__(self, status, msg):
        super().__init__("%s (status %d)" % (msg, status))
        self.status = status


class ConvergenceErr(IIError):
    """Errors::ConvergenceErr of the reference (lib/Tools/ErrorsHandling/
    convergence_errors.cpp); the reference exits the process, here it raises."""


class NotConverged(IIError):
    """The multishift solver's `throw "Not converged."`
    (lib/Solver/multiShiftSolver_CG.cpp:163)."""


class SolverOutputC(ctypes.Structure):
    _fields_ = [("diff", ctypes.c_double), ("iterations", ctypes.c_int),
                ("timing_ms", ctypes.c_double), ("true_residual", ctypes.c_double),
                ("restarts", ctypes.c_int), ("inner_iterations", ctypes.c_int),
                ("recursion_restarts", ctypes.c_int)]


class PropParamsC(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxiter", ctypes.c_int), ("precision", ctypes.c_int),
                ("batched", ctypes.c_int), ("delta", ctypes.c_double),
                ("use_graph", ctypes.c_int), ("use_learned", ctypes.c_int),
                ("method", ctypes.c_int), ("restart_factor", ctypes.c_double),
                ("nkrylov", ctypes.c_int), ("northo", ctypes.c_int)]


II_METHOD = {"cg": 0, "bicgstab": 1, "fgcr": 2}


class TrainParamsC(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_int), ("epochs", ctypes.c_int), ("batch", ctypes.c_int),
                ("lr", ctypes.c_double), ("pair_tol", ctypes.c_double),
                ("seed", ctypes.c_ulonglong), ("keep_weights", ctypes.c_int)]


class TrainReportC(ctypes.Structure):
    _fields_ = [("loss_first", ctypes.c_double), ("loss_last", ctypes.c_double),
                ("pairs_ms", ctypes.c_double), ("train_ms", ctypes.c_double),
                ("steps", ctypes.c_int)]


class IOInfoC(ctypes.Structure):
    _fields_ = [("plaquette", ctypes.c_double), ("link_trace", ctypes.c_double),
                ("header_plaquette", ctypes.c_double), ("header_link_trace", ctypes.c_double),
                ("checksum", ctypes.c_uint), ("header_checksum", ctypes.c_uint),
                ("bits", ctypes.c_int), ("is3x2", ctypes.c_int)]


II_IO = {"Binary": 0, "NERSC": 1, "ILDG": 2}

_lib = None


def load():
    """Loads the device library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("iroiro_amd: %s is not built; run __graft_entry__.build() "
                          "(there is no CPU fallback)" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cd, cl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    so = ctypes.POINTER(SolverOutputC)
    sig = {
        "ii_context_create": [ctypes.POINTER(vp), ip, ci],
        "ii_context_destroy": [vp],
        "ii_device_info": [vp, ctypes.c_char_p, ctypes.c_size_t, dp],
        "ii_gauge_upload": [vp, dp],
        "ii_gauge_invalidate": [vp],
        "ii_gauge_plaquette": [vp, dp],
        "ii_gauge_download": [vp, dp],
        "ii_gauge_unit": [vp],
        "ii_gauge_heatbath": [vp, cd, ctypes.c_ulonglong, ci, ci],
        "ii_gauge_apply_apbc": [vp],
        "ii_dirac_mult_eo": [vp, cd, dp, dp],
        "ii_dirac_mult_oe": [vp, cd, dp, dp],
        "ii_dirac_mult_eo_dag": [vp, cd, dp, dp],
        "ii_dirac_mult_oe_dag": [vp, cd, dp, dp],
        "ii_dirac_mult": [vp, cd, dp, dp],
        "ii_dirac_mult_dag": [vp, cd, dp, dp],
        "ii_dirac_gamma5": [vp, dp, dp, cl],
        "ii_fopr_ddagd_mult": [vp, cd, dp, dp],
        "ii_fopr_ddagd_mult_noalloc": [vp, cd, dp, dp],
        "ii_dirac_full_mult": [vp, cd, dp, dp, ci],
        "ii_dirac_mult_eo_prec": [vp, cd, ci, ci, ci, ci, dp, dp],
        "ii_dirac_md_force": [vp, cd, dp, dp, dp],
        "ii_dirac_full_md_force": [vp, cd, dp, dp, dp],
        "ii_dirac_full_md_force_pm": [vp, dp, dp, dp, ci],
        "ii_traceless_antihermite": [vp, dp, dp],
        "ii_action_nf2_md_force": [vp, cd, dp, cd, ci, dp, so],
        "ii_bench_md_force": [vp, ci, dp, dp],
        "ii_rational_approx": [vp, ci, ci, ci, ci, cd, cd, dp, dp, dp, dp, dp, dp, dp],
        "ii_rational_solve": [vp, cd, dp, dp, cd, dp, dp, ci, cd, ci, ci, so],
        "ii_action_nf_md_force": [vp, cd, dp, ci, dp, dp, ci, cd, ci, ci, dp, so],
        "ii_gauge_read": [vp, ctypes.c_char_p, ci, ci, ctypes.POINTER(IOInfoC)],
        "ii_gauge_write": [vp, ctypes.c_char_p, ci, ctypes.POINTER(IOInfoC)],
        "ii_meson_correlators": [vp, vp, vp, ci, ci, dp, dp],
        "ii_solver_cg": [vp, cd, dp, dp, cd, ci, so],
        "ii_multishift_cg": [vp, cd, dp, dp, dp, ci, cd, ci, so],
        "ii_eo_invert": [vp, cd, dp, dp, cd, ci, so],
        "ii_multishift_cg_full": [vp, cd, dp, dp, dp, ci, cd, ci, so],
        "ii_multishift_mixed": [vp, cd, dp, dp, dp, ci, cd, ci, ci, so],
        "ii_qprop_multishift": [vp, cd, ip, dp, ci, cd, ci, vp, so],
        "ii_field_create": [vp, ci, ci, ctypes.POINTER(vp)],
        "ii_field_destroy": [vp],
        "ii_field_upload": [vp, ci, dp],
        "ii_field_download": [vp, ci, dp],
        "ii_field_point_sources": [vp, ip],
        "ii_field_gaussian": [vp, ci, ctypes.c_ulonglong, ci],
        "ii_propagator_solve": [vp, cd, ctypes.POINTER(PropParamsC), vp, vp, so],
        "ii_net_init": [vp, cd, cd, ctypes.c_ulonglong],
        "ii_net_set_weights": [vp, ci, fp],
        "ii_net_get_weights": [vp, ci, fp],
        "ii_net_layer_apply": [vp, ci, ci, dp, dp],
        "ii_net_layer_wgrad": [vp, ci, dp, dp, fp],
        "ii_net_loss_grad": [vp, ci, dp, dp, dp, fp],
        "ii_net_apply": [vp, dp, dp],
        "ii_net_apply_batched": [vp, ci, dp, dp],
        "ii_net_train": [vp, cd, ctypes.POINTER(TrainParamsC), ctypes.POINTER(TrainReportC)],
        "ii_net_save": [vp, ctypes.c_char_p],
        "ii_net_load": [vp, ctypes.c_char_p],
        "ii_net_bench": [vp, ci, dp],
        "ii_bench_net_batched": [vp, ci, ci, dp],
        "ii_net_tune": [vp, ci, ip, ip, ctypes.c_char_p, ctypes.c_size_t],
        "ii_fgcr_solve": [vp, cd, dp, dp, cd, ci, ci, ci, so],
        "ii_autotune": [vp, ctypes.c_char_p, ctypes.c_size_t],
        "ii_tune_save": [vp, ctypes.c_char_p],
        "ii_tune_load": [vp, ctypes.c_char_p],
        "ii_bench_hop": [vp, ci, ci, ci, dp],
        "ii_bench_copy": [vp, ctypes.c_size_t, ci, dp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ci
    L.ii_last_error.argtypes = [vp]
    L.ii_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def as_f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d doubles, got %d" % (n, a.size))
    return a


def check(status, ctx_handle, what):
    if status == II_OK:
        return
    msg = load().ii_last_error(ctx_handle)
    msg = msg.decode() if msg else what
    if status == II_ERR_NOT_CONVERGED:
        raise ConvergenceErr(status, "%s: %s" % (what, msg))
    raise IIError(status, "%s: %s" % (what, msg))
