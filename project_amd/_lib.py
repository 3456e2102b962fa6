"""ctypes binding of libuppe.so (the C ABI declared in include/uppe.h).

The product path has exactly one backend: the HIP kernels in this library.  If the library
is missing or fails to load, every solver call raises; nothing falls back to the CPU.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libuppe.so")

c_int, c_double, c_void_p, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
P = ctypes.POINTER

UPPE_NORM_LINF, UPPE_NORM_L2 = 0, 1   # uppe3d_set_error_norm


class Uppe3DDesc(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nt", c_int), ("dt", c_double),
                ("omega0", c_double), ("a_scale", c_double), ("n2", c_double),
                ("n4", c_double), ("Ui", c_double), ("rho_nt", c_double),
                ("tau_rec", c_double), ("tau_c", c_double), ("sigma_drude", c_double),
                ("av_coef", c_double), ("use_shock", c_int), ("lut_n", c_int),
                ("lut_x0", c_double), ("lut_x1", c_double)]


class Uppe3DStatus(ctypes.Structure):
    _fields_ = [("z", c_double), ("h", c_double), ("err", c_double), ("attempts", c_int),
                ("accepted", c_int), ("rejected", c_int), ("done", c_int),
                ("nonfinite", c_int), ("fail_count", c_int), ("cur", c_int),
                ("plasma_groups", c_int)]


class UppeRf3DDesc(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("nt", c_int), ("lut_n", c_int), ("dt", c_double),
                ("i_coef", c_double), ("chi3_coef", c_double), ("chi5_coef", c_double),
                ("av_coef", c_double), ("rate_rec", c_double), ("rho_nt", c_double),
                ("Ui", c_double), ("e_charge", c_double), ("m_e", c_double), ("tau_c", c_double),
                ("coupling", c_double)]


class Uppe1DDesc(ctypes.Structure):
    _fields_ = [("nt", c_int), ("dt", c_double), ("omega0", c_double), ("gamma", c_double),
                ("use_ss", c_int), ("use_raman", c_int), ("f_R", c_double),
                ("use_plasma", c_int), ("K", c_int), ("sigma_K", c_double), ("Ui", c_double),
                ("rho_nt", c_double), ("tau_rec", c_double), ("tau_c", c_double),
                ("sigma_drude", c_double), ("area", c_double), ("avalanche", c_int),
                ("L", c_double), ("tol", c_double), ("max_step", c_double),
                ("min_step", c_double), ("max_hist", c_int), ("max_attempts", c_int)]


class Uppe1DStatus(ctypes.Structure):
    _fields_ = [("z", c_double), ("h", c_double), ("err", c_double), ("attempts", c_int),
                ("accepted", c_int), ("rejected", c_int), ("forced", c_int), ("nhist", c_int),
                ("done", c_int), ("nonfinite", c_int), ("fsal_hits", c_int),
                ("overflow", c_int)]


class Uppe1DRecord(ctypes.Structure):
    _fields_ = [("hs", c_void_p), ("hs_cap", c_int), ("ckpt", c_void_p),
                ("ckpt_every", c_int), ("ckpt_cap", c_int), ("replay_hs", c_void_p),
                ("replay_M", c_int)]


class Uppe1DSlice(ctypes.Structure):
    _fields_ = [("L", c_double), ("gamma", c_double), ("tol", c_double),
                ("max_step", c_double), ("min_step", c_double)]


# name -> (restype, argtypes); must match include/uppe.h
SIGNATURES = {
    "uppe_last_error": (ctypes.c_char_p, []),
    "uppe3d_workspace_size": (c_int, [P(Uppe3DDesc), P(c_size_t)]),
    "uppe3d_plan_create": (c_int, [P(Uppe3DDesc), c_void_p, c_size_t] + [c_void_p] * 10
                           + [c_void_p, P(c_void_p)]),
    "uppe3d_plan_destroy": (c_int, [c_void_p]),
    "uppe3d_load_field": (c_int, [c_void_p, c_void_p, c_void_p]),
    "uppe3d_begin": (c_int, [c_void_p, c_double, c_double, c_double, c_double, c_void_p]),
    "uppe3d_attempt": (c_int, [c_void_p, c_void_p]),
    "uppe3d_attempt_timed": (c_int, [c_void_p, c_int, c_void_p, P(ctypes.c_float),
                                     P(ctypes.c_float)]),
    "uppe3d_replay_attempt": (c_int, [c_void_p, c_double, c_int, c_void_p]),
    "uppe3d_run": (c_int, [c_void_p, c_int, c_int, c_void_p, P(Uppe3DStatus)]),
    "uppe3d_status_get": (c_int, [c_void_p, c_void_p, P(Uppe3DStatus)]),
    "uppe3d_store_spectrum": (c_int, [c_void_p, c_void_p, c_void_p]),
    "uppe3d_store_field": (c_int, [c_void_p, c_void_p, c_void_p]),
    "uppe3d_state_ptr": (c_int, [c_void_p, c_void_p, P(c_void_p)]),
    "uppe3d_get_state": (c_int, [c_void_p, c_void_p, c_void_p]),
    "uppe3d_set_state": (c_int, [c_void_p, c_void_p, c_void_p]),
    "uppe3d_diag_workspace_size": (c_int, [P(Uppe3DDesc), P(c_size_t)]),
    "uppe3d_diagnostics": (c_int, [c_void_p, c_double, c_double, c_double, c_double, c_double,
                                   c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "uppe3d_field_at_time": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "uppe3d_apply_mirror": (c_int, [c_void_p, c_double, c_double, c_double, c_double, c_double,
                                    c_double, c_void_p]),
    "uppe3d_set_error_norm": (c_int, [c_void_p, c_int]),
    "uppe3d_set_fused": (c_int, [c_void_p, c_int]),
    "uppe3d_attempt_path": (c_int, [c_void_p, P(c_int)]),
    "uppe3d_set_stage_record": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "uppe3d_adjoint_workspace_size": (c_int, [P(Uppe3DDesc), P(c_size_t)]),
    "uppe3d_adjoint_step": (c_int, [c_void_p, c_double] + [c_void_p] * 7),
    "uppe3d_adjoint_check": (c_int, [c_void_p, c_void_p, P(c_int)]),
    "uppe_rf3d_workspace_size": (c_int, [P(UppeRf3DDesc), P(c_size_t)]),
    "uppe_rf3d_plan_create": (c_int, [P(UppeRf3DDesc), c_void_p, c_size_t] + [c_void_p] * 9
                              + [c_void_p, P(c_void_p)]),
    "uppe_rf3d_plan_destroy": (c_int, [c_void_p]),
    "uppe_rf3d_load": (c_int, [c_void_p, c_void_p, c_double, c_void_p]),
    "uppe_rf3d_attempt": (c_int, [c_void_p, c_double, c_void_p, P(c_double)]),
    "uppe_rf3d_accept": (c_int, [c_void_p]),
    "uppe_rf3d_store_field": (c_int, [c_void_p, c_void_p, c_void_p]),
    "uppe_rf3d_store_spectrum": (c_int, [c_void_p, c_void_p, c_void_p]),
    "uppe1d_workspace_size": (c_int, [P(Uppe1DDesc), P(c_size_t)]),
    "uppe1d_propagate": (c_int, [P(Uppe1DDesc)] + [c_void_p] * 5
                         + [c_void_p] * 5 + [P(Uppe1DRecord), c_void_p, P(Uppe1DStatus)]),
    "uppe1d_propagate_noalloc": (c_int, [P(Uppe1DDesc)] + [c_void_p] * 5
                                 + [c_void_p] * 4 + [P(Uppe1DRecord), c_void_p,
                                                     P(Uppe1DStatus)]),
    "uppe1d_slices_workspace_size": (c_int, [P(Uppe1DDesc), c_int, P(c_size_t)]),
    "uppe1d_propagate_slices": (c_int, [P(Uppe1DDesc), c_int, P(Uppe1DSlice)] + [c_void_p] * 5
                                + [c_void_p] * 6 + [c_void_p, P(Uppe1DStatus)]),
    "uppe1d_batch_workspace_size": (c_int, [c_int, P(Uppe1DDesc), P(c_size_t)]),
    "uppe1d_batch_prepare": (c_int, [c_int, P(Uppe1DDesc)] + [c_void_p] * 5
                             + [c_void_p] * 5 + [P(Uppe1DRecord), c_void_p]),
    "uppe1d_batch_launch": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int,
                                    c_void_p]),
    "uppe1d_batch_status": (c_int, [c_void_p, c_int, c_void_p, P(Uppe1DStatus)]),
    "uppe1d_adjoint_workspace_size": (c_int, [c_int, c_int, P(c_size_t)]),
    "uppe1d_adjoint": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
}

_lib = None


class UppeError(RuntimeError):
    pass


def load():
    """Loads libuppe.so once; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UppeError("libuppe.so not built: run `python -m project_amd.build` "
                            "(the HIP extension is the only backend)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(lib, name):
                continue   # reported by missing_symbols(); calling it raises AttributeError
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def missing_symbols():
    lib = load()
    return [n for n in SIGNATURES if not hasattr(lib, n)]


def check(rc, what=""):
    if rc != 0:
        msg = load().uppe_last_error().decode(errors="replace")
        raise UppeError("%s failed (status %d): %s" % (what, rc, msg))


def ptr(t):
    """Device pointer of a torch tensor (or None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(torch_stream=None):
    import torch
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def host_ptr(a):
    """Pointer to a contiguous float64 NumPy array (kept alive by the caller)."""
    return ctypes.c_void_p(a.ctypes.data)
