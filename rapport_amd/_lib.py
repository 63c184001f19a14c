"""ctypes binding of librapport_amd.so (C ABI in include/rapport_amd.h).

The library is built in-tree by ``rapport_amd.build`` (``__graft_entry__.build``).
There is no fallback: if the library is missing or cannot be loaded, every GPU
entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librapport_amd.so")

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
P = c_void_p  # device pointers travel as void*


class PlasmaParams(ctypes.Structure):
    """fp_plasma_params."""
    _fields_ = [("Nr", c_int), ("Nt", c_int), ("dt", c_double), ("field_to_I", c_double),
                ("beta_g", c_double), ("beta_s", c_double), ("na", c_double),
                ("inv_tau_r", c_double), ("inv_tau_ste", c_double), ("enable_ste", c_int),
                ("table_size", c_int), ("log_min_I", c_double), ("inv_log_step", c_double),
                ("Imin", c_double), ("Imax", c_double)]


class NLParams(ctypes.Structure):
    """fp_nl_params."""
    _fields_ = [("Nr", c_int), ("Nt", c_int), ("dz", c_double), ("invE2", c_double),
                ("Imin", c_double), ("Imax", c_double), ("photo_pref", c_double),
                ("rho_max", c_double), ("plasma_pref", c_double), ("pc_re", c_double),
                ("pc_im", c_double), ("ste_pref", c_double), ("kerr_pref", c_double),
                ("f_R", c_double), ("en_kerr_inst", c_int), ("en_kerr_raman", c_int),
                ("en_pi_loss", c_int), ("en_plasma_absorb", c_int), ("n_knots", c_int)]


class NLTables(ctypes.Structure):
    """fp_nl_tables (float64 device pointers)."""
    _fields_ = [("tw", P), ("mA", P), ("mB", P), ("mC", P), ("Rf", P), ("knots", P),
                ("pcoef", P), ("rw", P)]


class NLTablesF32(ctypes.Structure):
    """fp_nl_tables_f32."""
    _fields_ = [("tw", P), ("mA", P), ("mB", P), ("mC", P), ("Rf", P), ("knots", P),
                ("inv_h", P), ("pcoef_s", P), ("rw", P)]


class BatchInstance(ctypes.Structure):
    """fp_batch_instance (one instance of the batched mixed-precision z-step)."""
    _fields_ = [("Nr", c_int), ("Nt", c_int), ("y_scale_exp", c_int), ("enable_ste", c_int),
                ("Rk", ctypes.c_float), ("iRk", ctypes.c_float), ("Yh", P), ("Yl", P),
                ("spec", P), ("w0", P), ("w1", P), ("rho", P), ("rho_s", P), ("phase", P),
                ("mask_r", P), ("slots", P), ("W_lut", P), ("Ws_lut", P), ("tw", P),
                ("status", P), ("pp", PlasmaParams), ("np", NLParams), ("nt", NLTablesF32),
                ("zc_lo", c_int), ("zc_hi", c_int)]


class Rate0DParams(ctypes.Structure):
    """fp_rate0d_params (0-D re-integration over intensity histories)."""
    _fields_ = [("n_rows", c_int), ("Nt", c_int), ("na", c_double), ("inv_tau_r", c_double),
                ("beta_g", c_double), ("beta_s", c_double), ("avalanche", c_int),
                ("trapping", c_int), ("ste", c_int), ("n_knots", c_int), ("Imin", c_double),
                ("Imax", c_double)]


class RecordDst(ctypes.Structure):
    """fp_record_dst (history entries of one record)."""
    _fields_ = [("fluence", P), ("rho", P), ("rho_s", P), ("ipeak", P), ("absorbed", P),
                ("bins", P), ("rho_at", P), ("it_snap", c_int), ("rho_onaxis", P),
                ("rho_s_onaxis", P), ("I_onaxis", P), ("r_stride", c_int), ("t_stride", c_int),
                ("Nt_sub", c_int), ("rho_cube", P), ("rho_s_cube", P), ("I_cube", P), ("r_w", P),
                ("partial", P), ("imax", P), ("dedz_photo", P), ("dedz_plasma", P)]


# name -> argtypes (restype is int for all but fp_last_error)
_SIGS = {
    "fp_version": [],
    "fp_plasma_scan": [P, P, P, ctypes.POINTER(PlasmaParams), P, P, P, P],
    "fp_plasma_scan_spec": [P, P, P, P, ctypes.POINTER(PlasmaParams), P, P, P, P, P],
    "fp_fft_rows": [P, P, c_int, c_int, c_int, P, P],
    "fp_hankel_gemm": [P, P, P, c_int, c_int, c_double, P, P, P],
    "fp_half_linear": [P, P, c_int, c_int, P, P, c_double, c_double, P, P, P],
    "fp_nl_split": [P, P, P, P, P, ctypes.POINTER(NLParams), ctypes.POINTER(NLTables), P],
    "fp_nl_step": [P, P, P, P, c_int, c_int, ctypes.POINTER(NLParams), ctypes.POINTER(NLTables), P],
    "fp_record_rows": [P, P, P, P, ctypes.POINTER(NLParams), ctypes.POINTER(NLTables), c_double,
                       c_int, P, P, P, P],
    "fp_record_history": [P, c_int, P, P, P, ctypes.POINTER(NLParams), ctypes.POINTER(NLTables),
                          c_double, c_int, P, P, ctypes.POINTER(RecordDst), P],
    "fp_record_history_noalloc": [P, c_int, P, P, P, ctypes.POINTER(NLParams),
                                  ctypes.POINTER(NLTables), c_double, c_int, P, P,
                                  ctypes.POINTER(RecordDst), P],
    "fp_absmax_f32": [P, ctypes.c_longlong, P, P],
    "fp_fft_rows_c64": [P, P, c_int, c_int, c_int, P, P],
    "fp_plasma_scan_spec_c64": [P, P, P, P, ctypes.POINTER(PlasmaParams), P, P, P, P, P, P],
    "fp_hankel_gemm_mixed": [P, P, c_int, P, P, c_int, c_int, ctypes.c_float, P, P, P, P, P, P],
    "fp_hankel_gemm_band": [P, P, P, c_int, c_int, c_double, P, P, c_int, c_int, c_int, P],
    "fp_hankel_gemm_mixed_band": [P, P, c_int, P, P, c_int, c_int, ctypes.c_float, P, P, P, P, P,
                                  c_int, c_int, c_int, P],
    "fp_nl_split_c64": [P, P, P, P, P, ctypes.POINTER(NLParams), ctypes.POINTER(NLTablesF32), P],
    "fp_nl_step_c64": [P, P, P, P, c_int, c_int, ctypes.POINTER(NLParams),
                       ctypes.POINTER(NLTablesF32), P, P],
    "fp_adj_nl_step_c64": [P, P, P, P, P, P, P, P, ctypes.POINTER(NLParams),
                           ctypes.POINTER(NLTablesF32), P, P],
    "fp_adj_plasma_c64": [P, P, P, P, P, P, ctypes.POINTER(PlasmaParams), P, P, P, P, P],
    "fp_batch_plan_create": [ctypes.POINTER(BatchInstance), c_int, ctypes.POINTER(c_void_p)],
    "fp_batch_plan_destroy": [P],
    "fp_batch_set_config": [P, c_int, c_int, c_int, c_int],
    "fp_batch_set_grid": [P, c_int, c_int],
    "fp_batch_set_merged": [P, c_int],
    "fp_batch_set_merged_passes": [P, c_int],
    "fp_batch_pass": [P, c_int, P],
    "fp_batch_step_mixed": [P, P],
    "fp_batch_launches": [P],
    "fp_rate0d_batch": [P, P, P, P, ctypes.POINTER(Rate0DParams), P, P, P, P],
    "fp_shock_propagate": [P, P, P, P, c_int, c_int, c_int, c_int, P, P],
}

_lib = None


def load():
    """Load the library (once). Raises if it is missing -- no CPU fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rapport_amd.build` "
                           "(or __graft_entry__.build()); there is no fallback path")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    lib.fp_last_error.argtypes = []
    lib.fp_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def declared_symbols():
    """Every entry point declared in include/rapport_amd.h."""
    return sorted(list(_SIGS) + ["fp_last_error"])


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.fp_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed (code {rc}): {msg}")
    return rc


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a GPU tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)
