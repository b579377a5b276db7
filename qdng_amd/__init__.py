"""qdng_amd: MI355X grid-propagation path of QDng behind a C ABI (libqdx.so).

This module is a thin ctypes binding of include/qdx.h and include/qdx_io.h.  Everything numeric
runs in libqdx.so on the GPU; there is no CPU fallback: if the library is
missing or no HIP device is present, the calls raise.

Arrays follow QDng's layout: complex128, dimension 0 fastest, multistate
vectors state-major.  In numpy terms a wave function on an (N0, N1, N2) grid
with S states is an array of shape (S, N2, N1, N0) in C order.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqdx.so")

FP64 = 0
FP32_BULK = 1
COMPACT = 2      # FP64 arithmetic, Chebyshev terms stored in 40 bits per component
_PREC = {"fp64": FP64, "fp32": FP32_BULK, "fp32_bulk": FP32_BULK, "compact": COMPACT, FP64: FP64,
         FP32_BULK: FP32_BULK, COMPACT: COMPACT}

_lib = None


class QdxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"qdx error {code}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("ndims", C.c_int32), ("N", C.c_int32 * 4), ("xmin", C.c_double * 4),
                ("xmax", C.c_double * 4)]


class ChebyInfo(C.Structure):
    _fields_ = [("order", C.c_int32), ("nz", C.c_int32), ("offset", C.c_double),
                ("scaling", C.c_double), ("prec", C.c_double), ("emin", C.c_double),
                ("emax", C.c_double), ("offset_i", C.c_double)]


class StepStats(C.Structure):
    _fields_ = [("n_hpsi", C.c_int64), ("norm2", C.c_double), ("nconv", C.c_double)]


EXPORTS = [
    "qdx_last_error", "qdx_version", "qdx_ctx_create", "qdx_ctx_synchronize", "qdx_ctx_destroy",
    "qdx_wf_create", "qdx_wf_upload", "qdx_wf_download", "qdx_wf_copy", "qdx_wf_norm",
    "qdx_wf_overlap", "qdx_wf_normalize", "qdx_wf_destroy", "qdx_ham_create",
    "qdx_ham_add_kinetic", "qdx_ham_add_potential", "qdx_ham_add_dipole", "qdx_ham_add_gobbler", "qdx_ham_bounds",
    "qdx_ham_apply", "qdx_ham_expec", "qdx_ham_expec_noalloc", "qdx_ham_destroy", "qdx_bessel_j0", "qdx_cheby_create",
    "qdx_cheby_info_get", "qdx_cheby_params", "qdx_cheby_coeffs", "qdx_gspo_create",
    "qdx_prop_apply", "qdx_prop_run", "qdx_prop_set_direction", "qdx_prop_exponent", "qdx_prop_destroy", "qdx_ctx_profile", "qdx_ctx_kernel_time",
    "qdx_batch_create", "qdx_batch_footprint", "qdx_batch_submit", "qdx_batch_run",
    "qdx_batch_job_state", "qdx_batch_job_status", "qdx_batch_fetch", "qdx_batch_release", "qdx_batch_usage",
    "qdx_batch_serve", "qdx_batch_set_graphs", "qdx_batch_graph_stats", "qdx_batch_destroy",
    "qdx_ctx_set_radix_plan", "qdx_ctx_get_radix_plan", "qdx_batch_get_tiling", "qdx_batch_set_tiling",
    "qdx_batch_autotune", "qdx_batch_set_autotune", "qdx_batch_tuning_save", "qdx_batch_tuning_load",
    "qdx_batch_set_lds_schedule", "qdx_batch_get_lds_schedule", "qdx_batch_autotune_lds",
    "qdx_ctx_set_line_points", "qdx_ctx_get_line_points", "qdx_cheby_create_imag", "qdx_bessel_i0",
    "qdx_wf_axpy", "qdx_wf_expec_position", "qdx_wf_expec_momentum", "qdx_eigen_imag", "qdx_eigen_ac", "qdx_sil_create", "qdx_gspo_create_imag", "qdx_arnoldi_create",
    # include/qdx_io.h
    "qdx_meta_create", "qdx_meta_read", "qdx_meta_write", "qdx_meta_size", "qdx_meta_key", "qdx_meta_get",
    "qdx_meta_set", "qdx_meta_set_int", "qdx_meta_set_double", "qdx_meta_destroy", "qdx_file_meta",
    "qdx_file_read", "qdx_file_write", "qdx_gridop_load", "qdx_gridop_save", "qdx_laser_load",
    "qdx_laser_save", "qdx_propagation_meta", "qdx_wf_load", "qdx_wf_save", "qdx_ham_add_potential_file",
    "qdx_ham_add_dipole_file", "qdx_wf_info", "qdx_io_set_format", "qdx_io_get_format", "qdx_file_format",
    # optimal control (include/qdx.h)
    "qdx_ham_laser_len", "qdx_ham_laser_get", "qdx_ham_laser_set", "qdx_proj_create", "qdx_proj_apply",
    "qdx_proj_expec", "qdx_proj_destroy", "qdx_oct_run",
]

K_CHEBY_ROW, K_LINE_PASS, K_CHEBY_INIT, K_GSPO_ROW, K_BATCH_ROW, K_BATCH_COL, K_BATCH_LINE = range(7)


class OctOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("operator_target", C.c_int32), ("phase", C.c_int32),
                ("ntargets", C.c_int32), ("iterations", C.c_int32), ("nsteps", C.c_int32),
                ("conv", C.c_double), ("alpha", C.c_double)]


OCT_METHODS = {"krotov": 0, "rabitz": 1, "rabitzfb": 2}
OCT_ROW_FIXED, OCT_ROW_PER_TARGET = 5, 5


class JobDesc(C.Structure):
    _fields_ = [("ham", C.c_void_p), ("ops", C.POINTER(C.c_void_p)), ("nops", C.c_int32), ("nsteps", C.c_int32),
                ("dt", C.c_double), ("order", C.c_int32), ("first_step", C.c_int32), ("prec", C.c_double),
                ("psi0", C.POINTER(C.c_double))]


class ServeResult(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("finish_s", C.c_double), ("n_hpsi", C.c_int64),
                ("failed_step", C.c_int64)]


class Tiling(C.Structure):
    _fields_ = [("row_threads", C.c_int32), ("row_lines", C.c_int32), ("col_width", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TuneResult(C.Structure):
    _fields_ = [("tiling", Tiling), ("radix0", C.c_int32 * 16), ("nst0", C.c_int32),
                ("radix1", C.c_int32 * 16), ("nst1", C.c_int32),
                ("row_us_default", C.c_double), ("row_us", C.c_double),
                ("col_us_default", C.c_double), ("col_us", C.c_double), ("candidates", C.c_int32)]

    def as_dict(self):
        return {"tiling": self.tiling.as_dict(), "radix0": list(self.radix0[:self.nst0]),
                "radix1": list(self.radix1[:self.nst1]), "row_us_default": self.row_us_default,
                "row_us": self.row_us, "col_us_default": self.col_us_default, "col_us": self.col_us,
                "candidates": self.candidates}


def lib():
    """Load libqdx.so (built by qdng_amd/build.py); raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python qdng_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    i32 = C.c_int32
    L.qdx_last_error.restype = C.c_char_p
    L.qdx_version.restype = C.c_char_p
    sig = {
        "qdx_ctx_create": [C.c_int, C.POINTER(vp)],
        "qdx_ctx_synchronize": [vp],
        "qdx_ctx_destroy": [vp],
        "qdx_wf_create": [vp, C.POINTER(Grid), i32, C.POINTER(vp)],
        "qdx_wf_upload": [vp, i32, dp],
        "qdx_wf_download": [vp, i32, dp],
        "qdx_wf_copy": [vp, vp],
        "qdx_wf_norm": [vp, dp, dp],
        "qdx_wf_overlap": [vp, vp, dp],
        "qdx_wf_normalize": [vp],
        "qdx_wf_destroy": [vp],
        "qdx_ham_create": [vp, C.POINTER(Grid), i32, i32, i32, C.POINTER(vp)],
        "qdx_ham_laser_len": [vp, C.POINTER(C.c_int64)],
        "qdx_ham_laser_get": [vp, dp, C.c_int64],
        "qdx_ham_laser_set": [vp, dp, C.c_int64],
        "qdx_proj_create": [C.POINTER(vp), i32, i32, i32, C.POINTER(vp)],
        "qdx_proj_apply": [vp, vp, vp],
        "qdx_proj_expec": [vp, vp, dp],
        "qdx_proj_destroy": [vp],
        "qdx_oct_run": [vp, C.POINTER(OctOpts), dp, C.c_int64, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), vp,
                        dp, C.POINTER(i32), C.POINTER(i32)],
        "qdx_ham_add_kinetic": [vp, i32, dp],
        "qdx_ham_add_potential": [vp, i32, i32, dp],
        "qdx_ham_add_dipole": [vp, i32, i32, dp, dp, i32, C.c_double],
        "qdx_ham_add_gobbler": [vp, i32, i32, dp, dp, i32, C.c_double, i32],
        "qdx_ham_bounds": [vp, dp, dp],
        "qdx_ham_apply": [vp, vp, vp, C.c_int64],
        "qdx_ham_expec": [vp, vp, C.c_int64, dp],
        "qdx_ham_expec_noalloc": [vp, vp, C.c_int64, dp],
        "qdx_ham_destroy": [vp],
        "qdx_bessel_j0": [i32, C.c_double, dp, C.POINTER(i32)],
        "qdx_cheby_create": [vp, C.c_double, i32, C.c_double, C.c_double, i32, C.POINTER(vp)],
        "qdx_cheby_info_get": [vp, C.POINTER(ChebyInfo)],
        "qdx_cheby_params": [vp, C.c_double, i32, C.c_double, C.c_double, C.POINTER(ChebyInfo), dp, i32],
        "qdx_cheby_coeffs": [vp, dp],
        "qdx_gspo_create": [C.POINTER(vp), i32, C.c_double, i32, C.POINTER(vp)],
        "qdx_prop_apply": [vp, vp, C.c_int64, C.POINTER(StepStats)],
        "qdx_prop_set_direction": [vp, i32],
        "qdx_prop_exponent": [vp, dp],
        "qdx_prop_run": [vp, vp, C.c_int64, C.c_int64, vp, dp, dp, C.POINTER(i32), C.POINTER(StepStats)],
        "qdx_prop_destroy": [vp],
        "qdx_ctx_profile": [vp, i32],
        "qdx_ctx_kernel_time": [vp, i32, C.POINTER(C.c_int64), dp],
        "qdx_batch_create": [vp, i32, C.c_int64, C.POINTER(vp)],
        "qdx_batch_footprint": [vp, C.POINTER(JobDesc), C.POINTER(C.c_int64)],
        "qdx_batch_submit": [vp, C.POINTER(JobDesc), C.POINTER(C.c_int64)],
        "qdx_batch_run": [vp, i32],
        "qdx_batch_job_state": [vp, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int64)],
        "qdx_batch_fetch": [vp, C.c_int64, dp],
        "qdx_batch_job_status": [vp, C.c_int64, C.POINTER(C.c_int64)],
        "qdx_batch_release": [vp, C.c_int64],
        "qdx_batch_usage": [vp, C.POINTER(i32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
        "qdx_batch_serve": [vp, C.POINTER(JobDesc), dp, i32, i32, C.POINTER(dp), C.POINTER(ServeResult)],
        "qdx_batch_set_graphs": [vp, i32],
        "qdx_ctx_set_radix_plan": [vp, i32, i32, C.POINTER(i32), i32],
        "qdx_ctx_get_radix_plan": [vp, i32, i32, C.POINTER(i32), C.POINTER(i32)],
        "qdx_ctx_set_line_points": [vp, i32],
        "qdx_cheby_create_imag": [vp, C.c_double, i32, C.c_double, C.c_double, C.POINTER(vp)],
        "qdx_bessel_i0": [i32, C.c_double, dp, C.POINTER(i32)],
        "qdx_wf_axpy": [vp, C.c_double, C.c_double, vp],
        "qdx_wf_expec_position": [vp, i32, dp],
        "qdx_wf_expec_momentum": [vp, i32, dp],
        "qdx_sil_create": [vp, C.c_double, i32, C.c_double, i32, C.POINTER(vp)],
        "qdx_gspo_create_imag": [C.POINTER(vp), i32, C.c_double, C.POINTER(vp)],
        "qdx_arnoldi_create": [vp, C.c_double, i32, C.c_double, i32, C.POINTER(vp)],
        "qdx_eigen_imag": [vp, vp, vp, i32, i32, C.c_double, i32, C.POINTER(vp), dp, dp, C.POINTER(i32)],
        "qdx_eigen_ac": [vp, vp, vp, i32, i32, C.c_double, i32, i32, i32, C.POINTER(vp), dp, dp, C.POINTER(i32), dp],
        "qdx_ctx_get_line_points": [vp, C.POINTER(i32)],
        "qdx_batch_get_tiling": [vp, i32, i32, C.POINTER(Tiling)],
        "qdx_batch_set_tiling": [vp, i32, i32, C.POINTER(Tiling)],
        "qdx_batch_autotune": [vp, i32, i32, C.POINTER(TuneResult)],
        "qdx_batch_set_autotune": [vp, i32],
        "qdx_batch_set_lds_schedule": [vp, i32, i32],
        "qdx_batch_get_lds_schedule": [vp, C.POINTER(i32), C.POINTER(i32)],
        "qdx_batch_autotune_lds": [vp, dp],
        "qdx_batch_tuning_save": [vp, C.c_char_p],
        "qdx_batch_tuning_load": [vp, C.c_char_p],
        "qdx_batch_graph_stats": [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
        "qdx_batch_destroy": [vp],
        "qdx_meta_create": [C.POINTER(vp)],
        "qdx_meta_read": [C.c_char_p, C.POINTER(vp)],
        "qdx_meta_write": [vp, C.c_char_p],
        "qdx_meta_size": [vp, C.POINTER(i32)],
        "qdx_meta_key": [vp, i32, C.c_char_p, i32],
        "qdx_meta_get": [vp, C.c_char_p, C.c_char_p, i32],
        "qdx_meta_set": [vp, C.c_char_p, C.c_char_p],
        "qdx_meta_set_int": [vp, C.c_char_p, C.c_int64],
        "qdx_meta_set_double": [vp, C.c_char_p, C.c_double],
        "qdx_meta_destroy": [vp],
        "qdx_file_meta": [C.c_char_p, C.POINTER(vp)],
        "qdx_file_read": [C.c_char_p, C.c_char_p, i32, vp, C.c_int64],
        "qdx_file_write": [C.c_char_p, C.c_char_p, i32, vp, vp, C.c_int64],
        "qdx_gridop_load": [C.c_char_p, i32, C.c_double, C.c_double, C.POINTER(Grid), dp, C.c_int64,
                            C.POINTER(C.c_int64)],
        "qdx_gridop_save": [C.c_char_p, C.POINTER(Grid), dp],
        "qdx_laser_load": [C.c_char_p, dp, C.POINTER(C.c_int64), dp, C.c_int64],
        "qdx_laser_save": [C.c_char_p, C.c_double, C.c_int64, dp],
        "qdx_io_set_format": [i32, i32, C.c_double],
        "qdx_wf_info": [vp, C.POINTER(Grid), C.POINTER(i32)],
        "qdx_io_get_format": [C.POINTER(i32), C.POINTER(i32), dp],
        "qdx_file_format": [C.c_char_p, C.POINTER(i32)],
        "qdx_propagation_meta": [C.c_char_p, C.c_int64, C.c_double, C.c_char_p, i32],
        "qdx_wf_load": [vp, C.c_char_p, i32, C.POINTER(vp)],
        "qdx_wf_save": [vp, C.c_char_p, i32],
        "qdx_ham_add_potential_file": [vp, i32, i32, C.c_char_p, i32, C.c_double, C.c_double],
        "qdx_ham_add_dipole_file": [vp, i32, i32, C.c_char_p, C.c_char_p, C.c_double],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def _chk(rc: int) -> None:
    if rc != 0:
        raise QdxError(rc, lib().qdx_last_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _enc(s) -> bytes:
    return os.fsencode(s)


# ---------------------------------------------------------------------------
# QDng files (include/qdx_io.h): host-only, no device needed

def _meta_to_dict(m) -> dict:
    n = C.c_int32()
    _chk(lib().qdx_meta_size(m, C.byref(n)))
    out = {}
    buf = C.create_string_buffer(4096)
    for i in range(n.value):
        _chk(lib().qdx_meta_key(m, i, buf, len(buf)))
        k = buf.value
        _chk(lib().qdx_meta_get(m, k, buf, len(buf)))
        out[k.decode()] = buf.value.decode()
    return out


def read_meta(path: str) -> dict:
    """key -> value of a QDng .meta file, in key order (qdx_meta_read)"""
    m = C.c_void_p()
    _chk(lib().qdx_meta_read(_enc(path), C.byref(m)))
    try:
        return _meta_to_dict(m)
    finally:
        lib().qdx_meta_destroy(m)


def file_meta(name: str) -> dict:
    """meta data of the file name (name.meta, else the name cut at its last '_'; qdx_file_meta)"""
    m = C.c_void_p()
    _chk(lib().qdx_file_meta(_enc(name), C.byref(m)))
    try:
        return _meta_to_dict(m)
    finally:
        lib().qdx_meta_destroy(m)


def _dict_to_meta(items: dict):
    m = C.c_void_p()
    _chk(lib().qdx_meta_create(C.byref(m)))
    try:
        for k, v in items.items():
            if isinstance(v, bool) or isinstance(v, str):
                _chk(lib().qdx_meta_set(m, k.encode(), str(v).encode()))
            elif isinstance(v, (int, np.integer)):
                _chk(lib().qdx_meta_set_int(m, k.encode(), int(v)))
            else:
                _chk(lib().qdx_meta_set_double(m, k.encode(), float(v)))
    except Exception:
        lib().qdx_meta_destroy(m)
        raise
    return m


def write_meta(path: str, items: dict) -> None:
    """write key = value lines in key order (qdx_meta_write); floats with 6 significant digits"""
    m = _dict_to_meta(items)
    try:
        _chk(lib().qdx_meta_write(m, _enc(path)))
    finally:
        lib().qdx_meta_destroy(m)


def read_file(name: str, suffix: str, count: int, dtype=np.complex128, counter: int = -1) -> np.ndarray:
    """count values of a raw data file name[_counter]suffix (qdx_file_read)"""
    out = np.zeros(int(count), dtype=dtype)
    _chk(lib().qdx_file_read(_enc(name), suffix.encode(), int(counter), out.ctypes.data_as(C.c_void_p),
                             out.nbytes))
    return out


def write_file(name: str, suffix: str, data: np.ndarray, meta: dict | None = None, counter: int = -1) -> None:
    """raw data file name[_counter]suffix and, with meta, name.meta (qdx_file_write)"""
    a = np.ascontiguousarray(data)
    m = _dict_to_meta(meta) if meta is not None else None
    try:
        _chk(lib().qdx_file_write(_enc(name), suffix.encode(), int(counter), m, a.ctypes.data_as(C.c_void_p),
                                  a.nbytes))
    finally:
        if m is not None:
            lib().qdx_meta_destroy(m)


def load_gridop(name: str, setzero: bool = False, offset: float = 0.0, scale: float = 1.0):
    """(N, xmin, xmax, values) of a QDng grid operator file (qdx_gridop_load)"""
    g = Grid()
    n = C.c_int64()
    _chk(lib().qdx_gridop_load(_enc(name), 0, 0.0, 1.0, C.byref(g), None, 0, C.byref(n)))
    v = np.zeros(n.value)
    _chk(lib().qdx_gridop_load(_enc(name), int(bool(setzero)), float(offset), float(scale), C.byref(g), _dp(v),
                               v.size, C.byref(n)))
    nd = g.ndims
    return list(g.N[:nd]), list(g.xmin[:nd]), list(g.xmax[:nd]), v


def save_gridop(name: str, N, xmin, xmax, values) -> None:
    g = make_grid(N, xmin, xmax)
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    if v.size != int(np.prod(N)):
        raise ValueError("values do not match the grid")
    _chk(lib().qdx_gridop_save(_enc(name), C.byref(g), _dp(v)))


FORMAT_BINARY, FORMAT_STREAM = 0, 1


def set_file_format(wf_format="binary", compress: bool = False, comptol: float = 1e-6) -> None:
    """Output format of wave functions, lasers and operators (QDng's global
    parameters wf_format, compress, comptol; qdx_io_set_format)."""
    fmt = {"binary": FORMAT_BINARY, "stream": FORMAT_STREAM}.get(wf_format, -1)
    _chk(lib().qdx_io_set_format(fmt, int(bool(compress)), float(comptol)))


def get_file_format():
    f, c, t = C.c_int32(), C.c_int32(), C.c_double()
    _chk(lib().qdx_io_get_format(C.byref(f), C.byref(c), C.byref(t)))
    return ("stream" if f.value == FORMAT_STREAM else "binary"), bool(c.value), t.value


def file_format(path: str) -> str:
    """"stream" or "binary": the format of an existing data file (qdx_file_format)."""
    f = C.c_int32()
    _chk(lib().qdx_file_format(_enc(path), C.byref(f)))
    return "stream" if f.value == FORMAT_STREAM else "binary"


def load_laser(name: str):
    """(dt, field) of a QDng laser file (qdx_laser_load)"""
    dt = C.c_double()
    nt = C.c_int64()
    _chk(lib().qdx_laser_load(_enc(name), C.byref(dt), C.byref(nt), None, 0))
    v = np.zeros(nt.value)
    _chk(lib().qdx_laser_load(_enc(name), C.byref(dt), C.byref(nt), _dp(v), v.size))
    return dt.value, v


def save_laser(name: str, dt: float, field) -> None:
    v = np.ascontiguousarray(field, dtype=np.float64).ravel()
    _chk(lib().qdx_laser_save(_enc(name), float(dt), v.size, _dp(v)))


def write_propagation_meta(directory: str, nt: int, dt: float, wf_base: str = "WF", wcycle: int = 1) -> None:
    """Propagation.meta as ProgPropa writes it (qdx_propagation_meta); directory '' or ending in '/'"""
    _chk(lib().qdx_propagation_meta(_enc(directory), int(nt), float(dt), wf_base.encode(), int(wcycle)))


def make_grid(N, xmin, xmax) -> Grid:
    if not 1 <= len(N) <= 4:
        raise QdxError(-1, "grids have 1..4 dimensions")
    g = Grid()
    g.ndims = len(N)
    for d in range(len(N)):
        g.N[d] = int(N[d])
        g.xmin[d] = float(xmin[d])
        g.xmax[d] = float(xmax[d])
    return g


def bessel_j0(n: int, x: float):
    out = np.zeros(n)
    nz = C.c_int32(0)
    _chk(lib().qdx_bessel_j0(int(n), float(x), _dp(out), C.byref(nz)))
    return out, nz.value


def bessel_i0(n: int, x: float):
    """I_k(x) exp(-x), k < n, and the underflow count (qdx_bessel_i0)"""
    out = np.zeros(n)
    nz = C.c_int32(0)
    _chk(lib().qdx_bessel_i0(int(n), float(x), _dp(out), C.byref(nz)))
    return out, nz.value


class Context:
    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        _chk(lib().qdx_ctx_create(int(device), C.byref(self.h)))
        self._destroy = lib().qdx_ctx_destroy

    def synchronize(self):
        _chk(lib().qdx_ctx_synchronize(self.h))

    def profile(self, on: bool = True) -> None:
        """Start (clearing the counters) or stop per-launch HIP-event timing."""
        _chk(lib().qdx_ctx_profile(self.h, int(bool(on))))

    def set_radix_plan(self, L: int, radix, precision="fp64") -> None:
        """radix sequence of the length-L line transforms (qdx_ctx_set_radix_plan)"""
        r = (C.c_int32 * len(radix))(*radix)
        _chk(lib().qdx_ctx_set_radix_plan(self.h, _PREC[precision], int(L), r, len(radix)))

    def radix_plan(self, L: int, precision="fp64"):
        r = (C.c_int32 * 16)()
        n = C.c_int32()
        _chk(lib().qdx_ctx_get_radix_plan(self.h, _PREC[precision], int(L), r, C.byref(n)))
        return list(r[:n.value])

    def set_line_points(self, points: int) -> None:
        """points per thread of the Chebyshev line kernels on grids with nd >= 2:
        8, 4 or 0 to choose (qdx_ctx_set_line_points)"""
        _chk(lib().qdx_ctx_set_line_points(self.h, int(points)))

    @property
    def line_points(self) -> int:
        n = C.c_int32()
        _chk(lib().qdx_ctx_get_line_points(self.h, C.byref(n)))
        return n.value

    def kernel_time(self, kernel_class: int):
        """(launches, total milliseconds) of one kernel class since profile(True)."""
        n = C.c_int64()
        ms = C.c_double()
        _chk(lib().qdx_ctx_kernel_time(self.h, int(kernel_class), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def close(self):
        if self.h:
            self._destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WaveFunction:
    """WFGridCartesian / WFMultistate on the device (complex128)."""

    def __init__(self, ctx: Context, N, xmin, xmax, nstates: int = 1):
        self.ctx = ctx
        self.N = [int(v) for v in N]
        self.grid = make_grid(N, xmin, xmax)
        self.S = int(nstates)
        self.h = C.c_void_p()
        _chk(lib().qdx_wf_create(ctx.h, C.byref(self.grid), self.S, C.byref(self.h)))
        self._destroy = lib().qdx_wf_destroy

    @property
    def size(self) -> int:
        return int(np.prod(self.N))

    def upload(self, psi: np.ndarray, state: int = -1) -> "WaveFunction":
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        want = self.size * (self.S if state < 0 else 1)
        if a.size != want:
            raise ValueError(f"expected {want} complex values, got {a.size}")
        _chk(lib().qdx_wf_upload(self.h, int(state), _dp(a)))
        return self

    def download(self, state: int = -1) -> np.ndarray:
        out = np.zeros(self.size * (self.S if state < 0 else 1), dtype=np.complex128)
        _chk(lib().qdx_wf_download(self.h, int(state), _dp(out)))
        return out

    def copy_from(self, other: "WaveFunction") -> None:
        _chk(lib().qdx_wf_copy(self.h, other.h))

    def norm(self):
        n = C.c_double()
        per = np.zeros(self.S)
        _chk(lib().qdx_wf_norm(self.h, C.byref(n), _dp(per)))
        return n.value, per

    def overlap(self, other: "WaveFunction") -> complex:
        out = np.zeros(2)
        _chk(lib().qdx_wf_overlap(self.h, other.h, _dp(out)))
        return complex(out[0], out[1])

    def normalize(self) -> None:
        _chk(lib().qdx_wf_normalize(self.h))

    def expec_position(self, dim: int = 0) -> np.ndarray:
        """<x_dim> of each state (qdx_wf_expec_position, GridPosition filter)"""
        out = np.zeros(self.S)
        _chk(lib().qdx_wf_expec_position(self.h, int(dim), _dp(out)))
        return out

    def expec_momentum(self, dim: int = 0) -> np.ndarray:
        """<p_dim> of each state (qdx_wf_expec_momentum, GridMomentum filter)"""
        out = np.zeros(self.S)
        _chk(lib().qdx_wf_expec_momentum(self.h, int(dim), _dp(out)))
        return out

    def axpy(self, a: complex, x: "WaveFunction") -> "WaveFunction":
        """self += a x (qdx_wf_axpy)"""
        _chk(lib().qdx_wf_axpy(self.h, float(complex(a).real), float(complex(a).imag), x.h))
        return self

    @classmethod
    def load(cls, ctx: Context, name: str, counter: int = -1) -> "WaveFunction":
        """Wave function from QDng files name[_counter].wf + name.meta
        (WFGridCartesian) or name-<i>... (WFMultistate); qdx_wf_load."""
        h = C.c_void_p()
        _chk(lib().qdx_wf_load(ctx.h, _enc(name), int(counter), C.byref(h)))
        g = Grid()
        S = C.c_int32()
        _chk(lib().qdx_wf_info(h, C.byref(g), C.byref(S)))
        self = cls.__new__(cls)
        self.ctx = ctx
        self.N = [int(g.N[d]) for d in range(g.ndims)]
        self.grid = g
        self.S = S.value
        self.h = h
        self._destroy = lib().qdx_wf_destroy
        return self

    def save(self, name: str, counter: int = -1) -> None:
        """Write in QDng's format (qdx_wf_save; binary unless set_file_format chose the stream)."""
        _chk(lib().qdx_wf_save(self.h, _enc(name), int(counter)))

    def __del__(self):
        try:
            if self.h:
                self._destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class Hamiltonian:
    """OSum (multistate=False) or OMultistate of grid terms."""

    def __init__(self, ctx: Context | None, N, xmin, xmax, nstates: int = 1,
                 multistate: bool = False, unity: bool = True):
        self.ctx = ctx
        self.N = [int(v) for v in N]
        self.grid = make_grid(N, xmin, xmax)
        self.S = int(nstates)
        self.h = C.c_void_p()
        _chk(lib().qdx_ham_create(ctx.h if ctx else None, C.byref(self.grid), self.S,
                                  int(bool(multistate)), int(bool(unity)), C.byref(self.h)))
        self._destroy = lib().qdx_ham_destroy

    @property
    def size(self) -> int:
        return int(np.prod(self.N))

    def add_kinetic(self, state: int, mass) -> "Hamiltonian":
        m = np.ascontiguousarray(np.atleast_1d(mass), dtype=np.float64)
        _chk(lib().qdx_ham_add_kinetic(self.h, int(state), _dp(m)))
        return self

    def add_potential(self, row: int, col: int, V) -> "Hamiltonian":
        v = np.ascontiguousarray(V, dtype=np.float64).ravel()
        if v.size != self.size:
            raise ValueError("potential size does not match the grid")
        _chk(lib().qdx_ham_add_potential(self.h, int(row), int(col), _dp(v)))
        return self

    def add_dipole(self, row: int, col: int, mu, laser, scale: float = 1.0) -> "Hamiltonian":
        m = np.ascontiguousarray(mu, dtype=np.float64).ravel()
        e = np.ascontiguousarray(laser, dtype=np.float64).ravel()
        if m.size != self.size:
            raise ValueError("dipole size does not match the grid")
        _chk(lib().qdx_ham_add_dipole(self.h, int(row), int(col), _dp(m), _dp(e), e.size, float(scale)))
        return self

    def add_gobbler(self, row: int, col: int, lp=None, rp=None, order: int = 20, gain: float = 1.0,
                    nip: bool = False) -> "Hamiltonian":
        """OGobbler: lp / rp map a dimension to its low-pass / high-pass edge
        (x units); nip: the negative imaginary potential -i v (qdx_ham_add_gobbler)"""
        a = np.full(len(self.N), np.nan)
        b = np.full(len(self.N), np.nan)
        for d, x in dict(lp or {}).items():
            a[int(d)] = float(x)
        for d, x in dict(rp or {}).items():
            b[int(d)] = float(x)
        _chk(lib().qdx_ham_add_gobbler(self.h, int(row), int(col), _dp(a), _dp(b), int(order), float(gain),
                                       int(bool(nip))))
        return self

    def add_potential_file(self, row: int, col: int, name: str, setzero: bool = False, offset: float = 0.0,
                           scale: float = 1.0) -> "Hamiltonian":
        """OGridPotential from QDng files name.op + name.meta (qdx_ham_add_potential_file)"""
        _chk(lib().qdx_ham_add_potential_file(self.h, int(row), int(col), _enc(name), int(bool(setzero)),
                                              float(offset), float(scale)))
        return self

    def add_dipole_file(self, row: int, col: int, mu_name: str, laser_name: str,
                        scale: float = 1.0) -> "Hamiltonian":
        """OGridDipole + laser from QDng files (qdx_ham_add_dipole_file)"""
        _chk(lib().qdx_ham_add_dipole_file(self.h, int(row), int(col), _enc(mu_name), _enc(laser_name),
                                           float(scale)))
        return self

    def bounds(self):
        a, b = C.c_double(), C.c_double()
        _chk(lib().qdx_ham_bounds(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def apply(self, dst: WaveFunction, src: WaveFunction, step: int = 0) -> None:
        _chk(lib().qdx_ham_apply(self.h, dst.h, src.h, int(step)))

    def expec(self, psi: WaveFunction, step: int = 0) -> float:
        out = C.c_double()
        _chk(lib().qdx_ham_expec(self.h, psi.h, int(step), C.byref(out)))
        return out.value

    def laser(self) -> np.ndarray:
        """Raw field samples of the first dipole term (Laser)."""
        n = C.c_int64()
        _chk(lib().qdx_ham_laser_len(self.h, C.byref(n)))
        out = np.zeros(n.value)
        _chk(lib().qdx_ham_laser_get(self.h, _dp(out), n.value))
        return out

    def set_laser(self, field) -> None:
        """Raw field samples of every dipole term (Laser::Set)."""
        f = np.ascontiguousarray(field, dtype=np.float64).ravel()
        _chk(lib().qdx_ham_laser_set(self.h, _dp(f), f.size))

    def cheby_params(self, dt: float, order: int = 0, prec: float = -1.0, scaling: float = 0.0):
        info = ChebyInfo()
        coeff = np.zeros(4096)
        _chk(lib().qdx_cheby_params(self.h, float(dt), int(order), float(prec), float(scaling),
                                    C.byref(info), _dp(coeff), coeff.size))
        return info, coeff[:info.order].copy()

    def __del__(self):
        try:
            if self.h:
                self._destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class _Propagator:
    h: C.c_void_p

    def apply(self, psi: WaveFunction, step: int = 0) -> StepStats:
        st = StepStats()
        _chk(lib().qdx_prop_apply(self.h, psi.h, int(step), C.byref(st)))
        return st

    def backward(self, on: bool = True) -> None:
        """OPropagator::Backward() / Forward(): real time with exponent +i dt."""
        _chk(lib().qdx_prop_set_direction(self.h, int(bool(on))))

    def exponent(self) -> complex:
        """OPropagator::Exponent(): c of exp(c H) for one step."""
        out = np.zeros(2)
        _chk(lib().qdx_prop_exponent(self.h, _dp(out)))
        return complex(out[0], out[1])

    def run(self, psi: WaveFunction, first_step: int, nsteps: int, psi0: WaveFunction | None = None,
            observables: bool = True):
        """ProgPropa step loop on the device; returns (norm2, autocorr, n_hpsi, stats)."""
        norm2 = np.zeros(max(nsteps, 1))
        ac = np.zeros(2 * max(nsteps, 1))
        nh = np.zeros(max(nsteps, 1), dtype=np.int32)
        st = StepStats()
        want_ac = observables and psi0 is not None
        _chk(lib().qdx_prop_run(self.h, psi.h, int(first_step), int(nsteps),
                                psi0.h if psi0 is not None else None,
                                _dp(norm2) if observables else None,
                                _dp(ac) if want_ac else None,
                                nh.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st)))
        return (norm2[:nsteps], (ac[0::2] + 1j * ac[1::2])[:nsteps], nh[:nsteps], st)

    def __del__(self):
        try:
            if self.h:
                self._destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class Cheby(_Propagator):
    """OCheby: real time, forward; or imag=True: imaginary time, backward
    (exponent -dt, qdx_cheby_create_imag; FP64)."""

    def __init__(self, H: Hamiltonian, dt: float, order: int = 0, prec: float = -1.0,
                 scaling: float = 0.0, precision="fp64", imag: bool = False):
        self.H = H
        self.h = C.c_void_p()
        if imag:
            if _PREC[precision] != FP64:
                raise QdxError(-1, "imaginary time needs FP64")
            _chk(lib().qdx_cheby_create_imag(H.h, float(dt), int(order), float(prec), float(scaling),
                                             C.byref(self.h)))
        else:
            _chk(lib().qdx_cheby_create(H.h, float(dt), int(order), float(prec), float(scaling),
                                        _PREC[precision], C.byref(self.h)))
        self._destroy = lib().qdx_prop_destroy
        self.info = ChebyInfo()
        _chk(lib().qdx_cheby_info_get(self.h, C.byref(self.info)))
        self.order = self.info.order

    def coeffs(self) -> np.ndarray:
        out = np.zeros(self.order)
        _chk(lib().qdx_cheby_coeffs(self.h, _dp(out)))
        return out


class SIL(_Propagator):
    """OSIL short iterative Lanczos (qdx_sil_create): real time forward, or
    imaginary time backward (imag=True); FP64."""

    def __init__(self, H: Hamiltonian, dt: float, order: int = 0, err: float = 0.0, imag: bool = False):
        self.H = H
        self.h = C.c_void_p()
        _chk(lib().qdx_sil_create(H.h, float(dt), int(order), float(err), int(bool(imag)), C.byref(self.h)))
        self._destroy = lib().qdx_prop_destroy


class Arnoldi(_Propagator):
    """OArnoldi (qdx_arnoldi_create): real time forward, or imaginary time
    backward (imag=True); FP64."""

    def __init__(self, H: Hamiltonian, dt: float, order: int = 0, conv: float = 0.0, imag: bool = False):
        self.H = H
        self.h = C.c_void_p()
        _chk(lib().qdx_arnoldi_create(H.h, float(dt), int(order), float(conv), int(bool(imag)), C.byref(self.h)))
        self._destroy = lib().qdx_prop_destroy


def eigen_imag(U, psi0: WaveFunction, nef: int, max_steps: int = 512, conv: float | None = None,
               diag: bool = True, H: Hamiltonian | None = None):
    """Eigenstates by imaginary-time relaxation (qdx_eigen_imag, ProgEigen
    method "imag"); conv defaults to QDng's 1e-5 with diag, 1e-8 without; H:
    the Hamiltonian of the energies (default U's own; GSPO needs it).
    Returns (wave functions, energies, raw energies, steps per state)."""
    if conv is None:
        conv = 1e-5 if diag else 1e-8
    hs = (C.c_void_p * nef)()
    e = np.zeros(nef)
    er = np.zeros(nef)
    st = np.zeros(nef, dtype=np.int32)
    _chk(lib().qdx_eigen_imag(U.h, H.h if H is not None else None, psi0.h, int(nef), int(max_steps), float(conv),
                              int(bool(diag)), hs, _dp(e), _dp(er), st.ctypes.data_as(C.POINTER(C.c_int32))))
    out = []
    for h in hs:
        w = WaveFunction.__new__(WaveFunction)
        w.ctx, w.N, w.grid, w.S = psi0.ctx, list(psi0.N), psi0.grid, psi0.S
        w.h = C.c_void_p(h)
        w._destroy = lib().qdx_wf_destroy
        out.append(w)
    return out, e, er, st


def eigen_ac(U, psi0: WaveFunction, nsteps: int = 512, win: bool = True, tol: float = 1.0, width: int = -1,
             diag: bool = False, H: Hamiltonian | None = None, max_ef: int = 256):
    """Eigenstates from the autocorrelation (qdx_eigen_ac, ProgEigen method
    "ac"): nsteps - 1 real-time steps of U, peaks of |FFT <psi(t)|psi0>|,
    one function integrated out per peak.  Returns (wave functions, energies,
    raw energies, autocorrelation)."""
    hs = (C.c_void_p * max_ef)()
    e = np.zeros(max_ef)
    er = np.zeros(max_ef)
    n = C.c_int32()
    ac = np.zeros(2 * nsteps)
    _chk(lib().qdx_eigen_ac(U.h, H.h if H is not None else None, psi0.h, int(nsteps), int(bool(win)), float(tol),
                            int(width), int(bool(diag)), int(max_ef), hs, _dp(e), _dp(er), C.byref(n), _dp(ac)))
    out = []
    for h in hs[:n.value]:
        w = WaveFunction.__new__(WaveFunction)
        w.ctx, w.N, w.grid, w.S = psi0.ctx, list(psi0.N), psi0.grid, psi0.S
        w.h = C.c_void_p(h)
        w._destroy = lib().qdx_wf_destroy
        out.append(w)
    return out, e[:n.value].copy(), er[:n.value].copy(), ac[0::2] + 1j * ac[1::2]


class GSPO(_Propagator):
    """OGSPO with split operators A, B, C, ... (each a Hamiltonian of one term kind)."""

    def __init__(self, ops, dt: float, precision="fp64", imag: bool = False):
        self.ops = list(ops)
        arr = (C.c_void_p * len(self.ops))(*[o.h.value for o in self.ops])
        self.h = C.c_void_p()
        if imag:   # imaginary time, backward (qdx_gspo_create_imag; FP64)
            if _PREC[precision] != FP64:
                raise QdxError(-1, "imaginary time needs FP64")
            _chk(lib().qdx_gspo_create_imag(arr, len(self.ops), float(dt), C.byref(self.h)))
        else:
            _chk(lib().qdx_gspo_create(arr, len(self.ops), float(dt), _PREC[precision], C.byref(self.h)))
        self._destroy = lib().qdx_prop_destroy


class Projection:
    """OProjection: P psi = sign sum_k <wf_k|psi> wf_k (or psi - that, inv)."""

    def __init__(self, wfs, positive: bool = True, inv: bool = False):
        arr = (C.c_void_p * len(wfs))(*[w.h.value for w in wfs])
        self.h = C.c_void_p()
        _chk(lib().qdx_proj_create(arr, len(wfs), int(bool(positive)), int(bool(inv)), C.byref(self.h)))
        self._destroy = lib().qdx_proj_destroy

    def apply(self, dst: WaveFunction, src: WaveFunction) -> None:
        _chk(lib().qdx_proj_apply(self.h, dst.h, src.h))

    def expec(self, psi: WaveFunction) -> float:
        out = C.c_double()
        _chk(lib().qdx_proj_expec(self.h, psi.h, C.byref(out)))
        return out.value

    def __del__(self):
        try:
            if self.h:
                self._destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


def oct_run(U, shape, psi_i, nsteps: int, psi_t=None, targets=None, method="krotov", phase=False,
            iterations: int = 50, conv: float = 1e-5, alpha: float = 1.0, gobbler=None):
    """ProgOCT::Run on the device (qdx_oct_run).  Returns (report rows,
    converged); the propagator's Hamiltonian then holds the optimised field."""
    o = OctOpts(OCT_METHODS[method], int(targets is not None), int(bool(phase)), len(psi_i), int(iterations),
                int(nsteps), float(conv), float(alpha))
    sh = np.ascontiguousarray(shape, dtype=np.float64).ravel()
    pi = (C.c_void_p * len(psi_i))(*[w.h.value for w in psi_i])
    pt = (C.c_void_p * len(psi_i))(*[w.h.value for w in psi_t]) if psi_t is not None else None
    tg = (C.c_void_p * len(psi_i))(*[t.h.value for t in targets]) if targets is not None else None
    rows = np.full((iterations + 1, OCT_ROW_FIXED + OCT_ROW_PER_TARGET * len(psi_i)), np.nan)
    nrows, conv_ = C.c_int32(), C.c_int32()
    _chk(lib().qdx_oct_run(U.h, C.byref(o), _dp(sh), sh.size, pi, pt, tg, gobbler.h if gobbler else None,
                           _dp(rows), C.byref(nrows), C.byref(conv_)))
    return rows[:nrows.value], bool(conv_.value)


class Batch:
    """Batched independent single-state Chebyshev jobs (qdx_batch_*).

    One tick advances every active job by one time step; jobs are admitted
    while they fit into the HBM budget (submit raises QdxError with code
    QDX_ERR_MEMORY = -4 otherwise)."""

    ERR_MEMORY = -4

    def __init__(self, ctx: Context, precision="fp32", hbm_budget: int = 8 << 30):
        self.ctx = ctx
        self.h = C.c_void_p()
        _chk(lib().qdx_batch_create(ctx.h, _PREC[precision], int(hbm_budget), C.byref(self.h)))
        self._destroy = lib().qdx_batch_destroy
        self._keep = {}
        self.sizes = {}

    @staticmethod
    def job(H=None, dt=1.0, nsteps=1, psi0=None, order=0, prec=-1.0, ops=None, first_step=0):
        """A job description: Chebyshev with Hamiltonian H (one state), or
        GSPO with split operators ops = [T, W]; first_step: the clock step
        it starts at (a restart from a snapshot)."""
        psi = np.ascontiguousarray(psi0, dtype=np.complex128).ravel()
        d = JobDesc()
        keep = [psi]
        if ops is not None:
            arr = (C.c_void_p * len(ops))(*[o.h.value for o in ops])
            keep.append(arr)
            keep.extend(ops)
            d.ham = None
            d.ops = C.cast(arr, C.POINTER(C.c_void_p))
            d.nops = len(ops)
        else:
            keep.append(H)
            d.ham = H.h.value
            d.nops = 0
        d.dt = float(dt)
        d.nsteps = int(nsteps)
        d.order = int(order)
        d.first_step = int(first_step)
        d.prec = float(prec)
        d.psi0 = psi.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        return d, keep

    def footprint(self, H=None, dt=1.0, nsteps=1, psi0=None, order=0, prec=-1.0, ops=None) -> int:
        d, keep = self.job(H, dt, nsteps, psi0, order, prec, ops)
        out = C.c_int64()
        _chk(lib().qdx_batch_footprint(self.h, C.byref(d), C.byref(out)))
        return out.value

    def submit(self, H=None, dt=1.0, nsteps=1, psi0=None, order=0, prec=-1.0, ops=None, first_step=0) -> int:
        d, keep = self.job(H, dt, nsteps, psi0, order, prec, ops, first_step)
        jid = C.c_int64()
        _chk(lib().qdx_batch_submit(self.h, C.byref(d), C.byref(jid)))
        self.sizes[jid.value] = keep[0].size
        return jid.value

    def try_submit(self, H=None, dt=1.0, nsteps=1, psi0=None, order=0, prec=-1.0, ops=None):
        """submit, or None when the job does not fit into the free budget"""
        try:
            return self.submit(H, dt, nsteps, psi0, order, prec, ops)
        except QdxError as e:
            if e.code == self.ERR_MEMORY:
                return None
            raise

    def serve(self, jobs, arrivals, max_active: int = 0, want_psi: bool = False):
        """Continuous batching over requests (kwargs dicts for job()) arriving
        at `arrivals` seconds; returns (admit_s, finish_s, n_hpsi[, psi list])."""
        n = len(jobs)
        descs = (JobDesc * max(n, 1))()
        keep = []
        for i, kw in enumerate(jobs):
            d, k = self.job(**kw)
            descs[i] = d
            keep.append(k)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        res = (ServeResult * max(n, 1))()
        outs = None
        psis = []
        if want_psi:
            outs = (C.POINTER(C.c_double) * max(n, 1))()
            for i, k in enumerate(keep):
                a = np.zeros(k[0].size, dtype=np.complex128)
                psis.append(a)
                outs[i] = a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        _chk(lib().qdx_batch_serve(self.h, descs, _dp(arr), n, int(max_active), outs, res))
        admit = np.array([res[i].admit_s for i in range(n)])
        finish = np.array([res[i].finish_s for i in range(n)])
        nh = np.array([res[i].n_hpsi for i in range(n)])
        # -1 per sound request, else the step at which it stopped (failed_step)
        self.last_failed = np.array([res[i].failed_step for i in range(n)], dtype=np.int64)
        return (admit, finish, nh, psis) if want_psi else (admit, finish, nh)

    def run(self, nticks: int = 1) -> None:
        _chk(lib().qdx_batch_run(self.h, int(nticks)))

    def state(self, jid: int, device: bool = True):
        """(steps_left, n_hpsi, steps_done); device=False skips the device read"""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        if device:
            _chk(lib().qdx_batch_job_state(self.h, int(jid), C.byref(a), C.byref(b), C.byref(c)))
        else:
            _chk(lib().qdx_batch_job_state(self.h, int(jid), C.byref(a), None, None))
        return a.value, b.value, c.value

    def failed_step(self, jid: int) -> int:
        """-1 for a sound job, else the step whose state's norm was not a
        number; the job stopped there (qdx_batch_job_status)"""
        v = C.c_int64()
        _chk(lib().qdx_batch_job_status(self.h, int(jid), C.byref(v)))
        return v.value

    def fetch(self, jid: int) -> np.ndarray:
        out = np.zeros(self.sizes[jid], dtype=np.complex128)
        _chk(lib().qdx_batch_fetch(self.h, int(jid), out.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def release(self, jid: int) -> None:
        _chk(lib().qdx_batch_release(self.h, int(jid)))
        self.sizes.pop(jid, None)

    def usage(self):
        n, b, t = C.c_int32(), C.c_int64(), C.c_int64()
        _chk(lib().qdx_batch_usage(self.h, C.byref(n), C.byref(b), C.byref(t)))
        return n.value, b.value, t.value

    def tiling(self, n0: int, n1: int) -> dict:
        """tile shape and LDS schedule that 2D jobs of shape n0 x n1 get"""
        t = Tiling()
        _chk(lib().qdx_batch_get_tiling(self.h, int(n0), int(n1), C.byref(t)))
        return t.as_dict()

    def set_tiling(self, n0: int, n1: int, row_threads: int, row_lines: int, col_width: int) -> None:
        t = Tiling(row_threads, row_lines, col_width)
        _chk(lib().qdx_batch_set_tiling(self.h, int(n0), int(n1), C.byref(t)))

    def lds_schedule(self):
        """(row, column) LDS buffers per line: 2 ping-pong, 1 in place"""
        r, c = C.c_int32(), C.c_int32()
        _chk(lib().qdx_batch_get_lds_schedule(self.h, C.byref(r), C.byref(c)))
        return r.value, c.value

    def set_lds_schedule(self, row_bufs: int, col_bufs: int) -> None:
        _chk(lib().qdx_batch_set_lds_schedule(self.h, int(row_bufs), int(col_bufs)))

    def autotune_lds(self):
        """time the four LDS schedules on the current jobs (they advance 12
        steps) and keep the fastest; returns ms per tick of (2,2), (1,2),
        (2,1), (1,1)"""
        ms = (C.c_double * 4)()
        _chk(lib().qdx_batch_autotune_lds(self.h, ms))
        return list(ms)

    def autotune(self, n0: int, n1: int) -> dict:
        """time the schedules of 2D jobs of shape n0 x n1 on the device and
        install the fastest (qdx_batch_autotune)"""
        r = TuneResult()
        _chk(lib().qdx_batch_autotune(self.h, int(n0), int(n1), C.byref(r)))
        return r.as_dict()

    def set_autotune(self, enable: bool) -> None:
        _chk(lib().qdx_batch_set_autotune(self.h, 1 if enable else 0))

    def save_tuning(self, path: str) -> None:
        _chk(lib().qdx_batch_tuning_save(self.h, os.fsencode(path)))

    def load_tuning(self, path: str) -> None:
        _chk(lib().qdx_batch_tuning_load(self.h, os.fsencode(path)))

    def set_graphs(self, enable: bool) -> None:
        """replay ticks from captured hipGraphs (default) or launch every
        kernel of a tick from the host"""
        _chk(lib().qdx_batch_set_graphs(self.h, 1 if enable else 0))

    def graph_stats(self):
        """(graphs captured, graph replays) so far"""
        c, r = C.c_int64(), C.c_int64()
        _chk(lib().qdx_batch_graph_stats(self.h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def __del__(self):
        try:
            if self.h:
                self._destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass
