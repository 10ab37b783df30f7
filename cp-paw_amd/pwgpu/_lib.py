"""ctypes binding of include/pwgpu.h (the library's C ABI).

The product path runs only on the HIP library; a missing or unloadable
libpwgpu.so raises immediately (there is no CPU fallback).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libpwgpu.so")

PW_OK = 0
PW_PREC_FP64 = 0
PW_PREC_MIXED = 1

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_cp = ctypes.c_char_p


class Request(ctypes.Structure):
    """struct pw_request of include/pwgpu.h."""
    _fields_ = [("rbas", ctypes.c_double * 9), ("epw_psi", ctypes.c_double), ("epw_rho", ctypes.c_double),
                ("nb", ctypes.c_int), ("nat", ctypes.c_int), ("nsp", ctypes.c_int), ("lnxx", ctypes.c_int),
                ("nbarepro", ctypes.c_int), ("prec", ctypes.c_int),
                ("isp", _c_int_p), ("lnx", _c_int_p), ("lox", _c_int_p),
                ("pro", _c_dbl_p), ("occ", _c_dbl_p), ("dov", _c_dbl_p), ("dh", _c_dbl_p),
                ("vofr", _c_dbl_p), ("r", _c_dbl_p), ("psi", _c_dbl_p),
                ("dt", ctypes.c_double), ("emass", ctypes.c_double), ("emasscg2", ctypes.c_double),
                ("friction", ctypes.c_double), ("etol", ctypes.c_double),
                ("nconv", ctypes.c_int), ("max_steps", ctypes.c_int), ("autotune", ctypes.c_int),
                ("rtol", ctypes.c_double)]


class Result(ctypes.Structure):
    """struct pw_result of include/pwgpu.h."""
    _fields_ = [("status", ctypes.c_int), ("steps", ctypes.c_int), ("converged", ctypes.c_int),
                ("etot", ctypes.c_double), ("ekin", ctypes.c_double), ("epot", ctypes.c_double),
                ("e1c", ctypes.c_double), ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_end", ctypes.c_double), ("footprint", ctypes.c_longlong), ("residual", ctypes.c_double)]

# name -> argument types (all return int status)
SIGNATURES = {
    "pw_version": [],
    "pw_last_error": [_cp, ctypes.c_int],
    "pw_malloc": [ctypes.POINTER(_vp), ctypes.c_longlong],
    "pw_free": [_vp],
    "pw_memcpy_htod": [_vp, _vp, ctypes.c_longlong],
    "pw_memcpy_dtoh": [_vp, _vp, ctypes.c_longlong],
    "pw_device_synchronize": [],
    "pw_plan_create": [_c_dbl_p, _c_dbl_p, ctypes.c_int,
                       ctypes.c_double, ctypes.c_double, ctypes.POINTER(_vp)],
    "pw_plan_destroy": [_vp],
    "pw_plan_geti4": [_vp, _cp, _c_int_p],
    "pw_plan_getr8": [_vp, _cp, _c_dbl_p],
    "pw_plan_getr8a": [_vp, _cp, ctypes.c_int, _c_dbl_p],
    "pw_plan_geti4a": [_vp, _cp, ctypes.c_int, _c_int_p],
    "pw_fft": [_vp, _cp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int],
    "pw_fft_noalloc": [_vp, _cp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_int],
    "pw_waves_sphere_setup": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                              _c_dbl_p, _c_dbl_p, _c_dbl_p],
    "pw_plan_set_stream": [_vp, _vp],
    "pw_waves_create": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                        ctypes.c_int, _c_int_p, ctypes.c_int, ctypes.POINTER(_vp)],
    "pw_waves_destroy": [_vp],
    "pw_waves_setr8": [_vp, _cp, ctypes.c_double],
    "pw_waves_setr8a": [_vp, _cp, ctypes.c_int, _c_dbl_p],
    "pw_waves_setc8a": [_vp, _cp, ctypes.c_int, _c_dbl_p],
    "pw_waves_getc8a": [_vp, _cp, ctypes.c_int, _c_dbl_p],
    "pw_waves_getr8a": [_vp, _cp, ctypes.c_int, _c_dbl_p],
    "pw_waves_getr8": [_vp, _cp, _c_dbl_p],
    "pw_waves_geti4": [_vp, _cp, _c_int_p],
    "pw_waves_devptr": [_vp, _cp, ctypes.POINTER(_vp)],
    "pw_waves_gramschmidt": [_vp],
    "pw_waves_projections": [_vp],
    "pw_waves_etot": [_vp],
    "pw_waves_propagate": [_vp],
    "pw_waves_orthogonalize": [_vp],
    "pw_waves_switch": [_vp],
    "pw_waves_step": [_vp],
    "pw_waves_set_stream": [_vp, _vp],
    "pw_waves_profile": [_vp, ctypes.c_int],
    "pw_waves_kernel_time": [_vp, _cp, _c_dbl_p, _c_int_p],
    "pw_waves_vpsi": [_vp, _vp, _vp],
    "pw_waves_density": [_vp, _vp, _vp],
    "pw_ortho_x": [ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_int_p],
    "pw_potential_tables": [_vp, ctypes.c_int, _c_int_p, _c_dbl_p],
    "pw_potential_functional": [_vp, ctypes.c_int],
    "pw_waves_set_potential": [_vp, _vp],
    "pw_vofrho": [_vp, ctypes.c_int, ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _vp, _vp, _c_dbl_p, _c_dbl_p,
                  _c_dbl_p],
    "pw_waves_autotune": [_vp],
    "pw_waves_footprint": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, ctypes.c_int,
                           _c_int_p, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)],
    "pw_server_create": [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(_vp)],
    "pw_server_destroy": [_vp],
    "pw_server_submit": [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)],
    "pw_server_wait": [_vp, ctypes.c_longlong, ctypes.POINTER(Result)],
    "pw_server_stats": [_vp, ctypes.POINTER(ctypes.c_longlong), _c_int_p, _c_int_p],
    "pw_server_pool_stats": [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                             ctypes.POINTER(ctypes.c_longlong)],
}

_lib = None


class PwError(RuntimeError):
    """Non-zero status from the library (the reference would ERROR$STOP)."""

    def __init__(self, code, msg):
        super().__init__(f"pwgpu error {code}: {msg}")
        self.code = code


def load(path=None):
    """Load libpwgpu.so and declare every entry point of include/pwgpu.h."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIBPATH
    if not os.path.exists(p):
        raise OSError(f"pwgpu HIP library not built: {p} (run __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().pw_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != PW_OK:
        raise PwError(rc, last_error())
    return rc


def dptr(a):
    """ctypes double pointer to a contiguous numpy float64/complex128 array."""
    return a.ctypes.data_as(_c_dbl_p)


def iptr(a):
    return a.ctypes.data_as(_c_int_p)
