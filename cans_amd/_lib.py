"""ctypes binding of libcans_hip.so (declared in include/cans_hip.h).

The shared library is built in-tree (``cans_amd/libcans_hip.so``) by
``cans_amd.build.build()``.  Importing this module never compiles anything; a
missing library raises immediately so that no caller can silently run without
the HIP path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libcans_hip.so")

CANS_OK, CANS_EINVAL, CANS_EHIP, CANS_ENOMEM, CANS_EUNSUPPORTED = 0, -1, -2, -3, -4
CANS_FP64, CANS_MIXED = 0, 1
CANS_FIELD_VELOCITY, CANS_FIELD_PRESSURE = 0, 1
HALO_ABSENT, HALO_LOWER, HALO_NONE, HALO_UPPER = -99, -1, 0, 1

_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_cp = C.c_char_p


class Helmholtz(C.Structure):
    """cans_helmholtz (include/cans_hip.h)."""
    _fields_ = [("plan", _vp), ("normfft", C.c_double), ("lambdaxy_d", _vp), ("a_d", _vp), ("b_d", _vp),
                ("c_d", _vp), ("rhsbx_d", _vp), ("rhsby_d", _vp), ("rhsbz_d", _vp), ("cbc", C.c_char * 6),
                ("c_or_f", C.c_char * 3)]


class StepParams(C.Structure):
    """cans_step_params (include/cans_hip.h)."""
    _fields_ = [("cbcvel", C.c_char * 18), ("bcvel", C.c_double * 18), ("cbcpre", C.c_char * 6),
                ("bcpre", C.c_double * 6), ("nb", C.c_int * 6), ("is_bound", C.c_int * 6),
                ("dl", C.c_double * 3), ("dli", C.c_double * 3), ("visc", C.c_double),
                ("is_forced", C.c_int * 3), ("velf", C.c_double * 3), ("bforce", C.c_double * 3),
                ("dzc_h", _vp), ("dzf_h", _vp), ("dzci_d", _vp), ("dzfi_d", _vp),
                ("grid_vol_ratio_c_d", _vp), ("grid_vol_ratio_f_d", _vp), ("plan", _vp),
                ("normfft", C.c_double), ("lambdaxy_d", _vp), ("a_d", _vp), ("b_d", _vp), ("c_d", _vp),
                ("rhsbx_d", _vp), ("rhsby_d", _vp), ("rhsbz_d", _vp), ("hvel", Helmholtz * 3),
                ("nscal", C.c_int), ("scal", _vp), ("is_buoyancy", C.c_int), ("gacc", C.c_double * 3),
                ("beta", C.c_double)]


class Scalar(C.Structure):
    """cans_scalar (include/cans_hip.h)."""
    _fields_ = [("val", _vp), ("dsdtrko", _vp), ("alpha", C.c_double), ("cbc", C.c_char * 6),
                ("bc", C.c_double * 6), ("source", C.c_double), ("is_forced", C.c_int), ("scalf", C.c_double),
                ("h", Helmholtz)]


class PoissonBox(C.Structure):
    """cans_poisson_box (include/cans_hip.h)."""
    _fields_ = [("n", C.c_int * 3), ("sy", C.c_long), ("sz", C.c_long), ("cbc", C.c_char * 6),
                ("lambdaxy_h", _vp), ("a_h", _vp), ("b_h", _vp), ("c_h", _vp), ("p_d", _vp)]


class State(C.Structure):
    """cans_state (include/cans_hip.h)."""
    _fields_ = [("u", _vp), ("v", _vp), ("w", _vp), ("p", _vp), ("pp", _vp),
                ("dudtrko", _vp), ("dvdtrko", _vp), ("dwdtrko", _vp)]


# name -> (restype, argtypes); the exported C-ABI
SIGNATURES = {
    "cans_ctx_create": (C.c_int, [_ip, C.c_int, C.c_int, _vp, C.POINTER(_vp)]),
    "cans_ctx_destroy": (C.c_int, [_vp]),
    "cans_ctx_layout": (C.c_int, [_vp, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "cans_ctx_sync": (C.c_int, [_vp]),
    "cans_last_error": (C.c_char_p, []),
    "cans_ctx_elem_size": (C.c_int, [_vp, C.c_int, _ip]),
    "cans_field_alloc": (C.c_int, [_vp, C.POINTER(_vp)]),
    "cans_field_alloc_kind": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "cans_field_free": (C.c_int, [_vp, _vp]),
    "cans_field_register": (C.c_int, [_vp, _vp, C.c_int]),
    "cans_ctx_set_numerics": (C.c_int, [_vp, C.c_int, C.c_int]),
    "cans_ctx_set_dtdma": (C.c_int, [_vp, C.c_int]),
    "cans_ctx_set_refinement": (C.c_int, [_vp, C.c_int, C.c_int]),
    "cans_ctx_set_mixed_increments": (C.c_int, [_vp, C.c_int]),
    "cans_field_upload": (C.c_int, [_vp, _vp, _dp]),
    "cans_field_download": (C.c_int, [_vp, _vp, _dp]),
    "cans_device_alloc": (C.c_int, [_vp, C.c_long, C.POINTER(_vp)]),
    "cans_device_free": (C.c_int, [_vp, _vp]),
    "cans_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_long]),
    "cans_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_long]),
    "cans_initgrid": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
    "cans_initgrid_natural": (C.c_int, [C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
    "cans_eigenvalues": (C.c_int, [C.c_int, _cp, C.c_char, _dp]),
    "cans_transform_matrix": (C.c_int, [C.c_int, _cp, C.c_char, C.c_int, _dp]),
    "cans_bc_rhs": (C.c_int, [_cp, _dp, _dp, _dp, C.c_char, _dp]),
    "cans_tridmatrix": (C.c_int, [_cp, C.c_int, _dp, _dp, C.c_char, _dp, _dp, _dp]),
    "cans_fftini": (C.c_int, [_vp, _ip, _cp, _cp, C.POINTER(_vp), _dp]),
    "cans_fftend": (C.c_int, [_vp]),
    "cans_fft_tuning": (C.c_int, [_vp, _ip]),
    "cans_fft_radix": (C.c_int, [_vp, _ip]),
    "cans_fft_set_radix": (C.c_int, [_vp, _ip]),
    "cans_solver": (C.c_int, [_vp, _ip, _ip, _vp, C.c_double, _vp, _vp, _vp, _vp, _cp, _cp, _vp]),
    "cans_solver_noalloc": (C.c_int, [_vp, _ip, _ip, _vp, C.c_double, _vp, _vp, _vp, _vp, _cp, _cp, _vp]),
    "cans_poisson_batch_create": (C.c_int, [_vp, C.c_int, C.POINTER(PoissonBox), _dp, C.POINTER(_vp)]),
    "cans_poisson_batch_solve": (C.c_int, [_vp]),
    "cans_poisson_batch_tuning": (C.c_int, [_vp, _ip]),
    "cans_poisson_batch_destroy": (C.c_int, [_vp]),
    "cans_solver_refined": (C.c_int, [_vp, _ip, _ip, _vp, C.c_double, _vp, _vp, _vp, _vp, _cp, _cp, _dp, _vp, _vp,
                                      C.c_int, _vp]),
    "cans_solve_helmholtz_batch": (C.c_int, [_vp, C.c_int, C.POINTER(Helmholtz), _dp, _ip, C.POINTER(_vp)]),
    "cans_solve_helmholtz": (C.c_int, [_vp, _ip, _ip, _vp, C.c_double, C.c_double, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _ip, _cp, _cp, _vp]),
    "cans_rk": (C.c_int, [_vp, _dp, _ip, _dp, _vp, _vp, _vp, _vp, C.c_double, C.c_double, _vp, _ip, _dp,
                          _dp, _dp, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dp]),
    "cans_rk_scal": (C.c_int, [_vp, _dp, _ip, _dp, _dp, _vp, _vp, _vp, C.c_double, C.c_double, _ip, _vp, _vp, _vp,
                               C.c_int, C.c_double, C.c_double, _dp, _vp, _vp, _dp]),
    "cans_bulk_forcing_s": (C.c_int, [_vp, _ip, C.c_int, C.c_double, _vp]),
    "cans_fillps": (C.c_int, [_vp, _ip, _dp, _vp, C.c_double, _vp, _vp, _vp, _vp]),
    "cans_correc": (C.c_int, [_vp, _ip, _dp, _vp, C.c_double, _vp, _vp, _vp, _vp]),
    "cans_updatep": (C.c_int, [_vp, _ip, _dp, _vp, _vp, C.c_double, _vp, _vp, C.c_int]),
    "cans_bounduvw": (C.c_int, [_vp, _cp, _ip, _dp, _ip, _ip, _dp, _dp, _dp, _vp, _vp, _vp, C.c_int,
                                C.c_int, C.c_int]),
    "cans_boundp": (C.c_int, [_vp, _cp, _ip, _dp, _ip, _ip, _dp, _dp, _vp, C.c_int]),
    "cans_updt_rhs_b": (C.c_int, [_vp, _cp, _cp, _ip, _ip, _vp, _vp, _vp, _vp, _dp]),
    "cans_chkdiv": (C.c_int, [_vp, _ip, _ip, _dp, _dp, _vp, _vp, _vp, _vp, _dp, _dp]),
    "cans_chkdt": (C.c_int, [_vp, _ip, _dp, _vp, _vp, C.c_double, C.c_double, _vp, _vp, _vp, C.c_int,
                             C.c_int, _dp]),
    "cans_bulk_mean": (C.c_int, [_vp, _ip, _vp, _vp, _dp]),
    "cans_rk3_substep": (C.c_int, [_vp, C.POINTER(StepParams), C.c_int, C.c_double, C.POINTER(State)]),
    "cans_rk3_step": (C.c_int, [_vp, C.POINTER(StepParams), C.c_double, C.POINTER(State)]),
    "cans_ctx_forcing": (C.c_int, [_vp, C.POINTER(_vp)]),
    "cans_mom_jt": (C.c_int, [_vp, _ip, _dp, C.c_double] + [_vp] * 12),
    "cans_rk3_substep_adjoint": (C.c_int, [_vp, C.POINTER(StepParams), C.c_int, C.c_double] + [_vp] * 16),
    "cans_rk3_substep_adjoint_scal": (C.c_int, [_vp, C.POINTER(StepParams), C.c_int, C.c_double] + [_vp] * 16
                                      + [C.POINTER(_vp)] * 5),
}

_lib = None


class CansError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise CansError("libcans_hip.so is not built: run cans_amd.build.build() (or __graft_entry__.build())")
        # One HIP runtime per process: PyTorch bundles its own libamdhip64, and the
        # Python host hands torch's streams and memory to this library, so torch's
        # runtime must be the one loaded first (the library then binds to it).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != CANS_OK:
        msg = lib().cans_last_error().decode(errors="replace")
        raise CansError("%s failed (%d): %s" % (what, rc, msg))


def ivec(seq):
    return (C.c_int * len(seq))(*[int(x) for x in seq])


def dvec(seq):
    return (C.c_double * len(seq))(*[float(x) for x in seq])
