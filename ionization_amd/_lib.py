"""
ctypes binding of libtdse_hip.so (declared in include/tdse_hip.h).

The library is loaded from the package directory only. A missing library or
a missing symbol raises ``DeviceError``: there is no CPU fallback anywhere in
this package.
"""

import ctypes
import os

from .exceptions import DeviceError

FLAG_NONFINITE = 1     # TDSE_FLAG_NONFINITE: a stored norm was NaN or infinite
FLAG_NORM_GROWTH = 2   # TDSE_FLAG_NORM_GROWTH: a stored norm grew by more than 0.1 %

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtdse_hip.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_size_t_p = ctypes.POINTER(ctypes.c_size_t)


class ShDesc(ctypes.Structure):
    _fields_ = [
        ("r_points", ctypes.c_int32),
        ("l_bound", ctypes.c_int32),
        ("chunks", ctypes.c_int32),
        ("rows", ctypes.c_int32),
        ("n_steps", ctypes.c_int32),
        ("store_every", ctypes.c_int32),
        ("n_basis", ctypes.c_int32),
        ("lb", ctypes.c_int32),
        ("delta_r", ctypes.c_double),
        ("corr0", ctypes.c_double),
        ("centrifugal", ctypes.c_double),
        ("norm_by_l", ctypes.c_int32),
        ("gauge", ctypes.c_int32),
        ("method", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class LineDesc(ctypes.Structure):
    _fields_ = [
        ("z_points", ctypes.c_int32),
        ("chunks", ctypes.c_int32),
        ("rows", ctypes.c_int32),
        ("n_steps", ctypes.c_int32),
        ("store_every", ctypes.c_int32),
        ("n_basis", ctypes.c_int32),
        ("renormalize", ctypes.c_int32),
        ("method", ctypes.c_int32),
        ("delta_z", ctypes.c_double),
        ("offdiag", ctypes.c_double),
        ("zq_first", ctypes.c_double),
        ("zq_step", ctypes.c_double),
        ("vel_factor", ctypes.c_double),
    ]


class CylDesc(ctypes.Structure):
    _fields_ = [
        ("z_points", ctypes.c_int32),
        ("rho_points", ctypes.c_int32),
        ("z_chunks", ctypes.c_int32),
        ("z_rows", ctypes.c_int32),
        ("rho_chunks", ctypes.c_int32),
        ("rho_rows", ctypes.c_int32),
        ("lanes", ctypes.c_int32),
        ("n_steps", ctypes.c_int32),
        ("store_every", ctypes.c_int32),
        ("n_basis", ctypes.c_int32),
        ("has_mask", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("z_prefactor", ctypes.c_double),
        ("rho_prefactor", ctypes.c_double),
        ("coulomb", ctypes.c_double),
        ("test_charge", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("mask_inner", ctypes.c_double),
        ("mask_outer", ctypes.c_double),
        ("mask_smoothness", ctypes.c_double),
        ("z_first", ctypes.c_double),
        ("z_step", ctypes.c_double),
        ("rho_first", ctypes.c_double),
        ("rho_step", ctypes.c_double),
    ]


# name -> (restype, argtypes); every name here is declared in include/tdse_hip.h
SIGNATURES = {
    "tdse_abi_version": (ctypes.c_int, []),
    "tdse_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "tdse_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "tdse_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tdse_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "tdse_pool_usage": (ctypes.c_int, [ctypes.c_void_p, c_size_t_p, c_size_t_p]),
    "tdse_copy_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, c_double_p]),
    "tdse_tdma": (ctypes.c_int, [ctypes.c_void_p] + [c_double_p] * 5 + [ctypes.c_int64]),
    "tdse_tdma_batched": (
        ctypes.c_int,
        [ctypes.c_void_p] + [c_double_p] * 5 + [c_int64_p, c_int64_p, ctypes.c_int64],
    ),
    "tdse_tdma_batched_noalloc": (
        ctypes.c_int,
        [ctypes.c_void_p] + [ctypes.c_void_p] * 5 + [c_int64_p, c_int64_p, ctypes.c_int64, ctypes.c_int64],
    ),
    "tdse_sh_add": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(ShDesc)] + [c_double_p] * 5 + [c_int32_p, c_int32_p],
    ),
    "tdse_line_add": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(LineDesc)] + [c_double_p] * 5 + [c_int32_p],
    ),
    "tdse_cyl_add": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(CylDesc)] + [c_double_p] * 7 + [c_int32_p],
    ),
    "tdse_set_hold": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "tdse_profile_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_double_p, c_int64_p]),
    "tdse_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "tdse_line_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "tdse_sim_state": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_int32_p, c_int32_p, c_int32_p],
    ),
    "tdse_download": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32] + [c_double_p] * 4,
    ),
    "tdse_download_norm_by_l": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_double_p]),
    "tdse_upload_state": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_double_p, ctypes.c_int32],
    ),
    "tdse_remove": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "tdse_sim_flags": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_int32_p]),
    "tdse_materialize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "tdse_last_run_ms": (ctypes.c_int, [ctypes.c_void_p, c_double_p]),
    "tdse_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "tdse_profile_read": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_int64_p, c_double_p, c_int64_p]),
}

_LIB = None


def load():
    """Load the library once; raise DeviceError if it or any declared symbol is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise DeviceError(
            f"{LIB_PATH} not found: build it with `python -m ionization_amd.csrc.build` "
            "(there is no CPU fallback)"
        )
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise DeviceError(f"could not load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        try:
            fn = getattr(lib, name)
        except AttributeError as e:
            raise DeviceError(f"{LIB_PATH} does not export {name}") from e
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc, ctx=None):
    if rc != 0:
        lib = load()
        msg = lib.tdse_last_error(ctx)
        raise DeviceError(f"tdse error {rc}: {msg.decode() if msg else ''}")


def dptr(a):
    """double* of a C-contiguous float64 / complex128 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(c_double_p)
