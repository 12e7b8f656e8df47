"""ctypes binding of the jhtx C ABI (``include/jhtx.h``), built in-tree as ``libjhtx.so``.

There is no fallback: if the library is missing or a call fails, an exception is raised.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjhtx.so")

# Every symbol include/jhtx.h declares.
EXPORTS = (
    "jhtx_plan_create",
    "jhtx_plan_create_tuned",
    "jhtx_plan_get_tuning",
    "jhtx_plan_workspace_bytes",
    "jhtx_synth",
    "jhtx_adjoint",
    "jhtx_adjoint_noalloc",
    "jhtx_plan_destroy",
    "jhtx_last_error",
    "jhtx_plan_query",
    "jhtx_plan_set_profiling",
    "jhtx_plan_read_profile",
    "jhtx_host_rec_tables",
    "jhtx_plan_create_rings",
    "jhtx_legendre_synth",
    "jhtx_legendre_adjoint",
    "jhtx_dfs_extend",
    "jhtx_dfs_fold",
    "jhtx_nufft_grid_fill",
    "jhtx_nufft_grid_gather",
    "jhtx_nufft_interp",
    "jhtx_nufft_spread",
    "jhtx_nufft_workspace_bytes",
    "jhtx_probe_peaks",
    "jhtx_group_create",
    "jhtx_group_workspace_bytes",
    "jhtx_group_synth",
    "jhtx_group_adjoint",
    "jhtx_group_query",
    "jhtx_group_destroy",
)

JHTX_OK = 0
JHTX_EINVAL = -1
JHTX_EUNSUPPORTED = -2
JHTX_EDEVICE = -3
JHTX_ENOMEM = -4

PREC_F64 = 0
PREC_MIXED = 1
NOALLOC_SCRATCH_BYTES = 64 << 20  # JHTX_NOALLOC_SCRATCH_BYTES
FLAG_BLUESTEIN_ALL = 1
GROUP_MAX = 16  # plans per ragged group
GROUP_NSIDE_MAX = 512


TUNING_FIELDS = ("synth_rings", "adj_rings", "block", "mx_synth_group", "mx_synth_occ",
                 "mx_adj_group", "mx_adj_occ", "mx_adj_cb", "mx_single", "mx_f16", "mx_adj_f16",
                 "mx_synth_fold")


class Tuning(ctypes.Structure):
    """``jhtx_tuning``: kernel shape choices of a plan (0 = library default)."""

    _fields_ = [(name, ctypes.c_int) for name in TUNING_FIELDS]


class JhtxError(RuntimeError):
    """A jhtx C call returned an error code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"jhtx error {code}: {msg}")
        self.code = code


_LIB = None


def load() -> ctypes.CDLL:
    """Load ``libjhtx.so`` once; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"jhtx HIP library not found at {LIB_PATH}; build it first "
            "(python -m jht_amd.build, or __graft_entry__.build())"
        )
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.jhtx_plan_create.argtypes = [ctypes.c_int] * 6 + [ctypes.POINTER(vp)]
    lib.jhtx_plan_create.restype = ctypes.c_int
    lib.jhtx_plan_create_tuned.argtypes = [ctypes.c_int] * 6 + [ctypes.POINTER(Tuning), ctypes.POINTER(vp)]
    lib.jhtx_plan_create_tuned.restype = ctypes.c_int
    lib.jhtx_plan_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    lib.jhtx_plan_get_tuning.restype = ctypes.c_int
    lib.jhtx_plan_workspace_bytes.argtypes = [vp]
    lib.jhtx_plan_workspace_bytes.restype = ctypes.c_size_t
    for name in ("jhtx_synth", "jhtx_adjoint"):
        fn = getattr(lib, name)
        fn.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp]
        fn.restype = ctypes.c_int
    lib.jhtx_plan_destroy.argtypes = [vp]
    lib.jhtx_plan_destroy.restype = None
    lib.jhtx_last_error.argtypes = []
    lib.jhtx_last_error.restype = ctypes.c_char_p
    lib.jhtx_plan_query.argtypes = [vp, ctypes.c_int]
    lib.jhtx_plan_query.restype = ctypes.c_int
    lib.jhtx_plan_set_profiling.argtypes = [vp, ctypes.c_int]
    lib.jhtx_plan_set_profiling.restype = ctypes.c_int
    lib.jhtx_plan_read_profile.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    lib.jhtx_plan_read_profile.restype = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    lib.jhtx_host_rec_tables.argtypes = [ctypes.c_int] * 3 + [dp] * 7
    lib.jhtx_host_rec_tables.restype = ctypes.c_int
    i, ll, d = ctypes.c_int, ctypes.c_longlong, ctypes.c_double
    sigs = {
        "jhtx_plan_create_rings": [i, vp, i, i, i, ctypes.POINTER(vp)],
        "jhtx_legendre_synth": [vp, vp, vp, i, vp],
        "jhtx_legendre_adjoint": [vp, vp, vp, i, vp, vp],
        "jhtx_dfs_extend": [vp, i, i, i, vp, vp],
        "jhtx_dfs_fold": [vp, i, i, i, vp, vp],
        "jhtx_nufft_grid_fill": [vp, i, i, vp, vp, vp, i, i, vp],
        "jhtx_nufft_grid_gather": [vp, i, i, vp, vp, vp, i, i, vp],
        "jhtx_nufft_interp": [vp, i, i, i, d, vp, ll, vp, vp, vp, vp, ctypes.c_size_t, vp],
        "jhtx_nufft_spread": [vp, vp, vp, ll, vp, i, i, i, d, vp, ctypes.c_size_t, vp],
        "jhtx_probe_peaks": [vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double), vp],
        "jhtx_adjoint_noalloc": [vp, vp, vp, i, vp],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.jhtx_nufft_workspace_bytes.argtypes = [ll, i, i]
    lib.jhtx_nufft_workspace_bytes.restype = ctypes.c_size_t
    lib.jhtx_group_create.argtypes = [vp, vp, i, ctypes.POINTER(vp)]
    lib.jhtx_group_create.restype = ctypes.c_int
    lib.jhtx_group_workspace_bytes.argtypes = [vp]
    lib.jhtx_group_workspace_bytes.restype = ctypes.c_size_t
    for name in ("jhtx_group_synth", "jhtx_group_adjoint"):
        fn = getattr(lib, name)
        fn.argtypes = [vp, vp, vp, vp, vp]
        fn.restype = ctypes.c_int
    lib.jhtx_group_query.argtypes = [vp, i]
    lib.jhtx_group_query.restype = ctypes.c_longlong
    lib.jhtx_group_destroy.argtypes = [vp]
    lib.jhtx_group_destroy.restype = None
    _LIB = lib
    return lib


def check(code: int) -> None:
    if code != JHTX_OK:
        msg = load().jhtx_last_error().decode(errors="replace")
        raise JhtxError(code, msg)
