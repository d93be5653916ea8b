"""ctypes binding of libbass_hip.so (declared in include/bass_hip.h).

The library is the only compute path: importing this module never falls back
to a CPU implementation, and every call raises if the library or a GPU is
missing.  PyTorch supplies device buffers and the current HIP stream.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbass_hip.so")

BASS_OK = 0
BASS_E_OVERFLOW = -1
BASS_E_ARG = -2
BASS_E_WORKSPACE = -3
BASS_E_HIP = -4
BASS_E_DEVICE = -5

MODE_GATE = 0
MODE_ROT = 1
MODE_MOVE = 2
LIFT_NO_SCAN = 16  # bass_lift mode flag: no output-count scan (a segmented truncate reads the output next)
LIFT_SPEC = 32  # bass_lift mode flag: the truncate's digit-1 pass against the predicted bucket
TRUNC_SPEC = 8  # bass_truncate in_flags bit: the input comes from a LIFT_SPEC lift issued just before
GT_COUNTED = 1  # bass_gate_truncate flag: the input count from the state-count slot
GT_SPEC = 2  # bass_gate_truncate flag: the speculative digit-1 pass

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t


class BassError(RuntimeError):
    """A libbass_hip call failed (message from bass_last_error)."""


class BassStats(ctypes.Structure):
    _fields_ = [
        ("nnz", ctypes.c_int64),
        ("total", ctypes.c_double),
        ("kept", ctypes.c_double),
        ("norm", ctypes.c_double),
        ("sum2", ctypes.c_double),
        ("sum4", ctypes.c_double),
        ("n_in", ctypes.c_int64),
        ("err", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


# bass_job (include/bass_hip.h) as a numpy record, so a batch's table is
# filled with vector assignments
JOB_DTYPE = np.dtype([("x", "<u8"), ("a", "<u8"), ("xo", "<u8"), ("ao", "<u8"), ("n", "<i8"), ("k", "<i8"),
                      ("cap_out", "<i8"), ("j", "<i4"), ("j2", "<i4"), ("renorm", "<i4"), ("key_bits", "<i4"),
                      ("g", "<f8", (32,)), ("key_map", "<i4", (64,)), ("scale", "<f8"),
                      ("cnt_in", "<u8"), ("cnt_out", "<u8"), ("nq", "<i4"), ("pad", "<i4")])
assert JOB_DTYPE.itemsize == 616

# bass_bplan_run operations
BOP_GATE1, BOP_GATE2, BOP_ROT1, BOP_ROT2, BOP_MOVE, BOP_DIAG2Q, BOP_SCALE, BOP_MOMENTS, BOP_RDM1Q, BOP_GATE12 = range(10)

# name -> (restype, argtypes); mirrors include/bass_hip.h
SIGNATURES = {
    "bass_abi_version": (_I, []),
    "bass_last_error": (ctypes.c_char_p, []),
    "bass_workspace_bytes": (_SZ, [_I64]),
    "bass_read_stats": (_I, [_P, _SZ, ctypes.POINTER(BassStats), _P]),
    "bass_gate": (_I, [_P, _P, _I64, _I, _I, _I, _P, _I, _I, _I, _P, _P, _I64, ctypes.POINTER(_I64), _P, _SZ,
                       _P]),
    "bass_densify": (_I, [_P, _P, _I64, _P, _P, _I64, _P, _SZ, _P]),
    "bass_lift": (_I, [_P, _P, _I64, _I, _I, _P, _I, _P, _P, _I64, ctypes.POINTER(_I64), _P, _SZ, _P]),
    "bass_lift_counted": (_I, [_P, _P, _I64, _I, _I, _P, _I, _P, _P, _I64, ctypes.POINTER(_I64), _P, _SZ, _P]),
    "bass_set_count": (_I, [_P, _SZ, _I64, _P]),
    "bass_diag2q_counted": (_I, [_P, _P, _I64, _I, _I, _P, _P, _SZ, _P]),
    "bass_diag2q_run": (_I, [_P, _P, _I64, _I, _P, _P, _I, _P, _SZ, _P]),
    "bass_counted_read": (_I, [_P, _SZ, _P, _I64, _P, _P]),
    "bass_truncate": (_I, [_P, _P, _I64, _I64, _P, _P, _I64, _I, _I64, _I, _P, _I, _P, _SZ, _P]),
    "bass_truncate_noalloc": (_I, [_P, _P, _I64, _I64, _P, _P, _I64, _I, _P, _I, _P, _P]),
    "bass_lift_truncate": (_I, [_P, _P, _I64, _I, _I, _P, _I, _I64, _P, _P, _I64, _P, _P, _I64, _I, _I, _P, _I, _P, _SZ,
                                _P]),
    "bass_gate_truncate": (_I, [_P, _P, _I64, _I, _I, _P, _I, _I64, _P, _P, _I64, _I, _I, _P, _I, _P, _SZ, _P]),
    "bass_diag2q": (_I, [_P, _P, _I64, _I, _I, _P, _P]),
    "bass_moments": (_I, [_P, _I64, _P, _SZ, _P]),
    "bass_scale": (_I, [_P, _I64, _D, _P]),
    "bass_gather": (_I, [_P, _P, _P, _I64, _P, _P, _P]),
    "bass_rdm1q": (_I, [_P, _P, _I64, _I, _P, _P, _SZ, _P]),
    "bass_rdm1q_ordered": (_I, [_P, _P, _I64, _I, _P, _P, _SZ, _P]),
    "bass_rdm2q_ordered": (_I, [_P, _P, _I64, _I, _P, _P, _SZ, _P]),
    "bass_ht_clear": (_I, [_P, _P, _I, _P]),
    "bass_ht_insert": (_I, [_P, _P, _I, _P, _I64, _I64, _P, _SZ, _P]),
    "bass_ht_lookup": (_I, [_P, _P, _I, _P, _I64, _P, _P]),
    "bass_radix_sort": (_I, [_P, _P, _I64, _I, _P, _SZ, _P]),
    "bass_sort_digit_bits": (_I, [_I]),
    "bass_tune": (_I, [_I, _I, _I]),
    "bass_sort_plan": (_I, [_I, _P]),
    "bass_merge_sorted": (_I, [_P, _P, _I64, _P, _P, _P, _SZ, _P]),
    "bass_dense_scatter": (_I, [_P, _P, _I64, _P, _I, _P, _I, _P]),
    "bass_dense_gate": (_I, [_P, _I, _I, _I, _I, _P, _P]),
    "bass_dense_vdot": (_I, [_P, _P, _I64, _P, _P, _SZ, _P]),
    "bass_batch_step": (_I, [_P, _I, _I, _I, _I, _P, _P, _I64, _P, _SZ, _P]),
    "bass_batch_move": (_I, [_P, _I, _P, _SZ, _P]),
    "bass_batch_diag2q": (_I, [_P, _I, _P, _SZ, _P]),
    "bass_batch_scale": (_I, [_P, _I, _P, _SZ, _P]),
    "bass_batch_moments": (_I, [_P, _I, _P, _SZ, _P]),
    "bass_batch_rdm1q": (_I, [_P, _I, _P, _P, _SZ, _P]),
    "bass_batch_stats": (_I, [_P, _SZ, _I, _P, _P]),
    "bass_bplan_ws_bytes": (_SZ, [_I, _I64]),
    "bass_bplan_create": (_I, [_I, _I64, _P, _P, _P, _SZ, _I, ctypes.POINTER(_P)]),
    "bass_bplan_destroy": (_I, [_P]),
    "bass_bplan_run": (_I, [_P, _I, _P, _I, _P, _P]),
    "bass_timing": (_I, [_I]),
    "bass_timing_read": (_I, [ctypes.c_char_p, ctypes.POINTER(_I64), ctypes.POINTER(_D)]),
}


def timing(enable):
    """Reset and enable/disable per-operation HIP-event timing."""
    load().bass_timing(1 if enable else 0)


def timing_read(name):
    """(launch count, summed ms) of one timed operation."""
    c, t = _I64(0), _D(0.0)
    load().bass_timing_read(name.encode(), ctypes.byref(c), ctypes.byref(t))
    return int(c.value), float(t.value)

_lib = None


def load():
    """Load the library (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    # torch first: the library must bind to the HIP runtime torch loads (the
    # one that owns the device memory and streams it is handed).  Loaded
    # first, it would pull in a second runtime that sees no device.
    import torch  # noqa: F401

    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    # the committed per-size-class tuning table, if any (layout only, never results)
    from . import tuning

    table = tuning.load_table()
    if table is not None:
        tuning.apply(L, table)
    _lib = L
    return L


def last_error():
    return load().bass_last_error().decode(errors="replace")


def check(rc, what):
    if rc != BASS_OK:
        msg = f"{what}: {last_error()} (code {rc})"
        if rc == BASS_E_OVERFLOW:
            raise RuntimeError("Temporary buffer overflow: " + msg)
        raise BassError(msg)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def mat_arg(G):
    """Complex matrix -> contiguous float64 (re, im) pairs, row-major."""
    g = np.ascontiguousarray(np.asarray(G, dtype=np.complex128))
    buf = np.ascontiguousarray(g.view(np.float64).ravel())
    return buf, buf.ctypes.data_as(ctypes.c_void_p)
