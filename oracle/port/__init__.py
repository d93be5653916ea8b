"""ctypes binding of the oracle's C++ 1-qubit step (oracle/port/cpu_port.cpp)."""

import ctypes
import os

import numpy as np

_LIB = None
PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_build", "libbass_cpu.so")


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(PATH):
            raise ImportError(f"{PATH} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(PATH)
        P = ctypes.c_void_p
        L.cpu_step_1q.restype = ctypes.c_int64
        L.cpu_step_1q.argtypes = [P, P, ctypes.c_int64, ctypes.c_int, P, ctypes.c_int64, P, P, P, P]
        L.cpu_gate_1q.restype = ctypes.c_int64
        L.cpu_gate_1q.argtypes = [P, P, ctypes.c_int64, ctypes.c_int, P, P, P]
        L.cpu_radix_sort.restype = None
        L.cpu_radix_sort.argtypes = [P, P, ctypes.c_int64]
        _LIB = L
    return _LIB


def step_1q(x, a, n, q, G, k, tx, ta):
    """One 1-qubit gate + top-k truncation + renormalisation in place on
    x (uint64, room for 2n) / a (complex128, room for 2n); tx / ta scratch of
    2n.  Returns (new count, kept / total, count after the merge)."""
    L = load()
    g = np.ascontiguousarray(np.asarray(G, dtype=np.complex128)).view(np.float64)
    gf = ctypes.c_double(1.0)
    mg = ctypes.c_int64(0)
    m = L.cpu_step_1q(x.ctypes.data, a.ctypes.data, int(n), int(q), g.ctypes.data, int(k), tx.ctypes.data,
                      ta.ctypes.data, ctypes.byref(gf), ctypes.byref(mg))
    return int(m), gf.value, int(mg.value)


def gate_1q(x, a, n, q, G, xo, ao):
    """Emit, 16-bit radix sort and merge of a 1-qubit gate on x[:n] / a[:n]
    into xo / ao (room for 2n entries); no truncation.  Returns the merged
    count."""
    L = load()
    g = np.ascontiguousarray(np.asarray(G, dtype=np.complex128)).view(np.float64)
    return int(L.cpu_gate_1q(x.ctypes.data, a.ctypes.data, int(n), int(q), g.ctypes.data, xo.ctypes.data,
                             ao.ctypes.data))


def radix_sort(x, a, n):
    """The reference's 16-bit LSD radix sort of x[:n] (uint64) carrying a
    (complex128), in place."""
    load().cpu_radix_sort(x.ctypes.data, a.ctypes.data, int(n))
