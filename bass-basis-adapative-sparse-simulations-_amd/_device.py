"""Device context: HIP stream, caller-owned workspace and state buffers.

PyTorch-ROCm only provides memory and the current stream; every computation
runs in libbass_hip.so.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L


def require_gpu():
    if not torch.cuda.is_available():
        raise L.BassError("no ROCm GPU visible: the sparse-state kernels run only on the device")


class Context:
    """One per device: workspace (grown on demand) and the stats block."""

    def __init__(self, device=None):
        require_gpu()
        self.lib = L.load()
        self.device = torch.device(device if device is not None else "cuda:0")
        self.ws = None
        self.ws_bytes = 0
        self.stats = L.BassStats()

    @property
    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def ensure_ws(self, max_entries):
        need = int(self.lib.bass_workspace_bytes(max(1, int(max_entries))))
        if need > self.ws_bytes:
            grow = max(need, int(self.ws_bytes * 1.5))
            self.ws = None
            # zeroed: the stats block keeps device error bits between calls
            # (cleared when read), so a new workspace must not start with
            # leftover bytes there
            self.ws = torch.zeros(grow, dtype=torch.uint8, device=self.device)
            self.ws_bytes = grow

    @property
    def wsp(self):
        return ctypes.c_void_p(self.ws.data_ptr())

    def read_stats(self, what="read_stats"):
        rc = self.lib.bass_read_stats(self.wsp, self.ws_bytes, ctypes.byref(self.stats), self.stream)
        L.check(rc, what)
        s = self.stats
        return _Stats(s.nnz, s.total, s.kept, s.norm, s.sum2, s.sum4, s.n_in, s.flags)

    def synchronize(self):
        torch.cuda.current_stream(self.device).synchronize()


class _Stats:
    __slots__ = ("nnz", "total", "kept", "norm", "sum2", "sum4", "n_in", "flags")

    def __init__(self, nnz, total, kept, norm, sum2, sum4, n_in=0, flags=0):
        self.n_in = int(n_in)
        self.flags = int(flags)
        self.nnz = int(nnz)
        self.total = float(total)
        self.kept = float(kept)
        self.norm = float(norm)
        self.sum2 = float(sum2)
        self.sum4 = float(sum4)


_CONTEXTS = {}


def context(device=None):
    key = str(torch.device(device if device is not None else "cuda:0"))
    c = _CONTEXTS.get(key)
    if c is None:
        c = Context(key)
        _CONTEXTS[key] = c
    return c


class Buffers:
    """A (keys, amplitudes) pair of device buffers with a capacity."""

    __slots__ = ("x", "a", "cap", "device")

    def __init__(self, cap, device):
        self.cap = max(1, int(cap))
        self.device = device
        self.x = torch.empty(self.cap, dtype=torch.int64, device=device)
        self.a = torch.empty(self.cap, dtype=torch.complex128, device=device)

    def ensure(self, cap, keep=0):
        if cap <= self.cap:
            return
        new = max(int(cap), 2 * self.cap)
        x = torch.empty(new, dtype=torch.int64, device=self.device)
        a = torch.empty(new, dtype=torch.complex128, device=self.device)
        if keep:
            x[:keep].copy_(self.x[:keep])
            a[:keep].copy_(self.a[:keep])
        self.x, self.a, self.cap = x, a, new

    @property
    def xp(self):
        return ctypes.c_void_p(self.x.data_ptr())

    @property
    def ap(self):
        return ctypes.c_void_p(self.a.data_ptr())


def to_host_keys(t, n):
    return t[:n].cpu().numpy().view(np.uint64)


def to_host_amps(t, n):
    return t[:n].cpu().numpy()
