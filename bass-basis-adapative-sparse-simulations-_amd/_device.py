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
        self._dev_index = self.device.index if self.device.index is not None else 0
        # the raw handle of torch's current stream without building a Stream
        # object (several microseconds per call otherwise)
        self._raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        self.ws = None
        self.ws_bytes = 0
        self._wsp = None
        # bass_workspace_bytes(n) = fixed + per * n (n >= 1): read once
        b1 = int(self.lib.bass_workspace_bytes(1))
        self._ws_per = int(self.lib.bass_workspace_bytes(2)) - b1
        self._ws_fixed = b1 - self._ws_per
        self.stats = L.BassStats()

    @property
    def stream(self):
        """torch's current stream on this device (the caller's stream)."""
        if self._raw_stream is not None:
            return ctypes.c_void_p(self._raw_stream(self._dev_index))
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def ensure_ws(self, max_entries, keep=False):
        """Grow the workspace to hold calls of max_entries entries.  keep=True:
        the workspace holds a preceding call's results that this call reads (a
        gate's histogram and per-tile sums), so it must not be replaced."""
        need = self._ws_fixed + self._ws_per * max(1, int(max_entries))
        if need > self.ws_bytes:
            if keep:
                raise RuntimeError(f"workspace of {self.ws_bytes} bytes holds the preceding gate's histogram; "
                                   f"this call needs {need}: size it before the gate")
            grow = max(need, int(self.ws_bytes * 1.5))
            self.ws = None
            # zeroed: the stats block keeps device error bits between calls
            # (cleared when read), so a new workspace must not start with
            # leftover bytes there
            self.ws = torch.zeros(grow, dtype=torch.uint8, device=self.device)
            self.ws_bytes = grow
            self._wsp = ctypes.c_void_p(self.ws.data_ptr())

    @property
    def wsp(self):
        return self._wsp

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

    __slots__ = ("x", "a", "cap", "device", "_xt", "_xp", "_at", "_ap")

    def __init__(self, cap, device):
        self.cap = max(1, int(cap))
        self.device = device
        self.x = torch.empty(self.cap, dtype=torch.int64, device=device)
        self.a = torch.empty(self.cap, dtype=torch.complex128, device=device)
        self._xt = self._at = None

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

    # device pointers, rebuilt only when a tensor is replaced
    @property
    def xp(self):
        if self._xt is not self.x:
            self._xt, self._xp = self.x, ctypes.c_void_p(self.x.data_ptr())
        return self._xp

    @property
    def ap(self):
        if self._at is not self.a:
            self._at, self._ap = self.a, ctypes.c_void_p(self.a.data_ptr())
        return self._ap


def to_host_keys(t, n):
    return t[:n].cpu().numpy().view(np.uint64)


def to_host_amps(t, n):
    return t[:n].cpu().numpy()
