"""MI355X (gfx950) sparse-state core of BASS (basis-adaptive sparse simulation).

Layout mirrors the reference's src/ tree for the hot path:
  core/         gates, SparseState + gate/sort/merge kernels, truncation kernels
  simulation/   BASS driver (+ hash, rotation, RDM kernels), FixedBasisSimulator
  utils/        circuit generators
  benchmarking/ fidelity and participation-ratio metrics
  csrc/         HIP kernels and the C ABI (include/bass_hip.h) -> libbass_hip.so

The package directory name is not a Python identifier; import it with
importlib.import_module("bass-basis-adapative-sparse-simulations-_amd").
"""

from . import _lib
from ._lib import LIB_PATH, BassError

__all__ = ["BassError", "LIB_PATH", "load_library", "BASS", "FixedBasisSimulator", "SparseState"]


def load_library():
    """Load libbass_hip.so (ImportError if it has not been built)."""
    return _lib.load()


def __getattr__(name):
    if name == "BASS":
        from .simulation.bass_simulator import BASS
        return BASS
    if name == "FixedBasisSimulator":
        from .simulation.simulator import FixedBasisSimulator
        return FixedBasisSimulator
    if name == "SparseState":
        from .core.sparse_state import SparseState
        return SparseState
    raise AttributeError(name)
