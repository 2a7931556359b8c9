"""pequod_amd -- MI355X-native implementation of pequod's time-step hot path:
the modal Helmholtz (PV) inversion and the CABARET advection step, in
hand-written HIP behind the C ABI of include/pequod_hip.h."""
from ._lib import LIB_PATH, PqError, load  # noqa: F401

__all__ = ["load", "LIB_PATH", "PqError"]
