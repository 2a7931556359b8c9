"""Ensembles of independent members of mixed grid sizes, layer counts and
stratifications on one GPU (pq_ens_* of the C ABI).

Members of the same size class (grid, leading dimensions and transform
precision) share one ragged batched inversion handle with per-member layer
counts and stratifications; the classes run concurrently, and the whole
inversion -- or one time step of all CABARET model members -- can be
captured into a HIP graph and replayed.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, qg
from ._lib import PREC, check, load
from .strat import Stratification

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


def _arrays(strat: Stratification):
    return (np.ascontiguousarray(strat.ss, dtype=np.float64),
            np.asfortranarray(strat.modes_to_z, dtype=np.float64),
            np.asfortranarray(strat.z_to_modes, dtype=np.float64))


class Ensemble:
    """pq_ens: add members (device arrays or CABARET models), finalize, then
    invert() / step(); capture() turns either into a replayed HIP graph."""

    INVERT, STEP = 0, 1

    def __init__(self, precision: str = "fp32"):
        h = _vp()
        check(load().pq_ens_create(ctypes.byref(h), PREC[precision]))
        self._h = h
        self.precision = precision
        self._keep = []
        self.members = []
        self.finalized = False

    def add(self, strat: Stratification, zeta, psi, g: int | None = None) -> int:
        """Member with device tensors zeta / psi of shape (L, ld2, ld1) whose
        leading g x g blocks are the views (g defaults to the tensor's
        smaller trailing extent)."""
        L, ld2, ld1 = zeta.shape
        if g is None:
            g = min(ld1, ld2)
        # the element type of every precision's fields (fp32auto: FP32 storage)
        es = _lib.IO_BYTES[self.precision]
        qg.check_views([zeta], es, g, L, "zeta")
        qg.check_views([psi], es, g, L, "psi")
        ss, m2z, z2m = _arrays(strat)
        self._keep += [ss, m2z, z2m, zeta, psi]
        idx = ctypes.c_int()
        zl1, zl2 = zeta.stride(1), zeta.stride(0) // zeta.stride(1)
        pl1, pl2 = psi.stride(1), psi.stride(0) // psi.stride(1)
        check(load().pq_ens_add(self._h, g, L, ss.ctypes.data_as(_dp), m2z.ctypes.data_as(_dp),
                                z2m.ctypes.data_as(_dp), zeta.data_ptr(), zl1, zl2, psi.data_ptr(),
                                pl1, pl2, ctypes.byref(idx)))
        self.members.append(("field", g, L))
        return idx.value

    def add_model(self, model, strat: Stratification) -> int:
        """A CABARET model member (fixed dt); stepped by step()."""
        ss, m2z, z2m = _arrays(strat)
        self._keep += [ss, m2z, z2m, model]
        idx = ctypes.c_int()
        check(load().pq_ens_add_model(self._h, model._h, ss.ctypes.data_as(_dp), m2z.ctypes.data_as(_dp),
                                      z2m.ctypes.data_as(_dp), ctypes.byref(idx)))
        self.members.append(("model", model.g, model.layers))
        return idx.value

    def finalize(self, tiles="auto") -> None:
        """Builds the size classes.  tiles: "auto" applies the autotuner's
        table per class, None keeps the built-in defaults."""
        check(load().pq_ens_finalize(self._h))
        self.finalized = True
        if tiles == "auto":
            from . import autotune
            for k in range(self.class_count):
                h = load().pq_ens_class_handle(self._h, k)
                g, L, n = self.class_shape(k)
                t = autotune.lookup(g, L, self.class_precision(k), n)
                if t:
                    load().pq_inv_set_tiles(h, *t)

    @property
    def class_count(self) -> int:
        return load().pq_ens_class_count(self._h)

    def class_precision(self, k: int) -> str:
        """Precision of size class k (fp32auto resolved per class)."""
        return _lib.PREC_NAME[load().pq_ens_class_precision(self._h, k)]

    def member_class(self, member: int):
        c, s = ctypes.c_int(), ctypes.c_int()
        check(load().pq_ens_member_class(self._h, member, ctypes.byref(c), ctypes.byref(s)))
        return c.value, s.value

    def class_shape(self, k: int):
        """(g, largest layer count, members) of size class k."""
        ms = [i for i in range(len(self.members)) if self.member_class(i)[0] == k]
        g = self.members[ms[0]][1]
        return g, max(self.members[i][2] for i in ms), len(ms)

    def invert(self) -> None:
        check(load().pq_ens_invert(self._h))

    def step(self, nsteps: int = 1) -> None:
        check(load().pq_ens_step(self._h, nsteps))

    def capture(self, what: int) -> None:
        check(load().pq_ens_capture(self._h, what))

    def synchronize(self) -> None:
        check(load().pq_ens_synchronize(self._h))

    def status(self) -> int:
        return load().pq_ens_status(self._h)

    @property
    def stream(self) -> int:
        return load().pq_ens_stream(self._h)

    def close(self):
        if getattr(self, "_h", None):
            try:
                load().pq_ens_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    __del__ = close
