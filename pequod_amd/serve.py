"""Request serving on one GPU (pq_srv_* of the C ABI).

Members stay resident in HBM: field members (a device PV field and its psi)
and CABARET model members (FP64 state owned by the engine).  Requests -- the
inversion of a field member, or n time steps of a model member -- can be
submitted at any time; a scheduler thread in the library batches every
waiting inversion of a grid size into one ragged launch whenever a handle of
that size is free and replays model members' captured step graphs on their
own streams, so requests that arrive while others run join the next
launches (continuous batching).  Members are admitted by device footprint
under a cap.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, qg
from ._lib import PREC, PqError, check, load
from .strat import Stratification

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


class Server:
    FIELD, MODEL = 0, 1

    def __init__(self, precision: str = "fp32", cap_bytes: float = 200e9, max_batch: int = 16):
        h = _vp()
        check(load().pq_srv_create(ctypes.byref(h), PREC[precision], int(cap_bytes), max_batch))
        self._h = h
        self.precision = precision
        self._keep = {}

    def open_field(self, strat: Stratification, zeta, psi, g: int | None = None) -> int:
        """Field member with device tensors zeta / psi of shape (L, ld2, ld1)
        whose leading g x g blocks are the views."""
        L, ld2, ld1 = zeta.shape
        g = g or min(ld1, ld2)
        es = _lib.IO_BYTES[self.precision]
        qg.check_views([zeta], es, g, L, "zeta")
        qg.check_views([psi], es, g, L, "psi")
        ss = np.ascontiguousarray(strat.ss, dtype=np.float64)
        m2z = np.asfortranarray(strat.modes_to_z, dtype=np.float64)
        z2m = np.asfortranarray(strat.z_to_modes, dtype=np.float64)
        idx = ctypes.c_int()
        check(load().pq_srv_open_field(
            self._h, g, L, ss.ctypes.data_as(_dp), m2z.ctypes.data_as(_dp), z2m.ctypes.data_as(_dp),
            zeta.data_ptr(), zeta.stride(1), zeta.stride(0) // zeta.stride(1), psi.data_ptr(),
            psi.stride(1), psi.stride(0) // psi.stride(1), ctypes.byref(idx)))
        self._keep[idx.value] = (zeta, psi)
        return idx.value

    def open_model(self, g: int, strat: Stratification, beta, force_wind, zeta_sdata, **kw) -> int:
        """Model member (arguments as qg.Model; fixed dt), started from the
        Fortran-ordered (g, g, L) interior field zeta_sdata."""
        p, keep = qg.model_params(g, strat, beta, force_wind, **kw)
        z = np.asfortranarray(zeta_sdata, dtype=np.float64)
        idx = ctypes.c_int()
        check(load().pq_srv_open_model(self._h, ctypes.byref(p), z.ctypes.data_as(_dp),
                                       ctypes.byref(idx)))
        self._keep[idx.value] = keep
        return idx.value

    def model_get(self, member: int, name: str, g: int, layers: int) -> np.ndarray:
        """A model member's field (names as qg.Model.get)."""
        m = load().pq_srv_model(self._h, member)
        assert m, f"member {member} is not a model"
        a = np.zeros(qg.field_shape(name, g, layers), order="F")
        check(load().pq_model_get(m, qg.FIELDS[name], a.ctypes.data_as(_dp), a.size))
        return a

    def submit(self, member: int, kind: int, steps: int = 0) -> int:
        r = ctypes.c_longlong()
        check(load().pq_srv_submit(self._h, member, kind, steps, ctypes.byref(r)))
        return r.value

    def poll(self, request: int):
        """Latency [s] of a completed request, else None; raises PqError for
        a failed request (code PQ_ERR_NAN when its inversion gave a NaN
        psiC)."""
        lat = ctypes.c_double()
        rc = load().pq_srv_poll(self._h, request, ctypes.byref(lat))
        if rc < 0:
            raise PqError(-rc, load().pq_last_error().decode(errors="replace"))
        return lat.value if rc == 1 else None

    def drain(self, timeout: float = 600.0) -> None:
        check(load().pq_srv_drain(self._h, timeout))

    def stats(self, kind: int = -1, first: int = 0) -> dict:
        """Latencies of the completed requests of `kind` (-1: all) numbered
        `first` or later."""
        out = np.zeros(8)
        check(load().pq_srv_stats(self._h, kind, first, out.ctypes.data_as(_dp)))
        return dict(count=int(out[0]), mean_ms=out[1] * 1e3, p50_ms=out[2] * 1e3, p95_ms=out[3] * 1e3,
                    p99_ms=out[4] * 1e3, max_ms=out[5] * 1e3, completed_per_s=out[6],
                    failed=int(out[7]))

    @property
    def used_bytes(self) -> int:
        return load().pq_srv_used_bytes(self._h)

    def close_member(self, member: int) -> None:
        check(load().pq_srv_close(self._h, member))
        self._keep.pop(member, None)

    def close(self):
        if getattr(self, "_h", None):
            try:
                load().pq_srv_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    __del__ = close
