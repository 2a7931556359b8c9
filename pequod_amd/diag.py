"""Diagnostics of a stream function on the device (pq_diag_* and
pq_psi_av_add_dev of the C ABI): energy(strat, psi, ekin, epot)
(qg3p_diagnostics.F90:778-863) and the time-mean accumulation
psi_av += dt * psi (qg3.F90:672-681).

Device fields are torch tensors of shape (L, ld2, ld1) whose leading
g x g block of every layer is the view: the sdata slice of a data array
held as a (L, G+2, G+2) tensor t is t[:, 1:, 1:]."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, load
from .strat import Stratification

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def _view(t, g: int):
    """(pointer, ld1, ld2) of the g x g view at the start of tensor t, whose
    last two dimensions are (ld2, ld1) with unit stride along ld1."""
    import torch
    if not t.is_cuda or t.dtype != torch.float64:
        raise ValueError("diagnostics take FP64 CUDA (HIP) tensors")
    if t.stride(-1) != 1:
        raise ValueError("the i index must have unit stride")
    ld1 = t.stride(-2)
    ld2 = t.stride(-3) // ld1 if t.dim() == 3 else t.shape[-2]
    if t.dim() == 3 and t.stride(-3) != ld1 * ld2:
        raise ValueError("layer stride must be a multiple of the column stride")
    if ld1 < g or ld2 < g or t.shape[-1] < g or t.shape[-2] < g:
        raise ValueError(f"view smaller than g={g}")
    return _vp(t.data_ptr()), ld1, ld2


class Energy:
    """Handle for the energy diagnostic of g x g x L fields."""

    def __init__(self, g: int, strat: Stratification | None = None, *, h=None, s2=None):
        h = np.ascontiguousarray(strat.h if h is None else h, dtype=np.float64)
        s2 = np.ascontiguousarray(strat.s2 if s2 is None else s2, dtype=np.float64)
        self.g, self.layers = g, len(h)
        d = _vp()
        check(load().pq_diag_create(ctypes.byref(d), g, self.layers, h.ctypes.data_as(_dp),
                                    s2.ctypes.data_as(_dp)))
        self._h = d

    def energy(self, psi: np.ndarray):
        """Host field: data(ld1, ld2, L) Fortran-ordered whose leading g x g
        block is the view.  Returns (ekin[L], epot)."""
        psi = np.asfortranarray(psi, dtype=np.float64)
        ekin = np.zeros(self.layers)
        epot = ctypes.c_double()
        check(load().pq_diag_energy(self._h, psi.ctypes.data_as(_dp), psi.shape[0], psi.shape[1],
                                    ekin.ctypes.data_as(_dp), ctypes.byref(epot)))
        return ekin, epot.value

    def energy_dev(self, psi, out, stream=None) -> None:
        """Device field psi (L, ld2, ld1); out: FP64 device tensor of L + 1
        values (ekin(1..L), epot).  Asynchronous on `stream` (a hipStream_t
        handle; default: the handle's own)."""
        p, ld1, ld2 = _view(psi, self.g)
        assert out.numel() >= self.layers + 1 and out.is_cuda
        check(load().pq_diag_energy_dev(self._h, p, ld1, ld2, _vp(out.data_ptr()),
                                        _vp(stream) if stream else None))

    @property
    def stream(self) -> int:
        return load().pq_diag_stream(self._h)

    def close(self):
        if getattr(self, "_h", None):
            try:
                load().pq_diag_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    __del__ = close


def psi_av_add(psi, psi_av, dt: float, g: int, stream=None) -> None:
    """psi_av += dt * psi over the g x g x L views of two FP64 device tensors
    (asynchronous on `stream`, default the null stream)."""
    p, ld1, ld2 = _view(psi, g)
    a, la1, la2 = _view(psi_av, g)
    L = psi.shape[0] if psi.dim() == 3 else 1
    check(load().pq_psi_av_add_dev(g, L, dt, p, ld1, ld2, a, la1, la2,
                                   _vp(stream) if stream else None))
