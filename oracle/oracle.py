"""ctypes front end of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package
(pequod_amd) never imports this module.

Arrays follow the Fortran layout of the reference (column-major, i fastest):
numpy arrays passed here are Fortran-ordered float64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so from pq_oracle.c with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i, d, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        sig = {
            "pqo_sint_create": (vp, [i]),
            "pqo_sint_destroy": (None, [vp]),
            "pqo_sint_apply": (None, [vp, _dp]),
            "pqo_helm_create": (vp, [i, i, _dp]),
            "pqo_helm_destroy": (None, [vp]),
            "pqo_helm_solve": (None, [vp, _dp, i, i]),
            "pqo_apply_dirichlet_bc": (None, [i, i, d, _dp, i, i]),
            "pqo_add_laplace_bc_rhs": (None, [i, i, d, _dp, i, i]),
            "pqo_mean": (d, [i, _dp, i]),
            "pqo_strat_modes": (i, [i, _dp, _dp, _dp, _dp, _dp]),
            "pqo_strat_l2": (None, [d, d, d, _dp, _dp]),
            "pqo_strat_l3": (None, [d, d, d, d, d, _dp, _dp, _dp, _dp]),
            "pqo_inv_create": (vp, [i, i, _dp, _dp, _dp]),
            "pqo_inv_destroy": (None, [vp]),
            "pqo_inv_run": (i, [vp, _dp, i, i, _dp, i, i]),
            "pqo_inv_corr": (_dp, [vp]),
            "pqo_inv_psic": (_dp, [vp]),
            "pqo_cabaret_cfl": (d, [i, i, d, _dp, _dp]),
            "pqo_cabaret_adapt_dt": (None, [i, i, d, d, d, _dp, _dp, _dp, _dp, _dp]),
            "pqo_cabaret_predictor": (None, [i, i, d, _dp, _dp, _dp]),
            "pqo_cabaret_beta": (None, [i, i, d, d, _dp, _dp, _dp, _dp]),
            "pqo_cabaret_forcing": (None, [i, i, d, _dp, d, d, _dp, _dp]),
            "pqo_cabaret_cell_boundaries_velocity": (None, [i, i, _dp, _dp, _dp]),
            "pqo_cabaret_cell_boundaries": (None, [i, i, d, _dp, _dp, _dp, _dp, _dp]),
            "pqo_cabaret_extrapolator": (None, [i, i, d, _dp, _dp, _dp, _dp, _dp, _dp]),
            "pqo_cabaret_flux": (None, [i, i, _dp, _dp, _dp, _dp, _dp, _dp]),
            "pqo_cabaret_corrector": (None, [i, i, d, _dp, _dp, _dp, _dp]),
            "pqo_rel_from_zeta_and_psi": (None, [i, i, _dp, _dp, i, d, _dp, _dp, _dp]),
            "pqo_apply_rel_bc": (None, [i, i, i, d, _dp, _dp]),
            "pqo_update_zeta_bc": (None, [i, i, _dp, _dp, _dp, _dp, _dp]),
            "pqo_wind_field": (None, [i, d, d, d, d, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1)
    return a.ctypes.data_as(_dp)


def f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


# --------------------------------------------------------------------------
def sint(x: np.ndarray) -> np.ndarray:
    """Unnormalised DST-I (FFTPACK sint) of a 1-D array."""
    x = np.array(x, dtype=np.float64)
    h = lib().pqo_sint_create(len(x))
    try:
        lib().pqo_sint_apply(h, _p(x))
    finally:
        lib().pqo_sint_destroy(h)
    return x


class Helm:
    """helm_cache + helm (qg3p_helmholtz.F90:292, 614)."""

    def __init__(self, g: int, ss):
        self.g = g
        self.ss = f64(ss)
        self.layers = len(self.ss)
        self.h = lib().pqo_helm_create(g, self.layers, _p(self.ss))
        if not self.h:
            raise ValueError("pqo_helm_create failed")

    def solve(self, zeta: np.ndarray) -> None:
        assert zeta.shape[0] >= self.g and zeta.shape[1] >= self.g
        lib().pqo_helm_solve(self.h, _p(zeta), zeta.shape[0], zeta.shape[1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().pqo_helm_destroy(self.h)
            self.h = None


def apply_dirichlet_bc(zeta: np.ndarray, bc: float, g: int | None = None):
    g = g or zeta.shape[0]
    L = zeta.shape[2] if zeta.ndim == 3 else 1
    lib().pqo_apply_dirichlet_bc(g, L, bc, _p(zeta), zeta.shape[0], zeta.shape[1])


def add_laplace_bc_rhs(rhs: np.ndarray, bc: float):
    g = rhs.shape[0]
    L = rhs.shape[2] if rhs.ndim == 3 else 1
    lib().pqo_add_laplace_bc_rhs(g, L, bc, _p(rhs), rhs.shape[0], rhs.shape[1])


def mean(psi: np.ndarray) -> float:
    psi = f64(psi)
    return lib().pqo_mean(psi.shape[0], _p(psi), psi.shape[0])


def strat_modes(s1, s2):
    s1, s2 = f64(s1), f64(s2)
    L = len(s1)
    ss = np.zeros(L)
    m2z = np.zeros((L, L), order="F")
    z2m = np.zeros((L, L), order="F")
    rc = lib().pqo_strat_modes(L, _p(s1), _p(s2), _p(ss), _p(m2z), _p(z2m))
    if rc != 0:
        raise ValueError(f"pqo_strat_modes failed ({rc})")
    return ss, m2z, z2m


def strat_l2(rd, h1, h2):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().pqo_strat_l2(rd, h1, h2, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def strat_l3(rd1, rd2, h1, h2, h3):
    o = [ctypes.c_double() for _ in range(4)]
    lib().pqo_strat_l3(rd1, rd2, h1, h2, h3, *[ctypes.byref(x) for x in o])
    return tuple(x.value for x in o)


class Inversion:
    """helmholtz_cache + pv_inversion (qg3p_aux.F90:152, 452)."""

    def __init__(self, g: int, ss, m2z, z2m):
        self.g = g
        self.ss, self.m2z, self.z2m = f64(ss), f64(m2z), f64(z2m)
        self.layers = len(self.ss)
        self.h = lib().pqo_inv_create(g, self.layers, _p(self.ss), _p(self.m2z), _p(self.z2m))
        if not self.h:
            raise ValueError("pqo_inv_create failed")

    def run(self, zeta: np.ndarray, psi: np.ndarray | None = None):
        """zeta: (>=g, >=g, L) Fortran array (the g x g block is used)."""
        g, L = self.g, self.layers
        if psi is None:
            psi = np.zeros((g, g, L), order="F")
        rc = lib().pqo_inv_run(self.h, _p(zeta), zeta.shape[0], zeta.shape[1],
                               _p(psi), psi.shape[0], psi.shape[1])
        return psi, rc

    def run_view(self, zeta_ptr, ldz1, ldz2, psi_ptr, ldp1, ldp2):
        return lib().pqo_inv_run(self.h, zeta_ptr, ldz1, ldz2, psi_ptr, ldp1, ldp2)

    @property
    def corr(self) -> np.ndarray:
        g, L = self.g, self.layers
        p = lib().pqo_inv_corr(self.h)
        return np.ctypeslib.as_array(p, shape=(g * g * L,)).reshape((g, g, L), order="F").copy()

    @property
    def psic(self) -> np.ndarray:
        p = lib().pqo_inv_psic(self.h)
        return np.ctypeslib.as_array(p, shape=(self.layers,)).copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().pqo_inv_destroy(self.h)
            self.h = None


def wind_field(g, tau, wind_shift=0.0, wind_tilt=0.0, wind_asym=1.0):
    fw = np.zeros((g, g), order="F")
    lib().pqo_wind_field(g, tau, wind_shift, wind_tilt, wind_asym, _p(fw))
    return fw
