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
# PQ_ORACLE_LIB points the checker at another build of pq_oracle.c: the
# mutation tests load copies with planted mistakes this way.
_LIB_OVERRIDE = os.environ.get("PQ_ORACLE_LIB")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


# Builds of pq_oracle.c: the checker (statement-order rounding) and the two
# builds the CPU baseline times (oracle/Makefile).
BUILDS = {"check": "liboracle.so", "o3": "liboracle_o3.so", "fast": "liboracle_fast.so"}


def build(force: bool = False) -> str:
    """Compile every build of pq_oracle.c with the committed Makefile."""
    if force or not all(os.path.exists(os.path.join(HERE, f)) for f in BUILDS.values()):
        args = ["make", "-s", "-C", HERE] + (["-B"] if force else []) + ["all"]
        subprocess.check_call(args)
    return LIB_PATH


_builds = {}


def lib(build_name: str = "check"):
    """The oracle library: the checker build, or one of the CPU-baseline
    builds ("o3", "fast"), which are timed, never used to check."""
    global _lib
    if build_name != "check":
        if build_name not in _builds:
            path = os.path.join(HERE, BUILDS[build_name])
            if not os.path.exists(path):
                build()
            L = ctypes.CDLL(path)
            i, vp = ctypes.c_int, ctypes.c_void_p
            for name, (res, args) in {"pqo_inv_create": (vp, [i, i, _dp, _dp, _dp]),
                                      "pqo_inv_destroy": (None, [vp]),
                                      "pqo_inv_run": (i, [vp, _dp, i, i, _dp, i, i])}.items():
                getattr(L, name).restype = res
                getattr(L, name).argtypes = args
            _builds[build_name] = L
        return _builds[build_name]
    if _lib is None:
        if _LIB_OVERRIDE:
            L = ctypes.CDLL(_LIB_OVERRIDE)
        else:
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
            "pqo_energy": (None, [i, i, i, _dp, _dp, _dp, _dp, _dp]),
            "pqo_phelm_create": (vp, [i, i, i, _dp]),
            "pqo_phelm_destroy": (None, [vp]),
            "pqo_phelm_solve": (None, [vp, _dp, i, i]),
            "pqo_chan_create": (vp, [i, i, i, _dp, _dp, _dp, i, d]),
            "pqo_chan_destroy": (None, [vp]),
            "pqo_chan_corr": (_dp, [vp]),
            "pqo_chan_inv_corr_a": (_dp, [vp]),
            "pqo_chan_psic": (_dp, [vp]),
            "pqo_zeta_boundary": (None, [i, i, i, _dp, i, d, _dp, _dp, _dp]),
            "pqo_chan_run": (i, [vp, _dp, i, i, _dp, i, i]),
            "pqo_zeta_step": (None, [i, i, d, _dp, _dp, _dp, _dp, d, d, _dp, _dp, _dp, _dp, _dp]),
            "pqo_ra_filter": (None, [i, i, _dp, _dp, _dp, _dp]),
            "pqo_psi_av_add": (None, [i, i, i, d, _dp, _dp]),
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


def energy(psi_data: np.ndarray, h, s2):
    """energy (qg3p_diagnostics.F90:778-863) of a data array (gi+2, gj+2, L)
    with halo; returns (ekin[L], epot)."""
    psi = f64(psi_data)
    gi, gj, L = psi.shape[0] - 2, psi.shape[1] - 2, psi.shape[2]
    h, s2 = f64(h), f64(np.resize(np.asarray(s2, dtype=np.float64), L))
    ekin = np.zeros(L)
    epot = ctypes.c_double()
    lib().pqo_energy(gi, gj, L, _p(h), _p(s2), _p(psi), _p(ekin), ctypes.byref(epot))
    return ekin, epot.value


def psi_av_add(psi_data: np.ndarray, psi_av: np.ndarray, dt: float) -> None:
    """psi_av (gi, gj, L) += dt * psi over all nodes (qg3.F90:672-681)."""
    psi = f64(psi_data)
    assert psi_av.flags.f_contiguous and psi_av.dtype == np.float64
    gi, gj, L = psi_av.shape
    lib().pqo_psi_av_add(gi, gj, L, dt, _p(psi), _p(psi_av))


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

    def __init__(self, g: int, ss, m2z, z2m, build_name: str = "check"):
        self.g = g
        self.ss, self.m2z, self.z2m = f64(ss), f64(m2z), f64(z2m)
        self.layers = len(self.ss)
        self._lib = lib(build_name)
        self.h = self._lib.pqo_inv_create(g, self.layers, _p(self.ss), _p(self.m2z), _p(self.z2m))
        if not self.h:
            raise ValueError("pqo_inv_create failed")

    def run(self, zeta: np.ndarray, psi: np.ndarray | None = None):
        """zeta: (>=g, >=g, L) Fortran array (the g x g block is used)."""
        g, L = self.g, self.layers
        if psi is None:
            psi = np.zeros((g, g, L), order="F")
        rc = self._lib.pqo_inv_run(self.h, _p(zeta), zeta.shape[0], zeta.shape[1],
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
            self._lib.pqo_inv_destroy(self.h)
            self.h = None


class PeriodicHelm:
    """helm on gi x gj nodes, walls at i = 1, gi, periodic in j (period gj)
    (qg3p_helmholtz.F90:423-610, 614-769, the non-PH1D path)."""

    def __init__(self, gi: int, gj: int, ss):
        self.ss = f64(np.atleast_1d(ss))
        self.gi, self.gj = gi, gj
        self.h = lib().pqo_phelm_create(gi, gj, len(self.ss), _p(self.ss))
        assert self.h

    def solve(self, zeta: np.ndarray) -> None:
        """In place on a Fortran-ordered (ld1, ld2, L) array."""
        assert zeta.flags.f_contiguous and zeta.dtype == np.float64
        z = zeta if zeta.ndim == 3 else zeta[:, :, None]
        lib().pqo_phelm_solve(self.h, _p(zeta), z.shape[0], z.shape[1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().pqo_phelm_destroy(self.h)
            self.h = None


def zeta_boundary(psi: np.ndarray, ss, bc_type: int, alpha: float):
    """(z_b, z_d) of a (gi, gj, L) modal field (qg3p_aux.F90:288-379)."""
    psi = f64(psi)
    gi, gj, L = psi.shape
    zb, zd = np.zeros(L), np.zeros(L)
    lib().pqo_zeta_boundary(gi, gj, L, _p(f64(ss)), bc_type, alpha, _p(psi), _p(zb), _p(zd))
    return zb, zd


class Channel:
    """pv_inversion for the periodic channel (qg3p_aux.F90:152-193, 233-379,
    452-479, 524-575)."""

    def __init__(self, gi: int, gj: int, ss, m2z, z2m, bc_type: int = 1, alpha: float = 0.0):
        self.ss, self.m2z, self.z2m = f64(ss), f64(m2z), f64(z2m)
        self.gi, self.gj, self.layers = gi, gj, len(self.ss)
        self.h = lib().pqo_chan_create(gi, gj, self.layers, _p(self.ss), _p(self.m2z),
                                       _p(self.z2m), bc_type, alpha)
        assert self.h

    def run(self, zeta: np.ndarray, psi: np.ndarray | None = None):
        """zeta, psi: Fortran-ordered (ld1, ld2, L) arrays whose leading
        gi x gj blocks are the views; returns (psi, status)."""
        zeta = f64(zeta)
        if psi is None:
            psi = np.zeros((self.gi, self.gj, self.layers), order="F")
        rc = lib().pqo_chan_run(self.h, _p(zeta), zeta.shape[0], zeta.shape[1], _p(psi),
                                psi.shape[0], psi.shape[1])
        return psi, rc

    def corr(self) -> np.ndarray:
        n = self.gi * self.gj * self.layers * 2
        return np.ctypeslib.as_array(lib().pqo_chan_corr(self.h), (n,)).reshape(
            (self.gi, self.gj, self.layers, 2), order="F").copy()

    def inv_corr_a(self) -> np.ndarray:
        n = 4 * max(self.layers - 1, 1)
        return np.ctypeslib.as_array(lib().pqo_chan_inv_corr_a(self.h), (n,)).reshape(
            (2, 2, -1), order="F").copy()

    def psic(self) -> np.ndarray:
        n = 2 * max(self.layers - 1, 1)
        return np.ctypeslib.as_array(lib().pqo_chan_psic(self.h), (n,)).reshape(
            (2, -1), order="F").copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().pqo_chan_destroy(self.h)
            self.h = None


def wind_field(g, tau, wind_shift=0.0, wind_tilt=0.0, wind_asym=1.0):
    fw = np.zeros((g, g), order="F")
    lib().pqo_wind_field(g, tau, wind_shift, wind_tilt, wind_asym, _p(fw))
    return fw


# --------------------------------------------------------------------------
# CABARET stages and the full step
def _g_L_node(a):
    return a.shape[0] - 2, (a.shape[2] if a.ndim == 3 else 1)


def cabaret_cfl(dt, ui, uj):
    g, L = ui.shape[0] - 3, ui.shape[2]
    return lib().pqo_cabaret_cfl(g, L, dt, _p(ui), _p(uj))


def cabaret_adapt_dt(target_cfl, max_dt, max_dt_increase, ui, uj, dt):
    g, L = ui.shape[0] - 3, ui.shape[2]
    d, a, u = ctypes.c_double(dt), ctypes.c_double(), ctypes.c_double()
    lib().pqo_cabaret_adapt_dt(g, L, target_cfl, max_dt, max_dt_increase, _p(ui), _p(uj),
                               ctypes.cast(ctypes.byref(d), _dp), ctypes.cast(ctypes.byref(a), _dp),
                               ctypes.cast(ctypes.byref(u), _dp))
    return d.value, a.value, u.value


def cabaret_predictor(dt, zeta_old, res, zeta):
    g, L = _g_L_node(zeta)
    lib().pqo_cabaret_predictor(g, L, dt, _p(zeta_old), _p(res), _p(zeta))


def cabaret_beta(dt, tw_a, beta, ui, zeta, res_beta):
    g, L = _g_L_node(zeta)
    lib().pqo_cabaret_beta(g, L, dt, tw_a, _p(f64(beta)), _p(ui), _p(zeta), _p(res_beta))


def cabaret_forcing(dt, fw, visc, visc_bot, rel, zeta):
    g, L = _g_L_node(zeta)
    lib().pqo_cabaret_forcing(g, L, dt, _p(fw), visc, visc_bot, _p(rel), _p(zeta))


def cabaret_cell_boundaries(tw_u, psi, ui, uj, uimo, ujmo):
    g, L = _g_L_node(psi)
    lib().pqo_cabaret_cell_boundaries(g, L, tw_u, _p(psi), _p(ui), _p(uj), _p(uimo), _p(ujmo))


def cabaret_cell_boundaries_velocity(psi, ui, uj):
    g, L = _g_L_node(psi)
    lib().pqo_cabaret_cell_boundaries_velocity(g, L, _p(psi), _p(ui), _p(uj))


def cabaret_extrapolator(dt, zeta, zeta_old, ui, uj, zfi, zfj):
    g, L = _g_L_node(zeta)
    lib().pqo_cabaret_extrapolator(g, L, dt, _p(zeta), _p(zeta_old), _p(ui), _p(uj), _p(zfi),
                                   _p(zfj))


def cabaret_flux(ui, uj, zfi, zfj, flux_i, flux_j):
    g, L = ui.shape[0] - 3, ui.shape[2]
    lib().pqo_cabaret_flux(g, L, _p(ui), _p(uj), _p(zfi), _p(zfj), _p(flux_i), _p(flux_j))


def cabaret_corrector(dt, flux_i, flux_j, res, zeta):
    g, L = _g_L_node(zeta)
    lib().pqo_cabaret_corrector(g, L, dt, _p(flux_i), _p(flux_j), _p(res), _p(zeta))


def rel_from_zeta_and_psi(s1, s2, bc_type, alpha, zeta, psi, rel):
    g, L = _g_L_node(zeta)
    lib().pqo_rel_from_zeta_and_psi(g, L, _p(f64(s1)), _p(f64(s2)), bc_type, alpha, _p(zeta),
                                    _p(psi), _p(rel))


def update_zeta_bc(s1, s2, rel, psi, zeta):
    g, L = _g_L_node(zeta)
    lib().pqo_update_zeta_bc(g, L, _p(f64(s1)), _p(f64(s2)), _p(rel), _p(psi), _p(zeta))


class State(ctypes.Structure):
    _fields_ = [("g", ctypes.c_int), ("layers", ctypes.c_int)] + [
        (n, _dp) for n in ("psi", "rel", "zeta", "zeta_old", "zeta_fi", "ui", "uimo", "flux_i",
                           "zeta_fj", "uj", "ujmo", "flux_j", "res", "res_beta", "force_wind")] + [
        (n, ctypes.c_double) for n in ("dt", "tweight_a", "tweight_u", "model_time")] + [
        ("psi_av", _dp), ("max_spin", ctypes.c_double), ("diag_time", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("beta", _dp), ("s1", _dp), ("s2", _dp), ("viscosity", ctypes.c_double),
                ("visc_bot", ctypes.c_double), ("bc_type", ctypes.c_int),
                ("alpha", ctypes.c_double), ("adapt_dt", ctypes.c_int),
                ("target_cfl", ctypes.c_double), ("max_dt", ctypes.c_double),
                ("max_dt_increase", ctypes.c_double)]


class Model:
    """The reference's CABARET time loop on one domain (qg3.F90:590-967),
    driving the oracle's stages; state arrays are Fortran-ordered numpy."""

    NODE = ("psi", "rel", "zeta", "zeta_old")
    FI = ("zeta_fi", "ui", "uimo", "flux_i")
    FJ = ("zeta_fj", "uj", "ujmo", "flux_j")

    def __init__(self, g, strat_s1, strat_s2, ss, m2z, z2m, beta, force_wind, *, viscosity,
                 visc_bot, dt, bc_type=1, alpha=0.0, adapt_dt=False, target_cfl=0.1,
                 max_dt=0.0, max_dt_increase=1.5):
        L = len(ss)
        self.g, self.layers = g, L
        self.a = {}
        for n in self.NODE:
            self.a[n] = np.zeros((g + 2, g + 2, L), order="F")
        for n in self.FI:
            self.a[n] = np.zeros((g + 3, g + 2, L), order="F")
        for n in self.FJ:
            self.a[n] = np.zeros((g + 2, g + 3, L), order="F")
        self.a["res"] = np.zeros((g, g, L), order="F")
        self.a["res_beta"] = np.zeros((g, g, L), order="F")
        self.a["force_wind"] = f64(force_wind)
        self.s1, self.s2, self.beta = f64(strat_s1), f64(strat_s2), f64(beta)
        self.inv = Inversion(g, ss, m2z, z2m)
        self.st = State(g, L, *[_p(self.a[n]) for n in (
            "psi", "rel", "zeta", "zeta_old", "zeta_fi", "ui", "uimo", "flux_i", "zeta_fj", "uj",
            "ujmo", "flux_j", "res", "res_beta", "force_wind")], dt, 0.5, 0.5, dt / 2.0)
        self.pr = Params(_p(self.beta), _p(self.s1), _p(self.s2), viscosity, visc_bot, bc_type,
                         alpha, int(adapt_dt), target_cfl, max_dt, max_dt_increase)
        L_ = lib()
        L_.pqo_step.restype = ctypes.c_int
        L_.pqo_step.argtypes = [ctypes.POINTER(State), ctypes.POINTER(Params), ctypes.c_void_p,
                                ctypes.c_int]

    def init(self, zeta_sdata):
        """START_PERTURB-style start (qg3.F90:393-425) from an interior field."""
        g, L = self.g, self.layers
        z = self.a["zeta"]
        z[:] = 0.0
        z[2:g, 2:g, :] = f64(zeta_sdata)[1:g - 1, 1:g - 1, :]
        off = 8 * (1 + (g + 2))  # element (2, 2): the sdata view
        rc = lib().pqo_inv_run(self.inv.h, ctypes.cast(z.ctypes.data + off, _dp), g + 2, g + 2,
                               ctypes.cast(self.a["psi"].ctypes.data + off, _dp), g + 2, g + 2)
        rel_from_zeta_and_psi(self.s1, self.s2, self.pr.bc_type, self.pr.alpha, z, self.a["psi"],
                              self.a["rel"])
        update_zeta_bc(self.s1, self.s2, self.a["rel"], self.a["psi"], z)
        self.a["zeta_old"][:] = z
        return rc

    def step(self, nsteps=1, final_partial=False):
        for k in range(nsteps):
            last = final_partial and k == nsteps - 1
            rc = lib().pqo_step(ctypes.byref(self.st), ctypes.byref(self.pr), self.inv.h,
                                0 if last else 1)
            if rc == 3:
                raise FloatingPointError("NaN psiC encountered")

    def scalars(self):
        return dict(dt=self.st.dt, tweight_a=self.st.tweight_a, tweight_u=self.st.tweight_u,
                    model_time=self.st.model_time)

    def average(self, max_spin=0.0):
        """Start the time mean of psi (qg3.F90:503, 672-681)."""
        g, L = self.g, self.layers
        self.a["psi_av"] = np.zeros((g, g, L), order="F")
        self.st.psi_av = _p(self.a["psi_av"])
        self.st.max_spin = max_spin
        self.st.diag_time = 0.0

    def psi_av(self):
        """psi_av / diagnostic_time (qg3.F90:886-898) and diagnostic_time."""
        return self.a["psi_av"] * (1.0 / self.st.diag_time), self.st.diag_time

    def energy(self, h):
        return energy(self.a["psi"], h, self.s2)


def zeta_step(T, beta, s1, s2, force_wind, viscosity, visc_bot, psi, rel, rel_p, src, dst):
    """dst = src + T F (zeta_centred / zeta_forward_euler, qg3p_aux.F90:1789-2202)."""
    g, L = _g_L_node(dst)
    lib().pqo_zeta_step(g, L, T, _p(f64(beta)), _p(f64(s1)), _p(f64(s2)), _p(f64(force_wind)),
                        viscosity, visc_bot, _p(psi), _p(rel), _p(rel_p), _p(src), _p(dst))


def ra_filter(phi_l, phi_m, phi_old_l, phi_old_m):
    """Robert-Asselin-Williams filter in place (qg3p_aux.F90:1716-1786)."""
    g, L = _g_L_node(phi_l)
    lib().pqo_ra_filter(g, L, _p(phi_l), _p(phi_m), _p(phi_old_l), _p(phi_old_m))


class LFState(ctypes.Structure):
    _fields_ = [("g", ctypes.c_int), ("layers", ctypes.c_int),
                ("psi", _dp * 2), ("rel", _dp * 2), ("zeta", _dp * 2), ("zeta_old", _dp * 2),
                ("force_wind", _dp), ("dt", ctypes.c_double), ("model_time", ctypes.c_double),
                ("leapfrog", ctypes.c_int)]


class LeapfrogModel:
    """The reference's time loop without CABARET (centred scheme with the
    Robert-Asselin-Williams filter, forward Euler start; qg3.F90:456-490,
    612-660) on one domain, driving the oracle's stages."""

    NAMES = ("psi", "rel", "zeta", "zeta_old")

    def __init__(self, g, strat_s1, strat_s2, ss, m2z, z2m, beta, force_wind, *, viscosity,
                 visc_bot, dt, bc_type=1, alpha=0.0):
        L = len(ss)
        self.g, self.layers = g, L
        self.a = {f"{n}{k}": np.zeros((g + 2, g + 2, L), order="F") for n in self.NAMES
                  for k in (1, 2)}
        self.a["force_wind"] = f64(force_wind)
        self.s1, self.s2, self.beta = f64(strat_s1), f64(strat_s2), f64(beta)
        self.inv = Inversion(g, ss, m2z, z2m)
        arr = lambda n: (_dp * 2)(_p(self.a[n + "1"]), _p(self.a[n + "2"]))
        self.st = LFState(g, L, arr("psi"), arr("rel"), arr("zeta"), arr("zeta_old"),
                          _p(self.a["force_wind"]), dt, 0.0, 1)
        self.pr = Params(_p(self.beta), _p(self.s1), _p(self.s2), viscosity, visc_bot, bc_type,
                         alpha, 0, 0.0, 0.0, 0.0)
        L_ = lib()
        for name in ("pqo_lf_init", "pqo_lf_step"):
            f = getattr(L_, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.POINTER(LFState), ctypes.POINTER(Params), ctypes.c_void_p]

    def init(self, zeta_sdata):
        g = self.g
        z = self.a["zeta1"]
        z[:] = 0.0
        z[2:g, 2:g, :] = f64(zeta_sdata)[1:g - 1, 1:g - 1, :]
        return lib().pqo_lf_init(ctypes.byref(self.st), ctypes.byref(self.pr), self.inv.h)

    def step(self, nsteps=1):
        for _ in range(nsteps):
            rc = lib().pqo_lf_step(ctypes.byref(self.st), ctypes.byref(self.pr), self.inv.h)
            if rc == 3:
                raise FloatingPointError("NaN psiC encountered")

    def scalars(self):
        return dict(dt=self.st.dt, model_time=self.st.model_time, leapfrog=self.st.leapfrog)
