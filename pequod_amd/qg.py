"""Host-side mirror of the reference's Helmholtz / PV-inversion / CABARET
interface, over the C ABI of libpequod_hip.so.

Device arrays are torch tensors used only as device memory.  A Fortran
array data(ii, jj, L) is held as a C-contiguous tensor of shape (L, jj, ii),
which has the same memory layout (i fastest).  Shapes on one domain of G
nodes per side: node fields (L, G+2, G+2); i-faces (L, G+2, G+3); j-faces
(L, G+3, G+2); residuals (L, G, G); wind (G, G).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import PREC, check, load
from .strat import Stratification

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def _ptr(t) -> int:
    return int(t.data_ptr())


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


PQ_STREAM_NULL = ctypes.c_void_p(-1)


def _stream_of(t):
    """torch's current stream on t's device as a C-ABI stream argument (the
    default stream has handle 0, which the ABI spells PQ_STREAM_NULL)."""
    import torch
    h = torch.cuda.current_stream(t.device).cuda_stream
    return ctypes.c_void_p(h) if h else PQ_STREAM_NULL


def check_views(tensors, elem_bytes: int, g: int, layers: int, what: str,
                require_cuda: bool = True) -> None:
    """Checks a batch of device views before their raw pointers and the
    first view's strides go to the C ABI, which reads and writes g x g x
    layers elements of elem_bytes each through every pointer with those
    strides: each tensor has shape (L, >= g, >= g) with L == layers, the
    element size of the handle's precision, unit innermost stride and the
    strides and device of the first.  Raises ValueError otherwise."""
    if len(tensors) == 0:
        raise ValueError(f"{what}: empty batch")
    t0 = tensors[0]
    for k, t in enumerate(tensors):
        name = f"{what}[{k}]"
        if require_cuda and not t.is_cuda:
            raise ValueError(f"{name}: not a device (HIP) tensor")
        if t.dim() != 3:
            raise ValueError(f"{name}: {t.dim()}-d tensor; views are (L, ld2, ld1)")
        if t.element_size() != elem_bytes:
            raise ValueError(f"{name}: {t.element_size()}-byte elements, the handle's precision "
                             f"stores {elem_bytes}-byte values")
        if t.shape[0] != layers or t.shape[1] < g or t.shape[2] < g:
            raise ValueError(f"{name}: shape {tuple(t.shape)} does not hold a {g} x {g} x {layers} view")
        if t.stride(2) != 1 or t.stride(0) % t.stride(1) != 0:
            raise ValueError(f"{name}: strides {t.stride()} are not a column-major (ld1, ld2) view")
        if t.stride() != t0.stride():
            raise ValueError(f"{name}: strides {t.stride()} differ from the batch's first {t0.stride()}")
        if t.device != t0.device:
            raise ValueError(f"{name}: on {t.device}, the batch's first on {t0.device}")


def _require_cuda(*ts):
    for t in ts:
        if not t.is_cuda:
            raise ValueError("device entry points need CUDA (HIP) tensors")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


# ----------------------------------------------------------------- helm
class HelmCache:
    """allocate(helm_c, decomp, ss) / helm(helm_c, zeta) / deallocate(helm_c)
    (qg3p_helmholtz.F90:292, 614, 389)."""

    def __init__(self, ii: int, jj: int, ss, precision: str = "fp64", periodic: bool = False):
        """periodic: the channel's decomposition (periodic in j, jj = 2 (ii - 1))."""
        self.ss = np.ascontiguousarray(ss, dtype=np.float64)
        self.g = ii
        self.layers = len(self.ss)
        h = _vp()
        create = load().pq_helm_create_periodic if periodic else load().pq_helm_create
        check(create(ctypes.byref(h), ii, jj, self.layers, _np_ptr(self.ss), PREC[precision]))
        self._h = h
        self.precision = precision

    def helm(self, zeta) -> None:
        """In place.  zeta: numpy (Fortran order, ii x jj x L) or a device
        tensor of shape (L, jj, ii)."""
        if isinstance(zeta, np.ndarray):
            assert zeta.dtype == np.float64 and zeta.flags.f_contiguous
            check(load().pq_helm_solve(self._h, _np_ptr(zeta), zeta.shape[0], zeta.shape[1]))
        else:
            _require_cuda(zeta)
            L, jj, ii = zeta.shape
            check(load().pq_helm_solve_dev(self._h, _ptr(zeta), ii, jj, _stream_of(zeta)))

    def deallocate(self):
        if getattr(self, "_h", None):
            try:
                load().pq_helm_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    __del__ = deallocate


# ------------------------------------------------------- PV inversion
class HelmholtzCache:
    """helmholtz_cache: allocate(hcache, decomp, strat, vbc)
    (qg3p_aux.F90:152-193) for a closed basin of g x g nodes; serves up to
    batch_max members per call."""

    def __init__(self, g: int, strat: Stratification, precision: str = "fp64",
                 batch_max: int = 1, tiles="auto"):
        self.g = g
        self.strat = strat
        self.layers = strat.layers
        self.batch_max = batch_max
        h = _vp()
        ss = np.ascontiguousarray(strat.ss, dtype=np.float64)
        m2z = np.asfortranarray(strat.modes_to_z, dtype=np.float64)
        z2m = np.asfortranarray(strat.z_to_modes, dtype=np.float64)
        if precision == "fp32auto":
            precision = _lib.PREC_NAME[load().pq_prec_fp32_auto(g, self.layers, _np_ptr(m2z), _np_ptr(z2m))]
        self.precision = precision
        check(load().pq_inv_create(ctypes.byref(h), g, self.layers, _np_ptr(ss), _np_ptr(m2z),
                                   _np_ptr(z2m), PREC[precision], batch_max))
        self._h = h
        # tile shapes: "auto" = the autotuner's table (built-in defaults when
        # the size class has no entry), None = built-in defaults, or (R, C)
        if tiles == "auto":
            from . import autotune
            autotune.apply(self)
        elif tiles is not None and not self.set_tiles(*tiles):
            raise ValueError(f"tile shape {tiles} not supported for g={g}, L={self.layers}")

    @property
    def stream(self) -> int:
        return load().pq_inv_stream(self._h)

    def run_host(self, zeta: np.ndarray, psi: np.ndarray) -> int:
        """Host drop-in: zeta, psi Fortran-ordered float64 views."""
        for a in (zeta, psi):
            assert a.dtype == np.float64 and a.flags.f_contiguous
        return check(load().pq_inv_run(self._h, _np_ptr(zeta), zeta.shape[0], zeta.shape[1],
                                       _np_ptr(psi), psi.shape[0], psi.shape[1]), allow_warn=True)

    def run_dev(self, zetas, psis, stream=None) -> None:
        """zetas / psis: lists of device tensors of shape (L, ld2, ld1) whose
        leading g x g blocks are the views (or views of such tensors offset
        to element (1, 1)), with the element type of the handle's precision
        and the same strides throughout each list."""
        b = len(zetas)
        if len(psis) != b or b > self.batch_max:
            raise ValueError(f"{b} zeta and {len(psis)} psi views for a handle of batch {self.batch_max}")
        es = _lib.IO_BYTES[self.precision]
        check_views(zetas, es, self.g, self.layers, "zeta")
        check_views(psis, es, self.g, self.layers, "psi")
        zt = (_vp * b)(*[_ptr(z) for z in zetas])
        pt = (_vp * b)(*[_ptr(p) for p in psis])
        z0, p0 = zetas[0], psis[0]
        ldz1, ldz2 = z0.stride(1), z0.stride(0) // z0.stride(1)
        ldp1, ldp2 = p0.stride(1), p0.stride(0) // p0.stride(1)
        s = _vp(stream) if stream is not None else _stream_of(z0)
        check(load().pq_inv_run_dev(self._h, b, zt, ldz1, ldz2, pt, ldp1, ldp2, s))

    def pass_times(self, zetas, psis, reps: int = 20):
        """Mean duration [ms] of the row-forward, column and row-inverse
        passes, each timed alone between HIP events on the handle's stream."""
        b = len(zetas)
        zt = (_vp * b)(*[_ptr(z) for z in zetas])
        pt = (_vp * b)(*[_ptr(p) for p in psis])
        z0, p0 = zetas[0], psis[0]
        ldz1, ldz2 = z0.stride(1), z0.stride(0) // z0.stride(1)
        ldp1, ldp2 = p0.stride(1), p0.stride(0) // p0.stride(1)
        ms = (ctypes.c_float * 3)()
        check(load().pq_inv_pass_times(self._h, b, zt, ldz1, ldz2, pt, ldp1, ldp2, reps, ms),
              allow_warn=True)
        return list(ms)

    def set_tiles(self, rows: int, cols: int) -> bool:
        """Select the tile shape (rows per row-pass tile, columns per
        column-pass workgroup); False (handle unchanged) if not supported."""
        return load().pq_inv_set_tiles(self._h, rows, cols) == _lib.PQ_OK

    def tiles(self):
        r, c = ctypes.c_int(), ctypes.c_int()
        check(load().pq_inv_get_tiles(self._h, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def status(self):
        psic = np.zeros(self.layers * self.batch_max)
        rc = load().pq_inv_status(self._h, _np_ptr(psic))
        return rc, psic.reshape(self.batch_max, self.layers)

    def deallocate(self):
        if getattr(self, "_h", None):
            try:
                load().pq_inv_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    __del__ = deallocate


class ChannelCache(HelmholtzCache):
    """helmholtz_cache of the periodic channel (qg3p_aux.F90:152-193,
    233-286): gi x gj nodes with gj = 2 (gi - 1), walls at i = 1 and i = gi,
    periodic in j; the baroclinic modes get the momentum-conserving
    correction for the velocity boundary condition bc_type (BC_PARTIAL_SLIP
    with alpha, BC_NO_SLIP, BC_FREE_SLIP)."""

    def __init__(self, gi: int, strat: Stratification, bc_type: int = 1, alpha: float = 0.0,
                 precision: str = "fp64", batch_max: int = 1, tiles=None):
        self.g = gi
        self.gj = 2 * (gi - 1)
        self.strat = strat
        self.layers = strat.layers
        self.batch_max = batch_max
        ss = np.ascontiguousarray(strat.ss, dtype=np.float64)
        m2z = np.asfortranarray(strat.modes_to_z, dtype=np.float64)
        z2m = np.asfortranarray(strat.z_to_modes, dtype=np.float64)
        if precision == "fp32auto":
            precision = _lib.PREC_NAME[load().pq_prec_fp32_choice(self.layers, _np_ptr(m2z), _np_ptr(z2m))]
        self.precision = precision
        h = _vp()
        check(load().pq_chan_create(ctypes.byref(h), gi, self.gj, self.layers, _np_ptr(ss),
                                    _np_ptr(m2z), _np_ptr(z2m), bc_type, alpha, PREC[precision],
                                    batch_max))
        self._h = h
        if tiles is not None and not self.set_tiles(*tiles):
            raise ValueError(f"tile shape {tiles} not supported for gi={gi}, L={self.layers}")

    def status(self):
        """(status, psiC pairs (batch, L, 2); the barotropic mode's are 0)."""
        psic = np.zeros(2 * self.layers * self.batch_max)
        rc = load().pq_inv_status(self._h, _np_ptr(psic))
        return rc, psic.reshape(self.batch_max, self.layers, 2)


def pv_inversion(zeta, hcache: HelmholtzCache, psi) -> None:
    """pv_inversion(strat, vbc, zeta, hcache, psi) (qg3p_aux.F90:452-479) on
    device node fields of shape (L, G+2, G+2): reads zeta's owned nodes
    (sdata) and writes psi's owned nodes."""
    _require_cuda(zeta, psi)
    zs = zeta[:, 1:, 1:]
    ps = psi[:, 1:, 1:]
    hcache.run_dev([zs], [ps])


# ------------------------------------------------------------ CABARET
def _L(t):
    return t.shape[0]


def cabaret_cfl(dt, ui, uj) -> float:
    """cabaret_cfl (qg3p_aux.F90:920-943)."""
    _require_cuda(ui, uj)
    g = ui.shape[2] - 3
    out = ctypes.c_double()
    check(load().pq_cabaret_cfl_dev(g, _L(ui), dt, _ptr(ui), _ptr(uj), ctypes.byref(out),
                                    _stream_of(ui)))
    return out.value


def cabaret_predictor(dt, zeta_old, res, zeta):
    g = zeta.shape[2] - 2
    check(load().pq_cabaret_predictor_dev(g, _L(zeta), dt, _ptr(zeta_old), _ptr(res), _ptr(zeta),
                                          _stream_of(zeta)))


def cabaret_beta(dt, tweight_a, beta, ui, zeta, res_beta):
    g = zeta.shape[2] - 2
    check(load().pq_cabaret_beta_dev(g, _L(zeta), dt, tweight_a, _ptr(beta), _ptr(ui), _ptr(zeta),
                                     _ptr(res_beta), _stream_of(zeta)))


def cabaret_forcing(dt, force_wind, viscosity, visc_bot, rel, zeta):
    g = zeta.shape[2] - 2
    check(load().pq_cabaret_forcing_dev(g, _L(zeta), dt, _ptr(force_wind), viscosity, visc_bot,
                                        _ptr(rel), _ptr(zeta), _stream_of(zeta)))


def cabaret_cell_boundaries(tweight_u, psi, ui, uj, uimo, ujmo):
    g = psi.shape[2] - 2
    check(load().pq_cabaret_cell_boundaries_dev(g, _L(psi), tweight_u, _ptr(psi), _ptr(ui),
                                                _ptr(uj), _ptr(uimo), _ptr(ujmo), _stream_of(psi)))


def cabaret_cell_boundaries_velocity(psi, ui, uj):
    g = psi.shape[2] - 2
    check(load().pq_cabaret_cell_boundaries_velocity_dev(g, _L(psi), _ptr(psi), _ptr(ui), _ptr(uj),
                                                         _stream_of(psi)))


def cabaret_extrapolator(dt, zeta, zeta_old, ui, uj, zeta_fi, zeta_fj):
    g = zeta.shape[2] - 2
    check(load().pq_cabaret_extrapolator_dev(g, _L(zeta), dt, _ptr(zeta), _ptr(zeta_old), _ptr(ui),
                                             _ptr(uj), _ptr(zeta_fi), _ptr(zeta_fj), None,
                                             _stream_of(zeta)))


def cabaret_flux(ui, uj, zeta_fi, zeta_fj, flux_i, flux_j):
    g = ui.shape[2] - 3
    check(load().pq_cabaret_flux_dev(g, _L(ui), _ptr(ui), _ptr(uj), _ptr(zeta_fi), _ptr(zeta_fj),
                                     _ptr(flux_i), _ptr(flux_j), _stream_of(ui)))


def cabaret_corrector(dt, flux_i, flux_j, res, zeta):
    g = zeta.shape[2] - 2
    check(load().pq_cabaret_corrector_dev(g, _L(zeta), dt, _ptr(flux_i), _ptr(flux_j), _ptr(res),
                                          _ptr(zeta), _stream_of(zeta)))


BC_PARTIAL_SLIP, BC_NO_SLIP, BC_FREE_SLIP = 1, 2, 3


def rel_from_zeta_and_psi(s1, s2, bc_type, alpha, zeta, psi, rel):
    g = zeta.shape[2] - 2
    check(load().pq_rel_from_zeta_and_psi_dev(g, _L(zeta), _ptr(s1), _ptr(s2), bc_type, alpha,
                                              _ptr(zeta), _ptr(psi), _ptr(rel), _stream_of(zeta)))


def update_zeta_bc(s1, s2, rel, psi, zeta):
    g = zeta.shape[2] - 2
    check(load().pq_update_zeta_bc_dev(g, _L(zeta), _ptr(s1), _ptr(s2), _ptr(rel), _ptr(psi),
                                       _ptr(zeta), _stream_of(zeta)))


def zeta_step(T, beta, s1, s2, force_wind, viscosity, visc_bot, psi, rel, rel_p, src, dst):
    """dst = src + T F on the interior (zeta_centred / zeta_forward_euler,
    qg3p_aux.F90:1789-2202); node fields (L, G+2, G+2), force_wind (G, G)
    as a (G, G) tensor in Fortran order transposed (j, i)."""
    g = dst.shape[2] - 2
    check(load().pq_zeta_step_dev(g, _L(dst), T, _ptr(beta), _ptr(s1), _ptr(s2), _ptr(force_wind),
                                  viscosity, visc_bot, _ptr(psi), _ptr(rel), _ptr(rel_p), _ptr(src),
                                  _ptr(dst), _stream_of(dst)))


def ra_filter(phi_l, phi_m, phi_old_l, phi_old_m):
    """Robert-Asselin-Williams filter in place (qg3p_aux.F90:1716-1786)."""
    g = phi_l.shape[2] - 2
    check(load().pq_ra_filter_dev(g, _L(phi_l), _ptr(phi_l), _ptr(phi_m), _ptr(phi_old_l),
                                  _ptr(phi_old_m), _stream_of(phi_l)))


# --------------------------------------------------------------- model
FIELDS = {"psi": 0, "zeta": 1, "rel": 2, "zeta_old": 3, "ui": 4, "uj": 5, "zeta_fi": 6,
          "zeta_fj": 7, "res": 8, "res_beta": 9, "uimo": 10, "ujmo": 11, "flux_i": 12,
          "flux_j": 13, "psi2": 14, "zeta2": 15, "rel2": 16, "zeta_old2": 17}
SCHEMES = {"cabaret": 0, "leapfrog": 1}


def field_shape(name: str, g: int, L: int):
    """Shape of a model field (the reference's array shapes)."""
    if name in ("psi", "zeta", "rel", "zeta_old", "psi2", "zeta2", "rel2", "zeta_old2"):
        return (g + 2, g + 2, L)
    if name in ("ui", "zeta_fi", "uimo", "flux_i"):
        return (g + 3, g + 2, L)
    if name in ("uj", "zeta_fj", "ujmo", "flux_j"):
        return (g + 2, g + 3, L)
    return (g, g, L)


def model_params(g: int, strat: Stratification, beta, force_wind: np.ndarray, *,
                 viscosity: float, visc_bot: float, dt: float, bc_type: int = BC_PARTIAL_SLIP,
                 alpha: float = 0.0, adapt_dt: bool = False, target_cfl: float = 0.1,
                 max_dt: float = 0.0, max_dt_increase: float = 1.5, precision: str = "fp64",
                 scheme: str = "cabaret"):
    """pq_model_params for a model; returns (params, arrays it points into),
    the arrays to be kept alive while the params are used."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (strat.s1, strat.s2, strat.ss)]
    keep += [np.asfortranarray(strat.modes_to_z), np.asfortranarray(strat.z_to_modes),
             np.ascontiguousarray(beta, dtype=np.float64),
             np.asfortranarray(force_wind, dtype=np.float64)]
    s1, s2, ss, m2z, z2m, b, fw = keep
    p = _lib.PqModelParams(g, strat.layers, _np_ptr(s1), _np_ptr(s2), _np_ptr(ss), _np_ptr(m2z),
                           _np_ptr(z2m), _np_ptr(b), _np_ptr(fw), viscosity, visc_bot, bc_type,
                           alpha, int(adapt_dt), target_cfl, max_dt, max_dt_increase, dt,
                           PREC[precision], SCHEMES[scheme])
    return p, keep


class Model:
    """Device-resident CABARET model: one time step in the order of the
    reference's time loop (qg3.F90:590-967)."""

    def __init__(self, g: int, strat: Stratification, beta, force_wind: np.ndarray, **kw):
        self.g, self.layers = g, strat.layers
        self.strat = strat
        p, self._keep = model_params(g, strat, beta, force_wind, **kw)
        h = _vp()
        check(load().pq_model_create(ctypes.byref(h), ctypes.byref(p)))
        self._h = h

    def init(self, zeta_sdata: np.ndarray) -> int:
        z = np.asfortranarray(zeta_sdata, dtype=np.float64)
        return check(load().pq_model_init(self._h, _np_ptr(z)), allow_warn=True)

    def step(self, nsteps: int = 1, final_partial: bool = False) -> None:
        """Enqueues the steps (asynchronous; see status)."""
        check(load().pq_model_step(self._h, nsteps, int(final_partial)))

    def status(self) -> int:
        """Waits for the model's work; raises PqError on a NaN psiC in any
        inversion since the last call (the reference aborts,
        qg3p_aux.F90:504-509) and returns PQ_WARN_LARGE after |psiC| >
        1e100, else PQ_OK."""
        return check(load().pq_model_status(self._h), allow_warn=True)

    def shape(self, name: str):
        return field_shape(name, self.g, self.layers)

    def get(self, name: str) -> np.ndarray:
        a = np.zeros(self.shape(name), order="F")
        check(load().pq_model_get(self._h, FIELDS[name], _np_ptr(a), a.size))
        return a

    def set(self, name: str, a: np.ndarray) -> None:
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.shape == self.shape(name)
        check(load().pq_model_set(self._h, FIELDS[name], _np_ptr(a), a.size))

    def scalars(self):
        out = np.zeros(4)
        check(load().pq_model_scalars(self._h, _np_ptr(out)))
        return dict(dt=out[0], tweight_a=out[1], tweight_u=out[2], model_time=out[3])

    @property
    def stream(self) -> int:
        return load().pq_model_stream(self._h)

    def average(self, enable: bool = True, max_spin: float = 0.0) -> None:
        """Time mean of psi from the first step whose model time exceeds
        max_spin (qg3.F90:672-681); enabling restarts it."""
        check(load().pq_model_average(self._h, int(enable), max_spin))

    def psi_av(self):
        """(psi_av / diagnostic_time as a (G, G, L) array, diagnostic_time)."""
        a = np.zeros((self.g, self.g, self.layers), order="F")
        t = ctypes.c_double()
        check(load().pq_model_psi_av(self._h, _np_ptr(a), a.size, ctypes.byref(t)))
        return a, t.value

    def write_snapshot(self, directory: str, index: int = -1, model_time_day: float = 0.0,
                       total_time_day: float = 0.0, l_scale: float = 1.0, t_scale: float = 1.0,
                       supp: bool = False) -> None:
        """The reference's snapshot data set qgL_G_s[.index] (and, for CABARET
        with supp, qgL_G_s_supp) in `directory`; index -1 gives the name
        a START_PICKUP start reads."""
        check(load().pq_model_write_snapshot(self._h, directory.encode(), index, model_time_day,
                                             total_time_day, l_scale, t_scale, int(supp)))

    def write_mean(self, directory: str, model_time_day: float = 0.0, total_time_day: float = 0.0,
                   l_scale: float = 1.0, t_scale: float = 1.0) -> None:
        check(load().pq_model_write_mean(self._h, directory.encode(), model_time_day, total_time_day,
                                         l_scale, t_scale))

    def read_checkpoint(self, directory: str):
        """START_PICKUP from `directory`; returns (model_time_day, total_time_day)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(load().pq_model_read_checkpoint(self._h, directory.encode(), ctypes.byref(a),
                                              ctypes.byref(b)))
        return a.value, b.value

    def energy(self, h=None):
        """energy(strat, psi, ekin, epot) of the current psi: (ekin[L], epot)."""
        h = np.ascontiguousarray(self.strat.h if h is None else h, dtype=np.float64)
        ekin = np.zeros(self.layers)
        epot = ctypes.c_double()
        check(load().pq_model_energy(self._h, _np_ptr(h), _np_ptr(ekin), ctypes.byref(epot)))
        return ekin, epot.value

    def close(self):
        if getattr(self, "_h", None):
            try:
                load().pq_model_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    __del__ = close
