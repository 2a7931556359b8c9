"""Loader for libpequod_hip.so, the HIP implementation behind the C ABI in
include/pequod_hip.h.  There is no fallback: if the library is missing or
cannot be loaded, every entry point raises."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpequod_hip.so")

PQ_OK, PQ_ERR_ARG, PQ_ERR_HIP, PQ_ERR_NAN, PQ_WARN_LARGE, PQ_ERR_NOMEM = range(6)
# fp64: FP64 storage and transforms; fp32: FP32 storage and transforms;
# mixed: FP64 storage, FP32 transforms; fp32s: FP32 storage, FP64 transforms
# fp32auto: FP32 storage, FP32 transforms when the projection is well
# conditioned (pq_prec_fp32_choice; on the matrix cores where fp32m exists
# for the grid, pq_prec_fp32_auto), else FP64 transforms
# fp32w: FP32 storage, FP64 arithmetic (FP32 rounding between FFT stages)
# fp32m: FP32 storage, row transforms on the matrix cores (fp16 hi/lo split
# operands, FP32 accumulation), g - 1 = 1024
PREC = {"fp64": 0, "fp32": 1, "mixed": 2, "fp32s": 3, "fp32auto": 4, "fp32w": 5, "fp32m": 6}
PREC_NAME = {v: k for k, v in PREC.items()}
# element type of the zeta / psi arrays of each precision
IO_BYTES = {"fp64": 8, "fp32": 4, "mixed": 8, "fp32s": 4, "fp32auto": 4, "fp32w": 4, "fp32m": 4}

_i, _d, _vp, _sz, _l = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_long
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); mirrors include/pequod_hip.h
SIGNATURES = {
    "pq_last_error": (ctypes.c_char_p, []),
    "pq_device_info": (_i, [ctypes.c_char_p, _i]),
    "pq_malloc": (_i, [ctypes.POINTER(_vp), _sz]),
    "pq_free": (_i, [_vp]),
    "pq_memcpy_h2d": (_i, [_vp, _vp, _sz]),
    "pq_memcpy_d2h": (_i, [_vp, _vp, _sz]),
    "pq_synchronize": (_i, []),
    "pq_mirror_register": (_i, [_vp, _sz]),
    "pq_mirror_unregister": (_i, [_vp]),
    "pq_mirror_device_ptr": (_vp, [_vp]),
    "pq_mirror_push": (_i, [_vp]),
    "pq_mirror_pull": (_i, [_vp]),
    "pq_helm_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _dp, _i]),
    "pq_helm_create_periodic": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _dp, _i]),
    "pq_helm_destroy": (_i, [_vp]),
    "pq_helm_solve": (_i, [_vp, _dp, _i, _i]),
    "pq_helm_solve_dev": (_i, [_vp, _vp, _l, _l, _vp]),
    "pq_inv_create": (_i, [ctypes.POINTER(_vp), _i, _i, _dp, _dp, _dp, _i, _i]),
    "pq_chan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _dp, _dp, _dp, _i, _d, _i, _i]),
    "pq_inv_destroy": (_i, [_vp]),
    "pq_inv_run": (_i, [_vp, _dp, _i, _i, _dp, _i, _i]),
    "pq_inv_run_dev": (_i, [_vp, _i, ctypes.POINTER(_vp), _l, _l, ctypes.POINTER(_vp), _l, _l, _vp]),
    "pq_inv_status": (_i, [_vp, _dp]),
    "pq_inv_stream": (_vp, [_vp]),
    "pq_inv_set_tiles": (_i, [_vp, _i, _i]),
    "pq_inv_set_member": (_i, [_vp, _i, _dp, _dp, _dp]),
    "pq_inv_set_member_layers": (_i, [_vp, _i, _i, _dp, _dp, _dp]),
    "pq_srv_create": (_i, [ctypes.POINTER(_vp), _i, _sz, _i]),
    "pq_srv_open_field": (_i, [_vp, _i, _i, _dp, _dp, _dp, _vp, _l, _l, _vp, _l, _l,
                               ctypes.POINTER(_i)]),
    "pq_srv_open_model": (_i, [_vp, _vp, _dp, ctypes.POINTER(_i)]),
    "pq_srv_model": (_vp, [_vp, _i]),
    "pq_srv_submit": (_i, [_vp, _i, _i, _i, ctypes.POINTER(ctypes.c_longlong)]),
    "pq_srv_poll": (_i, [_vp, ctypes.c_longlong, _dp]),
    "pq_srv_drain": (_i, [_vp, _d]),
    "pq_srv_stats": (_i, [_vp, _i, ctypes.c_longlong, _dp]),
    "pq_srv_used_bytes": (_sz, [_vp]),
    "pq_srv_close": (_i, [_vp, _i]),
    "pq_srv_destroy": (_i, [_vp]),
    "pq_inv_get_tiles": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "pq_inv_pass_times": (_i, [_vp, _i, ctypes.POINTER(_vp), _l, _l, ctypes.POINTER(_vp), _l, _l, _i,
                               ctypes.POINTER(ctypes.c_float)]),
    "pq_cabaret_cfl_dev": (_i, [_i, _i, _d, _vp, _vp, _dp, _vp]),
    "pq_cabaret_predictor_dev": (_i, [_i, _i, _d, _vp, _vp, _vp, _vp]),
    "pq_cabaret_beta_dev": (_i, [_i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp]),
    "pq_cabaret_forcing_dev": (_i, [_i, _i, _d, _vp, _d, _d, _vp, _vp, _vp]),
    "pq_cabaret_cell_boundaries_dev": (_i, [_i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pq_cabaret_cell_boundaries_velocity_dev": (_i, [_i, _i, _vp, _vp, _vp, _vp]),
    "pq_cabaret_extrapolator_dev": (_i, [_i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pq_cabaret_extrapolator_noalloc_dev": (_i, [_i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pq_cabaret_flux_dev": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pq_cabaret_corrector_dev": (_i, [_i, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    "pq_rel_from_zeta_and_psi_dev": (_i, [_i, _i, _vp, _vp, _i, _d, _vp, _vp, _vp, _vp]),
    "pq_update_zeta_bc_dev": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pq_zeta_step_dev": (_i, [_i, _i, _d, _vp, _vp, _vp, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pq_ra_filter_dev": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp]),
    "pq_model_create": (_i, [ctypes.POINTER(_vp), _vp]),
    "pq_model_destroy": (_i, [_vp]),
    "pq_model_init": (_i, [_vp, _dp]),
    "pq_model_step": (_i, [_vp, _i, _i]),
    "pq_model_status": (_i, [_vp]),
    "pq_model_get": (_i, [_vp, _i, _dp, _sz]),
    "pq_model_set": (_i, [_vp, _i, _dp, _sz]),
    "pq_model_scalars": (_i, [_vp, _dp]),
    "pq_model_stream": (_vp, [_vp]),
    "pq_model_average": (_i, [_vp, _i, _d]),
    "pq_model_write_snapshot": (_i, [_vp, ctypes.c_char_p, _i, _d, _d, _d, _d, _i]),
    "pq_model_write_mean": (_i, [_vp, ctypes.c_char_p, _d, _d, _d, _d]),
    "pq_model_read_checkpoint": (_i, [_vp, ctypes.c_char_p, _dp, _dp]),
    "pq_model_psi_av": (_i, [_vp, _dp, _sz, _dp]),
    "pq_model_energy": (_i, [_vp, _dp, _dp, _dp]),
    "pq_diag_create": (_i, [ctypes.POINTER(_vp), _i, _i, _dp, _dp]),
    "pq_diag_destroy": (_i, [_vp]),
    "pq_diag_stream": (_vp, [_vp]),
    "pq_diag_energy": (_i, [_vp, _dp, _i, _i, _dp, _dp]),
    "pq_diag_energy_dev": (_i, [_vp, _vp, _l, _l, _vp, _vp]),
    "pq_psi_av_add_dev": (_i, [_i, _i, _d, _vp, _l, _l, _vp, _l, _l, _vp]),
    "pq_ens_create": (_i, [ctypes.POINTER(_vp), _i]),
    "pq_ens_add": (_i, [_vp, _i, _i, _dp, _dp, _dp, _vp, _l, _l, _vp, _l, _l, ctypes.POINTER(_i)]),
    "pq_ens_add_model": (_i, [_vp, _vp, _dp, _dp, _dp, ctypes.POINTER(_i)]),
    "pq_ens_finalize": (_i, [_vp]),
    "pq_ens_class_count": (_i, [_vp]),
    "pq_ens_class_handle": (_vp, [_vp, _i]),
    "pq_ens_member_class": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "pq_ens_class_precision": (_i, [_vp, _i]),
    "pq_prec_fp32_choice": (_i, [_i, _dp, _dp]),
    "pq_prec_fp32_auto": (_i, [_i, _i, _dp, _dp]),
    "pq_ens_stream": (_vp, [_vp]),
    "pq_ens_invert": (_i, [_vp]),
    "pq_ens_step": (_i, [_vp, _i]),
    "pq_ens_capture": (_i, [_vp, _i]),
    "pq_ens_synchronize": (_i, [_vp]),
    "pq_ens_status": (_i, [_vp]),
    "pq_ens_destroy": (_i, [_vp]),
}


class PqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pequod_hip error {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libpequod_hip.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"{path} not found: build it with `python -m pequod_amd.build`")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, allow_warn: bool = False) -> int:
    if rc == PQ_OK or (allow_warn and rc == PQ_WARN_LARGE):
        return rc
    msg = load().pq_last_error().decode(errors="replace")
    raise PqError(rc, msg)


class PqModelParams(ctypes.Structure):
    """pq_model_params of include/pequod_hip.h."""
    _fields_ = [
        ("g", _i), ("layers", _i),
        ("s1", _dp), ("s2", _dp),
        ("ss", _dp), ("modes_to_z", _dp), ("z_to_modes", _dp),
        ("beta", _dp), ("force_wind", _dp),
        ("viscosity", _d), ("visc_bot", _d),
        ("bc_type", _i), ("alpha", _d),
        ("adapt_dt", _i), ("target_cfl", _d), ("max_dt", _d), ("max_dt_increase", _d),
        ("dt", _d), ("precision", _i), ("scheme", _i),
    ]
