"""The C ABI library loads and exports every function include/pequod_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pequod_hip.h")
LIB = os.path.join(ROOT, "pequod_amd", "libpequod_hip.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(pq_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    for must in ("pq_helm_create", "pq_helm_solve", "pq_inv_create", "pq_inv_run",
                 "pq_inv_run_dev", "pq_cabaret_extrapolator_dev", "pq_model_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libpequod_hip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\s[TW]\s(pq_\w+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    if not os.path.exists(LIB):
        pytest.skip("libpequod_hip.so not built")
    from pequod_amd import _lib
    L = _lib.load()
    for n in declared():
        assert hasattr(L, n)
    assert set(_lib.SIGNATURES) >= set(declared())
    # pure host entry points work without a device
    assert L.pq_last_error() is not None


def test_supported_sizes_rejected_without_device():
    """Argument validation happens before any device work."""
    if not os.path.exists(LIB):
        pytest.skip("libpequod_hip.so not built")
    from pequod_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    ss = (ctypes.c_double * 1)(0.0)
    rc = L.pq_helm_create(ctypes.byref(h), 100, 100, 1, ss, 0)
    assert rc == _lib.PQ_ERR_ARG
    assert b"unsupported" in L.pq_last_error()


def test_fp32_precision_choice_follows_projection_condition():
    """FP32 transforms up to kappa_inf = 32: the 3-layer (4.1), 6-layer
    (14.4) and 8-layer ensemble (26.2) stratifications; FP64 transforms for
    8 layers with depths growing 2.5x per layer (85.8)."""
    import ctypes
    import numpy as np
    from pequod_amd import _lib, synth
    L = _lib.load()
    dp = ctypes.POINTER(ctypes.c_double)

    def choice(st):
        m2z = np.asfortranarray(st.modes_to_z)
        z2m = np.asfortranarray(st.z_to_modes)
        return L.pq_prec_fp32_choice(st.layers, m2z.ctypes.data_as(dp), z2m.ctypes.data_as(dp))

    assert choice(synth.config("qg3_1023").stratification()) == _lib.PREC["fp32"]
    assert choice(synth.config("qg6_2047").stratification()) == _lib.PREC["fp32"]
    eight = [m for m in (synth.ensemble_member(k) for k in range(16)) if m.layers == 8][0]
    assert choice(eight.stratification()) == _lib.PREC["fp32"]
    assert choice(synth.steep_member(8, 2.5).stratification()) == _lib.PREC["fp32s"]


def test_kernel_units_have_distinct_kernel_symbols():
    """The inversion kernels are compiled once per transform length and
    linked into one library: a kernel whose symbol does not encode the
    length would be defined by every unit, and every unit's launch would
    resolve to one unit's kernel (compiled for another length).  The kernel
    symbols of any two units must therefore be disjoint."""
    build = os.path.join(ROOT, "pequod_amd", "_build")
    objs = sorted(f for f in os.listdir(build) if f.startswith("pq_inv_kernels_n")) \
        if os.path.isdir(build) else []
    if len(objs) < 2:
        pytest.skip("kernel unit objects not built")
    seen = {}
    for o in objs:
        out = subprocess.run(["nm", os.path.join(build, o)], capture_output=True, text=True,
                             check=True).stdout
        # kernel handles are the unit's weak objects ("V") in namespace pq
        kernels = set(re.findall(r"\sV\s(_ZN2pq\w+)$", out, flags=re.M))
        assert kernels, o
        for k in kernels:
            assert k not in seen, f"{k} defined by {seen.get(k)} and {o}"
            seen[k] = o
