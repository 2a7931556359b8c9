"""pq_cabaret_extrapolator_noalloc_dev, the extrapolator entry point that
allocates nothing (its work area is a fixed device array of the library's),
against pq_cabaret_extrapolator_dev through the C ABI: bit for bit on two
calls in a row, the second on a smaller grid (the work area keeps the first
call's values), and a grid beyond the work area is refused before any GPU
work."""
import ctypes

import numpy as np
import pytest

from pequod_amd._lib import PQ_ERR_ARG, load

from helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _fields(g, L, seed, dev):
    rng = np.random.default_rng(seed)
    node = lambda: np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))
    fi = lambda: np.asfortranarray(rng.standard_normal((g + 3, g + 2, L)))
    fj = lambda: np.asfortranarray(rng.standard_normal((g + 2, g + 3, L)))
    arrs = dict(z=node(), zo=node(), ui=fi(), uj=fj(), zfi=fi(), zfj=fj())
    return {k: to_dev(v, dev) for k, v in arrs.items()}


def _call(fn, g, L, dt, t, with_work, stream):
    args = [g, L, dt, t["z"].data_ptr(), t["zo"].data_ptr(), t["ui"].data_ptr(), t["uj"].data_ptr(),
            t["zfi"].data_ptr(), t["zfj"].data_ptr()]
    if with_work:
        args.append(None)
    return fn(*args, stream)


def test_extrapolator_noalloc_matches(dev):
    import torch
    lib = load()
    noalloc = lib.pq_cabaret_extrapolator_noalloc_dev
    ref = lib.pq_cabaret_extrapolator_dev
    stream = _vp(torch.cuda.current_stream(dev).cuda_stream or -1)
    for k, (g, L) in enumerate([(257, 3), (65, 2)]):
        print(f"extrapolator noalloc g={g} L={L}", flush=True)
        a = _fields(g, L, 40 + k, dev)
        b = {n: v.clone() for n, v in a.items()}
        assert _call(ref, g, L, 0.31, a, True, stream) == 0
        assert _call(noalloc, g, L, 0.31, b, False, stream) == 0
        torch.cuda.synchronize()
        for n in ("zfi", "zfj"):
            np.testing.assert_array_equal(from_dev(b[n]), from_dev(a[n]), err_msg=f"{n} g={g}")
        # the faces really changed
        assert not torch.equal(a["zfi"], _fields(g, L, 40 + k, dev)["zfi"])


def test_extrapolator_noalloc_rejects_oversize(dev):
    import torch
    lib = load()
    noalloc = lib.pq_cabaret_extrapolator_noalloc_dev
    # 2 (g+3)(g+2) L doubles above 2^23: g = 2049, one layer; no array is
    # touched, the size check comes first
    p = torch.zeros(8, dtype=torch.float64, device=dev)
    ptr = p.data_ptr()
    rc = noalloc(2049, 1, 0.1, ptr, ptr, ptr, ptr, ptr, ptr, None)
    assert rc == PQ_ERR_ARG
    assert b"work area" in lib.pq_last_error()
