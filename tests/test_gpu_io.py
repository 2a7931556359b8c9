"""The library's output / checkpoint files for device-resident models against
the oracle's restatement of the reference's format (byte for byte), and the
START_PICKUP start: a CABARET run continued from a checkpoint equals the
uninterrupted run bit for bit.
"""
import filecmp
import os

import numpy as np
import pytest

from oracle import io_format as F
from oracle import oracle as O
from pequod_amd import qg, synth
from pequod_amd._lib import PqError

pytestmark = pytest.mark.gpu


def _model(g=33, scheme="cabaret"):
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    m = qg.Model(g, st, nd["beta"], c.wind(), viscosity=nd["viscosity"], visc_bot=nd["visc_bot"],
                 dt=nd["dt"], bc_type=1, alpha=nd["alpha"], scheme=scheme)
    m.init(c.q())
    return m, st


def _same_files(a, b):
    return filecmp.cmp(a, b, shallow=False)


def test_cabaret_snapshot_bytes_match_oracle(dev, tmp_path):
    print("io cabaret snapshot", flush=True)
    g = 33
    m, st = _model(g)
    m.step(3)
    m.step(1, final_partial=True)   # the driver's output point
    lib_dir, ora_dir = tmp_path / "lib", tmp_path / "ora"
    lib_dir.mkdir()
    ora_dir.mkdir()
    m.write_snapshot(str(lib_dir), 7, 12.5, 112.5, 1.0e6, 5.0e5, supp=True)
    L = st.layers
    F.write_snapshot(str(ora_dir), g, L, 12.5, 112.5, m.get("psi"), m.get("zeta"), 1.0e6, 5.0e5, index=7)
    F.write_supp(str(ora_dir), g, L, 12.5, 112.5, m.scalars()["dt"], m.get("ui"), m.get("uj"),
                 m.get("zeta_fi"), m.get("zeta_fj"), m.get("res"), m.get("res_beta"), 1.0e6, 5.0e5,
                 index=7)
    for name in (f"qg{L}_{g}_s.7", f"qg{L}_{g}_s.7.d", f"qg{L}_{g}_s_supp.7", f"qg{L}_{g}_s_supp.7.d"):
        assert _same_files(lib_dir / name, ora_dir / name), name


def test_cabaret_pickup_continues_bitwise(dev, tmp_path):
    """N steps, then M steps, against N - 1 steps plus the step to the
    output point, a checkpoint, START_PICKUP in a fresh model and M steps."""
    print("io cabaret pickup", flush=True)
    N, M = 5, 4
    a, _ = _model()
    a.step(N + M)
    b, _ = _model()
    b.step(N - 1)
    b.step(1, final_partial=True)
    b.write_snapshot(str(tmp_path), -1, 3.0, 3.0, supp=True)
    c, _ = _model()
    c.step(2)  # state to be replaced by the checkpoint
    tm, tt = c.read_checkpoint(str(tmp_path))
    assert (tm, tt) == (3.0, 3.0)
    assert c.scalars()["model_time"] == c.scalars()["dt"] / 2
    c.step(M)
    for name in ("psi", "zeta", "zeta_old", "ui", "uj", "zeta_fi", "zeta_fj", "res", "res_beta"):
        np.testing.assert_array_equal(c.get(name), a.get(name), err_msg=name)


def test_leapfrog_snapshot_and_pickup(dev, tmp_path):
    """Both time levels in one record per layer, byte for byte; the pickup
    recomputes rel from psi (so the continuation agrees to round-off)."""
    print("io leapfrog", flush=True)
    g = 33
    a, st = _model(g, "leapfrog")
    a.step(4)
    a.write_snapshot(str(tmp_path), -1, 1.0, 2.0)
    L = st.layers
    od = tmp_path / "ora"
    od.mkdir()
    F.write_snapshot(str(od), g, L, 1.0, 2.0, a.get("psi"), a.get("zeta"), psi2=a.get("psi2"),
                     zeta2=a.get("zeta2"))
    for name in (f"qg{L}_{g}_s", f"qg{L}_{g}_s.d"):
        assert _same_files(tmp_path / name, od / name), name
    c, _ = _model(g, "leapfrog")
    c.read_checkpoint(str(tmp_path))
    a.step(6)
    c.step(6)
    for name in ("psi", "zeta", "psi2", "zeta2"):
        x, y = c.get(name), a.get(name)
        assert np.max(np.abs(x - y)) <= 1e-10 * np.max(np.abs(y)), name


def test_time_mean_file(dev, tmp_path):
    m, st = _model()
    m.average(True, 0.0)
    m.step(3)
    m.write_mean(str(tmp_path), 4.0, 5.0, 2.0, 3.0)
    mean, _ = m.psi_av()
    od = tmp_path / "ora"
    od.mkdir()
    F.write_mean(str(od), 33, st.layers, 4.0, 5.0, mean, 2.0, 3.0)
    for name in (f"qg{st.layers}_33_s_av", f"qg{st.layers}_33_s_av.d"):
        assert _same_files(tmp_path / name, od / name), name


def test_checkpoint_errors(dev, tmp_path):
    """Missing files, a record of the wrong size and supplementary times that
    do not match the snapshot's (qg3p_io.F90:253-296) are errors."""
    m, st = _model()
    with pytest.raises(PqError):
        m.read_checkpoint(str(tmp_path))
    m.step(1, final_partial=True)
    m.write_snapshot(str(tmp_path), -1, 1.0, 1.0, supp=True)
    L = st.layers
    supp = tmp_path / f"qg{L}_33_s_supp.d"
    raw = bytearray(supp.read_bytes())
    bad = F.record(1.5)  # model_time_day 1.5 instead of 1.0
    supp.write_bytes(bytes(bad) + bytes(raw[len(bad):]))
    with pytest.raises(PqError):
        m.read_checkpoint(str(tmp_path))
    snap = tmp_path / f"qg{L}_33_s.d"
    data = bytearray(snap.read_bytes())
    snap.write_bytes(bytes(data[:-20]))
    with pytest.raises(PqError):
        m.read_checkpoint(str(tmp_path))
