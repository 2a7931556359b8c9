"""Models with FP32 field storage: every CABARET / centred-scheme field is
held in FP32 on the device while the stage kernels compute in FP64 (each
read widens, each write rounds once), and the inversion runs at the handle's
precision.  Checked against the FP64 oracle's model with a stated relative
tolerance; ensemble members, checkpoint files and restarts of FP32 models
against the same model stepped alone, bit for bit.
"""
import numpy as np
import pytest

from oracle import io_format as F
from oracle import oracle as O
from pequod_amd import qg, synth
from pequod_amd._lib import PqError
from pequod_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

# max |x - oracle| / max |oracle| after 10 steps + a final partial step of
# the 3-layer physics at 65^2: the state (psi, zeta, zeta_old) within 1e-6;
# fields that difference it (rel, the tendencies res / res_beta, face
# values, velocities and fluxes) lose digits to cancellation, bounded by
# 3e-5.  "fp32" and "fp32w" add FP32 transforms to the inversion.
STATE = ("psi", "zeta", "zeta_old")


def _tol(name):
    return 1e-6 if name in STATE else 3e-5


def _setup(g=65, seed=0, **extra):
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    kw = dict(viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=1,
              alpha=nd["alpha"])
    kw.update(extra)
    return c, st, nd, kw, c.q(member=seed) if seed else c.q()


def _rel(a, b):
    den = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (den if den > 0 else 1.0)


@pytest.mark.parametrize("precision", ["fp32s", "fp32", "fp32w"])
def test_fp32_cabaret_model_matches_oracle(dev, precision):
    print(f"fp32 cabaret model {precision}", flush=True)
    c, st, nd, kw, q = _setup()
    gm = qg.Model(c.g, st, nd["beta"], c.wind(), precision=precision, **kw)
    om = O.Model(c.g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], c.wind(), **kw)
    gm.init(q)
    om.init(q)
    gm.step(10)
    om.step(10)
    gm.step(1, final_partial=True)
    om.step(1, final_partial=True)
    assert gm.scalars()["model_time"] == om.scalars()["model_time"]
    errs = {name: _rel(gm.get(name), om.a[name])
            for name in ("psi", "zeta", "zeta_old", "rel", "ui", "uj", "zeta_fi", "zeta_fj", "res",
                         "res_beta", "flux_i", "flux_j")}
    print(precision, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for name, e in errs.items():
        assert e <= _tol(name), (name, e)
    # the stored values are FP32 numbers
    psi = gm.get("psi")
    assert np.array_equal(psi, psi.astype(np.float32).astype(np.float64))


def test_fp32_leapfrog_model_matches_oracle(dev):
    print("fp32 leapfrog model", flush=True)
    c, st, nd, kw, q = _setup()
    gm = qg.Model(c.g, st, nd["beta"], c.wind(), precision="fp32", scheme="leapfrog", **kw)
    om = O.LeapfrogModel(c.g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"],
                         c.wind(), **kw)
    gm.init(q)
    om.init(q)
    gm.step(11)
    om.step(11)
    for name in ("psi", "zeta", "rel", "zeta_old"):
        for k, suf in ((1, ""), (2, "2")):
            e = _rel(gm.get(name + suf), om.a[f"{name}{k}"])
            assert e <= _tol(name), (name, k, e)


@pytest.mark.parametrize("graph", [False, True])
def test_fp32_models_in_ensemble_match_standalone(dev, graph):
    """One member per size class, so the batched class inversion is the
    standalone handle's; an FP64 ensemble refuses FP32 models and the
    reverse."""
    print(f"fp32 ensemble models graph={graph}", flush=True)
    sizes = [(33, 1), (65, 2), (49, 3)]

    def make(g, s):
        c, st, nd, kw, q = _setup(g, s)
        m = qg.Model(g, st, nd["beta"], c.wind(), precision="fp32", **kw)
        m.init(q)
        return m, st

    ens_models = [make(g, s) for g, s in sizes]
    ref_models = [make(g, s) for g, s in sizes]
    ens = Ensemble("fp32")
    for m, st in ens_models:
        ens.add_model(m, st)
    ens.finalize(tiles=None)
    if graph:
        ens.capture(Ensemble.STEP)
    ens.step(4)
    ens.synchronize()
    for (m, _), (r, _) in zip(ens_models, ref_models):
        r.step(4)
        for name in ("zeta", "psi", "zeta_fi", "ui", "res"):
            np.testing.assert_array_equal(m.get(name), r.get(name), err_msg=name)
        assert m.scalars()["model_time"] == r.scalars()["model_time"]
    e64 = Ensemble("fp64")
    with pytest.raises(PqError):
        e64.add_model(*make(33, 1))
    c, st, nd, kw, q = _setup(33)
    m64 = qg.Model(33, st, nd["beta"], c.wind(), **kw)
    m64.init(q)
    with pytest.raises(PqError):
        Ensemble("fp32").add_model(m64, st)


def test_fp32_model_files_and_restart(dev, tmp_path):
    """The snapshot files of an FP32 model hold its values widened to FP64
    (byte for byte the oracle's writer on them), and a START_PICKUP restart
    continues bit for bit."""
    print("fp32 model io", flush=True)
    g, N, M = 33, 5, 4
    c, st, nd, kw, q = _setup(g)

    def model():
        m = qg.Model(g, st, nd["beta"], c.wind(), precision="fp32", **kw)
        m.init(q)
        return m

    a = model()
    a.step(N + M)
    b = model()
    b.step(N - 1)
    b.step(1, final_partial=True)
    b.write_snapshot(str(tmp_path), -1, 3.0, 3.0, supp=True)
    od = tmp_path / "ora"
    od.mkdir()
    L = st.layers
    F.write_snapshot(str(od), g, L, 3.0, 3.0, b.get("psi"), b.get("zeta"))
    F.write_supp(str(od), g, L, 3.0, 3.0, b.scalars()["dt"], b.get("ui"), b.get("uj"),
                 b.get("zeta_fi"), b.get("zeta_fj"), b.get("res"), b.get("res_beta"))
    for name in (f"qg{L}_{g}_s.d", f"qg{L}_{g}_s_supp.d"):
        assert (tmp_path / name).read_bytes() == (od / name).read_bytes(), name
    r = model()
    r.read_checkpoint(str(tmp_path))
    r.step(M)
    for name in ("psi", "zeta", "zeta_old", "ui", "uj", "zeta_fi", "zeta_fj", "res", "res_beta"):
        np.testing.assert_array_equal(r.get(name), a.get(name), err_msg=name)


def test_fp32_model_time_mean_and_energy(dev):
    """The FP64 time mean and energy of an FP32 model (its psi read widened):
    against the FP64 oracle's model within 1e-6 (the FP32 state's error),
    and the energy exactly the oracle's energy of the model's own psi."""
    print("fp32 model diagnostics", flush=True)
    c, st, nd, kw, q = _setup()
    gm = qg.Model(c.g, st, nd["beta"], c.wind(), precision="fp32", **kw)
    om = O.Model(c.g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], c.wind(), **kw)
    gm.init(q)
    om.init(q)
    dt = nd["dt"]
    max_spin = gm.scalars()["model_time"] + 2.2 * dt
    gm.average(True, max_spin)
    om.average(max_spin)
    gm.step(7)
    om.step(7)
    a, t = gm.psi_av()
    b, tb = om.psi_av()
    assert t == tb
    assert _rel(a, b) <= 1e-6
    ekin, epot = gm.energy()
    ek2, ep2 = O.energy(om.a["psi"], st.h, st.s2)
    assert np.all(np.abs(ekin - ek2) <= 1e-6 * np.abs(ek2)) and abs(epot - ep2) <= 1e-6 * ep2
    ek3, ep3 = O.energy(gm.get("psi"), st.h, st.s2)
    assert np.all(np.abs(ekin - ek3) <= 1e-12 * np.abs(ek3)) and abs(epot - ep3) <= 1e-12 * ep3


def test_fp32_model_served_matches_standalone(dev):
    """An FP32 model member of the server advances exactly as the same
    model stepped alone (12 + 3 steps: the 10-step graph and single steps)."""
    from pequod_amd.serve import Server
    print("fp32 served model", flush=True)
    g = 65
    c, st, nd, kw, q = _setup(g, 3)
    srv = Server("fp32", cap_bytes=16e9, max_batch=4)
    mid = srv.open_model(g, st, nd["beta"], c.wind(), q, precision="fp32", **kw)
    for n in (12, 3):
        srv.submit(mid, Server.MODEL, n)
    srv.drain(120)
    assert srv.stats()["failed"] == 0
    ref = qg.Model(g, st, nd["beta"], c.wind(), precision="fp32", **kw)
    ref.init(q)
    ref.step(15)
    for name in ("zeta", "psi", "zeta_fi", "ui", "res"):
        np.testing.assert_array_equal(srv.model_get(mid, name, g, st.layers), ref.get(name), err_msg=name)
    import ctypes
    from pequod_amd._lib import check, load
    sc = np.zeros(4)
    check(load().pq_model_scalars(load().pq_srv_model(srv._h, mid),
                                  sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    assert sc[3] == ref.scalars()["model_time"]  # host clock advanced per replayed step
    ref.close()
    srv.close()
