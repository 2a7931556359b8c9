"""Parity of the HIP CABARET stages and time step with the FP64 oracle.

Stage kernels and the oracle both evaluate every statement of the
reference routine in the written order with no FP contraction, so the
stages must agree bit for bit.  The full step includes the PV inversion
(round-off level differences), so whole steps are compared with a
relative tolerance.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pequod_amd import qg, strat as S, synth

from helpers import from_dev, random_strat, rel_err, to_dev

pytestmark = pytest.mark.gpu


def _node(rng, g, L):
    return np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))


def _fi(rng, g, L):
    return np.asfortranarray(rng.standard_normal((g + 3, g + 2, L)))


def _fj(rng, g, L):
    return np.asfortranarray(rng.standard_normal((g + 2, g + 3, L)))


def _res(rng, g, L):
    return np.asfortranarray(rng.standard_normal((g, g, L)))


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev) if a.ndim == 1 else to_dev(a, dev)


def _sync():
    import torch
    torch.cuda.synchronize()


@pytest.fixture(params=[(9, 1), (17, 3), (33, 2), (65, 4)])
def gl(request):
    return request.param


def test_predictor_beta(dev, gl):
    g, L = gl
    rng = np.random.default_rng(1)
    zo, res, ui, z, rb = _node(rng, g, L), _res(rng, g, L), _fi(rng, g, L), _node(rng, g, L), _res(rng, g, L)
    beta = rng.uniform(0.1, 1.0, L)
    dt, twa = 0.37, 0.61
    tz, tzo, tres, tui, trb, tb = (_t(a, dev) for a in (z, zo, res, ui, rb, beta))
    qg.cabaret_predictor(dt, tzo, tres, tz)
    qg.cabaret_beta(dt, twa, tb, tui, tz, trb)
    _sync()
    O.cabaret_predictor(dt, zo, res, z)
    O.cabaret_beta(dt, twa, beta, ui, z, rb)
    np.testing.assert_array_equal(from_dev(tz), z)
    np.testing.assert_array_equal(from_dev(trb), rb)


def test_forcing(dev, gl):
    g, L = gl
    rng = np.random.default_rng(2)
    rel, z = _node(rng, g, L), _node(rng, g, L)
    fw = np.asfortranarray(rng.standard_normal((g, g)))
    tz, trel = _t(z, dev), _t(rel, dev)
    import torch
    tfw = torch.from_numpy(np.ascontiguousarray(fw.T)).to(dev)
    qg.cabaret_forcing(0.3, tfw, 0.02, 0.001, trel, tz)
    _sync()
    O.cabaret_forcing(0.3, fw, 0.02, 0.001, rel, z)
    np.testing.assert_array_equal(from_dev(tz), z)


def test_cell_boundaries(dev, gl):
    g, L = gl
    rng = np.random.default_rng(3)
    psi, ui, uj, uimo, ujmo = _node(rng, g, L), _fi(rng, g, L), _fj(rng, g, L), _fi(rng, g, L), _fj(rng, g, L)
    ts = [_t(a, dev) for a in (psi, ui, uj, uimo, ujmo)]
    qg.cabaret_cell_boundaries(0.4, *ts)
    _sync()
    O.cabaret_cell_boundaries(0.4, psi, ui, uj, uimo, ujmo)
    for t, a in zip(ts[1:], (ui, uj, uimo, ujmo)):
        np.testing.assert_array_equal(from_dev(t), a)
    ts2 = [_t(a, dev) for a in (psi, ui, uj)]
    qg.cabaret_cell_boundaries_velocity(*ts2)
    _sync()
    O.cabaret_cell_boundaries_velocity(psi, ui, uj)
    np.testing.assert_array_equal(from_dev(ts2[1]), ui)
    np.testing.assert_array_equal(from_dev(ts2[2]), uj)


@pytest.mark.parametrize("sign", ["mixed", "pos", "neg"])
def test_extrapolator_flux_corrector(dev, gl, sign):
    g, L = gl
    rng = np.random.default_rng(4)
    z, zo = _node(rng, g, L), _node(rng, g, L)
    ui, uj = _fi(rng, g, L), _fj(rng, g, L)
    if sign == "pos":
        ui, uj = np.abs(ui), np.abs(uj)
    elif sign == "neg":
        ui, uj = -np.abs(ui), -np.abs(uj)
    zfi, zfj = _fi(rng, g, L), _fj(rng, g, L)
    fli, flj, res = _fi(rng, g, L), _fj(rng, g, L), _res(rng, g, L)
    T = {k: _t(v, dev) for k, v in dict(z=z, zo=zo, ui=ui, uj=uj, zfi=zfi, zfj=zfj, fli=fli,
                                         flj=flj, res=res).items()}
    dt = 0.29
    qg.cabaret_extrapolator(dt, T["z"], T["zo"], T["ui"], T["uj"], T["zfi"], T["zfj"])
    qg.cabaret_flux(T["ui"], T["uj"], T["zfi"], T["zfj"], T["fli"], T["flj"])
    qg.cabaret_corrector(dt, T["fli"], T["flj"], T["res"], T["z"])
    _sync()
    O.cabaret_extrapolator(dt, z, zo, ui, uj, zfi, zfj)
    O.cabaret_flux(ui, uj, zfi, zfj, fli, flj)
    O.cabaret_corrector(dt, fli, flj, res, z)
    for k, a in dict(zfi=zfi, zfj=zfj, fli=fli, flj=flj, res=res, z=z).items():
        np.testing.assert_array_equal(from_dev(T[k]), a, err_msg=k)


@pytest.mark.parametrize("bc", [1, 2, 3])
def test_rel_and_zeta_bc(dev, gl, bc):
    g, L = gl
    rng = np.random.default_rng(5)
    z, psi, rel = _node(rng, g, L), _node(rng, g, L), _node(rng, g, L)
    s1, s2 = random_strat(L, rng)
    T = [_t(a, dev) for a in (s1, s2, z, psi, rel)]
    qg.rel_from_zeta_and_psi(T[0], T[1], bc, 0.08, T[2], T[3], T[4])
    qg.update_zeta_bc(T[0], T[1], T[4], T[3], T[2])
    _sync()
    O.rel_from_zeta_and_psi(s1, s2, bc, 0.08, z, psi, rel)
    O.update_zeta_bc(s1, s2, rel, psi, z)
    np.testing.assert_array_equal(from_dev(T[4]), rel)
    np.testing.assert_array_equal(from_dev(T[2]), z)


def test_cabaret_cfl_known_answers(dev):
    """unittest/test_cabaret_cfl.F90:17-36 on the GPU."""
    import torch
    g = 65
    ui = torch.ones((1, g + 2, g + 3), dtype=torch.float64, device=dev)
    uj = torch.zeros((1, g + 3, g + 2), dtype=torch.float64, device=dev)
    assert qg.cabaret_cfl(1.0, ui, uj) == 1.0
    assert qg.cabaret_cfl(2.0, ui, uj) == 2.0
    uj.fill_(2.0)
    assert qg.cabaret_cfl(1.0, ui, uj) == 2.0
    assert qg.cabaret_cfl(2.0, ui, uj) == 4.0
    ui[0, 9, 9] = 10.0  # Fortran u_i%data(10, 10, 1)
    assert qg.cabaret_cfl(1.0, ui, uj) == 10.0
    assert qg.cabaret_cfl(2.0, ui, uj) == 20.0
    uj[0, 9, 9] = 20.0
    assert qg.cabaret_cfl(1.0, ui, uj) == 20.0
    assert qg.cabaret_cfl(2.0, ui, uj) == 40.0


def _models(cfg_name, g_override=None, adapt=False):
    c = synth.config(cfg_name)
    if g_override:
        c.n = g_override - 2
    nd = c.nondim()
    st = nd["strat"]
    fw = c.wind()
    kw = dict(viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=1,
              alpha=nd["alpha"], adapt_dt=adapt, target_cfl=0.1, max_dt=86400.0 / c.t_scale,
              max_dt_increase=1.5)
    gm = qg.Model(c.g, st, nd["beta"], fw, **kw)
    om = O.Model(c.g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], fw, **kw)
    q = c.q()
    return c, gm, om, q


@pytest.mark.parametrize("adapt", [False, True])
def test_model_steps_match_oracle(dev, adapt):
    """10 full steps + a final partial step of the 3-layer 255^2 physics on a 65^2 grid
    (Courant number about 0.3), FP64, vs the oracle."""
    c, gm, om, q = _models("qg3_255", g_override=65, adapt=adapt)
    gm.init(q)
    om.init(q)
    for name in ("psi", "zeta", "rel"):
        assert rel_err(gm.get(name), om.a[name]) < 1e-11, name
    gm.step(10)
    om.step(10)
    gm.step(1, final_partial=True)
    om.step(1, final_partial=True)
    sg, so = gm.scalars(), om.scalars()
    for k in sg:
        assert abs(sg[k] - so[k]) <= 1e-12 * abs(so[k]), k
    for name in ("psi", "zeta", "zeta_old", "rel", "ui", "uj", "zeta_fi", "zeta_fj", "res",
                 "res_beta", "flux_i", "flux_j"):
        a, b = gm.get(name), om.a[name]
        den = np.max(np.abs(b))
        assert np.max(np.abs(a - b)) <= 1e-9 * (den if den > 0 else 1.0), name


@pytest.mark.parametrize("L,bc", [(3, 1), (1, 2), (4, 3)])
def test_model_step_is_the_stage_sequence(dev, L, bc):
    """The model's step (fused kernels: predictor + beta, rel + zeta_bc +
    cell boundaries, extrapolator + flux, corrector + zeta_old + face copy)
    against the stage entry points called in the reference's order
    (qg3.F90:782-960) on copies of the same state, with the inversion by an
    FP64 handle of the same stratification: bit for bit after 3 full steps
    and after a final partial step."""
    import torch
    print(f"model step vs stages L={L} bc={bc}", flush=True)
    g = 33
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    rng = np.random.default_rng(L + 10 * bc)
    if L == 3:
        st = nd["strat"]
    else:
        s1, s2 = random_strat(L, rng)
        st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    beta = np.full(L, float(np.asarray(nd["beta"]).ravel()[0]))
    fw = c.wind()
    kw = dict(viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=bc,
              alpha=nd["alpha"])
    gm = qg.Model(g, st, beta, fw, **kw)
    q = c.q()[:, :, :L] if L <= 3 else np.concatenate([c.q(), c.q(member=1)[:, :, :L - 3]], axis=2)
    gm.init(np.asfortranarray(q))
    names = ("psi", "zeta", "rel", "zeta_old", "ui", "uj", "zeta_fi", "zeta_fj", "res", "res_beta",
             "uimo", "ujmo", "flux_i", "flux_j")
    T = {n: to_dev(gm.get(n), dev) for n in names}
    hc = qg.HelmholtzCache(g, st, "fp64", tiles=None)
    t1 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    ts1, ts2, tb = t1(st.s1), t1(st.s2), t1(beta)
    tfw = torch.from_numpy(np.ascontiguousarray(fw.T)).to(dev)
    sc = gm.scalars()
    dt, tw_a, tw_u = sc["dt"], sc["tweight_a"], sc["tweight_u"]

    def stage_step(last):
        qg.cabaret_predictor(dt, T["zeta_old"], T["res"], T["zeta"])
        qg.cabaret_beta(dt, tw_a, tb, T["ui"], T["zeta"], T["res_beta"])
        qg.pv_inversion(T["zeta"], hc, T["psi"])
        qg.rel_from_zeta_and_psi(ts1, ts2, bc, kw["alpha"], T["zeta"], T["psi"], T["rel"])
        qg.update_zeta_bc(ts1, ts2, T["rel"], T["psi"], T["zeta"])
        qg.cabaret_forcing(dt, tfw, kw["viscosity"], kw["visc_bot"], T["rel"], T["zeta"])
        qg.cabaret_cell_boundaries(tw_u, T["psi"], T["ui"], T["uj"], T["uimo"], T["ujmo"])
        qg.cabaret_extrapolator(dt, T["zeta"], T["zeta_old"], T["ui"], T["uj"], T["zeta_fi"],
                                T["zeta_fj"])
        qg.cabaret_flux(T["ui"], T["uj"], T["zeta_fi"], T["zeta_fj"], T["flux_i"], T["flux_j"])
        if not last:
            qg.cabaret_corrector(dt, T["flux_i"], T["flux_j"], T["res"], T["zeta"])
            T["zeta_old"].copy_(T["zeta"])

    for k in range(3):
        stage_step(False)
    gm.step(3)
    torch.cuda.synchronize()
    for n in names:
        np.testing.assert_array_equal(gm.get(n), from_dev(T[n]), err_msg=f"full steps: {n}")
    stage_step(True)
    gm.step(1, final_partial=True)
    torch.cuda.synchronize()
    for n in names:
        np.testing.assert_array_equal(gm.get(n), from_dev(T[n]), err_msg=f"final partial step: {n}")
    hc.deallocate()
    gm.close()


@pytest.mark.parametrize("adapt", [False, True])
def test_model_status_nan(dev, adapt):
    """A NaN psiC in a model's inversion is reported like the reference's
    abort (qg3p_aux.F90:504-509): by Model.status() after asynchronous steps,
    and with adapt_dt by the next step itself."""
    from pequod_amd._lib import PQ_ERR_NAN, PqError
    c = synth.config("qg3_255")
    c.n = 31
    nd = c.nondim()
    kw = dict(viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=1,
              alpha=nd["alpha"], adapt_dt=adapt, target_cfl=0.1, max_dt=nd["dt"])
    m = qg.Model(c.g, nd["strat"], nd["beta"], c.wind(), **kw)
    m.init(c.q())
    m.step(2)
    assert m.status() == 0
    # the predictor starts from zeta_old (qg3p_aux.F90:993-1020)
    for name in ("zeta", "zeta_old"):
        z = m.get(name)
        z[c.g // 2, c.g // 3, 0] = np.nan
        m.set(name, z)
    if adapt:
        m.step(1)  # its inversion turns NaN; the next step's adapt_dt reads the status
        with pytest.raises(PqError) as e:
            m.step(1)
        assert e.value.code == PQ_ERR_NAN
    else:
        m.step(2)
        with pytest.raises(PqError) as e:
            m.status()
        assert e.value.code == PQ_ERR_NAN
        assert m.status() == 0  # cleared once reported
    m.close()
