"""Serving engine: field members (PV inversions) and CABARET model members
(n-step advances) served from one request stream with continuous batching.
Each result is the same in the mix as alone: inversions to the psiC
summation order of a ragged batch (a few FP32 roundings), model advances
exactly; admission follows the device-footprint cap."""
import time

import numpy as np
import pytest

from pequod_amd import qg, strat as S, synth
from pequod_amd._lib import PqError
from pequod_amd.serve import Server

from helpers import from_dev, random_strat, rel_err, to_dev

pytestmark = pytest.mark.gpu

FIELDS = [(33, 3, 1), (65, 2, 2), (33, 2, 3), (65, 5, 4), (129, 4, 5), (33, 6, 6), (97, 8, 7)]


def _field(g, L, seed):
    import torch
    rng = np.random.default_rng(seed)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((g, g, L)))
    zt = to_dev(z, "cuda", torch.float32)
    return st, z, zt, torch.zeros_like(zt)


def _model_args(g, seed):
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    return dict(g=g, strat=nd["strat"], beta=nd["beta"], force_wind=c.wind(),
                viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=1,
                alpha=nd["alpha"]), c.q(member=seed)


def test_serve_mix_matches_alone(dev):
    import torch
    print("serve mix", flush=True)
    srv = Server("fp32", cap_bytes=64e9, max_batch=4)
    fields = [_field(g, L, s) for g, L, s in FIELDS]
    fid = [srv.open_field(st, zt, pt) for st, _, zt, pt in fields]
    margs = [_model_args(33, 1), _model_args(65, 2)]
    mid = []
    for kw, q in margs:
        kw = dict(kw)
        g = kw.pop("g")
        mid.append(srv.open_model(g, kw.pop("strat"), kw.pop("beta"), kw.pop("force_wind"), q, **kw))
    steps = {mid[0]: [2, 3, 5], mid[1]: [7]}
    # interleaved submissions, some while earlier work is on the device
    reqs = []
    for rep in range(3):
        for k, m in enumerate(fid):
            reqs.append(srv.submit(m, Server.FIELD))
            if k == 2 and steps[mid[0]][rep:rep + 1]:
                reqs.append(srv.submit(mid[0], Server.MODEL, steps[mid[0]][rep]))
        if rep == 0:
            reqs.append(srv.submit(mid[1], Server.MODEL, 7))
        time.sleep(0.002)
    srv.drain(120)
    lat = [srv.poll(r) for r in reqs]
    assert all(x is not None and x > 0 for x in lat)
    st = srv.stats()
    assert st["count"] == len(reqs) and st["failed"] == 0
    assert st["p50_ms"] <= st["p95_ms"] <= st["p99_ms"] <= st["max_ms"]
    # inversions: as a standalone handle gives (same inputs, same precision)
    for (g, L, s), (strat, z, zt, pt) in zip(FIELDS, fields):
        hc = qg.HelmholtzCache(g, strat, "fp32", tiles=None)
        ref = torch.zeros_like(zt)
        hc.run_dev([zt], [ref])
        assert hc.status()[0] == 0
        d = float((pt.double() - ref.double()).abs().max() / ref.double().abs().max())
        assert d < 3e-7, (g, L, s, d)
    # model advances: exactly the standalone model's state after as many steps
    for m, (kw, q) in zip(mid, margs):
        kw = dict(kw)
        g = kw.pop("g")
        ref = qg.Model(g, kw.pop("strat"), kw.pop("beta"), kw.pop("force_wind"), **kw)
        ref.init(q)
        ref.step(sum(steps[m]))
        for name in ("zeta", "psi", "zeta_fi", "ui"):
            np.testing.assert_array_equal(srv.model_get(m, name, g, 3), ref.get(name), err_msg=name)
        ref.close()
    srv.close()


def test_serve_field_against_oracle(dev):
    """A served ragged batch of the ensemble workload's stratifications
    (2..8 layers, k^-3 PV) within 1e-6 of the FP64 oracle."""
    import torch
    from oracle import oracle as O
    srv = Server("fp32", cap_bytes=64e9, max_batch=8)
    cfgs, outs, refs = [], [], []
    for k in range(40):
        c = synth.ensemble_member(k)
        if c.layers in [x.layers for x in cfgs]:
            continue
        c.n = 127
        cfgs.append(c)
    keep = []
    for c in cfgs:
        st = c.stratification()
        q = c.q(member=c.extra["member"])
        refs.append(O.Inversion(c.g, st.ss, st.modes_to_z, st.z_to_modes).run(q)[0])
        zt = to_dev(q, "cuda", torch.float32)
        pt = torch.zeros_like(zt)
        keep.append((zt, pt))
        srv.submit(srv.open_field(st, zt, pt), Server.FIELD)
    srv.drain(60)
    for c, (zt, pt), ref in zip(cfgs, keep, refs):
        err = rel_err(from_dev(pt), ref)
        print(f"served layers {c.layers}: {err:.2e}", flush=True)
        assert err < 1e-6
    srv.close()


def test_serve_admission_by_footprint(dev):
    """An open that would take the reserved bytes above the cap fails and
    changes nothing; closing a member makes room again."""
    import torch
    st, _, zt, pt = _field(65, 3, 11)
    probe = Server("fp32", cap_bytes=1e12, max_batch=2)
    probe.open_field(st, zt, pt)
    one = probe.used_bytes
    probe.close()
    srv = Server("fp32", cap_bytes=one + 1000, max_batch=2)
    a = srv.open_field(st, zt, pt)
    assert srv.used_bytes == one
    zt2, pt2 = torch.zeros_like(zt), torch.zeros_like(pt)
    with pytest.raises(PqError) as e:
        srv.open_field(st, zt2, pt2)
    assert "cap" in str(e.value)
    assert srv.used_bytes == one
    srv.submit(a, Server.FIELD)
    srv.drain(30)
    srv.close_member(a)
    b = srv.open_field(st, zt2, pt2)
    srv.submit(b, Server.FIELD)
    srv.drain(30)
    assert srv.stats()["count"] == 2
    srv.close()


def test_serve_nan_member_fails(dev):
    """A served inversion whose psiC is NaN fails with PQ_ERR_NAN (the
    reference aborts, qg3p_aux.F90:504-509); the other members of its batch
    complete, and the member's next clean request succeeds.  A model
    member's advance whose inversions turn NaN fails the same way."""
    import ctypes

    import torch
    from pequod_amd._lib import PQ_ERR_NAN, load
    print("serve NaN member", flush=True)
    srv = Server("fp32", cap_bytes=64e9, max_batch=4)
    st, _, zt, pt = _field(33, 3, 21)
    st2, _, zt2, pt2 = _field(33, 2, 22)
    bad_z = zt.clone()
    bad_z[1, 5, 7] = float("nan")
    good = srv.open_field(st2, zt2, pt2)
    bad = srv.open_field(st, bad_z, torch.zeros_like(pt))
    r_bad, r_good = srv.submit(bad, Server.FIELD), srv.submit(good, Server.FIELD)
    srv.drain(30)
    with pytest.raises(PqError) as e:
        srv.poll(r_bad)
    assert e.value.code == PQ_ERR_NAN and "NaN" in str(e.value)
    assert srv.poll(r_good) is not None
    assert srv.stats()["failed"] == 1
    bad_z.copy_(zt)
    r = srv.submit(bad, Server.FIELD)
    srv.drain(30)
    assert srv.poll(r) is not None
    # a model member whose state turns NaN
    kw, q = _model_args(33, 3)
    kw = dict(kw)
    g = kw.pop("g")
    m = srv.open_model(g, kw.pop("strat"), kw.pop("beta"), kw.pop("force_wind"), q, **kw)
    r_ok = srv.submit(m, Server.MODEL, 3)
    srv.drain(30)
    assert srv.poll(r_ok) is not None
    # the predictor starts from zeta_old (qg3p_aux.F90:993-1020)
    for name in ("zeta", "zeta_old"):
        z = srv.model_get(m, name, g, 3)
        z[g // 2, g // 2, 1] = np.nan
        z = np.asfortranarray(z)
        rc = load().pq_model_set(load().pq_srv_model(srv._h, m), qg.FIELDS[name],
                                 z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), z.size)
        assert rc == 0
    r_nan = srv.submit(m, Server.MODEL, 12)
    srv.drain(30)
    with pytest.raises(PqError) as e:
        srv.poll(r_nan)
    assert e.value.code == PQ_ERR_NAN
    srv.close()
