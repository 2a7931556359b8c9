"""Ensembles of mixed sizes, layer counts and stratifications: the batched,
class-grouped inversion (eager and replayed from a HIP graph) gives what
each member's own handle gives (exactly for members with their class's
largest layer count; to the psiC summation order, which follows the column
tile, for the others), and a graph-replayed ensemble time step of CABARET
models gives exactly the standalone models' states."""
import numpy as np
import pytest

from pequod_amd import qg, strat as S, synth
from pequod_amd.ensemble import Ensemble

from helpers import random_strat, to_dev

pytestmark = pytest.mark.gpu

# (g, layers, seed): sizes 2^a and 3 * 2^a, classes with one and several members
MEMBERS = [(33, 3, 1), (65, 2, 2), (33, 3, 3), (49, 4, 4), (65, 2, 5), (33, 2, 6), (33, 3, 7)]


def _member(g, L, seed, dtype):
    import torch
    rng = np.random.default_rng(seed)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((g, g, L)))
    zt = to_dev(z, "cuda", dtype)
    return st, zt, torch.zeros_like(zt)


def _same(out, ref, prec, exact):
    """Ensemble vs alone: bitwise for full-layer members, else within a few
    roundings of the output type."""
    import torch
    if exact:
        return torch.equal(out, ref)
    d = float((out.double() - ref.double()).abs().max() / ref.double().abs().max())
    return d < (1e-13 if prec == "fp64" else 3e-7)


def _alone(g, st, prec, zt):
    import torch
    hc = qg.HelmholtzCache(g, st, prec, tiles=None)
    out = torch.zeros_like(zt)
    hc.run_dev([zt], [out])
    assert hc.status()[0] == 0
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_ensemble_inversion_matches_members(dev, prec):
    import torch
    print(f"ensemble inversion {prec}", flush=True)
    dt = torch.float32 if prec == "fp32" else torch.float64
    ens = Ensemble(prec)
    mem = [_member(g, L, s, dt) for g, L, s in MEMBERS]
    for st, zt, pt in mem:
        ens.add(st, zt, pt)
    ens.finalize(tiles=None)
    assert ens.class_count == 3  # one ragged class per grid size
    lmax = {g: max(L for gg, L, _ in MEMBERS if gg == g) for g, _, _ in MEMBERS}
    ens.invert()
    assert ens.status() == 0
    for (g, L, s), (st, zt, pt) in zip(MEMBERS, mem):
        assert _same(pt, _alone(g, st, prec, zt), prec, L == lmax[g]), (g, L, s)


def test_ensemble_graph_replays(dev):
    """Captured inversion replayed three times with new inputs in between
    (host-side work and synchronisation between replays); every replay
    equals the eager per-member result."""
    import torch
    dt = torch.float32
    ens = Ensemble("fp32")
    mem = [_member(g, L, s, dt) for g, L, s in MEMBERS]
    for st, zt, pt in mem:
        ens.add(st, zt, pt)
    ens.finalize(tiles=None)
    ens.capture(Ensemble.INVERT)
    rng = np.random.default_rng(99)
    for rep in range(3):
        for (g, L, s), (st, zt, pt) in zip(MEMBERS, mem):
            zt.copy_(torch.from_numpy(rng.standard_normal(tuple(zt.shape))).to(zt))
            pt.zero_()
        torch.cuda.synchronize()
        ens.invert()
        ens.synchronize()
        assert ens.status() == 0
        lmax = {g: max(L for gg, L, _ in MEMBERS if gg == g) for g, _, _ in MEMBERS}
        for (g, L, s), (st, zt, pt) in zip(MEMBERS, mem):
            assert _same(pt, _alone(g, st, "fp32", zt), "fp32", L == lmax[g]), (rep, g, L, s)


def _model(g, n_layers_seed):
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    m = qg.Model(g, st, nd["beta"], c.wind(), viscosity=nd["viscosity"], visc_bot=nd["visc_bot"],
                 dt=nd["dt"], bc_type=1, alpha=nd["alpha"])
    m.init(c.q(member=n_layers_seed))
    return m, st


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_model_step_matches_standalone(dev, graph):
    print(f"ensemble model step graph={graph}", flush=True)
    sizes = [(33, 1), (65, 2), (33, 3)]
    ens_models = [_model(g, s) for g, s in sizes]
    ref_models = [_model(g, s) for g, s in sizes]
    ens = Ensemble("fp64")
    for m, st in ens_models:
        ens.add_model(m, st)
    ens.finalize(tiles=None)
    if graph:
        ens.capture(Ensemble.STEP)
    ens.step(4)
    ens.synchronize()
    for (m, _), (r, _) in zip(ens_models, ref_models):
        r.step(4)
        for name in ("zeta", "psi", "zeta_fi", "ui"):
            np.testing.assert_array_equal(m.get(name), r.get(name), err_msg=name)
        assert m.scalars()["model_time"] == r.scalars()["model_time"]


def test_ensemble_fp32auto_within_1e6(dev):
    """Mixed ensemble members with 2..8 layers (the ensemble workload's
    stratifications, on small grids) plus an ill-conditioned 8-layer member
    (kappa_inf 85.8): the FP32-storage ensemble with the per-class transform
    choice stays within 1e-6 max relative psi error of the FP64 oracle for
    every member; the ensemble's classes use FP32 transforms, the steep
    member's class FP64 transforms."""
    import torch
    from oracle import oracle as O
    from helpers import from_dev, rel_err
    cfgs = []
    for k in range(40):
        c = synth.ensemble_member(k)
        if c.layers not in [x.layers for x in cfgs]:
            c.n = 127 if k % 2 else 191
            cfgs.append(c)
    assert sorted(c.layers for c in cfgs) == list(range(2, 9))
    cfgs.append(synth.steep_member(8, 2.5, n=191))
    ens = Ensemble("fp32auto")
    zs, ps, refs = [], [], []
    for c in cfgs:
        st = c.stratification()
        q = c.q(member=c.extra["member"])
        refs.append(O.Inversion(c.g, st.ss, st.modes_to_z, st.z_to_modes).run(q)[0])
        zs.append(to_dev(q, "cuda", torch.float32))
        ps.append(torch.zeros_like(zs[-1]))
        ens.add(st, zs[-1], ps[-1])
    ens.finalize()
    precs = {ens.class_precision(k) for k in range(ens.class_count)}
    assert precs == {"fp32", "fp32s"}, precs
    ens.invert()
    assert ens.status() == 0
    for c, p, ref in zip(cfgs, ps, refs):
        err = rel_err(from_dev(p), ref)
        print(f"layers {c.layers} g {c.g}: {err:.2e}", flush=True)
        assert err < 1e-6, (c.layers, err)
