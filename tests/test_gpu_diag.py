"""Parity of the device diagnostics (energy, time mean of psi) with the
FP64 oracle.

The energy reduction adds in a different order from the reference's loops
(partial sums per workgroup, then a fixed-order sum of the partials), so it
agrees with the oracle to FP64 rounding (relative 1e-12 here); it is
bit-identical from run to run.  The time-mean accumulation is elementwise
with the reference's rounding (product, then sum) and must agree bit for bit.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pequod_amd import diag, qg, synth
from pequod_amd._lib import PQ_ERR_NAN, PqError, load
from pequod_amd.ensemble import Ensemble

from helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def _strat(rng, L):
    h = rng.uniform(0.5, 2.0, L)
    s2 = rng.uniform(0.1, 3.0, L)
    s2[-1] = 0.0
    return h, s2


def _close(a, b, rtol=1e-12):
    a, b = np.atleast_1d(a), np.atleast_1d(b)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("g,L", [(5, 2), (9, 1), (65, 3), (100, 4), (257, 8), (1025, 3)])
def test_energy_dev_matches_oracle(dev, g, L):
    import torch
    print(f"energy g={g} L={L}", flush=True)
    rng = np.random.default_rng(g + 10 * L)
    psi = np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))
    h, s2 = _strat(rng, L)
    t = to_dev(psi, dev)
    e = diag.Energy(g, h=h, s2=s2)
    out = torch.zeros(L + 1, dtype=torch.float64, device=dev)
    e.energy_dev(t[:, 1:, 1:], out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ekin, epot = O.energy(psi, h, s2)
    assert _close(got[:L], ekin), (got[:L], ekin)
    assert _close(got[L], epot) if L > 1 else got[L] == 0.0
    # deterministic: the same bits again
    out2 = torch.zeros_like(out)
    e.energy_dev(t[:, 1:, 1:], out2)
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    e.close()


@pytest.mark.parametrize("case", KA["energy"]["cases"], ids=lambda c: c["name"])
def test_energy_known_answers_on_device(dev, case):
    """unittest/test_energy_2.F90's fields on a 9 x 9 grid (the answers do
    not depend on the grid size)."""
    ka = KA["energy"]
    g = 9
    s21, _ = O.strat_l2(ka["rd"], *ka["h"])
    psi = np.zeros((g + 2, g + 2, 2), order="F")
    i = np.arange(1, g + 1, dtype=np.float64)
    for k in range(2):
        psi[1:g + 1, :, k] = (case["a"][k] * i + case["b"][k])[:, None]
    e = diag.Energy(g, h=ka["h"], s2=[s21, 0.0])
    ekin, epot = e.energy(psi[1:, 1:, :])
    assert abs(epot - case["epot"]) <= 1e-14
    assert np.all(np.abs(ekin - case["ekin"]) <= 1e-14)
    e.close()


def test_energy_host_dropin_with_leading_dims(dev):
    """pq_diag_energy on a host g x g view inside a larger array."""
    g, L = 33, 3
    rng = np.random.default_rng(7)
    big = np.asfortranarray(rng.standard_normal((g + 9, g + 4, L)))
    h, s2 = _strat(rng, L)
    e = diag.Energy(g, h=h, s2=s2)
    ekin, epot = e.energy(big)
    data = np.zeros((g + 2, g + 2, L), order="F")
    data[1:g + 1, 1:g + 1, :] = big[:g, :g, :]
    ek2, ep2 = O.energy(data, h, s2)
    assert _close(ekin, ek2) and _close(epot, ep2)
    e.close()


def test_energy_nan_is_reported(dev):
    g, L = 17, 2
    psi = np.zeros((g, g, L), order="F")
    psi[3, 4, 1] = np.nan
    e = diag.Energy(g, h=[1.0, 1.0], s2=[1.0, 0.0])
    with pytest.raises(PqError) as ei:
        e.energy(psi)
    assert ei.value.code == PQ_ERR_NAN
    e.close()


def test_psi_av_add_bitwise(dev):
    g, L = 70, 3
    rng = np.random.default_rng(11)
    av = np.zeros((g, g, L), order="F")
    tav = to_dev(av, dev)
    for dt in (0.3, 0.7, 1.1):
        psi = np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))
        diag.psi_av_add(to_dev(psi, dev)[:, 1:, 1:], tav, dt, g)
        O.psi_av_add(psi, av, dt)
    np.testing.assert_array_equal(from_dev(tav), av)


def _model_pair(g=65):
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    fw = c.wind()
    kw = dict(viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=1,
              alpha=nd["alpha"])
    gm = qg.Model(g, st, nd["beta"], fw, **kw)
    om = O.Model(g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], fw, **kw)
    q = c.q()
    gm.init(q)
    om.init(q)
    return gm, om, st, nd["dt"]


def test_model_time_mean_and_energy_match_oracle(dev):
    """Averaging that starts after a spin-up of 3.2 steps: samples from the
    4th step on, weighted by dt; energy of the final psi."""
    gm, om, st, dt = _model_pair()
    t0 = gm.scalars()["model_time"]
    max_spin = t0 + 3.2 * dt
    gm.average(True, max_spin)
    om.average(max_spin)
    with pytest.raises(PqError):
        gm.psi_av()  # no samples yet
    gm.step(8)
    om.step(8)
    a, t = gm.psi_av()
    b, tb = om.psi_av()
    assert t == tb and abs(t - 5 * dt) <= 1e-12 * t
    den = np.max(np.abs(b))
    assert np.max(np.abs(a - b)) <= 1e-9 * den
    ekin, epot = gm.energy()
    ek2, ep2 = O.energy(om.a["psi"], st.h, st.s2)
    assert np.all(np.abs(ekin - ek2) <= 1e-7 * np.abs(ek2)) and abs(epot - ep2) <= 1e-7 * ep2
    # and exactly the oracle's energy of the device model's own psi
    ek3, ep3 = O.energy(gm.get("psi"), st.h, st.s2)
    assert _close(ekin, ek3) and _close(epot, ep3)


def test_ensemble_capture_with_time_mean(dev):
    """A captured ensemble step holds the accumulation: refused while the
    spin-up is not over; once it is, replays match a standalone model bit for
    bit (the ensemble's step is the standalone step)."""
    gm, _, st, dt = _model_pair(33)
    ref, _, _, _ = _model_pair(33)
    gm.average(True, gm.scalars()["model_time"] + 5 * dt)
    ens = Ensemble("fp64")
    ens.add_model(gm, st)
    ens.finalize(tiles=None)
    with pytest.raises(PqError):
        ens.capture(Ensemble.STEP)
    gm.average(True, 0.0)
    ref.average(True, 0.0)
    ens.capture(Ensemble.STEP)
    ens.step(3)
    ens.synchronize()
    ref.step(3)
    a, t = gm.psi_av()
    b, tb = ref.psi_av()
    assert t == tb
    np.testing.assert_array_equal(a, b)
    # switching averaging off makes the captured step stale
    gm.average(False)
    with pytest.raises(PqError):
        ens.step(1)
    ens.close()
