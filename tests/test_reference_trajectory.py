"""The oracle's CABARET model against the reference's own trajectory.

tests/golden/ref_*.npz hold states that pequod's Fortran routines reached in
this container (tests/golden/make_reference_fixtures.py builds them
unchanged and drives them in the order of qg3.F90's start and time loop).
The oracle restates those routines with its own FFT in the sine
transforms, so the two agree to accumulated round-off: every field after 1,
2 and 10 steps on 17^2 x 3 (three boundary conditions), norms and samples
on 33^2, and the 255^2-interior workload's 100 adaptive steps, dt by dt.
"""
import numpy as np
import pytest

from oracle import oracle as O

from helpers import load_reference_fixture

FIELDS = ["psi", "zeta", "rel", "zeta_old", "zeta_fi", "ui", "uimo", "flux_i", "zeta_fj", "uj",
          "ujmo", "flux_j", "res", "res_beta"]
NODE = ("psi", "zeta", "rel", "zeta_old")


def owned(name, a):
    """Node fields on their owned nodes 2..G+1: the halo cells are never
    read on one domain and the reference leaves some uninitialised."""
    return a[1:-1, 1:-1, :] if name in NODE else a


def oracle_model(inp, bc=1, adapt=False):
    st = inp["strat"]
    m = O.Model(inp["g"], st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, inp["beta"],
                inp["force_wind"], viscosity=inp["viscosity"], visc_bot=inp["visc_bot"],
                dt=inp["dt"], bc_type=bc, alpha=inp["alpha"], adapt_dt=adapt, target_cfl=0.1,
                max_dt=48.0 * inp["dt"], max_dt_increase=1.5)
    assert m.init(inp["zeta0"]) == 0
    return m


def rel_dev(a, b):
    """max over layers of max|a - b| / max|b|."""
    return max(np.max(np.abs(a[..., l] - b[..., l])) / max(np.max(np.abs(b[..., l])), 1e-300)
               for l in range(b.shape[-1]))


@pytest.mark.parametrize("bc", [1, 2, 3])
def test_oracle_matches_reference_fields_17(bc):
    fx, inp = load_reference_fixture(f"ref_traj_17_bc{bc}.npz")
    m = oracle_model(inp, bc)
    done = 0
    for s in (1, 2, 10):
        if f"s{s}_psi" not in fx:
            continue
        m.step(s - done)
        done = s
        sc = m.scalars()
        for k in ("dt", "tweight_a", "tweight_u", "model_time"):
            assert sc[k] == pytest.approx(float(fx[f"s{s}_{k}"]), rel=1e-14), (s, k)
        for name in FIELDS:
            key = f"s{s}_{name}"
            if key in fx:
                d = rel_dev(owned(name, m.a[name]), owned(name, fx[key]))
                assert d < 1e-13, (s, name, d)


def test_oracle_matches_reference_33():
    fx, inp = load_reference_fixture("ref_traj_33_bc1.npz")
    m = oracle_model(inp)
    m.step(10)
    for name in ("psi", "zeta", "zeta_fi", "ui", "res"):
        a = m.a[name]
        np.testing.assert_allclose(np.max(np.abs(a), axis=(0, 1)), fx[f"s10_{name}_linf"], rtol=1e-11)
        np.testing.assert_allclose(np.sqrt(np.sum(a * a, axis=(0, 1))), fx[f"s10_{name}_l2"],
                                   rtol=1e-11)
        assert rel_dev(a[::4, ::4, :], fx[f"s10_{name}_sub"]) < 1e-11


def test_oracle_matches_reference_255_interior_adaptive():
    """The 255^2-interior 3-layer workload for 100 steps with the reference's
    adaptive time step: dt = 1800 s gives Courant number ~1.5 on this field,
    so the reference's own default (adapt_dt, target Courant number 0.1)
    sets the steps.  Every dt and the final state agree."""
    fx, inp = load_reference_fixture("ref_traj_257_adapt.npz")
    m = oracle_model(inp, adapt=True)
    dts = []
    for _ in range(int(fx["nsteps"])):
        m.step(1)
        dts.append(m.scalars()["dt"])
    np.testing.assert_allclose(dts, fx["dts"], rtol=1e-12)
    assert m.scalars()["model_time"] == pytest.approx(float(fx["s100_model_time"]), rel=1e-13)
    for name in ("psi", "zeta", "zeta_fi", "ui", "res"):
        a = m.a[name]
        np.testing.assert_allclose(np.max(np.abs(a), axis=(0, 1)), fx[f"s100_{name}_linf"],
                                   rtol=1e-11)
        assert rel_dev(a[::16, ::16, :], fx[f"s100_{name}_sub"]) < 1e-11, name


def test_oracle_matches_reference_dt_sequence():
    """300 adaptive steps on 65^2 x 3: the same dt at every step, hence the
    same step count to any model time."""
    fx, inp = load_reference_fixture("ref_dt_65_adapt.npz")
    m = oracle_model(inp, adapt=True)
    dts = []
    for _ in range(int(fx["nsteps"])):
        m.step(1)
        dts.append(m.scalars()["dt"])
    np.testing.assert_allclose(dts, fx["dts"], rtol=1e-10)
    assert m.scalars()["model_time"] == pytest.approx(float(fx["model_time"]), rel=1e-12)
    for name in ("psi", "zeta"):
        assert rel_dev(m.a[name][::4, ::4, :], fx[f"s300_{name}_sub"]) < 1e-9, name
