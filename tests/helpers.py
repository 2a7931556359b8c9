"""Shared helpers for the tests (host side only)."""
import numpy as np


def rel_err(a, b):
    """max over layers of ||a_l - b_l||_inf / ||b_l||_inf."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 2:
        a, b = a[:, :, None], b[:, :, None]
    errs = []
    for l in range(b.shape[2]):
        den = np.max(np.abs(b[:, :, l]))
        errs.append(np.max(np.abs(a[:, :, l] - b[:, :, l])) / (den if den > 0 else 1.0))
    return max(errs)


def random_strat(L, rng):
    """Random positive stratification couplings (non-dimensional)."""
    s1 = np.zeros(L)
    s2 = np.zeros(L)
    for i in range(L - 1):
        s2[i] = rng.uniform(0.02, 0.3)
        s1[i + 1] = rng.uniform(0.02, 0.3)
    return s1, s2


def to_dev(a, dev, dtype=None):
    """Fortran-ordered numpy (ii, jj, L) -> torch tensor (L, jj, ii)."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0)))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev).contiguous()


def from_dev(t):
    """torch (L, jj, ii) -> Fortran-ordered numpy (ii, jj, L)."""
    return np.asfortranarray(t.detach().double().cpu().numpy().transpose(2, 1, 0))


def reference_case(g, member=0):
    """Non-dimensional inputs of the reference-trajectory fixtures
    (tests/golden/make_reference_fixtures.py): the 3-layer 10 km workload on
    g x g nodes with the seeded k^-3 PV field of `member`."""
    from pequod_amd import synth
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    return dict(g=g, L=3, dt=nd["dt"], beta=nd["beta"], viscosity=nd["viscosity"],
                visc_bot=nd["visc_bot"], alpha=nd["alpha"], strat=st, force_wind=c.wind(),
                zeta0=c.q(member=member))


def load_reference_fixture(name):
    """A fixture of tests/golden and its regenerated inputs (checked against
    the checksums stored with it)."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
    fx = dict(np.load(here))
    inp = reference_case(int(fx["g"]), int(fx["member"]))
    assert inp["dt"] == float(fx["in_dt"])
    assert np.sum(np.abs(inp["zeta0"])) == float(fx["in_q_abs_sum"])
    assert np.sum(np.abs(inp["force_wind"])) == float(fx["in_wind_abs_sum"])
    return fx, inp
