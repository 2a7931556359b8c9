"""Parity of the periodic channel (helm periodic in j, and pv_inversion with
the momentum-conserving correction) on the GPU with the FP64 oracle.

The oracle restates the reference's periodic path (poinit_j,
periodic_coefficients, periodic_ge, correction_function_momentum,
zeta_boundary, modes_to_z_correct_momentum); the GPU path replaces the
Thomas sweep plus periodic elimination along j with a periodic real FFT of
length gj = 2 (gi - 1) and applies the correction in the i-spectral domain.
Both solve the same discrete system, so FP64 results agree to round-off
(the closed basin's bound); the FP32 paths are held to 1e-6 on the seeded
k^-3 fields and 2e-6 on white noise, as in the closed basin.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pequod_amd import qg, strat as S, synth

from helpers import from_dev, random_strat, rel_err, to_dev

pytestmark = pytest.mark.gpu

SIZES = [9, 13, 17, 33, 65, 129, 257, 1025]


def fp64_tol(g, m2z=None):
    N = g - 1
    k = 1.0 if m2z is None else max(1.0, np.linalg.cond(m2z) / 4.0)
    return 4e-12 * max(1.0, (N / 256.0) ** 2) * k


def _strat(L, seed):
    rng = np.random.default_rng(seed)
    s1, s2 = random_strat(L, rng)
    return S.from_parameters(np.ones(L), s1[1:], s2[:-1])


@pytest.mark.parametrize("g", SIZES)
def test_periodic_helm_fp64_matches_oracle(dev, g):
    print(f"periodic helm fp64 g={g}", flush=True)
    gj = 2 * (g - 1)
    rng = np.random.default_rng(g)
    ss = np.array([0.0, 0.3, 1.7])
    z = np.asfortranarray(rng.standard_normal((g, gj, 3)))
    ref = z.copy(order="F")
    O.PeriodicHelm(g, gj, ss).solve(ref)
    out = z.copy(order="F")
    qg.HelmCache(g, gj, ss, periodic=True).helm(out)  # host drop-in
    assert rel_err(out[1:-1, :], ref[1:-1, :]) < fp64_tol(g)
    # boundary semantics: i = 1 keeps its input, i = gi is zeroed
    np.testing.assert_array_equal(out[0, :], ref[0, :])
    np.testing.assert_array_equal(out[-1, :], ref[-1, :])


@pytest.mark.parametrize("prec", ["fp32", "fp32w", "fp32s"])
@pytest.mark.parametrize("g", [9, 65, 257, 1025])
def test_periodic_helm_fp32_device(dev, g, prec):
    import torch
    print(f"periodic helm {prec} g={g}", flush=True)
    gj = 2 * (g - 1)
    rng = np.random.default_rng(300 + g)
    ss = np.array([0.0, 0.05, 0.4])
    z = np.asfortranarray(rng.standard_normal((g, gj, 3)).astype(np.float32).astype(np.float64))
    ref = z.copy(order="F")
    O.PeriodicHelm(g, gj, ss).solve(ref)
    t = to_dev(z, dev, torch.float32)
    qg.HelmCache(g, gj, ss, prec, periodic=True).helm(t)
    torch.cuda.synchronize()
    out = from_dev(t)
    assert rel_err(out[1:-1, :], ref[1:-1, :]) < 2e-6
    np.testing.assert_array_equal(out[0, :], ref[0, :])
    np.testing.assert_array_equal(out[-1, :], ref[-1, :])


@pytest.mark.parametrize("g", [17, 65, 257])
def test_periodic_helm_convergence_orders(dev, g):
    """The reference's periodic order test problem (unittest/test_helmholtz.F90:
    248-356) on the GPU: the error against cos(2 pi x) sin(pi y) equals the
    oracle's (the discretisation error, not round-off)."""
    gj = 2 * (g - 1)
    i = np.arange(1, g + 1)[:, None]
    j = np.arange(1, gj + 1)[None, :]
    x, y = (j - 1) / gj, (g - i) / (g - 1)
    s = np.cos(2 * np.pi * x) * np.sin(np.pi * y)
    for ss in (0.0, 0.5):
        if ss == 0.0:
            phi = -np.pi ** 2 * s * (1.0 / (g - 1) ** 2 + 4.0 / gj ** 2)
        else:
            phi = -s * (np.pi ** 2 / (g - 1) ** 2 + 4 * np.pi ** 2 / gj ** 2 + ss)
        a = np.asfortranarray(phi[:, :, None])
        b = a.copy(order="F")
        qg.HelmCache(g, gj, [ss], periodic=True).helm(a)
        O.PeriodicHelm(g, gj, [ss]).solve(b)
        a[[0, -1], :, 0] = 0.0
        b[[0, -1], :, 0] = 0.0
        ea, eb = np.max(np.abs(a[:, :, 0] - s)), np.max(np.abs(b[:, :, 0] - s))
        assert abs(ea - eb) <= 1e-3 * eb, (ss, ea, eb)


@pytest.mark.parametrize("g", SIZES)
@pytest.mark.parametrize("L", [1, 2, 3, 8])
def test_channel_fp64_matches_oracle(dev, g, L):
    print(f"channel fp64 g={g} L={L}", flush=True)
    gj = 2 * (g - 1)
    st = _strat(L, 40 + L)
    rng = np.random.default_rng(g * 100 + L)
    z = np.asfortranarray(rng.standard_normal((g, gj, L)))
    bc, alpha = 1, 0.25
    ref, rc = O.Channel(g, gj, st.ss, st.modes_to_z, st.z_to_modes, bc, alpha).run(z)
    assert rc == 0
    out = np.zeros((g, gj, L), order="F")
    hc = qg.ChannelCache(g, st, bc, alpha)
    assert hc.run_host(z, out) == 0
    assert rel_err(out, ref) < fp64_tol(g, st.modes_to_z)


@pytest.mark.parametrize("bc_type,alpha", [(1, 0.0), (1, 3.5), (2, 0.0), (3, 0.0)])
def test_channel_boundary_conditions_and_psic(dev, bc_type, alpha):
    """Every velocity boundary condition; psiC pairs as the oracle's."""
    g, L = 65, 4
    gj = 2 * (g - 1)
    st = _strat(L, 7)
    z = np.asfortranarray(np.random.default_rng(bc_type).standard_normal((g, gj, L)))
    och = O.Channel(g, gj, st.ss, st.modes_to_z, st.z_to_modes, bc_type, alpha)
    ref, _ = och.run(z)
    hc = qg.ChannelCache(g, st, bc_type, alpha)
    out = np.zeros((g, gj, L), order="F")
    hc.run_host(z, out)
    assert rel_err(out, ref) < fp64_tol(g, st.modes_to_z)
    rc, pc = hc.status()
    assert rc == 0
    opc = och.psic()  # (2, L-1)
    assert np.allclose(pc[0, 1:, :].T, opc, rtol=1e-10, atol=1e-12 * np.max(np.abs(opc)) + 1e-300)
    assert np.all(pc[0, 0, :] == 0.0)


def test_channel_device_batch_views(dev):
    """A batch of members on device views with leading dimensions larger
    than the channel (the sdata slices of (gi+2, gj+2, L) data arrays)."""
    import torch
    g, L, B = 33, 3, 3
    gj = 2 * (g - 1)
    st = _strat(L, 11)
    hc = qg.ChannelCache(g, st, 2, 0.0, batch_max=B)
    och = O.Channel(g, gj, st.ss, st.modes_to_z, st.z_to_modes, 2, 0.0)
    zs, ps, refs = [], [], []
    for b in range(B):
        z = np.asfortranarray(np.random.default_rng(b).standard_normal((g + 2, gj + 2, L)))
        refs.append(och.run(z[1:g + 1, 1:gj + 1, :])[0])
        zs.append(to_dev(z, dev))
        ps.append(torch.zeros_like(zs[-1]))
    hc.run_dev([t[:, 1:, 1:] for t in zs], [t[:, 1:, 1:] for t in ps])
    torch.cuda.synchronize()
    for b in range(B):
        out = from_dev(ps[b])
        assert rel_err(out[1:g + 1, 1:gj + 1, :], refs[b]) < fp64_tol(g, st.modes_to_z)
        # nothing outside the views is written
        assert np.all(out[0, :, :] == 0.0) and np.all(out[:, 0, :] == 0.0)
        assert np.all(out[g + 1, :, :] == 0.0) and np.all(out[:, gj + 1, :] == 0.0)


@pytest.mark.parametrize("prec", ["fp32", "fp32w", "fp32s", "mixed"])
@pytest.mark.parametrize("g,L", [(257, 3), (1025, 3), (513, 6)])
def test_channel_fp32_within_1e6(dev, prec, g, L):
    """Reduced-precision channels on seeded k^-3 PV fields (the closed
    basin's workload statistics) against the FP64 oracle."""
    import torch
    print(f"channel {prec} g={g} L={L}", flush=True)
    gj = 2 * (g - 1)
    c = synth.config("qg6_2047" if L == 6 else "qg3_1023")
    st = c.stratification()
    # two seeded k^-3 fields side by side along the period (each is zero on
    # its edge columns)
    q = np.zeros((g, gj, L), order="F")
    for half in range(2):
        f = synth.k3_field(g - 2, L, synth.seed_for(7, half), 1e-5)
        q[:, half * (g - 1):(half + 1) * (g - 1), :] = f[:, :g - 1, :]
    ref, _ = O.Channel(g, gj, st.ss, st.modes_to_z, st.z_to_modes, 1, 0.1).run(q)
    dt = torch.float64 if prec == "mixed" else torch.float32
    zt = to_dev(q, dev, dt)
    pt = torch.zeros_like(zt)
    hc = qg.ChannelCache(g, st, 1, 0.1, precision=prec)
    hc.run_dev([zt], [pt])
    torch.cuda.synchronize()
    assert rel_err(from_dev(pt), ref) < 1e-6
