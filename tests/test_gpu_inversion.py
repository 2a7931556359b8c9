"""Parity of the HIP Helmholtz solve and PV inversion with the FP64 oracle.

The oracle restates helm / pv_inversion (qg3p_helmholtz.F90:614-769,
qg3p_aux.F90:452-521) with FFTPACK-style sine transforms along i and the
Thomas sweep along j; the GPU path uses sine transforms along both axes.
Both solve the same discrete system, so FP64 results agree to round-off.
Tolerances: FP64 max relative error 4e-12 per layer (scaled by N^2 beyond
N = 256 and by the projection condition number); FP32 and mixed 1e-6
(the stated bound for the reduced-precision paths).
"""
import numpy as np
import pytest

from oracle import oracle as O
from pequod_amd import qg, strat as S, synth
from pequod_amd._lib import IO_BYTES, PQ_ERR_NAN, PqError

from helpers import from_dev, random_strat, rel_err, to_dev

pytestmark = pytest.mark.gpu

SIZES = [9, 13, 17, 25, 33, 49, 65, 97, 129, 257, 769, 1025]


def fp64_tol(g, m2z=None):
    """FP64 parity bound: both solvers carry round-off amplified by the
    condition number of the 5-point operator (~N^2 for the Poisson mode),
    so the bound grows with N^2 beyond N = 256, and by the condition number
    of the layer <-> mode projection."""
    N = g - 1
    k = 1.0 if m2z is None else max(1.0, np.linalg.cond(m2z) / 4.0)
    return 4e-12 * max(1.0, (N / 256.0) ** 2) * k


def residual(psi, rhs, ss):
    """Discrete residual of (Delta_5 - ss) psi = rhs on the interior,
    relative to the backward-error scale (|Delta_5| + ss) max|psi|, max over
    layers.  A backward-stable solve gives a few eps."""
    r = 0.0
    for l in range(psi.shape[2]):
        p = psi[:, :, l]
        lap = p[2:, 1:-1] + p[:-2, 1:-1] + p[1:-1, 2:] + p[1:-1, :-2] - 4 * p[1:-1, 1:-1]
        res = lap - ss[l] * p[1:-1, 1:-1] - rhs[1:-1, 1:-1, l]
        r = max(r, np.max(np.abs(res)) / ((8.0 + ss[l]) * np.max(np.abs(p))))
    return r


@pytest.mark.parametrize("g", SIZES)
@pytest.mark.parametrize("L", [1, 3])
def test_helm_fp64_matches_oracle(dev, g, L):
    print(f"helm fp64 g={g} L={L}", flush=True)
    rng = np.random.default_rng(g * 10 + L)
    ss = np.concatenate([[0.0], rng.uniform(0.01, 1.0, L - 1)])
    z = np.asfortranarray(rng.standard_normal((g, g, L)))
    ref = z.copy(order="F")
    O.Helm(g, ss).solve(ref)
    out = z.copy(order="F")
    qg.HelmCache(g, g, ss).helm(out)  # host drop-in path
    # interior to round-off, boundary semantics exactly
    assert rel_err(out[1:-1, 1:-1], ref[1:-1, 1:-1]) < fp64_tol(g)
    chk = out.copy(order="F")
    O.apply_dirichlet_bc(chk, 0.0)
    assert residual(chk, z, ss) < 1e-13
    for sl in [np.s_[0, :], np.s_[-1, :], np.s_[:, 0], np.s_[:, -1]]:
        np.testing.assert_array_equal(out[sl], ref[sl])


@pytest.mark.parametrize("g", [33, 129])
def test_helm_device_inplace(dev, g):
    import torch
    rng = np.random.default_rng(7)
    ss = np.array([0.0, 0.3])
    z = np.asfortranarray(rng.standard_normal((g, g, 2)))
    ref = z.copy(order="F")
    O.Helm(g, ss).solve(ref)
    t = to_dev(z, dev)
    qg.HelmCache(g, g, ss).helm(t)
    torch.cuda.synchronize()
    out = from_dev(t)
    assert rel_err(out[1:-1, 1:-1], ref[1:-1, 1:-1]) < 1e-12
    np.testing.assert_array_equal(out[0, :], ref[0, :])
    np.testing.assert_array_equal(out[:, -1], ref[:, -1])


@pytest.mark.parametrize("prec", ["fp32", "fp32w", "fp32s"])
@pytest.mark.parametrize("g", [9, 65, 257, 1025])
def test_helm_fp32_device_boundary(dev, g, prec):
    """FP32-storage helm on a device array: interior within the FP32 bound
    of the FP64 oracle, and helm's boundary semantics exactly (which edge
    values are zeroed and which are left as they were)."""
    import torch
    print(f"helm {prec} g={g}", flush=True)
    rng = np.random.default_rng(500 + g)
    ss = np.array([0.0, 0.05, 0.4])
    z = np.asfortranarray(rng.standard_normal((g, g, 3)).astype(np.float32).astype(np.float64))
    ref = z.copy(order="F")
    O.Helm(g, ss).solve(ref)
    t = to_dev(z, dev, torch.float32)
    qg.HelmCache(g, g, ss, prec).helm(t)
    torch.cuda.synchronize()
    out = from_dev(t)
    assert rel_err(out[1:-1, 1:-1], ref[1:-1, 1:-1]) < 2e-6
    # z is exact in FP32, so zeroed and untouched boundary values match exactly
    for sl in [np.s_[0, :], np.s_[-1, :], np.s_[:, 0], np.s_[:, -1]]:
        np.testing.assert_array_equal(out[sl], ref[sl])


@pytest.mark.parametrize("g", [17, 33, 65, 129, 257])
@pytest.mark.parametrize("p,q", [(1, 1), (3, 5), (2, 7)])
def test_helm_eigenfunction_closed_form(dev, g, p, q):
    """sin(p pi i/N) sin(q pi j/N) is an exact eigenfunction of the 5-point
    operator: psi = rhs / (lambda_p + lambda_q - ss)."""
    N = g - 1
    i = np.arange(g)
    f = np.sin(np.pi * p * i / N)[:, None] * np.sin(np.pi * q * i / N)[None, :]
    ss = np.array([0.0, 0.5])
    z = np.asfortranarray(np.stack([f, f], axis=2))
    qg.HelmCache(g, g, ss).helm(z)
    O.apply_dirichlet_bc(z, 0.0)
    lam = lambda k: -4 * np.sin(k * np.pi / (2 * N)) ** 2
    for l in range(2):
        exp = f / (lam(p) + lam(q) - ss[l])
        # round-off is bounded by eps * cond * |rhs| / |smallest eigenvalue|;
        # a wrong index or sign gives O(|exp|) errors
        scale = np.max(np.abs(f)) / abs(2 * lam(1) - ss[l])
        assert np.max(np.abs(z[:, :, l] - exp)) < 2e-16 * N * N * scale


def _oracle_inv(g, st, zeta):
    inv = O.Inversion(g, st.ss, st.modes_to_z, st.z_to_modes)
    psi, rc = inv.run(zeta)
    assert rc == 0
    return psi


@pytest.mark.parametrize("g", [9, 13, 33, 65, 97, 257, 769, 1025])
@pytest.mark.parametrize("L", [1, 2, 3, 5, 8])
def test_inversion_fp64_matches_oracle(dev, g, L):
    if g > 300 and L > 3:
        pytest.skip("large sizes covered with L <= 3")
    print(f"pv_inversion fp64 g={g} L={L}", flush=True)
    rng = np.random.default_rng(1000 + g + L)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((g, g, L)))
    ref = _oracle_inv(g, st, z)
    psi = np.zeros((g, g, L), order="F")
    hc = qg.HelmholtzCache(g, st, "fp64")
    hc.run_host(z, psi)
    assert rel_err(psi, ref) < fp64_tol(g, st.modes_to_z)


def test_inversion_sdata_views_and_batch(dev):
    """Device batch of 3 members passed as sdata views of (G+2)^2 arrays."""
    import torch
    g, L, B = 65, 3, 3
    rng = np.random.default_rng(5)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    hc = qg.HelmholtzCache(g, st, "fp64", batch_max=B)
    zs, ps, refs = [], [], []
    for b in range(B):
        full = np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))
        refs.append(_oracle_inv(g, st, np.asfortranarray(full[1:-1, 1:-1])))
        zs.append(to_dev(full, dev))
        ps.append(torch.zeros_like(zs[-1]))
    hc.run_dev([z[:, 1:, 1:] for z in zs], [p[:, 1:, 1:] for p in ps])
    rc, psic = hc.status()
    assert rc == 0
    for b in range(B):
        out = from_dev(ps[b])
        assert rel_err(out[1:-1, 1:-1], refs[b]) < 1e-12
        # halo untouched
        assert np.all(out[0, :, :] == 0) and np.all(out[-1, :, :] == 0)


@pytest.mark.parametrize("prec", ["fp32", "mixed", "fp32s", "fp32w"])
@pytest.mark.parametrize("cfg", ["qg3_255", "qg3_1023"])
def test_inversion_reduced_precision_k3(dev, prec, cfg):
    """Seeded k^-3 PV on the 3-layer 255^2 and 1023^2 workloads: max relative psi error <= 1e-6
    vs the FP64 oracle (the bound stated for the reduced-precision path), and
    the mass constraint: every output layer has a zero trapezoid mean."""
    import torch
    c = synth.config(cfg)
    st = c.stratification()
    q = c.q()
    g = c.g
    print(f"pv_inversion {prec} {cfg} g={g}", flush=True)
    ref = _oracle_inv(g, st, q)
    hc = qg.HelmholtzCache(g, st, prec)
    dt = torch.float32 if IO_BYTES[prec] == 4 else torch.float64
    zt = to_dev(q, dev, dt)
    pt = torch.zeros_like(zt)
    hc.run_dev([zt], [pt])
    assert hc.status()[0] == 0
    out = from_dev(pt)
    err = rel_err(out, ref)
    print(f"{prec} rel err {err:.3e}")
    assert err < 1e-6
    if prec == "fp32s":
        # FP64 transforms: only the FP32 roundings of input, scratch and output remain
        assert err < 3e-7
    if prec == "fp32w":
        # FP64 arithmetic, FP32 rounding between FFT stages
        assert err < 5e-7
    if prec in ("fp32", "mixed"):
        # FP32 transforms with the lowest input modes of the inverse
        # transforms summed in FP64 (measured 1.3e-7 .. 1.7e-7)
        assert err < 4e-7
    for l in range(st.layers):
        assert abs(O.mean(out[:, :, l])) < 1e-6 * np.max(np.abs(ref[:, :, l]))


@pytest.mark.parametrize("prec", ["fp32", "fp32s", "fp32w"])
@pytest.mark.parametrize("g", [9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769, 1025])
def test_inversion_fp32_all_sizes(dev, g, prec):
    """FP32 (odd-extension transforms) on every kernel unit up to 1024:
    random white-noise PV with 3 layers, error vs the FP64 oracle.  White
    noise is the hardest input for relative error (flat spectrum), so the
    bound is 2e-6 here."""
    import torch
    print(f"pv_inversion {prec} g={g}", flush=True)
    rng = np.random.default_rng(77 + g)
    L = 3
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((g, g, L)))
    ref = _oracle_inv(g, st, z)
    hc = qg.HelmholtzCache(g, st, prec)
    zt = to_dev(z, dev, torch.float32)
    pt = torch.zeros_like(zt)
    hc.run_dev([zt], [pt])
    assert hc.status()[0] == 0
    assert rel_err(from_dev(pt), ref) < 2e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32s", "fp32w"])
def test_inversion_tiles_do_not_change_results(dev, prec):
    """Every accepted tile shape gives the same psi (to the psiC summation
    order, which depends on the column tile)."""
    import torch
    g, L = 257, 3
    rng = np.random.default_rng(11)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((g, g, L)))
    dt = torch.float32 if IO_BYTES[prec] == 4 else torch.float64
    hc = qg.HelmholtzCache(g, st, prec, tiles=None)
    zt = to_dev(z, dev, dt)
    base = torch.zeros_like(zt)
    hc.run_dev([zt], [base])
    b = from_dev(base)
    tried = 0
    for R in (2, 4, 8, 16):
        for C in (1, 2, 4, 8):
            if not hc.set_tiles(R, C):
                continue
            tried += 1
            out = torch.zeros_like(zt)
            hc.run_dev([zt], [out])
            assert hc.status()[0] == 0
            assert rel_err(from_dev(out), b) < (1e-13 if prec == "fp64" else 1e-6), (R, C)
    assert tried >= 4


def test_inversion_nan_status(dev):
    g, L = 33, 2
    st = S.from_parameters(np.ones(L), [0.1], [0.2])
    z = np.asfortranarray(np.ones((g, g, L)))
    z[10, 10, 0] = np.nan
    psi = np.zeros((g, g, L), order="F")
    hc = qg.HelmholtzCache(g, st, "fp64")
    with pytest.raises(PqError) as e:
        hc.run_host(z, psi)
    assert e.value.code == PQ_ERR_NAN
    # the handle recovers for the next call
    z[10, 10, 0] = 1.0
    hc.run_host(z, psi)


def test_inversion_rejects_unsupported_size(dev):
    st = S.from_parameters(np.ones(2), [0.1], [0.2])
    with pytest.raises(PqError):
        qg.HelmholtzCache(100, st, "fp64")


@pytest.mark.parametrize("prec", ["fp64", "fp32s", "fp32", "fp32w"])
def test_inversion_six_layer_2047(dev, prec):
    """The 6-layer 2047^2 workload (dx 2.5 km, FP32 storage for the reduced
    paths): every FP32-storage mode stays within 1e-6 max relative psi error
    vs the FP64 oracle (FP64 within its condition-scaled bound); FP32
    transforms measure 3.4e-7 here."""
    import torch
    c = synth.config("qg6_2047")
    st = c.stratification()
    q = c.q()
    g = c.g
    print(f"pv_inversion {prec} 6-layer g={g}", flush=True)
    ref = _oracle_inv(g, st, q)
    hc = qg.HelmholtzCache(g, st, prec)
    dt = torch.float32 if IO_BYTES[prec] == 4 else torch.float64
    zt = to_dev(q, dev, dt)
    pt = torch.zeros_like(zt)
    hc.run_dev([zt], [pt])
    assert hc.status()[0] == 0
    out = from_dev(pt)
    err = rel_err(out, ref)
    print(f"{prec} rel err {err:.3e}", flush=True)
    bound = {"fp64": fp64_tol(g, st.modes_to_z), "fp32s": 1e-6, "fp32": 1e-6, "fp32w": 1e-6}[prec]
    assert err < bound


@pytest.mark.parametrize("g,L", [(1537, 1), (1537, 3), (3073, 1), (3073, 3), (4097, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_inversion_largest_lengths(dev, g, L, prec):
    """The transform lengths N = 1536, 3072 and 4096 (the radix-24 final
    stage and the largest LDS / register footprints) against the FP64
    oracle on a seeded k^-3 field: FP64 to round-off, FP32 within 1e-6."""
    import torch
    print(f"pv_inversion {prec} g={g} L={L}", flush=True)
    c = synth.config("qg3_1023")
    c.n = g - 2
    c.H, c.gprime = c.H[:L], c.gprime[:L - 1]
    st = c.stratification()
    q = c.q(member=g + L)
    ref = _oracle_inv(g, st, q)
    dt = torch.float64 if prec == "fp64" else torch.float32
    zt = to_dev(q, dev, dt)
    pt = torch.zeros_like(zt)
    hc = qg.HelmholtzCache(g, st, prec, tiles=None)
    hc.run_dev([zt], [pt])
    assert hc.status()[0] == 0
    err = rel_err(from_dev(pt), ref)
    print(f"  max rel err {err:.2e}", flush=True)
    assert err < (fp64_tol(g, st.modes_to_z) if prec == "fp64" else 1e-6)
