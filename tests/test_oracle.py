"""Pins the FP64 oracle (CPU) against the reference's own known answers
(tests/golden/known_answers.json, transcribed from its unit tests), closed
forms and invariants.  No GPU needed."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pequod_amd import strat as S, synth

from helpers import random_strat, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def _fe(a, b):
    """fe of test_tools.F90:86-94: |a - b| <= 10 max(spacing(a), epsilon)."""
    sp = np.spacing(abs(a)) if a != 0 else np.finfo(float).tiny
    return abs(a - b) <= KA["fne_factor"]["value"] * max(sp, np.finfo(float).eps)


# ------------------------------------------------------------ sine transform
def _dst_exact(x):
    """Brute-force DST-I in long double with exactly reduced angles."""
    n = len(x)
    N = n + 1
    j = np.arange(1, n + 1)
    out = np.zeros(n, dtype=np.longdouble)
    for k in range(1, n + 1):
        m = (j * k) % (2 * N)
        out[k - 1] = 2 * np.sum(np.asarray(x, np.longdouble) *
                                np.sin(np.pi * m.astype(np.longdouble) / N))
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 11, 15, 23, 31, 47, 63, 95, 127, 191, 255])
def test_sint_brute_force(n):
    x = np.random.default_rng(n).standard_normal(n)
    X = O.sint(x)
    ref = _dst_exact(x).astype(float)
    assert np.max(np.abs(X - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref))) * np.log2(n + 2)


@pytest.mark.parametrize("n", [7, 63, 767, 1023])
def test_sint_involution(n):
    """sint o sint = 2 (n+1) I (qg3p_helmholtz.F90:672-678 relies on it)."""
    x = np.random.default_rng(n).standard_normal(n)
    y = O.sint(O.sint(x)) / (2 * (n + 1))
    assert np.max(np.abs(y - x)) < 1e-13


def test_sint_linearity_and_basis():
    n = 31
    e = np.zeros(n)
    e[4] = 1.0
    X = O.sint(e)
    k = np.arange(1, n + 1)
    assert np.allclose(X, 2 * np.sin(np.pi * 5 * k / (n + 1)), atol=1e-14)


# ------------------------------------------------------------ helmholtz
@pytest.mark.parametrize("g", [9, 33, 129])
@pytest.mark.parametrize("ss", [0.0, 0.7])
def test_helm_eigenfunctions(g, ss):
    N = g - 1
    i = np.arange(g)
    for p, q in [(1, 1), (2, 5), (N - 1, 3)]:
        f = np.sin(np.pi * p * i / N)[:, None] * np.sin(np.pi * q * i / N)[None, :]
        z = np.asfortranarray(f[:, :, None].copy())
        O.Helm(g, [ss]).solve(z)
        O.apply_dirichlet_bc(z, 0.0)
        lam = lambda k: -4 * np.sin(k * np.pi / (2 * N)) ** 2
        exp = f / (lam(p) + lam(q) - ss)
        # round-off is bounded by eps * cond * |rhs| / |smallest eigenvalue|;
        # a wrong index or sign gives O(|exp|) errors
        scale = np.max(np.abs(f)) / abs(2 * lam(1) - ss)
        assert np.max(np.abs(z[:, :, 0] - exp)) < 2e-16 * N * N * scale


def test_helm_residual_random():
    """Discrete residual of the 5-point Helmholtz operator on the interior."""
    g, ss = 65, 0.3
    rng = np.random.default_rng(0)
    rhs = np.asfortranarray(rng.standard_normal((g, g, 1)))
    psi = rhs.copy(order="F")
    O.Helm(g, [ss]).solve(psi)
    O.apply_dirichlet_bc(psi, 0.0)
    p = psi[:, :, 0]
    lap = p[2:, 1:-1] + p[:-2, 1:-1] + p[1:-1, 2:] + p[1:-1, :-2] - 4 * p[1:-1, 1:-1]
    res = lap - ss * p[1:-1, 1:-1] - rhs[1:-1, 1:-1, 0]
    assert np.max(np.abs(res)) < 1e-12 * np.max(np.abs(rhs))


def _order_errors(ss, gs):
    m = []
    for g in gs:
        N = g - 1
        i = np.arange(1, g + 1)
        x = (i[None, :] - 1) / N   # j index -> x
        y = (g - i[:, None]) / N   # i index -> y
        s = np.sin(np.pi * x) * np.sin(np.pi * y)
        if ss == 0.0:
            phi = -np.pi ** 2 * s * (1.0 / N ** 2 + 1.0 / N ** 2)
        else:
            phi = -s * (np.pi ** 2 / N ** 2 + np.pi ** 2 / N ** 2 + ss)
        z = np.asfortranarray(phi[:, :, None])
        O.Helm(g, [ss]).solve(z)
        O.apply_dirichlet_bc(z, 0.0)
        m.append(np.max(np.abs(z[:, :, 0] - s)))
    return np.array(m)


def test_helm_convergence_orders():
    """The reference's order test (unittest/test_helmholtz.F90:48-136):
    Poisson order >= 2, Helmholtz (ss = 0.5) order >= 3.9."""
    ka = KA["helmholtz_order"]
    gs = ka["g"][:6]
    for ss, lim in [(0.0, ka["poisson_min_order"]), (ka["helmholtz_ss"], ka["helmholtz_min_order"])]:
        m = _order_errors(ss, gs)
        o = -np.log10(m[1:] / m[:-1]) / np.log10(np.array(gs[1:]) / np.array(gs[:-1]))
        assert np.all(o >= lim), (ss, o)


def test_laplace_bc_known_answer():
    """helm + add_laplace_bc_rhs reproduce psi = 1 (test_laplace_bc.F90:27-44)."""
    ka = KA["laplace_bc"]
    g = ka["grid"]
    phi0 = np.asfortranarray(np.ones((g, g, 1)))
    O.Helm(g, [0.0]).solve(phi0)
    O.apply_dirichlet_bc(phi0, 0.0)
    phi1 = np.asfortranarray(np.ones((g, g, 1)))
    O.add_laplace_bc_rhs(phi1, ka["bc"])
    O.Helm(g, [0.0]).solve(phi1)
    O.apply_dirichlet_bc(phi1, ka["bc"])
    d = phi1 - phi0
    assert _fe(d.min(), ka["expect_min"]) and _fe(d.max(), ka["expect_max"])
    # laplace_bc_rhs(decomp, 1) alone: zero field with the boundary RHS
    phi = np.asfortranarray(np.zeros((g, g, 1)))
    O.add_laplace_bc_rhs(phi, 1.0)
    O.Helm(g, [0.0]).solve(phi)
    O.apply_dirichlet_bc(phi, 1.0)
    assert np.allclose(phi, 1.0, atol=1e-14)


# ------------------------------------------------------------ mean
@pytest.mark.parametrize("case", KA["mean"]["cases"])
def test_mean_known_answers(case):
    n = case["n"]
    a = np.full((n, n), case["fill"])
    if case["zero_block"]:
        i0, i1, j0, j1 = case["zero_block"]
        a[i0 - 1:i1, j0 - 1:j1] = 0.0
    assert _fe(O.mean(a), case["expect"])


# ------------------------------------------------------------ stratification
def test_stratification_known_answers():
    ka = KA["stratification"]
    h = ka["h"]
    s21, s12 = O.strat_l2(ka["rd_bcl1"], h[0], h[1])
    ss, m2z, z2m = O.strat_modes([0.0, s12], [s21, 0.0])
    assert ss[0] == 0.0 and _fe(1.0 / np.sqrt(ss[1]), ka["two_layer"]["rd"][0])
    assert np.allclose(z2m @ np.ones(2), ka["two_layer"]["z_ones_to_modes"], atol=1e-15)
    a, b, c, d = O.strat_l3(ka["rd_bcl1"], ka["rd_bcl2"], *h)
    ss, m2z, z2m = O.strat_modes([0.0, b, d], [a, c, 0.0])
    # the oracle orders baroclinic modes by ascending ss (largest Rd first)
    rd = sorted(1.0 / np.sqrt(ss[1:]))
    assert all(_fe(x, y) or abs(x - y) < 1e-9 * y for x, y in zip(rd, ka["three_layer"]["rd"]))
    assert np.allclose(z2m @ np.ones(3), ka["three_layer"]["z_ones_to_modes"], atol=1e-15)
    assert np.allclose(m2z @ z2m, np.eye(3), atol=1e-14)


@pytest.mark.parametrize("L", [2, 3, 5, 8])
def test_strat_host_matches_oracle(L):
    """pequod_amd.strat (host FP64) vs the oracle's eigen-solver: same ss,
    and the same inversion operator M diag(.) D."""
    rng = np.random.default_rng(L)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    ss, m2z, z2m = O.strat_modes(s1, s2)
    assert np.allclose(st.ss, ss, rtol=1e-13, atol=1e-15)
    w = rng.standard_normal(L)
    assert np.allclose(st.modes_to_z @ np.diag(w) @ st.z_to_modes, m2z @ np.diag(w) @ z2m, atol=1e-12)
    # S = M diag(-ss) D
    assert np.allclose(st.modes_to_z @ np.diag(-st.ss) @ st.z_to_modes, st.s_mat, atol=1e-13)


# ------------------------------------------------------------ PV inversion
def _inv_case(g, L, seed):
    rng = np.random.default_rng(seed)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((g, g, L)))
    return st, z


def test_inversion_invariants():
    """pv_inversion: every psi layer has zero trapezoid mean; the modal
    fields satisfy the Helmholtz equation; independent of baroclinic mode
    order and eigenvector scaling (qg3p_diagnostics.F90:615-635)."""
    g, L = 33, 3
    st, z = _inv_case(g, L, 3)
    inv = O.Inversion(g, st.ss, st.modes_to_z, st.z_to_modes)
    psi, rc = inv.run(z)
    assert rc == 0
    for l in range(L):
        assert abs(O.mean(psi[:, :, l])) < 1e-13 * np.max(np.abs(psi))
    # permuted / rescaled modes give the same psi
    perm = [0, 2, 1]
    scale = np.array([1.0, -3.0, 0.25])
    m2z = st.modes_to_z[:, perm] * scale
    z2m = np.linalg.inv(m2z)
    psi2, _ = O.Inversion(g, st.ss[perm], m2z, z2m).run(z)
    assert rel_err(psi2, psi) < 1e-12
    # layer-space equation: Delta psi_l + (S psi)_l = zeta_l - c_l on the
    # interior, where the constants c_l come from the mass correction; so
    # the residual is constant in space per layer
    for l in range(L):
        p = psi[:, :, l]
        lap = p[2:, 1:-1] + p[:-2, 1:-1] + p[1:-1, 2:] + p[1:-1, :-2] - 4 * p[1:-1, 1:-1]
        sp = sum(st.s_mat[l, k] * psi[1:-1, 1:-1, k] for k in range(L))
        r = lap + sp - z[1:-1, 1:-1, l]
        assert np.max(np.abs(r - r.mean())) < 1e-11 * np.max(np.abs(z))


def test_inversion_barotropic_only():
    """One layer: psi = helm_0(zeta) - mean (qg3p_diagnostics.F90:631)."""
    g = 17
    rng = np.random.default_rng(9)
    z = np.asfortranarray(rng.standard_normal((g, g, 1)))
    psi, rc = O.Inversion(g, [0.0], np.ones((1, 1)), np.ones((1, 1))).run(z)
    h = z.copy(order="F")
    O.Helm(g, [0.0]).solve(h)
    O.apply_dirichlet_bc(h, 0.0)
    assert np.allclose(psi, h - O.mean(h[:, :, 0]), atol=1e-14)


# ------------------------------------------------------------ CABARET
def test_cabaret_cfl_known_answers():
    ka = KA["cabaret_cfl"]
    g = ka["grid"]["iis"]
    for c in ka["cases"]:
        ui = np.full((g + 3, g + 2, 1), c["u_i"], order="F")
        uj = np.full((g + 2, g + 3, 1), c["u_j"], order="F")
        if "u_i_at_10_10" in c:
            ui[9, 9, 0] = c["u_i_at_10_10"]
        if "u_j_at_10_10" in c:
            uj[9, 9, 0] = c["u_j_at_10_10"]
        assert _fe(O.cabaret_cfl(c["dt"], ui, uj), c["expect"])


def _small_model(g=33, adapt=False, wind=True):
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    fw = c.wind() if wind else np.zeros((g, g), order="F")
    m = O.Model(g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], fw,
                viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"],
                alpha=nd["alpha"], adapt_dt=adapt, target_cfl=0.1, max_dt=10.0)
    return c, nd, m


def test_cabaret_zero_state_stays_zero_without_wind():
    c, nd, m = _small_model(wind=False)
    m.init(np.zeros((m.g, m.g, m.layers), order="F"))
    m.step(5)
    for k, v in m.a.items():
        if k != "force_wind":
            assert np.all(v == 0.0), k


def test_cabaret_flux_form_conservation():
    """With beta, wind, viscosity and drag off, the corrector changes the
    interior PV sum only through fluxes at the boundary faces
    (manual/manual.tex:702-704)."""
    g, L = 33, 2
    rng = np.random.default_rng(11)
    z = np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))
    fli = np.asfortranarray(rng.standard_normal((g + 3, g + 2, L)))
    flj = np.asfortranarray(rng.standard_normal((g + 2, g + 3, L)))
    res = np.zeros((g, g, L), order="F")
    z0 = z.copy(order="F")
    dt = 0.37
    O.cabaret_corrector(dt, fli, flj, res, z)
    for l in range(L):
        dsum = (z[2:g, 2:g, l] - z0[2:g, 2:g, l]).sum()
        # interior nodes 3..g: faces 3 and g+1 bound the region
        bflux = (fli[g, 2:g, l] - fli[2, 2:g, l]).sum() + (flj[2:g, g, l] - flj[2:g, 2, l]).sum()
        assert abs(dsum + 0.5 * dt * bflux) < 1e-12 * np.abs(z).sum()


def test_cabaret_constant_field_preserved():
    """A spatially constant PV with uniform face values and zero
    velocity is stationary under extrapolator + flux + corrector."""
    g, L = 17, 1
    z = np.full((g + 2, g + 2, L), 3.0, order="F")
    zo = z.copy(order="F")
    ui = np.zeros((g + 3, g + 2, L), order="F")
    uj = np.zeros((g + 2, g + 3, L), order="F")
    zfi = np.full((g + 3, g + 2, L), 3.0, order="F")
    zfj = np.full((g + 2, g + 3, L), 3.0, order="F")
    fli, flj = np.zeros_like(zfi), np.zeros_like(zfj)
    res = np.zeros((g, g, L), order="F")
    O.cabaret_extrapolator(0.1, z, zo, ui, uj, zfi, zfj)
    assert np.all(zfi[2:g + 1, 2:g, :] == 3.0) and np.all(zfj[2:g, 2:g + 1, :] == 3.0)
    O.cabaret_flux(ui, uj, zfi, zfj, fli, flj)
    O.cabaret_corrector(0.1, fli, flj, res, z)
    assert np.all(z == 3.0)


def test_cabaret_uniform_advection_exact_shift():
    """Uniform flow u = 1 along i with dt = 1 (Courant number 1).  From a
    consistent state -- face values equal to the upwind node value and the
    residual equal to minus the flux divergence -- one CABARET step
    (predictor, extrapolator with limiter, flux, corrector) translates the
    profile by exactly one node: q^{n+1/2}_k = (q_k + q_{k-1}) / 2,
    q_f^{n+1} = 2 q^{n+1/2}_{k-1} - q_f^n = q_{k-1}, d_f = 0 so the limiter is
    inactive, and q^{n+1}_k = q_{k-1}.  A wrong upwind side, limiter sign or
    half-step factor breaks the exact shift."""
    g, L = 33, 1
    dt = 1.0
    ui = np.zeros((g + 3, g + 2, L), order="F")
    ui[:, 2:g, 0] = 1.0           # every i-face of the interior columns 3..g
    uj = np.zeros((g + 2, g + 3, L), order="F")
    prof = np.where(np.arange(1, g + 3) < 12, 1.0, 0.0)   # 1-based data index
    z = np.asfortranarray(np.repeat(prof[:, None], g + 2, axis=1)[:, :, None])
    zo = z.copy(order="F")
    zfi = np.zeros((g + 3, g + 2, L), order="F")
    zfi[1:g + 2, :, 0] = z[0:g + 1, :, 0]   # face i (1-based) = node i-1
    zfj = np.zeros((g + 2, g + 3, L), order="F")
    res = np.zeros((g, g, L), order="F")
    # res(i-1, j-1) = -(f(i+1) - f(i)) = q(i-1) - q(i) for data nodes i = 3..g
    res[1:g - 1, :, 0] = (prof[1:g - 1] - prof[2:g])[:, None]
    fli, flj = np.zeros_like(zfi), np.zeros_like(zfj)
    zh = z.copy(order="F")
    O.cabaret_predictor(dt, zo, res, zh)
    O.cabaret_extrapolator(dt, zh, zo, ui, uj, zfi, zfj)
    O.cabaret_flux(ui, uj, zfi, zfj, fli, flj)
    O.cabaret_corrector(dt, fli, flj, res, zh)
    col = zh[:, g // 2, 0]
    shifted = np.concatenate([[prof[0]], prof[:-1]])
    assert np.array_equal(col[3:g], shifted[3:g])
    assert np.array_equal(zfi[3:g, g // 2, 0], prof[2:g - 1])


# ------------------------------------------------------------ energy diagnostics
def _strat2(rd, h1, h2):
    s21, s12 = O.strat_l2(rd, h1, h2)
    return np.array([s21, 0.0])


def _rows_field(gi, gj, a, b):
    """Data array (gi+2, gj+2, L) whose layer k is a_k * i + b_k on the owned
    rows (i = 1..gi the global row) and on every column, halo included."""
    L = len(a)
    psi = np.zeros((gi + 2, gj + 2, L), order="F")
    i = np.arange(1, gi + 1, dtype=np.float64)
    for k in range(L):
        psi[1:gi + 1, :, k] = (a[k] * i + b[k])[:, None]
    return psi


@pytest.mark.parametrize("case", KA["energy"]["cases"], ids=lambda c: c["name"])
def test_energy_known_answers(case):
    """unittest/test_energy_2.F90 on its own 5 x 9 node grid."""
    ka = KA["energy"]
    s2 = _strat2(ka["rd"], *ka["h"])
    assert _fe(s2[0], 1.0)
    psi = _rows_field(ka["gi"], ka["gj"], case["a"], case["b"])
    ekin, epot = O.energy(psi, ka["h"], s2)
    assert _fe(epot, case["epot"]), (epot, case["epot"])
    assert all(_fe(x, y) for x, y in zip(ekin, case["ekin"])), (ekin, case["ekin"])


def _energy_numpy(psi, h, s2):
    """The same diagnostic written with array operations: trapezoid weights
    w_i w_j on the interface differences, and per cell the mean over the cell
    of the two centred gradient components."""
    gi, gj, L = psi.shape[0] - 2, psi.shape[1] - 2, psi.shape[2]
    p = psi[1:gi + 1, 1:gj + 1, :]
    wi = np.ones(gi)
    wi[[0, -1]] = 0.5
    wj = np.ones(gj)
    wj[[0, -1]] = 0.5
    H = np.sum(h)
    cff = 1.0 / ((gi - 1) * (gj - 1))
    epot = sum(0.5 * s2[k] * h[k] / H * np.sum(np.outer(wi, wj) * (p[:, :, k] - p[:, :, k + 1]) ** 2)
               for k in range(L - 1)) * cff
    ekin = np.zeros(L)
    for k in range(L):
        q = p[:, :, k]
        dj = 0.5 * ((q[1:, 1:] - q[1:, :-1]) + (q[:-1, 1:] - q[:-1, :-1]))
        di = 0.5 * ((q[1:, 1:] - q[:-1, 1:]) + (q[1:, :-1] - q[:-1, :-1]))
        ekin[k] = 0.5 * h[k] / H * np.sum(di ** 2 + dj ** 2) * cff
    return ekin, epot


@pytest.mark.parametrize("gi,gj,L", [(5, 9, 2), (9, 13, 3), (17, 6, 1), (33, 33, 5)])
def test_energy_matches_array_form(gi, gj, L):
    """A non-square grid tells i from j; random layers, depths and s2 tell
    the layer and interface indices apart; halo values must not enter."""
    rng = np.random.default_rng(gi * 100 + gj)
    psi = rng.standard_normal((gi + 2, gj + 2, L))
    psi = np.asfortranarray(psi)
    h = rng.uniform(0.5, 2.0, L)
    s2 = rng.uniform(0.1, 3.0, L)
    s2[-1] = 0.0
    ekin, epot = O.energy(psi, h, s2)
    ek2, ep2 = _energy_numpy(psi, h, s2)
    assert np.allclose(ekin, ek2, rtol=1e-13, atol=0) and abs(epot - ep2) <= 1e-13 * abs(ep2)
    # the halo ring is never read on one domain
    psi2 = psi.copy(order="F")
    psi2[0, :, :] = psi2[-1, :, :] = np.nan
    psi2[:, 0, :] = psi2[:, -1, :] = np.nan
    ek3, ep3 = O.energy(psi2, h, s2)
    assert np.array_equal(ek3, ekin) and ep3 == epot


def test_energy_plane_closed_form():
    """psi_k = a i + b j + c_k: ekin_k = (h_k / 2H)(a^2 + b^2) and
    epot = sum_k s2_k h_k / (2H) (c_k - c_k+1)^2, on any grid."""
    gi, gj, L = 12, 7, 3
    a, b = 0.3, -1.7
    c = np.array([2.0, -0.5, 1.0])
    h = np.array([1.0, 2.0, 4.0])
    s2 = np.array([0.8, 1.9, 0.0])
    i = np.arange(gi + 2)[:, None, None]
    j = np.arange(gj + 2)[None, :, None]
    psi = np.asfortranarray(a * i + b * j + c[None, None, :] + 0.0 * i * j)
    ekin, epot = O.energy(psi, h, s2)
    H = h.sum()
    assert np.allclose(ekin, h / (2 * H) * (a * a + b * b), rtol=1e-13)
    ep = sum(s2[k] * h[k] / (2 * H) * (c[k] - c[k + 1]) ** 2 for k in range(L - 1))
    assert abs(epot - ep) <= 1e-13 * ep


def test_psi_av_add_is_dt_weighted_sum():
    g, L = 6, 2
    rng = np.random.default_rng(3)
    av = np.zeros((g, g, L), order="F")
    ref = np.zeros((g, g, L))
    for dt in (0.5, 0.25, 1.0):
        psi = np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))
        O.psi_av_add(psi, av, dt)
        ref = ref + dt * psi[1:g + 1, 1:g + 1, :]
    assert np.array_equal(av, ref)


# ------------------------------------------------------------ periodic channel
def _per_order_errors(ss, gs):
    m = []
    for gi in gs:
        gj = 2 * (gi - 1)
        i = np.arange(1, gi + 1)[:, None]
        j = np.arange(1, gj + 1)[None, :]
        x = (j - 1) / gj
        y = (gi - i) / (gi - 1)
        s = np.cos(2 * np.pi * x) * np.sin(np.pi * y)
        if ss == 0.0:
            phi = -np.pi ** 2 * s * (1.0 / (gi - 1) ** 2 + 4.0 / gj ** 2)
        else:
            phi = -s * (np.pi ** 2 / (gi - 1) ** 2 + 4 * np.pi ** 2 / gj ** 2 + ss)
        z = np.asfortranarray(phi[:, :, None])
        O.PeriodicHelm(gi, gj, [ss]).solve(z)
        z[0, :, 0] = 0.0
        z[-1, :, 0] = 0.0
        m.append(np.max(np.abs(z[:, :, 0] - s)))
    return np.array(m)


def test_periodic_helm_convergence_orders():
    """unittest/test_helmholtz.F90:248-356: the channel (periodic in j,
    g_jj = 2 (g_ii - 1)) Poisson order >= 2 and Helmholtz (ss = 0.5)
    order >= 3.9."""
    ka = KA["helmholtz_periodic_order"]
    gs = ka["g"][:6]
    for ss, lim in [(0.0, ka["poisson_min_order"]), (ka["helmholtz_ss"], ka["helmholtz_min_order"])]:
        m = _per_order_errors(ss, gs)
        o = -np.log10(m[1:] / m[:-1]) / np.log10(np.array(gs[1:]) / np.array(gs[:-1]))
        assert np.all(o >= lim), (ss, o)


@pytest.mark.parametrize("gi,gj", [(9, 16), (17, 32), (12, 7)])
@pytest.mark.parametrize("ss", [0.0, 0.7])
def test_periodic_helm_eigenfunctions(gi, gj, ss):
    """sin(p pi (i-1) / (gi-1)) cos / sin(2 pi q (j-1) / gj) are exact discrete
    eigenfunctions: the solve divides by lambda_p + mu_q - ss."""
    N = gi - 1
    i = np.arange(gi)[:, None]
    j = np.arange(gj)[None, :]
    lam = -4 * np.sin(np.arange(N) * np.pi / (2 * N)) ** 2
    for p, q, f in [(1, 0, np.cos), (2, 1, np.cos), (N - 1, 3, np.sin), (3, gj // 2, np.cos)]:
        rhs = np.sin(np.pi * p * i / N) * f(2 * np.pi * q * j / gj)
        z = np.asfortranarray(rhs[:, :, None].copy())
        O.PeriodicHelm(gi, gj, [ss]).solve(z)
        z[0, :, 0] = z[-1, :, 0] = 0.0
        mu = -4 * np.sin(np.pi * q / gj) ** 2
        exp = rhs / (lam[p] + mu - ss)
        assert np.max(np.abs(z[:, :, 0] - exp)) < 1e-13 * max(1.0, np.max(np.abs(exp))), (p, q)


def _periodic_dense(gi, gj, ss):
    """Dense 5-point operator (Delta - ss) on the interior rows 2..gi-1,
    periodic in j, zero on the walls: unknown (i, j) at (i - 2) + (gi - 2) j."""
    n = gi - 2
    A = np.zeros((n * gj, n * gj))
    for j in range(gj):
        for i in range(n):
            k = i + n * j
            A[k, k] = -4.0 - ss
            if i > 0:
                A[k, k - 1] = 1.0
            if i < n - 1:
                A[k, k + 1] = 1.0
            A[k, i + n * ((j + 1) % gj)] += 1.0
            A[k, i + n * ((j - 1) % gj)] += 1.0
    return A


@pytest.mark.parametrize("gi,gj,ss", [(6, 7, 0.0), (7, 8, 0.4), (5, 3, 1.3)])
def test_periodic_helm_brute_force(gi, gj, ss):
    rng = np.random.default_rng(gi * gj)
    rhs = rng.standard_normal((gi, gj))
    z = np.asfortranarray(rhs[:, :, None].copy())
    O.PeriodicHelm(gi, gj, [ss]).solve(z)
    ref = np.linalg.solve(_periodic_dense(gi, gj, ss), rhs[1:-1, :].reshape(-1, order="F"))
    got = z[1:-1, :, 0].reshape(-1, order="F")
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
    # the wall rows: i = gi is zeroed, i = 1 keeps the input (helm's output
    # boundary semantics, qg3p_helmholtz.F90:757-761)
    assert np.all(z[-1, :, 0] == 0.0) and np.array_equal(z[0, :, 0], rhs[0, :])


def _channel_case(gi, gj, L, seed):
    rng = np.random.default_rng(seed)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((gi, gj, L)))
    return st, z


def _lap(p):
    """5-point Laplacian on the interior rows, periodic in j."""
    return (p[2:, :] + p[:-2, :] + np.roll(p[1:-1, :], 1, axis=1) + np.roll(p[1:-1, :], -1, axis=1)
            - 4 * p[1:-1, :])


@pytest.mark.parametrize("bc_type,alpha", [(1, 0.2), (2, 0.0), (3, 0.0)])
def test_channel_inversion_invariants(bc_type, alpha):
    """Channel pv_inversion: (i) each mode solves (Delta - ss_m) phi_m =
    (z_to_modes zeta)_m on the interior; (ii) every baroclinic mode's mean
    boundary relative PV vanishes on both walls (the momentum constraint,
    qg3p_aux.F90:233-286, 524-575); (iii) the barotropic mode is zero on
    the walls and every mode is constant along each wall."""
    gi, gj, L = 13, 24, 3
    st, z = _channel_case(gi, gj, L, 5)
    ch = O.Channel(gi, gj, st.ss, st.modes_to_z, st.z_to_modes, bc_type, alpha)
    psi, rc = ch.run(z)
    assert rc == 0
    phi = np.einsum("ml,ijl->ijm", st.z_to_modes, psi)
    zm = np.einsum("ml,ijl->ijm", st.z_to_modes, z)
    scale = np.max(np.abs(zm))
    for m in range(L):
        r = _lap(phi[:, :, m]) - st.ss[m] * phi[1:-1, :, m] - zm[1:-1, :, m]
        assert np.max(np.abs(r)) < 1e-11 * scale, m
        for w in (0, -1):
            assert np.ptp(phi[w, :, m]) < 1e-13 * np.max(np.abs(phi))
    assert np.max(np.abs(phi[[0, -1], :, 0])) < 1e-14 * np.max(np.abs(phi))
    zb, zd = O.zeta_boundary(phi, st.ss, bc_type, alpha)
    assert np.max(np.abs(zb[1:])) < 1e-12 * scale and np.max(np.abs(zd[1:])) < 1e-12 * scale
    if bc_type != 3:  # free slip: the homogeneous solution already satisfies it
        assert np.any(np.abs(ch.psic()) > 1e-3 * scale)  # the correction does something


def test_channel_correction_functions_closed_form():
    """The correction functions are uniform along the channel and solve the
    1D problem c(i-1) + c(i+1) - (2 + ss) c(i) = 0 between the walls:
    c1 = sinh(k (gi - i)) / sinh(k (gi - 1)), cosh k = 1 + ss / 2."""
    gi, gj, L = 17, 8, 3
    st, _ = _channel_case(gi, gj, L, 2)
    ch = O.Channel(gi, gj, st.ss, st.modes_to_z, st.z_to_modes, 1, 0.5)
    corr = ch.corr()
    i = np.arange(1, gi + 1)
    for m in range(1, L):
        k = np.arccosh(1 + st.ss[m] / 2)
        c1 = np.sinh(k * (gi - i)) / np.sinh(k * (gi - 1))
        c2 = np.sinh(k * (i - 1)) / np.sinh(k * (gi - 1))
        assert np.max(np.abs(corr[:, :, m, 0] - c1[:, None])) < 1e-13
        assert np.max(np.abs(corr[:, :, m, 1] - c2[:, None])) < 1e-13


def test_channel_brute_force():
    """The whole constrained problem as one dense system per baroclinic
    mode on a tiny channel: interior Helmholtz unknowns plus the two wall
    values, with the two boundary constraints (partial slip)."""
    gi, gj, L, alpha = 6, 5, 2, 0.7
    st, z = _channel_case(gi, gj, L, 8)
    psi, rc = O.Channel(gi, gj, st.ss, st.modes_to_z, st.z_to_modes, 1, alpha).run(z)
    phi = np.einsum("ml,ijl->ijm", st.z_to_modes, psi)
    zm = np.einsum("ml,ijl->ijm", st.z_to_modes, z)
    n = gi - 2
    cff = alpha / (alpha + 3.0)
    for m in range(L):
        ss = st.ss[m]
        A0 = _periodic_dense(gi, gj, ss)
        nu = n * gj + (2 if m > 0 else 0)
        A = np.zeros((nu, nu))
        A[:n * gj, :n * gj] = A0
        rhs = np.zeros(nu)
        rhs[:n * gj] = zm[1:-1, :, m].reshape(-1, order="F")
        if m > 0:
            top, bot = n * gj, n * gj + 1
            for j in range(gj):
                A[0 + n * j, top] += 1.0          # wall value enters row i = 2
                A[n - 1 + n * j, bot] += 1.0      # and row i = gi - 1
                # mean boundary PV: cff (-3.5 w + 4 phi_2 - 0.5 phi_3) - ss w
                A[top, 0 + n * j] += 4.0 * cff / gj
                A[top, 1 + n * j] += -0.5 * cff / gj
                A[bot, n - 1 + n * j] += 4.0 * cff / gj
                A[bot, n - 2 + n * j] += -0.5 * cff / gj
            A[top, top] = -3.5 * cff - ss
            A[bot, bot] = -3.5 * cff - ss
        x = np.linalg.solve(A, rhs)
        ref = np.zeros((gi, gj))
        ref[1:-1, :] = x[:n * gj].reshape((n, gj), order="F")
        if m > 0:
            ref[0, :], ref[-1, :] = x[n * gj], x[n * gj + 1]
        assert np.max(np.abs(phi[:, :, m] - ref)) < 1e-12 * np.max(np.abs(ref)), m


# ------------------------------------------------------------ leapfrog
def _lf_fields(g, L):
    return [np.zeros((g + 2, g + 2, L), order="F") for _ in range(4)]


def test_zeta_step_zonal_flow_is_steady():
    """psi, rel functions of i only: every Jacobian term cancels exactly."""
    g, L = 12, 2
    psi, rel, relp, src = _lf_fields(g, L)
    i = np.arange(g + 2, dtype=float)[:, None, None]
    psi += np.sin(0.3 * i) * np.array([1.0, -0.4])
    rel += np.cos(0.2 * i) * np.array([0.5, 0.1])
    src += np.random.default_rng(0).standard_normal(src.shape)
    dst = src.copy(order="F")
    O.zeta_step(0.7, [0.3, 0.2], [0.0, 0.4], [0.6, 0.0], np.zeros((g, g)), 0.0, 0.0,
                psi, rel, relp, src, dst)
    assert np.array_equal(dst, src)


@pytest.mark.parametrize("L", [1, 2, 3])
def test_zeta_step_uniform_flow_beta(L):
    """psi = a j in every layer, rel = 0: only beta_y(i) = beta ((G+1)/2 -
    (i-1)) is advected, 0.25 J = -a beta (the PV gradient points to small
    i); with layer couplings s2 psi_(l+1) = s2 b i added to the top layer's
    PV the tendency becomes a (s2 b - beta)."""
    g, a, b, T = 10, 0.3, 0.7, 2.0
    psi, rel, relp, src = _lf_fields(g, L)
    i = np.arange(1, g + 3, dtype=float)[:, None]
    j = np.arange(1, g + 3, dtype=float)[None, :]
    beta = np.linspace(0.2, 0.5, L)
    for l in range(L):
        psi[:, :, l] = a * j
    s1 = np.zeros(L)
    s2 = np.zeros(L)
    if L > 1:
        psi[:, :, 1] = b * i + 0 * j   # layer 2 varies along i only
        s2[0] = 0.9
    dst = src.copy(order="F")
    O.zeta_step(T, beta, s1, s2, np.zeros((g, g)), 0.0, 0.0, psi, rel, relp, src, dst)
    exp0 = a * (s2[0] * b - beta[0]) if L > 1 else -a * beta[0]
    assert np.allclose(dst[2:g, 2:g, 0], T * exp0, rtol=1e-13, atol=1e-14)


def test_zeta_step_wind_viscosity_drag():
    """Wind in the top layer only; viscosity times the 5-point Laplacian of
    rel_p (rel_p = i^2: Laplacian 2); bottom drag on rel_p in the bottom
    layer (one layer: inside the viscosity bracket, as the reference)."""
    g, T, nu, r = 9, 0.5, 0.3, 0.2
    fw = np.random.default_rng(1).standard_normal((g, g))
    i = np.arange(1, g + 3, dtype=float)[:, None]
    for L in (1, 3):
        psi, rel, relp, src = _lf_fields(g, L)
        relp += (i * i)[:, :, None]
        dst = src.copy(order="F")
        O.zeta_step(T, np.zeros(L), np.zeros(L), np.zeros(L), fw, nu, r, psi, rel, relp, src, dst)
        ii = i[2:g, :]
        top = fw[1:g - 1, 1:g - 1] + nu * 2.0 + (-nu * r * ii * ii if L == 1 else 0.0)
        assert np.allclose(dst[2:g, 2:g, 0], T * top, rtol=1e-13)
        if L == 3:
            assert np.allclose(dst[2:g, 2:g, 1], T * nu * 2.0, rtol=1e-13)
            assert np.allclose(dst[2:g, 2:g, 2], T * (nu * 2.0 - r * ii * ii), rtol=1e-13)
        # only the interior is written
        assert np.all(dst[[0, 1, g, g + 1], :, :] == 0.0) and np.all(dst[:, [0, 1, g, g + 1], :] == 0.0)


def test_ra_filter_weights():
    """Williams' filter with nu = 0.02, alpha = 0.5: the level n value moves
    by nu alpha / 2 d and the level n+1 value by -nu (1 - alpha) / 2 d,
    d = phi(n-1) - 2 phi(n) + phi(n+1); sequences linear in time are left
    unchanged; phi_old(m) receives the filtered level n; halo untouched."""
    g, L = 6, 2
    rng = np.random.default_rng(4)
    a, b, c = (np.asfortranarray(rng.standard_normal((g + 2, g + 2, L))) for _ in range(3))
    pl, pm, pol, pom = c.copy(order="F"), b.copy(order="F"), a.copy(order="F"), np.zeros_like(a)
    O.ra_filter(pl, pm, pol, pom)
    d = a - 2 * b + c
    s = np.s_[1:g + 1, 1:g + 1, :]
    assert np.allclose(pm[s], (b + 0.005 * d)[s], rtol=0, atol=1e-14)
    assert np.allclose(pl[s], (c - 0.005 * d)[s], rtol=0, atol=1e-14)
    assert np.array_equal(pom[s], pm[s])
    assert np.array_equal(pl[0], c[0]) and np.array_equal(pm[:, -1], b[:, -1])
    lin = [a, a + 0.5, a + 1.0]
    pl, pm, pol, pom = lin[2].copy(order="F"), lin[1].copy(order="F"), lin[0].copy(order="F"), np.zeros_like(a)
    O.ra_filter(pl, pm, pol, pom)
    assert np.allclose(pm[s], lin[1][s], atol=1e-14) and np.allclose(pl[s], lin[2][s], atol=1e-14)


def _smooth_q(g, L, amp=1e-2):
    x = np.linspace(0.0, 1.0, g)
    X, Y = np.meshgrid(x, x, indexing="ij")
    q = np.zeros((g, g, L), order="F")
    for l in range(L):
        q[:, :, l] = amp * (np.sin(np.pi * X) * np.sin(2 * np.pi * Y) * (1 + 0.3 * l)
                            + 0.5 * np.sin(3 * np.pi * X + l) * np.sin(np.pi * Y))
    return q


def _lf_and_cabaret(g=65, dt_frac=0.25):
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    kw = dict(viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"] * dt_frac, bc_type=1,
              alpha=nd["alpha"])
    fw = c.wind()
    lf = O.LeapfrogModel(g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], fw, **kw)
    cab = O.Model(g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], fw, **kw)
    return lf, cab, _smooth_q(g, st.layers)


def test_leapfrog_follows_cabaret():
    """Both schemes integrate the same equations: on a smooth 3-layer field
    (wind, beta, stretching, viscosity and drag all active) the PV change
    of the leapfrog model over 20 steps follows the CABARET model's per
    layer (correlation above 0.999, amplitude within 1 %); a wrong sign or
    layer index in the Jacobian, the coupling, the forcing or the filter
    breaks this."""
    lf, cab, q = _lf_and_cabaret()
    assert lf.init(q) == 0 and cab.init(q) == 0
    z0 = cab.a["zeta"].copy()
    n = 20
    lf.step(n - 1)   # the forward Euler start is the first step
    cab.step(n)
    assert abs(lf.scalars()["model_time"] - n * lf.st.dt) < 1e-9 * n * lf.st.dt
    zl = lf.a["zeta1"] if lf.st.leapfrog == 2 else lf.a["zeta2"]   # the level written last
    for l in range(q.shape[2]):
        dl = (zl - z0)[2:-2, 2:-2, l].ravel()
        dc = (cab.a["zeta"] - z0)[2:-2, 2:-2, l].ravel()
        corr = np.dot(dl, dc) / np.sqrt(np.dot(dl, dl) * np.dot(dc, dc))
        assert corr > 0.999, (l, corr)
        assert abs(np.linalg.norm(dl) / np.linalg.norm(dc) - 1) < 0.01, l


def test_leapfrog_second_order_in_time():
    """Free nonlinear evolution (beta, stretching and the Jacobian; no
    forcing, viscosity or drag) over a fixed time: halving dt cuts the
    error against a dt/16 run by about 4 (second order: the centred step
    uses psi and rel of the middle level, the filter's alpha = 0.5 keeps
    the order); a first-order slip (e.g. psi of the wrong level) gives 2."""
    g = 33
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]

    def run(dt, n):
        m = O.LeapfrogModel(g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"],
                            np.zeros((g, g)), viscosity=0.0, visc_bot=0.0, dt=dt, bc_type=2)
        assert m.init(_smooth_q(g, st.layers, 0.3)) == 0
        m.step(n - 1)
        return (m.a["zeta1"] if m.st.leapfrog == 2 else m.a["zeta2"]).copy()

    D, n = nd["dt"], 16
    ref = run(D / 16, n * 16)
    e = [np.max(np.abs(run(D / k, n * k) - ref)) for k in (1, 2, 4)]
    assert e[0] / e[1] > 3.4 and e[1] / e[2] > 3.4, e
