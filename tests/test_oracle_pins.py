"""Pins of the oracle's CABARET and boundary routines by what the mathematics
fixes, independently of the oracle's own statements: closed forms on
polynomial fields, the stencils' Taylor properties, the physical identities
the routines implement (PV = relative vorticity + S psi with the
stratification matrix S whose eigenvalues the reference's known answers
pin), hand-derived known answers of the limiter and its mirror and transpose
symmetries.  tests/test_oracle_mutants.py plants mistakes in copies of the
oracle and checks that these tests catch each of them.

Reference routines: qg3p_aux.F90:643-776 (update_zeta_bc,
rel_from_zeta_and_psi), 779-915 (apply_rel_bc), 920-990 (cabaret_cfl,
cabaret_adapt_dt), 1023-1061 (cabaret_beta), 1065-1164 (cabaret_forcing),
1168-1313 (cell boundaries), 1317-1544 (extrapolator); qg3.F90:253-279
(wind).  Index conventions (qg3p_base.F90:887-941, one domain): data nodes
1..G+2 with owned nodes 2..G+1 and interior 3..G; i-face i lies between
data nodes i-1 and i.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pequod_amd import strat as S, synth

G, L3 = 12, 3


def node(g, L, fill=None, rng=None):
    if rng is not None:
        return np.asfortranarray(rng.standard_normal((g + 2, g + 2, L)))
    return np.full((g + 2, g + 2, L), 0.0 if fill is None else fill, order="F")


def fi(g, L, fill=0.0):
    return np.full((g + 3, g + 2, L), fill, order="F")


def fj(g, L, fill=0.0):
    return np.full((g + 2, g + 3, L), fill, order="F")


def idx(shape):
    """1-based data index grids (i, j) of a 2-D shape."""
    i = np.arange(1, shape[0] + 1, dtype=np.float64)[:, None]
    j = np.arange(1, shape[1] + 1, dtype=np.float64)[None, :]
    return i + 0 * j, j + 0 * i


def strat3():
    return S.from_parameters(np.ones(3), [0.15, 0.07], [0.2, 0.11])


def s_apply(st, psi):
    """(S psi)_l at every node with the stratification matrix S."""
    return np.einsum("lm,ijm->ijl", st.s_mat, psi)


# ------------------------------------------------------------- S matrix
def test_s_matrix_is_pinned_by_known_answers():
    """The S of the identities below is the one whose eigenvalues the
    reference's test_stratification fixes (3-layer closed form): -ss are its
    eigenvalues, and S annihilates a barotropic (layer-uniform) column."""
    st = strat3()
    ev = np.sort(np.linalg.eigvals(st.s_mat).real)
    np.testing.assert_allclose(np.sort(-st.ss), ev, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(st.s_mat @ np.ones(3), 0.0, atol=1e-15)
    # the coupling of layer 1 to layer 2 is s2(1), of layer 3 to layer 2 s1(3)
    assert st.s_mat[0, 1] == st.s2[0] and st.s_mat[2, 1] == st.s1[2]


# --------------------------------------------------- cell boundaries
@pytest.mark.parametrize("tw", [0.0, 0.25])
def test_cell_boundaries_quadratic_psi(tw):
    """psi = a_l i^2 + b_l j^2 + c_l i j + d_l: the corner average of a
    quadratic is exact, so the face velocities are
    ui(i, j) = -(2 b j + c (i - 1/2)),  uj(i, j) = 2 a i + c (j - 1/2)
    on i-faces i = 3..G+1, j = 3..G and j-faces i = 3..G, j = 3..G+1, and
    u = (1 + tw) u_new - tw u_old with u_old <- u_new.  Every other face is
    left as it was."""
    g, L = G, 2
    a, b, c, d = [3.0, -1.0], [-2.0, 4.0], [5.0, 1.0], [7.0, -3.0]
    psi = node(g, L)
    I, J = idx(psi.shape[:2])
    for l in range(L):
        psi[:, :, l] = a[l] * I ** 2 + b[l] * J ** 2 + c[l] * I * J + d[l]
    rng = np.random.default_rng(0)
    ui, uj = fi(g, L, 777.0), fj(g, L, 777.0)
    uimo = np.asfortranarray(rng.integers(-9, 9, ui.shape).astype(np.float64))
    ujmo = np.asfortranarray(rng.integers(-9, 9, uj.shape).astype(np.float64))
    uimo0, ujmo0 = uimo.copy(order="F"), ujmo.copy(order="F")
    O.cabaret_cell_boundaries(tw, psi, ui, uj, uimo, ujmo)
    Ii, Ji = idx(ui.shape[:2])
    Ij, Jj = idx(uj.shape[:2])
    ri = (Ii >= 3) & (Ii <= g + 1) & (Ji >= 3) & (Ji <= g)
    rj = (Ij >= 3) & (Ij <= g) & (Jj >= 3) & (Jj <= g + 1)
    for l in range(L):
        ui_new = -(2 * b[l] * Ji + c[l] * (Ii - 0.5))
        uj_new = 2 * a[l] * Ij + c[l] * (Jj - 0.5)
        np.testing.assert_array_equal(uimo[:, :, l][ri], ui_new[ri])
        np.testing.assert_array_equal(ujmo[:, :, l][rj], uj_new[rj])
        np.testing.assert_array_equal(
            ui[:, :, l][ri], ((1 + tw) * ui_new - tw * uimo0[:, :, l])[ri])
        np.testing.assert_array_equal(
            uj[:, :, l][rj], ((1 + tw) * uj_new - tw * ujmo0[:, :, l])[rj])
        assert np.all(ui[:, :, l][~ri] == 777.0) and np.all(uj[:, :, l][~rj] == 777.0)
        np.testing.assert_array_equal(uimo[:, :, l][~ri], uimo0[:, :, l][~ri])
    # the velocity-only variant (pickup) gives the same u_new
    vi, vj = fi(g, L, 777.0), fj(g, L, 777.0)
    O.cabaret_cell_boundaries_velocity(psi, vi, vj)
    np.testing.assert_array_equal(vi[ri[:, :, None].repeat(L, 2)], uimo[ri[:, :, None].repeat(L, 2)])
    np.testing.assert_array_equal(vj[rj[:, :, None].repeat(L, 2)], ujmo[rj[:, :, None].repeat(L, 2)])


# ---------------------------------------------------------- apply_rel_bc
def _edges(g):
    """(west i=2, east i=G+1, south j=2, north j=G+1) boundary masks of the
    owned nodes, corners included, on a (G+2, G+2) array."""
    I, J = idx((g + 2, g + 2))
    own = (I >= 2) & (I <= g + 1) & (J >= 2) & (J <= g + 1)
    return own & (I == 2), own & (I == g + 1), own & (J == 2), own & (J == g + 1)


@pytest.mark.parametrize("bc", [1, 2, 3])
@pytest.mark.parametrize("axis", [0, 1])
def test_apply_rel_bc_wall_profiles(bc, axis):
    """psi = k x^2 with x the distance in nodes from the wall at index 2
    (varying along one axis only).  The partial-slip stencil
    -3.5 p0 + 4 p1 - 0.5 p2 annihilates constants and gives 3 for a unit
    slope and 2 for a unit curvature, so at the near wall
    rel = alpha/(alpha+3) * 2k and at the far wall (x = X - y, X = G - 1)
    rel = alpha/(alpha+3) * (2k - 6kX); no slip 2 (p1 - p0) gives 2k and
    2k(3 - 2G); free slip gives 0.  The walls parallel to the profile see a
    uniform psi and get 0; corners belong to the j walls (written last)."""
    g, L, k, alpha = G, 2, 3.0, 0.2
    cff = alpha / (alpha + 3.0)  # 1/16
    psi = node(g, L)
    I, J = idx(psi.shape[:2])
    x = (I if axis == 0 else J) - 2.0
    for l in range(L):
        psi[:, :, l] = (k + l) * x ** 2
    rel = node(g, L, 555.0)
    O.lib().pqo_apply_rel_bc(g, L, bc, alpha, O._p(psi), O._p(rel))
    west, east, south, north = _edges(g)
    near, far, par1, par2 = (west, east, south, north) if axis == 0 else (south, north, west, east)
    X = g - 1.0
    for l in range(L):
        kk = k + l
        if bc == 1:
            v_near, v_far = cff * 2 * kk, cff * (2 * kk - 6 * kk * X)
        elif bc == 2:
            v_near, v_far = 2 * kk, 2 * kk * (3 - 2 * g)
        else:
            v_near = v_far = 0.0
        r = rel[:, :, l]
        if axis == 0:  # corners belong to the j walls (par1, par2), which read 0
            np.testing.assert_array_equal(r[near & ~par1 & ~par2], v_near)
            np.testing.assert_array_equal(r[far & ~par1 & ~par2], v_far)
            np.testing.assert_array_equal(r[par1 | par2], 0.0)
        else:  # the profile's walls are the j walls: corners take their values
            np.testing.assert_array_equal(r[near], v_near)
            np.testing.assert_array_equal(r[far], v_far)
            np.testing.assert_array_equal(r[(par1 | par2) & ~near & ~far], 0.0)
        assert np.all(r[~(west | east | south | north)] == 555.0)


# ------------------------------------------- rel_from_zeta / update_zeta_bc
@pytest.mark.parametrize("L", [1, 2, 3, 5])
def test_rel_from_zeta_inverts_pv_definition(L):
    """PV = relative vorticity + S psi.  With zeta := R + S psi for any R,
    rel_from_zeta_and_psi returns R on the owned nodes off the walls (whose
    values apply_rel_bc sets); one layer copies zeta, halo included."""
    g = G
    rng = np.random.default_rng(10 + L)
    st = S.from_parameters(np.ones(L), rng.uniform(0.05, 0.3, L - 1), rng.uniform(0.05, 0.3, L - 1))
    psi, R = node(g, L, rng=rng), node(g, L, rng=rng)
    zeta = np.asfortranarray(R + s_apply(st, psi))
    rel = node(g, L, 555.0)
    O.rel_from_zeta_and_psi(st.s1, st.s2, 3, 0.0, zeta, psi, rel)
    I, J = idx(psi.shape[:2])
    inner = (I >= 3) & (I <= g) & (J >= 3) & (J <= g)
    halo = (I == 1) | (I == g + 2) | (J == 1) | (J == g + 2)
    for l in range(L):
        np.testing.assert_allclose(rel[:, :, l][inner], R[:, :, l][inner], rtol=0, atol=1e-13)
        if L == 1:
            np.testing.assert_array_equal(rel[:, :, l][halo], zeta[:, :, l][halo])
        else:
            assert np.all(rel[:, :, l][halo] == 555.0)


@pytest.mark.parametrize("L", [1, 2, 3, 5])
def test_update_zeta_bc_is_pv_definition_on_walls(L):
    """update_zeta_bc writes PV = rel + S psi on every owned wall node,
    corners included, and leaves the interior and the halo alone."""
    g = G
    rng = np.random.default_rng(20 + L)
    st = S.from_parameters(np.ones(L), rng.uniform(0.05, 0.3, L - 1), rng.uniform(0.05, 0.3, L - 1))
    psi, rel = node(g, L, rng=rng), node(g, L, rng=rng)
    zeta = node(g, L, 555.0)
    O.update_zeta_bc(st.s1, st.s2, rel, psi, zeta)
    west, east, south, north = _edges(g)
    wall = west | east | south | north
    pv = rel + s_apply(st, psi)
    for l in range(L):
        np.testing.assert_allclose(zeta[:, :, l][wall], pv[:, :, l][wall], rtol=0, atol=1e-13)
        assert np.all(zeta[:, :, l][~wall] == 555.0)


# ------------------------------------------------------------ cabaret_beta
def test_beta_cell_centre_velocity():
    """The beta term uses the node's velocity, the mean of its west and east
    i-faces: with ui = a i + b j (face indices) the node at data (i, j) moves
    with a (i + 1/2) + b j, so rb = dt beta_l (a (i + 1/2) + b j) on the
    interior nodes 3..G, zeta += (1 + tw) rb - tw rb_old and rb_old <- rb;
    the walls are left alone."""
    g, L = G, 2
    dt, tw, beta, a, b = 0.25, 0.5, [0.5, 2.0], 2.0, -3.0
    ui = fi(g, L)
    Ii, Ji = idx(ui.shape[:2])
    for l in range(L):
        ui[:, :, l] = a * Ii + b * Ji
    rng = np.random.default_rng(3)
    zeta = np.asfortranarray(rng.integers(-50, 50, (g + 2, g + 2, L)).astype(np.float64))
    rbo = np.asfortranarray(rng.integers(-50, 50, (g, g, L)).astype(np.float64))
    z0, rbo0 = zeta.copy(order="F"), rbo.copy(order="F")
    O.cabaret_beta(dt, tw, np.array(beta), ui, zeta, rbo)
    I, J = idx((g + 2, g + 2))
    inner = (I >= 3) & (I <= g) & (J >= 3) & (J <= g)
    for l in range(L):
        rb = dt * beta[l] * (a * (I + 0.5) + b * J)
        rb_s = rb[2:g, 2:g]  # res_beta(i, j) belongs to data node (i + 1, j + 1)
        np.testing.assert_array_equal(rbo[1:g - 1, 1:g - 1, l], rb_s)
        exp = z0[:, :, l] + (1 + tw) * rb
        exp[2:g, 2:g] -= tw * rbo0[1:g - 1, 1:g - 1, l]
        np.testing.assert_array_equal(zeta[:, :, l][inner], exp[inner])
        np.testing.assert_array_equal(zeta[:, :, l][~inner], z0[:, :, l][~inner])
        # res_beta outside the interior block is untouched
        m = np.ones((g, g), bool)
        m[1:g - 1, 1:g - 1] = False
        np.testing.assert_array_equal(rbo[:, :, l][m], rbo0[:, :, l][m])


# --------------------------------------------------------- cabaret_forcing
@pytest.mark.parametrize("L", [1, 2, 4])
def test_forcing_closed_forms(L):
    """rel = c (i^2 + j^2) has the 5-point Laplacian 4c; the wind
    force_wind(a, b) = a + 100 b (sdata indices) reaches data node
    (a + 1, b + 1) of the top layer only; the bottom drag -r rel acts on the
    last layer only (both on the one layer of a one-layer model):
    zeta += dt (F [top] - r rel [bottom] + nu 4c) on the interior 3..G."""
    g = G
    dt, nu, r, c = 0.5, 0.25, 0.125, 1.0
    I, J = idx((g + 2, g + 2))
    rel = np.zeros((g + 2, g + 2, L), order="F")
    for l in range(L):
        rel[:, :, l] = (c + l) * (I ** 2 + J ** 2)
    Ia, Jb = idx((g, g))
    fw = np.asfortranarray(Ia + 100.0 * Jb)
    zeta = node(g, L, 1.0)
    O.cabaret_forcing(dt, fw, nu, r, rel, zeta)
    inner = (I >= 3) & (I <= g) & (J >= 3) & (J <= g)
    wind = np.zeros((g + 2, g + 2))
    wind[1:g + 1, 1:g + 1] = fw
    for l in range(L):
        f = nu * 4 * (c + l) * np.ones_like(I)
        if l == 0:
            f = f + wind
        if l == L - 1:
            f = f - r * rel[:, :, l]
        np.testing.assert_allclose(zeta[:, :, l][inner], (1.0 + dt * f)[inner], rtol=1e-15, atol=0)
        assert np.all(zeta[:, :, l][~inner] == 1.0)


# -------------------------------------------------------------- wind field
def test_wind_symmetric_double_gyre():
    """No shift, tilt or asymmetry: the profile is 2 pi tau sin(2 pi y),
    y = 1/2 - (i - 1)/(G - 1), uniform along j; synth.wind (the GPU path's
    input) is the same field."""
    g, tau = 41, 0.3
    fw = O.wind_field(g, tau)
    y = 0.5 - np.arange(g) / (g - 1.0)
    exp = (2 * np.pi * tau * np.sin(2 * np.pi * y))[:, None] * np.ones((1, g))
    np.testing.assert_allclose(fw, exp, rtol=0, atol=1e-14)
    c = synth.config("qg3_255")
    c.n = g - 2
    np.testing.assert_array_equal(c.wind(), O.wind_field(g, c.nondim()["tau"]))


def test_wind_tilted_asymmetric_known_answers():
    """Shifted, tilted, asymmetric gyres: the dividing line y_mid = shift +
    tilt x (x = (j - 1)/(G - 1) - 1/2) carries zero stress, and the gyres'
    extremes, midway between y_mid and the walls y = -1/2 and 1/2, are
    -2 pi tau asym (south) and 2 pi tau / asym (north).  With G = 41,
    shift 0.1, tilt 0.2, those points are grid nodes at x = 0 and x = 1/2."""
    g, tau, shift, tilt, asym = 41, 0.3, 0.1, 0.2, 2.0
    fw = O.wind_field(g, tau, shift, tilt, asym)
    A = 2 * np.pi * tau

    def at(y, x):
        i = int(round((0.5 - y) * (g - 1))) + 1
        j = int(round((x + 0.5) * (g - 1))) + 1
        return fw[i - 1, j - 1]

    for x in (0.0, 0.5):
        ym = shift + tilt * x
        assert abs(at(ym, x)) < 1e-14
        assert at((ym + 0.5) / 2, x) == pytest.approx(A / asym, rel=1e-14)
        assert at((ym - 0.5) / 2, x) == pytest.approx(-A * asym, rel=1e-14)
        assert abs(at(0.5, x)) < 1e-14 and abs(at(-0.5, x)) < 1e-13
    # sign: negative south of the dividing line, positive north of it
    I, J = idx((g, g))
    y, x = 0.5 - (I - 1) / (g - 1), (J - 1) / (g - 1) - 0.5
    ym = shift + tilt * x
    assert np.all(fw[(y < ym - 1e-9) & (y > -0.5 + 1e-9)] < 0)
    assert np.all(fw[(y > ym + 1e-9) & (y < 0.5 - 1e-9)] > 0)


# ------------------------------------------------------- cfl / adapt_dt
def _faces_with_max(g, L, umax):
    """Face velocities whose largest |u| inside the cabaret_cfl ranges
    (i-faces i = 3..G+1, j = 3..G; j-faces i = 3..G, j = 3..G+1) is umax, with
    larger values just outside them that must be ignored."""
    ui, uj = fi(g, L, 0.01), fj(g, L, -0.02)
    ui[g, g - 1, L - 1] = -umax           # i-face (G+1, G): in range
    ui[1, 4, 0] = 50.0                    # i-face i = 2: out
    ui[5, g, 0] = -60.0                   # i-face j = G+1: out
    uj[3, 1, 0] = 70.0                    # j-face j = 2: out
    uj[g, 5, 0] = 80.0                    # j-face i = G+1: out
    return ui, uj


@pytest.mark.parametrize("case", ["cfl", "growth", "cap", "still"])
def test_adapt_dt_closed_form(case):
    """dt = min(target_cfl / max|u|, max_dt_increase * dt_old, max_dt) unless
    max|u| = 0; tweight_a = dt / (2 dt_old), tweight_u = dt / (dt_old + dt)."""
    g, L = G, 2
    umax = 0.0 if case == "still" else 2.0
    ui, uj = _faces_with_max(g, L, umax)
    if case == "still":
        ui[:], uj[:] = 0.0, 0.0
    dt_old = {"cfl": 0.04, "growth": 0.02, "cap": 0.04, "still": 0.03}[case]
    max_dt = 0.01 if case == "cap" else 1.0
    dt, twa, twu = O.cabaret_adapt_dt(0.1, max_dt, 1.5, ui, uj, dt_old)
    exp = {"cfl": 0.1 / 2.0, "growth": 1.5 * 0.02, "cap": 0.01, "still": 0.03}[case]
    assert dt == pytest.approx(exp, rel=1e-15)
    assert twa == pytest.approx(dt / (2 * dt_old), rel=1e-15)
    assert twu == pytest.approx(dt / (dt_old + dt), rel=1e-15)
    if case != "still":
        assert O.cabaret_cfl(0.5, ui, uj) == pytest.approx(1.0, rel=1e-15)


# ------------------------------------------------------- the limiter
def _extrap_state(g, L, rng):
    z, zo = node(g, L, rng=rng), node(g, L, rng=rng)
    ui = np.asfortranarray(rng.uniform(0.1, 1.0, (g + 3, g + 2, L)) * rng.choice([-1, 1], (g + 3, g + 2, L)))
    uj = np.asfortranarray(rng.uniform(0.1, 1.0, (g + 2, g + 3, L)) * rng.choice([-1, 1], (g + 2, g + 3, L)))
    zfi = np.asfortranarray(rng.standard_normal((g + 3, g + 2, L)))
    zfj = np.asfortranarray(rng.standard_normal((g + 2, g + 3, L)))
    return [z, zo, ui, uj, zfi, zfj]


def _run_extrap(dt, s):
    s = [a.copy(order="F") for a in s]
    O.cabaret_extrapolator(dt, *s)
    return s[4], s[5]


# (case, psi_n, psi_o, fi2 (upwind face), fi1 (the face), expected):
# f = 2 psi_n - fi2 bounded by max/min(fi1, fi2, psi_n) - d_f with
# d_f = -2 (psi_n - psi_o) - dt v (fi1 - fi2); dt v = 1/2 here.
LIMITER_CASES = [
    ("clip_max", 1.0, 1.0, 0.0, 0.0, 1.0),       # f = 2 above the cell's max 1
    ("shift_by_flux", 1.0, 1.0, 0.0, 0.5, 1.25),  # d_f = -1/4 lifts the max
    ("tendency", 1.0, 0.75, 1.0, 1.0, 1.5),       # d_f = -1/2 lifts the min above f
    ("free", 1.0, 1.0, 0.9, 1.2, 1.1),            # inside the bounds: f itself
    ("clip_min", -1.0, -1.0, 0.0, 0.0, -1.0),     # f = -2 below the min -1
]


@pytest.mark.parametrize("case", LIMITER_CASES, ids=[c[0] for c in LIMITER_CASES])
@pytest.mark.parametrize("direction", ["i+", "i-", "j+", "j-"])
def test_limiter_known_answers(case, direction):
    """Interior face (5, 6) (i-faces) or (6, 5) (j-faces) with uniform
    velocity +-1 and dt = 1/2: the upwind cell is the node behind the face
    in the flow direction and the upwind face the one behind that cell."""
    _, pn, po, f2, f1, exp = case
    g, L = G, 1
    z, zo = node(g, L, 0.0), node(g, L, 0.0)
    ui, uj, zfi, zfj = fi(g, L), fj(g, L), fi(g, L), fj(g, L)
    axis, sgn = direction[0], (1.0 if direction[1] == "+" else -1.0)
    f, u = (zfi, ui) if axis == "i" else (zfj, uj)
    u[:] = sgn
    # 0-based positions: the face, its upwind cell and the upwind face
    if axis == "i":
        face = (4, 5, 0)
        cell = (3, 5, 0) if sgn > 0 else (4, 5, 0)
        up = (3, 5, 0) if sgn > 0 else (5, 5, 0)
    else:
        face = (5, 4, 0)
        cell = (5, 3, 0) if sgn > 0 else (5, 4, 0)
        up = (5, 3, 0) if sgn > 0 else (5, 5, 0)
    z[cell], zo[cell], f[up], f[face] = pn, po, f2, f1
    out_i, out_j = _run_extrap(0.5, [z, zo, ui, uj, zfi, zfj])
    got = (out_i if axis == "i" else out_j)[face]
    assert got == exp


@pytest.mark.parametrize("direction", ["i+", "i-", "j+", "j-"])
def test_limiter_wall_faces(direction):
    """The first face from a wall, flow leaving the wall: the upwind value is
    the wall node psi_n itself (f = psi_n) and d_f = -(psi_n - psi_o), so
    with psi_n = 1, psi_o = 1/2 and the face at 2 the min bound
    min(2, 1, 1) + 1/2 = 3/2 applies."""
    g, L = G, 1
    z, zo = node(g, L, 0.0), node(g, L, 0.0)
    ui, uj, zfi, zfj = fi(g, L), fj(g, L), fi(g, L), fj(g, L)
    axis, sgn = direction[0], (1.0 if direction[1] == "+" else -1.0)
    f, u = (zfi, ui) if axis == "i" else (zfj, uj)
    u[:] = sgn
    # data node 2 / G+1 is the wall node; the faces next to it are 3 / G+1
    w = 1 if sgn > 0 else g  # 0-based wall node
    fc = 2 if sgn > 0 else g  # 0-based face
    if axis == "i":
        cell, face = (w, 6, 0), (fc, 6, 0)
    else:
        cell, face = (6, w, 0), (6, fc, 0)
    z[cell], zo[cell], f[face] = 1.0, 0.5, 2.0
    out_i, out_j = _run_extrap(0.5, [z, zo, ui, uj, zfi, zfj])
    assert (out_i if axis == "i" else out_j)[face] == 1.5


def _mirror_i(s, g):
    """Reflection i -> G+3-i of nodes; i-face i (between nodes i-1 and i)
    goes to face G+4-i, so both node and i-face arrays simply reverse along
    i, with ui negated."""
    z, zo, ui, uj, zfi, zfj = s
    r = lambda a: np.asfortranarray(a[::-1])
    return [r(z), r(zo), r(-ui), r(uj), r(zfi), r(zfj)]


def _transpose(s):
    """i <-> j: nodes transposed, i-faces become j-faces and back."""
    z, zo, ui, uj, zfi, zfj = s
    t = lambda a: np.asfortranarray(a.transpose(1, 0, 2))
    return [t(z), t(zo), t(uj), t(ui), t(zfj), t(zfi)]


def test_limiter_mirror_symmetry():
    """Reflecting the basin in i and reversing the flow commutes with the
    extrapolator, bit for bit (upwind branches, walls and bounds of
    both signs map onto each other)."""
    g, L = G, 2
    rng = np.random.default_rng(7)
    s = _extrap_state(g, L, rng)
    a_i, a_j = _run_extrap(0.3, s)
    m = _mirror_i(s, g)
    b_i, b_j = _run_extrap(0.3, m)
    # faces 3..G+1 (i) and the j-faces of nodes 3..G are what the routine writes
    np.testing.assert_array_equal(a_i[::-1], b_i)
    np.testing.assert_array_equal(a_j[::-1], b_j)


def test_limiter_transpose_symmetry():
    """Transposing the basin (i <-> j, i-faces <-> j-faces) commutes with the
    extrapolator, bit for bit."""
    g, L = G, 2
    rng = np.random.default_rng(8)
    s = _extrap_state(g, L, rng)
    a_i, a_j = _run_extrap(0.3, s)
    b_i, b_j = _run_extrap(0.3, _transpose(s))
    np.testing.assert_array_equal(a_i.transpose(1, 0, 2), b_j)
    np.testing.assert_array_equal(a_j.transpose(1, 0, 2), b_i)
