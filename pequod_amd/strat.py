"""Vertical structure on the host, in FP64 (the reference keeps this on the
host too): the stratification matrix S, its eigendecomposition and the
layer <-> mode conversion matrices (qg3p_diagnostics.F90:25-47, 210-422).

S is tridiagonal with positive off-diagonal products, so a diagonal
similarity makes it symmetric; numpy's symmetric eigensolver then gives the
modes.  The barotropic mode (ss = 0, eigenvector of ones) is moved first as
in sort_barotropic (qg3p_diagnostics.F90:389-418); the baroclinic modes are
ordered by ascending ss.  The inversion's result does not depend on the
eigenvector scaling, sign or the order of the baroclinic modes.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Stratification:
    layers: int
    h: np.ndarray          # non-dimensional layer thicknesses
    s1: np.ndarray         # coupling to the layer above (s1[0] unused)
    s2: np.ndarray         # coupling to the layer below (s2[L-1] unused)
    s_mat: np.ndarray      # L x L stratification matrix
    ss: np.ndarray         # modal Helmholtz constants (-eigenvalues of S)
    modes_to_z: np.ndarray  # L x L, column m is mode m
    z_to_modes: np.ndarray  # inverse of modes_to_z

    @property
    def rd(self) -> np.ndarray:
        """Deformation radii 1/sqrt(ss) of the baroclinic modes."""
        return 1.0 / np.sqrt(self.ss[1:])


def s_matrix(s1: np.ndarray, s2: np.ndarray) -> np.ndarray:
    """Form S as s_matrix does (qg3p_diagnostics.F90:342-370)."""
    L = len(s1)
    S = np.zeros((L, L))
    for i in range(L):
        if i > 0:
            S[i, i - 1] = s1[i]
            if i < L - 1:
                S[i, i] = -(s1[i] + s2[i])
                S[i, i + 1] = s2[i]
            else:
                S[i, i] = -s1[i]
        elif i < L - 1:
            S[i, i] = -s2[i]
            S[i, i + 1] = s2[i]
    return S


def from_parameters(h, s1, s2, l_scale: float | None = None) -> Stratification:
    """allocate_stratification_parameters (qg3p_diagnostics.F90:210-237):
    s1 has entries for layers 2..L and s2 for layers 1..L-1 (length L-1
    each), dimensional if l_scale is given."""
    h = np.asarray(h, dtype=np.float64)
    L = len(h)
    S1 = np.zeros(L)
    S2 = np.zeros(L)
    f = 1.0 if l_scale is None else l_scale ** 2
    if L > 1:
        S1[1:] = f * np.asarray(s1, dtype=np.float64)
        S2[:-1] = f * np.asarray(s2, dtype=np.float64)
    hn = h if l_scale is None else h / l_scale
    return _finish(L, hn, S1, S2)


def from_density(h, gprime, f0: float, dx: float) -> Stratification:
    """Non-dimensional stratification on a grid of spacing dx (the reference's
    l_scale) from layer depths h [m], reduced gravities g'_{i+1/2} [m s^-2]
    (length L-1) and f0 [s^-1]: s2(i) = f0^2 / (g'_{i+1/2} h_i),
    s1(i) = f0^2 / (g'_{i-1/2} h_i)."""
    h = np.asarray(h, dtype=np.float64)
    gp = np.asarray(gprime, dtype=np.float64)
    L = len(h)
    S1 = np.zeros(L)
    S2 = np.zeros(L)
    for i in range(L):
        if i < L - 1:
            S2[i] = f0 ** 2 / (gp[i] * h[i]) * dx ** 2
        if i > 0:
            S1[i] = f0 ** 2 / (gp[i - 1] * h[i]) * dx ** 2
    return _finish(L, h / dx, S1, S2)


def from_rd(rd, h, l_scale: float | None = None) -> Stratification:
    """allocate_stratification_l2 / _l3 from deformation radii
    (qg3p_diagnostics.F90:149-208, strat_l2 / strat_l3 at 256-321)."""
    h = np.asarray(h, dtype=np.float64)
    L = len(h)
    S1 = np.zeros(L)
    S2 = np.zeros(L)
    if L == 2:
        fsq = h[0] * h[1] / (2.0 * rd[0] ** 2)
        S2[0] = fsq * 2.0 / (h[0] * (h[0] + h[1]))
        S1[1] = fsq * 2.0 / (h[1] * (h[0] + h[1]))
    elif L == 3:
        lam1, lam2 = -1.0 / rd[0] ** 2, -1.0 / rd[1] ** 2
        b1, c1 = -(lam1 + lam2), lam1 * lam2
        w1 = 2.0 / (h[0] * (h[0] + h[1]))
        w2 = 2.0 / (h[1] * (h[0] + h[1]))
        w3 = 2.0 / (h[1] * (h[1] + h[2]))
        w4 = 2.0 / (h[2] * (h[1] + h[2]))
        ll = w1 * w3 + w1 * w4 + w2 * w4
        a2, b2, c2 = ll * (w3 + w4), -b1 * ll, c1 * (w1 + w2)
        f2 = (-b2 - np.sqrt(b2 * b2 - 4.0 * a2 * c2)) / (2.0 * a2)
        f1 = (b1 - f2 * (w3 + w4)) / (w1 + w2)
        S2[0], S1[1], S2[1], S1[2] = f1 * w1, f1 * w2, f2 * w3, f2 * w4
    elif L != 1:
        raise ValueError("from_rd supports 1, 2 or 3 layers")
    if l_scale is not None:
        S1 *= l_scale ** 2
        S2 *= l_scale ** 2
        h = h / l_scale
    return _finish(L, h, S1, S2)


def _finish(L, h, S1, S2) -> Stratification:
    S = s_matrix(S1, S2)
    if L == 1:
        one = np.ones((1, 1))
        return Stratification(1, h, S1, S2, S, np.zeros(1), one, one.copy())
    d = np.ones(L)
    for i in range(L - 1):
        if not (S2[i] > 0 and S1[i + 1] > 0):
            raise ValueError("stratification couplings must be positive")
        d[i + 1] = d[i] * np.sqrt(S2[i] / S1[i + 1])
    B = np.diag(d) @ S @ np.diag(1.0 / d)
    for i in range(L - 1):
        B[i, i + 1] = B[i + 1, i] = np.sqrt(S2[i] * S1[i + 1])
    w, V = np.linalg.eigh(B)
    ss = -w
    order = np.argsort(ss)
    ss = ss[order]
    V = (V / d[:, None])[:, order]
    ss[0] = 0.0
    V[:, 0] = 1.0
    m2z = np.asfortranarray(V)
    z2m = np.asfortranarray(np.linalg.inv(m2z))
    return Stratification(L, h, S1, S2, S, ss, m2z, z2m)
