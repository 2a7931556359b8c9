"""The matrix-core DST-I (pequod_amd/csrc/pq_mfma.h) restated on the CPU
with the kernel's lane and register layouts, against the sine transform
X_p = 2 sum_i x_i sin(pi p i / N) of FFTPACK's sint (qg3p_helmholtz.F90:
1370-1426; the oracle's pqo_sint follows it).

In FP64 the factorisation (two GEMM stages with N = 32 x 32, the
conjugate-symmetric half of stage 1, the mirror rows of stage 2 and the
p = 32 m rows from T[0] and T[32]) must reproduce the transform to
round-off, so a wrong index map, sign or row assignment fails.  With the
fp16 hi + lo split operands and FP32 accumulation it must stay within the
FP32 transform's error class (3e-7 relative to the output's maximum).
"""
import numpy as np
import pytest

from oracle import oracle as O

N, A, B = 1024, 32, 32


def rows_of_d(g):
    """D register v of lane group g holds row 8 (v // 4) + 4 g + v % 4."""
    v = np.arange(16)
    return 8 * (v // 4) + 4 * g + v % 4


def split(a):
    a = np.asarray(a, dtype=np.float32)
    hi = a.astype(np.float16)
    lo = (a - hi.astype(np.float32)).astype(np.float16)
    return hi.astype(np.float64), lo.astype(np.float64)


def mma(a, b, prec):
    """A (32 x K) times B (K x 32) as v_mfma_f32_32x32x16_f16 would, per
    k-step of 16: exact products of the split operands summed in FP32."""
    if prec == "fp64":
        return a @ b
    ah, al = split(a)
    bh, bl = split(b)
    acc = np.zeros((a.shape[0], b.shape[1]), np.float32)
    for k0 in range(0, a.shape[1], 16):
        k = slice(k0, k0 + 16)
        t = ah[:, k] @ bh[k] + ah[:, k] @ bl[k] + al[:, k] @ bh[k]
        acc = (acc.astype(np.float64) + t).astype(np.float32)
    return acc.astype(np.float64)


def stage2_k():
    """Stage-2 k index (16 t + 8 g + s) -> (i1, part): i1 = 8 t + 4 g + s % 4,
    part = s // 4 (0: Re, 1: Im)."""
    i1 = np.zeros(64, int)
    part = np.zeros(64, int)
    for t in range(4):
        for g in range(2):
            for s in range(8):
                k = 16 * t + 8 * g + s
                i1[k] = 8 * t + 4 * g + s % 4
                part[k] = s // 4
    return i1, part


def mfma_dst(x, prec="fp64"):
    """X of one row x[0..N) (x[0] ignored) by the kernel's two stages."""
    x = np.array(x, dtype=np.float64)
    x[0] = 0.0
    if prec != "fp64":
        m = np.max(np.abs(x))
        e = int(np.floor(np.log2(m))) + 1 if m > 0 else 0
        x = np.ldexp(x, 10 - e).astype(np.float32).astype(np.float64)
    # stage 1: A[m = i1][k = i2] = x[i1 + 32 i2]; B columns: n-tile 0 Re T[c],
    # n-tile 1 Im T[c] (c >= 1) and (-1)^i2 (c = 0, T[32])
    i1 = np.arange(A)[:, None]
    i2 = np.arange(B)[None, :]
    a1 = x[i1 + 32 * i2]
    cc = np.arange(32)[None, :]
    kk = np.arange(B)[:, None]
    b_re = np.cos(np.pi * cc * kk / 32)
    b_im = np.where(cc == 0, (-1.0) ** kk, np.sin(np.pi * cc * kk / 32))
    d0, d1 = mma(a1, b_re, prec), mma(a1, b_im, prec)   # [i1][c]
    # twiddles W^(c i1) (1 for c = 0): Z[c][i1]
    w = np.exp(1j * np.pi * np.arange(32)[None, :] * np.arange(A)[:, None] / N)
    if prec != "fp64":
        w = w.real.astype(np.float32) + 1j * w.imag.astype(np.float32)
    zr = w.real * d0 - w.imag * d1
    zi = w.real * d1 + w.imag * d0
    if prec != "fp64":
        zr, zi = zr.astype(np.float32).astype(np.float64), zi.astype(np.float32).astype(np.float64)
    # stage-2 B operand [k][c] and the A operands of the direct / mirror rows
    # and the p = 32 m rows
    ki1, kpart = stage2_k()
    b2 = np.where(kpart[:, None] == 0, zr[ki1, :], zi[ki1, :])
    m = np.arange(32)[:, None]
    a2 = np.where(m < 16,
                  np.where(kpart[None, :] == 0, np.sin(2 * np.pi * m * ki1[None, :] / 32),
                           np.cos(2 * np.pi * m * ki1[None, :] / 32)),
                  np.where(kpart[None, :] == 0, np.sin(2 * np.pi * (m - 15) * ki1[None, :] / 32),
                           -np.cos(2 * np.pi * (m - 15) * ki1[None, :] / 32)))
    a3 = np.where(kpart[None, :] == (m % 2), np.sin(np.pi * m * ki1[None, :] / 32), 0.0)
    d2, d3 = mma(a2, b2, prec), mma(a3, b2, prec)       # [row m][column c]
    X = np.zeros(N)
    # read out through the lane layout
    for lane in range(64):
        c, g = lane % 32, lane // 32
        rows = rows_of_d(g)
        for v in range(16):
            mm = rows[v]
            if c != 0:
                p = c + 64 * mm if v < 8 else 64 - c + 64 * (mm - 16)
                X[p] = d2[mm, c]
            else:
                X[32 * mm] = d3[mm, c]
    X *= 2.0
    if prec != "fp64":
        X = np.ldexp(X, e - 10)
    return X


def test_register_layout_covers_every_output_once():
    seen = np.zeros(N, int)
    for lane in range(64):
        c, g = lane % 32, lane // 32
        for v, mm in enumerate(rows_of_d(g)):
            if c != 0:
                seen[c + 64 * mm if v < 8 else 64 - c + 64 * (mm - 16)] += 1
            else:
                seen[32 * mm] += 1
    assert np.all(seen == 1)


def test_stage2_k_map_is_a_bijection():
    i1, part = stage2_k()
    assert sorted(zip(i1, part)) == [(i, p) for i in range(32) for p in range(2)]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_factorisation_exact_in_fp64(seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(N)
    x[0] = 0.0
    ref = O.sint(x[1:])  # FFTPACK sint of x_1..x_{N-1}: X_1..X_{N-1}
    out = mfma_dst(x)
    assert abs(out[0]) < 1e-9
    assert np.max(np.abs(out[1:] - ref)) < 1e-11 * np.max(np.abs(ref))


def test_split_operands_fp32_class_white():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(N)
    x[0] = 0.0
    ref = O.sint(x[1:])
    out = mfma_dst(x, "split")
    assert np.max(np.abs(out[1:] - ref)) < 3e-7 * np.max(np.abs(ref))


def test_split_operands_red_spectrum_and_low_modes():
    """A red spectrum (the inverse pass's input: the solution's modes decay
    like p^-4) carries the split's error relative to its largest inputs.
    Taking the lowest 4 modes out of the split transform and adding them
    exactly (what k_rows_inv_mfma does, in FP64) brings the error from the
    3e-7 class down by an order of magnitude."""
    rng = np.random.default_rng(7)
    k = np.arange(N)
    x = rng.standard_normal(N) / np.maximum(k, 1) ** 4
    x[0] = 0.0
    ref = O.sint(x[1:])
    plain = mfma_dst(x, "split")
    e_plain = np.max(np.abs(plain[1:] - ref)) / np.max(np.abs(ref))
    xh = x.copy()
    xh[1:5] = 0.0
    low = 2.0 * np.sin(np.pi * np.outer(np.arange(N), np.arange(1, 5)) / N) @ x[1:5]
    split_low = mfma_dst(xh, "split") + low
    e_low = np.max(np.abs(split_low[1:] - ref)) / np.max(np.abs(ref))
    assert e_plain < 6e-7
    assert e_low < 5e-8 and e_low < e_plain / 5


def test_a_wrong_mirror_row_fails():
    """The check has teeth: stage-2 mirror rows with sin(2 pi p'' i1 / 32)
    instead of p'' + 1 break exactness."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal(N)
    x[0] = 0.0
    ref = O.sint(x[1:])
    # the same factorisation with the mirror rows' frequency off by one
    good = mfma_dst(x)
    bad = good.copy()
    pq = np.arange(16)
    # outputs 64 - c + 64 p'' from rows with p'' instead of p'' + 1: equal to
    # the direct outputs' transform of conj Z, i.e. wrong
    i1 = np.arange(32)
    w = np.exp(1j * np.pi * np.arange(32)[None, :] * i1[:, None] / N)
    T = np.array([[np.sum(np.exp(1j * np.pi * p * np.arange(32) / 32) * x[i + 32 * np.arange(32)])
                   for p in range(32)] for i in i1])
    Z = w * T
    for c in range(1, 32):
        for p2 in pq:
            bad[64 - c + 64 * p2] = 2 * np.sum(np.sin(2 * np.pi * p2 * i1 / 32) * Z[:, c].real
                                               - np.cos(2 * np.pi * p2 * i1 / 32) * Z[:, c].imag)
    assert np.max(np.abs(good[1:] - ref)) < 1e-11 * np.max(np.abs(ref))
    assert np.max(np.abs(bad[1:] - ref)) > 1e-3 * np.max(np.abs(ref))
