"""The reference's output / checkpoint file format, restated with numpy.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the library's
writer and reader (pq_model_write_snapshot, pq_model_write_mean,
pq_model_read_checkpoint).

A data set is a plain-text meta-data file base[.index] and a binary data
file base[.index].d (open_parallel_header / open_parallel_data,
qg3p_io.F90:382-423; serial runs carry no process suffix).  The meta-data
starts with write_parallel_header (qg3p_io.F90:426-438) and lists the
variables; the data file holds one record per variable, in that order,
written by write_unformatted (qg3p_io.F90:653-732): a default (4-byte)
integer byte count, the 8-byte doubles in column-major order, and the byte
count again (manual/manual.tex:1755-1775).
"""
from __future__ import annotations

import os
import struct

import numpy as np


def fortran_e26_17e3(x: float) -> str:
    """Fortran's E26.17E3 edit descriptor (real_8_chr, qg3p_params.F90:12):
    [-]0.<17 digits>E<sign><3 digits>, right-justified in 26 columns."""
    if x == 0.0:
        mant = ("-" if np.signbit(x) else "") + "0." + "0" * 17 + "E+000"
    else:
        s = f"{abs(x):.16e}"            # d.dddddddddddddddde+kk: 17 significant digits
        digits = s[0] + s[2:18]
        k = int(s[19:]) + 1
        mant = ("-" if x < 0 else "") + "0." + digits + "E" + ("-" if k < 0 else "+") + f"{abs(k):03d}"
    return mant.rjust(26)


def meta_text(g: int, layers: int, l_scale: float, t_scale: float, variables) -> str:
    """write_parallel_header for one serial domain (n = 1, p_jj = p_ii = 1,
    p = 0, p_j = 1, p_ii - p_i + 1 = 1; ii_0 = jj_0 = 0), then the variable
    lines."""
    lines = ["1 1 1 0 1 1", f"{g} {g} {layers}", f"0 0 {g} {g}",
             fortran_e26_17e3(l_scale) + " " + fortran_e26_17e3(t_scale)]
    return "\n".join(lines + list(variables)) + "\n"


def record(a) -> bytes:
    """write_unformatted of one array or scalar."""
    a = np.asarray(a, dtype="<f8")
    data = np.asfortranarray(a).tobytes(order="F")
    n = struct.pack("<i", len(data))
    return n + data + n


def read_records(path: str):
    """All records of a data file as flat float64 arrays (read_unformatted
    checks: both byte counts equal and a whole number of doubles)."""
    out = []
    with open(path, "rb") as f:
        raw = f.read()
    pos = 0
    while pos < len(raw):
        (n,) = struct.unpack_from("<i", raw, pos)
        assert n % 8 == 0 and pos + 8 + n <= len(raw), "truncated record"
        data = np.frombuffer(raw, dtype="<f8", count=n // 8, offset=pos + 4).copy()
        (n2,) = struct.unpack_from("<i", raw, pos + 4 + n)
        assert n2 == n, "record byte counts differ"
        out.append(data)
        pos += 8 + n
    return out


def snapshot_variables(layers: int, leapfrog: bool):
    """write_s_header (qg3p_io.F90:441-471)."""
    shape = "  ii jj leapfrog" if leapfrog else "  ii jj"
    return (["model_time_day", "total_time_day"] + [f"psi{l}{shape}" for l in range(1, layers + 1)]
            + [f"zeta{l}{shape}" for l in range(1, layers + 1)])


def supp_variables(layers: int):
    """write_s_supp_header (qg3p_io.F90:474-511)."""
    v = ["model_time_day", "total_time_day", "dt"]
    for name, shape in [("ui", "flux_i_ii flux_i_jj"), ("uj", "flux_j_ii flux_j_jj"),
                        ("zeta_fi", "flux_i_ii flux_i_jj"), ("zeta_fj", "flux_j_ii flux_j_jj"),
                        ("res", "ii jj"), ("res_beta", "ii jj")]:
        v += [f"{name}{l}  {shape}" for l in range(1, layers + 1)]
    return v


def write_snapshot(directory, g, layers, t_model, t_total, psi, zeta, l_scale=1.0, t_scale=1.0,
                   index=None, psi2=None, zeta2=None):
    """Snapshot data set (qg3.F90:830-856): psi, zeta are (G+2, G+2, L) data
    arrays whose sdata (2:G+1, 2:G+1) is written per layer; with psi2 /
    zeta2 (the centred scheme) both levels share a record
    (write_unformatted_rank2_rank2)."""
    lf = psi2 is not None
    base = os.path.join(directory, f"qg{layers}_{g}_s" + ("" if index is None else f".{index}"))
    with open(base, "w") as f:
        f.write(meta_text(g, layers, l_scale, t_scale, snapshot_variables(layers, lf)))
    recs = [record(t_model), record(t_total)]
    for a, b in ((psi, psi2), (zeta, zeta2)):
        for l in range(layers):
            s = a[1:g + 1, 1:g + 1, l]
            if lf:
                both = np.concatenate([np.asfortranarray(s).ravel(order="F"),
                                       np.asfortranarray(b[1:g + 1, 1:g + 1, l]).ravel(order="F")])
                recs.append(record(both))
            else:
                recs.append(record(s))
    with open(base + ".d", "wb") as f:
        f.write(b"".join(recs))
    return base


def write_supp(directory, g, layers, t_model, t_total, dt, ui, uj, zeta_fi, zeta_fj, res, res_beta,
               l_scale=1.0, t_scale=1.0, index=None):
    """CABARET supplementary data set (qg3.F90:858-885): i-face arrays
    (G+3, G+2, L) as data(3:G+1, 3:G), j-face arrays (G+2, G+3, L) as
    data(3:G, 3:G+1), res / res_beta (G, G, L)."""
    base = os.path.join(directory, f"qg{layers}_{g}_s_supp" + ("" if index is None else f".{index}"))
    with open(base, "w") as f:
        f.write(meta_text(g, layers, l_scale, t_scale, supp_variables(layers)))
    recs = [record(t_model), record(t_total), record(dt)]
    for a in (ui, uj, zeta_fi, zeta_fj):
        fi = a.shape[0] == g + 3
        for l in range(layers):
            recs.append(record(a[2:g + 1, 2:g, l] if fi else a[2:g, 2:g + 1, l]))
    for a in (res, res_beta):
        for l in range(layers):
            recs.append(record(a[:, :, l]))
    with open(base + ".d", "wb") as f:
        f.write(b"".join(recs))
    return base


def write_mean(directory, g, layers, t_model, t_total, psi_av_mean, l_scale=1.0, t_scale=1.0):
    """Time-mean data set (qg3.F90:886-898; write_s_av_header,
    qg3p_io.F90:514-534): psi_av * (1 / diagnostic_time) per layer."""
    base = os.path.join(directory, f"qg{layers}_{g}_s_av")
    with open(base, "w") as f:
        f.write(meta_text(g, layers, l_scale, t_scale,
                          ["model_time_day", "total_time_day"]
                          + [f"psi{l}_av  ii jj" for l in range(1, layers + 1)]))
    recs = [record(t_model), record(t_total)] + [record(psi_av_mean[:, :, l]) for l in range(layers)]
    with open(base + ".d", "wb") as f:
        f.write(b"".join(recs))
    return base
