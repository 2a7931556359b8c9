"""The oracle's restatement of the reference's output file format
(oracle/io_format.py) against the format's specification: the Fortran
E26.17E3 edit descriptor, Fortran-style records, and the meta-data layout
(qg3p_io.F90:382-438, 653-732; manual/manual.tex:1665-1775).  No GPU."""
import os
import struct

import numpy as np
import pytest

from oracle import io_format as F


@pytest.mark.parametrize("x,s", [
    (0.75e6, "0.75000000000000000E+006"),
    (1.0, "0.10000000000000000E+001"),
    (-2.5e-3, "-0.25000000000000001E-002"),  # 2.5e-3 is 0.00250000000000000005...: rounded
    (0.0, "0.00000000000000000E+000"),
    (1e-300, "0.10000000000000000E-299"),
    (12345.678, "0.12345678000000000E+005"),
    (np.nextafter(1.0, 0.0), "0.99999999999999989E+000"),
    (3.0e5, "0.30000000000000000E+006"),
])
def test_e26_17e3(x, s):
    """[-]0.<17 digits>E<sign><3 digits>, right-justified in 26 columns;
    the 17 digits are the correctly rounded significant digits."""
    out = F.fortran_e26_17e3(x)
    assert len(out) == 26 and out.strip() == s
    assert float(out.replace("E", "e")) == x


def test_record_layout():
    """A 4-byte little-endian byte count, the doubles column-major, the
    byte count again."""
    a = np.arange(6, dtype=float).reshape((2, 3))
    r = F.record(a)
    assert len(r) == 4 + 48 + 4
    assert struct.unpack("<i", r[:4])[0] == 48 == struct.unpack("<i", r[-4:])[0]
    assert np.array_equal(np.frombuffer(r[4:-4], "<f8"), a.ravel(order="F"))
    assert F.record(2.5) == struct.pack("<i", 8) + struct.pack("<d", 2.5) + struct.pack("<i", 8)


def test_meta_layout():
    """Four header lines (6, 3, 4 integers / 2 reals) and the variables."""
    t = F.meta_text(65, 3, 1.0e6, 5.0e5, F.snapshot_variables(3, False)).splitlines()
    assert [len(x.split()) for x in t[:4]] == [6, 3, 4, 2]
    assert t[0].split() == ["1", "1", "1", "0", "1", "1"] and t[1].split() == ["65", "65", "3"]
    assert t[2].split() == ["0", "0", "65", "65"]
    assert [float(v.replace("E", "e")) for v in t[3].split()] == [1.0e6, 5.0e5]
    assert t[4:] == ["model_time_day", "total_time_day", "psi1  ii jj", "psi2  ii jj", "psi3  ii jj",
                     "zeta1  ii jj", "zeta2  ii jj", "zeta3  ii jj"]
    lf = F.snapshot_variables(2, True)
    assert lf[2] == "psi1  ii jj leapfrog"


def test_snapshot_round_trip(tmp_path):
    g, L = 9, 2
    rng = np.random.default_rng(0)
    psi, zeta, psi2, zeta2 = (np.asfortranarray(rng.standard_normal((g + 2, g + 2, L))) for _ in range(4))
    base = F.write_snapshot(str(tmp_path), g, L, 1.5, 11.5, psi, zeta, index=4)
    assert os.path.basename(base) == "qg2_9_s.4"
    recs = F.read_records(base + ".d")
    assert [r.size for r in recs] == [1, 1] + [g * g] * 4
    assert recs[0][0] == 1.5 and recs[1][0] == 11.5
    assert np.array_equal(recs[3].reshape((g, g), order="F"), psi[1:g + 1, 1:g + 1, 1])
    base = F.write_snapshot(str(tmp_path), g, L, 0.0, 0.0, psi, zeta, psi2=psi2, zeta2=zeta2)
    recs = F.read_records(base + ".d")
    assert [r.size for r in recs] == [1, 1] + [2 * g * g] * 4
    assert np.array_equal(recs[2][g * g:].reshape((g, g), order="F"), psi2[1:g + 1, 1:g + 1, 0])
