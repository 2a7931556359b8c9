"""The Fortran drop-in (pequod_amd/fortran) over the C ABI: the argument
path on the CPU, and allocate / helm / deallocate on a discrete
eigenfunction on the GPU (exact solution rhs / (lambda_p + lambda_q - ss))."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "_build", "helm_drop_in")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE),
                               reason="Fortran drop-in test program not built (no amdflang)")


@needs_exe
def test_fortran_binding_rejects_bad_grid():
    r = subprocess.run([EXE, "args"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "args ok" in r.stdout


@needs_exe
@pytest.mark.gpu
def test_fortran_helm_drop_in_eigenfunction():
    print("fortran helm drop-in 33x33x2", flush=True)
    r = subprocess.run([EXE, "solve"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    err = float(r.stdout.split()[-1])
    assert err < 1e-13


EXE_CAB = os.path.join(os.path.dirname(EXE), "cabaret_drop_in")


def _cabaret_inputs(g, L, bc, seed):
    """Random state of the shapes the reference passes (FP64), and the
    scalars of one step: dt, target_cfl, max_dt, max_dt_increase,
    viscosity, visc_bot, alpha."""
    import numpy as np
    rng = np.random.default_rng(seed)
    f = lambda *s: np.asfortranarray(rng.standard_normal(s))
    nd, fi, fj, rs = (g + 2, g + 2, L), (g + 3, g + 2, L), (g + 2, g + 3, L), (g, g, L)
    a = dict(s1=rng.uniform(0.05, 0.5, L), s2=rng.uniform(0.05, 0.5, L),
             beta=rng.uniform(0.1, 1.0, L), fw=f(g, g),
             zeta_old=f(*nd), zeta=f(*nd), psi=f(*nd), rel=f(*nd), res=f(*rs), res_beta=f(*rs),
             ui=f(*fi), uimo=f(*fi), zeta_fi=f(*fi), flux_i=f(*fi),
             uj=f(*fj), ujmo=f(*fj), zeta_fj=f(*fj), flux_j=f(*fj))
    sc = (0.29, 0.4, 2.0, 1.5, 0.03, 0.002, 0.7)
    return a, sc


def _write_inputs(path, g, L, bc, a, sc):
    import numpy as np
    with open(path, "wb") as fh:
        fh.write(np.array([g, L, bc], dtype="<i4").tobytes())
        fh.write(np.array(sc, dtype="<f8").tobytes())
        for k in ("s1", "s2", "beta", "fw", "zeta_old", "zeta", "psi", "rel", "res", "res_beta",
                  "ui", "uimo", "zeta_fi", "flux_i", "uj", "ujmo", "zeta_fj", "flux_j"):
            fh.write(np.asarray(a[k], dtype="<f8").tobytes(order="F"))


def _read_outputs(path, g, L):
    import numpy as np
    raw = np.fromfile(path, dtype="<f8")
    out = {"dt": raw[0], "tweight_a": raw[1], "tweight_u": raw[2], "cfl": raw[3]}
    pos = 4
    nd, fi, fj, rs = (g + 2, g + 2, L), (g + 3, g + 2, L), (g + 2, g + 3, L), (g, g, L)
    for k, s in (("zeta", nd), ("rel", nd), ("res", rs), ("res_beta", rs), ("ui", fi), ("uj", fj),
                 ("uimo", fi), ("ujmo", fj), ("zeta_fi", fi), ("zeta_fj", fj), ("flux_i", fi),
                 ("flux_j", fj)):
        n = int(np.prod(s))
        out[k] = raw[pos:pos + n].reshape(s, order="F")
        pos += n
    assert pos == raw.size
    return out


@pytest.mark.skipif(not os.path.exists(EXE_CAB), reason="Fortran drop-in test program not built (no amdflang)")
@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["copy", "resident"])
@pytest.mark.parametrize("g,L,bc", [(17, 1, 2), (33, 3, 1), (33, 4, 3)])
def test_fortran_cabaret_drop_in_step(tmp_path, mode, g, L, bc):
    """One CABARET step through qg_cabaret_hip with the reference's call
    sequence and argument types (qg3.F90:598-952, the inversion left out),
    host arrays copied per call or registered as resident device mirrors,
    against the oracle's stages in the same order: bit for bit."""
    import numpy as np
    from oracle import oracle as O
    print(f"fortran cabaret drop-in {mode} g={g} L={L} bc={bc}", flush=True)
    a, sc = _cabaret_inputs(g, L, bc, 100 * g + 10 * L + bc)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_inputs(fin, g, L, bc, a, sc)
    r = subprocess.run([EXE_CAB, mode, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "cabaret drop-in ok" in r.stdout
    got = _read_outputs(fout, g, L)
    # the oracle, in the same order
    dt, target_cfl, max_dt, inc, visc, visc_bot, alpha = sc
    s = {k: np.array(v, copy=True, order="F") for k, v in a.items()}
    dt, tw_a, tw_u = O.cabaret_adapt_dt(target_cfl, max_dt, inc, s["ui"], s["uj"], dt)
    O.cabaret_predictor(dt, s["zeta_old"], s["res"], s["zeta"])
    O.cabaret_beta(dt, tw_a, s["beta"], s["ui"], s["zeta"], s["res_beta"])
    O.rel_from_zeta_and_psi(s["s1"], s["s2"], bc, alpha, s["zeta"], s["psi"], s["rel"])
    O.update_zeta_bc(s["s1"], s["s2"], s["rel"], s["psi"], s["zeta"])
    cfl = O.cabaret_cfl(dt, s["ui"], s["uj"])
    O.cabaret_forcing(dt, s["fw"], visc, visc_bot, s["rel"], s["zeta"])
    O.cabaret_cell_boundaries(tw_u, s["psi"], s["ui"], s["uj"], s["uimo"], s["ujmo"])
    O.cabaret_extrapolator(dt, s["zeta"], s["zeta_old"], s["ui"], s["uj"], s["zeta_fi"], s["zeta_fj"])
    O.cabaret_flux(s["ui"], s["uj"], s["zeta_fi"], s["zeta_fj"], s["flux_i"], s["flux_j"])
    O.cabaret_corrector(dt, s["flux_i"], s["flux_j"], s["res"], s["zeta"])
    assert (got["dt"], got["tweight_a"], got["tweight_u"], got["cfl"]) == (dt, tw_a, tw_u, cfl)
    for k in ("zeta", "rel", "res", "res_beta", "ui", "uj", "uimo", "ujmo", "zeta_fi", "zeta_fj",
              "flux_i", "flux_j"):
        np.testing.assert_array_equal(got[k], s[k], err_msg=k)
