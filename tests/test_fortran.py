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
