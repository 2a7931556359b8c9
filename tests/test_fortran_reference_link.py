"""The Fortran drop-in modules against the reference's own modules: the
reference program qg3 (qg3.F90 and every module it uses, compiled unchanged
from /root/reference) with only the renamed import of qg_aux
(qg3p_aux.F90:13-17) switched from qg_helmholtz to qg_helmholtz_hip, plus
qg_cabaret_hip compiled against the reference's qg_parallel and
qg_diagnostics, links against libpequod_hip.so.  Compile and link only, in
a temporary directory: nothing built from the reference is kept, and the
test is skipped where the reference sources are absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
FC = "/opt/rocm/llvm/bin/amdflang"
LIB = os.path.join(ROOT, "pequod_amd", "libpequod_hip.so")
FDIR = os.path.join(ROOT, "pequod_amd", "fortran")

pytestmark = pytest.mark.skipif(
    not (os.path.isdir(REF) and os.path.exists(FC) and os.path.exists(LIB)
         and os.path.exists("/opt/conda/include/mpif.h")),
    reason="needs the reference sources, amdflang, MPICH and the built library")

FFLAGS = ["-cpp", "-O1", "-fdefault-real-8", "-fdefault-double-8", "-DNDEBUG", f"-I{REF}",
          "-I/opt/conda/include"]
REF_MODULES = ["qg3p_params.F90", "test_tools.F90", "qg3p_base.F90", "lapack_interfaces.F90",
               "qg3p_diagnostics.F90", "qg3p_laplace_bc.F90", "qg3p_helmholtz.F90",
               "qg3p_helmholtz_petsc.F90"]
SEAM_OLD = ("  use qg_helmholtz, only : solve_helmholtz => helm, &\n"
            "    & helmholtz_solver_cache => helm_cache\n"
            "  use qg_helmholtz, only : allocate, deallocate\n")
SEAM_NEW = ("  use qg_helmholtz_hip, only : solve_helmholtz => helm, &\n"
            "    & helmholtz_solver_cache => helm_cache\n"
            "  use qg_helmholtz_hip, only : allocate, deallocate\n")


def _fc(tmp, src, obj):
    r = subprocess.run([FC, *FFLAGS, "-module-dir", tmp, "-I", tmp, "-c", src, "-o", obj],
                       capture_output=True, text=True, cwd=tmp)
    assert r.returncode == 0, f"{src}:\n{r.stderr[-3000:]}"
    return obj


def test_reference_program_links_with_hip_helmholtz(tmp_path):
    tmp = str(tmp_path)
    objs = [_fc(tmp, os.path.join(REF, m), os.path.join(tmp, m.replace(".F90", ".o")))
            for m in REF_MODULES]
    for m in ("pequod_hip_c.F90", "qg_helmholtz_hip.F90", "qg_cabaret_hip.F90"):
        objs.append(_fc(tmp, os.path.join(FDIR, m), os.path.join(tmp, m.replace(".F90", ".o"))))
    aux = open(os.path.join(REF, "qg3p_aux.F90")).read()
    assert aux.count(SEAM_OLD) == 1
    swapped = aux.replace(SEAM_OLD, SEAM_NEW)
    # the one changed line pair is the seam; everything else is the reference's text
    diff = [(a, b) for a, b in zip(aux.splitlines(), swapped.splitlines()) if a != b]
    assert len(diff) == 2 and all("qg_helmholtz_hip" in b for _, b in diff)
    src = os.path.join(tmp, "qg3p_aux.F90")
    with open(src, "w") as f:
        f.write(swapped)
    objs.append(_fc(tmp, src, os.path.join(tmp, "qg3p_aux.o")))
    objs.append(_fc(tmp, os.path.join(REF, "qg3p_io.F90"), os.path.join(tmp, "qg3p_io.o")))
    objs.append(_fc(tmp, os.path.join(REF, "qg3.F90"), os.path.join(tmp, "qg3.o")))
    exe = os.path.join(tmp, "qg3")
    r = subprocess.run([FC, "-o", exe, *objs, f"-L{os.path.dirname(LIB)}", "-lpequod_hip",
                        f"-Wl,-rpath,{os.path.dirname(LIB)}", "-L/opt/conda/lib",
                        "-Wl,-rpath,/opt/conda/lib", "-lfmpich", "-lmpich", "-lmkl_rt"],
                       capture_output=True, text=True, cwd=tmp)
    assert r.returncode == 0, r.stderr[-3000:]
    # qg_aux's helm is the HIP one: the program calls the C ABI's solver
    und = subprocess.run(["nm", "-u", exe], capture_output=True, text=True).stdout
    assert "pq_helm_solve" in und and "pq_helm_create" in und
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libpequod_hip.so" in ldd
    shutil.rmtree(tmp, ignore_errors=True)
