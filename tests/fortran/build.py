"""Build the Fortran drop-in test programs (tests/fortran/_build/helm_drop_in
and cabaret_drop_in): the product modules pequod_amd/fortran/{pequod_hip_c,
qg_helmholtz_hip,qg_cabaret_hip}.F90 against the qg_parallel and
qg_diagnostics test scaffolds, linked to pequod_amd/libpequod_hip.so (rpath
relative to the binary, so the tree can move).  Returns the helm program's
path, or None when no Fortran compiler is installed."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "helm_drop_in")
EXE_CABARET = os.path.join(OUT, "cabaret_drop_in")


def _fc():
    for c in ("/opt/rocm/llvm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    return None


def build() -> str | None:
    fc = _fc()
    if fc is None:
        return None
    os.makedirs(OUT, exist_ok=True)
    fdir = os.path.join(ROOT, "pequod_amd", "fortran")
    srcs = [os.path.join(fdir, "pequod_hip_c.F90"),
            os.path.join(HERE, "qg_parallel_stub.F90"),
            os.path.join(HERE, "qg_diagnostics_stub.F90"),
            os.path.join(fdir, "qg_helmholtz_hip.F90"),
            os.path.join(fdir, "qg_cabaret_hip.F90"),
            os.path.join(HERE, "helm_drop_in.F90"),
            os.path.join(HERE, "cabaret_drop_in.F90")]
    objs = {}
    for s in srcs:
        name = os.path.splitext(os.path.basename(s))[0]
        o = os.path.join(OUT, name + ".o")
        subprocess.run([fc, "-cpp", "-O2", "-c", s, "-module-dir", OUT, "-I", OUT, "-o", o],
                       check=True, capture_output=True, text=True)
        objs[name] = o
    libdir = os.path.join(ROOT, "pequod_amd")
    rel = os.path.relpath(libdir, OUT)
    common = [objs["pequod_hip_c"], objs["qg_parallel_stub"], objs["qg_diagnostics_stub"]]
    for exe, mods in ((EXE, ["qg_helmholtz_hip", "helm_drop_in"]),
                      (EXE_CABARET, ["qg_cabaret_hip", "cabaret_drop_in"])):
        subprocess.run([fc, "-o", exe, *[objs[m] for m in mods][::-1], *common[::-1],
                        f"-L{libdir}", "-lpequod_hip", f"-Wl,-rpath,$ORIGIN/{rel}"],
                       check=True, capture_output=True, text=True)
    return EXE


if __name__ == "__main__":
    print(build())
