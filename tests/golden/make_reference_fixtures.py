"""Writes the reference-trajectory fixtures (tests/golden/ref_*.npz) by
running pequod's own Fortran routines in this container.

The reference's modules are compiled unchanged from /root/reference into a
temporary directory outside the repository, together with ref_driver.F90
(the driver's start and time-step order calling those routines).  Nothing
built here is kept or travels: only the fixtures (inputs and the states the
reference reached) are written.  Needs amdflang, MPICH (mpif.h,
libfmpich) and MKL (LAPACK), which this container has; run it as
    python tests/golden/make_reference_fixtures.py
The compiled reference's own qg_parameters fix layers = 3 and at most
257 x 257 nodes, so the cases are 3-layer grids of 17, 33, 65 and 257
nodes.  The fixtures hold the states only; the tests regenerate the inputs
with pequod_amd.synth (seeded, bit-reproducible).

Inputs are the hot path's synthetic workloads (pequod_amd.synth): the
seeded k^-3 PV field, the non-dimensional 3-layer stratification of
H = [300, 700, 3000] m, g' = [0.02, 0.01] m/s^2 at dx = 10 km, beta, the
wind (qg3.F90:253-279 with shift 0, tilt 0, asymmetry 1), viscosity and
bottom drag, as in the CPU-runnable 255^2 configuration.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
FC = "/opt/rocm/llvm/bin/amdflang"
# the reference's modules in dependency order (makefile:21-56)
MODULES = ["qg3p_params.F90", "test_tools.F90", "qg3p_base.F90", "lapack_interfaces.F90",
           "qg3p_diagnostics.F90", "qg3p_laplace_bc.F90", "qg3p_helmholtz.F90",
           "qg3p_helmholtz_petsc.F90", "qg3p_aux.F90"]
# IEEE arithmetic in statement order (the oracle's rounding), no fast-math
FFLAGS = ["-cpp", "-O2", "-ffp-contract=off", "-fdefault-real-8", "-fdefault-double-8",
          "-DNDEBUG", f"-I{REF}", "-I/opt/conda/include"]
LIBS = ["-L/opt/conda/lib", "-Wl,-rpath,/opt/conda/lib", "-lfmpich", "-lmpich", "-lmkl_rt"]

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def build(tmp: str) -> str:
    objs = []
    for src in MODULES + [os.path.join(HERE, "ref_driver.F90")]:
        path = src if os.path.isabs(src) else os.path.join(REF, src)
        obj = os.path.join(tmp, os.path.basename(src).replace(".F90", ".o"))
        subprocess.run([FC, *FFLAGS, "-module-dir", tmp, "-I", tmp, "-c", path, "-o", obj],
                       check=True, cwd=tmp)
        objs.append(obj)
    exe = os.path.join(tmp, "ref_driver")
    subprocess.run([FC, "-o", exe, *objs, *LIBS], check=True, cwd=tmp)
    return exe


FIELDS = [("psi", "nd"), ("zeta", "nd"), ("rel", "nd"), ("zeta_old", "nd"),
          ("zeta_fi", "fi"), ("ui", "fi"), ("uimo", "fi"), ("flux_i", "fi"),
          ("zeta_fj", "fj"), ("uj", "fj"), ("ujmo", "fj"), ("flux_j", "fj"),
          ("res", "rs"), ("res_beta", "rs")]


def shape(kind, g, L):
    return {"nd": (g + 2, g + 2, L), "fi": (g + 3, g + 2, L), "fj": (g + 2, g + 3, L),
            "rs": (g, g, L)}[kind]


def case_inputs(g: int, member: int = 0):
    """Non-dimensional inputs of the 3-layer 10 km workload on g x g nodes."""
    from pequod_amd import synth
    c = synth.config("qg3_255")
    c.n = g - 2
    nd = c.nondim()
    st = nd["strat"]
    return dict(g=g, L=3, dt=nd["dt"], beta=nd["beta"], viscosity=nd["viscosity"],
                visc_bot=nd["visc_bot"], alpha=nd["alpha"], h=st.h, s1=st.s1[1:], s2=st.s2[:-1],
                force_wind=c.wind(), zeta0=c.q(member=member))


def input_check(inp):
    """Input checksums stored with each fixture, so a test can tell that it
    regenerated the same inputs."""
    return dict(in_dt=inp["dt"], in_q_abs_sum=np.sum(np.abs(inp["zeta0"])),
                in_wind_abs_sum=np.sum(np.abs(inp["force_wind"])))


def run_case(exe, tmp, inp, nsteps, out_steps, adapt=0, bc_type=1, target_cfl=0.1,
             max_dt=None, max_dt_increase=1.5):
    g, L = inp["g"], inp["L"]
    max_dt = inp["dt"] * 48.0 if max_dt is None else max_dt
    fin, fout, fdt = (os.path.join(tmp, n) for n in ("in.bin", "out.bin", "dt.bin"))
    with open(fin, "wb") as f:
        np.array([g, nsteps, adapt, bc_type, len(out_steps), *out_steps], np.int32).tofile(f)
        np.array([inp["dt"], *inp["beta"], inp["viscosity"], inp["visc_bot"], inp["alpha"],
                  target_cfl, max_dt, max_dt_increase], np.float64).tofile(f)
        np.concatenate([inp["h"], inp["s1"], inp["s2"]]).astype(np.float64).tofile(f)
        np.asfortranarray(inp["force_wind"]).ravel(order="F").tofile(f)
        np.asfortranarray(inp["zeta0"]).ravel(order="F").tofile(f)
    subprocess.run([exe, fin, fout, fdt], check=True, cwd=tmp, capture_output=True)
    raw = np.fromfile(fout, dtype=np.uint8)
    dts = np.fromfile(fdt, dtype=np.float64)
    out, pos = {}, 0
    for _ in out_steps:
        step = int(raw[pos:pos + 4].view(np.int32)[0])
        pos += 4
        sc = raw[pos:pos + 32].view(np.float64)
        pos += 32
        rec = dict(dt=sc[0], tweight_a=sc[1], tweight_u=sc[2], model_time=sc[3])
        for name, kind in FIELDS:
            shp = shape(kind, g, L)
            n = int(np.prod(shp)) * 8
            rec[name] = raw[pos:pos + n].view(np.float64).reshape(shp, order="F").copy(order="F")
            pos += n
        out[step] = rec
    assert pos == raw.size
    return out, dts


def summary(a: np.ndarray, stride: int):
    """Per-layer max |.|, L2 norm, min, max and a strided subsample."""
    L = a.shape[2]
    return dict(linf=np.array([np.max(np.abs(a[:, :, l])) for l in range(L)]),
                l2=np.array([np.sqrt(np.sum(a[:, :, l] ** 2)) for l in range(L)]),
                vmin=a.min(axis=(0, 1)), vmax=a.max(axis=(0, 1)),
                sub=np.ascontiguousarray(a[::stride, ::stride, :]))


def main():
    tmp = tempfile.mkdtemp(prefix="pq_ref_")
    try:
        exe = build(tmp)
        # 1. 17^2 x 3, every field after 1, 2 and 10 steps (partial slip);
        # no slip and free slip: the state after 1 and 10 steps
        for bc in (1, 2, 3):
            inp = case_inputs(17)
            steps = [1, 2, 10] if bc == 1 else [1, 10]
            out, dts = run_case(exe, tmp, inp, 10, steps, bc_type=bc)
            keep = [f for f, _ in FIELDS] if bc == 1 else ["psi", "zeta", "rel", "zeta_fi",
                                                           "zeta_fj"]
            arrs = input_check(inp)
            for s_, rec in out.items():
                for k in ("dt", "tweight_a", "tweight_u", "model_time", *keep):
                    arrs[f"s{s_}_{k}"] = np.asarray(rec[k])
            np.savez_compressed(os.path.join(HERE, f"ref_traj_17_bc{bc}.npz"), g=17, L=3, bc=bc,
                                member=0, **arrs)
            print(f"ref_traj_17_bc{bc}: dt {dts[0]:.6g}")
        # 33^2 x 3 after 10 steps: norms and a subsample of the state
        inp = case_inputs(33)
        out, dts = run_case(exe, tmp, inp, 10, [10])
        arrs = input_check(inp)
        for name in ("psi", "zeta", "zeta_fi", "ui", "res"):
            for k, v in summary(out[10][name], 4).items():
                arrs[f"s10_{name}_{k}"] = v
        np.savez_compressed(os.path.join(HERE, "ref_traj_33_bc1.npz"), g=33, L=3, bc=1, member=0,
                            **arrs)
        # 2. 257^2 x 3 (the 255^2 interior workload), 100 steps with the
        # reference's adaptive time step (target Courant number 0.1)
        inp = case_inputs(257)
        out, dts = run_case(exe, tmp, inp, 100, [100], adapt=1)
        rec = out[100]
        arrs = dict(input_check(inp), dts=dts)
        for k in ("dt", "tweight_a", "tweight_u", "model_time"):
            arrs[f"s100_{k}"] = rec[k]
        for name in ("psi", "zeta", "zeta_fi", "ui", "res"):
            for k, v in summary(rec[name], 16).items():
                arrs[f"s100_{name}_{k}"] = v
        np.savez_compressed(os.path.join(HERE, "ref_traj_257_adapt.npz"), g=257, L=3, bc=1,
                            member=0, nsteps=100, **arrs)
        print(f"ref_traj_257_adapt: dt {dts[0]:.6g} .. {dts[-1]:.6g}")
        # 3. 65^2 x 3 with adapt_dt, 300 steps: the dt sequence
        inp = case_inputs(65, member=1)
        out, dts = run_case(exe, tmp, inp, 300, [300], adapt=1)
        rec = out[300]
        arrs = dict(input_check(inp), dts=dts)
        for name in ("psi", "zeta"):
            for k, v in summary(rec[name], 4).items():
                arrs[f"s300_{name}_{k}"] = v
        np.savez_compressed(os.path.join(HERE, "ref_dt_65_adapt.npz"), g=65, L=3, bc=1, member=1,
                            nsteps=300, model_time=rec["model_time"], **arrs)
        print(f"ref_dt_65_adapt: {len(dts)} steps, model time {rec['model_time']:.6g}")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
