"""Writes tests/golden/oracle_dt_1023_adapt.npz: the adaptive time steps of
the oracle's CABARET model over 1000 steps of the 3-layer 1023^2-interior
workload (dx 5 km, seeded k^-3 PV), the reference's adapt_dt with target
Courant number 0.1.  The oracle's adaptive trajectory is pinned against the
reference's own at 255^2 (tests/test_reference_trajectory.py); the
reference as built here is limited to 257 x 257 nodes, so at 1023^2 the
oracle's dt sequence is the fixture.  Run as
    python tests/golden/make_oracle_dt_fixture.py
(about two minutes on one core)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pequod_amd import synth  # noqa: E402

NSTEPS = 1000


def main():
    c = synth.config("qg3_1023")
    nd = c.nondim()
    st = nd["strat"]
    q, wind = c.q(), c.wind()
    m = O.Model(c.g, st.s1, st.s2, st.ss, st.modes_to_z, st.z_to_modes, nd["beta"], wind,
                viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=1,
                alpha=nd["alpha"], adapt_dt=True, target_cfl=0.1, max_dt=48.0 * nd["dt"],
                max_dt_increase=1.5)
    assert m.init(q) == 0
    dts = np.zeros(NSTEPS)
    for k in range(NSTEPS):
        m.step(1)
        dts[k] = m.scalars()["dt"]
    psi = m.a["psi"][1:-1, 1:-1, :]
    np.savez_compressed(os.path.join(HERE, "oracle_dt_1023_adapt.npz"), g=c.g, nsteps=NSTEPS,
                        dts=dts, model_time=m.scalars()["model_time"],
                        in_q_abs_sum=np.sum(np.abs(q)), in_dt=nd["dt"],
                        psi_linf=np.max(np.abs(psi), axis=(0, 1)),
                        psi_sub=np.ascontiguousarray(psi[::64, ::64, :]))
    print(f"dt {dts[0]:.6g} .. {dts[-1]:.6g}, model time {m.scalars()['model_time']:.6g}")


if __name__ == "__main__":
    main()
