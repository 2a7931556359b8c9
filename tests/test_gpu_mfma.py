"""PQ_PREC_FP32_MFMA ("fp32m"): the row sine transforms on the matrix cores
(fp16 hi/lo split operands, FP32 accumulation) against the FP64 oracle.

The oracle restates pv_inversion and helm (qg3p_aux.F90:452-521,
qg3p_helmholtz.F90:614-769).  The mode exists for transform length 1024
(g = 1025), the closed basin; the column pass is the FP32 mode's.
Tolerances: 4e-7 on the seeded k^-3 workload (measured 1.8e-7, the FP32
FFT path's figure), 2e-6 on white noise (flat spectrum, the hardest input
for relative error), as for the FP32 transforms.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pequod_amd import qg, strat as S, synth
from pequod_amd._lib import PqError

from helpers import from_dev, random_strat, rel_err, to_dev

pytestmark = pytest.mark.gpu

G = 1025


def _oracle_inv(g, st, zeta):
    psi, rc = O.Inversion(g, st.ss, st.modes_to_z, st.z_to_modes).run(zeta)
    assert rc == 0
    return psi


def test_mfma_k3_workload(dev):
    """3-layer 1023^2 seeded k^-3 PV: psi within 4e-7 of the FP64 oracle,
    zero trapezoid mean per layer, and the same error as the FP32 FFT
    path within a factor 1.5."""
    import torch
    c = synth.config("qg3_1023")
    st = c.stratification()
    q = c.q()
    print("pv_inversion fp32m qg3_1023", flush=True)
    ref = _oracle_inv(G, st, q)
    errs = {}
    for prec in ("fp32m", "fp32"):
        hc = qg.HelmholtzCache(G, st, prec)
        zt = to_dev(q, dev, torch.float32)
        pt = torch.zeros_like(zt)
        hc.run_dev([zt], [pt])
        assert hc.status()[0] == 0
        out = from_dev(pt)
        errs[prec] = rel_err(out, ref)
        if prec == "fp32m":
            for l in range(st.layers):
                assert abs(O.mean(out[:, :, l])) < 1e-6 * np.max(np.abs(ref[:, :, l]))
    print(f"rel err fp32m {errs['fp32m']:.3e} fp32 {errs['fp32']:.3e}")
    assert errs["fp32m"] < 4e-7
    assert errs["fp32m"] < 1.5 * errs["fp32"]


@pytest.mark.parametrize("L", [1, 3, 8])
def test_mfma_white_noise(dev, L):
    """White-noise PV, 1, 3 and 8 layers (the mass correction's mode loop),
    within 2e-6 of the oracle."""
    import torch
    print(f"pv_inversion fp32m white noise L={L}", flush=True)
    rng = np.random.default_rng(900 + L)
    s1, s2 = random_strat(L, rng)
    st = S.from_parameters(np.ones(L), s1[1:], s2[:-1])
    z = np.asfortranarray(rng.standard_normal((G, G, L)))
    ref = _oracle_inv(G, st, z)
    hc = qg.HelmholtzCache(G, st, "fp32m")
    zt = to_dev(z, dev, torch.float32)
    pt = torch.zeros_like(zt)
    hc.run_dev([zt], [pt])
    assert hc.status()[0] == 0
    assert rel_err(from_dev(pt), ref) < 2e-6


def test_mfma_batch_views(dev):
    """A batch of 8 members (more row blocks than the persistent grid has
    workgroups) as sdata views of (G+2)^2 arrays: each member as the oracle
    gives it, the halo untouched."""
    import torch
    c = synth.config("qg3_1023")
    st = c.stratification()
    B = 8
    print("pv_inversion fp32m batch 8 views", flush=True)
    hc = qg.HelmholtzCache(G, st, "fp32m", batch_max=B)
    zs, ps, refs = [], [], []
    for b in range(B):
        q = c.q(member=b)
        full = np.zeros((G + 2, G + 2, st.layers), order="F")
        full[1:-1, 1:-1] = q
        refs.append(_oracle_inv(G, st, q))
        zs.append(to_dev(full, dev, torch.float32))
        ps.append(torch.zeros_like(zs[-1]))
    hc.run_dev([z[:, 1:, 1:] for z in zs], [p[:, 1:, 1:] for p in ps])
    assert hc.status()[0] == 0
    for b in range(B):
        out = from_dev(ps[b])
        assert rel_err(out[1:-1, 1:-1], refs[b]) < 4e-7, b
        assert np.all(out[0, :, :] == 0) and np.all(out[-1, :, :] == 0)
        assert np.all(out[:, 0, :] == 0) and np.all(out[:, -1, :] == 0)


def test_mfma_helm_boundary(dev):
    """helm (mode 1) with FP32 storage on the matrix cores: interior within
    2e-6 of the oracle, and helm's boundary semantics exactly."""
    import torch
    print("helm fp32m", flush=True)
    rng = np.random.default_rng(1500)
    ss = np.array([0.0, 0.05, 0.4])
    z = np.asfortranarray(rng.standard_normal((G, G, 3)).astype(np.float32).astype(np.float64))
    ref = z.copy(order="F")
    O.Helm(G, ss).solve(ref)
    t = to_dev(z, dev, torch.float32)
    qg.HelmCache(G, G, ss, "fp32m").helm(t)
    torch.cuda.synchronize()
    out = from_dev(t)
    assert rel_err(out[1:-1, 1:-1], ref[1:-1, 1:-1]) < 2e-6
    for sl in [np.s_[0, :], np.s_[-1, :], np.s_[:, 0], np.s_[:, -1]]:
        np.testing.assert_array_equal(out[sl], ref[sl])


def test_mfma_rejects_other_shapes(dev):
    """Other transform lengths, other row tiles and the channel are refused."""
    st = synth.config("qg3_255").stratification()
    with pytest.raises(PqError):
        qg.HelmholtzCache(257, st, "fp32m")
    with pytest.raises(PqError):
        qg.ChannelCache(G, st, precision="fp32m")
    hc = qg.HelmholtzCache(G, st, "fp32m", tiles=None)
    assert hc.tiles()[0] == 16
    assert not hc.set_tiles(8, 2)
    assert hc.set_tiles(16, 1)


def test_mfma_model_steps(dev):
    """A CABARET model whose inversion runs on the matrix cores (FP32
    fields) follows the FP32-transform model: after 5 steps at Courant
    number 0.3, psi agrees to 1e-6 relative."""
    c = synth.config("qg3_1023")
    st = c.stratification()
    q = c.q()
    nd = c.nondim()
    print("model fp32m 1023^2 x 3, 5 steps", flush=True)
    kw = dict(viscosity=nd["viscosity"], visc_bot=nd["visc_bot"], dt=0.05 * nd["dt"], bc_type=1,
              alpha=nd["alpha"])
    out = {}
    for prec in ("fp32m", "fp32"):
        m = qg.Model(G, st, nd["beta"], c.wind(), precision=prec, **kw)
        m.init(q)
        m.step(5)
        assert m.status() == 0
        out[prec] = m.get("psi")
        m.close()
    assert rel_err(out["fp32m"][1:-1, 1:-1], out["fp32"][1:-1, 1:-1]) < 1e-6
