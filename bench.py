"""Benchmark of the PV (modal Helmholtz) inversion on one MI355X.

Metric: 3-layer 1023^2 Helmholtz inversions/s at <= 1e-6 relative psi error
vs FP64, with p50/p99 latency.  Workload: 3-layer QG, interior
1023 x 1023 (G = 1025), dx = 5 km, H = [300, 700, 3000] m,
g' = [0.02, 0.01] m/s^2, seeded k^-3 PV with RMS 1e-5 s^-1.

A step is one call of the inversion over a batch of B independent PV fields
already resident in HBM (inputs cycle through a ring of distinct buffers
larger than the 256 MB Infinity Cache).  value = B * steps / wall time.
The same run also reports: the plain FP64 drop-in rate, per-pass kernel
durations (HIP events) with the roofline of the dominant pass, single-call
p50/p99 latency, the max relative psi error of the measured path against
the FP64 oracle, CABARET steps/s of the device-resident FP64 model, and the
oracle's CPU rate on the host (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ensemble_rates(torch, dev, precision, log, members=64, reps=20):
    """Inversions/s of the 64-member mixed ensemble: the batched, graph-
    replayed ensemble in `precision` and the FP64 drop-in (one handle and
    one call per member), with the max relative psi error of the batched
    run against the drop-in over all members."""
    from pequod_amd import qg, synth
    from pequod_amd._lib import IO_BYTES
    from pequod_amd.ensemble import Ensemble

    cfgs = [synth.ensemble_member(k) for k in range(members)]
    dt = torch.float32 if IO_BYTES[precision] == 4 else torch.float64
    fields = []
    for c in cfgs:
        q = c.q(member=c.extra["member"])
        fields.append(torch.from_numpy(np.ascontiguousarray(q.transpose(2, 1, 0))).to(dev))
    log(f"[bench] ensemble: {members} members, "
        f"{len(set((c.n, c.layers) for c in cfgs))} size classes")
    ens = Ensemble(precision)
    zs = [f.to(dt) for f in fields]
    ps = [torch.zeros_like(z) for z in zs]
    for c, z, p in zip(cfgs, zs, ps):
        ens.add(c.stratification(), z, p)
    ens.finalize()
    ens.capture(Ensemble.INVERT)
    ens.invert()
    ens.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ens.invert()
    ens.synchronize()
    t_ens = (time.perf_counter() - t0) / reps
    assert ens.status() == 0
    # plain FP64 drop-in, member by member
    hcs = [qg.HelmholtzCache(c.g, c.stratification(), "fp64") for c in cfgs]
    outs = [torch.zeros_like(f) for f in fields]
    for h, f, o in zip(hcs, fields, outs):
        h.run_dev([f], [o])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for h, f, o in zip(hcs, fields, outs):
            h.run_dev([f], [o])
    torch.cuda.synchronize()
    t_64 = (time.perf_counter() - t0) / reps
    err = 0.0
    for p, o in zip(ps, outs):
        pd, od = p.double(), o
        for l in range(o.shape[0]):
            err = max(err, float((pd[l] - od[l]).abs().max() / od[l].abs().max()))
    chosen = sorted({ens.class_precision(k) for k in range(ens.class_count)})
    ens.close()
    return {"members": members, "precision": precision, "class_precisions": chosen,
            "batched_graph_inversions_per_s": round(members / t_ens, 1),
            "fp64_dropin_inversions_per_s": round(members / t_64, 1),
            "speedup_vs_fp64_dropin": round(t_64 / t_ens, 3),
            "max_rel_psi_err_vs_fp64": err}


def cfl_dt(c, q, cfl=0.3):
    """A fixed time step at Courant number `cfl` of the flow that the PV
    field q inverts to (FP64 inversion on the device; non-dimensional, grid
    spacing 1), capped at the configuration's dt.  The k^-3 fields at 5 km
    spacing exceed Courant number 1 at dt = 1800 s."""
    from pequod_amd import qg
    nd = c.nondim()
    m = qg.Model(c.g, nd["strat"], nd["beta"], c.wind(), viscosity=nd["viscosity"],
                 visc_bot=nd["visc_bot"], dt=nd["dt"], bc_type=1, alpha=nd["alpha"])
    m.init(q)
    psi = m.get("psi")[1:-1, 1:-1, :]
    m.close()
    umax = max(np.max(np.abs(np.diff(psi, axis=0))), np.max(np.abs(np.diff(psi, axis=1))))
    return min(nd["dt"], cfl / umax)


def ensemble_step_rates(torch, dev, precision, log, members=64, steps=200, ref_steps=20):
    """Steps/s of the 64-member mixed ensemble run as CABARET models: one
    ensemble step advances every member by one time step (predictor, PV
    inversion, corrector), fields stored in FP32, the inversions batched per
    size class, the whole step replayed from a HIP graph.  Against FP64
    models (FP64 fields and transforms) stepped one by one, the plain
    drop-in, timed over `ref_steps` steps.  Reports the max relative psi
    error of the FP32 members against FP64 models after the same steps."""
    from pequod_amd import qg, synth
    from pequod_amd.ensemble import Ensemble

    cfgs = [synth.ensemble_member(k) for k in range(members)]

    def make(c, prec, dt):
        nd = c.nondim()
        m = qg.Model(c.g, nd["strat"], nd["beta"], c.wind(), viscosity=nd["viscosity"],
                     visc_bot=nd["visc_bot"], dt=dt, bc_type=1, alpha=nd["alpha"],
                     precision=prec)
        m.init(c.q(member=c.extra["member"]))
        return m, nd["strat"]

    dts = [cfl_dt(c, c.q(member=c.extra["member"])) for c in cfgs]
    log(f"[bench] ensemble step: {members} CABARET members ({precision} models), {steps} steps, "
        f"dt {min(dts):.3g}..{max(dts):.3g} (non-dimensional)")
    mods = [make(c, precision, dt) for c, dt in zip(cfgs, dts)]
    ens = Ensemble(precision)
    for m, st in mods:
        ens.add_model(m, st)
    ens.finalize()
    ens.capture(Ensemble.STEP)
    ens.step(1)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.synchronize()
    t_ens = (time.perf_counter() - t0) / steps
    assert ens.status() == 0
    # FP64 drop-in: the same members as FP64 models, one pq_model_step each
    ref = [make(c, "fp64", dt)[0] for c, dt in zip(cfgs, dts)]
    for r in ref:
        r.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(ref_steps):
        for r in ref:
            r.step(1)
    torch.cuda.synchronize()
    t_64 = (time.perf_counter() - t0) / ref_steps
    for r in ref:
        r.step(steps - ref_steps)
    err = 0.0
    for (m, _), r in zip(mods, ref):
        a, b = m.get("psi"), r.get("psi")
        for l in range(a.shape[2]):
            err = max(err, float(np.max(np.abs(a[:, :, l] - b[:, :, l])) / np.max(np.abs(b[:, :, l]))))
        r.close()
    ens.close()
    for m, _ in mods:
        m.close()
    return {"members": members, "precision": precision, "steps": steps,
            "batched_graph_steps_per_s": round(1.0 / t_ens, 2),
            "member_steps_per_s": round(members / t_ens, 1),
            "fp64_dropin_steps_per_s": round(1.0 / t_64, 2),
            "speedup_vs_fp64_dropin": round(t_64 / t_ens, 3),
            "max_rel_psi_err_vs_fp64_after_steps": err,
            "dt_range_nondim": [round(min(dts), 5), round(max(dts), 5)]}


def accuracy(psi, q, st, psi_ref=None):
    """The error measures of one inversion result psi (Fortran (G, G, L),
    FP64 copy) of the PV field q under stratification st:
    - max relative psi error against the FP64 result psi_ref;
    - max relative q error: the PV that psi inverts to on the interior,
      Delta_5 psi + S psi, against q (per layer, over max |q|);
    - the discrete Helmholtz residual per mode m on the interior,
      (Delta_5 - ss_m) phi_m - zeta_m with phi = z_to_modes psi and
      zeta = z_to_modes q (the mass correction adds homogeneous solutions
      only), relative to max |zeta_m| and, as a backward error, to
      (8 + ss_m) max |phi_m| (round-off level for a backward-stable solve).
    Computed in FP64 on the host."""
    def lap(p):
        return p[2:, 1:-1] + p[:-2, 1:-1] + p[1:-1, 2:] + p[1:-1, :-2] - 4.0 * p[1:-1, 1:-1]

    L = st.layers
    out = {}
    if psi_ref is not None:
        out["max_rel_psi_err_vs_fp64"] = float(max(
            np.max(np.abs(psi[:, :, l] - psi_ref[:, :, l])) / np.max(np.abs(psi_ref[:, :, l]))
            for l in range(L)))
    qi = q[1:-1, 1:-1, :]
    pv = np.stack([lap(psi[:, :, l]) for l in range(L)], 2) + \
        np.einsum("lm,ijm->ijl", st.s_mat, psi[1:-1, 1:-1, :])
    out["max_rel_q_err"] = float(max(np.max(np.abs(pv[:, :, l] - qi[:, :, l])) /
                                     np.max(np.abs(qi[:, :, l])) for l in range(L)))
    phi = np.einsum("ml,ijl->ijm", st.z_to_modes, psi)
    zt = np.einsum("ml,ijl->ijm", st.z_to_modes, q)
    rr, rb = [], []
    for m in range(L):
        r = np.max(np.abs(lap(phi[:, :, m]) - st.ss[m] * phi[1:-1, 1:-1, m] - zt[1:-1, 1:-1, m]))
        rr.append(float(r / np.max(np.abs(zt[1:-1, 1:-1, m]))))
        rb.append(float(r / ((8.0 + st.ss[m]) * np.max(np.abs(phi[:, :, m])))))
    out["helmholtz_residual_per_mode"] = {"rel_to_rhs": rr, "backward": rb}
    return out


def cpu_baselines(O, g, st, q, log, seconds=4.0):
    """Inversions/s of the FP64 CPU port (the oracle's pv_inversion, same
    algorithm and loop order as the reference) on this host, for its -O3
    AVX2/FMA build and the -ffast-math build that mirrors the reference's
    optimised flags: one core, and one independent inversion per thread on
    every core the box gives this process (the reference has no threading;
    its throughput equivalent is one serial solve per core)."""
    import concurrent.futures as cf
    import platform

    cores = min(os.cpu_count() or 1, 16)
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass

    def rate(build, threads):
        invs = [O.Inversion(g, st.ss, st.modes_to_z, st.z_to_modes, build_name=build)
                for _ in range(threads)]
        invs[0].run(q)

        def worker(inv):
            n, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < seconds:
                inv.run(q)
                n += 1
            return n, time.perf_counter() - t0

        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            res = list(ex.map(worker, invs))
        return sum(n for n, _ in res) / max(t for _, t in res), sum(n for n, _ in res)

    variants = {}
    for build, flags in (("o3", "-O3 -march=x86-64-v3"), ("fast", "-O3 -march=x86-64-v3 -ffast-math")):
        r1, n1 = rate(build, 1)
        rn, nn = rate(build, cores)
        variants[build] = {"flags": flags, "one_core_per_s": round(r1, 3),
                           "all_cores_per_s": round(rn, 3)}
        log(f"[bench] CPU port {flags}: {r1:.2f}/s on one core, {rn:.2f}/s on {cores} threads")
    best = max(variants, key=lambda k: variants[k]["all_cores_per_s"])
    return {"value": variants[best]["all_cores_per_s"], "unit": "inversions/s", "cores": cores,
            "kind": "port",
            "sample": f"3-layer 1023^2 FP64 inversions, {seconds:.0f} s per build and thread count; "
                      f"value: the {variants[best]['flags']} build, one inversion per thread",
            "cpu_model": model, "nproc": os.cpu_count(), "variants": variants}


def large_rates(torch, dev, log, batch=2, reps=20):
    """The 6-layer 2047^2 workload (dx 2.5 km, H = [150, 250, 400, 700, 1200,
    2300] m, g' 0.03 -> 0.005): inversions/s in FP32 (storage and
    transforms), FP32 storage with FP64 transforms, and the FP64 drop-in, at
    batch `batch`; the reduced precisions' max relative psi error against the
    FP64 result of the same field (itself pinned to the oracle by the
    tests)."""
    from pequod_amd import qg, synth

    c = synth.config("qg6_2047")
    st = c.stratification()
    q = c.q()
    base = torch.from_numpy(np.ascontiguousarray(q.transpose(2, 1, 0))).to(dev)
    out, ref = {}, None
    log(f"[bench] 6-layer 2047^2: batch {batch}")
    for prec in ("fp64", "fp32", "fp32s"):
        dt = torch.float64 if prec == "fp64" else torch.float32
        hc = qg.HelmholtzCache(c.g, st, prec, batch_max=batch)
        zs = [(base * (1.0 + 0.01 * k)).to(dt) for k in range(batch)]
        ps = [torch.empty_like(z) for z in zs]
        hc.run_dev(zs, ps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            hc.run_dev(zs, ps)
        torch.cuda.synchronize()
        rate = batch * reps / (time.perf_counter() - t0)
        assert hc.status()[0] == 0
        r = {"inversions_per_s": round(rate, 1)}
        # the unscaled field's result (member 0 of the batch holds q itself)
        p0 = np.asfortranarray(ps[0].double().cpu().numpy().transpose(2, 1, 0))
        if prec == "fp64":
            ref = p0
            r.update(accuracy(p0, q, st))
        else:
            r.update(accuracy(p0, q, st, ref))
        out[prec] = r
        hc.deallocate()
    return {"workload": "6-layer QG, 2047^2 interior (G=2049), dx 2.5 km, PV inversion",
            "batch": batch, **out}


def serving_rates(torch, dev, log, rate=2000.0, window=1.0, steps=50, seed=11, load=0.5):
    """Serving leg: Poisson arrivals (mean `rate` per s over `window` s), 70 %
    inversions of field members and 30 % `steps`-step advances of CABARET
    model members, members resident on the device.  Field members: grids
    {127, 255, 511, 1023, 2047}^2 x 2..6 layers, FP32 storage; model members
    (FP32 fields and transforms): grids {127, 255, 511, 1023}^2 x {2, 4, 6}
    layers.  Returns
    latency percentiles per kind and the completed rate at the offered rate
    (a backlog forms when it exceeds the capacity, and latency is then
    queueing), and again at `load` times the completed rate measured there,
    a load the engine sustains."""
    from pequod_amd import qg, synth
    from pequod_amd.serve import Server

    srv = Server("fp32", cap_bytes=200e9, max_batch=16)
    fields, models = [], []
    keep = []
    for n in (127, 255, 511, 1023, 2047):
        for L in range(2, 7):
            c = synth.steep_member(L, 1.6, n=n, member=100 + 10 * L)
            zt = torch.from_numpy(np.ascontiguousarray(c.q(member=100 + L).transpose(2, 1, 0)))
            zt = zt.to(dev, torch.float32)
            pt = torch.zeros_like(zt)
            keep += [zt, pt]
            fields.append(srv.open_field(c.stratification(), zt, pt))
    for n in (127, 255, 511, 1023):
        for L in (2, 4, 6):
            c = synth.steep_member(L, 1.6, n=n, member=200 + L)
            nd = c.nondim()
            q = c.q(member=200 + L)
            models.append(srv.open_model(c.g, nd["strat"], nd["beta"], c.wind(), q,
                                         viscosity=nd["viscosity"], visc_bot=nd["visc_bot"],
                                         dt=cfl_dt(c, q), bc_type=1, alpha=nd["alpha"],
                                         precision="fp32"))
    # warm-up: one request per member; statistics start after it
    for m in fields:
        srv.submit(m, Server.FIELD)
    for m in models:
        first = srv.submit(m, Server.MODEL, 1)
    srv.drain(600)
    first += 1
    used = srv.used_bytes
    rng = np.random.default_rng(seed)

    def run(rate, window, first):
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * window * 1.5) + 10))
        arrivals = arrivals[arrivals < window]
        kinds = rng.random(arrivals.size) < 0.7
        targets = [int(rng.integers(len(fields))) if k else int(rng.integers(len(models))) for k in kinds]
        log(f"[bench] serving: {arrivals.size} requests over {window} s at {rate:.0f}/s "
            f"({len(fields)} field members, {len(models)} model members)")
        t0 = time.perf_counter()
        last = first
        tick = 10.0
        for n_sub, (a, inv, tgt) in enumerate(zip(arrivals, kinds, targets)):
            while time.perf_counter() - t0 < a:
                pass
            if inv:
                last = srv.submit(fields[tgt], Server.FIELD)
            else:
                last = srv.submit(models[tgt], Server.MODEL, steps)
            if a >= tick:  # progress of long windows
                log(f"[bench] serving: {n_sub + 1} requests submitted after {a:.0f} s")
                tick += 10.0
        # drain, reporting progress every 10 s (a backlog forms when the
        # offered rate exceeds the capacity)
        from pequod_amd._lib import PqError
        while True:
            try:
                srv.drain(10.0)
                break
            except PqError as e:
                if "pending" not in str(e) or time.perf_counter() - t0 > 1800:
                    raise
                log(f"[bench] serving: draining ({e}) at {time.perf_counter() - t0:.0f} s")
        makespan = time.perf_counter() - t0
        out = {"offered_per_s": round(rate, 1), "window_s": window, "requests": int(arrivals.size),
               "served_per_s": round(arrivals.size / makespan, 1)}
        for name, kind in (("inversion", Server.FIELD), (f"advance_{steps}", Server.MODEL), ("all", -1)):
            st = srv.stats(kind, first)
            out[name] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()
                         if k in ("count", "p50_ms", "p95_ms", "p99_ms", "max_ms", "failed")}
        return out, last + 1

    out, nxt = run(rate, window, first)
    out["hbm_used_gb"] = round(used / 1e9, 2)
    sustained_rate = load * out["served_per_s"]
    out["sustained"], _ = run(sustained_rate, max(window, 4.0), nxt)
    srv.close()
    return out


def channel_rates(torch, dev, c, st, log, batch=8, reps=50, check=True):
    """The periodic channel (3-layer, gi = 1025, period gj = 2048, partial
    slip): batched inversions/s in FP32 and FP64, the FP32 pass times, and
    the FP32 error against the FP64 oracle on two seeded k^-3 fields."""
    from pequod_amd import qg, synth
    g = c.g
    gj = 2 * (g - 1)
    L = st.layers
    q = np.zeros((g, gj, L), order="F")
    for half in range(2):
        f = synth.k3_field(g - 2, L, synth.seed_for(c.config_id, 500 + half), c.rms) * c.t_scale
        q[:, half * (g - 1):(half + 1) * (g - 1), :] = f[:, :g - 1, :]
    alpha = c.nondim()["alpha"]
    out = {"workload": f"{L}-layer channel {g} x {gj} (walls along i, periodic along j), partial slip",
           "batch_per_step": batch}
    for prec in ("fp32", "fp64"):
        dt = torch.float64 if prec == "fp64" else torch.float32
        base = torch.from_numpy(np.ascontiguousarray(q.transpose(2, 1, 0))).to(dev)
        zs = [base.to(dt) * (1.0 + 1e-3 * k) for k in range(batch)]
        ps = [torch.empty_like(zs[0]) for _ in range(batch)]
        hc = qg.ChannelCache(g, st, 1, alpha, precision=prec, batch_max=batch)
        for _ in range(3):
            hc.run_dev(zs, ps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            hc.run_dev(zs, ps)
        torch.cuda.synchronize()
        out[f"{prec}_inversions_per_s"] = round(batch * reps / (time.perf_counter() - t0), 1)
        assert hc.status()[0] == 0
        if prec == "fp32":
            out["fp32_pass_ms"] = [round(x, 5) for x in hc.pass_times(zs, ps, reps=10)]
            if check:
                from oracle import oracle as O
                ref, rc = O.Channel(g, gj, st.ss, st.modes_to_z, st.z_to_modes, 1, alpha).run(q)
                hc.run_dev([base.to(dt)], [ps[0]])
                torch.cuda.synchronize()
                o = ps[0].double().cpu().numpy().transpose(2, 1, 0)
                out["fp32_max_rel_psi_err_vs_fp64"] = float(max(
                    np.max(np.abs(o[:, :, l] - ref[:, :, l])) / np.max(np.abs(ref[:, :, l]))
                    for l in range(L)))
        hc.deallocate()
        log(f"[bench] channel {prec}: {out[f'{prec}_inversions_per_s']} inversions/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--precision", default="fp32m",
                    choices=["fp32m", "fp32", "fp32w", "fp32s", "mixed", "fp64"],
                    help="fp32m (default): FP32 storage, row transforms on the matrix cores (fp16 "
                         "hi/lo split operands, FP32 accumulation); fp32: FP32 storage and "
                         "transforms; fp32w: FP32 storage, FP64 arithmetic; fp32s: FP32 storage, "
                         "FP64 transforms; mixed: FP64 storage, FP32 transforms; fp64: plain drop-in")
    ap.add_argument("--ensemble-precision", default="fp32auto",
                    choices=["fp32auto", "fp32", "fp32w", "fp32s", "mixed", "fp64"],
                    help="precision of the batched 64-member ensemble leg")
    ap.add_argument("--ring", type=int, default=24, help="distinct input/output buffers")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cabaret", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-member mixed-size ensemble leg")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 6-layer 2047^2 leg")
    ap.add_argument("--no-ensemble-step", action="store_true",
                    help="skip the 64-member ensemble time-step (CABARET models) leg")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving leg (Poisson request stream, continuous batching)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="serving leg: requests per second")
    ap.add_argument("--serve-window", type=float, default=1.0, help="serving leg: arrival window [s]")
    ap.add_argument("--no-channel", action="store_true", help="skip the periodic channel leg")
    ap.add_argument("--no-latency", action="store_true",
                    help="skip the B=1 latency leg (every inversion launch then has the bench's batch)")
    ap.add_argument("--sweep-tiles", action="store_true",
                    help="time every tile shape of the inversion passes, record the best and exit")
    ap.add_argument("--tiles", default="auto", help="'auto' (tuned table), 'default', or R,C")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 10 device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)

    import torch

    from pequod_amd import qg, synth

    dev = torch.device("cuda:0")
    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 10 copies of {args.copy_probe} MiB")
        return
    c = synth.config("qg3_1023")
    st = c.stratification()
    g, L, B = c.g, st.layers, args.batch
    q = c.q()  # (G, G, L) Fortran order, non-dimensional

    def dev_field(a, dtype):
        return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).to(dev, dtype)

    from pequod_amd import autotune

    def make_cache(prec, batch):
        if args.tiles == "auto":
            return qg.HelmholtzCache(g, st, prec, batch_max=batch)
        if args.tiles == "default":
            return qg.HelmholtzCache(g, st, prec, batch_max=batch, tiles=None)
        return qg.HelmholtzCache(g, st, prec, batch_max=batch,
                                 tiles=tuple(int(x) for x in args.tiles.split(",")))

    from pequod_amd._lib import IO_BYTES

    def io_dtype(prec):
        return torch.float32 if IO_BYTES[prec] == 4 else torch.float64

    def run_rate(prec, steps, warmup, batch):
        dt = io_dtype(prec)
        hc = make_cache(prec, batch)
        base = dev_field(q, dt)
        nring = max(args.ring, batch)
        # distinct inputs: the seeded field with a per-buffer amplitude
        zs = [base * (1.0 + 0.01 * k) for k in range(nring)]
        ps = [torch.empty_like(base) for _ in range(nring)]
        torch.cuda.synchronize()

        def call(k):
            idx = [(k * batch + b) % nring for b in range(batch)]
            hc.run_dev([zs[i] for i in idx], [ps[i] for i in idx])

        for k in range(warmup):
            call(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            call(k)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rc, _ = hc.status()
        assert rc == 0, f"inversion status {rc}"
        return hc, zs, ps, (t1 - t0), base

    if args.sweep_tiles:
        for prec in ("fp64", "fp32", "fp32w", "fp32s"):
            for bt in (1, B):
                dt = io_dtype(prec)
                hc = qg.HelmholtzCache(g, st, prec, batch_max=bt, tiles=None)
                base = dev_field(q, dt)
                zs = [base * (1.0 + 0.01 * k) for k in range(bt)]
                ps = [torch.empty_like(base) for _ in range(bt)]
                res = autotune.sweep(hc, zs, ps, reps=10)
                for R, C, ms in res:
                    log(f"[sweep] {prec} B={bt} R={R} C={C} passes {[round(x, 4) for x in ms]} "
                        f"total {sum(ms):.4f} ms")
                e = autotune.record(g, L, prec, bt, res)
                log(f"[sweep] {prec} B={bt} best {e}")
                hc.deallocate()
        # size classes of the 64-member mixed ensemble (FP32 storage)
        from pequod_amd.ensemble import Ensemble
        cfgs = [synth.ensemble_member(k) for k in range(64)]
        ens = Ensemble("fp32auto")
        keep = []
        for cm in cfgs:
            zt = torch.from_numpy(np.ascontiguousarray(cm.q(member=cm.extra["member"]).transpose(2, 1, 0)))
            zt = zt.to(dev, torch.float32)
            keep += [zt, torch.zeros_like(zt)]
            ens.add(cm.stratification(), keep[-2], keep[-1])
        ens.finalize(tiles=None)
        autotune.sweep_ensemble(ens, log=log)
        ens.close()
        return

    log(f"[bench] 3-layer 1023^2 {args.precision} batch {B}: {args.steps} steps")
    hc, zs, ps, wall, base = run_rate(args.precision, args.steps, args.warmup, B)
    ms_per_step = 1000.0 * wall / args.steps
    value = B * args.steps / wall

    # per-pass kernel durations (HIP events on the handle's stream)
    torch.cuda.synchronize()
    passes = hc.pass_times(zs[:B], ps[:B], reps=20)
    el = IO_BYTES[args.precision]                 # zeta / psi element size
    el_t = 8 if args.precision == "fp64" else 4   # scratch element size
    names = ["k_rows_fwd", "k_cols", "k_rows_inv"]
    dom = int(np.argmax(passes))
    # algorithmic bytes per launch: one read + one write of the L x G x G
    # field per member (B_inv = 2 L G^2 s); the column pass moves the
    # scratch field (n^2 per layer) once each way.
    alg = {
        "k_rows_fwd": B * L * g * g * el + B * L * (g - 1) ** 2 * el_t,
        "k_cols": B * 2 * L * (g - 1) ** 2 * el_t,
        "k_rows_inv": B * L * (g - 1) ** 2 * el_t + B * L * g * g * el,
    }
    ach = alg[names[dom]] / (passes[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(f"{args.precision}_b{B}", {}).get(names[dom])
        except Exception:
            traffic = None

    # single-call latency (B = 1), HIP events on torch's stream
    p50 = p99 = None
    if not args.no_latency:
        hc1, zs1, ps1, _, _ = run_rate(args.precision, 5, 5, 1)
        lat = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for k in range(200):
            e0.record()
            hc1.run_dev([zs1[k % len(zs1)]], [ps1[k % len(ps1)]])
            e1.record()
            e1.synchronize()
            lat.append(e0.elapsed_time(e1))
        p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))

    # plain FP64 drop-in rate on the same workload, and the FP32 FFT path
    fp64_rate = fp32_rate = None
    if args.precision != "fp64":
        _, _, _, w64, _ = run_rate("fp64", max(20, args.steps // 4), 5, B)
        fp64_rate = B * max(20, args.steps // 4) / w64
    if args.precision not in ("fp32", "fp64"):
        _, _, _, w32, _ = run_rate("fp32", max(20, args.steps // 4), 5, B)
        fp32_rate = B * max(20, args.steps // 4) / w32

    # device-resident CABARET models (same grid, FP64 and FP32 fields) and
    # the centred scheme: steps/s at a fixed dt of Courant number 0.3
    cab = cab32 = lfr = None
    if not args.no_cabaret:
        nd = c.nondim()
        dt_run = cfl_dt(c, q)
        nst = 200

        def model_rate(**extra):
            m = qg.Model(g, st, nd["beta"], c.wind(), viscosity=nd["viscosity"],
                         visc_bot=nd["visc_bot"], dt=dt_run, bc_type=1, alpha=nd["alpha"], **extra)
            m.init(q)
            m.step(5)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            m.step(nst)
            torch.cuda.synchronize()  # device-wide: includes the model's stream
            rate = nst / (time.perf_counter() - t0)
            assert np.all(np.isfinite(m.get("psi")))
            m.close()
            return rate

        cab = model_rate()
        cab32 = model_rate(precision="fp32")
        lfr = model_rate(scheme="leapfrog")

    # 64-member mixed ensemble (sizes {255, 511, 767, 1023}^2, 2..8 layers,
    # per-member stratification): batched per size class and replayed from a
    # HIP graph, against the plain FP64 drop-in called member by member.
    ens = None
    if not args.no_ensemble:
        ens = ensemble_rates(torch, dev, args.ensemble_precision, log)

    ens_step = None
    if not args.no_ensemble_step:
        ens_step = ensemble_step_rates(torch, dev, args.ensemble_precision, log)

    large = None
    if not args.no_large:
        large = large_rates(torch, dev, log)

    chan = None
    if not args.no_channel:
        chan = channel_rates(torch, dev, c, st, log, check=not args.no_cpu)

    serving = None
    if not args.no_serving:
        serving = serving_rates(torch, dev, log, rate=args.serve_rate, window=args.serve_window)

    # CPU baseline (the FP64 port on the host's cores) on a bounded sample,
    # and the error measures of the measured path against the FP64 oracle
    # (the checker build) on the same input.
    cpu = None
    err = acc = acc_ref = None
    if not args.no_cpu:
        from oracle import oracle as O
        inv = O.Inversion(g, st.ss, st.modes_to_z, st.z_to_modes)
        ref, rc = inv.run(q)
        assert rc == 0
        cpu = cpu_baselines(O, g, st, q, log)
        # error of the measured path on the unscaled seeded field
        out = torch.empty_like(base)
        hc.run_dev([base], [out])
        torch.cuda.synchronize()
        o = np.asfortranarray(out.double().cpu().numpy().transpose(2, 1, 0))
        acc = accuracy(o, q, st, ref)
        acc_ref = accuracy(ref, q, st)
        err = acc["max_rel_psi_err_vs_fp64"]

    line = {
        "metric": "3-layer 1023^2 Helmholtz inversions/s at <=1e-6 rel err vs FP64; p50/p99 ms",
        "value": round(value, 1),
        "unit": "inversions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32m": "f32 storage; row transforms f16 hi+lo on MFMA, f32 accumulate",
                  "fp32": "f32", "fp32s": "f32 storage, f64 transforms",
                  "fp32w": "f32 storage, f64 arithmetic",
                  "mixed": "f64 storage, f32 transforms", "fp64": "f64"}[args.precision],
        "data": "synthetic (seeded k^-3 PV, rms 1e-5 s^-1)",
        "config": {"workload": "3-layer QG, 1023^2 interior (G=1025), dx 5 km, PV inversion",
                   "batch_per_step": B, "precision": args.precision, "ring_buffers": max(args.ring, B),
                   "tiles_RC": list(hc.tiles())},
        "p50_ms": None if p50 is None else round(p50, 4),
        "p99_ms": None if p99 is None else round(p99, 4),
        "max_rel_psi_err_vs_fp64": err,
        "max_rel_q_err": None if acc is None else acc["max_rel_q_err"],
        "helmholtz_residual_per_mode": None if acc is None else acc["helmholtz_residual_per_mode"],
        "fp64_oracle_accuracy": acc_ref,
        "fp64_dropin_inversions_per_s": None if fp64_rate is None else round(fp64_rate, 1),
        "fp32_fft_inversions_per_s": None if fp32_rate is None else round(fp32_rate, 1),
        "pass_ms": {n: round(t, 5) for n, t in zip(names, passes)},
        "cabaret_fp64_steps_per_s": None if cab is None else round(cab, 2),
        "cabaret_fp32_steps_per_s": None if cab32 is None else round(cab32, 2),
        "leapfrog_fp64_steps_per_s": None if lfr is None else round(lfr, 2),
        # one full CABARET step (inversion included) moves at least 21 L G^2
        # values: 10 field reads and 9 writes of the step's state plus the
        # inversion's read and write
        "cabaret_step_roofline": None if cab is None else {
            "bytes_per_step": {"fp64": 21 * L * g * g * 8, "fp32": 21 * L * g * g * 4},
            "achieved_GBs": {"fp64": round(21 * L * g * g * 8 * cab / 1e9, 1),
                             "fp32": round(21 * L * g * g * 4 * cab32 / 1e9, 1)},
            "frac": {"fp64": round(21 * L * g * g * 8 * cab / 1e9 / HBM_PEAK_GBS, 4),
                     "fp32": round(21 * L * g * g * 4 * cab32 / 1e9 / HBM_PEAK_GBS, 4)}},
        "ensemble_64_mixed": ens,
        "ensemble_64_mixed_steps": ens_step,
        "channel": chan,
        "qg6_2047": large,
        "serving": serving,
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
