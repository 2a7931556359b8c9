"""Tile autotuner for the inversion passes.

The row passes work on tiles of R rows (also the height of the scratch
blocks) and the column pass on C adjacent columns per workgroup; the best
shape depends on the transform length, the layer count, the precision and
the batch (LDS per workgroup sets the occupancy, C*R the contiguous run of
the column pass).  `sweep` times every accepted (R, C) on the device with
HIP events around each pass; `best` picks the fastest total.  Results of
earlier sweeps live in tiles.json next to this file (keyed by size class)
and are applied by `apply` when a handle is created.  Results do not depend
on the tiles (only the order of the FP64 psiC sum changes).
"""
from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tiles.json")

ROWS = (2, 4, 8, 16)
COLS = (1, 2, 4, 8)


def key(g: int, layers: int, precision: str, batch: int) -> str:
    """Size class: transform length, layers, precision and batch bucket."""
    bb = 1 if batch <= 1 else (8 if batch <= 8 else 64)
    return f"N{g - 1}_L{layers}_{precision}_B{bb}"


def _table() -> dict:
    try:
        with open(TABLE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(g: int, layers: int, precision: str, batch: int):
    """Tuned (R, C) for this size class, or None."""
    e = _table().get(key(g, layers, precision, batch))
    return None if e is None else (int(e["R"]), int(e["C"]))


def apply(hc) -> bool:
    """Set the tuned tiles on a HelmholtzCache if the table has its class."""
    t = lookup(hc.g, hc.layers, hc.precision, hc.batch_max)
    return bool(t) and hc.set_tiles(*t)


def sweep(hc, zetas, psis, reps: int = 10):
    """Time every accepted tile shape on device arrays (lists of tensors of
    the handle's batch).  Returns [(R, C, [ms_fwd, ms_cols, ms_inv])]."""
    r0 = hc.tiles()
    out = []
    try:
        for R in ROWS:
            for C in COLS:
                if not hc.set_tiles(R, C):
                    continue
                out.append((R, C, hc.pass_times(zetas, psis, reps=reps)))
    finally:
        hc.set_tiles(*r0)
    return out


def best(results):
    """(R, C) with the smallest total pass time."""
    R, C, _ = min(results, key=lambda r: sum(r[2]))
    return R, C


def record(g: int, layers: int, precision: str, batch: int, results) -> dict:
    """Store the best shape of a sweep in tiles.json; returns the entry."""
    R, C = best(results)
    ms = dict(((r, c), t) for r, c, t in results)[(R, C)]
    tab = _table()
    tab[key(g, layers, precision, batch)] = {"R": R, "C": C, "pass_ms": [round(x, 5) for x in ms]}
    with open(TABLE, "w") as f:
        json.dump(tab, f, indent=1, sort_keys=True)
    return tab[key(g, layers, precision, batch)]


def sweep_ensemble(ens, reps: int = 10, log=None):
    """Tile shapes of a finalized Ensemble's size classes, tuned in place:
    class by class (the others fixed), every accepted (R, C) is timed as
    the whole ensemble's inversion (wall time of `reps` eager runs, so the
    classes' overlap on the device counts), and the fastest is kept and
    stored in tiles.json under the class's key (grid, largest layer count,
    precision, member count bucket).  Returns {class: (R, C, ms)}."""
    import ctypes
    import time

    from ._lib import load

    lib = load()

    def run_ms():
        ens.invert()
        ens.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ens.invert()
        ens.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    out = {}
    tab = _table()
    for k in range(ens.class_count):
        h = lib.pq_ens_class_handle(ens._h, k)
        res = []
        for R in ROWS:
            for C in COLS:
                if lib.pq_inv_set_tiles(h, R, C) == 0:
                    res.append((run_ms(), R, C))
        ms, R, C = min(res)
        lib.pq_inv_set_tiles(h, R, C)
        g, L, n = ens.class_shape(k)
        prec = ens.class_precision(k)
        tab[key(g, L, prec, n)] = {"R": R, "C": C, "ensemble_ms": round(ms, 4)}
        out[k] = (R, C, ms)
        if log:
            log(f"[sweep] ensemble class {k} (g={g}, L<={L}, {n} members, {prec}): R={R} C={C} "
                f"{ms:.3f} ms")
    with open(TABLE, "w") as f:
        json.dump(tab, f, indent=1, sort_keys=True)
    return out
