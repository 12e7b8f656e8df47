"""In-house autotuner for the recursion kernels' shapes.

For one geometry (nside, lmax, spin), batch size and precision it times candidate kernel
shapes (``jhtx_tuning`` fields: rings per lane of the FP64 recursion kernels, maps per
recursion, register budget and ring-chunk width of the mixed-precision kernels) with the library's
own HIP-event stage timers, picks the fastest synthesis and adjoint shapes independently
(they are separate kernels), and registers the winner so later plans of that geometry use
it.  FP64 candidates give identical results (the same arithmetic in another shape); mixed
candidates, which include the FP16-split matrix-core kernels, all stay within the mixed
path's 1e-5 contract but differ in their last bits.
"""

from __future__ import annotations

import itertools
import json
import os

import numpy as np
import torch

from .healpix import Plan, _prec, alm_size, set_tuning

# candidate shapes per (precision, spin): synthesis fields, adjoint fields
_CANDIDATES = {
    ("f64", 0): ([{"synth_rings": r} for r in (1, 4, 8)], [{"adj_rings": r} for r in (4, 8, 1)]),
    ("f64", 2): ([{"synth_rings": r} for r in (1, 2, 4)], [{"adj_rings": r} for r in (4, 2, 1)]),
    ("mixed", 0): ([{"mx_synth_group": g} for g in (32, 16, 8)], [{"mx_adj_group": g} for g in (16, 8)]),
    ("mixed", 2): ([{"mx_synth_group": 32}] + [{"mx_synth_group": 16, "mx_f16": f} for f in (1, 2)] +
                   [{"mx_synth_group": g, "mx_synth_occ": o} for g, o in ((8, 2), (4, 1), (8, 1))],
                   [{"mx_adj_group": g, "mx_adj_occ": 2, "mx_adj_cb": 1, "mx_adj_f16": 1} for g in (32, 16)] +
                   [{"mx_adj_group": g, "mx_adj_occ": o, "mx_adj_cb": c, "mx_adj_f16": 2}
                    for g, o, c in ((16, 2, 1), (16, 2, 2), (8, 2, 2), (4, 1, 2), (8, 1, 2))]),
}


def _inputs(nside, lmax, spin, batch, device, seed=0):
    rng = np.random.default_rng(seed)
    K = alm_size(lmax)
    shape = (batch, K) if spin == 0 else (batch, 2, K)
    a = torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), device=device)
    npix = 12 * nside * nside
    maps = torch.empty((batch, npix) if spin == 0 else (batch, 2, npix), dtype=torch.float64, device=device)
    return a, maps, torch.empty_like(a)


def _time_plan(plan: Plan, a, maps, back, batch, steps):
    plan.synth(a, maps, batch)
    plan.adjoint(maps, back, batch)
    torch.cuda.synchronize(plan.device)
    plan.set_profiling(True)
    plan.read_profile()
    for _ in range(steps):
        plan.synth(a, maps, batch)
        plan.adjoint(maps, back, batch)
    prof = plan.read_profile()
    plan.set_profiling(False)
    return prof["synth_rec"][0] / steps, (prof["adj_rec"][0] + prof["adj_finish"][0]) / steps


def autotune(nside: int, lmax: int, spin: int = 0, batch: int = 1, precision: str = "f64",
             steps: int = 2, apply: bool = True, cache: str | None = None, device=None) -> dict:
    """Time the candidate shapes; return ``{"tuning": {...}, "synth_ms": {...}, "adj_ms": {...}}``.

    With ``apply`` the winner is registered (``set_tuning``) for this geometry, precision and
    batch.  ``cache`` names a JSON file: a stored result for the same key is reused
    instead of timing again, and new results are added to it.
    """
    prec = _prec(precision)
    key = f"{nside}/{lmax}/{spin}/{batch}/{precision}"
    if cache and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if key in stored:
            res = stored[key]
            if apply:
                set_tuning(nside, lmax, spin, prec, res["tuning"], batch)
            return res
    synth_c, adj_c = _CANDIDATES.get((precision, spin), ([{}], [{}]))
    if precision == "mixed" and batch == 1:
        # one-map calls of a mixed plan run the FP64 vector-FMA kernels (mx_single): their
        # rings per lane are the shapes to time (the FP64 matrix-core adjoint does not apply)
        synth_c, adj_c = _CANDIDATES[("f64", spin)]
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    a, maps, back = _inputs(nside, lmax, spin, batch, dev)
    t_syn, t_adj = {}, {}
    # synthesis and adjoint are separate kernels: candidate i of each is timed in one plan
    for sc, ac in itertools.zip_longest(synth_c, adj_c):
        plan = Plan(nside, lmax, spin, batch, prec, 0, dev, {**(sc or {}), **(ac or {})})
        ts, ta = _time_plan(plan, a, maps, back, batch, steps)
        del plan
        if sc is not None:
            t_syn[json.dumps(sc, sort_keys=True)] = ts
        if ac is not None:
            t_adj[json.dumps(ac, sort_keys=True)] = ta
    best_s = json.loads(min(t_syn, key=t_syn.get)) if t_syn else {}
    best_a = json.loads(min(t_adj, key=t_adj.get)) if t_adj else {}
    tuning = {**best_s, **best_a}
    res = {"tuning": tuning, "synth_ms": t_syn, "adj_ms": t_adj}
    if apply:
        set_tuning(nside, lmax, spin, prec, tuning, batch)
    if cache:
        stored = {}
        if os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        stored[key] = res
        with open(cache, "w") as f:
            json.dump(stored, f, indent=1, sort_keys=True)
    return res


__all__ = ["autotune"]
