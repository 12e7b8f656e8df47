"""Benchmark: HEALPix spin-2 synthesis + adjoint on one MI355X.

Workload (the metric's configuration): spin 2, nside = 2048, lmax = 4096, E/B a_lm with
C_l^EE = (l+1)^-2, C_l^BB = 0.1 C_l^EE, l < 2 zero (synthetic, seeded).  One step is one
synthesis S a -> (Q, U) followed by one adjoint S^T (Q, U) -> a', per map of the batch;
inputs are resident in HBM before the timed region.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision f64|mixed] [--batch B]
                    [--nside N --lmax L --spin S] [--no-cpu]

Roofline: the dominant kernel (recursion + contraction, synthesis or adjoint) is timed
live with HIP events recorded by the library on the call's stream; its algorithmic work is
26 flops (spin 2; 8 for spin 0) per (l, m, north-half ring) triple of the active triangle
l >= max(m, |s|), times 2 nside rings.  CPU baseline: the reference's fused masked scan
restated in numpy (oracle/jht_oracle.py) on a sample of orders m over every l and ring,
scaled by (lmax+1)/sample, plus the full-size numpy ring FFTs, single-threaded.
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

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def gaussian_alm(lmax, rng, lmin=0, amp=1.0):
    """Re, Im ~ N(0,1) sqrt(C_l/2) (m = 0 real with variance C_l), C_l = amp (l+1)^-2."""
    m = np.concatenate([np.full(lmax + 1 - mm, mm) for mm in range(lmax + 1)])
    ell = np.concatenate([np.arange(mm, lmax + 1) for mm in range(lmax + 1)])
    cl = amp * (ell + 1.0) ** -2
    re = rng.standard_normal(ell.size)
    im = rng.standard_normal(ell.size)
    a = np.where(m == 0, re * np.sqrt(cl), re * np.sqrt(cl / 2)) + 1j * np.where(m == 0, 0.0, im * np.sqrt(cl / 2))
    a[ell < lmin] = 0.0
    return a


def triples(nside, lmax, spin):
    """Active (l, m) pairs with l >= max(m, |spin|), times 2 nside north-half rings."""
    K = (lmax + 1) * (lmax + 2) // 2
    inactive = 0 if spin == 0 else (3 if lmax >= 1 else 1)  # (0,0), (1,0), (1,1)
    return (K - inactive) * 2 * nside


def cpu_baseline(nside, lmax, spin, rows=8, seed=1):
    """Time the reference algorithm (numpy restatement) on a bounded sample: the fused masked
    l-scan for `rows` orders m over all l and rings (synthesis + adjoint contractions),
    scaled to all lmax+1 orders, plus the full ring-FFT assembly in both directions."""
    from oracle import jht_oracle as O

    assert spin == 2, "the CPU baseline sample is written for the spin-2 workload"
    rng = np.random.default_rng(seed)
    geo = O.RingInfo(nside)
    T = 2 * nside
    x = geo.z[:T]
    M = lmax + 1
    ms = np.unique(np.linspace(0, lmax, rows).astype(np.int64))
    R = ms.size
    bp = rng.standard_normal((R, M)) + 1j * rng.standard_normal((R, M))
    bm = rng.standard_normal((R, M)) + 1j * rng.standard_normal((R, M))
    V = [rng.standard_normal((R, T)) + 1j * rng.standard_normal((R, T)) for _ in range(4)]
    t0 = time.perf_counter()
    O.sampled_spin2_pair(x, lmax, ms, bp, bm, *V)
    t_rec = (time.perf_counter() - t0) * (M / R)
    # ring FFT assembly, both directions, full size
    G = (rng.standard_normal((M, geo.nrings)) + 1j * rng.standard_normal((M, geo.nrings)))
    t0 = time.perf_counter()
    qu = O.ring_synthesis(G, G, geo)
    O.ring_analysis(qu, geo, lmax)
    O.ring_analysis(qu, geo, lmax, minus=True)
    t_fft = time.perf_counter() - t0
    del G, qu
    total = t_rec + t_fft
    return {
        "value": 1.0 / total,
        "unit": "maps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"numpy restatement of jht's fused masked l-scan (spin 2, both +-2 "
                   f"recursions, synthesis + adjoint contractions) for {R} of {M} orders m over "
                   f"all {M} l and {T} rings, scaled x{M / R:.0f} ({t_rec:.1f} s/map), plus "
                   f"full-size numpy ring FFTs both ways ({t_fft:.1f} s/map); 1 thread"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="f64", choices=["f64"])
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=4096)
    ap.add_argument("--spin", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=12)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported "
              "(each map is independent -- multi-GPU runs shard the batch, see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import jht_amd
    from jht_amd.healpix import get_plan

    nside, lmax, spin, B = args.nside, args.lmax, args.spin, args.batch
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(20261203)
    K = jht_amd.alm_size(lmax)
    if spin == 0:
        a = np.stack([gaussian_alm(lmax, rng) for _ in range(B)])
    else:
        a = np.stack([np.stack([gaussian_alm(lmax, rng, 2), gaussian_alm(lmax, rng, 2, 0.1)]) for _ in range(B)])
    alm = torch.tensor(a, device=dev)
    del a
    plan = get_plan(nside, lmax, spin, B)
    npix = 12 * nside * nside
    maps = torch.empty((B, npix) if spin == 0 else (B, 2, npix), dtype=torch.float64, device=dev)
    back = torch.empty_like(alm)

    def step():
        plan.synth(alm, maps, B)
        plan.adjoint(maps, back, B)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.set_profiling(True)
    plan.read_profile()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = plan.read_profile()
    plan.set_profiling(False)

    # size-independent check: <S a, v> == <a, S^T v>_G on the bench data itself
    G = torch.as_tensor(jht_amd.alm_metric_weight(lmax).copy(), device=dev)
    lhs = float((maps.double() * maps.double()).sum())
    rhs = float((G * (alm.conj() * back).real).sum())
    ident = abs(lhs - rhs) / abs(lhs)

    ms_step = dt / args.steps * 1e3
    value = B * args.steps / dt
    per = {k: (v[0] / max(v[1], 1), v[1]) for k, v in prof.items()}
    dom = max(("synth_rec", "adj_rec"), key=lambda k: per[k][0])
    flops = triples(nside, lmax, spin) * (26 if spin else 8) * B
    achieved = flops / (per[dom][0] * 1e-3) / 1e12
    out = {
        "metric": "maps/s synth+adjoint nside=2048 lmax=4096 (spin 2)",
        "value": round(value, 4),
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian E/B a_lm, C_l=(l+1)^-2, B=0.1 E, l<2 zero)",
        "config": {"workload": f"spin-{spin} synthesis+adjoint, nside={nside}, lmax={lmax}",
                   "nside": nside, "lmax": lmax, "spin": spin, "batch": B,
                   "precision": args.precision, "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 3),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None},
        "stages_ms": {k: round(v[0], 3) for k, v in per.items()},
        "adjoint_identity": {"lhs": lhs, "rhs": rhs, "rel": ident},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nside, lmax, spin, args.cpu_rows)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
