"""Benchmark: HEALPix spin-2 synthesis + adjoint on one MI355X.

Workload (the metric's configuration): spin 2, nside = 2048, lmax = 4096, E/B a_lm with
C_l^EE = (l+1)^-2, C_l^BB = 0.1 C_l^EE, l < 2 zero (synthetic, seeded).  One step is one
synthesis S a -> (Q, U) followed by one adjoint S^T (Q, U) -> a' for every map of the
batch; inputs are resident in HBM before the timed region.  `value` is maps/s of the
selected precision (default: mixed, batch 32); the FP64 path (batch 1) is timed beside it
and the mixed outputs of the first map are compared with the FP64 GPU outputs
(max|x - x64| / max|x64|).  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision f64|mixed] [--batch B]
                    [--nside N --lmax L --spin S] [--no-cpu] [--no-f64]
                    [--offgrid NPTS | --c4 | --peaks | --cpu-c2]

Roofline: the dominant recursion + contraction kernel (synthesis or adjoint; the adjoint's
chunk reduction is the separate stage adj_finish) is timed live with HIP events recorded by
the library on the call's stream around that kernel alone.  Work unit: one (l, m,
north-half ring) triple of the active triangle l >= max(m, |s|).  FP64 path: 26 flops per
triple per map (spin 2; 8 for spin 0) against the 78.6 TFLOP/s FP64 peak.  Mixed path: the
contraction, 16 flops per triple per map (spin 2; 4 for spin 0), against the 157.3 TFLOP/s
dense FP32 MFMA peak; the 16-map spin-2 kernels run it on the FP16 matrix cores as four FP16
products per FP32-equivalent product, against a quarter of the 2500 TFLOP/s dense FP16 peak; a
one-map mixed call runs the FP64 kernels and is priced as the FP64 path.  CPU baseline: the reference's
masked l-scans restated in numpy (oracle/jht_oracle.py), the synthesis scan and then the
adjoint scan, on a sample of orders m over every l and ring spread over every host core
(at most 16), scaled by (lmax+1)/sample, plus the full-size numpy ring FFTs; --cpu-c2 times
the spin-0 nside 1024 / lmax 2048 pair exactly (every order).
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
# HBM bytes per call of the recursion kernels, measured by rocprofv3 FETCH_SIZE and WRITE_SIZE
# runs of this script (each counter in its own run, scaled by the counters of a 1 GiB device
# copy in the same run: FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); keyed by (precision, nside, lmax,
# spin, batch, stage).  Not measured live: the JSON names the profile the value comes from, and
# a kernel change needs a new profile.
TRAFFIC_PROFILE = "profiles/kernel_stats_c3_spin2_mixed_batch32.md"
MEASURED_TRAFFIC = {
    ("mixed", 2048, 4096, 2, 32, "adj_rec"): 107.34e9,
    ("mixed", 2048, 4096, 2, 32, "synth_rec"): 156.73e9,
}
PEAK_FP32_MFMA_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_FP16_MFMA_TFLOPS = 2500.0  # MI355X dense FP16 matrix
# FP16 products per FP32-equivalent product of the FP16-split kernels: the 16-map forms issue
# all four (hi hi, lo hi, hi lo, lo lo), the 32-map spin-2 forms leave lo lo out; their ceiling
# for the FP32-equivalent contraction flops is the FP16 peak divided by that count
F16_SPLIT_PRODUCTS = {16: 4, 32: 3}
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


def cpu_cores():
    """Host threads the CPU baseline uses: the process's CPU set, at most 16 (one GPU's share
    of a node)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _cpu_rows(job):
    """One worker's share of the sampled scans: the synthesis scan and then the adjoint scan
    (the reference runs them as two jitted programs) over its orders m; returns seconds."""
    nside, lmax, spin, ms, seed = job
    from oracle import jht_oracle as O

    rng = np.random.default_rng(seed)
    x = O.RingInfo(nside).z[: 2 * nside]
    R, L1, T = len(ms), lmax + 1, 2 * nside

    def cplx(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    coef = cplx(R, L1) if spin == 0 else (cplx(R, L1), cplx(R, L1))
    V = [cplx(R, T) for _ in range(2 if spin == 0 else 4)]
    t0 = time.perf_counter()
    O.sampled_synthesis_scan(x, lmax, ms, spin, coef)
    O.sampled_adjoint_scan(x, lmax, ms, spin, V)
    return time.perf_counter() - t0


def _cpu_warm(_):
    return os.getpid()


def cpu_baseline(nside, lmax, spin, rows_per_core=2, all_orders=False, seed=1):
    """Time the reference algorithm (the oracle's numpy restatement of jht's masked l-scans)
    on every host core (a fork pool, started before this process touches the GPU).

    Orders m are dealt round-robin to the workers; each runs the synthesis scan and then the
    adjoint scan over its orders, every l and every north-half ring, masked as jht's lax.scan
    runs them (so every order costs the same).  The slowest worker's time, scaled by
    (lmax + 1) / orders timed, is the recursion time per map; the full-size numpy ring FFTs
    of both directions are timed in this process and added.  all_orders: time every order
    (no scaling)."""
    import multiprocessing

    from oracle import jht_oracle as O

    P = cpu_cores()
    M = lmax + 1
    ms = np.arange(M) if all_orders else np.unique(np.linspace(0, lmax, P * rows_per_core).astype(np.int64))
    chunks = [ms[k::P] for k in range(P) if len(ms[k::P])]
    with multiprocessing.get_context("fork").Pool(len(chunks)) as pool:
        pool.map(_cpu_warm, range(len(chunks)))
        times = pool.map(_cpu_rows, [(nside, lmax, spin, c, seed + k) for k, c in enumerate(chunks)])
    t_rec = max(times) * (M / ms.size)
    rng = np.random.default_rng(seed)
    geo = O.RingInfo(nside)
    G = rng.standard_normal((M, geo.nrings)) + 1j * rng.standard_normal((M, geo.nrings))
    t0 = time.perf_counter()
    out = O.ring_synthesis(G, G, geo)
    if spin == 0:
        O.ring_analysis(out.real.astype(complex), geo, lmax)
    else:
        O.ring_analysis(out, geo, lmax)
        O.ring_analysis(out, geo, lmax, minus=True)
    t_fft = time.perf_counter() - t0
    del G, out
    total = t_rec + t_fft
    what = "every" if all_orders else f"{ms.size} of {M}"
    return {
        "value": 1.0 / total,
        "unit": "maps/s",
        "cores": len(chunks),
        "kind": "port",
        "sample": (f"numpy restatement of jht's masked l-scans (spin {spin}: synthesis scan then adjoint "
                   f"scan{', both +-2 recursions in each' if spin else ''}) for {what} orders m over all "
                   f"{M} l and {2 * nside} rings, {len(chunks)} processes, slowest {max(times):.1f} s"
                   + ("" if all_orders else f", scaled x{M / ms.size:.1f}") +
                   f" ({t_rec:.1f} s/map); full-size numpy ring FFTs both ways on one core ({t_fft:.1f} s/map)"),
    }


def run(plan, alm, maps, back, B, steps, warmup):
    import torch

    def step():
        plan.synth(alm, maps, B)
        plan.adjoint(maps, back, B)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    plan.set_profiling(True)
    plan.read_profile()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    prof = plan.read_profile()
    plan.set_profiling(False)
    per = {k: (v[0] / max(v[1], 1), v[1]) for k, v in prof.items()}
    return dt, per


def bench_offgrid(args):
    """Off-grid pair (synthesis_general + adjoint_synthesis_general, FP64, epsilon 1e-10) on
    random points: points/s of the pair and the per-stage split (HIP events)."""
    import torch

    import jht_amd
    from jht_amd.offgrid import get_offgrid_plan

    lmax, spin, npts = args.lmax, args.spin, args.offgrid
    rng = np.random.default_rng(20261204)
    if spin == 0:
        a = gaussian_alm(lmax, rng)
    else:
        a = np.stack([gaussian_alm(lmax, rng, 2), gaussian_alm(lmax, rng, 2, 0.1)])
    alm = torch.tensor(a, device="cuda")
    loc = torch.tensor(np.stack([rng.uniform(0, np.pi, npts), rng.uniform(0, 2 * np.pi, npts)], axis=1), device="cuda")

    def step():
        f = jht_amd.synthesis_general(alm, loc, spin=spin, lmax=lmax)
        return jht_amd.adjoint_synthesis_general(f, loc, spin=spin, lmax=lmax)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    plan = get_offgrid_plan(lmax, spin, 1e-10)
    print(json.dumps({
        "metric": f"off-grid synth+adjoint points/s lmax={lmax} (spin {spin}), epsilon 1e-10",
        "value": round(npts / dt, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (seeded Gaussian a_lm, uniform random points)",
        "config": {"workload": "off-grid synthesis+adjoint", "lmax": lmax, "spin": spin, "npts": npts,
                   "rings": plan.ns, "grid": [plan.nk, plan.nm], "W": plan.W},
    }), flush=True)


def bench_peaks(args):
    """Measured device rates beside the published peaks the roofline uses: FP32 matrix
    (v_mfma_f32_16x16x4f32 loops), FP64 vector FMA and device-to-device copy (jhtx_probe_peaks);
    best of --steps runs."""
    import ctypes

    import torch

    from jht_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    buf = torch.empty(1 << 29, dtype=torch.uint8, device=dev)
    best = [0.0, 0.0, 0.0]
    for _ in range(max(1, args.steps)):
        out = (ctypes.c_double * 3)()
        _lib.check(lib.jhtx_probe_peaks(ctypes.c_void_p(buf.data_ptr()), buf.numel(), out,
                                        ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
        best = [max(b, o) for b, o in zip(best, out)]
    print(json.dumps({
        "metric": "measured device rates (best of runs) against the published MI355X peaks",
        "steps": args.steps,
        "fp32_mfma_tflops": {"measured": round(best[0], 2), "published": PEAK_FP32_MFMA_TFLOPS,
                             "frac": round(best[0] / PEAK_FP32_MFMA_TFLOPS, 4)},
        "fp64_fma_tflops": {"measured": round(best[1], 2), "published": PEAK_FP64_TFLOPS,
                            "frac": round(best[1] / PEAK_FP64_TFLOPS, 4)},
        "hbm_copy_gbs": {"measured": round(best[2], 1), "published": HBM_PEAK_GBS,
                         "frac": round(best[2] / HBM_PEAK_GBS, 4)},
    }), flush=True)


C4_INSTANCES = [(512, 1024, 0, 64), (128, 256, 0, 6), (256, 512, 0, 5), (512, 1024, 0, 5)]


def bench_c4(args):
    """Ragged batched workload: 64 spin-0 maps at nside 512 / lmax 1024 plus 16 spin-0 maps
    over nside 128 / 256 / 512 (lmax = 2 nside), mixed precision, synthesis + adjoint of
    every instance captured in one graph and replayed.  The graph holds the ragged group's
    five launches (one per stage over all instances; `value`), and for comparison a graph of
    each instance's own calls.  Reports maps/s of the replays, the unbatched FP64 path on the
    same maps (one map per call per geometry, weighted by the instance batches), and the
    largest error of the first map of each instance against FP64: the synthesis, and the
    adjoint of the same mixed map."""
    import torch

    from jht_amd import _lib
    from jht_amd.graph import TransformGraph
    from jht_amd.healpix import get_plan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    inst = [(ns, lm, sp, b, "mixed") for ns, lm, sp, b in C4_INSTANCES]
    rng = np.random.default_rng(20261205)
    alms = [torch.tensor(np.stack([gaussian_alm(lm, rng) for _ in range(b)]), device=dev)
            for ns, lm, sp, b in C4_INSTANCES]

    def replay_time(graph):
        for i, a in enumerate(alms):
            graph.alm[i].copy_(a)
        for _ in range(args.warmup):
            graph.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    per_inst = TransformGraph(inst, device=dev)
    dt_inst = replay_time(per_inst)
    del per_inst
    g = TransformGraph(inst, device=dev, fused=True)
    dt = replay_time(g)
    nmaps = sum(b for *_, b in C4_INSTANCES)
    t64, errs = 0.0, {}
    for i, (ns, lm, sp, b) in enumerate(C4_INSTANCES):
        p64 = get_plan(ns, lm, sp, 1, precision=_lib.PREC_F64)
        a0 = g.alm[i][:1].contiguous()
        m64 = torch.empty((1, 12 * ns * ns), dtype=torch.float64, device=dev)
        b64 = torch.empty_like(a0)
        d, _ = run(p64, a0, m64, b64, 1, max(2, args.steps // 2), 1)
        t64 += b * d / max(2, args.steps // 2)
        # FP64 adjoint of the mixed map, against the graph's mixed adjoint of it
        p64.adjoint(g.maps[i][:1].contiguous(), b64, 1)
        torch.cuda.synchronize()
        key = f"{ns}/{lm} x{b}"
        errs[key] = {"synthesis": float((g.maps[i][0] - m64[0]).abs().max() / m64[0].abs().max()),
                     "adjoint": float((g.back[i][0] - b64[0]).abs().max() / b64[0].abs().max())}
    print(json.dumps({
        "metric": "ragged batch maps/s: 64 spin-0 maps at nside 512 / lmax 1024 + 16 over nside 128/256/512, "
                  "mixed, synth+adjoint in one graph of one launch per stage",
        "value": round(nmaps / dt, 2), "unit": "maps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "dtype": "f64 recursion / f32 mfma contraction", "data": "synthetic (seeded Gaussian a_lm, C_l=(l+1)^-2)",
        "config": {"workload": "ragged spin-0 batch (C4)", "instances": [list(x) for x in C4_INSTANCES],
                   "launches": "ragged group: 5 kernel launches per replay (2 synthesis, 3 adjoint)"},
        "per_instance_graph_maps_per_s": round(nmaps / dt_inst, 2),
        "f64_unbatched_maps_per_s": round(nmaps / t64, 2), "speedup_vs_f64": round(t64 / dt, 3),
        "rel_err_vs_f64": errs,
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="mixed", choices=["f64", "mixed"])
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--nside", type=int, default=2048)
    ap.add_argument("--lmax", type=int, default=4096)
    ap.add_argument("--spin", type=int, default=2)
    ap.add_argument("--tuning", default="",
                    help="kernel shapes for the mixed plan, jhtx_tuning fields as key=value,... "
                         "(e.g. mx_synth_fold=1,mx_synth_group=32)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=2, help="orders m per host core in the CPU baseline sample")
    ap.add_argument("--cpu-c2", action="store_true",
                    help="instead of the headline transform pair, time the CPU baseline exactly on the "
                         "spin-0 nside 1024 / lmax 2048 pair (every order m, all host cores) and print "
                         "its own JSON line; no GPU work")
    ap.add_argument("--offgrid", type=int, default=0, metavar="NPTS",
                    help="instead of the headline transform pair, time the off-grid pair "
                         "synthesis_general + adjoint_synthesis_general at --lmax / --spin on "
                         "NPTS uniformly random points and print its own JSON line")
    ap.add_argument("--c4", action="store_true",
                    help="instead of the headline transform pair, time the ragged batch (64 spin-0 "
                         "maps at nside 512 / lmax 1024 + 16 over nside 128/256/512) captured in "
                         "one graph, against the unbatched FP64 path, and print its own JSON line")
    ap.add_argument("--peaks", action="store_true",
                    help="instead of the headline transform pair, measure the device's FP32 matrix, "
                         "FP64 FMA and copy rates (the roofline denominators) and print them beside "
                         "the published peaks")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one plain device copy of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known traffic)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus N > 1 is not supported "
              "(maps are independent: several GPUs would each take a share of the batch)",
              file=sys.stderr)
        sys.exit(2)

    if args.cpu_c2:
        r = cpu_baseline(1024, 2048, 0, all_orders=True)
        print(json.dumps({"metric": "CPU baseline maps/s synth+adjoint nside=1024 lmax=2048 (spin 0), every order m",
                          **r}), flush=True)
        return
    headline = not (args.offgrid > 0 or args.c4 or args.peaks)
    # the CPU baseline forks its workers before this process touches the GPU
    cpu = cpu_baseline(args.nside, args.lmax, args.spin, args.cpu_rows) if headline and not args.no_cpu else None

    import torch

    import jht_amd
    from jht_amd import _lib
    from jht_amd.healpix import get_plan

    if args.copy_probe_mb > 0:
        src = torch.ones(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.offgrid > 0:
        return bench_offgrid(args)
    if args.c4:
        return bench_c4(args)
    if args.peaks:
        return bench_peaks(args)

    nside, lmax, spin, B = args.nside, args.lmax, args.spin, args.batch
    prec = _lib.PREC_MIXED if args.precision == "mixed" else _lib.PREC_F64
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(20261203)
    if spin == 0:
        a = np.stack([gaussian_alm(lmax, rng) for _ in range(B)])
    else:
        a = np.stack([np.stack([gaussian_alm(lmax, rng, 2), gaussian_alm(lmax, rng, 2, 0.1)]) for _ in range(B)])
    alm = torch.tensor(a, device=dev)
    del a
    npix = 12 * nside * nside
    mshape = (B, npix) if spin == 0 else (B, 2, npix)
    maps = torch.empty(mshape, dtype=torch.float64, device=dev)
    back = torch.empty_like(alm)
    torch.cuda.synchronize()
    t_plan = time.perf_counter()
    tuning = {k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(",") if kv)} or None
    plan = get_plan(nside, lmax, spin, B, precision=prec, tuning=tuning)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t_plan
    dt, per = run(plan, alm, maps, back, B, args.steps, args.warmup)
    value = B * args.steps / dt

    out_extra = {}
    if not args.no_f64 and prec != _lib.PREC_F64:
        # FP64 reference path on the first map: timing and the mixed-vs-FP64 error
        p64 = get_plan(nside, lmax, spin, 1, precision=_lib.PREC_F64)
        m64 = torch.empty((1,) + mshape[1:], dtype=torch.float64, device=dev)
        b64 = torch.empty_like(alm[:1])
        a0 = alm[:1].contiguous()
        dt64, per64 = run(p64, a0, m64, b64, 1, max(2, args.steps // 2), 1)
        # same input map for both adjoints: the FP64 synthesis output
        bmx = torch.empty_like(alm)
        mx_in = maps.clone()
        mx_in[0] = m64[0]
        plan.adjoint(mx_in, bmx, B)
        torch.cuda.synchronize()
        e_s = float((maps[0] - m64[0]).abs().max() / m64[0].abs().max())
        e_a = float((bmx[0] - b64[0]).abs().max() / b64[0].abs().max())
        # adjoint error per l-band: max|b - b64| / max|b64| over the band's coefficients
        ell = torch.as_tensor(np.concatenate([np.arange(mm, lmax + 1) for mm in range(lmax + 1)]), device=dev)
        edges = [e for e in (0, 64, 256, 1024, 2048, 4096) if e <= lmax] + [lmax + 1]
        bands = {}
        for lo, hi in zip(edges[:-1], edges[1:]):
            sel = (ell >= lo) & (ell < hi)
            ref = b64[0][..., sel].abs().max()
            if float(ref) > 0:
                bands[f"{lo}-{hi - 1}"] = float((bmx[0][..., sel] - b64[0][..., sel]).abs().max() / ref)
        v64 = 1.0 / dt64 * max(2, args.steps // 2)
        out_extra = {"f64_maps_per_s": round(v64, 4), "speedup_vs_f64": round(value / v64, 3),
                     "rel_err_vs_f64": {"synthesis": e_s, "adjoint": e_a, "adjoint_by_l_band": bands},
                     "f64_stages_ms": {k: round(v[0], 3) for k, v in per64.items()}}
        del p64, m64, b64, bmx, mx_in

    # size-independent check: <S a, v> == <a, S^T v>_G with v = S a
    G = torch.as_tensor(jht_amd.alm_metric_weight(lmax).copy(), device=dev)
    lhs = float((maps[0] * maps[0]).sum())
    rhs = float((G * (alm[0].conj() * back[0]).real).sum())
    ident = abs(lhs - rhs) / abs(lhs)

    ms_step = dt / args.steps * 1e3
    dom = max(("synth_rec", "adj_rec"), key=lambda k: per[k][0])
    tri = triples(nside, lmax, spin)
    # one-map calls of a mixed plan run the FP64 vector-FMA recursion kernels (mx_single)
    fp64_kernels = prec == _lib.PREC_F64 or (B == 1 and plan.tuning["mx_single"] == 1)
    tun = plan.tuning
    # the 16- and 32-map kernels on the FP16 matrix cores (hi + lo split)
    tile16 = prec != _lib.PREC_F64 and B >= 16
    f16 = {"synth_rec": tile16 and tun["mx_f16"] == 1,
           "adj_rec": tile16 and tun["mx_adj_f16"] == 1 and tun["mx_adj_cb"] == 1}
    peak_basis = None
    if fp64_kernels:
        flops, peak, bound = tri * (26 if spin else 8) * B, PEAK_FP64_TFLOPS, "mfma"
    elif f16[dom]:
        group = tun["mx_adj_group"] if dom == "adj_rec" else tun["mx_synth_group"]
        nprod = F16_SPLIT_PRODUCTS[32 if (spin == 2 and group >= 32) else 16]
        flops, peak, bound = tri * 16 * B, round(PEAK_FP16_MFMA_TFLOPS / nprod, 1), "mfma"
        peak_basis = f"fp16 dense 2500 TFLOP/s / {nprod} fp16 products per fp32-equivalent product"
    else:
        flops, peak, bound = tri * (16 if spin else 4) * B, PEAK_FP32_MFMA_TFLOPS, "mfma"
    achieved = flops / (per[dom][0] * 1e-3) / 1e12
    out = {
        "metric": f"maps/s synth+adjoint nside={nside} lmax={lmax} (spin {spin}): mixed vs FP64 GPU at rel err <= 1e-5",
        "value": round(value, 4),
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if prec == _lib.PREC_F64 else
                  "f64 recursion and contraction / f32 ring fft" if fp64_kernels else
                  "f64 recursion / fp32 contraction as fp16 hi+lo split mfma" if f16["synth_rec"] and f16["adj_rec"]
                  else "f64 recursion / f32 mfma contraction"),
        "data": "synthetic (seeded Gaussian E/B a_lm, C_l=(l+1)^-2, B=0.1 E, l<2 zero)",
        "config": {"workload": f"spin-{spin} synthesis+adjoint, nside={nside}, lmax={lmax}",
                   "nside": nside, "lmax": lmax, "spin": spin, "batch": B,
                   "precision": args.precision, "parallelism": "single-gpu"},
        "roofline": {"bound": bound, "kernel": dom, "achieved": round(achieved, 3),
                     "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), **({"peak_basis": peak_basis} if peak_basis else {}),
                     "traffic": MEASURED_TRAFFIC.get((args.precision, nside, lmax, spin, B, dom)),
                     **({"traffic_source": TRAFFIC_PROFILE}
                        if (args.precision, nside, lmax, spin, B, dom) in MEASURED_TRAFFIC else {})},
        "stages_ms": {k: round(v[0], 3) for k, v in per.items()},
        "adjoint_identity": {"lhs": lhs, "rhs": rhs, "rel": ident},
        "plan_create_s": round(t_plan, 3),
    }
    out.update(out_extra)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
