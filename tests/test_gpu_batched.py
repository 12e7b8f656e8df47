"""Batched, ragged and graph-captured transforms, and the autotuner (GPU).

* 64 spin-0 maps at nside 512 / lmax 1024 (FP64) synthesised and adjoined by one captured
  graph, replayed three times with fresh inputs and eager work in between; every replay
  equals the eager batched call bit for bit, and single-map calls too.
* A ragged mixed-precision batch of 16 spin-0 maps over nside 128 / 256 / 512 with
  lmax = 2 nside in one graph, against the unbatched FP64 path (<= 1e-5).
* The autotuner times every candidate shape and its choice gives the same results.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def _gauss_batch(lmax, n, seed, spin=0):
    rng = np.random.default_rng(seed)
    if spin == 0:
        return torch.tensor(np.stack([O.gaussian_alm(lmax, rng) for _ in range(n)]), device="cuda")
    return torch.tensor(np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                                  for _ in range(n)]), device="cuda")


def test_graph_64_maps_fp64_replays_bitwise():
    import jht_amd
    from jht_amd.graph import TransformGraph

    nside, lmax, B = 512, 1024, 64
    g = TransformGraph([(nside, lmax, 0, B, "f64")])
    for rep in range(3):
        a = _gauss_batch(lmax, B, 100 + rep)
        g.alm[0].copy_(a)
        g.replay()
        # eager work between replays, on the same stream
        m_eager = jht_amd.synthesis_batch(a, nside, lmax, 0)
        b_eager = jht_amd.adjoint_synthesis_batch(m_eager, nside, lmax, 0)
        torch.cuda.synchronize()
        assert torch.equal(g.maps[0], m_eager), rep
        assert torch.equal(g.back[0], b_eager), rep
        for i in (0, 31, 63):
            assert torch.equal(g.maps[0][i], jht_amd.synthesis(a[i], nside, lmax, 0)), (rep, i)
        print(f"replay {rep}: bitwise equal to eager", flush=True)


def test_graph_ragged_mixed_batch():
    import jht_amd
    from jht_amd.graph import TransformGraph

    inst = [(128, 256, 0, 6, "mixed"), (256, 512, 0, 5, "mixed"), (512, 1024, 0, 5, "mixed")]
    g = TransformGraph(inst)
    inputs = [_gauss_batch(lmax, B, 7 + k) for k, (_, lmax, _, B, _) in enumerate(inst)]
    for k, a in enumerate(inputs):
        g.alm[k].copy_(a)
    for rep in range(2):
        g.replay()
        torch.cuda.synchronize()
        for k, (nside, lmax, _, B, _) in enumerate(inst):
            worst_s = worst_a = 0.0
            for i in range(B):
                s64 = jht_amd.synthesis(inputs[k][i], nside, lmax, 0)
                worst_s = max(worst_s, _rel(g.maps[k][i], s64))
                b64 = jht_amd.adjoint_synthesis(g.maps[k][i], nside, lmax, 0)
                worst_a = max(worst_a, _rel(g.back[k][i], b64))
            print(f"ragged rep {rep} nside {nside}: synth {worst_s:.2e} adj {worst_a:.2e}", flush=True)
            assert worst_s <= 1e-5 and worst_a <= 1e-5


@pytest.mark.parametrize("precision,spin,B", [("f64", 0, 1), ("f64", 2, 1), ("mixed", 0, 4), ("mixed", 2, 4),
                                              ("mixed", 2, 1)])
def test_autotune_choice_is_exact(precision, spin, B, tmp_path):
    import jht_amd
    from jht_amd import healpix as H
    from jht_amd.tune import autotune

    nside, lmax = 128, 256
    cache = str(tmp_path / "tune.json")
    H.clear_plans()
    try:
        a = _gauss_batch(lmax, B, 3, spin)
        ref_m = jht_amd.synthesis_batch(a, nside, lmax, spin, precision="f64")
        ref_b = jht_amd.adjoint_synthesis_batch(ref_m, nside, lmax, spin, precision="f64")
        res = autotune(nside, lmax, spin, B, precision, steps=1, cache=cache)
        print(f"autotune {precision} spin {spin}: {res['tuning']}", flush=True)
        assert res["tuning"] and all(v > 0 for v in res["synth_ms"].values())
        plan = H.get_plan(nside, lmax, spin, B, H._prec(precision))
        got = plan.tuning
        for k, v in res["tuning"].items():
            assert got[k] == v, (k, got[k], v)
        m = jht_amd.synthesis_batch(a, nside, lmax, spin, precision=precision)
        b = jht_amd.adjoint_synthesis_batch(ref_m, nside, lmax, spin, precision=precision)
        tol = 1e-12 if precision == "f64" else 1e-5
        assert _rel(m, ref_m) <= tol and _rel(b, ref_b) <= tol
        again = autotune(nside, lmax, spin, B, precision, cache=cache)  # served from the cache
        assert again["tuning"] == res["tuning"]
    finally:
        H.set_tuning(nside, lmax, spin, H._prec(precision), None, B)
        H.clear_plans()
