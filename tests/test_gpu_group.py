"""Ragged groups: several mixed-precision instances in one kernel launch per stage (GPU).

* Against each instance's own plan running the same FP32 matrix-core recursion kernels
  (tuning mx_f16 = mx_adj_f16 = 2, 16 maps per recursion): the recursions are the same
  code on the same chunks, the ring FFT differs only in its work-group shape (the fold's
  FP64 partial sums are grouped by the work-group width), so maps and a_lm agree to FP32
  rounding; and against the FP64 path at <= 1e-5.  Spin 0 and spin 2, ragged batches
  (1, 3, 5, 17 maps: partial 16-map groups), nside 16..128, both azimuth modes.
* The BASELINE C4 set -- 64 spin-0 maps at nside 512 / lmax 1024 plus 6 + 5 + 5 maps at
  nside 128 / 256 / 512 (lmax = 2 nside) -- in ONE captured graph of five kernel launches:
  three replays with fresh inputs and eager work in between equal the eager group calls
  bit for bit, and sampled maps of every instance match the unbatched FP64 path <= 1e-5.
* One launch per stage, counted by rocprofv3's kernel trace of a group synthesis +
  adjoint (tests/_group_trace.py): exactly one dispatch of each of the five group
  kernels and no one-plan transform kernel.
"""

from __future__ import annotations

import csv
import glob
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FP32_TUNING = dict(mx_f16=2, mx_adj_f16=2, mx_synth_group=16, mx_adj_group=16, mx_adj_cb=1, mx_single=2)


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def _alm(lmax, n, seed, spin):
    rng = np.random.default_rng(seed)
    if spin == 0:
        return torch.tensor(np.stack([O.gaussian_alm(lmax, rng) for _ in range(n)]), device="cuda")
    return torch.tensor(np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                                  for _ in range(n)]), device="cuda")


def _buffers(inst):
    from jht_amd.healpix import alm_size

    alms, maps, back = [], [], []
    for nside, lmax, spin, B in inst:
        K, npix = alm_size(lmax), 12 * nside * nside
        ash = (B, K) if spin == 0 else (B, 2, K)
        msh = (B, npix) if spin == 0 else (B, 2, npix)
        maps.append(torch.zeros(msh, dtype=torch.float64, device="cuda"))
        back.append(torch.zeros(ash, dtype=torch.complex128, device="cuda"))
    return maps, back


@pytest.mark.parametrize("mode", ["unrolled", "looped"])
@pytest.mark.parametrize("spin", [0, 2])
def test_group_matches_plans_and_fp64(spin, mode):
    import jht_amd
    from jht_amd import _lib
    from jht_amd.graph import RaggedGroup
    from jht_amd.healpix import Plan

    inst = [(16, 32, spin, 3), (64, 128, spin, 17), (32, 80, spin, 5), (128, 200, spin, 1)]
    jht_amd.set_azimuth_fft_mode(mode)
    try:
        g = RaggedGroup([(n, l, s, B, "mixed") for n, l, s, B in inst])
        assert g.query(2) == 2 and g.query(3) == 3
        alms = [_alm(l, B, 40 + k, s) for k, (n, l, s, B) in enumerate(inst)]
        maps, back = _buffers(inst)
        g.synth(alms, maps)
        g.adjoint(maps, back)
        torch.cuda.synchronize()
        flags = _lib.FLAG_BLUESTEIN_ALL if mode == "looped" else 0
        for k, (nside, lmax, s, B) in enumerate(inst):
            p = Plan(nside, lmax, s, B, _lib.PREC_MIXED, flags, tuning=FP32_TUNING)
            m_ref = torch.empty_like(maps[k])
            b_ref = torch.empty_like(back[k])
            p.synth(alms[k], m_ref, B)
            p.adjoint(maps[k], b_ref, B)
            torch.cuda.synchronize()
            es, ea = _rel(maps[k], m_ref), _rel(back[k], b_ref)
            worst_s = worst_a = 0.0
            for i in {0, B - 1}:
                s64 = jht_amd.synthesis(alms[k][i], nside, lmax, s)
                worst_s = max(worst_s, _rel(maps[k][i], s64))
                b64 = jht_amd.adjoint_synthesis(maps[k][i], nside, lmax, s)
                worst_a = max(worst_a, _rel(back[k][i], b64))
            print(f"spin {s} {mode} nside {nside} lmax {lmax} batch {B}: vs plan synth {es:.1e} adj {ea:.1e}; "
                  f"vs FP64 synth {worst_s:.1e} adj {worst_a:.1e}", flush=True)
            assert es <= 2e-6 and ea <= 2e-6
            assert worst_s <= 1e-5 and worst_a <= 1e-5
    finally:
        jht_amd.set_azimuth_fft_mode("unrolled")


def test_group_c4_graph_replays_bitwise_and_fp64():
    import jht_amd
    from jht_amd.graph import RaggedGroup, TransformGraph

    inst = [(512, 1024, 0, 64), (128, 256, 0, 6), (256, 512, 0, 5), (512, 1024, 0, 5)]
    tg = TransformGraph([(n, l, s, B, "mixed") for n, l, s, B in inst], fused=True)
    assert isinstance(tg.group, RaggedGroup)
    eager = RaggedGroup([(n, l, s, B, "mixed") for n, l, s, B in inst])
    for rep in range(3):
        alms = [_alm(l, B, 300 + 10 * rep + k, s) for k, (n, l, s, B) in enumerate(inst)]
        for k, a in enumerate(alms):
            tg.alm[k].copy_(a)
        tg.replay()
        # eager work on the same stream between replays
        maps, back = _buffers(inst)
        eager.synth(alms, maps)
        eager.adjoint(maps, back)
        torch.cuda.synchronize()
        for k in range(len(inst)):
            assert torch.equal(tg.maps[k], maps[k]), (rep, k)
            assert torch.equal(tg.back[k], back[k]), (rep, k)
        print(f"C4 replay {rep}: bitwise equal to the eager group calls", flush=True)
    for k, (nside, lmax, s, B) in enumerate(inst):
        worst_s = worst_a = 0.0
        for i in sorted({0, B // 2, B - 1}):
            s64 = jht_amd.synthesis(alms[k][i], nside, lmax, s)
            worst_s = max(worst_s, _rel(tg.maps[k][i], s64))
            b64 = jht_amd.adjoint_synthesis(tg.maps[k][i], nside, lmax, s)
            worst_a = max(worst_a, _rel(tg.back[k][i], b64))
        print(f"C4 nside {nside} batch {B}: vs FP64 synth {worst_s:.2e} adj {worst_a:.2e}", flush=True)
        assert worst_s <= 1e-5 and worst_a <= 1e-5


def test_group_rejects_bad_calls():
    from jht_amd import _lib
    from jht_amd.graph import RaggedGroup

    g = RaggedGroup([(16, 32, 0, 2, "mixed"), (32, 64, 0, 1, "mixed")])
    maps, back = _buffers([(16, 32, 0, 2), (32, 64, 0, 1)])
    with pytest.raises(ValueError):
        g.synth([torch.zeros((2, 10), dtype=torch.complex128, device="cuda")] * 2, maps)
    lib = _lib.load()
    vp = _lib.ctypes.c_void_p
    # mixed spins and an FP64 plan are refused by the C ABI too
    from jht_amd.healpix import Plan

    p0 = Plan(16, 32, 0, 1, _lib.PREC_MIXED, workspace=False)
    p2 = Plan(16, 32, 2, 1, _lib.PREC_MIXED, workspace=False)
    p64 = Plan(16, 32, 0, 1, _lib.PREC_F64, workspace=False)
    big = Plan(1024, 64, 0, 1, _lib.PREC_MIXED, workspace=False)
    for plans, code in (([p0, p2], _lib.JHTX_EINVAL), ([p0, p64], _lib.JHTX_EINVAL),
                        ([big], _lib.JHTX_EUNSUPPORTED)):
        h = vp()
        arr = (vp * len(plans))(*[p._h.value for p in plans])
        bat = (_lib.ctypes.c_int * len(plans))(*([1] * len(plans)))
        assert lib.jhtx_group_create(arr, bat, len(plans), _lib.ctypes.byref(h)) == code
        assert not h.value
    h = vp()
    arr = (vp * 1)(p0._h.value)
    assert lib.jhtx_group_create(arr, (_lib.ctypes.c_int * 1)(2), 1, _lib.ctypes.byref(h)) == _lib.JHTX_EINVAL


@pytest.mark.skipif(shutil.which("rocprofv3") is None, reason="rocprofv3 not on PATH")
def test_group_one_launch_per_stage(tmp_path):
    out = tmp_path / "trace"
    env = dict(os.environ, TMPDIR=str(tmp_path))
    cmd = ["rocprofv3", "--kernel-trace", "--output-format", "csv", "-d", str(out), "-o", "run", "--",
           sys.executable, os.path.join(ROOT, "tests", "_group_trace.py")]
    print("running", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=150)
    assert "group trace done" in res.stdout, res.stdout[-2000:] + res.stderr[-2000:]
    files = glob.glob(str(out / "**" / "*kernel_trace.csv"), recursive=True)
    assert files, res.stdout[-2000:] + res.stderr[-2000:]
    names = []
    for f in files:
        with open(f) as fh:
            names += [row["Kernel_Name"] for row in csv.DictReader(fh)]
    group = {}
    for n in names:
        for k in ("k_group_synth", "k_group_ring_synth", "k_group_ring_adjoint", "k_group_adj", "k_group_finish"):
            if k + "<" in n or k + "(" in n:
                group[k] = group.get(k, 0) + 1
    print("group dispatches:", group, flush=True)
    assert group == {"k_group_synth": 1, "k_group_ring_synth": 1, "k_group_ring_adjoint": 1,
                     "k_group_adj": 1, "k_group_finish": 1}
    one_plan = [n for n in names if any(k in n for k in ("k_synth_", "k_adj_mx_", "k_ring_synth", "k_ring_adjoint",
                                                          "k_adj_finish"))]
    assert not one_plan, one_plan[:5]
