"""GPU parity: the HIP operators (through the C ABI) against the float64 oracle.

Gates mirror the reference's: operator parity <= 1e-12 relative (max|x - ref| / max|ref|,
tests/test_vectorized.py:32, tests/test_gpu.py:37), the adjoint inner-product identity
<= 1e-12 (tests/test_grad.py:45,105-131), looped (chirp-z) == direct at atol 1e-12
(tests/test_azimuth_fft.py:28), and input validation (tests/test_healpix.py:138-156).
"""

from __future__ import annotations

import contextlib
import warnings

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _jht():
    import jht_amd

    return jht_amd


@contextlib.contextmanager
def _tuned(nside, lmax, spin, precision, batch=None, **fields):
    """Plans of this geometry and precision use the given jhtx_tuning fields inside the
    block (set_tuning: the library's only kernel switch)."""
    from jht_amd import healpix as H

    prec = H._prec(precision)
    H.clear_plans()
    H.set_tuning(nside, lmax, spin, prec, fields, batch)
    try:
        yield
    finally:
        H.set_tuning(nside, lmax, spin, prec, None, batch)
        H.clear_plans()


def _alm(lmax, rng, spin):
    if spin == 0:
        return O.random_alm(lmax, rng)
    return np.stack([O.random_alm(lmax, rng, 2), O.random_alm(lmax, rng, 2)])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


CASES = [(1, 2), (2, 4), (4, 7), (8, 12), (8, 20), (16, 24), (16, 40), (32, 64), (64, 128), (48, 80)]


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("nside,lmax", CASES)
def test_synthesis_matches_oracle(nside, lmax, spin):
    j = _jht()
    rng = np.random.default_rng(100 * nside + lmax + spin)
    a = _alm(lmax, rng, spin)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = j.synthesis(torch.tensor(a, device="cuda"), nside, lmax, spin).cpu().numpy()
    ref = O.synthesis(a, nside, lmax, spin)
    print(f"synth nside={nside} lmax={lmax} spin={spin}: rel {_rel(out, ref):.2e}", flush=True)
    assert _rel(out, ref) <= TOL


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("nside,lmax", CASES)
def test_adjoint_matches_oracle(nside, lmax, spin):
    j = _jht()
    rng = np.random.default_rng(200 * nside + lmax + spin)
    npix = 12 * nside * nside
    v = rng.standard_normal(npix if spin == 0 else (2, npix))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = j.adjoint_synthesis(torch.tensor(v, device="cuda"), nside, lmax, spin).cpu().numpy()
    ref = O.adjoint_synthesis(v, nside, lmax, spin)
    print(f"adj nside={nside} lmax={lmax} spin={spin}: rel {_rel(out, ref):.2e}", flush=True)
    assert _rel(out, ref) <= TOL


@pytest.mark.parametrize("spin", [0, 2])
def test_adjoint_identity_gpu(spin):
    j = _jht()
    nside, lmax = 32, 48
    rng = np.random.default_rng(7 + spin)
    npix = 12 * nside * nside
    a = _alm(lmax, rng, spin)
    v = rng.standard_normal(npix if spin == 0 else (2, npix))
    Sa = j.synthesis(torch.tensor(a, device="cuda"), nside, lmax, spin).cpu().numpy()
    b = j.adjoint_synthesis(torch.tensor(v, device="cuda"), nside, lmax, spin).cpu().numpy()
    G = O.alm_metric_weight(lmax)
    lhs = float(np.sum(Sa * v))
    rhs = float(np.sum(G * np.real(np.conj(a) * b)))
    assert abs(lhs - rhs) / abs(lhs) <= TOL


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("nside,lmax", [(8, 12), (16, 24), (32, 48), (64, 128)])
def test_looped_matches_unrolled(nside, lmax, spin):
    j = _jht()
    rng = np.random.default_rng(300 + nside + spin)
    a = torch.tensor(_alm(lmax, rng, spin), device="cuda")
    npix = 12 * nside * nside
    v = torch.tensor(rng.standard_normal(npix if spin == 0 else (2, npix)), device="cuda")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        u_s = j.synthesis(a, nside, lmax, spin).cpu().numpy()
        u_a = j.adjoint_synthesis(v, nside, lmax, spin).cpu().numpy()
        j.set_azimuth_fft_mode("looped")
        try:
            l_s = j.synthesis(a, nside, lmax, spin).cpu().numpy()
            l_a = j.adjoint_synthesis(v, nside, lmax, spin).cpu().numpy()
        finally:
            j.set_azimuth_fft_mode("unrolled")
    assert np.allclose(l_s, u_s, rtol=0, atol=1e-12)
    assert np.allclose(l_a, u_a, rtol=0, atol=1e-12)


def test_batch_equals_single():
    j = _jht()
    nside, lmax = 16, 32
    rng = np.random.default_rng(11)
    a = np.stack([O.random_alm(lmax, rng) for _ in range(3)])
    out = j.synthesis_batch(torch.tensor(a, device="cuda"), nside, lmax, 0).cpu().numpy()
    for i in range(3):
        ref = j.synthesis(torch.tensor(a[i], device="cuda"), nside, lmax, 0).cpu().numpy()
        assert np.array_equal(out[i], ref)


def test_input_validation_gpu():
    j = _jht()
    nside, lmax = 16, 24
    K = O.alm_size(lmax)
    with pytest.raises(ValueError, match="alm shape"):
        j.synthesis(np.zeros(K - 1, dtype=np.complex128), nside, lmax, spin=0)
    with pytest.raises(ValueError, match="alm shape"):
        j.synthesis(np.zeros(K, dtype=np.complex128), nside, lmax, spin=2)
    with pytest.raises(ValueError, match="map shape"):
        j.adjoint_synthesis(np.zeros(12 * nside**2 - 1), nside, lmax, spin=0)
    with pytest.raises(ValueError, match="map shape"):
        j.adjoint_synthesis(np.zeros(12 * nside**2), nside, lmax, spin=2)
    with pytest.raises(NotImplementedError):
        j.synthesis(np.zeros((2, K), dtype=np.complex128), nside, lmax, spin=1)


@pytest.mark.parametrize("spin,synth_r,adj_r,block", [
    (0, 4, 4, 64), (0, 8, 8, 64), (0, 1, 1, 64), (2, 2, 2, 64), (2, 4, 4, 64), (2, 4, 4, 256), (0, 8, 8, 256)])
def test_lane_variants(spin, synth_r, adj_r, block):
    """Every rings-per-lane instantiation and multi-chunk ring reductions (forced through
    the plan's tuning) against the oracle."""
    j = _jht()

    nside, lmax = 128, 200
    with _tuned(nside, lmax, spin, "f64", synth_rings=synth_r, adj_rings=adj_r, block=block):
        rng = np.random.default_rng(17 + spin + synth_r)
        a = _alm(lmax, rng, spin)
        npix = 12 * nside * nside
        v = rng.standard_normal(npix if spin == 0 else (2, npix))
        out = j.synthesis(torch.tensor(a, device="cuda"), nside, lmax, spin).cpu().numpy()
        b = j.adjoint_synthesis(torch.tensor(v, device="cuda"), nside, lmax, spin).cpu().numpy()
    e1 = _rel(out, O.synthesis(a, nside, lmax, spin))
    e2 = _rel(b, O.adjoint_synthesis(v, nside, lmax, spin))
    print(f"variant spin={spin} R={synth_r}/{adj_r} block={block}: synth {e1:.2e} adj {e2:.2e}", flush=True)
    assert e1 <= TOL and e2 <= TOL


MIXED_TOL = 1e-5  # the mixed-precision contract (FP64 recursion, FP32 MFMA contraction)


@pytest.mark.parametrize("spin,batch", [(0, 5), (0, 8), (0, 9), (2, 3), (2, 4), (2, 6)])
@pytest.mark.parametrize("nside,lmax", [(16, 32), (64, 128), (128, 256)])
def test_mixed_matches_f64(nside, lmax, spin, batch):
    """Batched mixed precision against the FP64 path (and the FP64 path against the
    oracle for the first map), synthesis and adjoint."""
    j = _jht()
    rng = np.random.default_rng(1000 + nside + spin * 7 + batch)
    a = np.stack([O.gaussian_alm(lmax, rng) if spin == 0 else
                  np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                  for _ in range(batch)])
    at = torch.tensor(a, device="cuda")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        s64 = j.synthesis_batch(at, nside, lmax, spin, precision="f64")
        smx = j.synthesis_batch(at, nside, lmax, spin, precision="mixed")
        b64 = j.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="f64")
        bmx = j.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="mixed")
    s64, smx, b64, bmx = (t.cpu().numpy() for t in (s64, smx, b64, bmx))
    es = max(_rel(smx[i], s64[i]) for i in range(batch))
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch))
    print(f"mixed nside={nside} lmax={lmax} spin={spin} B={batch}: synth {es:.2e} adj {ea:.2e}", flush=True)
    assert es <= MIXED_TOL and ea <= MIXED_TOL
    if nside <= 64:
        assert _rel(s64[0], O.synthesis(a[0], nside, lmax, spin)) <= TOL


@pytest.mark.parametrize("single", [1, 0])
@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("nside,lmax", [(16, 32), (64, 128), (256, 400)])
def test_mixed_single_map(nside, lmax, spin, single):
    """One-map calls of a mixed plan: the FP64 vector-FMA recursion kernels on complex64
    ring coefficients with the FP32 ring FFT (single 1, the default) and the 4-map
    matrix-core kernels (single 0), against the FP64 path and, on the small grids, the
    oracle."""
    j = _jht()
    from jht_amd import healpix as H

    with _tuned(nside, lmax, spin, "mixed", mx_single=1 if single else 2):
        rng = np.random.default_rng(4100 + nside + spin)
        a = (O.gaussian_alm(lmax, rng) if spin == 0 else
             np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)]))
        at = torch.tensor(a[None], device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s64 = j.synthesis_batch(at, nside, lmax, spin, precision="f64")
            smx = j.synthesis_batch(at, nside, lmax, spin, precision="mixed")
            b64 = j.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="f64")
            bmx = j.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="mixed")
            plan = H.get_plan(nside, lmax, spin, 1, H._prec("mixed"))
            assert plan.tuning["mx_single"] == (1 if single else 2)
    s64, smx, b64, bmx = (t[0].cpu().numpy() for t in (s64, smx, b64, bmx))
    es, ea = _rel(smx, s64), _rel(bmx, b64)
    print(f"mixed one-map single={single} nside={nside} lmax={lmax} spin={spin}: synth {es:.2e} adj {ea:.2e}",
          flush=True)
    assert es <= MIXED_TOL and ea <= MIXED_TOL
    if nside <= 64:
        v = s64
        assert _rel(smx, O.synthesis(a, nside, lmax, spin)) <= MIXED_TOL
        assert _rel(bmx, O.adjoint_synthesis(v, nside, lmax, spin)) <= MIXED_TOL


@pytest.mark.parametrize("f16", [1, 0])
@pytest.mark.parametrize("nside,lmax,batch,amp", [(16, 32, 16, 1.0), (64, 128, 17, 1.0), (128, 256, 16, 1e-30),
                                                  (128, 256, 16, 1e25), (256, 400, 32, 1.0)])
def test_mixed_f16_split_synthesis(nside, lmax, batch, amp, f16):
    """The 16-map spin-2 synthesis on the FP16 matrix cores (f16 1: hi + lo halves of every
    operand after power-of-two scales) and on the FP32 ones (f16 0) against the FP64 path:
    ragged map groups, inputs far from unit scale (the per-map scale pass), and an all-zero
    map, which must synthesise to exact zeros."""
    j = _jht()
    from jht_amd import healpix as H

    with _tuned(nside, lmax, 2, "mixed", mx_f16=1 if f16 else 2):
        rng = np.random.default_rng(4400 + nside + batch)
        a = np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                      for _ in range(batch)]) * amp
        a[1] = 0.0
        at = torch.tensor(a, device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s64 = j.synthesis_batch(at, nside, lmax, 2, precision="f64").cpu().numpy()
            smx = j.synthesis_batch(at, nside, lmax, 2, precision="mixed").cpu().numpy()
            plan = H.get_plan(nside, lmax, 2, batch, H._prec("mixed"))
            assert plan.tuning["mx_f16"] == (1 if f16 else 2)
    assert np.all(smx[1] == 0.0)
    es = max(_rel(smx[i], s64[i]) for i in range(batch) if i != 1)
    print(f"mixed f16={f16} nside={nside} lmax={lmax} B={batch} amp={amp:g}: synth {es:.2e}", flush=True)
    assert es <= MIXED_TOL


@pytest.mark.parametrize("group,fold", [(16, 1), (32, 1), (16, 2)])
@pytest.mark.parametrize("nside,lmax,batch,amp", [(1, 2, 17, 1.0), (4, 8, 33, 1.0), (16, 32, 16, 1.0),
                                                  (64, 128, 17, 1.0), (128, 256, 33, 1e-30), (128, 256, 32, 1e25),
                                                  (256, 400, 40, 1.0), (512, 640, 32, 1.0)])
def test_mixed_fold_synthesis(nside, lmax, batch, amp, group, fold):
    """The spin-2 FP16-split synthesis with the N/S parity folded into B (fold 1: one
    recursion sign per wave), 16 maps x 128 rings and 32 maps x 256 rings per work-group,
    and the even/odd-accumulator form (fold 2, 16 maps), against the FP64 path: ragged map
    groups, grids from 1 to 4 ring chunks, inputs far from unit scale, an all-zero map
    (exact zeros)."""
    j = _jht()
    from jht_amd import healpix as H

    with _tuned(nside, lmax, 2, "mixed", mx_synth_fold=fold, mx_synth_group=group):
        rng = np.random.default_rng(4800 + nside + batch + group)
        a = np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                      for _ in range(batch)]) * amp
        a[1] = 0.0
        at = torch.tensor(a, device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s64 = j.synthesis_batch(at, nside, lmax, 2, precision="f64").cpu().numpy()
            smx = j.synthesis_batch(at, nside, lmax, 2, precision="mixed").cpu().numpy()
            plan = H.get_plan(nside, lmax, 2, batch, H._prec("mixed"))
            assert plan.tuning["mx_synth_fold"] == fold and plan.tuning["mx_synth_group"] == group
    assert np.all(smx[1] == 0.0)
    es = max(_rel(smx[i], s64[i]) for i in range(batch) if i != 1)
    print(f"mixed fold={fold} group={group} nside={nside} lmax={lmax} B={batch} amp={amp:g}: synth {es:.2e}", flush=True)
    assert es <= MIXED_TOL


@pytest.mark.parametrize("f16", [1, 0])
@pytest.mark.parametrize("nside,lmax,batch,amp", [(16, 32, 16, 1.0), (64, 128, 17, 1.0), (128, 256, 16, 1e-30),
                                                  (128, 256, 16, 1e25), (256, 400, 32, 1.0)])
def test_mixed_f16_split_synthesis_spin0(nside, lmax, batch, amp, f16):
    """The 16-map spin-0 tile synthesis on the FP16 matrix cores (f16 1) and on the FP32 ones
    (f16 0) against the FP64 path: ragged groups, inputs far from unit scale, an all-zero
    map (exact zeros)."""
    j = _jht()
    from jht_amd import healpix as H

    with _tuned(nside, lmax, 0, "mixed", mx_f16=1 if f16 else 2):
        rng = np.random.default_rng(4600 + nside + batch)
        a = np.stack([O.gaussian_alm(lmax, rng) for _ in range(batch)]) * amp
        a[1] = 0.0
        at = torch.tensor(a, device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s64 = j.synthesis_batch(at, nside, lmax, 0, precision="f64").cpu().numpy()
            smx = j.synthesis_batch(at, nside, lmax, 0, precision="mixed").cpu().numpy()
            plan = H.get_plan(nside, lmax, 0, batch, H._prec("mixed"))
            assert plan.tuning["mx_f16"] == (1 if f16 else 2)
    assert np.all(smx[1] == 0.0)
    es = max(_rel(smx[i], s64[i]) for i in range(batch) if i != 1)
    print(f"mixed spin-0 f16={f16} nside={nside} lmax={lmax} B={batch} amp={amp:g}: synth {es:.2e}", flush=True)
    assert es <= MIXED_TOL


@pytest.mark.parametrize("f16", [1, 0])
@pytest.mark.parametrize("nside,lmax,batch,amp", [(1, 2, 17, 1.0), (4, 8, 17, 1.0), (16, 32, 16, 1.0),
                                                  (64, 128, 17, 1.0), (128, 256, 16, 1e-30), (128, 256, 16, 1e25),
                                                  (512, 640, 32, 1.0)])
def test_mixed_f16_split_adjoint(nside, lmax, batch, amp, f16):
    """The 16-map spin-2 column-block adjoint on the FP16 matrix cores (f16 1: hi + lo halves;
    the recursion values scaled by the plan's power of two, the V values by one per map and
    work-group) and on the FP32 ones (f16 0) against the FP64 adjoint: ragged groups, tiny
    grids, inputs far from unit scale, an all-zero map (exact zeros), and a grid with 4 ring
    chunks."""
    j = _jht()
    from jht_amd import healpix as H

    with _tuned(nside, lmax, 2, "mixed", mx_adj_f16=1 if f16 else 2):
        rng = np.random.default_rng(4500 + nside + batch)
        v = rng.standard_normal((batch, 2, 12 * nside * nside)) * amp
        v[1] = 0.0
        vt = torch.tensor(v, device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            b64 = j.adjoint_synthesis_batch(vt, nside, lmax, 2, precision="f64").cpu().numpy()
            bmx = j.adjoint_synthesis_batch(vt, nside, lmax, 2, precision="mixed").cpu().numpy()
            plan = H.get_plan(nside, lmax, 2, batch, H._prec("mixed"))
            assert plan.tuning["mx_adj_f16"] == (1 if f16 else 2)
    assert np.all(bmx[1] == 0.0)
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch) if i != 1)
    print(f"mixed adjoint f16={f16} nside={nside} lmax={lmax} B={batch} amp={amp:g}: {ea:.2e}", flush=True)
    assert ea <= MIXED_TOL


@pytest.mark.parametrize("f16", [1, 0])
@pytest.mark.parametrize("nside,lmax,batch,amp", [(16, 32, 16, 1.0), (64, 128, 17, 1.0), (128, 256, 16, 1e-30),
                                                  (128, 256, 16, 1e25), (512, 640, 32, 1.0)])
def test_mixed_f16_split_adjoint_spin0(nside, lmax, batch, amp, f16):
    """The 16-map spin-0 column-block adjoint on the FP16 matrix cores (f16 1) and on the
    FP32 ones (f16 0) against the FP64 adjoint: ragged groups, inputs far from unit scale,
    an all-zero map (exact zeros), and a grid with 4 ring chunks."""
    j = _jht()
    from jht_amd import healpix as H

    with _tuned(nside, lmax, 0, "mixed", mx_adj_f16=1 if f16 else 2):
        rng = np.random.default_rng(4700 + nside + batch)
        v = rng.standard_normal((batch, 12 * nside * nside)) * amp
        v[1] = 0.0
        vt = torch.tensor(v, device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            b64 = j.adjoint_synthesis_batch(vt, nside, lmax, 0, precision="f64").cpu().numpy()
            bmx = j.adjoint_synthesis_batch(vt, nside, lmax, 0, precision="mixed").cpu().numpy()
            plan = H.get_plan(nside, lmax, 0, batch, H._prec("mixed"))
            assert plan.tuning["mx_adj_f16"] == (1 if f16 else 2)
    assert np.all(bmx[1] == 0.0)
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch) if i != 1)
    print(f"mixed spin-0 adjoint f16={f16} nside={nside} lmax={lmax} B={batch} amp={amp:g}: {ea:.2e}", flush=True)
    assert ea <= MIXED_TOL


@pytest.mark.parametrize("synth_group,synth_occ,adj_group,adj_occ",
                         [(8, 1, 8, 1), (8, 2, 4, 1), (4, 1, 8, 2), (4, 1, 4, 1), (8, 3, 8, 2), (8, 2, 16, 2),
                          (16, 2, 16, 2), (16, 2, 4, 1)])
def test_mixed_group_variants(synth_group, synth_occ, adj_group, adj_occ):
    """The 8-maps-per-recursion spin-2 kernels (both register budgets), forced through the
    plan's tuning overrides, against the FP64 path; batch 11 leaves a ragged last group."""
    j = _jht()
    from jht_amd import healpix as H

    nside, lmax, batch = 64, 128, 11
    with _tuned(nside, lmax, 2, "mixed", mx_synth_group=synth_group, mx_synth_occ=synth_occ,
                mx_adj_group=adj_group, mx_adj_occ=adj_occ):
        rng = np.random.default_rng(77 + synth_group + synth_occ + adj_group + adj_occ)
        a = np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                      for _ in range(batch)])
        at = torch.tensor(a, device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s64 = j.synthesis_batch(at, nside, lmax, 2, precision="f64")
            smx = j.synthesis_batch(at, nside, lmax, 2, precision="mixed")
            b64 = j.adjoint_synthesis_batch(s64, nside, lmax, 2, precision="f64")
            bmx = j.adjoint_synthesis_batch(s64, nside, lmax, 2, precision="mixed")
    s64, smx, b64, bmx = (t.cpu().numpy() for t in (s64, smx, b64, bmx))
    es = max(_rel(smx[i], s64[i]) for i in range(batch))
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch))
    print(f"mixed groups synth={synth_group}/{synth_occ} adj={adj_group}/{adj_occ}: synth {es:.2e} "
          f"adj {ea:.2e}", flush=True)
    assert es <= MIXED_TOL and ea <= MIXED_TOL


@pytest.mark.parametrize("batch,cb", [(9, 1), (18, 1), (18, 0), (32, 1)])
def test_mixed_adjoint_ring_chunks(batch, cb):
    """Spin-2 mixed adjoint on a grid of 4 ring chunks (T = 1024 rings), 8-map (batch 9) and
    16-map (batch 18, a ragged last group; batch 32, two full groups) recursions, the 16-map
    one in column-block form (cb 1) and with the cross-wave reduction (cb 0), against the
    FP64 adjoint."""
    j = _jht()

    nside, lmax = 512, 640
    with _tuned(nside, lmax, 2, "mixed", mx_adj_cb=1 if cb else 2):
        rng = np.random.default_rng(500 + batch)
        v = torch.tensor(rng.standard_normal((batch, 2, 12 * nside * nside)), device="cuda")
        b64 = j.adjoint_synthesis_batch(v, nside, lmax, 2, precision="f64").cpu().numpy()
        bmx = j.adjoint_synthesis_batch(v, nside, lmax, 2, precision="mixed").cpu().numpy()
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch))
    print(f"mixed adjoint batch={batch} cb={cb}: {ea:.2e}", flush=True)
    assert ea <= MIXED_TOL


@pytest.mark.parametrize("nside,lmax,batch", [(32, 40, 7), (64, 128, 32), (512, 640, 40), (128, 300, 33)])
def test_mixed_adjoint_32_maps(nside, lmax, batch):
    """The 32-map spin-2 adjoint (one recursion sign per wave, hi / lo planes, double-buffered
    tiles) against the 16-map column-block kernel (the same split and scales, another
    summation order: <= 5e-6 of the largest coefficient) and against the FP64 adjoint, on one
    and four ring chunks and with ragged groups."""
    j = _jht()
    from jht_amd import _lib
    from jht_amd.healpix import Plan

    rng = np.random.default_rng(700 + batch)
    v = torch.tensor(rng.standard_normal((batch, 2, 12 * nside * nside)), device="cuda")
    out = {}
    for g in (32, 16):
        p = Plan(nside, lmax, 2, batch, _lib.PREC_MIXED, 0,
                 tuning=dict(mx_adj_group=g, mx_adj_cb=1, mx_adj_f16=1, mx_single=2))
        assert p.tuning["mx_adj_group"] == g
        b = torch.empty((batch, 2, O.alm_size(lmax)), dtype=torch.complex128, device="cuda")
        p.adjoint(v, b, batch)
        torch.cuda.synchronize()
        out[g] = b.cpu().numpy()
    b64 = j.adjoint_synthesis_batch(v, nside, lmax, 2, precision="f64").cpu().numpy()
    e16 = max(_rel(out[32][i], out[16][i]) for i in range(batch))
    ea = max(_rel(out[32][i], b64[i]) for i in range(batch))
    print(f"32-map adjoint nside={nside} lmax={lmax} batch={batch}: vs 16-map {e16:.2e}, vs FP64 {ea:.2e}", flush=True)
    assert e16 <= 5e-6
    assert ea <= MIXED_TOL


@pytest.mark.parametrize("nside,lmax,batch", [(1, 2, 32), (2, 3, 32), (2, 5, 33), (8, 33, 32), (8, 40, 64)])
def test_mixed_32_map_kernels_small_lmax(nside, lmax, batch):
    """The 32-map spin-2 kernels (k_synth_s2q; k_adj_s2p with its tile layout of partials) at
    tiny band limits, against FP64: columns of a single 32-l tile and of two, tiles running
    past lmax (their padding rows must not reach the reduction), lmin = lmax (one step), and
    a ragged second map group."""
    j = _jht()
    with _tuned(nside, lmax, 2, "mixed", mx_synth_group=32, mx_adj_group=32):
        rng = np.random.default_rng(1700 + lmax + batch)
        a = np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                      for _ in range(batch)])
        at = torch.tensor(a, device="cuda")
        v = torch.tensor(rng.standard_normal((batch, 2, 12 * nside * nside)), device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            s64 = j.synthesis_batch(at, nside, lmax, 2, precision="f64").cpu().numpy()
            smx = j.synthesis_batch(at, nside, lmax, 2, precision="mixed").cpu().numpy()
            b64 = j.adjoint_synthesis_batch(v, nside, lmax, 2, precision="f64").cpu().numpy()
            bmx = j.adjoint_synthesis_batch(v, nside, lmax, 2, precision="mixed").cpu().numpy()
        from jht_amd import healpix as H
        t = H.get_plan(nside, lmax, 2, batch, H._prec("mixed")).tuning
        assert t["mx_synth_group"] == 32 and t["mx_adj_group"] == 32 and t["mx_f16"] == 1 and t["mx_synth_fold"] == 1
    es = max(_rel(smx[i], s64[i]) for i in range(batch))
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch))
    print(f"32-map kernels nside={nside} lmax={lmax} B={batch}: synth {es:.2e} adj {ea:.2e}", flush=True)
    assert es <= MIXED_TOL and ea <= MIXED_TOL


@pytest.mark.parametrize("group", [32, 16, 8])
@pytest.mark.parametrize("nside,lmax", [(64, 128), (256, 300)])
def test_mixed_spin0_groups(nside, lmax, group):
    """Spin-0 mixed synthesis and adjoint with 32 maps per recursion (the tile synthesis
    with 4 column blocks; the adjoint takes 16), 16 (the 16x16x4 tile synthesis, the
    column-block adjoint) and 8 (the rank-1 synthesis, the per-wave-tile adjoint), batch 18
    (a ragged group), against the FP64 path."""
    j = _jht()
    from jht_amd import healpix as H

    with _tuned(nside, lmax, 0, "mixed", mx_synth_group=group, mx_adj_group=min(group, 16)):
        batch = 18
        rng = np.random.default_rng(90 + group + nside)
        a = np.stack([O.gaussian_alm(lmax, rng, 0) for _ in range(batch)])
        at = torch.tensor(a, device="cuda")
        s64 = j.synthesis_batch(at, nside, lmax, 0, precision="f64")
        smx = j.synthesis_batch(at, nside, lmax, 0, precision="mixed")
        b64 = j.adjoint_synthesis_batch(s64, nside, lmax, 0, precision="f64")
        bmx = j.adjoint_synthesis_batch(s64, nside, lmax, 0, precision="mixed")
    s64, smx, b64, bmx = (x.cpu().numpy() for x in (s64, smx, b64, bmx))
    es = max(_rel(smx[i], s64[i]) for i in range(batch))
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch))
    print(f"mixed spin-0 group {group} nside={nside} lmax={lmax}: synth {es:.2e} adj {ea:.2e}", flush=True)
    assert es <= MIXED_TOL and ea <= MIXED_TOL


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("nside,lmax", [(1, 2), (2, 5), (4, 8), (8, 23)])
def test_mixed_16map_kernels_tiny_grids(nside, lmax, spin):
    """The 16-map mixed kernels (tile syntheses, column-block adjoints) on grids where a
    work-group holds only a few valid rings (T = 2 .. 16), batch 17 (one full group and a
    one-map ragged group), against the FP64 path."""
    j = _jht()
    from jht_amd import healpix as H

    H.clear_plans()
    try:
        batch = 17
        rng = np.random.default_rng(300 + 10 * nside + spin)
        if spin:
            a = np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
                          for _ in range(batch)])
        else:
            a = np.stack([O.gaussian_alm(lmax, rng, 0) for _ in range(batch)])
        at = torch.tensor(a, device="cuda")
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")  # lmax above the 3 nside - 1 ceiling
            s64 = j.synthesis_batch(at, nside, lmax, spin, precision="f64")
            smx = j.synthesis_batch(at, nside, lmax, spin, precision="mixed")
            b64 = j.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="f64")
            bmx = j.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="mixed")
    finally:
        H.clear_plans()
    s64, smx, b64, bmx = (x.cpu().numpy() for x in (s64, smx, b64, bmx))
    es = max(_rel(smx[i], s64[i]) for i in range(batch))
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch))
    print(f"16-map kernels nside={nside} lmax={lmax} spin={spin}: synth {es:.2e} adj {ea:.2e}", flush=True)
    assert es <= MIXED_TOL and ea <= MIXED_TOL


def test_mixed_adjoint_column_block_bitwise():
    """The FP32 column-block 16-map adjoint keeps the arithmetic of the cross-wave form (FP32
    over each 64-ring block, FP64 over the blocks in order): identical bits, ragged batch."""
    j = _jht()

    nside, lmax, batch = 256, 400, 18
    rng = np.random.default_rng(808)
    v = torch.tensor(rng.standard_normal((batch, 2, 12 * nside * nside)), device="cuda")
    out = {}
    for cb in (1, 2):
        with _tuned(nside, lmax, 2, "mixed", mx_adj_f16=2, mx_adj_cb=cb):
            out[cb] = j.adjoint_synthesis_batch(v, nside, lmax, 2, precision="mixed").cpu()
    assert torch.equal(out[1], out[2])


@pytest.mark.parametrize("nside,lmax", [(8, 12), (64, 128), (128, 300), (256, 200)])
def test_f64_adjoint_ring_shapes_agree(nside, lmax):
    """Spin-2 FP64 adjoint with 1, 2 and 4 rings per lane (different ring-chunk counts and
    in-lane sums) against each other and, on the small grids, the oracle (lmax above and
    below the ring count)."""
    j = _jht()

    rng = np.random.default_rng(900 + nside)
    v = rng.standard_normal((2, 12 * nside * nside))
    vt = torch.tensor(v, device="cuda")
    outs = {}
    for r in (1, 2, 4):
        with _tuned(nside, lmax, 2, "f64", adj_rings=r):
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                outs[r] = j.adjoint_synthesis(vt, nside, lmax, 2).cpu().numpy()
    e = max(_rel(outs[r], outs[1]) for r in (2, 4))
    print(f"f64 adjoint rings per lane nside={nside} lmax={lmax}: {e:.2e}", flush=True)
    assert e <= TOL
    if nside <= 64:
        assert _rel(outs[4], O.adjoint_synthesis(v, nside, lmax, 2)) <= TOL


@pytest.mark.parametrize("nside,lmax", [(600, 700), (1100, 1200)])
def test_mixed_split_polar_rings(nside, lmax):
    """Mixed spin-2 plans transform the polar rings with i >= 513 (N = 4 i > 2048) as four
    length-i chirp-z transforms (Lq = 2048 up to i = 1024, 4096 above) with a radix-4 combine;
    the FP64 path keeps one chirp-z transform per ring.  nside 600 is not a power of two (its
    belt rings are chirp-z rings as well, not split) and splits only in the 2048 class; nside
    1100 splits in both.  Synthesis and adjoint against FP64, every map <= 1e-5."""
    j = _jht()

    rng = np.random.default_rng(6000 + nside)
    batch = 3
    a = np.stack([np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)]) for _ in range(batch)])
    v = rng.standard_normal((batch, 2, 12 * nside * nside))
    at, vt = torch.tensor(a, device="cuda"), torch.tensor(v, device="cuda")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        s64 = j.synthesis_batch(at, nside, lmax, 2, precision="f64").cpu().numpy()
        smx = j.synthesis_batch(at, nside, lmax, 2, precision="mixed").cpu().numpy()
        b64 = j.adjoint_synthesis_batch(vt, nside, lmax, 2, precision="f64").cpu().numpy()
        bmx = j.adjoint_synthesis_batch(vt, nside, lmax, 2, precision="mixed").cpu().numpy()
    es = max(_rel(smx[i], s64[i]) for i in range(batch))
    ea = max(_rel(bmx[i], b64[i]) for i in range(batch))
    print(f"split polar rings nside={nside} lmax={lmax}: synth {es:.2e} adjoint {ea:.2e}", flush=True)
    assert es <= MIXED_TOL and ea <= MIXED_TOL
