"""Parity at the headline size: nside 2048, lmax 4096, spin 2 (and spin 0) on the GPU.

(a) The FP64 Legendre stage through the C ABI (jhtx_legendre_synth / jhtx_legendre_adjoint
    on a HEALPix plan): F[c][ring][m] on every ring, north and south, and the adjoint's
    b_lm for sampled orders m, against the oracle's long-double lambda (pinned against
    mpmath to 1e-15 in tests/test_oracle.py), <= 5e-12 per column (max over the column's
    rings).  The oracle's float64 form -- the reference's own arithmetic -- is compared too,
    and its distance from the long-double one printed: at lmax 4096 it reaches 3e-11 to
    5e-11 per column (the float64 lgamma seed at high m, sqrt(1 - x^2) and the
    log-renormalised recursion next to the poles).  The FP64 forward recursion's own
    rounding next to the poles leaves ~2e-12 on the m = 0 columns, hence 5e-12.
(b) Single (l, m) E or B modes, 16 maps in one call of the default mixed plan (the 16-map
    FP16-split kernels) and 4 maps through the FP64 path, against the closed form on every
    pixel (tests/_closed_form.py); reference: tests/test_highL.py:76-101 (random E/B at
    2048, the single mode l = 1400, m = 700).
(c) Batches of 16 and 32 Gaussian E/B maps (C_l = (l+1)^-2, B = 0.1 E; 32 is the benchmark's
    batch and its 32-map kernels): mixed against FP64, synthesis per map and the adjoint per
    l-band, each <= 1e-5.
(d) Amplitudes spread over 60 decades across the maps of a group and a beam-smoothed
    (steep) spectrum, at nside 1024 / lmax 2048: the per-map scales of the FP16 split.

Relative error is max|x - ref| / max|ref| (tests/test_vectorized.py:32); per l-band it is
taken over the band's coefficients.
"""

from __future__ import annotations

import math
import warnings

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O

from _closed_form import PixelGeometry

pytestmark = pytest.mark.gpu

NSIDE, LMAX = 2048, 4096
SAMPLED_M = (0, 1, 700, 2048, 3500, 4095, 4096)
MIXED_TOL = 1e-5
BAND_EDGES = (0, 64, 256, 1024, 2048, 4096)


@pytest.fixture(autouse=True)
def _release_plans():
    warnings.simplefilter("ignore")  # lmax = 2 nside is above the 1.5 nside ceiling
    yield
    from jht_amd import healpix as H

    H.clear_plans()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _rel(a, b) -> float:
    if isinstance(a, torch.Tensor):
        return float((a - b).abs().max() / b.abs().max())
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _gaussian_alm_gpu(lmax, batch, spin, seed, amps=None, beam_sigma=0.0):
    """(batch, [2,] K) Gaussian a_lm made on the GPU: Re, Im ~ N(0,1) sqrt(C_l / 2) (m = 0
    real with variance C_l), C_l = amp (l+1)^-2 exp(-l (l+1) sigma^2); spin 2: (E, B) with
    C_l^BB = 0.1 C_l^EE and l < 2 zero."""
    dev = "cuda"
    ell = torch.as_tensor(O.alm_l_of_index(lmax), device=dev, dtype=torch.float64)
    m0 = torch.as_tensor(O.alm_m_of_index(lmax) == 0, device=dev)
    cl = (ell + 1.0) ** -2 * torch.exp(-ell * (ell + 1.0) * beam_sigma**2)
    sd = torch.where(m0, cl.sqrt(), (0.5 * cl).sqrt())
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    ncomp = 1 if spin == 0 else 2
    shape = (batch, ncomp, ell.numel())
    re = torch.randn(shape, generator=g, device=dev, dtype=torch.float64) * sd
    im = torch.where(m0, 0.0, torch.randn(shape, generator=g, device=dev, dtype=torch.float64) * sd)
    a = torch.complex(re, im)
    if spin:
        a[:, 1] *= math.sqrt(0.1)
        a[..., ell < 2] = 0.0
    if amps is not None:
        a *= amps.to(torch.complex128)[:, None, None]
    return a if spin else a[:, 0].contiguous()


def _band_errors(b, ref, lmax):
    """{band: max|b - ref| / max|ref| over the band's coefficients} (last axis K)."""
    ell = torch.as_tensor(O.alm_l_of_index(lmax), device=b.device)
    edges = [e for e in BAND_EDGES if e <= lmax] + [lmax + 1]
    out = {}
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (ell >= lo) & (ell < hi)
        r = ref[..., sel].abs().max()
        if float(r) > 0:
            out[f"{lo}-{hi - 1}"] = float((b[..., sel] - ref[..., sel]).abs().max() / r)
    return out


# --------------------------------------------------------------------------------------- #
# (a) FP64 Legendre stage, sampled orders
# --------------------------------------------------------------------------------------- #
@pytest.mark.parametrize("spin", [2, 0])
def test_legendre_stage_f64_sampled_orders(spin):
    from jht_amd import _lib
    from jht_amd import healpix as H

    lib = _lib.load()
    geo = O.RingInfo(NSIDE)
    ncomp = 1 if spin == 0 else 2
    plan = H.Plan(NSIDE, LMAX, spin, 1, _lib.PREC_F64)
    a = _gaussian_alm_gpu(LMAX, 1, spin, seed=31)
    F = torch.empty((1, ncomp, geo.nrings, LMAX + 1), dtype=torch.complex128, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    print(f"legendre spin {spin}: synth", flush=True)
    _lib.check(lib.jhtx_legendre_synth(plan._h, a.data_ptr(), F.data_ptr(), 1, stream))
    g = torch.Generator(device="cuda")
    g.manual_seed(32)
    Fin = torch.complex(torch.randn(F.shape, generator=g, device="cuda", dtype=torch.float64),
                        torch.randn(F.shape, generator=g, device="cuda", dtype=torch.float64))
    b = torch.empty_like(a)
    print(f"legendre spin {spin}: adjoint", flush=True)
    _lib.check(lib.jhtx_legendre_adjoint(plan._h, Fin.data_ptr(), b.data_ptr(), 1, plan.workspace.data_ptr(),
                                         stream))
    torch.cuda.synchronize()
    ms = torch.as_tensor(SAMPLED_M, device="cuda")
    Fs = F[0][..., ms].cpu().numpy()  # (ncomp, nrings, len(SAMPLED_M))
    Fi = Fin[0][..., ms].cpu().numpy()
    ah = a[0].cpu().numpy().reshape(ncomp, -1)
    bh = b[0].cpu().numpy().reshape(ncomp, -1)
    del F, Fin
    worst_s = worst_a = 0.0
    for i, m in enumerate(SAMPLED_M):
        lmin = max(m, spin)
        cb = O.alm_column_base(m, LMAX)
        col = slice(cb + lmin - m, cb + LMAX - m + 1)  # l = lmin .. lmax of column m
        res = {}
        for precise in (True, False):
            if spin == 0:
                lam = O.normalized_legendre(geo.z, m, LMAX, precise)  # (L - m + 1, nrings)
                res[precise] = ([lam.T @ ah[0][col]], [lam @ Fi[0][:, i]])
            else:
                lp = O.spin_weighted_lambda(geo.z, m, 2, LMAX, precise)
                lm = O.spin_weighted_lambda(geo.z, m, -2, LMAX, precise)
                E, B = ah[0][col], ah[1][col]
                p, n = lp @ Fi[0][:, i], lm @ Fi[1][:, i]
                res[precise] = ([lp.T @ (-(E + 1j * B)), lm.T @ (-(E - 1j * B))], [-(p + n) / 2, 1j * (p - n) / 2])
        got_s = [Fs[c][:, i] for c in range(ncomp)]
        got_a = [bh[c][col] for c in range(ncomp)]

        def err(x, y):
            return max(_rel(u, v) for u, v in zip(x, y))

        es, ea = err(got_s, res[True][0]), err(got_a, res[True][1])
        print(f"legendre f64 spin {spin} m={m}: GPU vs long-double oracle F {es:.2e} adjoint {ea:.2e}; "
              f"GPU vs float64 oracle F {err(got_s, res[False][0]):.2e} adjoint {err(got_a, res[False][1]):.2e}; "
              f"float64 vs long-double oracle F {err(res[False][0], res[True][0]):.2e}", flush=True)
        if spin and m < 2:  # entries l < |s| of the column are zero
            assert np.all(bh[:, cb : cb + 2 - m] == 0)
        worst_s, worst_a = max(worst_s, es), max(worst_a, ea)
    assert worst_s <= 5e-12 and worst_a <= 5e-12


# --------------------------------------------------------------------------------------- #
# (b) single modes against the closed form, every pixel
# --------------------------------------------------------------------------------------- #
MODES_S2 = [(2, 0, "E"), (2, 2, "B"), (100, 1, "E"), (1400, 700, "E"), (1400, 700, "B"), (4096, 0, "B"),
            (4096, 4096, "E"), (3000, 2500, "B"), (4096, 700, "E"), (2048, 2047, "B"), (3500, 100, "E"),
            (1024, 512, "B"), (2600, 1300, "E"), (4000, 3999, "B"), (500, 499, "E"), (3333, 1111, "B")]


@pytest.mark.parametrize("precision,tol,modes", [("mixed", MIXED_TOL, MODES_S2),
                                                 ("f64", 1e-12, [MODES_S2[i] for i in (3, 4, 6, 7)])])
def test_single_modes_spin2(precision, tol, modes):
    import jht_amd

    B, K = len(modes), O.alm_size(LMAX)
    alm = torch.zeros((B, 2, K), dtype=torch.complex128, device="cuda")
    vals = []
    for i, (l, m, c) in enumerate(modes):
        v = complex(0.7, -0.4) if m > 0 else complex(0.7, 0.0)
        vals.append(v)
        alm[i, 0 if c == "E" else 1, O.alm_column_base(m, LMAX) + l - m] = v
    print(f"single modes spin 2 {precision} batch {B}", flush=True)
    out = jht_amd.synthesis_batch(alm, NSIDE, LMAX, 2, precision=precision)
    pg = PixelGeometry(NSIDE)
    worst = 0.0
    for i, (l, m, c) in enumerate(modes):
        aE, aB = (vals[i], 0j) if c == "E" else (0j, vals[i])
        e = _rel(out[i], pg.spin2(l, m, aE, aB, precise=precision == "f64"))
        print(f"  mode l={l} m={m} {c}: rel {e:.2e}", flush=True)
        worst = max(worst, e)
    assert worst <= tol


# --------------------------------------------------------------------------------------- #
# (c) the benchmark's batch: mixed against FP64, adjoint per l-band
# --------------------------------------------------------------------------------------- #
@pytest.mark.parametrize("B", [16, 32])
def test_mixed_batch_against_f64_by_band(B):
    """Batch 16 runs the 16-map kernels (four split products), batch 32 the benchmark's
    default 32-map kernels (three products, leading zero tiles of the adjoint skipped)."""
    import jht_amd
    from jht_amd import healpix as H

    sel = [0, B - 1]
    alm = _gaussian_alm_gpu(LMAX, B, 2, seed=20261203)
    print(f"batch {B} mixed synthesis", flush=True)
    smx = jht_amd.synthesis_batch(alm, NSIDE, LMAX, 2, precision="mixed")
    s64 = jht_amd.synthesis_batch(alm[sel].contiguous(), NSIDE, LMAX, 2, precision="f64")
    es = [_rel(smx[i], s64[k]) for k, i in enumerate(sel)]
    print(f"  synthesis vs f64: {es}", flush=True)
    smx[sel] = s64  # the same adjoint input for both paths on the compared maps
    print(f"batch {B} mixed adjoint", flush=True)
    bmx = jht_amd.adjoint_synthesis_batch(smx, NSIDE, LMAX, 2, precision="mixed")
    del smx
    b64 = jht_amd.adjoint_synthesis_batch(s64, NSIDE, LMAX, 2, precision="f64")
    worst = 0.0
    for k, i in enumerate(sel):
        bands = _band_errors(bmx[i], b64[k], LMAX)
        print(f"  adjoint map {i} by l-band: {bands}", flush=True)
        worst = max(worst, max(bands.values()))
    del bmx, b64, s64, alm
    H.clear_plans()  # a batch-32 plan at this size holds about 75 GB of workspace
    assert max(es) <= MIXED_TOL and worst <= MIXED_TOL


# --------------------------------------------------------------------------------------- #
# (d) amplitudes over 60 decades, steep spectrum
# --------------------------------------------------------------------------------------- #
@pytest.mark.parametrize("spin", [2, 0])
def test_mixed_amplitude_decades_and_steep_spectrum(spin):
    """16 maps in one group with amplitudes 1e-30 .. 1e30 and a beam-smoothed spectrum
    (C_l down by e^-23 ~ 1e-10 more at lmax): the synthesis of every map <= 1e-5; the adjoint
    of white-noise maps with the same amplitudes <= 1e-5 in every l-band; the adjoint of the
    smoothed maps <= 1e-5 over each map's coefficients.  (Per band the last is only reported:
    the high-l coefficients of a smooth map are cancellation residues far below the
    contraction's rounding of its large terms, in FP32 as in any other 24-bit arithmetic.)"""
    import jht_amd

    nside, lmax, B = 1024, 2048, 16
    npix = 12 * nside * nside
    amps = torch.tensor([10.0 ** (-30 + 4 * i) for i in range(B)], device="cuda", dtype=torch.float64)
    a = _gaussian_alm_gpu(lmax, B, spin, seed=71 + spin, amps=amps, beam_sigma=math.sqrt(23.0) / lmax)
    print(f"decades spin {spin}: synthesis", flush=True)
    s64 = jht_amd.synthesis_batch(a, nside, lmax, spin, precision="f64")
    smx = jht_amd.synthesis_batch(a, nside, lmax, spin, precision="mixed")
    es = [_rel(smx[i], s64[i]) for i in range(B)]
    print(f"  synthesis per map: {[f'{e:.1e}' for e in es]}", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(72 + spin)
    shape = (B, npix) if spin == 0 else (B, 2, npix)
    v = torch.randn(shape, generator=g, device="cuda", dtype=torch.float64)
    v *= amps.view((B,) + (1,) * (v.dim() - 1))
    print(f"decades spin {spin}: adjoint of white maps", flush=True)
    bw64 = jht_amd.adjoint_synthesis_batch(v, nside, lmax, spin, precision="f64")
    bwmx = jht_amd.adjoint_synthesis_batch(v, nside, lmax, spin, precision="mixed")
    worst_w = 0.0
    for i in range(B):
        bands = _band_errors(bwmx[i], bw64[i], lmax)
        worst_w = max(worst_w, max(bands.values()))
    print(f"  white adjoint worst band: {worst_w:.2e}", flush=True)
    print(f"decades spin {spin}: adjoint of smoothed maps", flush=True)
    bs64 = jht_amd.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="f64")
    bsmx = jht_amd.adjoint_synthesis_batch(s64, nside, lmax, spin, precision="mixed")
    ea = [_rel(bsmx[i], bs64[i]) for i in range(B)]
    print(f"  smoothed adjoint per map: {[f'{e:.1e}' for e in ea]}", flush=True)
    print(f"  smoothed adjoint by band (map 8): {_band_errors(bsmx[8], bs64[8], lmax)}", flush=True)
    assert max(es) <= MIXED_TOL and worst_w <= MIXED_TOL and max(ea) <= MIXED_TOL
