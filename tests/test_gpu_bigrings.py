"""Single-mode spin-0 maps at nside 2048 / lmax 4096 (GPU) against their closed form on
every pixel: every ring-FFT launch class of the large grid (cap rings with 4096 < N < 8192
through the L = 16384 chirp-z class, power-of-two cap rings, the 8192-point belt rings,
shifted and unshifted) and the polar cap rings where the exponent-tracked seeds of high m
act.

For a_lm = a on one (l, m > 0) only, f(theta_r, phi_j) = 2 Re(a lambda_lm(theta_r)
e^{i m phi_j}) (tests/_closed_form.py, lambda from the oracle's recursion).  FP64: three
modes in one call, <= 1e-12; mixed: 16 modes in one call, so the default 16-map kernels
run, <= 1e-5 (max |x - x_ref| / max |x_ref| per map).
"""

from __future__ import annotations

import warnings

import pytest
import torch

from oracle import jht_oracle as O

from _closed_form import PixelGeometry

pytestmark = pytest.mark.gpu

NSIDE, LMAX = 2048, 4096
MODES = [(3000, 2500), (4096, 700), (1800, 1799), (4096, 4096), (2, 0), (4096, 0), (1, 1), (100, 3),
         (2500, 2048), (4095, 4000), (1024, 1), (3333, 3000), (700, 699), (4096, 1), (2047, 1024), (3900, 3899)]


@pytest.mark.parametrize("precision,tol,modes", [("f64", 1e-12, MODES[:3]), ("mixed", 1e-5, MODES)])
def test_single_modes_big_rings(precision, tol, modes):
    import jht_amd
    from jht_amd import healpix as H

    B, K = len(modes), O.alm_size(LMAX)
    alm = torch.zeros((B, K), dtype=torch.complex128, device="cuda")
    vals = []
    for i, (l, m) in enumerate(modes):
        v = complex(0.7, -0.4) if m > 0 else complex(0.7, 0.0)
        vals.append(v)
        alm[i, O.alm_column_base(m, LMAX) + l - m] = v
    print(f"single modes spin 0 {precision} batch {B}", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = jht_amd.synthesis_batch(alm, NSIDE, LMAX, 0, precision=precision)
    pg = PixelGeometry(NSIDE)
    worst = 0.0
    for i, (l, m) in enumerate(modes):
        want = pg.spin0(l, m, vals[i], precise=precision == "f64")
        e = float((out[i] - want).abs().max() / want.abs().max())
        print(f"  mode l={l} m={m} {precision}: rel {e:.2e}", flush=True)
        worst = max(worst, e)
    del out, pg
    H.clear_plans()
    torch.cuda.empty_cache()
    assert worst <= tol
