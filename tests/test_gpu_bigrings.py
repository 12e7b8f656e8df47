"""Single-mode maps at nside 2048 (GPU): the ring-FFT launch classes that only exist on
large grids (cap rings with 4096 < N < 8192 through the L = 16384 chirp-z class, the
8192-point belt rings) against the closed form of one (l, m) mode.

For spin 0 and a_lm = a on one (l, m > 0) only, the map on ring r is
    f(theta_r, phi_j) = 2 Re(a lambda_lm(theta_r) e^{i m phi_j}),
with lambda from the oracle's recursion (pinned against scipy and mpmath in
tests/test_oracle.py).  Checked on rings of every class, relative to the map's peak
(max |x - x_ref| / max |x_ref|): FP64 <= 1e-12, mixed <= 1e-5.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O

pytestmark = pytest.mark.gpu

NSIDE, LMAX = 2048, 4096


def _rings():
    # north cap rings in the L = 16384 chirp-z class (i > 1024, N = 4i not a power of two),
    # a power-of-two cap ring, belt rings (N = 8192), the equator, and their mirrors
    caps = [1100, 1500, 1999, 2047, 1024]
    belt = [2048, 3000, 4095]
    north = [i - 1 for i in caps + belt]
    return north + [4 * NSIDE - 2 - r for r in north if r < 4 * NSIDE - 2 - r]


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("mixed", 1e-5)])
@pytest.mark.parametrize("l,m", [(3000, 2500), (4096, 700), (1800, 1799)])
def test_single_mode_big_rings(precision, tol, l, m):
    import jht_amd

    geo = O.RingInfo(NSIDE)
    a = 0.7 - 0.4j
    alm = np.zeros(O.alm_size(LMAX), dtype=np.complex128)
    alm[O.alm_column_base(m, LMAX) + (l - m)] = a
    out = jht_amd.synthesis_batch(torch.tensor(alm[None], device="cuda"), NSIDE, LMAX, 0, precision=precision)[0]
    rings = _rings()
    lam_all = O.normalized_legendre(geo.z, m, l)[l - m]  # lambda_lm on every ring
    peak = 2.0 * abs(a) * float(np.max(np.abs(lam_all)))  # the map's max |value| (the metric's scale)
    lam = lam_all[rings]
    worst = 0.0
    for r, lv in zip(rings, lam):
        N = int(geo.npix_ring[r])
        phi = geo.phi0[r] + 2 * np.pi * np.arange(N) / N
        want = 2.0 * np.real(a * lv * np.exp(1j * m * phi))
        got = out[int(geo.startpix[r]) : int(geo.startpix[r]) + N].cpu().numpy()
        worst = max(worst, float(np.max(np.abs(got - want))))
    e = worst / peak  # max |x - x_ref| / max |x_ref| over the map, as the reference's tests measure
    print(f"single mode l={l} m={m} {precision}: rel {e:.2e} over {len(rings)} rings", flush=True)
    assert e <= tol
