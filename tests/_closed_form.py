"""Closed-form single-mode maps, built on the GPU in float64 (test infrastructure).

For one (l, m) mode the synthesis is known in closed form from lambda_lm(theta_r) (the
oracle's recursion, pinned against scipy, mpmath and the explicit Wigner sum in
tests/test_oracle.py) and e^{i m phi}:

* spin 0: f = 2 Re(a lambda e^{i m phi}) for m > 0, Re(a) lambda for m = 0;
* spin 2: Q + iU = Fp e^{i m phi} + conj(Fm) e^{-i m phi} for m > 0 and Fp for m = 0, with
  Fp = bp lambda^{+2}, Fm = bm lambda^{-2}, bp = -(aE + i aB), bm = -(aE - i aB)
  (src/jht/healpix.py:309-330: D_k = sum_{m = k} Fp_m + sum_{m >= 1, -m = k} conj(Fm_m),
  map = ifft(D) N).

Pixel j of ring r sits at phi = pi (2 j + shift_r) / N_r (shift 1 for the cap rings and
the shifted belt rings), so m phi is reduced exactly, as the integer m (2 j + shift)
mod 2 N_r, before the float64 cos / sin.  ``precise`` takes lambda from the oracle's
long-double recursion (for the FP64 gates at lmax ~ 4096, where the float64 form's own
rounding reaches 1e-10 near the poles).
"""

from __future__ import annotations

import numpy as np
import torch

from oracle import jht_oracle as O


class PixelGeometry:
    """Per-pixel ring index and phase numerator of a RING map, on the device."""

    def __init__(self, nside: int, device="cuda"):
        geo = O.RingInfo(nside)
        self.geo = geo
        N = torch.as_tensor(geo.npix_ring, device=device)
        start = torch.as_tensor(geo.startpix, device=device)
        shift = torch.as_tensor(np.rint(geo.phi0 * geo.npix_ring / np.pi).astype(np.int64), device=device)
        self.ring = torch.repeat_interleave(torch.arange(geo.nrings, device=device), N)
        j = torch.arange(geo.npix, device=device) - start[self.ring]
        self.num = 2 * j + shift[self.ring]  # phi = pi num / N
        self.den = 2 * N[self.ring]
        self.device = device

    def cos_sin(self, m: int):
        k = torch.remainder(m * self.num, self.den)
        ang = (2.0 * np.pi) * k.to(torch.float64) / self.den.to(torch.float64)
        return torch.cos(ang), torch.sin(ang)

    def spin0(self, l: int, m: int, a: complex, precise: bool = False) -> torch.Tensor:
        lam = O.normalized_legendre(self.geo.z, m, l, precise)[l - m]
        lam = torch.as_tensor(lam, device=self.device)[self.ring]
        if m == 0:
            return a.real * lam
        c, s = self.cos_sin(m)
        return 2.0 * lam * (a.real * c - a.imag * s)

    def spin2(self, l: int, m: int, aE: complex, aB: complex, precise: bool = False) -> torch.Tensor:
        """(2, npix): Q, U of the single mode."""
        bp = -(aE + 1j * aB)
        bm = -(aE - 1j * aB)
        lam = lambda s: O.spin_weighted_lambda(self.geo.z, m, s, l, precise)[-1]  # noqa: E731
        lp = torch.as_tensor(lam(2), device=self.device)[self.ring]
        if m == 0:
            return torch.stack([bp.real * lp, bp.imag * lp])
        lm = torch.as_tensor(lam(-2), device=self.device)[self.ring]
        c, s = self.cos_sin(m)
        fpr, fpi = bp.real * lp, bp.imag * lp
        fmr, fmi = bm.real * lm, bm.imag * lm
        q = (fpr + fmr) * c - (fpi + fmi) * s
        u = (fpr - fmr) * s + (fpi - fmi) * c
        return torch.stack([q, u])
