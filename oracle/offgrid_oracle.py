"""CPU float64 oracle for the off-grid (arbitrary-point) synthesis pair.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/`` may import it.
The product path (``jht_amd.offgrid``) never calls it.

A numpy restatement (written fresh, not copied) of jrcheshire/jht v0.2.0's
``synthesis_general`` / ``adjoint_synthesis_general``: ``src/jht/offgrid.py:1-243`` (a
Legendre step on a Clenshaw-Curtis theta grid, the double-Fourier-sphere extension to a
2 pi-periodic theta and a theta FFT) and ``src/jht/_nufft.py:1-181`` (a 2-D type-2 NUFFT
with the exponential-of-semicircle kernel, and its transpose, type 1).

How it is pinned (``tests/test_offgrid_oracle.py``), none of which retypes a formula:

* the NUFFT against the exact non-uniform DFT (``nudft2`` / ``nudft2_adjoint``, a plain
  double sum over modes) at the kernel's stated accuracy;
* the whole synthesis at HEALPix pixel centres against the on-grid oracle synthesis
  (itself pinned by scipy's ``sph_harm_y``, the explicit Wigner sum and brute-force Y);
* the whole synthesis against a brute-force sum of spin-weighted harmonics at random
  points, built from the on-grid oracle's Wigner-d columns for m < 0 as well as m >= 0
  (no double-Fourier-sphere identity involved);
* the strict adjoint identity <S a, v>_pts = sum G Re(conj(a) (S^T v)) in the metric
  G = 2 - delta_m0, and type 1 as the exact transpose of type 2.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

from . import jht_oracle as O

TWO_PI = 2.0 * np.pi

# requested accuracy -> (kernel width W, oversampling sigma): the reference's table of
# achievable accuracies (src/jht/_nufft.py:33-39)
KERNEL_DB = {1e-4: (6, 2.0), 1e-6: (8, 2.0), 1e-8: (10, 2.0), 1e-10: (14, 2.0), 1e-12: (16, 2.0)}


def kernel_params(eps: float):
    """(W, sigma): the cheapest entry whose accuracy meets eps, else the most accurate
    (``_nufft.py:42-47``)."""
    ok = [e for e in KERNEL_DB if e <= eps * 1.0000001]
    return KERNEL_DB[max(ok)] if ok else KERNEL_DB[min(KERNEL_DB)]


def es_beta(W: int, sigma: float) -> float:
    """ES shape parameter beta = pi W (1 - 1/(2 sigma)) (``_nufft.py:66-67``)."""
    return float(np.pi * W * (1.0 - 0.5 / sigma))


def es_kernel(t, beta: float, W: int):
    """psi(t) = exp(beta (sqrt(1 - (2t/W)^2) - 1)) on |t| < W/2, else 0 (``_nufft.py:70-85``)."""
    t = np.asarray(t, dtype=np.float64)
    z2 = (2.0 * t / W) ** 2
    inside = z2 < 1.0
    return np.where(inside, np.exp(beta * (np.sqrt(np.where(inside, 1.0 - z2, 1.0)) - 1.0)), 0.0)


def es_correction(N: int, n: int, W: int, beta: float) -> np.ndarray:
    """Fourier transform of psi at the centred modes of an n-point grid, by Gauss-Legendre
    quadrature with 2W + 24 nodes over the support (``_nufft.py:88-95``)."""
    ks = np.arange(-(N // 2), -(N // 2) + N, dtype=np.float64)
    t, wq = np.polynomial.legendre.leggauss(2 * W + 24)
    t, wq = t * (W / 2.0), wq * (W / 2.0)
    psi = np.exp(beta * (np.sqrt(np.clip(1.0 - (2.0 * t / W) ** 2, 0.0, 1.0)) - 1.0))
    return (psi[None, :] * np.cos(ks[:, None] * (TWO_PI / n) * t[None, :]) * wq[None, :]).sum(1)


class NufftPlan:
    """Tables of a 2-D NUFFT with Nk x Nm centred modes (``_nufft.py:98-151``)."""

    def __init__(self, Nk: int, Nm: int, eps: float):
        W, sigma = kernel_params(eps)
        self.Nk, self.Nm, self.W = Nk, Nm, W
        self.nk = O.next_smooth_even(sigma * Nk)
        self.nm = O.next_smooth_even(sigma * Nm)
        self.beta = es_beta(W, sigma)
        self.corr_k = es_correction(Nk, self.nk, W, self.beta)
        self.corr_m = es_correction(Nm, self.nm, W, self.beta)
        self.kbins = np.arange(-(Nk // 2), -(Nk // 2) + Nk) % self.nk
        self.mbins = np.arange(-(Nm // 2), -(Nm // 2) + Nm) % self.nm


def stencil(coord, n: int, W: int, beta: float):
    """(cell index mod n, weight) of the W-wide window around each point: cells
    ceil(xi - W/2) + 0..W-1 with xi = coord n / 2 pi (``_nufft.py:154-165``)."""
    xi = np.asarray(coord, dtype=np.float64) * n / TWO_PI
    base = np.ceil(xi - W / 2.0).astype(np.int64)
    ell = base[:, None] + np.arange(W)[None, :]
    return ell % n, es_kernel(xi[:, None] - ell, beta, W)


def nufft2d2(plan: NufftPlan, coeffs, x, y):
    """Type 2: f_j = sum_{k,m} c[k,m] e^{i(k x_j + m y_j)}, centred modes (``_nufft.py:154-170``)."""
    c = np.asarray(coeffs) / plan.corr_k[:, None] / plan.corr_m[None, :]
    C = np.zeros((plan.nk, plan.nm), dtype=np.complex128)
    C[plan.kbins[:, None], plan.mbins[None, :]] = c
    g = np.fft.ifft2(C) * (plan.nk * plan.nm)
    lk, wk = stencil(x, plan.nk, plan.W, plan.beta)
    lm, wm = stencil(y, plan.nm, plan.W, plan.beta)
    gg = g[lk[:, :, None], lm[:, None, :]]
    return np.sum(wk[:, :, None] * wm[:, None, :] * gg, axis=(1, 2))


def nufft2d1(plan: NufftPlan, vals, x, y):
    """Type 1, the transpose of type 2: c[k,m] = sum_j v_j e^{-i(k x_j + m y_j)} (``_nufft.py:173-181``)."""
    lk, wk = stencil(x, plan.nk, plan.W, plan.beta)
    lm, wm = stencil(y, plan.nm, plan.W, plan.beta)
    contrib = (wk[:, :, None] * wm[:, None, :]) * np.asarray(vals)[:, None, None]
    g = np.zeros((plan.nk, plan.nm), dtype=np.complex128)
    np.add.at(g, (np.broadcast_to(lk[:, :, None], contrib.shape), np.broadcast_to(lm[:, None, :], contrib.shape)), contrib)
    C = np.fft.fft2(g)
    return C[plan.kbins[:, None], plan.mbins[None, :]] / plan.corr_k[:, None] / plan.corr_m[None, :]


def nudft2(coeffs, x, y):
    """Exact type 2 by the double sum (the NUFFT's pin)."""
    Nk, Nm = coeffs.shape
    k = np.arange(-(Nk // 2), -(Nk // 2) + Nk)
    m = np.arange(-(Nm // 2), -(Nm // 2) + Nm)
    ek = np.exp(1j * np.outer(np.asarray(x), k))
    em = np.exp(1j * np.outer(np.asarray(y), m))
    return np.einsum("jk,km,jm->j", ek, coeffs, em)


def nudft2_adjoint(vals, x, y, Nk: int, Nm: int):
    """Exact type 1 by the double sum."""
    k = np.arange(-(Nk // 2), -(Nk // 2) + Nk)
    m = np.arange(-(Nm // 2), -(Nm // 2) + Nm)
    ek = np.exp(-1j * np.outer(np.asarray(x), k))
    em = np.exp(-1j * np.outer(np.asarray(y), m))
    return np.einsum("j,jk,jm->km", np.asarray(vals), ek, em)


# --------------------------------------------------------------------------- #
# Legendre step on the Clenshaw-Curtis grid and the double-Fourier-sphere map
# --------------------------------------------------------------------------- #


def cc_rings(lmax: int):
    """Clenshaw-Curtis rings on [0, pi] with both poles: n_s = next 2,3,5-smooth even
    size >= lmax + 1, plus one (``offgrid.py:72-75``)."""
    ns = O.next_smooth_even(lmax + 1) + 1
    return ns, np.arange(ns) * np.pi / (ns - 1)


def _cb(m: int, lmax: int) -> int:
    """idx(l, m) - l in the healpy packing idx(l, m) = m(2 lmax + 1 - m)/2 + l."""
    return m * (2 * lmax + 1 - m) // 2


def _lam(x, m: int, spin: int, lmax: int) -> np.ndarray:
    """lambda^spin_{l,m}(x), rows l = max(m,|spin|)..lmax."""
    if spin == 0:
        return O.normalized_legendre(x, m, lmax)
    return O.spin_weighted_lambda(x, m, spin, lmax)


def legendre_synth(c_tri, x, lmax: int, spin: int) -> np.ndarray:
    """C[m, j] = sum_l c_lm lambda^spin_lm(x_j) (``_recursion.py:372-398`` synth_contract)."""
    C = np.zeros((lmax + 1, np.size(x)), dtype=np.complex128)
    for m in range(lmax + 1):
        lmin = max(m, abs(spin))
        if lmin > lmax:
            continue
        cb = _cb(m, lmax)
        C[m] = _lam(x, m, spin, lmax).T @ c_tri[cb + lmin : cb + lmax + 1]
    return C


def legendre_adjoint(Ccot, x, lmax: int, spin: int) -> np.ndarray:
    """b_lm = sum_j lambda^spin_lm(x_j) Ccot[m, j] (``_recursion.py:401-423`` adjoint_contract)."""
    b = np.zeros(O.alm_size(lmax), dtype=np.complex128)
    for m in range(lmax + 1):
        lmin = max(m, abs(spin))
        if lmin > lmax:
            continue
        cb = _cb(m, lmax)
        b[cb + lmin : cb + lmax + 1] = _lam(x, m, spin, lmax) @ Ccot[m]
    return b


def _dfs_tables(lmax: int, spin: int, ns: int):
    nd = 2 * ns - 2
    mv = np.arange(-lmax, lmax + 1)
    sign_m = ((-1.0) ** np.abs(mv)) * ((-1.0) ** spin)
    south = np.arange(ns, nd)
    return nd, sign_m, south, nd - south


def coeffs_to_Fkm(Cpos, Cneg, lmax: int, spin: int):
    """(+spin, -spin) per-m ring coefficients -> 2-D Fourier coefficients F[k, m'] of the
    2 pi-periodic extension F(-theta, phi) = (-1)^s F(theta, phi + pi) (``offgrid.py:124-135``)."""
    ns = Cpos.shape[1]
    nd, sign_m, south, refl = _dfs_tables(lmax, spin, ns)
    Cfull = np.zeros((2 * lmax + 1, ns), dtype=np.complex128)
    Cfull[lmax:] = Cpos
    Cfull[:lmax] = np.conj(Cneg[1:][::-1])
    Dd = np.zeros((nd, 2 * lmax + 1), dtype=np.complex128)
    Dd[:ns] = Cfull.T
    Dd[south] = sign_m[None, :] * Cfull.T[refl]
    return np.fft.fftshift(np.fft.fft(Dd, axis=0), axes=0) / nd


def Fkm_to_coeffs(G, lmax: int, spin: int, ns: int):
    """Hermitian transpose of coeffs_to_Fkm (``offgrid.py:138-153``)."""
    nd, sign_m, south, refl = _dfs_tables(lmax, spin, ns)
    dDd = np.fft.ifft(np.fft.ifftshift(G, axes=0), axis=0)
    dT = dDd[:ns].copy()
    np.add.at(dT, refl, sign_m[None, :] * dDd[south])
    dC = dT.T
    Cpos_cot = dC[lmax:].copy()
    Cneg_cot = np.zeros((lmax + 1, ns), dtype=np.complex128)
    Cneg_cot[1:] = np.conj(dC[:lmax][::-1])
    return Cpos_cot, Cneg_cot


def _channels(alm, lmax: int, spin: int):
    """(+spin, -spin) channel coefficients: a itself for spin 0; -(E + iB) and
    -(-1)^s (E - iB) otherwise (``offgrid.py:98-110``)."""
    if spin == 0:
        return alm, alm
    s_sign = (-1.0) ** spin
    return -(alm[0] + 1j * alm[1]), -s_sign * (alm[0] - 1j * alm[1])


def _check_spin(spin: int):
    if spin not in (0, 1, 2, 3):
        raise NotImplementedError(f"spin={spin} unsupported (0, 1, 2, 3)")


def synthesis_general(alm, loc, spin: int, lmax: int, epsilon: float = 1e-10):
    """a_lm -> field at points loc = (theta, phi) (``offgrid.py:165-205``)."""
    _check_spin(spin)
    alm = np.asarray(alm, dtype=np.complex128)
    loc = np.asarray(loc, dtype=np.float64)
    ns, theta = cc_rings(lmax)
    x = np.cos(theta)
    a, b = _channels(alm, lmax, spin)
    Cpos = legendre_synth(a, x, lmax, spin)
    Cneg = Cpos if spin == 0 else legendre_synth(b, x, lmax, -spin)
    F = coeffs_to_Fkm(Cpos, Cneg, lmax, spin)
    plan = NufftPlan(2 * ns - 2, 2 * lmax + 1, epsilon)
    f = nufft2d2(plan, F, loc[:, 0], loc[:, 1])
    return f.real if spin == 0 else np.stack([f.real, f.imag])


def adjoint_synthesis_general(field, loc, spin: int, lmax: int, epsilon: float = 1e-10):
    """Exact transpose of synthesis_general in the (2 - delta_m0) metric (``offgrid.py:208-243``)."""
    _check_spin(spin)
    loc = np.asarray(loc, dtype=np.float64)
    field = np.asarray(field, dtype=np.float64)
    ns, theta = cc_rings(lmax)
    x = np.cos(theta)
    fc = field.astype(np.complex128) if spin == 0 else field[0] + 1j * field[1]
    plan = NufftPlan(2 * ns - 2, 2 * lmax + 1, epsilon)
    G = nufft2d1(plan, fc, loc[:, 0], loc[:, 1])
    Cp, Cn = Fkm_to_coeffs(G, lmax, spin, ns)
    p = legendre_adjoint(Cp, x, lmax, spin)
    n = legendre_adjoint(Cn, x, lmax, -spin)
    g = O.alm_metric_weight(lmax)
    if spin == 0:
        return (p + n) / g
    s_sign = (-1.0) ** spin
    return np.stack([-(p + s_sign * n), 1j * (p - s_sign * n)]) / g


def brute_force_field(alm, loc, spin: int, lmax: int):
    """Direct sum of spin-weighted harmonics at each point, negative m included:
    spin 0: f = sum_{l, m>=0} G_m Re(a_lm lambda_lm(theta) e^{i m phi});
    spin s > 0: Q + iU = sum_{l, all m} sa_lm lambda^s_{l,m}(theta) e^{i m phi} with
    sa_{l,-m} = (-1)^(s+m) conj(-sa_lm) (the reality condition of Q, U, from
    conj(sY_lm) = (-1)^(s+m) -sY_{l,-m})."""
    loc = np.asarray(loc, dtype=np.float64)
    th, ph = loc[:, 0], loc[:, 1]
    x = np.cos(th)
    if spin == 0:
        f = np.zeros(th.size)
        for m in range(lmax + 1):
            cb = _cb(m, lmax)
            lam = O.normalized_legendre(x, m, lmax)
            s = lam.T @ alm[cb + m : cb + lmax + 1]
            f += (1.0 if m == 0 else 2.0) * np.real(s * np.exp(1j * m * ph))
        return f
    ap, am = _channels(np.asarray(alm), lmax, spin)  # +s a_lm and -s a_lm, m >= 0
    z = np.zeros(th.size, dtype=np.complex128)
    for m in range(-lmax, lmax + 1):
        mm = abs(m)
        lmin = max(mm, spin)
        if lmin > lmax:
            continue
        cb = _cb(mm, lmax)
        if m >= 0:
            c = ap[cb + lmin : cb + lmax + 1]
        else:
            c = ((-1.0) ** (mm + spin)) * np.conj(am[cb + lmin : cb + lmax + 1])
        # lambda^{s}_{l,m} = (-1)^m sqrt((2l+1)/4pi) d^l_{-m,s}, for either sign of m
        d = O.wigner_d_column(x, -m, spin, lmax)
        l = np.arange(lmin, lmax + 1, dtype=np.float64)
        lam = ((-1.0) ** mm) * np.sqrt((2.0 * l + 1.0) / O.FOUR_PI)[:, None] * d
        z += (lam.T @ c) * np.exp(1j * m * ph)
    return np.stack([z.real, z.imag])


@lru_cache(maxsize=None)
def healpix_centres(nside: int):
    """(theta, phi) of every RING pixel centre."""
    geo = O.RingInfo(nside)
    th, ph = [], []
    for r in range(geo.nrings):
        N = int(geo.npix_ring[r])
        th.append(np.full(N, np.arccos(geo.z[r])))
        ph.append(geo.phi0[r] + TWO_PI * np.arange(N) / N)
    return np.stack([np.concatenate(th), np.concatenate(ph)], axis=1)
