"""CPU checks of the caller layer's host parts (no GPU): ring weights, the real-DOF
isometry, band power and the differentiable CG solve.

Pins: the ring weights against the oracle's restatement and their defining property (the
m = 0 quadrature is exact for even Legendre polynomials to 2 nside, ``weights.py:60-83``);
``||T a||_2 = ||a||_w`` and ``T^-1 T = I`` on the active modes (``masked.py:142-176``);
band power against an explicit per-l sum; CG against ``numpy.linalg.solve`` and its
implicit backward against ``torch.autograd.gradcheck`` (finite differences).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O


def _c():
    from jht_amd import callers

    return callers


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 16, 32])
def test_ring_weights_match_oracle_and_quadrature(nside):
    C = _c()
    w = C.ring_weights(nside)
    assert w.shape == (2 * nside,)
    assert np.allclose(w, O.ring_weights(nside), rtol=0, atol=1e-13)
    geo = O.RingInfo(nside)
    T = 2 * nside
    mult = np.where(np.arange(T) == T - 1, 1.0, 2.0)
    from numpy.polynomial.legendre import legval

    for ell in range(0, 2 * nside + 1, 2):
        c = np.zeros(ell + 1)
        c[ell] = 1.0
        s = float(np.sum(mult * geo.npix_ring[:T] * (1.0 + w) * legval(geo.z[:T], c)))
        assert abs(s - (geo.npix if ell == 0 else 0.0)) <= 1e-9 * geo.npix, ell


@pytest.mark.parametrize("use_weights", [True, False])
def test_pixel_weights_total_solid_angle(use_weights):
    C = _c()
    for nside in (2, 8, 16):
        w = C.pixel_weights(nside, use_weights)
        assert w.shape == (12 * nside * nside,)
        assert abs(w.sum() - 4 * np.pi) <= 1e-12 * 4 * np.pi
        assert np.allclose(w, O.pixel_weights(nside, use_weights), rtol=1e-13, atol=0)


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("lmax", [0, 1, 2, 7, 20])
def test_real_dof_isometry(lmax, spin):
    C = _c()
    rng = np.random.default_rng(lmax + 10 * spin)
    K = O.alm_size(lmax)
    ell = O.alm_l_of_index(lmax)
    m = O.alm_m_of_index(lmax)

    def one():
        a = rng.standard_normal(K) + 1j * rng.standard_normal(K)
        a[m == 0] = a[m == 0].real
        a[ell < spin] = 0.0
        return a

    a = one() if spin == 0 else np.stack([one(), one()])
    x = C.alm_to_real(torch.as_tensor(a), lmax, spin)
    n0 = int(np.sum((m == 0) & (ell >= spin)))
    npos = int(np.sum((m > 0) & (ell >= spin)))
    assert x.shape == (C.n_dof(lmax, spin),)
    assert C.n_dof(lmax, spin) == (n0 + 2 * npos) * (1 if spin == 0 else 2)
    G = O.alm_metric_weight(lmax)
    wnorm = float(np.sum(G * np.abs(a) ** 2))
    assert abs(float(torch.sum(x * x)) - wnorm) <= 1e-12 * max(wnorm, 1.0)
    back = C.real_to_alm(x, lmax, spin).numpy()
    assert np.max(np.abs(back - a), initial=0.0) <= 1e-15 * max(1.0, np.max(np.abs(a), initial=0.0))


def test_bandpower_explicit_sum():
    C = _c()
    lmax = 12
    rng = np.random.default_rng(3)
    a = O.random_alm(lmax, rng)
    cl = C.bandpower(torch.as_tensor(a), lmax).numpy()
    ell = O.alm_l_of_index(lmax)
    m = O.alm_m_of_index(lmax)
    for L in range(lmax + 1):
        sel = ell == L
        ref = (np.abs(a[sel & (m == 0)]) ** 2).sum() + 2 * (np.abs(a[sel & (m > 0)]) ** 2).sum()
        assert abs(cl[L] - ref / (2 * L + 1)) <= 1e-14 * max(ref, 1.0)
    a2 = np.stack([a, 2 * a])
    cl2 = C.bandpower(torch.as_tensor(a2), lmax, spin=2).numpy()
    assert np.allclose(cl2[0], cl) and np.allclose(cl2[1], 4 * cl)


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    return Q @ Q.T + n * np.eye(n)


def test_cg_matches_dense_solve():
    C = _c()
    n = 40
    A = torch.as_tensor(_spd(n, 1))
    b = torch.as_tensor(np.random.default_rng(2).standard_normal(n))
    x = C.cg(lambda v, M: M @ v, b, (A,), tol=1e-13, maxiter=500)
    ref = np.linalg.solve(A.numpy(), b.numpy())
    assert np.max(np.abs(x.numpy() - ref)) <= 1e-10 * np.max(np.abs(ref))
    # maxiter = 0 returns the starting point; a converged start stops at once
    assert torch.all(C.cg(lambda v, M: M @ v, b, (A,), maxiter=0) == 0)


def test_cg_implicit_backward_gradcheck():
    C = _c()
    n = 6
    A0 = torch.as_tensor(_spd(n, 4))
    d = torch.rand(n, dtype=torch.float64) + 0.5
    b = torch.randn(n, dtype=torch.float64)
    b.requires_grad_(True)
    d.requires_grad_(True)

    def op(v, D):
        return A0 @ v + D * v

    def f(bb, DD):
        return C.cg(op, bb, (DD,), tol=1e-14, maxiter=200)

    assert torch.autograd.gradcheck(f, (b, d), eps=1e-6, atol=1e-7, rtol=1e-6)
