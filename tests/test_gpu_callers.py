"""GPU checks of the callers (row f) and of the operators' autograd, run on the HIP path.

Gates follow the reference's tests: analysis against the oracle's analysis (1e-12) and
its round trip at lmax = nside (``tests/test_analysis.py``); the real-DOF adjoint identity
(OP_TOL 1e-10, ``tests/test_wiener.py:234``); the Wiener mean against a dense solve built
from the oracle's synthesis matrix (SOLVE_TOL 1e-8, ``:250-267``); Wiener -> deconvolve
limits (``:347-369``); constrained-realization moments within a 4-sigma Monte-Carlo budget
(``:374-419``); finite-difference gradients through CG (FD_TOL 1e-2, ``:422-446``); and
``torch.autograd.gradcheck`` of both operators (finite differences, the reference's
``tests/test_grad.py:176-194`` at eps 1e-6).
"""

from __future__ import annotations

import warnings

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O

pytestmark = pytest.mark.gpu

SOLVE_TOL = 1e-8
LIMIT_TOL = 1e-7
FD_TOL = 1e-2


def _j():
    import jht_amd

    return jht_amd


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _alm(lmax, spin, seed):
    rng = np.random.default_rng(seed)
    if spin == 0:
        return O.random_alm(lmax, rng)
    return np.stack([O.random_alm(lmax, rng, 2), O.random_alm(lmax, rng, 2)])


def _theta(nside):
    geo = O.RingInfo(nside)
    return np.repeat(np.arccos(geo.z), geo.npix_ring)


def _cap(nside, th_cut):
    m = np.ones(12 * nside * nside)
    m[_theta(nside) < th_cut] = 0.0
    return m


def _signal_cl(lmax, spin):
    base = 1.0 / (np.arange(lmax + 1) + 1.0) ** 2 + 1e-3
    return base if spin == 0 else (base, 0.5 * base)


def _dense_S(nside, lmax, spin):
    """Synthesis matrix in the real-DOF basis from the oracle: columns S(T^-1 e_k)."""
    j = _j()
    nx = j.n_dof(lmax, spin)
    cols = []
    for k in range(nx):
        e = torch.zeros(nx, dtype=torch.float64)
        e[k] = 1.0
        a = j.real_to_alm(e, lmax, spin).numpy()
        cols.append(np.asarray(O.synthesis(a, nside, lmax, spin)).ravel())
    return np.stack(cols, axis=1)


def _prior_diag(cl, lmax, spin):
    ell, m = O.alm_l_of_index(lmax), O.alm_m_of_index(lmax)
    act = ell >= spin
    e = np.concatenate([ell[act & (m == 0)], ell[act & (m > 0)], ell[act & (m > 0)]])

    def one(c):
        return 1.0 / np.asarray(c)[e]

    return one(cl) if spin == 0 else np.concatenate([one(cl[0]), one(cl[1])])


@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("use_weights", [True, False])
@pytest.mark.parametrize("niter", [0, 3])
def test_analysis_matches_oracle(spin, use_weights, niter):
    j = _j()
    nside, lmax = 16, 24
    rng = np.random.default_rng(5 + spin)
    v = rng.standard_normal(12 * nside * nside if spin == 0 else (2, 12 * nside * nside))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        got = j.analysis(torch.tensor(v, device="cuda"), nside, lmax, spin, niter, use_weights).cpu().numpy()
    ref = O.analysis(v, nside, lmax, spin, niter, use_weights)
    print(f"analysis spin={spin} w={use_weights} niter={niter}: rel {_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) <= 1e-12


@pytest.mark.parametrize("spin", [0, 2])
def test_analysis_round_trip(spin):
    """Band-limited map at lmax = nside: weighted niter=3 recovers a to ~1e-13
    (``docs/accuracy.md``: the weighted round trip at lmax = nside)."""
    j = _j()
    nside = lmax = 32
    a = _alm(lmax, spin, 11)
    m = j.synthesis(torch.tensor(a, device="cuda"), nside, lmax, spin)
    back = j.analysis(m, nside, lmax, spin, niter=3).cpu().numpy()
    err = _rel(back, a)
    print(f"round trip spin={spin}: {err:.2e}", flush=True)
    assert err <= 1e-10
    bare = j.analysis(m, nside, lmax, spin, niter=0).cpu().numpy()
    assert _rel(bare, a) > err  # iterations improve on the bare estimator


@pytest.mark.parametrize("spin", [0, 2])
def test_real_dof_adjoint_identity_and_grad(spin):
    j = _j()
    nside, lmax = 8, 12
    rng = np.random.default_rng(21 + spin)
    nx = j.n_dof(lmax, spin)
    x = torch.tensor(rng.standard_normal(nx), device="cuda", requires_grad=True)
    v = torch.tensor(rng.standard_normal(12 * nside * nside if spin == 0 else (2, 12 * nside * nside)), device="cuda")
    y = j.synthesis_real(x, nside, lmax, spin)
    lhs = float((y * v).sum())
    Stv = j.alm_to_real(j.adjoint_synthesis(v, nside, lmax, spin), lmax, spin)
    rhs = float((x.detach() * Stv).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    (y * v).sum().backward()  # VJP of the real-linear map is its Euclidean transpose
    assert torch.allclose(x.grad, Stv, rtol=0, atol=1e-12 * float(Stv.abs().max()))


@pytest.mark.parametrize("spin", [0, 2])
def test_operator_gradcheck(spin):
    """Finite differences (eps 1e-6) against both backward rules, complex and real inputs."""
    j = _j()
    nside, lmax = 2, 3
    a = torch.tensor(_alm(lmax, spin, 31 + spin), device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: j.synthesis(t, nside, lmax, spin), (a,), eps=1e-6, atol=1e-7, rtol=1e-6)
    v = torch.tensor(np.random.default_rng(40 + spin).standard_normal(
        48 if spin == 0 else (2, 48)), device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: j.adjoint_synthesis(t, nside, lmax, spin), (v,), eps=1e-6, atol=1e-7, rtol=1e-6)


@pytest.mark.parametrize("spin", [0, 2])
def test_wiener_mean_matches_dense_solve(spin):
    j = _j()
    nside, lmax = 4, 6
    mask = _cap(nside, 0.4)
    ninv = mask * (0.5 + np.random.default_rng(9).random(12 * nside * nside))
    a_true = _alm(lmax, spin, 3)
    m = np.asarray(O.synthesis(a_true, nside, lmax, spin))
    cl = _signal_cl(lmax, spin)
    a_w = j.wiener(torch.tensor(m, device="cuda"), cl, nside, lmax, spin, inv_noise=ninv,
                   max_iter=500, tol=1e-12).cpu().numpy()
    S = _dense_S(nside, lmax, spin)
    nflat = ninv if spin == 0 else np.tile(ninv, 2)
    A = S.T @ (nflat[:, None] * S) + np.diag(_prior_diag(cl, lmax, spin))
    b = S.T @ (nflat * m.ravel())
    x = np.linalg.solve(A, b)
    a_d = j.real_to_alm(torch.tensor(x), lmax, spin).numpy()
    err = np.max(np.abs(a_w - a_d))
    print(f"wiener spin={spin}: |a - a_dense| = {err:.2e}", flush=True)
    assert err <= SOLVE_TOL * (1.0 + np.max(np.abs(a_d)))


@pytest.mark.parametrize("spin", [0, 2])
def test_wiener_limits_to_deconvolve(spin):
    j = _j()
    nside, lmax = 8, 8
    mask = _cap(nside, 0.4)
    a_true = _alm(lmax, spin, 7)
    m = np.asarray(O.synthesis(a_true, nside, lmax, spin))
    mm = torch.tensor(m * (mask if spin == 0 else mask[None, :]), device="cuda")

    def const_cl(val):
        c = np.full(lmax + 1, val)
        return c if spin == 0 else (c, c)

    a_w = j.wiener(mm, const_cl(1e14), nside, lmax, spin, mask=mask, max_iter=800, tol=1e-12).cpu().numpy()
    a_d = j.deconvolve(mm, mask, nside, lmax, spin, reg=0.0, max_iter=800, tol=1e-12).cpu().numpy()
    assert np.max(np.abs(a_w - a_d)) <= LIMIT_TOL * (1.0 + np.max(np.abs(a_d)))
    reg = 0.5
    a_w2 = j.wiener(mm, const_cl(1.0 / reg), nside, lmax, spin, mask=mask, max_iter=800, tol=1e-12).cpu().numpy()
    a_d2 = j.deconvolve(mm, mask, nside, lmax, spin, reg=reg, max_iter=800, tol=1e-12).cpu().numpy()
    assert np.max(np.abs(a_w2 - a_d2)) <= SOLVE_TOL * (1.0 + np.max(np.abs(a_d2)))


def test_deconvolve_recovers_band_limited_truth():
    j = _j()
    nside, lmax = 16, 12
    mask = _cap(nside, 0.3)
    a_true = _alm(lmax, 0, 12)
    m = torch.tensor(np.asarray(O.synthesis(a_true, nside, lmax, 0)), device="cuda")
    a = j.deconvolve(m * torch.tensor(mask, device="cuda"), mask, nside, lmax, 0, max_iter=400, tol=1e-12).cpu().numpy()
    assert _rel(a, a_true) <= 1e-6


def test_constrained_realization_moments():
    """Sample mean -> Wiener mean, sample covariance -> (A_x + D)^-1, 4-sigma budgets."""
    j = _j()
    nside, lmax, spin = 8, 4, 0
    n_draws = 400
    mask = _cap(nside, 0.4)
    ninv = mask * 1.0
    cl = _signal_cl(lmax, spin)
    m = torch.tensor(np.asarray(O.synthesis(_alm(lmax, spin, 8), nside, lmax, spin)), device="cuda")
    S = _dense_S(nside, lmax, spin)
    A = S.T @ (ninv[:, None] * S) + np.diag(_prior_diag(cl, lmax, spin))
    Sigma = np.linalg.inv(A)
    x_hat = np.linalg.solve(A, S.T @ (ninv * m.cpu().numpy()))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1234)
    X = np.stack([j.alm_to_real(j.constrained_realization(m, cl, nside, lmax, gen, spin, inv_noise=ninv,
                                                          max_iter=200, tol=1e-10), lmax, spin).cpu().numpy()
                  for _ in range(n_draws)])
    mean_err = float(np.linalg.norm(X.mean(0) - x_hat))
    assert mean_err <= 4.0 * np.sqrt(np.trace(Sigma) / n_draws)
    cov_err = float(np.linalg.norm(np.cov(X.T) - Sigma, "fro"))
    assert cov_err <= 4.0 * np.sqrt((np.trace(Sigma) ** 2 + np.linalg.norm(Sigma, "fro") ** 2) / n_draws)


def test_wiener_differentiable():
    j = _j()
    nside, lmax, spin = 8, 6, 0
    mask = _cap(nside, 0.4)
    ninv = torch.tensor(mask * (0.5 + np.random.default_rng(9).random(12 * nside * nside)), device="cuda")
    m = torch.tensor(np.asarray(O.synthesis(_alm(lmax, spin, 10), nside, lmax, spin)), device="cuda")
    cl0 = torch.tensor(_signal_cl(lmax, spin), device="cuda")

    def loss(cl, mm=m):
        a = j.wiener(mm, cl, nside, lmax, spin, inv_noise=ninv, max_iter=200, tol=1e-11)
        return (a.real ** 2 + a.imag ** 2).sum()

    c = cl0.clone().requires_grad_(True)
    g = torch.autograd.grad(loss(c), c)[0].cpu().numpy()
    assert np.all(np.isfinite(g))
    for i in (1, 3):
        h = 1e-4 * float(cl0[i])
        cp, cm = cl0.clone(), cl0.clone()
        cp[i] += h
        cm[i] -= h
        fd = (float(loss(cp)) - float(loss(cm))) / (2.0 * h)
        print(f"wiener grad cl[{i}]: {g[i]:.6e} fd {fd:.6e}", flush=True)
        assert abs(g[i] - fd) <= FD_TOL * abs(fd) + 1e-6
    mm = m.clone().requires_grad_(True)
    gm = torch.autograd.grad(loss(cl0, mm), mm)[0]
    assert torch.isfinite(gm).all()
    # data gradient against a finite difference along a random direction
    d = torch.tensor(np.random.default_rng(2).standard_normal(m.shape), device="cuda")
    h = 1e-3
    fd = (float(loss(cl0, m + h * d)) - float(loss(cl0, m - h * d))) / (2 * h)
    assert abs(float((gm * d).sum()) - fd) <= FD_TOL * abs(fd) + 1e-6


def test_pseudo_alm_is_analysis_of_masked_map():
    j = _j()
    nside, lmax = 16, 16
    mask = _cap(nside, 0.5)
    m = torch.tensor(np.random.default_rng(1).standard_normal(12 * nside * nside), device="cuda")
    got = j.pseudo_alm(m, mask, nside, lmax, 0, niter=0, use_weights=False).cpu().numpy()
    ref = (4 * np.pi / (12 * nside * nside)) * O.adjoint_synthesis(m.cpu().numpy() * mask, nside, lmax, 0)
    assert _rel(got, ref) <= 1e-12
