"""Callers of the transform pair, run unchanged on the MI355X operators.

Torch host compositions over :func:`jht_amd.synthesis` / :func:`jht_amd.adjoint_synthesis`,
with jht's names, argument meaning and results:

* ring quadrature weights and the iterated analysis -- ``src/jht/weights.py:60-106``,
  ``src/jht/_analysis.py:49-63``;
* the real-DOF isometry ``T`` and the real-DOF transforms / band power --
  ``src/jht/masked.py:117-207``, ``src/jht/diff.py:55-138``;
* masked pseudo-a_lm, cut-sky deconvolution, the Wiener filter and constrained
  realizations solved by conjugate gradients in the reference's coordinates --
  ``src/jht/masked.py:77-439``.

All heavy work is the two HIP operators; the solvers add only vector updates on the GPU.
The CG solve is a ``torch.autograd.Function`` with an implicit backward (one more solve
with the symmetric operator), so gradients flow to the data and to the prior spectrum.
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np
import torch

from .healpix import (
    RingInfo,
    _prep,
    adjoint_synthesis,
    alm_metric_weight,
    alm_size,
    synthesis,
)

__all__ = [
    "ring_weights",
    "pixel_weights",
    "bare_analysis",
    "analysis",
    "map2alm",
    "n_dof",
    "alm_to_real",
    "real_to_alm",
    "synthesis_real",
    "analysis_real",
    "bandpower",
    "pseudo_alm",
    "deconvolve",
    "wiener",
    "constrained_realization",
    "cg",
]

_SQRT2 = float(np.sqrt(2.0))


# --------------------------------------------------------------------------- #
# ring weights -- src/jht/weights.py:60-106
# --------------------------------------------------------------------------- #
@lru_cache(maxsize=None)
def ring_weights(nside: int) -> np.ndarray:
    """Per-ring weight deviations ``w_i`` over the north half + equator (``2 nside``).

    Minimum-norm solution of ``sum_i c_i n_i (1 + w_i) P_l(z_i) = Npix delta_l0`` for even
    ``l <= 2 nside`` (``c_i = 2`` for N/S pairs, 1 for the equator): the m = 0 colatitude
    quadrature is exact for those Legendre polynomials (``weights.py:60-83``).
    """
    from numpy.polynomial.legendre import legvander

    geo = RingInfo(nside)
    T = 2 * nside
    z = geo.z[:T]
    npr = geo.npix_ring[:T].astype(np.float64)
    mult = np.full(T, 2.0)
    mult[-1] = 1.0
    P_even = legvander(z, 2 * nside)[:, ::2]
    A = ((mult * npr)[:, None] * P_even).T
    rhs = -A.sum(axis=1)
    rhs[0] += geo.npix
    w, *_ = np.linalg.lstsq(A, rhs, rcond=None)
    w.setflags(write=False)
    return w


@lru_cache(maxsize=None)
def _pixel_weights_np(nside: int, use_weights: bool) -> np.ndarray:
    geo = RingInfo(nside)
    base = 4.0 * np.pi / geo.npix
    if not use_weights:
        out = np.full(geo.npix, base)
    else:
        w = ring_weights(nside)
        r = np.arange(geo.nrings)
        north = np.where(r < 2 * nside, r, 4 * nside - 2 - r)
        out = base * (1.0 + np.repeat(w[north], geo.npix_ring))
    out.setflags(write=False)
    return out


@lru_cache(maxsize=None)
def _pixel_weights_dev(nside: int, use_weights: bool, device: str) -> torch.Tensor:
    return torch.as_tensor(_pixel_weights_np(nside, use_weights).copy(), device=device)


def pixel_weights(nside: int, use_weights: bool = True, device=None):
    """Per-pixel quadrature weights in RING order: ``(4pi/Npix)(1 + w_ring)``, or the
    uniform ``4pi/Npix`` with ``use_weights=False`` (``weights.py:86-106``).  numpy on the
    host, or a cached float64 tensor when ``device`` is given."""
    if device is None:
        return _pixel_weights_np(nside, use_weights)
    return _pixel_weights_dev(nside, use_weights, str(torch.device(device)))


# --------------------------------------------------------------------------- #
# analysis -- src/jht/_analysis.py:49-63
# --------------------------------------------------------------------------- #
def _maps_dev(maps) -> torch.Tensor:
    """Maps as a float64 GPU tensor (autograd history is kept)."""
    return _prep(maps, False)[0]


def _wvec(nside: int, spin: int, use_weights: bool, device) -> torch.Tensor:
    w = pixel_weights(nside, use_weights, device)
    return w if spin == 0 else w[None, :]


def bare_analysis(maps, nside: int, lmax: int, spin: int = 0, use_weights: bool = True) -> torch.Tensor:
    """``A0 = S^T W``: the weighted adjoint, the non-iterated estimator (``_analysis.py:49-52``)."""
    v = _maps_dev(maps)
    return adjoint_synthesis(_wvec(nside, spin, use_weights, v.device) * v, nside, lmax, spin)


def analysis(maps, nside: int, lmax: int, spin: int = 0, niter: int = 3,
             use_weights: bool = True) -> torch.Tensor:
    """Approximate inverse of :func:`synthesis` by ``niter`` Jacobi steps
    ``a <- a + A0 (map - S a)`` (``_analysis.py:55-63``); ``niter = 0`` is the bare estimator."""
    v = _maps_dev(maps)
    a = bare_analysis(v, nside, lmax, spin, use_weights)
    for _ in range(niter):
        a = a + bare_analysis(v - synthesis(a, nside, lmax, spin), nside, lmax, spin, use_weights)
    return a


map2alm = analysis


# --------------------------------------------------------------------------- #
# the real-DOF isometry T -- src/jht/masked.py:117-207
# --------------------------------------------------------------------------- #
@lru_cache(maxsize=None)
def _dof_layout(lmax: int, spin: int):
    """Flat indices of the active ``m = 0`` and ``m > 0`` modes (``l >= |spin|``) and the
    per-channel DOF count ``n0 + 2 n+`` (``masked.py:117-139``)."""
    ms = np.concatenate([np.full(lmax + 1 - m, m) for m in range(lmax + 1)])
    ls = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    active = ls >= abs(spin)
    idx = np.arange(alm_size(lmax))
    k0 = idx[active & (ms == 0)]
    kp = idx[active & (ms > 0)]
    ell = np.concatenate([ls[active & (ms == 0)], ls[active & (ms > 0)], ls[active & (ms > 0)]])
    return k0, kp, int(k0.size + 2 * kp.size), ell


@lru_cache(maxsize=None)
def _dof_index(lmax: int, spin: int, device: str):
    k0, kp, nx, ell = _dof_layout(lmax, spin)
    d = torch.device(device)
    return torch.as_tensor(k0, device=d), torch.as_tensor(kp, device=d), nx, torch.as_tensor(ell, device=d)


def n_dof(lmax: int, spin: int = 0) -> int:
    """Length of the real DOF vector (both E and B channels for spin 2)."""
    nx = _dof_layout(lmax, spin)[2]
    return nx if spin == 0 else 2 * nx


def _as_tensor(x, dtype) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x if x.dtype == dtype else x.to(dtype)
    return torch.as_tensor(np.asarray(x)).to(dtype)


def alm_to_real(alm, lmax: int, spin: int = 0) -> torch.Tensor:
    """``T``: healpy a_lm -> real ``x = [a_l0, sqrt2 Re a_lm>0, sqrt2 Im a_lm>0]`` per channel,
    so that ``||x||_2 = ||a||_w`` (``masked.py:148-165``).  Works on any device."""
    a = _as_tensor(alm, torch.complex128)
    k0, kp, _, _ = _dof_index(lmax, spin, str(a.device))

    def one(c):
        return torch.cat([c[k0].real, _SQRT2 * c[kp].real, _SQRT2 * c[kp].imag])

    return one(a) if spin == 0 else torch.cat([one(a[0]), one(a[1])])


def real_to_alm(x, lmax: int, spin: int = 0) -> torch.Tensor:
    """``T^-1``: real DOF vector -> healpy a_lm, ``l < |spin|`` modes zero (``masked.py:168-176``)."""
    xv = _as_tensor(x, torch.float64)
    k0, kp, nx, _ = _dof_index(lmax, spin, str(xv.device))
    K = alm_size(lmax)
    n0, npos = k0.numel(), kp.numel()

    def one(v):
        re = torch.zeros(K, dtype=torch.float64, device=v.device)
        im = torch.zeros(K, dtype=torch.float64, device=v.device)
        re = re.index_put((k0,), v[:n0]).index_put((kp,), v[n0:n0 + npos] / _SQRT2)
        im = im.index_put((kp,), v[n0 + npos:n0 + 2 * npos] / _SQRT2)
        return torch.complex(re, im)

    if spin == 0:
        return one(xv)
    return torch.stack([one(xv[:nx]), one(xv[nx:])])


# --------------------------------------------------------------------------- #
# real-DOF transforms and band power -- src/jht/diff.py:55-138
# --------------------------------------------------------------------------- #
def synthesis_real(x, nside: int, lmax: int, spin: int = 0) -> torch.Tensor:
    """``S o T^-1``: real DOF vector -> map, a plain real-linear map (``diff.py:55-63``)."""
    xv = _as_tensor(x, torch.float64)
    if not xv.is_cuda:
        xv = xv.to("cuda")
    return synthesis(real_to_alm(xv, lmax, spin), nside, lmax, spin)


def analysis_real(maps, nside: int, lmax: int, spin: int = 0, niter: int = 3,
                  use_weights: bool = True) -> torch.Tensor:
    """``T o analysis``: map -> real DOF vector (``diff.py:66-75``)."""
    return alm_to_real(analysis(maps, nside, lmax, spin, niter, use_weights), lmax, spin)


@lru_cache(maxsize=None)
def _ell_norm(lmax: int, device: str):
    ells = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    d = torch.device(device)
    return (torch.as_tensor(ells, device=d), torch.as_tensor(alm_metric_weight(lmax).copy(), device=d),
            1.0 / (2.0 * torch.arange(lmax + 1, device=d, dtype=torch.float64) + 1.0))


def bandpower(alm, lmax: int, spin: int = 0) -> torch.Tensor:
    """``C_l = (1/(2l+1)) sum_m (2 - delta_m0) |a_lm|^2``; ``(lmax+1,)`` or ``(2, lmax+1)``
    for (EE, BB) (``diff.py:120-138``)."""
    a = _as_tensor(alm, torch.complex128)
    ells, G, norm = _ell_norm(lmax, str(a.device))

    def one(c):
        p = G * (c.real ** 2 + c.imag ** 2)
        cl = torch.zeros(lmax + 1, dtype=torch.float64, device=c.device).index_add(0, ells, p)
        return cl * norm

    return one(a) if spin == 0 else torch.stack([one(a[0]), one(a[1])])


# --------------------------------------------------------------------------- #
# conjugate gradients with an implicit backward
# --------------------------------------------------------------------------- #
def _cg_iterate(apply_A, b: torch.Tensor, x0, tol: float, maxiter: int):
    """Plain CG, stopping when ``|r|^2 <= tol^2 |b|^2`` or after ``maxiter`` steps: the rule
    of ``jax.scipy.sparse.linalg.cg`` with ``atol = 0`` and no preconditioner (the solver
    the reference calls, ``masked.py:302,356,438``)."""
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = b - apply_A(x) if x0 is not None else b.clone()
    p = r.clone()
    gamma = torch.dot(r, r)
    stop = (tol * tol) * torch.dot(b, b)
    k = 0
    # one host read per iteration decides termination (the solves are operator-bound)
    while k < maxiter and bool(gamma > stop):
        Ap = apply_A(p)
        alpha = gamma / torch.dot(p, Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        gnew = torch.dot(r, r)
        p = r + (gnew / gamma) * p
        gamma = gnew
        k += 1
    return x, k


class _SPDSolve(torch.autograd.Function):
    """``y = A(theta)^{-1} b`` for a symmetric positive-definite operator.

    Backward: ``lam = A^{-1} g`` (one more CG solve); ``grad_b = lam`` and, for the operator
    parameters theta, ``grad_theta = -d/dtheta <lam, A(theta) y>`` with y, lam held fixed.
    """

    @staticmethod
    def forward(ctx, b, x0, tol, maxiter, op, *params):
        with torch.no_grad():
            y, iters = _cg_iterate(lambda v: op(v, *params), b, x0, tol, maxiter)
        ctx.op, ctx.tol, ctx.maxiter = op, tol, maxiter
        ctx.save_for_backward(y, *params)
        ctx.iters = iters
        return y

    @staticmethod
    def backward(ctx, g):
        y, *params = ctx.saved_tensors
        op = ctx.op
        with torch.no_grad():
            lam, _ = _cg_iterate(lambda v: op(v, *params), g.contiguous(), None, ctx.tol, ctx.maxiter)
        grads = [None] * len(params)
        need = [i for i, p in enumerate(params) if ctx.needs_input_grad[5 + i]]
        if need:
            with torch.enable_grad():
                ps = [p.detach().requires_grad_(i in need) for i, p in enumerate(params)]
                s = -torch.dot(lam, op(y, *ps))
                got = torch.autograd.grad(s, [ps[i] for i in need], allow_unused=True)
            for i, gi in zip(need, got):
                grads[i] = gi
        grad_b = lam if ctx.needs_input_grad[0] else None
        return (grad_b, None, None, None, None, *grads)


def cg(op, b: torch.Tensor, params=(), x0=None, tol: float = 1e-8, maxiter: int = 200) -> torch.Tensor:
    """Solve ``op(x, *params) = b`` by CG (``op`` symmetric positive definite in the plain
    Euclidean product), differentiable in ``b`` and in the tensors ``params``."""
    return _SPDSolve.apply(b, x0, tol, maxiter, op, *params)


# --------------------------------------------------------------------------- #
# masked operators -- src/jht/masked.py:77-371
# --------------------------------------------------------------------------- #
def _weight_pix(nside: int, spin: int, device, *, inv_noise=None, mask=None,
                use_weights: bool = True) -> torch.Tensor:
    """Per-pixel weight: ``inv_noise`` (N^-1, mask folded in) if given, else ``mask * W``
    (``masked.py:77-96``)."""
    if inv_noise is not None:
        w = _as_tensor(inv_noise, torch.float64).to(device)
    elif mask is not None:
        w = _as_tensor(mask, torch.float64).to(device) * pixel_weights(nside, use_weights, device)
    else:
        raise ValueError("provide either inv_noise (per-pixel N^-1) or mask")
    return w if spin == 0 else w[None, :]


def _prior_sqrt(signal_cl, lmax: int, spin: int, device) -> torch.Tensor:
    """x-space whitening ``pw = sqrt(Cl)`` per DOF, exactly 0 where ``Cl <= 0``
    (``masked.py:210-231``); differentiable in Cl on the positive entries."""
    ell = _dof_index(lmax, spin, str(torch.device(device)))[3]

    def one(cl):
        c = _as_tensor(cl, torch.float64).to(device)
        if tuple(c.shape) != (lmax + 1,):
            raise ValueError(
                f"signal_cl channel must have shape (lmax+1,) = ({lmax + 1},), got {tuple(c.shape)}")
        c = c[ell]
        pos = c > 0
        return torch.where(pos, torch.sqrt(torch.where(pos, c, torch.ones_like(c))), torch.zeros_like(c))

    if spin == 0:
        return one(signal_cl)
    c_ee, c_bb = signal_cl
    return torch.cat([one(c_ee), one(c_bb)])


def _apply_Ax(x: torch.Tensor, w_pix: torch.Tensor, nside: int, lmax: int, spin: int) -> torch.Tensor:
    """``A_x x = T S^T diag(w) S T^-1 x`` (``masked.py:234-247``)."""
    a = real_to_alm(x, lmax, spin)
    return alm_to_real(adjoint_synthesis(w_pix * synthesis(a, nside, lmax, spin), nside, lmax, spin), lmax, spin)


def _rhs_x(maps: torch.Tensor, w_pix: torch.Tensor, nside: int, lmax: int, spin: int) -> torch.Tensor:
    """``b_x = T S^T (w m)`` (``masked.py:277-279``)."""
    return alm_to_real(adjoint_synthesis(w_pix * maps, nside, lmax, spin), lmax, spin)


def pseudo_alm(maps, mask, nside: int, lmax: int, spin: int = 0, niter: int = 3,
               use_weights: bool = True) -> torch.Tensor:
    """Masked pseudo-a_lm: :func:`analysis` of the zero-filled map ``M * map``
    (``masked.py:101-114``)."""
    v = _maps_dev(maps)
    msk = _as_tensor(mask, torch.float64).to(v.device)
    return analysis((msk if spin == 0 else msk[None, :]) * v, nside, lmax, spin, niter, use_weights)


def deconvolve(maps, mask, nside: int, lmax: int, spin: int = 0, *, max_iter: int = 200,
               tol: float = 1e-8, reg: float = 0.0, use_weights: bool = True, x0=None) -> torch.Tensor:
    """Cut-sky deconvolution: CG on ``(S^T M W S + reg) a = S^T M W m`` in the real-DOF
    metric (``masked.py:285-318``)."""
    v = _maps_dev(maps)
    w_pix = _weight_pix(nside, spin, v.device, mask=mask, use_weights=use_weights)
    b_x = _rhs_x(v, w_pix, nside, lmax, spin)

    def op(x, w):
        return _apply_Ax(x, w, nside, lmax, spin) + reg * x

    x_init = None if x0 is None else alm_to_real(_prep(x0, True)[0], lmax, spin)
    x = cg(op, b_x, (w_pix,), x0=x_init, tol=tol, maxiter=max_iter)
    return real_to_alm(x, lmax, spin)


def _whitened_op(y, w_pix, pw, nside: int, lmax: int, spin: int) -> torch.Tensor:
    """``(P A_x P + I) y``, ``P = diag(pw)`` (``masked.py:250-268``)."""
    return pw * _apply_Ax(pw * y, w_pix, nside, lmax, spin) + y


def _whiten_x0(x0, pw, lmax: int, spin: int) -> torch.Tensor:
    xv = alm_to_real(_prep(x0, True)[0], lmax, spin)
    return torch.where(pw > 0, xv / torch.where(pw > 0, pw, torch.ones_like(pw)), torch.zeros_like(xv))


def wiener(maps, signal_cl, nside: int, lmax: int, spin: int = 0, *, inv_noise=None, mask=None,
           max_iter: int = 200, tol: float = 1e-8, use_weights: bool = True, x0=None) -> torch.Tensor:
    """Wiener-filter mean: ``(S^T N^-1 S + C^-1) a = S^T N^-1 m`` solved by CG in
    prior-whitened coordinates ``x = P y`` (``masked.py:324-371``).  Differentiable in
    ``maps`` and ``signal_cl``."""
    v = _maps_dev(maps)
    w_pix = _weight_pix(nside, spin, v.device, inv_noise=inv_noise, mask=mask, use_weights=use_weights)
    pw = _prior_sqrt(signal_cl, lmax, spin, v.device)
    b_y = pw * _rhs_x(v, w_pix, nside, lmax, spin)

    def op(y, w, p):
        return _whitened_op(y, w, p, nside, lmax, spin)

    y_init = None if x0 is None else _whiten_x0(x0, pw.detach(), lmax, spin)
    y = cg(op, b_y, (w_pix, pw), x0=y_init, tol=tol, maxiter=max_iter)
    return real_to_alm(pw * y, lmax, spin)


def constrained_realization(maps, signal_cl, nside: int, lmax: int, key=None, spin: int = 0, *,
                            inv_noise=None, mask=None, max_iter: int = 200, tol: float = 1e-8,
                            use_weights: bool = True, x0=None) -> torch.Tensor:
    """One posterior draw ``a ~ N(a_wiener, (S^T N^-1 S + C^-1)^-1)`` (``masked.py:374-439``).

    ``key`` is a ``torch.Generator`` on the maps' device or an integer seed (the reference
    takes a ``jax.random`` key; its threefry draws are not reproduced, only their
    distribution).  The whitened RHS gains ``P T S^T(sqrt(w) omega) + xi`` on the modes
    with ``Cl > 0``."""
    v = _maps_dev(maps)
    w_pix = _weight_pix(nside, spin, v.device, inv_noise=inv_noise, mask=mask, use_weights=use_weights)
    pw = _prior_sqrt(signal_cl, lmax, spin, v.device)
    b_x = _rhs_x(v, w_pix, nside, lmax, spin)
    if isinstance(key, torch.Generator):
        gen = key
    else:
        gen = torch.Generator(device=v.device)
        gen.manual_seed(0 if key is None else int(key))
    omega = torch.randn(v.shape, generator=gen, dtype=torch.float64, device=v.device)
    s1 = alm_to_real(adjoint_synthesis(torch.sqrt(w_pix) * omega, nside, lmax, spin), lmax, spin)
    xi = torch.randn(b_x.shape, generator=gen, dtype=torch.float64, device=v.device)
    rhs = pw * (b_x + s1) + torch.where(pw > 0, xi, torch.zeros_like(xi))

    def op(y, w, p):
        return _whitened_op(y, w, p, nside, lmax, spin)

    y_init = None if x0 is None else _whiten_x0(x0, pw.detach(), lmax, spin)
    y = cg(op, rhs, (w_pix, pw), x0=y_init, tol=tol, maxiter=max_iter)
    return real_to_alm(pw * y, lmax, spin)
