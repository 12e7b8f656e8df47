"""Off-grid spherical-harmonic synthesis at arbitrary points on the HIP kernels.

``synthesis_general(alm, loc, spin=, lmax=, epsilon=)`` evaluates the band-limited field
at points ``loc = (theta, phi)``; ``adjoint_synthesis_general`` is its exact transpose in
the ``(2 - delta_m0)`` metric.  Same signatures, errors and conventions as
``src/jht/offgrid.py:165-243``; the accuracy tier is the NUFFT ``epsilon`` (default
1e-10), not the on-grid 1e-12.

Pipeline (``src/jht/offgrid.py:124-160``, ``src/jht/_nufft.py:154-181``), FP64 throughout:

1. Legendre step on a Clenshaw-Curtis ring set (both poles) -- the library's recursion
   kernels on a ring-set plan (``jhtx_plan_create_rings``, ``jhtx_legendre_synth``);
2. the double-Fourier-sphere extension to a 2 pi-periodic theta (``jhtx_dfs_extend``);
3. an FFT along theta (library FFT);
4. the deconvolved oversampled grid (``jhtx_nufft_grid_fill``) and its 2-D inverse FFT
   (library FFT);
5. the exponential-of-semicircle stencil sum at every point (``jhtx_nufft_interp``: points
   sorted into grid tiles, each tile and its halo staged in LDS once), which also returns
   d/dtheta and d/dphi for the pointing gradient.

The adjoint runs the transposes in reverse (``jhtx_nufft_spread``: points sorted into grid
tiles, each tile accumulated in LDS, then added to the grid; 2-D
FFT, ``jhtx_nufft_grid_gather``, inverse theta FFT, ``jhtx_dfs_fold``,
``jhtx_legendre_adjoint``).  Both functions are differentiable in ``alm`` / ``field`` and
in ``loc`` through torch.autograd.  Spins 0, 1, 2 and 3; for s > 0 the -s channel carries
(-1)^s (``src/jht/offgrid.py:98-121``).
"""

from __future__ import annotations

import threading

import numpy as np
import torch

from . import _lib
from .healpix import _adjoint_backward_alm, _metric_tensors, _prep, _synth_backward, alm_size

TWO_PI = 2.0 * np.pi

# requested accuracy -> (ES kernel width W, oversampling sigma): the reference's table of
# achievable accuracies (src/jht/_nufft.py:33-39)
_KERNELS = {1e-4: (6, 2.0), 1e-6: (8, 2.0), 1e-8: (10, 2.0), 1e-10: (14, 2.0), 1e-12: (16, 2.0)}


def _kernel_params(eps: float):
    """Cheapest (W, sigma) whose accuracy meets eps; the most accurate if none does."""
    ok = [e for e in _KERNELS if e <= eps * 1.0000001]
    return _KERNELS[max(ok)] if ok else _KERNELS[min(_KERNELS)]


def _smooth_even(n: float) -> int:
    """Smallest even 2,3,5-smooth integer >= n (an FFT-friendly size)."""
    n = int(np.ceil(n))
    n += n % 2
    while True:
        r = n
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return n
        n += 2


def _es_correction(N: int, n: int, W: int, beta: float) -> np.ndarray:
    """Fourier transform of the ES kernel at the N centred modes of an n-point grid
    (Gauss-Legendre, 2W + 24 nodes over the support)."""
    ks = np.arange(-(N // 2), -(N // 2) + N, dtype=np.float64)
    t, wq = np.polynomial.legendre.leggauss(2 * W + 24)
    t, wq = t * (W / 2.0), wq * (W / 2.0)
    psi = np.exp(beta * (np.sqrt(np.clip(1.0 - (2.0 * t / W) ** 2, 0.0, 1.0)) - 1.0))
    return (psi[None, :] * np.cos(ks[:, None] * (TWO_PI / n) * t[None, :]) * wq[None, :]).sum(1)


class OffgridPlan:
    """Device tables of one (lmax, spin, epsilon): the Clenshaw-Curtis ring-set plan and
    the NUFFT grid sizes, kernel parameters and deconvolution factors."""

    def __init__(self, lmax: int, spin: int, epsilon: float, device=None):
        lib = _lib.load()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.lmax, self.spin, self.epsilon = lmax, spin, epsilon
        self.ncomp = 1 if spin == 0 else 2
        self.K = alm_size(lmax)
        self.ns = _smooth_even(lmax + 1) + 1  # Clenshaw-Curtis rings on [0, pi]
        theta = np.arange(self.ns) * np.pi / (self.ns - 1)
        north = np.ascontiguousarray(theta[: (self.ns + 1) // 2])
        self.Nk, self.Nm = 2 * self.ns - 2, 2 * lmax + 1
        self.W, sigma = _kernel_params(epsilon)
        self.nk, self.nm = _smooth_even(sigma * self.Nk), _smooth_even(sigma * self.Nm)
        self.beta = float(np.pi * self.W * (1.0 - 0.5 / sigma))
        corr_k = _es_correction(self.Nk, self.nk, self.W, self.beta)
        corr_m = _es_correction(self.Nm, self.nm, self.W, self.beta)
        h = _lib.ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.jhtx_plan_create_rings(self.ns, north.ctypes.data, lmax, spin, 1, _lib.ctypes.byref(h)))
            self._h = h
            nbytes = int(lib.jhtx_plan_workspace_bytes(h))
            self.workspace = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=self.device)
            # theta-FFT normalisation 1/Nk folded into the row factors
            self.sk = torch.as_tensor(1.0 / (self.Nk * corr_k), dtype=torch.float64, device=self.device)
            self.sm = torch.as_tensor(1.0 / corr_m, dtype=torch.float64, device=self.device)
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
            except Exception:  # interpreter shutdown
                pass
            self._lib.jhtx_plan_destroy(h)
            self._h = None

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _point_workspace(self, npts: int):
        """Scratch for the tile sort of the points (the library allocates nothing)."""
        nws = int(self._lib.jhtx_nufft_workspace_bytes(npts, self.nk, self.nm))
        return torch.empty(max(nws, 8), dtype=torch.uint8, device=self.device), nws

    def grid(self, alm: torch.Tensor) -> torch.Tensor:
        """alm ([2,] K complex128) -> the oversampled grid g (nk x nm) of the type-2 NUFFT."""
        s, lib = self._stream(), self._lib
        F = torch.empty((self.ncomp, self.ns, self.lmax + 1), dtype=torch.complex128, device=self.device)
        _lib.check(lib.jhtx_legendre_synth(self._h, alm.data_ptr(), F.data_ptr(), 1, s))
        D = torch.empty((self.Nk, self.Nm), dtype=torch.complex128, device=self.device)
        _lib.check(lib.jhtx_dfs_extend(F.data_ptr(), self.ns, self.lmax, self.spin, D.data_ptr(), s))
        X = torch.fft.fft(D, dim=0).contiguous()  # the FFT along rows may return other strides
        del D, F
        C = torch.empty((self.nk, self.nm), dtype=torch.complex128, device=self.device)
        _lib.check(lib.jhtx_nufft_grid_fill(X.data_ptr(), self.Nk, self.lmax, self.sk.data_ptr(), self.sm.data_ptr(),
                                            C.data_ptr(), self.nk, self.nm, s))
        del X
        return torch.fft.ifft2(C, norm="forward").contiguous()  # unnormalised inverse DFT

    def interp(self, g: torch.Tensor, loc: torch.Tensor, deriv: bool):
        """(Re f, Im f or None, dloc [4, npts] or None) at the points."""
        npts = loc.shape[0]
        re = torch.empty(npts, dtype=torch.float64, device=self.device)
        im = torch.empty(npts, dtype=torch.float64, device=self.device) if self.spin else None
        d = torch.empty((4, npts), dtype=torch.float64, device=self.device) if deriv else None
        ws, nws = self._point_workspace(npts)
        _lib.check(self._lib.jhtx_nufft_interp(g.data_ptr(), self.nk, self.nm, self.W, self.beta, loc.data_ptr(), npts,
                                               re.data_ptr(), im.data_ptr() if im is not None else None,
                                               d.data_ptr() if d is not None else None, ws.data_ptr(), nws,
                                               self._stream()))
        return re, im, d

    def adjoint(self, re: torch.Tensor, im, loc: torch.Tensor) -> torch.Tensor:
        """Field values at the points -> the raw Legendre adjoint ([2,] K complex128, before
        the metric): the transposes of interp, grid and grid_fill, in reverse."""
        s, lib = self._stream(), self._lib
        g = torch.zeros((self.nk, self.nm), dtype=torch.complex128, device=self.device)
        npts = loc.shape[0]
        ws, nws = self._point_workspace(npts)
        _lib.check(lib.jhtx_nufft_spread(re.data_ptr(), im.data_ptr() if im is not None else None, loc.data_ptr(),
                                         npts, g.data_ptr(), self.nk, self.nm, self.W, self.beta,
                                         ws.data_ptr(), nws, s))
        C = torch.fft.fft2(g).contiguous()
        del g
        Y = torch.empty((self.Nk, self.Nm), dtype=torch.complex128, device=self.device)
        _lib.check(lib.jhtx_nufft_grid_gather(C.data_ptr(), self.Nk, self.lmax, self.sk.data_ptr(), self.sm.data_ptr(),
                                              Y.data_ptr(), self.nk, self.nm, s))
        del C
        dD = torch.fft.ifft(Y, dim=0, norm="forward").contiguous()  # unnormalised: the FFT's transpose
        V = torch.empty((self.ncomp, self.ns, self.lmax + 1), dtype=torch.complex128, device=self.device)
        _lib.check(lib.jhtx_dfs_fold(dD.data_ptr(), self.ns, self.lmax, self.spin, V.data_ptr(), s))
        del dD, Y
        out = torch.empty((self.ncomp, self.K), dtype=torch.complex128, device=self.device)
        _lib.check(lib.jhtx_legendre_adjoint(self._h, V.data_ptr(), out.data_ptr(), 1, self.workspace.data_ptr(), s))
        return out


_PLANS: dict = {}
_LOCK = threading.Lock()


def get_offgrid_plan(lmax: int, spin: int, epsilon: float, device=None) -> OffgridPlan:
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    key = (lmax, spin, float(epsilon), str(dev))
    with _LOCK:
        p = _PLANS.get(key)
        if p is None:
            p = OffgridPlan(lmax, spin, epsilon, dev)
            _PLANS[key] = p
        return p


def clear_offgrid_plans() -> None:
    with _LOCK:
        _PLANS.clear()


def _fields(plan: OffgridPlan, alm: torch.Tensor, loc: torch.Tensor, deriv: bool):
    """alm -> (field [npts] or [2, npts], dloc [4, npts] or None)."""
    re, im, d = plan.interp(plan.grid(alm), loc, deriv)
    return (re if plan.spin == 0 else torch.stack([re, im])), d


def _loc_grad(plan: OffgridPlan, w: torch.Tensor, d: torch.Tensor) -> torch.Tensor:
    """sum over components of w_c(j) d field_c(j) / d(theta_j, phi_j) -> (npts, 2)."""
    if plan.spin == 0:
        return torch.stack([w * d[0], w * d[2]], dim=1)
    return torch.stack([w[0] * d[0] + w[1] * d[1], w[0] * d[2] + w[1] * d[3]], dim=1)


def _adjoint_raw(plan: OffgridPlan, field: torch.Tensor, loc: torch.Tensor) -> torch.Tensor:
    """The strict adjoint in the (2 - delta_m0) metric: the library's Legendre adjoint
    returns spin 0 sum lambda V and spin s (-(p + t n) / 2, i (p - t n) / 2), t = (-1)^s;
    the off-grid transpose is spin 0 (p + n) / G and spin s (-(p + t n), i (p - t n)) / G
    (src/jht/offgrid.py:113-121, :240-243)."""
    if plan.spin == 0:
        b = plan.adjoint(field, None, loc)[0]
    else:
        b = plan.adjoint(field[0].contiguous(), field[1].contiguous(), loc)
    _, G = _metric_tensors(plan.lmax, str(b.device))
    return b / G if plan.spin == 0 else b * (2.0 / G)


class _SynthGeneralFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alm, loc, plan):
        need_loc = ctx.needs_input_grad[1]
        f, d = _fields(plan, alm, loc, need_loc)
        ctx.plan = plan
        ctx.save_for_backward(loc, d if need_loc else None)
        return f

    @staticmethod
    def backward(ctx, gf):
        plan = ctx.plan
        loc, d = ctx.saved_tensors
        gf = gf.contiguous()
        ga = gl = None
        if ctx.needs_input_grad[0]:
            b = _AdjGeneralFn.apply(gf, loc, plan)
            ga = _synth_backward(b, plan.lmax, plan.spin)
        if ctx.needs_input_grad[1]:
            gl = _loc_grad(plan, gf, d)
        return ga, gl, None


class _AdjGeneralFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, field, loc, plan):
        ctx.plan = plan
        ctx.save_for_backward(field, loc)
        return _adjoint_raw(plan, field, loc)

    @staticmethod
    def backward(ctx, h):
        plan = ctx.plan
        field, loc = ctx.saved_tensors
        a = _adjoint_backward_alm(h.contiguous(), plan.lmax).contiguous()
        need_loc = ctx.needs_input_grad[1]
        if ctx.needs_input_grad[0] and not need_loc:
            return _SynthGeneralFn.apply(a, loc, plan), None, None
        # d/d loc_j of sum Re(conj(h) S^T v) = sum_c v_c(j) d[S a]_c(j) / d loc_j
        f, d = _fields(plan, a, loc, True)
        gfield = f if ctx.needs_input_grad[0] else None
        return gfield, _loc_grad(plan, field, d), None


def _check_spin(spin: int):
    if spin not in (0, 1, 2, 3):
        raise NotImplementedError(f"spin={spin} unsupported (0, 1, 2, 3)")


def _loc_npts(loc) -> int:
    shape = tuple(loc.shape) if hasattr(loc, "shape") else np.shape(loc)
    if len(shape) != 2 or shape[-1] != 2:
        raise ValueError(f"loc must have shape (npts, 2) [theta, phi]; got {shape}")
    return int(shape[0])


def _loc_tensor(loc) -> torch.Tensor:
    t, _ = _prep(loc, False)
    return t


def synthesis_general(alm, loc, *, spin: int = 0, lmax: int, epsilon: float = 1e-10) -> torch.Tensor:
    """``a_lm -> field`` at arbitrary points ``loc`` (``src/jht/offgrid.py:165-205``).

    alm: healpy-packed, ``(alm_size(lmax),)`` for spin 0 or ``(2, alm_size(lmax))`` (E, B)
    for spin 2.  loc: ``(npts, 2)`` real, ``loc[:, 0] = theta``, ``loc[:, 1] = phi``; angles
    are not validated, both are evaluated on the 2 pi-periodic extension (theta through
    F(-theta, phi) = (-1)^spin F(theta, phi + pi)).  Returns ``(npts,)`` for spin 0 or
    ``(2, npts)`` (Q, U).  Differentiable in alm and in loc.
    """
    _check_spin(spin)
    _loc_npts(loc)
    shape = tuple(alm.shape) if hasattr(alm, "shape") else np.shape(alm)
    expect = (alm_size(lmax),) if spin == 0 else (2, alm_size(lmax))
    if tuple(shape) != expect:
        raise ValueError(
            f"alm shape {tuple(shape)} != expected {expect} for lmax={lmax}, spin={spin} "
            "(a wrong-size alm would otherwise be silently clamped by the gather)"
        )
    loc_t = _loc_tensor(loc)
    a, single = _prep(alm, True)
    plan = get_offgrid_plan(lmax, spin, epsilon, a.device)
    out = _SynthGeneralFn.apply(a, loc_t, plan)
    return out.to(torch.float32) if single else out


def adjoint_synthesis_general(field, loc, *, spin: int = 0, lmax: int, epsilon: float = 1e-10) -> torch.Tensor:
    """Exact transpose of :func:`synthesis_general` (``src/jht/offgrid.py:208-243``):
    ``<S_g a, v>_pts == <a, S_g^T v>_G`` with G = 2 - delta_m0.  field: ``(npts,)`` for
    spin 0 or ``(2, npts)`` real (Q, U) planes.  Differentiable in field and in loc."""
    _check_spin(spin)
    npts = _loc_npts(loc)
    if (isinstance(field, torch.Tensor) and field.is_complex()) or (
            not isinstance(field, torch.Tensor) and np.iscomplexobj(field)):
        raise TypeError(
            "field must be real ((npts,) for spin=0; (2, npts) real Q/U planes for spin>0); "
            "a complex field would be silently mis-folded"
        )
    shape = tuple(field.shape) if hasattr(field, "shape") else np.shape(field)
    expect = (npts,) if spin == 0 else (2, npts)
    if tuple(shape) != expect:
        raise ValueError(f"field shape {tuple(shape)} != expected {expect} for {npts} points")
    loc_t = _loc_tensor(loc)
    v, single = _prep(field, False)
    plan = get_offgrid_plan(lmax, spin, epsilon, v.device)
    out = _AdjGeneralFn.apply(v, loc_t, plan)
    return out.to(torch.complex64) if single else out


def synthesis_general_real(x, loc, *, spin: int = 0, lmax: int, epsilon: float = 1e-10) -> torch.Tensor:
    """``S_g o T^-1``: real DOF vector -> field at the points (``src/jht/diff.py:79-92``).

    ``x`` has length ``n_dof(lmax, spin)``; the result is that of
    :func:`synthesis_general`.  A plain real-linear map, differentiable in ``x`` and ``loc``.
    """
    from .callers import real_to_alm

    return synthesis_general(real_to_alm(x, lmax, spin), loc, spin=spin, lmax=lmax, epsilon=epsilon)


def adjoint_synthesis_general_real(field, loc, *, spin: int = 0, lmax: int, epsilon: float = 1e-10) -> torch.Tensor:
    """``T o S_g^T``: field at the points -> real DOF vector (``src/jht/diff.py:95-107``), the
    exact Euclidean transpose of :func:`synthesis_general_real` (T is an isometry in the
    (2 - delta_m0) metric)."""
    from .callers import alm_to_real

    return alm_to_real(adjoint_synthesis_general(field, loc, spin=spin, lmax=lmax, epsilon=epsilon), lmax, spin)


__all__ = ["synthesis_general", "adjoint_synthesis_general", "synthesis_general_real",
           "adjoint_synthesis_general_real", "OffgridPlan", "get_offgrid_plan", "clear_offgrid_plans"]
