"""HEALPix (RING) on-grid synthesis and its exact adjoint on MI355X.

Drop-in for ``jht.synthesis`` / ``jht.adjoint_synthesis`` (src/jht/healpix.py:363-412)
with the same call signatures, healpy m-major a_lm packing, orthonormal Y_lm with the
Condon-Shortley phase and HEALPix/COSMO (Q, U).  Arrays are torch tensors on the GPU
(numpy input is accepted and moved to ``cuda:0``); the heavy lifting is done by the
hand-written HIP library ``libjhtx.so`` through its C ABI.

Differentiation: each operator is a ``torch.autograd.Function`` whose backward is the
other operator (torch's complex-gradient convention, d/dRe + i d/dIm):

* synthesis backward:  grad_a = G * S^T(g) for m > 0, G = 2 - delta_m0; at m = 0
  grad_a = S^T(g) for spin 0 and, for spin 2, (bE + i bB, bB - i bE) with b = S^T(g)
  (the conjugate of jht's native JAX VJP, tests/test_grad.py:134-148);
* adjoint backward:    grad_v = S(h / G) with the m = 0 entries of h replaced by Re(h).
"""

from __future__ import annotations

import threading
import warnings
from functools import lru_cache

import numpy as np
import torch

from . import _lib

__all__ = [
    "RingInfo",
    "adjoint_synthesis",
    "adjoint_synthesis_batch",
    "alm_column_base",
    "alm_metric_weight",
    "alm_size",
    "synthesis",
    "synthesis_batch",
    "get_azimuth_fft_mode",
    "set_azimuth_fft_mode",
    "enable_looped_fft",
]


# --------------------------------------------------------------------------- #
# a_lm layout (healpy) -- src/jht/healpix.py:49-74
# --------------------------------------------------------------------------- #
def alm_size(lmax: int) -> int:
    return (lmax + 1) * (lmax + 2) // 2


def alm_column_base(m: int, lmax: int) -> int:
    """Start index of the contiguous column ``a_{l,m}, l = m..lmax``."""
    return m * (2 * lmax + 1 - m) // 2 + m


@lru_cache(maxsize=None)
def _m_of_index(lmax: int) -> np.ndarray:
    out = np.concatenate([np.full(lmax + 1 - m, m, dtype=np.int64) for m in range(lmax + 1)])
    out.setflags(write=False)
    return out


@lru_cache(maxsize=None)
def alm_metric_weight(lmax: int) -> np.ndarray:
    """Per-mode weight ``2 - delta_{m0}``: the metric in which ``adjoint_synthesis`` is the
    strict transpose of ``synthesis`` (healpix.py:58-74)."""
    w = np.where(_m_of_index(lmax) == 0, 1.0, 2.0)
    w.setflags(write=False)
    return w


# --------------------------------------------------------------------------- #
# RING geometry (host numpy) -- src/jht/healpix.py:107-140
# --------------------------------------------------------------------------- #
class RingInfo:
    """Per-ring HEALPix RING geometry: z = cos(theta), ring length, phi0, start pixel."""

    def __init__(self, nside: int):
        ring = np.arange(1, 4 * nside, dtype=np.int64)
        n_r = np.empty(ring.shape, dtype=np.int64)
        z = np.empty(ring.shape, dtype=np.float64)
        phi0 = np.empty(ring.shape, dtype=np.float64)
        north = ring < nside
        south = ring > 3 * nside
        belt = ~north & ~south
        i = ring[north]
        n_r[north] = 4 * i
        z[north] = 1.0 - i**2 / (3.0 * nside**2)
        phi0[north] = np.pi / (4.0 * i)
        j = 4 * nside - ring[south]
        n_r[south] = 4 * j
        z[south] = -(1.0 - j**2 / (3.0 * nside**2))
        phi0[south] = np.pi / (4.0 * j)
        k = ring[belt]
        n_r[belt] = 4 * nside
        z[belt] = 4.0 / 3.0 - 2.0 * k / (3.0 * nside)
        phi0[belt] = np.where((k - nside) % 2 == 0, np.pi / (4.0 * nside), 0.0)
        self.nside = nside
        self.npix = 12 * nside**2
        self.z = z
        self.npix_ring = n_r
        self.phi0 = phi0
        self.startpix = np.concatenate([[0], np.cumsum(n_r)[:-1]])
        self.nrings = len(ring)


# --------------------------------------------------------------------------- #
# azimuth-FFT mode (src/jht/_azimuth.py:56-89): "looped" routes every ring DFT of
# length > 32 through the chirp-z path; part of the plan cache key.
# --------------------------------------------------------------------------- #
_LEGAL_MODES = ("unrolled", "looped")
_AZIMUTH_FFT_MODE = "unrolled"


def set_azimuth_fft_mode(mode: str) -> None:
    if mode not in _LEGAL_MODES:
        raise ValueError(f"unknown azimuth FFT mode {mode!r}; expected one of {_LEGAL_MODES}")
    global _AZIMUTH_FFT_MODE
    _AZIMUTH_FFT_MODE = mode


def get_azimuth_fft_mode() -> str:
    return _AZIMUTH_FFT_MODE


def enable_looped_fft() -> None:
    set_azimuth_fft_mode("looped")


# --------------------------------------------------------------------------- #
# plans
# --------------------------------------------------------------------------- #
class Plan:
    """A jhtx plan on one device plus its workspace (allocated once, reused per call).

    The workspace belongs to the plan, so calls on one Plan object must not overlap: keep a
    plan (and the cached ones :func:`get_plan` returns) on one stream, or create a private
    Plan per stream.
    """

    def __init__(self, nside: int, lmax: int, spin: int, max_batch: int = 1,
                 precision: int = _lib.PREC_F64, flags: int = 0, device=None, tuning=None,
                 workspace: bool = True):
        lib = _lib.load()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.nside, self.lmax, self.spin = nside, lmax, spin
        self.max_batch, self.precision, self.flags = max_batch, precision, flags
        self.ncomp = 1 if spin == 0 else 2
        self.K = alm_size(lmax)
        self.npix = 12 * nside * nside
        h = _lib.ctypes.c_void_p()
        with torch.cuda.device(self.device):
            t = None
            if tuning:
                unknown = set(tuning) - set(_lib.TUNING_FIELDS)
                if unknown:
                    raise ValueError(f"unknown tuning fields {sorted(unknown)}")
                t = _lib.Tuning(**{k: int(v) for k, v in tuning.items()})
            _lib.check(lib.jhtx_plan_create_tuned(nside, lmax, spin, precision, max_batch, flags,
                                                  _lib.ctypes.byref(t) if t is not None else None,
                                                  _lib.ctypes.byref(h)))
            self._h = h
            # workspace=False: the plan only lends its tables (e.g. to a RaggedGroup)
            nbytes = int(lib.jhtx_plan_workspace_bytes(h)) if workspace else 0
            self.workspace = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=self.device)
        self._has_workspace = workspace
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

    @property
    def tuning(self) -> dict:
        """The kernel shapes this plan uses (every field filled in)."""
        t = _lib.Tuning()
        _lib.check(self._lib.jhtx_plan_get_tuning(self._h, _lib.ctypes.byref(t)))
        return {k: getattr(t, k) for k in _lib.TUNING_FIELDS}

    STAGES = ("synth_rec", "synth_fft", "adj_fft", "adj_rec", "adj_finish")

    def set_profiling(self, on: bool) -> None:
        _lib.check(self._lib.jhtx_plan_set_profiling(self._h, 1 if on else 0))

    def read_profile(self) -> dict:
        """{stage: (summed ms, launches)} of the calls since the last read (HIP events)."""
        ms = (_lib.ctypes.c_double * len(self.STAGES))()
        cnt = (_lib.ctypes.c_int * len(self.STAGES))()
        _lib.check(self._lib.jhtx_plan_read_profile(self._h, ms, cnt))
        return {name: (ms[i], cnt[i]) for i, name in enumerate(self.STAGES)}

    def _need_workspace(self):
        if not self._has_workspace:
            raise RuntimeError("this plan was created without a workspace (workspace=False)")

    def synth(self, alm: torch.Tensor, out: torch.Tensor, batch: int) -> torch.Tensor:
        self._need_workspace()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.jhtx_synth(self._h, alm.data_ptr(), out.data_ptr(), batch,
                                        self.workspace.data_ptr(), stream))
        return out

    def adjoint(self, maps: torch.Tensor, out: torch.Tensor, batch: int) -> torch.Tensor:
        self._need_workspace()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.jhtx_adjoint(self._h, maps.data_ptr(), out.data_ptr(), batch,
                                          self.workspace.data_ptr(), stream))
        return out


_PLANS: dict = {}
_PLAN_LOCK = threading.Lock()
_WARNED: set = set()


def _warn_once(nside: int, lmax: int, spin: int) -> None:
    """The reference's once-per-geometry band-ceiling warning (healpix.py:201-207)."""
    key = (nside, lmax, spin)
    if key in _WARNED:
        return
    _WARNED.add(key)
    if 2 * lmax > 3 * nside:
        warnings.warn(
            f"jht_amd: lmax={lmax} exceeds the HEALPix band-limit ceiling 1.5*nside "
            f"(= {1.5 * nside:g} for nside={nside}); the transform is validated up to "
            "the ceiling, accuracy above it is not.",
            stacklevel=4,
        )


# kernel shapes chosen for a geometry, precision and batch (jht_amd.tune.autotune or
# set_tuning); batch None: any batch without an entry of its own
_TUNED: dict = {}


def set_tuning(nside: int, lmax: int, spin: int, precision: int, tuning, batch: int | None = None) -> None:
    """Use ``tuning`` (a dict of ``jhtx_tuning`` fields, or None for the defaults) for the
    plans of this geometry and precision created from now on for calls of ``batch`` maps
    (``None``: every batch that has no tuning of its own).  A winner timed at one batch is
    not applied to another: the best maps-per-recursion group depends on it."""
    key = (nside, lmax, spin, precision, batch)
    with _PLAN_LOCK:
        if tuning:
            _TUNED[key] = dict(tuning)
        else:
            _TUNED.pop(key, None)
        for k in [k for k in _PLANS if k[:4] == key[:4]]:
            del _PLANS[k]


def _tuning_for(nside: int, lmax: int, spin: int, precision: int, batch: int):
    t = _TUNED.get((nside, lmax, spin, precision, batch))
    return t if t is not None else _TUNED.get((nside, lmax, spin, precision, None))


def get_plan(nside: int, lmax: int, spin: int, batch: int = 1, precision: int = _lib.PREC_F64,
             device=None, tuning=None) -> Plan:
    if spin not in (0, 2):
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    _warn_once(nside, lmax, spin)
    if not torch.cuda.is_available():
        raise RuntimeError("jht_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    flags = _lib.FLAG_BLUESTEIN_ALL if _AZIMUTH_FFT_MODE == "looped" else 0
    if tuning is None:
        tuning = _tuning_for(nside, lmax, spin, precision, batch)
    tkey = tuple(sorted(tuning.items())) if tuning else ()
    key = (nside, lmax, spin, precision, flags, dev, tkey)
    with _PLAN_LOCK:
        p = _PLANS.get(key)
        if p is None or p.max_batch < batch:
            mb = max(batch, p.max_batch if p is not None else 1)
            p = None  # release the smaller workspace first
            _PLANS.pop(key, None)
            p = Plan(nside, lmax, spin, mb, precision, flags, torch.device("cuda", dev), tuning)
            _PLANS[key] = p
        return p


def clear_plans() -> None:
    """Release every cached plan (device tables and workspaces)."""
    with _PLAN_LOCK:
        _PLANS.clear()


# --------------------------------------------------------------------------- #
# raw (non-differentiable) batched calls
# --------------------------------------------------------------------------- #
def _as_device(x, dtype) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        t = x
        if t.device.type != "cuda":
            t = t.to("cuda")
    else:
        t = torch.as_tensor(np.asarray(x), device="cuda")
    if t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


_PRECISIONS = {"f64": _lib.PREC_F64, "mixed": _lib.PREC_MIXED}


def _prec(precision) -> int:
    if precision not in _PRECISIONS:
        raise ValueError(f"unknown precision {precision!r}; expected one of {tuple(_PRECISIONS)}")
    return _PRECISIONS[precision]


def _synth_raw(alm: torch.Tensor, nside: int, lmax: int, spin: int, precision: str = "f64") -> torch.Tensor:
    """alm (B, [2,] K) complex128 on GPU -> (B, [2,] npix) float64."""
    B = alm.shape[0]
    plan = get_plan(nside, lmax, spin, B, precision=_prec(precision), device=alm.device)
    shape = (B, plan.npix) if spin == 0 else (B, 2, plan.npix)
    out = torch.empty(shape, dtype=torch.float64, device=alm.device)
    return plan.synth(alm, out, B)


def _adjoint_raw(maps: torch.Tensor, nside: int, lmax: int, spin: int, precision: str = "f64") -> torch.Tensor:
    B = maps.shape[0]
    plan = get_plan(nside, lmax, spin, B, precision=_prec(precision), device=maps.device)
    shape = (B, plan.K) if spin == 0 else (B, 2, plan.K)
    out = torch.empty(shape, dtype=torch.complex128, device=maps.device)
    return plan.adjoint(maps, out, B)


@lru_cache(maxsize=None)
def _metric_tensors(lmax: int, device: str):
    m0 = torch.as_tensor(_m_of_index(lmax) == 0, device=device)
    G = torch.as_tensor(alm_metric_weight(lmax).copy(), device=device, dtype=torch.float64)
    return m0, G


def _synth_backward(b: torch.Tensor, lmax: int, spin: int) -> torch.Tensor:
    """torch gradient w.r.t. a from b = S^T(g) (see the module docstring)."""
    m0, G = _metric_tensors(lmax, str(b.device))
    grad = b * G
    if spin != 0:  # E, B at m = 0 (any spin weight: the off-grid pair also takes 1 and 3)
        bE = b[..., 0, :].real
        bB = b[..., 1, :].real
        gE = torch.where(m0, torch.complex(bE, bB), grad[..., 0, :])
        gB = torch.where(m0, torch.complex(bB, -bE), grad[..., 1, :])
        grad = torch.stack([gE, gB], dim=-2)
    return grad


def _adjoint_backward_alm(h: torch.Tensor, lmax: int) -> torch.Tensor:
    m0, G = _metric_tensors(lmax, str(h.device))
    return torch.where(m0, torch.complex(h.real, torch.zeros_like(h.real)), h / G)


class _SynthesisFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alm, nside, lmax, spin, precision="f64"):
        ctx.geom = (nside, lmax, spin, precision)
        return _synth_raw(alm, nside, lmax, spin, precision)

    @staticmethod
    def backward(ctx, g):
        nside, lmax, spin, precision = ctx.geom
        b = _AdjointFn.apply(g.contiguous(), nside, lmax, spin, precision)
        return _synth_backward(b, lmax, spin), None, None, None, None


class _AdjointFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, nside, lmax, spin, precision="f64"):
        ctx.geom = (nside, lmax, spin, precision)
        return _adjoint_raw(maps, nside, lmax, spin, precision)

    @staticmethod
    def backward(ctx, h):
        nside, lmax, spin, precision = ctx.geom
        a = _adjoint_backward_alm(h.contiguous(), lmax)
        return _SynthesisFn.apply(a.contiguous(), nside, lmax, spin, precision), None, None, None, None


_X64_WARNED = [False]


def _prep(x, complex_in: bool):
    """Move to the GPU in double precision; remember the caller's precision."""
    single = False
    if isinstance(x, torch.Tensor):
        single = x.dtype in (torch.float32, torch.complex64)
    else:
        a = np.asarray(x)
        single = a.dtype in (np.float32, np.complex64)
    if single and not _X64_WARNED[0]:
        _X64_WARNED[0] = True
        warnings.warn(
            "jht_amd: single-precision input -- the FP64 operators run in float64 and the "
            "result is returned in the input's precision (~1e-7 accuracy tier).",
            stacklevel=3,
        )
    return _as_device(x, torch.complex128 if complex_in else torch.float64), single


# --------------------------------------------------------------------------- #
# public transforms -- src/jht/healpix.py:363-412
# --------------------------------------------------------------------------- #
def synthesis(alm, nside: int, lmax: int, spin: int = 0) -> torch.Tensor:
    """``S: a_lm -> map`` on the HEALPix RING grid.

    alm: complex, ``(alm_size(lmax),)`` for spin 0 or ``(2, alm_size(lmax))`` (E, B) for
    spin 2.  Returns a real ``(12 nside**2,)`` map, or ``(2, 12 nside**2)`` (Q, U).
    """
    shape = tuple(alm.shape) if hasattr(alm, "shape") else np.shape(alm)
    expect = (alm_size(lmax),) if spin == 0 else (2, alm_size(lmax))
    if tuple(shape) != expect:
        raise ValueError(
            f"alm shape {tuple(shape)} != expected {expect} for lmax={lmax}, spin={spin} "
            "(a wrong-size alm would otherwise be silently clamped by the gather)"
        )
    if spin not in (0, 2):
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    a, single = _prep(alm, True)
    out = _SynthesisFn.apply(a.unsqueeze(0), nside, lmax, spin)[0]
    return out.to(torch.float32) if single else out


def adjoint_synthesis(m, nside: int, lmax: int, spin: int = 0) -> torch.Tensor:
    """``S^T = Y^H : map -> a_lm``, the exact transpose of :func:`synthesis`.

    ``<S a, v>_map == <a, S^T v>_alm`` in the ``(2 - delta_m0)``-weighted inner product;
    equals ``(Npix / 4pi) * healpy.map2alm(v, iter=0, use_weights=False)``.
    """
    shape = tuple(m.shape) if hasattr(m, "shape") else np.shape(m)
    npix = 12 * nside * nside
    expect = (npix,) if spin == 0 else (2, npix)
    if tuple(shape) != expect:
        raise ValueError(
            f"map shape {tuple(shape)} != expected {expect} for nside={nside}, spin={spin} "
            "(a wrong-size map would otherwise be silently clamped by the gather)"
        )
    if spin not in (0, 2):
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    v, single = _prep(m, False)
    out = _AdjointFn.apply(v.unsqueeze(0), nside, lmax, spin)[0]
    return out.to(torch.complex64) if single else out


def synthesis_batch(alm, nside: int, lmax: int, spin: int = 0, precision: str = "f64") -> torch.Tensor:
    """Batched :func:`synthesis`: ``(B, K)`` / ``(B, 2, K)`` -> ``(B, npix)`` / ``(B, 2, npix)``.

    ``precision="mixed"`` keeps the l-recursion in FP64 and runs the contraction over l on
    the FP32 matrix cores for groups of maps (relative error ~1e-6 against the FP64 path).
    """
    expect = (alm_size(lmax),) if spin == 0 else (2, alm_size(lmax))
    if tuple(alm.shape[1:]) != expect or len(alm.shape) != len(expect) + 1:
        raise ValueError(f"alm shape {tuple(alm.shape)} != expected (B, *{expect})")
    if spin not in (0, 2):
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    a, _ = _prep(alm, True)
    return _SynthesisFn.apply(a, nside, lmax, spin, precision)


def adjoint_synthesis_batch(m, nside: int, lmax: int, spin: int = 0, precision: str = "f64") -> torch.Tensor:
    """Batched :func:`adjoint_synthesis` (``precision`` as in :func:`synthesis_batch`)."""
    npix = 12 * nside * nside
    expect = (npix,) if spin == 0 else (2, npix)
    if tuple(m.shape[1:]) != expect or len(m.shape) != len(expect) + 1:
        raise ValueError(f"map shape {tuple(m.shape)} != expected (B, *{expect})")
    if spin not in (0, 2):
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    v, _ = _prep(m, False)
    return _AdjointFn.apply(v, nside, lmax, spin, precision)
