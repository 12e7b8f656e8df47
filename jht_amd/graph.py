"""Ragged batches of transforms: one launch per stage, replayed as one captured HIP graph.

A :class:`RaggedGroup` runs the transforms of several mixed-precision instances -- each a
geometry (nside, lmax), a batch of maps, one spin for all -- with ONE kernel launch per
stage (``jhtx_group_*`` in the C ABI): synthesis is 2 launches (recursion, ring FFT) and
the adjoint 3 (ring FFT, recursion, chunk reduction), whatever the number of instances.
The reference batches one geometry by ``jax.vmap`` (scripts/gpu_check.py:150-153).

A :class:`TransformGraph` holds the instances' static input/output buffers and captures
synthesis followed by the adjoint of every instance into one ``torch.cuda.CUDAGraph``:
either each instance's own calls (its plan's kernels, any precision) or, with
``fused=True``, the group's five launches.  ``replay()`` then runs the whole ragged batch
with one graph launch: no per-call host work, no allocation.  The captured calls are the
library's ordinary entry points on the capture stream (they only launch kernels), so
results are bitwise those of eager calls.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _lib
from .healpix import Plan, _prec, _tuning_for, alm_size, get_azimuth_fft_mode


@dataclass(frozen=True)
class Instance:
    nside: int
    lmax: int
    spin: int = 0
    batch: int = 1
    precision: str = "f64"


class RaggedGroup:
    """Several mixed-precision instances (one spin, nside <= 512, at most 16) transformed with
    one kernel launch per stage.

    The contraction runs on the FP32 matrix cores, 16 maps per FP64 recursion (an instance
    with fewer maps leaves columns empty).  Each instance gets a private plan that only lends
    its tables; the group owns one workspace, so calls on one group must not overlap.
    """

    def __init__(self, instances, device=None):
        self.instances = [i if isinstance(i, Instance) else Instance(*i) for i in instances]
        n = len(self.instances)
        if not 1 <= n <= _lib.GROUP_MAX:
            raise ValueError(f"a RaggedGroup takes 1..{_lib.GROUP_MAX} instances, got {n}")
        spins = {ins.spin for ins in self.instances}
        if len(spins) != 1:
            raise ValueError(f"a RaggedGroup needs one spin, got {sorted(spins)}")
        for ins in self.instances:
            if ins.precision != "mixed":
                raise ValueError(f"a RaggedGroup runs mixed-precision instances only, got {ins.precision!r}")
            if ins.nside > _lib.GROUP_NSIDE_MAX:
                raise ValueError(f"a RaggedGroup takes nside <= {_lib.GROUP_NSIDE_MAX}, got {ins.nside}")
        self.spin = spins.pop()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        lib = _lib.load()
        flags = _lib.FLAG_BLUESTEIN_ALL if get_azimuth_fft_mode() == "looped" else 0
        self.plans = [Plan(ins.nside, ins.lmax, ins.spin, ins.batch, _lib.PREC_MIXED, flags, self.device,
                           workspace=False) for ins in self.instances]
        handles = (_lib.ctypes.c_void_p * n)(*[p._h.value for p in self.plans])
        batches = (_lib.ctypes.c_int * n)(*[ins.batch for ins in self.instances])
        h = _lib.ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.jhtx_group_create(handles, batches, n, _lib.ctypes.byref(h)))
            self._h = h
            nbytes = int(lib.jhtx_group_workspace_bytes(h))
            self.workspace = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=self.device)
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                torch.cuda.synchronize(self.device)
            except Exception:  # interpreter shutdown
                pass
            self._lib.jhtx_group_destroy(h)
            self._h = None

    def query(self, what: int) -> int:
        """``jhtx_group_query``: 2 / 3 launches per synthesis / adjoint call, 4 + s work-groups of stage s."""
        return int(self._lib.jhtx_group_query(self._h, what))

    def _ptrs(self, tensors, shapes, dtype):
        if len(tensors) != len(self.instances):
            raise ValueError(f"expected {len(self.instances)} tensors, got {len(tensors)}")
        for t, shp in zip(tensors, shapes):
            if tuple(t.shape) != shp or t.dtype != dtype or not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"expected a contiguous {dtype} tensor of shape {shp} on {self.device}, "
                                 f"got {tuple(t.shape)} {t.dtype} on {t.device}")
        return (_lib.ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])

    def _shapes(self):
        a, m = [], []
        for ins in self.instances:
            K, npix = alm_size(ins.lmax), 12 * ins.nside * ins.nside
            a.append((ins.batch, K) if ins.spin == 0 else (ins.batch, 2, K))
            m.append((ins.batch, npix) if ins.spin == 0 else (ins.batch, 2, npix))
        return a, m

    def synth(self, alms, maps):
        """``maps[i] = S alms[i]`` for every instance (2 kernel launches on the current stream)."""
        ash, msh = self._shapes()
        ain = self._ptrs(alms, ash, torch.complex128)
        mout = self._ptrs(maps, msh, torch.float64)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.jhtx_group_synth(self._h, ain, mout, self.workspace.data_ptr(), stream))
        return maps

    def adjoint(self, maps, alms):
        """``alms[i] = S^T maps[i]`` for every instance (3 kernel launches on the current stream)."""
        ash, msh = self._shapes()
        min_ = self._ptrs(maps, msh, torch.float64)
        aout = self._ptrs(alms, ash, torch.complex128)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(self._lib.jhtx_group_adjoint(self._h, min_, aout, self.workspace.data_ptr(), stream))
        return alms


class TransformGraph:
    """``instances``: iterable of :class:`Instance` or ``(nside, lmax, spin, batch, precision)``.

    Fill ``alm[i]`` (``(batch, K)`` or ``(batch, 2, K)`` complex128) in place, call
    :meth:`replay`, read ``maps[i] = S alm[i]`` and, with ``adjoint=True``,
    ``back[i] = S^T maps[i]``.

    Each instance gets a private plan (tables and workspace) that no other call uses, so a
    replay may run on one stream while eager transforms of the same geometry run on another.
    With ``fused=True`` the instances (mixed precision, one spin, nside <= 512) run as one
    :class:`RaggedGroup`: the graph then holds five kernel launches in all.
    """

    def __init__(self, instances, adjoint: bool = True, device=None, fused: bool = False):
        self.instances = [i if isinstance(i, Instance) else Instance(*i) for i in instances]
        if not self.instances:
            raise ValueError("TransformGraph needs at least one instance")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.adjoint = adjoint
        self.fused = fused
        self.group = RaggedGroup(self.instances, self.device) if fused else None
        self.plans, self.alm, self.maps, self.back = [], [], [], []
        flags = _lib.FLAG_BLUESTEIN_ALL if get_azimuth_fft_mode() == "looped" else 0
        for ins in self.instances:
            prec = _prec(ins.precision)
            plan = None if fused else Plan(ins.nside, ins.lmax, ins.spin, ins.batch, prec, flags, self.device,
                                           _tuning_for(ins.nside, ins.lmax, ins.spin, prec, ins.batch))
            K, npix = alm_size(ins.lmax), 12 * ins.nside * ins.nside
            ashape = (ins.batch, K) if ins.spin == 0 else (ins.batch, 2, K)
            mshape = (ins.batch, npix) if ins.spin == 0 else (ins.batch, 2, npix)
            self.plans.append(plan)
            self.alm.append(torch.zeros(ashape, dtype=torch.complex128, device=self.device))
            self.maps.append(torch.zeros(mshape, dtype=torch.float64, device=self.device))
            self.back.append(torch.zeros(ashape, dtype=torch.complex128, device=self.device) if adjoint else None)
        # one eager pass first (first-use kernel attributes are set outside the capture)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self._launch()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch()

    def _launch(self):
        if self.fused:
            self.group.synth(self.alm, self.maps)
            if self.adjoint:
                self.group.adjoint(self.maps, self.back)
            return
        for ins, plan, a, m, b in zip(self.instances, self.plans, self.alm, self.maps, self.back):
            plan.synth(a, m, ins.batch)
            if self.adjoint:
                plan.adjoint(m, b, ins.batch)

    def replay(self) -> None:
        """Run every instance's synthesis (and adjoint) on the current stream."""
        self.graph.replay()


__all__ = ["Instance", "RaggedGroup", "TransformGraph"]
