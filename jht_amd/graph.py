"""Ragged batches of transforms replayed as one captured HIP graph.

A :class:`TransformGraph` holds several instances -- each a geometry (nside, lmax, spin),
a batch of maps and a precision -- with their plans, workspaces and static input/output
buffers, and captures synthesis followed by the adjoint of every instance into one
``torch.cuda.CUDAGraph``.  ``replay()`` then launches the whole ragged batch with one
graph launch: no per-call host work, no allocation.  The captured calls are the library's
ordinary entry points on the capture stream (they only launch kernels), so results are
bitwise those of eager calls.
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


class TransformGraph:
    """``instances``: iterable of :class:`Instance` or ``(nside, lmax, spin, batch, precision)``.

    Fill ``alm[i]`` (``(batch, K)`` or ``(batch, 2, K)`` complex128) in place, call
    :meth:`replay`, read ``maps[i] = S alm[i]`` and, with ``adjoint=True``,
    ``back[i] = S^T maps[i]``.

    Each instance gets a private plan (tables and workspace) that no other call uses, so a
    replay may run on one stream while eager transforms of the same geometry run on another.
    """

    def __init__(self, instances, adjoint: bool = True, device=None):
        self.instances = [i if isinstance(i, Instance) else Instance(*i) for i in instances]
        if not self.instances:
            raise ValueError("TransformGraph needs at least one instance")
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.adjoint = adjoint
        self.plans, self.alm, self.maps, self.back = [], [], [], []
        flags = _lib.FLAG_BLUESTEIN_ALL if get_azimuth_fft_mode() == "looped" else 0
        for ins in self.instances:
            prec = _prec(ins.precision)
            plan = Plan(ins.nside, ins.lmax, ins.spin, ins.batch, prec, flags, self.device,
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
        for ins, plan, a, m, b in zip(self.instances, self.plans, self.alm, self.maps, self.back):
            plan.synth(a, m, ins.batch)
            if self.adjoint:
                plan.adjoint(m, b, ins.batch)

    def replay(self) -> None:
        """Run every instance's synthesis (and adjoint) on the current stream."""
        self.graph.replay()


__all__ = ["Instance", "TransformGraph"]
