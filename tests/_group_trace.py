"""One ragged-group synthesis + adjoint (3 spin-0 instances), for the kernel-trace count in
tests/test_gpu_group.py (run under rocprofv3 there; not a test module itself)."""

from __future__ import annotations

import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import torch  # noqa: E402

from jht_amd.graph import RaggedGroup  # noqa: E402
from jht_amd.healpix import alm_size  # noqa: E402


def main():
    inst = [(32, 64, 0, 3), (64, 128, 0, 17), (16, 32, 0, 1)]
    g = RaggedGroup([(n, l, s, B, "mixed") for n, l, s, B in inst])
    gen = torch.Generator(device="cuda").manual_seed(5)
    alms = [torch.randn((B, alm_size(l)), dtype=torch.complex128, device="cuda", generator=gen)
            for n, l, s, B in inst]
    maps = [torch.zeros((B, 12 * n * n), dtype=torch.float64, device="cuda") for n, l, s, B in inst]
    back = [torch.zeros_like(a) for a in alms]
    torch.cuda.synchronize()
    g.synth(alms, maps)
    g.adjoint(maps, back)
    torch.cuda.synchronize()
    print("group trace done", flush=True)


if __name__ == "__main__":
    main()
