"""The committed golden fixtures (tests/golden/*.npz) through the GPU operators.

c1_spin0_synth.npz is the first benchmark configuration's own input: nside 64, lmax 128,
spin 0, the C_l = (l+1)^-2 Gaussian a_lm (seed 20261203) and its RING map.  The four
nside 8 / 16 fixtures hold white a_lm and N(0,1) maps with their synthesis and adjoint.
Each runs through jht_amd.synthesis / adjoint_synthesis (the C ABI) and must match the
fixture to 1e-12 relative, the reference's gate (tests/test_healpix.py:78-90).
"""

from __future__ import annotations

import glob
import os
import warnings

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_golden_fixture_on_gpu(path):
    import jht_amd

    d = np.load(path)
    nside, lmax, spin = int(d["nside"]), int(d["lmax"]), int(d["spin"])
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        out = jht_amd.synthesis(torch.tensor(d["alm"], device="cuda"), nside, lmax, spin).cpu().numpy()
        e = _rel(out, d["map"])
        print(f"{os.path.basename(path)}: synthesis rel {e:.2e}", flush=True)
        assert e <= 1e-12
        if "v" in d.files:
            b = jht_amd.adjoint_synthesis(torch.tensor(d["v"], device="cuda"), nside, lmax, spin).cpu().numpy()
            e = _rel(b, d["adj"])
            print(f"{os.path.basename(path)}: adjoint rel {e:.2e}", flush=True)
            assert e <= 1e-12
