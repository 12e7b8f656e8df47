"""Write the golden fixtures under tests/golden/ from the float64 oracle.

Each fixture holds the seeded inputs and the oracle's outputs for one small configuration:

* c1_spin0_synth.npz   -- nside 64, lmax 128, spin 0: the C_l = (l+1)^-2 Gaussian a_lm
                          (seed 20261203) and its RING map (the BASELINE config-1 case);
* s0_n8_l12.npz, s0_n16_l24.npz, s2_n8_l12.npz, s2_n16_l24.npz
                       -- white a_lm (tests/test_healpix.py:34-44 recipe) and N(0,1) maps
                          with synthesis and adjoint outputs, spin 0 and spin 2.

The oracle was pinned independently before these were written (tests/test_oracle.py).
Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import jht_oracle as O  # noqa: E402


def main() -> None:
    rng = np.random.default_rng(20261203)
    a = O.gaussian_alm(128, rng)
    np.savez_compressed(os.path.join(HERE, "c1_spin0_synth.npz"), nside=64, lmax=128, spin=0,
                        seed=20261203, alm=a, map=O.synthesis(a, 64, 128, 0))
    for spin in (0, 2):
        for nside, lmax in ((8, 12), (16, 24)):
            rng = np.random.default_rng(1000 * spin + nside)
            if spin == 0:
                a = O.random_alm(lmax, rng)
            else:
                a = np.stack([O.random_alm(lmax, rng, 2), O.random_alm(lmax, rng, 2)])
            npix = 12 * nside * nside
            v = rng.standard_normal(npix if spin == 0 else (2, npix))
            np.savez_compressed(
                os.path.join(HERE, f"s{spin}_n{nside}_l{lmax}.npz"), nside=nside, lmax=lmax, spin=spin,
                seed=1000 * spin + nside, alm=a, map=O.synthesis(a, nside, lmax, spin), v=v,
                adj=O.adjoint_synthesis(v, nside, lmax, spin))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
