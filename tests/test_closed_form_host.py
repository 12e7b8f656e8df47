"""The GPU tests' closed-form single-mode maps (tests/_closed_form.py), evaluated on the CPU,
against the oracle's full synthesis of the same one-mode a_lm: spin 0 and spin 2 (E and B),
m = 0, interior m and sectoral modes, on grids whose rings cover every class (caps,
shifted and unshifted belt rings)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import jht_oracle as O

from _closed_form import PixelGeometry


@pytest.mark.parametrize("nside,lmax", [(8, 16), (16, 40)])
def test_closed_form_single_modes_match_oracle(nside, lmax):
    pg = PixelGeometry(nside, "cpu")
    K = O.alm_size(lmax)
    for l, m in [(2, 0), (7, 0), (5, 3), (16, 16), (lmax, 1), (lmax, lmax - 1)]:
        idx = O.alm_column_base(m, lmax) + (l - m)
        val = 0.7 - 0.4j if m > 0 else 0.7 + 0.0j
        a = np.zeros(K, complex)
        a[idx] = val
        ref = O.synthesis(a, nside, lmax, 0)
        got = pg.spin0(l, m, val).numpy()
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), (l, m)
        for aE, aB in [(val, 0j), (0j, val)]:
            A = np.zeros((2, K), complex)
            A[0, idx], A[1, idx] = aE, aB
            ref2 = O.synthesis(A, nside, lmax, 2)
            got2 = pg.spin2(l, m, complex(aE), complex(aB)).numpy()
            assert np.max(np.abs(got2 - ref2)) <= 1e-13 * np.max(np.abs(ref2)), (l, m, aE, aB)
