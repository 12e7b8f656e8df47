"""jhtx_adjoint_noalloc -- the adjoint with its ring coefficients and chunk partials in the
library's static per-GPU scratch array -- against jhtx_adjoint with a caller workspace,
through the C ABI.  Two calls in a row, the second with a smaller batch and new inputs, must
give jhtx_adjoint's results bit for bit (the same kernels on the same inputs), on an FP64
spin-2 plan and mixed spin-2 plans (the 16-map FP16-split kernels, then a ragged group; the
32-map kernel, whose first-stored-l table also lives in the scratch, then 7 maps);
the existing entry point keeps working between them; a call above the scratch size is
refused before any launch.
"""

from __future__ import annotations

import warnings

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _maps(rng, batch, nside, spin):
    npix = 12 * nside * nside
    shape = (batch, npix) if spin == 0 else (batch, 2, npix)
    return torch.tensor(rng.standard_normal(shape), device="cuda")


def _adjoint_ws(lib, plan, maps, batch, stream):
    from jht_amd import _lib

    out = torch.full((batch, plan.ncomp, plan.K) if plan.spin else (batch, plan.K), np.nan,
                     dtype=torch.complex128, device="cuda")
    _lib.check(lib.jhtx_adjoint(plan._h, maps.data_ptr(), out.data_ptr(), batch, plan.workspace.data_ptr(), stream))
    return out


def _adjoint_noalloc(lib, plan, maps, batch, stream):
    from jht_amd import _lib

    out = torch.full((batch, plan.ncomp, plan.K) if plan.spin else (batch, plan.K), np.nan,
                     dtype=torch.complex128, device="cuda")
    _lib.check(lib.jhtx_adjoint_noalloc(plan._h, maps.data_ptr(), out.data_ptr(), batch, stream))
    return out


@pytest.mark.parametrize("precision,nside,lmax,batches", [("f64", 64, 128, (3, 1)), ("mixed", 128, 256, (16, 5)),
                                                      ("mixed", 64, 128, (32, 7))])
def test_adjoint_noalloc_matches_workspace_entry(precision, nside, lmax, batches):
    from jht_amd import _lib
    from jht_amd import healpix as H

    lib = _lib.load()
    spin = 2
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        plan = H.Plan(nside, lmax, spin, batches[0], H._prec(precision))
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(2026 + nside)
    for call, batch in enumerate(batches):
        maps = _maps(rng, batch, nside, spin)
        print(f"noalloc {precision} nside={nside} lmax={lmax} call {call}: batch {batch}", flush=True)
        got = _adjoint_noalloc(lib, plan, maps, batch, stream)
        want = _adjoint_ws(lib, plan, maps, batch, stream)
        torch.cuda.synchronize()
        assert torch.isfinite(torch.view_as_real(want)).all()
        assert torch.equal(got, want), (precision, call)
        # the workspace entry point on other inputs in between, then the scratch one again
        other = _adjoint_ws(lib, plan, _maps(rng, batch, nside, spin), batch, stream)
        again = _adjoint_noalloc(lib, plan, maps, batch, stream)
        torch.cuda.synchronize()
        assert torch.isfinite(torch.view_as_real(other)).all()
        assert torch.equal(again, want), (precision, call)
    del plan
    H.clear_plans()


def test_adjoint_noalloc_refuses_above_scratch():
    from jht_amd import _lib
    from jht_amd import healpix as H

    lib = _lib.load()
    nside, lmax, spin = 1024, 2048, 2  # ring coefficients alone: 2 x 4095 x 2049 x 16 B = 268 MB
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        plan = H.Plan(nside, lmax, spin, 1, _lib.PREC_F64)
    maps = torch.zeros((1, 2, 12 * nside * nside), dtype=torch.float64, device="cuda")
    out = torch.full((1, 2, plan.K), 7.0, dtype=torch.complex128, device="cuda")
    rc = lib.jhtx_adjoint_noalloc(plan._h, maps.data_ptr(), out.data_ptr(), 1, torch.cuda.current_stream().cuda_stream)
    msg = lib.jhtx_last_error().decode()
    print(f"refused: rc {rc}: {msg}", flush=True)
    assert rc == _lib.JHTX_EINVAL and "bytes of scratch" in msg
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())  # nothing ran
    del plan
    H.clear_plans()
