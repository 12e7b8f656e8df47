"""Off-grid synthesis at arbitrary points (GPU): jht_amd.synthesis_general and
adjoint_synthesis_general against the oracle (oracle/offgrid_oracle.py, a restatement of
src/jht/offgrid.py + src/jht/_nufft.py pinned in tests/test_offgrid_oracle.py).

* parity with the oracle, spins 0-3, random points plus the poles and grid nodes, at
  the 1e-10 and 1e-12 kernel tiers (the two run the same algorithm: they agree far below
  the tier);
* at HEALPix pixel centres the field equals the on-grid GPU synthesis (nside 128,
  lmax 256) to the NUFFT tier;
* the strict adjoint identity <S a, v> = sum G Re(conj(a) S^T v) at lmax 512, 2e5 points;
* autograd: the alm gradient through the adjoint bridge, and the pointing gradient
  d field / d loc of both operators against central finite differences;
* a full-size case (lmax 1024, 1e6 points, spin 2) against brute-force harmonic sums at
  16 of the points;
* the reference's argument errors (offgrid.py:188-240).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import jht_oracle as O
from oracle import offgrid_oracle as G

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _alm(lmax, spin, rng):
    if spin == 0:
        return O.random_alm(lmax, rng)
    return np.stack([O.random_alm(lmax, rng, spin), O.random_alm(lmax, rng, spin)])


def _loc(n, rng, extra=True):
    pts = np.stack([rng.uniform(0.0, np.pi, n), rng.uniform(0.0, 2 * np.pi, n)], axis=1)
    if extra:
        pts = np.concatenate([pts, [[0.0, 0.0], [np.pi, 1.0], [np.pi / 2, 0.0], [0.3, 2 * np.pi - 1e-15]]])
    return pts


@pytest.mark.parametrize("spin", [0, 1, 2, 3])
@pytest.mark.parametrize("lmax,eps", [(12, 1e-10), (40, 1e-12), (7, 1e-6)])
def test_synthesis_matches_oracle(spin, lmax, eps):
    import jht_amd

    rng = np.random.default_rng(100 + lmax + spin)
    alm, loc = _alm(lmax, spin, rng), _loc(300, rng)
    got = jht_amd.synthesis_general(torch.tensor(alm, device="cuda"), torch.tensor(loc, device="cuda"),
                                    spin=spin, lmax=lmax, epsilon=eps)
    want = G.synthesis_general(alm, loc, spin, lmax, eps)
    e = _rel(got, want)
    print(f"synthesis_general spin {spin} lmax {lmax} eps {eps}: rel {e:.2e}", flush=True)
    assert got.shape == want.shape and e < 1e-12


@pytest.mark.parametrize("spin", [0, 1, 2, 3])
@pytest.mark.parametrize("lmax", [12, 40])
def test_adjoint_matches_oracle(spin, lmax):
    import jht_amd

    rng = np.random.default_rng(200 + lmax + spin)
    loc = _loc(300, rng)
    v = rng.standard_normal(loc.shape[0]) if spin == 0 else rng.standard_normal((2, loc.shape[0]))
    got = jht_amd.adjoint_synthesis_general(torch.tensor(v, device="cuda"), torch.tensor(loc, device="cuda"),
                                            spin=spin, lmax=lmax)
    want = G.adjoint_synthesis_general(v, loc, spin, lmax)
    e = _rel(got, want)
    print(f"adjoint_synthesis_general spin {spin} lmax {lmax}: rel {e:.2e}", flush=True)
    assert got.shape == want.shape and e < 1e-12


@pytest.mark.parametrize("spin", [0, 2])
def test_pixel_centres_match_on_grid_gpu(spin):
    import jht_amd

    nside, lmax = 128, 256
    rng = np.random.default_rng(300 + spin)
    alm = torch.tensor(O.gaussian_alm(lmax, rng) if spin == 0 else
                       np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)]), device="cuda")
    loc = torch.tensor(G.healpix_centres(nside), device="cuda")
    got = jht_amd.synthesis_general(alm, loc, spin=spin, lmax=lmax, epsilon=1e-12)
    want = jht_amd.synthesis(alm, nside, lmax, spin)
    e = _rel(got, want)
    print(f"pixel centres nside {nside} spin {spin}: rel {e:.2e}", flush=True)
    assert e < 1e-10


@pytest.mark.parametrize("spin", [0, 2])
def test_adjoint_identity_large(spin):
    import jht_amd

    lmax, npts = 512, 200_000
    rng = np.random.default_rng(400 + spin)
    alm = torch.tensor(_alm(lmax, spin, rng), device="cuda")
    loc = torch.tensor(_loc(npts, rng, extra=False), device="cuda")
    v = torch.tensor(rng.standard_normal(npts if spin == 0 else (2, npts)), device="cuda")
    Sa = jht_amd.synthesis_general(alm, loc, spin=spin, lmax=lmax)
    STv = jht_amd.adjoint_synthesis_general(v, loc, spin=spin, lmax=lmax)
    Gm = torch.tensor(O.alm_metric_weight(lmax), device="cuda")
    lhs = float((Sa * v).sum())
    rhs = float((Gm * (alm.conj() * STv).real).sum())
    print(f"adjoint identity spin {spin}: {lhs:.12e} vs {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


@pytest.mark.parametrize("spin", [0, 1, 2, 3])
def test_alm_gradient_bridge(spin):
    """d/d alm of 0.5 |S a - d|^2 equals G (S^T r) with the spin-2 m = 0 convention of the
    on-grid operators; checked against central differences along a random direction."""
    import jht_amd

    lmax = 24
    rng = np.random.default_rng(500 + spin)
    a0 = _alm(lmax, spin, rng)
    loc = torch.tensor(_loc(400, rng), device="cuda")
    d = jht_amd.synthesis_general(torch.tensor(_alm(lmax, spin, rng), device="cuda"), loc, spin=spin, lmax=lmax)

    def loss(a):
        r = jht_amd.synthesis_general(a, loc, spin=spin, lmax=lmax) - d
        return 0.5 * (r * r).sum()

    a = torch.tensor(a0, device="cuda", requires_grad=True)
    loss(a).backward()
    du = _alm(lmax, spin, rng)
    h = 1e-6
    lp = float(loss(torch.tensor(a0 + h * du, device="cuda")))
    lm = float(loss(torch.tensor(a0 - h * du, device="cuda")))
    fd = (lp - lm) / (2 * h)
    an = float((a.grad.conj() * torch.tensor(du, device="cuda")).real.sum())
    print(f"alm gradient spin {spin}: fd {fd:.10e} autograd {an:.10e}", flush=True)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("spin", [0, 1, 2])
def test_pointing_gradients(spin):
    """d/d loc of <w, S a>(loc) and of <c, S^T v>(loc) against central differences."""
    import jht_amd

    lmax, n = 20, 50
    rng = np.random.default_rng(600 + spin)
    alm = torch.tensor(_alm(lmax, spin, rng), device="cuda")
    loc0 = _loc(n, rng, extra=False)
    loc0[:, 0] = np.clip(loc0[:, 0], 0.2, np.pi - 0.2)
    w = torch.tensor(rng.standard_normal(n if spin == 0 else (2, n)), device="cuda")
    c = torch.tensor(_alm(lmax, spin, rng), device="cuda")

    def f_syn(L):
        return (w * jht_amd.synthesis_general(alm, L, spin=spin, lmax=lmax)).sum()

    def f_adj(L):
        return (c.conj() * jht_amd.adjoint_synthesis_general(w, L, spin=spin, lmax=lmax)).real.sum()

    for f in (f_syn, f_adj):
        L = torch.tensor(loc0, device="cuda", requires_grad=True)
        f(L).backward()
        g = L.grad.cpu().numpy()
        h = 1e-6
        worst = 0.0
        for j in (0, 7, 23, 41):
            for k in (0, 1):
                e = np.zeros_like(loc0)
                e[j, k] = h
                fd = (float(f(torch.tensor(loc0 + e, device="cuda"))) - float(f(torch.tensor(loc0 - e, device="cuda")))) / (2 * h)
                worst = max(worst, abs(fd - g[j, k]) / max(1.0, abs(fd)))
        print(f"pointing gradient spin {spin} {f.__name__}: worst rel {worst:.2e}", flush=True)
        assert worst < 1e-6


def test_full_size_spin2_against_brute_force():
    import jht_amd

    lmax, npts = 1024, 1_000_000
    rng = np.random.default_rng(700)
    alm = np.stack([O.gaussian_alm(lmax, rng, 2), O.gaussian_alm(lmax, rng, 2, 0.1)])
    loc = _loc(npts, rng, extra=False)
    out = jht_amd.synthesis_general(torch.tensor(alm, device="cuda"), torch.tensor(loc, device="cuda"), spin=2, lmax=lmax)
    assert out.shape == (2, npts) and bool(torch.isfinite(out).all())
    idx = rng.choice(npts, 16, replace=False)
    want = G.brute_force_field(alm, loc[idx], 2, lmax)
    e = float(np.max(np.abs(out[:, idx].cpu().numpy() - want)) / np.max(np.abs(want)))
    print(f"full size lmax {lmax} {npts} points: rel {e:.2e} at 16 points", flush=True)
    assert e < 1e-9


def test_errors_match_reference():
    import jht_amd

    lmax = 4
    a = torch.zeros(O.alm_size(lmax), dtype=torch.complex128, device="cuda")
    loc = torch.zeros((3, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="loc must have shape"):
        jht_amd.synthesis_general(a, torch.zeros(3, device="cuda"), lmax=lmax)
    with pytest.raises(ValueError, match="alm shape"):
        jht_amd.synthesis_general(a[:-1], loc, lmax=lmax)
    with pytest.raises(NotImplementedError):
        jht_amd.synthesis_general(a, loc, spin=5, lmax=lmax)
    with pytest.raises(NotImplementedError):
        jht_amd.synthesis_general(torch.zeros((2, O.alm_size(lmax)), dtype=torch.complex128, device="cuda"), loc,
                                  spin=4, lmax=lmax)
    with pytest.raises(TypeError, match="field must be real"):
        jht_amd.adjoint_synthesis_general(torch.zeros(3, dtype=torch.complex128, device="cuda"), loc, lmax=lmax)
    with pytest.raises(ValueError, match="field shape"):
        jht_amd.adjoint_synthesis_general(torch.zeros(4, dtype=torch.float64, device="cuda"), loc, lmax=lmax)
    # no points: empty field / zero alm
    empty = torch.zeros((0, 2), dtype=torch.float64, device="cuda")
    assert jht_amd.synthesis_general(a, empty, lmax=lmax).shape == (0,)
    assert float(jht_amd.adjoint_synthesis_general(torch.zeros(0, dtype=torch.float64, device="cuda"), empty,
                                                   lmax=lmax).abs().max()) == 0.0


@pytest.mark.parametrize("W", [6, 14, 16])
def test_binned_and_direct_nufft_paths_agree(W):
    """jhtx_nufft_interp / jhtx_nufft_spread with the tile-sort workspace against the direct
    (per-point gather / global atomic) paths, through the C ABI, on a random grid with
    points clustered near the poles and on tile borders."""
    from jht_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(800 + W)
    nk, nm, npts, beta = 200, 150, 20000, 2.3 * W
    g = torch.tensor(rng.standard_normal((nk, nm)) + 1j * rng.standard_normal((nk, nm)), device="cuda")
    th = np.concatenate([rng.uniform(0, np.pi, npts - 200), rng.uniform(0, 1e-3, 100),
                         np.full(100, 2 * np.pi * 32 / nk)])
    loc = torch.tensor(np.stack([th, rng.uniform(-1.0, 7.0, npts)], axis=1), device="cuda")
    nws = int(lib.jhtx_nufft_workspace_bytes(npts, nk, nm))
    ws = torch.empty(nws, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    outs = []
    for w, n in ((None, 0), (ws.data_ptr(), nws)):
        re = torch.empty(npts, dtype=torch.float64, device="cuda")
        im = torch.empty_like(re)
        d = torch.empty((4, npts), dtype=torch.float64, device="cuda")
        _lib.check(lib.jhtx_nufft_interp(g.data_ptr(), nk, nm, W, beta, loc.data_ptr(), npts, re.data_ptr(),
                                         im.data_ptr(), d.data_ptr(), w, n, s))
        grid = torch.zeros((nk, nm), dtype=torch.complex128, device="cuda")
        _lib.check(lib.jhtx_nufft_spread(re.data_ptr(), im.data_ptr(), loc.data_ptr(), npts, grid.data_ptr(),
                                         nk, nm, W, beta, w, n, s))
        outs.append((re, im, d, grid))
    for a, b in zip(outs[0], outs[1]):
        assert _rel(b, a) < 1e-13
    # an undersized workspace is refused
    with pytest.raises(_lib.JhtxError):
        _lib.check(lib.jhtx_nufft_spread(outs[0][0].data_ptr(), None, loc.data_ptr(), npts, outs[0][3].data_ptr(),
                                         nk, nm, W, beta, ws.data_ptr(), nws - 4, s))


@pytest.mark.parametrize("spin", [0, 2])
def test_real_dof_offgrid_pair_is_a_transpose(spin):
    """synthesis_general_real / adjoint_synthesis_general_real (src/jht/diff.py:79-107):
    <S x, v> = <x, S^T v> in plain real inner products."""
    import jht_amd

    lmax = 20
    rng = np.random.default_rng(900 + spin)
    n = jht_amd.n_dof(lmax, spin)
    x = torch.tensor(rng.standard_normal(n), device="cuda")
    loc = torch.tensor(_loc(300, rng), device="cuda")
    v = torch.tensor(rng.standard_normal(loc.shape[0] if spin == 0 else (2, loc.shape[0])), device="cuda")
    Sx = jht_amd.synthesis_general_real(x, loc, spin=spin, lmax=lmax)
    STv = jht_amd.adjoint_synthesis_general_real(v, loc, spin=spin, lmax=lmax)
    lhs, rhs = float((Sx * v).sum()), float((x * STv).sum())
    print(f"real-DOF off-grid spin {spin}: {lhs:.12e} vs {rhs:.12e}", flush=True)
    assert STv.shape == (n,) and abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))
