"""CPU checks of the native library and the host logic (no GPU calls).

* libjhtx.so loads and exports every function include/jhtx.h declares;
* the plan's host tables (jhtx_host_rec_tables) reproduce RingInfo's colatitudes bit for
  bit, and a numpy model of the kernels' rescaled, exponent-tracked recursion run on those
  tables reproduces the oracle's per-m ring coefficients (both hemispheres) to 1e-12 --
  this validates the algebra the HIP kernels implement (alpha/gamma/G scaling, seeds,
  2^512 exponent steps, even/odd N/S combination) independently of the device;
* the Python layer's validation errors and warnings match the reference's.
"""

from __future__ import annotations

import ctypes
import os
import re
import warnings

import numpy as np
import pytest

from oracle import jht_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from jht_amd import _lib

    return _lib.load()


def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "jhtx.h")).read()
    names = set(re.findall(r"\b(jhtx_[a-z_0-9]+)\s*\(", hdr))
    assert names, "no declarations found"
    lib = _lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    from jht_amd import _lib as L

    assert set(L.EXPORTS) == names


def host_tables(nside, lmax, spin):
    lib = _lib()
    T, K, M = 2 * nside, O.alm_size(lmax), lmax + 1
    out = {k: np.zeros(n) for k, n in dict(x=T, lsin=T, lcos=T, alpha=K, gamma=K, G=K,
                                           seed=(M if spin == 0 else 8 * M)).items()}
    ptr = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    rc = lib.jhtx_host_rec_tables(nside, lmax, spin, ptr(out["x"]), ptr(out["lsin"]), ptr(out["lcos"]),
                                  ptr(out["alpha"]), ptr(out["gamma"]), ptr(out["G"]), ptr(out["seed"]))
    assert rc == 0
    return out


def _seed_split(L, sign):
    """The kernel's seed_split: exp(L) = sign * mu * 2^(512 ex)."""
    e = np.minimum(np.floor((L * 1.4426950408889634 + 256.0) / 512.0), 0.0)
    n = 512.0 * e
    r = (L - n * 6.93147180369123816490e-01) - n * 1.90821492927058770002e-10
    return sign * np.exp(r), e.astype(int)


def model_synth(tab, nside, lmax, spin, alm, ms):
    """Numpy model of k_synth_s0 / k_synth_s2 for the orders ``ms``: returns
    {m: (north (T,), south (T-1,))} for spin 0, {m: (FpN, FpS, FmN, FmS)} for spin 2."""
    T = 2 * nside
    x = tab["x"]
    res = {}
    for m in ms:
        cb = m * (2 * lmax + 1 - m) // 2
        lmin = m if spin == 0 else max(m, 2)
        nrec = 1 if spin == 0 else 2
        mu = np.zeros((nrec, T))
        mup = np.zeros((nrec, T))
        ex = np.zeros((nrec, T), int)
        if spin == 0:
            L = tab["seed"][m] + m * tab["lsin"]
            s, ex[0] = _seed_split(L, -1.0 if m & 1 else 1.0)
            mup[0] = -s
        else:
            for k in range(2):
                sp = tab["seed"][8 * m + 4 * k: 8 * m + 4 * k + 4]
                L = np.full(T, sp[0])
                if sp[1] > 0:
                    L = L + sp[1] * tab["lcos"]
                if sp[2] > 0:
                    L = L + sp[2] * tab["lsin"]
                s, ex[k] = _seed_split(L, sp[3])
                mup[k] = -s
        acc = np.zeros((2, 4 if spin else 1, T), complex)  # [class][channel]
        for j, l in enumerate(range(lmin, lmax + 1)):
            idx = cb + l
            al = 0.0 if l == lmin else tab["alpha"][idx]
            ga = 0.0 if (l == lmin or spin == 0) else tab["gamma"][idx]
            c = np.stack([al * x + ga, al * x - ga])[:nrec]
            mn = c * mu - mup
            mup, mu = mu, mn
            f = (ex == 0).astype(float)
            g = tab["G"][idx]
            if spin == 0:
                acc[j & 1, 0] += alm[idx] * g * mn[0] * f[0]
            else:
                aE, aB = alm[0][idx] * g, alm[1][idx] * g
                vp, vm = mn[0] * f[0], mn[1] * f[1]
                acc[j & 1] += np.stack([aE * vp, aB * vp, aE * vm, aB * vm])
            if j % 8 == 7:
                big = (ex < 0) & (np.abs(mu) > 2.0**256)
                mu = np.where(big, mu * 2.0**-512, mu)
                mup = np.where(big, mup * 2.0**-512, mup)
                ex = ex + big
        c0, c1 = (acc[1], acc[0]) if (lmin - m) & 1 else (acc[0], acc[1])
        if spin == 0:
            res[m] = ((c0 + c1)[0], (c0 - c1)[0][: T - 1])
        else:
            PE, PB, ME, MB = c0 + c1
            dE, dB = (c0 - c1)[2], (c0 - c1)[3]
            qE, qB = (c0 - c1)[0], (c0 - c1)[1]
            res[m] = (-PE - 1j * PB, (-dE - 1j * dB)[: T - 1], -ME + 1j * MB, (-qE + 1j * qB)[: T - 1])
    return res


@pytest.mark.parametrize("nside,lmax,spin", [(8, 12, 0), (32, 64, 0), (64, 192, 0), (8, 12, 2), (32, 64, 2), (64, 160, 2)])
def test_kernel_model_matches_oracle(nside, lmax, spin):
    tab = host_tables(nside, lmax, spin)
    geo = O.RingInfo(nside)
    assert np.array_equal(tab["x"], geo.z[: 2 * nside])  # bitwise
    rng = np.random.default_rng(nside + lmax + spin)
    T = 2 * nside
    if spin == 0:
        alm = O.random_alm(lmax, rng)
        plan = O.RecursionPlan(geo.z[:T], 0, lmax)
        Ftot, Fsig = O.synth_contract_eo(plan, O.tri_to_dense(alm, lmax))
        ms = sorted({0, 1, 2, lmax // 3, lmax // 2, lmax - 1, lmax})
        mod = model_synth(tab, nside, lmax, 0, alm, ms)
        scale = np.max(np.abs(Ftot))
        for m in ms:
            north, south = mod[m]
            assert np.max(np.abs(north - Ftot[m])) <= 1e-12 * scale, m
            sref = ((-1.0) ** m) * Fsig[m][: T - 1]
            assert np.max(np.abs(south - sref)) <= 1e-12 * scale, m
    else:
        alm = np.stack([O.random_alm(lmax, rng, 2), O.random_alm(lmax, rng, 2)])
        pp, pm = O.RecursionPlan(geo.z[:T], 2, lmax), O.RecursionPlan(geo.z[:T], -2, lmax)
        aE, aB = O.tri_to_dense(alm[0], lmax), O.tri_to_dense(alm[1], lmax)
        ref = O.synth_contract_spin2_ns(pp, pm, -(aE + 1j * aB), -(aE - 1j * aB))
        ms = sorted({0, 1, 2, 3, lmax // 3, lmax // 2, lmax - 1, lmax})
        mod = model_synth(tab, nside, lmax, 2, alm, ms)
        scale = max(np.max(np.abs(r)) for r in ref)
        for m in ms:
            sg = (-1.0) ** m
            want = (ref[0][m], sg * ref[1][m][: T - 1], ref[2][m], sg * ref[3][m][: T - 1])
            for got, w in zip(mod[m], want):
                assert np.max(np.abs(got - w)) <= 1e-12 * scale, m


def test_kernel_model_deep_underflow():
    """High m near the pole: seeds far below the double range (the 2^512 exponent steps
    and the dropped contributions while ex < 0)."""
    nside, lmax = 256, 700
    tab = host_tables(nside, lmax, 0)
    geo = O.RingInfo(nside)
    T = 2 * nside
    rng = np.random.default_rng(5)
    alm = O.random_alm(lmax, rng)
    ms = [600, 650, 700]
    mod = model_synth(tab, nside, lmax, 0, alm, ms)
    for m in ms:
        lam = O.normalized_legendre(geo.z[:T], m, lmax)
        cb = O.alm_column_base(m, lmax)
        ref = alm[cb: cb + lmax + 1 - m] @ lam
        scale = max(np.max(np.abs(ref)), 1e-300)
        assert np.max(np.abs(mod[m][0] - ref)) <= 1e-12 * max(scale, 1e-3), m
    assert np.min(tab["seed"][600] + 600 * tab["lsin"]) < -745  # seeds really underflow


def test_python_validation_without_gpu():
    import jht_amd

    K = jht_amd.alm_size(24)
    with pytest.raises(ValueError, match="alm shape"):
        jht_amd.synthesis(np.zeros(K - 1, complex), 16, 24, 0)
    with pytest.raises(ValueError, match="map shape"):
        jht_amd.adjoint_synthesis(np.zeros(12 * 16**2), 16, 24, 2)
    with pytest.raises(ValueError, match="unknown azimuth FFT mode"):
        jht_amd.set_azimuth_fft_mode("bogus")
    assert jht_amd.get_azimuth_fft_mode() == "unrolled"
    G = jht_amd.alm_metric_weight(24)
    m = O.alm_m_of_index(24)
    assert np.all(G[m == 0] == 1) and np.all(G[m > 0] == 2)
    geo, ref = jht_amd.RingInfo(16), O.RingInfo(16)
    assert np.array_equal(geo.z, ref.z) and np.array_equal(geo.startpix, ref.startpix)


def test_ceiling_warning_once_per_geometry():
    from jht_amd import healpix as H

    H._WARNED.discard((4, 7, 0))
    with pytest.warns(UserWarning, match="ceiling"):
        H._warn_once(4, 7, 0)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        H._warn_once(4, 7, 0)  # second time: silent
        H._warn_once(4, 6, 0)  # at the ceiling: silent


def test_tuning_validation_without_gpu():
    """An out-of-range jhtx_tuning field is rejected before any device work."""
    from jht_amd import _lib as L

    lib = _lib()
    h = ctypes.c_void_p()
    bad = L.Tuning(synth_rings=3)  # spin 0 allows 1, 4, 8
    rc = lib.jhtx_plan_create_tuned(16, 24, 0, 0, 1, 0, ctypes.byref(bad), ctypes.byref(h))
    assert rc == L.JHTX_EINVAL and not h.value
    assert b"tuning" in lib.jhtx_last_error()
    bad = L.Tuning(mx_adj_cb=3)
    rc = lib.jhtx_plan_create_tuned(16, 24, 2, 1, 1, 0, ctypes.byref(bad), ctypes.byref(h))
    assert rc == L.JHTX_EINVAL and not h.value
    bad = L.Tuning(mx_synth_group=32, mx_synth_fold=2)  # 32 maps per spin-2 recursion: only folded
    rc = lib.jhtx_plan_create_tuned(16, 24, 2, 1, 1, 0, ctypes.byref(bad), ctypes.byref(h))
    assert rc == L.JHTX_EINVAL and not h.value
    bad = L.Tuning(mx_single=3)
    rc = lib.jhtx_plan_create_tuned(16, 24, 2, 1, 1, 0, ctypes.byref(bad), ctypes.byref(h))
    assert rc == L.JHTX_EINVAL and not h.value
    bad = L.Tuning(mx_f16=3)
    rc = lib.jhtx_plan_create_tuned(16, 24, 2, 1, 1, 0, ctypes.byref(bad), ctypes.byref(h))
    assert rc == L.JHTX_EINVAL and not h.value
    bad = L.Tuning(mx_adj_f16=3)
    rc = lib.jhtx_plan_create_tuned(16, 24, 0, 1, 1, 0, ctypes.byref(bad), ctypes.byref(h))
    assert rc == L.JHTX_EINVAL and not h.value
    for field in ("mx_synth_fold",):
        bad = L.Tuning(**{field: 3})
        rc = lib.jhtx_plan_create_tuned(16, 24, 2, 1, 1, 0, ctypes.byref(bad), ctypes.byref(h))
        assert rc == L.JHTX_EINVAL and not h.value, field


def test_offgrid_entry_validation_without_gpu():
    """The off-grid C entry points reject bad arguments before any device work, and the
    host-side NUFFT tables of jht_amd.offgrid match the oracle's (the reference's
    epsilon table, 2,3,5-smooth grid sizes, ES deconvolution factors)."""
    from jht_amd import _lib as L
    from jht_amd import offgrid as OG
    from oracle import offgrid_oracle as G

    lib = _lib()
    h = ctypes.c_void_p()
    dp = ctypes.POINTER(ctypes.c_double)

    def rings(theta, nrings, spin=0):
        arr = (ctypes.c_double * len(theta))(*theta)
        return lib.jhtx_plan_create_rings(nrings, ctypes.cast(arr, ctypes.c_void_p), 8, spin, 1, ctypes.byref(h))

    half = np.pi / 2
    assert rings([0.0, 0.5, half], 4) == L.JHTX_EINVAL  # even ring count
    assert rings([0.0, 0.5, 1.0], 5) == L.JHTX_EINVAL   # last north ring is not the equator
    assert b"equator" in lib.jhtx_last_error()
    assert rings([0.5, 0.2, half], 5) == L.JHTX_EINVAL  # not sorted
    assert rings([0.0, 0.5, half], 5, spin=4) == L.JHTX_EUNSUPPORTED
    assert not h.value
    assert lib.jhtx_dfs_extend(None, 5, 8, 0, None, None) == L.JHTX_EINVAL
    assert lib.jhtx_nufft_interp(None, 64, 64, 14, 1.0, None, 10, None, None, None, None, 0, None) == L.JHTX_EINVAL
    assert lib.jhtx_nufft_spread(None, None, None, 10, None, 64, 64, 7, 1.0, None, 0, None) == L.JHTX_EINVAL
    assert lib.jhtx_nufft_workspace_bytes(1000, 64, 64) == 4 * (2 * 4 + 1 + 3000)
    # host tables
    for lmax, eps in ((12, 1e-10), (100, 1e-6), (7, 3e-9)):
        ns = OG._smooth_even(lmax + 1) + 1
        ns_o, _ = G.cc_rings(lmax)
        assert ns == ns_o
        W, sigma = OG._kernel_params(eps)
        assert (W, sigma) == G.kernel_params(eps)
        p = G.NufftPlan(2 * ns - 2, 2 * lmax + 1, eps)
        assert (OG._smooth_even(sigma * (2 * ns - 2)), OG._smooth_even(sigma * (2 * lmax + 1))) == (p.nk, p.nm)
        np.testing.assert_array_equal(OG._es_correction(p.Nk, p.nk, W, p.beta), p.corr_k)


def test_offgrid_python_validation_without_gpu():
    import jht_amd

    K = jht_amd.alm_size(4)
    with pytest.raises(NotImplementedError):
        jht_amd.synthesis_general(np.zeros(K, complex), np.zeros((2, 2)), spin=7, lmax=4)
    with pytest.raises(NotImplementedError, match="0, 1, 2, 3"):
        jht_amd.synthesis_general(np.zeros((2, K), complex), np.zeros((2, 2)), spin=4, lmax=4)
    with pytest.raises(ValueError, match="loc must have shape"):
        jht_amd.synthesis_general(np.zeros(K, complex), np.zeros(3), lmax=4)
    with pytest.raises(TypeError, match="field must be real"):
        jht_amd.adjoint_synthesis_general(np.zeros(3, complex), np.zeros((3, 2)), lmax=4)


def test_reference_public_names_exist():
    """Every public name of the reference package (src/jht/__init__.py:66-106) on the
    transform path and its callers has a counterpart in jht_amd.  The XLA compilation cache
    (enable_compilation_cache, src/jht/_cache.py) has none: nothing here is compiled at run
    time."""
    import jht_amd

    names = ["synthesis", "adjoint_synthesis", "synthesis_general", "adjoint_synthesis_general",
             "analysis", "map2alm", "bare_analysis", "ring_weights", "pixel_weights", "pseudo_alm",
             "deconvolve", "wiener", "constrained_realization", "synthesis_real", "analysis_real",
             "synthesis_general_real", "adjoint_synthesis_general_real", "bandpower", "alm_to_real",
             "real_to_alm", "n_dof", "alm_size", "alm_metric_weight",
             "set_azimuth_fft_mode", "get_azimuth_fft_mode", "enable_looped_fft", "__version__"]
    missing = [n for n in names if not hasattr(jht_amd, n)]
    assert not missing, missing


def test_tuning_struct_matches_header():
    """The ctypes jhtx_tuning mirror has the header's fields in the header's order (the
    struct crosses the C ABI by layout)."""
    import re

    from jht_amd import _lib as L

    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "jhtx.h")).read()
    body = re.search(r"typedef struct jhtx_tuning \{(.*?)\} jhtx_tuning;", hdr, re.S).group(1)
    fields = [f.strip() for decl in re.findall(r"int ([^;]+);", body) for f in decl.split(",")]
    assert tuple(fields) == L.TUNING_FIELDS
    assert [f[0] for f in L.Tuning._fields_] == fields
    assert ctypes.sizeof(L.Tuning) == 4 * len(fields)


def test_ragged_group_validates_instances_on_the_host():
    """RaggedGroup refuses what the group kernels cannot run before it touches a device:
    mixed spins, FP64 instances, nside above 512, more than 16 instances, none."""
    import pytest

    from jht_amd import _lib as L
    from jht_amd.graph import RaggedGroup

    bad = [
        [(16, 32, 0, 1, "mixed"), (16, 32, 2, 1, "mixed")],
        [(16, 32, 0, 1, "f64")],
        [(1024, 64, 0, 1, "mixed")],
        [(16, 32, 0, 1, "mixed")] * (L.GROUP_MAX + 1),
        [],
    ]
    for inst in bad:
        with pytest.raises(ValueError):
            RaggedGroup(inst)
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "jhtx.h")).read()
    assert "n <= 16" in hdr and "nside <= 512" in hdr
    assert L.GROUP_MAX == 16 and L.GROUP_NSIDE_MAX == 512
