"""CPU float64 oracle for the HEALPix on-grid synthesis / adjoint pair.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``jht_amd``) never calls it and fails loudly when
the HIP library is missing.

It is a numpy restatement (written fresh, not copied) of the reference
algorithm in jrcheshire/jht v0.2.0, which cannot be imported here (its only
runtime dependency, jax/jaxlib 0.9.2, is absent and there is no network).  Each
function cites the reference file:line it follows.  Two independent
restatements are provided so they can pin each other:

* the *fused* form -- one l-major scan over all (m, north-half ring) with the
  branch-free log-renormalised recursion and even/odd (N/S) accumulators, as in
  ``src/jht/_recursion.py:257-572`` and the ring-group FFT assembly of
  ``src/jht/healpix.py:184-357``;
* the *eager* form -- per-m recursion, per-ring FFT, as in
  ``src/jht/_reference.py:22-143``.

Plus the chirp-z (Bluestein) cap path of ``src/jht/_azimuth.py:95-232`` and the
analysis / ring-weight drivers of ``src/jht/_analysis.py`` / ``weights.py``.

How each function is pinned (beyond fused == eager) is listed in the test
files ``tests/test_oracle.py``: closed forms for l <= 2, scipy
``sph_harm_y`` to l = 512, the explicit Wigner-d sum, mpmath, brute-force dense
Y synthesis at small nside, the adjoint identity, and monopole/constant-map
closed forms.
"""

from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

FOUR_PI = 4.0 * np.pi

# --------------------------------------------------------------------------- #
# a_lm packing (healpy m-major) -- src/jht/healpix.py:49-101
# --------------------------------------------------------------------------- #


def alm_size(lmax: int) -> int:
    """Number of packed coefficients, ``healpix.py:49-50``."""
    return (lmax + 1) * (lmax + 2) // 2


def alm_column_base(m: int, lmax: int) -> int:
    """Flat index of ``a_{m,m}``; the column ``l = m..lmax`` is contiguous (``healpix.py:53-55``)."""
    return m * (2 * lmax + 1 - m) // 2 + m


def alm_m_of_index(lmax: int) -> np.ndarray:
    return np.concatenate([np.full(lmax + 1 - m, m, dtype=np.int64) for m in range(lmax + 1)])


def alm_l_of_index(lmax: int) -> np.ndarray:
    return np.concatenate([np.arange(m, lmax + 1, dtype=np.int64) for m in range(lmax + 1)])


def alm_metric_weight(lmax: int) -> np.ndarray:
    """``G = 2 - delta_{m0}`` per packed mode (``healpix.py:58-74``)."""
    return np.where(alm_m_of_index(lmax) == 0, 1.0, 2.0)


def tri_to_dense(alm: np.ndarray, lmax: int) -> np.ndarray:
    """(K,) -> (M, lmax+1), zero where l < m (``healpix.py:77-101,260-261``)."""
    M = lmax + 1
    out = np.zeros((M, lmax + 1), dtype=np.complex128)
    for m in range(M):
        b = alm_column_base(m, lmax)
        out[m, m:] = alm[b : b + (lmax + 1 - m)]
    return out


def dense_to_tri(dense: np.ndarray, lmax: int) -> np.ndarray:
    """(M, lmax+1) -> (K,) (``healpix.py:263-264``)."""
    M = lmax + 1
    out = np.empty(alm_size(lmax), dtype=dense.dtype)
    for m in range(M):
        b = alm_column_base(m, lmax)
        out[b : b + (lmax + 1 - m)] = dense[m, m:]
    return out


# --------------------------------------------------------------------------- #
# RING geometry -- src/jht/healpix.py:107-140
# --------------------------------------------------------------------------- #


class RingInfo:
    """HEALPix RING geometry: per ring length, z = cos(theta), phi0, start pixel."""

    def __init__(self, nside: int):
        nside = int(nside)
        ring = np.arange(1, 4 * nside, dtype=np.int64)  # 1-based ring number
        n_r = np.empty(ring.shape, dtype=np.int64)
        z = np.empty(ring.shape, dtype=np.float64)
        phi0 = np.empty(ring.shape, dtype=np.float64)
        cap_n = ring < nside
        cap_s = ring > 3 * nside
        belt = ~cap_n & ~cap_s
        i = ring[cap_n]
        n_r[cap_n] = 4 * i
        z[cap_n] = 1.0 - i**2 / (3.0 * nside**2)
        phi0[cap_n] = np.pi / (4.0 * i)
        j = 4 * nside - ring[cap_s]
        n_r[cap_s] = 4 * j
        z[cap_s] = -(1.0 - j**2 / (3.0 * nside**2))
        phi0[cap_s] = np.pi / (4.0 * j)
        k = ring[belt]
        n_r[belt] = 4 * nside
        z[belt] = 4.0 / 3.0 - 2.0 * k / (3.0 * nside)
        phi0[belt] = np.where((k - nside) % 2 == 0, np.pi / (4.0 * nside), 0.0)
        self.nside = nside
        self.npix = 12 * nside * nside
        self.nrings = int(ring.size)
        self.z = z
        self.npix_ring = n_r
        self.phi0 = phi0
        self.startpix = np.concatenate([[0], np.cumsum(n_r)[:-1]]).astype(np.int64)


# --------------------------------------------------------------------------- #
# per-m recursions (validated reference form) -- src/jht/_recursion.py:37-235
# --------------------------------------------------------------------------- #


def legendre_recurrence_coeffs(m: int, lmax: int):
    """a_l, b_l for l = m+1..lmax (``_recursion.py:37-50``)."""
    l = np.arange(m + 1, lmax + 1, dtype=np.float64)
    a = np.sqrt((2.0 * l - 1.0) * (2.0 * l + 1.0) / ((l - m) * (l + m)))
    b = np.sqrt((2.0 * l + 1.0) * ((l - 1.0) ** 2 - m * m) / ((2.0 * l - 3.0) * (l - m) * (l + m)))
    return a, b


def sectoral_seed_logabs(m: int) -> float:
    """theta-independent part of log|lambda_mm| (``_recursion.py:53-56``)."""
    k = np.arange(1, m + 1, dtype=np.float64)
    return float(-0.5 * np.log(FOUR_PI) + 0.5 * np.sum(np.log((2.0 * k + 1.0) / (2.0 * k))))


def _renorm_step(raw, p, s):
    """Pair-norm log-renormalisation shared by every recursion (``_recursion.py:104-113``)."""
    r = np.sqrt(raw * raw + p * p)
    r = np.where(r == 0.0, 1.0, r)
    return raw / r, p / r, s + np.log(r)


def normalized_legendre(x, m: int, lmax: int, precise: bool = False) -> np.ndarray:
    """lambda_{l,m}(theta), l = m..lmax, shape (lmax-m+1, T) (``_recursion.py:59-117``).

    ``precise=True`` runs the same recursion and seed in long double (64-bit mantissa) and
    rounds the result to float64 once: the high-precision check of the float64 form."""
    dt = np.longdouble if precise else np.float64
    x = np.asarray(x, dtype=np.float64).astype(dt)
    if precise:  # 1 - x is exact: no cancellation next to the poles
        sin_t = np.sqrt(np.clip((1.0 - x) * (1.0 + x), 0.0, 1.0))
    else:
        sin_t = np.sqrt(np.clip(1.0 - x * x, 0.0, 1.0))
    if precise:
        k = np.arange(1, m + 1, dtype=dt)
        c = -0.5 * np.log(4 * np.longdouble(np.pi)) + 0.5 * np.sum(np.log((2 * k + 1) / (2 * k)))
    else:
        c = sectoral_seed_logabs(m)
    if m == 0:
        s = np.full(x.shape, c, dtype=dt)
    else:
        with np.errstate(divide="ignore"):
            s = c + m * np.log(sin_t)
    p = np.full(x.shape, 1.0 if m % 2 == 0 else -1.0, dtype=dt)
    q = np.zeros_like(x)
    out = [p * np.exp(s)]
    if precise:
        l = np.arange(m + 1, lmax + 1, dtype=dt)
        a = np.sqrt((2 * l - 1) * (2 * l + 1) / ((l - m) * (l + m)))
        b = np.sqrt((2 * l + 1) * ((l - 1) ** 2 - m * m) / ((2 * l - 3) * (l - m) * (l + m)))
    else:
        a, b = legendre_recurrence_coeffs(m, lmax)
    for i in range(a.size):
        raw = a[i] * x * p - b[i] * q
        p, q, s = _renorm_step(raw, p, s)
        out.append(p * np.exp(s))
    return np.array(out).astype(np.float64)


def wigner_recurrence_coeffs(M: int, Mp: int, lmin: int, lmax: int):
    """A_l, C_l, B_l for l = lmin..lmax-1 (each produces d^{l+1}) (``_recursion.py:136-143``)."""
    l = np.arange(lmin, lmax, dtype=np.float64)
    D = l * np.sqrt(((l + 1.0) ** 2 - M * M) * ((l + 1.0) ** 2 - Mp * Mp))
    A = (2.0 * l + 1.0) * l * (l + 1.0) / D
    C = -(2.0 * l + 1.0) * M * Mp / D
    B = (l + 1.0) * np.sqrt((l * l - M * M) * (l * l - Mp * Mp)) / D
    return A, C, B


@lru_cache(maxsize=None)
def _exact_log_fac(M: int, Mp: int, lmin: int, k0: int) -> float:
    """The seed's log-factorial ratio evaluated in mpmath (dps 40)."""
    import mpmath as mp

    with mp.workdps(40):
        lg = mp.loggamma
        v = mp.mpf(0.5) * (lg(lmin + M + 1) + lg(lmin - M + 1) + lg(lmin + Mp + 1) + lg(lmin - Mp + 1)) - (
            lg(k0 + 1) + lg(lmin + M - k0 + 1) + lg(lmin - Mp - k0 + 1) + lg(k0 - M + Mp + 1))
        return mp.nstr(v, 30)


def wigner_seed_params(M: int, Mp: int, lmin: int, precise: bool = False):
    """Single-term seed of d^{lmin}_{M,M'}: (sign, log_fac, pow_cos, pow_sin) (``_recursion.py:166-178``).

    ``precise=False`` is the reference's arithmetic: eight float64 ``lgamma`` values of
    arguments up to 2 lmin, whose rounding leaves up to ~1e-11 relative error in the seed
    near lmin = 4096 (``tests/test_oracle.py`` measures it against mpmath).  ``precise=True``
    takes the same log-factorial ratio from mpmath, as a long double."""
    k0 = max(0, M - Mp)
    if precise:
        log_fac = np.longdouble(_exact_log_fac(M, Mp, lmin, k0))
    else:
        lg = math.lgamma
        log_fac = 0.5 * (lg(lmin + M + 1) + lg(lmin - M + 1) + lg(lmin + Mp + 1) + lg(lmin - Mp + 1)) - (
            lg(k0 + 1) + lg(lmin + M - k0 + 1) + lg(lmin - Mp - k0 + 1) + lg(k0 - M + Mp + 1)
        )
    sign = -1.0 if (k0 % 2) else 1.0
    pow_cos = 2 * lmin - 2 * k0 + M - Mp
    pow_sin = 2 * k0 - M + Mp
    return sign, log_fac, pow_cos, pow_sin


def wigner_seed(x, M: int, Mp: int, lmin: int, precise: bool = False):
    """(sign, log|d^{lmin}_{M,M'}(theta)|) (``_recursion.py:166-188``); ``precise``: in long
    double from the exact log-factorial ratio."""
    dt = np.longdouble if precise else np.float64
    x = np.asarray(x, dtype=np.float64).astype(dt)
    sign, log_fac, pc, ps = wigner_seed_params(M, Mp, lmin, precise)
    cos_h = np.sqrt(np.clip((1.0 + x) / 2.0, 0.0, 1.0))
    sin_h = np.sqrt(np.clip((1.0 - x) / 2.0, 0.0, 1.0))
    out = np.full(x.shape, log_fac, dtype=dt)
    with np.errstate(divide="ignore"):
        if pc > 0:
            out = out + pc * np.log(cos_h)
        if ps > 0:
            out = out + ps * np.log(sin_h)
    return sign, out


def wigner_d_column(x, M: int, Mp: int, lmax: int, precise: bool = False) -> np.ndarray:
    """d^l_{M,M'}(theta), l = max(|M|,|M'|)..lmax (``_recursion.py:191-221``).

    ``precise=True`` runs the same recursion with every quantity in long double (64-bit
    mantissa) from the exact seed; the result is rounded to float64 once.  It is the
    high-precision check of the float64 restatement, whose log-scale carry alone leaves
    ~|log| 1e-16 relative error (5e-13 where the seed is e^-2800)."""
    dt = np.longdouble if precise else np.float64
    x = np.asarray(x, dtype=np.float64).astype(dt)
    lmin = max(abs(M), abs(Mp))
    if lmax < lmin:
        raise ValueError(f"lmax={lmax} < lmin={lmin}")
    sign, s = wigner_seed(x, M, Mp, lmin, precise)
    p = np.full(x.shape, sign, dtype=dt)
    q = np.zeros_like(x)
    out = [p * np.exp(s)]
    if precise:
        l = np.arange(lmin, lmax, dtype=dt)
        D = l * np.sqrt(((l + 1) ** 2 - M * M) * ((l + 1) ** 2 - Mp * Mp))
        A = (2 * l + 1) * l * (l + 1) / D
        C = -(2 * l + 1) * M * Mp / D
        B = (l + 1) * np.sqrt((l * l - M * M) * (l * l - Mp * Mp)) / D
    else:
        A, C, B = wigner_recurrence_coeffs(M, Mp, lmin, lmax)
    for i in range(A.size):
        raw = (A[i] * x + C[i]) * p - B[i] * q
        p, q, s = _renorm_step(raw, p, s)
        out.append(p * np.exp(s))
    return np.array(out).astype(np.float64)


def spin_weighted_lambda(x, m: int, spin: int, lmax: int, precise: bool = False) -> np.ndarray:
    """_s lambda_{l,m}(theta) = (-1)^m sqrt((2l+1)/4pi) d^l_{-m,s}, l = max(m,|s|)..lmax
    (``_recursion.py:224-235``).  ``precise``: see :func:`wigner_d_column`."""
    lmin = max(m, abs(spin))
    if precise:
        d = wigner_d_column(x, -m, spin, lmax, True).astype(np.longdouble)
        l = np.arange(lmin, lmax + 1, dtype=np.longdouble)
        pref = ((-1.0) ** m) * np.sqrt((2 * l + 1) / (4 * np.longdouble(np.pi)))
        return (pref[:, None] * d).astype(np.float64)
    d = wigner_d_column(x, -m, spin, lmax)
    l = np.arange(lmin, lmax + 1, dtype=np.float64)
    pref = ((-1.0) ** m) * np.sqrt((2.0 * l + 1.0) / FOUR_PI)
    return pref[:, None] * d


# --------------------------------------------------------------------------- #
# fused all-m recursion plan + N/S-symmetric contractions -- _recursion.py:257-572
# --------------------------------------------------------------------------- #


class RecursionPlan:
    """Dense (L1, M) tables, l-major, indexed by the l each entry *produces*
    (``_recursion.py:257-335``)."""

    def __init__(self, x, spin: int, lmax: int):
        x = np.asarray(x, dtype=np.float64)
        M = L1 = lmax + 1
        s_abs = abs(spin)
        lmin = np.maximum(np.arange(M), s_abs)
        A = np.zeros((L1, M))
        B = np.zeros((L1, M))
        C = np.zeros((L1, M))
        seed_sign = np.zeros(M)
        seed_log = np.zeros((M, x.size))
        if spin == 0:
            pref = np.ones((L1, M))
            sin_t = np.sqrt(np.clip(1.0 - x * x, 0.0, 1.0))
            with np.errstate(divide="ignore"):
                log_sin = np.log(sin_t)
            for m in range(M):
                a, b = legendre_recurrence_coeffs(m, lmax)
                A[m + 1 :, m] = a
                B[m + 1 :, m] = b
                seed_sign[m] = 1.0 if m % 2 == 0 else -1.0
                c = sectoral_seed_logabs(m)
                seed_log[m] = c if m == 0 else c + m * log_sin
        else:
            l = np.arange(L1, dtype=np.float64)
            pref = ((-1.0) ** np.arange(M))[None, :] * np.sqrt((2.0 * l + 1.0) / FOUR_PI)[:, None]
            for m in range(M):
                lm = int(lmin[m])
                if lm < lmax:
                    Ac, Cc, Bc = wigner_recurrence_coeffs(-m, spin, lm, lmax)
                    A[lm + 1 :, m] = Ac
                    C[lm + 1 :, m] = Cc
                    B[lm + 1 :, m] = Bc
                seed_sign[m], seed_log[m] = wigner_seed(x, -m, spin, lm)
        self.A, self.B, self.C, self.pref = A, B, C, pref
        self.seed_sign, self.seed_log = seed_sign, seed_log
        self.lmin = lmin
        self.spin, self.lmax, self.x = spin, lmax, x


class _Scan:
    """Carry of the all-m recursion.  Row m only changes once l >= lmin(m), so each
    step touches rows ``[0, nact)`` -- element-wise identical to the reference's
    masked update (``_rec_step``, ``_recursion.py:352-369``), half the work."""

    def __init__(self, plan: RecursionPlan):
        M, T = plan.seed_log.shape
        self.plan = plan
        self.p = np.zeros((M, T))
        self.q = np.zeros((M, T))
        self.s = np.zeros((M, T))

    def step(self, l: int):
        """Advance to degree l; returns (nact, lam[:nact]) with lam = 0 for l < lmin."""
        pl = self.plan
        nact = min(l + 1, pl.lmin.size) if l >= abs(pl.spin) else 0
        if nact == 0:
            return 0, None
        sl = slice(0, nact)
        p, q, s = self.p[sl], self.q[sl], self.s[sl]
        lmin = pl.lmin[sl]
        seed_rows = lmin == l
        rec_rows = lmin < l
        x = pl.x[None, :]
        A = pl.A[l, sl][:, None]
        B = pl.B[l, sl][:, None]
        C = pl.C[l, sl][:, None]
        raw = (A * x + C) * p - B * q
        rp, rq, rs = _renorm_step(raw, p, s)
        rc = rec_rows[:, None]
        sd = seed_rows[:, None]
        new_p = np.where(sd, pl.seed_sign[sl][:, None], np.where(rc, rp, p))
        new_q = np.where(sd, 0.0, np.where(rc, rq, q))
        new_s = np.where(sd, pl.seed_log[sl], np.where(rc, rs, s))
        self.p[sl], self.q[sl], self.s[sl] = new_p, new_q, new_s
        lam = new_p * np.exp(new_s)
        return nact, lam


def synth_contract_eo(plan: RecursionPlan, alm_dense: np.ndarray):
    """Spin-0 half-grid synthesis: (F_total, F_signed) (``_recursion.py:441-469``)."""
    M, T = plan.seed_log.shape
    Ftot = np.zeros((M, T), dtype=np.complex128)
    Fsig = np.zeros((M, T), dtype=np.complex128)
    sc = _Scan(plan)
    with np.errstate(under="ignore"):
        for l in range(plan.lmax + 1):
            nact, lam = sc.step(l)
            if nact == 0:
                continue
            c = (alm_dense[:nact, l] * plan.pref[l, :nact])[:, None] * lam
            Ftot[:nact] += c
            if l % 2:
                Fsig[:nact] -= c
            else:
                Fsig[:nact] += c
    return Ftot, Fsig


def adjoint_contract_eo(plan: RecursionPlan, Vn: np.ndarray, Vs: np.ndarray) -> np.ndarray:
    """Spin-0 half-grid adjoint: b_{l,m} = pref sum_t lam (Vn + (-1)^l Vs) (``_recursion.py:472-497``)."""
    M, T = plan.seed_log.shape
    b = np.zeros((M, plan.lmax + 1), dtype=np.complex128)
    sc = _Scan(plan)
    with np.errstate(under="ignore"):
        for l in range(plan.lmax + 1):
            nact, lam = sc.step(l)
            if nact == 0:
                continue
            par = -1.0 if l % 2 else 1.0
            tn = np.einsum("mt,mt->m", lam, Vn[:nact])
            ts = np.einsum("mt,mt->m", lam, Vs[:nact])
            b[:nact, l] = plan.pref[l, :nact] * (tn + par * ts)
    return b


def synth_contract_spin2_ns(plan_p, plan_m, bp_dense, bm_dense):
    """Spin-2 half-grid synthesis (``_recursion.py:512-543``): (FpN, FpS, FmN, FmS)."""
    M, T = plan_p.seed_log.shape
    FpN, FpS, FmN, FmS = (np.zeros((M, T), dtype=np.complex128) for _ in range(4))
    sp, sm = _Scan(plan_p), _Scan(plan_m)
    with np.errstate(under="ignore"):
        for l in range(plan_p.lmax + 1):
            nact, lam_p = sp.step(l)
            _, lam_m = sm.step(l)
            if nact == 0:
                continue
            par = -1.0 if l % 2 else 1.0
            cp = (bp_dense[:nact, l] * plan_p.pref[l, :nact])[:, None]
            cm = (bm_dense[:nact, l] * plan_p.pref[l, :nact])[:, None]
            FpN[:nact] += cp * lam_p
            FpS[:nact] += par * cp * lam_m
            FmN[:nact] += cm * lam_m
            FmS[:nact] += par * cm * lam_p
    return FpN, FpS, FmN, FmS


def adjoint_contract_spin2_ns(plan_p, plan_m, Vpn, Sp, Vmn, Sm):
    """Spin-2 half-grid adjoint (``_recursion.py:546-572``): (p2, m2) dense (M, lmax+1)."""
    M, T = plan_p.seed_log.shape
    p2 = np.zeros((M, plan_p.lmax + 1), dtype=np.complex128)
    m2 = np.zeros_like(p2)
    sp, sm = _Scan(plan_p), _Scan(plan_m)
    with np.errstate(under="ignore"):
        for l in range(plan_p.lmax + 1):
            nact, lam_p = sp.step(l)
            _, lam_m = sm.step(l)
            if nact == 0:
                continue
            par = -1.0 if l % 2 else 1.0
            pr = plan_p.pref[l, :nact]
            p2[:nact, l] = pr * (np.einsum("mt,mt->m", lam_p, Vpn[:nact]) + par * np.einsum("mt,mt->m", lam_m, Sp[:nact]))
            m2[:nact, l] = pr * (np.einsum("mt,mt->m", lam_m, Vmn[:nact]) + par * np.einsum("mt,mt->m", lam_p, Sm[:nact]))
    return p2, m2


# --------------------------------------------------------------------------- #
# ring-group azimuth assembly -- src/jht/healpix.py:143-173, 222-357
# --------------------------------------------------------------------------- #


@lru_cache(maxsize=16)
def _geometry(nside: int, lmax: int):
    geo = RingInfo(nside)
    m = np.arange(lmax + 1)
    phase = np.exp(1j * m[:, None] * geo.phi0[None, :])  # (M, nrings)
    T = 2 * nside
    south_src = np.arange(T - 1)
    south_tgt = 4 * nside - 2 - south_src
    sign_m = ((-1.0) ** m)[:, None]
    return geo, phase, T, south_src, south_tgt, sign_m


def _build_full(FN, FS, geo, T, south_src, south_tgt, sign_m):
    """North half + reflected south with (-1)^m (``healpix.py:250-253``)."""
    F = np.zeros((FN.shape[0], geo.nrings), dtype=np.complex128)
    F[:, :T] = FN
    F[:, south_tgt] = sign_m * FS[:, south_src]
    return F


def _fold_south(V, T, south_src, south_tgt, sign_m):
    """(M, nrings) -> (Vn, Vs = (-1)^m V_south, equator column zero) (``healpix.py:255-258``)."""
    Vs = np.zeros((V.shape[0], T), dtype=np.complex128)
    Vs[:, south_src] = V[:, south_tgt]
    return V[:, :T], sign_m * Vs


def ring_synthesis(Gp: np.ndarray, Gm: np.ndarray, geo: RingInfo) -> np.ndarray:
    """Per ring: D_k = sum_{m = k mod N} Gp_m + sum_{m>=1, -m = k mod N} conj(Gm_m);
    ring = ifft(D) * N (``healpix.py:285-290,323-329``).  Returns complex (npix,)."""
    M = Gp.shape[0]
    m = np.arange(M)
    out = np.empty(geo.npix, dtype=np.complex128)
    for N in np.unique(geo.npix_ring):
        rings = np.flatnonzero(geo.npix_ring == N)
        D = np.zeros((rings.size, N), dtype=np.complex128)
        np.add.at(D, (slice(None), m % N), Gp[:, rings].T)
        np.add.at(D, (slice(None), (-m[1:]) % N), np.conj(Gm[1:, rings].T))
        vals = np.fft.ifft(D, axis=1) * N
        for j, r in enumerate(rings):
            s = geo.startpix[r]
            out[s : s + N] = vals[j]
    return out


def ring_analysis(maps_c: np.ndarray, geo: RingInfo, lmax: int, minus: bool = False) -> np.ndarray:
    """Per ring: W = fft(ring); V[m] = W[m mod N] (or conj(W[-m mod N]) when ``minus``)
    (``healpix.py:299-302,343-347``) -- before the e^{-i m phi0} phase.  (M, nrings)."""
    M = lmax + 1
    m = np.arange(M)
    V = np.empty((M, geo.nrings), dtype=np.complex128)
    for N in np.unique(geo.npix_ring):
        rings = np.flatnonzero(geo.npix_ring == N)
        pix = geo.startpix[rings][:, None] + np.arange(N)[None, :]
        W = np.fft.fft(maps_c[pix], axis=1)
        if minus:
            V[:, rings] = np.conj(W[:, (-m) % N]).T
        else:
            V[:, rings] = W[:, m % N].T
    return V


# --------------------------------------------------------------------------- #
# public operators (fused form) -- src/jht/healpix.py:363-412
# --------------------------------------------------------------------------- #

_PLAN_CACHE: dict = {}


def _plans(nside: int, lmax: int, spin: int):
    key = (nside, lmax, spin)
    if key not in _PLAN_CACHE:
        geo = RingInfo(nside)
        x = geo.z[: 2 * nside]
        if spin == 0:
            _PLAN_CACHE[key] = (RecursionPlan(x, 0, lmax),)
        else:
            _PLAN_CACHE[key] = (RecursionPlan(x, 2, lmax), RecursionPlan(x, -2, lmax))
        if len(_PLAN_CACHE) > 6:
            _PLAN_CACHE.pop(next(iter(_PLAN_CACHE)))
    return _PLAN_CACHE[key]


def synthesis(alm, nside: int, lmax: int, spin: int = 0) -> np.ndarray:
    """S: a_lm -> RING map (fused form of ``healpix.py:277-291,309-330``)."""
    if spin not in (0, 2):
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    alm = np.asarray(alm, dtype=np.complex128)
    geo, phase, T, ssrc, stgt, sgn = _geometry(nside, lmax)
    if spin == 0:
        (plan,) = _plans(nside, lmax, 0)
        Ftot, Fsig = synth_contract_eo(plan, tri_to_dense(alm, lmax))
        G = _build_full(Ftot, Fsig, geo, T, ssrc, stgt, sgn) * phase
        return np.real(ring_synthesis(G, G, geo))
    plan_p, plan_m = _plans(nside, lmax, 2)
    aE = tri_to_dense(alm[0], lmax)
    aB = tri_to_dense(alm[1], lmax)
    FpN, FpS, FmN, FmS = synth_contract_spin2_ns(plan_p, plan_m, -(aE + 1j * aB), -(aE - 1j * aB))
    Fp = _build_full(FpN, FpS, geo, T, ssrc, stgt, sgn) * phase
    Fm = _build_full(FmN, FmS, geo, T, ssrc, stgt, sgn) * phase
    qu = ring_synthesis(Fp, Fm, geo)
    return np.stack([qu.real, qu.imag])


def adjoint_synthesis(maps, nside: int, lmax: int, spin: int = 0) -> np.ndarray:
    """S^T: RING map -> a_lm (fused form of ``healpix.py:293-305,332-355``)."""
    if spin not in (0, 2):
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    maps = np.asarray(maps, dtype=np.float64)
    geo, phase, T, ssrc, stgt, sgn = _geometry(nside, lmax)
    cph = np.conj(phase)
    if spin == 0:
        (plan,) = _plans(nside, lmax, 0)
        V = ring_analysis(maps.astype(np.complex128), geo, lmax) * cph
        Vn, Vs = _fold_south(V, T, ssrc, stgt, sgn)
        return dense_to_tri(adjoint_contract_eo(plan, Vn, Vs), lmax)
    plan_p, plan_m = _plans(nside, lmax, 2)
    qu = maps[0] + 1j * maps[1]
    Vp = ring_analysis(qu, geo, lmax) * cph
    Vm = ring_analysis(qu, geo, lmax, minus=True) * cph
    Vpn, Sp = _fold_south(Vp, T, ssrc, stgt, sgn)
    Vmn, Sm = _fold_south(Vm, T, ssrc, stgt, sgn)
    p2, m2 = adjoint_contract_spin2_ns(plan_p, plan_m, Vpn, Sp, Vmn, Sm)
    aE = 0.5 * (-p2 - m2)
    aB = 0.5 * (1j * p2 - 1j * m2)
    return np.stack([dense_to_tri(aE, lmax), dense_to_tri(aB, lmax)])


# --------------------------------------------------------------------------- #
# eager per-m / per-ring form -- src/jht/_reference.py:22-143
# --------------------------------------------------------------------------- #


def synthesis_reference(alm, nside: int, lmax: int, spin: int = 0) -> np.ndarray:
    """Per-m recursion over the full ring set, then one FFT per ring."""
    geo = RingInfo(nside)
    x = geo.z
    mm = np.arange(lmax + 1)
    alm = np.asarray(alm, dtype=np.complex128)
    if spin == 0:
        F = np.zeros((lmax + 1, geo.nrings), dtype=np.complex128)
        for m in range(lmax + 1):
            b = alm_column_base(m, lmax)
            F[m] = alm[b : b + lmax + 1 - m] @ normalized_legendre(x, m, lmax)
        out = []
        for r in range(geo.nrings):
            N = int(geo.npix_ring[r])
            G = F[:, r] * np.exp(1j * mm * geo.phi0[r])
            D = np.zeros(N, dtype=np.complex128)
            np.add.at(D, mm % N, G)
            np.add.at(D, (-mm[1:]) % N, np.conj(G[1:]))
            out.append(np.real(np.fft.ifft(D) * N))
        return np.concatenate(out)
    if spin != 2:
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    Fp = np.zeros((lmax + 1, geo.nrings), dtype=np.complex128)
    Fm = np.zeros_like(Fp)
    for m in range(lmax + 1):
        lmin = max(m, 2)
        b = alm_column_base(m, lmax) + (lmin - m)
        e = alm_column_base(m, lmax) + (lmax + 1 - m)
        aE, aB = alm[0][b:e], alm[1][b:e]
        Fp[m] = -(aE + 1j * aB) @ spin_weighted_lambda(x, m, 2, lmax)
        Fm[m] = -(aE - 1j * aB) @ spin_weighted_lambda(x, m, -2, lmax)
    q, u = [], []
    for r in range(geo.nrings):
        N = int(geo.npix_ring[r])
        ph = np.exp(1j * mm * geo.phi0[r])
        D = np.zeros(N, dtype=np.complex128)
        np.add.at(D, mm % N, Fp[:, r] * ph)
        np.add.at(D, (-mm[1:]) % N, np.conj(Fm[1:, r] * ph[1:]))
        v = np.fft.ifft(D) * N
        q.append(v.real)
        u.append(v.imag)
    return np.stack([np.concatenate(q), np.concatenate(u)])


def adjoint_synthesis_reference(maps, nside: int, lmax: int, spin: int = 0) -> np.ndarray:
    geo = RingInfo(nside)
    x = geo.z
    mm = np.arange(lmax + 1)
    maps = np.asarray(maps, dtype=np.float64)
    if spin == 0:
        V = np.zeros((lmax + 1, geo.nrings), dtype=np.complex128)
        for r in range(geo.nrings):
            N = int(geo.npix_ring[r])
            s = geo.startpix[r]
            V[:, r] = np.fft.fft(maps[s : s + N])[mm % N] * np.exp(-1j * mm * geo.phi0[r])
        return np.concatenate([normalized_legendre(x, m, lmax) @ V[m] for m in range(lmax + 1)])
    if spin != 2:
        raise NotImplementedError(f"spin={spin} unsupported (only 0 and 2)")
    Vp = np.zeros((lmax + 1, geo.nrings), dtype=np.complex128)
    Vm = np.zeros_like(Vp)
    for r in range(geo.nrings):
        N = int(geo.npix_ring[r])
        s = geo.startpix[r]
        W = np.fft.fft(maps[0][s : s + N] + 1j * maps[1][s : s + N])
        ph = np.exp(-1j * mm * geo.phi0[r])
        Vp[:, r] = W[mm % N] * ph
        Vm[:, r] = np.conj(W[(-mm) % N]) * ph
    aE, aB = [], []
    for m in range(lmax + 1):
        lmin = max(m, 2)
        p2 = spin_weighted_lambda(x, m, 2, lmax) @ Vp[m]
        m2 = spin_weighted_lambda(x, m, -2, lmax) @ Vm[m]
        pad = np.zeros(lmin - m, dtype=np.complex128)
        aE.append(np.concatenate([pad, 0.5 * (-p2 - m2)]))
        aB.append(np.concatenate([pad, 0.5 * (1j * p2 - 1j * m2)]))
    return np.stack([np.concatenate(aE), np.concatenate(aB)])


# --------------------------------------------------------------------------- #
# chirp-z (Bluestein) pruned DFT -- src/jht/_azimuth.py:150-191, _nufft.py:_next_size
# --------------------------------------------------------------------------- #


def next_smooth_even(n: float) -> int:
    """Smallest even 2,3,5-smooth integer >= n (``_nufft.py:_next_size``)."""
    n = int(np.ceil(n))
    n += n % 2
    while True:
        r = n
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return n
        n += 2


def chirp(sq: np.ndarray, N: int, sign: int) -> np.ndarray:
    """exp(sign * i pi (sq mod 2N) / N) with exact integer reduction (``_azimuth.py:175-178``)."""
    r = np.mod(np.asarray(sq, dtype=np.int64), 2 * N)
    return np.exp(sign * 1j * np.pi * r / N)


def bluestein_pruned_dft(x: np.ndarray, N: int, Q: int, sign: int, L: int | None = None) -> np.ndarray:
    """X[q] = sum_{p<P} x[p] exp(sign 2 pi i p q / N), q < Q, via a length-L
    circular convolution (``_azimuth.py:150-191``).  ``x`` is (..., P)."""
    x = np.asarray(x, dtype=np.complex128)
    P = x.shape[-1]
    if L is None:
        L = next_smooth_even(P + Q - 1)
    p = np.arange(P, dtype=np.int64)
    q = np.arange(Q, dtype=np.int64)
    a = np.zeros(x.shape[:-1] + (L,), dtype=np.complex128)
    a[..., :P] = x * chirp(p * p, N, sign)
    k = np.zeros(L, dtype=np.complex128)
    k[:Q] = chirp(q * q, N, -sign)
    neg = np.arange(1, P, dtype=np.int64)
    k[L - neg] = chirp(neg * neg, N, -sign)
    conv = np.fft.ifft(np.fft.fft(a, axis=-1) * np.fft.fft(k), axis=-1)
    return conv[..., :Q] * chirp(q * q, N, sign)


def ring_synthesis_bluestein(Gp: np.ndarray, Gm: np.ndarray, geo: RingInfo) -> np.ndarray:
    """Cap rings through the pruned chirp-z DFT of the unfolded spectrum, belt rings
    through a plain FFT (looped mode, ``_azimuth.py:197-211`` + ``healpix.py:284-290``)."""
    M = Gp.shape[0]
    out = ring_synthesis(Gp, Gm, geo)  # belt values come from here
    cap = np.flatnonzero(geo.npix_ring < 4 * geo.nside)
    mnz = (np.arange(M) > 0).astype(np.float64)
    n_out = int(geo.npix_ring[cap].max()) if cap.size else 0
    L = next_smooth_even(M + n_out - 1) if cap.size else 0
    for r in cap:
        N = int(geo.npix_ring[r])
        A = bluestein_pruned_dft(Gp[:, r], N, N, +1, L)
        B = bluestein_pruned_dft(Gm[:, r] * mnz, N, N, +1, L)
        s = geo.startpix[r]
        out[s : s + N] = A + np.conj(B)
    return out


# --------------------------------------------------------------------------- #
# ring weights and analysis -- src/jht/weights.py:60-106, _analysis.py:49-63
# --------------------------------------------------------------------------- #


def ring_weights(nside: int) -> np.ndarray:
    """Minimum-norm ring-weight deviations w_i over the north half + equator
    (``weights.py:60-83``): sum_i c_i n_i (1 + w_i) P_l(z_i) = Npix delta_l0, even l <= 2 nside."""
    from numpy.polynomial.legendre import legvander

    geo = RingInfo(nside)
    T = 2 * nside
    z = geo.z[:T]
    n = geo.npix_ring[:T].astype(np.float64)
    c = np.full(T, 2.0)
    c[-1] = 1.0
    P = legvander(z, 2 * nside)[:, ::2]
    A = ((c * n)[:, None] * P).T
    rhs = -A.sum(axis=1)
    rhs[0] += geo.npix
    w, *_ = np.linalg.lstsq(A, rhs, rcond=None)
    return w


def pixel_weights(nside: int, use_weights: bool = True) -> np.ndarray:
    """(4 pi / Npix)(1 + w_ring) per pixel (``weights.py:86-106``)."""
    geo = RingInfo(nside)
    base = FOUR_PI / geo.npix
    if not use_weights:
        return np.full(geo.npix, base)
    w = ring_weights(nside)
    r = np.arange(geo.nrings)
    half = np.where(r < 2 * nside, r, 4 * nside - 2 - r)
    return base * (1.0 + np.repeat(w[half], geo.npix_ring))


def bare_analysis(maps, nside, lmax, spin=0, use_weights=True):
    """A0 = S^T W (``_analysis.py:49-52``)."""
    w = pixel_weights(nside, use_weights)
    return adjoint_synthesis(w * np.asarray(maps), nside, lmax, spin)


def analysis(maps, nside, lmax, spin=0, niter=3, use_weights=True):
    """Jacobi refinement a <- a + A0 (map - S a) (``_analysis.py:55-63``)."""
    a = bare_analysis(maps, nside, lmax, spin, use_weights)
    for _ in range(niter):
        res = np.asarray(maps) - synthesis(a, nside, lmax, spin)
        a = a + bare_analysis(res, nside, lmax, spin, use_weights)
    return a


# --------------------------------------------------------------------------- #
# synthetic inputs used by the tests and the bench (the BASELINE recipe)
# --------------------------------------------------------------------------- #


def gaussian_alm(lmax: int, rng, lmin: int = 0, amp: float = 1.0) -> np.ndarray:
    """Re, Im ~ N(0,1) sqrt(C_l/2), C_l = amp (l+1)^-2; m = 0 real with variance C_l;
    l < lmin zero."""
    ell = alm_l_of_index(lmax)
    m = alm_m_of_index(lmax)
    cl = amp * (ell + 1.0) ** -2
    K = alm_size(lmax)
    re = rng.standard_normal(K)
    im = rng.standard_normal(K)
    a = np.where(m == 0, re * np.sqrt(cl), re * np.sqrt(cl / 2)) + 1j * np.where(m == 0, 0.0, im * np.sqrt(cl / 2))
    a[ell < lmin] = 0.0
    return a.astype(np.complex128)


def random_alm(lmax: int, rng, lmin: int = 0) -> np.ndarray:
    """The reference tests' white alm: (N(0,1) + i N(0,1))/sqrt 2, m=0 real, l<lmin zero
    (``tests/test_healpix.py:34-44``)."""
    K = alm_size(lmax)
    a = (rng.standard_normal(K) + 1j * rng.standard_normal(K)) / np.sqrt(2)
    a[: lmax + 1] = a[: lmax + 1].real
    ell = alm_l_of_index(lmax)
    a[ell < lmin] = 0.0
    return a.astype(np.complex128)


# --------------------------------------------------------------------------- #
# row-sampled masked scan (CPU baseline timing) -- the reference's _rec_step
# (_recursion.py:352-369) over every l for a subset of orders m, exactly as the fused
# lax.scan runs it (inactive steps masked, not skipped), spin 2 (_recursion.py:512-572).
# --------------------------------------------------------------------------- #


class _RowTables:
    """Per-row recursion tables of the sampled scans: the Legendre recursion and sectoral
    seed for spin 0 (``_recursion.py:37-56``), the Wigner-d recursion and single-term seed
    with the (-1)^m sqrt((2l+1)/4pi) prefactor for spin +-2 (``:136-188``)."""

    def __init__(self, x, spin, lmax, ms):
        ms = np.asarray(ms, dtype=np.int64)
        L1, R = lmax + 1, ms.size
        self.A = np.zeros((L1, R))
        self.B = np.zeros((L1, R))
        self.C = np.zeros((L1, R))
        self.seed_sign = np.zeros(R)
        self.seed_log = np.zeros((R, x.size))
        self.lmin = np.maximum(ms, abs(spin))
        sin_t = np.sqrt(np.clip(1.0 - x * x, 0.0, 1.0))
        for i, m in enumerate(ms):
            lm = int(self.lmin[i])
            if spin == 0:
                if lm < lmax:
                    self.A[lm + 1:, i], self.B[lm + 1:, i] = legendre_recurrence_coeffs(int(m), lmax)
                self.seed_sign[i] = -1.0 if m % 2 else 1.0
                with np.errstate(divide="ignore"):
                    self.seed_log[i] = sectoral_seed_logabs(int(m)) + (m * np.log(sin_t) if m else 0.0)
            else:
                if lm < lmax:
                    Ac, Cc, Bc = wigner_recurrence_coeffs(-int(m), spin, lm, lmax)
                    self.A[lm + 1:, i], self.C[lm + 1:, i], self.B[lm + 1:, i] = Ac, Cc, Bc
                self.seed_sign[i], self.seed_log[i] = wigner_seed(x, -int(m), spin, lm)
        l = np.arange(L1, dtype=np.float64)
        if spin == 0:
            self.pref = np.ones((L1, R))
        else:
            self.pref = ((-1.0) ** ms)[None, :] * np.sqrt((2.0 * l + 1.0) / FOUR_PI)[:, None]


def _masked_step(tb, l, x, p, q, s):
    A = tb.A[l][:, None]
    B = tb.B[l][:, None]
    C = tb.C[l][:, None]
    raw = (A * x[None, :] + C) * p - B * q
    rp, rq, rs = _renorm_step(raw, p, s)
    seed = (tb.lmin == l)[:, None]
    rec = (tb.lmin < l)[:, None]
    act = (tb.lmin <= l)[:, None]
    p2 = np.where(seed, tb.seed_sign[:, None], np.where(rec, rp, p))
    q2 = np.where(seed, 0.0, np.where(rec, rq, q))
    s2 = np.where(seed, tb.seed_log, np.where(rec, rs, s))
    lam = np.where(act, p2 * np.exp(s2), 0.0)
    return p2, q2, s2, lam


def sampled_synthesis_scan(x, lmax, ms, spin, coef):
    """The fused synthesis scan restricted to the orders ``ms``, every step masked as the
    reference's lax.scan runs it (spin 0: ``_recursion.py:441-469``; spin 2, both +-2
    recursions in one scan: ``:512-543``).  coef: spin 0 the (R, L1) a_lm rows; spin 2 the
    pair (bp, bm).  Returns the accumulators (spin 0: Ftot, Fsig; spin 2: FpN, FpS, FmN, FmS)."""
    x = np.asarray(x, dtype=np.float64)
    R, T = len(ms), x.size
    tabs = [_RowTables(x, s, lmax, ms) for s in ((0,) if spin == 0 else (2, -2))]
    st = [[np.zeros((R, T)) for _ in range(3)] for _ in tabs]
    F = [np.zeros((R, T), complex) for _ in range(2 if spin == 0 else 4)]
    with np.errstate(under="ignore", divide="ignore", invalid="ignore"):
        for l in range(lmax + 1):
            lams = []
            for k, tb in enumerate(tabs):
                *st[k], lam = _masked_step(tb, l, x, *st[k])
                lams.append(lam)
            par = -1.0 if l % 2 else 1.0
            pr = tabs[0].pref[l]
            if spin == 0:
                c = (coef[:, l] * pr)[:, None] * lams[0]
                F[0] += c
                F[1] += par * c
            else:
                cp = (coef[0][:, l] * pr)[:, None]
                cm = (coef[1][:, l] * pr)[:, None]
                F[0] += cp * lams[0]
                F[1] += par * cp * lams[1]
                F[2] += cm * lams[1]
                F[3] += par * cm * lams[0]
    return F


def sampled_adjoint_scan(x, lmax, ms, spin, V):
    """The fused adjoint scan restricted to the orders ``ms``, masked as the reference runs
    it (spin 0: ``_recursion.py:472-497``, V = (Vn, Vs); spin 2: ``:546-572``,
    V = (Vpn, Sp, Vmn, Sm), each (R, T)).  Returns the dense (R, L1) rows (spin 2: p2, m2)."""
    x = np.asarray(x, dtype=np.float64)
    R, T = len(ms), x.size
    tabs = [_RowTables(x, s, lmax, ms) for s in ((0,) if spin == 0 else (2, -2))]
    st = [[np.zeros((R, T)) for _ in range(3)] for _ in tabs]
    out = [np.zeros((R, lmax + 1), complex) for _ in range(1 if spin == 0 else 2)]
    with np.errstate(under="ignore", divide="ignore", invalid="ignore"):
        for l in range(lmax + 1):
            lams = []
            for k, tb in enumerate(tabs):
                *st[k], lam = _masked_step(tb, l, x, *st[k])
                lams.append(lam)
            par = -1.0 if l % 2 else 1.0
            pr = tabs[0].pref[l]
            dot = lambda a, b: np.einsum("rt,rt->r", a, b)  # noqa: E731
            if spin == 0:
                out[0][:, l] = pr * (dot(lams[0], V[0]) + par * dot(lams[0], V[1]))
            else:
                out[0][:, l] = pr * (dot(lams[0], V[0]) + par * dot(lams[1], V[1]))
                out[1][:, l] = pr * (dot(lams[1], V[2]) + par * dot(lams[0], V[3]))
    return out
