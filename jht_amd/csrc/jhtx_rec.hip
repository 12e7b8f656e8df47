// FP64 recursion + contraction kernels: the on-the-fly scaled associated-Legendre (spin 0)
// and Wigner-d (spin +-2) recursions over l for one order m and a tile of north-half
// rings, contracted against a_lm (synthesis) or reduced over rings (adjoint).
//
// Follows the algorithm of src/jht/_recursion.py:441-572 (synth_contract_eo,
// adjoint_contract_eo, synth_contract_spin2_ns, adjoint_contract_spin2_ns): one pass over
// the north half + equator gives both hemispheres through
//     lambda^s_lm(pi - theta) = (-1)^(l+m) lambda^-s_lm(theta).
// Differences of method (same operator, same tolerance):
//   * the three-term recursion runs on mu_l = lambda_l / G_l with the coefficient of
//     mu_{l-2} scaled to exactly 1, so each step is one FMA pair; G_l (and the spin-2
//     prefactor (-1)^m sqrt((2l+1)/4pi)) is folded into a_lm when it is staged;
//   * instead of jht's per-step log-renormalisation (_recursion.py:358-368) the seed is
//     carried as mantissa x 2^(512 ex) and rescaled by exact powers of two; while ex < 0
//     the true value is below 2^-256 and its contribution is dropped;
//   * only the active triangle l >= max(m, |s|) is iterated (no masked steps);
//   * lanes own rings (R per lane); the even/odd (l - m) accumulators give the north
//     value as their sum and the (-1)^m-reflected south value as their difference.
#include <vector>

#include "jhtx_internal.h"

namespace jhtx {

namespace {

// Ring coefficients F / V are complex128 in FP64 plans and complex64 in mixed plans (whose
// ring FFT runs in FP32).  The recursion and its sums stay FP64 in both; only the value
// stored to or loaded from F / V changes type.
__device__ __forceinline__ double2 ld_c(const double2& v) { return v; }
__device__ __forceinline__ double2 ld_c(const float2& v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ void st_c(double2* q, double2 v) { *q = v; }
__device__ __forceinline__ void st_c(float2* q, double2 v) { *q = make_float2((float)v.x, (float)v.y); }

constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 32 significant bits
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kLog2e = 1.44269504088896338700e+00;
constexpr int CH = 64;  // l values staged per LDS chunk

// exp(L) = sign * mu * 2^(512 ex), mu in [2^-256, 2^256), ex <= 0.
__device__ __forceinline__ void seed_split(double L, double sign, double& mu, int& ex) {
  if (!(L > -1.0e300)) {  // exact pole (never on the HEALPix grid)
    mu = 0.0;
    ex = 0;
    return;
  }
  int e = (int)floor((L * kLog2e + 256.0) * (1.0 / 512.0));
  if (e > 0) e = 0;
  const double n = (double)(kScaleStep * e);
  const double r = (L - n * kLn2Hi) - n * kLn2Lo;
  mu = sign * exp(r);
  ex = e;
}

// XCD-aware decode of the 1-D block index: consecutive physical blocks go round-robin
// over the 8 XCDs, so all ring chunks of one m are placed on the same XCD (shared L2 for
// the a_lm column and coefficients) and each XCD sweeps m in increasing order (longest
// columns first).
// Work-groups are dealt round-robin to the 8 XCDs (blockIdx & 7); consecutive orders m
// alternate between XCDs so cheap high-m and costly low-m columns spread evenly.
// GROUPED (the adjoint kernels): each XCD takes groups of kMGroup consecutive m, the
// groups alternating between XCDs, and within a group the work-groups of one ring chunk
// run back to back -- they read adjacent m of the same ring rows of V, so those cache
// lines are shared in the XCD's L2 instead of being fetched once per m (7x less fetch).
template <bool GROUPED>
__device__ __forceinline__ bool decode_block(const RecArgs& A, int& m, int& chunk) {
  const int P = blockIdx.x;
  const int xcd = P & 7;
  const int slot = P >> 3;
  if (GROUPED) {
    const int mi = slot % kMGroup;
    chunk = (slot / kMGroup) % A.nchunk;
    const int g = slot / (kMGroup * A.nchunk);
    m = (g * 8 + xcd) * kMGroup + mi;
  } else {
    chunk = slot % A.nchunk;
    m = (slot / A.nchunk) * 8 + xcd;
  }
  return m < A.M;
}

enum { FULL = 0, MASKED = 1, RECONLY = 2 };

__device__ __forceinline__ double2 operator+(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 operator-(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// ---------------------------------------------------------------------------------- //
// spin 0 synthesis
// ---------------------------------------------------------------------------------- //
template <int R, int MODE>
__device__ __forceinline__ void s0_synth_block(const double* s_al, const double2* s_a, int j0,
                                               const double (&x)[R], double (&mu)[R],
                                               double (&mup)[R], const double (&f)[R],
                                               double2 (&c0)[R], double2 (&c1)[R]) {
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const double al = s_al[j0 + jj];
    const double2 a = s_a[j0 + jj];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double c = al * x[r];
      const double mn = fma(c, mu[r], -mup[r]);
      mup[r] = mu[r];
      mu[r] = mn;
      if (MODE != RECONLY) {
        const double v = (MODE == MASKED) ? mn * f[r] : mn;
        double2& acc = (jj & 1) ? c1[r] : c0[r];
        acc.x = fma(a.x, v, acc.x);
        acc.y = fma(a.y, v, acc.y);
      }
    }
  }
}

template <int R>
__device__ __forceinline__ void rescale(double (&mu)[R], double (&mup)[R], int (&ex)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) rescale_pair(mu[r], mup[r], ex[r]);
}

template <int R, typename C>
__global__ __launch_bounds__(256, 2) void k_synth_s0(RecArgs A, const double2* __restrict__ alm,
                                                  C* __restrict__ F) {
  __shared__ double s_al[CH];
  __shared__ double2 s_a[CH];
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const double2* ab = alm + (long long)b * A.K;
  double x[R], mu[R], mup[R], f[R];
  int ex[R];
  double2 c0[R], c1[R];
  const int tb = (chunk * (int)blockDim.x + tid) * R;
  const double sgn = (m & 1) ? -1.0 : 1.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = tb + r;
    c0[r] = make_double2(0.0, 0.0);
    c1[r] = make_double2(0.0, 0.0);
    x[r] = 0.0;
    mu[r] = 0.0;
    mup[r] = 0.0;
    ex[r] = 0;
    if (t < A.T) {
      x[r] = A.x[t];
      const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
      double s;
      seed_split(L, sgn, s, ex[r]);
      mup[r] = -s;  // the first step (alpha = 0) yields mu = seed at l = m
    }
  }
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      double al = 0.0;
      double2 a = make_double2(0.0, 0.0);
      if (l <= lmax) {
        const long long idx = cb + l;
        al = (l == m) ? 0.0 : A.alpha[idx];
        const double g = A.G[idx];
        const double2 v = ab[idx];
        a = make_double2(v.x * g, v.y * g);
      }
      s_al[tid] = al;
      s_a[tid] = a;
    }
    __syncthreads();
    const int nst = min(CH, (lmax - l0 + 1 + 7) & ~7);
    for (int j0 = 0; j0 < nst; j0 += 8) {
      bool deep = false, alldeep = true;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        deep |= ex[r] < 0;
        alldeep &= ex[r] < 0;
      }
      if (__ballot(deep) == 0ull) {
        s0_synth_block<R, FULL>(s_al, s_a, j0, x, mu, mup, f, c0, c1);
      } else if (__ballot(!alldeep) == 0ull) {
        s0_synth_block<R, RECONLY>(s_al, s_a, j0, x, mu, mup, f, c0, c1);
        rescale<R>(mu, mup, ex);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) f[r] = (ex[r] == 0) ? 1.0 : 0.0;
        s0_synth_block<R, MASKED>(s_al, s_a, j0, x, mu, mup, f, c0, c1);
        rescale<R>(mu, mup, ex);
      }
    }
  }
  C* Fb = F + (long long)b * A.nrings * A.M;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = tb + r;
    if (t < A.T) {
      st_c(&Fb[(long long)t * A.M + m], c0[r] + c1[r]);
      if (t < A.T - 1) st_c(&Fb[(long long)(A.nrings - 1 - t) * A.M + m], c0[r] - c1[r]);
    }
  }
}

// ---------------------------------------------------------------------------------- //
// spin 2 synthesis
// ---------------------------------------------------------------------------------- //
struct S2Acc {
  double2 pe, pb, me, mb;  // sum aE G mu+, aB G mu+, aE G mu-, aB G mu-
};

template <int R, int MODE>
__device__ __forceinline__ void s2_synth_block(const double2* s_ag, const double2* s_e,
                                               const double2* s_b, int j0, const double (&x)[R],
                                               double (&pp)[R], double (&ppp)[R], double (&pm)[R],
                                               double (&pmp)[R], const double (&fp)[R],
                                               const double (&fm)[R], S2Acc (&c0)[R],
                                               S2Acc (&c1)[R]) {
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const double2 ag = s_ag[j0 + jj];
    const double2 e = s_e[j0 + jj];
    const double2 bb = s_b[j0 + jj];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double cp = fma(ag.x, x[r], ag.y);
      const double cm = fma(ag.x, x[r], -ag.y);
      const double np = fma(cp, pp[r], -ppp[r]);
      const double nm = fma(cm, pm[r], -pmp[r]);
      ppp[r] = pp[r];
      pp[r] = np;
      pmp[r] = pm[r];
      pm[r] = nm;
      if (MODE != RECONLY) {
        const double vp = (MODE == MASKED) ? np * fp[r] : np;
        const double vm = (MODE == MASKED) ? nm * fm[r] : nm;
        S2Acc& a = (jj & 1) ? c1[r] : c0[r];
        a.pe.x = fma(e.x, vp, a.pe.x);
        a.pe.y = fma(e.y, vp, a.pe.y);
        a.pb.x = fma(bb.x, vp, a.pb.x);
        a.pb.y = fma(bb.y, vp, a.pb.y);
        a.me.x = fma(e.x, vm, a.me.x);
        a.me.y = fma(e.y, vm, a.me.y);
        a.mb.x = fma(bb.x, vm, a.mb.x);
        a.mb.y = fma(bb.y, vm, a.mb.y);
      }
    }
  }
}

__device__ __forceinline__ double spin2_seed_L(const double* sp, double lc, double ls) {
  // sp = {centred log_fac, pow_cos, pow_sin, sign}; lc, ls: centred log cos, log sin of
  // theta/2 (the host tables), so L = log_fac + pc log cos(theta/2) + ps log sin(theta/2)
  double L = sp[0];
  if (sp[1] > 0.0) L = __dadd_rn(L, __dmul_rn(sp[1], lc));
  if (sp[2] > 0.0) L = __dadd_rn(L, __dmul_rn(sp[2], ls));
  return L;
}

template <int R, typename C>
__global__ __launch_bounds__(256) void k_synth_s2(RecArgs A, const double2* __restrict__ alm,
                                                  C* __restrict__ F) {
  __shared__ double2 s_ag[CH];
  __shared__ double2 s_e[CH];
  __shared__ double2 s_b[CH];
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int lmax = A.lmax;
  const int lmin = max(m, A.sabs);
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const double2* aE = alm + (long long)b * 2 * A.K;
  const double2* aB = aE + A.K;
  double x[R], pp[R], ppp[R], pm[R], pmp[R], fp[R], fm[R];
  int ep[R], em[R];
  S2Acc c0[R], c1[R];
  const int tb = (chunk * (int)blockDim.x + tid) * R;
  const double* sd = A.seed + 8 * m;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = tb + r;
    c0[r].pe = c0[r].pb = c0[r].me = c0[r].mb = make_double2(0.0, 0.0);
    c1[r] = c0[r];
    x[r] = 0.0;
    pp[r] = ppp[r] = pm[r] = pmp[r] = 0.0;
    ep[r] = em[r] = 0;
    if (t < A.T && lmin <= lmax) {
      x[r] = A.x[t];
      const double lc = A.lcos[t], ls = A.lsin[t];
      double s;
      seed_split(spin2_seed_L(sd, lc, ls), sd[3], s, ep[r]);
      ppp[r] = -s;
      seed_split(spin2_seed_L(sd + 4, lc, ls), sd[7], s, em[r]);
      pmp[r] = -s;
    }
  }
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      double2 ag = make_double2(0.0, 0.0), e = ag, bb = ag;
      if (l <= lmax) {
        const long long idx = cb + l;
        if (l != lmin) ag = make_double2(A.alpha[idx], A.gamma[idx]);
        const double g = A.G[idx];
        const double2 ve = aE[idx], vb = aB[idx];
        e = make_double2(ve.x * g, ve.y * g);
        bb = make_double2(vb.x * g, vb.y * g);
      }
      s_ag[tid] = ag;
      s_e[tid] = e;
      s_b[tid] = bb;
    }
    __syncthreads();
    const int nst = min(CH, (lmax - l0 + 1 + 7) & ~7);
    for (int j0 = 0; j0 < nst; j0 += 8) {
      bool deep = false, alldeep = true;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        deep |= (ep[r] < 0) | (em[r] < 0);
        alldeep &= (ep[r] < 0) & (em[r] < 0);
      }
      if (__ballot(deep) == 0ull) {
        s2_synth_block<R, FULL>(s_ag, s_e, s_b, j0, x, pp, ppp, pm, pmp, fp, fm, c0, c1);
      } else if (__ballot(!alldeep) == 0ull) {
        s2_synth_block<R, RECONLY>(s_ag, s_e, s_b, j0, x, pp, ppp, pm, pmp, fp, fm, c0, c1);
        rescale<R>(pp, ppp, ep);
        rescale<R>(pm, pmp, em);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          fp[r] = (ep[r] == 0) ? 1.0 : 0.0;
          fm[r] = (em[r] == 0) ? 1.0 : 0.0;
        }
        s2_synth_block<R, MASKED>(s_ag, s_e, s_b, j0, x, pp, ppp, pm, pmp, fp, fm, c0, c1);
        rescale<R>(pp, ppp, ep);
        rescale<R>(pm, pmp, em);
      }
    }
  }
  // bp = -(aE + i aB), bm = -(-1)^s (aE - i aB); south values carry (-1)^(l-m)
  const long long ringM = (long long)A.M;
  C* Fp = F + (long long)b * 2 * A.nrings * ringM;
  C* Fm = Fp + (long long)A.nrings * ringM;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = tb + r;
    if (t >= A.T) continue;
    if ((lmin - m) & 1) {  // classes are counted from lmin; the reflection needs (l - m)
      const S2Acc tmp = c0[r];
      c0[r] = c1[r];
      c1[r] = tmp;
    }
    const double2 PE = c0[r].pe + c1[r].pe, PB = c0[r].pb + c1[r].pb;
    const double2 ME = c0[r].me + c1[r].me, MB = c0[r].mb + c1[r].mb;
    // -PE - i PB  and  -ME + i MB
    const double sg = A.ssign;
    st_c(&Fp[(long long)t * ringM + m], make_double2(-PE.x + PB.y, -PE.y - PB.x));
    st_c(&Fm[(long long)t * ringM + m], make_double2(sg * (-ME.x - MB.y), sg * (-ME.y + MB.x)));
    if (t < A.T - 1) {
      const long long rs = (long long)(A.nrings - 1 - t) * ringM + m;
      const double2 dE = c0[r].me - c1[r].me, dB = c0[r].mb - c1[r].mb;  // from mu-
      const double2 qE = c0[r].pe - c1[r].pe, qB = c0[r].pb - c1[r].pb;  // from mu+
      st_c(&Fp[rs], make_double2(-dE.x + dB.y, -dE.y - dB.x));
      st_c(&Fm[rs], make_double2(sg * (-qE.x - qB.y), sg * (-qE.y + qB.x)));
    }
  }
}

// ---------------------------------------------------------------------------------- //
// adjoint: per-lane partial sums over the lane's rings for a block of l values, then a
// work-group reduction through LDS into one partial per (ring chunk, l, m).
// ---------------------------------------------------------------------------------- //
constexpr int NV = 16;  // double2 partials per lane per block

// Moves a double between lanes with a DPP row control (two 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 16 lanes of a DPP row, every lane receiving the total: row rotations by 8,
// 4, 2 and 1 lanes (each step doubles the set of lanes a value covers).
__device__ __forceinline__ double row_sum16(double v) {
  v += dpp_f64<0x128>(v);
  v += dpp_f64<0x124>(v);
  v += dpp_f64<0x122>(v);
  v += dpp_f64<0x121>(v);
  return v;
}

// Block reduction over the work-group's lanes (nthr a multiple of 64, at most 256) of the
// NV double2 values P[v] every lane holds, into out0/out1 (value v = jl * ncomp + comp).
// LDS red[v * 256 + tid] is written lane-contiguously; the barrier that publishes it also
// ORs the next block's activity flag.  Thread (v = tid / 16, s = tid % 16) then loads lanes
// s, s + 16, ... of value v in one batch of independent reads (consecutive threads read
// consecutive lanes), sums them as a tree, and the 16 partials of a value are summed
// across the DPP row.  A second barrier retires the reads before the next block
// overwrites red.  The floating-point order is fixed: results do not depend on the batch
// or the launch.  (Halving the LDS traffic first with gfx950 lane swaps measured no faster
// at nside 1024-2048: the reduction is not LDS-bound.)
constexpr int kRedSize = NV * 256;  // double2 elements of red

// Zeroes the slots of absent lanes (work-groups under 256 threads) so the reduction reads
// every slot unconditionally.  Published by the first barrier of wg_reduce_store.
__device__ __forceinline__ void red_zero_tail(double2* red, int tid, int nthr) {
  for (int L = nthr + tid; L < 256; L += nthr) {
#pragma unroll
    for (int v = 0; v < NV; ++v) red[v * 256 + L] = make_double2(0.0, 0.0);
  }
}

__device__ __forceinline__ int wg_reduce_store(double2 (&P)[NV], double2* red, double2* out0,
                                               double2* out1, int ncomp, int nl, int act_next,
                                               int tid, int nthr) {
#pragma unroll
  for (int v = 0; v < NV; ++v) red[v * 256 + tid] = P[v];
  const int flag = __syncthreads_or(act_next);
  const int sidx = tid & 15;
  for (int v = tid >> 4; v < NV; v += nthr >> 4) {
    double2 q[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) q[k] = red[v * 256 + sidx + 16 * k];
#pragma unroll
    for (int w = 8; w > 0; w >>= 1) {
#pragma unroll
      for (int k = 0; k < w; ++k) q[k] = q[k] + q[k + w];
    }
    const double2 acc = make_double2(row_sum16(q[0].x), row_sum16(q[0].y));
    const int jl = v / ncomp, comp = v - jl * ncomp;
    if (sidx == 0 && jl < nl) (comp == 0 ? out0 : out1)[jl] = acc;
  }
  __syncthreads();
  return flag;
}

template <int R, int MODE>
__device__ __forceinline__ void s0_adj_block(const double* s_al, int j0, const double (&x)[R],
                                             double (&mu)[R], double (&mup)[R],
                                             const double (&f)[R], const double2 (&U0)[R],
                                             const double2 (&U1)[R], double2 (&P)[NV]) {
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const double al = s_al[j0 + jj];
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double c = al * x[r];
      const double mn = fma(c, mu[r], -mup[r]);
      mup[r] = mu[r];
      mu[r] = mn;
      if (MODE != RECONLY) {
        const double v = (MODE == MASKED) ? mn * f[r] : mn;
        const double2 U = (jj & 1) ? U1[r] : U0[r];
        acc.x = fma(v, U.x, acc.x);
        acc.y = fma(v, U.y, acc.y);
      }
    }
    P[(j0 & 8) + jj] = acc;
  }
}

template <int R, typename C>
__global__ __launch_bounds__(256, 2) void k_adj_s0(RecArgs A, const C* __restrict__ V,
                                                double2* __restrict__ part) {
  __shared__ double s_al[CH];
  __shared__ double2 red[kRedSize];
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int b = blockIdx.y;
  const int batch = gridDim.y;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const C* Vb = V + (long long)b * A.nrings * ringM;
  double2* pb = part + ((long long)chunk * batch + b) * A.K + cb;
  double x[R], mu[R], mup[R], f[R];
  int ex[R];
  double2 U0[R], U1[R];
  const int tb = (chunk * nthr + tid) * R;
  const double sgn = (m & 1) ? -1.0 : 1.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = tb + r;
    x[r] = mu[r] = mup[r] = 0.0;
    ex[r] = 0;
    U0[r] = U1[r] = make_double2(0.0, 0.0);
    if (t < A.T) {
      x[r] = A.x[t];
      const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
      double s;
      seed_split(L, sgn, s, ex[r]);
      mup[r] = -s;
      const double2 vn = ld_c(Vb[(long long)t * ringM + m]);
      double2 vs = make_double2(0.0, 0.0);
      if (t < A.T - 1) vs = ld_c(Vb[(long long)(A.nrings - 1 - t) * ringM + m]);
      U0[r] = vn + vs;  // (l - m) even
      U1[r] = vn - vs;  // (l - m) odd
    }
  }
  double2 P[NV];
  auto active = [&]() {
    bool a = false;
#pragma unroll
    for (int r = 0; r < R; ++r) a |= (ex[r] == 0) && (tb + r < A.T);
    return (int)a;
  };
  red_zero_tail(red, tid, nthr);
  int wg_active = __syncthreads_or(active());
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      s_al[tid] = (l <= lmax && l != m) ? A.alpha[cb + l] : 0.0;
    }
    __syncthreads();
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += NV) {
      const int nl = min(NV, lmax - (l0 + jb) + 1);
      if (!wg_active) {
        s0_adj_block<R, RECONLY>(s_al, jb, x, mu, mup, f, U0, U1, P);
        rescale<R>(mu, mup, ex);
        s0_adj_block<R, RECONLY>(s_al, jb + 8, x, mu, mup, f, U0, U1, P);
        rescale<R>(mu, mup, ex);
        if (tid < nl) pb[l0 + jb + tid] = make_double2(0.0, 0.0);
        wg_active = __syncthreads_or(active());
        continue;
      }
      bool deep = false;
#pragma unroll
      for (int r = 0; r < R; ++r) deep |= ex[r] < 0;
      if (__ballot(deep) == 0ull) {
        s0_adj_block<R, FULL>(s_al, jb, x, mu, mup, f, U0, U1, P);
        s0_adj_block<R, FULL>(s_al, jb + 8, x, mu, mup, f, U0, U1, P);
      } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int r = 0; r < R; ++r) f[r] = (ex[r] == 0) ? 1.0 : 0.0;
          s0_adj_block<R, MASKED>(s_al, jb + 8 * h, x, mu, mup, f, U0, U1, P);
          rescale<R>(mu, mup, ex);
        }
      }
      wg_active = wg_reduce_store(P, red, pb + l0 + jb, nullptr, 1, nl, active(), tid, nthr);
    }
  }
}

template <int R, int MODE>
__device__ __forceinline__ void s2_adj_block(const double2* s_ag, int j0, const double (&x)[R],
                                             double (&pp)[R], double (&ppp)[R], double (&pm)[R],
                                             double (&pmp)[R], const double (&fp)[R],
                                             const double (&fm)[R], const double2 (&Vpn)[R],
                                             const double2 (&Vps)[R], const double2 (&Vmn)[R],
                                             const double2 (&Vms)[R], double2 (&P)[NV]) {
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const double2 ag = s_ag[j0 + jj];
    double2 a2 = make_double2(0.0, 0.0), b2 = a2;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double cp = fma(ag.x, x[r], ag.y);
      const double cm = fma(ag.x, x[r], -ag.y);
      const double np = fma(cp, pp[r], -ppp[r]);
      const double nm = fma(cm, pm[r], -pmp[r]);
      ppp[r] = pp[r];
      pp[r] = np;
      pmp[r] = pm[r];
      pm[r] = nm;
      if (MODE != RECONLY) {
        const double vp = (MODE == MASKED) ? np * fp[r] : np;
        const double vm = (MODE == MASKED) ? nm * fm[r] : nm;
        // p2 += mu+ Vpn + s mu- Vps ; m2 += mu- Vmn + s mu+ Vms ; s = (-1)^(l-m)
        const double svm = (jj & 1) ? -vm : vm;
        const double svp = (jj & 1) ? -vp : vp;
        a2.x = fma(vp, Vpn[r].x, a2.x);
        a2.y = fma(vp, Vpn[r].y, a2.y);
        a2.x = fma(svm, Vps[r].x, a2.x);
        a2.y = fma(svm, Vps[r].y, a2.y);
        b2.x = fma(vm, Vmn[r].x, b2.x);
        b2.y = fma(vm, Vmn[r].y, b2.y);
        b2.x = fma(svp, Vms[r].x, b2.x);
        b2.y = fma(svp, Vms[r].y, b2.y);
      }
    }
    P[2 * jj] = a2;
    P[2 * jj + 1] = b2;
  }
}

template <int R, typename C>
__global__ __launch_bounds__(256) void k_adj_s2(RecArgs A, const C* __restrict__ V,
                                                double2* __restrict__ part) {
  __shared__ double2 s_ag[CH];
  __shared__ double2 red[kRedSize];
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int b = blockIdx.y;
  const int batch = gridDim.y;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const int lmax = A.lmax;
  const int lmin = max(m, A.sabs);
  if (lmin > lmax) return;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const C* Vp = V + (long long)b * 2 * A.nrings * ringM;
  const C* Vm = Vp + (long long)A.nrings * ringM;
  double2* p_p2 = part + (((long long)chunk * batch + b) * 2 + 0) * A.K + cb;
  double2* p_m2 = part + (((long long)chunk * batch + b) * 2 + 1) * A.K + cb;
  double x[R], pp[R], ppp[R], pm[R], pmp[R], fp[R], fm[R];
  int ep[R], em[R];
  double2 Vpn[R], Vps[R], Vmn[R], Vms[R];
  const int tb = (chunk * nthr + tid) * R;
  const double* sd = A.seed + 8 * m;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = tb + r;
    x[r] = pp[r] = ppp[r] = pm[r] = pmp[r] = 0.0;
    ep[r] = em[r] = 0;
    Vpn[r] = Vps[r] = Vmn[r] = Vms[r] = make_double2(0.0, 0.0);
    if (t < A.T) {
      x[r] = A.x[t];
      const double lc = A.lcos[t], ls = A.lsin[t];
      double s;
      seed_split(spin2_seed_L(sd, lc, ls), sd[3], s, ep[r]);
      ppp[r] = -s;
      seed_split(spin2_seed_L(sd + 4, lc, ls), sd[7], s, em[r]);
      pmp[r] = -s;
      Vpn[r] = ld_c(Vp[(long long)t * ringM + m]);
      Vmn[r] = ld_c(Vm[(long long)t * ringM + m]);
      if (t < A.T - 1) {
        // (-1)^(l-m) = (-1)^(l-lmin) (-1)^(lmin-m): fold the second factor into the data
        const double sg = ((lmin - m) & 1) ? -1.0 : 1.0;
        const long long rs = (long long)(A.nrings - 1 - t) * ringM + m;
        const double2 up = ld_c(Vp[rs]), um = ld_c(Vm[rs]);
        Vps[r] = make_double2(sg * up.x, sg * up.y);
        Vms[r] = make_double2(sg * um.x, sg * um.y);
      }
    }
  }
  double2 P[NV];
  auto active = [&]() {
    bool a = false;
#pragma unroll
    for (int r = 0; r < R; ++r) a |= ((ep[r] == 0) || (em[r] == 0)) && (tb + r < A.T);
    return (int)a;
  };
  red_zero_tail(red, tid, nthr);
  int wg_active = __syncthreads_or(active());
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      double2 ag = make_double2(0.0, 0.0);
      if (l <= lmax && l != lmin) ag = make_double2(A.alpha[cb + l], A.gamma[cb + l]);
      s_ag[tid] = ag;
    }
    __syncthreads();
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += 8) {
      const int nl = min(8, lmax - (l0 + jb) + 1);
      if (!wg_active) {
        s2_adj_block<R, RECONLY>(s_ag, jb, x, pp, ppp, pm, pmp, fp, fm, Vpn, Vps, Vmn, Vms, P);
        rescale<R>(pp, ppp, ep);
        rescale<R>(pm, pmp, em);
        if (tid < nl) {
          p_p2[l0 + jb + tid] = make_double2(0.0, 0.0);
          p_m2[l0 + jb + tid] = make_double2(0.0, 0.0);
        }
        wg_active = __syncthreads_or(active());
        continue;
      }
      bool deep = false;
#pragma unroll
      for (int r = 0; r < R; ++r) deep |= (ep[r] < 0) | (em[r] < 0);
      if (__ballot(deep) == 0ull) {
        s2_adj_block<R, FULL>(s_ag, jb, x, pp, ppp, pm, pmp, fp, fm, Vpn, Vps, Vmn, Vms, P);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          fp[r] = (ep[r] == 0) ? 1.0 : 0.0;
          fm[r] = (em[r] == 0) ? 1.0 : 0.0;
        }
        s2_adj_block<R, MASKED>(s_ag, jb, x, pp, ppp, pm, pmp, fp, fm, Vpn, Vps, Vmn, Vms, P);
        rescale<R>(pp, ppp, ep);
        rescale<R>(pm, pmp, em);
      }
      wg_active = wg_reduce_store(P, red, p_p2 + l0 + jb, p_m2 + l0 + jb, 2, nl, active(), tid, nthr);
    }
  }
}

// Sum the ring-chunk partials, apply G_l (and for spin s != 0 the E/B algebra
// aE = (-p2 - (-1)^s m2)/2, aB = i (p2 - (-1)^s m2)/2, healpix.py:353-354 for s = 2),
// write packed a_lm.
// P: the chunk partials' type -- double2 (FP64 path) or float2 (mixed path, whose partials
// are sums of FP32 matrix-core tiles).  The chunk sum is taken in FP64 in chunk order
// (fixed).  One thread per (map, packed index): the grid runs over the flat triangle,
// so short high-m columns waste no threads, and the chunk loads are issued 8 at a time.
// Entries with l < |s| are zero; they sit in the columns m < |s|.
// lfirst (or null): lfirst[c * M + m] is the first l at which ring chunk c's partial of column
// m was stored; below it the partial is zero and is neither stored nor read.
// idx(l, m) = col_base(m) + l; column m holds l = m .. lmax, so it starts at col_base(m) + m
__device__ __forceinline__ long long col_base(int m, int lmax) { return (long long)m * (2 * lmax + 1 - m) / 2; }

// the order m of packed index idx: the largest m with col_base(m) + m = m (2 lmax + 3 - m) / 2
// at or below idx (the root of that quadratic, then corrected by at most a step or two)
__device__ __forceinline__ int m_of_index(long long idx, int lmax) {
  const double a = 2.0 * lmax + 3.0;
  int m = (int)(0.5 * (a - sqrt(fmax(a * a - 8.0 * (double)idx, 0.0))));
  m = max(0, min(m, lmax));
  while (m > 0 && col_base(m, lmax) + m > idx) --m;
  while (m < lmax && col_base(m + 1, lmax) + m + 1 <= idx) ++m;
  return m;
}

template <typename P>
__device__ __forceinline__ void adj_finish_body(const RecArgs& A, int spin, int nchunk, const P* __restrict__ part,
                                                const int* __restrict__ lfirst, double2* __restrict__ alm,
                                                long long idx, int b, int batch) {
  if (idx >= A.K) return;
  const int ncomp = spin ? 2 : 1;
  const double g = A.G[idx];
  // part[((c * batch + b) * ncomp + comp) * len + pos]: len = K and pos = idx, or, with the
  // first-stored-l table (the 32-map spin-2 adjoint), its tile layout: pos = 32 tile_base(m)
  // + l - lmin (nothing stored below lmin)
  int mi = 0, li = 0, lmin = 0;
  long long len = A.K, pos = idx;
  if (lfirst) {
    mi = m_of_index(idx, A.lmax);
    li = (int)(idx - col_base(mi, A.lmax));
    lmin = mi > A.sabs ? mi : A.sabs;
    len = tile_part_len(A.lmax, A.sabs);
    pos = 32 * tile_base(mi, A.lmax, A.sabs) + (li - lmin);
  }
  const P* p0 = part + ((long long)b * ncomp) * len + pos;
  const long long cstride = (long long)batch * ncomp * len;
  // chunk c holds a stored partial at this index
  auto stored = [&](int c) { return !lfirst || (li >= lmin && li >= lfirst[(long long)c * A.M + mi]); };
  double2 s0 = make_double2(0.0, 0.0), s1 = s0;
  int c = 0;
  for (; c + 8 <= nchunk; c += 8) {
    P u[8], v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const bool ok = stored(c + k);
      u[k] = ok ? p0[(c + k) * cstride] : P{};
      if (spin) v[k] = ok ? p0[(c + k) * cstride + len] : P{};
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      s0.x += (double)u[k].x;
      s0.y += (double)u[k].y;
      if (spin) {
        s1.x += (double)v[k].x;
        s1.y += (double)v[k].y;
      }
    }
  }
  for (; c < nchunk; ++c) {
    if (!stored(c)) continue;
    const P u = p0[c * cstride];
    s0.x += (double)u.x;
    s0.y += (double)u.y;
    if (spin) {
      const P v = p0[c * cstride + len];
      s1.x += (double)v.x;
      s1.y += (double)v.y;
    }
  }
  if (spin == 0) {
    alm[(long long)b * A.K + idx] = make_double2(s0.x * g, s0.y * g);
    return;
  }
  double2* aE = alm + (long long)b * 2 * A.K;
  double2* aB = aE + A.K;
  bool below = false;  // l < |s|: column m < |s| at l = m .. |s| - 1
  for (int m = 0; m < A.sabs && m <= A.lmax; ++m) {
    const long long cb = (long long)m * (2 * A.lmax + 1 - m) / 2;
    below |= (idx >= cb + m) && (idx < cb + A.sabs);
  }
  if (below) {
    aE[idx] = make_double2(0.0, 0.0);
    aB[idx] = make_double2(0.0, 0.0);
    return;
  }
  const double2 p2 = make_double2(s0.x * g, s0.y * g);
  const double2 m2 = make_double2(s1.x * g * A.ssign, s1.y * g * A.ssign);  // the -spin channel's (-1)^s
  aE[idx] = make_double2(0.5 * (-p2.x - m2.x), 0.5 * (-p2.y - m2.y));
  const double2 d = p2 - m2;
  aB[idx] = make_double2(-0.5 * d.y, 0.5 * d.x);
}

template <typename P>
__global__ __launch_bounds__(256) void k_adj_finish(RecArgs A, int spin, int nchunk, const P* __restrict__ part,
                                                    const int* __restrict__ lfirst, double2* __restrict__ alm) {
  adj_finish_body<P>(A, spin, nchunk, part, lfirst, alm, (long long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y,
                     gridDim.y);
}

// Ragged groups: the chunk reduction of every plan in one launch (per plan: ceil(K / 256)
// work-groups per map).
struct GroupFin {
  RecArgs A;
  long long part_off;
  int batch, nchunk, spin, kblocks;
};

__global__ __launch_bounds__(256) void k_group_finish(const GroupFin* __restrict__ G, GroupCall C) {
  int local;
  const int i = group_member(C, local);
  const RecArgs A = G[i].A;
  const int kb = G[i].kblocks;
  const float2* part = reinterpret_cast<const float2*>(C.ws + G[i].part_off);
  adj_finish_body<float2>(A, G[i].spin, G[i].nchunk, part, nullptr, reinterpret_cast<double2*>(C.out[i]),
                          (long long)(local % kb) * 256 + threadIdx.x, local / kb, G[i].batch);
}

RecArgs make_args(const jhtx_plan* p, bool adjoint) {
  RecArgs a;
  a.alpha = p->rec.alpha;
  a.gamma = p->rec.gamma;
  a.G = p->rec.G;
  a.x = p->rec.x;
  a.lsin = p->rec.lsin;
  a.lcos = p->rec.lcos;
  a.seed = p->rec.seed;
  a.lmax = p->lmax;
  a.M = p->M;
  a.T = p->T;
  a.nside = p->nside;
  a.nrings = p->nrings;
  a.nchunk = adjoint ? p->adj_nchunk : p->nchunk;
  a.K = p->K;
  a.Mf = p->M;
  a.sabs = p->spin < 0 ? -p->spin : p->spin;
  a.ssign = (p->spin & 1) ? -1.0 : 1.0;
  return a;
}

}  // namespace

namespace {

template <typename C>
hipError_t synth_rec_t(const jhtx_plan* p, const double2* alm, C* F, int batch, hipStream_t s) {
  const RecArgs a = make_args(p, false);
  const int Mpad = (p->M + 8 * kMGroup - 1) / (8 * kMGroup) * (8 * kMGroup);
  dim3 grid(Mpad * p->nchunk, batch);
  dim3 block(p->rec_block);
  if (p->spin == 0) {
    switch (p->rec_R) {
      case 8: hipLaunchKernelGGL((k_synth_s0<8, C>), grid, block, 0, s, a, alm, F); break;
      case 4: hipLaunchKernelGGL((k_synth_s0<4, C>), grid, block, 0, s, a, alm, F); break;
      default: hipLaunchKernelGGL((k_synth_s0<1, C>), grid, block, 0, s, a, alm, F); break;
    }
  } else {
    switch (p->rec_R) {
      case 4: hipLaunchKernelGGL((k_synth_s2<4, C>), grid, block, 0, s, a, alm, F); break;
      case 2: hipLaunchKernelGGL((k_synth_s2<2, C>), grid, block, 0, s, a, alm, F); break;
      default: hipLaunchKernelGGL((k_synth_s2<1, C>), grid, block, 0, s, a, alm, F); break;
    }
  }
  return hipGetLastError();
}

// the vector-FMA adjoint kernels (the FP64 matrix-core form reads complex128 V only)
template <typename C>
hipError_t adj_rec_t(const jhtx_plan* p, const C* V, double2* part, int batch, hipStream_t s) {
  const RecArgs a = make_args(p, true);
  const int Mpad = (p->M + 8 * kMGroup - 1) / (8 * kMGroup) * (8 * kMGroup);
  dim3 grid(Mpad * p->adj_nchunk, batch);
  dim3 block(p->adj_block);
  if (p->spin == 0) {
    switch (p->adj_R) {
      case 8: hipLaunchKernelGGL((k_adj_s0<8, C>), grid, block, 0, s, a, V, part); break;
      case 4: hipLaunchKernelGGL((k_adj_s0<4, C>), grid, block, 0, s, a, V, part); break;
      default: hipLaunchKernelGGL((k_adj_s0<1, C>), grid, block, 0, s, a, V, part); break;
    }
  } else {
    switch (p->adj_R) {
      case 4: hipLaunchKernelGGL((k_adj_s2<4, C>), grid, block, 0, s, a, V, part); break;
      case 2: hipLaunchKernelGGL((k_adj_s2<2, C>), grid, block, 0, s, a, V, part); break;
      default: hipLaunchKernelGGL((k_adj_s2<1, C>), grid, block, 0, s, a, V, part); break;
    }
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_synth_rec(const jhtx_plan* p, const double2* alm, double2* F, int batch,
                            hipStream_t s) {
  return synth_rec_t(p, alm, F, batch, s);
}

hipError_t launch_synth_rec(const jhtx_plan* p, const double2* alm, float2* F, int batch,
                            hipStream_t s) {
  return synth_rec_t(p, alm, F, batch, s);
}

// the recursion + contraction kernel of the FP64 adjoint; its chunk partials are summed by
// launch_adj_finish (adj_rec_chunks chunks)
hipError_t launch_adj_rec(const jhtx_plan* p, const double2* V, double2* part, int batch, hipStream_t s) {
  return adj_rec_t(p, V, part, batch, s);
}

// complex64 V (a mixed plan's FP32 ring FFT output), FP64 recursion and partials
hipError_t launch_adj_rec(const jhtx_plan* p, const float2* V, double2* part, int batch, hipStream_t s) {
  return adj_rec_t(p, V, part, batch, s);
}

// ring chunks (partials per map and component) of the FP64 adjoint
int adj_rec_chunks(const jhtx_plan* p) { return p->adj_nchunk; }

hipError_t launch_adj_finish(const jhtx_plan* p, const double2* part, double2* alm, int batch,
                             int nchunk, hipStream_t s) {
  const RecArgs a = make_args(p, true);
  hipLaunchKernelGGL(k_adj_finish<double2>, dim3((unsigned)((p->K + 255) / 256), batch), dim3(256), 0, s, a, p->spin,
                     nchunk, part, (const int*)nullptr, alm);
  return hipGetLastError();
}

hipError_t launch_adj_finish(const jhtx_plan* p, const float2* part, const int* lfirst, double2* alm, int batch,
                             int nchunk, hipStream_t s) {
  const RecArgs a = make_args(p, true);
  hipLaunchKernelGGL(k_adj_finish<float2>, dim3((unsigned)((p->K + 255) / 256), batch), dim3(256), 0, s, a, p->spin,
                     nchunk, part, lfirst, alm);
  return hipGetLastError();
}


int group_fin_build(const GroupLayout& L, void** desc, int* start) {
  std::vector<GroupFin> h(L.n);
  start[0] = 0;
  for (int i = 0; i < L.n; ++i) {
    const jhtx_plan* p = L.plan[i];
    h[i].A = make_args(p, true);
    h[i].part_off = L.part_off[i];
    h[i].batch = L.batch[i];
    h[i].nchunk = group_adj_chunks(p);
    h[i].spin = p->spin;
    h[i].kblocks = (int)((p->K + 255) / 256);
    const long long n = (long long)start[i] + (long long)h[i].kblocks * L.batch[i];
    if (n > 0x7fffffffLL) {
      set_error("ragged group: too many reduction work-groups for one launch");
      return JHTX_EINVAL;
    }
    start[i + 1] = (int)n;
  }
  if (hipMalloc(desc, sizeof(GroupFin) * L.n) != hipSuccess) return JHTX_ENOMEM;
  if (hipMemcpy(*desc, h.data(), sizeof(GroupFin) * L.n, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("ragged group: descriptor upload failed");
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

hipError_t group_fin_launch(const void* desc, GroupCall C, hipStream_t s) {
  hipLaunchKernelGGL(k_group_finish, dim3((unsigned)C.start[C.n]), dim3(256), 0, s,
                     reinterpret_cast<const GroupFin*>(desc), C);
  return hipGetLastError();
}

}  // namespace jhtx
