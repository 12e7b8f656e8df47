// Variable-length ring FFTs: the azimuthal half of the transform pair.
//
// Synthesis (src/jht/healpix.py:283-291 spin 0, :319-330 spin 2): per ring of N pixels
//   D_k = sum_{m = k mod N} Fp_m e^{i m phi0} + sum_{m >= 1, -m = k mod N} conj(Fm_m e^{i m phi0})
//   ring_j = sum_k D_k e^{+2 pi i j k / N}           (ifft(D) * N)
// Adjoint (healpix.py:293-305, :332-355): W = unnormalised forward DFT of the ring,
//   V+_m = W[m mod N] e^{-i m phi0},  V-_m = conj(W[-m mod N]) e^{-i m phi0}.
// phi0 is either 0 or pi/N (a half-sample shift), so e^{i m phi0} = (-1)^q e^{i pi k / N}
// for m = k + q N and the phase is applied once per bin.
//
// One work-group handles a mirror ring pair (r, 4 nside - 2 - r): the pair shares N and
// phi0.  Spin 0 packs the two real rings into one complex transform.  Transforms:
//   N <= 32                 direct DFT through an LDS root table,
//   N a power of two        radix-4/2 Stockham FFT, 16 elements per thread in registers,
//                           exchanges through LDS (re and im separately if LDS is short),
//   otherwise               chirp-z (Bluestein, _azimuth.py:150-191) at the next power of
//                           two L >= 2N-1 with the exact k^2 mod 2N chirp reduction and a
//                           kernel spectrum precomputed at plan creation.
#include <vector>
#include <algorithm>

#include "jhtx_internal.h"

namespace jhtx {

namespace {

constexpr int kDirectMax = 32;

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// e^{i pi t}
__device__ __forceinline__ double2 cispi(double t) {
  double s, c;
  sincospi(t, &s, &c);
  return make_double2(c, s);
}
// e^{s i pi n^2 / N} with the exact integer reduction n^2 mod 2N
__device__ __forceinline__ double2 chirp(long long n, int N, int s) {
  const long long r = (n * n) % (2LL * N);
  const double2 c = cispi((double)r / (double)N);
  return s > 0 ? c : cconj(c);
}

struct FftArgs {
  int nside, lmax, M, nrings, spin;
  long long npix;
  const double2* tw;  // exp(-2 pi i k / twL)
  int twL;
  const int* pairs;
  const long long* bofs;  // per pair p: offset of its Bluestein kernel spectrum, -1 if none
  const double2* bspec;
  int lds_doubles;
  int bluestein_all;
};

// transform kind for a ring of length N
__host__ __device__ inline int ring_kind(int N, int bluestein_all, int* L) {
  if (N <= kDirectMax) {
    *L = N;
    return 0;
  }
  if (((N & (N - 1)) == 0) && !bluestein_all) {
    *L = N;
    return 1;
  }
  int l = 128;
  while (l < 2 * N - 1) l <<= 1;
  *L = l;
  return 2;
}

__host__ __device__ inline int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

// --------------------------------------------------------------------------------- //
// power-of-two FFT, cyclic register layout: x[e] holds element tid + nt*e (nt = L/16)
// --------------------------------------------------------------------------------- //
template <int R>
__device__ __forceinline__ void fft_pass(double2 (&x)[16], int tid, int nt, int Ns, int dir,
                                         const double2* __restrict__ tw, int twL) {
  constexpr int NB = 16 / R;
  if (tid >= nt) return;
  const int step = twL / (Ns * R);
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) {
    const int j = tid + bb * nt;
    const int k = j & (Ns - 1);
    double2 v[R];
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = x[bb + t * NB];
    if (Ns > 1) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        double2 w = tw[t * k * step];
        if (dir > 0) w.y = -w.y;
        v[t] = cmul(v[t], w);
      }
    }
    if (R == 2) {
      x[bb] = cadd(v[0], v[1]);
      x[bb + NB] = csub(v[0], v[1]);
    } else {
      const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
      const double2 a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
      // dir < 0 (forward): a3 = -i d ; dir > 0: a3 = +i d
      const double2 a3 = dir < 0 ? make_double2(d.y, -d.x) : make_double2(-d.y, d.x);
      x[bb] = cadd(a0, a2);
      x[bb + NB] = cadd(a1, a3);
      x[bb + 2 * NB] = csub(a0, a2);
      x[bb + 3 * NB] = csub(a1, a3);
    }
  }
}

// move pass outputs (at Stockham positions) back to the cyclic layout through LDS
template <int R>
__device__ __forceinline__ void fft_exchange(double2 (&x)[16], int tid, int nt, int Ns, int lgNs,
                                             double* lds, bool split) {
  constexpr int NB = 16 / R;
  constexpr int lgR = (R == 2) ? 1 : 2;
  if (!split) {
    double2* l2 = reinterpret_cast<double2*>(lds);
    if (tid < nt) {
#pragma unroll
      for (int bb = 0; bb < NB; ++bb) {
        const int j = tid + bb * nt;
        const int base = ((j >> lgNs) << (lgNs + lgR)) + (j & (Ns - 1));
#pragma unroll
        for (int u = 0; u < R; ++u) l2[base + u * Ns] = x[bb + u * NB];
      }
    }
    __syncthreads();
    if (tid < nt) {
#pragma unroll
      for (int e = 0; e < 16; ++e) x[e] = l2[tid + nt * e];
    }
    __syncthreads();
  } else {
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      if (tid < nt) {
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
          const int j = tid + bb * nt;
          const int base = ((j >> lgNs) << (lgNs + lgR)) + (j & (Ns - 1));
#pragma unroll
          for (int u = 0; u < R; ++u) lds[base + u * Ns] = part ? x[bb + u * NB].y : x[bb + u * NB].x;
        }
      }
      __syncthreads();
      if (tid < nt) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          if (part)
            x[e].y = lds[tid + nt * e];
          else
            x[e].x = lds[tid + nt * e];
        }
      }
      __syncthreads();
    }
  }
}

// In-place FFT of length L = 2^lgL >= 16; dir -1 forward (e^{-2 pi i}), +1 backward.
// Starts with a barrier so LDS may still hold the input the caller loaded from.
__device__ void fft_pow2(double2 (&x)[16], int lgL, int tid, int dir, const double2* tw, int twL,
                         double* lds, int lds_doubles) {
  const int L = 1 << lgL;
  const int nt = L >> 4;
  const bool split = (2 * L > lds_doubles);
  __syncthreads();
  int Ns = 1, lgNs = 0, rem = lgL;
  while (rem > 0) {
    if (rem & 1) {
      fft_pass<2>(x, tid, nt, Ns, dir, tw, twL);
      rem -= 1;
      if (rem) fft_exchange<2>(x, tid, nt, Ns, lgNs, lds, split);
      Ns <<= 1;
      lgNs += 1;
    } else {
      fft_pass<4>(x, tid, nt, Ns, dir, tw, twL);
      rem -= 2;
      if (rem) fft_exchange<4>(x, tid, nt, Ns, lgNs, lds, split);
      Ns <<= 2;
      lgNs += 2;
    }
  }
}

// --------------------------------------------------------------------------------- //
// length-N DFT  X_j = sum_k in_k e^{s 2 pi i j k / N}, in = LDS double2[N] at offset 0.
// Returns the double2 offset of the output (N values) in LDS.  Ends with a barrier.
// --------------------------------------------------------------------------------- //
__device__ int ring_dft(const FftArgs& A, int N, int kind, int L, int s, long long bofs,
                        double* lds, int tid, int nthr) {
  double2* l2 = reinterpret_cast<double2*>(lds);
  if (kind == 0) {
    double2* roots = l2 + N;
    double2* out = l2 + 2 * N;
    for (int q = tid; q < N; q += nthr) roots[q] = cispi((double)(2 * s * q) / (double)N);
    __syncthreads();
    for (int j = tid; j < N; j += nthr) {
      double2 acc = make_double2(0.0, 0.0);
      for (int k = 0; k < N; ++k) {
        const double2 w = roots[(j * k) % N];
        const double2 v = l2[k];
        acc.x = fma(v.x, w.x, fma(-v.y, w.y, acc.x));
        acc.y = fma(v.x, w.y, fma(v.y, w.x, acc.y));
      }
      out[j] = acc;
    }
    __syncthreads();
    return 2 * N;  // double2 offset of out
  }
  const int lgL = ilog2(L);
  const int nt = L >> 4;
  double2 x[16];
  if (kind == 1) {
    if (tid < nt) {
#pragma unroll
      for (int e = 0; e < 16; ++e) x[e] = l2[tid + nt * e];
    }
    fft_pow2(x, lgL, tid, s < 0 ? -1 : 1, A.tw, A.twL, lds, A.lds_doubles);
    if (tid < nt) {
#pragma unroll
      for (int e = 0; e < 16; ++e) l2[tid + nt * e] = x[e];
    }
    __syncthreads();
    return 0;
  }
  // Bluestein: a_k = in_k c_k, b_n = conj(c_n), c_n = e^{s i pi n^2/N}; X_j = c_j (a * b)_j
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int i = tid + nt * e;
      x[e] = (i < N) ? cmul(l2[i], chirp(i, N, s)) : make_double2(0.0, 0.0);
    }
  }
  fft_pow2(x, lgL, tid, -1, A.tw, A.twL, lds, A.lds_doubles);
  if (tid < nt) {
    const double2* B = A.bspec + bofs;  // spectrum of e^{-i pi n^2/N}; conj for s < 0
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const double2 bv = B[tid + nt * e];
      x[e] = cmul(x[e], s > 0 ? bv : cconj(bv));
    }
  }
  fft_pow2(x, lgL, tid, +1, A.tw, A.twL, lds, A.lds_doubles);
  const double invL = 1.0 / (double)L;
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int i = tid + nt * e;
      if (i < N) l2[i] = cmul(cscale(x[e], invL), chirp(i, N, s));
    }
  }
  __syncthreads();
  return 0;
}

// --------------------------------------------------------------------------------- //
// synthesis fold: D_k (k < N) into LDS dst; combine -> dst_k += i D_k
// --------------------------------------------------------------------------------- //
__device__ void ring_fold(const double2* __restrict__ Fp, const double2* __restrict__ Fm, int M,
                          int N, bool sh, bool real0, bool combine, double2* dst,
                          double2* scratch, int tid, int nthr) {
  if (N <= nthr) {
    const int S = N * (nthr / N);
    if (tid < S) {
      double2 ap = make_double2(0.0, 0.0), am = ap;
      const int km = (N - tid % N) % N;
      for (int m = tid; m < M; m += S) {
        const int q = m / N;
        const double sg = (sh && (q & 1)) ? -1.0 : 1.0;
        double2 v = Fp[m];
        if (m == 0 && real0) v.y = 0.0;
        ap.x += sg * v.x;
        ap.y += sg * v.y;
        if (m >= 1) {
          const int qq = (m + km) / N;
          const double sm = (sh && (qq & 1)) ? -1.0 : 1.0;
          const double2 w = Fm[m];
          am.x += sm * w.x;
          am.y -= sm * w.y;
        }
      }
      scratch[tid] = ap;
      scratch[S + tid] = am;
    }
    __syncthreads();
    for (int k = tid; k < N; k += nthr) {
      double2 acc = make_double2(0.0, 0.0);
      for (int j = k; j < S; j += N) acc = cadd(acc, scratch[j]);
      for (int j = (N - k) % N; j < S; j += N) acc = cadd(acc, scratch[S + j]);
      if (sh) acc = cmul(acc, cispi((double)k / (double)N));
      if (combine) {
        const double2 d = dst[k];
        dst[k] = make_double2(d.x - acc.y, d.y + acc.x);
      } else {
        dst[k] = acc;
      }
    }
    __syncthreads();
    return;
  }
  for (int k = tid; k < N; k += nthr) {
    double2 acc = make_double2(0.0, 0.0);
    int q = 0;
    for (int m = k; m < M; m += N, ++q) {
      const double sg = (sh && (q & 1)) ? -1.0 : 1.0;
      double2 v = Fp[m];
      if (m == 0 && real0) v.y = 0.0;
      acc.x += sg * v.x;
      acc.y += sg * v.y;
    }
    q = 1;
    for (int m = N - k; m < M; m += N, ++q) {
      const double sg = (sh && (q & 1)) ? -1.0 : 1.0;
      const double2 w = Fm[m];
      acc.x += sg * w.x;
      acc.y -= sg * w.y;
    }
    if (sh) acc = cmul(acc, cispi((double)k / (double)N));
    if (combine) {
      const double2 d = dst[k];
      dst[k] = make_double2(d.x - acc.y, d.y + acc.x);
    } else {
      dst[k] = acc;
    }
  }
  __syncthreads();
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_ring_synth(FftArgs A, const double2* __restrict__ F,
                                                      double* __restrict__ map) {
  extern __shared__ double lds[];
  double2* l2 = reinterpret_cast<double2*>(lds);
  const int p = A.pairs[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int rn = p;
  const bool pair = p < 2 * A.nside - 1;
  const int rs = 4 * A.nside - 2 - p;
  const int N = ring_len(A.nside, rn);
  const bool sh = ring_shifted(A.nside, rn);
  int L;
  const int kind = ring_kind(N, A.bluestein_all, &L);
  const long long bofs = A.bofs[p];
  const long long ringM = A.M;
  double2* scratch = l2 + N;  // fold partials (only when N <= BLOCK)
  if (A.spin == 0) {
    const double2* Fb = F + (long long)b * A.nrings * ringM;
    ring_fold(Fb + rn * ringM, Fb + rn * ringM, A.M, N, sh, true, false, l2, scratch, tid, BLOCK);
    if (pair) ring_fold(Fb + rs * ringM, Fb + rs * ringM, A.M, N, sh, true, true, l2, scratch, tid, BLOCK);
    const int o = ring_dft(A, N, kind, L, +1, bofs, lds, tid, BLOCK);
    const double2* out = l2 + o;
    double* mb = map + (long long)b * A.npix;
    const long long sn = ring_start(A.nside, rn);
    const long long ss = ring_start(A.nside, rs);
    for (int j = tid; j < N; j += BLOCK) {
      const double2 z = out[j];
      mb[sn + j] = z.x;
      if (pair) mb[ss + j] = z.y;
    }
  } else {
    const double2* Fpb = F + (long long)b * 2 * A.nrings * ringM;
    const double2* Fmb = Fpb + (long long)A.nrings * ringM;
    double* Q = map + (long long)b * 2 * A.npix;
    double* U = Q + A.npix;
    for (int h = 0; h < (pair ? 2 : 1); ++h) {
      const int r = h ? rs : rn;
      ring_fold(Fpb + r * ringM, Fmb + r * ringM, A.M, N, sh, false, false, l2, scratch, tid, BLOCK);
      const int o = ring_dft(A, N, kind, L, +1, bofs, lds, tid, BLOCK);
      const double2* out = l2 + o;
      const long long st = ring_start(A.nside, r);
      for (int j = tid; j < N; j += BLOCK) {
        const double2 z = out[j];
        Q[st + j] = z.x;
        U[st + j] = z.y;
      }
      __syncthreads();
    }
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_ring_adjoint(FftArgs A, const double* __restrict__ map,
                                                        double2* __restrict__ V) {
  extern __shared__ double lds[];
  double2* l2 = reinterpret_cast<double2*>(lds);
  const int p = A.pairs[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int rn = p;
  const bool pair = p < 2 * A.nside - 1;
  const int rs = 4 * A.nside - 2 - p;
  const int N = ring_len(A.nside, rn);
  const bool sh = ring_shifted(A.nside, rn);
  int L;
  const int kind = ring_kind(N, A.bluestein_all, &L);
  const long long bofs = A.bofs[p];
  const long long ringM = A.M;
  const long long sn = ring_start(A.nside, rn);
  const long long ss = ring_start(A.nside, rs);
  if (A.spin == 0) {
    const double* mb = map + (long long)b * A.npix;
    for (int j = tid; j < N; j += BLOCK) l2[j] = make_double2(mb[sn + j], pair ? mb[ss + j] : 0.0);
    __syncthreads();
    const int o = ring_dft(A, N, kind, L, -1, bofs, lds, tid, BLOCK);
    const double2* W = l2 + o;
    double2* Vn = V + ((long long)b * A.nrings + rn) * ringM;
    double2* Vs = V + ((long long)b * A.nrings + rs) * ringM;
    for (int m = tid; m < A.M; m += BLOCK) {
      const int k = m % N, q = m / N;
      const double2 wk = W[k], wr = cconj(W[(N - k) % N]);
      const double sg = (sh && (q & 1)) ? -0.5 : 0.5;
      double2 t = make_double2(sg, 0.0);
      if (sh) t = cscale(cispi(-(double)k / (double)N), sg);
      Vn[m] = cmul(cadd(wk, wr), t);
      if (pair) {
        const double2 d = csub(wk, wr);  // -i d / 2 is the south ring's spectrum
        Vs[m] = cmul(make_double2(d.y, -d.x), t);
      }
    }
  } else {
    const double* Q = map + (long long)b * 2 * A.npix;
    const double* U = Q + A.npix;
    for (int h = 0; h < (pair ? 2 : 1); ++h) {
      const int r = h ? rs : rn;
      const long long st = h ? ss : sn;
      __syncthreads();
      for (int j = tid; j < N; j += BLOCK) l2[j] = make_double2(Q[st + j], U[st + j]);
      __syncthreads();
      const int o = ring_dft(A, N, kind, L, -1, bofs, lds, tid, BLOCK);
      const double2* W = l2 + o;
      double2* Vp = V + (((long long)b * 2 + 0) * A.nrings + r) * ringM;
      double2* Vm = V + (((long long)b * 2 + 1) * A.nrings + r) * ringM;
      for (int m = tid; m < A.M; m += BLOCK) {
        const int k = m % N, q = m / N;
        const double sg = (sh && (q & 1)) ? -1.0 : 1.0;
        double2 t = make_double2(sg, 0.0);
        if (sh) t = cscale(cispi(-(double)k / (double)N), sg);
        Vp[m] = cmul(W[k], t);
        Vm[m] = cmul(cconj(W[(N - k) % N]), t);
      }
    }
  }
}

// Kernel spectrum of a Bluestein ring: FFT_L of b_n = e^{-i pi n^2 / N} (n < N, mirrored).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bluestein_spectrum(FftArgs A) {
  extern __shared__ double lds[];
  const int p = A.pairs[blockIdx.x];
  const long long bofs = A.bofs[p];
  if (bofs < 0) return;
  const int tid = threadIdx.x;
  const int N = ring_len(A.nside, p);
  int L;
  ring_kind(N, A.bluestein_all, &L);
  const int lgL = ilog2(L);
  const int nt = L >> 4;
  double2 x[16];
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int i = tid + nt * e;
      int n = -1;
      if (i < N) n = i;
      else if (i > L - N) n = L - i;
      x[e] = (n >= 0) ? chirp(n, N, -1) : make_double2(0.0, 0.0);
    }
  }
  fft_pow2(x, lgL, tid, -1, A.tw, A.twL, lds, A.lds_doubles);
  if (tid < nt) {
    double2* B = const_cast<double2*>(A.bspec) + bofs;
#pragma unroll
    for (int e = 0; e < 16; ++e) B[tid + nt * e] = x[e];
  }
}


FftArgs make_fft_args(const jhtx_plan* p, const FftClass& c) {
  FftArgs a;
  a.nside = p->nside;
  a.lmax = p->lmax;
  a.M = p->M;
  a.nrings = p->nrings;
  a.spin = p->spin;
  a.npix = p->npix;
  a.tw = p->fft.tw;
  a.twL = p->fft.lmax_fft;
  a.pairs = c.pairs;
  a.bofs = p->fft.bofs;
  a.bspec = p->fft.bspec;
  a.lds_doubles = c.lds_doubles;
  a.bluestein_all = (p->flags & JHTX_FLAG_BLUESTEIN_ALL) ? 1 : 0;
  return a;
}

constexpr int kBlocks[3] = {64, 256, 1024};
constexpr int kLdsCap[3] = {2048, 8192, 16384};  // doubles: 16, 64, 128 KiB

}  // namespace

int build_fft_tables(jhtx_plan* p) {
  const int nside = p->nside;
  const int npairs = 2 * nside;
  const int ball = (p->flags & JHTX_FLAG_BLUESTEIN_ALL) ? 1 : 0;
  std::vector<int> cls_pairs[3];
  int need[3] = {0, 0, 0};
  int twL = 16;
  std::vector<long long> bofs(npairs, -1);
  long long blen = 0;
  for (int q = 0; q < npairs; ++q) {
    const int N = ring_len(nside, q);
    int L;
    const int kind = ring_kind(N, ball, &L);
    if (L > 16384) {
      set_error("ring transform length " + std::to_string(L) + " (nside=" + std::to_string(nside) +
                ") exceeds the 16384-point single-work-group FFT");
      return JHTX_EUNSUPPORTED;
    }
    const int c = (L <= 1024) ? 0 : (L <= 4096 ? 1 : 2);
    cls_pairs[c].push_back(q);
    const int blk = kBlocks[c];
    int nd = (kind == 0) ? 6 * N : 2 * L;
    if (N <= blk) nd = std::max(nd, 2 * N + 4 * blk);
    nd = std::max(nd, 2 * N);
    need[c] = std::max(need[c], nd);
    if (kind >= 1) twL = std::max(twL, L);
    if (kind == 2) {
      bofs[q] = blen;
      blen += L;
    }
  }
  // twiddles exp(-2 pi i k / twL), accurate to the last bit
  std::vector<double2> tw(twL);
  for (int k = 0; k < twL; ++k) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)twL;
    tw[k] = make_double2((double)cosl(a), (double)sinl(a));
  }
  hipError_t e;
  e = hipMalloc(&p->fft.tw, sizeof(double2) * twL);
  if (e != hipSuccess) return JHTX_ENOMEM;
  hipMemcpy(p->fft.tw, tw.data(), sizeof(double2) * twL, hipMemcpyHostToDevice);
  p->fft.lmax_fft = twL;
  e = hipMalloc(&p->fft.bofs, sizeof(long long) * npairs);
  if (e != hipSuccess) return JHTX_ENOMEM;
  hipMemcpy(p->fft.bofs, bofs.data(), sizeof(long long) * npairs, hipMemcpyHostToDevice);
  p->fft.bspec_len = blen;
  if (blen > 0) {
    e = hipMalloc(&p->fft.bspec, sizeof(double2) * blen);
    if (e != hipSuccess) return JHTX_ENOMEM;
  }
  p->fft.ncls = 0;
  for (int c = 0; c < 3; ++c) {
    FftClass& fc = p->fft.cls[c];
    fc.block = kBlocks[c];
    fc.npairs = (int)cls_pairs[c].size();
    fc.lds_doubles = std::min(std::max(need[c], 64), kLdsCap[c]);
    fc.pairs = nullptr;
    if (fc.npairs) {
      e = hipMalloc(&fc.pairs, sizeof(int) * fc.npairs);
      if (e != hipSuccess) return JHTX_ENOMEM;
      hipMemcpy(fc.pairs, cls_pairs[c].data(), sizeof(int) * fc.npairs, hipMemcpyHostToDevice);
      p->fft.ncls++;
    }
  }
  // dynamic LDS above 64 KiB must be opted into per kernel
  const int big = kLdsCap[2] * (int)sizeof(double);
  hipFuncSetAttribute((const void*)k_ring_synth<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_ring_adjoint<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  hipFuncSetAttribute((const void*)k_bluestein_spectrum<1024>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
  // Bluestein kernel spectra (one launch per class; pairs without one return at once)
  if (blen > 0) {
    for (int c = 0; c < 3; ++c) {
      const FftClass& fc = p->fft.cls[c];
      if (!fc.npairs) continue;
      const FftArgs a = make_fft_args(p, fc);
      const size_t sh = sizeof(double) * fc.lds_doubles;
      if (c == 0)
        hipLaunchKernelGGL(k_bluestein_spectrum<64>, dim3(fc.npairs), dim3(64), sh, 0, a);
      else if (c == 1)
        hipLaunchKernelGGL(k_bluestein_spectrum<256>, dim3(fc.npairs), dim3(256), sh, 0, a);
      else
        hipLaunchKernelGGL(k_bluestein_spectrum<1024>, dim3(fc.npairs), dim3(1024), sh, 0, a);
    }
    e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      set_error(std::string("Bluestein spectrum kernel failed: ") + hipGetErrorString(e));
      return JHTX_EDEVICE;
    }
  }
  return JHTX_OK;
}

void free_fft_tables(jhtx_plan* p) {
  if (p->fft.tw) hipFree(p->fft.tw);
  if (p->fft.bofs) hipFree(p->fft.bofs);
  if (p->fft.bspec) hipFree(p->fft.bspec);
  for (int c = 0; c < 3; ++c)
    if (p->fft.cls[c].pairs) hipFree(p->fft.cls[c].pairs);
  p->fft = FftTables();
}

hipError_t launch_ring_synth(const jhtx_plan* p, const double2* F, double* map, int batch,
                             hipStream_t s) {
  for (int c = 0; c < 3; ++c) {
    const FftClass& fc = p->fft.cls[c];
    if (!fc.npairs) continue;
    const FftArgs a = make_fft_args(p, fc);
    const size_t sh = sizeof(double) * fc.lds_doubles;
    const dim3 grid(fc.npairs, batch);
    if (c == 0)
      hipLaunchKernelGGL(k_ring_synth<64>, grid, dim3(64), sh, s, a, F, map);
    else if (c == 1)
      hipLaunchKernelGGL(k_ring_synth<256>, grid, dim3(256), sh, s, a, F, map);
    else
      hipLaunchKernelGGL(k_ring_synth<1024>, grid, dim3(1024), sh, s, a, F, map);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_ring_adjoint(const jhtx_plan* p, const double* map, double2* V, int batch,
                               hipStream_t s) {
  for (int c = 0; c < 3; ++c) {
    const FftClass& fc = p->fft.cls[c];
    if (!fc.npairs) continue;
    const FftArgs a = make_fft_args(p, fc);
    const size_t sh = sizeof(double) * fc.lds_doubles;
    const dim3 grid(fc.npairs, batch);
    if (c == 0)
      hipLaunchKernelGGL(k_ring_adjoint<64>, grid, dim3(64), sh, s, a, map, V);
    else if (c == 1)
      hipLaunchKernelGGL(k_ring_adjoint<256>, grid, dim3(256), sh, s, a, map, V);
    else
      hipLaunchKernelGGL(k_ring_adjoint<1024>, grid, dim3(1024), sh, s, a, map, V);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace jhtx
