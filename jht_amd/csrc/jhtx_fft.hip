// Variable-length ring FFTs: the azimuthal half of the transform pair.
//
// Synthesis (src/jht/healpix.py:283-291 spin 0, :319-330 spin 2): per ring of N pixels
//   D_k = sum_{m = k mod N} Fp_m e^{i m phi0} + sum_{m >= 1, -m = k mod N} conj(Fm_m e^{i m phi0})
//   ring_j = sum_k D_k e^{+2 pi i j k / N}           (ifft(D) * N)
// Adjoint (healpix.py:293-305, :332-355): W = unnormalised forward DFT of the ring,
//   V+_m = W[m mod N] e^{-i m phi0},  V-_m = conj(W[-m mod N]) e^{-i m phi0}.
// phi0 is either 0 or pi/N (a half-sample shift), so e^{i m phi0} = (-1)^q e^{i pi k / N}
// for m = k + q N: the sign per term, the phase once per bin, read from a table (the twiddle
// table for power-of-two rings, whose length 2N divides; folded into the chirp of chirp-z
// rings, e^{i pi (k^2 + k) / N} with an exact integer exponent).
//
// One work-group handles a mirror ring pair (r, 4 nside - 2 - r): the pair shares N and
// phi0.  Spin 0 packs the two real rings into one complex transform.  Transforms:
//   N <= 32                 direct DFT through an LDS root table,
//   N a power of two        radix-4/2 Stockham FFT, 16 elements per thread in registers,
//                           exchanges through LDS (re and im separately if LDS is short),
//   otherwise               chirp-z (Bluestein, _azimuth.py:150-191) at the next power of
//                           two L >= 2N-1 with the exact k^2 mod 2N chirp reduction and a
//                           kernel spectrum precomputed at plan creation.
// The transform arithmetic is templated on the scalar type T: FP64 plans use double, the
// mixed-precision plans (whose ring coefficients F / V are FP32) use float.  Folds sum in
// FP64 and every table is computed in FP64 (then rounded once for the float path).
#include <vector>
#include <algorithm>
#include <cstdlib>

#include "jhtx_internal.h"

namespace jhtx {

namespace {

constexpr int kDirectMax = 32;

template <typename T> struct CxOf;
template <> struct CxOf<double> { using type = double2; };
template <> struct CxOf<float> { using type = float2; };
template <typename T> using cx = typename CxOf<T>::type;

__device__ __forceinline__ double2 mk2(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ float2 mk2(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk2(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk2(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return mk2(fmaT(a.x, b.x, -a.y * b.y), fmaT(a.x, b.y, a.y * b.x));
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return mk2(a.x, -a.y); }
template <typename C, typename S> __device__ __forceinline__ C cscale(C a, S s) { return mk2(a.x * s, a.y * s); }
// FP64 -> complex of scalar type T (identity for double)
template <typename T> __device__ __forceinline__ cx<T> to_cx(double2 v) { return mk2((T)v.x, (T)v.y); }
// e^{i pi t} in FP64 (only used in loops that are not unrolled: register pressure)
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
// e^{i pi (n^2 + n) / N}: the chirp times the half-sample shift e^{i pi n / N}
__device__ __forceinline__ double2 chirp_shifted(long long n, int N) {
  const long long r = (n * n + n) % (2LL * N);
  return cispi((double)r / (double)N);
}

struct FftArgs {
  int nside, lmax, M, nrings, spin;
  long long npix;
  const double2* tw;  // exp(-2 pi i k / twL)
  const float2* twf;  // the same rounded to FP32 (mixed plans)
  int twL;
  const int* pairs;
  const long long* bofs;  // per pair p: offset of its Bluestein kernel spectrum, -1 if none
  const double2* bspec;
  const float2* bspecf;
  const long long* cofs;  // per pair p: offset of its chirp table c_n = e^{i pi n^2/N}, n < N
  const double2* chirps;
  const float2* chirpsf;
  long long chirp_shift;  // offset of the shifted chirp tables
  const long long* qofs;  // per pair p: offset of its split-transform tables in qtab, -1 if none
  const float2* qtab;
  int lds_bytes;
  int bluestein_all;
  long long Mf;  // row stride of F / V (M unless the synthesis's caller pads F's rows)
};

template <typename T> __device__ __forceinline__ const cx<T>* tw_of(const FftArgs& A);
template <> __device__ __forceinline__ const double2* tw_of<double>(const FftArgs& A) { return A.tw; }
template <> __device__ __forceinline__ const float2* tw_of<float>(const FftArgs& A) { return A.twf; }
template <typename T> __device__ __forceinline__ const cx<T>* bspec_of(const FftArgs& A);
template <> __device__ __forceinline__ const double2* bspec_of<double>(const FftArgs& A) { return A.bspec; }
template <> __device__ __forceinline__ const float2* bspec_of<float>(const FftArgs& A) { return A.bspecf; }
template <typename T> __device__ __forceinline__ const cx<T>* chirps_of(const FftArgs& A);
template <> __device__ __forceinline__ const double2* chirps_of<double>(const FftArgs& A) { return A.chirps; }
template <> __device__ __forceinline__ const float2* chirps_of<float>(const FftArgs& A) { return A.chirpsf; }

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
// power-of-two FFT, cyclic register layout: x[e] holds element tid + nt*e (nt = L/E)
// --------------------------------------------------------------------------------- //
template <int R, int E, typename T>
__device__ __forceinline__ void fft_pass(cx<T> (&x)[E], int tid, int nt, int Ns, int dir,
                                         const cx<T>* __restrict__ tw, int twL) {
  constexpr int NB = E / R;
  if (tid >= nt) return;
  const int step = twL / (Ns * R);
#pragma unroll
  for (int bb = 0; bb < NB; ++bb) {
    const int j = tid + bb * nt;
    const int k = j & (Ns - 1);
    cx<T> v[R];
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = x[bb + t * NB];
    if (Ns > 1) {
      // one table load per butterfly; w^2, w^3, ... by complex products (<= 2 ulp each)
      cx<T> w = tw[k * step];
      if (dir > 0) w.y = -w.y;
      v[1] = cmul(v[1], w);
      if (R >= 4) {
        const cx<T> w2 = cmul(w, w);
        const cx<T> w3 = cmul(w2, w);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
        if (R == 8) {
          const cx<T> w4 = cmul(w2, w2);
          v[4] = cmul(v[4], w4);
          v[5] = cmul(v[5], cmul(w4, w));
          v[6] = cmul(v[6], cmul(w4, w2));
          v[7] = cmul(v[7], cmul(w4, w3));
        }
      }
    }
    if (R == 8) {
      // two 4-point DFTs (even, odd inputs), combined with W8^k = e^{dir i pi k / 4}
      const cx<T> e0 = cadd(v[0], v[4]), e1 = csub(v[0], v[4]);
      const cx<T> e2 = cadd(v[2], v[6]), ed = csub(v[2], v[6]);
      const cx<T> e3 = dir < 0 ? mk2(ed.y, -ed.x) : mk2(-ed.y, ed.x);
      const cx<T> E0 = cadd(e0, e2), E1 = cadd(e1, e3), E2 = csub(e0, e2), E3 = csub(e1, e3);
      const cx<T> o0 = cadd(v[1], v[5]), o1 = csub(v[1], v[5]);
      const cx<T> o2 = cadd(v[3], v[7]), od = csub(v[3], v[7]);
      const cx<T> o3 = dir < 0 ? mk2(od.y, -od.x) : mk2(-od.y, od.x);
      const cx<T> O0 = cadd(o0, o2), O1 = cadd(o1, o3), O2 = csub(o0, o2), O3 = csub(o1, o3);
      const T c = (T)0.70710678118654752440;  // cos(pi/4)
      const T sd = dir < 0 ? -c : c;          // sin(dir pi/4)
      // W8^1 O1, W8^2 O2 = dir i O2, W8^3 O3
      const cx<T> t1 = mk2(c * O1.x - sd * O1.y, c * O1.y + sd * O1.x);
      const cx<T> t2 = dir < 0 ? mk2(O2.y, -O2.x) : mk2(-O2.y, O2.x);
      const cx<T> t3 = mk2(-c * O3.x - sd * O3.y, -c * O3.y + sd * O3.x);
      x[bb] = cadd(E0, O0);
      x[bb + NB] = cadd(E1, t1);
      x[bb + 2 * NB] = cadd(E2, t2);
      x[bb + 3 * NB] = cadd(E3, t3);
      x[bb + 4 * NB] = csub(E0, O0);
      x[bb + 5 * NB] = csub(E1, t1);
      x[bb + 6 * NB] = csub(E2, t2);
      x[bb + 7 * NB] = csub(E3, t3);
    } else if (R == 2) {
      x[bb] = cadd(v[0], v[1]);
      x[bb + NB] = csub(v[0], v[1]);
    } else {
      const cx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
      const cx<T> a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
      // dir < 0 (forward): a3 = -i d ; dir > 0: a3 = +i d
      const cx<T> a3 = dir < 0 ? mk2(d.y, -d.x) : mk2(-d.y, d.x);
      x[bb] = cadd(a0, a2);
      x[bb + NB] = cadd(a1, a3);
      x[bb + 2 * NB] = csub(a0, a2);
      x[bb + 3 * NB] = csub(a1, a3);
    }
  }
}

// move pass outputs (at Stockham positions) back to the cyclic layout through LDS
template <int R, int E, typename T>
__device__ __forceinline__ void fft_exchange(cx<T> (&x)[E], int tid, int nt, int Ns, int lgNs,
                                             void* lds, bool split) {
  constexpr int NB = E / R;
  constexpr int lgR = (R == 2) ? 1 : (R == 4 ? 2 : 3);
  if (!split) {
    cx<T>* l2 = reinterpret_cast<cx<T>*>(lds);
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
      for (int e = 0; e < E; ++e) x[e] = l2[tid + nt * e];
    }
    __syncthreads();
  } else {
    T* l1 = reinterpret_cast<T*>(lds);
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      if (tid < nt) {
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) {
          const int j = tid + bb * nt;
          const int base = ((j >> lgNs) << (lgNs + lgR)) + (j & (Ns - 1));
#pragma unroll
          for (int u = 0; u < R; ++u) l1[base + u * Ns] = part ? x[bb + u * NB].y : x[bb + u * NB].x;
        }
      }
      __syncthreads();
      if (tid < nt) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          if (part)
            x[e].y = l1[tid + nt * e];
          else
            x[e].x = l1[tid + nt * e];
        }
      }
      __syncthreads();
    }
  }
}

// Compile-time pass sequence for L = 2^LG with radix 8 where it fits: with REM = log2 of
// the radix still to apply, radix 8 when REM = 0 mod 3, radix 4 when REM = 2 or 1 mod 3
// (radix 2 for REM = 1), so 2^13 and 2^14 take 5 passes instead of 7; every pass but the
// last ends with an exchange through LDS.  Radix 8 needs E >= 8 elements per thread.
template <int E, int LG, int LGNS, typename T>
__device__ __forceinline__ void fft_passes(cx<T> (&x)[E], int tid, int dir, const cx<T>* tw,
                                           int twL, void* lds, bool split) {
  constexpr int REM = LG - LGNS;
  if constexpr (REM > 0) {
    constexpr int R = (REM == 1) ? 2 : ((REM % 3 == 0 && E % 8 == 0) ? 8 : 4);
    constexpr int lgR = (R == 2) ? 1 : (R == 4 ? 2 : 3);
    constexpr int nt = (1 << LG) / E;
    constexpr int Ns = 1 << LGNS;
    fft_pass<R, E, T>(x, tid, nt, Ns, dir, tw, twL);
    if constexpr (REM > lgR) {
      fft_exchange<R, E, T>(x, tid, nt, Ns, LGNS, lds, split);
      fft_passes<E, LG, LGNS + lgR, T>(x, tid, dir, tw, twL, lds, split);
    }
  }
}

// In-place FFT of length L = 2^lgL with LGLO <= lgL <= LGHI (the range of one launch
// class); dir -1 forward (e^{-2 pi i}), +1 backward.  Starts with a barrier so LDS may
// still hold the input the caller loaded from.
template <int E, int LG, int LGHI, typename T>
__device__ __forceinline__ void fft_dispatch(cx<T> (&x)[E], int lgL, int tid, int dir,
                                             const cx<T>* tw, int twL, void* lds, bool split) {
  if constexpr (LG < LGHI) {
    if (lgL != LG) {
      fft_dispatch<E, LG + 1, LGHI, T>(x, lgL, tid, dir, tw, twL, lds, split);
      return;
    }
  }
  fft_passes<E, LG, 0, T>(x, tid, dir, tw, twL, lds, split);
}

template <int E, int LGLO, int LGHI, typename T>
__device__ __forceinline__ void fft_pow2(cx<T> (&x)[E], int lgL, int tid, int dir,
                                         const cx<T>* tw, int twL, void* lds, int lds_bytes) {
  const bool split = (2 * (1 << lgL) * (int)sizeof(T) > lds_bytes);
  __syncthreads();
  fft_dispatch<E, LGLO, LGHI, T>(x, lgL, tid, dir, tw, twL, lds, split);
}

// transform-length range of the launch class (threads, elements per thread)
constexpr int lg_hi(int block, int e) {
  int l = 0;
  while ((1 << l) < block * e) ++l;
  return l;
}
constexpr int lg_lo(int block, int e) { return block == 64 ? 6 : (block == 256 ? lg_hi(block, e) - 1 : lg_hi(block, e)); }
static_assert(lg_lo(1024, 16) == 14 && lg_hi(1024, 16) == 14, "the 1024-thread class is L = 16384 only");

// --------------------------------------------------------------------------------- //
// length-N DFT  X_j = sum_k in_k e^{s 2 pi i j k / N}, in = LDS cx<T>[N] at offset 0.
// Returns the complex offset of the output (N values) in LDS.  Ends with a barrier.
// --------------------------------------------------------------------------------- //
// Chirp-z rings with the half-sample shift (sh) take the shifted chirp where the shift
// belongs: the synthesis multiplies bin k of its input by e^{i pi k / N}, the adjoint bin j
// of its output by e^{-i pi j / N}.
template <int E, int LGLO, int LGHI, int KIND, typename T>
__device__ __forceinline__ int ring_dft(const FftArgs& A, int N, int L, int s, long long bofs,
                                        const cx<T>* __restrict__ ch, bool sh, void* lds, int tid, int nthr) {
  cx<T>* l2 = reinterpret_cast<cx<T>*>(lds);
  if constexpr (KIND == 0) {
    cx<T>* roots = l2 + N;
    cx<T>* out = l2 + 2 * N;
    for (int q = tid; q < N; q += nthr) roots[q] = to_cx<T>(cispi((double)(2 * s * q) / (double)N));
    __syncthreads();
    for (int j = tid; j < N; j += nthr) {
      cx<T> acc = mk2((T)0, (T)0);
      for (int k = 0; k < N; ++k) {
        const cx<T> w = roots[(j * k) % N];
        const cx<T> v = l2[k];
        acc.x = fmaT(v.x, w.x, fmaT(-v.y, w.y, acc.x));
        acc.y = fmaT(v.x, w.y, fmaT(v.y, w.x, acc.y));
      }
      out[j] = acc;
    }
    __syncthreads();
    return 2 * N;  // complex offset of out
  }
  const int lgL = ilog2(L);
  const int nt = L / E;
  cx<T> x[E];
  if constexpr (KIND == 1) {
    if (tid < nt) {
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = l2[tid + nt * e];
    }
    fft_pow2<E, LGLO, LGHI, T>(x, lgL, tid, s < 0 ? -1 : 1, tw_of<T>(A), A.twL, lds, A.lds_bytes);
    if (tid < nt) {
#pragma unroll
      for (int e = 0; e < E; ++e) l2[tid + nt * e] = x[e];
    }
    __syncthreads();
    return 0;
  }
  // Bluestein: a_k = in_k c_k, b_n = conj(c_n), c_n = e^{s i pi n^2/N}; X_j = c_j (a * b)_j
  const cx<T>* cpre = (sh && s > 0) ? ch + A.chirp_shift : ch;
  const cx<T>* cpost = (sh && s < 0) ? ch + A.chirp_shift : ch;
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + nt * e;
      if (i < N) {
        const cx<T> c = cpre[i];
        x[e] = cmul(l2[i], s > 0 ? c : cconj(c));
      } else {
        x[e] = mk2((T)0, (T)0);
      }
      if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
  // forward FFT, pointwise product with the kernel spectrum, backward FFT: one call site.
  // tid passes through an empty asm register constraint (no instruction) in every
  // iteration, so the FFT's thread-index arithmetic is recomputed per pass instead of
  // being hoisted out of the loop and held (spilled) across both passes.
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    int tp = tid;
    asm volatile("" : "+v"(tp));
    fft_pow2<E, LGLO, LGHI, T>(x, lgL, tp, pass ? +1 : -1, tw_of<T>(A), A.twL, lds, A.lds_bytes);
    if (pass == 0 && tp < nt) {
      const cx<T>* B = bspec_of<T>(A) + bofs;  // spectrum of e^{-i pi n^2/N}; conj for s < 0
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const cx<T> bv = B[tp + nt * e];
        x[e] = cmul(x[e], s > 0 ? bv : cconj(bv));
        if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);  // bound in-flight loads
      }
    }
  }
  const T invL = (T)1 / (T)L;
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + nt * e;
      if (i < N) {
        const cx<T> c = cpost[i];
        l2[i] = cmul(cscale(x[e], invL), s > 0 ? c : cconj(c));
      }
      if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
  __syncthreads();
  return 0;
}

// --------------------------------------------------------------------------------- //
// synthesis fold: D_k (k < N) into LDS dst; combine -> dst_k += i D_k.  Sums in FP64.
// --------------------------------------------------------------------------------- //
// e^{i pi k / N}, the half-sample shift of bin k: phase 1 reads it from the twiddle table
// (exp(-2 pi i k tstep / twL) with twL = 2 N tstep, conjugated), phase 2 computes it
__device__ __forceinline__ double2 bin_phase(int phase, const double2* __restrict__ tw, int tstep, int k, int N) {
  if (phase == 1) {
    const double2 w = tw[k * tstep];
    return make_double2(w.x, -w.y);
  }
  return cispi((double)k / (double)N);
}

template <typename T>
__device__ __forceinline__ double2 ld64(const cx<T>* p, int m) {
  const cx<T> v = p[m];
  return make_double2((double)v.x, (double)v.y);
}

// Loads are issued in groups (clamped, always-valid addresses; each value's validity applied
// after) so that a thread waits for one memory latency per group rather than per term: the
// large-ring launch classes run one work-group per CU and cannot hide a latency per term.
template <typename T>
// sh: the ring is shifted (signs (-1)^q per term); phase: how bin k's e^{i pi k / N} is
// applied (0 not here -- unshifted, or carried by a chirp-z ring's chirp; 1, 2: bin_phase).
__device__ __forceinline__ void ring_fold(const cx<T>* __restrict__ Fp, const cx<T>* __restrict__ Fm, int M,
                          int N, bool sh, int phase, const double2* __restrict__ tw, int tstep,
                          bool real0, bool combine, cx<T>* dst, double2* scratch, int tid, int nthr) {
  constexpr int G = sizeof(T) == 4 ? 8 : 4;  // terms in flight per thread and field
  if (N <= nthr) {
    const int S = N * (nthr / N);
    if (tid < S) {
      double2 ap = make_double2(0.0, 0.0), am = ap;
      const int km = (N - tid % N) % N;
      for (int m0 = tid; m0 < M; m0 += G * S) {
        cx<T> fp[G], fm[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int mc = min(m0 + u * S, M - 1);
          fp[u] = Fp[mc];
          fm[u] = Fm[mc];
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int m = m0 + u * S;
          if (m < M) {
            const int q = m / N;
            const double sg = (sh && (q & 1)) ? -1.0 : 1.0;
            double2 v = make_double2((double)fp[u].x, (double)fp[u].y);
            if (m == 0 && real0) v.y = 0.0;
            ap.x += sg * v.x;
            ap.y += sg * v.y;
            if (m >= 1) {
              const int qq = (m + km) / N;
              const double sm = (sh && (qq & 1)) ? -1.0 : 1.0;
              am.x += sm * (double)fm[u].x;
              am.y -= sm * (double)fm[u].y;
            }
          }
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
      if (phase) acc = cmul(acc, bin_phase(phase, tw, tstep, k, N));
      if (combine) {
        const cx<T> d = dst[k];
        dst[k] = mk2((T)(d.x - (T)acc.y), (T)(d.y + (T)acc.x));
      } else {
        dst[k] = to_cx<T>(acc);
      }
    }
    __syncthreads();
    return;
  }
  // N > nthr: bin k = tid + nthr e sums Fp at m = k + q N (sign (-1)^q when shifted) and
  // conj(Fm) at m = N - k + (q - 1) N (q >= 1, the same sign).  With M <= 2 N every bin has
  // at most two terms of each field, and the terms of G bins are loaded at once.
  if (M <= 2 * N) {
    const double s1 = sh ? -1.0 : 1.0;  // the q = 1 sign
    for (int k0 = tid; k0 < N; k0 += G * nthr) {
      cx<T> p0[G], p1[G], q0[G], q1[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int kc = min(k0 + u * nthr, N - 1);
        p0[u] = Fp[min(kc, M - 1)];
        p1[u] = Fp[min(kc + N, M - 1)];
        q0[u] = Fm[min(N - kc, M - 1)];
        q1[u] = Fm[min(2 * N - kc, M - 1)];
      }
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int k = k0 + u * nthr;
        if (k >= N) break;
        double2 acc = make_double2(0.0, 0.0);
        if (k < M) {  // m = k (q = 0)
          acc.x += (double)p0[u].x;
          acc.y += (k == 0 && real0) ? 0.0 : (double)p0[u].y;
        }
        if (k + N < M) {  // m = k + N (q = 1)
          acc.x += s1 * (double)p1[u].x;
          acc.y += s1 * (double)p1[u].y;
        }
        if (N - k < M) {  // m = N - k (q = 1)
          acc.x += s1 * (double)q0[u].x;
          acc.y -= s1 * (double)q0[u].y;
        }
        if (2 * N - k < M) {  // m = 2 N - k (q = 2)
          acc.x += (double)q1[u].x;
          acc.y -= (double)q1[u].y;
        }
        if (phase) acc = cmul(acc, bin_phase(phase, tw, tstep, k, N));
        if (combine) {
          const cx<T> d = dst[k];
          dst[k] = mk2((T)(d.x - (T)acc.y), (T)(d.y + (T)acc.x));
        } else {
          dst[k] = to_cx<T>(acc);
        }
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
      double2 v = ld64<T>(Fp, m);
      if (m == 0 && real0) v.y = 0.0;
      acc.x += sg * v.x;
      acc.y += sg * v.y;
    }
    q = 1;
    for (int m = N - k; m < M; m += N, ++q) {
      const double sg = (sh && (q & 1)) ? -1.0 : 1.0;
      const double2 w = ld64<T>(Fm, m);
      acc.x += sg * w.x;
      acc.y -= sg * w.y;
    }
    if (phase) acc = cmul(acc, bin_phase(phase, tw, tstep, k, N));
    if (combine) {
      const cx<T> d = dst[k];
      dst[k] = mk2((T)(d.x - (T)acc.y), (T)(d.y + (T)acc.x));
    } else {
      dst[k] = to_cx<T>(acc);
    }
  }
  __syncthreads();
}

// LDS: ring data as cx<T>[N] from offset 0; the fold's FP64 partials follow at a 16-byte
// aligned offset (only for N <= BLOCK)
template <typename T>
__device__ __forceinline__ double2* fold_scratch(void* lds, int N) {
  const size_t off = ((size_t)N * sizeof(cx<T>) + 15) & ~(size_t)15;
  return reinterpret_cast<double2*>(reinterpret_cast<char*>(lds) + off);
}

// A spin-2 power-of-two ring of N = BLOCK E samples with M <= 2 N, from registers to
// registers: thread tid folds bins k = tid + BLOCK e straight into x[e] (the cyclic layout
// the transform takes), and after the transform x[e] holds sample tid + BLOCK e, stored from
// there.  Only the transform's exchanges use LDS.  The fold is ring_fold's M <= 2 N case,
// term for term (bin k sums Fp at k and k + N, conj(Fm) at N - k and 2 N - k), so the
// samples are those of the staged path bit for bit.
template <int BLOCK, int E, int LG, typename T>
__device__ __forceinline__ void ring_synth_direct(const FftArgs& A, const cx<T>* __restrict__ Fp,
                                                  const cx<T>* __restrict__ Fm, int N, bool sh, int phase,
                                                  int tstep, void* lds, double* __restrict__ Qr,
                                                  double* __restrict__ Ur) {
  constexpr int G = 8;  // bins per load group
  static_assert(E % G == 0, "load groups of 8 bins");
  const int tid = threadIdx.x;
  const int M = A.M;
  const bool wide = M > N;  // the terms at k + N and 2 N - k exist for some bins
  const double s1 = sh ? -1.0 : 1.0;  // the q = 1 sign
  cx<T> x[E];
  if (!wide) {
    // M <= N: bin k takes Fp at k (k < M) and conj(Fm) at N - k (N - k < M) only; all E
    // bins' loads are issued before the first is used
    cx<T> p0[E], q0[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = tid + BLOCK * e;
      p0[e] = Fp[min(k, M - 1)];
      q0[e] = Fm[min(N - k, M - 1)];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = tid + BLOCK * e;
      double2 acc = make_double2(0.0, 0.0);
      if (k < M) {
        acc.x += (double)p0[e].x;
        acc.y += (double)p0[e].y;
      }
      if (N - k < M) {
        acc.x += s1 * (double)q0[e].x;
        acc.y -= s1 * (double)q0[e].y;
      }
      if (phase) acc = cmul(acc, bin_phase(phase, A.tw, tstep, k, N));
      x[e] = to_cx<T>(acc);
    }
  }
#pragma unroll
  for (int e0 = 0; e0 < E && wide; e0 += G) {
    cx<T> p0[G], p1[G], q0[G], q1[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int k = tid + BLOCK * (e0 + u);
      p0[u] = Fp[min(k, M - 1)];
      q0[u] = Fm[min(N - k, M - 1)];
      if (wide) {
        p1[u] = Fp[min(k + N, M - 1)];
        q1[u] = Fm[min(2 * N - k, M - 1)];
      }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const int k = tid + BLOCK * (e0 + u);
      double2 acc = make_double2(0.0, 0.0);
      if (k < M) {  // m = k (q = 0)
        acc.x += (double)p0[u].x;
        acc.y += (double)p0[u].y;
      }
      if (wide && k + N < M) {  // m = k + N (q = 1)
        acc.x += s1 * (double)p1[u].x;
        acc.y += s1 * (double)p1[u].y;
      }
      if (N - k < M) {  // m = N - k (q = 1)
        acc.x += s1 * (double)q0[u].x;
        acc.y -= s1 * (double)q0[u].y;
      }
      if (wide && 2 * N - k < M) {  // m = 2 N - k (q = 2)
        acc.x += (double)q1[u].x;
        acc.y -= (double)q1[u].y;
      }
      if (phase) acc = cmul(acc, bin_phase(phase, A.tw, tstep, k, N));
      x[e0 + u] = to_cx<T>(acc);
    }
  }
  fft_pow2<E, LG, LG, T>(x, LG, tid, +1, tw_of<T>(A), A.twL, lds, A.lds_bytes);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int j = tid + BLOCK * e;
    __builtin_nontemporal_store((double)x[e].x, &Qr[j]);
    __builtin_nontemporal_store((double)x[e].y, &Ur[j]);
  }
}

// Work-group per mirror ring pair; template on the transform kind, spin and scalar type so
// each instantiation holds one code path (register pressure).
// Ring pair p, map b; spin 2 runs one ring per work-group (half 0 the north ring, 1 the
// south one), spin 0 packs the pair into one complex transform.  LGLO..LGHI: the power-of-two
// transform lengths the work-group shape covers.
template <int BLOCK, int E, int LGLO, int LGHI, int SPIN, int KIND, typename T>
__device__ __forceinline__ void ring_synth_body(const FftArgs& A, const cx<T>* __restrict__ F,
                                                double* __restrict__ map, int p, int b, int half) {
  extern __shared__ double lds[];
  cx<T>* l2 = reinterpret_cast<cx<T>*>(lds);
  const int tid = threadIdx.x;
  const int rn = p;
  const bool pair = p < 2 * A.nside - 1;
  const int rs = 4 * A.nside - 2 - p;
  if (SPIN == 2 && half && !pair) return;  // the equator has no mirror
  const int N = ring_len(A.nside, rn);
  const bool sh = ring_shifted(A.nside, rn);
  int L;
  ring_kind(N, A.bluestein_all, &L);
  const long long bofs = A.bofs[p];
  const cx<T>* ch = (A.cofs[p] >= 0) ? chirps_of<T>(A) + A.cofs[p] : nullptr;
  const long long ringM = A.Mf;
  double2* scratch = fold_scratch<T>(lds, N);  // fold partials (only when N <= BLOCK)
  // the bin phase of a shifted ring: in the chirp (chirp-z), the twiddle table (power of
  // two), computed (direct DFT)
  const int phase = (!sh || KIND == 2) ? 0 : (KIND == 1 ? 1 : 2);
  const int tstep = (KIND == 1) ? A.twL / (2 * N) : 0;
  if constexpr (SPIN == 0) {
    const cx<T>* Fb = F + (long long)b * A.nrings * ringM;
    ring_fold<T>(Fb + rn * ringM, Fb + rn * ringM, A.M, N, sh, phase, A.tw, tstep, true, false, l2, scratch, tid,
                 BLOCK);
    if (pair)
      ring_fold<T>(Fb + rs * ringM, Fb + rs * ringM, A.M, N, sh, phase, A.tw, tstep, true, true, l2, scratch, tid,
                   BLOCK);
    const int o = ring_dft<E, LGLO, LGHI, KIND, T>(A, N, L, +1, bofs, ch, sh, lds, tid, BLOCK);
    const cx<T>* out = l2 + o;
    double* mb = map + (long long)b * A.npix;
    const long long sn = ring_start(A.nside, rn);
    const long long ss = ring_start(A.nside, rs);
    for (int j = tid; j < N; j += BLOCK) {
      const cx<T> z = out[j];
      // streamed: the map is not read again by this launch
      __builtin_nontemporal_store((double)z.x, &mb[sn + j]);
      if (pair) __builtin_nontemporal_store((double)z.y, &mb[ss + j]);
    }
  } else {
    const cx<T>* Fpb = F + (long long)b * 2 * A.nrings * ringM;
    const cx<T>* Fmb = Fpb + (long long)A.nrings * ringM;
    double* Q = map + (long long)b * 2 * A.npix;
    double* U = Q + A.npix;
    const int r = half ? rs : rn;
    if constexpr (KIND == 1 && LGLO == LGHI) {
      if (N == BLOCK * E && A.M <= 2 * N) {
        ring_synth_direct<BLOCK, E, LGLO, T>(A, Fpb + r * ringM, Fmb + r * ringM, N, sh, phase, tstep, lds,
                                             Q + ring_start(A.nside, r), U + ring_start(A.nside, r));
        return;
      }
    }
    ring_fold<T>(Fpb + r * ringM, Fmb + r * ringM, A.M, N, sh, phase, A.tw, tstep, false, false, l2, scratch, tid,
                 BLOCK);
    const int o = ring_dft<E, LGLO, LGHI, KIND, T>(A, N, L, +1, bofs, ch, sh, lds, tid, BLOCK);
    const cx<T>* out = l2 + o;
    const long long st = ring_start(A.nside, r);
    for (int j = tid; j < N; j += BLOCK) {
      const cx<T> z = out[j];
      __builtin_nontemporal_store((double)z.x, &Q[st + j]);
      __builtin_nontemporal_store((double)z.y, &U[st + j]);
    }
  }
}

template <int BLOCK, int E, int SPIN, int KIND, typename T>
__global__ __launch_bounds__(BLOCK) void k_ring_synth(FftArgs A, const cx<T>* __restrict__ F,
                                                      double* __restrict__ map) {
  const int slot = blockIdx.x;  // spin 2: 2 pair slot + member
  const int p = A.pairs[SPIN == 2 ? (slot >> 1) : slot];
  ring_synth_body<BLOCK, E, lg_lo(BLOCK, E), lg_hi(BLOCK, E), SPIN, KIND, T>(A, F, map, p, blockIdx.y, slot & 1);
}

template <int BLOCK, int E, int LGLO, int LGHI, int SPIN, int KIND, typename T>
__device__ __forceinline__ void ring_adjoint_body(const FftArgs& A, const double* __restrict__ map,
                                                  cx<T>* __restrict__ V, int p, int b, int half) {
  extern __shared__ double lds[];
  cx<T>* l2 = reinterpret_cast<cx<T>*>(lds);
  const int tid = threadIdx.x;
  const int rn = p;
  const bool pair = p < 2 * A.nside - 1;
  const int rs = 4 * A.nside - 2 - p;
  if (SPIN == 2 && half && !pair) return;  // the equator has no mirror
  const int N = ring_len(A.nside, rn);
  const bool sh = ring_shifted(A.nside, rn);
  int L;
  ring_kind(N, A.bluestein_all, &L);
  const long long bofs = A.bofs[p];
  const cx<T>* ch = (A.cofs[p] >= 0) ? chirps_of<T>(A) + A.cofs[p] : nullptr;
  const long long ringM = A.Mf;
  const long long sn = ring_start(A.nside, rn);
  const long long ss = ring_start(A.nside, rs);
  // e^{-i pi k / N} of a shifted ring: in the chirp (chirp-z), the twiddle table (power of
  // two), computed (direct DFT).  A shifted chirp-z ring's output bin j carries e^{-i pi j / N}
  // already, so the conjugated mirror bin N - k carries -e^{-i pi k / N} for k > 0: nflip.
  const int phase = (!sh || KIND == 2) ? 0 : (KIND == 1 ? 1 : 2);
  const int tstep = (KIND == 1) ? A.twL / (2 * N) : 0;
  const bool nflip = KIND == 2 && sh;
  if constexpr (SPIN == 0) {
    const double* mb = map + (long long)b * A.npix;
    // the map is read once: streamed past the caches, 8 loads per ring in flight per thread
    for (int j0 = tid; j0 < N; j0 += 8 * BLOCK) {
      double vn[8], vs[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int jc = min(j0 + u * BLOCK, N - 1);
        vn[u] = __builtin_nontemporal_load(&mb[sn + jc]);
        vs[u] = pair ? __builtin_nontemporal_load(&mb[ss + jc]) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (j0 + u * BLOCK < N) l2[j0 + u * BLOCK] = mk2((T)vn[u], (T)vs[u]);
    }
    __syncthreads();
    const int o = ring_dft<E, LGLO, LGHI, KIND, T>(A, N, L, -1, bofs, ch, sh, lds, tid, BLOCK);
    const cx<T>* W = l2 + o;
    cx<T>* Vn = V + ((long long)b * A.nrings + rn) * ringM;
    cx<T>* Vs = V + ((long long)b * A.nrings + rs) * ringM;
    for (int m = tid; m < A.M; m += BLOCK) {
      const int k = m % N, q = m / N;
      const cx<T> wk = W[k];
      cx<T> wr = cconj(W[(N - k) % N]);
      if (nflip && k) wr = mk2(-wr.x, -wr.y);
      const double sg = (sh && (q & 1)) ? -0.5 : 0.5;
      double2 t64 = make_double2(sg, 0.0);
      if (phase) t64 = cscale(cconj(bin_phase(phase, A.tw, tstep, k, N)), sg);
      const cx<T> t = to_cx<T>(t64);
      Vn[m] = cmul(cadd(wk, wr), t);
      if (pair) {
        const cx<T> d = csub(wk, wr);  // -i d / 2 is the south ring's spectrum
        Vs[m] = cmul(mk2(d.y, -d.x), t);
      }
    }
  } else {
    const double* Q = map + (long long)b * 2 * A.npix;
    const double* U = Q + A.npix;
    const bool h = half;
    const int r = h ? rs : rn;
    const long long st = h ? ss : sn;
    int o = 0;
    bool direct = false;
    if constexpr (KIND == 1 && LGLO == LGHI) {
      // a power-of-two ring of N = BLOCK E samples: thread tid loads samples tid + BLOCK e
      // straight into x[e], the cyclic layout the transform takes (no LDS staging)
      if (N == BLOCK * E) {
        direct = true;
        cx<T> x[E];
#pragma unroll
        for (int e0 = 0; e0 < E; e0 += 8) {  // 8 loads per field in flight per thread
          double vq[8], vu[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int j = tid + BLOCK * (e0 + u);
            vq[u] = __builtin_nontemporal_load(&Q[st + j]);
            vu[u] = __builtin_nontemporal_load(&U[st + j]);
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) x[e0 + u] = mk2((T)vq[u], (T)vu[u]);
        }
        fft_pow2<E, LGLO, LGHI, T>(x, LGLO, tid, -1, tw_of<T>(A), A.twL, lds, A.lds_bytes);
#pragma unroll
        for (int e = 0; e < E; ++e) l2[tid + BLOCK * e] = x[e];
        __syncthreads();
      }
    }
    if (!direct) {
      for (int j0 = tid; j0 < N; j0 += 8 * BLOCK) {  // 8 loads per field in flight per thread
        double vq[8], vu[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int jc = min(j0 + u * BLOCK, N - 1);
          vq[u] = __builtin_nontemporal_load(&Q[st + jc]);
          vu[u] = __builtin_nontemporal_load(&U[st + jc]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (j0 + u * BLOCK < N) l2[j0 + u * BLOCK] = mk2((T)vq[u], (T)vu[u]);
      }
      __syncthreads();
      o = ring_dft<E, LGLO, LGHI, KIND, T>(A, N, L, -1, bofs, ch, sh, lds, tid, BLOCK);
    }
    const cx<T>* W = l2 + o;
    cx<T>* Vp = V + (((long long)b * 2 + 0) * A.nrings + r) * ringM;
    cx<T>* Vm = V + (((long long)b * 2 + 1) * A.nrings + r) * ringM;
    for (int m = tid; m < A.M; m += BLOCK) {
      const int k = m % N, q = m / N;
      const double sg = (sh && (q & 1)) ? -1.0 : 1.0;
      double2 t64 = make_double2(sg, 0.0);
      if (phase) t64 = cscale(cconj(bin_phase(phase, A.tw, tstep, k, N)), sg);
      const cx<T> t = to_cx<T>(t64);
      Vp[m] = cmul(W[k], t);
      cx<T> wr = cconj(W[(N - k) % N]);
      if (nflip && k) wr = mk2(-wr.x, -wr.y);
      Vm[m] = cmul(wr, t);
    }
  }
}

template <int BLOCK, int E, int SPIN, int KIND, typename T>
__global__ __launch_bounds__(BLOCK) void k_ring_adjoint(FftArgs A, const double* __restrict__ map,
                                                        cx<T>* __restrict__ V) {
  const int slot = blockIdx.x;  // spin 2: 2 pair slot + member
  const int p = A.pairs[SPIN == 2 ? (slot >> 1) : slot];
  ring_adjoint_body<BLOCK, E, lg_lo(BLOCK, E), lg_hi(BLOCK, E), SPIN, KIND, T>(A, map, V, p, blockIdx.y, slot & 1);
}

// Chirp table c_n = e^{i pi (n^2 mod 2N) / N}, n < N, of each Bluestein ring pair, and its
// shifted copy e^{i pi ((n^2 + n) mod 2N) / N}.
__global__ void k_chirp_table(FftArgs A) {
  const int p = A.pairs[blockIdx.x];
  const long long o = A.cofs[p];
  if (o < 0) return;
  const int N = ring_len(A.nside, p);
  double2* c = const_cast<double2*>(A.chirps) + o;
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    c[n] = chirp(n, N, +1);
    c[A.chirp_shift + n] = chirp_shifted(n, N);
  }
}

// Kernel spectrum of a Bluestein ring: FFT_L of b_n = e^{-i pi n^2 / N} (n < N, mirrored),
// always in FP64.
template <int BLOCK, int E>
__global__ __launch_bounds__(BLOCK) void k_bluestein_spectrum(FftArgs A) {
  extern __shared__ double lds[];
  const int p = A.pairs[blockIdx.x];
  const long long bofs = A.bofs[p];
  if (bofs < 0) return;
  const int tid = threadIdx.x;
  const int N = ring_len(A.nside, p);
  const double2* ch = A.chirps + A.cofs[p];
  int L;
  ring_kind(N, A.bluestein_all, &L);
  const int lgL = ilog2(L);
  const int nt = L / E;
  double2 x[E];
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = tid + nt * e;
      int n = -1;
      if (i < N) n = i;
      else if (i > L - N) n = L - i;
      x[e] = (n >= 0) ? cconj(ch[n]) : make_double2(0.0, 0.0);
    }
  }
  fft_pow2<E, lg_lo(BLOCK, E), lg_hi(BLOCK, E), double>(x, lgL, tid, -1, A.tw, A.twL, lds, A.lds_bytes);
  if (tid < nt) {
    double2* B = const_cast<double2*>(A.bspec) + bofs;
#pragma unroll
    for (int e = 0; e < E; ++e) B[tid + nt * e] = x[e];
  }
}

// FP64 table -> FP32 table (the float path's twiddles, chirps and kernel spectra)
__global__ void k_round_f32(const double2* __restrict__ in, float2* __restrict__ out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = make_float2((float)in[i].x, (float)in[i].y);
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
  a.twf = p->fft.twf;
  a.twL = p->fft.lmax_fft;
  a.pairs = c.pairs;
  a.bofs = p->fft.bofs;
  a.bspec = p->fft.bspec;
  a.bspecf = p->fft.bspecf;
  a.cofs = p->fft.cofs;
  a.chirps = p->fft.chirps;
  a.chirpsf = p->fft.chirpsf;
  a.chirp_shift = p->fft.chirp_shift;
  a.qofs = p->fft.qofs;
  a.qtab = p->fft.qtab;
  a.lds_bytes = c.lds_bytes;
  a.bluestein_all = (p->flags & JHTX_FLAG_BLUESTEIN_ALL) ? 1 : 0;
  a.Mf = p->M;
  return a;
}

// launch classes: (threads, elements per thread) = (64,16) L<=1024, (256,16) L<=4096,
// (512,16) L<=8192, (512,32) L<=16384; crossed with the transform kind (0 direct,
// 1 power of two, 2 chirp-z): class index = 3 * length class + kind
constexpr int kNlen = 4;
constexpr int kBlocks[kNlen] = {64, 256, 512, 512};
constexpr int kElems[kNlen] = {16, 16, 16, 32};
constexpr int kLdsCap[kNlen] = {16384, 65536, 131072, 131072};  // bytes

inline int class_of(int L) { return (L <= 1024) ? 0 : (L <= 4096) ? 1 : (L <= 8192) ? 2 : 3; }

template <int BLOCK, int E, int SPIN, int KIND, typename T>
void launch_one(bool synth, const FftArgs& a, dim3 grid, size_t sh, hipStream_t s,
                const void* in, void* out) {
  if (sh > 65536) {
    static const bool once = [] {
      (void)hipFuncSetAttribute((const void*)k_ring_synth<BLOCK, E, SPIN, KIND, T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
      (void)hipFuncSetAttribute((const void*)k_ring_adjoint<BLOCK, E, SPIN, KIND, T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 163840);
      return true;
    }();
    (void)once;
  }
  if (synth)
    hipLaunchKernelGGL((k_ring_synth<BLOCK, E, SPIN, KIND, T>), grid, dim3(BLOCK), sh, s, a,
                       (const cx<T>*)in, (double*)out);
  else
    hipLaunchKernelGGL((k_ring_adjoint<BLOCK, E, SPIN, KIND, T>), grid, dim3(BLOCK), sh, s, a,
                       (const double*)in, (cx<T>*)out);
}

// threads of a launch class: the FP32 path runs the L <= 16384 chirp-z class as
// 1024 threads x 16 elements (the FP64 path as 512 x 32: its data would not fit 1024 lanes)
inline int class_block(int lc, bool f32) { return (lc == 3 && f32) ? 1024 : kBlocks[lc]; }

template <int SPIN, typename T>
void launch_class(bool synth, int lc, int kind, const FftArgs& a, dim3 grid, size_t sh,
                  hipStream_t s, const void* in, void* out) {
  if constexpr (sizeof(T) == 4) {  // FP64 at 1024 x 16 spills (128-register cap): 512 x 32
    if (lc == 3) {
      launch_one<1024, 16, SPIN, 2, T>(synth, a, grid, sh, s, in, out);
      return;
    }
  }
  switch (lc * 3 + kind) {
    case 0: launch_one<64, 16, SPIN, 0, T>(synth, a, grid, sh, s, in, out); break;
    case 1: launch_one<64, 16, SPIN, 1, T>(synth, a, grid, sh, s, in, out); break;
    case 2: launch_one<64, 16, SPIN, 2, T>(synth, a, grid, sh, s, in, out); break;
    case 4: launch_one<256, 16, SPIN, 1, T>(synth, a, grid, sh, s, in, out); break;
    case 5: launch_one<256, 16, SPIN, 2, T>(synth, a, grid, sh, s, in, out); break;
    case 7: launch_one<512, 16, SPIN, 1, T>(synth, a, grid, sh, s, in, out); break;
    case 8: launch_one<512, 16, SPIN, 2, T>(synth, a, grid, sh, s, in, out); break;
    default: launch_one<512, 32, SPIN, 2, T>(synth, a, grid, sh, s, in, out); break;
  }
}

void launch_spectrum(int lc, const FftArgs& a, dim3 grid, size_t sh) {
  switch (lc) {
    case 0: hipLaunchKernelGGL((k_bluestein_spectrum<64, 16>), grid, dim3(64), sh, 0, a); break;
    case 1: hipLaunchKernelGGL((k_bluestein_spectrum<256, 16>), grid, dim3(256), sh, 0, a); break;
    case 2: hipLaunchKernelGGL((k_bluestein_spectrum<512, 16>), grid, dim3(512), sh, 0, a); break;
    default: hipLaunchKernelGGL((k_bluestein_spectrum<512, 32>), grid, dim3(512), sh, 0, a); break;
  }
}

// ---------------------------------------------------------------------------------- //
// Chirp-z transforms of the polar rings split four ways (mixed-precision spin-2 plans).  A
// polar ring has N = 4 i pixels.  With bin k = 4 k' + r and pixel j = j1 + i j2 (j1 < i,
// j2 < 4), e^{2 pi i j k / N} = e^{2 pi i j1 k' / i} e^{2 pi i j1 r / N} i^{j2 r}, so
//   synthesis  x_{j1 + i j2} = sum_r i^{j2 r} u_{j1}^r Y_r[j1],
//              Y_r[j1] = sum_k' D_{4 k' + r} e^{i pi 4 k' / N} e^{2 pi i j1 k' / i},
//              u_j = e^{i pi (2 j + 1) / N}  (the ring's half-sample shift, 1/N of a turn per
//              bin, split as e^{i pi 4 k' / N} before the sub-transform and u's 1 after it);
//   adjoint    W'_{k1 + i k2} = e^{-i pi k2 / 4} sum_s (-i)^{k2 s} (w_{k1}^*)^s Z'_s[k1],
//              Z'_s[k1] = e^{-i pi k1 / N} sum_j' x_{4 j' + s} e^{-2 pi i k1 j' / i},
//              w_k = e^{2 pi i k / N},
// with W'_k = W_k e^{-i pi k / N} the shifted spectrum.  The length-i sub-transforms are
// chirp-z transforms of length Lq = 2048 or 4096 (the next power of two >= 2 i - 1) where
// the ring's own would need 8192 or 16384: a quarter of the tables, a fifth fewer butterflies,
// and 16 or 32 KB of LDS per work-group instead of 64 or 128 KB.  The sub-transforms run in
// turn; each is folded into the four output quarters (synthesis) or output bins (adjoint)
// as it completes, so no LDS holds the intermediate results.
// Per ring pair p the tables at qtab + qofs[p] are: S (Lq, the kernel spectrum, FFT of
// e^{-i pi n^2 / i}), then i entries each of
//   T1_k = e^{i pi (4 k^2 + 4 k) / N}, T2_j = e^{i pi 4 j^2 / N} (= e^{i pi j^2 / i}),
//   T3_j = e^{i pi (2 j + 1) / N}, T4_k = e^{i pi (4 k^2 + k) / N}, T5_k = e^{i pi 2 k / N},
// every exponent reduced exactly mod 2 N in integers.
// ---------------------------------------------------------------------------------- //
constexpr int kQuadMinI = 513;  // polar rings with i >= this (N > 2048) run split

__host__ __device__ inline int quad_len(int i) {  // Lq: power of two >= 2 i - 1
  int l = 128;
  while (l < 2 * i - 1) l <<= 1;
  return l;
}

// LDS behind a split ring's FFT exchange: sub-transforms 1..3's inputs, i <= Lq / 2 each
inline size_t quad_stash_bytes(int Lq) { return 3 * (size_t)(Lq / 2) * sizeof(float2); }

struct QuadTabs {
  const float2 *S, *T1, *T2, *T3, *T4, *T5;
};

__device__ __forceinline__ QuadTabs quad_tabs(const FftArgs& A, int p, int i, int Lq) {
  const float2* base = A.qtab + A.qofs[p];
  QuadTabs t;
  t.S = base;
  t.T1 = base + Lq;
  t.T2 = t.T1 + i;
  t.T3 = t.T2 + i;
  t.T4 = t.T3 + i;
  t.T5 = t.T4 + i;
  return t;
}

// e^{i pi (e mod 2N) / N} for an exponent given exactly as an integer
__device__ __forceinline__ double2 cispi_int(long long e, int N) {
  const long long r = e % (2LL * N);
  return cispi((double)r / (double)N);
}

// z times i^n (n mod 4)
__device__ __forceinline__ float2 rot_i(float2 z, int n) {
  switch (n & 3) {
    case 0: return z;
    case 1: return make_float2(-z.y, z.x);
    case 2: return make_float2(-z.x, -z.y);
    default: return make_float2(z.y, -z.x);
  }
}

// Plan-time tables of a split ring, in FP64 (rounded to FP32 afterwards)
template <int BLOCK, int E>
__global__ __launch_bounds__(BLOCK) void k_quad_tables(FftArgs A, double2* __restrict__ q64) {
  extern __shared__ double lds[];
  const int p = A.pairs[blockIdx.x];
  const int tid = threadIdx.x;
  const int N = ring_len(A.nside, p), i = N >> 2;
  const int Lq = quad_len(i);
  const int lgL = ilog2(Lq);
  double2* S = q64 + A.qofs[p];
  double2* T = S + Lq;
  // the kernel spectrum: forward FFT of b_n = e^{-i pi n^2 / i} (n < i, mirrored)
  const int nt = Lq / E;
  double2 x[E];
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = tid + nt * e;
      int n = -1;
      if (k < i) n = k;
      else if (k > Lq - i) n = Lq - k;
      x[e] = (n >= 0) ? chirp(n, i, -1) : make_double2(0.0, 0.0);
    }
  }
  fft_pow2<E, lg_lo(BLOCK, E), lg_hi(BLOCK, E), double>(x, lgL, tid, -1, A.tw, A.twL, lds, A.lds_bytes);
  if (tid < nt) {
#pragma unroll
    for (int e = 0; e < E; ++e) S[tid + nt * e] = x[e];
  }
  for (int k = tid; k < i; k += BLOCK) {
    const long long kk = k;
    T[k] = cispi_int(4 * kk * kk + 4 * kk, N);
    T[i + k] = cispi_int(4 * kk * kk, N);
    T[2 * i + k] = cispi_int(2 * kk + 1, N);
    T[3 * i + k] = cispi_int(4 * kk * kk + kk, N);
    T[4 * i + k] = cispi_int(2 * kk, N);
  }
}

// folded spectrum D_k of one spin-2 ring with M <= 2 N (no half-sample phase), FP64 sums
__device__ __forceinline__ double2 fold_bin2(const float2* __restrict__ Fp, const float2* __restrict__ Fm, int k,
                                             int N, int M) {
  // terms: Fp at m = k (+), m = k + N (-); conj(Fm) at m = N - k (-), m = 2 N - k (+)
  const float2 p0 = Fp[min(k, M - 1)], p1 = Fp[min(k + N, M - 1)];
  const float2 q0 = Fm[min(N - k, M - 1)], q1 = Fm[min(2 * N - k, M - 1)];
  double2 acc = make_double2(0.0, 0.0);
  if (k < M) {
    acc.x += (double)p0.x;
    acc.y += (double)p0.y;
  }
  if (k + N < M) {
    acc.x -= (double)p1.x;
    acc.y -= (double)p1.y;
  }
  if (N - k < M) {
    acc.x -= (double)q0.x;
    acc.y += (double)q0.y;
  }
  if (2 * N - k < M) {
    acc.x += (double)q1.x;
    acc.y -= (double)q1.y;
  }
  return acc;
}

// one length-Lq chirp-z convolution in registers: forward FFT, the kernel spectrum (conj for
// the adjoint), inverse FFT; the 1/Lq scale is left to the caller
template <int E, int LG>
__device__ __forceinline__ void quad_convolve(const FftArgs& A, float2 (&x)[E], const float2* __restrict__ S,
                                              bool adjoint, void* lds, int tid) {
  constexpr int nt = (1 << LG) / E;
#pragma unroll 1
  for (int pass = 0; pass < 2; ++pass) {
    int tp = tid;
    asm volatile("" : "+v"(tp));
    fft_pow2<E, LG, LG, float>(x, LG, tp, pass ? +1 : -1, A.twf, A.twL, lds, A.lds_bytes);
    if (pass == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float2 s = S[tp + nt * e];
        x[e] = cmul(x[e], adjoint ? cconj(s) : s);
        if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);  // bound in-flight loads
      }
    }
  }
}

// Synthesis of one polar ring (map b, half 0 north / 1 south), split four ways
// 4 waves per SIMD (at most 128 VGPRs), 1024 / BLOCK work-groups per CU: the phases of one
// ring's work-group overlap another's
template <int BLOCK, int E, int LG>
__global__ __launch_bounds__(BLOCK, 4) void k_ring_synth_q(FftArgs A, const float2* __restrict__ F,
                                                        double* __restrict__ map) {
  extern __shared__ double lds[];
  constexpr int L = 1 << LG, nt = L / E, NY = E / 2;
  static_assert(nt == BLOCK, "one element column per thread");
  const int slot = blockIdx.x;
  const int p = A.pairs[slot >> 1], half = slot & 1, b = blockIdx.y;
  const int tid = threadIdx.x;
  const int N = ring_len(A.nside, p), i = N >> 2;
  const QuadTabs tb = quad_tabs(A, p, i, L);
  const int r = half ? 4 * A.nside - 2 - p : p;
  const long long ringM = A.Mf;
  const float2* Fp = F + ((long long)b * 2 * A.nrings + r) * ringM;
  const float2* Fm = Fp + (long long)A.nrings * ringM;
  const float invL = 1.f / (float)L;
  float2 X[4][NY], up[NY];  // output quarters j2 and u_j1^r (j1 = tid + nt e)
#pragma unroll
  for (int e = 0; e < NY; ++e) {
    up[e] = make_float2(1.f, 0.f);
#pragma unroll
    for (int q = 0; q < 4; ++q) X[q][e] = make_float2(0.f, 0.f);
  }
  // Every F entry of the ring read once: element k takes bins 4 k .. 4 k + 3 (the four
  // sub-transforms' inputs, contiguous in F) times T1_k.  Sub-transform 0's stay in registers,
  // 1..3 go to this thread's own LDS slots behind the FFT exchange (no other thread reads
  // them).  Read per sub-transform, each pass fetched every 4th entry of the same lines and
  // the lines came from HBM up to four times.
  float2* const stash = reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + A.lds_bytes);
  constexpr int H = L / 2;  // i <= L / 2
  float2 x0[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = tid + nt * e;
    // loads at a clamped element with no branch around them (validity after): a branch
    // would make the element's loads wait apart from the next element's
    const int kc = min(k, i - 1);
    const float2 t1 = tb.T1[kc];
    double2 d[4];
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) d[r4] = fold_bin2(Fp, Fm, 4 * kc + r4, N, A.M);
    x0[e] = make_float2(0.f, 0.f);
    if (k < i) {
#pragma unroll
      for (int r4 = 0; r4 < 4; ++r4) {
        const float2 v = cmul(make_float2((float)d[r4].x, (float)d[r4].y), t1);
        if (r4 == 0) x0[e] = v;
        else stash[(r4 - 1) * H + k] = v;
      }
    }
    if ((e & 1) == 1) __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll 1
  for (int r4 = 0; r4 < 4; ++r4) {
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = tid + nt * e;
      x[e] = (r4 == 0) ? x0[e] : ((k < i) ? stash[(r4 - 1) * H + k] : make_float2(0.f, 0.f));
    }
    quad_convolve<E, LG>(A, x, tb.S, false, lds, tid);
#pragma unroll
    for (int e = 0; e < NY; ++e) {
      const int j1 = tid + nt * e;
      const int jc = min(j1, i - 1);  // table loads with no branch around them
      const float2 t2 = tb.T2[jc], t3 = tb.T3[jc];
      if (j1 < i) {
        const float2 y = cmul(cmul(cscale(x[e], invL), t2), up[e]);
#pragma unroll
        for (int q = 0; q < 4; ++q) X[q][e] = cadd(X[q][e], rot_i(y, q * r4));
        up[e] = cmul(up[e], t3);
      }
    }
  }
  double* Q = map + (long long)b * 2 * A.npix;
  double* U = Q + A.npix;
  const long long st = ring_start(A.nside, r);
#pragma unroll
  for (int e = 0; e < NY; ++e) {
    const int j1 = tid + nt * e;
    if (j1 < i) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        __builtin_nontemporal_store((double)X[q][e].x, &Q[st + j1 + (long long)i * q]);
        __builtin_nontemporal_store((double)X[q][e].y, &U[st + j1 + (long long)i * q]);
      }
    }
  }
}

// Adjoint of one polar ring, split four ways; writes V+ and V- of every order m directly
template <int BLOCK, int E, int LG>
__global__ __launch_bounds__(BLOCK, 4) void k_ring_adjoint_q(FftArgs A, const double* __restrict__ map,
                                                          float2* __restrict__ V) {
  extern __shared__ double lds[];
  constexpr int L = 1 << LG, nt = L / E, NY = E / 2;
  static_assert(nt == BLOCK, "one element column per thread");
  const int slot = blockIdx.x;
  const int p = A.pairs[slot >> 1], half = slot & 1, b = blockIdx.y;
  const int tid = threadIdx.x;
  const int N = ring_len(A.nside, p), i = N >> 2;
  const QuadTabs tb = quad_tabs(A, p, i, L);
  const int r = half ? 4 * A.nside - 2 - p : p;
  const double* Q = map + (long long)b * 2 * A.npix;
  const double* U = Q + A.npix;
  const long long st = ring_start(A.nside, r);
  const float invL = 1.f / (float)L;
  float2 W[4][NY], up[NY];  // output bins k1 + i k2 and (w_k1^*)^s (k1 = tid + nt e)
#pragma unroll
  for (int e = 0; e < NY; ++e) {
    up[e] = make_float2(1.f, 0.f);
#pragma unroll
    for (int q = 0; q < 4; ++q) W[q][e] = make_float2(0.f, 0.f);
  }
  // Every sample of the ring read once: element jp takes samples 4 jp .. 4 jp + 3 (the four
  // sub-transforms' inputs; two 16-byte loads each of Q and U, as a ring starts at an even
  // pixel) times conj(T2_jp).  Sub-transform 0's stay in registers, 1..3 go to this thread's
  // own LDS slots behind the FFT exchange (no other thread reads them).  Read per
  // sub-transform, each pass fetched every 4th sample of the same lines and the lines came
  // from HBM up to four times.
  float2* const stash = reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + A.lds_bytes);
  constexpr int H = L / 2;  // i <= L / 2
  const double2* Q2 = reinterpret_cast<const double2*>(Q + st);
  const double2* U2 = reinterpret_cast<const double2*>(U + st);
  float2 x0[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int jp = tid + nt * e;
    const long long jc = 2 * (long long)min(jp, i - 1);  // clamped: validity after
    const double2 q01 = Q2[jc], q23 = Q2[jc + 1], u01 = U2[jc], u23 = U2[jc + 1];
    const float2 t2c = tb.T2[min(jp, i - 1)];
    x0[e] = make_float2(0.f, 0.f);
    if (jp < i) {
      const float2 t2 = cconj(t2c);
      x0[e] = cmul(make_float2((float)q01.x, (float)u01.x), t2);
      stash[jp] = cmul(make_float2((float)q01.y, (float)u01.y), t2);
      stash[H + jp] = cmul(make_float2((float)q23.x, (float)u23.x), t2);
      stash[2 * H + jp] = cmul(make_float2((float)q23.y, (float)u23.y), t2);
    }
    if ((e & 1) == 1) __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll 1
  for (int s4 = 0; s4 < 4; ++s4) {
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int jp = tid + nt * e;
      x[e] = (s4 == 0) ? x0[e] : ((jp < i) ? stash[(s4 - 1) * H + jp] : make_float2(0.f, 0.f));
    }
    quad_convolve<E, LG>(A, x, tb.S, true, lds, tid);
#pragma unroll
    for (int e = 0; e < NY; ++e) {
      const int k1 = tid + nt * e;
      const int kc = min(k1, i - 1);  // table loads with no branch around them
      const float2 t4 = tb.T4[kc], t5 = tb.T5[kc];
      if (k1 < i) {
        const float2 z = cmul(cmul(cscale(x[e], invL), cconj(t4)), up[e]);
#pragma unroll
        for (int q = 0; q < 4; ++q) W[q][e] = cadd(W[q][e], rot_i(z, -q * s4));
        up[e] = cmul(up[e], cconj(t5));
      }
    }
  }
  // the quarter phases e^{-i pi k2 / 4}, then V+ at m = k + q N and V- at m = (N - k) mod N
  // + q N (the conjugated mirror bin: -1 for k > 0, as the shift is carried by the bins)
  const float c4 = 0.70710678118654752440f;
  const float2 ph[4] = {make_float2(1.f, 0.f), make_float2(c4, -c4), make_float2(0.f, -1.f), make_float2(-c4, -c4)};
  const long long ringM = A.Mf;
  float2* Vp = V + (((long long)b * 2 + 0) * A.nrings + r) * ringM;
  float2* Vm = V + (((long long)b * 2 + 1) * A.nrings + r) * ringM;
#pragma unroll
  for (int e = 0; e < NY; ++e) {
    const int k1 = tid + nt * e;
    if (k1 < i) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = k1 + i * q;
        const float2 wk = cmul(W[q][e], ph[q]);
        for (int m = k, qq = 0; m < A.M; m += N, ++qq) Vp[m] = (qq & 1) ? make_float2(-wk.x, -wk.y) : wk;
        const float2 wr = (k != 0) ? make_float2(-wk.x, wk.y) : cconj(wk);  // -conj(wk) for k > 0
        for (int m = (N - k) % N, qq = 0; m < A.M; m += N, ++qq) Vm[m] = (qq & 1) ? make_float2(-wr.x, -wr.y) : wr;
      }
    }
  }
}

// ---------------------------------------------------------------------------------- //
// ragged groups: every ring of every plan in one launch of 256-thread work-groups
// ---------------------------------------------------------------------------------- //
// Per plan (device-resident).  Work-group shape 256 x 16 covers transform lengths 64..4096
// (nside <= 512); each work-group takes its ring's kind at run time.
constexpr int kGroupBlock = 256, kGroupE = 16, kGroupLgLo = 6, kGroupLgHi = 12;
static_assert((1 << kGroupLgHi) == kGroupBlock * kGroupE, "the group shape holds L = 4096");

struct GroupFft {
  FftArgs A;
  long long f_off;
};

// Work-group order within a plan: map b, then ring slots from the equator to the poles
// (largest transforms first, so the cheap polar rings fill the launch's tail).
__device__ __forceinline__ void group_ring_slot(const FftArgs& A, int spin, int local, int& p, int& b, int& half) {
  const int per = (spin == 2 ? 4 : 2) * A.nside;  // spin 2: a work-group per ring of each pair
  const int slot = per - 1 - local % per;
  b = local / per;
  p = spin == 2 ? (slot >> 1) : slot;
  half = slot & 1;
}

template <int SPIN>
__global__ __launch_bounds__(kGroupBlock) void k_group_ring_synth(const GroupFft* __restrict__ G, GroupCall C) {
  int local;
  const int i = group_member(C, local);
  const FftArgs A = G[i].A;
  int p, b, half;
  group_ring_slot(A, SPIN, local, p, b, half);
  const float2* F = reinterpret_cast<const float2*>(C.ws + G[i].f_off);
  double* map = reinterpret_cast<double*>(C.out[i]);
  int L;
  switch (ring_kind(ring_len(A.nside, p), A.bluestein_all, &L)) {
    case 0:
      ring_synth_body<kGroupBlock, kGroupE, kGroupLgLo, kGroupLgHi, SPIN, 0, float>(A, F, map, p, b, half);
      break;
    case 1:
      ring_synth_body<kGroupBlock, kGroupE, kGroupLgLo, kGroupLgHi, SPIN, 1, float>(A, F, map, p, b, half);
      break;
    default:
      ring_synth_body<kGroupBlock, kGroupE, kGroupLgLo, kGroupLgHi, SPIN, 2, float>(A, F, map, p, b, half);
      break;
  }
}

template <int SPIN>
__global__ __launch_bounds__(kGroupBlock) void k_group_ring_adjoint(const GroupFft* __restrict__ G, GroupCall C) {
  int local;
  const int i = group_member(C, local);
  const FftArgs A = G[i].A;
  int p, b, half;
  group_ring_slot(A, SPIN, local, p, b, half);
  const double* map = reinterpret_cast<const double*>(C.in[i]);
  float2* V = reinterpret_cast<float2*>(C.ws + G[i].f_off);
  int L;
  switch (ring_kind(ring_len(A.nside, p), A.bluestein_all, &L)) {
    case 0:
      ring_adjoint_body<kGroupBlock, kGroupE, kGroupLgLo, kGroupLgHi, SPIN, 0, float>(A, map, V, p, b, half);
      break;
    case 1:
      ring_adjoint_body<kGroupBlock, kGroupE, kGroupLgLo, kGroupLgHi, SPIN, 1, float>(A, map, V, p, b, half);
      break;
    default:
      ring_adjoint_body<kGroupBlock, kGroupE, kGroupLgLo, kGroupLgHi, SPIN, 2, float>(A, map, V, p, b, half);
      break;
  }
}

}  // namespace

int build_fft_tables(jhtx_plan* p) {
  const int nside = p->nside;
  const int npairs = 2 * nside;
  const int ball = (p->flags & JHTX_FLAG_BLUESTEIN_ALL) ? 1 : 0;
  // the ring kernels of mixed-precision plans run in FP32 (complex = 8 bytes)
  const bool f32 = (p->precision == JHTX_PREC_MIXED);
  const int cs = f32 ? 8 : 16;
  constexpr int kNcls = 3 * kNlen;
  std::vector<int> cls_pairs[kNcls];
  int need[kNcls] = {0}, need64[kNcls] = {0};
  int twL = 16;
  std::vector<long long> bofs(npairs, -1), cofs(npairs, -1), qofs(npairs, -1);
  std::vector<int> qpairs[2];
  long long blen = 0, clen = 0, qlen = 0;
  // the polar chirp-z rings of mixed-precision spin-2 plans with i >= kQuadMinI run split
  const bool quad = f32 && p->spin == 2 && !ball;
  for (int q = 0; q < npairs; ++q) {
    const int N = ring_len(nside, q);
    int L;
    const int kind = ring_kind(N, ball, &L);
    const int lc = class_of(L);
    if (L > 16384 || (kind == 1 && 16 * N > kLdsCap[lc])) {
      set_error("ring transform length " + std::to_string(L) + " (nside=" + std::to_string(nside) +
                ") exceeds the single-work-group ring FFT (nside <= 2048)");
      return JHTX_EUNSUPPORTED;
    }
    if (kind >= 1) twL = std::max(twL, L);
    if (quad && kind == 2 && q + 1 < nside && q + 1 >= kQuadMinI) {
      const int i = q + 1, Lq = quad_len(i);
      qpairs[Lq == 2048 ? 0 : 1].push_back(q);
      qofs[q] = qlen;
      qlen += Lq + 5LL * i;
      continue;
    }
    const int c = 3 * lc + kind;
    cls_pairs[c].push_back(q);
    const int blk = class_block(lc, f32);
    // bytes: direct DFT in + roots + out; FFT data (split exchanges if above the cap);
    // the fold's FP64 partials behind the ring data when N <= BLOCK
    int nd = (kind == 0) ? 3 * N * cs : L * cs;
    if (N <= blk) nd = std::max(nd, ((N * cs + 15) & ~15) + 2 * blk * 16);
    nd = std::max(nd, N * cs);
    need[c] = std::max(need[c], nd);
    need64[c] = std::max(need64[c], (kind == 0) ? 3 * N * 16 : L * 16);
    if (kind == 1 && ring_shifted(nside, q)) twL = std::max(twL, 2 * N);  // bin phases from the table
    if (kind == 2) {
      bofs[q] = blen;
      blen += L;
      cofs[q] = clen;
      clen += N;
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
  (void)hipMemcpy(p->fft.tw, tw.data(), sizeof(double2) * twL, hipMemcpyHostToDevice);
  if (f32) {
    std::vector<float2> twf(twL);
    for (int k = 0; k < twL; ++k) twf[k] = make_float2((float)tw[k].x, (float)tw[k].y);
    e = hipMalloc(&p->fft.twf, sizeof(float2) * twL);
    if (e != hipSuccess) return JHTX_ENOMEM;
    (void)hipMemcpy(p->fft.twf, twf.data(), sizeof(float2) * twL, hipMemcpyHostToDevice);
  }
  p->fft.lmax_fft = twL;
  e = hipMalloc(&p->fft.bofs, sizeof(long long) * npairs);
  if (e != hipSuccess) return JHTX_ENOMEM;
  (void)hipMemcpy(p->fft.bofs, bofs.data(), sizeof(long long) * npairs, hipMemcpyHostToDevice);
  p->fft.bspec_len = blen;
  if (blen > 0) {
    e = hipMalloc(&p->fft.bspec, sizeof(double2) * blen);
    if (e != hipSuccess) return JHTX_ENOMEM;
    if (f32) {
      e = hipMalloc(&p->fft.bspecf, sizeof(float2) * blen);
      if (e != hipSuccess) return JHTX_ENOMEM;
    }
  }
  e = hipMalloc(&p->fft.cofs, sizeof(long long) * npairs);
  if (e != hipSuccess) return JHTX_ENOMEM;
  (void)hipMemcpy(p->fft.cofs, cofs.data(), sizeof(long long) * npairs, hipMemcpyHostToDevice);
  p->fft.chirp_shift = clen;  // the shifted tables follow the plain ones
  if (clen > 0) {
    e = hipMalloc(&p->fft.chirps, sizeof(double2) * 2 * clen);
    if (e != hipSuccess) return JHTX_ENOMEM;
    if (f32) {
      e = hipMalloc(&p->fft.chirpsf, sizeof(float2) * 2 * clen);
      if (e != hipSuccess) return JHTX_ENOMEM;
    }
  }
  p->fft.ncls = 0;
  for (int c = 0; c < kNcls; ++c) {
    FftClass& fc = p->fft.cls[c];
    fc.lc = c / 3;
    fc.kind = c % 3;
    fc.block = class_block(fc.lc, f32);
    fc.npairs = (int)cls_pairs[c].size();
    fc.lds_bytes = std::min(std::max(need[c], 512), kLdsCap[fc.lc]);
    fc.lds64_bytes = std::min(std::max(need64[c], 512), kLdsCap[fc.lc]);
    fc.pairs = nullptr;
    if (fc.npairs) {
      e = hipMalloc(&fc.pairs, sizeof(int) * fc.npairs);
      if (e != hipSuccess) return JHTX_ENOMEM;
      (void)hipMemcpy(fc.pairs, cls_pairs[c].data(), sizeof(int) * fc.npairs, hipMemcpyHostToDevice);
      p->fft.ncls++;
    }
  }
  // chirp tables, then the Bluestein kernel spectra (FP64), then their FP32 copies
  if (blen > 0) {
    const int big = kLdsCap[2];
    (void)hipFuncSetAttribute((const void*)k_bluestein_spectrum<512, 16>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
    (void)hipFuncSetAttribute((const void*)k_bluestein_spectrum<512, 32>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
    for (int c = 0; c < kNcls; ++c) {
      const FftClass& fc = p->fft.cls[c];
      if (!fc.npairs || fc.kind != 2) continue;
      hipLaunchKernelGGL(k_chirp_table, dim3(fc.npairs), dim3(256), 0, 0, make_fft_args(p, fc));
    }
    for (int c = 0; c < kNcls; ++c) {
      const FftClass& fc = p->fft.cls[c];
      if (!fc.npairs || fc.kind != 2) continue;
      FftArgs a = make_fft_args(p, fc);
      a.lds_bytes = fc.lds64_bytes;
      launch_spectrum(fc.lc, a, dim3(fc.npairs), fc.lds64_bytes);
    }
    if (f32) {
      hipLaunchKernelGGL(k_round_f32, dim3(1024), dim3(256), 0, 0, (const double2*)p->fft.chirps, p->fft.chirpsf,
                         2 * clen);
      hipLaunchKernelGGL(k_round_f32, dim3(1024), dim3(256), 0, 0, (const double2*)p->fft.bspec, p->fft.bspecf, blen);
    }
    e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      set_error(std::string("Bluestein spectrum kernel failed: ") + hipGetErrorString(e));
      return JHTX_EDEVICE;
    }
  }
  // split polar rings: launch classes (Lq = 2048: 256 threads, 4096: 512 threads, 8 elements
  // each) and their tables, computed in FP64 and rounded once
  e = hipMalloc(&p->fft.qofs, sizeof(long long) * npairs);
  if (e != hipSuccess) return JHTX_ENOMEM;
  (void)hipMemcpy(p->fft.qofs, qofs.data(), sizeof(long long) * npairs, hipMemcpyHostToDevice);
  for (int c = 0; c < 2; ++c) {
    FftClass& fc = p->fft.qcls[c];
    fc.lc = 2 + c;
    fc.kind = 3;
    fc.block = c ? 512 : 256;
    fc.npairs = (int)qpairs[c].size();
    fc.lds_bytes = (c ? 4096 : 2048) * (int)sizeof(float2);
    fc.pairs = nullptr;
    if (fc.npairs) {
      e = hipMalloc(&fc.pairs, sizeof(int) * fc.npairs);
      if (e != hipSuccess) return JHTX_ENOMEM;
      (void)hipMemcpy(fc.pairs, qpairs[c].data(), sizeof(int) * fc.npairs, hipMemcpyHostToDevice);
    }
  }
  if (qlen > 0) {
    double2* q64 = nullptr;
    e = hipMalloc(&q64, sizeof(double2) * qlen);
    if (e != hipSuccess) return JHTX_ENOMEM;
    e = hipMalloc(&p->fft.qtab, sizeof(float2) * qlen);
    if (e != hipSuccess) {
      (void)hipFree(q64);
      return JHTX_ENOMEM;
    }
    for (int c = 0; c < 2; ++c) {
      const FftClass& fc = p->fft.qcls[c];
      if (!fc.npairs) continue;
      FftArgs a = make_fft_args(p, fc);
      a.lds_bytes = 4096 * (int)sizeof(double2);  // the FP64 spectrum FFT, Lq <= 4096
      hipLaunchKernelGGL((k_quad_tables<256, 16>), dim3(fc.npairs), dim3(256), a.lds_bytes, 0, a, q64);
    }
    hipLaunchKernelGGL(k_round_f32, dim3(1024), dim3(256), 0, 0, (const double2*)q64, p->fft.qtab, qlen);
    e = hipDeviceSynchronize();
    (void)hipFree(q64);
    if (e != hipSuccess) {
      set_error(std::string("split ring table kernel failed: ") + hipGetErrorString(e));
      return JHTX_EDEVICE;
    }
  }
  return JHTX_OK;
}

void free_fft_tables(jhtx_plan* p) {
  if (p->fft.tw) (void)hipFree(p->fft.tw);
  if (p->fft.twf) (void)hipFree(p->fft.twf);
  if (p->fft.bofs) (void)hipFree(p->fft.bofs);
  if (p->fft.bspec) (void)hipFree(p->fft.bspec);
  if (p->fft.bspecf) (void)hipFree(p->fft.bspecf);
  if (p->fft.cofs) (void)hipFree(p->fft.cofs);
  if (p->fft.chirps) (void)hipFree(p->fft.chirps);
  if (p->fft.chirpsf) (void)hipFree(p->fft.chirpsf);
  for (int c = 0; c < 12; ++c)
    if (p->fft.cls[c].pairs) (void)hipFree(p->fft.cls[c].pairs);
  for (int c = 0; c < 2; ++c)
    if (p->fft.qcls[c].pairs) (void)hipFree(p->fft.qcls[c].pairs);
  if (p->fft.qofs) (void)hipFree(p->fft.qofs);
  if (p->fft.qtab) (void)hipFree(p->fft.qtab);
  p->fft = FftTables();
}

static hipError_t launch_ring(const jhtx_plan* p, bool synth, const void* in, void* out, int batch,
                              hipStream_t s, long long ld) {
  const bool f32 = (p->precision == JHTX_PREC_MIXED);
  for (int c = 0; c < 12; ++c) {
    const FftClass& fc = p->fft.cls[c];
    if (!fc.npairs) continue;
    FftArgs a = make_fft_args(p, fc);
    if (ld) a.Mf = ld;
    const size_t sh = fc.lds_bytes;
    const dim3 grid(p->spin == 2 ? 2 * fc.npairs : fc.npairs, batch);  // spin 2: a work-group per ring
    if (p->spin == 0) {
      if (f32)
        launch_class<0, float>(synth, fc.lc, fc.kind, a, grid, sh, s, in, out);
      else
        launch_class<0, double>(synth, fc.lc, fc.kind, a, grid, sh, s, in, out);
    } else {
      if (f32)
        launch_class<2, float>(synth, fc.lc, fc.kind, a, grid, sh, s, in, out);
      else
        launch_class<2, double>(synth, fc.lc, fc.kind, a, grid, sh, s, in, out);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int c = 0; c < 2; ++c) {  // split polar rings (mixed spin-2 plans)
    const FftClass& fc = p->fft.qcls[c];
    if (!fc.npairs) continue;
    FftArgs a = make_fft_args(p, fc);
    if (ld) a.Mf = ld;
    const dim3 grid(2 * fc.npairs, batch);
    const size_t sh = fc.lds_bytes + quad_stash_bytes(c ? 4096 : 2048);
    if (synth) {
      if (c)
        hipLaunchKernelGGL((k_ring_synth_q<512, 8, 12>), grid, dim3(512), sh, s, a, (const float2*)in, (double*)out);
      else
        hipLaunchKernelGGL((k_ring_synth_q<256, 8, 11>), grid, dim3(256), sh, s, a, (const float2*)in, (double*)out);
    } else {
      if (c)
        hipLaunchKernelGGL((k_ring_adjoint_q<512, 8, 12>), grid, dim3(512), sh, s, a, (const double*)in, (float2*)out);
      else
        hipLaunchKernelGGL((k_ring_adjoint_q<256, 8, 11>), grid, dim3(256), sh, s, a, (const double*)in, (float2*)out);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_ring_synth(const jhtx_plan* p, const void* F, double* map, int batch,
                             hipStream_t s, long long ld) {
  return launch_ring(p, true, F, map, batch, s, ld);
}

hipError_t launch_ring_adjoint(const jhtx_plan* p, const double* map, void* V, int batch,
                               hipStream_t s) {
  return launch_ring(p, false, map, V, batch, s, 0);
}


// ---- ragged groups (host) ----
int group_fft_build(const GroupLayout& L, void** desc, int* start, int* lds_bytes) {
  std::vector<GroupFft> h(L.n);
  int lds = 512;
  start[0] = 0;
  for (int i = 0; i < L.n; ++i) {
    const jhtx_plan* p = L.plan[i];
    FftClass none;
    h[i].A = make_fft_args(p, none);
    h[i].A.pairs = nullptr;  // the group kernels enumerate every ring pair
    h[i].f_off = L.f_off[i];
    const int ball = (p->flags & JHTX_FLAG_BLUESTEIN_ALL) ? 1 : 0;
    for (int q = 0; q < 2 * p->nside; ++q) {
      const int N = ring_len(p->nside, q);
      int Lt;
      const int kind = ring_kind(N, ball, &Lt);
      if (Lt > kGroupBlock * kGroupE) {
        set_error("ragged group: ring transform length " + std::to_string(Lt) + " above " +
                  std::to_string(kGroupBlock * kGroupE) + " (nside <= " + std::to_string(kGroupNside) + ")");
        return JHTX_EUNSUPPORTED;
      }
      // LDS as for the one-plan launch classes, at the group's 256-thread shape (complex64)
      int nd = (kind == 0) ? 3 * N * 8 : Lt * 8;
      if (N <= kGroupBlock) nd = std::max(nd, ((N * 8 + 15) & ~15) + 2 * kGroupBlock * 16);
      lds = std::max(lds, std::max(nd, N * 8));
    }
    const long long n = (long long)start[i] + (long long)(p->spin == 2 ? 4 : 2) * p->nside * L.batch[i];
    if (n > 0x7fffffffLL) {
      set_error("ragged group: too many ring work-groups for one launch");
      return JHTX_EINVAL;
    }
    start[i + 1] = (int)n;
  }
  for (int i = 0; i < L.n; ++i) h[i].A.lds_bytes = lds;
  *lds_bytes = lds;
  if (hipMalloc(desc, sizeof(GroupFft) * L.n) != hipSuccess) return JHTX_ENOMEM;
  if (hipMemcpy(*desc, h.data(), sizeof(GroupFft) * L.n, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("ragged group: descriptor upload failed");
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

hipError_t group_fft_launch(bool synth, const GroupLayout& L, const void* desc, GroupCall C, int lds_bytes,
                            hipStream_t s) {
  const GroupFft* G = reinterpret_cast<const GroupFft*>(desc);
  const dim3 grid((unsigned)C.start[C.n]);
  if (synth) {
    if (L.spin == 0)
      hipLaunchKernelGGL(k_group_ring_synth<0>, grid, dim3(kGroupBlock), lds_bytes, s, G, C);
    else
      hipLaunchKernelGGL(k_group_ring_synth<2>, grid, dim3(kGroupBlock), lds_bytes, s, G, C);
  } else {
    if (L.spin == 0)
      hipLaunchKernelGGL(k_group_ring_adjoint<0>, grid, dim3(kGroupBlock), lds_bytes, s, G, C);
    else
      hipLaunchKernelGGL(k_group_ring_adjoint<2>, grid, dim3(kGroupBlock), lds_bytes, s, G, C);
  }
  return hipGetLastError();
}

}  // namespace jhtx
