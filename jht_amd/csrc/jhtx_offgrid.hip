// Off-grid (arbitrary-point) synthesis: the stages around the Legendre step and the
// library 2-D FFTs, for one field in FP64.
//
// Follows src/jht/offgrid.py:124-160 and src/jht/_nufft.py:154-181:
//   * dfs_extend / dfs_fold: the double-Fourier-sphere map from the per-ring coefficients
//     of a Clenshaw-Curtis ring set on [0, pi] (both poles, ns rings) to a 2 pi-periodic
//     theta grid of nd = 2 ns - 2 rings, and its exact transpose;
//   * grid_fill / grid_gather: the centred Fourier modes (Nk x Nm, Nm = 2 lmax + 1) placed
//     on the oversampled nk x nm grid with the kernel deconvolution folded in, and back;
//   * interp (type 2) and spread (type 1): the W x W exponential-of-semicircle stencil
//     psi(t) = exp(beta (sqrt(1 - (2t/W)^2) - 1)), |t| < W/2, around each point, with
//     interp optionally returning d/dtheta and d/dphi (the pointing gradient).
// Layouts: ring coefficients F[c][ring][m] (c = +spin, -spin channel; spin 0 has one);
// theta-grid arrays D[j][m' + lmax]; NUFFT grids [row k][column m] row-major; loc[j] =
// (theta_j, phi_j).  Spreading adds with hardware FP64 atomics into a zeroed grid.
#include "jhtx_internal.h"

namespace jhtx {
namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ long long gid() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }

__device__ __forceinline__ int wrap(long long v, int n) {
  long long r = v % n;
  return (int)(r < 0 ? r + n : r);
}

// D[j][m' + lmax] from F: rows j < ns are the rings themselves, rows j >= ns the
// reflected rings nd - j times (-1)^(|m'| + spin); column m' >= 0 reads the +spin channel
// at m', m' < 0 the conjugate of the -spin channel at -m'.
__global__ void k_dfs_extend(const double2* __restrict__ F, int ns, int lmax, int spin, int ncomp,
                             double2* __restrict__ D) {
  const int M = lmax + 1, Nm = 2 * lmax + 1, nd = 2 * ns - 2;
  const long long i = gid();
  if (i >= (long long)nd * Nm) return;
  const int j = (int)(i / Nm), col = (int)(i - (long long)j * Nm);
  const int mp = col - lmax, am = mp < 0 ? -mp : mp;
  const bool south = j >= ns;
  const int jj = south ? nd - j : j;
  const double2* Fc = F + (mp < 0 ? (long long)(ncomp - 1) * ns * M : 0);
  double2 v = Fc[(long long)jj * M + am];
  if (mp < 0) v.y = -v.y;
  const double sg = (south && ((am + spin) & 1)) ? -1.0 : 1.0;
  D[i] = make_double2(sg * v.x, sg * v.y);
}

// Transpose of k_dfs_extend: V[c][ring][m].  Ring jj collects row jj and, for 0 < jj <
// ns - 1, its reflection nd - jj.  Spin 0 (one channel feeding both signs of m') sums
// the two channels.
__global__ void k_dfs_fold(const double2* __restrict__ D, int ns, int lmax, int spin, int ncomp,
                           double2* __restrict__ V) {
  const int M = lmax + 1, Nm = 2 * lmax + 1, nd = 2 * ns - 2;
  const long long i = gid();
  if (i >= (long long)ns * M) return;
  const int jj = (int)(i / M), m = (int)(i - (long long)jj * M);
  const bool refl = jj >= 1 && jj <= ns - 2;
  const double sg = ((m + spin) & 1) ? -1.0 : 1.0;
  double2 p = D[(long long)jj * Nm + lmax + m];
  if (refl) {
    const double2 q = D[(long long)(nd - jj) * Nm + lmax + m];
    p = make_double2(p.x + sg * q.x, p.y + sg * q.y);
  }
  double2 n = make_double2(0.0, 0.0);
  if (m >= 1) {
    n = D[(long long)jj * Nm + lmax - m];
    if (refl) {
      const double2 q = D[(long long)(nd - jj) * Nm + lmax - m];
      n = make_double2(n.x + sg * q.x, n.y + sg * q.y);
    }
    n.y = -n.y;
  }
  if (ncomp == 1) {
    V[i] = make_double2(p.x + n.x, p.y + n.y);
  } else {
    V[i] = p;
    V[(long long)ns * M + i] = n;
  }
}

// C[i][jm] of the oversampled grid: mode k = i (i < Nk/2) or i - nk (i >= nk - Nk/2), mode
// m' = jm (jm <= lmax) or jm - nm (jm >= nm - lmax), from X[k mod Nk][m' + lmax] scaled by
// sk[k + Nk/2] sm[m' + lmax]; every other cell is zero.
__global__ void k_grid_fill(const double2* __restrict__ X, int Nk, int lmax, const double* __restrict__ sk,
                            const double* __restrict__ sm, double2* __restrict__ C, int nk, int nm) {
  const long long i = gid();
  if (i >= (long long)nk * nm) return;
  const int r = (int)(i / nm), c = (int)(i - (long long)r * nm);
  const int h = Nk / 2, Nm = 2 * lmax + 1;
  double2 out = make_double2(0.0, 0.0);
  const bool rk = r < h || r >= nk - h;
  const bool cm = c <= lmax || c >= nm - lmax;
  if (rk && cm) {
    const int k = r < h ? r : r - nk;
    const int mp = c <= lmax ? c : c - nm;
    const double s = sk[k + h] * sm[mp + lmax];
    const double2 v = X[(long long)(k < 0 ? k + Nk : k) * Nm + mp + lmax];
    out = make_double2(s * v.x, s * v.y);
  }
  C[i] = out;
}

// Transpose of k_grid_fill: X[k mod Nk][m' + lmax] = s C[cell of (k, m')].
__global__ void k_grid_gather(const double2* __restrict__ C, int Nk, int lmax, const double* __restrict__ sk,
                              const double* __restrict__ sm, double2* __restrict__ X, int nk, int nm) {
  const int Nm = 2 * lmax + 1, h = Nk / 2;
  const long long i = gid();
  if (i >= (long long)Nk * Nm) return;
  const int kr = (int)(i / Nm), col = (int)(i - (long long)kr * Nm);
  const int k = kr < h ? kr : kr - Nk;
  const int mp = col - lmax;
  const double s = sk[k + h] * sm[col];
  const double2 v = C[(long long)(k < 0 ? k + nk : k) * nm + (mp < 0 ? mp + nm : mp)];
  X[i] = make_double2(s * v.x, s * v.y);
}

// ES weights of the W cells base + o (o < W) around xi, and optionally d psi / d xi.
template <int W, bool DERIV>
__device__ __forceinline__ void es_weights(double xi, double beta, long long& base, double (&w)[W],
                                           double (&dw)[W]) {
  base = (long long)ceil(xi - 0.5 * W);
#pragma unroll
  for (int o = 0; o < W; ++o) {
    const double t = xi - (double)(base + o);
    const double z = t * (2.0 / W);
    const double u = 1.0 - z * z;
    double v = 0.0, d = 0.0;
    if (u > 0.0) {
      const double r = sqrt(u);
      v = exp(beta * (r - 1.0));
      if (DERIV) d = -v * beta * z * (2.0 / W) / r;
    }
    w[o] = v;
    if (DERIV) dw[o] = d;
  }
}

// Type 2 at point j: f = sum_{o1,o2} wk[o1] wm[o2] g[row][col]; with DERIV also
// df/dtheta = (nk / 2 pi) sum dwk wm g and df/dphi = (nm / 2 pi) sum wk dwm g.
// dloc layout [4][npts]: Re d/dtheta, Im d/dtheta, Re d/dphi, Im d/dphi.
template <int W, bool DERIV>
__global__ __launch_bounds__(256) void k_interp(const double2* __restrict__ g, int nk, int nm, double beta,
                                                const double* __restrict__ loc, long long npts,
                                                double* __restrict__ re, double* __restrict__ im,
                                                double* __restrict__ dloc) {
  const long long j = gid();
  if (j >= npts) return;
  const double xk = loc[2 * j] * (nk / kTwoPi), xm = loc[2 * j + 1] * (nm / kTwoPi);
  long long bk, bm;
  double wk[W], dwk[W], wm[W], dwm[W];
  es_weights<W, DERIV>(xk, beta, bk, wk, dwk);
  es_weights<W, DERIV>(xm, beta, bm, wm, dwm);
  int cols[W];
#pragma unroll
  for (int o = 0; o < W; ++o) cols[o] = wrap(bm + o, nm);
  double2 f = make_double2(0.0, 0.0), ft = f, fp = f;
  for (int o1 = 0; o1 < W; ++o1) {
    const double2* row = g + (long long)wrap(bk + o1, nk) * nm;
    double2 s = make_double2(0.0, 0.0), sd = s;
#pragma unroll
    for (int o2 = 0; o2 < W; ++o2) {
      const double2 v = row[cols[o2]];
      s.x = fma(wm[o2], v.x, s.x);
      s.y = fma(wm[o2], v.y, s.y);
      if (DERIV) {
        sd.x = fma(dwm[o2], v.x, sd.x);
        sd.y = fma(dwm[o2], v.y, sd.y);
      }
    }
    f.x = fma(wk[o1], s.x, f.x);
    f.y = fma(wk[o1], s.y, f.y);
    if (DERIV) {
      ft.x = fma(dwk[o1], s.x, ft.x);
      ft.y = fma(dwk[o1], s.y, ft.y);
      fp.x = fma(wk[o1], sd.x, fp.x);
      fp.y = fma(wk[o1], sd.y, fp.y);
    }
  }
  re[j] = f.x;
  if (im) im[j] = f.y;
  if (DERIV) {
    const double ck = nk / kTwoPi, cm = nm / kTwoPi;
    dloc[j] = ck * ft.x;
    dloc[npts + j] = ck * ft.y;
    dloc[2 * npts + j] = cm * fp.x;
    dloc[3 * npts + j] = cm * fp.y;
  }
}

// Type 1 at point j: g[row][col] += wk[o1] wm[o2] (re_j + i im_j).
template <int W>
__global__ __launch_bounds__(256) void k_spread(const double* __restrict__ re, const double* __restrict__ im,
                                                const double* __restrict__ loc, long long npts, double2* g,
                                                int nk, int nm, double beta) {
  const long long j = gid();
  if (j >= npts) return;
  const double xk = loc[2 * j] * (nk / kTwoPi), xm = loc[2 * j + 1] * (nm / kTwoPi);
  long long bk, bm;
  double wk[W], wm[W], unused[W];
  es_weights<W, false>(xk, beta, bk, wk, unused);
  es_weights<W, false>(xm, beta, bm, wm, unused);
  const double vr = re[j], vi = im ? im[j] : 0.0;
  for (int o1 = 0; o1 < W; ++o1) {
    double* row = reinterpret_cast<double*>(g + (long long)wrap(bk + o1, nk) * nm);
    const double ar = wk[o1] * vr, ai = wk[o1] * vi;
#pragma unroll
    for (int o2 = 0; o2 < W; ++o2) {
      const int c = wrap(bm + o2, nm);
      unsafeAtomicAdd(row + 2 * c, wm[o2] * ar);
      if (vi != 0.0) unsafeAtomicAdd(row + 2 * c + 1, wm[o2] * ai);
    }
  }
}

int blocks_for(long long n) { return (int)((n + 255) / 256); }

bool valid_width(int W) { return W == 6 || W == 8 || W == 10 || W == 14 || W == 16; }

int launch_status(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + " launch failed: " + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

template <bool DERIV>
void launch_interp(int W, dim3 grid, hipStream_t s, const double2* g, int nk, int nm, double beta,
                   const double* loc, long long npts, double* re, double* im, double* dloc) {
  switch (W) {
    case 6: hipLaunchKernelGGL((k_interp<6, DERIV>), grid, dim3(256), 0, s, g, nk, nm, beta, loc, npts, re, im, dloc); break;
    case 8: hipLaunchKernelGGL((k_interp<8, DERIV>), grid, dim3(256), 0, s, g, nk, nm, beta, loc, npts, re, im, dloc); break;
    case 10: hipLaunchKernelGGL((k_interp<10, DERIV>), grid, dim3(256), 0, s, g, nk, nm, beta, loc, npts, re, im, dloc); break;
    case 14: hipLaunchKernelGGL((k_interp<14, DERIV>), grid, dim3(256), 0, s, g, nk, nm, beta, loc, npts, re, im, dloc); break;
    default: hipLaunchKernelGGL((k_interp<16, DERIV>), grid, dim3(256), 0, s, g, nk, nm, beta, loc, npts, re, im, dloc); break;
  }
}

}  // namespace
}  // namespace jhtx

using namespace jhtx;

extern "C" {

int jhtx_dfs_extend(const void* F, int nrings, int lmax, int spin, void* D, void* stream) {
  if (!F || !D || nrings < 3 || nrings % 2 == 0 || lmax < 0 || (spin != 0 && spin != 2)) {
    set_error("jhtx_dfs_extend: invalid arguments");
    return JHTX_EINVAL;
  }
  const long long n = (long long)(2 * nrings - 2) * (2 * lmax + 1);
  hipLaunchKernelGGL(k_dfs_extend, dim3(blocks_for(n)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(F), nrings, lmax, spin, spin ? 2 : 1,
                     reinterpret_cast<double2*>(D));
  return launch_status("jhtx_dfs_extend");
}

int jhtx_dfs_fold(const void* D, int nrings, int lmax, int spin, void* F, void* stream) {
  if (!F || !D || nrings < 3 || nrings % 2 == 0 || lmax < 0 || (spin != 0 && spin != 2)) {
    set_error("jhtx_dfs_fold: invalid arguments");
    return JHTX_EINVAL;
  }
  const long long n = (long long)nrings * (lmax + 1);
  hipLaunchKernelGGL(k_dfs_fold, dim3(blocks_for(n)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(D), nrings, lmax, spin, spin ? 2 : 1,
                     reinterpret_cast<double2*>(F));
  return launch_status("jhtx_dfs_fold");
}

int jhtx_nufft_grid_fill(const void* X, int Nk, int lmax, const double* sk, const double* sm, void* C,
                         int nk, int nm, void* stream) {
  if (!X || !C || !sk || !sm || Nk < 2 || Nk % 2 || lmax < 0 || nk < Nk || nm < 2 * lmax + 1) {
    set_error("jhtx_nufft_grid_fill: invalid arguments");
    return JHTX_EINVAL;
  }
  const long long n = (long long)nk * nm;
  hipLaunchKernelGGL(k_grid_fill, dim3(blocks_for(n)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(X), Nk, lmax, sk, sm, reinterpret_cast<double2*>(C), nk, nm);
  return launch_status("jhtx_nufft_grid_fill");
}

int jhtx_nufft_grid_gather(const void* C, int Nk, int lmax, const double* sk, const double* sm, void* X,
                           int nk, int nm, void* stream) {
  if (!X || !C || !sk || !sm || Nk < 2 || Nk % 2 || lmax < 0 || nk < Nk || nm < 2 * lmax + 1) {
    set_error("jhtx_nufft_grid_gather: invalid arguments");
    return JHTX_EINVAL;
  }
  const long long n = (long long)Nk * (2 * lmax + 1);
  hipLaunchKernelGGL(k_grid_gather, dim3(blocks_for(n)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     reinterpret_cast<const double2*>(C), Nk, lmax, sk, sm, reinterpret_cast<double2*>(X), nk, nm);
  return launch_status("jhtx_nufft_grid_gather");
}

int jhtx_nufft_interp(const void* grid, int nk, int nm, int W, double beta, const double* loc, long long npts,
                      double* re, double* im, double* dloc, void* stream) {
  if (npts == 0 && grid && valid_width(W)) return JHTX_OK;  // no points (their buffers may be NULL)
  if (!grid || !loc || !re || npts < 0 || !valid_width(W) || nk < W || nm < W) {
    set_error("jhtx_nufft_interp: invalid arguments");
    return JHTX_EINVAL;
  }
  const dim3 g(blocks_for(npts));
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2* gr = reinterpret_cast<const double2*>(grid);
  if (dloc)
    launch_interp<true>(W, g, s, gr, nk, nm, beta, loc, npts, re, im, dloc);
  else
    launch_interp<false>(W, g, s, gr, nk, nm, beta, loc, npts, re, im, dloc);
  return launch_status("jhtx_nufft_interp");
}

int jhtx_nufft_spread(const double* re, const double* im, const double* loc, long long npts, void* grid, int nk,
                      int nm, int W, double beta, void* stream) {
  if (npts == 0 && grid && valid_width(W)) return JHTX_OK;  // no points (their buffers may be NULL)
  if (!grid || !loc || !re || npts < 0 || !valid_width(W) || nk < W || nm < W) {
    set_error("jhtx_nufft_spread: invalid arguments");
    return JHTX_EINVAL;
  }
  const dim3 g(blocks_for(npts));
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double2* gr = reinterpret_cast<double2*>(grid);
  switch (W) {
    case 6: hipLaunchKernelGGL(k_spread<6>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
    case 8: hipLaunchKernelGGL(k_spread<8>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
    case 10: hipLaunchKernelGGL(k_spread<10>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
    case 14: hipLaunchKernelGGL(k_spread<14>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
    default: hipLaunchKernelGGL(k_spread<16>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
  }
  return launch_status("jhtx_nufft_spread");
}

}  // extern "C"
