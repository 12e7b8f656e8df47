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

// Binned spreading.  The grid is cut into kTile x kTile cell tiles; a point belongs to the
// tile holding its stencil's first cell.  A counting sort groups the points by tile
// (k_tile_count: per-tile counts and each point's rank; k_tile_offsets: exclusive prefix
// sum; k_tile_scatter: the permutation), then one work-group per tile spreads its points
// into an LDS copy of the tile plus a W - 1 halo with LDS FP64 atomics and adds the
// non-zero cells to the grid once (tiles and halos of neighbours overlap: global atomics).
constexpr int kTile = 32;

__global__ void k_fill_int(int* __restrict__ p, long long n, int v) {
  const long long i = gid();
  if (i < n) p[i] = v;
}

__device__ __forceinline__ int tile_of(double x, int n, int W, int& first) {
  first = wrap((long long)ceil(x - 0.5 * W), n);
  return first / kTile;
}

__global__ void k_tile_count(const double* __restrict__ loc, long long npts, int nk, int nm, int W,
                             int* __restrict__ counts, int* __restrict__ tile, int* __restrict__ rank) {
  const long long j = gid();
  if (j >= npts) return;
  int r0, c0;
  const int tr = tile_of(loc[2 * j] * (nk / kTwoPi), nk, W, r0);
  const int tc = tile_of(loc[2 * j + 1] * (nm / kTwoPi), nm, W, c0);
  const int t = tr * ((nm + kTile - 1) / kTile) + tc;
  tile[j] = t;
  rank[j] = atomicAdd(counts + t, 1);
}

// Exclusive prefix sum of counts[0, n) into offsets[0, n] by one work-group of 1024 threads:
// each thread sums a contiguous run, the run totals are scanned in LDS, then each run is
// written out.
__global__ __launch_bounds__(1024) void k_tile_offsets(const int* __restrict__ counts, int n, int* __restrict__ offsets) {
  __shared__ long long part[1024];
  const int tid = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = tid * per, e = min(n, b + per);
  long long s = 0;
  for (int i = b; i < e; ++i) s += counts[i];
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const long long v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  long long run = part[tid] - s;  // exclusive start of this thread's run
  for (int i = b; i < e; ++i) {
    offsets[i] = (int)run;
    run += counts[i];
  }
  if (tid == 1023) offsets[n] = (int)part[1023];
}

__global__ void k_tile_scatter(const int* __restrict__ tile, const int* __restrict__ rank, long long npts,
                               const int* __restrict__ offsets, int* __restrict__ perm) {
  const long long j = gid();
  if (j >= npts) return;
  perm[offsets[tile[j]] + rank[j]] = (int)j;
}

template <int W>
__global__ __launch_bounds__(256) void k_spread_tiled(const double* __restrict__ re, const double* __restrict__ im,
                                                      const double* __restrict__ loc, const int* __restrict__ perm,
                                                      const int* __restrict__ offsets, double2* g, int nk, int nm,
                                                      double beta) {
  constexpr int E = kTile + W - 1;  // tile + halo, cells per side
  __shared__ double s_re[E * E], s_im[E * E];
  const int ntc = (nm + kTile - 1) / kTile;
  const int t = blockIdx.x;
  const int p0 = offsets[t], p1 = offsets[t + 1];
  if (p0 == p1) return;
  const int tr = t / ntc, tc = t - tr * ntc;
  const int row0 = tr * kTile, col0 = tc * kTile;
  for (int i = threadIdx.x; i < E * E; i += blockDim.x) s_re[i] = s_im[i] = 0.0;
  __syncthreads();
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    const int j = perm[q];
    long long bk, bm;
    double wk[W], wm[W], unused[W];
    es_weights<W, false>(loc[2 * j] * (nk / kTwoPi), beta, bk, wk, unused);
    es_weights<W, false>(loc[2 * j + 1] * (nm / kTwoPi), beta, bm, wm, unused);
    // first stencil cell inside the tile: local offsets are in [0, kTile)
    const int lr = wrap(bk, nk) - row0, lc = wrap(bm, nm) - col0;
    const double vr = re[j], vi = im ? im[j] : 0.0;
    for (int o1 = 0; o1 < W; ++o1) {
      const int base = (lr + o1) * E + lc;
      const double ar = wk[o1] * vr, ai = wk[o1] * vi;
#pragma unroll
      for (int o2 = 0; o2 < W; ++o2) {
        unsafeAtomicAdd(s_re + base + o2, wm[o2] * ar);
        if (vi != 0.0) unsafeAtomicAdd(s_im + base + o2, wm[o2] * ai);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < E * E; i += blockDim.x) {
    const double vr = s_re[i], vi = s_im[i];
    if (vr == 0.0 && vi == 0.0) continue;
    const int r = i / E, c = i - r * E;
    double* cell = reinterpret_cast<double*>(g + (long long)wrap(row0 + r, nk) * nm + wrap(col0 + c, nm));
    unsafeAtomicAdd(cell, vr);
    if (vi != 0.0) unsafeAtomicAdd(cell + 1, vi);
  }
}

// Binned interpolation: one work-group per tile stages the tile plus its W - 1 halo from
// the grid into LDS once (rows and columns wrap), then each thread evaluates the stencils
// of the tile's points from LDS.
template <int W, bool DERIV>
__global__ __launch_bounds__(256) void k_interp_tiled(const double2* __restrict__ g, int nk, int nm, double beta,
                                                      const double* __restrict__ loc, const int* __restrict__ perm,
                                                      const int* __restrict__ offsets, long long npts,
                                                      double* __restrict__ re, double* __restrict__ im,
                                                      double* __restrict__ dloc) {
  constexpr int E = kTile + W - 1;
  __shared__ double2 s_g[E * E];
  const int ntc = (nm + kTile - 1) / kTile;
  const int t = blockIdx.x;
  const int p0 = offsets[t], p1 = offsets[t + 1];
  if (p0 == p1) return;
  const int tr = t / ntc, tc = t - tr * ntc;
  const int row0 = tr * kTile, col0 = tc * kTile;
  for (int i = threadIdx.x; i < E * E; i += blockDim.x) {
    const int r = i / E, c = i - r * E;
    s_g[i] = g[(long long)wrap(row0 + r, nk) * nm + wrap(col0 + c, nm)];
  }
  __syncthreads();
  for (int q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
    const int j = perm[q];
    long long bk, bm;
    double wk[W], dwk[W], wm[W], dwm[W];
    es_weights<W, DERIV>(loc[2 * j] * (nk / kTwoPi), beta, bk, wk, dwk);
    es_weights<W, DERIV>(loc[2 * j + 1] * (nm / kTwoPi), beta, bm, wm, dwm);
    const int lr = wrap(bk, nk) - row0, lc = wrap(bm, nm) - col0;
    double2 f = make_double2(0.0, 0.0), ft = f, fp = f;
    // bounded unroll: a full unroll of the stencil rows hoists all W x W cell loads
#pragma unroll 2
    for (int o1 = 0; o1 < W; ++o1) {
      const double2* row = s_g + (lr + o1) * E + lc;
      double2 sr = make_double2(0.0, 0.0), sd = sr;
#pragma unroll
      for (int o2 = 0; o2 < W; ++o2) {
        const double2 v = row[o2];
        sr.x = fma(wm[o2], v.x, sr.x);
        sr.y = fma(wm[o2], v.y, sr.y);
        if (DERIV) {
          sd.x = fma(dwm[o2], v.x, sd.x);
          sd.y = fma(dwm[o2], v.y, sd.y);
        }
      }
      f.x = fma(wk[o1], sr.x, f.x);
      f.y = fma(wk[o1], sr.y, f.y);
      if (DERIV) {
        ft.x = fma(dwk[o1], sr.x, ft.x);
        ft.y = fma(dwk[o1], sr.y, ft.y);
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
}

int blocks_for(long long n) { return (int)((n + 255) / 256); }

long long tile_count(int nk, int nm) { return (long long)((nk + kTile - 1) / kTile) * ((nm + kTile - 1) / kTile); }

size_t binning_bytes(long long npts, int nk, int nm) {
  return sizeof(int) * (size_t)(2 * tile_count(nk, nm) + 1 + 3 * npts);
}

// Counting sort of the points by grid tile into the workspace: [counts nt][offsets nt + 1]
// [tile npts][rank npts][perm npts] (int32).  perm lists the points of tile t at
// offsets[t] .. offsets[t + 1].
int bin_points(const double* loc, long long npts, int nk, int nm, int W, void* workspace, size_t bytes,
               hipStream_t s, const char* who, const int** perm_out, const int** offsets_out, long long* nt_out) {
  const long long nt = tile_count(nk, nm);
  if (nt > 0x7fffffffLL) {
    set_error(std::string(who) + ": grid too large for binning");
    return JHTX_EINVAL;
  }
  if (bytes < binning_bytes(npts, nk, nm)) {
    set_error(std::string(who) + ": workspace smaller than jhtx_nufft_workspace_bytes");
    return JHTX_EINVAL;
  }
  int* counts = reinterpret_cast<int*>(workspace);
  int* offsets = counts + nt;
  int* tile = offsets + nt + 1;
  int* rank = tile + npts;
  int* perm = rank + npts;
  const dim3 g(blocks_for(npts));
  hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(nt)), dim3(256), 0, s, counts, nt, 0);
  hipLaunchKernelGGL(k_tile_count, g, dim3(256), 0, s, loc, npts, nk, nm, W, counts, tile, rank);
  hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(1024), 0, s, counts, (int)nt, offsets);
  hipLaunchKernelGGL(k_tile_scatter, g, dim3(256), 0, s, tile, rank, npts, offsets, perm);
  *perm_out = perm;
  *offsets_out = offsets;
  *nt_out = nt;
  return JHTX_OK;
}

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
  if (!F || !D || nrings < 3 || nrings % 2 == 0 || lmax < 0 || spin < 0 || spin > 3) {
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
  if (!F || !D || nrings < 3 || nrings % 2 == 0 || lmax < 0 || spin < 0 || spin > 3) {
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
                      double* re, double* im, double* dloc, void* workspace, size_t workspace_bytes, void* stream) {
  if (npts == 0 && grid && valid_width(W)) return JHTX_OK;  // no points (their buffers may be NULL)
  if (!grid || !loc || !re || npts < 0 || npts > 0x7fffffffLL || !valid_width(W) || nk < W || nm < W) {
    set_error("jhtx_nufft_interp: invalid arguments");
    return JHTX_EINVAL;
  }
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2* gr = reinterpret_cast<const double2*>(grid);
  if (!workspace) {
    const dim3 g(blocks_for(npts));
    if (dloc)
      launch_interp<true>(W, g, s, gr, nk, nm, beta, loc, npts, re, im, dloc);
    else
      launch_interp<false>(W, g, s, gr, nk, nm, beta, loc, npts, re, im, dloc);
    return launch_status("jhtx_nufft_interp");
  }
  const int* perm;
  const int* offsets;
  long long nt;
  int rc = bin_points(loc, npts, nk, nm, W, workspace, workspace_bytes, s, "jhtx_nufft_interp", &perm, &offsets, &nt);
  if (rc) return rc;
  const dim3 gt((unsigned)nt);
#define JHTX_INTERP_TILED(WW)                                                                                   \
  if (dloc)                                                                                                    \
    hipLaunchKernelGGL((k_interp_tiled<WW, true>), gt, dim3(256), 0, s, gr, nk, nm, beta, loc, perm, offsets, npts, \
                       re, im, dloc);                                                                          \
  else                                                                                                         \
    hipLaunchKernelGGL((k_interp_tiled<WW, false>), gt, dim3(256), 0, s, gr, nk, nm, beta, loc, perm, offsets,     \
                       npts, re, im, dloc);
  switch (W) {
    case 6: JHTX_INTERP_TILED(6) break;
    case 8: JHTX_INTERP_TILED(8) break;
    case 10: JHTX_INTERP_TILED(10) break;
    case 14: JHTX_INTERP_TILED(14) break;
    default: JHTX_INTERP_TILED(16) break;
  }
#undef JHTX_INTERP_TILED
  return launch_status("jhtx_nufft_interp");
}

size_t jhtx_nufft_workspace_bytes(long long npts, int nk, int nm) {
  if (npts < 0 || nk < 1 || nm < 1) return 0;
  return binning_bytes(npts, nk, nm);
}

int jhtx_nufft_spread(const double* re, const double* im, const double* loc, long long npts, void* grid, int nk,
                      int nm, int W, double beta, void* workspace, size_t workspace_bytes, void* stream) {
  if (npts == 0 && grid && valid_width(W)) return JHTX_OK;  // no points (their buffers may be NULL)
  if (!grid || !loc || !re || npts < 0 || npts > 0x7fffffffLL || !valid_width(W) || nk < W || nm < W) {
    set_error("jhtx_nufft_spread: invalid arguments");
    return JHTX_EINVAL;
  }
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double2* gr = reinterpret_cast<double2*>(grid);
  const dim3 g(blocks_for(npts));
  if (!workspace) {  // direct: W^2 global atomics per point
    switch (W) {
      case 6: hipLaunchKernelGGL(k_spread<6>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
      case 8: hipLaunchKernelGGL(k_spread<8>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
      case 10: hipLaunchKernelGGL(k_spread<10>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
      case 14: hipLaunchKernelGGL(k_spread<14>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
      default: hipLaunchKernelGGL(k_spread<16>, g, dim3(256), 0, s, re, im, loc, npts, gr, nk, nm, beta); break;
    }
    return launch_status("jhtx_nufft_spread");
  }
  const int* perm;
  const int* offsets;
  long long nt;
  int rc = bin_points(loc, npts, nk, nm, W, workspace, workspace_bytes, s, "jhtx_nufft_spread", &perm, &offsets, &nt);
  if (rc) return rc;
  const dim3 gt((unsigned)nt);
  switch (W) {
    case 6: hipLaunchKernelGGL(k_spread_tiled<6>, gt, dim3(256), 0, s, re, im, loc, perm, offsets, gr, nk, nm, beta); break;
    case 8: hipLaunchKernelGGL(k_spread_tiled<8>, gt, dim3(256), 0, s, re, im, loc, perm, offsets, gr, nk, nm, beta); break;
    case 10: hipLaunchKernelGGL(k_spread_tiled<10>, gt, dim3(256), 0, s, re, im, loc, perm, offsets, gr, nk, nm, beta); break;
    case 14: hipLaunchKernelGGL(k_spread_tiled<14>, gt, dim3(256), 0, s, re, im, loc, perm, offsets, gr, nk, nm, beta); break;
    default: hipLaunchKernelGGL(k_spread_tiled<16>, gt, dim3(256), 0, s, re, im, loc, perm, offsets, gr, nk, nm, beta); break;
  }
  return launch_status("jhtx_nufft_spread");
}

}  // extern "C"
