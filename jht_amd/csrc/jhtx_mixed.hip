// Mixed-precision batched recursion + contraction: the l-recursion stays in FP64 on the
// vector ALUs (one ring per lane, exponent-tracked exactly as in the FP64 kernels); the
// contraction over l runs on the FP32 matrix cores for a group of MG maps at once (the
// 16-map spin-2 synthesis on the FP16 ones, with every operand split into hi + lo halves:
// k_synth_tile_s2h).
//
// Synthesis, F[t][col] = sum_l lambda_l(t) a~_l[col]: v_mfma_f32_4x4x1f32 (16 blocks of 4x4)
// takes A from lane 4b+i and B from lane 4b+j, so the recursion value of the lane's ring is
// the A operand as it stands, B is the lane's column (component lane%4) of G_l a_lm, and each
// instruction accumulates an outer product over one l.  Even and odd (l - lmin) go to
// separate accumulators (north = sum, south = difference, as in the FP64 kernels).
//
// Adjoint, b_l[col] = sum_t lambda_l(t) U[t][col]: each wave writes its FP32 recursion values
// for 32 consecutive l to an LDS tile; v_mfma_f32_16x16x4f32 then contracts 16 same-parity l
// rows against 4 rings per instruction, so the matrix core performs the ring reduction.  The
// per-wave results are summed over the work-group in FP64 and written as one FP32 partial per
// ring chunk (they are sums of FP32 tiles); the chunk reduction then finishes the sum in FP64.
#include <vector>

#include "jhtx_internal.h"

namespace jhtx {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kLog2e = 1.44269504088896338700e+00;
constexpr int CH = 64;

__device__ __forceinline__ void seed_split(double L, double sign, double& mu, int& ex) {
  if (!(L > -1.0e300)) {
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

__device__ __forceinline__ double spin2_seed_L(const double* sp, double lc, double ls) {
  double L = sp[0];
  if (sp[1] > 0.0) L = __dadd_rn(L, __dmul_rn(sp[1], lc));
  if (sp[2] > 0.0) L = __dadd_rn(L, __dmul_rn(sp[2], ls));
  return L;
}

// Work-groups are dealt round-robin to the 8 XCDs (blockIdx & 7); consecutive orders m
// alternate between XCDs so cheap high-m and costly low-m columns spread evenly.
// GROUPED (the adjoint kernels): each XCD takes groups of kMGroup consecutive m, the
// groups alternating between XCDs, and within a group the work-groups of one ring chunk
// run back to back -- they read adjacent m of the same ring rows of V, so those cache
// lines are shared in the XCD's L2 instead of being fetched once per m (7x less fetch).
// P: the work-group's slot, blockIdx.x of a one-plan launch (a ragged group launch passes
// the slot within its plan's range, which starts at a multiple of 8, so P & 7 is still the
// XCD).
template <bool GROUPED>
__device__ __forceinline__ bool decode_block(const RecArgs& A, int P, int& m, int& chunk) {
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

template <bool GROUPED>
__device__ __forceinline__ bool decode_block(const RecArgs& A, int& m, int& chunk) {
  return decode_block<GROUPED>(A, (int)blockIdx.x, m, chunk);
}

// Tiled order: each XCD runs tiles of MB consecutive orders m x CB consecutive ring chunks
// (m fastest), the MB-blocks of m alternating between XCDs.  The work-groups resident on
// an XCD at one time then share the a_lm columns of their m across CB chunks (one HBM read
// per CB chunks instead of per chunk) and write F[ring][m] for MB adjacent m (runs of
// 8 MB bytes per ring).  The grid holds Mpad x ceil(nchunk / CB) x CB slots, Mpad a
// multiple of 8 MB; decode_tile<kMGroup, 1> is decode_block<true>.
template <int MB, int CB>
__device__ __forceinline__ bool decode_tile(const RecArgs& A, int P, int& m, int& chunk) {
  const int xcd = P & 7;
  const int slot = P >> 3;
  const int in = slot % (MB * CB), blk = slot / (MB * CB);
  const int ncb = (A.nchunk + CB - 1) / CB;
  chunk = (blk % ncb) * CB + in / MB;
  m = ((blk / ncb) * 8 + xcd) * MB + in % MB;
  return m < A.M && chunk < A.nchunk;
}

// work order (decode_tile) of the 32-map spin-2 adjoint, the 16-map column-block adjoints
// and the 16-map spin-0 FP16-split synthesis
constexpr int kA_MB = 8, kA_CB = 2;

// ring coefficients of the mixed path are complex64 (the FP32 ring FFT consumes them)
__device__ __forceinline__ float2 f2(double x, double y) { return make_float2((float)x, (float)y); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

enum { FULL = 0, MASKED = 1, RECONLY = 2 };

// Stage G_l a_lm of MG spin-2 maps for the CH orders l0..l0+CH-1 into LDS as FP32.  Thread
// t takes l = l0 + t % CH and (map, E/B) pairs t / CH, t / CH + 256 / CH, ...: a wave reads
// 64 consecutive coefficients of one column (coalesced), G_l is loaded once, and all of the
// thread's loads are issued before the first use (one memory latency per chunk).
template <int MG>
__device__ __forceinline__ void load_alm_s2(const RecArgs& A, const double2* __restrict__ alm, long long cb,
                                            int l0, int lmax, int mg, int batch, int tid,
                                            double2 (&a)[2 * MG * CH / 256], double& g) {
  constexpr int NP = 2 * MG * CH / 256;  // (map, E/B) pairs per thread
  static_assert(NP * 256 == 2 * MG * CH, "256 threads cover the chunk exactly");
  const int j = tid % CH, pr0 = tid / CH;
  const int l = l0 + j;
  const bool lok = l <= lmax;
  g = lok ? A.G[cb + l] : 0.0;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    const int pr = pr0 + u * (256 / CH);  // pair = 2 map + E/B
    const int b = mg * MG + (pr >> 1);
    a[u] = (lok && b < batch) ? alm[((long long)b * 2 + (pr & 1)) * A.K + cb + l] : make_double2(0.0, 0.0);
  }
}

// the tile kernel's layout: s_a[j * AS + 4 map + comp], comp = (E.re, E.im, B.re, B.im)
template <int MG, int AS>
__device__ __forceinline__ void store_alm_s2(const double2 (&a)[2 * MG * CH / 256], double g, int tid, float* s_a) {
  constexpr int NP = 2 * MG * CH / 256;
  const int j = tid % CH, pr0 = tid / CH;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    const int pr = pr0 + u * (256 / CH);
    float* d = s_a + j * AS + 4 * (pr >> 1) + 2 * (pr & 1);
    d[0] = (float)(a[u].x * g);
    d[1] = (float)(a[u].y * g);
  }
}

// the rank-1 kernel's layout: s_a[j][comp][map]
template <int MG>
__device__ __forceinline__ void stage_alm_s2_cm(const RecArgs& A, const double2* __restrict__ alm, long long cb,
                                                int l0, int lmax, int mg, int batch, int tid, float* s_a) {
  constexpr int NP = 2 * MG * CH / 256;
  double2 a[NP];
  double g;
  load_alm_s2<MG>(A, alm, cb, l0, lmax, mg, batch, tid, a, g);
  const int j = tid % CH, pr0 = tid / CH;
#pragma unroll
  for (int u = 0; u < NP; ++u) {
    const int pr = pr0 + u * (256 / CH);
    const int map = pr >> 1, c0 = 2 * (pr & 1);
    s_a[(j * 4 + c0) * MG + map] = (float)(a[u].x * g);
    s_a[(j * 4 + c0 + 1) * MG + map] = (float)(a[u].y * g);
  }
}

// ---------------------------------------------------------------------------------- //
// spin 2 synthesis, MG = 4 NB maps per work-group (grid.y = map group)
// ---------------------------------------------------------------------------------- //
// Eight l steps of the +-2 recursions and their rank-1 MFMA updates.  The operands of the
// eight steps are already in registers and MODE is fixed for the group, so the body is
// straight-line code: no LDS wait or branch between the steps.
template <int MODE, int NB>
__device__ __forceinline__ void s2_steps8(const double2 (&ag)[8], const floatx4 (&bq)[8][NB], double x,
                                          double fp, double fm, double& pp, double& ppp, double& pm,
                                          double& pmp, floatx4 (&accP)[2][4 * NB],
                                          floatx4 (&accM)[2][4 * NB]) {
#pragma unroll
  for (int jj = 0; jj < 8; ++jj) {
    const double cp = fma(ag[jj].x, x, ag[jj].y);
    const double cm = fma(ag[jj].x, x, -ag[jj].y);
    const double np = fma(cp, pp, -ppp);
    const double nm = fma(cm, pm, -pmp);
    ppp = pp;
    pp = np;
    pmp = pm;
    pm = nm;
    if (MODE != RECONLY) {
      // mask in FP64: a lane still below 2^-256 may hold a mantissa above the FP32 range
      const float ap = (float)((MODE == MASKED) ? np * fp : np);
      const float am = (float)((MODE == MASKED) ? nm * fm : nm);
#pragma unroll
      for (int g = 0; g < 4 * NB; ++g) {
        accP[jj & 1][g] = __builtin_amdgcn_mfma_f32_4x4x1f32(ap, bq[jj][g >> 2][g & 3], accP[jj & 1][g], 0, 0, 0);
        accM[jj & 1][g] = __builtin_amdgcn_mfma_f32_4x4x1f32(am, bq[jj][g >> 2][g & 3], accM[jj & 1][g], 0, 0, 0);
      }
    }
  }
}

template <int NB, int MINB>
__global__ __launch_bounds__(256, MINB) void k_synth_mx_s2(RecArgs A, const double2* __restrict__ alm,
                                                     float2* __restrict__ F, int batch) {
  constexpr int MG = 4 * NB;
  __shared__ double2 s_ag[CH];
  __shared__ __align__(16) float s_a[CH][4][MG];  // [l][component: aE.re aE.im aB.re aB.im][map]
  __shared__ float s_ep[4][2][2][64][4];          // [wave][class][sign][ring][component]
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * 256 + tid;
  const bool valid = t < A.T;
  double x = 0.0, pp = 0.0, ppp = 0.0, pm = 0.0, pmp = 0.0;
  int ep = 0, em = 0;
  if (valid && lmin <= lmax) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m;
    const double lc = A.lcos[t], ls = A.lsin[t];
    double s;
    seed_split(spin2_seed_L(sd, lc, ls), sd[3], s, ep);
    ppp = -s;
    seed_split(spin2_seed_L(sd + 4, lc, ls), sd[7], s, em);
    pmp = -s;
  }
  floatx4 accP[2][MG], accM[2][MG];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int g = 0; g < MG; ++g) accP[c][g] = accM[c][g] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      double2 ag = make_double2(0.0, 0.0);
      if (l <= lmax && l != lmin) ag = make_double2(A.alpha[cb + l], A.gamma[cb + l]);
      s_ag[tid] = ag;
    }
    stage_alm_s2_cm<MG>(A, alm, cb, l0, lmax, mg, batch, tid, &s_a[0][0][0]);
    __syncthreads();
    const int nst = min(CH, (lmax - l0 + 1 + 7) & ~7);
    int j0 = 0;
    // Below 2^-256 on every lane: recursion only.  A lane's exponent never decreases, so
    // once a wave leaves this phase it does not return to it.
    while (j0 < nst && __ballot(valid && !((ep < 0) && (em < 0))) == 0ull) {
      double2 ag[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) ag[jj] = s_ag[j0 + jj];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const double np = fma(fma(ag[jj].x, x, ag[jj].y), pp, -ppp);
        const double nm = fma(fma(ag[jj].x, x, -ag[jj].y), pm, -pmp);
        ppp = pp;
        pp = np;
        pmp = pm;
        pm = nm;
      }
      rescale_pair(pp, ppp, ep);
      rescale_pair(pm, pmp, em);
      j0 += 8;
    }
    for (; j0 < nst; j0 += 8) {
      const int anydeep = __ballot(valid && ((ep < 0) || (em < 0))) != 0ull;
      const double fp = (ep == 0) ? 1.0 : 0.0, fm = (em == 0) ? 1.0 : 0.0;
      double2 ag[8];
      floatx4 bq[8][NB];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        ag[jj] = s_ag[j0 + jj];
#pragma unroll
        for (int n = 0; n < NB; ++n) bq[jj][n] = *reinterpret_cast<const floatx4*>(&s_a[j0 + jj][lane & 3][4 * n]);
      }
      // masks are exact 1/0 factors (1 on every lane outside the deep region)
      s2_steps8<MASKED, NB>(ag, bq, x, fp, fm, pp, ppp, pm, pmp, accP, accM);
      if (anydeep) {
        rescale_pair(pp, ppp, ep);
        rescale_pair(pm, pmp, em);
      }
    }
  }
  // epilogue: lane l holds rings 4(l/4)+r, component l%4; gather a ring's 4 components
  const int swap = (lmin - m) & 1;  // classes counted from lmin; reflection needs (l - m)
  const long long ringM = A.M;
#pragma unroll
  for (int g = 0; g < MG; ++g) {
    const int b = mg * MG + g;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        s_ep[wave][c ^ swap][0][(lane & ~3) + r][lane & 3] = accP[c][g][r];
        s_ep[wave][c ^ swap][1][(lane & ~3) + r][lane & 3] = accM[c][g][r];
      }
    wave_sync();
    if (valid && b < batch && lmin <= lmax) {
      double P0[4], P1[4], M0[4], M1[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        P0[k] = s_ep[wave][0][0][lane][k];
        P1[k] = s_ep[wave][1][0][lane][k];
        M0[k] = s_ep[wave][0][1][lane][k];
        M1[k] = s_ep[wave][1][1][lane][k];
      }
      // PE = (P[0] + i P[1]), PB = (P[2] + i P[3]); Fp = -PE - i PB ; Fm = -ME + i MB
      float2* Fp = F + (long long)b * 2 * A.nrings * ringM;
      float2* Fm = Fp + (long long)A.nrings * ringM;
      const double pe_r = P0[0] + P1[0], pe_i = P0[1] + P1[1], pb_r = P0[2] + P1[2], pb_i = P0[3] + P1[3];
      const double me_r = M0[0] + M1[0], me_i = M0[1] + M1[1], mb_r = M0[2] + M1[2], mb_i = M0[3] + M1[3];
      Fp[(long long)t * ringM + m] = f2(-pe_r + pb_i, -pe_i - pb_r);
      Fm[(long long)t * ringM + m] = f2(-me_r - mb_i, -me_i + mb_r);
      if (t < A.T - 1) {
        const long long rs = (long long)(A.nrings - 1 - t) * ringM + m;
        const double de_r = M0[0] - M1[0], de_i = M0[1] - M1[1], db_r = M0[2] - M1[2], db_i = M0[3] - M1[3];
        const double qe_r = P0[0] - P1[0], qe_i = P0[1] - P1[1], qb_r = P0[2] - P1[2], qb_i = P0[3] - P1[3];
        Fp[rs] = f2(-de_r + db_i, -de_i - db_r);
        Fm[rs] = f2(-qe_r - qb_i, -qe_i + qb_r);
      }
    } else if (valid && b < batch) {
      float2* Fp = F + (long long)b * 2 * A.nrings * ringM;
      float2* Fm = Fp + (long long)A.nrings * ringM;
      Fp[(long long)t * ringM + m] = Fm[(long long)t * ringM + m] = f2(0.0, 0.0);
      if (t < A.T - 1) {
        const long long rs = (long long)(A.nrings - 1 - t) * ringM + m;
        Fp[rs] = Fm[rs] = f2(0.0, 0.0);
      }
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------------- //
// spin 2 synthesis on 16x16x4 tiles, MG = 4 NCB maps per work-group (grid.y = map group)
// ---------------------------------------------------------------------------------- //
// Lane = (ring, sign): lanes 0-31 run the +2 recursion of 32 rings, lanes 32-63 the -2
// recursion of the same rings, so a wave holds one recursion per lane.  Every 8 l-steps
// the wave's FP32 values go to an LDS tile [step][lane]; v_mfma_f32_16x16x4f32 then
// contracts 4 same-parity steps (K) of 16 (ring, sign) rows (M) against 16 columns
// (4 maps x 4 components, N) per instruction: 4 row blocks x 2 parities x NCB column
// blocks = 8 NCB instructions per 8 steps, 32 cycles each, for 32 rings x MG maps.  Against
// the rank-1 4x4x1 form this issues a quarter of the matrix instructions and runs half the
// FP64 instructions per lane, for the same matrix-core work.
constexpr int kTileTS = 72;  // tile row stride (floats): A reads at rows 2k+p, bank 16k+i

template <int NCB>
__device__ __forceinline__ void synth_tile_s2_body(const RecArgs& A, const double2* __restrict__ alm,
                                                   float2* __restrict__ F, int batch, int P, int mg) {
  constexpr int MG = 4 * NCB;
  constexpr int AS = 4 * MG + 8;  // s_a row stride (floats): B reads at rows 2k+p, bank 16k+j
  constexpr int RINGS = 128;      // rings per work-group (4 waves x 32)
  // loop: s_a [CH][AS] and the wave tiles; epilogue: per-wave staging of D (aliases both)
  constexpr int LOOP_FLOATS = CH * AS + 4 * 2 * 8 * kTileTS;  // s_a + double-buffered wave tiles
  constexpr int EPI_FLOATS = 4 * 2 * 64 * 16;
  __shared__ double2 s_ag[CH][2];  // [l][sign]: (alpha, +gamma) and (alpha, -gamma)
  __shared__ __align__(16) float s_buf[LOOP_FLOATS > EPI_FLOATS ? LOOP_FLOATS : EPI_FLOATS];
  float* s_a = s_buf;                  // [l][16 cb + 4 g + component]
  float* tiles = s_buf + CH * AS;      // [wave][buffer][step][lane]
  int m, chunk;
  if (!decode_block<true>(A, P, m, chunk)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r32 = lane & 31, sg = lane >> 5;  // ring within the wave, recursion sign
  const int ii = lane & 15, kk = lane >> 4;   // MFMA lane decomposition
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  const long long cb0 = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * RINGS + wave * 32 + r32;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid && lmin <= lmax) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m + 4 * sg;
    double sv;
    seed_split(spin2_seed_L(sd, A.lcos[t], A.lsin[t]), sd[3], sv, ex);
    mup = -sv;
  }
  floatx4 D[2][4][NCB];  // [parity][row block][column block]
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int c = 0; c < NCB; ++c) D[p][rb][c] = floatx4{0.f, 0.f, 0.f, 0.f};
  float* const tile0 = tiles + wave * 2 * 8 * kTileTS;
  float* const tile1 = tile0 + 8 * kTileTS;
  // the next chunk's coefficients are loaded into registers while this chunk computes
  constexpr int NP = 2 * MG * CH / 256;
  double2 pf_a[NP], pf_ag = make_double2(0.0, 0.0);
  double pf_g = 0.0;
  auto prefetch = [&](int l0n) {
    if (tid < 2 * CH) {
      const int l = l0n + (tid >> 1);
      pf_ag = make_double2(0.0, 0.0);
      if (l <= lmax && l != lmin)
        pf_ag = make_double2(A.alpha[cb0 + l], (tid & 1) ? -A.gamma[cb0 + l] : A.gamma[cb0 + l]);
    }
    load_alm_s2<MG>(A, alm, cb0, l0n, lmax, mg, batch, tid, pf_a, pf_g);
  };
  if (lmin <= lmax) prefetch(lmin);
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < 2 * CH) s_ag[tid >> 1][tid & 1] = pf_ag;
    store_alm_s2<MG, AS>(pf_a, pf_g, tid, s_a);
    __syncthreads();
    if (l0 + CH <= lmax) prefetch(l0 + CH);
    const int nst = min(CH, (lmax - l0 + 1 + 7) & ~7);
    int j0 = 0;
    // every lane below 2^-256: recursion only (a lane's exponent never decreases)
    while (j0 < nst && __ballot(valid && ex == 0) == 0ull) {
      double2 ag[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) ag[jj] = s_ag[j0 + jj][sg];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const double n = fma(fma(ag[jj].x, x, ag[jj].y), mu, -mup);
        mup = mu;
        mu = n;
      }
      rescale_pair(mu, mup, ex);
      j0 += 8;
    }
    // Software pipeline over the 8-step groups: the matrix instructions of group j and the
    // FP64 recursion of group j + 1 (into the other tile buffer) form one straight-line
    // block that the scheduler is free to order (measured: 79.0 -> 75.9 ms at nside 2048,
    // 16 maps; forcing a strict 1 matrix : 2 vector interleave measured slower, 81.6 ms).
    // The rescale is branch-free (a no-op for lanes at exponent 0), so no branch splits the
    // block; the arithmetic and its order are those of the plain loop.
    auto rec_group = [&](int jg, float* tl) {
      const double f = (ex == 0) ? 1.0 : 0.0;
      double2 ag[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) ag[jj] = s_ag[jg + jj][sg];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const double n = fma(fma(ag[jj].x, x, ag[jj].y), mu, -mup);
        mup = mu;
        mu = n;
        // the 1/0 mask in FP64 first: a lane below 2^-256 may hold a mantissa above FP32 range
        tl[jj * kTileTS + lane] = (float)(n * f);
      }
      rescale_pair(mu, mup, ex);
    };
    auto mfma_group = [&](int jg, const float* tl) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        float bv[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) bv[c] = s_a[(jg + 2 * kk + p) * AS + 16 * c + ii];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const float av = tl[(2 * kk + p) * kTileTS + 16 * rb + ii];
#pragma unroll
          for (int c = 0; c < NCB; ++c)
            D[p][rb][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[c], D[p][rb][c], 0, 0, 0);
        }
      }
    };
    if (j0 < nst) {
      rec_group(j0, tile0);
      wave_sync();
      bool odd = false;  // which buffer holds group j0
      for (; j0 + 8 < nst; j0 += 8) {
        mfma_group(j0, odd ? tile1 : tile0);
        rec_group(j0 + 8, odd ? tile0 : tile1);
        wave_sync();
        odd = !odd;
      }
      mfma_group(j0, odd ? tile1 : tile0);
      wave_sync();
    }
  }
  // epilogue, one column block (4 maps) at a time: D[i][j] sits in lane j + 16 (i / 4),
  // register i % 4; row i of block rb is lane 16 rb + i of the recursion (sign = rb / 2)
  __syncthreads();  // the staging below aliases s_a and the tiles
  float* ep = s_buf + wave * 2 * 64 * 16;  // [parity][row 0..63][column 0..15]
  const int swap = (lmin - m) & 1;         // classes counted from lmin; reflection needs (l - m)
  const long long ringM = A.M;
  const int tt = chunk * RINGS + wave * 32 + r32;  // this lane's output ring
  const int half = lane >> 5;                      // this lane's maps: 2 half, 2 half + 1
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) ep[((p ^ swap) * 64 + 16 * rb + 4 * kk + r) * 16 + ii] = D[p][rb][c][r];
    wave_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int g = 2 * half + h;       // map within the column block
      const int b = mg * MG + 4 * c + g;
      if (tt < A.T && b < batch) {
        float2* Fp = F + (long long)b * 2 * A.nrings * ringM;
        float2* Fm = Fp + (long long)A.nrings * ringM;
        const long long rs = (long long)(A.nrings - 1 - tt) * ringM + m;
        if (lmin <= lmax) {
          double P0[4], P1[4], M0[4], M1[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            P0[k] = ep[(0 * 64 + r32) * 16 + 4 * g + k];
            P1[k] = ep[(1 * 64 + r32) * 16 + 4 * g + k];
            M0[k] = ep[(0 * 64 + 32 + r32) * 16 + 4 * g + k];
            M1[k] = ep[(1 * 64 + 32 + r32) * 16 + 4 * g + k];
          }
          // PE = (P[0] + i P[1]), PB = (P[2] + i P[3]); Fp = -PE - i PB ; Fm = -ME + i MB
          const double pe_r = P0[0] + P1[0], pe_i = P0[1] + P1[1], pb_r = P0[2] + P1[2], pb_i = P0[3] + P1[3];
          const double me_r = M0[0] + M1[0], me_i = M0[1] + M1[1], mb_r = M0[2] + M1[2], mb_i = M0[3] + M1[3];
          Fp[(long long)tt * ringM + m] = f2(-pe_r + pb_i, -pe_i - pb_r);
          Fm[(long long)tt * ringM + m] = f2(-me_r - mb_i, -me_i + mb_r);
          if (tt < A.T - 1) {
            const double de_r = M0[0] - M1[0], de_i = M0[1] - M1[1], db_r = M0[2] - M1[2], db_i = M0[3] - M1[3];
            const double qe_r = P0[0] - P1[0], qe_i = P0[1] - P1[1], qb_r = P0[2] - P1[2], qb_i = P0[3] - P1[3];
            Fp[rs] = f2(-de_r + db_i, -de_i - db_r);
            Fm[rs] = f2(-qe_r - qb_i, -qe_i + qb_r);
          }
        } else {
          Fp[(long long)tt * ringM + m] = Fm[(long long)tt * ringM + m] = make_float2(0.f, 0.f);
          if (tt < A.T - 1) Fp[rs] = Fm[rs] = make_float2(0.f, 0.f);
        }
      }
    }
    wave_sync();
  }
}

template <int NCB, int MINB>
__global__ __launch_bounds__(256, MINB) void k_synth_tile_s2(RecArgs A, const double2* __restrict__ alm,
                                                             float2* __restrict__ F, int batch) {
  synth_tile_s2_body<NCB>(A, alm, F, batch, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------------- //
// spin 0 synthesis, MG = 8 maps: instruction q covers maps 2q, 2q+1 (lane%4 = 2 map + re/im)
// ---------------------------------------------------------------------------------- //
template <int MG>
__global__ __launch_bounds__(256) void k_synth_mx_s0(RecArgs A, const double2* __restrict__ alm,
                                                     float2* __restrict__ F, int batch) {
  constexpr int NQ = MG / 2;
  __shared__ double s_al[CH];
  static_assert(NQ % 4 == 0, "operands are read as float4");
  __shared__ __align__(16) float s_a[CH][4][NQ];  // [l][lane%4][q]
  __shared__ float s_ep[4][2][64][4];  // [wave][class][ring][lane%4]
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * 256 + tid;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid) {
    x = A.x[t];
    const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
    double s;
    seed_split(L, (m & 1) ? -1.0 : 1.0, s, ex);
    mup = -s;
  }
  floatx4 acc[2][NQ];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[c][q] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      s_al[tid] = (l <= lmax && l != m) ? A.alpha[cb + l] : 0.0;
    }
    for (int i = tid; i < CH * 4 * NQ; i += 256) {
      const int j = i / (4 * NQ), c4 = (i / NQ) & 3, q = i % NQ;
      const int l = l0 + j, b = mg * MG + 2 * q + (c4 >> 1);
      float v = 0.f;
      if (l <= lmax && b < batch) {
        const long long idx = cb + l;
        const double2 a = alm[(long long)b * A.K + idx];
        v = (float)(((c4 & 1) ? a.y : a.x) * A.G[idx]);
      }
      s_a[j][c4][q] = v;
    }
    __syncthreads();
    const int nst = min(CH, (lmax - l0 + 1 + 7) & ~7);
    int j0 = 0;
    // every lane below 2^-256: recursion only (a lane's exponent never decreases)
    while (j0 < nst && __ballot(valid && ex == 0) == 0ull) {
      double al[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) al[jj] = s_al[j0 + jj];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const double mn = fma(al[jj] * x, mu, -mup);
        mup = mu;
        mu = mn;
      }
      rescale_pair(mu, mup, ex);
      j0 += 8;
    }
    for (; j0 < nst; j0 += 8) {
      const int anydeep = __ballot(valid && ex < 0) != 0ull;
      const double f = (ex == 0) ? 1.0 : 0.0;
      double al[8];
      floatx4 bq[8][NQ / 4];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        al[jj] = s_al[j0 + jj];
#pragma unroll
        for (int n = 0; n < NQ / 4; ++n) bq[jj][n] = *reinterpret_cast<const floatx4*>(&s_a[j0 + jj][lane & 3][4 * n]);
      }
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const double mn = fma(al[jj] * x, mu, -mup);
        mup = mu;
        mu = mn;
        // the 1/0 mask in FP64 first: a lane below 2^-256 may hold a mantissa above FP32 range
        const float a = (float)(mn * f);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
          acc[jj & 1][q] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, bq[jj][q >> 2][q & 3], acc[jj & 1][q], 0, 0, 0);
      }
      if (anydeep) rescale_pair(mu, mup, ex);
    }
  }
  const long long ringM = A.M;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) s_ep[wave][c][(lane & ~3) + r][lane & 3] = acc[c][q][r];
    wave_sync();
    if (valid) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int b = mg * MG + 2 * q + h;
        if (b >= batch) continue;
        const double c0r = s_ep[wave][0][lane][2 * h], c0i = s_ep[wave][0][lane][2 * h + 1];
        const double c1r = s_ep[wave][1][lane][2 * h], c1i = s_ep[wave][1][lane][2 * h + 1];
        float2* Fb = F + (long long)b * A.nrings * ringM;
        Fb[(long long)t * ringM + m] = f2(c0r + c1r, c0i + c1i);
        if (t < A.T - 1) Fb[(long long)(A.nrings - 1 - t) * ringM + m] = f2(c0r - c1r, c0i - c1i);
      }
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------------- //
// spin 0 synthesis on 16x16x4 tiles, 16 or 32 maps per work-group (grid.y = map group)
// ---------------------------------------------------------------------------------- //
// Lane = ring: a wave runs the recursion of 64 rings.  Every 8 l-steps the wave's FP32
// values go to an LDS tile [step][lane]; v_mfma_f32_16x16x4f32 contracts 4 same-parity
// steps (K) of 16 rings (M) against 8 maps x (re, im) (N): 2 parities x 4 row blocks x NCB
// column blocks = 8 NCB instructions per 8 steps for 64 rings x 8 NCB maps, a quarter of
// the instructions of the rank-1 4x4x1 form, with the recursion shared by 16 or 32 maps.
// ---------------------------------------------------------------------------------- //
// FP16 split of the 16-map spin-2 tile synthesis.  Every FP32 operand x is carried as two
// FP16 halves, hi = fp16(x) and lo = fp16(x - hi) (|x - hi - lo| <= 2^-22 |x|), after a
// power-of-two scale: sA for the recursion values (from the bound |mu_l| <= sqrt((2l+1)/4pi)
// / |G_l|), sB[map] for G_l a_lm (from max |a_lm| of the map and max |G|), both putting
// the operands at or below 2^12.  v_mfma_f32_16x16x32_f16 takes 8 halves per lane; each
// l step fills 4 of them with the four products hi hi, lo hi, hi lo, lo lo, so one
// instruction contracts 8 same-parity steps: half the matrix instructions of the FP32
// 16x16x4 form, at about the same cycles per instruction.  Accumulation stays FP32; the
// scales are undone in FP64 in the epilogue.
// ---------------------------------------------------------------------------------- //
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

// Partial maxima of |Re|, |Im| over a map's coefficients: block (x, y) covers map y with
// stride kScaleBlocks * 256.
__global__ __launch_bounds__(256) void k_alm_absmax(const double2* __restrict__ alm, long long n_per_map,
                                                    double* __restrict__ part) {
  __shared__ double red[256];
  const double2* a = alm + (long long)blockIdx.y * n_per_map;
  double mx = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n_per_map; i += (long long)kScaleBlocks * 256) {
    const double2 v = a[i];
    mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)blockIdx.y * kScaleBlocks + blockIdx.x] = red[0];
}
static_assert(kScaleBlocks == 256, "k_alm_scale reduces one partial per thread");

// sB[map] = 2^e with 2^e max|a_lm| max|G| <= 2^12 (1 for an all-zero or non-finite map)
__global__ __launch_bounds__(256) void k_alm_scale(const double* __restrict__ part, double gmax,
                                                   float* __restrict__ sB) {
  __shared__ double red[256];
  red[threadIdx.x] = part[(long long)blockIdx.x * kScaleBlocks + threadIdx.x];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double bound = red[0] * gmax;
    int e = 1;
    if (bound > 0.0 && bound < 1e300) (void)frexp(4096.0 / bound, &e);
    sB[blockIdx.x] = ldexpf(1.0f, min(max(e - 1, -120), 120));
  }
}

// hi = v with its mantissa cut to FP16's 10 bits (an integer mask), lo = v - hi (exact),
// both rounded to FP16 by one packed conversion: |v - hi - lo| <= 2^-21 |v| in FP16's normal
// range.  The mask also keeps the compiler from folding a preceding double -> float rounding
// into a software double -> half conversion (about 35 integer instructions per value).
__device__ __forceinline__ uint32_t split_f16(float v) {
  const float hi = __uint_as_float(__float_as_uint(v) & 0xffffe000u);
  const float lo = v - hi;
  return __builtin_bit_cast(uint32_t, halfx2{(_Float16)hi, (_Float16)lo});
}

// The same pair with hi rounded to nearest (|lo| <= 2^-11 |v|, either sign).  Where one
// operand of a product is split this way, the left-out lo.lo term has a random sign rather
// than the sign of the product (split_f16's lo always has the sign of v).
__device__ __forceinline__ uint32_t split_f16_rn(float v) {
  const _Float16 hi = (_Float16)v;
  const float lo = v - (float)hi;
  return __builtin_bit_cast(uint32_t, halfx2{hi, (_Float16)lo});
}

// split_f16_rn's pair in two instructions: the conversion writes hi to the lower half, and a
// mixed-precision FMA, v * 1 - hi (exact in FP32, one rounding to FP16), writes lo to the
// upper half.  Written out because the compiler converts hi back to FP32, subtracts and
// converts again (three more instructions per value).
__device__ __forceinline__ uint32_t split_f16_mix(float v) {
  uint32_t w = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
  asm("v_fma_mixhi_f16 %0, %1, 1.0, -%0 op_sel_hi:[0,0,1]" : "+v"(w) : "v"(v));
  return w;
}

// (hi, hi) and (lo, lo) from a split_f16 pair
__device__ __forceinline__ uint32_t dup_lo(uint32_t w) { return (w & 0xffffu) | (w << 16); }
__device__ __forceinline__ uint32_t dup_hi(uint32_t w) { return (w >> 16) | (w & 0xffff0000u); }

template <int NCB, int MINB>
__global__ __launch_bounds__(256, MINB) void k_synth_tile_s2h(RecArgs A, const double2* __restrict__ alm,
                                                              const float* __restrict__ sB, double sA,
                                                              float2* __restrict__ F, int batch) {
  constexpr int MG = 4 * NCB;
  constexpr int NCOL = 4 * MG;
  constexpr int ASH = 2 * NCOL + 8;  // s_b row stride (dwords): 64-bit B reads, 4 rows apart per lane group
  constexpr int TSH = 68;            // tile row stride (dwords): A reads 4 rows apart per lane group
  constexpr int GS = 16;             // l steps per group
  constexpr int RINGS = 128;         // rings per work-group (4 waves x 32)
  constexpr int LOOP_DW = CH * ASH + 4 * 2 * GS * TSH;
  constexpr int EPI_FLOATS = 4 * 2 * 64 * 16;
  static_assert(CH % GS == 0, "groups tile the chunk");
  __shared__ double2 s_ag[CH][2];
  __shared__ __align__(16) uint32_t s_buf[LOOP_DW > EPI_FLOATS ? LOOP_DW : EPI_FLOATS];
  uint32_t* s_b = s_buf;              // [l][2 column + {hi hi, lo lo}]
  uint32_t* tiles = s_buf + CH * ASH;  // [wave][buffer][step][lane]: (hi, lo)
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r32 = lane & 31, sg = lane >> 5;
  const int ii = lane & 15, kk = lane >> 4;
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  const long long cb0 = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * RINGS + wave * 32 + r32;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid && lmin <= lmax) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m + 4 * sg;
    double sv;
    seed_split(spin2_seed_L(sd, A.lcos[t], A.lsin[t]), sd[3], sv, ex);
    mup = -sv;
  }
  floatx4 D[2][4][NCB];  // [parity][row block][column block]
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int c = 0; c < NCB; ++c) D[p][rb][c] = floatx4{0.f, 0.f, 0.f, 0.f};
  uint32_t* const tile0 = tiles + wave * 2 * GS * TSH;
  uint32_t* const tile1 = tile0 + GS * TSH;
  // the maps' scales (published by the first barrier of the loop)
  __shared__ float s_sc[MG];
  if (tid < MG) s_sc[tid] = (mg * MG + tid < batch) ? sB[mg * MG + tid] : 0.f;
  constexpr int NP = 2 * MG * CH / 256;
  double2 pf_a[NP], pf_ag = make_double2(0.0, 0.0);
  double pf_g = 0.0;
  auto prefetch = [&](int l0n) {
    if (tid < 2 * CH) {
      const int l = l0n + (tid >> 1);
      pf_ag = make_double2(0.0, 0.0);
      if (l <= lmax && l != lmin)
        pf_ag = make_double2(A.alpha[cb0 + l], (tid & 1) ? -A.gamma[cb0 + l] : A.gamma[cb0 + l]);
    }
    load_alm_s2<MG>(A, alm, cb0, l0n, lmax, mg, batch, tid, pf_a, pf_g);
  };
  auto store_b = [&]() {
    const int j = tid % CH, pr0 = tid / CH;
#pragma unroll
    for (int u = 0; u < NP; ++u) {
      const int pr = pr0 + u * (256 / CH);
      uint32_t* d = s_b + j * ASH + 2 * (4 * (pr >> 1) + 2 * (pr & 1));
      const double gs = pf_g * (double)s_sc[pr >> 1];
      const uint32_t sr = split_f16((float)(pf_a[u].x * gs)), si = split_f16((float)(pf_a[u].y * gs));
      *reinterpret_cast<uint4*>(d) = make_uint4(dup_lo(sr), dup_hi(sr), dup_lo(si), dup_hi(si));
    }
  };
  if (lmin <= lmax) prefetch(lmin);
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < 2 * CH) s_ag[tid >> 1][tid & 1] = pf_ag;
    store_b();
    __syncthreads();
    if (l0 + CH <= lmax) prefetch(l0 + CH);
    const int nst = min(CH, (lmax - l0 + 1 + GS - 1) & ~(GS - 1));
    int j0 = 0;
    // every lane below 2^-256: recursion only (a lane's exponent never decreases)
    while (j0 < nst && __ballot(valid && ex == 0) == 0ull) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double2 ag[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) ag[jj] = s_ag[j0 + 8 * h + jj][sg];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double n = fma(fma(ag[jj].x, x, ag[jj].y), mu, -mup);
          mup = mu;
          mu = n;
        }
        rescale_pair(mu, mup, ex);
      }
      j0 += GS;
    }
    auto rec_group = [&](int jg, uint32_t* tl) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double f = (ex == 0) ? sA : 0.0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          double2 ag[4];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) ag[jj] = s_ag[jg + 8 * h + 4 * q + jj][sg];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            const double n = fma(fma(ag[jj].x, x, ag[jj].y), mu, -mup);
            mup = mu;
            mu = n;
            // the 1/0 mask in FP64 first: a lane below 2^-256 may hold a mantissa above range
            tl[(8 * h + 4 * q + jj) * TSH + lane] = split_f16((float)(n * f));
          }
        }
        rescale_pair(mu, mup, ex);
      }
    };
    // lane group kk takes steps 4 kk + p and 4 kk + 2 + p: A halves (hi, lo, hi, lo) of each,
    // B halves (hi, hi, lo, lo)
    auto mfma_group = [&](int jg, const uint32_t* tl) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        halfx8 bv[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
          const uint2 b0 = *reinterpret_cast<const uint2*>(s_b + (jg + 4 * kk + p) * ASH + 2 * (16 * c + ii));
          const uint2 b1 = *reinterpret_cast<const uint2*>(s_b + (jg + 4 * kk + 2 + p) * ASH + 2 * (16 * c + ii));
          bv[c] = __builtin_bit_cast(halfx8, make_uint4(b0.x, b0.y, b1.x, b1.y));
        }
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const uint32_t a0 = tl[(4 * kk + p) * TSH + 16 * rb + ii];
          const uint32_t a1 = tl[(4 * kk + 2 + p) * TSH + 16 * rb + ii];
          const halfx8 av = __builtin_bit_cast(halfx8, make_uint4(a0, a0, a1, a1));
#pragma unroll
          for (int c = 0; c < NCB; ++c)
            D[p][rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv[c], D[p][rb][c], 0, 0, 0);
        }
      }
    };
    if (j0 < nst) {
      rec_group(j0, tile0);
      wave_sync();
      bool odd = false;
      for (; j0 + GS < nst; j0 += GS) {
        mfma_group(j0, odd ? tile1 : tile0);
        rec_group(j0 + GS, odd ? tile0 : tile1);
        wave_sync();
        odd = !odd;
      }
      mfma_group(j0, odd ? tile1 : tile0);
      wave_sync();
    }
  }
  // epilogue (as k_synth_tile_s2), with the scales undone in FP64
  __syncthreads();
  float* ep = reinterpret_cast<float*>(s_buf) + wave * 2 * 64 * 16;
  const int swap = (lmin - m) & 1;
  const long long ringM = A.M;
  const int tt = chunk * RINGS + wave * 32 + r32;
  const int half = lane >> 5;
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) ep[((p ^ swap) * 64 + 16 * rb + 4 * kk + r) * 16 + ii] = D[p][rb][c][r];
    wave_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int g = 2 * half + h;
      const int b = mg * MG + 4 * c + g;
      if (tt < A.T && b < batch) {
        float2* Fp = F + (long long)b * 2 * A.nrings * ringM;
        float2* Fm = Fp + (long long)A.nrings * ringM;
        const long long rs = (long long)(A.nrings - 1 - tt) * ringM + m;
        if (lmin <= lmax) {
          const double inv = 1.0 / (sA * (double)sB[b]);  // a power of two: exact
          double P0[4], P1[4], M0[4], M1[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            P0[k] = inv * ep[(0 * 64 + r32) * 16 + 4 * g + k];
            P1[k] = inv * ep[(1 * 64 + r32) * 16 + 4 * g + k];
            M0[k] = inv * ep[(0 * 64 + 32 + r32) * 16 + 4 * g + k];
            M1[k] = inv * ep[(1 * 64 + 32 + r32) * 16 + 4 * g + k];
          }
          const double pe_r = P0[0] + P1[0], pe_i = P0[1] + P1[1], pb_r = P0[2] + P1[2], pb_i = P0[3] + P1[3];
          const double me_r = M0[0] + M1[0], me_i = M0[1] + M1[1], mb_r = M0[2] + M1[2], mb_i = M0[3] + M1[3];
          Fp[(long long)tt * ringM + m] = f2(-pe_r + pb_i, -pe_i - pb_r);
          Fm[(long long)tt * ringM + m] = f2(-me_r - mb_i, -me_i + mb_r);
          if (tt < A.T - 1) {
            const double de_r = M0[0] - M1[0], de_i = M0[1] - M1[1], db_r = M0[2] - M1[2], db_i = M0[3] - M1[3];
            const double qe_r = P0[0] - P1[0], qe_i = P0[1] - P1[1], qb_r = P0[2] - P1[2], qb_i = P0[3] - P1[3];
            Fp[rs] = f2(-de_r + db_i, -de_i - db_r);
            Fm[rs] = f2(-qe_r - qb_i, -qe_i + qb_r);
          }
        } else {
          Fp[(long long)tt * ringM + m] = Fm[(long long)tt * ringM + m] = make_float2(0.f, 0.f);
          if (tt < A.T - 1) Fp[rs] = Fm[rs] = make_float2(0.f, 0.f);
        }
      }
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------------- //
// Spin-2 synthesis on the FP16 matrix cores with the N/S parity folded into the B operand.
// Wave w runs one recursion sign (sg = w & 1) for the 64 rings of ring block w >> 1, so a
// work-group of NW waves covers 32 NW rings.  The lanes of sign sg hold lambda^sg of their
// north ring, which enters two sums per map: X = sum_l bX_l lambda^sg_l, the north ring of
// the field X (bp for +2, bm for -2), and Y = sum_l (-1)^(l+m) bY_l lambda^sg_l, the south
// ring of the other field (lambda^s(pi - theta) = (-1)^(l+m) lambda^-s(theta)).  B_sg
// therefore carries [X.re, X.im, s Y.re, s Y.im] per map with s = (-1)^(l+m) folded in:
// no even / odd accumulators (4 registers per 16 x 16 block instead of 8), and one
// v_mfma_f32_16x16x32_f16 contracts 8 consecutive steps.  Operands are split into FP16
// hi + lo halves after power-of-two scales, as in k_synth_tile_s2h: A halves (hi, lo, hi,
// lo) of two steps from the wave's tile, B halves (hi, hi, lo, lo) of the same steps,
// staged in LDS so that one 16-byte read is the B operand.  Work order decode_tile<MB, CB>,
// launched as <8, 2> like k_synth_s2q: groups of 8 consecutive orders m x 2 ring chunks per
// XCD (16 x 1 before: 53.3 -> 51.5 ms for 16 maps at nside 2048).
// Instantiated for 16 maps (4 waves, 128 rings, 3 work-groups per CU); 32 maps per
// recursion run k_synth_s2q below.
// ---------------------------------------------------------------------------------- //
constexpr int kP_GS = 8;   // l steps per pipeline group: one matrix K block
constexpr int kP_CH = 32;  // l steps per staged chunk of B
constexpr int kP_TS = 72;  // tile row stride (dwords): 2 kP_TS = 16 mod 64, conflict-free A reads

template <int NCB, int NW, int MINB, int MB, int CB>
__global__ __launch_bounds__(64 * NW, MINB) void k_synth_s2p(RecArgs A, const double2* __restrict__ alm,
                                                            const float* __restrict__ sB, double sA,
                                                            float2* __restrict__ F, int batch) {
  constexpr int MG = 4 * NCB;            // maps per work-group
  constexpr int NCOL = 4 * MG;           // B columns per sign
  constexpr int NT = 64 * NW;            // threads
  constexpr int RINGS = 32 * NW;         // rings per work-group
  constexpr int BROW = 4 * NCOL + 4;     // dwords per (sign, step pair) row of s_b; +4 rotates banks
  constexpr int NLP = kP_CH / 2;         // step pairs per chunk
  constexpr int SB_DW = 2 * NLP * BROW;  // s_b: [sign][step pair][column][(hh, ll) of 2 steps]
  constexpr int TILE_DW = NW * 2 * kP_GS * kP_TS;  // [wave][buffer][step][lane]: (hi, lo)
  constexpr int EP_DW = NW * 64 * 16;    // epilogue: [wave][ring][16 columns] floats
  constexpr int BUF_DW = (SB_DW + TILE_DW > EP_DW) ? SB_DW + TILE_DW : EP_DW;
  constexpr int NE = kP_CH * MG / NT;    // staged (l, map) entries per thread
  static_assert(NE * NT == kP_CH * MG, "the threads cover a chunk's entries exactly");
  static_assert(kP_CH % kP_GS == 0, "groups tile the chunk");
  // one LDS block, (alpha, gamma) first: its per-step reads then take immediate offsets
  struct Smem {
    double2 ag[kP_CH][2];  // (alpha, +gamma), (alpha, -gamma)
    float sc[MG];
    __align__(16) uint32_t buf[BUF_DW];
  };
  __shared__ Smem sm;
  auto& s_ag = sm.ag;
  float* const s_sc = sm.sc;
  uint32_t* const s_buf = sm.buf;
  uint32_t* const s_b = s_buf;
  int m, chunk;
  if (!decode_tile<MB, CB>(A, (int)blockIdx.x, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int sg = wave & 1;
  const int ii = lane & 15, kk = lane >> 4;
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  const long long cb0 = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * RINGS + 64 * (wave >> 1) + lane;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid && lmin <= lmax) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m + 4 * sg;
    double sv;
    seed_split(spin2_seed_L(sd, A.lcos[t], A.lsin[t]), sd[3], sv, ex);
    mup = -sv;
  }
  floatx4 D[4][NCB];  // [row block][column block]
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int c = 0; c < NCB; ++c) D[rb][c] = floatx4{0.f, 0.f, 0.f, 0.f};
  uint32_t* const tile0 = s_buf + SB_DW + wave * 2 * kP_GS * kP_TS;
  uint32_t* const tile1 = tile0 + kP_GS * kP_TS;
  if (tid < MG) s_sc[tid] = (mg * MG + tid < batch) ? sB[mg * MG + tid] : 0.f;
  // staging: entry e = g kP_CH + j (map g, step j) -- consecutive threads read consecutive l
  double2 pf_e[NE], pf_b[NE], pf_ag = make_double2(0.0, 0.0);
  double pf_g[NE];
  auto prefetch = [&](int l0n) {
    if (tid < kP_CH) {
      const int l = l0n + tid;
      pf_ag = (l <= lmax && l != lmin) ? make_double2(A.alpha[cb0 + l], A.gamma[cb0 + l]) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + u * NT, j = e % kP_CH, g = e / kP_CH;
      const int l = l0n + j, b = mg * MG + g;
      const bool ok = l <= lmax && b < batch;
      pf_g[u] = ok ? A.G[cb0 + l] : 0.0;
      pf_e[u] = ok ? alm[((long long)b * 2 + 0) * A.K + cb0 + l] : make_double2(0.0, 0.0);
      pf_b[u] = ok ? alm[((long long)b * 2 + 1) * A.K + cb0 + l] : make_double2(0.0, 0.0);
    }
  };
  auto store_b = [&](int l0c) {
#pragma unroll
    for (int u = 0; u < NE; ++u) {
      const int e = tid + u * NT, j = e % kP_CH, g = e / kP_CH;
      const double gs = pf_g[u] * (double)s_sc[g];
      const double er = pf_e[u].x * gs, ei = pf_e[u].y * gs, br = pf_b[u].x * gs, bi = pf_b[u].y * gs;
      // bp = -(E + iB), bm = -(E - iB)
      const uint32_t w[4] = {split_f16((float)(bi - er)), split_f16((float)(-ei - br)),
                             split_f16((float)(-er - bi)), split_f16((float)(br - ei))};
      // s = (-1)^(l+m): flip the sign bits of both halves of the folded copies
      const uint32_t neg = (((l0c + j + m) & 1) != 0) ? 0x80008000u : 0u;
      uint32_t* row = s_b + (j >> 1) * BROW + 4 * (4 * g) + 2 * (j & 1);
      // sign +: [bp.re, bp.im, s bm.re, s bm.im]; sign -: [bm.re, bm.im, s bp.re, s bp.im]
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const uint32_t wp = (c < 2) ? w[c] : (w[c] ^ neg);
        const uint32_t wm = (c < 2) ? w[c + 2] : (w[c - 2] ^ neg);
        *reinterpret_cast<uint2*>(row + 4 * c) = make_uint2(dup_lo(wp), dup_hi(wp));
        *reinterpret_cast<uint2*>(row + NLP * BROW + 4 * c) = make_uint2(dup_lo(wm), dup_hi(wm));
      }
    }
  };
  // 8 steps of the wave's recursion into tile tl (the 1/0 mask in FP64 first: a lane below
  // 2^-256 may hold a mantissa above FP32's range)
  auto rec_group = [&](int jg, uint32_t* tl) {
    const double f = (ex == 0) ? sA : 0.0;
    double2 ag[kP_GS];
#pragma unroll
    for (int jj = 0; jj < kP_GS; ++jj) ag[jj] = s_ag[jg + jj][sg];
#pragma unroll
    for (int jj = 0; jj < kP_GS; ++jj) {
      const double n = fma(fma(ag[jj].x, x, ag[jj].y), mu, -mup);
      mup = mu;
      mu = n;
      tl[jj * kP_TS + lane] = split_f16((float)(n * f));
    }
    rescale_pair(mu, mup, ex);
  };
  // lane group kk takes steps 2 kk, 2 kk + 1 of the group: A halves (hi, lo, hi, lo) of each
  // from the tile, B halves (hi, hi, lo, lo) of each from one 16-byte read of s_b
  auto mfma_group = [&](int jg, const uint32_t* tl) {
    halfx8 av[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const uint32_t a0 = tl[(2 * kk) * kP_TS + 16 * rb + ii];
      const uint32_t a1 = tl[(2 * kk + 1) * kP_TS + 16 * rb + ii];
      av[rb] = __builtin_bit_cast(halfx8, make_uint4(a0, a0, a1, a1));
    }
    const uint32_t* brow = s_b + sg * NLP * BROW + ((jg >> 1) + kk) * BROW + 4 * ii;
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
      const halfx8 bv = __builtin_bit_cast(halfx8, *reinterpret_cast<const uint4*>(brow + 64 * c));
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) D[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[rb], bv, D[rb][c], 0, 0, 0);
    }
  };
  if (lmin <= lmax) prefetch(lmin);
  for (int l0 = lmin; l0 <= lmax; l0 += kP_CH) {
    __syncthreads();
    if (tid < kP_CH) {
      s_ag[tid][0] = pf_ag;
      s_ag[tid][1] = make_double2(pf_ag.x, -pf_ag.y);
    }
    store_b(l0);
    __syncthreads();
    if (l0 + kP_CH <= lmax) prefetch(l0 + kP_CH);
    const int nst = min(kP_CH, (lmax - l0 + 1 + kP_GS - 1) & ~(kP_GS - 1));
    int j0 = 0;
    // every lane below 2^-256: recursion only (a lane's exponent never decreases)
    while (j0 < nst && __ballot(valid && ex == 0) == 0ull) {
      double2 ag[kP_GS];
#pragma unroll
      for (int jj = 0; jj < kP_GS; ++jj) ag[jj] = s_ag[j0 + jj][sg];
#pragma unroll
      for (int jj = 0; jj < kP_GS; ++jj) {
        const double n = fma(fma(ag[jj].x, x, ag[jj].y), mu, -mup);
        mup = mu;
        mu = n;
      }
      rescale_pair(mu, mup, ex);
      j0 += kP_GS;
    }
    if (j0 < nst) {
      rec_group(j0, tile0);
      wave_sync();
      bool odd = false;
      for (; j0 + kP_GS < nst; j0 += kP_GS) {
        mfma_group(j0, odd ? tile1 : tile0);
        rec_group(j0 + kP_GS, odd ? tile0 : tile1);
        wave_sync();
        odd = !odd;
      }
      mfma_group(j0, odd ? tile1 : tile0);
    }
  }
  // epilogue: per column block, the wave's D through LDS to (ring, map) rows; the scales
  // are undone in FP64
  __syncthreads();
  float* ep = reinterpret_cast<float*>(s_buf) + wave * 64 * 16;
  const long long ringM = A.M;
  const int g4 = ii >> 2;  // map within the column block, for the lane's column
  (void)g4;
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) ep[(16 * rb + 4 * kk + r) * 16 + ii] = D[rb][c][r];
    wave_sync();
    if (valid) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int b = mg * MG + 4 * c + g;
        if (b < batch) {
          float2* Fx = F + ((long long)b * 2 + sg) * A.nrings * ringM;       // the field of this sign
          float2* Fy = F + ((long long)b * 2 + (sg ^ 1)) * A.nrings * ringM; // the other field
          const long long rn = (long long)t * ringM + m;
          const long long rs = (long long)(A.nrings - 1 - t) * ringM + m;
          if (lmin <= lmax) {
            const double inv = 1.0 / (sA * (double)s_sc[4 * c + g]);  // a power of two: exact
            const float* e = ep + lane * 16 + 4 * g;
            Fx[rn] = f2(inv * e[0], inv * e[1]);
            if (t < A.T - 1) Fy[rs] = f2(inv * e[2], inv * e[3]);
          } else {
            Fx[rn] = make_float2(0.f, 0.f);
            if (t < A.T - 1) Fy[rs] = make_float2(0.f, 0.f);
          }
        }
      }
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------------- //
// Spin-2 synthesis for 32 maps per recursion with hi / lo planes (the layout of the 32-map
// adjoint k_adj_s2p, transposed).  The ring / sign / parity-fold arrangement is that of
// k_synth_s2p (with 8 waves): wave w runs recursion sign w & 1 for the 64 rings of ring block
// w >> 1 and contracts them against all 128 B columns (32 maps x [X.re, X.im, s Y.re,
// s Y.im]).  What changes is the K index of a matrix instruction: 32 consecutive l steps.
// The recursion values of a 32-step chunk go to the wave's tile as FP16 hi and lo planes
// ([ring][step], 8 steps per 16-byte store), G_l a_lm to B planes ([column][step]) likewise,
// and each lane's A and B operands are single 16-byte reads, so the split products are one
// instruction each: hi.hi, hi.lo and lo.hi over 32 steps (three instructions where
// k_synth_s2p issues four; lo.lo, below 2^-21 of a product, is left out, and B's hi is
// rounded to nearest so the dropped term has no sign bias).  The B planes are
// double-buffered, one barrier per chunk; the matrix instructions of chunk j and the
// recursion of chunk j + 1 (into the tile the wave has just read into registers) form one
// block, interleaved by column block.  Work order decode_tile<kQ_MB, kQ_CB>: each XCD takes
// groups of 8 consecutive orders m x 2 ring chunks, so its 32 resident work-groups are 8
// orders x 4 chunks and each order's a_lm column is streamed by four work-groups at once
// (two with 16 orders per group): fetch 136 -> 109 GB per call at nside 2048, 32 maps.  The
// 8-byte stores F[ring][m] of a group are 64-byte runs of the padded rows.
// ---------------------------------------------------------------------------------- //
constexpr int kQ_MB = 8, kQ_CB = 2;  // work order of k_synth_s2q: orders m x ring chunks
constexpr int kQ_CH = 32;  // l steps per chunk: one matrix K block
constexpr int kQ_PS = 72;  // plane row stride (halves): hi 0..31, lo 32..63, pad; 144 bytes

// the value of lane ^ 1 (a quad permutation [1, 0, 3, 2])
__device__ __forceinline__ float swap_pair(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xf, 0xf, false));
}

// f(integral_constant<int, I>) for I = B .. E - 1, unrolled at compile time
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// alpha_j x + gamma_j of recursion step j (0..31) of a 32-step chunk.  Lane i of every 16-lane
// row holds (alpha, gamma) of steps i (cf[0]) and 16 + i (cf[1]), so each coefficient is a
// row broadcast from lane j % 16 (row_newbcast, a 64-bit DPP source operand): one move and one
// FMA with no LDS read per step (a per-step LDS read of the coefficients, shared by the whole
// wave, queued behind the operand reads and held up the recursion).  The sources come from an
// LDS read, never from a vector instruction just before (the DPP read hazard), and the
// product is the same single-rounding FMA as fma(alpha, x, gamma).
template <int J>
__device__ __forceinline__ double ag_step(const double2 (&cf)[2], double x) {
  double t;
  asm("v_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf" : "=v"(t) : "v"(cf[J >> 4].y), "i"(J & 15));
  asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
      : "+v"(t)
      : "v"(cf[J >> 4].x), "v"(x), "i"(J & 15));
  return t;
}

// (hi, hi) and (lo, lo) halves of two FP32 values: hi by mantissa mask (truncation)
__device__ __forceinline__ void split2_f16(float a, float b, uint32_t& hi, uint32_t& lo) {
  const float ha = __uint_as_float(__float_as_uint(a) & 0xffffe000u);
  const float hb = __uint_as_float(__float_as_uint(b) & 0xffffe000u);
  hi = __builtin_bit_cast(uint32_t, halfx2{(_Float16)ha, (_Float16)hb});
  lo = __builtin_bit_cast(uint32_t, halfx2{(_Float16)(a - ha), (_Float16)(b - hb)});
}
// the same with hi rounded to nearest (lo of either sign)
__device__ __forceinline__ void split2_f16_rn(float a, float b, uint32_t& hi, uint32_t& lo) {
  const halfx2 h = halfx2{(_Float16)a, (_Float16)b};
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, halfx2{(_Float16)(a - (float)h[0]), (_Float16)(b - (float)h[1])});
}

template <int MB, int CB>
__global__ __launch_bounds__(512, 1) void k_synth_s2q(RecArgs A, const double2* __restrict__ alm,
                                                     const float* __restrict__ sB, double sA,
                                                     float2* __restrict__ F, int batch) {
  constexpr int NCB = 8, NW = 8, MG = 32, NCOL = 4 * MG, NT = 64 * NW, RINGS = 32 * NW;
  struct Smem {
    float sc[MG];
    __align__(16) uint16_t bp[2][2][NCOL][kQ_PS];  // [buffer][sign][column][hi | lo]
    __align__(16) uint16_t tp[NW][64][kQ_PS];      // [wave][ring][hi | lo]
  };
  static_assert(sizeof(uint16_t) * NW * 64 * kQ_PS >= sizeof(float) * NW * 64 * 16, "the epilogue fits the tiles");
  __shared__ Smem sm;
  int m, chunk;
  if (!decode_tile<MB, CB>(A, (int)blockIdx.x, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int sg = wave & 1;
  const int ii = lane & 15, kk = lane >> 4;
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  const long long cb0 = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * RINGS + 64 * (wave >> 1) + lane;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid && lmin <= lmax) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m + 4 * sg;
    double sv;
    seed_split(spin2_seed_L(sd, A.lcos[t], A.lsin[t]), sd[3], sv, ex);
    mup = -sv;
  }
  floatx4 D[4][NCB];  // [row block][column block]
#pragma unroll
  for (int rb = 0; rb < 4; ++rb)
#pragma unroll
    for (int c = 0; c < NCB; ++c) D[rb][c] = floatx4{0.f, 0.f, 0.f, 0.f};
  if (tid < MG) sm.sc[tid] = (mg * MG + tid < batch) ? sB[mg * MG + tid] : 0.f;
  // staging: thread (step pair jp = tid % 16, map g = tid / 16) takes steps 2 jp, 2 jp + 1
  const int sjp = tid & 15, sgm = tid >> 4;
  double2 pf_e[2], pf_b[2];
  double pf_g[2];
  bool pf_ok[2];
  // loads at clamped (l, map), with no branch around them (a branch leaves the compiler's
  // wait counting unsure which loads are in flight, and it then waits for all of them); the
  // coefficients past lmax or the batch are zeroed where they are used
  auto prefetch = [&](int l0n) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int l = l0n + 2 * sjp + e, b = mg * MG + sgm;
      pf_ok[e] = l <= lmax && b < batch;
      const long long lc = cb0 + min(l, lmax), bc = min(b, batch - 1);
      pf_g[e] = A.G[lc];
      pf_e[e] = alm[(bc * 2 + 0) * A.K + lc];
      pf_b[e] = alm[(bc * 2 + 1) * A.K + lc];
    }
  };
  // this lane's (alpha, gamma) of steps lane % 16 (cf[0]) and 16 + lane % 16 (cf[1]) of a
  // 32-step chunk, straight from the tables, half a chunk or more before their use (no
  // branch around the loads); prep_cf then zeroes the steps past lmax and the seed step at
  // lmin (whose zero coefficients turn the seed into the first value), and negates gamma for
  // the sign - recursion
  double2 cf[2];
  auto load_cf = [&](int h, int l0c) {
    const long long lc = cb0 + min(l0c + 16 * h + (lane & 15), lmax);
    cf[h] = make_double2(A.alpha[lc], A.gamma[lc]);
  };
  auto prep_cf = [&](int h, int l0c) {
    const int l = l0c + 16 * h + (lane & 15);
    const bool ok = l <= lmax && l != lmin;
    cf[h] = ok ? make_double2(cf[h].x, sg ? -cf[h].y : cf[h].y) : make_double2(0.0, 0.0);
  };
  auto store_b = [&](int l0c, int buf) {
    const double gsc = (double)sm.sc[sgm];
    float v[4][2];  // [bp.re, bp.im, bm.re, bm.im][step]
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const double gs = pf_ok[e] ? pf_g[e] * gsc : 0.0;
      const double er = pf_e[e].x * gs, ei = pf_e[e].y * gs, br = pf_b[e].x * gs, bi = pf_b[e].y * gs;
      // bp = -(E + iB), bm = -(E - iB)
      v[0][e] = (float)(bi - er), v[1][e] = (float)(-ei - br), v[2][e] = (float)(-er - bi), v[3][e] = (float)(br - ei);
    }
    uint32_t h[4], l[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) split2_f16_rn(v[c][0], v[c][1], h[c], l[c]);
    // The folded copies carry s = (-1)^(l+m), opposite for the two steps of a word.  The split
    // rounds to nearest, so it is odd: s x splits into x's halves with their signs flipped,
    // one XOR per word instead of a split.
    const uint32_t sm_ = (((l0c + 2 * sjp + m) & 1) != 0) ? 0x00008000u : 0x80000000u;
    // sign +: [bp.re, bp.im, s bm.re, s bm.im]; sign -: [bm.re, bm.im, s bp.re, s bp.im]
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t fl = (c >= 2) ? sm_ : 0u;
      uint16_t* const rp = &sm.bp[buf][0][4 * sgm + c][2 * sjp];
      uint16_t* const rm = &sm.bp[buf][1][4 * sgm + c][2 * sjp];
      *reinterpret_cast<uint32_t*>(rp) = h[c] ^ fl;
      *reinterpret_cast<uint32_t*>(rp + kQ_CH) = l[c] ^ fl;
      *reinterpret_cast<uint32_t*>(rm) = h[c ^ 2] ^ fl;
      *reinterpret_cast<uint32_t*>(rm + kQ_CH) = l[c ^ 2] ^ fl;
    }
  };
  uint16_t* const trow = &sm.tp[wave][lane][0];  // this lane's ring row
  // 8 recursion steps (group q of the chunk whose coefficients are in cf) into the tile; groups
  // where no lane of the wave is active (every value below 2^-256) store zeros and only advance
  // the recursion
  auto rec8 = [&](auto qc, bool& act) {
    constexpr int q = decltype(qc)::value;
    const bool on = __ballot(valid && ex == 0) != 0ull;
    const bool deep = __ballot(valid && ex < 0) != 0ull;  // exponents never decrease
    act |= on;
    const double f = (ex == 0) ? sA : 0.0;
    uint32_t hw[4], lw[4];
    static_for<0, 4>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      const double n0 = fma(ag_step<8 * q + 2 * p>(cf, x), mu, -mup);
      const double n1 = fma(ag_step<8 * q + 2 * p + 1>(cf, x), n0, -mu);
      mup = n0;
      mu = n1;
      split2_f16((float)(n0 * f), (float)(n1 * f), hw[p], lw[p]);
    });
    *reinterpret_cast<uint4*>(trow + 8 * q) = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    *reinterpret_cast<uint4*>(trow + kQ_CH + 8 * q) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    if (deep) rescale_pair(mu, mup, ex);
  };
  const int nch = (lmin <= lmax) ? (lmax - lmin + kQ_CH) / kQ_CH : 0;
  bool act_cur = false;
  if (nch > 0) {
    // chunk 0's B and (alpha, gamma) and chunk 1's (alpha, gamma); then chunk 1's a_lm in flight
    prefetch(lmin);
    load_cf(0, lmin);
    load_cf(1, lmin);
    __syncthreads();  // sm.sc
    store_b(lmin, 0);
    prefetch(lmin + kQ_CH);
    __syncthreads();
    prep_cf(0, lmin);
    prep_cf(1, lmin);
    static_for<0, 4>([&](auto qc) { rec8(qc, act_cur); });
    load_cf(0, lmin + kQ_CH);  // chunk 1's, recursed in the first loop iteration
    load_cf(1, lmin + kQ_CH);
    wave_sync();
  }
  const uint16_t* const arow = &sm.tp[wave][ii][8 * kk];  // A operands: ring 16 rb + ii, steps 8 kk..
#pragma unroll 1
  for (int jc = 0; jc < nch; ++jc) {
    const int b = jc & 1;
    const int l0 = lmin + kQ_CH * jc;
    const bool more = jc + 1 < nch;
    __syncthreads();  // chunk jc's B visible; buffer b ^ 1 free
    if (more) store_b(l0 + kQ_CH, b ^ 1);
    prefetch(l0 + 2 * kQ_CH);  // chunk jc + 2's a_lm (clamped loads past lmax)
    // this chunk's A operands, read before the recursion of chunk jc + 1 overwrites the tile
    uint4 ah[4], al[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      ah[rb] = *reinterpret_cast<const uint4*>(arow + 16 * rb * kQ_PS);
      al[rb] = *reinterpret_cast<const uint4*>(arow + 16 * rb * kQ_PS + kQ_CH);
    }
    const uint16_t* brow = &sm.bp[b][sg][ii][8 * kk];
    bool act_next = false;
    const bool mma = act_cur;
    // column block c: its three matrix instructions per row block, block c + 1's B read, then
    // 4 recursion steps of chunk jc + 1 (a group of 8 stored after every second block)
    uint4 bq[2][2];  // [block parity][hi / lo]: no register copies between blocks
    bq[0][0] = bq[0][1] = bq[1][0] = bq[1][1] = make_uint4(0u, 0u, 0u, 0u);
    if (mma) {
      bq[0][0] = *reinterpret_cast<const uint4*>(brow);
      bq[0][1] = *reinterpret_cast<const uint4*>(brow + kQ_CH);
    }
    static_for<0, NCB>([&](auto cc) {
      constexpr int c = decltype(cc)::value;
      if (mma) {
        const halfx8 bh = __builtin_bit_cast(halfx8, bq[c & 1][0]);
        const halfx8 bl = __builtin_bit_cast(halfx8, bq[c & 1][1]);
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const halfx8 xh = __builtin_bit_cast(halfx8, ah[rb]);
          const halfx8 xl = __builtin_bit_cast(halfx8, al[rb]);
          D[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bh, D[rb][c], 0, 0, 0);
          D[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bl, D[rb][c], 0, 0, 0);
          D[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, bh, D[rb][c], 0, 0, 0);
        }
      }
      // block c + 1's B, read after this block's matrix instructions were issued: the
      // recursion steps below cover its latency (LDS reads complete in order, so a read
      // issued before them would hold up this block's instructions)
      if (mma && c + 1 < NCB) {
        bq[(c + 1) & 1][0] = *reinterpret_cast<const uint4*>(brow + 16 * (c + 1) * kQ_PS);
        bq[(c + 1) & 1][1] = *reinterpret_cast<const uint4*>(brow + 16 * (c + 1) * kQ_PS + kQ_CH);
      }
      if constexpr ((c & 1) != 0) {
        // chunk jc + 1's coefficients: each half made ready before its first group; reloaded
        // with chunk jc + 2's once its last group has run
        if constexpr (c == 1 || c == 5) prep_cf(c >> 2, l0 + kQ_CH);
        if (more) rec8(std::integral_constant<int, (c >> 1)>{}, act_next);
        if constexpr (c == 3 || c == 7) load_cf(c >> 2, l0 + 2 * kQ_CH);
      }
      __builtin_amdgcn_sched_barrier(0);
    });
    wave_sync();
    act_cur = act_next;
  }
  // epilogue: per column block, the wave's D through LDS to (ring, map) rows; the scales
  // are undone in FP64
  __syncthreads();
  float* ep = reinterpret_cast<float*>(&sm.tp[0][0][0]) + wave * 64 * 16;
  const long long ringM = A.Mf;  // F's row stride
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int r = 0; r < 4; ++r) ep[(16 * rb + 4 * kk + r) * 16 + ii] = D[rb][c][r];
    wave_sync();
    if (valid) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int bm = mg * MG + 4 * c + g;
        if (bm < batch) {
          float2* Fx = F + ((long long)bm * 2 + sg) * A.nrings * ringM;        // the field of this sign
          float2* Fy = F + ((long long)bm * 2 + (sg ^ 1)) * A.nrings * ringM;  // the other field
          const long long rn = (long long)t * ringM + m;
          const long long rs = (long long)(A.nrings - 1 - t) * ringM + m;
          if (lmin <= lmax) {
            const double inv = 1.0 / (sA * (double)sm.sc[4 * c + g]);  // a power of two: exact
            const float* e = ep + lane * 16 + 4 * g;
            Fx[rn] = f2(inv * e[0], inv * e[1]);
            if (t < A.T - 1) Fy[rs] = f2(inv * e[2], inv * e[3]);
          } else {
            Fx[rn] = make_float2(0.f, 0.f);
            if (t < A.T - 1) Fy[rs] = make_float2(0.f, 0.f);
          }
        }
      }
    }
    wave_sync();
  }
}

// Same software pipeline as k_synth_tile_s2: the matrix instructions of group j and the
// recursion of group j + 1 (into the other tile buffer) form one straight-line block.

// G_l a_lm of MG maps for l0..l0+CH-1: thread t takes l = l0 + t % CH and maps t / CH + 4u
// (a wave reads 64 consecutive coefficients of one column), all loads issued at once
template <int MG>
__device__ __forceinline__ void load_alm_s0(const RecArgs& A, const double2* __restrict__ alm, long long cb, int l0,
                                            int lmax, int mg, int batch, int tid, double2 (&a)[MG / 4], double& g) {
  const int j = tid % CH, b0 = tid / CH;
  const int l = l0 + j;
  const bool lok = l <= lmax;
  g = lok ? A.G[cb + l] : 0.0;
#pragma unroll
  for (int u = 0; u < MG / 4; ++u) {
    const int b = mg * MG + b0 + 4 * u;
    a[u] = (lok && b < batch) ? alm[(long long)b * A.K + cb + l] : make_double2(0.0, 0.0);
  }
}

// s_a[j * AS + 2 map + (re, im)]
template <int MG, int AS>
__device__ __forceinline__ void store_alm_s0(const double2 (&a)[MG / 4], double g, int tid, float* s_a) {
  const int j = tid % CH, b0 = tid / CH;
#pragma unroll
  for (int u = 0; u < MG / 4; ++u) {
    float* d = s_a + j * AS + 2 * (b0 + 4 * u);
    d[0] = (float)(a[u].x * g);
    d[1] = (float)(a[u].y * g);
  }
}

// NCB column blocks of 8 maps: 2 (16 maps, 64 accumulator registers) or 4 (32 maps, 128)
template <int NCB>
__device__ __forceinline__ void synth_tile_s0_body(const RecArgs& A, const double2* __restrict__ alm,
                                                   float2* __restrict__ F, int batch, int P, int mg) {
  constexpr int MG = 8 * NCB, AS = 16 * NCB + 8, RINGS = 256;
  static_assert(AS % 64 == 8 || AS % 64 == 40, "B reads at rows 2k+p land on banks 16k+j");
  constexpr int EPS = 17;  // epilogue staging row stride (floats)
  constexpr int LOOP_FLOATS = CH * AS + 4 * 2 * 8 * kTileTS;  // s_a + double-buffered wave tiles
  constexpr int EPI_FLOATS = 4 * 2 * 64 * EPS;                // per wave [parity][ring][column]
  __shared__ double s_al[CH];
  __shared__ __align__(16) float s_buf[LOOP_FLOATS > EPI_FLOATS ? LOOP_FLOATS : EPI_FLOATS];
  float* s_a = s_buf;              // [l][2 map + (re, im)]
  float* tiles = s_buf + CH * AS;  // [wave][buffer][step][lane]
  int m, chunk;
  if (!decode_block<true>(A, P, m, chunk)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ii = lane & 15, kk = lane >> 4;  // MFMA lane decomposition
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * RINGS + tid;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid) {
    x = A.x[t];
    const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
    double sv;
    seed_split(L, (m & 1) ? -1.0 : 1.0, sv, ex);
    mup = -sv;
  }
  floatx4 D[2][4][NCB];  // [parity][row block][column block]
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int c = 0; c < NCB; ++c) D[p][rb][c] = floatx4{0.f, 0.f, 0.f, 0.f};
  float* const tile0 = tiles + wave * 2 * 8 * kTileTS;
  float* const tile1 = tile0 + 8 * kTileTS;
  // the next chunk's coefficients are loaded into registers while this chunk computes
  double2 pf_a[MG / 4];
  double pf_g = 0.0, pf_al = 0.0;
  auto prefetch = [&](int l0n) {
    if (tid < CH) {
      const int l = l0n + tid;
      pf_al = (l <= lmax && l != m) ? A.alpha[cb + l] : 0.0;
    }
    load_alm_s0<MG>(A, alm, cb, l0n, lmax, mg, batch, tid, pf_a, pf_g);
  };
  prefetch(m);
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) s_al[tid] = pf_al;
    store_alm_s0<MG, AS>(pf_a, pf_g, tid, s_a);
    __syncthreads();
    if (l0 + CH <= lmax) prefetch(l0 + CH);
    const int nst = min(CH, (lmax - l0 + 1 + 7) & ~7);
    int j0 = 0;
    // every lane below 2^-256: recursion only (a lane's exponent never decreases)
    while (j0 < nst && __ballot(valid && ex == 0) == 0ull) {
      double al[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) al[jj] = s_al[j0 + jj];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const double n = fma(al[jj] * x, mu, -mup);
        mup = mu;
        mu = n;
      }
      rescale_pair(mu, mup, ex);
      j0 += 8;
    }
    auto rec_group = [&](int jg, float* tl) {
      const double f = (ex == 0) ? 1.0 : 0.0;
      double al[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) al[jj] = s_al[jg + jj];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const double n = fma(al[jj] * x, mu, -mup);
        mup = mu;
        mu = n;
        // the 1/0 mask in FP64 first: a lane below 2^-256 may hold a mantissa above FP32 range
        tl[jj * kTileTS + lane] = (float)(n * f);
      }
      rescale_pair(mu, mup, ex);
    };
    auto mfma_group = [&](int jg, const float* tl) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        float bv[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) bv[c] = s_a[(jg + 2 * kk + p) * AS + 16 * c + ii];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const float av = tl[(2 * kk + p) * kTileTS + 16 * rb + ii];
#pragma unroll
          for (int c = 0; c < NCB; ++c)
            D[p][rb][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[c], D[p][rb][c], 0, 0, 0);
        }
      }
    };
    if (j0 < nst) {
      rec_group(j0, tile0);
      wave_sync();
      bool odd = false;  // which buffer holds group j0
      for (; j0 + 8 < nst; j0 += 8) {
        mfma_group(j0, odd ? tile1 : tile0);
        rec_group(j0 + 8, odd ? tile0 : tile1);
        wave_sync();
        odd = !odd;
      }
      mfma_group(j0, odd ? tile1 : tile0);
      wave_sync();
    }
  }
  // epilogue, one column block (8 maps) at a time: D[i][j] sits in lane j + 16 (i / 4),
  // register i % 4; row i of block rb is ring 16 rb + i of the wave, column j = 2 map + re/im
  __syncthreads();  // the staging below aliases s_a and the tiles
  float* ep = s_buf + wave * 2 * 64 * EPS;  // [parity][ring][column]
  const long long ringM = A.M;
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) ep[(p * 64 + 16 * rb + 4 * kk + r) * EPS + ii] = D[p][rb][c][r];
    wave_sync();
    if (valid) {
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        const int b = mg * MG + 8 * c + h;
        if (b >= batch) continue;
        const double c0r = ep[lane * EPS + 2 * h], c0i = ep[lane * EPS + 2 * h + 1];
        const double c1r = ep[(64 + lane) * EPS + 2 * h], c1i = ep[(64 + lane) * EPS + 2 * h + 1];
        float2* Fb = F + (long long)b * A.nrings * ringM;
        Fb[(long long)t * ringM + m] = f2(c0r + c1r, c0i + c1i);
        if (t < A.T - 1) Fb[(long long)(A.nrings - 1 - t) * ringM + m] = f2(c0r - c1r, c0i - c1i);
      }
    }
    wave_sync();
  }
}

template <int NCB, int MINB>
__global__ __launch_bounds__(256, MINB) void k_synth_tile_s0(RecArgs A, const double2* __restrict__ alm,
                                                             float2* __restrict__ F, int batch) {
  synth_tile_s0_body<NCB>(A, alm, F, batch, blockIdx.x, blockIdx.y);
}

// k_synth_tile_s0<NCB, MINB> on the FP16 matrix cores, with the operand split, scales and
// instruction layout of k_synth_tile_s2h: 16-step groups, A halves (hi, lo, hi, lo) from the
// wave tile, B halves (hi, hi, lo, lo) staged in LDS, four products per l step.
template <int NCB, int MINB>
__global__ __launch_bounds__(256, MINB) void k_synth_tile_s0h(RecArgs A, const double2* __restrict__ alm,
                                                              const float* __restrict__ sB, double sA,
                                                              float2* __restrict__ F, int batch) {
  constexpr int MG = 8 * NCB, RINGS = 256;
  constexpr int NCOL = 2 * MG;
  constexpr int ASH = 2 * NCOL + 8;  // s_b row stride (dwords): 64-bit B reads, 4 rows apart per lane group
  constexpr int TSH = 68;            // tile row stride (dwords): A reads 4 rows apart per lane group
  constexpr int GS = 16;
  constexpr int EPS = 17;
  constexpr int LOOP_DW = CH * ASH + 4 * 2 * GS * TSH;
  constexpr int EPI_FLOATS = 4 * 2 * 64 * EPS;
  static_assert(CH % GS == 0, "groups tile the chunk");
  __shared__ double s_al[CH];
  __shared__ float s_sc[MG];
  __shared__ __align__(16) uint32_t s_buf[LOOP_DW > EPI_FLOATS ? LOOP_DW : EPI_FLOATS];
  uint32_t* s_b = s_buf;               // [l][2 column + {hi hi, lo lo}]
  uint32_t* tiles = s_buf + CH * ASH;  // [wave][buffer][step][lane]: (hi, lo)
  int m, chunk;
  if (!decode_tile<kA_MB, kA_CB>(A, (int)blockIdx.x, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ii = lane & 15, kk = lane >> 4;
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const int t = chunk * RINGS + tid;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid) {
    x = A.x[t];
    const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
    double sv;
    seed_split(L, (m & 1) ? -1.0 : 1.0, sv, ex);
    mup = -sv;
  }
  if (tid < MG) s_sc[tid] = (mg * MG + tid < batch) ? sB[mg * MG + tid] : 0.f;
  floatx4 D[2][4][NCB];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
      for (int c = 0; c < NCB; ++c) D[p][rb][c] = floatx4{0.f, 0.f, 0.f, 0.f};
  uint32_t* const tile0 = tiles + wave * 2 * GS * TSH;
  uint32_t* const tile1 = tile0 + GS * TSH;
  double2 pf_a[MG / 4];
  double pf_g = 0.0, pf_al = 0.0;
  auto prefetch = [&](int l0n) {
    if (tid < CH) {
      const int l = l0n + tid;
      pf_al = (l <= lmax && l != m) ? A.alpha[cb + l] : 0.0;
    }
    load_alm_s0<MG>(A, alm, cb, l0n, lmax, mg, batch, tid, pf_a, pf_g);
  };
  // s_b[j][2 (2 map + re/im) + {hi hi, lo lo}]
  auto store_b = [&]() {
    const int j = tid % CH, b0 = tid / CH;
#pragma unroll
    for (int u = 0; u < MG / 4; ++u) {
      const int map = b0 + 4 * u;
      const double gs = pf_g * (double)s_sc[map];
      const uint32_t sr = split_f16((float)(pf_a[u].x * gs)), si = split_f16((float)(pf_a[u].y * gs));
      *reinterpret_cast<uint4*>(s_b + j * ASH + 4 * map) = make_uint4(dup_lo(sr), dup_hi(sr), dup_lo(si), dup_hi(si));
    }
  };
  prefetch(m);
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) s_al[tid] = pf_al;
    store_b();
    __syncthreads();
    if (l0 + CH <= lmax) prefetch(l0 + CH);
    const int nst = min(CH, (lmax - l0 + 1 + GS - 1) & ~(GS - 1));
    int j0 = 0;
    while (j0 < nst && __ballot(valid && ex == 0) == 0ull) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double al[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) al[jj] = s_al[j0 + 8 * h + jj];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double n = fma(al[jj] * x, mu, -mup);
          mup = mu;
          mu = n;
        }
        rescale_pair(mu, mup, ex);
      }
      j0 += GS;
    }
    auto rec_group = [&](int jg, uint32_t* tl) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double f = (ex == 0) ? sA : 0.0;
        double al[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) al[jj] = s_al[jg + 8 * h + jj];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double n = fma(al[jj] * x, mu, -mup);
          mup = mu;
          mu = n;
          // the 1/0 mask in FP64 first: a lane below 2^-256 may hold a mantissa above range
          tl[(8 * h + jj) * TSH + lane] = split_f16((float)(n * f));
        }
        rescale_pair(mu, mup, ex);
      }
    };
    auto mfma_group = [&](int jg, const uint32_t* tl) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        halfx8 bv[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
          const uint2 b0 = *reinterpret_cast<const uint2*>(s_b + (jg + 4 * kk + p) * ASH + 2 * (16 * c + ii));
          const uint2 b1 = *reinterpret_cast<const uint2*>(s_b + (jg + 4 * kk + 2 + p) * ASH + 2 * (16 * c + ii));
          bv[c] = __builtin_bit_cast(halfx8, make_uint4(b0.x, b0.y, b1.x, b1.y));
        }
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const uint32_t a0 = tl[(4 * kk + p) * TSH + 16 * rb + ii];
          const uint32_t a1 = tl[(4 * kk + 2 + p) * TSH + 16 * rb + ii];
          const halfx8 av = __builtin_bit_cast(halfx8, make_uint4(a0, a0, a1, a1));
#pragma unroll
          for (int c = 0; c < NCB; ++c)
            D[p][rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv[c], D[p][rb][c], 0, 0, 0);
        }
      }
    };
    if (j0 < nst) {
      rec_group(j0, tile0);
      wave_sync();
      bool odd = false;
      for (; j0 + GS < nst; j0 += GS) {
        mfma_group(j0, odd ? tile1 : tile0);
        rec_group(j0 + GS, odd ? tile0 : tile1);
        wave_sync();
        odd = !odd;
      }
      mfma_group(j0, odd ? tile1 : tile0);
      wave_sync();
    }
  }
  // epilogue (as k_synth_tile_s0), with the scales undone in FP64
  __syncthreads();
  float* ep = reinterpret_cast<float*>(s_buf) + wave * 2 * 64 * EPS;
  const long long ringM = A.M;
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) ep[(p * 64 + 16 * rb + 4 * kk + r) * EPS + ii] = D[p][rb][c][r];
    wave_sync();
    if (valid) {
#pragma unroll
      for (int h = 0; h < 8; ++h) {
        const int b = mg * MG + 8 * c + h;
        if (b >= batch) continue;
        const double inv = 1.0 / (sA * (double)s_sc[8 * c + h]);  // a power of two: exact
        const double c0r = inv * ep[lane * EPS + 2 * h], c0i = inv * ep[lane * EPS + 2 * h + 1];
        const double c1r = inv * ep[(64 + lane) * EPS + 2 * h], c1i = inv * ep[(64 + lane) * EPS + 2 * h + 1];
        float2* Fb = F + (long long)b * A.nrings * ringM;
        Fb[(long long)t * ringM + m] = f2(c0r + c1r, c0i + c1i);
        if (t < A.T - 1) Fb[(long long)(A.nrings - 1 - t) * ringM + m] = f2(c0r - c1r, c0i - c1i);
      }
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------------- //
// adjoint: per-wave FP32 tile of 32 l x 64 rings (per recursion sign), 16x16x4 MFMA
// ---------------------------------------------------------------------------------- //
// 16x16x4 f32 layouts: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k,
// D[i][j] at lane j + 16 (i / 4), register i % 4.
constexpr int TL = 32;  // l values per tile (16 rows of each parity)
// Tile row stride in floats: A-operand reads take rows 2i+p (i = lane%16) and ring column
// 4q + lane/16; with a stride of 66 the bank (4i + lane/16) mod 64 is distinct per lane.
constexpr int TS = 66;

// Sum the per-wave D tiles (rows = l, cols = 16) over the work-group in FP64 and store the
// chunk partials.  red aliases the wave tiles.  Caller provides the column -> output map.
template <int NSIGN>
__device__ __forceinline__ void adj_tile_mfma(const float* tile, int lane, const float (&Bv)[NSIGN][2][16],
                                              floatx4 (&D)[2]) {
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    D[p] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int sg = 0; sg < NSIGN; ++sg) {
      const float* ts = tile + sg * TL * TS;
      const int row = 2 * (lane & 15) + p;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const float a = ts[row * TS + 4 * q + (lane >> 4)];
        D[p] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Bv[sg][p][q], D[p], 0, 0, 0);
      }
    }
  }
}

// spin 2: column block cb2, column j = 4 g + c, c in {p2.re, p2.im, m2.re, m2.im}, map
// 4 cb2 + g of the group; MG = 4 NCB maps share one recursion.
//
// One B operand per (block, sign, ring) serves both parities: in every lane one of the
// north / south terms of B is zero (sign +: the south term for c < 2, the north term for
// c >= 2; sign -: the reverse), so B_odd = sigma B_even with sigma = +-1 fixed by the column
// and opposite for the two signs.  The odd rows of the sign - tile are stored negated, which
// makes the odd-parity accumulator equal sigma_+(column) times the true one; the sign is
// applied when the tile result is reduced.
//
// Tile layout [sign][row][TS2]: row 16 p + i holds step 2 i + p of the tile (parity p), and
// ring r of the wave sits in column 16 (r % 4) + r / 4.  The A operand of lane (i, k) for
// the 16 MFMAs of one (parity, sign) is then rings 4 q + k, q = 0..15: 16 contiguous floats,
// four 16-byte reads.  With TS2 = 68 the rows of the 16 lanes of a read group start 4 banks
// apart (conflict-free), and a recursion step writes one row of 64 consecutive floats.
constexpr int TS2 = 68;
__device__ __forceinline__ int s2_tile_row(int j) { return 16 * (j & 1) + (j >> 1); }
__device__ __forceinline__ int s2_tile_col(int r) { return 16 * (r & 3) + (r >> 2); }

// The ring values a column-block adjoint lane holds for the whole l loop: component ri of
// field `field` of map bcol at chunk ring 4 (q0 + q) + lane / 16 (n[q]) and at its mirror
// (s[q]), q < NQ, as raw bits.  All 2 NQ loads are issued before any is used: the addresses
// are clamped to valid rings and maps and the caller applies each value's validity (a load
// behind a branch waits for the one before it: 128 memory latencies per work-group).
template <typename W, int NQ>
__device__ __forceinline__ void load_ring_values(const RecArgs& A, const float2* __restrict__ V, int bcol,
                                                 int batch, int ncomp, int field, int ri, int m, int chunk,
                                                 int lane, int q0, W (&n)[NQ], W (&s)[NQ]) {
  const long long ringM = A.M;
  const float* Vf = reinterpret_cast<const float*>(
      V + ((long long)min(bcol, batch - 1) * ncomp + field) * A.nrings * ringM);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int tr = min(chunk * 256 + 4 * (q0 + q) + (lane >> 4), A.T - 1);
    n[q] = __builtin_bit_cast(W, Vf[2 * ((long long)tr * ringM + m) + ri]);
    s[q] = __builtin_bit_cast(W, Vf[2 * ((long long)(A.nrings - 1 - tr) * ringM + m) + ri]);
  }
}

// validity of the ring pair of q: the ring exists (ok), and it has a mirror (south: not the
// equator)
__device__ __forceinline__ void ring_valid(const RecArgs& A, int bcol, int batch, int chunk, int q, int lane,
                                           bool& ok, bool& south) {
  const int tr = chunk * 256 + 4 * q + (lane >> 4);
  ok = bcol < batch && tr < A.T;
  south = ok && tr < A.T - 1;
}

// act: wave-uniform bit sg set when the sign-sg tile holds a nonzero value (an all-zero
// sign tile contributes nothing and is skipped).  The operands of the next (parity, sign)
// step are read while the MFMAs of the current one run.
template <int NCB>
__device__ __forceinline__ void adj_tile_mfma_s2(const float* tile, int lane, const float (&Bv)[NCB][2][16],
                                                 floatx4 (&D)[NCB][2], int act) {
  const float* base = tile + (lane & 15) * TS2 + 16 * (lane >> 4);
  floatx4 a[2][4];
  // step s = 2 p + sg
  auto load = [&](int s, floatx4(&dst)[4]) {
    const floatx4* src = reinterpret_cast<const floatx4*>(base + ((s & 1) * TL + 16 * (s >> 1)) * TS2);
#pragma unroll
    for (int n = 0; n < 4; ++n) dst[n] = src[n];
  };
  load(0, a[0]);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s + 1 < 4) load(s + 1, a[(s + 1) & 1]);
    const int p = s >> 1, sg = s & 1;
    if (sg == 0) {
#pragma unroll
      for (int c = 0; c < NCB; ++c) D[c][p] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
    if ((act >> sg) & 1) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const float av = a[s & 1][q >> 2][q & 3];
#pragma unroll
        for (int c = 0; c < NCB; ++c)
          D[c][p] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, Bv[c][sg][q], D[c][p], 0, 0, 0);
      }
    }
  }
}

// NW waves per work-group cover 64 NW rings (one partial per chunk of rings); MINB: the
// register budget, in work-groups of this size per CU (NW * MINB = 8 caps it at 256 VGPRs)
template <int NCB, int MINB, int NW>
__global__ __launch_bounds__(64 * NW, MINB) void k_adj_mx_s2(RecArgs A, const float2* __restrict__ V,
                                                   float2* __restrict__ part, int batch) {
  constexpr int MG = 4 * NCB;
  constexpr int RINGS = 64 * NW;
  __shared__ double2 s_ag[CH];
  __shared__ __align__(16) float tiles[NW][2][TL][TS2];  // [wave][sign][row][col]; reused for the wave sums
  __shared__ int s_act[NW];
  // the tile results of a wave, [2 i + p][16 NCB cols + 2 pad], overwrite the wave's own tile
  constexpr int NC = 16 * NCB + 2;
  static_assert(2 * 16 * NC <= 2 * TL * TS2, "the tile results fit in the wave's tile");
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  if (lmin > lmax) return;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const int t = chunk * RINGS + tid;
  const bool valid = t < A.T;
  double x = 0.0, pp = 0.0, ppp = 0.0, pm = 0.0, pmp = 0.0;
  int ep = 0, em = 0;
  if (valid) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m;
    const double lc = A.lcos[t], ls = A.lsin[t];
    double s;
    seed_split(spin2_seed_L(sd, lc, ls), sd[3], s, ep);
    ppp = -s;
    seed_split(spin2_seed_L(sd + 4, lc, ls), sd[7], s, em);
    pmp = -s;
  }
  // B operands: ring 4q + lane/16 of this wave, column lane%16 -> map 4 c + g, component cc
  float Bv[NCB][2][16];  // [block][sign][ring], even parity (the odd one is +-B, see above)
  {
    // mu+: p2 <- Vpn, m2 <- s Vms ; mu-: p2 <- s Vps, m2 <- Vmn.  Component cc of the
    // lane's column: sign + reads Vp north (cc < 2) or Vm south (cc >= 2), sign - reads Vm
    // north (cc >= 2) or Vp south (cc < 2): one double each.
    const double par0 = ((lmin - m) & 1) ? -1.0 : 1.0;  // (-1)^(l-m) at even (l - lmin)
    const int g = (lane & 15) >> 2, cc = lane & 3;
    const bool lo = cc < 2;
    const int ri = cc & 1;  // re / im
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
      const int b = mg * MG + 4 * c + g;
      const float* Vp = reinterpret_cast<const float*>(V + ((long long)b * 2 + 0) * A.nrings * ringM);
      const float* Vm = reinterpret_cast<const float*>(V + ((long long)b * 2 + 1) * A.nrings * ringM);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int tr = chunk * RINGS + wave * 64 + 4 * q + (lane >> 4);
        double bp = 0.0, bm = 0.0;
        if (tr < A.T && b < batch) {
          const long long rn = 2 * ((long long)tr * ringM + m) + ri;
          const bool south = tr < A.T - 1;  // the equator ring has no mirror
          const long long rs = 2 * ((long long)(A.nrings - 1 - tr) * ringM + m) + ri;
          if (lo) {
            bp = (double)Vp[rn];
            bm = south ? par0 * (double)Vp[rs] : 0.0;
          } else {
            bp = south ? par0 * (double)Vm[rs] : 0.0;
            bm = (double)Vm[rn];
          }
        }
        Bv[c][0][q] = (float)bp;
        Bv[c][1][q] = (float)bm;
      }
    }
  }
  float2* p_p2 = part + (((long long)chunk * batch) * 2) * A.K;  // offset per map below
  float* mytile = &tiles[wave][0][0][0];
  const int tcol = s2_tile_col(lane);
  // the next chunk's (alpha, gamma) are loaded into registers while this chunk computes
  double2 pf_ag = make_double2(0.0, 0.0);
  auto prefetch = [&](int l0n) {
    const int l = l0n + tid;
    pf_ag = make_double2(0.0, 0.0);
    if (tid < CH && l <= lmax && l != lmin) pf_ag = make_double2(A.alpha[cb + l], A.gamma[cb + l]);
  };
  prefetch(lmin);
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) s_ag[tid] = pf_ag;
    __syncthreads();
    if (l0 + CH <= lmax) prefetch(l0 + CH);
#pragma unroll 1
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += TL) {
      // recursion for TL steps, values to the wave's tile; act collects which sign tiles
      // hold a value above 2^-256 (wave-uniform)
      int act = 0;
#pragma unroll 1
      for (int j0 = 0; j0 < TL; j0 += 8) {
        const bool deep = valid && ((ep < 0) || (em < 0));
        const int anydeep = __ballot(deep) != 0ull;
        act |= (__ballot(valid && ep == 0) != 0ull ? 1 : 0) | (__ballot(valid && em == 0) != 0ull ? 2 : 0);
        const double fp = (ep == 0 && valid) ? 1.0 : 0.0, fm = (em == 0 && valid) ? 1.0 : 0.0;
        double2 ag[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) ag[jj] = s_ag[jb + j0 + jj];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double cp = fma(ag[jj].x, x, ag[jj].y);
          const double cm = fma(ag[jj].x, x, -ag[jj].y);
          const double np = fma(cp, pp, -ppp);
          const double nm = fma(cm, pm, -pmp);
          ppp = pp;
          pp = np;
          pmp = pm;
          pm = nm;
          mytile[(0 * TL + s2_tile_row(j0 + jj)) * TS2 + tcol] = (float)(np * fp);
          // odd rows (odd l - lmin) of the sign - tile negated: one B serves both parities
          mytile[(1 * TL + s2_tile_row(j0 + jj)) * TS2 + tcol] = (float)(nm * ((jj & 1) ? -fm : fm));
        }
        if (anydeep) {
          rescale_pair(pp, ppp, ep);
          rescale_pair(pm, pmp, em);
        }
      }
      wave_sync();
      floatx4 D[NCB][2];
      adj_tile_mfma_s2<NCB>(mytile, lane, Bv, D, act);
      // the wave's tile results replace its tile (its own reads of the tile are complete);
      // rows 2 i + p with stride NC = 16 NCB + 2 make the stores below and the 8-byte reads
      // of the reduction conflict-free
      if (lane == 0) s_act[wave] = act;
#pragma unroll
      for (int c = 0; c < NCB; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 4 * (lane >> 4) + r, jc = lane & 15;
            // odd parity: sigma_+ = -1 for the m2 columns (component lane%4 >= 2)
            const float sig = (p == 1 && (lane & 2)) ? -1.0f : 1.0f;
            mytile[(2 * i + p) * NC + 16 * c + jc] = sig * D[c][p][r];
          }
      __syncthreads();
      int any = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) any |= s_act[w];
      // thread -> (p, i, g, h) with l = l0 + jb + 2 i + p fastest: sums the waves in FP64 for
      // components 2h, 2h + 1 (p2 for h = 0, m2 for h = 1) of map g; a work-group whose
      // tiles are all zero stores zeros
      for (int r = tid; r < 2 * 16 * MG * 2; r += RINGS) {
        const int p = r & 1, i = (r >> 1) & 15, g = (r >> 5) % MG, h = r / (32 * MG);
        const int l = l0 + jb + 2 * i + p;
        const int b = mg * MG + g;
        if (l <= lmax && b < batch) {
          double s0 = 0.0, s1 = 0.0;
          if (any) {
            for (int w = 0; w < NW; ++w) {
              const float2 v = *reinterpret_cast<const float2*>(&tiles[w][0][0][0] + (2 * i + p) * NC + 4 * g + 2 * h);
              s0 += (double)v.x;
              s1 += (double)v.y;
            }
          }
          p_p2[(long long)b * 2 * A.K + (long long)h * A.K + cb + l] = make_float2((float)s0, (float)s1);
        }
      }
      __syncthreads();
    }
  }
}

// Column-block form of the 16-map spin-2 adjoint: 256 rings, 4 waves, one chunk partial per
// 256 rings.  Wave w runs the recursion of rings 64 w .. 64 w + 63 of the chunk into its
// tile, as in k_adj_mx_s2; after a barrier it contracts column block w (maps 4 w .. 4 w + 3
// of the group) over the 256 rings, reading all four tiles.  Every wave then does the same
// matrix work and skips the same all-zero (ring block, sign) tiles, and its accumulators are
// the chunk partials of its columns.
__device__ __forceinline__ void adj_mx_s2_cb_body(const RecArgs& A, const float2* __restrict__ V,
                                                  float2* __restrict__ part, int batch, int P, int mg) {
  constexpr int NW = 4, MG = 16, RINGS = 256;
  __shared__ double2 s_ag[CH];
  __shared__ __align__(16) float tiles[NW][2][TL][TS2];  // [ring block][sign][row][col]
  __shared__ int s_act[NW];
  int m, chunk;
  if (!decode_block<true>(A, P, m, chunk)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  if (lmin > lmax) return;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const int t = chunk * RINGS + tid;
  const bool valid = t < A.T;
  double x = 0.0, pp = 0.0, ppp = 0.0, pm = 0.0, pmp = 0.0;
  int ep = 0, em = 0;
  if (valid) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m;
    const double lc = A.lcos[t], ls = A.lsin[t];
    double s;
    seed_split(spin2_seed_L(sd, lc, ls), sd[3], s, ep);
    ppp = -s;
    seed_split(spin2_seed_L(sd + 4, lc, ls), sd[7], s, em);
    pmp = -s;
  }
  // B operands of column block cw: chunk ring 4 q + lane/16 (q = 0..63),
  // column lane%16 -> map 4 cw + (lane%16)/4, component lane%4 (see k_adj_mx_s2)
  const int cw = wave;
  const int g = (lane & 15) >> 2, cc = lane & 3;
  const int bcol = mg * MG + 4 * cw + g;  // the map of this lane's column
  float Bv[2][64];                          // [sign][q]
  {
    const float par0 = ((lmin - m) & 1) ? -1.0f : 1.0f;
    const bool lo = cc < 2;
    // sign + column (p): V+ north and par0 V+ south; sign - column (n): the reverse
    load_ring_values(A, V, bcol, batch, 2, lo ? 0 : 1, cc & 1, m, chunk, lane, 0, Bv[0], Bv[1]);
#pragma unroll
    for (int q = 0; q < 64; ++q) {
      bool ok, south;
      ring_valid(A, bcol, batch, chunk, q, lane, ok, south);
      const float vn = ok ? Bv[0][q] : 0.f;
      const float vs = south ? par0 * Bv[1][q] : 0.f;
      Bv[0][q] = lo ? vn : vs;
      Bv[1][q] = lo ? vs : vn;
      // materialised here: otherwise the compiler recomputes them from the raw values in
      // the l loop and holds more registers there
      asm volatile("" : "+v"(Bv[0][q]), "+v"(Bv[1][q]));
    }
  }
  float* pcol = reinterpret_cast<float*>(part + (((long long)chunk * batch + bcol) * 2 + (cc >> 1)) * A.K) + (cc & 1);
  // odd parity: sigma_+ = -1 for the m2 columns
  const float sig_odd = (cc >= 2) ? -1.0f : 1.0f;
  float* mytile = &tiles[wave][0][0][0];
  const int tcol = s2_tile_col(lane);
  const float* abase = &tiles[0][0][0][0] + (lane & 15) * TS2 + 16 * (lane >> 4);
  double2 pf_ag = make_double2(0.0, 0.0);
  auto prefetch = [&](int l0n) {
    const int l = l0n + tid;
    pf_ag = make_double2(0.0, 0.0);
    if (tid < CH && l <= lmax && l != lmin) pf_ag = make_double2(A.alpha[cb + l], A.gamma[cb + l]);
  };
  prefetch(lmin);
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) s_ag[tid] = pf_ag;
    __syncthreads();
    if (l0 + CH <= lmax) prefetch(l0 + CH);
#pragma unroll 1
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += TL) {
      int act = 0;
#pragma unroll 1
      for (int j0 = 0; j0 < TL; j0 += 8) {
        const bool deep = valid && ((ep < 0) || (em < 0));
        const int anydeep = __ballot(deep) != 0ull;
        act |= (__ballot(valid && ep == 0) != 0ull ? 1 : 0) | (__ballot(valid && em == 0) != 0ull ? 2 : 0);
        const double fp = (ep == 0 && valid) ? 1.0 : 0.0, fm = (em == 0 && valid) ? 1.0 : 0.0;
        double2 ag[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) ag[jj] = s_ag[jb + j0 + jj];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double cp = fma(ag[jj].x, x, ag[jj].y);
          const double cm = fma(ag[jj].x, x, -ag[jj].y);
          const double np = fma(cp, pp, -ppp);
          const double nm = fma(cm, pm, -pmp);
          ppp = pp;
          pp = np;
          pmp = pm;
          pm = nm;
          mytile[(0 * TL + s2_tile_row(j0 + jj)) * TS2 + tcol] = (float)(np * fp);
          // odd rows (odd l - lmin) of the sign - tile negated: one B serves both parities
          mytile[(1 * TL + s2_tile_row(j0 + jj)) * TS2 + tcol] = (float)(nm * ((jj & 1) ? -fm : fm));
        }
        if (anydeep) {
          rescale_pair(pp, ppp, ep);
          rescale_pair(pm, pmp, em);
        }
      }
      if (lane == 0) s_act[wave] = act;
      __syncthreads();
      int actv[4];  // this half's ring blocks
#pragma unroll
      for (int w = 0; w < 4; ++w) actv[w] = s_act[w];
      // D[p][rb]: parity p, ring block rb.  Each ring block accumulates in FP32 over its 64
      // rings (both signs) and the four are summed in FP64, in ring-block order: the
      // arithmetic of the per-wave tiles and FP64 cross-wave sum of k_adj_mx_s2 (a single
      // FP32 accumulation over the 256 rings loses up to 8x accuracy in the high-l bands).
      floatx4 D[2][4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) D[p][rb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
          if (!((actv[rb] >> sg) & 1)) continue;
          floatx4 a[2][4];
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const floatx4* src =
                reinterpret_cast<const floatx4*>(abase + ((rb * 2 + sg) * TL + 16 * p) * TS2);
#pragma unroll
            for (int n = 0; n < 4; ++n) a[p][n] = src[n];
          }
#pragma unroll
          for (int qq = 0; qq < 16; ++qq)
#pragma unroll
            for (int p = 0; p < 2; ++p)
              D[p][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[p][qq >> 2][qq & 3], Bv[sg][16 * rb + qq],
                                                              D[p][rb], 0, 0, 0);
        }
      // D[i][j] at lane j + 16 (i / 4), register i % 4: row i -> l = l0 + jb + 2 i + p
      double v[2][4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          v[p][r] = 0.0;
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) v[p][r] += (double)D[p][rb][r];
        }
      if (bcol < batch) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int l = l0 + jb + 2 * (4 * (lane >> 4) + r) + p;
            const double t = v[p][r];
            if (l <= lmax) pcol[2 * (cb + l)] = (p == 1) ? sig_odd * (float)t : (float)t;
          }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256, 2) void k_adj_mx_s2_cb(RecArgs A, const float2* __restrict__ V,
                                                        float2* __restrict__ part, int batch) {
  adj_mx_s2_cb_body(A, V, part, batch, blockIdx.x, blockIdx.y);
}

// k_adj_mx_s2_cb<1> on the FP16 matrix cores.  The tiles hold (hi, lo) FP16 pairs of the
// recursion values (scale sA) instead of floats, the B operands (hi, lo) pairs of the lane's
// V values; v_mfma_f32_16x16x32_f16 contracts 8 rings per instruction (A halves hi, lo, hi, lo
// of a ring, B halves hi, hi, lo, lo: the four products), half the instructions of the FP32
// form.  B's scale is local: each work-group scales every map's V values of its rings by a
// power of two that puts their maximum at or below 2^12, and undoes it (with sA) before the
// FP64 sums are stored, so the chunk partials are those of the FP32 kernel up to rounding.
__global__ __launch_bounds__(256, 2) void k_adj_mx_s2_cbh(RecArgs A, const float2* __restrict__ V, double sA,
                                                         float2* __restrict__ part, int batch) {
  constexpr int NW = 4, MG = 16, RINGS = 256;
  __shared__ double2 s_ag[CH];
  __shared__ __align__(16) uint32_t tiles[NW][2][TL][TS2];  // [ring block][sign][row][col]: (hi, lo)
  __shared__ int s_act[NW];
  int m, chunk;
  if (!decode_tile<kA_MB, kA_CB>(A, (int)blockIdx.x, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  if (lmin > lmax) return;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const int t = chunk * RINGS + tid;
  const bool valid = t < A.T;
  double x = 0.0, pp = 0.0, ppp = 0.0, pm = 0.0, pmp = 0.0;
  int ep = 0, em = 0;
  if (valid) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m;
    const double lc = A.lcos[t], ls = A.lsin[t];
    double s;
    seed_split(spin2_seed_L(sd, lc, ls), sd[3], s, ep);
    ppp = -s;
    seed_split(spin2_seed_L(sd + 4, lc, ls), sd[7], s, em);
    pmp = -s;
  }
  // B operands of column block cw: chunk ring 4 q + lane/16 (q = 0..63), column lane%16 ->
  // map 4 cw + (lane%16)/4, component lane%4 (as k_adj_mx_s2_cb)
  const int cw = wave;
  const int g = (lane & 15) >> 2, cc = lane & 3;
  const int bcol = mg * MG + 4 * cw + g;
  uint32_t Bv[2][64];  // [sign][q]: (hi, lo) of the scaled value
  double inv;          // 1 / (sA sB) of this lane's map
  {
    const float par0 = ((lmin - m) & 1) ? -1.0f : 1.0f;
    const bool lo = cc < 2;
    // the values go into Bv as float bits (sign + column: V+ north and par0 V+ south; sign -:
    // the reverse), then are split in place once the maximum is known
    load_ring_values(A, V, bcol, batch, 2, lo ? 0 : 1, cc & 1, m, chunk, lane, 0, Bv[0], Bv[1]);
    float mx = 0.f;
#pragma unroll
    for (int q = 0; q < 64; ++q) {
      bool ok, south;
      ring_valid(A, bcol, batch, chunk, q, lane, ok, south);
      const float vn = ok ? __uint_as_float(Bv[0][q]) : 0.f;
      const float vs = south ? par0 * __uint_as_float(Bv[1][q]) : 0.f;
      const float bp = lo ? vn : vs, bm = lo ? vs : vn;
      mx = fmaxf(mx, fmaxf(fabsf(bp), fabsf(bm)));
      Bv[0][q] = __float_as_uint(bp);
      Bv[1][q] = __float_as_uint(bm);
    }
    // the map's maximum over its 16 lanes (4 components x 4 ring residues): lanes that
    // differ in bits 0, 1, 4 and 5
    mx = fmaxf(mx, __shfl_xor(mx, 1));
    mx = fmaxf(mx, __shfl_xor(mx, 2));
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    int e = 1;
    if (mx > 0.f && mx < 3.0e38f) (void)frexpf(4096.f / mx, &e);
    e = min(max(e - 1, -120), 120);
    const float sB = ldexpf(1.f, e);
    inv = 1.0 / (sA * (double)sB);
#pragma unroll
    for (int q = 0; q < 64; ++q) {
      Bv[0][q] = split_f16(__uint_as_float(Bv[0][q]) * sB);
      Bv[1][q] = split_f16(__uint_as_float(Bv[1][q]) * sB);
    }
  }
  float* pcol = reinterpret_cast<float*>(part + (((long long)chunk * batch + bcol) * 2 + (cc >> 1)) * A.K) + (cc & 1);
  const float sig_odd = (cc >= 2) ? -1.0f : 1.0f;
  uint32_t* mytile = &tiles[wave][0][0][0];
  const int tcol = s2_tile_col(lane);
  const uint32_t* abase = &tiles[0][0][0][0] + (lane & 15) * TS2 + 16 * (lane >> 4);
  double2 pf_ag = make_double2(0.0, 0.0);
  auto prefetch = [&](int l0n) {
    const int l = l0n + tid;
    pf_ag = make_double2(0.0, 0.0);
    if (tid < CH && l <= lmax && l != lmin) pf_ag = make_double2(A.alpha[cb + l], A.gamma[cb + l]);
  };
  prefetch(lmin);
  for (int l0 = lmin; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) s_ag[tid] = pf_ag;
    __syncthreads();
    if (l0 + CH <= lmax) prefetch(l0 + CH);
#pragma unroll 1
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += TL) {
      int act = 0;
#pragma unroll 1
      for (int j0 = 0; j0 < TL; j0 += 8) {
        const bool deep = valid && ((ep < 0) || (em < 0));
        const int anydeep = __ballot(deep) != 0ull;
        act |= (__ballot(valid && ep == 0) != 0ull ? 1 : 0) | (__ballot(valid && em == 0) != 0ull ? 2 : 0);
        const double fp = (ep == 0 && valid) ? sA : 0.0, fm = (em == 0 && valid) ? sA : 0.0;
        double2 ag[8];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) ag[jj] = s_ag[jb + j0 + jj];
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double cp = fma(ag[jj].x, x, ag[jj].y);
          const double cm = fma(ag[jj].x, x, -ag[jj].y);
          const double np = fma(cp, pp, -ppp);
          const double nm = fma(cm, pm, -pmp);
          ppp = pp;
          pp = np;
          pmp = pm;
          pm = nm;
          mytile[(0 * TL + s2_tile_row(j0 + jj)) * TS2 + tcol] = split_f16((float)(np * fp));
          // odd rows (odd l - lmin) of the sign - tile negated: one B serves both parities
          mytile[(1 * TL + s2_tile_row(j0 + jj)) * TS2 + tcol] = split_f16((float)(nm * ((jj & 1) ? -fm : fm)));
        }
        if (anydeep) {
          rescale_pair(pp, ppp, ep);
          rescale_pair(pm, pmp, em);
        }
      }
      if (lane == 0) s_act[wave] = act;
      __syncthreads();
      int actv[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) actv[w] = s_act[w];
      // byte selectors of (hi, hi) and (lo, lo) from a (hi, lo) word.  pz is always 0, but
      // the compiler cannot prove it: the B operands built from the loop-invariant Bv then
      // stay in the loop instead of being hoisted as 256 more registers
      const uint32_t pz = (uint32_t)(l0 + jb > (1 << 30));
      const uint32_t sel_hh = 0x01000100u | pz, sel_ll = 0x03020302u | pz;
      // D[p][rb]: parity p, ring block rb, FP32 over the block's 64 rings, FP64 over blocks
      floatx4 D[2][4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) D[p][rb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
          if (!((actv[rb] >> sg) & 1)) continue;
          uint4 a[2][4];  // this lane's row: the 16 rings 4 n + lane/16 of the block, (hi, lo)
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const uint4* src = reinterpret_cast<const uint4*>(abase + ((rb * 2 + sg) * TL + 16 * p) * TS2);
#pragma unroll
            for (int n = 0; n < 4; ++n) a[p][n] = src[n];
          }
          // instruction pair n: rings 4 (4 n + i) + k, i = 0..3, of lane group k.  A lane's K
          // slice is the 4 rings' (hi, lo) words as the tile holds them (no copies); the pair
          // multiplies it by the (hi, hi) and by the (lo, lo) words of the rings' V values,
          // which is the split's four products at two instructions per 16 rings
#pragma unroll
          for (int n = 0; n < 4; ++n) {
            uint32_t hh[4], ll[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const uint32_t w = Bv[sg][16 * rb + 4 * n + i];
              hh[i] = __builtin_amdgcn_perm(w, w, sel_hh);
              ll[i] = __builtin_amdgcn_perm(w, w, sel_ll);
            }
            const halfx8 bh = __builtin_bit_cast(halfx8, make_uint4(hh[0], hh[1], hh[2], hh[3]));
            const halfx8 bl = __builtin_bit_cast(halfx8, make_uint4(ll[0], ll[1], ll[2], ll[3]));
#pragma unroll
            for (int p = 0; p < 2; ++p) {
              const halfx8 av = __builtin_bit_cast(halfx8, a[p][n]);
              D[p][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bh, D[p][rb], 0, 0, 0);
              D[p][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bl, D[p][rb], 0, 0, 0);
            }
          }
        }
      if (bcol < batch) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int l = l0 + jb + 2 * (4 * (lane >> 4) + r) + p;
            double v = 0.0;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) v += (double)D[p][rb][r];
            v *= inv;
            if (l <= lmax) pcol[2 * (cb + l)] = (p == 1) ? sig_odd * (float)v : (float)v;
          }
      }
      __syncthreads();
    }
  }
}

// The spin-2 adjoint for 32 maps per recursion: 8 waves over 256 rings, wave w running ONE
// recursion sign (w & 1) for the 64 rings of ring block w >> 1 and contracting column block
// w (maps 4 w .. 4 w + 3 of the group, components p.re, p.im, n.re, n.im) over all eight
// (ring block, sign) tiles, with the N/S parity fold, the local V scale and the FP16 hi + lo
// split of k_adj_mx_s2_cbh.  The contraction's K index is the ring: the tiles hold the hi
// and lo halves of the recursion values in separate planes ([row][ring]), and each lane's
// B operands are the hi and lo halves of its V values packed ring-pairwise, built once per
// work-group, so the split products come from one instruction each with no per-step operand
// rebuilding: hi.hi, hi.lo and lo.hi over 32 rings each.  lo.lo is left out: with both hi
// halves rounded to nearest the lo halves are below 2^-11 of their values, so the product is
// below 2^-22 of a.b, and its sign is random (a quarter fewer matrix instructions).  The tiles
// are double-buffered (one barrier per 32 l between the recursion of tile j + 1 and the
// contraction of tile j), and the two waves of each SIMD (w and w + 4) run the two halves in
// opposite order: one issues FP64 recursion while the other issues matrix instructions.
constexpr int kQ_RS = 72;  // plane row stride (halves): 144 bytes, rows 4 banks apart

// ring (within the 256-ring chunk) of entry t of a lane's B operands: t = 16 rb + 8 sl + e
// is ring 64 rb + 32 sl + 8 k4 + e, k4 = lane / 16 (the lane's K block of 8 rings)
__device__ __forceinline__ int q_ring(int t, int lane) {
  return 64 * (t >> 4) + 32 * ((t >> 3) & 1) + 8 * (lane >> 4) + (t & 7);
}

// Tiles before the first tile in which any of the work-group's recursions is active (values
// at or above 2^-256: activity only grows with l) contribute exactly zero: their partials are
// not stored, and lfirst[chunk * M + m] records the first stored l for the chunk reduction.
// Work order decode_tile<kA_MB, kA_CB>: each XCD takes groups of 8 consecutive orders m x 2
// ring chunks (adjacent orders read adjacent V entries).  16 x 1, 8 x 1, 8 x 4 and 16 x 2
// measured slower, although 8 x 2 fetches more (V lines shared by 8 orders instead of 16).
__global__ __launch_bounds__(512, 1) void k_adj_s2p(RecArgs A, const float2* __restrict__ V, double sA,
                                                   float2* __restrict__ part, int* __restrict__ lfirst, int batch) {
  constexpr int NW = 8, MG = 32, RINGS = 256;
  // one LDS block, (alpha, gamma) first: its per-step reads then take immediate offsets
  struct Smem {
    double2 ag[2][2][TL];  // [buffer][sign][step]: (alpha, +-gamma)
    int act[2][NW];        // [buffer][ring block 2 + sign]
    __align__(16) uint16_t tp[2][4][2][2][TL][kQ_RS];  // [buffer][ring block][sign][hi / lo][row][ring]
  };
  __shared__ Smem sm;
  auto& s_ag = sm.ag;
  auto& s_act = sm.act;
  auto& tp = sm.tp;
  int m, chunk;
  if (!decode_tile<kA_MB, kA_CB>(A, (int)blockIdx.x, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int sgw = wave & 1, rbw = wave >> 1;  // this wave's recursion: sign, ring block
  const int lmax = A.lmax;
  const int lmin = max(m, 2);
  if (lmin > lmax) {
    if (mg == 0 && tid == 0) lfirst[(long long)chunk * A.M + m] = lmax + 1;
    return;
  }
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const int t = chunk * RINGS + 64 * rbw + lane;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid) {
    x = A.x[t];
    const double* sd = A.seed + 8 * m + 4 * sgw;
    double sv;
    seed_split(spin2_seed_L(sd, A.lcos[t], A.lsin[t]), sd[3], sv, ex);
    mup = -sv;
  }
  // B operands of column block cw = wave: column lane % 16 -> map 4 cw + (lane % 16) / 4,
  // component lane % 4; Bq[sign][2 rb + sl][hi / lo] holds the halves of the lane's 8 rings
  // q_ring(16 rb + 8 sl + e), e = 0..7, in ring pairs
  const int g = (lane & 15) >> 2, cc = lane & 3;
  const int bcol = mg * MG + 4 * wave + g;
  uint4 Bq[2][8][2];
  int esc;  // 1 / (sA sB) of this lane's map is 2^esc (both scales are powers of two)
  {
    const float par0 = ((lmin - m) & 1) ? -1.0f : 1.0f;
    const bool lo = cc < 2;
    const float* Vs =
        reinterpret_cast<const float*>(V + ((long long)min(bcol, batch - 1) * 2 + (lo ? 0 : 1)) * A.nrings * ringM);
    float bf[2][64];  // [sign][t]; every load in flight before any is used (clamped addresses)
#pragma unroll
    for (int q = 0; q < 64; ++q) {
      const int tr = min(chunk * RINGS + q_ring(q, lane), A.T - 1);
      bf[0][q] = Vs[2 * ((long long)tr * ringM + m) + (cc & 1)];                   // north ring
      bf[1][q] = Vs[2 * ((long long)(A.nrings - 1 - tr) * ringM + m) + (cc & 1)];  // its mirror
    }
    float mx = 0.f;
#pragma unroll
    for (int q = 0; q < 64; ++q) {
      const int tr = chunk * RINGS + q_ring(q, lane);
      const bool ok = bcol < batch && tr < A.T;
      const bool south = ok && tr < A.T - 1;  // the equator ring has no mirror
      const float vn = ok ? bf[0][q] : 0.f;
      const float vs = south ? par0 * bf[1][q] : 0.f;
      // sign + column (p): V+ north and par0 V+ south; sign - column (n): the reverse
      bf[0][q] = lo ? vn : vs;
      bf[1][q] = lo ? vs : vn;
      mx = fmaxf(mx, fmaxf(fabsf(bf[0][q]), fabsf(bf[1][q])));
    }
    // the map's maximum over its 16 lanes (4 components x 4 K blocks): bits 0, 1, 4, 5
    mx = fmaxf(mx, __shfl_xor(mx, 1));
    mx = fmaxf(mx, __shfl_xor(mx, 2));
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    int e = 1;
    if (mx > 0.f && mx < 3.0e38f) (void)frexpf(4096.f / mx, &e);
    e = min(max(e - 1, -120), 120);
    const float sB = ldexpf(1.f, e);
    esc = -(ilogb(sA) + ilogb((double)sB));
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        uint32_t w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) w[k] = split_f16_mix(bf[s][8 * u + k] * sB);
        // hi halves of rings 2i, 2i + 1 -> dword i; lo halves likewise
        Bq[s][u][0] = make_uint4(__builtin_amdgcn_perm(w[1], w[0], 0x05040100u), __builtin_amdgcn_perm(w[3], w[2], 0x05040100u),
                                 __builtin_amdgcn_perm(w[5], w[4], 0x05040100u), __builtin_amdgcn_perm(w[7], w[6], 0x05040100u));
        Bq[s][u][1] = make_uint4(__builtin_amdgcn_perm(w[1], w[0], 0x07060302u), __builtin_amdgcn_perm(w[3], w[2], 0x07060302u),
                                 __builtin_amdgcn_perm(w[5], w[4], 0x07060302u), __builtin_amdgcn_perm(w[7], w[6], 0x07060302u));
      }
  }
  // this lane's partial vector (map bcol, component cc / 2) in the tile layout: tile jt of
  // order m at 32 (tile_base(m) + jt); the lane's rows 8 k + o of a tile (k = lane / 16)
  float2* const pvec = part + (((long long)chunk * batch + bcol) * 2 + (cc >> 1)) * tile_part_len(lmax, 2) +
                       32 * tile_base(m, lmax, 2) + 8 * (lane >> 4);
  const float sig_odd = (cc >= 2) ? -1.0f : 1.0f;
  const int ntile = (lmax - lmin + TL) / TL;
  int jf = ntile;  // first active tile (work-group uniform)
  // 32 recursion steps of this wave's sign into its planes of buffer b; odd rows of the
  // sign - tile negated (one B serves both parities, as in k_adj_mx_s2_cbh)
  // this lane's coefficients of the tile in buffer b (steps lane % 16 and 16 + lane % 16),
  // broadcast per step by ag_step
  double2 cf[2];
  auto load_cf = [&](int b) {
    cf[0] = s_ag[b][sgw][lane & 15];
    cf[1] = s_ag[b][sgw][16 + (lane & 15)];
  };
  static_assert(TL == 32, "ag_step covers 32 steps");
  auto rec_tile = [&](int b) {
    uint16_t* ph = &tp[b][rbw][sgw][0][0][lane];
    uint16_t* pl = &tp[b][rbw][sgw][1][0][lane];
    int act = 0;
    static_for<0, TL / 8>([&](auto gc) {
      constexpr int j0 = 8 * decltype(gc)::value;
      const int anydeep = __ballot(valid && ex < 0) != 0ull;
      act |= __ballot(valid && ex == 0) != 0ull;
      const double f = (ex == 0 && valid) ? sA : 0.0;
      const double fo = sgw ? -f : f;  // odd steps
      static_for<0, 8>([&](auto jc) {
        constexpr int jj = decltype(jc)::value;
        const double n = fma(ag_step<j0 + jj>(cf, x), mu, -mup);
        mup = mu;
        mu = n;
        uint32_t w = split_f16_mix((float)(n * ((jj & 1) ? fo : f)));
        // one packed conversion, then its two halves stored (ds_write_b16 / _d16_hi): left
        // visible, the compiler converts each half separately
        asm volatile("" : "+v"(w));
        const int row = s2_tile_row(j0 + jj) * kQ_RS;
        ph[row] = (uint16_t)w;
        pl[row] = (uint16_t)(w >> 16);
      });
      if (anydeep) rescale_pair(mu, mup, ex);
    });
    if (lane == 0) s_act[b][wave] = act;
  };
  // A operands: lane (row i = lane % 16, K block k4 = lane / 16) reads 8 rings of a plane row
  const int a_off = (lane & 15) * kQ_RS + 8 * (lane >> 4);
  auto contract_tile = [&](int jt, int b) {
    int actv[NW];
    bool any = false;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      actv[w] = s_act[b][w];
      any |= actv[w] != 0;
    }
    if (any && jf == ntile) jf = jt;
    floatx4 D[2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) D[p][rb] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (any) {
      // 16 units u = (ring block rb, sign sg, 32-ring slice sl) of 6 matrix instructions
      // each; unit u + 1's A operands are read after unit u's instructions are issued (LDS
      // reads complete in order: read before them they would hold them up)
      uint4 ah[2][2], al[2][2];  // [buffer][parity]
      auto load_unit = [&](int u, int buf) {
        const int rb = u >> 2, sg = (u >> 1) & 1, sl = u & 1;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const int o = a_off + 16 * p * kQ_RS + 32 * sl;
          ah[buf][p] = *reinterpret_cast<const uint4*>(&tp[b][rb][sg][0][0][o]);
          al[buf][p] = *reinterpret_cast<const uint4*>(&tp[b][rb][sg][1][0][o]);
        }
      };
      load_unit(0, 0);
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int rb = u >> 2, sg = (u >> 1) & 1, sl = u & 1;
        if (actv[2 * rb + sg]) {
          const halfx8 bh = __builtin_bit_cast(halfx8, Bq[sg][2 * rb + sl][0]);
          const halfx8 bl = __builtin_bit_cast(halfx8, Bq[sg][2 * rb + sl][1]);
#pragma unroll
          for (int p = 0; p < 2; ++p) {
            const halfx8 xh = __builtin_bit_cast(halfx8, ah[u & 1][p]);
            const halfx8 xl = __builtin_bit_cast(halfx8, al[u & 1][p]);
            D[p][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bh, D[p][rb], 0, 0, 0);
            D[p][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bl, D[p][rb], 0, 0, 0);
            D[p][rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, bh, D[p][rb], 0, 0, 0);
          }
        }
        if (u + 1 < 16) load_unit(u + 1, (u + 1) & 1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // D[i][j] at lane j + 16 (i / 4), register i % 4: row i -> l = lmin + TL jt + 2 i + p, so
    // this lane holds rows 8 k + o, o = 2 r + p = 0..7, of column j.  The four ring blocks'
    // sums are added in FP32 (rounding 2^-24 of the sum, far below the split's 2^-22) and
    // scaled by an exact power of two.  The re lane (cc even) and its im partner (lane ^ 1)
    // swap halves so that each writes four (re, im) pairs, 32 contiguous bytes: the re lane
    // rows 0..3, the im lane rows 4..7.  Rows past lmax land in the tile's padding, which the
    // chunk reduction does not read.
    if (bcol < batch && jf <= jt) {
      float v[8];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float u = ldexpf((D[p][0][r] + D[p][1][r]) + (D[p][2][r] + D[p][3][r]), esc);
          v[2 * r + p] = (p == 1) ? sig_odd * u : u;
        }
      const bool odd = (lane & 1) != 0;
      float w[4];  // the partner's rows o (to the re lane) or 4 + o (to the im lane)
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        const float lo = swap_pair(v[o]), hi = swap_pair(v[4 + o]);
        w[o] = odd ? hi : lo;
      }
      float4* const dst = reinterpret_cast<float4*>(pvec + 32 * jt + (odd ? 4 : 0));
      dst[0] = odd ? make_float4(w[0], v[4], w[1], v[5]) : make_float4(v[0], w[0], v[1], w[1]);
      dst[1] = odd ? make_float4(w[2], v[6], w[3], v[7]) : make_float4(v[2], w[2], v[3], w[3]);
    }
  };
  // (alpha, gamma) of step j of tile jt (zero past lmax, and at lmin: the seed step)
  auto ag_at = [&](int jt, int j) {
    const int l = lmin + TL * jt + j;
    return (l <= lmax && l != lmin) ? make_double2(A.alpha[cb + l], A.gamma[cb + l]) : make_double2(0.0, 0.0);
  };
  auto stage_ag = [&](int b, double2 v) {
    s_ag[b][0][tid] = v;
    s_ag[b][1][tid] = make_double2(v.x, -v.y);
  };
  if (tid < TL) {
    stage_ag(0, ag_at(0, tid));
    stage_ag(1, ag_at(1, tid));
  }
  double2 pf = (tid < TL) ? ag_at(2, tid) : make_double2(0.0, 0.0);
  __syncthreads();
  load_cf(0);
  rec_tile(0);
  __syncthreads();
  const bool rec_first = wave < 4;  // waves w and w + 4 share a SIMD
  for (int jt = 0; jt < ntile; ++jt) {
    const int b = jt & 1;
    // s_ag[b] held tile jt, read by the recursion before the last barrier: now tile jt + 2
    if (tid < TL) {
      stage_ag(b, pf);
      pf = ag_at(jt + 3, tid);
    }
    const bool more = jt + 1 < ntile;
    if (more) load_cf(b ^ 1);  // tile jt + 1's coefficients
    if (rec_first) {
      if (more) rec_tile(b ^ 1);
      contract_tile(jt, b);
    } else {
      contract_tile(jt, b);
      if (more) rec_tile(b ^ 1);
    }
    __syncthreads();
  }
  if (mg == 0 && tid == 0) lfirst[(long long)chunk * A.M + m] = lmin + TL * jf;  // above lmax if none
}

// spin 0: columns j = 2 g + (re, im), MG = 8 maps; U0 = Vn + Vs (even l - m), U1 = Vn - Vs
__global__ __launch_bounds__(256) void k_adj_mx_s0(RecArgs A, const float2* __restrict__ V,
                                                   float2* __restrict__ part, int batch) {
  constexpr int MG = 8;
  __shared__ double s_al[CH];
  __shared__ float tiles[4][TL][TS];  // [wave][l][ring] (one sign)
  __shared__ double redbuf[4][2][16][16];
  int m, chunk;
  if (!decode_block<true>(A, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const int t = chunk * 256 + tid;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid) {
    x = A.x[t];
    const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
    double s;
    seed_split(L, (m & 1) ? -1.0 : 1.0, s, ex);
    mup = -s;
  }
  float Bv[1][2][16];
  {
    const int g = (lane & 15) >> 1, c = lane & 1, b = mg * MG + g;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int tr = chunk * 256 + wave * 64 + 4 * q + (lane >> 4);
      double vn = 0.0, vs = 0.0;
      if (tr < A.T && b < batch) {
        const float2* Vb = V + (long long)b * A.nrings * ringM;
        const float2 n = Vb[(long long)tr * ringM + m];
        vn = c ? n.y : n.x;
        if (tr < A.T - 1) {
          const float2 s = Vb[(long long)(A.nrings - 1 - tr) * ringM + m];
          vs = c ? s.y : s.x;
        }
      }
      Bv[0][0][q] = (float)(vn + vs);
      Bv[0][1][q] = (float)(vn - vs);
      if ((q & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
  float* mytile = &tiles[wave][0][0];
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      s_al[tid] = (l <= lmax && l != m) ? A.alpha[cb + l] : 0.0;
    }
    __syncthreads();
#pragma unroll 1
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += TL) {
#pragma unroll 1
      for (int j0 = 0; j0 < TL; j0 += 8) {
        const int anydeep = __ballot(valid && ex < 0) != 0ull;
        const double f = (ex == 0 && valid) ? 1.0 : 0.0;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double c = s_al[jb + j0 + jj] * x;
          const double mn = fma(c, mu, -mup);
          mup = mu;
          mu = mn;
          mytile[(j0 + jj) * TS + lane] = (float)(mn * f);
        }
        if (anydeep) rescale_pair(mu, mup, ex);
      }
      wave_sync();
      floatx4 D[2];
      adj_tile_mfma<1>(mytile, lane, Bv, D);
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int r = 0; r < 4; ++r) redbuf[wave][p][4 * (lane >> 4) + r][lane & 15] = (double)D[p][r];
      __syncthreads();
      if (tid < 2 * 16 * MG) {
        const int p = tid / (16 * MG), i = (tid / MG) % 16, g = tid % MG;
        const int l = l0 + jb + 2 * i + p;
        const int b = mg * MG + g;
        if (l <= lmax && b < batch) {
          double re = 0.0, im = 0.0;
          for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            re += redbuf[w][p][i][2 * g];
            im += redbuf[w][p][i][2 * g + 1];
          }
          part[((long long)chunk * batch + b) * A.K + cb + l] = make_float2((float)re, (float)im);
        }
      }
      __syncthreads();
    }
  }
}

// Column-block form of the spin-0 adjoint for 16 maps (4 waves, 256 rings per work-group).
// Wave w runs the recursion of rings 64 w .. 64 w + 63 of the chunk into its tile (layout
// of the spin-2 tiles: row 16 p + i holds step 2 i + p, ring r in column 16 (r % 4) + r / 4);
// after a barrier it contracts parity p = w / 2 of column block c = w % 2 (maps 8 c .. 8 c + 7,
// re / im) over all 256 rings against U_p = V_north +- V_south.  Every wave does the same
// matrix work and skips the same all-zero ring-block tiles; no cross-wave reduction.  Each
// 64-ring block accumulates in FP32 (two chains of 32 consecutive rings) and the chains
// are summed in FP64.
__device__ __forceinline__ void adj_mx_s0_cb_body(const RecArgs& A, const float2* __restrict__ V,
                                                  float2* __restrict__ part, int batch, int P, int mg) {
  constexpr int NW = 4, MG = 16, RINGS = 256;
  __shared__ double s_al[CH];
  __shared__ __align__(16) float tiles[NW][TL][TS2];  // [ring block][row][col]
  __shared__ int s_act[NW];
  int m, chunk;
  if (!decode_block<true>(A, P, m, chunk)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const int t = chunk * RINGS + tid;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid) {
    x = A.x[t];
    const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
    double sv;
    seed_split(L, (m & 1) ? -1.0 : 1.0, sv, ex);
    mup = -sv;
  }
  // this wave's output: parity pw, column block cw; lane column j = lane % 16 -> map
  // 8 cw + j / 2, re / im = j % 2; B rows: chunk ring 4 q + lane / 16, q = 0..63
  const int pw = wave >> 1, cw = wave & 1;
  const int j = lane & 15;
  const int bcol = mg * MG + 8 * cw + (j >> 1);
  float Bv[64];
#pragma unroll
  for (int h = 0; h < 4; ++h) {  // four quarters of 32 loads each: fewer registers in flight
    float vn[16], vs[16];
    load_ring_values(A, V, bcol, batch, 1, 0, j & 1, m, chunk, lane, 16 * h, vn, vs);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      bool ok, south;  // south: not the equator, which has no mirror
      ring_valid(A, bcol, batch, chunk, 16 * h + q, lane, ok, south);
      const double n = ok ? (double)vn[q] : 0.0, sv = south ? (double)vs[q] : 0.0;
      Bv[16 * h + q] = (float)(pw ? n - sv : n + sv);
      // materialised here: otherwise the compiler recomputes it from both raw values in the
      // l loop and holds twice the registers there (occupancy 3 -> 2)
      asm volatile("" : "+v"(Bv[16 * h + q]));
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the next quarter's loads behind this one's use
  }
  float* pcol = reinterpret_cast<float*>(part + ((long long)chunk * batch + bcol) * A.K) + (j & 1);
  float* mytile = &tiles[wave][0][0];
  const int tcol = s2_tile_col(lane);
  const float* abase = &tiles[0][0][0] + (16 * pw + (lane & 15)) * TS2 + 16 * (lane >> 4);
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      s_al[tid] = (l <= lmax && l != m) ? A.alpha[cb + l] : 0.0;
    }
    __syncthreads();
#pragma unroll 1
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += TL) {
      int act = 0;
#pragma unroll 1
      for (int j0 = 0; j0 < TL; j0 += 8) {
        const int anydeep = __ballot(valid && ex < 0) != 0ull;
        act |= __ballot(valid && ex == 0) != 0ull;
        const double f = (ex == 0 && valid) ? 1.0 : 0.0;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double c = s_al[jb + j0 + jj] * x;
          const double mn = fma(c, mu, -mup);
          mup = mu;
          mu = mn;
          mytile[s2_tile_row(j0 + jj) * TS2 + tcol] = (float)(mn * f);
        }
        if (anydeep) rescale_pair(mu, mup, ex);
      }
      if (lane == 0) s_act[wave] = act;
      __syncthreads();
      floatx4 D[NW][2];
#pragma unroll
      for (int rb = 0; rb < NW; ++rb) D[rb][0] = D[rb][1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int rb = 0; rb < NW; ++rb) {
        if (!s_act[rb]) continue;
        floatx4 a[4];
        const floatx4* src = reinterpret_cast<const floatx4*>(abase + rb * TL * TS2);
#pragma unroll
        for (int n = 0; n < 4; ++n) a[n] = src[n];
        // two chains over the contiguous halves of the block (rings 0-31, 32-63), issued
        // alternately: a strided split (every other group of 4 rings) sums nearly in-phase
        // terms at high l, where consecutive rings advance the recursion by about pi/2,
        // and measured 2-3x the error in the l >= 1024 bands
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) {
          D[rb][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[qq >> 2][qq & 3], Bv[16 * rb + qq], D[rb][0], 0, 0, 0);
          D[rb][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2 + (qq >> 2)][qq & 3], Bv[16 * rb + 8 + qq], D[rb][1],
                                                          0, 0, 0);
        }
      }
      // D[i][j] at lane j + 16 (i / 4), register i % 4: row i -> l = l0 + jb + 2 i + pw
      if (bcol < batch) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int l = l0 + jb + 2 * (4 * (lane >> 4) + r) + pw;
          double v = 0.0;
#pragma unroll
          for (int rb = 0; rb < NW; ++rb) v += (double)D[rb][0][r] + (double)D[rb][1][r];
          if (l <= lmax) pcol[2 * (cb + l)] = (float)v;
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256, 2) void k_adj_mx_s0_cb(RecArgs A, const float2* __restrict__ V,
                                                         float2* __restrict__ part, int batch) {
  adj_mx_s0_cb_body(A, V, part, batch, blockIdx.x, blockIdx.y);
}

// k_adj_mx_s0_cb on the FP16 matrix cores, with the split, the local V scale and the
// instruction layout of k_adj_mx_s2_cbh: 8 rings per instruction, the two chains over the
// block halves (rings 0-31, 32-63) kept.
__global__ __launch_bounds__(256, 2) void k_adj_mx_s0_cbh(RecArgs A, const float2* __restrict__ V, double sA,
                                                         float2* __restrict__ part, int batch) {
  constexpr int NW = 4, MG = 16, RINGS = 256;
  __shared__ double s_al[CH];
  __shared__ __align__(16) uint32_t tiles[NW][TL][TS2];  // [ring block][row][col]: (hi, lo)
  __shared__ int s_act[NW];
  int m, chunk;
  if (!decode_tile<kA_MB, kA_CB>(A, (int)blockIdx.x, m, chunk)) return;
  const int mg = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lmax = A.lmax;
  const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
  const long long ringM = A.M;
  const int t = chunk * RINGS + tid;
  const bool valid = t < A.T;
  double x = 0.0, mu = 0.0, mup = 0.0;
  int ex = 0;
  if (valid) {
    x = A.x[t];
    const double L = __dadd_rn(A.seed[m], __dmul_rn((double)m, A.lsin[t]));
    double sv;
    seed_split(L, (m & 1) ? -1.0 : 1.0, sv, ex);
    mup = -sv;
  }
  // this wave's output: parity pw, column block cw; lane column j = lane % 16 -> map
  // 8 cw + j / 2, re / im = j % 2; B rows: chunk ring 4 q + lane / 16, q = 0..63
  const int pw = wave >> 1, cw = wave & 1;
  const int j = lane & 15;
  const int bcol = mg * MG + 8 * cw + (j >> 1);
  uint32_t Bv[64];  // (hi, lo) of the scaled value
  double inv;
  {
    float mx = 0.f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // two halves of 64 loads each: fewer registers in flight
      float vn[32], vs[32];
      load_ring_values(A, V, bcol, batch, 1, 0, j & 1, m, chunk, lane, 32 * h, vn, vs);
#pragma unroll
      for (int q = 0; q < 32; ++q) {
        bool ok, south;  // south: not the equator, which has no mirror
        ring_valid(A, bcol, batch, chunk, 32 * h + q, lane, ok, south);
        const double n = ok ? (double)vn[q] : 0.0, sv = south ? (double)vs[q] : 0.0;
        const float uf = (float)(pw ? n - sv : n + sv);
        mx = fmaxf(mx, fabsf(uf));
        Bv[32 * h + q] = __float_as_uint(uf);
      }
      __builtin_amdgcn_sched_barrier(0);  // keep the second half's loads behind the first's use
    }
    // the map's maximum over its 8 lanes (re / im x 4 ring residues): bits 0, 4 and 5
    mx = fmaxf(mx, __shfl_xor(mx, 1));
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    int e = 1;
    if (mx > 0.f && mx < 3.0e38f) (void)frexpf(4096.f / mx, &e);
    e = min(max(e - 1, -120), 120);
    const float sB = ldexpf(1.f, e);
    inv = 1.0 / (sA * (double)sB);
#pragma unroll
    for (int q = 0; q < 64; ++q) Bv[q] = split_f16(__uint_as_float(Bv[q]) * sB);
  }
  float* pcol = reinterpret_cast<float*>(part + ((long long)chunk * batch + bcol) * A.K) + (j & 1);
  uint32_t* mytile = &tiles[wave][0][0];
  const int tcol = s2_tile_col(lane);
  const uint32_t* abase = &tiles[0][0][0] + (16 * pw + (lane & 15)) * TS2 + 16 * (lane >> 4);
  for (int l0 = m; l0 <= lmax; l0 += CH) {
    __syncthreads();
    if (tid < CH) {
      const int l = l0 + tid;
      s_al[tid] = (l <= lmax && l != m) ? A.alpha[cb + l] : 0.0;
    }
    __syncthreads();
#pragma unroll 1
    for (int jb = 0; jb < CH && l0 + jb <= lmax; jb += TL) {
      int act = 0;
#pragma unroll 1
      for (int j0 = 0; j0 < TL; j0 += 8) {
        const int anydeep = __ballot(valid && ex < 0) != 0ull;
        act |= __ballot(valid && ex == 0) != 0ull;
        const double f = (ex == 0 && valid) ? sA : 0.0;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
          const double c = s_al[jb + j0 + jj] * x;
          const double mn = fma(c, mu, -mup);
          mup = mu;
          mu = mn;
          mytile[s2_tile_row(j0 + jj) * TS2 + tcol] = split_f16((float)(mn * f));
        }
        if (anydeep) rescale_pair(mu, mup, ex);
      }
      if (lane == 0) s_act[wave] = act;
      __syncthreads();
      // byte selectors of (hi, hi) and (lo, lo); pz is always 0 but not provably so, which
      // keeps the operands built from the loop-invariant Bv inside the loop
      const uint32_t pz = (uint32_t)(l0 + jb > (1 << 30));
      const uint32_t sel_hh = 0x01000100u | pz, sel_ll = 0x03020302u | pz;
      floatx4 D[NW][2];
#pragma unroll
      for (int rb = 0; rb < NW; ++rb) D[rb][0] = D[rb][1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int rb = 0; rb < NW; ++rb) {
        if (!s_act[rb]) continue;
        uint4 a[4];  // this lane's row: rings 4 n + lane/16, n = 0..15, (hi, lo)
        const uint4* src = reinterpret_cast<const uint4*>(abase + rb * TL * TS2);
#pragma unroll
        for (int n = 0; n < 4; ++n) a[n] = src[n];
        // instruction n: rings 4 (2 n) + k and 4 (2 n + 1) + k; chain n / 4 (block halves)
#pragma unroll
        for (int n = 0; n < 8; ++n) {
          const uint32_t w0 = Bv[16 * rb + 2 * n], w1 = Bv[16 * rb + 2 * n + 1];
          const halfx8 bq = __builtin_bit_cast(
              halfx8, make_uint4(__builtin_amdgcn_perm(w0, w0, sel_hh), __builtin_amdgcn_perm(w0, w0, sel_ll),
                                 __builtin_amdgcn_perm(w1, w1, sel_hh), __builtin_amdgcn_perm(w1, w1, sel_ll)));
          const uint4 av4 = a[n >> 1];
          const uint32_t a0 = (n & 1) ? av4.z : av4.x, a1 = (n & 1) ? av4.w : av4.y;
          D[rb][n >> 2] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
              __builtin_bit_cast(halfx8, make_uint4(a0, a0, a1, a1)), bq, D[rb][n >> 2], 0, 0, 0);
        }
      }
      // D[i][j] at lane j + 16 (i / 4), register i % 4: row i -> l = l0 + jb + 2 i + pw
      if (bcol < batch) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int l = l0 + jb + 2 * (4 * (lane >> 4) + r) + pw;
          double v = 0.0;
#pragma unroll
          for (int rb = 0; rb < NW; ++rb) v += (double)D[rb][0][r] + (double)D[rb][1][r];
          if (l <= lmax) pcol[2 * (cb + l)] = (float)(v * inv);
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------------- //
// ragged groups: the FP32 matrix-core kernels of several plans in one launch per stage
// ---------------------------------------------------------------------------------- //
// Per plan (device-resident): recursion arguments of the two stages (their ring-chunk
// widths differ for spin 2) and the workspace offsets.
struct GroupRec {
  RecArgs synth, adj;
  long long f_off, part_off;
  int batch, synth_gx, adj_gx;
};

// 16 maps per recursion (the last group of a plan ragged): k_synth_tile_s0<2, 2> or
// k_synth_tile_s2<4, 2> per work-group.
template <int SPIN>
__global__ __launch_bounds__(256, 2) void k_group_synth(const GroupRec* __restrict__ G, GroupCall C) {
  int local;
  const int i = group_member(C, local);
  const RecArgs A = G[i].synth;
  const int gx = G[i].synth_gx, batch = G[i].batch;
  float2* F = reinterpret_cast<float2*>(C.ws + G[i].f_off);
  const double2* alm = reinterpret_cast<const double2*>(C.in[i]);
  if constexpr (SPIN == 0)
    synth_tile_s0_body<2>(A, alm, F, batch, local % gx, local / gx);
  else
    synth_tile_s2_body<4>(A, alm, F, batch, local % gx, local / gx);
}

// k_adj_mx_s0_cb / k_adj_mx_s2_cb per work-group (16 maps, 256 rings)
template <int SPIN>
__global__ __launch_bounds__(256, 2) void k_group_adj(const GroupRec* __restrict__ G, GroupCall C) {
  int local;
  const int i = group_member(C, local);
  const RecArgs A = G[i].adj;
  const int gx = G[i].adj_gx, batch = G[i].batch;
  const float2* V = reinterpret_cast<const float2*>(C.ws + G[i].f_off);
  float2* part = reinterpret_cast<float2*>(C.ws + G[i].part_off);
  if constexpr (SPIN == 0)
    adj_mx_s0_cb_body(A, V, part, batch, local % gx, local / gx);
  else
    adj_mx_s2_cb_body(A, V, part, batch, local % gx, local / gx);
}

}  // namespace

RecArgs make_args_mixed(const jhtx_plan* p) {
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
  a.nchunk = (p->T + 255) / 256;
  a.K = p->K;
  a.Mf = p->M;
  a.sabs = p->spin;  // 0 or 2 (HEALPix plans)
  a.ssign = 1.0;
  return a;
}

// ring chunks of the mixed adjoint: every kernel's work-group covers 256 rings
int mixed_adj_chunks(const jhtx_plan* p) { return (p->T + 255) / 256; }

int mixed_group(const jhtx_plan* p, bool adjoint);

// the 32-map spin-2 adjoint k_adj_s2p runs for this plan
static bool adj_s2p(const jhtx_plan* p) {
  return p->spin == 2 && mixed_group(p, true) == 32 && p->mx_adj_cb && p->mx_adj_f16;
}

// padded rows where the synthesis runs k_synth_s2q (mixed spin 2, 32 maps per recursion,
// FP16 split, parity fold; a jhtx_synth call always has the scale workspace).  Padding the
// adjoint's V the same way (read by k_adj_s2p) measured slower, 69.1 -> 69.8 ms.
long long synth_ring_ld(const jhtx_plan* p, int batch) {
  const bool s2q = p->precision == JHTX_PREC_MIXED && !(p->mx_single && batch == 1) && p->spin == 2 &&
                   mixed_group(p, false) == 32 && p->mx_f16 && p->mx_synth_fold;
  return s2q ? padded_ld(p->M) : p->M;
}

long long adj_part_len(const jhtx_plan* p) {
  if (p->precision != JHTX_PREC_MIXED || !adj_s2p(p)) return p->K;
  return std::max<long long>(p->K, tile_part_len(p->lmax, 2));
}

size_t adj_lfirst_bytes(const jhtx_plan* p) {
  if (p->precision != JHTX_PREC_MIXED || !adj_s2p(p)) return 0;
  return ((sizeof(int) * (size_t)mixed_adj_chunks(p) * p->M) + 255) & ~(size_t)255;
}

int mixed_group(const jhtx_plan* p, bool adjoint) {
  if (p->spin == 0) {
    const int g = adjoint ? p->mx_adj_group : p->mx_synth_group;
    return (!adjoint && g >= 32) ? 32 : (g >= 16 ? 16 : 8);
  }
  return adjoint ? p->mx_adj_group : p->mx_synth_group;
}

hipError_t launch_synth_mixed(const jhtx_plan* p, const double2* alm, float2* F, int batch,
                              void* scale_ws, hipStream_t s, long long ld) {
  const RecArgs a = make_args_mixed(p);
  const int Mpad = (p->M + 8 * kMGroup - 1) / (8 * kMGroup) * (8 * kMGroup);
  int MG = mixed_group(p, false);
  // the spin-2 32-map form is the parity-folded FP16 kernel, which needs the scale pass's
  // workspace (jhtx_legendre_synth has none): otherwise 16 maps per recursion
  if (p->spin == 2 && MG == 32 && !(p->mx_f16 && p->mx_synth_fold && scale_ws)) MG = 16;
  dim3 grid(Mpad * a.nchunk, (batch + MG - 1) / MG);
  if (p->spin == 2 && (MG == 16 || MG == 32) && p->mx_f16 && p->mx_synth_fold && scale_ws) {
    // parity folded into B: 16 maps x 128 rings (k_synth_s2p, 4 waves) or 32 maps x 256 rings
    // (k_synth_s2q, 8 waves, hi / lo planes)
    double* part = reinterpret_cast<double*>(scale_ws);
    float* sB = reinterpret_cast<float*>(part + (size_t)kScaleBlocks * batch);
    hipLaunchKernelGGL(k_alm_absmax, dim3(kScaleBlocks, batch), dim3(256), 0, s, alm, 2 * p->K, part);
    hipLaunchKernelGGL(k_alm_scale, dim3(batch), dim3(256), 0, s, part, p->mx_gmax, sB);
    RecArgs at = a;
    const int rings = (MG == 32) ? 256 : 128;
    at.nchunk = (p->T + rings - 1) / rings;
    if (MG == 32) {
      // Mpad is a multiple of 8 kQ_MB; the grid holds Mpad x ceil(nchunk / kQ_CB) x kQ_CB slots
      static_assert((8 * kMGroup) % (8 * kQ_MB) == 0, "Mpad covers whole order groups");
      if (ld) at.Mf = ld;
      const dim3 gq(Mpad * ((at.nchunk + kQ_CB - 1) / kQ_CB) * kQ_CB, (batch + MG - 1) / MG);
      hipLaunchKernelGGL((k_synth_s2q<kQ_MB, kQ_CB>), gq, dim3(512), 0, s, at, alm, sB, p->mx_sA, F, batch);
    } else {
      const dim3 gt(Mpad * ((at.nchunk + kQ_CB - 1) / kQ_CB) * kQ_CB, (batch + MG - 1) / MG);
      hipLaunchKernelGGL((k_synth_s2p<4, 4, 3, kQ_MB, kQ_CB>), gt, dim3(256), 0, s, at, alm, sB, p->mx_sA, F, batch);
    }
  } else if (p->spin == 2 && MG == 16) {
    RecArgs at = a;
    at.nchunk = (p->T + 127) / 128;
    dim3 gt(Mpad * at.nchunk, (batch + 15) / 16);
    if (p->mx_f16 && scale_ws) {
      double* part = reinterpret_cast<double*>(scale_ws);
      float* sB = reinterpret_cast<float*>(part + (size_t)kScaleBlocks * batch);
      hipLaunchKernelGGL(k_alm_absmax, dim3(kScaleBlocks, batch), dim3(256), 0, s, alm, 2 * p->K, part);
      hipLaunchKernelGGL(k_alm_scale, dim3(batch), dim3(256), 0, s, part, p->mx_gmax, sB);
      hipLaunchKernelGGL((k_synth_tile_s2h<4, 2>), gt, dim3(256), 0, s, at, alm, sB, p->mx_sA, F, batch);
    } else {
      hipLaunchKernelGGL((k_synth_tile_s2<4, 2>), gt, dim3(256), 0, s, at, alm, F, batch);
    }
  } else if (p->spin == 0 && MG == 32)
    hipLaunchKernelGGL((k_synth_tile_s0<4, 2>), grid, dim3(256), 0, s, a, alm, F, batch);
  else if (p->spin == 0 && MG == 16 && p->mx_f16 && scale_ws) {
    double* part = reinterpret_cast<double*>(scale_ws);
    float* sB = reinterpret_cast<float*>(part + (size_t)kScaleBlocks * batch);
    hipLaunchKernelGGL(k_alm_absmax, dim3(kScaleBlocks, batch), dim3(256), 0, s, alm, p->K, part);
    hipLaunchKernelGGL(k_alm_scale, dim3(batch), dim3(256), 0, s, part, p->mx_gmax, sB);
    hipLaunchKernelGGL((k_synth_tile_s0h<2, 2>), dim3(Mpad * ((a.nchunk + kA_CB - 1) / kA_CB) * kA_CB, grid.y), dim3(256), 0,
                       s, a, alm, sB, p->mx_sA, F, batch);
  } else if (p->spin == 0 && MG == 16)
    hipLaunchKernelGGL((k_synth_tile_s0<2, 2>), grid, dim3(256), 0, s, a, alm, F, batch);
  else if (p->spin == 0)
    hipLaunchKernelGGL((k_synth_mx_s0<8>), grid, dim3(256), 0, s, a, alm, F, batch);
  else if (MG == 8 && p->mx_synth_occ == 3)
    hipLaunchKernelGGL((k_synth_mx_s2<2, 3>), grid, dim3(256), 0, s, a, alm, F, batch);
  else if (MG == 8 && p->mx_synth_occ == 2)
    hipLaunchKernelGGL((k_synth_mx_s2<2, 2>), grid, dim3(256), 0, s, a, alm, F, batch);
  else if (MG == 8)
    hipLaunchKernelGGL((k_synth_mx_s2<2, 1>), grid, dim3(256), 0, s, a, alm, F, batch);
  else
    hipLaunchKernelGGL((k_synth_mx_s2<1, 1>), grid, dim3(256), 0, s, a, alm, F, batch);
  return hipGetLastError();
}

hipError_t launch_adj_mixed(const jhtx_plan* p, const float2* V, float2* part, int* lfirst, int batch,
                            hipStream_t s) {
  RecArgs a = make_args_mixed(p);
  a.nchunk = mixed_adj_chunks(p);
  const int Mpad = (p->M + 8 * kMGroup - 1) / (8 * kMGroup) * (8 * kMGroup);
  int MG = mixed_group(p, true);
  // 32 maps per adjoint recursion: the FP16-split column-block kernel only
  if (MG == 32 && !adj_s2p(p)) MG = 16;
  dim3 grid(Mpad * a.nchunk, (batch + MG - 1) / MG);
  static_assert((8 * kMGroup) % (8 * kA_MB) == 0, "Mpad covers whole order groups");
  if (MG == 32)
    hipLaunchKernelGGL(k_adj_s2p, dim3(Mpad * ((a.nchunk + kA_CB - 1) / kA_CB) * kA_CB, (batch + 31) / 32), dim3(512), 0,
                       s, a, V, p->mx_sA, part, lfirst, batch);
  else if (p->spin == 0 && MG == 16 && p->mx_adj_f16)
    hipLaunchKernelGGL(k_adj_mx_s0_cbh, dim3(Mpad * ((a.nchunk + kA_CB - 1) / kA_CB) * kA_CB, grid.y), dim3(256), 0, s, a,
                       V, p->mx_sA, part, batch);
  else if (p->spin == 0 && MG == 16)
    hipLaunchKernelGGL(k_adj_mx_s0_cb, grid, dim3(256), 0, s, a, V, part, batch);
  else if (p->spin == 0)
    hipLaunchKernelGGL(k_adj_mx_s0, grid, dim3(256), 0, s, a, V, part, batch);
  else if (MG == 16 && p->mx_adj_cb && p->mx_adj_f16)
    hipLaunchKernelGGL(k_adj_mx_s2_cbh, dim3(Mpad * ((a.nchunk + kA_CB - 1) / kA_CB) * kA_CB, grid.y), dim3(256), 0, s, a,
                       V, p->mx_sA, part, batch);
  else if (MG == 16 && p->mx_adj_cb)
    hipLaunchKernelGGL(k_adj_mx_s2_cb, grid, dim3(256), 0, s, a, V, part, batch);
  else if (MG == 16)
    hipLaunchKernelGGL((k_adj_mx_s2<4, 2, 4>), grid, dim3(256), 0, s, a, V, part, batch);
  else if (MG == 8 && p->mx_adj_occ == 2)
    hipLaunchKernelGGL((k_adj_mx_s2<2, 2, 4>), grid, dim3(256), 0, s, a, V, part, batch);
  else if (MG == 8)
    hipLaunchKernelGGL((k_adj_mx_s2<2, 1, 4>), grid, dim3(256), 0, s, a, V, part, batch);
  else
    hipLaunchKernelGGL((k_adj_mx_s2<1, 1, 4>), grid, dim3(256), 0, s, a, V, part, batch);
  return hipGetLastError();
}


// ---- ragged groups (host) ----
int group_adj_chunks(const jhtx_plan* p) { return (p->T + 255) / 256; }

int group_rec_build(const GroupLayout& L, void** desc, int* start_synth, int* start_adj) {
  std::vector<GroupRec> h(L.n);
  start_synth[0] = start_adj[0] = 0;
  for (int i = 0; i < L.n; ++i) {
    const jhtx_plan* p = L.plan[i];
    const int Mpad = (p->M + 8 * kMGroup - 1) / (8 * kMGroup) * (8 * kMGroup);
    const int groups = (L.batch[i] + 15) / 16;
    GroupRec& g = h[i];
    g.synth = make_args_mixed(p);
    g.synth.nchunk = (p->T + (p->spin == 2 ? 127 : 255)) / (p->spin == 2 ? 128 : 256);
    g.adj = make_args_mixed(p);
    g.adj.nchunk = group_adj_chunks(p);
    g.f_off = L.f_off[i];
    g.part_off = L.part_off[i];
    g.batch = L.batch[i];
    g.synth_gx = Mpad * g.synth.nchunk;  // multiples of 8: every plan's range starts on XCD 0
    g.adj_gx = Mpad * g.adj.nchunk;
    const long long ns = (long long)start_synth[i] + (long long)g.synth_gx * groups;
    const long long na = (long long)start_adj[i] + (long long)g.adj_gx * groups;
    if (ns > 0x7fffffffLL || na > 0x7fffffffLL) {
      set_error("ragged group: too many work-groups for one launch");
      return JHTX_EINVAL;
    }
    start_synth[i + 1] = (int)ns;
    start_adj[i + 1] = (int)na;
  }
  if (hipMalloc(desc, sizeof(GroupRec) * L.n) != hipSuccess) return JHTX_ENOMEM;
  if (hipMemcpy(*desc, h.data(), sizeof(GroupRec) * L.n, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("ragged group: descriptor upload failed");
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

hipError_t group_rec_launch(bool synth, const GroupLayout& L, const void* desc, GroupCall C, hipStream_t s) {
  const GroupRec* G = reinterpret_cast<const GroupRec*>(desc);
  const dim3 grid((unsigned)C.start[C.n]);
  if (synth) {
    if (L.spin == 0)
      hipLaunchKernelGGL(k_group_synth<0>, grid, dim3(256), 0, s, G, C);
    else
      hipLaunchKernelGGL(k_group_synth<2>, grid, dim3(256), 0, s, G, C);
  } else {
    if (L.spin == 0)
      hipLaunchKernelGGL(k_group_adj<0>, grid, dim3(256), 0, s, G, C);
    else
      hipLaunchKernelGGL(k_group_adj<2>, grid, dim3(256), 0, s, G, C);
  }
  return hipGetLastError();
}

}  // namespace jhtx
