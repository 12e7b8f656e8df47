// Internal definitions shared by the jhtx HIP sources (plan layout, kernel argument
// blocks, launchers).  Not part of the public C ABI (include/jhtx.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <string>

#include "../../include/jhtx.h"

namespace jhtx {

// Exponent-tracking constants of the on-the-fly recursion.  A recursion value is
// stored as mu * 2^(kScaleStep * ex) with ex <= 0; while ex < 0 the true value is below
// 2^-256 and its contribution is dropped (it is below double precision of any output).
constexpr double kBig = 0x1p256;
constexpr double kRescale = 0x1p-512;
constexpr int kScaleStep = 512;

// One rescale check of a recursion pair: once the value has climbed past 2^256 while its
// exponent is still negative, both terms are scaled by 2^-512 and the exponent advances.
// Branch-free (lanes that stay put are multiplied by exactly 1.0), so divergent lanes cost
// no exec-mask juggling inside the unrolled recursion.
__device__ __forceinline__ void rescale_pair(double& mu, double& mup, int& ex) {
  const bool up = (ex < 0) & (__builtin_fabs(mu) > kBig);
  const double f = up ? kRescale : 1.0;
  mu *= f;
  mup *= f;
  ex += up ? 1 : 0;
}
// consecutive orders m per XCD group in the recursion kernels' work-group order
constexpr int kMGroup = 16;

// Partial layout of the 32-map spin-2 adjoint: order m holds ntile(m) = (lmax - lmin + 32) / 32
// tiles of 32 consecutive l from lmin = max(m, s), each tile 256-byte aligned (32 complex64),
// so that a work-group's tile is written as whole lines.  tile_base(m) is the number of tiles
// of the orders below m, in closed form: sum_{x=lo}^{hi} floor(x / 32) = S(hi) - S(lo - 1)
// with S(x) = 32 q (q - 1) / 2 + (r + 1) q for x = 32 q + r.
__host__ __device__ inline long long floor32_prefix(long long x) {
  if (x < 0) return 0;
  const long long q = x / 32, r = x % 32;
  return 32 * (q * (q - 1) / 2) + (r + 1) * q;
}
__host__ __device__ inline long long tile_base(int m, int lmax, int s) {
  long long t = (long long)(m < s ? m : s) * ((lmax - s + 32) / 32);  // orders below s start at l = s
  if (m > s) t += floor32_prefix(lmax + 32 - s) - floor32_prefix(lmax + 32 - m);
  return t;
}
// complex64 entries per (chunk, map, component) partial vector of that layout
__host__ __device__ inline long long tile_part_len(int lmax, int s) { return 32 * tile_base(lmax + 1, lmax, s); }

// Per-plan device tables for the recursion kernels.
struct RecTables {
  // triangle-packed by the healpy index idx(l, m) = m(2 lmax + 1 - m)/2 + l
  double* alpha = nullptr;  // mu_l = (alpha_l x + gamma_l) mu_{l-1} - mu_{l-2}
  double* gamma = nullptr;  // spin 2 only (the -2 recursion uses -gamma)
  double* G = nullptr;      // lambda_l = G_l mu_l
  // north half + equator rings, t < T = 2 nside
  double* x = nullptr;      // cos(theta_t)
  double* lsin = nullptr;   // spin 0: log sin(theta); spin 2: log sin(theta/2)
  double* lcos = nullptr;   // spin 2: log cos(theta/2)
  // per-m seed parameters: spin 0 -> 1 double (theta-independent log|lambda_mm|);
  // spin 2 -> 8 doubles (log_fac, pow_cos, pow_sin, sign) for s = +2 then s = -2
  double* seed = nullptr;
};

// Argument block for the recursion kernels (passed by value).
struct RecArgs {
  const double* alpha;
  const double* gamma;
  const double* G;
  const double* x;
  const double* lsin;
  const double* lcos;
  const double* seed;
  int lmax, M, T, nside, nrings, nchunk;
  long long K;
  int sabs;      // |spin| (the spin-weighted recursions start at l = max(m, |spin|))
  double ssign;  // (-1)^spin: the -spin channel's factor, -(-1)^s (E - iB)
  long long Mf;  // row stride of F: M, or synth_ring_ld() on the 32-map synthesis's path
};

// Ring-FFT launch class: ring pairs whose transform length fits one work-group shape.
struct FftClass {
  int lc = 0, kind = 0;    // length class and transform kind (0 direct, 1 pow2, 2 chirp-z)
  int block = 64;          // threads per work-group
  int lds_bytes = 0;       // dynamic LDS of the ring kernels (bytes)
  int lds64_bytes = 0;     // the same for the FP64 plan-time spectrum kernel
  int npairs = 0;
  int* pairs = nullptr;    // device list of ring-pair indices p (0 <= p < 2 nside)
};

struct FftTables {
  double2* tw = nullptr;  // exp(-2 pi i k / lmax_fft), k < lmax_fft
  float2* twf = nullptr;  // FP32 copies of tw, bspec and chirps (mixed-precision plans)
  float2* bspecf = nullptr;
  float2* chirpsf = nullptr;
  int lmax_fft = 0;       // twiddle table length: the largest power-of-two transform length,
                          // at least 2N for a shifted power-of-two ring of N pixels
  FftClass cls[12];
  int ncls = 0;
  long long* bofs = nullptr;  // device, per ring pair: offset of its Bluestein kernel spectrum or -1
  double2* bspec = nullptr;   // device, concatenated kernel spectra
  long long bspec_len = 0;
  long long* cofs = nullptr;  // device, per ring pair: offset of its chirp table or -1
  // device, concatenated chirp tables e^{i pi n^2 / N}, then (from offset chirp_shift) the
  // same tables with the half-sample ring shift folded in, e^{i pi (n^2 + n) / N}
  double2* chirps = nullptr;
  long long chirp_shift = 0;
  // polar rings of mixed-precision spin-2 plans transformed as four length-N/4 chirp-z
  // transforms: launch classes for Lq = 2048 and 4096, per ring pair the offset of its tables
  // in qtab (-1 if it is not split)
  FftClass qcls[2];
  long long* qofs = nullptr;
  float2* qtab = nullptr;
};

}  // namespace jhtx

struct jhtx_plan {
  int nside, lmax, spin, precision, max_batch, flags;
  int device;    // the HIP device current at creation (its tables live there)
  int ring_set;  // 1: an arbitrary symmetric ring set (jhtx_plan_create_rings), no ring FFT
  int M, T, nrings, ncomp, nchunk;
  long long npix, K;
  int rec_block;  // threads per recursion work-group (synthesis)
  int rec_R;      // rings per lane (synthesis)
  int adj_block, adj_R, adj_nchunk;  // the same for the adjoint recursion
  int block_cap;                     // work-group size cap of the FP64 recursion kernels
  int mx_synth_group;  // maps per work-group of the mixed spin-2 synthesis (4 or 8)
  int mx_synth_occ;    // register budget of the 8-map synthesis: 1 or 2 work-groups per CU
  int mx_adj_group;    // the same for the mixed spin-2 adjoint
  int mx_adj_occ;      // register budget of the 8-map adjoint: 1 or 2 work-groups per CU
  int mx_adj_cb;       // 16-map adjoint in column-block form (0 or 1)
  int mx_single;       // mixed plans: one-map calls run the FP64 vector-FMA recursion kernels
                       // on complex64 F / V (0 or 1)
  int mx_f16;          // 16-map synthesis on the FP16 matrix cores, operands split
                       // into hi + lo halves (0 or 1)
  int mx_adj_f16;      // the same for the 16-map column-block adjoints (0 or 1)
  int mx_synth_fold;   // spin-2 FP16-split synthesis with the N/S parity folded into B (0 or 1)
  double mx_sA;        // power-of-two scale of the recursion values for the FP16 split
  double mx_gmax;      // max |G| over the triangle (bounds the staged a_lm G)
  jhtx::RecTables rec;
  jhtx::FftTables fft;
  // stage timing (jhtx_plan_set_profiling)
  int profiling;
  void* prof;  // jhtx::Profile*
  // workspace layout (bytes): [F or V: batch*ncomp*nrings*M double2 (FP64) / float2 (mixed)]
  // [chunk partials: nchunk*batch*ncomp*K double2 (FP64 kernels) or float2 (mixed kernels)]
  // [scales: kScaleBlocks*batch double partial maxima, then batch float scales (FP16 split)]
  size_t ws_F_bytes, ws_part_bytes, ws_scale_bytes, ws_bytes;
};

namespace jhtx {

void set_error(const std::string& msg);

// launchers (return hipError_t)
hipError_t launch_synth_rec(const jhtx_plan* p, const double2* alm, double2* F, int batch,
                            hipStream_t s);
hipError_t launch_synth_rec(const jhtx_plan* p, const double2* alm, float2* F, int batch,
                            hipStream_t s);
hipError_t launch_adj_rec(const jhtx_plan* p, const double2* V, double2* part, int batch, hipStream_t s);
hipError_t launch_adj_rec(const jhtx_plan* p, const float2* V, double2* part, int batch, hipStream_t s);
int adj_rec_chunks(const jhtx_plan* p);
hipError_t launch_adj_finish(const jhtx_plan* p, const double2* part, double2* alm, int batch,
                             int nchunk, hipStream_t s);
// lfirst: the mixed adjoint's first-stored-l table (or null), adj_lfirst_bytes() before the partials
hipError_t launch_adj_finish(const jhtx_plan* p, const float2* part, const int* lfirst, double2* alm, int batch,
                             int nchunk, hipStream_t s);
// ld: F's row stride (0: M)
hipError_t launch_synth_mixed(const jhtx_plan* p, const double2* alm, float2* F, int batch,
                              void* scale_ws, hipStream_t s, long long ld = 0);
constexpr int kScaleBlocks = 256;  // partial maxima per map of the FP16 split's scale pass
hipError_t launch_adj_mixed(const jhtx_plan* p, const float2* V, float2* part, int* lfirst, int batch,
                            hipStream_t s);
// Row stride of the ring coefficients F in the workspace of a jhtx_synth call.  The 32-map
// spin-2 synthesis writes F one order m per work-group, groups of 8 consecutive m per XCD:
// with rows rounded up to 16 entries (128 bytes) each group's columns are one aligned
// 64-byte run of a row and never straddle two lines.  M everywhere else
// (jhtx_legendre_synth keeps the documented stride M).
long long synth_ring_ld(const jhtx_plan* p, int batch);
inline long long padded_ld(long long M) { return (M + 15) / 16 * 16; }
// bytes of the mixed adjoint's first-stored-l table (0 where its kernel stores every partial);
// it sits at the start of the partials region
size_t adj_lfirst_bytes(const jhtx_plan* p);
// complex64 entries per (chunk, map, component) of the mixed adjoint's partials: the tile
// layout's length for the 32-map spin-2 adjoint, K otherwise
long long adj_part_len(const jhtx_plan* p);
// ring chunks (partials per map) of the mixed adjoint
int mixed_adj_chunks(const jhtx_plan* p);
// F / V: complex128 (FP64 plans) or complex64 (mixed-precision plans)
// ld: the row stride of F (0: M)
hipError_t launch_ring_synth(const jhtx_plan* p, const void* F, double* map, int batch,
                             hipStream_t s, long long ld = 0);
hipError_t launch_ring_adjoint(const jhtx_plan* p, const double* map, void* V, int batch,
                               hipStream_t s);

// FFT plan building (host): fills p->fft
int build_fft_tables(jhtx_plan* p);
void free_fft_tables(jhtx_plan* p);

// ---- Ragged groups: the transforms of several plans, one kernel launch per stage ----
// A group call runs five stages (synthesis recursion, synthesis ring FFT; adjoint ring FFT,
// adjoint recursion, chunk reduction), each as ONE launch whose work-groups cover every
// plan: work-group w belongs to plan i when start[i] <= w < start[i + 1], and its index
// within the plan's range plays the role of the plan's own launch coordinates.  Ranges of
// the recursion stages start at multiples of 8, so the XCD-aware order inside each plan
// (work-group & 7 = XCD) is kept.
constexpr int kGroupMax = 16;   // plans per group
constexpr int kGroupNside = 512;  // largest nside: every ring transform fits 256 x 16 elements

// host-side layout of a group, in launch order (plans sorted by decreasing lmax)
struct GroupLayout {
  int n = 0, spin = 0;
  const jhtx_plan* plan[kGroupMax] = {};
  int batch[kGroupMax] = {};
  long long f_off[kGroupMax] = {};     // bytes: the plan's F / V in the group workspace
  long long part_off[kGroupMax] = {};  // bytes: its adjoint chunk partials
};

// Kernel argument of every group launch (by value): the stage's work-group ranges and the
// call's buffers, per plan in launch order.
struct GroupCall {
  int n;
  int start[kGroupMax + 1];
  const void* in[kGroupMax];
  void* out[kGroupMax];
  char* ws;
};

// Plan of work-group w of a group launch, and w's index within that plan's range.
__device__ __forceinline__ int group_member(const GroupCall& C, int& local) {
  const int w = blockIdx.x;
  int i = 0;
  while (i + 1 < C.n && w >= C.start[i + 1]) ++i;
  local = w - C.start[i];
  return i;
}

// Stage builders (upload per-plan descriptors, fill the stage's work-group ranges) and
// launchers; *desc is a device array freed with hipFree.
int group_rec_build(const GroupLayout& L, void** desc, int* start_synth, int* start_adj);
hipError_t group_rec_launch(bool synth, const GroupLayout& L, const void* desc, GroupCall C, hipStream_t s);
int group_fft_build(const GroupLayout& L, void** desc, int* start, int* lds_bytes);
hipError_t group_fft_launch(bool synth, const GroupLayout& L, const void* desc, GroupCall C, int lds_bytes,
                            hipStream_t s);
int group_fin_build(const GroupLayout& L, void** desc, int* start);
hipError_t group_fin_launch(const void* desc, GroupCall C, hipStream_t s);
// ring chunks of the group adjoint recursion (256 rings each) and of the synthesis
int group_adj_chunks(const jhtx_plan* p);

// ring geometry helpers (host and device)
__host__ __device__ inline int ring_len(int nside, int r /*0-based*/) {
  const int i = r + 1;
  if (i < nside) return 4 * i;
  if (i > 3 * nside) return 4 * (4 * nside - i);
  return 4 * nside;
}
// true when the ring's first pixel sits at phi0 = pi / N (half-sample shift)
__host__ __device__ inline bool ring_shifted(int nside, int r) {
  const int i = r + 1;
  if (i < nside || i > 3 * nside) return true;
  return ((i - nside) % 2) == 0;
}
__host__ __device__ inline long long ring_start(int nside, int r) {
  const long long i = r + 1;
  const long long ns = nside;
  if (i < ns) return 2 * i * (i - 1);
  if (i <= 3 * ns) return 2 * ns * (ns - 1) + (i - ns) * 4 * ns;
  const long long j = 4 * ns - i;  // mirror of north cap ring j
  return 12 * ns * ns - 2 * j * (j + 1);
}

}  // namespace jhtx
