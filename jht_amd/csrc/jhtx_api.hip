// C-ABI entry points and plan construction for the jhtx library (see include/jhtx.h).
//
// Plan creation replaces jht's cached `_prepare` (src/jht/healpix.py:184-357): it builds
// the RING geometry of the north half + equator (healpix.py:107-140, same floating-point
// expressions so the colatitudes are bit-identical), the per-m seeds
// (_recursion.py:53-56, :166-188) and the triangle-packed, rescaled recursion
// coefficients (_recursion.py:37-50, :136-143), then uploads them once.
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <cstdlib>

#include "jhtx_internal.h"

#include <initializer_list>
#include <mutex>

namespace jhtx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

namespace {

template <class T>
int upload(T** dst, const std::vector<T>& src) {
  if (src.empty()) {
    *dst = nullptr;
    return JHTX_OK;
  }
  if (hipMalloc(dst, sizeof(T) * src.size()) != hipSuccess) {
    set_error("hipMalloc failed for a plan table");
    return JHTX_ENOMEM;
  }
  if (hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("hipMemcpy failed for a plan table");
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

// z = cos(theta) of the 0-based north ring t < 2 nside, as RingInfo computes it.
double ring_z(int nside, int t) {
  const long long i = t + 1;
  if (i < nside) {
    const double num = (double)(i * i);
    const double den = 3.0 * (double)((long long)nside * nside);
    return 1.0 - num / den;
  }
  return 4.0 / 3.0 - (2.0 * (double)i) / (3.0 * (double)nside);
}

// ln 2 / 2: the Wigner seed tables are centred on cos(theta/2) = sin(theta/2) = 1/sqrt 2
constexpr long double kHalfLn2 = 0.346573590279972654708616060729088284L;

// Wigner-d single-term seed parameters of d^{lmin}_{M,M'} (_recursion.py:166-188):
// log|d| = log_fac + pc log cos(theta/2) + ps log sin(theta/2).  Near the equator, where
// high-m seeds live, the three terms are ~10^3 each and cancel to ~10^2, so the constant is
// stored centred, log_fac - (pc + ps) ln2/2, and the ring tables hold log cos(theta/2) +
// ln2/2 and log sin(theta/2) + ln2/2 (host_rec_tables_x): the kernels' L = c + pc lc + ps ls
// then adds small terms.  log_fac comes from long-double lgamma (the float64 lgamma sum of
// the reference's formula is off by up to ~1e-11 relative at lmin ~ 4096).
void wigner_seed_params(int M, int Mp, int lmin, double out[4]) {
  const int k0 = std::max(0, M - Mp);
  auto lg = [](long double v) { return lgammal(v); };
  const long double log_fac =
      0.5L * (lg(lmin + M + 1) + lg(lmin - M + 1) + lg(lmin + Mp + 1) + lg(lmin - Mp + 1)) -
      (lg(k0 + 1) + lg(lmin + M - k0 + 1) + lg(lmin - Mp - k0 + 1) + lg(k0 - M + Mp + 1));
  const int pc = 2 * lmin - 2 * k0 + M - Mp;  // power of cos(theta/2)
  const int ps = 2 * k0 - M + Mp;             // power of sin(theta/2)
  out[0] = (double)(log_fac - (long double)(pc + ps) * kHalfLn2);
  out[1] = (double)pc;
  out[2] = (double)ps;
  out[3] = (k0 % 2) ? -1.0 : 1.0;
}

// Host tables of the recursion (no device calls; also exported for CPU validation).
struct HostRec {
  std::vector<double> x, lsin, lcos, alpha, gamma, G, seed;
};

// Tables over the north-half rings x[t] = cos(theta_t), t < T (the last one the equator).
// The log of a factor that is exactly zero (a ring on a pole) is stored as -1e300, which the
// seed split reads as an exact zero seed, while a zero power of it still gives 0.
void host_rec_tables_x(const std::vector<double>& xs, int lmax, int spin, HostRec& h) {
  const int M = lmax + 1, T = (int)xs.size();
  const long long K = (long long)(lmax + 1) * (lmax + 2) / 2;
  std::vector<double>& x = h.x;
  std::vector<double>& lsin = h.lsin;
  std::vector<double>& lcos = h.lcos;
  std::vector<double>& alpha = h.alpha;
  std::vector<double>& gamma = h.gamma;
  std::vector<double>& G = h.G;
  std::vector<double>& seed = h.seed;
  x = xs;
  lsin.assign(T, 0.0);
  for (int t = 0; t < T; ++t) {
    const long double z = x[t];
    if (spin == 0) {
      // sin^2 theta = (1 - z)(1 + z): 1 - z is exact, where 1 - z*z in double loses ~7e-10
      // relative on the ring nearest a pole at nside 2048
      const long double s2 = std::min(std::max((1.0L - z) * (1.0L + z), 0.0L), 1.0L);
      lsin[t] = s2 > 0.0L ? (double)(0.5L * logl(s2)) : -1.0e300;
    }
  }
  if (spin != 0) {
    // log cos(theta/2) + ln2/2 and log sin(theta/2) + ln2/2 in long double (the seed
    // constants are centred to match: wigner_seed_params)
    lcos.resize(T);
    auto centred_log = [](long double v) {
      return v > 0.0L ? (double)(0.5L * logl(v) + kHalfLn2) : -1.0e300;
    };
    for (int t = 0; t < T; ++t) {
      const long double z = x[t];
      lcos[t] = centred_log(std::min(std::max((1.0L + z) / 2.0L, 0.0L), 1.0L));
      lsin[t] = centred_log(std::min(std::max((1.0L - z) / 2.0L, 0.0L), 1.0L));
    }
  }
  alpha.assign(K, 0.0);
  G.assign(K, 0.0);
  if (spin != 0) gamma.assign(K, 0.0);
  const long double four_pi = 4.0L * 3.14159265358979323846264338327950288L;
  if (spin == 0) {
    seed.resize(M);
    long double acc = 0.0L;  // sum_{k<=m} log((2k+1)/(2k))
    for (int m = 0; m < M; ++m) {
      if (m > 0) acc += logl((2.0L * m + 1.0L) / (2.0L * m));
      seed[m] = (double)(-0.5L * logl(four_pi) + 0.5L * acc);
    }
    std::vector<long double> g(lmax + 2);
    for (int m = 0; m < M; ++m) {
      const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
      const long double mm = m;
      g[m] = 1.0L;
      if (m + 1 <= lmax) g[m + 1] = 1.0L;
      for (int l = m + 2; l <= lmax; ++l) {
        const long double L = l;
        const long double b = sqrtl((2 * L + 1) * ((L - 1) * (L - 1) - mm * mm) /
                                    ((2 * L - 3) * (L - mm) * (L + mm)));
        g[l] = b * g[l - 2];
      }
      G[cb + m] = 1.0;
      for (int l = m + 1; l <= lmax; ++l) {
        const long double L = l;
        const long double a = sqrtl((2 * L - 1) * (2 * L + 1) / ((L - mm) * (L + mm)));
        alpha[cb + l] = (double)(a * g[l - 1] / g[l]);
        G[cb + l] = (double)g[l];
      }
    }
  } else {
    // spin-weighted (|s| = 1, 2, 3; on-grid plans use 2): the +|s| and -|s| recursions
    const int sa = spin < 0 ? -spin : spin;
    const long double ss2 = (long double)sa * sa;
    seed.resize(8 * (size_t)M);
    std::vector<long double> g(lmax + 2);
    for (int m = 0; m < M; ++m) {
      const int lmin = std::max(m, sa);
      wigner_seed_params(-m, sa, lmin, &seed[8 * m]);
      wigner_seed_params(-m, -sa, lmin, &seed[8 * m + 4]);
      if (lmin > lmax) continue;
      const long long cb = (long long)m * (2 * lmax + 1 - m) / 2;
      const long double mm = m;
      // d^{l'+1} = (A x + C) d^{l'} - B d^{l'-1}; coefficients of the produced degree l = l'+1
      auto coefs = [&](int l, long double& A, long double& B, long double& C) {
        const long double lp = l - 1;
        const long double D = lp * sqrtl(((lp + 1) * (lp + 1) - mm * mm) * ((lp + 1) * (lp + 1) - ss2));
        A = (2 * lp + 1) * lp * (lp + 1) / D;
        C = (long double)sa * mm * (2 * lp + 1) / D;  // -(2l'+1) M M' / D with M = -m, M' = +|s|
        B = (lp + 1) * sqrtl((lp * lp - mm * mm) * (lp * lp - ss2)) / D;
      };
      g[lmin] = 1.0L;
      if (lmin + 1 <= lmax) g[lmin + 1] = 1.0L;
      for (int l = lmin + 2; l <= lmax; ++l) {
        long double A, B, C;
        coefs(l, A, B, C);
        g[l] = B * g[l - 2];
      }
      const long double sg = (m & 1) ? -1.0L : 1.0L;
      for (int l = lmin; l <= lmax; ++l) {
        const long double pref = sg * sqrtl((2.0L * l + 1.0L) / four_pi);
        G[cb + l] = (double)(g[l] * pref);
        if (l > lmin) {
          long double A, B, C;
          coefs(l, A, B, C);
          alpha[cb + l] = (double)(A * g[l - 1] / g[l]);
          gamma[cb + l] = (double)(C * g[l - 1] / g[l]);
        }
      }
    }
  }
}

// The HEALPix north-half rings t < 2 nside.
void host_rec_tables(int nside, int lmax, int spin, HostRec& h) {
  std::vector<double> xs(2 * nside);
  for (int t = 0; t < 2 * nside; ++t) xs[t] = ring_z(nside, t);
  host_rec_tables_x(xs, lmax, spin, h);
}

int upload_rec_tables(jhtx_plan* p, const HostRec& h) {
  int rc;
  if ((rc = upload(&p->rec.alpha, h.alpha))) return rc;
  if ((rc = upload(&p->rec.G, h.G))) return rc;
  if ((rc = upload(&p->rec.gamma, h.gamma))) return rc;
  if ((rc = upload(&p->rec.x, h.x))) return rc;
  if ((rc = upload(&p->rec.lsin, h.lsin))) return rc;
  if ((rc = upload(&p->rec.lcos, h.lcos))) return rc;
  if ((rc = upload(&p->rec.seed, h.seed))) return rc;
  return JHTX_OK;
}

int build_rec_tables(jhtx_plan* p) {
  HostRec h;
  host_rec_tables(p->nside, p->lmax, p->spin, h);
  // FP16 split of the mixed synthesis: the recursion computes mu_l = lambda_l / G_l with
  // |lambda_l| <= sqrt((2l+1)/4pi), so |mu| <= sqrt((2l+1)/4pi) / |G_l|; the power-of-two
  // scale puts that bound at or below 2^12 (FP16 holds 65504)
  double bound = 0.0, gmax = 0.0;
  for (int m = 0; m <= p->lmax; ++m) {
    const long long cb = (long long)m * (2 * p->lmax + 1 - m) / 2;
    for (int l = m; l <= p->lmax; ++l) {
      const double g = std::fabs(h.G[cb + l]);
      gmax = std::max(gmax, g);
      if (g > 0.0) bound = std::max(bound, std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI)) / g);
    }
  }
  int e = 0;
  std::frexp(bound > 0.0 ? 4096.0 / bound : 1.0, &e);
  p->mx_sA = std::ldexp(1.0, std::min(std::max(e - 1, -120), 120));
  p->mx_gmax = gmax;
  return upload_rec_tables(p, h);
}

void free_rec_tables(jhtx_plan* p) {
  double* ptrs[] = {p->rec.alpha, p->rec.gamma, p->rec.G, p->rec.x,
                    p->rec.lsin,  p->rec.lcos,  p->rec.seed};
  for (double* q : ptrs)
    if (q) hipFree(q);
  p->rec = RecTables();
}

struct Profile {
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
};

constexpr int kStages = 5;  // profiled stages (include/jhtx.h)

// Bracket one stage with events when profiling is on.
struct StageTimer {
  Profile* pr;
  hipStream_t s;
  int stage;
  hipEvent_t a = nullptr;
  StageTimer(const jhtx_plan* p, hipStream_t s_, int st) : pr(nullptr), s(s_), stage(st) {
    if (p->profiling && p->prof && ((Profile*)p->prof)->recs.size() < kStages * 4096) {
      pr = (Profile*)p->prof;
      a = pr->get();
      (void)hipEventRecord(a, s);
    }
  }
  void done() {
    if (!pr) return;
    hipEvent_t b = pr->get();
    (void)hipEventRecord(b, s);
    pr->recs.push_back({stage, a, b});
    pr = nullptr;
  }
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// ---------------------------------------------------------------------------------- //
// Scratch of jhtx_adjoint_noalloc: a file-scope device array, one copy per GPU, held until
// the process ends.  Its name is a host-side stand-in: host code uses only the device
// address hipGetSymbolAddress returns for each device, resolved once per device when a
// plan is created there (plan set-up, with that device current).
// ---------------------------------------------------------------------------------- //
static __device__ __align__(256) unsigned char scratch[JHTX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices];          // device address of scratch per device
hipError_t g_scratch_err[kMaxDevices];     // result of the lookup (hipSuccess once resolved)
bool g_scratch_done[kMaxDevices];

// Called at plan creation with the plan's device current.
void resolve_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_done[dev]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[dev] = e;
  g_scratch_ptr[dev] = (e == hipSuccess) ? ptr : nullptr;
  g_scratch_done[dev] = true;
}

// Bytes of scratch an adjoint call of `batch` maps uses: the ring coefficients V, then the
// recursion's chunk partials (the workspace layout of jhtx_adjoint, sized for this batch).
size_t adjoint_scratch_bytes(const jhtx_plan* p, int batch, size_t* part_offset);

// A one-map call of a mixed plan that runs the FP64 recursion kernels (mx_single).
bool single_route(const jhtx_plan* p, int batch) {
  return p->precision == JHTX_PREC_MIXED && p->mx_single && batch == 1;
}

// The synthesis recursion kernel (stage 0) into F of the precision's type.  scale_ws: the
// workspace's scale region, or null where the call has no workspace (then no FP16 split).
hipError_t launch_synth_stage(const jhtx_plan* p, const void* alm, void* F, int batch, void* scale_ws,
                              hipStream_t s, long long ld = 0) {
  const double2* a = reinterpret_cast<const double2*>(alm);
  if (p->precision != JHTX_PREC_MIXED) return launch_synth_rec(p, a, reinterpret_cast<double2*>(F), batch, s);
  if (single_route(p, batch)) return launch_synth_rec(p, a, reinterpret_cast<float2*>(F), batch, s);
  return launch_synth_mixed(p, a, reinterpret_cast<float2*>(F), batch, scale_ws, s, ld);
}

// The adjoint recursion kernel (stage 3), then the sum of its ring-chunk partials into
// alm (stage 4).  F: ring coefficients of the precision's type; part: the workspace's
// partials.
hipError_t launch_adjoint_rec_finish(const jhtx_plan* p, const void* F, void* part, void* alm, int batch,
                                     hipStream_t s) {
  hipError_t e;
  double2* out = reinterpret_cast<double2*>(alm);
  if (p->precision == JHTX_PREC_MIXED && !single_route(p, batch)) {
    // the first-stored-l table (if the kernel keeps one), then the partials
    const size_t lfb = adj_lfirst_bytes(p);
    int* lfirst = lfb ? reinterpret_cast<int*>(part) : nullptr;
    float2* part32 = reinterpret_cast<float2*>(reinterpret_cast<char*>(part) + lfb);
    StageTimer t3(p, s, 3);
    e = launch_adj_mixed(p, reinterpret_cast<const float2*>(F), part32, lfirst, batch, s);
    t3.done();
    if (e != hipSuccess) return e;
    StageTimer t4(p, s, 4);
    e = launch_adj_finish(p, part32, lfirst, out, batch, mixed_adj_chunks(p), s);
    t4.done();
  } else {
    double2* part64 = reinterpret_cast<double2*>(part);
    StageTimer t3(p, s, 3);
    e = (p->precision == JHTX_PREC_MIXED) ? launch_adj_rec(p, reinterpret_cast<const float2*>(F), part64, batch, s)
                                           : launch_adj_rec(p, reinterpret_cast<const double2*>(F), part64, batch, s);
    t3.done();
    if (e != hipSuccess) return e;
    StageTimer t4(p, s, 4);
    e = launch_adj_finish(p, part64, out, batch, adj_rec_chunks(p), s);
    t4.done();
  }
  return e;
}

size_t adjoint_scratch_bytes(const jhtx_plan* p, int batch, size_t* part_offset) {
  const bool mixed = p->precision == JHTX_PREC_MIXED;
  const size_t f_elem = mixed ? sizeof(float2) : sizeof(double2);
  const size_t v_bytes = align256(f_elem * (size_t)batch * p->ncomp * p->nrings * p->M);
  size_t part;
  if (mixed && !single_route(p, batch))
    part = adj_lfirst_bytes(p) + sizeof(float2) * (size_t)mixed_adj_chunks(p) * batch * p->ncomp * adj_part_len(p);
  else
    part = sizeof(double2) * (size_t)adj_rec_chunks(p) * batch * p->ncomp * p->K;
  if (part_offset) *part_offset = v_bytes;
  return v_bytes + align256(part);
}

int check_call(const jhtx_plan* p, const void* in, void* out, int batch, void* ws) {
  if (!p) {
    set_error("null plan");
    return JHTX_EINVAL;
  }
  if (!in || !out || !ws) {
    set_error("null buffer pointer");
    return JHTX_EINVAL;
  }
  if (batch < 1 || batch > p->max_batch) {
    set_error("batch " + std::to_string(batch) + " outside [1, " + std::to_string(p->max_batch) + "]");
    return JHTX_EINVAL;
  }
  return JHTX_OK;
}

// Kernel shapes (defaults, then an explicit jhtx_tuning) and workspace sizes of a plan whose
// geometry fields (M, T, nrings, ncomp, K) are set.  Nothing but the tuning argument changes
// which kernels a plan runs.
int configure_plan(jhtx_plan* p, const jhtx_tuning* tuning) {
  // rings per lane of the FP64 recursion kernels; small grids use one ring per lane.
  // Defaults from sweeps at nside 2048 / lmax 4096 (synthesis: occupancy wins; adjoint:
  // rings per lane amortise the per-lane ring reduction)
  if (p->spin == 0) {
    p->rec_R = 1;
    // 8 rings per lane amortise the adjoint's cross-lane reduction over twice the work:
    // 9.31 -> 8.15 ms at nside 2048 / lmax 4096 (2 waves per SIMD, 30 spilled registers)
    p->adj_R = p->T >= 1024 ? 8 : 1;
  } else {
    p->rec_R = 1;
    p->adj_R = p->T >= 1024 ? 4 : (p->T >= 512 ? 2 : 1);
  }
  // maps per recursion of the mixed kernels follow the batch the plan is sized for (a partly
  // empty group costs matrix-core time, a larger one amortises the FP64 recursion over more
  // maps); spin 0 can also run 32 maps per recursion, by tuning: at 2 waves per SIMD instead
  // of 3 it measured +3 % at nside 2048 / batch 32 and -2 % on the ragged nside-512 batch
  const int group = p->max_batch >= 16 ? 16 : (p->max_batch >= 5 ? 8 : 4);
  // spin 2: 32 maps per synthesis recursion (the parity-folded FP16 kernel over 256 rings)
  // once the plan holds 32: 101.2 ms per 32 maps against 2 x 59.2 ms at nside 2048 / lmax 4096
  p->mx_synth_group = group;  // raised to 32 below for spin 2 when the plan holds 32 maps
  p->mx_synth_occ = 2;  // register budget of the 8-map kernels, in work-groups per CU
  p->mx_adj_group = group;
  p->mx_adj_occ = 2;
  // the 16-map adjoint in column-block form (1) or with per-wave tiles summed across waves (0)
  p->mx_adj_cb = 1;
  // One-map calls of a mixed plan: a matrix-core group of 4 maps would leave 3/4 of its
  // columns empty, and one 16x16x4 instruction per 64 (ring, l) pairs then costs 4x the
  // two FP64 FMAs per pair of the vector contraction.  Such calls run the FP64 recursion
  // kernels on complex64 F / V instead (the ring FFT stays FP32).
  p->mx_single = 1;
  // 16-map synthesis on the FP16 matrix cores (hi + lo halves of every operand, all four
  // products: half the matrix instructions of FP32): spin 2 77.5 -> 62.1 ms at nside 2048 /
  // lmax 4096, batch 16, error against FP64 4.7e-7 -> 1.0e-6
  p->mx_f16 = 1;
  // the same split in the 16-map column-block adjoint: spin 2 81.0 -> 59.6 ms at nside 2048
  // / lmax 4096, batch 16, error against FP64 1.5e-7 -> 1.6e-7, default on.  Spin 0: 26.4 ->
  // 23.9 ms, but its l >= 2048 band error rises from 2.5e-6 to 6.1e-6: off by default
  p->mx_adj_f16 = p->spin == 2 ? 1 : 0;
  // spin-2 FP16-split synthesis with the N/S parity folded into B: 60.9 -> 59.2 ms at
  // nside 2048 / lmax 4096, batch 16 (3 waves per SIMD instead of 2), and the 32-map form
  p->mx_synth_fold = 1;
  int block_cap = 256;
  if (tuning) {
    auto in = [](int v, std::initializer_list<int> ok) {
      for (int o : ok)
        if (v == o) return true;
      return false;
    };
    const jhtx_tuning& t = *tuning;
    const bool ok =
        (t.synth_rings == 0 || (p->spin == 0 ? in(t.synth_rings, {1, 4, 8}) : in(t.synth_rings, {1, 2, 4}))) &&
        (t.adj_rings == 0 || (p->spin == 0 ? in(t.adj_rings, {1, 4, 8}) : in(t.adj_rings, {1, 2, 4}))) &&
        (t.block == 0 || in(t.block, {64, 128, 192, 256})) &&
        (t.mx_synth_group == 0 || in(t.mx_synth_group, {4, 8, 16, 32})) &&
        (t.mx_synth_fold == 0 || in(t.mx_synth_fold, {1, 2})) &&
        (t.mx_synth_occ == 0 || in(t.mx_synth_occ, {1, 2, 3})) &&
        (t.mx_adj_group == 0 || in(t.mx_adj_group, {4, 8, 16, 32})) && (t.mx_adj_occ == 0 || in(t.mx_adj_occ, {1, 2})) &&
        (t.mx_adj_cb == 0 || in(t.mx_adj_cb, {1, 2})) && (t.mx_single == 0 || in(t.mx_single, {1, 2})) &&
        (t.mx_f16 == 0 || in(t.mx_f16, {1, 2})) && (t.mx_adj_f16 == 0 || in(t.mx_adj_f16, {1, 2}));
    if (!ok) {
      set_error("jhtx_tuning field out of range");
      return JHTX_EINVAL;
    }
    if (t.synth_rings) p->rec_R = t.synth_rings;
    if (t.adj_rings) p->adj_R = t.adj_rings;
    if (t.block) block_cap = t.block;
    if (t.mx_synth_group) p->mx_synth_group = t.mx_synth_group;
    if (t.mx_synth_occ) p->mx_synth_occ = t.mx_synth_occ;
    if (t.mx_adj_group) p->mx_adj_group = t.mx_adj_group;
    if (t.mx_adj_occ) p->mx_adj_occ = t.mx_adj_occ;
    if (t.mx_adj_cb) p->mx_adj_cb = (t.mx_adj_cb == 1);
    if (t.mx_single) p->mx_single = (t.mx_single == 1);
    if (t.mx_f16) p->mx_f16 = (t.mx_f16 == 1);
    if (t.mx_adj_f16) p->mx_adj_f16 = (t.mx_adj_f16 == 1);
    if (t.mx_synth_fold) p->mx_synth_fold = (t.mx_synth_fold == 1);
    // 32 maps per recursion: spin 0's tile kernel, or spin 2's parity-folded FP16 kernel
    if (p->mx_synth_group == 32 && p->spin == 2 && !(p->mx_synth_fold && p->mx_f16)) {
      set_error("jhtx_tuning: 32 maps per spin-2 synthesis recursion needs mx_synth_fold = 1 and mx_f16 = 1");
      return JHTX_EINVAL;
    }
  }
  if (p->spin == 2 && p->max_batch >= 32 && p->mx_synth_fold && p->mx_f16 && !(tuning && tuning->mx_synth_group))
    p->mx_synth_group = 32;
  // spin 2: 32 maps per adjoint recursion (one sign per wave, double-buffered tiles; the
  // bits of the 16-map kernel) once the plan holds 32: 111.8 -> 107.3 ms per 32 maps at
  // nside 2048 / lmax 4096
  if (p->spin == 2 && p->max_batch >= 32 && p->mx_adj_cb && p->mx_adj_f16 && !(tuning && tuning->mx_adj_group))
    p->mx_adj_group = 32;
  p->block_cap = block_cap;
  auto shape = [&](int R, int& block, int& nchunk) {
    const int lanes = (p->T + R - 1) / R;
    block = std::min(block_cap, std::max(64, ((lanes + 63) / 64) * 64));
    nchunk = (p->T + block * R - 1) / (block * R);
  };
  shape(p->rec_R, p->rec_block, p->nchunk);
  shape(p->adj_R, p->adj_block, p->adj_nchunk);
  // ring coefficients F / V: complex128, or complex64 for the mixed path (FP32 ring FFT)
  const size_t f_elem = (p->precision == JHTX_PREC_MIXED) ? sizeof(float2) : sizeof(double2);
  // the mixed plans' F rows padded for the 32-map synthesis (synth_ring_ld)
  const long long fld = (p->precision == JHTX_PREC_MIXED) ? jhtx::padded_ld(p->M) : p->M;
  p->ws_F_bytes = align256(f_elem * (size_t)p->max_batch * p->ncomp * p->nrings * fld);
  // the mixed kernels use one ring per lane and 256-lane work-groups (a mixed plan's one-map
  // calls through the FP64 kernels keep FP64 partials)
  size_t part_bytes;
  if (p->precision == JHTX_PREC_MIXED) {
    part_bytes = jhtx::adj_lfirst_bytes(p) + sizeof(float2) * (size_t)jhtx::mixed_adj_chunks(p) * p->max_batch *
                                                 p->ncomp * jhtx::adj_part_len(p);
    if (p->mx_single) part_bytes = std::max<size_t>(part_bytes, sizeof(double2) * (size_t)p->adj_nchunk * p->ncomp * p->K);
  } else {
    part_bytes = sizeof(double2) * (size_t)p->adj_nchunk * p->max_batch * p->ncomp * p->K;
  }
  p->ws_part_bytes = align256(part_bytes);
  p->ws_scale_bytes = (p->precision == JHTX_PREC_MIXED)
                          ? align256((sizeof(double) * jhtx::kScaleBlocks + sizeof(float)) * (size_t)p->max_batch)
                          : 0;
  p->ws_bytes = p->ws_F_bytes + p->ws_part_bytes + p->ws_scale_bytes;
  return JHTX_OK;
}

}  // namespace
}  // namespace jhtx

using namespace jhtx;

extern "C" {

const char* jhtx_last_error(void) { return g_err.c_str(); }

int jhtx_plan_create(int nside, int lmax, int spin, int precision, int max_batch, int flags,
                     jhtx_plan** out) {
  return jhtx_plan_create_tuned(nside, lmax, spin, precision, max_batch, flags, nullptr, out);
}

int jhtx_plan_get_tuning(const jhtx_plan* p, jhtx_tuning* t) {
  if (!p || !t) {
    set_error("null plan or tuning pointer");
    return JHTX_EINVAL;
  }
  t->synth_rings = p->rec_R;
  t->adj_rings = p->adj_R;
  t->block = p->block_cap;
  t->mx_synth_group = p->mx_synth_group;
  t->mx_synth_occ = p->mx_synth_occ;
  t->mx_adj_group = p->mx_adj_group;
  t->mx_adj_occ = p->mx_adj_occ;
  t->mx_single = p->mx_single ? 1 : 2;
  t->mx_f16 = p->mx_f16 ? 1 : 2;
  t->mx_adj_f16 = p->mx_adj_f16 ? 1 : 2;
  t->mx_adj_cb = p->mx_adj_cb ? 1 : 2;
  t->mx_synth_fold = p->mx_synth_fold ? 1 : 2;
  return JHTX_OK;
}

int jhtx_plan_create_tuned(int nside, int lmax, int spin, int precision, int max_batch, int flags,
                           const jhtx_tuning* tuning, jhtx_plan** out) {
  g_err.clear();
  if (!out) {
    set_error("null plan output pointer");
    return JHTX_EINVAL;
  }
  *out = nullptr;
  if (spin != 0 && spin != 2) {
    set_error("spin=" + std::to_string(spin) + " unsupported (only 0 and 2)");
    return JHTX_EUNSUPPORTED;
  }
  if (nside < 1 || nside > 8192 || lmax < 0 || lmax > 32767 || max_batch < 1) {
    set_error("invalid nside/lmax/max_batch");
    return JHTX_EINVAL;
  }
  if (precision != JHTX_PREC_F64 && precision != JHTX_PREC_MIXED) {
    set_error("precision " + std::to_string(precision) + " unknown");
    return JHTX_EINVAL;
  }
  jhtx_plan* p = new jhtx_plan();  // value-initialised: every field starts at zero / null
  (void)hipGetDevice(&p->device);
  resolve_scratch(p->device);
  p->nside = nside;
  p->lmax = lmax;
  p->spin = spin;
  p->precision = precision;
  p->max_batch = max_batch;
  p->flags = flags;
  p->M = lmax + 1;
  p->T = 2 * nside;
  p->nrings = 4 * nside - 1;
  p->ncomp = spin ? 2 : 1;
  p->npix = 12LL * nside * nside;
  p->K = (long long)(lmax + 1) * (lmax + 2) / 2;
  int rc = configure_plan(p, tuning);
  if (rc != JHTX_OK) {
    delete p;
    return rc;
  }
  rc = build_rec_tables(p);
  if (rc == JHTX_OK) rc = build_fft_tables(p);
  if (rc != JHTX_OK) {
    free_rec_tables(p);
    free_fft_tables(p);
    delete p;
    return rc;
  }
  *out = p;
  return JHTX_OK;
}

int jhtx_plan_create_rings(int nrings, const double* theta, int lmax, int spin, int max_batch,
                           jhtx_plan** out) {
  g_err.clear();
  if (!out) {
    set_error("null plan output pointer");
    return JHTX_EINVAL;
  }
  *out = nullptr;
  if (spin < 0 || spin > 3) {
    set_error("spin=" + std::to_string(spin) + " unsupported (0, 1, 2, 3)");
    return JHTX_EUNSUPPORTED;
  }
  if (!theta || nrings < 1 || nrings % 2 == 0 || nrings > 65535 || lmax < 0 || lmax > 32767 ||
      max_batch < 1) {
    set_error("invalid ring set (nrings must be odd, 1..65535) or lmax/max_batch");
    return JHTX_EINVAL;
  }
  const int T = (nrings + 1) / 2;
  const double half_pi = 1.57079632679489661923;
  for (int t = 0; t < T; ++t) {
    const bool in_range = theta[t] >= 0.0 && theta[t] <= half_pi + 1e-12;
    const bool sorted = t == 0 || theta[t] >= theta[t - 1];
    if (!in_range || !sorted) {
      set_error("theta must be nondecreasing in [0, pi/2]");
      return JHTX_EINVAL;
    }
  }
  if (std::fabs(theta[T - 1] - half_pi) > 1e-12) {
    set_error("the last north-half ring must be the equator (theta = pi/2)");
    return JHTX_EINVAL;
  }
  jhtx_plan* p = new jhtx_plan();
  (void)hipGetDevice(&p->device);
  p->ring_set = 1;
  p->lmax = lmax;
  p->spin = spin;
  p->precision = JHTX_PREC_F64;
  p->max_batch = max_batch;
  p->M = lmax + 1;
  p->T = T;
  p->nrings = nrings;
  p->ncomp = spin ? 2 : 1;
  p->K = (long long)(lmax + 1) * (lmax + 2) / 2;
  int rc = configure_plan(p, nullptr);
  if (rc != JHTX_OK) {
    delete p;
    return rc;
  }
  // the caller owns the ring coefficients: the workspace holds the adjoint's partials only
  p->ws_F_bytes = 0;
  p->ws_bytes = p->ws_part_bytes;
  std::vector<double> xs(T);
  for (int t = 0; t < T; ++t) xs[t] = std::cos(theta[t]);
  HostRec h;
  host_rec_tables_x(xs, lmax, spin, h);
  rc = upload_rec_tables(p, h);
  if (rc != JHTX_OK) {
    free_rec_tables(p);
    delete p;
    return rc;
  }
  *out = p;
  return JHTX_OK;
}

int jhtx_legendre_synth(jhtx_plan* p, const void* alm, void* F, int batch, void* stream) {
  static char dummy;
  int rc = check_call(p, alm, F, batch, &dummy);
  if (rc) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  StageTimer t0(p, s, 0);
  hipError_t e = launch_synth_stage(p, alm, F, batch, nullptr, s);
  t0.done();
  if (e != hipSuccess) {
    set_error(std::string("Legendre synthesis launch failed: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

int jhtx_legendre_adjoint(jhtx_plan* p, const void* F, void* alm, int batch, void* ws, void* stream) {
  int rc = check_call(p, F, alm, batch, ws);
  if (rc) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  void* part = reinterpret_cast<char*>(ws) + p->ws_F_bytes;
  hipError_t e = launch_adjoint_rec_finish(p, F, part, alm, batch, s);
  if (e != hipSuccess) {
    set_error(std::string("Legendre adjoint launch failed: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

size_t jhtx_plan_workspace_bytes(const jhtx_plan* p) { return p ? p->ws_bytes : 0; }

int jhtx_synth(jhtx_plan* p, const void* alm, void* map, int batch, void* ws, void* stream) {
  int rc = check_call(p, alm, map, batch, ws);
  if (rc) return rc;
  if (p->ring_set) {
    set_error("plan has no HEALPix geometry (a ring-set plan): use jhtx_legendre_synth");
    return JHTX_EINVAL;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  void* F = ws;
  void* scale_ws = reinterpret_cast<char*>(ws) + p->ws_F_bytes + p->ws_part_bytes;
  const long long ld = synth_ring_ld(p, batch);  // F's row stride, the same for both stages
  StageTimer t0(p, s, 0);
  hipError_t e = launch_synth_stage(p, alm, F, batch, scale_ws, s, ld);
  t0.done();
  if (e == hipSuccess) {
    StageTimer t1(p, s, 1);
    e = launch_ring_synth(p, F, reinterpret_cast<double*>(map), batch, s, ld);
    t1.done();
  }
  if (e != hipSuccess) {
    set_error(std::string("synthesis launch failed: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

int jhtx_adjoint(jhtx_plan* p, const void* map, void* alm, int batch, void* ws, void* stream) {
  int rc = check_call(p, map, alm, batch, ws);
  if (rc) return rc;
  if (p->ring_set) {
    set_error("plan has no HEALPix geometry (a ring-set plan): use jhtx_legendre_adjoint");
    return JHTX_EINVAL;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double2* V = reinterpret_cast<double2*>(ws);
  double2* part = reinterpret_cast<double2*>(reinterpret_cast<char*>(ws) + p->ws_F_bytes);
  StageTimer t2(p, s, 2);
  hipError_t e = launch_ring_adjoint(p, reinterpret_cast<const double*>(map), V, batch, s);
  t2.done();
  if (e == hipSuccess) {
    e = launch_adjoint_rec_finish(p, V, part, alm, batch, s);
  }
  if (e != hipSuccess) {
    set_error(std::string("adjoint launch failed: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

int jhtx_adjoint_noalloc(jhtx_plan* p, const void* map, void* alm, int batch, void* stream) {
  static char dummy;
  int rc = check_call(p, map, alm, batch, &dummy);
  if (rc) return rc;
  if (p->ring_set) {
    set_error("plan has no HEALPix geometry (a ring-set plan): use jhtx_legendre_adjoint");
    return JHTX_EINVAL;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != p->device) {
    set_error("jhtx_adjoint_noalloc: the plan's device (" + std::to_string(p->device) + ") is not current");
    return JHTX_EINVAL;
  }
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (dev < kMaxDevices && g_scratch_done[dev] && g_scratch_err[dev] == hipSuccess) base = g_scratch_ptr[dev];
    else if (dev < kMaxDevices && g_scratch_done[dev]) {
      set_error(std::string("jhtx_adjoint_noalloc: scratch address unavailable: ") +
                hipGetErrorString(g_scratch_err[dev]));
      return JHTX_EDEVICE;
    }
  }
  if (!base) {
    set_error("jhtx_adjoint_noalloc: no scratch resolved for device " + std::to_string(dev));
    return JHTX_EDEVICE;
  }
  size_t part_off = 0;
  const size_t need = adjoint_scratch_bytes(p, batch, &part_off);
  if (need > (size_t)JHTX_NOALLOC_SCRATCH_BYTES) {
    set_error("jhtx_adjoint_noalloc: this call needs " + std::to_string(need) + " bytes of scratch, above the " +
              std::to_string((size_t)JHTX_NOALLOC_SCRATCH_BYTES) +
              "-byte maximum (use jhtx_adjoint with a workspace, or fewer maps per call)");
    return JHTX_EINVAL;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  void* V = base;
  void* part = reinterpret_cast<char*>(base) + part_off;
  StageTimer t2(p, s, 2);
  hipError_t e = launch_ring_adjoint(p, reinterpret_cast<const double*>(map), V, batch, s);
  t2.done();
  if (e == hipSuccess) e = launch_adjoint_rec_finish(p, V, part, alm, batch, s);
  if (e != hipSuccess) {
    set_error(std::string("adjoint launch failed: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

int jhtx_host_rec_tables(int nside, int lmax, int spin, double* x, double* lsin, double* lcos,
                         double* alpha, double* gamma, double* G, double* seed) {
  if (nside < 1 || lmax < 0 || (spin != 0 && spin != 2)) {
    set_error("invalid nside/lmax/spin");
    return JHTX_EINVAL;
  }
  HostRec h;
  host_rec_tables(nside, lmax, spin, h);
  auto put = [](double* dst, const std::vector<double>& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), sizeof(double) * v.size());
  };
  put(x, h.x);
  put(lsin, h.lsin);
  put(lcos, h.lcos);
  put(alpha, h.alpha);
  put(gamma, h.gamma);
  put(G, h.G);
  put(seed, h.seed);
  return JHTX_OK;
}

int jhtx_plan_set_profiling(jhtx_plan* p, int enable) {
  if (!p) {
    set_error("null plan");
    return JHTX_EINVAL;
  }
  if (enable && !p->prof) p->prof = new Profile();
  p->profiling = enable ? 1 : 0;
  return JHTX_OK;
}

int jhtx_plan_read_profile(jhtx_plan* p, double* ms, int* counts) {
  if (!p || !ms || !counts) {
    set_error("null argument");
    return JHTX_EINVAL;
  }
  for (int i = 0; i < kStages; ++i) {
    ms[i] = 0.0;
    counts[i] = 0;
  }
  Profile* pr = (Profile*)p->prof;
  if (!pr) return JHTX_OK;
  for (auto& r : pr->recs) {
    if (hipEventSynchronize(r.b) != hipSuccess) {
      set_error("hipEventSynchronize failed");
      return JHTX_EDEVICE;
    }
    float t = 0.f;
    (void)hipEventElapsedTime(&t, r.a, r.b);
    ms[r.stage] += t;
    counts[r.stage] += 1;
    pr->pool.push_back(r.a);
    pr->pool.push_back(r.b);
  }
  pr->recs.clear();
  return JHTX_OK;
}

void jhtx_plan_destroy(jhtx_plan* p) {
  if (!p) return;
  if (p->prof) {
    Profile* pr = (Profile*)p->prof;
    for (auto& r : pr->recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pr->pool) (void)hipEventDestroy(e);
    delete pr;
  }
  free_rec_tables(p);
  free_fft_tables(p);
  delete p;
}

int jhtx_plan_query(const jhtx_plan* p, int what) {
  if (!p) return -1;
  switch (what) {
    case 0: return p->nside;
    case 1: return p->lmax;
    case 2: return p->spin;
    case 3: return p->precision;
    case 4: return p->max_batch;
    case 5: return p->ncomp;
    case 6: return p->nrings;
    case 7: return (int)(p->npix & 0x7fffffff);
    case 8: return p->fft.ncls;
    case 9: return p->ring_set;
    default: return -1;
  }
}

}  // extern "C"
