/*
 * jhtx -- MI355X (gfx950) HEALPix on-grid spherical-harmonic synthesis and its exact
 * adjoint, spin 0 and spin 2, behind a plain C ABI (no torch or HIP types beyond an
 * opaque stream handle).
 *
 * Conventions are those of jrcheshire/jht v0.2.0:
 *   - a_lm in healpy m-major packing, idx(l,m) = m(2 lmax + 1 - m)/2 + l,
 *     K = (lmax+1)(lmax+2)/2 complex128 (interleaved re, im) per component;
 *   - orthonormal Y_lm with the Condon-Shortley phase, map = sum a_lm Y_lm;
 *   - spin 2: input (E, B), output (Q, U) in the HEALPix/COSMO convention;
 *   - maps are RING ordered float64, npix = 12 nside^2 per component.
 *
 * Buffers are caller-owned DEVICE pointers.  A call never allocates: the caller passes a
 * workspace of jhtx_plan_workspace_bytes(plan) bytes.  Calls are ordered on the given
 * stream.  A plan is immutable after creation and may be used concurrently from several
 * streams provided each call gets its own workspace.  Every function returns 0 on success
 * or a negative JHTX_E* code; jhtx_last_error() then describes the failure (thread-local).
 */
#ifndef JHTX_H
#define JHTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jhtx_plan jhtx_plan;

enum {
  JHTX_OK = 0,
  JHTX_EINVAL = -1,       /* bad argument (nside, lmax, spin, batch, null pointer) */
  JHTX_EUNSUPPORTED = -2, /* spin not in {0, 2} (reference raises NotImplementedError) */
  JHTX_EDEVICE = -3,      /* a HIP runtime call failed */
  JHTX_ENOMEM = -4        /* plan tables could not be allocated */
};

enum {
  JHTX_PREC_F64 = 0,   /* FP64 recursion and FP64 contraction: <= 1e-12 vs float64 jht */
  JHTX_PREC_MIXED = 1  /* FP64 recursion, FP32 contraction on MFMA (batched) */
};

/* Flags for jhtx_plan_create. */
enum {
  JHTX_FLAG_BLUESTEIN_ALL = 1 /* route every ring DFT through the chirp-z path */
};

/* Replaces jht's cached `_prepare(nside, lmax, spin, fft_mode)` closure pair
 * (src/jht/healpix.py:184-357): builds the geometry, the triangle-packed recursion
 * coefficients and the seed tables once, on the device.  max_batch bounds the `batch`
 * argument of the calls (the workspace is sized for it). */
int jhtx_plan_create(int nside, int lmax, int spin, int precision, int max_batch, int flags,
                     jhtx_plan** plan);

/* Kernel shape choices of a plan (no counterpart in the reference: jht leaves tiling to
 * XLA).  A zero field keeps the library's default; jht_amd's autotuner times candidates
 * and passes the winner here.  This struct is the only way to change a plan's kernels: the
 * library reads no environment variables.  Allowed values:
 *   synth_rings, adj_rings   rings per lane of the FP64 recursion kernels: spin 0 {1,4,8},
 *                            spin 2 {1,2,4}
 *   block                    FP64 recursion work-group size cap {64,128,192,256}
 *   mx_synth_group           maps per recursion, mixed synthesis {4,8,16,32} (32: spin 0,
 *                            or spin 2 with mx_synth_fold): spin 2 uses the value; spin 0
 *                            runs tile kernels for 16 and 32, else 8
 *   mx_synth_occ             its register budget in work-groups per CU {1,2,3} (spin 2)
 *   mx_adj_group, mx_adj_occ the same for the mixed adjoint {4,8,16,32}, {1,2} (32: spin 2 with
 *                            mx_adj_cb = mx_adj_f16 = 1, else 16)
 *   mx_adj_cb                16-map mixed spin-2 adjoint in column-block form (each wave
 *                            contracts 4 maps over the work-group's rings; the same bits)
 *                            {1 on (default), 2 off: per-wave tiles and a cross-wave sum}
 *   mx_single                one-map calls of a mixed plan run the FP64 vector-FMA
 *                            recursion kernels on complex64 ring coefficients, with the
 *                            FP32 ring FFT {1 on (default), 2 off: the 4-map matrix-core
 *                            kernels}.
 *   mx_f16                   16-map mixed synthesis (spin 2 and spin 0) on the FP16 matrix cores, each
 *                            FP32 operand split into hi + lo halves after a power-of-two
 *                            scale {1 on (default), 2 off: FP32 16x16x4 tiles}
 *   mx_adj_f16               the same split for the 16-map mixed adjoint in column-block
 *                            form {1 on, 2 off}; default on for spin 2, off for spin 0
 *                            (there its l >= 2048 band error rises from 2.5e-6 to 6.1e-6)
 *   mx_synth_fold            spin-2 FP16-split synthesis with the north/south parity folded
 *                            into the B operand (one recursion sign per wave, no even/odd
 *                            accumulators; 16 or 32 maps per recursion) {1 on (default), 2 off} */
typedef struct jhtx_tuning {
  int synth_rings, adj_rings, block;
  int mx_synth_group, mx_synth_occ;
  int mx_adj_group, mx_adj_occ;
  int mx_adj_cb;
  int mx_single;
  int mx_f16;
  int mx_adj_f16;
  int mx_synth_fold;
} jhtx_tuning;

/* jhtx_plan_create with explicit kernel shapes (tuning may be NULL: all defaults).  An
 * out-of-range field fails with JHTX_EINVAL. */
int jhtx_plan_create_tuned(int nside, int lmax, int spin, int precision, int max_batch, int flags,
                           const jhtx_tuning* tuning, jhtx_plan** plan);

/* The kernel shapes a plan uses, every field filled in. */
int jhtx_plan_get_tuning(const jhtx_plan* plan, jhtx_tuning* tuning);

/* Bytes of device workspace a call on this plan needs (for max_batch maps). */
size_t jhtx_plan_workspace_bytes(const jhtx_plan* plan);

/* Replaces jht.synthesis(alm, nside, lmax, spin) (src/jht/healpix.py:363-387) and the
 * jitted closure it dispatches to (healpix.py:280-291 spin 0, :312-330 spin 2).
 * alm: batch x ncomp x K complex128 (interleaved); map: batch x ncomp x npix float64. */
int jhtx_synth(jhtx_plan* plan, const void* alm, void* map, int batch, void* workspace,
               void* stream);

/* Replaces jht.adjoint_synthesis(m, nside, lmax, spin) (src/jht/healpix.py:390-412) and
 * its closure (healpix.py:293-305 spin 0, :332-355 spin 2): the exact transpose of
 * jhtx_synth.  map: batch x ncomp x npix float64; alm: batch x ncomp x K complex128. */
int jhtx_adjoint(jhtx_plan* plan, const void* map, void* alm, int batch, void* workspace,
                 void* stream);

/* jhtx_adjoint without a workspace argument.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation).  The call allocates nothing and calls no
 * allocating routine: the adjoint's ring coefficients and chunk partials go to a
 * library-owned static device array of JHTX_NOALLOC_SCRATCH_BYTES per GPU, whose address is
 * resolved when a plan is created on that GPU and which the GPU holds until the process
 * ends.  A call needs what jhtx_adjoint uses of its workspace for `batch` maps: the ring
 * coefficients (batch ncomp nrings (lmax+1) complex128, complex64 on mixed plans) and the
 * recursion's ring-chunk partials (nchunk batch ncomp K, FP64 or FP32 pairs, nchunk =
 * ceil(2 nside / 256) on mixed plans); above the maximum it fails with JHTX_EINVAL and says
 * how much it needed (e.g. nside 512, lmax 1024, spin 0, one FP64 map: 42 MB).  Every call
 * on one GPU shares that array and it keeps the last call's contents (nothing in it needs to
 * start at zero): calls that overlap on one GPU -- two streams, or two host threads -- race
 * and return wrong results.  The plan's device must be current.  Results equal
 * jhtx_adjoint's bit for bit. */
#define JHTX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int jhtx_adjoint_noalloc(jhtx_plan* plan, const void* map, void* alm, int batch, void* stream);

/* Releases the plan's device tables (the plan must not be in use on any stream). */
void jhtx_plan_destroy(jhtx_plan* plan);

/* ---------------------------------------------------------------------------------- *
 * Ragged groups: the transforms of several mixed-precision plans -- different nside, lmax
 * and batch, one spin -- run as ONE kernel launch per stage.  The reference batches by
 * jax.vmap over one geometry (scripts/gpu_check.py:150-153: vmap of jht.synthesis over
 * realisations); a group also mixes geometries.  Every stage is a single launch whose
 * work-groups cover all plans: synthesis = recursion + ring FFT (2 launches), adjoint =
 * ring FFT + recursion + chunk reduction (3 launches), so a ragged batch captures into a
 * HIP graph as five kernel nodes.  The contraction runs on the FP32 matrix cores (16 maps
 * per recursion, FP32 in, FP32 accumulate, FP64 recursion); results agree with the plans'
 * own FP32 matrix-core kernels (tuning mx_f16 = mx_adj_f16 = 2, 16-map groups) to FP32
 * rounding and with the FP64 path to <= 1e-5.
 * ---------------------------------------------------------------------------------- */
typedef struct jhtx_group jhtx_group;

/* A group over plans[0..n): n <= 16 HEALPix plans of precision JHTX_PREC_MIXED, one spin
 * (0 or 2), nside <= 512 (every ring transform fits one 256 x 16 work-group), all on the
 * current device; batch[i] (1 <= batch[i] <= max_batch of plan i) maps of plan i per call.
 * The group uses the plans' tables, which must outlive it; it keeps no pointer to the
 * arrays passed. */
int jhtx_group_create(jhtx_plan* const* plans, const int* batch, int n, jhtx_group** group);

/* Bytes of device workspace a group call needs (ring coefficients and chunk partials of
 * every plan). */
size_t jhtx_group_workspace_bytes(const jhtx_group* group);

/* jhtx_synth of every plan: alm[i] (batch[i] x ncomp x K complex128) -> map[i]
 * (batch[i] x ncomp x npix float64), in 2 kernel launches on `stream`. */
int jhtx_group_synth(jhtx_group* group, const void* const* alm, void* const* map, void* workspace,
                     void* stream);

/* jhtx_adjoint of every plan: map[i] -> alm[i], in 3 kernel launches on `stream`. */
int jhtx_group_adjoint(jhtx_group* group, const void* const* map, void* const* alm, void* workspace,
                       void* stream);

/* Group introspection: 0 plans, 1 spin, 2 kernel launches per synthesis call, 3 per adjoint
 * call, 4 + s work-groups of stage s (s = 0..4 as in jhtx_plan_set_profiling). -1 if
 * unknown. */
long long jhtx_group_query(const jhtx_group* group, int what);

/* Releases the group's descriptors (not its plans). */
void jhtx_group_destroy(jhtx_group* group);

/* Description of the last failure on the calling thread ("" if none). */
const char* jhtx_last_error(void);

/* Stage timing (no counterpart in the reference; for benchmarks and profilers): when
 * enabled, every call brackets each kernel stage with HIP events on the call's stream.
 * Stages: 0 synthesis recursion, 1 synthesis ring FFT, 2 adjoint ring FFT,
 * 3 adjoint recursion, 4 adjoint chunk reduction (the sum of the recursion's ring-chunk
 * partials).  jhtx_plan_read_profile synchronises the recorded events, writes the summed
 * milliseconds and launch counts per stage (arrays of 5) and clears them.  Up to 4096
 * calls are recorded between reads. */
int jhtx_plan_set_profiling(jhtx_plan* plan, int enable);
int jhtx_plan_read_profile(jhtx_plan* plan, double* ms, int* counts);

/* Host-side recursion tables, computed without touching a device (no counterpart in the
 * reference; for validating the plan tables on a CPU).  Array lengths: x, lsin, lcos
 * 2 nside (lcos spin 2 only); alpha, G, gamma K (gamma spin 2 only); seed lmax+1 (spin 0)
 * or 8 (lmax+1) (spin 2).  Any pointer may be NULL to skip that table. */
int jhtx_host_rec_tables(int nside, int lmax, int spin, double* x, double* lsin, double* lcos,
                         double* alpha, double* gamma, double* G, double* seed);

/* Plan introspection: 0 nside, 1 lmax, 2 spin, 3 precision, 4 max_batch, 5 ncomp,
 * 6 nrings, 7 npix (low 31 bits), 8 number of ring-FFT launch classes, 9 1 for a ring-set
 * plan. -1 if unknown. */
int jhtx_plan_query(const jhtx_plan* plan, int what);

/* ---------------------------------------------------------------------------------- *
 * Off-grid synthesis (jht.synthesis_general / adjoint_synthesis_general, FP64, one field).
 * The transform is Legendre step on a Clenshaw-Curtis ring set -> double-Fourier-sphere
 * extension -> theta FFT -> deconvolved oversampled grid -> 2-D inverse FFT -> ES-kernel
 * interpolation at the points (src/jht/offgrid.py:124-205, src/jht/_nufft.py:154-170),
 * and its transpose.  The two plain FFTs are the caller's (library FFTs); these entry
 * points are every other stage.
 * ---------------------------------------------------------------------------------- */

/* Replaces build_recursion_plan(x_cc, +-spin, lmax) as called by the off-grid _prepare
 * (src/jht/offgrid.py:72-78): an FP64 plan over an arbitrary ring set symmetric about the
 * equator.  theta: the (nrings + 1) / 2 north-half colatitudes, nondecreasing in [0, pi/2],
 * the last the equator; rings on the pole are allowed.  spin 0, 1, 2 or 3 (HEALPix plans:
 * 0, 2).  Workspace: the adjoint's partials only.  jhtx_synth / jhtx_adjoint refuse such
 * a plan (no HEALPix geometry). */
int jhtx_plan_create_rings(int nrings, const double* theta, int lmax, int spin, int max_batch,
                           jhtx_plan** plan);

/* Replaces synth_contract (src/jht/_recursion.py:372-398) as used by the off-grid
 * _alm_to_Fkm (offgrid.py:147-151): F[b][c][ring][m] = sum_l a^c_lm lambda^{s_c}_lm(ring),
 * complex128, with c the +spin / -spin channel (spin s > 0: a^+ = -(E + iB),
 * a^- = -(-1)^s (E - iB); spin 0: one channel).  Works on any plan (HEALPix plans give the coefficients before the
 * ring FFT); no workspace. */
int jhtx_legendre_synth(jhtx_plan* plan, const void* alm, void* F, int batch, void* stream);

/* Replaces adjoint_contract (src/jht/_recursion.py:401-423) as used by the off-grid
 * _Fkm_to_alm (offgrid.py:154-160): the transpose of jhtx_legendre_synth, p = sum_ring
 * lambda^{+s} F^+, n = sum_ring lambda^{-s} F^-, returned as the HEALPix adjoint returns
 * them: spin 0 p; spin s (E, B) = (-(p + t n) / 2, i (p - t n) / 2), t = (-1)^s. */
int jhtx_legendre_adjoint(jhtx_plan* plan, const void* F, void* alm, int batch, void* workspace,
                          void* stream);

/* Replaces the DFS construction of _coeffs_to_Fkm (src/jht/offgrid.py:124-134, without
 * its FFT): from F[c][ring][m] (nrings Clenshaw-Curtis rings, both poles) to
 * D[j][m' + lmax], j < nd = 2 nrings - 2, m' = -lmax..lmax: D[j] = Cfull[j] for j < nrings,
 * (-1)^(|m'| + spin) Cfull[nd - j] beyond, Cfull[m' >= 0] = F^+[m'], Cfull[m' < 0] =
 * conj(F^-[-m']). */
int jhtx_dfs_extend(const void* F, int nrings, int lmax, int spin, void* D, void* stream);

/* Replaces the transpose _Fkm_to_coeffs (src/jht/offgrid.py:138-153, without its inverse
 * FFT): D[nd][2 lmax + 1] -> F[c][ring][m]; for spin 0 the two channels are summed. */
int jhtx_dfs_fold(const void* D, int nrings, int lmax, int spin, void* F, void* stream);

/* Replaces nufft2d2's mode placement (src/jht/_nufft.py:157-164): the Nk x (2 lmax + 1)
 * centred modes X[k mod Nk][m' + lmax] (an unshifted FFT along k) onto the zeroed
 * oversampled nk x nm grid, times sk[k + Nk/2] sm[m' + lmax] (device arrays: the
 * deconvolution and normalisation factors). */
int jhtx_nufft_grid_fill(const void* X, int Nk, int lmax, const double* sk, const double* sm, void* C,
                         int nk, int nm, void* stream);

/* Replaces nufft2d1's mode extraction (src/jht/_nufft.py:179-181): the transpose of
 * jhtx_nufft_grid_fill. */
int jhtx_nufft_grid_gather(const void* C, int Nk, int lmax, const double* sk, const double* sm, void* X,
                           int nk, int nm, void* stream);

/* Bytes of the optional workspace of jhtx_nufft_interp / jhtx_nufft_spread for npts points
 * on an nk x nm grid (no counterpart in the reference): a counting sort of the points
 * into 32 x 32-cell grid tiles. */
size_t jhtx_nufft_workspace_bytes(long long npts, int nk, int nm);

/* Replaces _stencil + the stencil sum of nufft2d2 (src/jht/_nufft.py:140-170): f_j =
 * sum of the W x W ES-weighted grid values around loc_j = (theta_j, phi_j) (grid
 * complex128 nk x nm, the 2 pi torus).  re / im receive Re f, Im f (im may be NULL).
 * dloc, when not NULL, receives [4][npts]: Re, Im of df/dtheta, then of df/dphi (the
 * pointing derivative).  W in {6, 8, 10, 14, 16}.  npts = 0 is a no-op (point buffers may
 * then be NULL).  With a workspace of jhtx_nufft_workspace_bytes bytes the points are
 * sorted into grid tiles and each tile (plus halo) is read from the grid once into LDS;
 * with workspace NULL each point gathers its stencil from the grid. */
int jhtx_nufft_interp(const void* grid, int nk, int nm, int W, double beta, const double* loc,
                      long long npts, double* re, double* im, double* dloc, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Replaces the spreading of nufft2d1 (src/jht/_nufft.py:173-178): adds (re_j + i im_j) times
 * the ES weights around loc_j into grid (which the caller zeroes); im may be NULL; npts = 0
 * is a no-op.  With a workspace of jhtx_nufft_workspace_bytes bytes the points are sorted
 * into grid tiles and each tile is accumulated in LDS before one addition to the grid;
 * with workspace NULL every stencil cell is a global atomic addition.  Either way the
 * summation order, hence the last bits, may vary between runs. */
int jhtx_nufft_spread(const double* re, const double* im, const double* loc, long long npts, void* grid,
                      int nk, int nm, int W, double beta, void* workspace, size_t workspace_bytes,
                      void* stream);

/* Measured device rates for roofline denominators (no counterpart in the reference; for
 * benchmarks): out[0] FP32 matrix TFLOP/s (v_mfma_f32_16x16x4f32 loops), out[1] FP64
 * vector-FMA TFLOP/s, out[2] device-to-device copy GB/s (bytes read + written).  buf is
 * device scratch of at least 512 MiB, overwritten; the call synchronises the stream. */
int jhtx_probe_peaks(void* buf, size_t bytes, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* JHTX_H */
