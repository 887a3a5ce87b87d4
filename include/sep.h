/*
 * sep.h -- C ABI of libsep, the MI355X-native filter-diagonalization library
 * (semi-empirical pseudopotential `filter` program: the C files of filter/ in the
 * reference).  Plain C types only; the library owns device memory, kernels
 * and streams, the caller owns every host array it passes.
 *
 * Conventions (follow the reference):
 *   grid        x fastest: g = nx*(ny*jz + jy) + jx           (filter/init.c:138-141)
 *   complex     interleaved {re, im} doubles                   (zomplex, filter/fd.h:17-19)
 *   states      state-major real FP64, m x N ("psitot")        (filter/main.c:46,86)
 *   products    every inner product is sum(...) * dv          (filter/norm.c, energy.c)
 * Every entry point returns SEP_OK (0) or an SEP_ERR_* code and never exits
 * the process (the reference calls nerror()/exit, filter/nerror.c:9-16);
 * sep_last_error() describes the last failure on a context.
 * A context is used by one host thread at a time; calls are synchronous.
 */
#ifndef SEP_H
#define SEP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SEP_OK 0
#define SEP_ERR_ARG 1         /* bad argument / inconsistent sizes */
#define SEP_ERR_UNSUPPORTED 2 /* grid length the HIP FFT has no plan for */
#define SEP_ERR_HIP 3         /* HIP runtime error / no device */
#define SEP_ERR_IO 4          /* missing or malformed input file */
#define SEP_ERR_NOMEM 5       /* device or host allocation failed */
#define SEP_ERR_NUMERIC 6     /* factorisation failed (e.g. Cholesky) */

#define SEP_PREC_FP64 0 /* complex128 filter, FP64 everywhere (parity path) */
#define SEP_PREC_FP32 1 /* complex64 filter; with sep_ortho_fp32 + sep_refine the FP32 pipeline */

typedef struct sep_ctx sep_ctx;

/* Grid of one nanocrystal box.  Replaces the grid part of par_st/long_st
 * (filter/fd.h:24-37) as set up by init() (filter/init.c:95-144). */
typedef struct {
  long nx, ny, nz;
  double dx, dy, dz;       /* spacing, bohr */
  double xmin, ymin, zmin; /* origin (grid point 0), bohr */
  double ekinmax;          /* kinetic clamp in Ha (Ekinmax = 10, filter/init.c:31) */
  int device;              /* HIP device ordinal */
} sep_grid_desc;

/* ---------------------------------------------------------------- context */
/* Allocates the per-grid device state (twiddles, k^2 axis tables).
 * No counterpart function in the reference: main() allocates these in
 * filter/main.c:27-51 and plans FFTW at filter/main.c:58-59. */
int sep_open(const sep_grid_desc *grid, sep_ctx **ctx);
int sep_close(sep_ctx *ctx);
const char *sep_last_error(const sep_ctx *ctx);
/* 1 when the library can transform a line of length n on the GPU. */
int sep_fft_length_supported(long n);

/* -------------------------------------------------------------- potential */
/* Replaces the potl array filled by init() (filter/init.c:157-182). */
int sep_set_potential(sep_ctx *ctx, const double *potl);
/* Device build of V(r) = sum_a scale_a * interp_{type a}(|r - R_a|) + expot(|r|)
 * (filter/init.c:161-182, filter/interpolate.c:11-23, expot filter/init.c:255-257;
 * the per-atom scale is filterStrainCube/interpolate.c:20-21, 1 = plain filter).
 * xyz: natom x 3 (bohr, already recentred as read_conf does, read.c:23-30);
 * tables: ntype rows of `stride` doubles (vr = r grid, pot = shifted table,
 * read.c:106), npot/dr per type (read.c:107).  expot = {e, x0, kT} or NULL.
 * potl_out (N, optional) receives V; vmin/vmax (optional) its extrema
 * (filter/init.c:196-199). */
int sep_build_potential(sep_ctx *ctx, long natom, const double *xyz, const int *type,
                        const double *scale, int ntype, const long *npot, const double *vr,
                        const double *pot, long stride, const double *dr,
                        const double *expot, double *potl_out, double *vmin, double *vmax);

/* Table interpolation used by sep_build_potential and sep_hf_adjoint on this
 * context.  SEP_INTERP_LINEAR (default) is the reference's piecewise-linear
 * interpolate() (filter/interpolate.c:11-23), a function of |r - R| with
 * kinks at every table row.  SEP_INTERP_CUBIC is an extension with no
 * counterpart in the reference: the cubic Hermite interpolant of the same
 * rows (Catmull-Rom tangents, one-sided at the table ends, same segment and
 * cutoff), C^1 in |r - R|, for derivative checks and fits that need V(r)
 * smooth in the atom positions; it differs from the linear V by O(dr^2 v''). */
#define SEP_INTERP_LINEAR 0
#define SEP_INTERP_CUBIC 1
int sep_set_interp(sep_ctx *ctx, int interp);

/* ---------------------------------------------------------------- H . psi */
/* phi <- T phi + V phi on `batch` complex vectors (in != out allowed).
 * Replaces hamiltonian(), filter/hamiltonian.c:11-23 (with kinetic(), :27-44). */
int sep_apply_h(sep_ctx *ctx, const double *in_c128, double *out_c128, long batch);
/* One Newton step: out = dE_1 (H in) - (2 + zm1 + vmin dE_1) in.
 * Replaces hnorm()/hamiltonian_norm(), filter/hnorm.c:11-39. */
int sep_hnorm(sep_ctx *ctx, const double *in_c128, double *out_c128, long batch, double zm1,
              double vmin, double dE_1);
/* Device-pointer variants (pointers from hipMalloc / torch): same semantics,
 * precision SEP_PREC_FP64 = complex128, SEP_PREC_FP32 = complex64 data. */
int sep_apply_h_dev(sep_ctx *ctx, const void *in_dev, void *out_dev, long batch, int precision);

/* No counterpart in the reference: sep_apply_h_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and result as
 * sep_apply_h_dev, but the FFT spectrum buffer (batch * N complex elements of
 * the precision) comes from a static device array of SEP_NOALLOC_SCRATCH_BYTES
 * that each GPU holds from library load to process end; no hipMalloc,
 * hipFree or allocating library call is made.  batch * N * (16 | 8) bytes
 * beyond that limit returns SEP_ERR_ARG.  Every call on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (other host threads,
 * other contexts' streams) race.  The potential must be set. */
#define SEP_NOALLOC_SCRATCH_BYTES (32L << 20)
#define SEP_MAX_DEVICES 64
int sep_apply_h_dev_noalloc(sep_ctx *ctx, const void *in_dev, void *out_dev, long batch,
                            int precision);

/* ---------------------------------------------------------- spectral range */
/* Power iteration from the fixed-seed state (energy.c:57,61): returns
 * dE = 1.1 (E - vmin) (energy.c:76), the converged E and the iteration count.
 * Replaces get_energy_range(), filter/energy.c:53-83. */
int sep_energy_range(sep_ctx *ctx, double vmin, double *dE, double *emax, long *iters);
/* sep_energy_range with the iteration trace of filter/energy.c:65-74
 * (ene-init.dat): log[3 i .. 3 i + 2] = (previous energy, energy, norm of
 * H psi) of iteration i, for i < min(iters, maxlog). */
int sep_energy_range_log(sep_ctx *ctx, double vmin, double *dE, double *emax, long *iters,
                         double *log, long maxlog);

/* ------------------------------------------------- random states (host) */
/* Numerical Recipes ran2 state that filter/rand.c:28-66 keeps in statics. */
typedef struct {
  long idum2, iy, iv[32];
} sep_rng;
void sep_rng_reset(sep_rng *st);
/* ran_nrc(), filter/rand.c:28-66 (bit-exact). */
double sep_ran_nrc(sep_rng *st, long *idum);
/* init_psi(), filter/init.c:236-251: uniform [-1,1] real draw, normalised. */
void sep_init_psi(sep_rng *st, long *idum, long ngrid, double dv, double *psi);

/* ------------------------------------------- Newton coefficients (host) */
/* samp_points_ashkenazy(), filter/coeff.c:82-134: nc Leja-type points on
 * [-2,2]; returns the capacity estimate fk (coeff.c:295-297). */
double sep_leja_points(long nc, double *zn);
/* coefficient(), filter/coeff.c:16-63: an (ms x nc, real parts; the
 * reference's imaginary parts are identically 0) of the Gaussian
 * sqrt(dt/pi) exp(-dt (E - El)^2) at the points mapped to [vmin, vmin + dE]. */
int sep_newton_coeffs(long nc, long ms, const double *zn, const double *el, double dE,
                      double vmin, double dt, double *an);

/* ------------------------------------------------------------ the filter */
/* Filters ns real random states (ns x N, each normalised) for ms targets:
 * psitot[jns*ms + ie] = sum_j an[ie][j] Re psi_j(jns), psi_j = prod (X - z_k) psi_0,
 * X = (H - vmin) 4/dE - 2, then normalize_all (filter.c:33).
 * Replaces filtering()/filter(), filter/filter.c:11-83 (the OpenMP loop of
 * filter/main.c:92-98 becomes one batched launch sequence).  The result
 * stays resident in the context (see sep_states_*); psitot (optional,
 * (ns*ms) x N host) receives a copy and ene (optional, ns*ms) the Rayleigh
 * quotients of energy_all (filter.c:36, energy.c:28-49).
 * precision: SEP_PREC_FP64 or SEP_PREC_FP32 (complex64 recurrence/FFTs,
 * FP64 accumulators). */
int sep_filter(sep_ctx *ctx, const double *psi0, long ns, const double *an, const double *zn,
               long nc, long ms, double vmin, double dE, int precision, double *psitot,
               double *ene);

/* Same as sep_filter with the random states already on the device
 * (psi0_dev: ns x N FP64, hipMalloc'd); nothing is copied back.  This is the
 * entry a caller uses when the inputs are resident in HBM (no counterpart
 * in the reference, whose states always live in host memory). */
int sep_filter_dev(sep_ctx *ctx, const void *psi0_dev, long ns, const double *an,
                   const double *zn, long nc, long ms, double vmin, double dE, int precision);

/* ------------------------------------------------- dense device contractions */
/* On device buffers of the context's grid (N = nx ny nz points per state,
 * state-major; FP64 when precision = SEP_PREC_FP64, FP32 otherwise), on the
 * matrix cores (FP64: rocBLAS dgemm; FP32: this library's MFMA kernels):
 *   sep_gram_dev    C[i][j] = sum_g A[i][g] B[j][g]   (C: host, m x mb, FP64)
 *   sep_rotate_dev  Out[j][g] = sum_i M[i][j] A[i][g] (M: host, m x mo row-major)
 * They are the projected-H / overlap contractions and the Ritz rotation of
 * filter/Hmat.c:30-46 and :61-78 (there scalar loops). */
int sep_gram_dev(sep_ctx *ctx, int precision, const void *A_dev, long m, const void *B_dev,
                 long mb, double *C);
int sep_rotate_dev(sep_ctx *ctx, int precision, const void *A_dev, long m, const double *M,
                   long mo, void *Out_dev);

/* ------------------------------------------------- batched uneven grids
 * sep_filter_dev for ndots nanocrystals at once, each with its own context
 * (grid, potential, stream) and its own random states, coefficients and
 * sizes; all contexts on one device.  mode SEP_BATCH_INTERLEAVE enqueues one
 * Newton step of every unfinished grid per round on the grids' streams, so
 * that kernels of small grids run beside those of large ones;
 * SEP_BATCH_GRAPH captures each grid's whole step sequence into one HIP
 * graph and launches the graphs on the grids' streams.  Results as
 * sep_filter_dev, per context.  No counterpart in the reference, which runs
 * one nanocrystal per process (filter/main.c). */
#define SEP_BATCH_INTERLEAVE 0
#define SEP_BATCH_GRAPH 1
int sep_filter_batch(int ndots, sep_ctx *const *ctxs, const void *const *psi0_dev, const long *ns,
                     const double *const *an, const double *const *zn, const long *nc,
                     const long *ms, const double *vmin, const double *dE, int precision,
                     int mode);

/* ------------------------------------------------------------ measurement */
/* Times the H.psi passes on nvec resident complex vectors (random data) with
 * HIP events on the context's stream: kernel_ms[0..3] = average duration of
 * pass A (xy forward), pass B (z pencils with k^2), pass C (xy backward +
 * Newton epilogue) and one accumulation launch (K = 64 Newton vectors,
 * ms targets, 2 nvec real states); kernel_ms[4] = passes A+B+C timed as one
 * span; kernel_ms[5] = one Newton step of the filter loop's split form
 * (the spectral passes and the kernel that fuses pass C with the next step's
 * pass A), 0 where the grid has no split form.  kernel_ms holds 6 values.  No counterpart in the reference (it only prints clock() totals,
 * filter/main.c:63-66,91-100). */
int sep_bench_hpsi(sep_ctx *ctx, int precision, long nvec, long ms, long iters,
                   double *kernel_ms);
/* HBM bandwidth probe: *gbs = the best rate (GB/s, bytes read + bytes
 * written) of a plain copy of nbytes (>= 1 MiB) between two device buffers,
 * 16 bytes per lane, over iters timed launches at each of four grid sizes --
 * the achievable bandwidth beside the nominal 8 TB/s.  No counterpart in the
 * reference. */
int sep_bench_copy(sep_ctx *ctx, long nbytes, long iters, double *gbs);

/* Live kernel timing of the filter loop.  While on, sep_filter/sep_filter_dev
 * record HIP events on the context's stream around the three H.psi passes of
 * one Newton step per ring block, and around that block's accumulation
 * launch (one host wait per block).  sep_set_timing resets the sums;
 * sep_get_timing returns avg_ms[0..3] = passes A, B, C and the accumulation
 * of a full ring block (where the loop runs pass C fused with the next
 * step's pass A, avg_ms[0] = 0, avg_ms[1] covers the spectral passes -- the
 * row + line family's y lines, z pencils and y lines -- and avg_ms[2] the
 * fused kernel), and
 * samples[0..1] = the number of sampled steps and
 * of sampled full blocks.  No counterpart in the reference (it only prints
 * clock() totals, filter/main.c:63-66,91-100). */
int sep_set_timing(sep_ctx *ctx, int on);
/* Frees the filter's ring of Newton vectors (up to 160 GB; the context
 * otherwise keeps it for the next filter until sep_close).  The resident
 * states and everything else stay; the next filter allocates a ring again.
 * No counterpart in the reference, which frees nothing before it exits. */
int sep_trim(sep_ctx *ctx);
/* Largest number of filter lanes. */
#define SEP_MAX_LANES 4
/* Filter lanes: sep_filter / sep_filter_dev split the random states into
 * `lanes` groups of whole complex vectors (two states each) that run their
 * Newton steps and accumulations on streams of their own, so that one
 * group's kernels overlap another's -- the device counterpart of the
 * reference's OpenMP loop over random states (filter/main.c:94).  Results
 * are bitwise those of one lane.  0 (default) = automatic: 2 lanes from 4
 * complex vectors on; 1 = off; at most SEP_MAX_LANES and one per complex
 * vector.  Lanes apply to the split-step loop (sep_autotune's families on
 * grids with a split form) and are off while live timing is on.  No
 * counterpart in the reference's function interface. */
int sep_set_filter_lanes(sep_ctx *ctx, int lanes);
/* Autotuner: times the H.psi passes of this grid in `precision` on nvec
 * resident random vectors with the plane kernels (when the line length has
 * them) and with the row + strided-line kernels, and keeps the faster family
 * for this context and precision.  ms_plane (-1 when unavailable) and
 * ms_lines (optional) get the measured time of one Newton step's H.psi in
 * the filter loop's split form (A + B + C where a family has none).
 * No counterpart in the reference (FFTW plans with FFTW_ESTIMATE,
 * filter/filter.c:22-23).  Uses the
 * context's state and ring buffers as scratch: call it before filtering. */
int sep_autotune(sep_ctx *ctx, int precision, long nvec, double *ms_plane, double *ms_lines);
/* H.psi kernel families of sep_set_hpsi_family. */
#define SEP_HPSI_PLANE 0         /* whole-plane kernels (passes A and C, C fused with A) */
#define SEP_HPSI_PLANE_CHUNKED 1 /* same, the fused kernel transposing through half a \
                                    plane of LDS so that two workgroups share a CU */
#define SEP_HPSI_LINES 2         /* row + strided-line kernels */
#define SEP_HPSI_PLANE_SWIZZLED 3 /* whole-plane kernels, the fused kernel on the \
                                     xor-swizzled LDS placement that SEP_HPSI_PLANE \
                                     replaces at complex64 128 x 128 */
#define SEP_HPSI_NFAMILY 4
/* Selects the H.psi kernel family of this context in `precision`.
 * SEP_ERR_UNSUPPORTED if the grid has no such form (plane families need
 * nx == ny and a plane length; the chunked and swizzled forms exist for
 * complex64 128 x 128 planes).  All families give the same H.psi to
 * rounding. */
int sep_set_hpsi_family(sep_ctx *ctx, int precision, int family);
int sep_get_hpsi_family(sep_ctx *ctx, int precision, int *family);
/* Autotuner over every family this grid has and over the number of filter
 * lanes (1 .. SEP_MAX_LANES, see sep_set_filter_lanes): times the filter
 * loop's Newton step (H.psi in split form plus the lane split) on nvec
 * resident random vectors, keeps the fastest family and lane count, and
 * returns the time per family (ms[SEP_HPSI_NFAMILY], -1 where the grid has
 * no such form; may be NULL), the chosen family and lane count.  Uses the
 * context's state and ring buffers as scratch: call it before filtering.
 * No counterpart in the reference (FFTW_ESTIMATE plans, filter/filter.c:22-23). */
int sep_autotune_ex(sep_ctx *ctx, int precision, long nvec, double *ms, int *family, int *lanes);
int sep_get_timing(sep_ctx *ctx, double *avg_ms, long *samples);

/* -------------------------------------------------- resident state block */
/* Upload/download m states (m x N FP64) to/from the context. */
int sep_states_upload(sep_ctx *ctx, const double *psi, long m);
int sep_states_download(sep_ctx *ctx, double *psi, long m);
/* normalize_all(), filter/norm.c:40-54, on the first m resident states. */
int sep_normalize_all(sep_ctx *ctx, long m);
/* energy_all(), filter/energy.c:28-49: ene[i] = <psi_i|H|psi_i> dv. */
int sep_energy_all(sep_ctx *ctx, long m, double *ene);

/* ---------------------------------------------------------------- subspace */
/* Orthonormalises the m resident states (thin SVD: double-double Gram,
 * shifted double-double Cholesky R, SVD of R on the device), keeps the columns with S_i/S_0 >= cutoff (1e-10 in
 * the reference), then normalize_all.  Replaces portho(), filter/ortho.c:11-52
 * (+ normalize_all at filter/main.c:112).  S (optional, m) gets the singular
 * values in descending order, *kept the new m. */
int sep_ortho(sep_ctx *ctx, long m, double cutoff, double *S, long *kept);
/* Rayleigh-Ritz on the m resident (orthonormal) states: H_ij = dv <psi_i|H psi_j>,
 * eigen-decomposition, psi <- psi Q, normalize_all; eval ascending.
 * Replaces Hmatreal(), filter/Hmat.c:11-86 (+ normalize_all, main.c:122). */
int sep_rayleigh_ritz(sep_ctx *ctx, long m, double *eval);
/* sigma_n = sqrt|<psi|H^2 psi> dv - (<psi|H psi> dv)^2| of the m resident states.
 * Replaces calc_sigma_E(), filter/energy.c:88-119. */
int sep_sigma(sep_ctx *ctx, long m, double *sige);

/* ------------------------------------------- Hellmann-Feynman adjoint of V(r)
 * Contractions of state densities of the resident states against the
 * derivatives of the forward V(r) of sep_build_potential (same atom and table
 * arguments): for each pair p = (i, j) of resident-state indices,
 *   dR[p][a][c]    = dv sum_g psi_i psi_j dV/dR_{a,c}       (natom x 3)
 *   dscale[p][a]   = dv sum_g psi_i psi_j dV/ds_a           (natom)
 *   dtab[p][t][k]  = dv sum_g psi_i psi_j dV/dpot_t[k]      (ntype x stride)
 * With weights (npairs), the pairs are first combined into the single
 * density sum_p w_p psi_i psi_j and each output holds one slice.  With i == j
 * these are the Hellmann-Feynman derivatives dE/d(.) of a converged state;
 * with i != j the couplings <psi_i|dV/dR_a|psi_j> that
 * nonAdiabaticCoupling/nonadiabatic.c:143-301 evaluates on the CPU (there
 * with a central-difference table derivative, pseudopotential.c:6-58; here
 * the exact derivative of the piecewise-linear interpolation,
 * filter/interpolate.c:11-23).  Outputs are optional (NULL skips).  dtab is
 * accumulated with atomic adds: reproducible to rounding, not bitwise. */
int sep_hf_adjoint(sep_ctx *ctx, long npairs, const long *pairs, const double *weights,
                   long natom, const double *xyz, const int *type, const double *scale,
                   int ntype, const long *npot, const double *vr, const double *pot, long stride,
                   const double *dr, double *dR, double *dscale, double *dtab);

/* ------------------------------------------- FP32 subspace + FP64 refinement
 * The SEP_PREC_FP32 pipeline: sep_filter (complex64) -> sep_ortho_fp32 ->
 * sep_rayleigh_ritz (FP64) -> sep_refine (FP64).  No counterpart in the
 * reference, which is FP64 throughout (filter/ortho.c, Hmat.c, energy.c). */
/* Smallest singular-value ratio sep_ortho_fp32 keeps: its FP64 Gram matrix
 * carries eigenvalue rounding of ~1e-16 l_0, i.e. singular-value ratios are
 * resolved down to ~1e-8. */
#define SEP_FP32_SVD_FLOOR 1.0e-7
/* portho() (filter/ortho.c:11-52) for the FP32 pipeline: Gram matrix of the
 * resident states (FP64 dgemm), its eigen-decomposition on the device, basis
 * A W S^-1 (FP64 dgemm), then FP64 Cholesky-QR (a second pass only if the
 * first Gram matrix, scaled to unit diagonal, is not within 0.5 of the
 * identity in Frobenius norm).  Directions with
 * S_i/S_0 below max(cutoff, SEP_FP32_SVD_FLOOR) are dropped.  The states of a
 * complex64 filter carry a noise floor near 1e-5 of S_0, so the floor drops
 * only exactly dependent directions. */
int sep_ortho_fp32(sep_ctx *ctx, long m, double cutoff, double *S, long *kept);
/* Shift (Ha) of the kinetic preconditioner of sep_refine. */
#define SEP_REFINE_SHIFT 1.0
/* One FP64 correction pass on m resident Ritz pairs with values eval:
 * residuals r_i = H psi_i - e_i psi_i, preconditioned t_i = (T + shift)^-1 r_i
 * (T the kinetic energy, applied in reciprocal space), Rayleigh-Ritz on
 * span{psi, t} in FP64 (the generalized problem on its eigen-whitened basis,
 * which leaves the Ritz states orthonormal to rounding), and the m pairs of
 * smallest sigma are kept (eval and
 * sige overwritten, ascending in energy).  Returns SEP_ERR_NUMERIC if the
 * enlarged basis holds fewer than m directions. */
int sep_refine(sep_ctx *ctx, long m, double *eval, double *sige);
/* Further correction passes (as sep_refine's) on the Ritz pairs of the
 * first m resident pairs (eval ascending, sige their variances) whose
 * energies lie in [lo, hi]: those kw pairs take `passes` passes of their own
 * (each Rayleigh-Ritz on [psi, P r, psi_prev] keeping the pairs inside the
 * window first) and are written back in place, with their eval and sige;
 * the other pairs are left as they were, and all m pairs end ascending in
 * energy (reordered only if a refined level crossed a neighbour outside the
 * window).  Needs room for m + 4 kw resident states.  Converges the levels
 * of a window of the FP32 pipeline at a cost in the window's size, not m's.
 * No counterpart in the reference. */
int sep_refine_window(sep_ctx *ctx, long m, double *eval, double *sige, double lo, double hi,
                      int passes, long *kw);

/* ------------------------------------------------------ spin-orbit spinors */
/* The reference's spin-orbit filter (spinors/filterSpinors), FP64.  A spinor
 * is four real rows of N doubles: up.re, up.im, dn.re, dn.im (the
 * reference's zomplex[2 ngrid], up then down, de-interleaved).  The kinetic
 * clamp of that program is Ekinmax = 20 Ha (spinors/filterSpinors/init.c:23):
 * open the context with ekinmax = 20 for its decks. */
/* init_list (spinors/filterSpinors/init.c:114-183): the grid points within
 * Rnlcut^2 = 0.49 * 6 ln 10 of each atom, with Vr = Vso exp(-r^2 / 0.49), r,
 * r^2 and the l = 1 harmonics; sigma = dx (init.c:96-97).  vso[a] is the
 * atom's spin-orbit strength (read.c:33-51; 0 for the passivating
 * ligands): atoms with Vso = 0 contribute exactly nothing and get no list.
 * nentries (optional) gets the number of (atom, point) entries. */
int sep_so_setup(sep_ctx *ctx, long natom, const double *xyz, const double *vso,
                 long *nentries);
/* hamiltonian() (spinors/filterSpinors/hamiltonian.c:8-50, nonlocal_potential
 * :79-192, the flagkb = 0 form): out = (T + V + V_so) in for nspinor device
 * spinors (in != out). */
int sep_apply_h_spinor_dev(sep_ctx *ctx, const void *in_dev, void *out_dev, long nspinor);
/* get_energy_range() (spinors/filterSpinors/energy.c:53-112), fixed seed
 * -874917403: Vmin = Emin - 0.2 |Emin|, dE = 1.2 Emax - Vmin; iters (optional,
 * 2 values) gets the iterations of the Emin and Emax loops. */
int sep_energy_range_spinor(sep_ctx *ctx, double *vmin, double *dE, long *iters);
/* filtering() for every random spinor (spinors/filterSpinors/filter.c:5-83,
 * main.c:84-107): nspinor device spinors, an (ms x nc, real: the Newton
 * coefficients of the scalar filter, coeff.c is the same), zn (nc).  The
 * ms * nspinor filtered spinors, each normalized, become the resident state
 * block, spinor jns * ms + jms as the reference's psitot. */
int sep_filter_spinor_dev(sep_ctx *ctx, const void *psi0_dev, long nspinor, const double *an,
                          const double *zn, long nc, long ms, double vmin, double dE);
/* The same on nspinor host spinors (4 N doubles each), for callers without
 * device memory of their own (the filter_spinors driver). */
int sep_filter_spinor(sep_ctx *ctx, const double *psi0, long nspinor, const double *an,
                      const double *zn, long nc, long ms, double vmin, double dE);
/* energy_all() / calc_sigma_E() (spinors/filterSpinors/energy.c:28-49,
 * 117-147) on the first m resident spinors. */
int sep_energy_spinor(sep_ctx *ctx, long m, double *ene);
int sep_sigma_spinor(sep_ctx *ctx, long m, double *sige);
/* Smallest S_i / S_0 sep_ortho_spinor resolves: it takes singular values
 * from the FP64 complex Gram matrix, whose eigenvalues are exact to about
 * 1e-16 of the largest. */
#define SEP_SPINOR_SVD_FLOOR 1.0e-7
/* portho() (spinors/filterSpinors/ortho.c:5-45): orthonormal basis of the
 * span of the m resident spinors, directions with S_i / S_0 below
 * max(cutoff, SEP_SPINOR_SVD_FLOOR) dropped; S (optional, m values) gets the
 * singular values, kept the basis size.  The reference's zgesvd resolves
 * S_i / S_0 down to its 1e-10 cutoff; this floor drops the directions
 * between 1e-10 and 1e-7 of the largest as well. */
int sep_ortho_spinor(sep_ctx *ctx, long m, double cutoff, double *S, long *kept);
/* Hmat() (spinors/filterSpinors/Hmat.c:5-72): the complex Hermitian
 * projected H on the m resident orthonormal spinors, its eigenvalues
 * ascending into eval, the resident spinors replaced by the Ritz spinors
 * (normalized, main.c:146). */
int sep_rayleigh_ritz_spinor(sep_ctx *ctx, long m, double *eval);

#ifdef __cplusplus
}
#endif
#endif
