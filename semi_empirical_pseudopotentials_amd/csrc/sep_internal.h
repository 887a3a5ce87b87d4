// sep_internal.h -- declarations shared by the libsep translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/sep.h"
#include "fft_core.h"
#include "sep_host.h"

namespace sep {

// Device-resident description of one grid.
struct GridDev {
  int nx, ny, nz;
  long N;
  void *tw64[3];  // exp(-2 pi i m/n) per axis, complex128
  void *tw32[3];  // same, complex64
  void *twsep32;  // per-thread-index twiddles of the separable fused kernel (null if none)
  double *kxx, *kyy, *kzz;  // filter/init.c:125-135 axis tables (1/N folded in)
  double ekin;              // Ekinmax / N (filter/init.c:137)
  // pden > 0 turns pass B into the kinetic preconditioner: the spectrum is
  // multiplied by pnum / (kin + pden) instead of kin (kin with 1/N folded in)
  double pnum, pden;
  // per precision (SEP_PREC_FP64 = 0, SEP_PREC_FP32 = 1): use the row/line
  // kernels even where a plane kernel exists (sep_autotune, or the test hook
  // SEP_FORCE_LINE_KERNELS)
  bool force_lines[2];
  // the filter loop runs pass C and the next pass A as separate kernels
  // (test hook SEP_SEPARATE_PASSES) instead of the fused plane kernel
  bool no_fuse;
  // per precision: the form of the fused plane kernel (SEP_FUSED_*)
  int fused_form[2];
};

// Spin-orbit neighbourhood lists on the device (sep_so_setup,
// spinors/filterSpinors/init.c:114-183), atoms with Vso != 0.
struct SoLists {
  int natom = 0;          // atoms with a list
  long nent = 0;          // entries over all atoms
  long *off = nullptr;    // [natom + 1] first entry of each atom
  int *jxyz = nullptr;    // [nent] grid index of each entry
  double *geo = nullptr;  // [nent][9] r, r2, Vr, y10 (re, im), y11 (re, im), y1-1 (re, im)
  long npts = 0;          // grid points in some list
  int *pidx = nullptr;    // [npts] their grid indices
  long *poff = nullptr;   // [npts + 1] into pent
  int *pent = nullptr;    // [nent] entries covering each point, atoms in order
  double c = 0.0;         // dv / (2 sqrt(2 pi sigma^2)), hamiltonian.c:98
  double s1 = 0.0;        // sigma_1 = sqrt(1/2) / sigma, init.c:97
  long maxn = 0;          // longest list
  // per atom the n x n matrix K(p, q) of the pair kernel (k_so_kmat), row q
  // at kmat + koff[a] + q n; null where it would take more than 8 GB
  double *kmat = nullptr;
  long *koff = nullptr;
  int *pmap = nullptr;  // [N] grid point -> its index u < npts, or -1
};
struct SoHost {
  std::vector<long> off, poff;
  std::vector<int> jxyz, pidx, pent;
  std::vector<double> geo;
  long maxn = 0;
};
int so_build_host(int nx, int ny, int nz, const double d[3], const double x0[3], long natom,
                  const double *xyz, const double *vso, SoHost &h);
// out_s += alpha V_so in_s for nsp spinors: component c of spinor s at
// in + s ss + c cs (out likewise); contrib holds nent * nsp * 4 doubles
int so_apply_dev(const SoLists &so, const double *in, long cs, long ss, double *out, long ocs,
                 long oss, long nsp, double alpha, double *contrib, hipStream_t st);
// V_so in_s gathered per covered point: g[(u nsp + s) 4 + c] for u < npts
// (the sums of so_apply_dev, without the add); contrib as there
int so_gather_dev(const SoLists &so, const double *in, long cs, long ss, long nsp, double *contrib,
                  double *g, hipStream_t st);
// Builds so.kmat / so.koff from the uploaded lists (off: the host copy of
// so.off); leaves them null where they would exceed 8 GB.
int so_kmat_build(SoLists &so, const std::vector<long> &off, hipStream_t st);
// Builds so.pmap over the N grid points; leaves it null when it does not fit
int so_pmap_build(SoLists &so, long N, hipStream_t st);
int spinor_times_i_dev(const double *x, double *out, long N, long m, hipStream_t st);
int spinor_gather_rows_dev(const double *src, double *dst, long N, long nsp, long ms,
                           hipStream_t st);
int axpby_dev(double a, const double *x, double b, double *y, long n, hipStream_t st);

bool hop_supported(int n);
bool plane_supported(int nx, int ny, int prec);
// Forms of the fused plane kernel (pass C + next pass A):
//   SEP_FUSED_PLANE     the fastest whole-plane form the grid has (the
//                       separable placement at complex64 128^2, else the
//                       swizzled one)
//   SEP_FUSED_CHUNKED   half a plane of LDS, two workgroups per CU
//   SEP_FUSED_SWIZZLED  the whole-plane kernel on the swizzled placement
enum { SEP_FUSED_PLANE = 0, SEP_FUSED_CHUNKED = 1, SEP_FUSED_SWIZZLED = 2 };
bool plane_fused_form_supported(int n, int prec, int form);
// Bytes of the separable fused kernel's twiddle table for planes of n (0 if
// that kernel does not exist for n / prec), and its host construction.
long plane_sep_tw_bytes(int n, int prec);
void plane_sep_tw_host(void *tab);

template <typename R>
int hpsi_prepare(const GridDev &gd);
template <typename R>
int hpsi_apply(const GridDev &gd, const VecView<R> &in, const VecView<R> &out, cx<R> *tb,
               const R *pot, int nvec, double alpha, double beta, double *part, long *nblk,
               hipStream_t st, hipEvent_t *ev = nullptr);
// Split Newton step (sep_hpsi.hip): pass A alone, then per step pass B and
// pass C fused with the next step's pass A (fwd_next); tb carries the
// spectrum of the newest vector between calls.
template <typename R>
bool hpsi_fused_supported(const GridDev &gd);
template <typename R>
int hpsi_fwd(const GridDev &gd, const VecView<R> &in, cx<R> *tb, int nvec, hipStream_t st);
// spinor filter step: psi_j += alpha (V_so psi_{j-1}) inside the fused
// kernel's epilogue (FP64 whole-plane family, not the last step): map[g] the
// row of g (4 doubles per spinor) in g, or -1
struct SoEpi {
  const int *map;
  const double *g;
  long nsp;
  double alpha;
};
template <typename R>
int hpsi_step(const GridDev &gd, const VecView<R> &in, const VecView<R> &out, cx<R> *tb,
              const R *pot, int nvec, double alpha, double beta, bool fwd_next, hipStream_t st,
              hipEvent_t *ev = nullptr, const SoEpi *so = nullptr);
// whether this grid's H.psi runs the whole-plane kernel family at prec
bool hpsi_plane_family(const GridDev &gd, int prec);

// ---- cubic Hermite interpolation of a uniform radial table (the smooth
// option of the differentiable path, SEP_INTERP_CUBIC).  On segment i
// (r_i <= r < r_(i+1), t = (r - r_i) / (r_(i+1) - r_i)) the value is
//   h00(t) p_i + h10(t) m_i + h01(t) p_(i+1) + h11(t) m_(i+1)
// with the Catmull-Rom tangents m_i = (p_(i+1) - p_(i-1)) / 2 (one-sided,
// p_1 - p_0 and p_(n-1) - p_(n-2), at the table ends), which is C^1 in r and
// reproduces quadratics exactly away from the ends.  As weights on the four
// table rows i-1 .. i+2: w[0] = -h10/2, w[1] = h00 - h11/2, w[2] = h01 +
// h10/2, w[3] = h11/2 (rows outside the table clamp to the end rows and
// cubic_edge() adds the one-sided corrections); dw = dw/dt.
__host__ __device__ inline void cubic_weights(double t, double *w, double *dw) {
  const double t2 = t * t, t3 = t2 * t;
  const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0, h10 = t3 - 2.0 * t2 + t;
  const double h01 = -2.0 * t3 + 3.0 * t2, h11 = t3 - t2;
  const double d00 = 6.0 * t2 - 6.0 * t, d10 = 3.0 * t2 - 4.0 * t + 1.0;
  const double d01 = -6.0 * t2 + 6.0 * t, d11 = 3.0 * t2 - 2.0 * t;
  w[0] = -0.5 * h10;
  w[1] = h00 - 0.5 * h11;
  w[2] = h01 + 0.5 * h10;
  w[3] = 0.5 * h11;
  dw[0] = -0.5 * d10;
  dw[1] = d00 - 0.5 * d11;
  dw[2] = d01 + 0.5 * d10;
  dw[3] = 0.5 * d11;
}
// Row weights of segment i of an n-row table after the end corrections:
// at i = 0 the tangent m_0 = p_1 - p_0 (not (p_1 - p_(-1)) / 2), at
// i + 1 = n - 1 the tangent m_(n-1) = p_(n-1) - p_(n-2).  Fills rows[4]
// (table indices) and wr[4] (weights; w or dw from cubic_weights).
__host__ __device__ inline void cubic_rows(long i, long n, const double *w, long *rows,
                                           double *wr) {
  // central-tangent weights on rows i-1, i, i+1, i+2
  rows[0] = i - 1, rows[1] = i, rows[2] = i + 1, rows[3] = i + 2;
  wr[0] = w[0], wr[1] = w[1], wr[2] = w[2], wr[3] = w[3];
  if (i == 0) {  // m_0 = p_1 - p_0 = 2 * (central form with p_-1 := 2 p_0 - p_1) / 2
    // w0 * p_-1 with p_-1 = 2 p_0 - p_1
    wr[1] += 2.0 * wr[0];
    wr[2] -= wr[0];
    wr[0] = 0.0;
    rows[0] = 0;
  }
  if (i + 2 > n - 1) {  // p_(n) := 2 p_(n-1) - p_(n-2)
    wr[2] += 2.0 * wr[3];
    wr[1] -= wr[3];
    wr[3] = 0.0;
    rows[3] = n - 1;
  }
}
// The value correction that cubic_rows applies at the table ends, for the
// forward kernel's clamped-index form (0 away from the ends).
__host__ __device__ inline double cubic_edge(long i, long n, const double *w, const double *p) {
  double c = 0.0;
  if (i == 0) c += w[0] * (p[0] - p[1]);               // p_-1 = 2 p_0 - p_1 vs p_0
  if (i + 2 > n - 1) c += w[3] * (p[n - 1] - p[n - 2]);  // p_n = 2 p_(n-1) - p_(n-2) vs p_(n-1)
  return c;
}

// ---- element-wise / reduction kernels (sep_kernels.hip)
int build_potential_dev(const GridDev &gd, double xmin, double ymin, double zmin, double dx,
                        double dy, double dz, long natom, const double *xyz_d, const int *type_d,
                        const double *scale_d, const long *npot_d, const double *vr_d,
                        const double *pot_d, long stride, const double *dr_d, double ex_e,
                        double ex_x0, double ex_t, int interp, double *potl_d, hipStream_t st);
// out[i] = sum_g a[i*lda + g] * b[i*ldb + g] for i < m (deterministic).
int dots_dev(const double *a, long lda, const double *b, long ldb, long m, long n, double *out_d,
             double *scratch, long scratch_len, hipStream_t st);
int dots_dev_f(const float *a, long lda, const float *b, long ldb, long m, long n, double *out_d,
               double *scratch, long scratch_len, hipStream_t st);
// x[i*ld + g] *= s[i]
// ---- Hellmann-Feynman adjoint (sep_adjoint.hip)
int pair_density_dev(const double *states, long N, const long *pairs_d, const double *w_d, int np,
                     double *rho, hipStream_t st);
int hf_pairs_per_pass(long maxnpot);
int hf_adjoint_dev(const GridDev &gd, double xmin, double ymin, double zmin, double dx, double dy,
                   double dz, double dv, int natom, const double *xyz, const int *type,
                   const double *scale, const long *npot, long maxnpot, const double *vr,
                   const double *pot, long stride, const double *dr, int ntype, const double *rho,
                   int np, int nsplit, int interp, double *work, double *htab, double *dR,
                   double *dscale, hipStream_t st);


int sub_scaled_rows_dev(double *y, long ldy, const double *x, long ldx, long m, long n,
                        const double *s_d, hipStream_t st);
// out_i = y_i - s_i x_i (row i, n entries) for m rows
int diff_scaled_rows_dev(double *out, long ldo, const double *y, long ldy, const double *x,
                         long ldx, long m, long n, const double *s_d, hipStream_t st);
int scale_rows_dev(double *x, long ld, long m, long n, const double *s_d, hipStream_t st);
int convert_d2f(const double *in, float *out, long n, hipStream_t st);
int convert_f2d(const float *in, double *out, long n, hipStream_t st);
// subspace-stage helpers on n x n row-major matrices (sep_kernels.hip)
int symmetrize_dev(double *A, long n, double scale, hipStream_t st, long blk = 0);
int eigvec_cols_dev(const double *V, long n, long k, const double *sc_d, bool reverse, double *M,
                    hipStream_t st);
int zero_lower_dev(double *A, long n, hipStream_t st);

// Newton accumulation on the matrix cores (sep_gemm.hip):
// acc[(s*ms + ie)*N + g] += sum_{k<cnt} anT[(j0 + k)*ms + ie] * ring row (s, k)
// ring row (state s, Newton vector k) at ring + s * sstride + k * slot
template <typename R>
int accum_mfma_dev(const R *ring, long slot, long sstride, int cnt, const double *anT_d, long j0,
                   int ms, double *acc, int ns, long N, hipStream_t st);

// The same for the complex64 filter with the coefficients packed as FP32
// (accum_pack_f32 from an[ms][nc]; accum_pack_len floats); returns
// SEP_ERR_UNSUPPORTED (nothing launched) where the shape does not allow it
// (N % 256, slot / sstride % 4, j0 % 16).
long accum_pack_len(long nc, int ms);
void accum_pack_f32(const double *an, long nc, int ms, float *cf);
int accum_mfma_packed_dev(const float *ring, long slot, long sstride, int cnt, const float *cf_d,
                          long nc, long j0, int ms, double *acc, int ns, long N, hipStream_t st);

// ---- GEMMs on the matrix cores (sep_gemm.hip); row-major operands whose
// rows run along the contraction index (state-major states, or the rows of
// a small matrix).
// C[i][j] = sum_k A[i*lda + k] B[j*ldb + k]  (m x mb into C_d, row-major),
// deterministic split-K through work (gram_work_len doubles); sym: m == mb
// and only the 128-tiles on and above the diagonal are formed (the caller
// mirrors them, symmetrize_dev with blk = 128)
template <typename R>
int gram_dev(const R *A, long lda, long m, const R *B, long ldb, long mb, long K, bool sym,
             double *C_d, double *work, long work_len, hipStream_t st);
long gram_work_len(long m, long mb, long K);
// Out[j*ldo + g] = sum_i M[i*ldm + j] A[i*lda + g]  (mo rows out, N columns),
// Out must not overlap A
template <typename R>
int combine_dev(const R *A, long lda, long m, const double *M_d, long ldm, long mo, long N, R *Out,
                long ldo, hipStream_t st);
// G = A A^T as double-double (hi + lo), m x m row-major: the upper triangle
// (16 x 16 tiles on and above the diagonal); entries below it are undefined
int gram_dd_dev(const double *A, long m, long N, double *hi_d, double *lo_d, double *work,
                long work_len, hipStream_t st);
long gram_dd_work_len(long m, long N);

}  // namespace sep
