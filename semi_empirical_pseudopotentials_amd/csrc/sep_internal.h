// sep_internal.h -- declarations shared by the libsep translation units.
#pragma once
#include <hip/hip_runtime.h>

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
};

bool hop_supported(int n);
bool plane_supported(int nx, int ny, int prec);

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
template <typename R>
int hpsi_step(const GridDev &gd, const VecView<R> &in, const VecView<R> &out, cx<R> *tb,
              const R *pot, int nvec, double alpha, double beta, bool fwd_next, hipStream_t st,
              hipEvent_t *ev = nullptr);

// ---- element-wise / reduction kernels (sep_kernels.hip)
int build_potential_dev(const GridDev &gd, double xmin, double ymin, double zmin, double dx,
                        double dy, double dz, long natom, const double *xyz_d, const int *type_d,
                        const double *scale_d, const long *npot_d, const double *vr_d,
                        const double *pot_d, long stride, const double *dr_d, double ex_e,
                        double ex_x0, double ex_t, double *potl_d, hipStream_t st);
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
                   int np, int nsplit, double *work, double *htab, double *dR, double *dscale,
                   hipStream_t st);


int sub_scaled_rows_dev(double *y, long ldy, const double *x, long ldx, long m, long n,
                        const double *s_d, hipStream_t st);
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
