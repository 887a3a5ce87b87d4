// sep_kernels.hip -- non-FFT kernels of libsep: V(r) assembly, deterministic
// per-state reductions, row scaling and precision conversion.
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "sep_internal.h"

namespace sep {

// ------------------------------------------------------------ V(r) assembly
// filter/init.c:161-182 (sum over atoms in file order, + expot), interpolate
// of filter/interpolate.c:11-23 with the per-atom scale of
// filterStrainCube/interpolate.c:20-21.  Contraction is off so each term is
// rounded exactly as the reference source writes it.  interp = SEP_INTERP_CUBIC
// replaces the linear segment by the cubic Hermite interpolant of
// cubic_weights() (same segment, same cutoff), a C^1 function of |r - R|.
__global__ void __launch_bounds__(256) k_build_potential(
    int nx, int ny, int nz, double xmin, double ymin, double zmin, double dx, double dy, double dz,
    long natom, const double *__restrict__ xyz, const int *__restrict__ type,
    const double *__restrict__ scale, const long *__restrict__ npot, const double *__restrict__ vr,
    const double *__restrict__ pot, long stride, const double *__restrict__ dr, double ex_e,
    double ex_x0, double ex_t, int interp, double *__restrict__ potl) {
#pragma clang fp contract(off)
  const long N = (long)nx * ny * nz;
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double sx[256], sy[256], sz[256], ss[256];
  __shared__ int st[256];
  // grid coordinates by repeated addition (filter/init.c:157-159)
  int jx = 0, jy = 0, jz = 0;
  double px = xmin, py = ymin, pz = zmin;
  if (g < N) {
    jx = (int)(g % nx);
    jy = (int)((g / nx) % ny);
    jz = (int)(g / ((long)nx * ny));
    for (int i = 0; i < jx; i++) px += dx;
    for (int i = 0; i < jy; i++) py += dy;
    for (int i = 0; i < jz; i++) pz += dz;
  }
  double sum = 0.0;
  for (long a0 = 0; a0 < natom; a0 += 256) {
    __syncthreads();
    long a = a0 + threadIdx.x;
    if (a < natom) {
      sx[threadIdx.x] = xyz[3 * a];
      sy[threadIdx.x] = xyz[3 * a + 1];
      sz[threadIdx.x] = xyz[3 * a + 2];
      ss[threadIdx.x] = scale ? scale[a] : 1.0;
      st[threadIdx.x] = type[a];
    }
    __syncthreads();
    const int na = (int)((natom - a0) < 256 ? (natom - a0) : 256);
    if (g < N) {
      for (int k = 0; k < na; k++) {
        double ddx = px - sx[k], ddy = py - sy[k], ddz = pz - sz[k];
        double del = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        const int t = st[k];
        const double drt = dr[t];
        long i = (long)(del / drt);
        if (i > npot[t] - 2) continue;  // adds 0.0 in the reference
        const double *p = pot + t * stride, *v = vr + t * stride;
        if (interp == SEP_INTERP_CUBIC) {
          double w[4], dw[4];
          cubic_weights((del - v[i]) / (v[i + 1] - v[i]), w, dw);
          const long n = npot[t];
          sum += ss[k] * (w[0] * p[i > 0 ? i - 1 : 0] + w[1] * p[i] + w[2] * p[i + 1] +
                          w[3] * p[i + 2 < n ? i + 2 : n - 1] + cubic_edge(i, n, w, p));
          continue;
        }
        double aa = ss[k] * (p[i + 1] - p[i]) / (v[i + 1] - v[i]);
        double bb = ss[k] * p[i] - v[i] * aa;
        sum += aa * del + bb;
      }
    }
  }
  if (g < N) {
    double rex = sqrt(px * px + py * py + pz * pz);
    double potEx = ex_e * (1.0 / (exp((rex - ex_x0) / ex_t) + 1.0) - 1.0);
    potl[g] = sum + potEx;
  }
}

int build_potential_dev(const GridDev &gd, double xmin, double ymin, double zmin, double dx,
                        double dy, double dz, long natom, const double *xyz_d, const int *type_d,
                        const double *scale_d, const long *npot_d, const double *vr_d,
                        const double *pot_d, long stride, const double *dr_d, double ex_e,
                        double ex_x0, double ex_t, int interp, double *potl_d, hipStream_t st) {
  dim3 grid((unsigned)((gd.N + 255) / 256));
  hipLaunchKernelGGL(k_build_potential, grid, dim3(256), 0, st, gd.nx, gd.ny, gd.nz, xmin, ymin,
                     zmin, dx, dy, dz, natom, xyz_d, type_d, scale_d, npot_d, vr_d, pot_d, stride,
                     dr_d, ex_e, ex_x0, ex_t, interp, potl_d);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// ------------------------------------------------ deterministic row dots
// stage 1: part[i*nb + b] = sum over block b's chunk; stage 2: fixed-order sum
template <typename R>
__global__ void __launch_bounds__(256) k_dot_part(const R *a, long lda, const R *b, long ldb,
                                                  long n, long chunk, double *part) {
  const long i = blockIdx.y;
  const long b0 = (long)blockIdx.x * chunk;
  long b1 = b0 + chunk;
  if (b1 > n) b1 = n;
  double s = 0.0;
  for (long g = b0 + threadIdx.x; g < b1; g += 256) s += (double)a[i * lda + g] * (double)b[i * ldb + g];
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[i * gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void k_dot_final(const double *part, long nb, long m, double *out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (long b = 0; b < nb; b++) s += part[i * nb + b];
  out[i] = s;
}

template <typename R>
static int dots_impl(const R *a, long lda, const R *b, long ldb, long m, long n, double *out_d,
                     double *scratch, long scratch_len, hipStream_t st) {
  long chunk = 16384;
  long nb = (n + chunk - 1) / chunk;
  if (nb * m > scratch_len) return SEP_ERR_NOMEM;
  hipLaunchKernelGGL(k_dot_part<R>, dim3((unsigned)nb, (unsigned)m), dim3(256), 0, st, a, lda, b,
                     ldb, n, chunk, scratch);
  hipLaunchKernelGGL(k_dot_final, dim3((unsigned)((m + 127) / 128)), dim3(128), 0, st, scratch, nb,
                     m, out_d);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
int dots_dev(const double *a, long lda, const double *b, long ldb, long m, long n, double *out_d,
             double *scratch, long scratch_len, hipStream_t st) {
  return dots_impl<double>(a, lda, b, ldb, m, n, out_d, scratch, scratch_len, st);
}
int dots_dev_f(const float *a, long lda, const float *b, long ldb, long m, long n, double *out_d,
               double *scratch, long scratch_len, hipStream_t st) {
  return dots_impl<float>(a, lda, b, ldb, m, n, out_d, scratch, scratch_len, st);
}

__global__ void k_scale_rows(double *x, long ld, long n, const double *s) {
  const long i = blockIdx.y;
  const double f = s[i];
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    x[i * ld + g] *= f;
}
int scale_rows_dev(double *x, long ld, long m, long n, const double *s_d, hipStream_t st) {
  unsigned nbx = (unsigned)((n + 255) / 256);
  if (nbx > 4096) nbx = 4096;
  hipLaunchKernelGGL(k_scale_rows, dim3(nbx, (unsigned)m), dim3(256), 0, st, x, ld, n, s_d);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// y_i -= s_i x_i for m rows of n points (residuals H psi - e psi)
// out_i = y_i - s_i x_i for the m rows (out may be y)
__global__ void k_sub_scaled_rows(double *out, long ldo, const double *y, long ldy,
                                  const double *x, long ldx, long n, const double *s) {
  const long i = blockIdx.y;
  const double f = s[i];
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    out[i * ldo + g] = y[i * ldy + g] - f * x[i * ldx + g];
}
int diff_scaled_rows_dev(double *out, long ldo, const double *y, long ldy, const double *x,
                         long ldx, long m, long n, const double *s_d, hipStream_t st) {
  unsigned nbx = (unsigned)((n + 255) / 256);
  if (nbx > 4096) nbx = 4096;
  hipLaunchKernelGGL(k_sub_scaled_rows, dim3(nbx, (unsigned)m), dim3(256), 0, st, out, ldo, y, ldy,
                     x, ldx, n, s_d);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
int sub_scaled_rows_dev(double *y, long ldy, const double *x, long ldx, long m, long n,
                        const double *s_d, hipStream_t st) {
  return diff_scaled_rows_dev(y, ldy, y, ldy, x, ldx, m, n, s_d, st);
}

__global__ void k_d2f(const double *in, float *out, long n) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    out[g] = (float)in[g];
}
__global__ void k_f2d(const float *in, double *out, long n) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    out[g] = (double)in[g];
}
int convert_d2f(const double *in, float *out, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_d2f, dim3(4096), dim3(256), 0, st, in, out, n);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
int convert_f2d(const float *in, double *out, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_f2d, dim3(4096), dim3(256), 0, st, in, out, n);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// ---- small dense matrices of the subspace stage (n <= a few thousand)

// A <- scale * (A + A^T) / 2 for an n x n row-major matrix; the (i > j)
// thread of each pair writes both entries, the diagonal is only scaled
// A <- scale (A + A^T) / 2 for a row-major n x n matrix.  With blk > 0 only
// the blk x blk blocks on and above the block diagonal hold values: an
// element of a block below it takes its mirror above, unaveraged.
__global__ void k_symmetrize(double *A, long n, double scale, long blk) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long i = (long)blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= n || j >= n || j > i) return;
  if (i == j) {
    A[i * n + i] *= scale;
    return;
  }
  const double a = (blk > 0 && i / blk != j / blk) ? A[j * n + i] * scale
                                                   : 0.5 * (A[i * n + j] + A[j * n + i]) * scale;
  A[i * n + j] = a;
  A[j * n + i] = a;
}
int symmetrize_dev(double *A, long n, double scale, hipStream_t st, long blk) {
  const dim3 blk2(32, 8), grd((unsigned)((n + 31) / 32), (unsigned)((n + 7) / 8));
  hipLaunchKernelGGL(k_symmetrize, grd, blk2, 0, st, A, n, scale, blk);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// M[i * k + j] = V[c(j) * n + i] * (sc ? sc[j] : 1), c(j) = j or n - 1 - j:
// eigenvector columns (column-major V, as the eigensolver leaves them) to the
// row-major coefficient matrix of combine_dev, through a 32 x 32 LDS tile
__global__ void k_eigvec_cols(const double *V, long n, long k, const double *sc, int reverse,
                              double *M) {
  __shared__ double t[32][33];
  const long i0 = (long)blockIdx.x * 32, j0 = (long)blockIdx.y * 32;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {  // r: output column j0 + r
    const long j = j0 + r, i = i0 + threadIdx.x;
    if (j < k && i < n) t[r][threadIdx.x] = V[(reverse ? n - 1 - j : j) * n + i];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {  // r: row i0 + r
    const long i = i0 + r, j = j0 + threadIdx.x;
    if (i < n && j < k) M[i * k + j] = t[threadIdx.x][r] * (sc ? sc[j] : 1.0);
  }
}
int eigvec_cols_dev(const double *V, long n, long k, const double *sc_d, bool reverse, double *M,
                    hipStream_t st) {
  const dim3 blk(32, 8), grd((unsigned)((n + 31) / 32), (unsigned)((k + 31) / 32));
  hipLaunchKernelGGL(k_eigvec_cols, grd, blk, 0, st, V, n, k, sc_d, reverse ? 1 : 0, M);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// zero the strictly lower triangle (i > j) of an n x n row-major matrix
__global__ void k_zero_lower(double *A, long n) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long i = (long)blockIdx.y * blockDim.y + threadIdx.y;
  if (i < n && j < i) A[i * n + j] = 0.0;
}
int zero_lower_dev(double *A, long n, hipStream_t st) {
  const dim3 blk(32, 8), grd((unsigned)((n + 31) / 32), (unsigned)((n + 7) / 8));
  hipLaunchKernelGGL(k_zero_lower, grd, blk, 0, st, A, n);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

}  // namespace sep
