// sep_gemm.hip -- matrix-core (MFMA) contractions:
//   Newton accumulation  acc[s][ie] += sum_k a[ie][k] psi_k[s]  (filter/filter.c:74-78)
//   Gram / projected H   C[i][j] = sum_g A[i][g] B[j][g]     (filter/Hmat.c:30-46,
//                                                           the SVD of ortho.c:32)
//   rotation             Out[j][g] = sum_i M[i][j] A[i][g]   (filter/Hmat.c:61-78)
// plus the double-double Gram that decides the SVD cutoff in FP64 exactly
// enough to resolve S_i/S_0 = 1e-10 (filter/ortho.c:41-46).  The Gram and
// rotation kernels serve FP32 operands; the FP64 subspace stage calls
// rocBLAS dgemm for those two plain products.
//
// MFMA 16x16x4 operand layout (wave64, lane l):  A[i = l%16][k = l/16],
// B[k = l/16][j = l%16]; accumulator element r < 4 of lane l holds
// D[i][j = l%16] with i = 4(l/16) + r for f32 but i = l/16 + 4r for f64
// (measured on gfx950).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sep_internal.h"

namespace sep {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename R>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = v4d;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  using acc_t = v4f;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ------------------------------------------------------------- split-K Gram
// Block: 256 threads, 64x64 tile of C, K chunk `kc` of the grid dimension.
// Wave w owns the 32x32 quadrant (w/2, w%2) = 2x2 MFMA tiles.
constexpr int GB = 64, GK = 32;

template <typename R>
__global__ void __launch_bounds__(256) k_gram(const R *__restrict__ A, long m,
                                              const R *__restrict__ B, long mb, long N, long kc,
                                              double *__restrict__ work) {
  __shared__ R As[GB][GK + 1];
  __shared__ R Bs[GB][GK + 1];
  const long i0 = (long)blockIdx.x * GB, j0 = (long)blockIdx.y * GB;
  const long g0 = (long)blockIdx.z * kc;
  long g1 = g0 + kc;
  if (g1 > N) g1 = N;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = w >> 1, wj = w & 1;
  typename Mfma<R>::acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = typename Mfma<R>::acc_t{0, 0, 0, 0};
  const int lr = t >> 2, lc = (t & 3) * 8;  // loader: row, 8 consecutive k
  for (long gk = g0; gk < g1; gk += GK) {
    __syncthreads();
    {
      const long ia = i0 + lr, jb = j0 + lr;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const long g = gk + lc + q;
        As[lr][lc + q] = (ia < m && g < g1) ? A[ia * N + g] : (R)0;
        Bs[lr][lc + q] = (jb < mb && g < g1) ? B[jb * N + g] : (R)0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      R a[2], b[2];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        a[x] = As[32 * wi + 16 * x + (lane & 15)][kk + (lane >> 4)];
        b[x] = Bs[32 * wj + 16 * x + (lane & 15)][kk + (lane >> 4)];
      }
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = Mfma<R>::op(a[x], b[y], acc[x][y]);
    }
  }
  double *wz = work + (long)blockIdx.z * m * mb;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long i = i0 + 32 * wi + 16 * x + Mfma<R>::row(lane, r);
        const long j = j0 + 32 * wj + 16 * y + (lane & 15);
        if (i < m && j < mb) wz[i * mb + j] = (double)acc[x][y][r];
      }
}

__global__ void k_sum_chunks(const double *work, long nchunk, long len, double *out) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (long c = 0; c < nchunk; c++) s += work[c * len + e];
    out[e] = s;
  }
}

static long gram_chunk(long N) {
  long kc = 65536;
  while (kc > 4096 && (N + kc - 1) / kc < 8) kc /= 2;
  return kc;
}

long gram_work_len(long m, long mb, long N) {
  long kc = gram_chunk(N);
  return ((N + kc - 1) / kc) * m * mb;
}

template <typename R>
int gram_dev(const R *A, long m, const R *B, long mb, long N, double *C_d, double *work,
             long work_len, hipStream_t st) {
  long kc = gram_chunk(N);
  long nch = (N + kc - 1) / kc;
  if (nch * m * mb > work_len) return SEP_ERR_NOMEM;
  dim3 grid((unsigned)((m + GB - 1) / GB), (unsigned)((mb + GB - 1) / GB), (unsigned)nch);
  hipLaunchKernelGGL(k_gram<R>, grid, dim3(256), 0, st, A, m, B, mb, N, kc, work);
  long len = m * mb;
  unsigned nb = (unsigned)((len + 255) / 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_sum_chunks, dim3(nb), dim3(256), 0, st, work, nch, len, C_d);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
template int gram_dev<float>(const float *, long, const float *, long, long, double *, double *,
                             long, hipStream_t);

// ----------------------------------------------------------------- rotation
// Out[j][g] = sum_i M[i*ldm + j] A[i][g]; block tile 64 (j) x 128 (g),
// reduction chunks of 16 states; wave w owns g columns [32w, 32w+32).
constexpr int CJ = 64, CG = 128, CI = 16;

template <typename R>
__global__ void __launch_bounds__(256) k_combine(const R *__restrict__ A, long m,
                                                 const double *__restrict__ M, long ldm, long mo,
                                                 long N, R *__restrict__ Out) {
  __shared__ R Ms[CI][CJ + 1];  // Ms[i][j]
  __shared__ R As[CI][CG + 4];
  const long j0 = (long)blockIdx.y * CJ, g0 = (long)blockIdx.x * CG;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  typename Mfma<R>::acc_t acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = typename Mfma<R>::acc_t{0, 0, 0, 0};
  for (long i0 = 0; i0 < m; i0 += CI) {
    __syncthreads();
    // A tile: 16 x 128, 8 per thread along g
    {
      const int r = t >> 4, c = (t & 15) * 8;
      const long i = i0 + r;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const long g = g0 + c + q;
        As[r][c + q] = (i < m && g < N) ? A[i * N + g] : (R)0;
      }
    }
    // M tile: 16 x 64, 4 per thread
    {
      const int r = t >> 4, c = (t & 15) * 4;
      const long i = i0 + r;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long j = j0 + c + q;
        Ms[r][c + q] = (i < m && j < mo) ? (R)M[i * ldm + j] : (R)0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < CI; kk += 4) {
      R a[4], b[2];
#pragma unroll
      for (int x = 0; x < 4; x++) a[x] = Ms[kk + (lane >> 4)][16 * x + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 2; y++) b[y] = As[kk + (lane >> 4)][32 * w + 16 * y + (lane & 15)];
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = Mfma<R>::op(a[x], b[y], acc[x][y]);
    }
  }
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long j = j0 + 16 * x + Mfma<R>::row(lane, r);
        const long g = g0 + 32 * w + 16 * y + (lane & 15);
        if (j < mo && g < N) Out[j * N + g] = acc[x][y][r];
      }
}

template <typename R>
int combine_dev(const R *A, long m, const double *M_d, long ldm, long mo, long N, R *Out,
                hipStream_t st) {
  dim3 grid((unsigned)((N + CG - 1) / CG), (unsigned)((mo + CJ - 1) / CJ));
  hipLaunchKernelGGL(k_combine<R>, grid, dim3(256), 0, st, A, m, M_d, ldm, mo, N, Out);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
template int combine_dev<float>(const float *, long, const double *, long, long, long, float *,
                                hipStream_t);

// ------------------------------------------------ Newton accumulation GEMM
// acc[s][ie][g] += sum_{k < cnt} anT[j0 + k][ie] * ring[k][s][g]
// (filter/filter.c:74-78 for cnt stored Newton vectors at once) as a matrix
// product on the matrix cores: targets x K times K x points.  Block: 256
// threads, 32 targets x 256 points; wave w owns 64 points (4 MFMA column
// tiles) x 32 targets (2 row tiles).  The K-block partial is formed in R
// (FP64 for the FP64 filter, FP32 for the complex64 filter) and added to the
// FP64 accumulators.
template <typename R>
__global__ void __launch_bounds__(256) k_accum_mfma(const R *__restrict__ ring, long slot,
                                                    long sstride, int cnt,
                                                    const double *__restrict__ anT, long j0,
                                                    int ms, double *__restrict__ acc, long N) {
  const long s = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long g0 = (long)blockIdx.x * 256 + w * 64;
  const int ie0 = blockIdx.z * 32;
  const int kq = lane >> 4, li = lane & 15;
  const R *rp = ring + s * sstride;
  typename Mfma<R>::acc_t d[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) d[x][y] = typename Mfma<R>::acc_t{0, 0, 0, 0};
#pragma unroll 4
  for (int k = 0; k < cnt; k += 4) {
    const int kk = k + kq;
    const bool kin = kk < cnt;
    R a[2], b[4];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const int ie = ie0 + 16 * x + li;
      a[x] = (kin && ie < ms) ? (R)anT[(j0 + kk) * ms + ie] : (R)0;
    }
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const long g = g0 + 16 * y + li;
      b[y] = (kin && g < N) ? rp[(long)kk * slot + g] : (R)0;
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) d[x][y] = Mfma<R>::op(a[x], b[y], d[x][y]);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ie = ie0 + 16 * x + Mfma<R>::row(lane, r);
        const long g = g0 + 16 * y + li;
        if (ie < ms && g < N) acc[(s * ms + ie) * N + g] += (double)d[x][y][r];
      }
}

// The same product with 16-byte loads: lane li of a wave loads VW = 16 /
// sizeof(R) consecutive points of one ring row per instruction, and MFMA
// column j = li of y-tile y stands for point g0 + (y / VW) 16 VW + VW li +
// y % VW, so that a wave's 64 points of one row come in 4 / VW loads of
// 16 x 16 bytes and its accumulator rows are updated in VW-point runs.
// Needs N % 256 == 0 (every block full).
template <typename R>
__global__ void __launch_bounds__(256) k_accum_mfma_v(const R *__restrict__ ring, long slot,
                                                      long sstride, int cnt, const double *__restrict__ anT,
                                                      long j0, int ms, double *__restrict__ acc,
                                                      long N) {
  constexpr int VW = 16 / sizeof(R);
  constexpr int NH = 4 / VW;  // 16-byte loads per row and k-step
  using V = typename std::conditional<sizeof(R) == 4, float4, double2>::type;
  const long s = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long g0 = (long)blockIdx.x * 256 + w * 64;
  const int ie0 = blockIdx.z * 32;
  const int kq = lane >> 4, li = lane & 15;
  const R *rp = ring + s * sstride + g0 + VW * li;
  typename Mfma<R>::acc_t d[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) d[x][y] = typename Mfma<R>::acc_t{0, 0, 0, 0};
  // one k-step: 4 ring rows (kq) x 64 points, 2 x 4 MFMA tiles
  auto load = [&](int kk, bool kin, R *a, R *b) {
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const int ie = ie0 + 16 * x + li;
      a[x] = (kin && ie < ms) ? (R)anT[(j0 + kk) * ms + ie] : (R)0;
    }
#pragma unroll
    for (int h = 0; h < NH; h++) {
      V v{};
      if (kin) v = *reinterpret_cast<const V *>(rp + (long)kk * slot + h * 16 * VW);
      const R *vp = reinterpret_cast<const R *>(&v);
#pragma unroll
      for (int q = 0; q < VW; q++) b[h * VW + q] = vp[q];
    }
  };
  auto mma = [&](const R *a, const R *b) {
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) d[x][y] = Mfma<R>::op(a[x], b[y], d[x][y]);
  };
  // U k-steps per iteration with all their loads issued first (more bytes in
  // flight per wave), then the remainder one step at a time
  constexpr int U = 4;
  int k = 0;
  for (; k + 4 * U <= cnt; k += 4 * U) {
    R a[U][2], b[U][4];
#pragma unroll
    for (int u = 0; u < U; u++) load(k + 4 * u + kq, true, a[u], b[u]);
#pragma unroll
    for (int u = 0; u < U; u++) mma(a[u], b[u]);
  }
  for (; k < cnt; k += 4) {
    R a[2], b[4];
    load(k + kq, k + kq < cnt, a, b);
    mma(a, b);
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int ie = ie0 + 16 * x + Mfma<R>::row(lane, r);
      if (ie >= ms) continue;
      double *ap = acc + (s * ms + ie) * N + g0 + VW * li;
#pragma unroll
      for (int h = 0; h < NH; h++) {
        double2 *p2 = reinterpret_cast<double2 *>(ap + h * 16 * VW);
#pragma unroll
        for (int q = 0; q < VW / 2; q++) {
          double2 o = p2[q];
          o.x += (double)d[x][h * VW + 2 * q][r];
          o.y += (double)d[x][h * VW + 2 * q + 1][r];
          p2[q] = o;
        }
      }
    }
}

template <typename R>
int accum_mfma_dev(const R *ring, long slot, long sstride, int cnt, const double *anT_d, long j0,
                   int ms, double *acc, int ns, long N, hipStream_t st) {
  dim3 grid((unsigned)((N + 255) / 256), (unsigned)ns, (unsigned)((ms + 31) / 32));
  if (N % 256 == 0 && slot % 4 == 0 && sstride % 4 == 0)
    hipLaunchKernelGGL(k_accum_mfma_v<R>, grid, dim3(256), 0, st, ring, slot, sstride, cnt, anT_d,
                       j0, ms, acc, N);
  else
    hipLaunchKernelGGL(k_accum_mfma<R>, grid, dim3(256), 0, st, ring, slot, sstride, cnt, anT_d, j0,
                       ms, acc, N);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
template int accum_mfma_dev<double>(const double *, long, long, int, const double *, long, int,
                                    double *,
                                    int, long, hipStream_t);
template int accum_mfma_dev<float>(const float *, long, long, int, const double *, long, int,
                                   double *,
                                   int, long, hipStream_t);

// The complex64 filter's accumulation with the coefficients packed as FP32:
// for target group tz (32 targets) and Newton rows 16 kb .. 16 kb + 15,
// cf[tz][kb][kq][li][u][x] = a[32 tz + 16 x + li][16 kb + 4 u + kq], so that
// the 4 k-steps x 2 targets one lane needs per 16 rows are 32 contiguous
// bytes (two 16-byte loads, in place of eight scattered FP64 loads and
// conversions: 22.0 -> 17.5 ms per 512-row block at 128^3, measured).
long accum_pack_len(long nc, int ms) { return (long)((ms + 31) / 32) * ((nc + 15) / 16) * 16 * 32; }

void accum_pack_f32(const double *an, long nc, int ms, float *cf) {
  const long nkp = (nc + 15) / 16 * 16;
  for (int tz = 0; tz < (ms + 31) / 32; tz++)
    for (long k = 0; k < nkp; k++)
      for (int il = 0; il < 32; il++) {
        const int ie = 32 * tz + il, x = il / 16, li = il % 16;
        const long kb = k / 16, u = (k % 16) / 4, kq = k % 4;
        cf[(long)tz * nkp * 32 + ((kb * 4 + kq) * 16 + li) * 8 + u * 2 + x] =
            (k < nc && ie < ms) ? (float)an[(long)ie * nc + k] : 0.0f;
      }
}

__global__ void __launch_bounds__(256) k_accum_mfma_p(const float *__restrict__ ring, long slot,
                                                      long sstride, int cnt,
                                                      const float *__restrict__ cf, long nkp,
                                                      long j0, int ms, double *__restrict__ acc,
                                                      long N) {
  const long s = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long g0 = (long)blockIdx.x * 256 + w * 64;
  const int ie0 = blockIdx.z * 32;
  const int kq = lane >> 4, li = lane & 15;
  const float *rp = ring + s * sstride + g0 + 4 * li;
  const float *cb = cf + (long)blockIdx.z * nkp * 32 + (long)(kq * 16 + li) * 8;
  v4f d[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) d[x][y] = v4f{0, 0, 0, 0};
  // 16 Newton rows: ring rows k + 4u + kq (u < 4) of 64 points, 2 x 4 MFMA tiles each
  auto rows16 = [&](int k, bool full) {
    v4f v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int kk = k + 4 * u + kq;
      v[u] = (full || kk < cnt) ? *reinterpret_cast<const v4f *>(rp + (long)kk * slot)
                                : v4f{0, 0, 0, 0};
    }
    const v4f *cp = reinterpret_cast<const v4f *>(cb + ((j0 + k) / 16) * 512);
    const v4f c0 = cp[0], c1 = cp[1];
    const float a[4][2] = {{c0[0], c0[1]}, {c0[2], c0[3]}, {c1[0], c1[1]}, {c1[2], c1[3]}};
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) d[x][y] = Mfma<float>::op(a[u][x], v[u][y], d[x][y]);
  };
  int k = 0;
  for (; k + 16 <= cnt; k += 16) rows16(k, true);
  if (k < cnt) rows16(k, false);
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int ie = ie0 + 16 * x + Mfma<float>::row(lane, r);
      if (ie >= ms) continue;
      double2 *p2 = reinterpret_cast<double2 *>(acc + (s * ms + ie) * N + g0 + 4 * li);
#pragma unroll
      for (int q = 0; q < 2; q++) {
        double2 o = p2[q];
        o.x += (double)d[x][2 * q][r];
        o.y += (double)d[x][2 * q + 1][r];
        p2[q] = o;
      }
    }
}

int accum_mfma_packed_dev(const float *ring, long slot, long sstride, int cnt, const float *cf_d,
                          long nc, long j0, int ms, double *acc, int ns, long N, hipStream_t st) {
  if (N % 256 != 0 || slot % 4 != 0 || sstride % 4 != 0 || j0 % 16 != 0 || j0 + cnt > nc)
    return SEP_ERR_UNSUPPORTED;
  const long nkp = (nc + 15) / 16 * 16;
  dim3 grid((unsigned)(N / 256), (unsigned)ns, (unsigned)((ms + 31) / 32));
  hipLaunchKernelGGL(k_accum_mfma_p, grid, dim3(256), 0, st, ring, slot, sstride, cnt, cf_d, nkp,
                     j0, ms, acc, N);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// ---------------------------------------------------- double-double Gram
// Each thread owns one (i, j) of a 16x16 tile and accumulates
// sum_g A[i][g] A[j][g] over its K chunk as an unevaluated sum hi + lo
// (TwoProduct by fma, TwoSum).  Output per chunk: hi, lo.
// Contraction must stay off here: fusing a*b into hi + p would destroy the
// error terms of TwoProduct/TwoSum.
__device__ __forceinline__ void dd_add_prod(double &hi, double &lo, double a, double b) {
#pragma clang fp contract(off)
  double p = a * b;
  double pe = __builtin_fma(a, b, -p);
  double s = hi + p;
  double bb = s - hi;
  double se = (hi - (s - bb)) + (p - bb);
  hi = s;
  lo += se + pe;
}
__device__ __forceinline__ void dd_norm(double &hi, double &lo) {
#pragma clang fp contract(off)
  double s = hi + lo;
  lo = lo - (s - hi);
  hi = s;
}

__global__ void __launch_bounds__(256) k_gram_dd(const double *__restrict__ A, long m, long N,
                                                 long kc, double *__restrict__ whi,
                                                 double *__restrict__ wlo) {
  constexpr int KT = 128;
  __shared__ double Ai[16][KT + 1], Aj[16][KT + 1];
  // tiles strictly below the diagonal are left to the caller, who mirrors
  // the upper ones (G is symmetric)
  if (blockIdx.y < blockIdx.x) return;
  const long i0 = (long)blockIdx.x * 16, j0 = (long)blockIdx.y * 16;
  const long g0 = (long)blockIdx.z * kc;
  long g1 = g0 + kc;
  if (g1 > N) g1 = N;
  const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
  double hi = 0.0, lo = 0.0;
  for (long gk = g0; gk < g1; gk += KT) {
    __syncthreads();
    for (int e = threadIdx.x; e < 16 * KT; e += 256) {
      const int r = e / KT, c = e % KT;
      const long g = gk + c;
      Ai[r][c] = (i0 + r < m && g < g1) ? A[(i0 + r) * N + g] : 0.0;
      Aj[r][c] = (j0 + r < m && g < g1) ? A[(j0 + r) * N + g] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int c = 0; c < KT; c++) dd_add_prod(hi, lo, Ai[ti][c], Aj[tj][c]);
    dd_norm(hi, lo);
  }
  const long i = i0 + ti, j = j0 + tj;
  if (i < m && j < m) {
    whi[(long)blockIdx.z * m * m + i * m + j] = hi;
    wlo[(long)blockIdx.z * m * m + i * m + j] = lo;
  }
}

__global__ void k_sum_chunks_dd(const double *whi, const double *wlo, long nchunk, long len,
                                double *ohi, double *olo) {
#pragma clang fp contract(off)
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (long)gridDim.x * blockDim.x) {
    double hi = 0.0, lo = 0.0;
    for (long c = 0; c < nchunk; c++) {
      double x = whi[c * len + e];
      double s = hi + x, bb = s - hi;
      lo += (hi - (s - bb)) + (x - bb);
      hi = s;
      lo += wlo[c * len + e];
    }
    double s = hi + lo;
    lo = lo - (s - hi);
    ohi[e] = s;
    olo[e] = lo;
  }
}

long gram_dd_work_len(long m, long N) { return 2 * gram_work_len(m, m, N); }

int gram_dd_dev(const double *A, long m, long N, double *hi_d, double *lo_d, double *work,
                long work_len, hipStream_t st) {
  long kc = gram_chunk(N);
  long nch = (N + kc - 1) / kc;
  if (2 * nch * m * m > work_len) return SEP_ERR_NOMEM;
  double *whi = work, *wlo = work + nch * m * m;
  dim3 grid((unsigned)((m + 15) / 16), (unsigned)((m + 15) / 16), (unsigned)nch);
  hipLaunchKernelGGL(k_gram_dd, grid, dim3(256), 0, st, A, m, N, kc, whi, wlo);
  long len = m * m;
  unsigned nb = (unsigned)((len + 255) / 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_sum_chunks_dd, dim3(nb), dim3(256), 0, st, whi, wlo, nch, len, hi_d, lo_d);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

}  // namespace sep
