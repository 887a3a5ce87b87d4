// sep_gemm.hip -- matrix-core (MFMA) contractions:
//   Newton accumulation  acc[s][ie] += sum_k a[ie][k] psi_k[s]  (filter/filter.c:74-78)
//   Gram / projected H   C[i][j] = sum_g A[i][g] B[j][g]     (filter/Hmat.c:30-46,
//                                                           the SVD of ortho.c:32)
//   rotation             Out[j][g] = sum_i M[i][j] A[i][g]   (filter/Hmat.c:61-78)
// plus the double-double Gram that decides the SVD cutoff in FP64 exactly
// enough to resolve S_i/S_0 = 1e-10 (filter/ortho.c:41-46).  The Gram and
// rotation kernels serve both the FP64 subspace stage (f64 MFMA 16x16x4)
// and FP32 operands (f32 MFMA 16x16x4).
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
// C[i][j] = sum_k A[i*lda + k] B[j*ldb + k] for row-major operands whose rows
// run along the contraction index (states over grid points, or the rows of
// a small matrix).  Block: 256 threads = 2 x 2 waves, 128 x 128 tile of C;
// wave (wi, wj) owns the 64 x 64 quadrant = 4 x 4 MFMA tiles.  The
// contraction is permuted inside each 16-wide k block: lane group h = l/16
// takes k0 + 4h + s at MFMA step s (s < 4) for both operands, so a lane
// reads 4 consecutive elements of each of its 8 rows (one or two 16-byte
// loads) instead of 4 rows x 1 element -- the product is unchanged, the
// order of its partial sums is the k permutation.  Split K: chunk z of kc
// contraction indices writes its partial tile to work[z]; k_sum_chunks adds
// the chunks in fixed order (deterministic).  Block order is XCD-grouped:
// hardware dispatch places block b on XCD b % 8, and all tiles of one chunk
// are given to the same XCD, so their shared A/B rows meet in its L2.
constexpr int GT = 128;  // block tile
constexpr int GKB = 16;  // contraction block per MFMA group

template <typename R>
struct Vec4Load {
  // four consecutive elements at p (16-byte aligned when VEC)
  template <bool VEC>
  static __device__ __forceinline__ void load(const R *p, R (&v)[4]) {
    if constexpr (VEC && sizeof(R) == 4) {
      const float4 q = *reinterpret_cast<const float4 *>(p);
      v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
    } else if constexpr (VEC && sizeof(R) == 8) {
      const double2 q0 = *reinterpret_cast<const double2 *>(p);
      const double2 q1 = *reinterpret_cast<const double2 *>(p + 2);
      v[0] = q0.x, v[1] = q0.y, v[2] = q1.x, v[3] = q1.y;
    } else {
#pragma unroll
      for (int s = 0; s < 4; s++) v[s] = p[s];
    }
  }
};

// tile t of the T(T+1)/2 tiles on and above the diagonal of a T x T grid
__device__ __forceinline__ void upper_tile(int t, int T, int &bi, int &bj) {
  int i = 0;
  while (t >= T - i) {
    t -= T - i;
    i++;
  }
  bi = i;
  bj = i + t;
}

template <typename R, bool VEC, int TJ, int ST>
__global__ void __launch_bounds__(256) k_gram_t(const R *__restrict__ A, long lda, long m,
                                                const R *__restrict__ B, long ldb, long mb,
                                                long K, long kc, int tj, int ntiles, int nch,
                                                int sym, double *__restrict__ work) {
  // wave tile 64 x 16 TJ, block tile 128 x 32 TJ (2 x 2 waves)
  constexpr int BW = 32 * TJ;
  // XCD-grouped decode: block b -> (chunk, tile)
  const int b = blockIdx.x, xcd = b & 7, slot = b >> 3;
  const int cgrp = slot / ntiles, tile = slot - cgrp * ntiles;
  const int z = xcd + 8 * cgrp;
  if (z >= nch) return;
  int bi, bj;
  if (sym)
    upper_tile(tile, tj, bi, bj);
  else
    bi = tile / tj, bj = tile - (tile / tj) * tj;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, h = lane >> 4;
  const long i0 = (long)bi * GT + 64 * (w >> 1), j0 = (long)bj * BW + 16 * TJ * (w & 1);
  const long g0 = (long)z * kc;
  const long g1 = g0 + kc < K ? g0 + kc : K;
  // rows outside the operands read row 0 (their products land in C
  // entries that are never stored)
  const R *ap[4], *bp[TJ];
#pragma unroll
  for (int x = 0; x < 4; x++) {
    const long i = i0 + 16 * x + li;
    ap[x] = A + (i < m ? i : 0) * lda + 4 * h;
  }
#pragma unroll
  for (int y = 0; y < TJ; y++) {
    const long j = j0 + 16 * y + li;
    bp[y] = B + (j < mb ? j : 0) * ldb + 4 * h;
  }
  typename Mfma<R>::acc_t acc[4][TJ];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < TJ; y++) acc[x][y] = typename Mfma<R>::acc_t{0, 0, 0, 0};
  // ST-stage register pipeline: the loads of block n + ST - 1 are issued
  // before the MFMAs of block n
  R a[ST][4][4], bb[ST][TJ][4];
  auto load = [&](int st, long g) {
#pragma unroll
    for (int x = 0; x < 4; x++) Vec4Load<R>::template load<VEC>(ap[x] + g, a[st][x]);
#pragma unroll
    for (int y = 0; y < TJ; y++) Vec4Load<R>::template load<VEC>(bp[y] + g, bb[st][y]);
  };
  auto step = [&](int st) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < TJ; y++)
          acc[x][y] = Mfma<R>::op(a[st][x][s], bb[st][y][s], acc[x][y]);
  };
  long g = g0;
  const long gfull = g0 + ((g1 - g0) / GKB) * GKB;
  if (g < gfull) {
    // Invariant at the loop head: stage u < ST - 1 holds block g + u GKB.
    // The body is ST straight-line steps with no branch, so the wait before
    // the MFMAs of a stage covers only that stage's loads (the compiler
    // drains every outstanding load at a loop head whose incoming paths
    // differ).  Loads past the last full block re-read it (never used).
    const long glast = gfull - GKB;
#pragma unroll
    for (int st = 0; st < ST - 1; st++) load(st, g + st * GKB < gfull ? g + st * GKB : glast);
    for (; g + ST * GKB <= gfull;) {
#pragma unroll
      for (int u = 0; u < ST; u++) {
        const long gl = g + (ST - 1) * GKB;
        load((u + ST - 1) % ST, gl < gfull ? gl : glast);
        step(u);
        g += GKB;
      }
    }
    const long rest = (gfull - g) / GKB;  // < ST blocks, in stages 0 .. rest - 1
#pragma unroll
    for (int u = 0; u < ST - 1; u++)
      if (u < rest) step(u);
    g = gfull;
  }
  if (g < g1) {  // partial last block, element-wise with zero fill
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const bool in = g + 4 * h + s < g1;
#pragma unroll
      for (int x = 0; x < 4; x++) a[0][x][s] = in ? ap[x][g + s] : (R)0;
#pragma unroll
      for (int y = 0; y < TJ; y++) bb[0][y][s] = in ? bp[y][g + s] : (R)0;
    }
    step(0);
  }
  double *wz = work + (long)z * m * mb;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < TJ; y++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const long i = i0 + 16 * x + Mfma<R>::row(lane, r);
        const long j = j0 + 16 * y + li;
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

// Shape of one Gram launch.  Block tile GT x bw; chunk count a multiple of 8
// (one group per XCD) giving >= ~2048 blocks, chunks of at least 4096
// contraction indices.
struct GramShape {
  int tj, ntiles, nch;
  long kc;
};
static GramShape gram_shape(long m, long mb, long K, bool sym, int bw) {
  GramShape g;
  const int ti = (int)((m + GT - 1) / GT);
  if (sym) bw = GT;  // square tiles on the diagonal
  g.tj = (int)((mb + bw - 1) / bw);
  g.ntiles = sym ? ti * (ti + 1) / 2 : ti * g.tj;
  long want = (2048 + g.ntiles - 1) / g.ntiles;
  long maxch = (K + 4095) / 4096;
  if (want > maxch) want = maxch;
  if (want < 1) want = 1;
  want = (want + 7) / 8 * 8;
  g.kc = (K + want - 1) / want;
  g.kc = (g.kc + GKB - 1) / GKB * GKB;
  g.nch = (int)((K + g.kc - 1) / g.kc);
  return g;
}

// kernel configuration per precision (wave tile width TJ, pipeline depth
// ST), from sweeps at m = 512, N = 128^3 on MI355X (TFLOP/s, f64 / f32,
// box-to-box spread ~10 %): TJ 4: ST 2 42 / 58-77, ST 3 45-46 / 66-71, ST 4
// 44-46 / 57-70; TJ 2 (two waves per SIMD): 23-32 / 48-52.  A register-only
// probe of v_mfma_f64_16x16x4f64 (no memory traffic) issues at 36-48
// TFLOP/s on this part (more independent accumulators per wave run slower),
// while FP64 VALU FMAs reach 65: the f64 Gram at 45 runs at the matrix
// cores' practical FP64 rate, not at a memory limit.
template <typename R>
struct GramCfg;
template <>
struct GramCfg<double> {
  static constexpr int TJ = 4, ST = 3;
};
template <>
struct GramCfg<float> {
  static constexpr int TJ = 4, ST = 2;
};

long gram_work_len(long m, long mb, long K) {
  long n = 0;
  for (int bw : {32, 64, 128})
    for (bool sym : {false, true}) {
      const GramShape g = gram_shape(m, mb, K, sym, bw);
      if (g.nch > n) n = g.nch;
    }
  return n * m * mb;
}

template <typename R, int TJ, int ST>
int gram_launch(const R *A, long lda, long m, const R *B, long ldb, long mb, long K, bool sym,
                double *C_d, double *work, long work_len, hipStream_t st) {
  if (sym && (m != mb || TJ != 4)) return SEP_ERR_ARG;
  const GramShape g = gram_shape(m, mb, K, sym, 32 * TJ);
  if ((long)g.nch * m * mb > work_len) return SEP_ERR_NOMEM;
  const int ngrp = (g.nch + 7) / 8;
  const unsigned nblk = (unsigned)(8 * ngrp * g.ntiles);
  const bool vec = lda % 4 == 0 && ldb % 4 == 0 && (reinterpret_cast<uintptr_t>(A) % 16) == 0 &&
                   (reinterpret_cast<uintptr_t>(B) % 16) == 0;
  if (vec)
    hipLaunchKernelGGL((k_gram_t<R, true, TJ, ST>), dim3(nblk), dim3(256), 0, st, A, lda, m, B,
                       ldb, mb, K, g.kc, g.tj, g.ntiles, g.nch, sym ? 1 : 0, work);
  else
    hipLaunchKernelGGL((k_gram_t<R, false, TJ, ST>), dim3(nblk), dim3(256), 0, st, A, lda, m, B,
                       ldb, mb, K, g.kc, g.tj, g.ntiles, g.nch, sym ? 1 : 0, work);
  const long len = m * mb;
  unsigned nb = (unsigned)((len + 255) / 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_sum_chunks, dim3(nb), dim3(256), 0, st, work, (long)g.nch, len, C_d);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

template <typename R>
int gram_dev(const R *A, long lda, long m, const R *B, long ldb, long mb, long K, bool sym,
             double *C_d, double *work, long work_len, hipStream_t st) {
  if (sym && m != mb) return SEP_ERR_ARG;
  return gram_launch<R, GramCfg<R>::TJ, GramCfg<R>::ST>(A, lda, m, B, ldb, mb, K, sym, C_d, work,
                                                        work_len, st);
}
template int gram_dev<float>(const float *, long, long, const float *, long, long, long, bool,
                             double *, double *, long, hipStream_t);
template int gram_dev<double>(const double *, long, long, const double *, long, long, long, bool,
                              double *, double *, long, hipStream_t);

// ----------------------------------------------------------------- rotation
// Out[j*ldo + g] = sum_i M[i*ldm + j] A[i*lda + g] for i < m, j < mo, g < N:
// D = M^T A on the matrix cores, M^T the first MFMA operand (16 j x 4 i), A
// the second (4 i x 16 g).  Wave tile 64 (j) x 64 (g) = 4 x 4 MFMA tiles;
// block = 4 waves on one 64-point g tile and 4 consecutive j tiles, so A's
// rows are shared through the CU's cache.  The contraction over i is
// permuted as in the Gram (lane group h takes i0 + 4h + s at step s), and
// each lane reads two consecutive points per A row: column c of g-tile
// pair (2p, 2p + 1) stands for points g0 + 32p + 2c and + 1, so a lane's
// two results land in one 16-byte store.
template <typename R, bool VEC, int TG, int ST>
__global__ void __launch_bounds__(256) k_combine_t(const R *__restrict__ A, long lda, long m,
                                                   const double *__restrict__ M, long ldm,
                                                   long mo, long N, R *__restrict__ Out,
                                                   long ldo) {
  // wave tile 64 (j) x 16 TG (g); TG even
  constexpr int NP = TG / 2;  // point pairs per lane
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, h = lane >> 4;
  const long g0 = (long)blockIdx.x * 16 * TG;
  const long j0 = ((long)blockIdx.y * 4 + w) * 64;
  if (j0 >= mo) return;
  typename Mfma<R>::acc_t acc[4][TG];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < TG; y++) acc[x][y] = typename Mfma<R>::acc_t{0, 0, 0, 0};
  // this lane's points: g0 + 32p + 2li + e (p < NP, e < 2)
  long gp[NP];
  bool gin[NP][2];
#pragma unroll
  for (int p = 0; p < NP; p++) {
    gp[p] = g0 + 32 * p + 2 * li;
    gin[p][0] = gp[p] < N;
    gin[p][1] = gp[p] + 1 < N;
  }
  const bool full_g = g0 + 16 * TG <= N;
  // FP32: MFMA row r of j-tile x stands for output state j0 + 4 r + x, so
  // that a lane's four M entries per contraction index are consecutive (two
  // 16-byte loads instead of four scattered ones: 47-55 -> 63-72 TFLOP/s);
  // FP64 keeps j0 + 16 x + r (the permuted form measured 27-34 against
  // 34-50 TFLOP/s)
  constexpr bool PERM = sizeof(R) == 4;
  const long jl = j0 + 4 * li;
  const bool jfull = PERM && jl + 3 < mo && (ldm & 1) == 0 &&
                     (reinterpret_cast<uintptr_t>(M) & 15) == 0;
  R mv[ST][4][4], av[ST][TG][4];  // mv[x][s]: M^T(j, i), av[y][s]: A(i, g)
  auto load = [&](int st, long i0) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const long i = i0 + 4 * h + s;
      const bool iin = i < m;
      if (PERM) {
        const double *mr = M + (iin ? i : 0) * ldm + jl;
        if (jfull) {
          const double2 q0 = *reinterpret_cast<const double2 *>(mr);
          const double2 q1 = *reinterpret_cast<const double2 *>(mr + 2);
          mv[st][0][s] = iin ? (R)q0.x : (R)0;
          mv[st][1][s] = iin ? (R)q0.y : (R)0;
          mv[st][2][s] = iin ? (R)q1.x : (R)0;
          mv[st][3][s] = iin ? (R)q1.y : (R)0;
        } else {
#pragma unroll
          for (int x = 0; x < 4; x++) mv[st][x][s] = (iin && jl + x < mo) ? (R)mr[x] : (R)0;
        }
      } else {
#pragma unroll
        for (int x = 0; x < 4; x++) {
          const long j = j0 + 16 * x + li;
          mv[st][x][s] = (iin && j < mo) ? (R)M[i * ldm + j] : (R)0;
        }
      }
      const R *ar = A + (iin ? i : 0) * lda;
#pragma unroll
      for (int p = 0; p < NP; p++) {
        R e0, e1;
        if (VEC && full_g) {
          if constexpr (sizeof(R) == 8) {
            const double2 q = *reinterpret_cast<const double2 *>(ar + gp[p]);
            e0 = q.x, e1 = q.y;
          } else {
            const float2 q = *reinterpret_cast<const float2 *>(ar + gp[p]);
            e0 = q.x, e1 = q.y;
          }
        } else {
          e0 = gin[p][0] ? ar[gp[p]] : (R)0;
          e1 = gin[p][1] ? ar[gp[p] + 1] : (R)0;
        }
        av[st][2 * p][s] = iin ? e0 : (R)0;
        av[st][2 * p + 1][s] = iin ? e1 : (R)0;
      }
    }
  };
  // FP64: the 16-row sub-tiles x of this wave's j tile that hold any state
  // (all four unless the wave covers the tail of mo, as the second wave does
  // for a 65-state window); the rest are not multiplied
  const int xact = PERM ? 4 : (int)((mo - j0 + 15) / 16 < 4 ? (mo - j0 + 15) / 16 : 4);
  auto step = [&](int st) {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int x = 0; x < 4; x++)
        if (x < xact) {
#pragma unroll
          for (int y = 0; y < TG; y++)
            acc[x][y] = Mfma<R>::op(mv[st][x][s], av[st][y][s], acc[x][y]);
        }
  };
  // the same branch-free pipeline as the Gram's: at the loop head stage
  // u < ST - 1 holds i-block i0 + u GKB; blocks past the end re-read block 0
  long i0 = 0;
#pragma unroll
  for (int st = 0; st < ST - 1; st++) load(st, st * GKB < m ? st * GKB : 0);
  for (; i0 + ST * GKB <= m;) {
#pragma unroll
    for (int u = 0; u < ST; u++) {
      const long il = i0 + (ST - 1) * GKB;
      load((u + ST - 1) % ST, il < m ? il : 0);
      step(u);
      i0 += GKB;
    }
  }
  {
    // the rest: stages u < ST - 1 hold blocks i0 + u GKB (the final block
    // may be partial, load() zero-fills i >= m); blocks beyond them are
    // loaded and multiplied one at a time
    const long rest = (m - i0 + GKB - 1) / GKB;
#pragma unroll
    for (int u = 0; u < ST - 1; u++)
      if (u < rest) step(u);
    for (long ib = i0 + (ST - 1) * GKB; ib < m; ib += GKB) {
      load(0, ib);
      step(0);
    }
  }
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const long j = PERM ? j0 + 4 * Mfma<R>::row(lane, r) + x : j0 + 16 * x + Mfma<R>::row(lane, r);
      if (j >= mo) continue;
      R *op = Out + j * ldo;
#pragma unroll
      for (int p = 0; p < NP; p++) {
        const R e0 = (R)acc[x][2 * p][r], e1 = (R)acc[x][2 * p + 1][r];
        if (VEC && full_g) {
          if constexpr (sizeof(R) == 8)
            *reinterpret_cast<double2 *>(op + gp[p]) = double2{e0, e1};
          else
            *reinterpret_cast<float2 *>(op + gp[p]) = float2{e0, e1};
        } else {
          if (gin[p][0]) op[gp[p]] = e0;
          if (gin[p][1]) op[gp[p] + 1] = e1;
        }
      }
    }
}

// from the same sweeps (TFLOP/s, f64 / f32): TG 4: ST 1 34-50 / 47-55, ST 2
// 37-41 / 39-47, ST 3 40 / 40; TG 2: ST 1-3 34-45 / 37
template <typename R>
struct CombineCfg;
template <>
struct CombineCfg<double> {
  static constexpr int TG = 4, ST = 1;
};
template <>
struct CombineCfg<float> {
  static constexpr int TG = 4, ST = 1;
};

template <typename R, int TG, int ST>
int combine_launch(const R *A, long lda, long m, const double *M_d, long ldm, long mo, long N,
                   R *Out, long ldo, hipStream_t st) {
  dim3 grid((unsigned)((N + 16 * TG - 1) / (16 * TG)), (unsigned)((mo + 255) / 256));
  const bool vec = lda % 2 == 0 && ldo % 2 == 0 &&
                   (reinterpret_cast<uintptr_t>(A) % (2 * sizeof(R))) == 0 &&
                   (reinterpret_cast<uintptr_t>(Out) % (2 * sizeof(R))) == 0;
  if (vec)
    hipLaunchKernelGGL((k_combine_t<R, true, TG, ST>), grid, dim3(256), 0, st, A, lda, m, M_d, ldm,
                       mo, N, Out, ldo);
  else
    hipLaunchKernelGGL((k_combine_t<R, false, TG, ST>), grid, dim3(256), 0, st, A, lda, m, M_d,
                       ldm, mo, N, Out, ldo);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

template <typename R>
int combine_dev(const R *A, long lda, long m, const double *M_d, long ldm, long mo, long N, R *Out,
                long ldo, hipStream_t st) {
  return combine_launch<R, CombineCfg<R>::TG, CombineCfg<R>::ST>(A, lda, m, M_d, ldm, mo, N, Out,
                                                                 ldo, st);
}
template int combine_dev<float>(const float *, long, long, const double *, long, long, long,
                                float *, long, hipStream_t);
template int combine_dev<double>(const double *, long, long, const double *, long, long, long,
                                 double *, long, hipStream_t);

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

// contraction chunk of the double-double Gram: >= 8 chunks where N allows
static long dd_chunk(long N) {
  long kc = 65536;
  while (kc > 4096 && (N + kc - 1) / kc < 8) kc /= 2;
  return kc;
}

long gram_dd_work_len(long m, long N) {
  const long kc = dd_chunk(N);
  return 2 * ((N + kc - 1) / kc) * m * m;
}

int gram_dd_dev(const double *A, long m, long N, double *hi_d, double *lo_d, double *work,
                long work_len, hipStream_t st) {
  long kc = dd_chunk(N);
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
