// sep_spinor.hip -- the nonlocal spin-orbit potential of the reference's
// spinor filter (spinors/filterSpinors/hamiltonian.c:79-192, the matrix
// form, flagkb = 0) and the spinor helpers of the subspace stage.
//
// A spinor is stored as four real rows of N points: up.re, up.im, dn.re,
// dn.im.  The kinetic and local parts act on the two complex components as
// the scalar H.psi passes (two complex vectors per spinor: re/im planar);
// this file adds V_so psi.  Per atom a with Vso != 0, over the grid points
// p, q within Rnlcut of it (init.c:114-183):
//   K(p,q)   = (V_p + V_q)/2 exp(-((r_p - r_q) sigma_1)^2) / ((r2_p + r2_q)/2)
//   S_ms(p)  = sum_q K(p,q) conj(Y_m(q)) psi_s(q),  m in {11, 10, 1-1}
//   up(p)   += c [Y11 S_11,up - Y1-1 S_1-1,up + sqrt2 (Y10 S_11,dn + Y1-1 S_10,dn)]
//   dn(p)   += c [-Y11 S_11,dn - Y1-1 S_1-1,dn + sqrt2 (Y10 S_1-1,up + Y11 S_10,up)]
// with c = dv / (2 sqrt(2 pi sigma^2)), Y_m(p) the l = 1 harmonics at p.
// k_so_contrib computes each atom's contribution at each of its points
// (one workgroup per atom and spinor, the q points staged in LDS tiles);
// k_so_gather adds them into the output per grid point, atoms in order, as
// the reference's loop over atoms accumulates psinl -- no atomics, so the
// result is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "sep_internal.h"

namespace sep {

static constexpr int SO_T = 256;   // threads per workgroup
static constexpr int SO_QT = 128;  // q points per LDS tile

// The rounding of the spin-orbit term is spelled out (explicit fma, no
// other contraction), so that the kernel that evaluates the pair kernel
// inline and the one that reads it precomputed agree bit for bit.

// K(p, q) = (V_p + V_q)/2 exp(-((r_p - r_q) s1)^2) / ((r2_p + r2_q)/2)
__device__ __forceinline__ double so_pair(double rp, double r2p, double vp, double rq, double r2q,
                                          double vq, double s1) {
#pragma clang fp contract(off)
  const double rr = (rp - rq) * s1;
  return 0.5 * (vp + vq) * exp(-(rr * rr)) / (0.5 * (r2p + r2q));
}

// The 12 products conj(Y_m) psi_s of one point (g: its 9 geometry words;
// y11 at g[5], y10 at g[3], y1-1 at g[7]): w[6 s + 2 m + {re, im}], m = 11,
// 10, 1-1; s = up, dn.
__device__ __forceinline__ void so_point_w(const double *g, double ur, double ui, double dr,
                                           double di, double *w) {
#pragma clang fp contract(off)
  const int yo[3] = {5, 3, 7};
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const double yr = g[yo[m]], yi = g[yo[m] + 1];
    w[2 * m] = fma(yr, ur, yi * ui);
    w[2 * m + 1] = fma(yr, ui, -(yi * ur));
    w[6 + 2 * m] = fma(yr, dr, yi * di);
    w[6 + 2 * m + 1] = fma(yr, di, -(yi * dr));
  }
}

// The four outputs (up re, up im, dn re, dn im) of point p from its 12 sums
// S (hamiltonian.c:160-185).
__device__ __forceinline__ void so_point_out(const double *g, double c, const double *S,
                                             double *o) {
#pragma clang fp contract(off)
  const double t10r = c * g[3], t10i = c * g[4];
  const double t11r = c * g[5], t11i = c * g[6];
  const double t1_r = c * g[7], t1_i = c * g[8];
  const double s2 = 1.4142135623730950488;
  // S index: up 11 = 0, up 10 = 2, up 1-1 = 4; dn 11 = 6, dn 10 = 8, dn 1-1 = 10
#define CRE(tr, ti, k) fma((tr), S[k], -((ti) * S[(k) + 1]))
#define CIM(tr, ti, k) fma((ti), S[k], (tr) * S[(k) + 1])
  o[0] = fma(s2, CRE(t10r, t10i, 6) + CRE(t1_r, t1_i, 8), CRE(t11r, t11i, 0) - CRE(t1_r, t1_i, 4));
  o[1] = fma(s2, CIM(t10r, t10i, 6) + CIM(t1_r, t1_i, 8), CIM(t11r, t11i, 0) - CIM(t1_r, t1_i, 4));
  o[2] = fma(s2, CRE(t10r, t10i, 4) + CRE(t11r, t11i, 2), -CRE(t11r, t11i, 6) - CRE(t1_r, t1_i, 10));
  o[3] = fma(s2, CIM(t10r, t10i, 4) + CIM(t11r, t11i, 2), -CIM(t11r, t11i, 6) - CIM(t1_r, t1_i, 10));
#undef CRE
#undef CIM
}

__global__ void __launch_bounds__(SO_T) k_so_contrib(SoLists so, const double *in, long cs, long ss,
                                                     double *contrib, long nsp) {
  const int a = blockIdx.x;
  const long s = blockIdx.y;
  const long o0 = so.off[a], n = so.off[a + 1] - o0;
  const double *ps = in + s * ss;
  __shared__ double qr[SO_QT], qr2[SO_QT], qv[SO_QT];
  __shared__ __align__(16) double qw[SO_QT][12];
  const double s1 = so.s1, c = so.c;
  for (long p0 = 0; p0 < n; p0 += SO_T) {
    const long p = p0 + threadIdx.x;
    const bool act = p < n;
    double rp = 0.0, r2p = 1.0, vp = 0.0;
    if (act) {
      const double *g = so.geo + (o0 + p) * 9;
      rp = g[0];
      r2p = g[1];
      vp = g[2];
    }
    // S[2 (3 s + m) + {0, 1}]: m = 11, 10, 1-1; s = up, dn
    double S[12];
#pragma unroll
    for (int k = 0; k < 12; k++) S[k] = 0.0;
    for (long q0 = 0; q0 < n; q0 += SO_QT) {
      __syncthreads();  // the previous tile is consumed
      for (int t = threadIdx.x; t < SO_QT; t += SO_T) {
        const long q = q0 + t;
        if (q < n) {
          const double *g = so.geo + (o0 + q) * 9;
          const long j = so.jxyz[o0 + q];
          qr[t] = g[0];
          qr2[t] = g[1];
          qv[t] = g[2];
          so_point_w(g, ps[j], ps[cs + j], ps[2 * cs + j], ps[3 * cs + j], qw[t]);
        }
      }
      __syncthreads();
      const int qn = (int)(n - q0 < SO_QT ? n - q0 : SO_QT);
      if (act) {
        for (int t = 0; t < qn; t++) {
          const double K = so_pair(rp, r2p, vp, qr[t], qr2[t], qv[t], s1);
#pragma unroll
          for (int k = 0; k < 12; k++) S[k] = fma(K, qw[t][k], S[k]);
        }
      }
    }
    if (act) so_point_out(so.geo + (o0 + p) * 9, c, S, contrib + ((o0 + p) * nsp + s) * 4);
  }
}

__global__ void k_so_gather(SoLists so, const double *contrib, long nsp, double *out, long cs,
                            long ss, double alpha) {
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long s = blockIdx.y;
  if (u >= so.npts) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long e = so.poff[u]; e < so.poff[u + 1]; e++) {
    const double *c = contrib + ((long)so.pent[e] * nsp + s) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += c[k];
  }
  double *o = out + s * ss + so.pidx[u];
#pragma unroll
  for (int k = 0; k < 4; k++) o[k * cs] += alpha * acc[k];
}

// The pair kernel K(p, q) = (V_p + V_q)/2 exp(-((r_p - r_q) s1)^2) / ((r2_p +
// r2_q)/2) of each atom's points, the expression k_so_contrib evaluates in its
// inner loop, once at set-up: row q of atom a at kmat + koff[a] + q n (the
// lanes p of one q contiguous).  It depends on the geometry only.
__global__ void __launch_bounds__(SO_T) k_so_kmat(SoLists so) {
  const int a = blockIdx.x;
  const long o0 = so.off[a], n = so.off[a + 1] - o0;
  double *km = so.kmat + so.koff[a];
  const double s1 = so.s1;
  for (long p = threadIdx.x; p < n; p += SO_T) {
    const double *gp = so.geo + (o0 + p) * 9;
    const double rp = gp[0], r2p = gp[1], vp = gp[2];
    for (long q = 0; q < n; q++) {
      const double *gq = so.geo + (o0 + q) * 9;
      km[q * n + p] = so_pair(rp, r2p, vp, gq[0], gq[1], gq[2], s1);
    }
  }
}

// k_so_contrib with K read from the matrices of k_so_kmat: KT threads, NP
// points p per thread.  A 1-D grid over ceil(natom / 8) * 8 * nsp blocks:
// block b runs on XCD b % 8 (round-robin dispatch), and the blocks of one
// atom share that XCD and sit close in the dispatch order, so the atom's K
// is read from HBM once and from that XCD's L2 by its other spinors.
// Measured at 96^3 with 465 atoms (244-277 points each), 8 spinors: KT =
// 256, NP = 1 209 us per call (26 TFLOP/s FP64); KT = 128, NP = 2 (half the
// LDS bytes per FMA) 348 us; one wave per spinor with four spinors at 64
// shared points per workgroup (K from L2 a quarter as often) 216 us; the
// 12 products per q written to memory first (44 us) and read back at
// wave-uniform addresses through scalar loads, with no LDS (62 VGPRs,
// occupancy 8) 246 us; K of the next 8 q's loaded while the current 8 are
// consumed 199 against 208 us (not kept: no gain in the filter's time) --
// neither the LDS bytes per FMA nor the L2 stream of K sets its pace.  Each
// point's sums run over q in the order of k_so_contrib: the two kernels
// agree bit for bit.
template <int KT, int NP>
__global__ void __launch_bounds__(KT) k_so_contrib_k(SoLists so, const double *in, long cs,
                                                     long ss, double *contrib, long nsp) {
  const long b = blockIdx.x;
  const long x = b % 8, j = b / 8;
  const long s = j % nsp, a = (j / nsp) * 8 + x;
  if (a >= so.natom) return;  // padding blocks of the last group of eight
  const long o0 = so.off[a], n = so.off[a + 1] - o0;
  const double *km = so.kmat + so.koff[a];
  const double *ps = in + s * ss;
  __shared__ __align__(16) double qw[SO_QT][12];
  const double c = so.c;
  for (long p0 = 0; p0 < n; p0 += KT * NP) {
    // points past the list read row entries of the last point (in bounds)
    // and are not written
    long pc[NP];
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const long p = p0 + threadIdx.x + i * KT;
      pc[i] = p < n ? p : n - 1;
    }
    double S[NP][12];
#pragma unroll
    for (int i = 0; i < NP; i++)
#pragma unroll
      for (int k = 0; k < 12; k++) S[i][k] = 0.0;
    for (long q0 = 0; q0 < n; q0 += SO_QT) {
      __syncthreads();  // the previous tile is consumed
      for (int t = threadIdx.x; t < SO_QT; t += KT) {
        const long q = q0 + t;
        if (q < n) {
          const double *g = so.geo + (o0 + q) * 9;
          const long jq = so.jxyz[o0 + q];
          so_point_w(g, ps[jq], ps[cs + jq], ps[2 * cs + jq], ps[3 * cs + jq], qw[t]);
        }
      }
      __syncthreads();
      const int qn = (int)(n - q0 < SO_QT ? n - q0 : SO_QT);
      const double *kq = km + q0 * n;
#pragma unroll 4
      for (int t = 0; t < qn; t++) {
        double K[NP];
#pragma unroll
        for (int i = 0; i < NP; i++) K[i] = kq[t * n + pc[i]];
        double w[12];
#pragma unroll
        for (int k = 0; k < 12; k++) w[k] = qw[t][k];
#pragma unroll
        for (int i = 0; i < NP; i++)
#pragma unroll
          for (int k = 0; k < 12; k++) S[i][k] = fma(K[i], w[k], S[i][k]);
      }
    }
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const long p = p0 + threadIdx.x + i * KT;
      if (p < n) so_point_out(so.geo + (o0 + p) * 9, c, S[i], contrib + ((o0 + p) * nsp + s) * 4);
    }
  }
}

// k_so_gather's sums without the add: g[(u nsp + s) 4 + c], the entries of
// point u in atom order
__global__ void k_so_gather_g(SoLists so, const double *contrib, long nsp, double *g) {
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long s = blockIdx.y;
  if (u >= so.npts) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long e = so.poff[u]; e < so.poff[u + 1]; e++) {
    const double *c = contrib + ((long)so.pent[e] * nsp + s) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += c[k];
  }
  double *o = g + (u * nsp + s) * 4;
#pragma unroll
  for (int k = 0; k < 4; k++) o[k] = acc[k];
}

// so.pmap: -1 everywhere, then u at the grid index of covered point u
__global__ void k_so_pmap_fill(int *pmap, long n) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    pmap[g] = -1;
}
__global__ void k_so_pmap_set(SoLists so) {
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < so.npts) so.pmap[so.pidx[u]] = (int)u;
}

int so_pmap_build(SoLists &so, long N, hipStream_t st) {
  if (so.natom == 0 || so.npts == 0) return SEP_OK;
  if (hipMalloc(&so.pmap, sizeof(int) * N) != hipSuccess) {
    (void)hipGetLastError();
    so.pmap = nullptr;  // the spinor filter then runs its steps unfused
    return SEP_OK;
  }
  hipLaunchKernelGGL(k_so_pmap_fill, dim3(1024), dim3(256), 0, st, so.pmap, N);
  hipLaunchKernelGGL(k_so_pmap_set, dim3((unsigned)((so.npts + 255) / 256)), dim3(256), 0, st, so);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

int so_kmat_build(SoLists &so, const std::vector<long> &off, hipStream_t st) {
  if (so.natom == 0) return SEP_OK;
  std::vector<long> koff((size_t)so.natom + 1, 0);
  so.maxn = 0;
  for (int a = 0; a < so.natom; a++) {
    const long n = off[a + 1] - off[a];
    koff[a + 1] = koff[a] + n * n;
    so.maxn = std::max(so.maxn, n);
  }
  const size_t bytes = sizeof(double) * (size_t)koff[so.natom];
  if (bytes > ((size_t)8 << 30)) return SEP_OK;  // the inline kernel instead
  const char *env = getenv("SEP_SO_KMAT");  // 0: always the inline kernel
  if (env && atoi(env) == 0) return SEP_OK;
  if (hipMalloc(&so.kmat, bytes) != hipSuccess) {
    (void)hipGetLastError();
    so.kmat = nullptr;
    return SEP_OK;  // the inline kernel instead
  }
  if (hipMalloc(&so.koff, sizeof(long) * koff.size()) != hipSuccess ||
      hipMemcpy(so.koff, koff.data(), sizeof(long) * koff.size(), hipMemcpyHostToDevice) !=
          hipSuccess)
    return SEP_ERR_HIP;
  hipLaunchKernelGGL(k_so_kmat, dim3((unsigned)so.natom), dim3(SO_T), 0, st, so);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

static void so_contrib_launch(const SoLists &so, const double *in, long cs, long ss, long nsp,
                              double *contrib, hipStream_t st) {
  if (so.kmat) {
    const long nb = (long)((so.natom + 7) / 8) * 8 * nsp;
    hipLaunchKernelGGL((k_so_contrib_k<256, 1>), dim3((unsigned)nb), dim3(256), 0, st, so, in, cs,
                       ss, contrib, nsp);
  } else {
    hipLaunchKernelGGL(k_so_contrib, dim3((unsigned)so.natom, (unsigned)nsp), dim3(SO_T), 0, st,
                       so, in, cs, ss, contrib, nsp);
  }
}

int so_gather_dev(const SoLists &so, const double *in, long cs, long ss, long nsp, double *contrib,
                  double *g, hipStream_t st) {
  if (so.natom == 0 || nsp < 1) return SEP_OK;
  so_contrib_launch(so, in, cs, ss, nsp, contrib, st);
  hipLaunchKernelGGL(k_so_gather_g, dim3((unsigned)((so.npts + 255) / 256), (unsigned)nsp),
                     dim3(256), 0, st, so, contrib, nsp, g);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

int so_apply_dev(const SoLists &so, const double *in, long cs, long ss, double *out, long ocs,
                 long oss, long nsp, double alpha, double *contrib, hipStream_t st) {
  if (so.natom == 0 || nsp < 1) return SEP_OK;
  so_contrib_launch(so, in, cs, ss, nsp, contrib, st);
  hipLaunchKernelGGL(k_so_gather, dim3((unsigned)((so.npts + 255) / 256), (unsigned)nsp), dim3(256),
                     0, st, so, contrib, nsp, out, ocs, oss, alpha);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// Neighbourhood lists on the host (init.c:114-183), atoms with Vso != 0
// only (the others contribute exact zeros).  Grid coordinates are
// accumulated as the reference's init.c:467-469 does.
int so_build_host(int nx, int ny, int nz, const double d[3], const double x0[3], long natom,
                  const double *xyz, const double *vso, SoHost &h) {
  const double rc2 = 0.49 * 6.0 * std::log(10.0);  // init.c:99
  const double eps = 1.0e-20;                        // EPSDX, fd.h
  const double tmp1 = 0.5 * std::sqrt(3.0 / 3.14159265358979323846);
  const double tmp2 = 0.5 * std::sqrt(3.0 / 6.28318530717958647692);
  const double tmp49 = 1.0 / 0.49;
  const int n3[3] = {nx, ny, nz};
  std::vector<double> ax[3];
  for (int k = 0; k < 3; k++) {
    ax[k].resize(n3[k]);
    double v = x0[k];
    for (int j = 0; j < n3[k]; j++, v += d[k]) ax[k][j] = v;
  }
  h = SoHost();
  h.off.push_back(0);
  for (long a = 0; a < natom; a++) {
    if (vso[a] == 0.0) continue;
    const double R[3] = {xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]};
    // index window containing every point within the cutoff (then tested exactly)
    int lo[3], hi[3];
    const double rc = std::sqrt(rc2);
    for (int k = 0; k < 3; k++) {
      lo[k] = std::max(0, (int)std::floor((R[k] - rc - x0[k]) / d[k]) - 2);
      hi[k] = std::min(n3[k] - 1, (int)std::ceil((R[k] + rc - x0[k]) / d[k]) + 2);
    }
    for (int jz = lo[2]; jz <= hi[2]; jz++) {
      const double dz = ax[2][jz] - R[2], dzeps = dz + eps;
      for (int jy = lo[1]; jy <= hi[1]; jy++) {
        const double dy = ax[1][jy] - R[1], dyeps = dy + eps;
        for (int jx = lo[0]; jx <= hi[0]; jx++) {
          const double dx = ax[0][jx] - R[0], dxeps = dx + eps;
          const double dr2 = dx * dx + dy * dy + dz * dz;
          if (!(dr2 < rc2)) continue;
          h.jxyz.push_back((int)(((long)jz * ny + jy) * nx + jx));
          double g[9];
          g[0] = std::sqrt(dr2);
          g[1] = dr2 > eps ? dr2 : 1.0 / eps;
          g[2] = vso[a] * std::exp(-tmp49 * dr2);
          double r_1 = 1.0 / std::sqrt(dx * dx + dy * dy + dzeps * dzeps);
          g[3] = tmp1 * dzeps * r_1;
          g[4] = 0.0;
          r_1 = 1.0 / std::sqrt(dxeps * dxeps + dy * dy + dz * dz);
          g[5] = -tmp2 * dxeps * r_1;  // y11.re
          g[7] = tmp2 * dxeps * r_1;   // y1-1.re
          r_1 = 1.0 / std::sqrt(dx * dx + dyeps * dyeps + dz * dz);
          g[6] = -tmp2 * dyeps * r_1;  // y11.im
          g[8] = -tmp2 * dyeps * r_1;  // y1-1.im
          h.geo.insert(h.geo.end(), g, g + 9);
        }
      }
    }
    h.off.push_back((long)h.jxyz.size());
  }
  const long nent = (long)h.jxyz.size();
  // inverse map: entries per grid point, atoms in order (entries of one
  // atom are ascending in grid index, so a stable sort by grid index keeps
  // the atom order within each point)
  std::vector<int> ord(nent);
  for (long e = 0; e < nent; e++) ord[e] = (int)e;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return h.jxyz[x] < h.jxyz[y]; });
  h.pent = ord;
  h.poff.clear();
  h.pidx.clear();
  for (long e = 0; e < nent; e++) {
    if (e == 0 || h.jxyz[ord[e]] != h.jxyz[ord[e - 1]]) {
      h.pidx.push_back(h.jxyz[ord[e]]);
      h.poff.push_back(e);
    }
  }
  h.poff.push_back(nent);
  h.maxn = 0;
  for (size_t a = 0; a + 1 < h.off.size(); a++) h.maxn = std::max<long>(h.maxn, h.off[a + 1] - h.off[a]);
  return SEP_OK;
}

// ------------------------------------------------------------- helpers
// out[s] = i x[s] for spinors: (-up.im, up.re, -dn.im, dn.re)
__global__ void k_spinor_times_i(const double *x, double *out, long N, long m) {
  const long n4 = 4 * N;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < m * N;
       t += (long)gridDim.x * blockDim.x) {
    const long s = t / N, g = t % N;
    const double *a = x + s * n4;
    double *o = out + s * n4;
    o[g] = -a[N + g];
    o[N + g] = a[g];
    o[2 * N + g] = -a[3 * N + g];
    o[3 * N + g] = a[2 * N + g];
  }
}
int spinor_times_i_dev(const double *x, double *out, long N, long m, hipStream_t st) {
  hipLaunchKernelGGL(k_spinor_times_i, dim3(2048), dim3(256), 0, st, x, out, N, m);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// filter accumulators, real rows [(4 s + c) ms + e] -> spinors [(s ms + e)][c]
__global__ void k_spinor_gather_rows(const double *src, double *dst, long N, long nsp, long ms) {
  const long total = nsp * ms * 4 * N;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const long g = t % N, rest = t / N;
    const long c = rest % 4, se = rest / 4;
    const long s = se / ms, e = se % ms;
    dst[t] = src[((4 * s + c) * ms + e) * N + g];
  }
}
int spinor_gather_rows_dev(const double *src, double *dst, long N, long nsp, long ms,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_spinor_gather_rows, dim3(4096), dim3(256), 0, st, src, dst, N, nsp, ms);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// y = a x + b y over n doubles
__global__ void k_axpby(double a, const double *x, double b, double *y, long n) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    y[t] = a * x[t] + b * y[t];
}
int axpby_dev(double a, const double *x, double b, double *y, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby, dim3(2048), dim3(256), 0, st, a, x, b, y, n);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

}  // namespace sep
