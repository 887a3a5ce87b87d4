// sep_adjoint.hip -- Hellmann-Feynman adjoint of V(r): contractions of state
// densities rho_p(g) = psi_i(g) psi_j(g) against the derivatives of the local
// potential V(g) = sum_a s_a interp_{t(a)}(|r_g - R_a|) (filter/init.c:161-182,
// filter/interpolate.c:11-23, strain scale of filterStrainCube/interpolate.c:20-21)
// with respect to the atom positions R_a, the per-atom scales s_a and the
// tabulated values pot_t[k]:
//
//   dR[p][a]    = dv sum_g rho_p(g) dV(g)/dR_a
//               = dv sum_g rho_p(g) (-aa_a(g)) (r_g - R_a) / |r_g - R_a|
//   dscale[p][a]= dv sum_g rho_p(g) interp_{t(a)}(|r_g - R_a|)
//   dtab[p][t][k] = dv sum_g rho_p(g) dV(g)/dpot_t[k]
//
// With rho_n = psi_n^2 these are dE_n/d(.) for a converged eigenpair
// (Hellmann-Feynman); with i != j the couplings <psi_i|dV/dR_a|psi_j> of
// nonAdiabaticCoupling/nonadiabatic.c:143-301.  The derivative is that of the
// forward model: segment k = (long)(del / dr_t), zero beyond row npot_t - 2,
// linear interp = (1 - w) pot[k] + w pot[k+1] with w = (del - vr[k]) /
// (vr[k+1] - vr[k]), or (SEP_INTERP_CUBIC) the cubic Hermite interpolant of
// cubic_weights()/cubic_rows() on rows k-1 .. k+2, whose derivative in del
// is continuous.
//
// Work split: one block per (atom, slice of the atom's bounding box of grid
// points within the table range); each block reduces its partial sums over
// the box slice in fixed order (written to a work array and summed in split
// order by a second kernel), and accumulates its table histogram in LDS,
// flushed with global atomic adds (so dtab is reproducible to rounding only).
#include <hip/hip_runtime.h>

#include "sep_internal.h"

namespace sep {

constexpr int HF_THREADS = 256;

template <int PB>
__global__ void __launch_bounds__(HF_THREADS) k_hf_atom(
    int nx, int ny, int nz, double xmin, double ymin, double zmin, double dx, double dy,
    double dz, const double *__restrict__ xyz, const int *__restrict__ type,
    const double *__restrict__ scale, const long *__restrict__ npot, const double *__restrict__ vr,
    const double *__restrict__ pot, long stride, const double *__restrict__ dr,
    const double *__restrict__ rho, long N, int np, int p0, int ntype, int nsplit, int interp,
    double *__restrict__ work, double *__restrict__ htab) {
  extern __shared__ double hist[];  // [PB][npot_t]
  const int a = blockIdx.x, split = blockIdx.y;
  const int t = type[a];
  const long nt = npot[t];
  const double drt = dr[t];
  const double rmax = (double)(nt - 1) * drt;
  const double Rx = xyz[3 * a], Ry = xyz[3 * a + 1], Rz = xyz[3 * a + 2];
  const double sa = scale ? scale[a] : 1.0;
  const double *vt = vr + t * stride, *pt = pot + t * stride;
  for (long k = threadIdx.x; k < (long)PB * nt; k += HF_THREADS) hist[k] = 0.0;
  __syncthreads();
  // bounding box of grid points within rmax of the atom
  auto lo = [](double c, double r, double o, double d) { return (int)ceil((c - r - o) / d); };
  auto hi = [](double c, double r, double o, double d) { return (int)floor((c + r - o) / d); };
  const int x0 = max(0, lo(Rx, rmax, xmin, dx)), x1 = min(nx - 1, hi(Rx, rmax, xmin, dx));
  const int y0 = max(0, lo(Ry, rmax, ymin, dy)), y1 = min(ny - 1, hi(Ry, rmax, ymin, dy));
  const int z0 = max(0, lo(Rz, rmax, zmin, dz)), z1 = min(nz - 1, hi(Rz, rmax, zmin, dz));
  const long Lx = x1 >= x0 ? x1 - x0 + 1 : 0, Ly = y1 >= y0 ? y1 - y0 + 1 : 0,
             Lz = z1 >= z0 ? z1 - z0 + 1 : 0;
  const long tot = Lx * Ly * Lz;
  const long chunk = (tot + nsplit - 1) / nsplit;
  const long b0 = (long)split * chunk, b1 = min(tot, b0 + chunk);
  double acc[PB][4];
#pragma unroll
  for (int q = 0; q < PB; q++) acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0.0;
  for (long b = b0 + threadIdx.x; b < b1; b += HF_THREADS) {
    const int ix = x0 + (int)(b % Lx);
    const int iy = y0 + (int)((b / Lx) % Ly);
    const int iz = z0 + (int)(b / (Lx * Ly));
    const double ddx = (xmin + ix * dx) - Rx, ddy = (ymin + iy * dy) - Ry,
                 ddz = (zmin + iz * dz) - Rz;
    const double del = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    const long k = (long)(del / drt);
    if (k > nt - 2) continue;
    const double span = vt[k + 1] - vt[k];
    const double w = (del - vt[k]) / span;
    // value, d(value)/d(del) without the scale, and the table-row weights
    long rows[4] = {k, k + 1, k + 1, k + 1};
    double wr[4] = {1.0 - w, w, 0.0, 0.0};
    double val, dval;
    if (interp == SEP_INTERP_CUBIC) {
      double cw[4], cd[4], dr4[4];
      cubic_weights(w, cw, cd);
      cubic_rows(k, nt, cw, rows, wr);
      cubic_rows(k, nt, cd, rows, dr4);
      val = dval = 0.0;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        val += wr[c] * pt[rows[c]];
        dval += dr4[c] * pt[rows[c]];
      }
      dval /= span;
    } else {
      val = (1.0 - w) * pt[k] + w * pt[k + 1];
      dval = (pt[k + 1] - pt[k]) / span;
    }
    const double g = del > 0.0 ? -sa * dval / del : 0.0;
    const long gi = ((long)iz * ny + iy) * nx + ix;
#pragma unroll
    for (int q = 0; q < PB; q++) {
      if (p0 + q >= np) break;
      const double r = rho[(long)(p0 + q) * N + gi];
      acc[q][0] += r * g * ddx;
      acc[q][1] += r * g * ddy;
      acc[q][2] += r * g * ddz;
      acc[q][3] += r * val;
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (wr[c] != 0.0) atomicAdd(&hist[q * nt + rows[c]], r * sa * wr[c]);
    }
  }
  // block reduction of the partial sums (wave shuffles, then LDS across waves)
  __shared__ double red[HF_THREADS / 64][PB * 4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < PB; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      double v = acc[q][c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[wv][q * 4 + c] = v;
    }
  __syncthreads();
  if (threadIdx.x < PB * 4) {
    double v = 0.0;
    for (int w2 = 0; w2 < HF_THREADS / 64; w2++) v += red[w2][threadIdx.x];
    work[((long)a * nsplit + split) * PB * 4 + threadIdx.x] = v;
  }
  // table histogram -> global (only rows this block touched are non-zero)
  for (long k = threadIdx.x; k < (long)PB * nt; k += HF_THREADS) {
    const int q = (int)(k / nt);
    const double v = hist[k];
    if (v != 0.0 && p0 + q < np)
      atomicAdd(&htab[((long)(p0 + q) * ntype + t) * stride + (k - (long)q * nt)], v);
  }
}

// Sum the per-split partials in split order and scale by dv.
template <int PB>
__global__ void k_hf_reduce(const double *__restrict__ work, int natom, int nsplit, int np, int p0,
                            double dv, double *__restrict__ dR, double *__restrict__ dscale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // (atom, q)
  if (i >= (long)natom * PB) return;
  const int a = (int)(i / PB), q = (int)(i % PB);
  if (p0 + q >= np) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int sp = 0; sp < nsplit; sp++)
    for (int c = 0; c < 4; c++) s[c] += work[((long)a * nsplit + sp) * PB * 4 + q * 4 + c];
  const long o = (long)(p0 + q) * natom + a;
  if (dR)
    for (int c = 0; c < 3; c++) dR[o * 3 + c] = s[c] * dv;
  if (dscale) dscale[o] = s[3] * dv;
}

// rho_p(g) = psi_i(g) psi_j(g) for np pairs, or, with weights, the single
// density rho(g) = sum_p w_p psi_i(g) psi_j(g) (summed in pair order).
__global__ void k_pair_density(const double *__restrict__ states, long N,
                               const long *__restrict__ pairs, const double *__restrict__ wts,
                               int np, double *__restrict__ rho) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < N;
       g += (long)gridDim.x * blockDim.x) {
    if (wts) {
      double s = 0.0;
      for (int p = 0; p < np; p++)
        s += wts[p] * states[pairs[2 * p] * N + g] * states[pairs[2 * p + 1] * N + g];
      rho[g] = s;
    } else {
      for (int p = 0; p < np; p++)
        rho[(long)p * N + g] = states[pairs[2 * p] * N + g] * states[pairs[2 * p + 1] * N + g];
    }
  }
}

int pair_density_dev(const double *states, long N, const long *pairs_d, const double *w_d, int np,
                     double *rho, hipStream_t st) {
  hipLaunchKernelGGL(k_pair_density, dim3(2048), dim3(256), 0, st, states, N, pairs_d, w_d, np, rho);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

template <int PB>
static int hf_pass(const GridDev &gd, double xmin, double ymin, double zmin, double dx, double dy,
                   double dz, double dv, int natom, const double *xyz, const int *type,
                   const double *scale, const long *npot, long maxnpot, const double *vr,
                   const double *pot, long stride, const double *dr, int ntype, const double *rho,
                   int np, int p0, int nsplit, int interp, double *work, double *htab, double *dR,
                   double *dscale, hipStream_t st) {
  const size_t lds = sizeof(double) * PB * (size_t)maxnpot;
  static bool attr = false;
  if (!attr) {
    hipFuncSetAttribute((const void *)k_hf_atom<PB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        160 * 1024 - 4096);
    attr = true;
  }
  hipLaunchKernelGGL(k_hf_atom<PB>, dim3((unsigned)natom, (unsigned)nsplit), dim3(HF_THREADS), lds,
                     st, gd.nx, gd.ny, gd.nz, xmin, ymin, zmin, dx, dy, dz, xyz, type, scale, npot,
                     vr, pot, stride, dr, rho, gd.N, np, p0, ntype, nsplit, interp, work, htab);
  const long nthr = (long)natom * PB;
  hipLaunchKernelGGL(k_hf_reduce<PB>, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, work,
                     natom, nsplit, np, p0, dv, dR, dscale);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// Largest pair batch whose LDS histogram (PB x maxnpot doubles) fits 152 KB.
int hf_pairs_per_pass(long maxnpot) {
  const long cap = (160 * 1024 - 8192) / 8;
  for (int pb = 8; pb > 1; pb /= 2)
    if ((long)pb * maxnpot <= cap) return pb;
  return maxnpot <= cap ? 1 : 0;
}

int hf_adjoint_dev(const GridDev &gd, double xmin, double ymin, double zmin, double dx, double dy,
                   double dz, double dv, int natom, const double *xyz, const int *type,
                   const double *scale, const long *npot, long maxnpot, const double *vr,
                   const double *pot, long stride, const double *dr, int ntype, const double *rho,
                   int np, int nsplit, int interp, double *work, double *htab, double *dR,
                   double *dscale, hipStream_t st) {
  const int pb = hf_pairs_per_pass(maxnpot);
  if (pb == 0) return SEP_ERR_ARG;
  for (int p0 = 0; p0 < np; p0 += pb) {
    int rc;
    switch (pb) {
      case 8:
        rc = hf_pass<8>(gd, xmin, ymin, zmin, dx, dy, dz, dv, natom, xyz, type, scale, npot,
                        maxnpot, vr, pot, stride, dr, ntype, rho, np, p0, nsplit, interp, work, htab,
                        dR, dscale, st);
        break;
      case 4:
        rc = hf_pass<4>(gd, xmin, ymin, zmin, dx, dy, dz, dv, natom, xyz, type, scale, npot,
                        maxnpot, vr, pot, stride, dr, ntype, rho, np, p0, nsplit, interp, work, htab,
                        dR, dscale, st);
        break;
      case 2:
        rc = hf_pass<2>(gd, xmin, ymin, zmin, dx, dy, dz, dv, natom, xyz, type, scale, npot,
                        maxnpot, vr, pot, stride, dr, ntype, rho, np, p0, nsplit, interp, work, htab,
                        dR, dscale, st);
        break;
      default:
        rc = hf_pass<1>(gd, xmin, ymin, zmin, dx, dy, dz, dv, natom, xyz, type, scale, npot,
                        maxnpot, vr, pot, stride, dr, ntype, rho, np, p0, nsplit, interp, work, htab,
                        dR, dscale, st);
    }
    if (rc != SEP_OK) return rc;
  }
  return SEP_OK;
}

}  // namespace sep
