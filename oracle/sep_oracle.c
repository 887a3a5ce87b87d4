/*
 * sep_oracle.c -- CPU restatement of the reference `filter` program.
 *
 * TEST INFRASTRUCTURE ONLY: see sep_oracle.h.  This file is the checker the
 * GPU product is compared against and the timed CPU baseline; the product
 * never links it.
 *
 * Parity status of each function (how it is pinned, see tests/):
 *   or_fft1d/or_fft3d     numpy.fft on random data; plane-wave closed forms
 *   or_ksqr               closed form (k_j = j*2pi/L mirrored, clamp)
 *   or_build_potential    numpy brute-force sum on small grids; reference
 *                         In13_P16 run.dat Vmin/Vmax
 *   or_hamiltonian        numpy restatement; plane-wave eigenvalues
 *   or_ran_nrc            numpy int64 restatement (bitwise); with or_init_psi
 *                         and the filter, the reference's In13_P16
 *                         eval-filt-*.dat at the recovered seed (1.9e-9 Ha)
 *   or_energy_range       constant-V closed form; In13_P16 run.dat dE, dt
 *   or_leja_points        brute-force greedy on tiny nc; point-set properties
 *   or_newton_coeffs      mpmath divided differences; interpolation identity
 *   or_filter_state       scalar Newton polynomial on plane waves
 *   or_ortho              numpy SVD (span, singular values, cutoff)
 *   or_hmat/or_sigma      scipy eigh; sigma = 0 for exact eigenvectors
 *   or_run                In13_P16 reference eval.dat row by row at the
 *                         recovered seed (sigma < 1e-2: 5.6e-9 Ha)
 */
#define _GNU_SOURCE
#include "sep_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define OR_PI 3.14159265358979323846
#define OR_TWOPI 6.28318530717958647692

static char g_err[512];
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char *or_last_error(void) { return g_err; }

typedef struct {
  double re, im;
} zc;

/* ========================================================================= */
/* FFT: mixed-radix Stockham autosort (own code; replaces FFTW3).            */
/* Pass with stride Ns, radix r: for j < n/r, t = j % Ns:                    */
/*   v_k = in[j + k n/r] * w^{k t}, w = exp(sign 2 pi i/(Ns r));             */
/*   out[(j/Ns) Ns r + t + k Ns] = DFT_r(v)_k                                */
/* ========================================================================= */
typedef struct {
  long n;
  int nf;
  int fac[64];
  zc *tw;        /* concatenated per-pass twiddles, Ns*r entries per pass */
  long twoff[64];
  zc *root[64];  /* r-th roots for the generic butterfly (per pass) */
} of_plan;

static int of_plan_init(of_plan *p, long n) {
  memset(p, 0, sizeof *p);
  p->n = n;
  long m = n;
  while (m > 1) {
    int r;
    if (m % 4 == 0)
      r = 4;
    else if (m % 2 == 0)
      r = 2;
    else if (m % 3 == 0)
      r = 3;
    else if (m % 5 == 0)
      r = 5;
    else {
      r = 7;
      while (m % r) r += 2;
    }
    if (p->nf >= 64) return -1;
    p->fac[p->nf++] = r;
    m /= r;
  }
  long tot = 0, Ns = 1;
  for (int f = 0; f < p->nf; f++) {
    p->twoff[f] = tot;
    tot += Ns * p->fac[f];
    Ns *= p->fac[f];
  }
  p->tw = (zc *)malloc(sizeof(zc) * (tot > 0 ? tot : 1));
  Ns = 1;
  for (int f = 0; f < p->nf; f++) {
    int r = p->fac[f];
    for (long t = 0; t < Ns; t++)
      for (int k = 0; k < r; k++) {
        double a = -OR_TWOPI * (double)(k * t) / (double)(Ns * r);
        p->tw[p->twoff[f] + t * r + k].re = cos(a);
        p->tw[p->twoff[f] + t * r + k].im = sin(a);
      }
    if (r >= 5) {
      p->root[f] = (zc *)malloc(sizeof(zc) * r);
      for (int k = 0; k < r; k++) {
        p->root[f][k].re = cos(-OR_TWOPI * k / r);
        p->root[f][k].im = sin(-OR_TWOPI * k / r);
      }
    }
    Ns *= r;
  }
  return 0;
}

static void of_plan_free(of_plan *p) {
  free(p->tw);
  for (int f = 0; f < p->nf; f++) free(p->root[f]);
}

/* one line, in -> result in `a` (ping-pong with `b`) */
static void of_exec(const of_plan *p, zc *a, zc *b, int sign) {
  long n = p->n, Ns = 1;
  zc *in = a, *out = b;
  for (int f = 0; f < p->nf; f++) {
    int r = p->fac[f];
    long nr = n / r;
    const zc *tw = p->tw + p->twoff[f];
    for (long j = 0; j < nr; j++) {
      long t = j % Ns;
      zc v[64];
      zc *vv = v;
      zc *heap = NULL;
      if (r > 64) vv = heap = (zc *)malloc(sizeof(zc) * r);
      for (int k = 0; k < r; k++) {
        zc x = in[j + k * nr];
        zc w = tw[t * r + k];
        if (sign > 0) w.im = -w.im;
        vv[k].re = x.re * w.re - x.im * w.im;
        vv[k].im = x.re * w.im + x.im * w.re;
      }
      double s = (double)sign;
      if (r == 2) {
        zc a0 = vv[0], a1 = vv[1];
        vv[0].re = a0.re + a1.re; vv[0].im = a0.im + a1.im;
        vv[1].re = a0.re - a1.re; vv[1].im = a0.im - a1.im;
      } else if (r == 4) {
        zc a0 = vv[0], a1 = vv[1], a2 = vv[2], a3 = vv[3];
        zc s02 = {a0.re + a2.re, a0.im + a2.im}, d02 = {a0.re - a2.re, a0.im - a2.im};
        zc s13 = {a1.re + a3.re, a1.im + a3.im}, d13 = {a1.re - a3.re, a1.im - a3.im};
        /* w = exp(sign*2pi i/4) = sign*i; y1 = d02 + w d13, y3 = d02 - w d13 */
        zc wd = {-s * d13.im, s * d13.re};
        vv[0].re = s02.re + s13.re; vv[0].im = s02.im + s13.im;
        vv[2].re = s02.re - s13.re; vv[2].im = s02.im - s13.im;
        vv[1].re = d02.re + wd.re;  vv[1].im = d02.im + wd.im;
        vv[3].re = d02.re - wd.re;  vv[3].im = d02.im - wd.im;
      } else if (r == 3) {
        const double c3 = 0.86602540378443864676; /* sin(2pi/3) */
        zc a0 = vv[0], a1 = vv[1], a2 = vv[2];
        zc t1 = {a1.re + a2.re, a1.im + a2.im};
        zc t2 = {a0.re - 0.5 * t1.re, a0.im - 0.5 * t1.im};
        zc d = {a1.re - a2.re, a1.im - a2.im};
        /* i * sign * c3 * d */
        zc t3 = {-s * c3 * d.im, s * c3 * d.re};
        vv[0].re = a0.re + t1.re; vv[0].im = a0.im + t1.im;
        vv[1].re = t2.re + t3.re; vv[1].im = t2.im + t3.im;
        vv[2].re = t2.re - t3.re; vv[2].im = t2.im - t3.im;
      } else {
        /* generic odd radix: direct DFT with the r-th roots */
        zc y[64];
        zc *yy = y, *heap2 = NULL;
        if (r > 64) yy = heap2 = (zc *)malloc(sizeof(zc) * r);
        const zc *rt = p->root[f];
        for (int q = 0; q < r; q++) {
          double sr = 0.0, si = 0.0;
          for (int k = 0; k < r; k++) {
            zc w = rt[(long)q * k % r];
            if (sign > 0) w.im = -w.im;
            sr += vv[k].re * w.re - vv[k].im * w.im;
            si += vv[k].re * w.im + vv[k].im * w.re;
          }
          yy[q].re = sr;
          yy[q].im = si;
        }
        for (int q = 0; q < r; q++) vv[q] = yy[q];
        free(heap2);
      }
      long base = (j / Ns) * Ns * r + t;
      for (int k = 0; k < r; k++) out[base + k * Ns] = vv[k];
      free(heap);
    }
    zc *tmp = in;
    in = out;
    out = tmp;
    Ns *= r;
  }
  if (in != a) memcpy(a, in, sizeof(zc) * n);
}

int or_fft1d(long n, double *data, int sign) {
  of_plan p;
  if (of_plan_init(&p, n)) return -1;
  zc *b = (zc *)malloc(sizeof(zc) * n);
  of_exec(&p, (zc *)data, b, sign);
  free(b);
  of_plan_free(&p);
  return 0;
}

int or_fft3d(long nx, long ny, long nz, double *data, int sign) {
  zc *d = (zc *)data;
  of_plan px, py, pz;
  if (of_plan_init(&px, nx) || of_plan_init(&py, ny) || of_plan_init(&pz, nz)) {
    set_err("fft plan failed");
    return -1;
  }
  long nmax = nx > ny ? nx : ny;
  if (nz > nmax) nmax = nz;
  /* x lines (contiguous) */
#pragma omp parallel
  {
    zc *a = (zc *)malloc(sizeof(zc) * nmax), *b = (zc *)malloc(sizeof(zc) * nmax);
#pragma omp for
    for (long l = 0; l < ny * nz; l++) {
      memcpy(a, d + l * nx, sizeof(zc) * nx);
      of_exec(&px, a, b, sign);
      memcpy(d + l * nx, a, sizeof(zc) * nx);
    }
    /* y lines */
#pragma omp for
    for (long l = 0; l < nx * nz; l++) {
      long jx = l % nx, jz = l / nx;
      zc *base = d + jz * nx * ny + jx;
      for (long jy = 0; jy < ny; jy++) a[jy] = base[jy * nx];
      of_exec(&py, a, b, sign);
      for (long jy = 0; jy < ny; jy++) base[jy * nx] = a[jy];
    }
    /* z lines */
#pragma omp for
    for (long l = 0; l < nx * ny; l++) {
      zc *base = d + l;
      for (long jz = 0; jz < nz; jz++) a[jz] = base[jz * nx * ny];
      of_exec(&pz, a, b, sign);
      for (long jz = 0; jz < nz; jz++) base[jz * nx * ny] = a[jz];
    }
    free(a);
    free(b);
  }
  of_plan_free(&px);
  of_plan_free(&py);
  of_plan_free(&pz);
  return 0;
}

/* ========================================================================= */
/* Grid setup                                                                */
/* ========================================================================= */

/* filter/init.c:124-144: per-axis tables with 1/N folded in (in the
 * reference's multiplication order), sum, clamp at Ekinmax/N. */
int or_ksqr(long nx, long ny, long nz, double dx, double dy, double dz,
            double ekinmax, double *ksqr) {
  double nx_1 = 1.0 / (double)nx, ny_1 = 1.0 / (double)ny, nz_1 = 1.0 / (double)nz;
  double dkx = OR_TWOPI / ((double)nx * dx);
  double dky = OR_TWOPI / ((double)ny * dy);
  double dkz = OR_TWOPI / ((double)nz * dz);
  double *kx = (double *)calloc(nx, sizeof(double));
  double *ky = (double *)calloc(ny, sizeof(double));
  double *kz = (double *)calloc(nz, sizeof(double));
  double mx = 1.0, my = 1.0, mz = 1.0;
  for (long j = 1; j <= nx / 2; j++) {
    double q = (double)j * dkx;
    kx[j] = kx[nx - j] = 0.5 * (q * q) * nx_1 * ny_1 * nz_1 / mx;
  }
  for (long j = 1; j <= ny / 2; j++) {
    double q = (double)j * dky;
    ky[j] = ky[ny - j] = 0.5 * (q * q) * ny_1 * nx_1 * nz_1 / my;
  }
  for (long j = 1; j <= nz / 2; j++) {
    double q = (double)j * dkz;
    kz[j] = kz[nz - j] = 0.5 * (q * q) * nz_1 * nx_1 * ny_1 / mz;
  }
  double emax = ekinmax * (ny_1 * nx_1 * nz_1);
  for (long jz = 0; jz < nz; jz++)
    for (long jy = 0; jy < ny; jy++)
      for (long jx = 0; jx < nx; jx++) {
        long g = nx * (ny * jz + jy) + jx;
        double v = kx[jx] + ky[jy] + kz[jz];
        if (v > emax) v = emax;
        ksqr[g] = v;
      }
  free(kx);
  free(ky);
  free(kz);
  return 0;
}

/* filter/interpolate.c:11-23; the scale factor is
 * filterStrainCube/interpolate.c:20-21 (scale 1 reproduces the plain form). */
/* Table interpolation of the V(r) builder and its adjoint: 0 = the
 * reference's linear interpolate() (filter/interpolate.c:11-23), 1 = the
 * cubic Hermite extension (libsep's SEP_INTERP_CUBIC: Catmull-Rom tangents,
 * one-sided at both table ends, same segment and cutoff).  Written here from
 * the Hermite basis and the tangents directly. */
static int g_interp = 0;
void or_set_interp(int interp) { g_interp = interp; }

/* tangent (per row step) of an n-row table at row k */
static double herm_tangent(const double *p, long n, long k) {
  if (k == 0) return p[1] - p[0];
  if (k == n - 1) return p[n - 1] - p[n - 2];
  return 0.5 * (p[k + 1] - p[k - 1]);
}
/* value and d/dt on segment i at t in [0, 1) */
static void herm_eval(const double *p, long n, long i, double t, double *val, double *dval) {
  double m0 = herm_tangent(p, n, i), m1 = herm_tangent(p, n, i + 1);
  double a = p[i], b = p[i + 1];
  /* Horner form of h00 a + h10 m0 + h01 b + h11 m1 */
  double c3 = 2.0 * a + m0 - 2.0 * b + m1;
  double c2 = -3.0 * a - 2.0 * m0 + 3.0 * b - m1;
  *val = ((c3 * t + c2) * t + m0) * t + a;
  *dval = (3.0 * c3 * t + 2.0 * c2) * t + m0;
}
/* d(value)/d(p_(i-1+q)), q = 0..3, on segment i at t: the Hermite basis
 * times the derivative of each tangent with respect to each row (rows
 * outside the table get weight 0) */
static void herm_row_weights(long n, long i, double t, double *wq) {
  double t2 = t * t, t3 = t2 * t;
  double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t;
  double h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
  for (int q = 0; q < 4; q++) wq[q] = 0.0;
  wq[1] += h00; /* p_i */
  wq[2] += h01; /* p_(i+1) */
  if (i == 0) { /* m0 = p1 - p0 */
    wq[2] += h10;
    wq[1] -= h10;
  } else { /* m0 = (p_(i+1) - p_(i-1)) / 2 */
    wq[2] += 0.5 * h10;
    wq[0] -= 0.5 * h10;
  }
  if (i + 1 == n - 1) { /* m1 = p_(n-1) - p_(n-2) */
    wq[2] += h11;
    wq[1] -= h11;
  } else { /* m1 = (p_(i+2) - p_i) / 2 */
    wq[3] += 0.5 * h11;
    wq[1] -= 0.5 * h11;
  }
}

double or_interpolate(double r, double dr, const double *vr, const double *pot,
                      long stride, long n, long j, double scale) {
  long i = (long)(r / dr);
  if (i > (n - 2)) return 0.0;
  if (g_interp == 1) {
    const double *p = pot + j * stride, *v = vr + j * stride;
    double val, dval;
    herm_eval(p, n, i, (r - v[i]) / (v[i + 1] - v[i]), &val, &dval);
    return scale * val;
  }
  const double *p = pot + j * stride, *v = vr + j * stride;
  double a = scale * (p[i + 1] - p[i]) / (v[i + 1] - v[i]);
  double b = scale * p[i] - v[i] * a;
  return a * r + b;
}

/* filter/init.c:157-182 (grid coordinates by repeated addition, :157-159),
 * expot filter/init.c:255-257. */
int or_build_potential(long nx, long ny, long nz, double xmin, double ymin,
                       double zmin, double dx, double dy, double dz,
                       long natom, const double *rx, const double *ry,
                       const double *rz, const long *type, const double *scale,
                       const long *npot, const double *vr, const double *pot,
                       long stride, const double *dr, double ex_e,
                       double ex_x0, double ex_t, double *potl) {
  double *vx = (double *)malloc(sizeof(double) * nx);
  double *vy = (double *)malloc(sizeof(double) * ny);
  double *vz = (double *)malloc(sizeof(double) * nz);
  double c;
  long j;
  for (j = 0, c = xmin; j < nx; j++, c += dx) vx[j] = c;
  for (j = 0, c = ymin; j < ny; j++, c += dy) vy[j] = c;
  for (j = 0, c = zmin; j < nz; j++, c += dz) vz[j] = c;
#pragma omp parallel for schedule(static)
  for (long jz = 0; jz < nz; jz++)
    for (long jy = 0; jy < ny; jy++)
      for (long jx = 0; jx < nx; jx++) {
        double rex = sqrt(vx[jx] * vx[jx] + vy[jy] * vy[jy] + vz[jz] * vz[jz]);
        double potEx = ex_e * (1.0 / (exp((rex - ex_x0) / ex_t) + 1.0) - 1.0);
        double sum = 0.0;
        for (long ie = 0; ie < natom; ie++) {
          double ddx = vx[jx] - rx[ie], ddy = vy[jy] - ry[ie], ddz = vz[jz] - rz[ie];
          double del = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
          long t = type[ie];
          sum += or_interpolate(del, dr[t], vr, pot, stride, npot[t], t,
                                scale ? scale[ie] : 1.0);
        }
        potl[nx * (ny * jz + jy) + jx] = sum + potEx;
      }
  free(vx);
  free(vy);
  free(vz);
  return 0;
}

/* ========================================================================= */
/* Hellmann-Feynman adjoint of V(r): derivatives of the forward model above  */
/* (filter/init.c:161-182 + interpolate.c:11-23 with the strain scale)        */
/* contracted with one density rho on the grid:                              */
/*   dR[a]     = dv sum_g rho dV/dR_a,  dV/dR_a = -s_a slope (r - R_a)/|r-R_a| */
/*   dscale[a] = dv sum_g rho interp(|r - R_a|)                              */
/*   dtab[t][k]= dv sum_g rho dV/dpot_t[k]                                   */
/* Brute force over every (grid point, atom); the product computes only the  */
/* atoms' bounding boxes on the GPU.                                         */
/* ========================================================================= */
int or_hf_adjoint(long nx, long ny, long nz, double xmin, double ymin, double zmin,
                  double dx, double dy, double dz, long natom, const double *rx,
                  const double *ry, const double *rz, const long *type,
                  const double *scale, const long *npot, const double *vr,
                  const double *pot, long stride, long ntype, const double *dr,
                  const double *rho, double dv, double *dR, double *dscale,
                  double *dtab) {
  double *vx = (double *)malloc(sizeof(double) * nx);
  double *vy = (double *)malloc(sizeof(double) * ny);
  double *vz = (double *)malloc(sizeof(double) * nz);
  double c;
  long j;
  for (j = 0, c = xmin; j < nx; j++, c += dx) vx[j] = c;
  for (j = 0, c = ymin; j < ny; j++, c += dy) vy[j] = c;
  for (j = 0, c = zmin; j < nz; j++, c += dz) vz[j] = c;
  memset(dtab, 0, sizeof(double) * ntype * stride);
#pragma omp parallel for schedule(dynamic)
  for (long a = 0; a < natom; a++) {
    const long t = type[a];
    const double s = scale ? scale[a] : 1.0;
    const double *v = vr + t * stride, *p = pot + t * stride;
    double gx = 0.0, gy = 0.0, gz = 0.0, gs = 0.0;
    double *h = (double *)calloc(stride, sizeof(double));
    for (long jz = 0; jz < nz; jz++)
      for (long jy = 0; jy < ny; jy++)
        for (long jx = 0; jx < nx; jx++) {
          double ddx = vx[jx] - rx[a], ddy = vy[jy] - ry[a], ddz = vz[jz] - rz[a];
          double del = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
          long i = (long)(del / dr[t]);
          if (i > npot[t] - 2) continue;
          double r = rho[nx * (ny * jz + jy) + jx];
          if (g_interp == 1) {
            double tt = (del - v[i]) / (v[i + 1] - v[i]), val, dval, wq[4];
            herm_eval(p, npot[t], i, tt, &val, &dval);
            dval /= (v[i + 1] - v[i]);
            if (del > 0.0) {
              double f = -r * s * dval / del;
              gx += f * ddx;
              gy += f * ddy;
              gz += f * ddz;
            }
            gs += r * val;
            herm_row_weights(npot[t], i, tt, wq);
            for (int q = 0; q < 4; q++)
              if (wq[q] != 0.0) h[i - 1 + q] += r * s * wq[q];
            continue;
          }
          double slope = (p[i + 1] - p[i]) / (v[i + 1] - v[i]);
          double w = (del - v[i]) / (v[i + 1] - v[i]);
          if (del > 0.0) {
            double f = -r * s * slope / del;
            gx += f * ddx;
            gy += f * ddy;
            gz += f * ddz;
          }
          gs += r * (p[i] + slope * (del - v[i]));
          h[i] += r * s * (1.0 - w);
          h[i + 1] += r * s * w;
        }
    dR[3 * a] = gx * dv;
    dR[3 * a + 1] = gy * dv;
    dR[3 * a + 2] = gz * dv;
    dscale[a] = gs * dv;
#pragma omp critical
    for (long k = 0; k < stride; k++) dtab[t * stride + k] += h[k] * dv;
    free(h);
  }
  free(vx);
  free(vy);
  free(vz);
  return 0;
}

/* ========================================================================= */
/* H application: filter/hamiltonian.c:11-44                                 */
/* ========================================================================= */
static void ham_one(long nx, long ny, long nz, const double *ksqr,
                    const double *potl, zc *phi, zc *tmp) {
  long N = nx * ny * nz;
  memcpy(tmp, phi, sizeof(zc) * N); /* fftwpsi <- phi (kinetic :31) */
  or_fft3d(nx, ny, nz, (double *)tmp, -1);
  for (long j = 0; j < N; j++) {
    tmp[j].re *= ksqr[j];
    tmp[j].im *= ksqr[j];
  }
  or_fft3d(nx, ny, nz, (double *)tmp, +1);
  /* phi = T psi + potl * psi, psi = the copy of phi taken at :16 */
  for (long i = 0; i < N; i++) {
    double pr = phi[i].re, pi = phi[i].im;
    phi[i].re = tmp[i].re + potl[i] * pr;
    phi[i].im = tmp[i].im + potl[i] * pi;
  }
}

int or_hamiltonian(long nx, long ny, long nz, const double *ksqr,
                   const double *potl, double *phi, long batch) {
  long N = nx * ny * nz;
  zc *tmp = (zc *)malloc(sizeof(zc) * N);
  for (long b = 0; b < batch; b++) ham_one(nx, ny, nz, ksqr, potl, (zc *)phi + b * N, tmp);
  free(tmp);
  return 0;
}

/* ========================================================================= */
/* Random numbers: filter/rand.c:11-66 (L'Ecuyer with Bays-Durham shuffle).  */
/* The reference keeps idum2/iy/iv in statics; here they live in or_rng.     */
/* ========================================================================= */
#define R_IM1 2147483563L
#define R_IM2 2147483399L
#define R_AM (1.0 / (double)R_IM1)
#define R_IMM1 (R_IM1 - 1)
#define R_IA1 40014L
#define R_IA2 40692L
#define R_IQ1 53668L
#define R_IQ2 52774L
#define R_IR1 12211L
#define R_IR2 3791L
#define R_NTAB 32
#define R_NDIV (1 + R_IMM1 / R_NTAB)
#define R_EPS 1.2e-7
#define R_RNMX (1.0 - R_EPS)

void or_rng_reset(or_rng *st) {
  st->idum2 = 123456789L;
  st->iy = 0;
  memset(st->iv, 0, sizeof st->iv);
}

double or_ran_nrc(or_rng *st, long *idum) {
  long j, k;
  double temp;
  if (*idum <= 0) {
    if (-(*idum) < 1)
      *idum = 1;
    else
      *idum = -(*idum);
    st->idum2 = (*idum);
    for (j = R_NTAB + 7; j >= 0; j--) {
      k = (*idum) / R_IQ1;
      *idum = R_IA1 * (*idum - k * R_IQ1) - k * R_IR1;
      if (*idum < 0) *idum += R_IM1;
      if (j < R_NTAB) st->iv[j] = *idum;
    }
    st->iy = st->iv[0];
  }
  k = (*idum) / R_IQ1;
  *idum = R_IA1 * (*idum - k * R_IQ1) - k * R_IR1;
  if (*idum < 0) *idum += R_IM1;
  k = st->idum2 / R_IQ2;
  st->idum2 = R_IA2 * (st->idum2 - k * R_IQ2) - k * R_IR2;
  if (st->idum2 < 0) st->idum2 += R_IM2;
  j = st->iy / R_NDIV;
  st->iy = st->iv[j] - st->idum2;
  st->iv[j] = *idum;
  if (st->iy < 1) st->iy += R_IMM1;
  if ((temp = R_AM * st->iy) > R_RNMX) return R_RNMX;
  return temp;
}

/* filter/init.c:236-251 with normalize() of filter/norm.c:23-36 (im = 0). */
void or_init_psi(or_rng *st, long *idum, long ngrid, double dv, double *psi_re) {
  long tidum = *idum;
  for (long g = 0; g < ngrid; g++) psi_re[g] = (-1.0 + 2.0 * or_ran_nrc(st, &tidum));
  double nrm = 0.0;
  for (long g = 0; g < ngrid; g++) nrm += psi_re[g] * psi_re[g];
  nrm = sqrt(nrm * dv);
  for (long g = 0; g < ngrid; g++) psi_re[g] /= nrm;
  *idum = tidum;
}

/* ========================================================================= */
/* Spectral range: filter/energy.c:11-24 (energy), :53-83 (power iteration)  */
/* ========================================================================= */
static double cnorm(const zc *p, long N, double dv) {
  double s = 0.0;
  for (long i = 0; i < N; i++) s += p[i].re * p[i].re + p[i].im * p[i].im;
  return sqrt(s * dv);
}

int or_energy_range(long nx, long ny, long nz, const double *ksqr,
                    const double *potl, double dv, double vmin, double *dE,
                    double *emax, long *iters) {
  long N = nx * ny * nz;
  zc *psi = (zc *)calloc(N, sizeof(zc)), *phi = (zc *)calloc(N, sizeof(zc));
  zc *tmp = (zc *)malloc(sizeof(zc) * N);
  double *re = (double *)malloc(sizeof(double) * N);
  or_rng st;
  or_rng_reset(&st);
  long idum = -874917403L;
  or_init_psi(&st, &idum, N, dv, re);
  for (long i = 0; i < N; i++) {
    psi[i].re = re[i];
    psi[i].im = 0.0;
  }
  double ene = 0.0 + 0.1, ene_old = 0.0;
  long i;
  for (i = 0; (fabs((ene - ene_old) / ene) > 1.0e-4) && (i < 500); i++) {
    ham_one(nx, ny, nz, ksqr, potl, psi, tmp); /* psi <- H psi */
    double nrm = cnorm(psi, N, dv);
    for (long g = 0; g < N; g++) {
      psi[g].re /= nrm;
      psi[g].im /= nrm;
    }
    ene_old = ene;
    memcpy(phi, psi, sizeof(zc) * N);
    ham_one(nx, ny, nz, ksqr, potl, phi, tmp);
    double e = 0.0;
    for (long g = 0; g < N; g++) e += psi[g].re * phi[g].re + psi[g].im * phi[g].im;
    ene = e * dv;
  }
  *dE = 1.1 * (ene - vmin);
  *emax = ene;
  *iters = i;
  free(psi);
  free(phi);
  free(tmp);
  free(re);
  return 0;
}

/* ========================================================================= */
/* Newton interpolation: filter/coeff.c                                      */
/* ========================================================================= */

/* samp_points_ashkenazy, filter/coeff.c:82-134: greedy maximal log-product
 * over nc3 = 32 nc candidates; the last nc3/8 candidates sit at 0 because
 * the imaginary range is 0 (coeff.c:85,91-98).  Strict '>' keeps the first
 * maximum.  Returns fk (coeff.c:295-297); zn gets the real parts. */
double or_leja_points(long nc, double *zn) {
  long nc3 = 32 * nc, imfrac = nc3 / 8;
  double min = -2.0, max = 2.0, range = 0.0;
  double minim = range * min, maxim = range * max;
  double dsre = (max - min) / (double)(nc3 - 1 - imfrac);
  double dsim = (maxim - minim) / (double)(imfrac - 1);
  zc *samp = (zc *)calloc(nc3, sizeof(zc));
  double *veca = (double *)calloc(nc3, sizeof(double));
  zc *point = (zc *)calloc(nc, sizeof(zc));
  for (long j = 0; j < nc3 - imfrac; j++) samp[j].re = min + (double)(j)*dsre;
  for (long j = nc3 - imfrac; j < nc3; j++) samp[j].im = minim + (double)(j - nc3 + imfrac) * dsim;
  point[0] = samp[0];
  for (long k = 0; k < nc3; k++) {
    double dre = samp[k].re - point[0].re, dim = samp[k].im - point[0].im;
    double del = dre * dre + dim * dim;
    veca[k] = (del < 1.0e-10) ? -1.0e30 : log(del);
  }
  for (long j = 1; j < nc; j++) {
    long kmax = 0;
    double fkmax = -2.0e30;
    for (long k = 0; k < nc3; k++)
      if (veca[k] > fkmax) {
        fkmax = veca[k];
        kmax = k;
      }
    point[j] = samp[kmax];
#pragma omp parallel for schedule(static)
    for (long k = 0; k < nc3; k++) {
      double dre = samp[k].re - point[j].re, dim = samp[k].im - point[j].im;
      double del = dre * dre + dim * dim;
      if (del < 1.0e-10)
        veca[k] = -1.0e30;
      else
        veca[k] += log(del);
    }
  }
  double fk = 1.0;
  for (long j = 0; j < nc; j++) fk *= (point[j].re * point[j].re + point[j].im * point[j].im);
  fk = sqrt(fk);
  fk = pow(fk, 1.0 / (double)nc);
  for (long j = 0; j < nc; j++) zn[j] = point[j].re;
  free(samp);
  free(veca);
  free(point);
  return fk;
}

/* coefficient, filter/coeff.c:16-63: divided differences of
 * f(x) = sqrt(dt/pi) exp(-dt (x - El)^2) at x = (z+2)/scale + Vmin,
 * scale = 4/dE.  Only the real parts are non-zero (the imaginary
 * recursion starts at 0 and stays 0).  an: ms x nc row-major. */
int or_newton_coeffs(long nc, long ms, const double *zn, const double *el,
                     double dE, double vmin, double dt, double *an) {
  double scale = (2.0 - (-2.0)) / dE;
#define FILT(x) (sqrt(dt / OR_PI) * exp(-((x) * (x)) * dt))
#pragma omp parallel for schedule(static)
  for (long ie = 0; ie < ms; ie++) {
    double *a = an + ie * nc;
    double x = (zn[0] + 2.0) / scale + vmin;
    a[0] = FILT(x - el[ie]);
    if (nc > 1) {
      x = (zn[1] + 2.0) / scale + vmin;
      a[1] = (FILT(x - el[ie]) - a[0]) / (zn[1] - zn[0]);
    }
    for (long j = 2; j < nc; j++) {
      double res = 1.0, sumre = 0.0;
      for (long i = 1; i < j; i++) {
        res *= (zn[j] - zn[i - 1]);
        sumre += a[i] * res;
      }
      res *= (zn[j] - zn[j - 1]);
      x = (zn[j] + 2.0) / scale + vmin;
      a[j] = (FILT(x - el[ie]) - a[0] - sumre) / res;
    }
  }
#undef FILT
  return 0;
}

/* ========================================================================= */
/* Filter: filter/filter.c:59-83 with hnorm/hamiltonian_norm, hnorm.c:11-39   */
/* ========================================================================= */
int or_filter_state(long nx, long ny, long nz, const double *ksqr,
                    const double *potl, const double *an, const double *zn,
                    long nc, long ms, double dE_1, double vmin,
                    const double *psi0, double *out) {
  long N = nx * ny * nz;
  zc *psin = (zc *)malloc(sizeof(zc) * N), *psim1 = (zc *)malloc(sizeof(zc) * N);
  zc *tmp = (zc *)malloc(sizeof(zc) * N);
  for (long i = 0; i < N; i++) {
    psin[i].re = psi0[i];
    psin[i].im = 0.0;
  }
  for (long ie = 0; ie < ms; ie++)
    for (long i = 0; i < N; i++) out[ie * N + i] = an[ie * nc + 0] * psin[i].re;
  for (long j = 1; j < nc; j++) {
    memcpy(psim1, psin, sizeof(zc) * N);
    /* hamiltonian_norm: psin <- T psin; psin += potl * psim1 */
    memcpy(tmp, psin, sizeof(zc) * N);
    or_fft3d(nx, ny, nz, (double *)tmp, -1);
    for (long g = 0; g < N; g++) {
      tmp[g].re *= ksqr[g];
      tmp[g].im *= ksqr[g];
    }
    or_fft3d(nx, ny, nz, (double *)tmp, +1);
    double zm1 = zn[j - 1];
    double c = (2.0 + zm1 + vmin * dE_1);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
      double hr = tmp[i].re + potl[i] * psim1[i].re;
      double hi = tmp[i].im + potl[i] * psim1[i].im;
      psin[i].re = dE_1 * hr - c * psim1[i].re;
      psin[i].im = dE_1 * hi - c * psim1[i].im;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++)
      for (long ie = 0; ie < ms; ie++) out[ie * N + i] += (an[ie * nc + j] * psin[i].re);
  }
  free(psin);
  free(psim1);
  free(tmp);
  return 0;
}

/* or_filter_state for ns random states at once, one state per thread
 * (filter/main.c:92-98 runs the states in parallel the same way); out is
 * [ns][ms][N]. */
int or_filter_states(long nx, long ny, long nz, const double *ksqr, const double *potl,
                     const double *an, const double *zn, long nc, long ms, double dE_1,
                     double vmin, long ns, const double *psi0, double *out) {
  const long N = nx * ny * nz;
  int rc = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (long s = 0; s < ns; s++) {
    int r = or_filter_state(nx, ny, nz, ksqr, potl, an, zn, nc, ms, dE_1, vmin, psi0 + s * N,
                            out + s * ms * N);
    if (r) {
#pragma omp critical
      rc = r;
    }
  }
  return rc;
}

/* filter/norm.c:40-54 */
void or_normalize_all(double *psi, double dv, long m, long ngrid) {
#pragma omp parallel for schedule(static)
  for (long ie = 0; ie < m; ie++) {
    double *p = psi + ie * ngrid, nrm = 0.0;
    for (long i = 0; i < ngrid; i++) nrm += p[i] * p[i];
    if (nrm != 0.0) nrm = 1.0 / sqrt(nrm * dv);
    for (long i = 0; i < ngrid; i++) p[i] *= nrm;
  }
}

/* filter/energy.c:28-49 */
int or_energy_all(long nx, long ny, long nz, const double *ksqr,
                  const double *potl, double dv, const double *psi, long m,
                  double *ene) {
  long N = nx * ny * nz;
  zc *phi = (zc *)malloc(sizeof(zc) * N), *tmp = (zc *)malloc(sizeof(zc) * N);
  for (long ie = 0; ie < m; ie++) {
    for (long i = 0; i < N; i++) {
      phi[i].re = psi[ie * N + i];
      phi[i].im = 0.0;
    }
    ham_one(nx, ny, nz, ksqr, potl, phi, tmp);
    double e = 0.0;
    for (long i = 0; i < N; i++) e += psi[ie * N + i] * phi[i].re;
    ene[ie] = e * dv;
  }
  free(phi);
  free(tmp);
  return 0;
}

/* ========================================================================= */
/* Dense FP64 helpers                                                        */
/* ========================================================================= */

/* Cyclic Jacobi eigensolver (independent of the product's tridiagonal QL).
 * A row-major symmetric n x n; w ascending; V[:, k] eigenvector k. */
int or_jacobi_eig(long n, const double *A0, double *w, double *V) {
  double *A = (double *)malloc(sizeof(double) * n * n);
  memcpy(A, A0, sizeof(double) * n * n);
  for (long i = 0; i < n; i++)
    for (long j = 0; j < n; j++) V[i * n + j] = (i == j) ? 1.0 : 0.0;
  double fro = 0.0;
  for (long i = 0; i < n * n; i++) fro += A[i] * A[i];
  fro = sqrt(fro);
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (long p = 0; p < n; p++)
      for (long q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
    if (sqrt(off) <= 1e-17 * fro || off == 0.0) break;
    for (long p = 0; p < n - 1; p++)
      for (long q = p + 1; q < n; q++) {
        double apq = A[p * n + q];
        if (apq == 0.0) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (long k = 0; k < n; k++) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (long k = 0; k < n; k++) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (long k = 0; k < n; k++) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  /* sort ascending */
  long *idx = (long *)malloc(sizeof(long) * n);
  for (long i = 0; i < n; i++) idx[i] = i;
  for (long i = 1; i < n; i++) {
    long x = idx[i], j = i;
    while (j > 0 && A[idx[j - 1] * n + idx[j - 1]] > A[x * n + x]) {
      idx[j] = idx[j - 1];
      j--;
    }
    idx[j] = x;
  }
  double *Vs = (double *)malloc(sizeof(double) * n * n);
  for (long k = 0; k < n; k++) {
    w[k] = A[idx[k] * n + idx[k]];
    for (long i = 0; i < n; i++) Vs[i * n + k] = V[i * n + idx[k]];
  }
  memcpy(V, Vs, sizeof(double) * n * n);
  free(Vs);
  free(idx);
  free(A);
  return 0;
}

/* ========================================================================= */
/* ortho: filter/ortho.c:11-52 (dgesvd 'O','N' + cutoff S_i/S_0 < 1e-10).    */
/* Householder QR of the N x m column-major matrix, one-sided Jacobi SVD of  */
/* R, left vectors U = Q U_R written back in place.                          */
/* ========================================================================= */
long or_ortho(double *psi, long ngrid, long m, double *S) {
  long N = ngrid;
  if (m <= 0) return 0;
  double *vs = (double *)calloc((size_t)N * m, sizeof(double)); /* reflectors */
  double *R = (double *)calloc(m * m, sizeof(double));
  double *beta = (double *)calloc(m, sizeof(double));
  for (long k = 0; k < m; k++) {
    double *x = psi + k * N;
    double nrm = 0.0;
    for (long i = k; i < N; i++) nrm += x[i] * x[i];
    nrm = sqrt(nrm);
    double alpha = (x[k] >= 0.0) ? -nrm : nrm;
    double *v = vs + k * N;
    for (long i = k; i < N; i++) v[i] = x[i];
    v[k] -= alpha;
    double vn = 0.0;
    for (long i = k; i < N; i++) vn += v[i] * v[i];
    beta[k] = (vn > 0.0) ? 2.0 / vn : 0.0;
    /* apply H = I - beta v v^T to columns k..m-1 */
#pragma omp parallel for schedule(dynamic, 1)
    for (long c = k; c < m; c++) {
      double *y = psi + c * N, d = 0.0;
      for (long i = k; i < N; i++) d += v[i] * y[i];
      d *= beta[k];
      for (long i = k; i < N; i++) y[i] -= d * v[i];
    }
    for (long r = 0; r <= k; r++) R[r * m + k] = psi[k * N + r];
  }
  /* one-sided Jacobi on B = R (columns): B V = U_R Sigma */
  double *B = (double *)malloc(sizeof(double) * m * m); /* column-major: B[c*m + r] */
  for (long c = 0; c < m; c++)
    for (long r = 0; r < m; r++) B[c * m + r] = R[r * m + c];
  for (int sweep = 0; sweep < 60; sweep++) {
    double maxc = 0.0;
    for (long p = 0; p < m - 1; p++)
      for (long q = p + 1; q < m; q++) {
        double *bp = B + p * m, *bq = B + q * m;
        double a = 0.0, b = 0.0, g = 0.0;
        for (long r = 0; r < m; r++) {
          a += bp[r] * bp[r];
          b += bq[r] * bq[r];
          g += bp[r] * bq[r];
        }
        if (g == 0.0 || a == 0.0 || b == 0.0) continue;
        double cth = fabs(g) / sqrt(a * b);
        if (cth > maxc) maxc = cth;
        if (cth < 1e-15) continue;
        double zeta = (b - a) / (2.0 * g);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (long r = 0; r < m; r++) {
          double xp = bp[r], xq = bq[r];
          bp[r] = c * xp - s * xq;
          bq[r] = s * xp + c * xq;
        }
      }
    if (maxc < 1e-15) break;
  }
  double *sv = (double *)malloc(sizeof(double) * m);
  long *idx = (long *)malloc(sizeof(long) * m);
  for (long c = 0; c < m; c++) {
    double s2 = 0.0;
    for (long r = 0; r < m; r++) s2 += B[c * m + r] * B[c * m + r];
    sv[c] = sqrt(s2);
    idx[c] = c;
  }
  for (long i = 1; i < m; i++) { /* descending */
    long x = idx[i], j = i;
    while (j > 0 && sv[idx[j - 1]] < sv[x]) {
      idx[j] = idx[j - 1];
      j--;
    }
    idx[j] = x;
  }
  /* U = Q [U_R; 0], Q = H_0 H_1 ... H_{m-1} */
  for (long k = 0; k < m; k++) {
    long c = idx[k];
    S[k] = sv[c];
    double *u = psi + k * N;
    memset(u, 0, sizeof(double) * N);
    if (sv[c] > 0.0)
      for (long r = 0; r < m; r++) u[r] = B[c * m + r] / sv[c];
    else
      u[k] = 1.0; /* any unit vector; never kept */
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (long k = 0; k < m; k++) {
    double *u = psi + k * N;
    for (long h = m - 1; h >= 0; h--) {
      double *v = vs + h * N, d = 0.0;
      for (long i = h; i < N; i++) d += v[i] * u[i];
      d *= beta[h];
      for (long i = h; i < N; i++) u[i] -= d * v[i];
    }
  }
  long cutoff = m;
  for (long i = 0; i < m; i++)
    if ((S[i] / S[0]) < 1.0e-10) {
      cutoff = i;
      break;
    }
  free(vs);
  free(R);
  free(beta);
  free(B);
  free(sv);
  free(idx);
  return cutoff;
}

/* Hmatreal, filter/Hmat.c:11-86: H_ij = dv <Re H psi_i | psi_j>; dsyev('V',
 * 'U') on the row-major array read column-major uses the entries with
 * i >= j (Hmat.c:55); psi_new_j = sum_i V[i][j] psi_i (Hmat.c:61-78). */
int or_hmat(long nx, long ny, long nz, const double *ksqr, const double *potl,
            double dv, double *psi, long m, double *eval) {
  long N = nx * ny * nz;
  double *H = (double *)calloc(m * m, sizeof(double));
  zc *phi = (zc *)malloc(sizeof(zc) * N), *tmp = (zc *)malloc(sizeof(zc) * N);
  for (long i = 0; i < m; i++) {
    for (long g = 0; g < N; g++) {
      phi[g].re = psi[i * N + g];
      phi[g].im = 0.0;
    }
    ham_one(nx, ny, nz, ksqr, potl, phi, tmp);
#pragma omp parallel for schedule(static)
    for (long j = 0; j < m; j++) {
      double s = 0.0;
      for (long g = 0; g < N; g++) s += phi[g].re * psi[j * N + g];
      H[i * m + j] = s * dv;
    }
  }
  double *S = (double *)malloc(sizeof(double) * m * m);
  for (long a = 0; a < m; a++)
    for (long b = 0; b < m; b++) {
      long hi = a > b ? a : b, lo = a > b ? b : a;
      S[a * m + b] = H[hi * m + lo];
    }
  double *V = (double *)malloc(sizeof(double) * m * m);
  or_jacobi_eig(m, S, eval, V);
#pragma omp parallel
  {
    double *t = (double *)malloc(sizeof(double) * m);
#pragma omp for schedule(static)
    for (long g = 0; g < N; g++) {
      for (long j = 0; j < m; j++) t[j] = psi[j * N + g];
      for (long j = 0; j < m; j++) {
        double s = 0.0;
        for (long i = 0; i < m; i++) s += V[i * m + j] * t[i];
        psi[j * N + g] = s;
      }
    }
    free(t);
  }
  free(H);
  free(S);
  free(V);
  free(phi);
  free(tmp);
  return 0;
}

/* calc_sigma_E, filter/energy.c:88-119 (two H applications). */
int or_sigma(long nx, long ny, long nz, const double *ksqr, const double *potl,
             double dv, const double *psi, long m, double *sige) {
  long N = nx * ny * nz;
  zc *phi = (zc *)malloc(sizeof(zc) * N), *tmp = (zc *)malloc(sizeof(zc) * N);
  for (long s = 0; s < m; s++) {
    const double *p = psi + s * N;
    for (long g = 0; g < N; g++) {
      phi[g].re = p[g];
      phi[g].im = 0.0;
    }
    ham_one(nx, ny, nz, ksqr, potl, phi, tmp);
    double e = 0.0;
    for (long g = 0; g < N; g++) e += p[g] * phi[g].re;
    e *= dv;
    ham_one(nx, ny, nz, ksqr, potl, phi, tmp);
    double e2 = 0.0;
    for (long g = 0; g < N; g++) e2 += p[g] * phi[g].re;
    e2 *= dv;
    e2 -= e * e;
    sige[s] = sqrt(fabs(e2));
  }
  free(phi);
  free(tmp);
  return 0;
}

/* ========================================================================= */
/* Whole program: input parsing (filter/init.c:12-50, read.c:11-176),        */
/* setup (init.c:54-233), pipeline (main.c:11-162)                           */
/* ========================================================================= */
#define OR_NTYPE 17
#define OR_NPOT 8192
static const char *or_sym[OR_NTYPE] = {"Cd", "Se", "In", "As", "Si", "H",  "Zn", "S", "P1",
                                       "P2", "P3", "P4", "Te", "Cdz", "Sez", "Ga", "P"};

static long or_type_of(const char *s) { /* read.c:126-152 */
  for (long j = 0; j < OR_NTYPE; j++)
    if (!strcmp(s, or_sym[j])) return j;
  return -1;
}

static FILE *open_in(const char *dir, const char *name, const char *mode) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  return fopen(path, mode);
}

int or_run(const char *indir, const char *outdir, long idum, long nthreads) {
  if (nthreads > 0) omp_set_num_threads((int)nthreads);
  FILE *pf = open_in(indir, "input.par", "r");
  if (!pf) {
    set_err("cannot open input.par");
    return -1;
  }
  long nx, ny, nz, ms, ns, nc, nth;
  double Elmin, Elmax;
  if (fscanf(pf, "%ld %ld %ld %ld %ld %ld %lg %lg %ld", &nx, &ny, &nz, &ms, &ns, &nc, &Elmin,
             &Elmax, &nth) != 9) {
    fclose(pf);
    set_err("input.par: expected 9 tokens");
    return -1;
  }
  fclose(pf);
  long N = nx * ny * nz;
  /* conf.par (read.c:11-39): recentre on the unweighted mean */
  pf = open_in(indir, "conf.par", "r");
  if (!pf) {
    set_err("cannot open conf.par");
    return -1;
  }
  long natom;
  if (fscanf(pf, "%ld", &natom) != 1) {
    fclose(pf);
    set_err("conf.par: natom");
    return -1;
  }
  double *rx = malloc(sizeof(double) * natom), *ry = malloc(sizeof(double) * natom),
         *rz = malloc(sizeof(double) * natom), *scale = malloc(sizeof(double) * natom);
  long *type = malloc(sizeof(long) * natom);
  double xd = 0, yd = 0, zd = 0;
  for (long i = 0; i < natom; i++) {
    char sym[16];
    if (fscanf(pf, "%15s %lf %lf %lf", sym, &rx[i], &ry[i], &rz[i]) != 4) {
      fclose(pf);
      set_err("conf.par: atom %ld", i);
      return -1;
    }
    type[i] = or_type_of(sym);
    if (type[i] < 0) {
      fclose(pf);
      set_err("atom type %s not in current list", sym);
      return -1;
    }
    xd += rx[i];
    yd += ry[i];
    zd += rz[i];
    scale[i] = 1.0;
  }
  fclose(pf);
  xd /= (double)natom;
  yd /= (double)natom;
  zd /= (double)natom;
  for (long i = 0; i < natom; i++) {
    rx[i] -= xd;
    ry[i] -= yd;
    rz[i] -= zd;
  }
  /* optional per-atom scale (filterStrainCube/init.c:478-486, column 7) */
  pf = open_in(indir, "strain.par", "r");
  if (pf) {
    for (long i = 0; i < natom; i++) {
      int t1;
      char at[16];
      double a4, a5, v, vr0;
      if (fscanf(pf, "%d %15s %lg %lg %lg %lg %lg", &t1, at, &a4, &a5, &v, &vr0, &scale[i]) != 7) {
        fclose(pf);
        set_err("strain.par: row %ld", i);
        return -1;
      }
    }
    fclose(pf);
  }
  /* expot.par (init.c:66-81) */
  double ex_e = 0.0, ex_x0 = 0.0, ex_t = 1.0;
  pf = open_in(indir, "expot.par", "r");
  if (pf) {
    if (fscanf(pf, "%lg %lg %lg", &ex_e, &ex_x0, &ex_t) != 3) {
      fclose(pf);
      set_err("expot.par");
      return -1;
    }
    fclose(pf);
  }
  /* box (init.c:90-92, size.c:30-44) */
  double ext[3];
  double *rr[3] = {rx, ry, rz};
  for (int a = 0; a < 3; a++) {
    double dr2 = 0.0;
    for (long i = 0; i < natom - 1; i++)
      for (long j = i + 1; j < natom; j++) {
        double d = rr[a][i] - rr[a][j];
        double d2 = d * d;
        if (d2 > dr2) dr2 = d2;
      }
    ext[a] = rint(0.5 * sqrt(dr2) + 5.0);
  }
  double xmin = -ext[0], ymin = -ext[1], zmin = -ext[2];
  double dx = (ext[0] - xmin) / (double)nx, dy = (ext[1] - ymin) / (double)ny,
         dz = (ext[2] - zmin) / (double)nz;
  double dv = dx * dy * dz;
  /* pot tables (read.c:43-122) */
  double *vr = calloc((size_t)OR_NTYPE * OR_NPOT, sizeof(double));
  double *pot = calloc((size_t)OR_NTYPE * OR_NPOT, sizeof(double));
  long npot[OR_NTYPE];
  double dr[OR_NTYPE];
  double la[OR_NTYPE] = {0}, lb[OR_NTYPE] = {0};
  la[8] = 0.64;
  la[9] = -0.384;
  la[10] = 0.64;
  la[11] = -0.20;
  lb[8] = lb[9] = lb[10] = lb[11] = 2.2287033;
  for (long j = 0; j < OR_NTYPE; j++) {
    char name[32];
    snprintf(name, sizeof name, "pot%s.par", or_sym[j]);
    npot[j] = 0;
    dr[j] = 0.0;
    pf = open_in(indir, name, "r");
    if (pf) {
      long n = 0;
      while (n < OR_NPOT &&
             fscanf(pf, "%lg %lg", &vr[j * OR_NPOT + n], &pot[j * OR_NPOT + n]) == 2)
        n++;
      fclose(pf);
      npot[j] = n;
    } else {
      npot[j] = npot[0];
      for (long i = 0; i < npot[j]; i++) {
        vr[j * OR_NPOT + i] = vr[i];
        double r = vr[j * OR_NPOT + i];
        pot[j * OR_NPOT + i] = (la[j] * exp(-(r * r) / lb[j]));
      }
    }
    for (long i = 0; i < npot[j]; i++) {
      pot[j * OR_NPOT + i] -= pot[j * OR_NPOT + npot[j] - 1];
      dr[j] = vr[j * OR_NPOT + 1] - vr[j * OR_NPOT + 0];
    }
  }
  for (long i = 0; i < natom; i++)
    if (npot[type[i]] < 2) {
      set_err("no usable pot table for atom type %s", or_sym[type[i]]);
      return -1;
    }
  double *potl = malloc(sizeof(double) * N), *ksqr = malloc(sizeof(double) * N);
  or_build_potential(nx, ny, nz, xmin, ymin, zmin, dx, dy, dz, natom, rx, ry, rz, type, scale,
                     npot, vr, pot, OR_NPOT, dr, ex_e, ex_x0, ex_t, potl);
  or_ksqr(nx, ny, nz, dx, dy, dz, 10.0, ksqr);
  double Vmin = 1.0e10, Vmax = -1.0e10;
  for (long g = 0; g < N; g++) {
    if (Vmax < potl[g]) Vmax = potl[g];
    if (Vmin > potl[g]) Vmin = potl[g];
  }
  double el[4096];
  if (ms > 4096) {
    set_err("ms too large for the oracle");
    return -1;
  }
  {
    double del = (Elmax - Elmin) / (double)(ms - 1);
    for (long j = 0; j < ms; j++) el[j] = Elmin + (double)(j)*del;
  }
  /* spectral range (energy.c:53-83), dt (main.c:70) */
  double dE, emax;
  long iters;
  or_energy_range(nx, ny, nz, ksqr, potl, dv, Vmin, &dE, &emax, &iters);
  double dE_1 = 4.0 / dE;
  double dt = ((double)nc / (2.5 * dE)) * ((double)nc / (2.5 * dE));
  double *zn = malloc(sizeof(double) * nc), *an = malloc(sizeof(double) * nc * ms);
  or_leja_points(nc, zn);
  or_newton_coeffs(nc, ms, zn, el, dE, Vmin, dt, an);
  /* random states (main.c:79-89) */
  long mstot = ms * ns;
  double *psitot = malloc(sizeof(double) * (size_t)mstot * N);
  double *r0 = malloc(sizeof(double) * (size_t)ns * N);
  or_rng st;
  or_rng_reset(&st);
  long id = idum;
  for (long jns = 0; jns < ns; jns++) or_init_psi(&st, &id, N, dv, r0 + jns * N);
  double tf0 = omp_get_wtime();
  /* random states are independent (main.c:92-98): one per thread */
#pragma omp parallel for schedule(dynamic, 1)
  for (long jns = 0; jns < ns; jns++) {
    double *blk = psitot + (size_t)jns * ms * N;
    or_filter_state(nx, ny, nz, ksqr, potl, an, zn, nc, ms, dE_1, Vmin, r0 + jns * N, blk);
    or_normalize_all(blk, dv, ms, N);
    double ene[4096];
    or_energy_all(nx, ny, nz, ksqr, potl, dv, blk, ms, ene);
    char name[64];
    snprintf(name, sizeof name, "eval-filt-0-%ld.dat", jns);
    FILE *pe = open_in(outdir, name, "w");
    if (pe) {
      for (long j = 0; j < ms; j++) fprintf(pe, "%ld %.16g %.16g\n", j, ene[j], el[j]);
      fclose(pe);
    }
  }
  double tfilter = omp_get_wtime() - tf0;
  double *S = malloc(sizeof(double) * mstot);
  long cutoff = or_ortho(psitot, N, mstot, S);
  or_normalize_all(psitot, dv, cutoff, N);
  double *eval = malloc(sizeof(double) * cutoff), *sige = malloc(sizeof(double) * cutoff);
  or_hmat(nx, ny, nz, ksqr, potl, dv, psitot, cutoff, eval);
  or_normalize_all(psitot, dv, cutoff, N);
  FILE *pp = open_in(outdir, "psi.dat", "wb");
  if (pp) {
    fwrite(psitot, sizeof(double), (size_t)cutoff * N, pp);
    fclose(pp);
  }
  or_sigma(nx, ny, nz, ksqr, potl, dv, psitot, cutoff, sige);
  FILE *pv = open_in(outdir, "eval.dat", "w");
  if (pv) {
    for (long j = 0; j < cutoff; j++) fprintf(pv, "%ld %.16g %g\n", j, eval[j], sige[j]);
    fclose(pv);
  }
  FILE *ps = open_in(outdir, "oracle_summary.dat", "w");
  if (ps) {
    fprintf(ps, "nx %ld\nny %ld\nnz %ld\nnatom %ld\n", nx, ny, nz, natom);
    fprintf(ps, "xd %.17g\nyd %.17g\nzd %.17g\n", ext[0], ext[1], ext[2]);
    fprintf(ps, "dx %.17g\ndy %.17g\ndz %.17g\ndv %.17g\n", dx, dy, dz, dv);
    fprintf(ps, "Vmin %.17g\nVmax %.17g\n", Vmin, Vmax);
    fprintf(ps, "Emax %.17g\niters %ld\ndE %.17g\ndt %.17g\n", emax, iters, dE, dt);
    fprintf(ps, "cutoff %ld\nS0 %.17g\n", cutoff, S[0]);
    fprintf(ps, "t_filter %.6f\n", tfilter);
    fclose(ps);
  }
  free(rx); free(ry); free(rz); free(scale); free(type);
  free(vr); free(pot); free(potl); free(ksqr);
  free(zn); free(an); free(psitot); free(r0); free(S); free(eval); free(sige);
  return 0;
}
