/*
 * sep_host.c -- host C parts of libsep:
 *   random states        filter/rand.c:28-66, filter/init.c:236-251
 *   Newton coefficients  filter/coeff.c:16-134 (kept on the host in FP64)
 *   dense FP64 algebra   replaces MKL dsyev_/dgesvd_ (filter/Hmat.c:55,
 *                        filter/ortho.c:32) for the small projected matrices
 *   deck reader          filter/init.c:12-122, filter/read.c:11-176
 */
#define _GNU_SOURCE
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sep.h"
#include "../../include/sep_deck.h"
#include "sep_host.h"

/* ===================================================== random numbers */
/* L'Ecuyer two-generator combination with a 32-entry Bays-Durham shuffle;
 * constants as filter/rand.c:11-24. */
enum {
  RN_NTAB = 32,
};
static const long RN_M1 = 2147483563L, RN_M2 = 2147483399L;
static const long RN_A1 = 40014L, RN_A2 = 40692L;
static const long RN_Q1 = 53668L, RN_Q2 = 52774L; /* m / a */
static const long RN_R1 = 12211L, RN_R2 = 3791L;  /* m % a */

void sep_rng_reset(sep_rng *st) {
  st->idum2 = 123456789L;
  st->iy = 0;
  memset(st->iv, 0, sizeof st->iv);
}

/* x <- a x mod m without overflow (Schrage) */
static inline long schrage(long x, long a, long q, long r, long m) {
  long k = x / q;
  x = a * (x - k * q) - k * r;
  return x < 0 ? x + m : x;
}

double sep_ran_nrc(sep_rng *st, long *idum) {
  const long mm1 = RN_M1 - 1, ndiv = 1 + mm1 / RN_NTAB;
  if (*idum <= 0) { /* (re)seed: 8 warm-ups, then fill the table */
    *idum = (-(*idum) < 1) ? 1 : -(*idum);
    st->idum2 = *idum;
    for (int j = RN_NTAB + 7; j >= 0; j--) {
      *idum = schrage(*idum, RN_A1, RN_Q1, RN_R1, RN_M1);
      if (j < RN_NTAB) st->iv[j] = *idum;
    }
    st->iy = st->iv[0];
  }
  *idum = schrage(*idum, RN_A1, RN_Q1, RN_R1, RN_M1);
  st->idum2 = schrage(st->idum2, RN_A2, RN_Q2, RN_R2, RN_M2);
  long j = st->iy / ndiv;
  st->iy = st->iv[j] - st->idum2;
  st->iv[j] = *idum;
  if (st->iy < 1) st->iy += mm1;
  double t = (1.0 / (double)RN_M1) * st->iy;
  const double rnmx = 1.0 - 1.2e-7;
  return t > rnmx ? rnmx : t;
}

void sep_init_psi(sep_rng *st, long *idum, long ngrid, double dv, double *psi) {
  long id = *idum;
  for (long g = 0; g < ngrid; g++) psi[g] = -1.0 + 2.0 * sep_ran_nrc(st, &id);
  double s = 0.0;
  for (long g = 0; g < ngrid; g++) s += psi[g] * psi[g];
  double nrm = sqrt(s * dv);
  for (long g = 0; g < ngrid; g++) psi[g] /= nrm;
  *idum = id;
}

/* ================================================ Newton interpolation */
double sep_leja_points(long nc, double *zn) {
  /* Candidates: 28 nc points evenly on [-2, 2] followed by 4 nc points at
   * the origin (the reference's imaginary range is 0, coeff.c:85,91-98).
   * Greedy: each new point maximises sum_k log|c - z_k|^2 over the
   * candidates, first maximum wins; candidates within 1e-5 of a chosen
   * point are retired (coeff.c:109-123). */
  const long n3 = 32 * nc, nzero = n3 / 8, nreal = n3 - nzero;
  const double step = 4.0 / (double)(nreal - 1);
  double *cre = (double *)malloc(sizeof(double) * n3);
  double *score = (double *)malloc(sizeof(double) * n3);
  double *pre = (double *)malloc(sizeof(double) * (nc > 0 ? nc : 1));
  if (!cre || !score || !pre) {
    free(cre); free(score); free(pre);
    return 0.0;
  }
  for (long k = 0; k < nreal; k++) cre[k] = -2.0 + (double)k * step;
  for (long k = nreal; k < n3; k++) cre[k] = 0.0;
  pre[0] = cre[0];
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n3; k++) {
    double d = cre[k] - pre[0];
    double d2 = d * d;
    score[k] = d2 < 1.0e-10 ? -1.0e30 : log(d2);
  }
  for (long j = 1; j < nc; j++) {
    long best = 0;
    double bval = -2.0e30;
#pragma omp parallel
    {
      long lb = -1;
      double lv = -2.0e30;
#pragma omp for schedule(static) nowait
      for (long k = 0; k < n3; k++)
        if (score[k] > lv) {
          lv = score[k];
          lb = k;
        }
#pragma omp critical
      {
        if (lb >= 0 && (lv > bval || (lv == bval && lb < best))) {
          bval = lv;
          best = lb;
        }
      }
    }
    pre[j] = cre[best];
    const double z = pre[j];
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n3; k++) {
      double d = cre[k] - z;
      double d2 = d * d;
      if (d2 < 1.0e-10)
        score[k] = -1.0e30;
      else
        score[k] += log(d2);
    }
  }
  double prod = 1.0;
  for (long j = 0; j < nc; j++) prod *= pre[j] * pre[j];
  double fk = pow(sqrt(prod), 1.0 / (double)nc);
  memcpy(zn, pre, sizeof(double) * nc);
  free(cre);
  free(score);
  free(pre);
  return fk;
}

int sep_newton_coeffs(long nc, long ms, const double *zn, const double *el, double dE,
                      double vmin, double dt, double *an) {
  if (nc < 1 || ms < 1 || !(dE > 0.0)) return SEP_ERR_ARG;
  const double scale = 4.0 / dE, pref = sqrt(dt / 3.14159265358979323846);
#pragma omp parallel for schedule(dynamic, 1)
  for (long ie = 0; ie < ms; ie++) {
    double *a = an + ie * nc;
    /* divided differences of f at E_j = (z_j + 2)/scale + vmin, in the
     * order of filter/coeff.c:29-53 */
    for (long j = 0; j < nc; j++) {
      double x = (zn[j] + 2.0) / scale + vmin - el[ie];
      double fj = pref * exp(-(x * x) * dt);
      if (j == 0) {
        a[0] = fj;
        continue;
      }
      double w = 1.0, acc = 0.0;
      for (long i = 1; i < j; i++) {
        w *= (zn[j] - zn[i - 1]);
        acc += a[i] * w;
      }
      w *= (zn[j] - zn[j - 1]);
      a[j] = (fj - a[0] - acc) / w;
    }
  }
  return SEP_OK;
}

/* =========================================================== dense FP64 */
/* ------------------------------------------- double-double Cholesky */
typedef struct {
  double hi, lo;
} dd_t;
static inline dd_t dd_two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  dd_t r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static inline dd_t dd_norm2(double hi, double lo) {
  double s = hi + lo;
  dd_t r = {s, lo - (s - hi)};
  return r;
}
static inline dd_t dd_add(dd_t a, dd_t b) {
  dd_t s = dd_two_sum(a.hi, b.hi);
  return dd_norm2(s.hi, s.lo + a.lo + b.lo);
}
static inline dd_t dd_mul(dd_t a, dd_t b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  return dd_norm2(p, e + a.hi * b.lo + a.lo * b.hi);
}
static inline dd_t dd_neg(dd_t a) {
  dd_t r = {-a.hi, -a.lo};
  return r;
}
static inline dd_t dd_div(dd_t a, dd_t b) {
  double q1 = a.hi / b.hi;
  dd_t q1b = dd_mul((dd_t){q1, 0.0}, b);
  dd_t r = dd_add(a, dd_neg(q1b));
  double q2 = r.hi / b.hi;
  return dd_norm2(q1, q2);
}
static inline dd_t dd_sqrt(dd_t a) {
  double x = sqrt(a.hi);
  dd_t xx = dd_mul((dd_t){x, 0.0}, (dd_t){x, 0.0});
  dd_t r = dd_add(a, dd_neg(xx));
  return dd_norm2(x, r.hi / (2.0 * x));
}

int sep_la_chol_dd(long n, const double *hi, const double *lo, double shift, double *Rout) {
  dd_t *R = (dd_t *)malloc(sizeof(dd_t) * n * n);
  for (long i = 0; i < n * n; i++) {
    R[i].hi = hi[i];
    R[i].lo = lo ? lo[i] : 0.0;
  }
  for (long i = 0; i < n; i++) R[i * n + i] = dd_add(R[i * n + i], (dd_t){shift, 0.0});
  int status = 0;
  for (long j = 0; j < n; j++) {
    dd_t d = R[j * n + j];
    for (long k = 0; k < j; k++) d = dd_add(d, dd_neg(dd_mul(R[k * n + j], R[k * n + j])));
    if (!(d.hi > 0.0)) {
      status = (int)(j + 1);
      break;
    }
    dd_t rjj = dd_sqrt(d);
    R[j * n + j] = rjj;
#pragma omp parallel for schedule(static) if (n - j > 64)
    for (long c = j + 1; c < n; c++) {
      dd_t s = R[j * n + c];
      for (long k = 0; k < j; k++) s = dd_add(s, dd_neg(dd_mul(R[k * n + j], R[k * n + c])));
      R[j * n + c] = dd_div(s, rjj);
    }
  }
  for (long i = 0; i < n; i++)
    for (long j = 0; j < n; j++) Rout[i * n + j] = (j >= i && !status) ? R[i * n + j].hi : 0.0;
  free(R);
  return status;
}

/* ============================================================ deck reader */
static const char *const type_sym[SEP_NTYPE] = {"Cd", "Se", "In", "As", "Si", "H",
                                                "Zn", "S",  "P1", "P2", "P3", "P4",
                                                "Te", "Cdz", "Sez", "Ga", "P"};

const char *sep_type_symbol(int t) { return (t >= 0 && t < SEP_NTYPE) ? type_sym[t] : "?"; }

static FILE *open_rel(const char *dir, const char *name) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", dir ? dir : ".", name);
  return fopen(path, "r");
}

void sep_deck_free(sep_deck *d) {
  free(d->xyz);
  free(d->type);
  free(d->sym);
  free(d->scale);
  free(d->vr);
  free(d->pot);
  memset(d, 0, sizeof *d);
}

static int deck_read(const char *dir, sep_deck *d, char *err, int errlen);

int sep_deck_read(const char *dir, sep_deck *d, char *err, int errlen) {
  const int rc = deck_read(dir, d, err, errlen);
  if (rc != SEP_OK) sep_deck_free(d);  // a refused deck holds nothing
  return rc;
}

static int deck_read(const char *dir, sep_deck *d, char *err, int errlen) {
  memset(d, 0, sizeof *d);
  FILE *f = open_rel(dir, "input.par");
  if (!f) f = open_rel(dir, "filter_input.par");
  if (!f) {
    snprintf(err, errlen, "cannot open input.par");
    return SEP_ERR_IO;
  }
  int n = fscanf(f, "%ld %ld %ld %ld %ld %ld %lg %lg %ld", &d->nx, &d->ny, &d->nz, &d->ms,
                 &d->ns, &d->nc, &d->Elmin, &d->Elmax, &d->nthreads);
  fclose(f);
  if (n != 9) {
    snprintf(err, errlen, "input.par: expected 9 leading tokens, read %d", n);
    return SEP_ERR_IO;
  }
  if (d->nx < 2 || d->ny < 2 || d->nz < 2 || d->ms < 1 || d->ns < 1 || d->nc < 1) {
    snprintf(err, errlen, "input.par: bad sizes");
    return SEP_ERR_ARG;
  }
  /* conf.par */
  f = open_rel(dir, "conf.par");
  if (!f) {
    snprintf(err, errlen, "cannot open conf.par");
    return SEP_ERR_IO;
  }
  if (fscanf(f, "%ld", &d->natom) != 1 || d->natom < 1) {
    fclose(f);
    snprintf(err, errlen, "conf.par: bad natom");
    return SEP_ERR_IO;
  }
  const long na = d->natom;
  d->xyz = (double *)malloc(sizeof(double) * 3 * na);
  d->type = (int *)malloc(sizeof(int) * na);
  d->sym = (char(*)[4])calloc(na, 4);
  d->scale = (double *)malloc(sizeof(double) * na);
  double mean[3] = {0, 0, 0};
  for (long i = 0; i < na; i++) {
    char s[64];
    if (fscanf(f, "%63s %lf %lf %lf", s, &d->xyz[3 * i], &d->xyz[3 * i + 1], &d->xyz[3 * i + 2]) !=
        4) {
      fclose(f);
      snprintf(err, errlen, "conf.par: atom %ld unreadable (natom mismatch?)", i);
      return SEP_ERR_IO;
    }
    int t = -1;
    for (int k = 0; k < SEP_NTYPE; k++)
      if (!strcmp(s, type_sym[k])) t = k;
    if (t < 0) {
      fclose(f);
      snprintf(err, errlen, "atom type %s not in current list", s);
      return SEP_ERR_IO;
    }
    d->type[i] = t;
    snprintf(d->sym[i], sizeof d->sym[i], "%.3s", s);
    d->scale[i] = 1.0;
    for (int a = 0; a < 3; a++) mean[a] += d->xyz[3 * i + a];
  }
  fclose(f);
  for (int a = 0; a < 3; a++) mean[a] /= (double)na;
  for (long i = 0; i < na; i++)
    for (int a = 0; a < 3; a++) d->xyz[3 * i + a] -= mean[a];
  /* strain.par */
  f = open_rel(dir, "strain.par");
  if (f) {
    for (long i = 0; i < na; i++) {
      int idx;
      char at[64];
      double a4, a5, v, v0;
      if (fscanf(f, "%d %63s %lg %lg %lg %lg %lg", &idx, at, &a4, &a5, &v, &v0, &d->scale[i]) != 7) {
        fclose(f);
        snprintf(err, errlen, "strain.par: row %ld unreadable", i);
        return SEP_ERR_IO;
      }
    }
    fclose(f);
    d->has_strain = 1;
  }
  /* expot.par */
  d->expot[0] = 0.0;
  d->expot[1] = 0.0;
  d->expot[2] = 1.0;
  f = open_rel(dir, "expot.par");
  if (f) {
    if (fscanf(f, "%lg %lg %lg", &d->expot[0], &d->expot[1], &d->expot[2]) != 3) {
      fclose(f);
      snprintf(err, errlen, "expot.par: expected e x0 kT");
      return SEP_ERR_IO;
    }
    fclose(f);
    d->has_expot = 1;
  }
  /* seed.par */
  f = open_rel(dir, "seed.par");
  if (f) {
    if (fscanf(f, "%ld", &d->seed) == 1 && d->seed != 0) d->has_seed = 1;
    fclose(f);
  }
  /* pot tables */
  d->vr = (double *)calloc((size_t)SEP_NTYPE * SEP_NPOT, sizeof(double));
  d->pot = (double *)calloc((size_t)SEP_NTYPE * SEP_NPOT, sizeof(double));
  const double lig_a[SEP_NTYPE] = {[8] = 0.64, [9] = -0.384, [10] = 0.64, [11] = -0.20};
  const double lig_b = 2.2287033;
  int from_file[SEP_NTYPE];
  for (int j = 0; j < SEP_NTYPE; j++) {
    char name[32];
    snprintf(name, sizeof name, "pot%s.par", type_sym[j]);
    double *vr = d->vr + (size_t)j * SEP_NPOT, *pot = d->pot + (size_t)j * SEP_NPOT;
    f = open_rel(dir, name);
    from_file[j] = f != NULL;
    if (f) {
      long k = 0;
      while (k < SEP_NPOT && fscanf(f, "%lg %lg", &vr[k], &pot[k]) == 2) k++;
      double extra[2];
      // more rows than the reference's table slot holds (filter/read.c:90-92
      // would write them into the next type's table): refused, not truncated
      const int over = k == SEP_NPOT && fscanf(f, "%lg %lg", &extra[0], &extra[1]) == 2;
      fclose(f);
      if (over) {
        snprintf(err, errlen, "%s: more than %d rows", name, SEP_NPOT);
        return SEP_ERR_IO;
      }
      d->npot[j] = k;
    } else {
      d->npot[j] = d->npot[0];
      for (long k = 0; k < d->npot[j]; k++) {
        vr[k] = d->vr[k];
        double b = (j >= 8 && j <= 11) ? lig_b : 1.0;
        pot[k] = lig_a[j] * exp(-(vr[k] * vr[k]) / b);
      }
    }
    if (d->npot[j] > 0) {
      double last = pot[d->npot[j] - 1];
      for (long k = 0; k < d->npot[j]; k++) pot[k] -= last;
    }
    d->dr[j] = d->npot[j] > 1 ? vr[1] - vr[0] : 0.0;
  }
  for (long i = 0; i < na; i++)
    if (d->npot[d->type[i]] < 2 || (!from_file[d->type[i]] && (d->type[i] < 8 || d->type[i] > 11))) {
      snprintf(err, errlen, "no pot table for atom type %s (and no potCd.par grid)",
               type_sym[d->type[i]]);
      return SEP_ERR_IO;
    }
  /* box, filter/init.c:90-122 with get_dot_ligand_size_z (filter/size.c:30-44) */
  double half[3];
  for (int a = 0; a < 3; a++) {
    double lo = d->xyz[a], hi = d->xyz[a];
    for (long i = 1; i < na; i++) {
      double c = d->xyz[3 * i + a];
      if (c < lo) lo = c;
      if (c > hi) hi = c;
    }
    double ext = hi - lo;
    half[a] = rint(0.5 * sqrt(ext * ext) + 5.0);
  }
  d->xd = half[0];
  d->yd = half[1];
  d->zd = half[2];
  d->xmin = -d->xd;
  d->ymin = -d->yd;
  d->zmin = -d->zd;
  d->dx = (d->xd - d->xmin) / (double)d->nx;
  d->dy = (d->yd - d->ymin) / (double)d->ny;
  d->dz = (d->zd - d->zmin) / (double)d->nz;
  d->dv = d->dx * d->dy * d->dz;
  return SEP_OK;
}
