/*
 * filter_spinors_main.c -- the spin-orbit filter program on libsep: a
 * drop-in for the reference's spinors/filterSpinors (main.c:4-195) with the
 * matrix form of the nonlocal spin-orbit term (input.par flagkb = 0).
 * Reads input.par (nx ny nz dx ms ns nc Elmin Elmax nthreads flagkb,
 * init.c:5-21), conf.par, pot<Sym>.par and an optional seed.par from the
 * working directory (or --dir); writes eval.dat, psi.dat (complex spinors,
 * zomplex[2 ngrid] per state, up then down, main.c:160-162),
 * eval-filt-0-<jns>.dat, Egrid.dat, conf.dat, pot<Sym>.dat and list.dat.
 *
 * Options (the reference ignores argv; these are extensions):
 *   --dir D      deck / output directory (default .)
 *   --seed S     ran_nrc seed (overrides seed.par; 0 = clock as main.c:81)
 *   --device N   HIP device ordinal
 * Failures write the message to a file named `error` and exit 1.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <sys/time.h>

#include "../../include/sep.h"

#define NTYPE 16   /* ist.natomtype, init.c:88 */
#define NPOT 8192  /* ist.npot, init.c:22 */

static const char *g_dir = ".";

static FILE *out_file(const char *name, const char *mode) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", g_dir, name);
  return fopen(path, mode);
}

static void die(const char *msg) {
  FILE *pf = out_file("error", "w");
  if (pf) {
    fprintf(pf, "%s\n", msg);
    fclose(pf);
  }
  fprintf(stderr, "filter_spinors: %s\n", msg);
  exit(1);
}

// host buffers: a failed allocation ends the program with a message
static void *xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) die("out of host memory");
  return p;
}

static void *xcalloc(size_t n, size_t size) {
  void *p = calloc(n ? n : 1, size);
  if (!p) die("out of host memory");
  return p;
}

static double wall(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

/* Atom types and their numbers (read.c:180-229). */
static const char *const TYPE_NAME[NTYPE] = {"Cd", "Se", "In", "As", "Si", "H",   "Zn", "S",
                                             "P1", "P2", "P3", "P4", "Te", "Cdz", "Sez", "Ga"};

static int type_of(const char *sym) {
  for (int t = 0; t < NTYPE; t++)
    if (!strcmp(sym, TYPE_NAME[t])) return t;
  return -1;
}

/* Spin-orbit strength per type (read.c:33-51, the As and Ga values of its
 * "GaAs" lines, which come last and win). */
static double vso_of(int t) {
  const double a = 6.057 / 0.529177;
  if (t == 2) return 0.48 * a * a * a / 8.0 / 27.2114; /* In */
  if (t == 3) return 3.15164354;                       /* As */
  if (t == 15) return 1.15454347;                      /* Ga */
  return 0.0;
}

/* Largest |r_i - r_j| along one axis (size.c:21-32). */
static double span(const double *r, long n, int axis) {
  double lo = r[axis], hi = r[axis];
  for (long i = 1; i < n; i++) {
    if (r[3 * i + axis] < lo) lo = r[3 * i + axis];
    if (r[3 * i + axis] > hi) hi = r[3 * i + axis];
  }
  return hi - lo;
}

int main(int argc, char **argv) {
  long seed_opt = 0;
  int have_seed_opt = 0, device = 0;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--dir") && i + 1 < argc)
      g_dir = argv[++i];
    else if (!strcmp(argv[i], "--seed") && i + 1 < argc) {
      seed_opt = atol(argv[++i]);
      have_seed_opt = 1;
    } else if (!strcmp(argv[i], "--device") && i + 1 < argc)
      device = atoi(argv[++i]);
    else {
      char msg[300];
      snprintf(msg, sizeof msg, "unknown option or missing value: %.250s", argv[i]);
      die(msg);
    }
  }
  time_t now = time(NULL);
  printf("This calculation began at: %s", ctime(&now));
  char err[512];
  /* input.par (init.c:5-21) */
  long nx, ny, nz, ms, ns, nc, nthreads, flagkb;
  double dx, Elmin, Elmax;
  {
    FILE *pf = out_file("input.par", "r");
    if (!pf) die("cannot read input.par");
    if (fscanf(pf, "%ld %ld %ld %lg %ld %ld %ld %lg %lg %ld %ld", &nx, &ny, &nz, &dx, &ms, &ns, &nc,
               &Elmin, &Elmax, &nthreads, &flagkb) != 11)
      die("input.par: expected nx ny nz dx ms ns nc Elmin Elmax nthreads flagkb");
    fclose(pf);
  }
  if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0.0) || ms < 1 || ns < 1 || nc < 2)
    die("input.par: bad sizes");
  if (flagkb != 0)
    die("the Kleinman-Bylander form (flagkb = 1) is not supported: use flagkb = 0, the matrix form");
  const double dy = dx, dz = dx;
  /* conf.par (init.c:23-45, read.c:3-78): recentred on the mean position */
  long natom = 0;
  FILE *pc = out_file("conf.par", "r");
  if (!pc || fscanf(pc, "%ld", &natom) != 1 || natom < 1) die("cannot read conf.par");
  double *xyz = (double *)xmalloc(sizeof(double) * 3 * natom);
  int *typ = (int *)xmalloc(sizeof(int) * natom);
  double *vso = (double *)xmalloc(sizeof(double) * natom);
  long nsemicond = 0;
  double mean[3] = {0.0, 0.0, 0.0};
  for (long i = 0; i < natom; i++) {
    char s[64];
    if (fscanf(pc, "%63s %lf %lf %lf", s, &xyz[3 * i], &xyz[3 * i + 1], &xyz[3 * i + 2]) != 4)
      die("conf.par: short atom list");
    typ[i] = type_of(s);
    if (typ[i] < 0) {
      snprintf(err, sizeof err, "atom type %s not in current list", s);
      die(err);
    }
    vso[i] = vso_of(typ[i]);
    if (typ[i] < 8 || typ[i] > 11) nsemicond++;
    for (int a = 0; a < 3; a++) mean[a] += xyz[3 * i + a];
  }
  fclose(pc);
  for (int a = 0; a < 3; a++) mean[a] /= (double)natom;
  for (long i = 0; i < natom; i++)
    for (int a = 0; a < 3; a++) xyz[3 * i + a] -= mean[a];
  {
    FILE *pf = out_file("conf.dat", "w");
    if (pf) {
      for (long i = 0; i < natom; i++)
        fprintf(pf, "%s %g %g %g %d %g\n", TYPE_NAME[typ[i]], xyz[3 * i], xyz[3 * i + 1],
                xyz[3 * i + 2], typ[i], vso[i]);
      fclose(pf);
    }
  }
  printf("nsemiconf = %ld\n", nsemicond);
  /* the box: the dot plus 5 bohr, at least the input sizes (init.c:46-79) */
  const double xd = rint(0.5 * span(xyz, natom, 0) + 5.0);
  const double yd = rint(0.5 * span(xyz, natom, 1) + 5.0);
  const double zd = rint(0.5 * span(xyz, natom, 2) + 5.0);
  printf("xd = %g yd = %g zd = %g\n", xd, yd, zd);
  long t = (long)((2.0 * xd) / dx);
  if (t > nx) nx = t;
  t = (long)((2.0 * yd) / dy);
  if (t > ny) ny = t;
  t = (long)((2.0 * zd) / dz) - 1;
  if (t > nz) nz = t;
  const double xmin = -((double)nx * dx) / 2.0, ymin = -((double)ny * dy) / 2.0,
               zmin = -((double)nz * dz) / 2.0;
  printf("new xd = %g yd = %g zd = %g\n", -xmin, -ymin, -zmin);
  const long N = nx * ny * nz;
  const double dv = dx * dy * dz;
  printf("nx = %ld  ny = %ld  nz = %ld npot = %d\n", nx, ny, nz, NPOT);
  printf("ms = %ld  ns = %ld natom = %ld nc = %ld\n", ms, ns, natom, nc);
  printf("Elmin = %g  Elmax = %g  Ekinmax = %g\n", Elmin, Elmax, 20.0);
  printf("ngrid = %ld nspin = 2\nthreads = %ld\n", N, nthreads);
  printf("Rnlcut = %g\n", sqrt(0.49 * 6.0 * log(10.0)));
  {
    const long dims[3] = {nx, ny, nz};
    for (int a = 0; a < 3; a++)
      if (!sep_fft_length_supported(dims[a])) {
        snprintf(err, sizeof err, "grid length %ld has no HIP FFT plan", dims[a]);
        die(err);
      }
  }
  /* pseudopotential tables (read.c:81-173): pot<Sym>.par where it exists,
   * else the passivating ligands' a exp(-r^2 / b) on the first type's r grid;
   * each table shifted to 0 at its last row */
  static const double la[NTYPE] = {0, 0, 0, 0, 0, 0, 0, 0, 0.64, -0.384, 0.64, -0.384, 0, 0, 0, 0};
  static const double lb[NTYPE] = {0, 0, 0, 0, 0, 0, 0, 0, 2.2287033, 2.2287033, 2.2287033,
                                   2.2287033, 0, 0, 0, 0};
  double *vr = (double *)xcalloc((size_t)NTYPE * NPOT, sizeof(double));
  double *pot = (double *)xcalloc((size_t)NTYPE * NPOT, sizeof(double));
  long npot[NTYPE];
  double drt[NTYPE];
  int used[NTYPE] = {0};
  for (long i = 0; i < natom; i++) used[typ[i]] = 1;
  for (int j = 0; j < NTYPE; j++) {
    char name[32];
    snprintf(name, sizeof name, "pot%s.par", TYPE_NAME[j]);
    FILE *pf = out_file(name, "r");
    npot[j] = 0;
    if (pf) {
      while (npot[j] < NPOT &&
             fscanf(pf, "%lg %lg", &vr[j * NPOT + npot[j]], &pot[j * NPOT + npot[j]]) == 2)
        npot[j]++;
      double extra;
      const int more = fscanf(pf, "%lg", &extra) == 1;
      fclose(pf);
      if (more) {
        snprintf(err, sizeof err, "%s holds more than %d rows", name, NPOT);
        die(err);
      }
    } else if (j >= 8 && j <= 11) {
      npot[j] = npot[0];
      for (long i = 0; i < npot[j]; i++) {
        vr[j * NPOT + i] = vr[i];
        pot[j * NPOT + i] = la[j] * exp(-vr[i] * vr[i] / lb[j]);
      }
    }
    if (used[j] && npot[j] < 2) {
      snprintf(err, sizeof err, "no potential table for atom type %s (%s)", TYPE_NAME[j], name);
      die(err);
    }
    if (npot[j] < 2) { /* unused type: a zero table */
      npot[j] = 2;
      vr[j * NPOT] = 0.0;
      vr[j * NPOT + 1] = 1.0;
      pot[j * NPOT] = pot[j * NPOT + 1] = 0.0;
      drt[j] = 1.0;
      continue;
    }
    const double last = pot[j * NPOT + npot[j] - 1];
    for (long i = 0; i < npot[j]; i++) pot[j * NPOT + i] -= last;
    drt[j] = vr[j * NPOT + 1] - vr[j * NPOT];
    snprintf(name, sizeof name, "pot%s.dat", TYPE_NAME[j]);
    FILE *po = out_file(name, "w");
    if (po) {
      for (long i = 0; i < npot[j]; i++) fprintf(po, "%g %g\n", vr[j * NPOT + i], pot[j * NPOT + i]);
      fclose(po);
    }
  }

  sep_grid_desc gdesc = {nx, ny, nz, dx, dy, dz, xmin, ymin, zmin, 20.0, device};
  sep_ctx *ctx = NULL;
  if (sep_open(&gdesc, &ctx) != SEP_OK) die("sep_open failed");
#define CHECK(call)                                                              \
  do {                                                                           \
    int rc_ = (call);                                                            \
    if (rc_ != SEP_OK) {                                                         \
      snprintf(err, sizeof err, "%s: %s (%d)", #call, sep_last_error(ctx), rc_); \
      die(err);                                                                  \
    }                                                                            \
  } while (0)
  /* V(r) on the grid (init.c:470-497) */
  double Vmin, Vmax;
  CHECK(sep_build_potential(ctx, natom, xyz, typ, NULL, NTYPE, npot, vr, pot, NPOT, drt, NULL,
                            NULL, &Vmin, &Vmax));
  printf("dV = %g vmin = %g Vmax = %g\n", Vmax - Vmin, Vmin, Vmax);
  /* the spin-orbit lists of the first nsemicond atoms (init.c:114-183) */
  long nent = 0;
  CHECK(sep_so_setup(ctx, nsemicond, xyz, vso, &nent));
  {
    /* list.dat: grid points within Rnlcut of each of those atoms */
    const double rc2 = 0.49 * 6.0 * log(10.0);
    FILE *pf = out_file("list.dat", "w");
    for (long a = 0; a < nsemicond && pf; a++) {
      long cnt = 0;
      for (long k = 0; k < nz; k++) {
        const double ddz = zmin + (double)k * dz - xyz[3 * a + 2];
        if (ddz * ddz >= rc2) continue;
        for (long j = 0; j < ny; j++) {
          const double ddy = ymin + (double)j * dy - xyz[3 * a + 1];
          if (ddz * ddz + ddy * ddy >= rc2) continue;
          for (long i = 0; i < nx; i++) {
            const double ddx = xmin + (double)i * dx - xyz[3 * a];
            if (ddx * ddx + ddy * ddy + ddz * ddz < rc2) cnt++;
          }
        }
      }
      fprintf(pf, "%ld %ld\n", a, cnt);
    }
    if (pf) fclose(pf);
  }
  /* energy targets (init.c:506-513) */
  double *el = (double *)xmalloc(sizeof(double) * ms);
  {
    const double del = ms > 1 ? (Elmax - Elmin) / (double)(ms - 1) : 0.0;
    FILE *pf = out_file("Egrid.dat", "w");
    for (long j = 0; j < ms; j++) {
      el[j] = Elmin + (double)j * del;
      if (pf) fprintf(pf, "%g\n", el[j]);
    }
    if (pf) fclose(pf);
  }
  /* spectral range (main.c:66-73) */
  double t0 = wall(), c0 = (double)clock();
  double vmin, dE;
  long iters[2];
  CHECK(sep_energy_range_spinor(ctx, &vmin, &dE, iters));
  printf("done calculate energy range, CPU time (sec) %g, wall run time (sec) %g\n",
         ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
  /* Newton interpolation (main.c:75-80) */
  const double dt = ((double)nc / (2.5 * dE)) * ((double)nc / (2.5 * dE));
  printf("nc = %ld dt = %g dE = %g\n", nc, dt, dE);
  printf("energy range: Vmin = %.17g dE = %.17g\n", vmin, dE);
  fflush(stdout);
  double *zn = (double *)xmalloc(sizeof(double) * nc);
  double *an = (double *)xmalloc(sizeof(double) * nc * ms);
  sep_leja_points(nc, zn);
  if (sep_newton_coeffs(nc, ms, zn, el, dE, vmin, dt, an) != SEP_OK) die("coefficient failed");
  /* random spinors: init_psi on the up, then the down component of each,
   * real parts random, each component normalized (main.c:81-96,
   * init.c:546-563) */
  long idum, seed = 0;
  if (have_seed_opt) {
    seed = seed_opt;
  } else {
    FILE *ps = out_file("seed.par", "r");
    if (ps) {
      if (fscanf(ps, "%ld", &seed) != 1) seed = 0;
      fclose(ps);
    }
  }
  if (seed == 0) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    srandom((unsigned)(tv.tv_sec ^ tv.tv_usec));
    idum = -random();
  } else {
    idum = seed < 0 ? seed : -seed;
  }
  printf("seed = %ld\n", idum);
  fflush(stdout);
  double *psi0 = (double *)xcalloc((size_t)ns * 4 * N, sizeof(double));
  sep_rng rng;
  sep_rng_reset(&rng);
  for (long j = 0; j < ns; j++) {
    sep_init_psi(&rng, &idum, N, dv, psi0 + (4 * j + 0) * N); /* up.re */
    sep_init_psi(&rng, &idum, N, dv, psi0 + (4 * j + 2) * N); /* dn.re */
  }
  /* filter (main.c:98-107, filter.c:5-48) */
  t0 = wall();
  c0 = (double)clock();
  CHECK(sep_filter_spinor(ctx, psi0, ns, an, zn, nc, ms, vmin, dE));
  CHECK(sep_trim(ctx));
  printf("done calculating filter, CPU time (sec) %g, wall run time (sec) %g\n",
         ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
  const long mstot = ms * ns;
  {
    double *ene = (double *)xmalloc(sizeof(double) * mstot);
    CHECK(sep_energy_spinor(ctx, mstot, ene));
    for (long jns = 0; jns < ns; jns++) { /* filter.c:24-28 */
      char name[64];
      snprintf(name, sizeof name, "eval-filt-0-%ld.dat", jns);
      FILE *pf = out_file(name, "w");
      if (pf) {
        for (long j = 0; j < ms; j++) fprintf(pf, "%ld %.16g %.16g\n", j, ene[jns * ms + j], el[j]);
        fclose(pf);
      }
    }
    free(ene);
  }
  fflush(stdout);
  /* ortho (main.c:117-124, ortho.c:5-45) */
  t0 = wall();
  c0 = (double)clock();
  double *S = (double *)xmalloc(sizeof(double) * mstot);
  long kept = 0;
  CHECK(sep_ortho_spinor(ctx, mstot, 1.0e-10, S, &kept));
  printf("cutoff is %ld\n", kept);
  printf("mstot ortho = %ld\n", kept);
  printf("done calculating ortho, CPU time (sec) %g, wall run time (sec) %g\n",
         ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
  /* Rayleigh-Ritz (main.c:129-137, Hmat.c:5-72) */
  t0 = wall();
  c0 = (double)clock();
  double *eval = (double *)xmalloc(sizeof(double) * kept);
  double *sige = (double *)xmalloc(sizeof(double) * kept);
  CHECK(sep_rayleigh_ritz_spinor(ctx, kept, eval));
  printf("done calculating Hmat, CPU time (sec) %g, wall run time (sec) %g\n",
         ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
  fflush(stdout);
  /* psi.dat: zomplex[2 ngrid] per state, up then down (main.c:139-142) */
  {
    double *rows = (double *)xmalloc(sizeof(double) * 4 * N * kept);
    double *z = (double *)xmalloc(sizeof(double) * 4 * N);
    CHECK(sep_states_download(ctx, rows, 4 * kept));
    FILE *pf = out_file("psi.dat", "wb");
    if (!pf) die("cannot write psi.dat");
    for (long s = 0; s < kept; s++) {
      const double *r = rows + 4 * N * s;
      for (long c = 0; c < 2; c++)
        for (long g = 0; g < N; g++) {
          z[2 * (c * N + g)] = r[(2 * c) * N + g];
          z[2 * (c * N + g) + 1] = r[(2 * c + 1) * N + g];
        }
      fwrite(z, sizeof(double), 4 * N, pf);
    }
    fclose(pf);
    free(rows);
    free(z);
  }
  /* sigma and eval.dat (main.c:144-152) */
  CHECK(sep_sigma_spinor(ctx, kept, sige));
  {
    FILE *pf = out_file("eval.dat", "w");
    if (!pf) die("cannot write eval.dat");
    for (long j = 0; j < kept; j++) fprintf(pf, "%ld %.16g %g\n", j, eval[j], sige[j]);
    fclose(pf);
  }
  sep_close(ctx);
  now = time(NULL);
  printf("This calculation ended at: %s", ctime(&now));
  free(xyz); free(typ); free(vso); free(vr); free(pot); free(el); free(zn); free(an);
  free(psi0); free(S); free(eval); free(sige);
  return 0;
}
