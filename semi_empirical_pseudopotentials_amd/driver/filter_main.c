/*
 * filter_main.c -- the `filter` program on libsep: a drop-in for the
 * reference's filter.x (filter/main.c:11-162).  Reads input.par, conf.par,
 * pot*.par (+ optional expot.par, strain.par, seed.par) from the working
 * directory (or --dir), writes eval.dat, psi.dat and the reference's side
 * files eval-filt-0-<jns>.dat, Egrid.dat, func.dat, coeff.dat, conf.dat,
 * pot<Sym>.dat, pS.dat, localPot.cube, ene-init.dat, prop0.dat, hmat.dat
 * (psi-filt-0-<jns>.dat with --psi-filt).
 *
 * Options (the reference ignores argv; these are extensions):
 *   --dir D            deck / output directory (default .)
 *   --seed S           ran_nrc seed (overrides seed.par; 0 = clock as main.c:79)
 *   --precision P      fp64 (default) or fp32: complex64 filter, FP64-Gram ortho,
 *                      FP64 Rayleigh-Ritz and sigma, then FP64 preconditioned
 *                      correction passes (sep_refine_window) on the pairs with
 *                      energies in [Elmin, Elmax] of input.par
 *   --refine-passes P  number of those passes (default 10; 0 = none)
 *   --device N         HIP device ordinal
 *   --lanes L          filter lanes (groups of random states on concurrent
 *                      streams): 0 = automatic (default), 1 = one stream, up
 *                      to SEP_MAX_LANES; results do not depend on it
 *   --psi-filt         also write psi-filt-0-<jns>.dat (filter/filter.c:44-48)
 *   --resume-filt      skip the random states and the filter: read the filtered
 *                      states from psi-filt-0-<jns>.dat of an earlier --psi-filt
 *                      run and continue at ortho (the reading that
 *                      filter/main.c:102-106 holds commented out)
 * Failures write the message to a file named `error` and exit 1, as the
 * reference's nerror() does (filter/nerror.c:9-16).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <sys/time.h>

#include "../../include/sep.h"
#include "../../include/sep_deck.h"

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
  fprintf(stderr, "filter: %s\n", msg);
  exit(1);
}

// host buffers: a failed allocation ends the program with a message
static void *xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) die("out of host memory");
  return p;
}

static double wall(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static void stamp(void) {
  time_t t = time(NULL);
  fputs(ctime(&t), stdout);
}
static void separation(void) {
  printf("\n******************************************************************************\n\n");
}

/* check_function, filter/coeff.c:138-163: Newton interpolant vs the exact
 * Gaussian on [-2, 2] for the first target. */
static void write_func(const double *an0, const double *zn, long nc, double dE, double vmin,
                       double dt, double el0) {
  FILE *pf = out_file("func.dat", "w");
  if (!pf) return;
  for (double x = -2.0; x < 2.0; x += 0.01) {
    double xn = 1.0, f = an0[0];
    for (long j = 1; j < nc; j++) {
      xn *= (x - zn[j - 1]);
      f += an0[j] * xn;
    }
    double xu = (x + 2.0) * dE / 4.0 + vmin;
    double g = sqrt(dt / 3.14159265358979323846) * exp(-((xu - el0) * (xu - el0)) * dt);
    fprintf(pf, "%g %g %g\n", xu, f, g);
  }
  fclose(pf);
}

/* conf.dat (filter/read.c:32-36): recentred positions and type index. */
static void write_conf(const sep_deck *d) {
  FILE *pw = out_file("conf.dat", "w");
  if (!pw) return;
  for (long i = 0; i < d->natom; i++)
    fprintf(pw, "%s %g %g %g %d\n", d->sym[i], d->xyz[3 * i], d->xyz[3 * i + 1],
            d->xyz[3 * i + 2], d->type[i]);
  fclose(pw);
}

/* pot<Sym>.dat (filter/read.c:108-116): every type's shifted table. */
static void write_pots(const sep_deck *d) {
  for (int j = 0; j < SEP_NTYPE; j++) {
    char name[32];
    snprintf(name, sizeof name, "pot%s.dat", sep_type_symbol(j));
    FILE *pf = out_file(name, "w");
    if (!pf) continue;
    for (long k = 0; k < d->npot[j]; k++)
      fprintf(pf, "%g %g\n", d->vr[j * SEP_NPOT + k], d->pot[j * SEP_NPOT + k]);
    fclose(pf);
  }
}

/* pS.dat (filter/init.c:185-190): the external potential along z,
 * e (1 / (exp((|z| - x0) / kT) + 1) - 1) (init.c:255-257). */
static void write_ps(const sep_deck *d) {
  FILE *pf = out_file("pS.dat", "w");
  if (!pf) return;
  double z = d->zmin;
  for (long jz = 0; jz < d->nz; jz++, z += d->dz) {
    const double e = d->expot[0] * (1.0 / (exp((fabs(z) - d->expot[1]) / d->expot[2]) + 1.0) - 1.0);
    fprintf(pf, "% .6f % .8f\n", z, e);
  }
  fclose(pf);
}

/* localPot.cube (filter/write.c:31-119, called at init.c:192): header,
 * the atoms of conf.par as read there (2-character symbols, positions as in
 * the file), then V with the reference's index nz (ny iX + iY) + iZ. */
static void write_cube(const sep_deck *d, const double *V) {
  char path[4096];
  snprintf(path, sizeof path, "%s/conf.par", g_dir);
  FILE *pc = fopen(path, "r");
  FILE *pf = out_file("localPot.cube", "w");
  if (!pc || !pf) {
    if (pc) fclose(pc);
    if (pf) fclose(pf);
    return;
  }
  long na = 0;
  if (fscanf(pc, "%ld", &na) != 1) na = 0;
  fprintf(pf, "CUBE FILE\n");
  fprintf(pf, "OUTER LOOP: X, MIDDLE LOOP: Y, INNER LOOP: Z\n");
  fprintf(pf, "%5li%12.6f%12.6f%12.6f\n", na, d->xmin, d->ymin, d->zmin);
  fprintf(pf, "%5li%12.6f%12.6f%12.6f\n", d->nx, d->dx, 0.0, 0.0);
  fprintf(pf, "%5li%12.6f%12.6f%12.6f\n", d->ny, 0.0, d->dy, 0.0);
  fprintf(pf, "%5li%12.6f%12.6f%12.6f\n", d->nz, 0.0, 0.0, d->dz);
  static const struct { const char *s; int z; } Z[] = {
      {"Cd", 48}, {"S", 16}, {"Se", 34}, {"Zn", 30}, {"Te", 52}, {"C", 6}, {"Si", 14},
      {"Cs", 55}, {"Pb", 82}, {"I", 53}, {"In", 49}, {"As", 33}, {"Ga", 31}, {"P", 15}};
  char line[80], sym[10];
  double x, y, z;
  if (!fgets(line, sizeof line, pc)) line[0] = 0;
  while (fgets(line, sizeof line, pc)) {
    if (sscanf(line, "%2s %lf %lf %lf", sym, &x, &y, &z) != 4) continue;
    int at = 1;
    for (size_t k = 0; k < sizeof Z / sizeof Z[0]; k++)
      if (!strcmp(sym, Z[k].s)) at = Z[k].z;
    fprintf(pf, "%5i%12.6f%12.6f%12.6f%12.6f\n", at, 0.0, x, y, z);
  }
  for (long ix = 0; ix < d->nx; ix++)
    for (long iy = 0; iy < d->ny; iy++) {
      const long ixy = d->nz * (d->ny * ix + iy);
      for (long iz = 0; iz < d->nz; iz++) {
        fprintf(pf, "%g ", V[ixy + iz]);
        if (iz % 6 == 5) fprintf(pf, "\n");
      }
      fprintf(pf, "\n");
    }
  fclose(pc);
  fclose(pf);
}

/* hmat.dat (filter/Hmat.c:28-80): the progress lines of the reference's
 * Rayleigh-Ritz (its rows, its eigensolve, its rotation of every 1000th
 * grid point). */
static void write_hmat_log(long mstot, long ngrid) {
  FILE *pg = out_file("hmat.dat", "w");
  if (!pg) return;
  fprintf(pg, "calculate the H matrix\n");
  for (long i = 0; i < mstot; i++) fprintf(pg, "finshed row %ld\n", i);
  fprintf(pg, "diagonalize the new Hamiltonian H\n");
  fprintf(pg, "copy the new function into psitot\n");
  for (long g = 0; g < ngrid; g += 1000) fprintf(pg, "finished grid point %ld\n", g);
  fprintf(pg, "free memory\n");
  fclose(pg);
}

int main(int argc, char **argv) {
  long seed_opt = 0;
  int have_seed_opt = 0, prec = SEP_PREC_FP64, device = 0, psi_filt = 0, resume = 0, lanes = 0;
  int refine_passes = 10;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--dir") && i + 1 < argc)
      g_dir = argv[++i];
    else if (!strcmp(argv[i], "--seed") && i + 1 < argc) {
      seed_opt = atol(argv[++i]);
      have_seed_opt = 1;
    } else if (!strcmp(argv[i], "--precision") && i + 1 < argc) {
      const char *p = argv[++i];
      if (!strcmp(p, "fp64"))
        prec = SEP_PREC_FP64;
      else if (!strcmp(p, "fp32"))
        prec = SEP_PREC_FP32;
      else
        die("--precision must be fp64 or fp32");
    } else if (!strcmp(argv[i], "--device") && i + 1 < argc)
      device = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--lanes") && i + 1 < argc) {
      char *end = NULL;
      const long v = strtol(argv[++i], &end, 10);
      if (!end || *end || v < 0 || v > SEP_MAX_LANES) {
        char msg[64];
        snprintf(msg, sizeof msg, "--lanes must be an integer in 0..%d", SEP_MAX_LANES);
        die(msg);
      }
      lanes = (int)v;
    } else if (!strcmp(argv[i], "--refine-passes") && i + 1 < argc) {
      char *end = NULL;
      const long v = strtol(argv[++i], &end, 10);
      if (!end || *end || v < 0 || v > 1000) die("--refine-passes must be an integer in 0..1000");
      refine_passes = (int)v;
    } else if (!strcmp(argv[i], "--psi-filt"))
      psi_filt = 1;
    else if (!strcmp(argv[i], "--resume-filt"))
      resume = 1;
    else {
      char msg[300];
      snprintf(msg, sizeof msg, "unknown option or missing value: %.250s", argv[i]);
      die(msg);
    }
  }
  stamp();
  separation();
  char err[512];
  sep_deck d;
  if (sep_deck_read(g_dir, &d, err, sizeof err) != SEP_OK) die(err);
  const long N = d.nx * d.ny * d.nz, ms = d.ms, ns = d.ns, nc = d.nc;
  /* filter/init.c:40-47 */
  printf("Number of atoms in the system, natom = %ld\n", d.natom);
  printf("Number of grid points used: nx = %ld  ny = %ld  nz = %ld\n", d.nx, d.ny, d.nz);
  printf("Number of states per filter, ms = %ld\nNumber of filter cycles, ns = %ld\n", ms, ns);
  printf("Length of Newton interpolation used, nc = %ld\n", nc);
  printf("Target energy for HOMO state = % .4f\n", d.Elmin);
  printf("Target energy for LUMO state = % .4f\n", d.Elmax);
  printf("Kinetic energy maximum = % .2f\n", 10.0);
  printf("Number of openMP threads used = %ld\n", d.nthreads);
  printf("Precision: %s\n", prec == SEP_PREC_FP64 ? "fp64" : "fp32 filter/ortho + fp64 Rayleigh-Ritz and refinement");
  if (prec == SEP_PREC_FP32)
    printf("FP64 correction passes on the pairs in [Elmin, Elmax]: %d\n", refine_passes);
  if (d.has_expot)
    printf("expot.par was detected: e = % .6f x0 = % .6f kT = % .6f\n", d.expot[0], d.expot[1],
           d.expot[2]);
  else
    printf("No expot.par file detected in cwd - setting external potential to 0!\n");
  if (d.has_strain) printf("Reading strain scale from strain.par...\n");
  printf("Box (quadrant) dimensions: xd = %.2f yd = %.2f zd = %.2f\n", d.xd, d.yd, d.zd);
  printf("Grid point spacing: dx = %.4f dy = %.4f dz = %.4f dv = %.6f dr = %.6f\n", d.dx, d.dy,
         d.dz, d.dv, sqrt(d.dx * d.dx + d.dy * d.dy + d.dz * d.dz));
  fflush(stdout);
  {
    const long dims[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; a++)
      if (!sep_fft_length_supported(dims[a])) {
        snprintf(err, sizeof err, "grid length %ld has no HIP FFT plan", dims[a]);
        die(err);
      }
  }

  sep_grid_desc gdesc = {d.nx, d.ny, d.nz, d.dx, d.dy, d.dz, d.xmin, d.ymin, d.zmin, 10.0, device};
  sep_ctx *ctx = NULL;
  int rc = sep_open(&gdesc, &ctx);
  if (rc != SEP_OK) {
    snprintf(err, sizeof err, "sep_open failed (%d)", rc);
    die(err);
  }
#define CHECK(call)                                                          \
  do {                                                                       \
    int rc_ = (call);                                                        \
    if (rc_ != SEP_OK) {                                                     \
      snprintf(err, sizeof err, "%s: %s (%d)", #call, sep_last_error(ctx), rc_); \
      die(err);                                                              \
    }                                                                        \
  } while (0)

  CHECK(sep_set_filter_lanes(ctx, lanes));
  /* V(r) on the device (filter/init.c:157-199) */
  int *ti = (int *)xmalloc(sizeof(int) * d.natom);
  for (long i = 0; i < d.natom; i++) ti[i] = d.type[i];
  double Vmin, Vmax;
  double *Vhost = (double *)xmalloc(sizeof(double) * N);
  if (!Vhost) die("potential buffer");
  CHECK(sep_build_potential(ctx, d.natom, d.xyz, ti, d.has_strain ? d.scale : NULL, SEP_NTYPE,
                            d.npot, d.vr, d.pot, SEP_NPOT, d.dr, d.expot, Vhost, &Vmin, &Vmax));
  free(ti);
  write_conf(&d);
  write_pots(&d);
  write_ps(&d);
  write_cube(&d, Vhost);
  free(Vhost);
  printf("Potential energy statistics: dV = %.5f Vmin = % .5f Vmax = %.5f\n", Vmax - Vmin, Vmin,
         Vmax);
  const double PIE = 3.14159265358979323846;
  printf("dT = %g\n", 0.5 * PIE * PIE / (d.dx * d.dx) + 0.5 * PIE * PIE / (d.dy * d.dy) +
                          0.5 * PIE * PIE / (d.dz * d.dz));
  /* energy targets, filter/init.c:209-214 */
  double *el = (double *)xmalloc(sizeof(double) * ms);
  {
    double del = ms > 1 ? (d.Elmax - d.Elmin) / (double)(ms - 1) : 0.0;
    FILE *pf = out_file("Egrid.dat", "w");
    for (long j = 0; j < ms; j++) {
      el[j] = d.Elmin + (double)j * del;
      if (pf) fprintf(pf, "%g\n", el[j]);
    }
    if (pf) fclose(pf);
  }
  /* spectral range (filter/main.c:63-66) */
  double t0 = wall(), c0 = (double)clock();
  double dE, Emax;
  long iters;
  {
    double *elog = (double *)xmalloc(sizeof(double) * 3 * 500);
    CHECK(sep_energy_range_log(ctx, Vmin, &dE, &Emax, &iters, elog, 500));
    FILE *pf = out_file("ene-init.dat", "w");  /* filter/energy.c:65-74 */
    if (pf) {
      for (long i = 0; i < iters && i < 500; i++)
        fprintf(pf, "%ld %.16g %.16g %.16g\n", i, elog[3 * i], elog[3 * i + 1], elog[3 * i + 2]);
      fclose(pf);
    }
    free(elog);
  }
  printf("dE = %g\n", dE);
  printf("done calculate energy range, CPU time (sec) %g, wall run time (sec) %g\n",
         ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
  /* Newton interpolation (main.c:70-74) */
  const double dt = ((double)nc / (2.5 * dE)) * ((double)nc / (2.5 * dE));
  printf("nc = %ld dt = %g dE = %g\n", nc, dt, dE);
  fflush(stdout);
  double *zn = (double *)xmalloc(sizeof(double) * nc);
  double *an = (double *)xmalloc(sizeof(double) * nc * ms);
  sep_leja_points(nc, zn);
  if (sep_newton_coeffs(nc, ms, zn, el, dE, Vmin, dt, an) != SEP_OK) die("coefficient failed");
  FILE *pc = out_file("coeff.dat", "w");
  if (pc) fclose(pc);
  write_func(an, zn, nc, dE, Vmin, dt, el[0]);
  const long mstot = ms * ns;
  double *psi0 = NULL, *ene = NULL;
  if (resume) { /* the filtered states of an earlier run, ns files of ms states */
    double *psitot = (double *)xmalloc(sizeof(double) * mstot * N);
    if (!psitot) die("psi-filt buffer");
    for (long jns = 0; jns < ns; jns++) {
      char name[64];
      snprintf(name, sizeof name, "psi-filt-0-%ld.dat", jns);
      FILE *pf = out_file(name, "rb");
      if (!pf) die("--resume-filt: a psi-filt-0-<jns>.dat file is missing");
      size_t got = fread(psitot + jns * ms * N, sizeof(double), (size_t)(ms * N), pf);
      int extra = fgetc(pf) != EOF;
      fclose(pf);
      if (got != (size_t)(ms * N) || extra) die("--resume-filt: psi-filt file of the wrong size");
    }
    CHECK(sep_states_upload(ctx, psitot, mstot));
    free(psitot);
    printf("read %ld filtered states from psi-filt-0-*.dat\n", mstot);
    fflush(stdout);
  } else {
    /* random states (main.c:79-89) */
    long idum;
    long s = have_seed_opt ? seed_opt : (d.has_seed ? d.seed : 0);
    if (s == 0) {
      struct timeval tv;
      gettimeofday(&tv, NULL);
      srandom((unsigned)(tv.tv_sec ^ tv.tv_usec));
      idum = -random();
    } else {
      idum = s < 0 ? s : -s;
    }
    printf("seed = %ld\n", idum);
    fflush(stdout);
    psi0 = (double *)xmalloc(sizeof(double) * ns * N);
    sep_rng rng;
    sep_rng_reset(&rng);
    for (long j = 0; j < ns; j++) sep_init_psi(&rng, &idum, N, d.dv, psi0 + j * N);
    /* filter (main.c:91-100) */
    ene = (double *)xmalloc(sizeof(double) * mstot);
    double *psif = psi_filt ? (double *)xmalloc(sizeof(double) * mstot * N) : NULL;
    t0 = wall();
    c0 = (double)clock();
    CHECK(sep_filter(ctx, psi0, ns, an, zn, nc, ms, Vmin, dE, prec, psif, ene));
    CHECK(sep_trim(ctx)); /* one filter per run: its ring is not needed again */
    {
      /* prop0.dat (filter/filter.c:69-81): every 100th Newton step; the file
       * is rewritten per random state, the last one remains */
      FILE *pf = out_file("prop0.dat", "w");
      if (pf) {
        for (long j = 100; j < nc; j += 100) fprintf(pf, "%ld %ld\n", j, ns - 1);
        fclose(pf);
      }
    }
    printf("done calculating filter, CPU time (sec) %g, wall run time (sec) %g\n",
           ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
    for (long jns = 0; jns < ns; jns++) { /* filter.c:38-48 */
      char name[64];
      snprintf(name, sizeof name, "eval-filt-0-%ld.dat", jns);
      FILE *pf = out_file(name, "w");
      if (pf) {
        for (long j = 0; j < ms; j++) fprintf(pf, "%ld %.16g %.16g\n", j, ene[jns * ms + j], el[j]);
        fclose(pf);
      }
      if (psif) {
        snprintf(name, sizeof name, "psi-filt-0-%ld.dat", jns);
        pf = out_file(name, "wb");
        if (pf) {
          fwrite(psif + jns * ms * N, sizeof(double), ms * N, pf);
          fclose(pf);
        }
      }
    }
    free(psif);
  }
  /* ortho (main.c:109-114) */
  t0 = wall();
  c0 = (double)clock();
  long kept = 0;
  double *S = (double *)xmalloc(sizeof(double) * mstot);
  if (prec == SEP_PREC_FP32)
    CHECK(sep_ortho_fp32(ctx, mstot, 1.0e-10, S, &kept));
  else
    CHECK(sep_ortho(ctx, mstot, 1.0e-10, S, &kept));
  printf("cutoff is %ld\n", kept);
  printf("mstot ortho = %ld\n", kept);
  printf("done calculating ortho, CPU time (sec) %g, wall run time (sec) %g\n",
         ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
  /* Rayleigh-Ritz (main.c:120-126) */
  t0 = wall();
  c0 = (double)clock();
  double *eval = (double *)xmalloc(sizeof(double) * kept);
  double *sige = (double *)xmalloc(sizeof(double) * kept);
  CHECK(sep_rayleigh_ritz(ctx, kept, eval));
  write_hmat_log(kept, N);
  if (prec == SEP_PREC_FP32) {
    /* the Ritz pairs' variances, then the window's pairs refined in place */
    CHECK(sep_sigma(ctx, kept, sige));
    if (refine_passes > 0) {
      long kw = 0;
      CHECK(sep_refine_window(ctx, kept, eval, sige, d.Elmin, d.Elmax, refine_passes, &kw));
      printf("refined %ld Ritz pairs in [% .4f, % .4f] with %d FP64 correction passes\n", kw,
             d.Elmin, d.Elmax, refine_passes);
    }
  }
  printf("done calculating Hmat, CPU time (sec) %g, wall run time (sec) %g\n",
         ((double)clock() - c0) / CLOCKS_PER_SEC, wall() - t0);
  fflush(stdout);
  /* psi.dat (main.c:129-131) */
  {
    double *psi = (double *)xmalloc(sizeof(double) * kept * N);
    if (!psi) die("psi.dat buffer");
    CHECK(sep_states_download(ctx, psi, kept));
    FILE *pf = out_file("psi.dat", "wb");
    if (!pf) die("cannot write psi.dat");
    fwrite(psi, sizeof(double), kept * N, pf);
    fclose(pf);
    free(psi);
  }
  /* sigma and eval.dat (main.c:135-142) */
  if (prec != SEP_PREC_FP32) CHECK(sep_sigma(ctx, kept, sige));  /* fp32: computed above */
  {
    FILE *pf = out_file("eval.dat", "w");
    if (!pf) die("cannot write eval.dat");
    for (long j = 0; j < kept; j++) fprintf(pf, "%ld %.16g %g\n", j, eval[j], sige[j]);
    fclose(pf);
  }
  sep_close(ctx);
  sep_deck_free(&d);
  free(el); free(zn); free(an); free(psi0); free(ene); free(S); free(eval); free(sige);
  separation();
  stamp();
  return 0;
}
