/*
 * couplings_main.c -- derivative couplings between quasiparticle states on
 * libsep, in place of the reference's nonAdiabaticCoupling program
 * (nonAdiabaticCoupling/main.c:25-147, nonadiabatic.c:9-301): for every
 * non-passivation atom a and every pair of selected states r, s
 *
 *   adiabatic:  V_rs = dv sum_g psi_r psi_s dV/dR_a / (E_s - E_r), r != s,
 *               V_sr = -V_rs, V_rr = 0                  (nonadiabatic.c:9-93)
 *   diabatic:   V_rs = V_sr = dv sum_g psi_r psi_s dV/dR_a, with r == s
 *                                                        (nonadiabatic.c:143-235)
 *
 * all atoms and pairs of a set in sep_hf_adjoint calls on the device.
 *
 * Inputs, from the working directory (or --dir): the filter deck the states
 * were computed on (input.par, conf.par, pot*.par, optional expot.par and
 * strain.par, read by sep_deck_read, so grid, atom positions and V(r) are the
 * filter's own), and the state set eval.par + psi.par (combine.py's output;
 * --eval / --psi name other files, e.g. the filter's eval.dat / psi.dat).
 * State selection as nonAdiabaticCoupling/read.c:93-113,347-375: converged
 * states (sigma < --sigma-cutoff, default 0.01) below --fermi-energy
 * (default -0.175 Ha) are holes, above it electrons; the --max-holes nearest
 * the Fermi energy below it and the --max-elecs nearest above it (default
 * 1000 each), each set in ascending energy.
 *
 * Outputs: Vij.dat (holes) and Vab.dat (electrons), or Vij-diabatic.dat and
 * Vab-diabatic.dat, with the reference's line formats
 * (nonadiabatic.c:96-123, 240-292), and holeElecEigenstates.dat listing the
 * selected states (write.c:29-33).  Degenerate pairs give infinite adiabatic
 * couplings, as in the reference.  Passivation atoms (P1..P4,
 * read.c:251-258) get no lines.
 *
 * Differences: dV/dR is the exact derivative of the filter's piecewise
 * linear V(r) (filter/interpolate.c:11-23), where the reference program
 * differentiates the first 1024 rows of each pot file by central differences
 * and interpolates that (pseudopotential.c:6-118); its own input.par keys
 * (calcType, gridSize, ...) are replaced by the filter deck and the options
 * below.  Failures write `error` and exit 1, as the filter driver does.
 *
 * Options: --dir D, --eval F, --psi F, --state-type adiabatic|diabatic,
 *          --sigma-cutoff S, --fermi-energy E, --max-holes N, --max-elecs N,
 *          --device N
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sep.h"
#include "../../include/sep_deck.h"

static const char *g_dir = ".";

static FILE *dir_file(const char *name, const char *mode) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s", g_dir, name);
  return fopen(path, mode);
}

static void die(const char *msg) {
  FILE *pf = dir_file("error", "w");
  if (pf) {
    fprintf(pf, "%s\n", msg);
    fclose(pf);
  }
  fprintf(stderr, "couplings: %s\n", msg);
  exit(1);
}

// host buffers: a failed allocation ends the program with a message
static void *xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) die("out of host memory");
  return p;
}

static double num_arg(const char *opt, const char *v) {
  char *end = NULL;
  const double x = strtod(v, &end);
  if (end == v || *end) {
    char m[256];
    snprintf(m, sizeof m, "%s needs a number, got '%s'", opt, v);
    die(m);
  }
  return x;
}

static int is_passivation(const char *sym) {
  return !strcmp(sym, "P1") || !strcmp(sym, "P2") || !strcmp(sym, "P3") || !strcmp(sym, "P4");
}

typedef struct {
  long index; /* line of the state file */
  double e, sigma;
} qp_state;

/* Couplings of one state set (n states resident from row `first`): writes
 * `name` with one line per (atom, r, s). */
static void couple_set(sep_ctx *ctx, const sep_deck *d, const qp_state *qs, long n, long first,
                       int diabatic, const char *name) {
  char err[512];
  if (n < 1) return;
  /* pairs r <= s (the diagonal only for diabatic) */
  long np = 0;
  long *pairs = (long *)xmalloc(sizeof(long) * 2 * n * (n + 1) / 2);
  long *pidx = (long *)xmalloc(sizeof(long) * n * n); /* (r, s) -> pair, or -1 */
  for (long r = 0; r < n; r++)
    for (long s = 0; s < n; s++) pidx[r * n + s] = -1;
  for (long r = 0; r < n; r++)
    for (long s = r + (diabatic ? 0 : 1); s < n; s++) {
      pairs[2 * np] = first + r;
      pairs[2 * np + 1] = first + s;
      pidx[r * n + s] = np++;
    }
  const long na = d->natom;
  double *dR = NULL;
  if (np > 0) {
    dR = (double *)xmalloc(sizeof(double) * np * na * 3);
    if (!pairs || !pidx || !dR) die("couplings: out of memory");
    int rc = sep_hf_adjoint(ctx, np, pairs, NULL, na, d->xyz, d->type,
                            d->has_strain ? d->scale : NULL, SEP_NTYPE, d->npot, d->vr, d->pot,
                            SEP_NPOT, d->dr, dR, NULL, NULL);
    if (rc != SEP_OK) {
      snprintf(err, sizeof err, "sep_hf_adjoint: %s (%d)", sep_last_error(ctx), rc);
      die(err);
    }
  }
  FILE *pf = dir_file(name, "w");
  if (!pf) die("cannot write the coupling file");
  for (long a = 0; a < na; a++) {
    if (is_passivation(d->sym[a])) continue;
    const double *R = d->xyz + 3 * a;
    for (long r = 0; r < n; r++)
      for (long s = 0; s < n; s++) {
        double v[3] = {0.0, 0.0, 0.0};
        const long lo = r < s ? r : s, hi = r < s ? s : r;
        const long p = pidx[lo * n + hi];
        if (p >= 0) {
          double f = 1.0;
          if (!diabatic) f = (r < s ? 1.0 : -1.0) / (qs[hi].e - qs[lo].e); /* V_sr = -V_rs */
          for (int c = 0; c < 3; c++) v[c] = f * dR[(p * na + a) * 3 + c];
        }
        const double mag = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (diabatic)
          fprintf(pf, "%ld %s %lg %lg %lg %lg %ld % .8f %ld % .8f %lg %lg %lg %lg\n", a, d->sym[a],
                  R[0], R[1], R[2], sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]), r, qs[r].e, s,
                  qs[s].e, v[0], v[1], v[2], mag);
        else
          fprintf(pf, "%ld %s %lg %lg %lg %ld % .8f %ld % .8f %lg %lg %lg %lg\n", a, d->sym[a],
                  R[0], R[1], R[2], r, qs[r].e, s, qs[s].e, v[0], v[1], v[2], mag);
      }
  }
  fclose(pf);
  free(dR);
  free(pairs);
  free(pidx);
}

int main(int argc, char **argv) {
  const char *eval_name = "eval.par", *psi_name = "psi.par";
  int diabatic = 0, device = 0;
  double sigma_cut = 0.01, fermi = -0.175;
  long max_holes = 1000, max_elecs = 1000;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--dir") && i + 1 < argc)
      g_dir = argv[++i];
    else if (!strcmp(argv[i], "--eval") && i + 1 < argc)
      eval_name = argv[++i];
    else if (!strcmp(argv[i], "--psi") && i + 1 < argc)
      psi_name = argv[++i];
    else if (!strcmp(argv[i], "--state-type") && i + 1 < argc) {
      const char *t = argv[++i];
      if (!strcmp(t, "diabatic"))
        diabatic = 1;
      else if (strcmp(t, "adiabatic"))
        die("--state-type must be adiabatic or diabatic");
    } else if (!strcmp(argv[i], "--sigma-cutoff") && i + 1 < argc)
      sigma_cut = num_arg("--sigma-cutoff", argv[++i]);
    else if (!strcmp(argv[i], "--fermi-energy") && i + 1 < argc)
      fermi = num_arg("--fermi-energy", argv[++i]);
    else if (!strcmp(argv[i], "--max-holes") && i + 1 < argc)
      max_holes = (long)num_arg("--max-holes", argv[++i]);
    else if (!strcmp(argv[i], "--max-elecs") && i + 1 < argc)
      max_elecs = (long)num_arg("--max-elecs", argv[++i]);
    else if (!strcmp(argv[i], "--device") && i + 1 < argc)
      device = atoi(argv[++i]);
    else {
      char m[256];
      snprintf(m, sizeof m, "unknown option %s", argv[i]);
      die(m);
    }
  }
  char err[512];
  sep_deck d;
  if (sep_deck_read(g_dir, &d, err, sizeof err) != SEP_OK) die(err);
  const long N = d.nx * d.ny * d.nz;
  /* the state list */
  FILE *pf = dir_file(eval_name, "r");
  if (!pf) die("cannot open the eigenvalue file (eval.par)");
  long cap = 1024, nst = 0;
  qp_state *all = (qp_state *)xmalloc(sizeof(qp_state) * cap);
  long idx;
  double e, sg;
  while (fscanf(pf, "%ld %lg %lg", &idx, &e, &sg) == 3) {
    if (nst == cap) {
      cap *= 2;
      qp_state *grown = (qp_state *)realloc(all, sizeof(qp_state) * cap);
      if (!grown) die("out of host memory");
      all = grown;
    }
    all[nst].index = nst;
    all[nst].e = e;
    all[nst].sigma = sg;
    nst++;
  }
  fclose(pf);
  if (nst == 0) die("the eigenvalue file lists no states");
  for (long i = 1; i < nst; i++)
    if (all[i].e < all[i - 1].e) die("the eigenvalue file is not in ascending energy");
  /* holes: converged, below the Fermi energy, the max_holes highest;
   * electrons: converged, above it, the max_elecs lowest */
  qp_state *holes = (qp_state *)xmalloc(sizeof(qp_state) * nst);
  qp_state *elecs = (qp_state *)xmalloc(sizeof(qp_state) * nst);
  long nh = 0, ne = 0;
  for (long i = nst - 1; i >= 0 && nh < max_holes; i--)
    if (all[i].sigma < sigma_cut && all[i].e < fermi) holes[nh++] = all[i];
  for (long i = 0; i < nst && ne < max_elecs; i++)
    if (all[i].sigma < sigma_cut && all[i].e > fermi) elecs[ne++] = all[i];
  for (long i = 0; i < nh / 2; i++) { /* ascending */
    qp_state t = holes[i];
    holes[i] = holes[nh - 1 - i];
    holes[nh - 1 - i] = t;
  }
  printf("states %ld: %ld holes, %ld electrons (sigma < %g, Fermi energy %g)\n", nst, nh, ne,
         sigma_cut, fermi);
  pf = dir_file("holeElecEigenstates.dat", "w"); /* write.c:29-33 */
  if (pf) {
    for (long i = 0; i < nh; i++) fprintf(pf, "%ld h %.10f %.10f\n", i, holes[i].e, holes[i].sigma);
    for (long i = 0; i < ne; i++) fprintf(pf, "%ld e %.10f %.10f\n", i, elecs[i].e, elecs[i].sigma);
    fclose(pf);
  }
  if (nh + ne == 0) die("no converged states on either side of the Fermi energy");
  /* the selected states from the state file: holes then electrons */
  double *psi = (double *)xmalloc(sizeof(double) * (nh + ne) * N);
  if (!psi) die("state buffer");
  pf = dir_file(psi_name, "rb");
  if (!pf) die("cannot open the state file (psi.par)");
  if (fseek(pf, 0, SEEK_END) || ftell(pf) != (long)(nst * N * sizeof(double)))
    die("the state file's size does not match the eigenvalue file and the grid");
  for (long k = 0; k < nh + ne; k++) {
    const long i = k < nh ? holes[k].index : elecs[k - nh].index;
    if (fseek(pf, i * N * (long)sizeof(double), SEEK_SET) ||
        fread(psi + k * N, sizeof(double), N, pf) != (size_t)N)
      die("short read of the state file");
  }
  fclose(pf);
  sep_grid_desc gdesc = {d.nx, d.ny, d.nz, d.dx, d.dy, d.dz, d.xmin, d.ymin, d.zmin, 10.0, device};
  sep_ctx *ctx = NULL;
  if (sep_open(&gdesc, &ctx) != SEP_OK) die("sep_open failed");
  int rc = sep_states_upload(ctx, psi, nh + ne);
  if (rc != SEP_OK) {
    snprintf(err, sizeof err, "sep_states_upload: %s (%d)", sep_last_error(ctx), rc);
    die(err);
  }
  free(psi);
  couple_set(ctx, &d, holes, nh, 0, diabatic, diabatic ? "Vij-diabatic.dat" : "Vij.dat");
  couple_set(ctx, &d, elecs, ne, nh, diabatic, diabatic ? "Vab-diabatic.dat" : "Vab.dat");
  sep_close(ctx);
  sep_deck_free(&d);
  free(all);
  free(holes);
  free(elecs);
  printf("done\n");
  return 0;
}
