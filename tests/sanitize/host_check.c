/* host_check.c -- drives the host C of libsep (csrc/sep_host.c) under
 * AddressSanitizer / UndefinedBehaviorSanitizer (built by
 * tests/test_host_sanitize.py with gcc only; no HIP): the deck reader on every
 * directory given on the command line (good and malformed decks), the
 * random-state generator, the Leja points and Newton coefficients, and the
 * double-double Cholesky on a positive definite and an indefinite matrix.
 * Prints one line per deck: "<dir> <status> <natom> <nx>" and exits 0 unless
 * a library result is inconsistent (the sanitizers abort on their own). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sep.h"
#include "sep_deck.h"
#include "sep_host.h"

int main(int argc, char **argv) {
  int bad = 0;
  for (int i = 1; i < argc; i++) {
    sep_deck d;
    memset(&d, 0, sizeof d);
    char err[256] = "";
    const int rc = sep_deck_read(argv[i], &d, err, (int)sizeof err);
    printf("%s %d %ld %ld %s\n", argv[i], rc, rc == SEP_OK ? d.natom : -1L,
           rc == SEP_OK ? d.nx : -1L, rc == SEP_OK ? "" : err);
    if (rc == SEP_OK) {
      /* touch every array the reader filled */
      double s = 0.0;
      for (long a = 0; a < d.natom; a++) s += d.xyz[3 * a] + d.scale[a] + d.type[a];
      for (int t = 0; t < SEP_NTYPE; t++)
        for (long k = 0; k < d.npot[t]; k++) s += d.vr[t * SEP_NPOT + k] + d.pot[t * SEP_NPOT + k];
      if (!isfinite(s)) bad = 1;
      sep_deck_free(&d);
    } else if (d.xyz || d.type || d.sym || d.scale || d.vr || d.pot) {
      bad = 1; /* a refused deck must hold nothing (LeakSanitizer checks the rest) */
    }
  }
  /* random states */
  sep_rng st;
  sep_rng_reset(&st);
  long idum = -1001;
  double *psi = malloc(sizeof(double) * 4096);
  sep_init_psi(&st, &idum, 4096, 0.125, psi);
  double n2 = 0.0;
  for (int g = 0; g < 4096; g++) n2 += psi[g] * psi[g] * 0.125;
  if (fabs(n2 - 1.0) > 1e-12) bad = 1;
  free(psi);
  /* Leja points and Newton coefficients */
  const long ncs[3] = {1, 16, 1024};
  for (int k = 0; k < 3; k++) {
    const long nc = ncs[k];
    double *zn = malloc(sizeof(double) * nc), *an = malloc(sizeof(double) * 2 * nc);
    const double el[2] = {-0.2, 0.1};
    sep_leja_points(nc, zn);
    if (sep_newton_coeffs(nc, 2, zn, el, 12.0, -1.0, (nc / 30.0) * (nc / 30.0), an) != SEP_OK) bad = 1;
    for (long j = 0; j < 2 * nc; j++)
      if (!isfinite(an[j])) bad = 1;
    free(zn);
    free(an);
  }
  /* double-double Cholesky */
  const double spd[9] = {4, 1, 0, 1, 3, 1, 0, 1, 2}, ind[9] = {1, 2, 0, 2, 1, 0, 0, 0, 1};
  const double lo[9] = {0};
  double R[9];
  if (sep_la_chol_dd(3, spd, lo, 0.0, R) != 0) bad = 1;
  if (sep_la_chol_dd(3, ind, lo, 0.0, R) == 0) bad = 1;
  printf("host checks %s\n", bad ? "FAILED" : "ok");
  return bad;
}
