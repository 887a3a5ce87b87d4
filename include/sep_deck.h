/*
 * sep_deck.h -- input-deck reader of libsep (host C).  Reads the files the
 * reference `filter` program reads from its working directory and derives
 * the box and grid exactly as it does, so the filter driver is a drop-in.
 *
 *   input.par   nx ny nz ms ns nc Elmin Elmax nthreads [ignored tokens...]
 *               (filter/init.c:12-50; filter_input.par is accepted only when
 *               input.par is absent)
 *   conf.par    natom, then natom lines "Sym x y z" in bohr (filter/read.c:11-39),
 *               recentred on the unweighted mean (read.c:23-30)
 *   pot<Sym>.par  "r v" rows, at most 8192 (filter/read.c:43-122): shifted
 *               by the last value; missing P1..P4 tables synthesised as
 *               a exp(-r^2/b) on potCd.par's r grid (read.c:96-101)
 *   expot.par   optional "e x0 kT" (filter/init.c:66-81)
 *   strain.par  optional per-atom "i sym a4 a5 vol volRef scale"
 *               (filterStrainCube/init.c:478-486)
 *   seed.par    optional extension: the ran_nrc seed idum; absent or 0 means
 *               seed from the clock as filter/main.c:79 does
 */
#ifndef SEP_DECK_H
#define SEP_DECK_H

#ifdef __cplusplus
extern "C" {
#endif

#define SEP_NTYPE 17   /* atom types, filter/read.c:126-152 */
#define SEP_NPOT 8192  /* table rows, filter/init.c:30 */

typedef struct {
  /* input.par */
  long nx, ny, nz, ms, ns, nc, nthreads;
  double Elmin, Elmax;
  /* conf.par (recentred), per-atom type index and scale */
  long natom;
  double *xyz;   /* natom x 3 */
  int *type;     /* natom */
  char (*sym)[4];
  double *scale; /* 1, or column 7 of strain.par */
  int has_strain;
  /* expot.par */
  double expot[3]; /* e, x0, kT; {0,0,1} when absent */
  int has_expot;
  /* pot tables: SEP_NTYPE rows of SEP_NPOT */
  long npot[SEP_NTYPE];
  double dr[SEP_NTYPE];
  double *vr, *pot;
  /* box and grid (filter/init.c:90-122) */
  double xd, yd, zd;
  double xmin, ymin, zmin, dx, dy, dz, dv;
  /* seed.par */
  long seed;
  int has_seed;
} sep_deck;

/* Reads the deck in `dir`.  Returns SEP_OK or SEP_ERR_IO / SEP_ERR_ARG with a
 * message in err; a refused deck holds no allocation (sep_deck_free on it
 * is harmless).  A pot table longer than SEP_NPOT rows is refused (the
 * reference, filter/read.c:90-92, would write past its table slot). */
int sep_deck_read(const char *dir, sep_deck *d, char *err, int errlen);
void sep_deck_free(sep_deck *d);
/* Symbol of atom type t (filter/read.c:155-176). */
const char *sep_type_symbol(int t);

#ifdef __cplusplus
}
#endif
#endif
