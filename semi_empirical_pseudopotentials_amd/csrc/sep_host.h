/* sep_host.h -- host-side C parts of libsep: the double-double Cholesky that
 * decides the ortho cutoff (the reference's SVD is Intel MKL's,
 * filter/ortho.c:32) and helpers shared by the C-ABI layer.  Internal to the
 * library. */
#ifndef SEP_HOST_H
#define SEP_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* Upper Cholesky of G = hi + lo (double-double, n x n row-major) + shift*I,
 * carried out in double-double; R (rounded to double) gets the factor.
 * Returns 0 or k+1 for a non-positive pivot k. */
int sep_la_chol_dd(long n, const double *hi, const double *lo, double shift, double *R);

#ifdef __cplusplus
}
#endif
#endif
