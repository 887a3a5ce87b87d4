/* sep_host.h -- host-side C parts of libsep: small dense FP64 linear algebra
 * (the reference links Intel MKL for these, filter/ortho.c:32, Hmat.c:55) and
 * helpers shared by the C-ABI layer.  Internal to the library. */
#ifndef SEP_HOST_H
#define SEP_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* One-sided Jacobi SVD of a square R (n x n row-major): S descending,
 * U (n x n row-major) columns = left singular vectors (zero column for a
 * zero singular value), V (optional) columns = right singular vectors. */
void sep_la_svd_jacobi(long n, const double *R, double *U, double *S, double *V);
/* Upper Cholesky of G = hi + lo (double-double, n x n row-major) + shift*I,
 * carried out in double-double; R (rounded to double) gets the factor.
 * Returns 0 or k+1 for a non-positive pivot k. */
int sep_la_chol_dd(long n, const double *hi, const double *lo, double shift, double *R);

#ifdef __cplusplus
}
#endif
#endif
