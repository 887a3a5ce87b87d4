"""MI355X-native filter diagonalization for the semi-empirical pseudopotential
method: libsep (HIP kernels behind a C ABI, include/sep.h), the `filter`
driver (bin/filter) and this thin Python mirror of the C ABI."""
from ._lib import (SEP_BATCH_GRAPH, SEP_BATCH_INTERLEAVE, SEP_HPSI_LINES, SEP_HPSI_NFAMILY,
                   SEP_HPSI_PLANE, SEP_HPSI_PLANE_CHUNKED, SEP_HPSI_PLANE_SWIZZLED,
                   SEP_INTERP_CUBIC, SEP_INTERP_LINEAR,
                   SEP_MAX_LANES, SEP_PREC_FP32, SEP_PREC_FP64, Context, SepError,
                   fft_length_supported, filter_batch, init_psi, leja_points, lib, newton_coeffs,
                   ran_nrc_stream)

__all__ = ["Context", "SepError", "SEP_PREC_FP64", "SEP_PREC_FP32", "SEP_BATCH_INTERLEAVE",
           "SEP_BATCH_GRAPH", "SEP_MAX_LANES", "SEP_HPSI_PLANE", "SEP_HPSI_PLANE_CHUNKED",
           "SEP_HPSI_LINES", "SEP_HPSI_PLANE_SWIZZLED", "SEP_HPSI_NFAMILY", "SEP_INTERP_LINEAR", "SEP_INTERP_CUBIC", "lib",
           "init_psi", "leja_points", "newton_coeffs", "ran_nrc_stream", "fft_length_supported",
           "filter_batch"]
