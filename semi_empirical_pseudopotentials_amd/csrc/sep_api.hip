// sep_api.cpp -- the C ABI of libsep (include/sep.h): context, device memory
// and the host-side orchestration of the filter-diagonalization pipeline.
#include <hip/hip_runtime.h>
#include <rocsolver/rocsolver.h>

#include <mutex>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <vector>

#include "fft_core.h"
#include "sep_internal.h"

using namespace sep;

struct sep_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  GridDev gd{};
  double dx = 0, dy = 0, dz = 0, dv = 0, xmin = 0, ymin = 0, zmin = 0, ekinmax = 10.0;
  double *pot64 = nullptr;
  float *pot32 = nullptr;
  bool has_pot = false;
  double vmin = 0, vmax = 0;
  // resident state block (FP64, m x N) and a second block of the same shape
  double *states = nullptr, *states2 = nullptr;
  long cap1 = 0, cap2 = 0;
  // FFT intermediate and general scratch
  void *tb = nullptr;
  size_t tb_bytes = 0;
  double *scr = nullptr;
  size_t scr_len = 0;
  // radial-table interpolation of the V(r) builder and its adjoint
  int interp = SEP_INTERP_LINEAR;
  // split-K partial sums of the matrix-core Gram (gram_dev)
  double *gwork = nullptr;
  size_t gwork_bytes = 0;
  void *ring = nullptr;
  size_t ring_bytes = 0;
  // a filter pass is using the ring (outside one the ring holds nothing
  // live: the filtered states are in the state block)
  bool ring_busy = false;
  // live kernel timing of the filter loop (sep_set_timing): one sampled
  // Newton step per ring block (passes A, B, C) and that block's accumulation
  bool timing = false;
  hipEvent_t tev[6] = {};
  double tsum[4] = {0, 0, 0, 0};
  long tcount = 0, tacc_count = 0;
  // dense eigen/Cholesky solves of the subspace stage (rocSOLVER on st),
  // created on first use; dws holds their eigenvalues, off-diagonal and info
  rocblas_handle blas = nullptr;
  void *dws = nullptr;
  size_t dws_bytes = 0;
  // all-zero potential: pass C then returns the inverse transform alone
  // (the kinetic preconditioner of sep_refine)
  void *zpot = nullptr;
  size_t zpot_bytes = 0;
  // filter lanes: concurrent groups of random states on streams of their own
  // (sep_set_filter_lanes; 0 = automatic); lane 0 runs on st
  int lanes = 0;
  hipStream_t lst[SEP_MAX_LANES - 1] = {};
  hipEvent_t lev[SEP_MAX_LANES] = {};
  // spin-orbit lists of the spinor H (sep_so_setup) and the per-entry
  // contributions of one application
  SoLists so;
  double *so_contrib = nullptr;
  size_t so_contrib_bytes = 0;
  std::string err;
};

static int fail(sep_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, x)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail((c), SEP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)
#define RCHK(x)               \
  do {                        \
    int r_ = (x);             \
    if (r_ != SEP_OK) return r_; \
  } while (0)

// hipMalloc that, when the memory is not there, frees the filter's ring if
// no filter pass is using it (it is sized from the free memory, up to 60 %
// of it, and kept between filters) and tries once more.
static hipError_t dev_malloc(sep_ctx *c, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess || !c->ring || c->ring_busy) return e;
  (void)hipGetLastError();  // the failed allocation must not be reported by a later call
  hipStreamSynchronize(c->st);
  hipFree(c->ring);
  c->ring = nullptr;
  c->ring_bytes = 0;
  return hipMalloc(p, bytes);
}
// Marks the contexts' rings as in use for a scope (filter passes).
struct RingBusy {
  std::vector<sep_ctx *> cs;
  void add(sep_ctx *c) {
    c->ring_busy = true;
    cs.push_back(c);
  }
  ~RingBusy() {
    for (sep_ctx *c : cs) c->ring_busy = false;
  }
};

static int ensure(sep_ctx *c, void **p, size_t *cap, size_t bytes) {
  if (*cap >= bytes && *p) return SEP_OK;
  if (*p) {  // work queued on the old buffer finishes before it is freed
    hipStreamSynchronize(c->st);
    hipFree(*p);
  }
  *p = nullptr;
  *cap = 0;
  if (bytes == 0) return SEP_OK;
  hipError_t e = dev_malloc(c, p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, SEP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  *cap = bytes;
  return SEP_OK;
}

static int ensure_states(sep_ctx *c, long m, int which) {
  size_t bytes = sizeof(double) * (size_t)m * c->gd.N;
  size_t cap = sizeof(double) * (size_t)(which == 1 ? c->cap1 : c->cap2) * c->gd.N;
  void **p = (void **)(which == 1 ? &c->states : &c->states2);
  if (cap >= bytes && *p) return SEP_OK;
  RCHK(ensure(c, p, &cap, bytes));
  (which == 1 ? c->cap1 : c->cap2) = m;
  return SEP_OK;
}
// ensure_states that keeps the first `keep` rows of the buffer
static int grow_states(sep_ctx *c, long m, int which, long keep) {
  long &cap = which == 1 ? c->cap1 : c->cap2;
  double *&p = which == 1 ? c->states : c->states2;
  if (p && cap >= m) return SEP_OK;
  double *q = nullptr;
  HIPCHK(c, dev_malloc(c, (void **)&q, sizeof(double) * (size_t)m * c->gd.N));
  if (p && keep > 0)
    HIPCHK(c, hipMemcpyAsync(q, p, sizeof(double) * (size_t)(keep < cap ? keep : cap) * c->gd.N,
                             hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (p) hipFree(p);
  p = q;
  cap = m;
  return SEP_OK;
}
static int ensure_scr(sep_ctx *c, size_t len) {
  size_t cap = c->scr_len * sizeof(double);
  RCHK(ensure(c, (void **)&c->scr, &cap, len * sizeof(double)));
  if (len > c->scr_len) c->scr_len = len;
  return SEP_OK;
}

// Temporary device allocation freed on scope exit.
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) hipFree(p);
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};
static int dev_alloc(sep_ctx *c, DevBuf &b, size_t bytes) {
  HIPCHK(c, dev_malloc(c, &b.p, bytes > 0 ? bytes : 8));
  return SEP_OK;
}
static int dev_upload(sep_ctx *c, DevBuf &b, const void *src, size_t bytes) {
  RCHK(dev_alloc(c, b, bytes));
  HIPCHK(c, hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  return SEP_OK;
}

extern "C" {

const char *sep_last_error(const sep_ctx *c) { return c ? c->err.c_str() : "null context"; }

// ------------------------------------------- allocation-free H application
// sep_apply_h_dev_noalloc takes its FFT spectrum buffer from this file-scope
// device array instead of the context's hipMalloc'd one.  Each GPU holds its
// own copy from module load to process end; host code reaches it only
// through the address hipGetSymbolAddress returns with that device current
// (g_noalloc_ptr), never through the array's name.
static __device__ __align__(256) unsigned char scratch[SEP_NOALLOC_SCRATCH_BYTES];
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[SEP_MAX_DEVICES] = {};

// Once per device, with that device current (called from sep_open).
static int noalloc_init(int device) {
  if (device < 0 || device >= SEP_MAX_DEVICES) return SEP_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return SEP_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SEP_ERR_HIP;
  g_noalloc_ptr[device] = p;
  return SEP_OK;
}

int sep_fft_length_supported(long n) { return hop_supported((int)n) ? 1 : 0; }

int sep_open(const sep_grid_desc *g, sep_ctx **out) {
  if (!g || !out) return SEP_ERR_ARG;
  *out = nullptr;
  for (long n : {g->nx, g->ny, g->nz})
    if (!hop_supported((int)n)) return SEP_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= g->device) return SEP_ERR_HIP;
  sep_ctx *c = new sep_ctx();
  c->device = g->device;
  // the second filter lane's stream is created with the context's own, so
  // that the runtime places the two on different hardware queues (streams
  // created later share queues with whatever else the process opened)
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
      hipStreamCreate(&c->lst[0]) != hipSuccess) {
    if (c->st) hipStreamDestroy(c->st);
    delete c;
    return SEP_ERR_HIP;
  }
  if (noalloc_init(c->device) != SEP_OK) {
    sep_close(c);
    return SEP_ERR_HIP;
  }
  c->gd.nx = (int)g->nx;
  c->gd.ny = (int)g->ny;
  c->gd.nz = (int)g->nz;
  c->gd.N = g->nx * g->ny * g->nz;
  c->dx = g->dx;
  c->dy = g->dy;
  c->dz = g->dz;
  c->dv = g->dx * g->dy * g->dz;
  c->xmin = g->xmin;
  c->ymin = g->ymin;
  c->zmin = g->zmin;
  c->ekinmax = g->ekinmax > 0 ? g->ekinmax : 10.0;
  // k^2 axis tables as filter/init.c:124-137 (1/N folded in, reference order)
  const double TWOPI = 6.28318530717958647692;
  const long n3[3] = {g->nx, g->ny, g->nz};
  const double d3[3] = {g->dx, g->dy, g->dz};
  const double nx_1 = 1.0 / (double)g->nx, ny_1 = 1.0 / (double)g->ny, nz_1 = 1.0 / (double)g->nz;
  double *tabs_d[3] = {nullptr, nullptr, nullptr};
  for (int a = 0; a < 3; a++) {
    const long n = n3[a];
    std::vector<double> k(n, 0.0);
    const double dk = TWOPI / ((double)n * d3[a]);
    for (long j = 1; j <= n / 2; j++) {
      double q = (double)j * dk;
      double v;
      if (a == 0)
        v = 0.5 * (q * q) * nx_1 * ny_1 * nz_1 / 1.0;
      else if (a == 1)
        v = 0.5 * (q * q) * ny_1 * nx_1 * nz_1 / 1.0;
      else
        v = 0.5 * (q * q) * nz_1 * nx_1 * ny_1 / 1.0;
      k[j] = k[n - j] = v;
    }
    if (hipMalloc(&tabs_d[a], sizeof(double) * n) != hipSuccess ||
        hipMemcpy(tabs_d[a], k.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
      sep_close(c);
      return SEP_ERR_HIP;
    }
    // twiddles exp(-2 pi i m / n)
    std::vector<double> t64(2 * n);
    std::vector<float> t32(2 * n);
    for (long m = 0; m < n; m++) {
      double ang = -TWOPI * (double)m / (double)n;
      t64[2 * m] = cos(ang);
      t64[2 * m + 1] = sin(ang);
      t32[2 * m] = (float)t64[2 * m];
      t32[2 * m + 1] = (float)t64[2 * m + 1];
    }
    if (hipMalloc(&c->gd.tw64[a], sizeof(double) * 2 * n) != hipSuccess ||
        hipMalloc(&c->gd.tw32[a], sizeof(float) * 2 * n) != hipSuccess ||
        hipMemcpy(c->gd.tw64[a], t64.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(c->gd.tw32[a], t32.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice) !=
            hipSuccess) {
      sep_close(c);
      return SEP_ERR_HIP;
    }
  }
  if (const long twb = plane_sep_tw_bytes(c->gd.nx, SEP_PREC_FP32)) {
    std::vector<char> tws(twb);
    plane_sep_tw_host(tws.data());
    if (hipMalloc(&c->gd.twsep32, twb) != hipSuccess ||
        hipMemcpy(c->gd.twsep32, tws.data(), twb, hipMemcpyHostToDevice) != hipSuccess) {
      sep_close(c);
      return SEP_ERR_HIP;
    }
  }
  c->gd.kxx = tabs_d[0];
  c->gd.kyy = tabs_d[1];
  c->gd.kzz = tabs_d[2];
  c->gd.ekin = c->ekinmax * (ny_1 * nx_1 * nz_1);
  c->gd.force_lines[0] = c->gd.force_lines[1] = getenv("SEP_FORCE_LINE_KERNELS") != nullptr;
  c->gd.no_fuse = getenv("SEP_SEPARATE_PASSES") != nullptr;
  *out = c;
  return SEP_OK;
}

int sep_close(sep_ctx *c) {
  if (!c) return SEP_OK;
  hipSetDevice(c->device);
  for (int a = 0; a < 3; a++) {
    if (c->gd.tw64[a]) hipFree(c->gd.tw64[a]);
    if (c->gd.tw32[a]) hipFree(c->gd.tw32[a]);
  }
  if (c->gd.twsep32) hipFree(c->gd.twsep32);
  for (void *p : {(void *)c->gd.kxx, (void *)c->gd.kyy, (void *)c->gd.kzz, (void *)c->pot64,
                  (void *)c->pot32, (void *)c->states, (void *)c->states2, c->tb, (void *)c->scr,
                  (void *)c->gwork, c->ring})
    if (p) hipFree(p);
  for (hipEvent_t e : c->tev)
    if (e) hipEventDestroy(e);
  for (void *p : {(void *)c->so.off, (void *)c->so.jxyz, (void *)c->so.geo, (void *)c->so.pidx,
                  (void *)c->so.poff, (void *)c->so.pent, (void *)c->so.kmat, (void *)c->so.koff,
                  (void *)c->so.pmap, (void *)c->so_contrib})
    if (p) hipFree(p);
  if (c->blas) rocblas_destroy_handle(c->blas);
  if (c->dws) hipFree(c->dws);
  if (c->zpot) hipFree(c->zpot);
  for (hipStream_t s : c->lst)
    if (s) hipStreamDestroy(s);
  for (hipEvent_t e : c->lev)
    if (e) hipEventDestroy(e);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
  return SEP_OK;
}

// ------------------------------------------------------------- potential
static int finish_potential(sep_ctx *c) {
  const long N = c->gd.N;
  size_t cap = 0;
  if (!c->pot32) RCHK(ensure(c, (void **)&c->pot32, &cap, sizeof(float) * N));
  RCHK(convert_d2f(c->pot64, c->pot32, N, c->st));
  std::vector<double> h(N);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->pot64, sizeof(double) * N, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  double vmin = 1.0e10, vmax = -1.0e10;  // filter/init.c:116-117,196-199
  for (long g = 0; g < N; g++) {
    if (vmax < h[g]) vmax = h[g];
    if (vmin > h[g]) vmin = h[g];
  }
  c->vmin = vmin;
  c->vmax = vmax;
  c->has_pot = true;
  return SEP_OK;
}

int sep_set_potential(sep_ctx *c, const double *potl) {
  if (!c || !potl) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  size_t cap = 0;
  if (!c->pot64) RCHK(ensure(c, (void **)&c->pot64, &cap, sizeof(double) * N));
  HIPCHK(c, hipMemcpyAsync(c->pot64, potl, sizeof(double) * N, hipMemcpyHostToDevice, c->st));
  return finish_potential(c);
}

int sep_build_potential(sep_ctx *c, long natom, const double *xyz, const int *type,
                        const double *scale, int ntype, const long *npot, const double *vr,
                        const double *pot, long stride, const double *dr, const double *expot,
                        double *potl_out, double *vmin, double *vmax) {
  if (!c || natom < 1 || !xyz || !type || ntype < 1 || !npot || !vr || !pot || !dr || stride < 2)
    return SEP_ERR_ARG;
  for (long a = 0; a < natom; a++)
    if (type[a] < 0 || type[a] >= ntype || npot[type[a]] < 2 || npot[type[a]] > stride ||
        !(dr[type[a]] > 0.0))
      return fail(c, SEP_ERR_ARG, "atom %ld: bad type or table", a);
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  size_t cap = 0;
  if (!c->pot64) RCHK(ensure(c, (void **)&c->pot64, &cap, sizeof(double) * N));
  DevBuf xyz_d, type_d, npot_d, vr_d, pot_d, dr_d, scale_d;
  RCHK(dev_upload(c, xyz_d, xyz, sizeof(double) * 3 * natom));
  RCHK(dev_upload(c, type_d, type, sizeof(int) * natom));
  RCHK(dev_upload(c, npot_d, npot, sizeof(long) * ntype));
  RCHK(dev_upload(c, vr_d, vr, sizeof(double) * ntype * stride));
  RCHK(dev_upload(c, pot_d, pot, sizeof(double) * ntype * stride));
  RCHK(dev_upload(c, dr_d, dr, sizeof(double) * ntype));
  if (scale) RCHK(dev_upload(c, scale_d, scale, sizeof(double) * natom));
  double ex_e = 0.0, ex_x0 = 0.0, ex_t = 1.0;
  if (expot) {
    ex_e = expot[0];
    ex_x0 = expot[1];
    ex_t = expot[2];
  }
  int rc = build_potential_dev(c->gd, c->xmin, c->ymin, c->zmin, c->dx, c->dy, c->dz, natom,
                               xyz_d.as<double>(), type_d.as<int>(), scale_d.as<double>(),
                               npot_d.as<long>(), vr_d.as<double>(), pot_d.as<double>(), stride,
                               dr_d.as<double>(), ex_e, ex_x0, ex_t, c->interp, c->pot64, c->st);
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (rc != SEP_OK) return fail(c, rc, "potential kernel launch failed");
  RCHK(finish_potential(c));
  if (potl_out)
    HIPCHK(c, hipMemcpy(potl_out, c->pot64, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (vmin) *vmin = c->vmin;
  if (vmax) *vmax = c->vmax;
  return SEP_OK;
}

}  // extern "C"

// ------------------------------------------------------------ H helpers
template <typename R>
static const R *pot_of(sep_ctx *c);
template <>
const double *pot_of<double>(sep_ctx *c) {
  return c->pot64;
}
template <>
const float *pot_of<float>(sep_ctx *c) {
  return c->pot32;
}

template <typename R>
static int apply_views(sep_ctx *c, const VecView<R> &in, const VecView<R> &out, int nvec,
                       double alpha, double beta, std::vector<double> *sums,
                       hipEvent_t *ev = nullptr) {
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  const long N = c->gd.N;
  RCHK(ensure(c, &c->tb, &c->tb_bytes, sizeof(cx<R>) * (size_t)nvec * N));
  double *part = nullptr;
  if (sums) {
    // enough for either kernel shape: rows kernel uses nrows/ROWS blocks per vector
    size_t need = (size_t)nvec * ((size_t)c->gd.ny * c->gd.nz + c->gd.nz) * 4;
    RCHK(ensure_scr(c, need));
    part = c->scr;
  }
  long nblk = 0;
  int rc = hpsi_apply<R>(c->gd, in, out, reinterpret_cast<cx<R> *>(c->tb), pot_of<R>(c), nvec,
                         alpha, beta, part, &nblk, c->st, ev);
  if (rc != SEP_OK) return fail(c, rc, "H.psi launch failed (grid %dx%dx%d)", c->gd.nx, c->gd.ny, c->gd.nz);
  if (sums) {
    std::vector<double> h((size_t)nblk * 4);
    HIPCHK(c, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const long per = nblk / nvec;
    sums->assign((size_t)nvec * 4, 0.0);
    for (int v = 0; v < nvec; v++)
      for (long b = 0; b < per; b++)
        for (int q = 0; q < 4; q++) (*sums)[v * 4 + q] += h[((size_t)v * per + b) * 4 + q];
  }
  return SEP_OK;
}

// H on m real states stored state-major at src (FP64), two per complex vector.
// dst may equal src.  Per-state sums: e[i] = sum psi*Hpsi, h2[i] = sum (Hpsi)^2.
static int apply_states(sep_ctx *c, double *src, double *dst, long m, double *e, double *h2) {
  const long N = c->gd.N;
  const long npair = m / 2;
  std::vector<double> sums;
  const bool want = e || h2;
  if (npair > 0) {
    // bound the FFT scratch: at most 64 complex vectors per launch
    for (long p0 = 0; p0 < npair; p0 += 64) {
      long np = npair - p0 < 64 ? npair - p0 : 64;
      VecView<double> in{src + 2 * p0 * N, src + 2 * p0 * N + N, nullptr, 2 * N};
      VecView<double> out{dst + 2 * p0 * N, dst + 2 * p0 * N + N, nullptr, 2 * N};
      RCHK(apply_views<double>(c, in, out, (int)np, 1.0, 0.0, want ? &sums : nullptr));
      if (want)
        for (long v = 0; v < np; v++) {
          if (e) {
            e[2 * (p0 + v)] = sums[v * 4 + 0];
            e[2 * (p0 + v) + 1] = sums[v * 4 + 1];
          }
          if (h2) {
            h2[2 * (p0 + v)] = sums[v * 4 + 2];
            h2[2 * (p0 + v) + 1] = sums[v * 4 + 3];
          }
        }
    }
  }
  if (m & 1) {
    VecView<double> in{src + (m - 1) * N, nullptr, nullptr, N};
    VecView<double> out{dst + (m - 1) * N, nullptr, nullptr, N};
    RCHK(apply_views<double>(c, in, out, 1, 1.0, 0.0, want ? &sums : nullptr));
    if (e) e[m - 1] = sums[0];
    if (h2) h2[m - 1] = sums[2];
  }
  return SEP_OK;
}

static int normalize_dev(sep_ctx *c, double *x, long m) {
  const long N = c->gd.N;
  if (m <= 0) return SEP_OK;
  const long nb = (N + 16383) / 16384;
  RCHK(ensure_scr(c, (size_t)nb * m + 2 * m));
  double *out = c->scr + nb * m;
  RCHK(dots_dev(x, N, x, N, m, N, out, c->scr, nb * m, c->st));
  std::vector<double> nrm(m);
  HIPCHK(c, hipMemcpyAsync(nrm.data(), out, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (long i = 0; i < m; i++) nrm[i] = nrm[i] != 0.0 ? 1.0 / sqrt(nrm[i] * c->dv) : 0.0;  // norm.c:50
  HIPCHK(c, hipMemcpyAsync(out + m, nrm.data(), sizeof(double) * m, hipMemcpyHostToDevice, c->st));
  RCHK(scale_rows_dev(x, N, m, N, out + m, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

extern "C" {

int sep_apply_h(sep_ctx *c, const double *in, double *out, long batch) {
  return sep_hnorm(c, in, out, batch, -2.0, 0.0, 1.0);
}

int sep_hnorm(sep_ctx *c, const double *in, double *out, long batch, double zm1, double vmin,
              double dE_1) {
  if (!c || !in || !out || batch < 1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  size_t bytes = sizeof(double) * 2 * (size_t)N * batch;
  void *d = nullptr;
  HIPCHK(c, hipMalloc(&d, bytes));
  int rc = SEP_OK;
  if (hipMemcpy(d, in, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = SEP_ERR_HIP;
  // alpha (T + V) psi - beta psi; hamiltonian(): zm1 = -2, vmin = 0, dE_1 = 1
  const double alpha = dE_1, beta = 2.0 + zm1 + vmin * dE_1;
  for (long b0 = 0; rc == SEP_OK && b0 < batch; b0 += 64) {
    long nb = batch - b0 < 64 ? batch - b0 : 64;
    cx<double> *p = reinterpret_cast<cx<double> *>(d) + b0 * N;
    VecView<double> v{nullptr, nullptr, p, N};
    rc = apply_views<double>(c, v, v, (int)nb, alpha, beta, nullptr);
  }
  if (rc == SEP_OK && hipMemcpy(out, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = SEP_ERR_HIP;
  hipFree(d);
  if (rc == SEP_ERR_HIP) return fail(c, rc, "hnorm: HIP copy failed");
  return rc;
}

int sep_apply_h_dev(sep_ctx *c, const void *in_dev, void *out_dev, long batch, int precision) {
  if (!c || !in_dev || !out_dev || batch < 1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  for (long b0 = 0; b0 < batch; b0 += 64) {
    long nb = batch - b0 < 64 ? batch - b0 : 64;
    if (precision == SEP_PREC_FP64) {
      VecView<double> vi{nullptr, nullptr, (cx<double> *)in_dev + b0 * N, N};
      VecView<double> vo{nullptr, nullptr, (cx<double> *)out_dev + b0 * N, N};
      RCHK(apply_views<double>(c, vi, vo, (int)nb, 1.0, 0.0, nullptr));
    } else {
      VecView<float> vi{nullptr, nullptr, (cx<float> *)in_dev + b0 * N, N};
      VecView<float> vo{nullptr, nullptr, (cx<float> *)out_dev + b0 * N, N};
      RCHK(apply_views<float>(c, vi, vo, (int)nb, 1.0, 0.0, nullptr));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

int sep_apply_h_dev_noalloc(sep_ctx *c, const void *in_dev, void *out_dev, long batch,
                            int precision) {
  if (!c || !in_dev || !out_dev || batch < 1) return SEP_ERR_ARG;
  if (precision != SEP_PREC_FP64 && precision != SEP_PREC_FP32)
    return fail(c, SEP_ERR_ARG, "unknown precision %d", precision);
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  const long N = c->gd.N;
  const size_t per = (precision == SEP_PREC_FP64 ? sizeof(cx<double>) : sizeof(cx<float>)) * (size_t)N;
  if ((size_t)batch * per > (size_t)SEP_NOALLOC_SCRATCH_BYTES)
    return fail(c, SEP_ERR_ARG,
                "apply_h_dev_noalloc: %ld vectors need %zu bytes of scratch, the library holds %ld",
                batch, (size_t)batch * per, (long)SEP_NOALLOC_SCRATCH_BYTES);
  void *tb = g_noalloc_ptr[c->device];
  if (!tb) return fail(c, SEP_ERR_HIP, "apply_h_dev_noalloc: scratch not initialised");
  HIPCHK(c, hipSetDevice(c->device));
  const long NB = 64;  // complex vectors per launch, as sep_apply_h_dev
  int rc = SEP_OK;
  for (long b0 = 0; rc == SEP_OK && b0 < batch; b0 += NB) {
    const int nb = (int)(batch - b0 < NB ? batch - b0 : NB);
    long nblk = 0;
    if (precision == SEP_PREC_FP64) {
      VecView<double> vi{nullptr, nullptr, (cx<double> *)in_dev + b0 * N, N};
      VecView<double> vo{nullptr, nullptr, (cx<double> *)out_dev + b0 * N, N};
      rc = hpsi_apply<double>(c->gd, vi, vo, reinterpret_cast<cx<double> *>(tb), c->pot64, nb, 1.0,
                              0.0, nullptr, &nblk, c->st);
    } else {
      VecView<float> vi{nullptr, nullptr, (cx<float> *)in_dev + b0 * N, N};
      VecView<float> vo{nullptr, nullptr, (cx<float> *)out_dev + b0 * N, N};
      rc = hpsi_apply<float>(c->gd, vi, vo, reinterpret_cast<cx<float> *>(tb), c->pot32, nb, 1.0,
                             0.0, nullptr, &nblk, c->st);
    }
  }
  if (rc != SEP_OK) return fail(c, rc, "apply_h_dev_noalloc: H.psi launch failed");
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

// ----------------------------------------------------------- power iteration
int sep_energy_range(sep_ctx *c, double vmin, double *dE, double *emax, long *iters) {
  return sep_energy_range_log(c, vmin, dE, emax, iters, nullptr, 0);
}

int sep_energy_range_log(sep_ctx *c, double vmin, double *dE, double *emax, long *iters,
                         double *log, long maxlog) {
  if (!c || (maxlog > 0 && !log)) return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  // fixed-seed start, filter/energy.c:57,61
  std::vector<double> re(N);
  sep_rng st;
  sep_rng_reset(&st);
  long idum = -874917403L;
  sep_init_psi(&st, &idum, N, c->dv, re.data());
  std::vector<double> h(2 * N, 0.0);
  for (long g = 0; g < N; g++) h[2 * g] = re[g];
  double *v = nullptr, *w = nullptr;
  HIPCHK(c, hipMalloc(&v, sizeof(double) * 2 * N));
  HIPCHK(c, hipMalloc(&w, sizeof(double) * 2 * N));
  HIPCHK(c, hipMemcpy(v, h.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice));
  VecView<double> V{nullptr, nullptr, (cx<double> *)v, N}, W{nullptr, nullptr, (cx<double> *)w, N};
  // w = H v.  Iteration i of the reference: psi <- H psi / |H psi|,
  // ene = <psi|H psi> dv.  H psi for the next iteration is the H psi of
  // the energy evaluation, so one application per iteration suffices.
  std::vector<double> sums;
  int rc = apply_views<double>(c, V, W, 1, 1.0, 0.0, &sums);
  double ene = 0.0 + 0.1, ene_old = 0.0;
  long i = 0;
  const long nb = (2 * N + 16383) / 16384;
  for (; rc == SEP_OK && (fabs((ene - ene_old) / ene) > 1.0e-4) && (i < 500); i++) {
    // |w|^2 over the complex vector = sums[2] + sums[3]
    double norma = sqrt((sums[2] + sums[3]) * c->dv);
    double inv = 1.0 / norma;
    if ((rc = ensure_scr(c, nb + 1)) != SEP_OK) break;
    if (hipMemcpy(c->scr + nb, &inv, sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      rc = SEP_ERR_HIP;
      break;
    }
    // v = w / norma (the normalised H psi)
    if (hipMemcpyAsync(v, w, sizeof(double) * 2 * N, hipMemcpyDeviceToDevice, c->st) != hipSuccess) {
      rc = SEP_ERR_HIP;
      break;
    }
    if ((rc = scale_rows_dev(v, 2 * N, 1, 2 * N, c->scr + nb, c->st)) != SEP_OK) break;
    ene_old = ene;
    if ((rc = apply_views<double>(c, V, W, 1, 1.0, 0.0, &sums)) != SEP_OK) break;
    ene = (sums[0] + sums[1]) * c->dv;
    if (i < maxlog) {  // the line of ene-init.dat (filter/energy.c:72)
      log[3 * i] = ene_old;
      log[3 * i + 1] = ene;
      log[3 * i + 2] = norma;
    }
  }
  hipFree(v);
  hipFree(w);
  if (rc != SEP_OK) return rc;
  if (dE) *dE = 1.1 * (ene - vmin);
  if (emax) *emax = ene;
  if (iters) *iters = i;
  return SEP_OK;
}

// -------------------------------------------------------------- states
int sep_states_upload(sep_ctx *c, const double *psi, long m) {
  if (!c || !psi || m < 1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(ensure_states(c, m, 1));
  HIPCHK(c, hipMemcpy(c->states, psi, sizeof(double) * m * c->gd.N, hipMemcpyHostToDevice));
  return SEP_OK;
}
int sep_states_download(sep_ctx *c, double *psi, long m) {
  if (!c || !psi || m < 1 || m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->st));
  HIPCHK(c, hipMemcpy(psi, c->states, sizeof(double) * m * c->gd.N, hipMemcpyDeviceToHost));
  return SEP_OK;
}
int sep_normalize_all(sep_ctx *c, long m) {
  if (!c || m < 0 || m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  return normalize_dev(c, c->states, m);
}
int sep_energy_all(sep_ctx *c, long m, double *ene) {
  if (!c || !ene || m < 1 || m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(ensure_states(c, m, 2));
  RCHK(apply_states(c, c->states, c->states2, m, ene, nullptr));
  for (long i = 0; i < m; i++) ene[i] *= c->dv;
  return SEP_OK;
}

int sep_sigma(sep_ctx *c, long m, double *sige) {
  if (!c || !sige || m < 1 || m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(ensure_states(c, m, 2));
  std::vector<double> e(m), h2(m);
  RCHK(apply_states(c, c->states, c->states2, m, e.data(), h2.data()));
  // <psi|H^2 psi> = |H psi|^2 for symmetric H (one application instead of two)
  for (long i = 0; i < m; i++) {
    double ev = e[i] * c->dv;
    double s2 = h2[i] * c->dv - ev * ev;
    sige[i] = sqrt(fabs(s2));
  }
  return SEP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- filter
// One filter pass (filter/filter.c:59-83 for all random states) in three
// stages, so that several grids can be interleaved or captured into graphs:
// start (allocations, uploads, zeroing), one Newton step per call (three
// H.psi passes, plus the accumulation of a full ring block), finish.
template <typename R>
struct FilterRun {
  sep_ctx *c = nullptr;
  long N = 0, ns = 0, nc = 0, ms = 0, K = 0, slot = 0, nvec = 0, j = 1;
  double vmin = 0.0, dE_1 = 0.0;
  std::vector<double> zn;
  double *an_d = nullptr;
  float *cf_d = nullptr;  // complex64 filter: coefficients packed for accum_mfma_packed_dev
  R *ring = nullptr;
  bool sampled = false;
  bool fused = false;  // split steps: tb carries the spectrum of the newest vector
  // spinors (4 real rows each: up.re, up.im, dn.re, dn.im): each step adds
  // dE_1 V_so psi_(j-1), and the finish leaves normalization to the caller
  bool so = false;
  // where this run's launches go: its stream, its slice of the spectrum
  // buffer and its filtered states (a lane owns a range of complex vectors)
  hipStream_t st = nullptr;
  cx<R> *tb = nullptr;
  double *acc = nullptr;
};

__global__ void k_zero(double *x, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = 0.0;
}
// zero fill by a kernel (also safe inside a captured graph)
static int zero_dev(void *p, size_t bytes, hipStream_t st) {
  hipLaunchKernelGGL(k_zero, dim3(2048), dim3(256), 0, st, static_cast<double *>(p),
                     (long)(bytes / sizeof(double)));
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

template <typename R>
static int filter_start(FilterRun<R> &f, sep_ctx *c, const double *psi0, bool psi0_on_device,
                        long ns, const double *an, const double *zn, long nc, long ms, double vmin,
                        double dE, bool so = false, double ring_cap = 0.0) {
  const long N = c->gd.N;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  f.c = c;
  f.N = N;
  f.ns = ns;
  f.nc = nc;
  f.ms = ms;
  f.vmin = vmin;
  f.dE_1 = 4.0 / dE;
  f.nvec = (ns + 1) / 2;
  const long S2 = 2 * f.nvec;
  // K Newton vectors are accumulated per pass over the ring; the
  // accumulators move 2 ms 8 / K bytes per real-state point per step.  The
  // ring holds up to 1024 vectors in up to 160 GB and 60 % of the free
  // device memory (the ring already held by this context counts as free):
  // at 128^3 with 16 states, 1024 complex64 vectors (137 GB) filter 4 %
  // faster than 512.  The test hook SEP_RING_DEPTH lowers the bound on K, so
  // that short filters span several ring blocks.
  // ring_cap > 0: the bound a batch of grids planned for this one
  size_t mfree = 0, mtotal = 0;
  if (ring_cap <= 0.0 && hipMemGetInfo(&mfree, &mtotal) != hipSuccess) mfree = 0;
  const double cap =
      ring_cap > 0.0 ? ring_cap : std::min(160.0e9, 0.6 * ((double)mfree + (double)c->ring_bytes));
  const char *kenv = getenv("SEP_RING_DEPTH");
  const long kmax = kenv && atol(kenv) >= 1 && atol(kenv) <= 1024 ? atol(kenv) : 1024;
  long K = nc < kmax ? nc : kmax;
  while (K > 8 && (double)K * S2 * N * sizeof(R) > cap) K /= 2;
  // ring layout [state][Newton vector][point]: row (s, k) at s * K * N + k * N,
  // so the accumulation walks each state's K rows within K * N elements;
  // where the allocation fails (memory the free-memory figure promised went
  // elsewhere meanwhile) the ring is halved until it fits
  int rc_ring = ensure(c, &c->ring, &c->ring_bytes, sizeof(R) * (size_t)K * S2 * N);
  while (rc_ring != SEP_OK && K > 8) {
    (void)hipGetLastError();  // the failed allocation must not be reported by a later launch
    K /= 2;
    rc_ring = ensure(c, &c->ring, &c->ring_bytes, sizeof(R) * (size_t)K * S2 * N);
  }
  RCHK(rc_ring);
  f.K = K;
  f.slot = N;
  RCHK(ensure(c, &c->tb, &c->tb_bytes, sizeof(cx<R>) * (size_t)f.nvec * N));
  f.ring = reinterpret_cast<R *>(c->ring);
  RCHK(ensure_states(c, ns * ms, 1));
  RCHK(zero_dev(c->states, sizeof(double) * ns * ms * N, c->st));
  f.st = c->st;
  f.tb = reinterpret_cast<cx<R> *>(c->tb);
  f.acc = c->states;
  // row (s, 0) <- psi_0 of state s; the padding state (odd ns) starts at zero
  const long ss = K * N;  // state stride of the ring
  if (ns < S2) RCHK(zero_dev(f.ring + ns * ss, sizeof(R) * N, c->st));
  const hipMemcpyKind kind = psi0_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (sizeof(R) == 8) {
    HIPCHK(c, hipMemcpy2DAsync(f.ring, sizeof(R) * ss, psi0, sizeof(double) * N,
                               sizeof(double) * N, ns, kind, c->st));
  } else {
    const double *src = psi0;
    if (!psi0_on_device) {
      RCHK(ensure_states(c, ns, 2));
      HIPCHK(c, hipMemcpyAsync(c->states2, psi0, sizeof(double) * ns * N, kind, c->st));
      src = c->states2;
    }
    for (long st = 0; st < ns; st++)
      RCHK(convert_d2f(src + st * N, (float *)f.ring + st * ss, N, c->st));
  }
  // coefficients transposed to [j][ie] for the accumulation kernel
  std::vector<double> anT((size_t)ms * nc);
  for (long ie = 0; ie < ms; ie++)
    for (long j = 0; j < nc; j++) anT[j * ms + ie] = an[ie * nc + j];
  HIPCHK(c, hipMalloc(&f.an_d, sizeof(double) * ms * nc));
  HIPCHK(c, hipMemcpyAsync(f.an_d, anT.data(), sizeof(double) * ms * nc, hipMemcpyHostToDevice, c->st));
  std::vector<float> cf;
  if (sizeof(R) == 4) {
    cf.resize(accum_pack_len(nc, (int)ms));
    accum_pack_f32(an, nc, (int)ms, cf.data());
    HIPCHK(c, hipMalloc(&f.cf_d, sizeof(float) * cf.size()));
    HIPCHK(c, hipMemcpyAsync(f.cf_d, cf.data(), sizeof(float) * cf.size(), hipMemcpyHostToDevice,
                             c->st));
  }
  HIPCHK(c, hipStreamSynchronize(c->st));  // anT and cf outlive the copies
  f.zn.assign(zn, zn + nc);
  f.j = 1;
  // pass A of psi_0; every later pass A runs fused with the previous pass C
  f.so = so;
  // spinors: V_so joins psi_j in the fused kernel's epilogue where the FP64
  // whole-plane family runs; elsewhere the steps run unfused
  const bool so_fused = sizeof(R) == 8 && c->so.pmap && c->so.natom > 0 &&
                        hpsi_plane_family(c->gd, SEP_PREC_FP64);
  f.fused = (!so || so_fused) && nc > 1 && !c->gd.no_fuse && hpsi_fused_supported<R>(c->gd);
  if (f.fused) {
    VecView<R> v0{f.ring, f.ring + ss, nullptr, 2 * ss};
    if (hpsi_fwd<R>(c->gd, v0, f.tb, (int)f.nvec, f.st) != SEP_OK)
      return fail(c, SEP_ERR_HIP, "filter: pass A launch failed");
  }
  return SEP_OK;
}

// acc += the block of cnt stored Newton vectors starting at Newton index j0
// (filter/filter.c:74-78): the packed-coefficient kernel for the complex64
// filter where the shape allows it, else the general one.
template <typename R>
static int accum_block(FilterRun<R> &f, long sstride, int cnt, long j0) {
  sep_ctx *c = f.c;
  if constexpr (sizeof(R) == 4) {
    if (f.cf_d) {
      const int rc = accum_mfma_packed_dev(reinterpret_cast<const float *>(f.ring), f.slot,
                                           sstride, cnt, f.cf_d, f.nc, j0, (int)f.ms, f.acc,
                                           (int)f.ns, f.N, f.st);
      if (rc != SEP_ERR_UNSUPPORTED)
        return rc == SEP_OK ? SEP_OK : fail(c, rc, "accumulation launch failed");
    }
  }
  const int rc = accum_mfma_dev<R>(f.ring, f.slot, sstride, cnt, f.an_d, j0, (int)f.ms, f.acc,
                                   (int)f.ns, f.N, f.st);
  return rc == SEP_OK ? SEP_OK : fail(c, rc, "accumulation launch failed");
}

// Newton step f.j (filter/filter.c:70-78, hnorm.c:19-24).  No host
// synchronisation unless live timing is on (which graphs must not use).
template <typename R>
static int filter_step(FilterRun<R> &f) {
  sep_ctx *c = f.c;
  const long j = f.j, K = f.K, slot = f.slot, N = f.N;
  const long sp = (j - 1) % K, sj = j % K;
  const long ss = K * N;  // state stride: vector v = states (2v, 2v+1)
  VecView<R> in{f.ring + sp * slot, f.ring + ss + sp * slot, nullptr, 2 * ss};
  VecView<R> out{f.ring + sj * slot, f.ring + ss + sj * slot, nullptr, 2 * ss};
  // psi_j = dE_1 H psi_{j-1} - (2 + z_{j-1} + Vmin dE_1) psi_{j-1}
  const double beta = 2.0 + f.zn[j - 1] + f.vmin * f.dE_1;
  const bool sample = c->timing && sj == K / 2;
  if (f.fused) {
    const bool fwd_next = j + 1 < f.nc;
    SoEpi se{}, *sep_ = nullptr;
    if constexpr (sizeof(R) == 8) {
      if (f.so && c->so.natom > 0) {
        // V_so psi_{j-1} per covered point, added in the epilogue; the last
        // step (no fused forward pass) adds it after the step as below
        const long nsp = f.ns / 4;
        const size_t nc4 = 4 * (size_t)c->so.nent * nsp;
        RCHK(ensure(c, (void **)&c->so_contrib, &c->so_contrib_bytes,
                    sizeof(double) * (nc4 + 4 * (size_t)c->so.npts * nsp)));
        if (fwd_next) {
          if (so_gather_dev(c->so, f.ring + sp * slot, ss, 4 * ss, nsp, c->so_contrib,
                            c->so_contrib + nc4, f.st) != SEP_OK)
            return fail(c, SEP_ERR_HIP, "filter: spin-orbit launch failed");
          se = SoEpi{c->so.pmap, c->so_contrib + nc4, nsp, f.dE_1};
          sep_ = &se;
        }
      }
    }
    if (hpsi_step<R>(c->gd, in, out, f.tb, pot_of<R>(c), (int)f.nvec, f.dE_1, beta, fwd_next,
                     f.st, sample ? c->tev : nullptr, sep_) != SEP_OK)
      return fail(c, SEP_ERR_HIP, "filter: H.psi step launch failed");
    if constexpr (sizeof(R) == 8) {
      if (f.so && c->so.natom > 0 && !fwd_next) {
        const long nsp = f.ns / 4;
        if (so_apply_dev(c->so, f.ring + sp * slot, ss, 4 * ss, f.ring + sj * slot, ss, 4 * ss,
                         nsp, f.dE_1, c->so_contrib, f.st) != SEP_OK)
          return fail(c, SEP_ERR_HIP, "filter: spin-orbit launch failed");
      }
    }
  } else {
    RCHK(apply_views<R>(c, in, out, (int)f.nvec, f.dE_1, beta, nullptr, sample ? c->tev : nullptr));
    if constexpr (sizeof(R) == 8) {
      if (f.so && c->so.natom > 0) {
        // spinor s, component q of slot k at ring + (4 s + q) ss + k slot
        const long nsp = f.ns / 4;
        RCHK(ensure(c, (void **)&c->so_contrib, &c->so_contrib_bytes,
                    sizeof(double) * 4 * (size_t)c->so.nent * nsp));
        if (so_apply_dev(c->so, f.ring + sp * slot, ss, 4 * ss, f.ring + sj * slot, ss, 4 * ss,
                         nsp, f.dE_1, c->so_contrib, f.st) != SEP_OK)
          return fail(c, SEP_ERR_HIP, "filter: spin-orbit launch failed");
      }
    }
  }
  f.sampled |= sample;
  if (sj == K - 1 || j == f.nc - 1) {
    if (f.sampled) hipEventRecord(c->tev[4], f.st);
    RCHK(accum_block<R>(f, ss, (int)(sj + 1), j - sj));
    if (f.sampled) {
      hipEventRecord(c->tev[5], f.st);
      HIPCHK(c, hipEventSynchronize(c->tev[5]));
      float t;
      for (int q = f.fused ? 1 : 0; q < 3; q++) {  // fused: pass A runs inside pass C
        hipEventElapsedTime(&t, c->tev[q], c->tev[q + 1]);
        c->tsum[q] += t;
      }
      c->tcount++;
      if (sj + 1 == K) {  // only full blocks time the accumulation of K vectors
        hipEventElapsedTime(&t, c->tev[4], c->tev[5]);
        c->tsum[3] += t;
        c->tacc_count++;
      }
    }
    f.sampled = false;
  }
  f.j++;
  return SEP_OK;
}

template <typename R>
static int filter_finish(FilterRun<R> &f, int rc) {
  sep_ctx *c = f.c;
  if (rc == SEP_OK && f.nc == 1)
    rc = accum_block<R>(f, f.K * f.N, 1, 0);
  hipStreamSynchronize(c->st);
  if (f.an_d) hipFree(f.an_d);
  if (f.cf_d) hipFree(f.cf_d);
  f.an_d = nullptr;
  f.cf_d = nullptr;
  if (rc != SEP_OK) return rc;
  if (f.so) return SEP_OK;  // spinors: normalized whole by the caller
  return normalize_dev(c, c->states, f.ns * f.ms);  // filter.c:33
}

// Number of lanes for a started run: the split-step (fused) loop only, no
// live timing (its events sample one stream), at most one lane per complex
// vector.  Automatic: two lanes from 4 complex vectors (8 random states) on
// -- measured at 128^3 (16 states, nc 4096): complex64 0.864 -> 0.813 s,
// complex128 2.078 -> 2.005 s; four lanes 0.854 s.  The second lane's stream
// is created by sep_open, further ones and the events on first use, all at
// default priority.
template <typename R>
static int filter_lanes(sep_ctx *c, const FilterRun<R> &f) {
  // spinors: one lane (the spin-orbit work buffer is shared)
  if (!f.fused || f.so || c->timing || f.nc <= 1) return 1;
  long L = c->lanes > 0 ? c->lanes : (f.nvec >= 4 ? 2 : 1);
  if (L > f.nvec) L = f.nvec;
  for (long l = 0; l < L; l++) {
    if (!c->lev[l] && hipEventCreateWithFlags(&c->lev[l], hipEventDisableTiming) != hipSuccess) {
      c->lev[l] = nullptr;
      return 1;
    }
    if (l > 0 && !c->lst[l - 1] && hipStreamCreate(&c->lst[l - 1]) != hipSuccess) {
      c->lst[l - 1] = nullptr;
      return 1;
    }
  }
  return (int)L;
}

template <typename R>
static int filter_impl(sep_ctx *c, const double *psi0, bool psi0_on_device, long ns,
                       const double *an, const double *zn, long nc, long ms, double vmin,
                       double dE, bool so = false) {
  RingBusy busy;
  busy.add(c);
  FilterRun<R> f;
  int rc = filter_start<R>(f, c, psi0, psi0_on_device, ns, an, zn, nc, ms, vmin, dE, so);
  const int L = rc == SEP_OK ? filter_lanes<R>(c, f) : 1;
  if (L <= 1) {
    while (rc == SEP_OK && f.j < nc) rc = filter_step<R>(f);
    return filter_finish<R>(f, rc);
  }
  // lanes: complex vectors [v0, v1) each, with their ring rows, spectrum
  // slice and filtered states, stepped in turn on their own streams after
  // the setup on st (fork), joined back into st before the finish
  std::vector<FilterRun<R>> lane((size_t)L, f);
  const long ss = f.K * f.N;
  HIPCHK(c, hipEventRecord(c->lev[0], c->st));
  for (int l = 0; l < L; l++) {
    const long v0 = l * f.nvec / L, v1 = (l + 1) * f.nvec / L;
    FilterRun<R> &g = lane[l];
    g.nvec = v1 - v0;
    g.ns = std::min(2 * v1, f.ns) - 2 * v0;
    g.ring = f.ring + 2 * v0 * ss;
    g.tb = f.tb + v0 * f.N;
    g.acc = f.acc + 2 * v0 * f.ms * f.N;
    g.st = l == 0 ? c->st : c->lst[l - 1];
    if (l > 0) HIPCHK(c, hipStreamWaitEvent(g.st, c->lev[0], 0));
  }
  while (rc == SEP_OK && lane[0].j < nc)
    for (int l = 0; l < L && rc == SEP_OK; l++) rc = filter_step<R>(lane[l]);
  for (int l = 1; l < L; l++) {  // join, also after a failed launch
    hipEventRecord(c->lev[l], lane[l].st);
    hipStreamWaitEvent(c->st, c->lev[l], 0);
  }
  return filter_finish<R>(f, rc);
}

// ------------------------------------------------- batched uneven grids
template <typename R>
static int filter_batch_impl(int nd, sep_ctx *const *ctxs, const void *const *psi0,
                             const long *ns, const double *const *an, const double *const *zn,
                             const long *nc, const long *ms, const double *vmin, const double *dE,
                             int mode) {
  RingBusy busy;
  for (int d = 0; d < nd; d++) busy.add(ctxs[d]);
  std::vector<FilterRun<R>> runs(nd);
  std::vector<hipGraphExec_t> execs(nd, nullptr);
  std::vector<hipGraph_t> graphs(nd, nullptr);
  int rc = SEP_OK;
  // One ring depth K for all grids: the rings share 60 % of the free memory
  // (the rings these contexts hold count as free) in proportion to their
  // bytes per Newton vector, instead of each taking 60 % of what the grids
  // before it left.  Rings larger than planned are released first, so the
  // planned ones fit.
  std::vector<double> per_k(nd);
  double avail = 0.0, sum_k = 0.0;
  {
    size_t mfree = 0, mtotal = 0;
    if (hipMemGetInfo(&mfree, &mtotal) == hipSuccess) avail = (double)mfree;
  }
  long kmax = 0;
  for (int d = 0; d < nd; d++) {
    per_k[d] = (double)(2 * ((ns[d] + 1) / 2)) * (double)ctxs[d]->gd.N * sizeof(R);
    sum_k += per_k[d];
    avail += (double)ctxs[d]->ring_bytes;
    kmax = std::max(kmax, nc[d]);
  }
  const char *kenv = getenv("SEP_RING_DEPTH");
  long K = std::min(kmax, (kenv && atol(kenv) >= 1 && atol(kenv) <= 1024) ? atol(kenv) : 1024L);
  const double total = std::min(160.0e9 * nd, 0.6 * avail);
  while (K > 8 && (double)K * sum_k > total) K /= 2;
  for (int d = 0; d < nd; d++) {
    sep_ctx *c = ctxs[d];
    if (c->ring && (double)c->ring_bytes > (double)K * per_k[d]) {
      hipStreamSynchronize(c->st);
      hipFree(c->ring);
      c->ring = nullptr;
      c->ring_bytes = 0;
    }
  }
  for (int d = 0; d < nd && rc == SEP_OK; d++) {
    rc = filter_start<R>(runs[d], ctxs[d], static_cast<const double *>(psi0[d]), true, ns[d],
                         an[d], zn[d], nc[d], ms[d], vmin[d], dE[d], false,
                         (double)K * per_k[d] * (1.0 + 1e-9));
    if (rc == SEP_OK) rc = hpsi_prepare<R>(ctxs[d]->gd);
  }
  if (rc == SEP_OK && mode == SEP_BATCH_GRAPH) {
    // each grid's whole step sequence as one graph on its own stream
    for (int d = 0; d < nd && rc == SEP_OK; d++) {
      sep_ctx *c = ctxs[d];
      const bool timing = c->timing;
      c->timing = false;
      if (hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        rc = fail(c, SEP_ERR_HIP, "batch: stream capture failed to start");
      } else {
        while (rc == SEP_OK && runs[d].j < nc[d]) rc = filter_step<R>(runs[d]);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->st, &g);
        graphs[d] = g;
        if (rc == SEP_OK && e != hipSuccess) rc = fail(c, SEP_ERR_HIP, "batch: capture failed");
        if (rc == SEP_OK && hipGraphInstantiate(&execs[d], g, nullptr, nullptr, 0) != hipSuccess)
          rc = fail(c, SEP_ERR_HIP, "batch: graph instantiation failed");
      }
      c->timing = timing;
    }
    for (int d = 0; d < nd && rc == SEP_OK; d++)
      if (hipGraphLaunch(execs[d], ctxs[d]->st) != hipSuccess)
        rc = fail(ctxs[d], SEP_ERR_HIP, "batch: graph launch failed");
  } else if (rc == SEP_OK) {
    // interleave: one Newton step of every unfinished grid per round, each
    // grid on its own stream(s) -- split into filter lanes as sep_filter_dev
    // splits it -- so that small grids fill what large ones leave
    std::vector<std::vector<FilterRun<R>>> lanes(nd);
    for (int d = 0; d < nd && rc == SEP_OK; d++) {
      sep_ctx *c = ctxs[d];
      const int L = filter_lanes<R>(c, runs[d]);
      if (L <= 1) {
        lanes[d].assign(1, runs[d]);
        continue;
      }
      const FilterRun<R> &f = runs[d];
      const long ss = f.K * f.N;
      if (hipEventRecord(c->lev[0], c->st) != hipSuccess) rc = fail(c, SEP_ERR_HIP, "batch: fork failed");
      lanes[d].assign((size_t)L, f);
      for (int l = 0; l < L && rc == SEP_OK; l++) {
        const long v0 = l * f.nvec / L, v1 = (l + 1) * f.nvec / L;
        FilterRun<R> &g = lanes[d][l];
        g.nvec = v1 - v0;
        g.ns = std::min(2 * v1, f.ns) - 2 * v0;
        g.ring = f.ring + 2 * v0 * ss;
        g.tb = f.tb + v0 * f.N;
        g.acc = f.acc + 2 * v0 * f.ms * f.N;
        g.st = l == 0 ? c->st : c->lst[l - 1];
        if (l > 0 && hipStreamWaitEvent(g.st, c->lev[0], 0) != hipSuccess)
          rc = fail(c, SEP_ERR_HIP, "batch: fork failed");
      }
    }
    bool left = true;
    while (rc == SEP_OK && left) {
      left = false;
      for (int d = 0; d < nd && rc == SEP_OK; d++)
        if (lanes[d][0].j < nc[d]) {
          for (size_t l = 0; l < lanes[d].size() && rc == SEP_OK; l++)
            rc = filter_step<R>(lanes[d][l]);
          left = true;
        }
    }
    for (int d = 0; d < nd; d++) {  // join each grid's lanes into its stream
      sep_ctx *c = ctxs[d];
      for (size_t l = 1; l < lanes[d].size(); l++) {
        hipEventRecord(c->lev[l], lanes[d][l].st);
        hipStreamWaitEvent(c->st, c->lev[l], 0);
      }
    }
  }
  int rc2 = SEP_OK;
  for (int d = 0; d < nd; d++) {
    const int r = filter_finish<R>(runs[d], rc);
    if (rc2 == SEP_OK) rc2 = r;
    if (execs[d]) hipGraphExecDestroy(execs[d]);
    if (graphs[d]) hipGraphDestroy(graphs[d]);
  }
  return rc != SEP_OK ? rc : rc2;
}

extern "C" int sep_filter_batch(int ndots, sep_ctx *const *ctxs, const void *const *psi0_dev,
                                const long *ns, const double *const *an,
                                const double *const *zn, const long *nc, const long *ms,
                                const double *vmin, const double *dE, int precision, int mode) {
  if (ndots < 1 || !ctxs || !psi0_dev || !ns || !an || !zn || !nc || !ms || !vmin || !dE)
    return SEP_ERR_ARG;
  if (mode != SEP_BATCH_INTERLEAVE && mode != SEP_BATCH_GRAPH) return SEP_ERR_ARG;
  for (int d = 0; d < ndots; d++) {
    sep_ctx *c = ctxs[d];
    if (!c || !psi0_dev[d] || ns[d] < 1 || !an[d] || !zn[d] || nc[d] < 1 || ms[d] < 1 ||
        !(dE[d] > 0.0))
      return SEP_ERR_ARG;
    if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
    for (int e = 0; e < d; e++)
      if (ctxs[e] == c) return fail(c, SEP_ERR_ARG, "batch: context %d appears twice", d);
    if (c->device != ctxs[0]->device)
      return fail(c, SEP_ERR_ARG, "batch: all contexts must be on one device");
  }
  HIPCHK(ctxs[0], hipSetDevice(ctxs[0]->device));
  if (precision == SEP_PREC_FP64)
    return filter_batch_impl<double>(ndots, ctxs, psi0_dev, ns, an, zn, nc, ms, vmin, dE, mode);
  if (precision == SEP_PREC_FP32)
    return filter_batch_impl<float>(ndots, ctxs, psi0_dev, ns, an, zn, nc, ms, vmin, dE, mode);
  return fail(ctxs[0], SEP_ERR_ARG, "unknown precision %d", precision);
}

extern "C" int sep_filter(sep_ctx *c, const double *psi0, long ns, const double *an,
                          const double *zn, long nc, long ms, double vmin, double dE,
                          int precision, double *psitot, double *ene) {
  if (!c || !psi0 || ns < 1 || !an || !zn || nc < 1 || ms < 1 || !(dE > 0.0)) return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  int rc;
  if (precision == SEP_PREC_FP64)
    rc = filter_impl<double>(c, psi0, false, ns, an, zn, nc, ms, vmin, dE);
  else if (precision == SEP_PREC_FP32)
    rc = filter_impl<float>(c, psi0, false, ns, an, zn, nc, ms, vmin, dE);
  else
    return fail(c, SEP_ERR_ARG, "unknown precision %d", precision);
  if (rc != SEP_OK) return rc;
  if (ene) RCHK(sep_energy_all(c, ns * ms, ene));
  if (psitot) RCHK(sep_states_download(c, psitot, ns * ms));
  return SEP_OK;
}

// ------------------------------------------- dense subspace matrices
// The subspace stage's n x n problems (n = kept states, up to 2 ns ms) are
// solved on the device by rocSOLVER in place of the reference's host LAPACK
// calls.  Matrices are row-major and symmetric on input, so the column-major
// view rocSOLVER takes is the same matrix.
static int dense_init(sep_ctx *c, long n) {
  if (!c->blas) {
    if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
      c->blas = nullptr;
      return fail(c, SEP_ERR_HIP, "rocblas_create_handle failed");
    }
    if (rocblas_set_stream(c->blas, c->st) != rocblas_status_success ||
        rocblas_set_atomics_mode(c->blas, rocblas_atomics_not_allowed) != rocblas_status_success)
      return fail(c, SEP_ERR_HIP, "rocblas_set_stream / atomics mode failed");
  }
  if (n > 0x7fffffffL) return fail(c, SEP_ERR_ARG, "dense solve of order %ld", n);
  return ensure(c, &c->dws, &c->dws_bytes, sizeof(double) * (2 * (size_t)n + 1));
}

// Eigen-decomposition of the symmetric n x n matrix A_d (FP64, divide and
// conquer): eigenvalues ascending into w (host), eigenvector j left in
// A_d[j * n .. j * n + n) (column-major).
static int dev_symeig(sep_ctx *c, long n, double *A_d, double *w, const char *what) {
  RCHK(dense_init(c, n));
  double *D = reinterpret_cast<double *>(c->dws), *E = D + n;
  rocblas_int *info_d = reinterpret_cast<rocblas_int *>(E + n);
  rocblas_status rs = rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper,
                                       (rocblas_int)n, A_d, (rocblas_int)n, D, E, info_d);
  if (rs != rocblas_status_success) return fail(c, SEP_ERR_HIP, "%s: rocsolver_dsyevd status %d", what, (int)rs);
  rocblas_int info = 0;
  HIPCHK(c, hipMemcpyAsync(w, D, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(&info, info_d, sizeof(info), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (info != 0) return fail(c, SEP_ERR_NUMERIC, "%s: eigensolver did not converge (%d)", what, (int)info);
  return SEP_OK;
}

// Singular values (descending, into S) and right singular vectors of the
// row-major m x m host matrix R on the device (rocSOLVER dgesvd):
// V[i * m + j] = component i of vector j.  The column-major view of R is
// R^T = V S U^T, so V is the left factor rocSOLVER returns for that view.
// Uses c->scr[0 .. 2 m^2).
static int dev_svd_right(sep_ctx *c, long m, const double *R, double *S, double *V) {
  RCHK(dense_init(c, m));
  RCHK(ensure_scr(c, 2 * (size_t)m * m));
  double *A_d = c->scr, *U_d = A_d + m * m;
  double *S_d = reinterpret_cast<double *>(c->dws), *E_d = S_d + m;
  rocblas_int *info_d = reinterpret_cast<rocblas_int *>(E_d + m);
  HIPCHK(c, hipMemcpyAsync(A_d, R, sizeof(double) * m * m, hipMemcpyHostToDevice, c->st));
  const rocblas_int n = (rocblas_int)m;
  rocblas_status rs = rocsolver_dgesvd(c->blas, rocblas_svect_singular, rocblas_svect_none, n, n,
                                       A_d, n, S_d, U_d, n, nullptr, 1, E_d, rocblas_outofplace,
                                       info_d);
  if (rs != rocblas_status_success) return fail(c, SEP_ERR_HIP, "ortho: rocsolver_dgesvd status %d", (int)rs);
  std::vector<double> u((size_t)m * m);
  rocblas_int info = 0;
  HIPCHK(c, hipMemcpyAsync(S, S_d, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(u.data(), U_d, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(&info, info_d, sizeof(info), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (info != 0) return fail(c, SEP_ERR_NUMERIC, "ortho: SVD did not converge (%d)", (int)info);
  for (long j = 0; j < m; j++)
    for (long i = 0; i < m; i++) V[i * m + j] = u[i + j * m];
  return SEP_OK;
}

// G_d (symmetric positive definite, row-major) <- R^-1 with R^T R = G, R
// upper triangular, zeros below the diagonal.  Column-major lower Cholesky
// L L^T = G is the row-major upper R = L^T; its column-major lower inverse
// L^-1 is the row-major upper R^-1.
static int dev_chol_inv(sep_ctx *c, long n, double *G_d, const char *what) {
  RCHK(dense_init(c, n));
  rocblas_int *info_d = reinterpret_cast<rocblas_int *>(c->dws);
  rocblas_int info[2] = {0, 0};
  rocblas_status rs = rocsolver_dpotrf(c->blas, rocblas_fill_lower, (rocblas_int)n, G_d,
                                       (rocblas_int)n, info_d);
  if (rs != rocblas_status_success) return fail(c, SEP_ERR_HIP, "%s: rocsolver_dpotrf status %d", what, (int)rs);
  HIPCHK(c, hipMemcpyAsync(&info[0], info_d, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (info[0] != 0) return fail(c, SEP_ERR_NUMERIC, "%s: Cholesky pivot %d not positive", what, (int)info[0]);
  rs = rocsolver_dtrtri(c->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, (rocblas_int)n, G_d,
                        (rocblas_int)n, info_d);
  if (rs != rocblas_status_success) return fail(c, SEP_ERR_HIP, "%s: rocsolver_dtrtri status %d", what, (int)rs);
  HIPCHK(c, hipMemcpyAsync(&info[1], info_d, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (info[1] != 0) return fail(c, SEP_ERR_NUMERIC, "%s: singular triangular factor (%d)", what, (int)info[1]);
  return zero_lower_dev(G_d, n, c->st) == SEP_OK ? SEP_OK : fail(c, SEP_ERR_HIP, "%s: zero_lower", what);
}

// C[i][j] = sum_k A[i*lda + k] B[j*ldb + k] (m x mb, row-major, device):
// the matrix-core Gram of sep_gemm.hip (f64 MFMA 16x16x4, deterministic
// split-K).  K = N for states.
static int sub_gram_k(sep_ctx *c, const double *A, long lda, long m, const double *B, long ldb,
                      long mb, long K, bool sym, double *C_d) {
  const long wl = gram_work_len(m, mb, K);
  RCHK(ensure(c, (void **)&c->gwork, &c->gwork_bytes, sizeof(double) * wl));
  const int rc = gram_dev<double>(A, lda, m, B, ldb, mb, K, sym, C_d, c->gwork, wl, c->st);
  return rc == SEP_OK ? SEP_OK : fail(c, rc, "Gram launch failed");
}
static int sub_gram(sep_ctx *c, const double *A, long m, const double *B, long mb, double *C_d) {
  return sub_gram_k(c, A, c->gd.N, m, B, c->gd.N, mb, c->gd.N, false, C_d);
}

// Symmetric C = A B^T, m x m (B = A for a Gram matrix, B = H A for a
// projected Hamiltonian, symmetric in exact arithmetic): only the 128-tiles
// on and above the diagonal are multiplied, those below take their mirror,
// and the whole is scaled.
static int sub_gram_sym(sep_ctx *c, const double *A, const double *B, long m, double scale,
                        double *C_d) {
  RCHK(sub_gram_k(c, A, c->gd.N, m, B, c->gd.N, m, c->gd.N, true, C_d));
  return symmetrize_dev(C_d, m, scale, c->st, 128) == SEP_OK ? SEP_OK
                                                             : fail(c, SEP_ERR_HIP, "symmetrize launch");
}

// Out[j][g] = sum_i M[i * ldm + j] A[i][g] (mo states out, Out != A): the
// matrix-core rotation of sep_gemm.hip.
static int sub_combine_k(sep_ctx *c, const double *A, long lda, long m, const double *M_d,
                         long ldm, long mo, long N, double *Out, long ldo) {
  const int rc = combine_dev<double>(A, lda, m, M_d, ldm, mo, N, Out, ldo, c->st);
  return rc == SEP_OK ? SEP_OK : fail(c, rc, "rotation launch failed");
}
static int sub_combine(sep_ctx *c, const double *A, long m, const double *M_d, long ldm, long mo,
                       double *Out) {
  return sub_combine_k(c, A, c->gd.N, m, M_d, ldm, mo, c->gd.N, Out, c->gd.N);
}

// Cholesky-QR in FP64 on the k resident states: G = A A^T, R = chol(G),
// R^-1 and A <- R^-T A, all on the device.  One pass leaves |A A^T - I| at
// about eps kappa(A)^2, so a second pass follows unless the first Gram
// matrix, scaled to unit diagonal, lies within 0.5 of the identity in
// Frobenius norm (then kappa(A)^2 < 3): true of a basis that is already
// orthonormal up to rounding, as the whitened states of the ortho stage are.
static int cholqr(sep_ctx *c, long k) {
  RCHK(ensure_states(c, k, 2));
  for (int pass = 0; pass < 2; pass++) {
    RCHK(ensure_scr(c, k * k));
    double *G_d = c->scr;
    RCHK(sub_gram_sym(c, c->states, c->states, k, 1.0, G_d));
    bool last = pass == 1;
    if (!last) {
      std::vector<double> g(k * k);
      HIPCHK(c, hipMemcpyAsync(g.data(), G_d, sizeof(double) * k * k, hipMemcpyDeviceToHost, c->st));
      HIPCHK(c, hipStreamSynchronize(c->st));
      double off = 0.0;
      for (long i = 0; i < k; i++)
        for (long j = 0; j < k; j++)
          if (i != j) off += g[i * k + j] * g[i * k + j] / (g[i * k + i] * g[j * k + j]);
      last = off < 0.25;  // false for NaN (a zero state): the factorisation reports it
    }
    RCHK(dev_chol_inv(c, k, G_d, "ortho: re-orthogonalisation"));
    RCHK(sub_combine(c, c->states, k, G_d, k, k, c->states2));
    std::swap(c->states, c->states2);
    std::swap(c->cap1, c->cap2);
    if (last) break;
  }
  return SEP_OK;
}

// --------------------------------------------------------------- ortho
extern "C" int sep_ortho(sep_ctx *c, long m, double cutoff, double *S, long *kept) {
  if (!c || m < 1 || m > c->cap1 || !kept) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  // 1. G = A A^T in double-double (resolves S_i/S_0 well below 1e-10)
  const long wl = gram_dd_work_len(m, N);
  RCHK(ensure_scr(c, wl + 2 * m * m));
  double *ghi_d = c->scr + wl, *glo_d = ghi_d + m * m;
  RCHK(gram_dd_dev(c->states, m, N, ghi_d, glo_d, c->scr, wl, c->st));
  std::vector<double> ghi(m * m), glo(m * m);
  HIPCHK(c, hipMemcpyAsync(ghi.data(), ghi_d, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(glo.data(), glo_d, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (long i = 1; i < m; i++)  // the lower triangle from the upper one
    for (long j = 0; j < i; j++) {
      ghi[i * m + j] = ghi[j * m + i];
      glo[i * m + j] = glo[j * m + i];
    }
  // 2. R = chol(G) in double-double; a shift at the double-double rounding
  //    level keeps exactly dependent states factorable
  double tr = 0.0;
  for (long i = 0; i < m; i++) tr += ghi[i * m + i];
  if (!(tr > 0.0)) return fail(c, SEP_ERR_NUMERIC, "ortho: all states are zero");
  std::vector<double> R(m * m);
  // the shift sits above the double-double rounding of G (~1e-27 tr) and is
  // removed again from the singular values below: (R^T R = G + shift I)
  double shift = 1.0e-24 * tr;
  int st = sep_la_chol_dd(m, ghi.data(), glo.data(), shift, R.data());
  for (int tries = 0; st && tries < 4; tries++) {
    shift *= 100.0;
    st = sep_la_chol_dd(m, ghi.data(), glo.data(), shift, R.data());
  }
  if (st) return fail(c, SEP_ERR_NUMERIC, "ortho: Cholesky pivot %d not positive", st);
  // 3. SVD of R (device): A = (A V S^-1) S V^T
  std::vector<double> V(m * m), s(m);
  RCHK(dev_svd_right(c, m, R.data(), s.data(), V.data()));
  for (long i = 0; i < m; i++) {
    double v = s[i] * s[i] - shift;
    s[i] = v > 0.0 ? sqrt(v) : 0.0;
  }
  long k = m;
  for (long i = 0; i < m; i++)
    if (s[i] / s[0] < cutoff) {  // filter/ortho.c:41-46
      k = i;
      break;
    }
  if (S) memcpy(S, s.data(), sizeof(double) * m);
  // 4. U_A = A (V_k S_k^-1), then two Cholesky-QR passes for orthonormality
  std::vector<double> M(m * k);
  for (long i = 0; i < m; i++)
    for (long j = 0; j < k; j++) M[i * k + j] = V[i * m + j] / (s[j] > 0.0 ? s[j] : 1.0);
  RCHK(ensure_states(c, m, 2));
  RCHK(ensure_scr(c, wl + 2 * m * m));
  double *M_d = c->scr + wl;
  HIPCHK(c, hipMemcpyAsync(M_d, M.data(), sizeof(double) * m * k, hipMemcpyHostToDevice, c->st));
  RCHK(sub_combine(c, c->states, m, M_d, k, k, c->states2));
  std::swap(c->states, c->states2);
  std::swap(c->cap1, c->cap2);
  RCHK(cholqr(c, k));
  *kept = k;
  return normalize_dev(c, c->states, k);  // filter/main.c:112
}

// --------------------------------------------------------- Rayleigh-Ritz
extern "C" int sep_rayleigh_ritz(sep_ctx *c, long m, double *eval) {
  if (!c || !eval || m < 1 || m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  RCHK(ensure_states(c, m, 2));
  RCHK(apply_states(c, c->states, c->states2, m, nullptr, nullptr));  // H psi_i
  RCHK(ensure_scr(c, 2 * m * m));
  double *P_d = c->scr, *M_d = P_d + m * m;
  RCHK(sub_gram(c, c->states, m, c->states2, m, P_d));
  RCHK(symmetrize_dev(P_d, m, c->dv, c->st));  // H_ij = dv <psi_i|H|psi_j>
  RCHK(dev_symeig(c, m, P_d, eval, "Rayleigh-Ritz"));
  RCHK(eigvec_cols_dev(P_d, m, m, nullptr, false, M_d, c->st));
  RCHK(sub_combine(c, c->states, m, M_d, m, m, c->states2));
  std::swap(c->states, c->states2);
  std::swap(c->cap1, c->cap2);
  return normalize_dev(c, c->states, m);  // filter/main.c:122
}

// Orthonormal basis of the span of the m resident states from their FP64
// Gram matrix: G = W L W^T on the device, directions with sqrt(l_i / l_0) <
// floor dropped, basis A W_k L_k^-1/2 left in c->states, then (polish)
// Cholesky-QR in FP64 to orthonormality at rounding level.  Without polish
// the basis is orthonormal only to the accuracy of G's small eigenpairs, for
// a caller that solves the generalized problem on it (rr_polished).  S gets
// sqrt(l) descending.
static int eig_basis(sep_ctx *c, long m, double floor, double *S, long *kept, bool polish) {
  const long N = c->gd.N;
  RCHK(ensure_states(c, m, 2));
  const double *A = c->states;
  RCHK(ensure_scr(c, 2 * m * m + m));
  double *G_d = c->scr, *M_d = G_d + m * m, *sc_d = M_d + m * m;
  RCHK(sub_gram_sym(c, A, A, m, 1.0, G_d));
  std::vector<double> w(m);
  RCHK(dev_symeig(c, m, G_d, w.data(), "ortho"));
  std::vector<double> s(m);
  for (long i = 0; i < m; i++) s[i] = sqrt(w[m - 1 - i] > 0.0 ? w[m - 1 - i] : 0.0);
  if (!(s[0] > 0.0)) return fail(c, SEP_ERR_NUMERIC, "ortho: all states are zero");
  long k = m;
  for (long i = 0; i < m; i++)
    if (s[i] / s[0] < floor) {
      k = i;
      break;
    }
  if (S) memcpy(S, s.data(), sizeof(double) * m);
  // basis j = A v_(m-1-j) / s_j: eigenvectors in descending order, scaled
  std::vector<double> sc(k);
  for (long j = 0; j < k; j++) sc[j] = 1.0 / s[j];
  HIPCHK(c, hipMemcpyAsync(sc_d, sc.data(), sizeof(double) * k, hipMemcpyHostToDevice, c->st));
  RCHK(eigvec_cols_dev(G_d, m, k, sc_d, true, M_d, c->st));
  RCHK(sub_combine(c, c->states, m, M_d, k, k, c->states2));
  std::swap(c->states, c->states2);
  std::swap(c->cap1, c->cap2);
  HIPCHK(c, hipStreamSynchronize(c->st));  // sc outlives the async copy
  if (polish) RCHK(cholqr(c, k));
  *kept = k;
  return normalize_dev(c, c->states, k);
}

// Rayleigh-Ritz on the span of k resident states Y that are orthonormal
// only approximately (an unpolished eig_basis): the generalized problem
// P x = e G x with G = Y Y^T and P = Y H Y^T, through G = R^T R and the
// standard problem R^-T P R^-1 = Q E Q^T.  The states become (R^-1 Q)^T Y --
// orthonormal to rounding, as after a Cholesky-QR pass followed by
// sep_rayleigh_ritz, with one combination of the k states instead of two.
// Energies ascending into eval; states normalized as filter/main.c:122.
static int rr_polished(sep_ctx *c, long k, double *eval) {
  RCHK(ensure_states(c, k, 2));
  RCHK(apply_states(c, c->states, c->states2, k, nullptr, nullptr));  // H y_i
  RCHK(ensure_scr(c, 3 * k * k));
  double *U_d = c->scr, *P_d = U_d + k * k, *W_d = P_d + k * k;
  RCHK(sub_gram_sym(c, c->states, c->states, k, 1.0, U_d));
  RCHK(sub_gram_sym(c, c->states, c->states2, k, 1.0, P_d));
  RCHK(dev_chol_inv(c, k, U_d, "Rayleigh-Ritz: Gram factor"));  // U = R^-1, row-major
  // row-major k x k products on the matrix cores:
  // W = P U = P^T U, then P' = U^T W = R^-T P R^-1 over P_d
  RCHK(sub_combine_k(c, U_d, k, k, P_d, k, k, k, W_d, k));
  RCHK(sub_combine_k(c, W_d, k, k, U_d, k, k, k, P_d, k));
  RCHK(symmetrize_dev(P_d, k, 1.0, c->st));
  RCHK(dev_symeig(c, k, P_d, eval, "Rayleigh-Ritz"));  // Q column-major in P_d = Q^T row-major
  // coefficients C = U Q = U (Q^T)^T, a Gram product of the rows of U and Q^T
  RCHK(sub_gram_k(c, U_d, k, k, P_d, k, k, k, false, W_d));
  RCHK(sub_combine(c, c->states, k, W_d, k, k, c->states2));
  std::swap(c->states, c->states2);
  std::swap(c->cap1, c->cap2);
  return normalize_dev(c, c->states, k);
}

// rr_polished that also leaves H applied to the Ritz vectors: the basis'
// H Y rotated with the same coefficients (no further H application), both
// blocks scaled to the normalized Ritz vectors.  After it c->states holds
// the k Ritz vectors and c->states2 (rows 0..k) their H psi; sige gets
// sqrt(|H psi|^2 dv - e^2).
static int rr_polished_h(sep_ctx *c, long k, double *eval, double *sige) {
  const long N = c->gd.N;
  RCHK(ensure_states(c, 2 * k, 2));
  RCHK(apply_states(c, c->states, c->states2, k, nullptr, nullptr));  // H y_i
  RCHK(ensure_scr(c, 3 * k * k));
  double *U_d = c->scr, *P_d = U_d + k * k, *W_d = P_d + k * k;
  RCHK(sub_gram_sym(c, c->states, c->states, k, 1.0, U_d));
  RCHK(sub_gram_sym(c, c->states, c->states2, k, 1.0, P_d));
  RCHK(dev_chol_inv(c, k, U_d, "Rayleigh-Ritz: Gram factor"));
  RCHK(sub_combine_k(c, U_d, k, k, P_d, k, k, k, W_d, k));
  RCHK(sub_combine_k(c, W_d, k, k, U_d, k, k, k, P_d, k));
  RCHK(symmetrize_dev(P_d, k, 1.0, c->st));
  RCHK(dev_symeig(c, k, P_d, eval, "Rayleigh-Ritz"));
  RCHK(sub_gram_k(c, U_d, k, k, P_d, k, k, k, false, W_d));
  // Ritz vectors into states2 rows k..2k, their H psi over the basis in states
  RCHK(sub_combine(c, c->states, k, W_d, k, k, c->states2 + k * N));
  RCHK(sub_combine(c, c->states2, k, W_d, k, k, c->states));
  HIPCHK(c, hipMemcpyAsync(c->states2, c->states2 + k * N, sizeof(double) * k * N,
                           hipMemcpyDeviceToDevice, c->st));
  std::swap(c->states, c->states2);  // states: Ritz vectors, states2: H psi
  std::swap(c->cap1, c->cap2);
  // normalization (filter/main.c:122) applied to both blocks, then sigma
  const long nb = (N + 16383) / 16384;
  RCHK(ensure_scr(c, (size_t)nb * k + 3 * (size_t)k));
  double *out = c->scr + nb * k;
  RCHK(dots_dev(c->states, N, c->states, N, k, N, out, c->scr, nb * k, c->st));
  std::vector<double> nrm(k), h2(k);
  HIPCHK(c, hipMemcpyAsync(nrm.data(), out, sizeof(double) * k, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (long i = 0; i < k; i++) nrm[i] = nrm[i] != 0.0 ? 1.0 / sqrt(nrm[i] * c->dv) : 0.0;
  HIPCHK(c, hipMemcpyAsync(out + k, nrm.data(), sizeof(double) * k, hipMemcpyHostToDevice, c->st));
  RCHK(scale_rows_dev(c->states, N, k, N, out + k, c->st));
  RCHK(scale_rows_dev(c->states2, N, k, N, out + k, c->st));
  RCHK(dots_dev(c->states2, N, c->states2, N, k, N, out + 2 * k, c->scr, nb * k, c->st));
  HIPCHK(c, hipMemcpyAsync(h2.data(), out + 2 * k, sizeof(double) * k, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (long i = 0; i < k; i++) sige[i] = sqrt(fabs(h2[i] * c->dv - eval[i] * eval[i]));
  return SEP_OK;
}

extern "C" int sep_ortho_fp32(sep_ctx *c, long m, double cutoff, double *S, long *kept) {
  if (!c || m < 1 || m > c->cap1 || !kept || !(cutoff >= 0.0)) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  return eig_basis(c, m, cutoff > SEP_FP32_SVD_FLOOR ? cutoff : SEP_FP32_SVD_FLOOR, S, kept, true);
}

// dst = P src for m real states (FP64), P = 1 / (T + shift) with T the
// clamped kinetic energy of the H.psi passes: forward FFT, multiply by
// 1 / (N (T + shift)), inverse FFT.  Pass B takes the preconditioner factor
// in place of T; a zero potential leaves pass C's epilogue the transform.
static int precond_states(sep_ctx *c, const double *src, double *dst, long m, double shift) {
  const long N = c->gd.N;
  if (!c->zpot) {
    RCHK(ensure(c, &c->zpot, &c->zpot_bytes, sizeof(double) * N));
    RCHK(zero_dev(c->zpot, sizeof(double) * N, c->st));
  }
  GridDev gd = c->gd;
  gd.pnum = 1.0 / ((double)N * (double)N);
  gd.pden = shift / (double)N;
  const double *zp = reinterpret_cast<const double *>(c->zpot);
  auto run = [&](VecView<double> in, VecView<double> out, long nv) -> int {
    RCHK(ensure(c, &c->tb, &c->tb_bytes, sizeof(cx<double>) * (size_t)nv * N));
    long nblk = 0;
    int rc = hpsi_apply<double>(gd, in, out, reinterpret_cast<cx<double> *>(c->tb), zp, (int)nv,
                                1.0, 0.0, nullptr, &nblk, c->st);
    return rc == SEP_OK ? SEP_OK : fail(c, rc, "preconditioner launch failed");
  };
  const long npair = m / 2;
  for (long p0 = 0; p0 < npair; p0 += 64) {  // two states per complex vector
    const long np = npair - p0 < 64 ? npair - p0 : 64;
    double *s0 = const_cast<double *>(src) + 2 * p0 * N, *d0 = dst + 2 * p0 * N;
    RCHK(run(VecView<double>{s0, s0 + N, nullptr, 2 * N}, VecView<double>{d0, d0 + N, nullptr, 2 * N},
             np));
  }
  if (m & 1) {
    double *s0 = const_cast<double *>(src) + (m - 1) * N, *d0 = dst + (m - 1) * N;
    RCHK(run(VecView<double>{s0, nullptr, nullptr, N}, VecView<double>{d0, nullptr, nullptr, N}, 1));
  }
  return SEP_OK;
}

// One FP64 correction pass on m resident Ritz pairs (values e): residuals
// r_i = H psi_i - e_i psi_i, preconditioned t_i = P r_i (P as in
// precond_states, shift SEP_REFINE_SHIFT), Rayleigh-Ritz in span{Psi, T}
// (orthonormalised from the FP64 Gram with singular-value floor 1e-7), and
// the m Ritz pairs of smallest energy variance kept, ascending in energy.
// One correction pass on the m resident Ritz pairs (eval ascending): a
// Rayleigh-Ritz on [Psi, P R] (2m states), or [Psi, P R, Psi_prev] (3m: the
// previous iterate of the same pairs, the momentum term of LOBPCG) when
// prev is given, of whose pairs m are kept: with lo > hi those of smallest
// variance, else those inside [lo, hi] first (smallest variance among
// them), then the rest by variance.  Kept pairs ascending in energy at the
// front of the state block; returns their count.
static int refine_pass(sep_ctx *c, long m, double *eval, double *sige, double lo, double hi,
                       long *kept, const double *prev = nullptr, bool have_h = false) {
  const long N = c->gd.N;
  const long nb = prev ? 3 : 2;  // blocks of the Rayleigh-Ritz basis
  // [Psi; H Psi] in states2 (2m rows), then r_i = H psi_i - e_i psi_i.
  // have_h: H Psi is already in states2 rows 0..m (the previous pass)
  RCHK(ensure_states(c, nb * m, 2));
  if (have_h)
    HIPCHK(c, hipMemcpyAsync(c->states2 + m * N, c->states2, sizeof(double) * m * N,
                             hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(c->states2, c->states, sizeof(double) * m * N, hipMemcpyDeviceToDevice,
                           c->st));
  if (!have_h) RCHK(apply_states(c, c->states, c->states2 + m * N, m, nullptr, nullptr));
  RCHK(ensure_scr(c, m));
  HIPCHK(c, hipMemcpyAsync(c->scr, eval, sizeof(double) * m, hipMemcpyHostToDevice, c->st));
  RCHK(sub_scaled_rows_dev(c->states2 + m * N, N, c->states, N, m, N, c->scr, c->st));
  std::swap(c->states, c->states2);
  std::swap(c->cap1, c->cap2);
  // t_i = P r_i into the second block of states2, copied back over r_i
  // (states2 is the former states buffer: grow it to nb m rows first)
  RCHK(ensure_states(c, nb * m, 2));
  RCHK(precond_states(c, c->states + m * N, c->states2 + m * N, m, SEP_REFINE_SHIFT));
  HIPCHK(c, hipMemcpyAsync(c->states + m * N, c->states2 + m * N, sizeof(double) * m * N,
                           hipMemcpyDeviceToDevice, c->st));
  RCHK(normalize_dev(c, c->states + m * N, m));
  if (prev)
    HIPCHK(c, hipMemcpyAsync(c->states + 2 * m * N, prev, sizeof(double) * m * N,
                             hipMemcpyDeviceToDevice, c->st));
  long k = 0;
  RCHK(eig_basis(c, nb * m, SEP_FP32_SVD_FLOOR, nullptr, &k, false));
  std::vector<double> ev(k), sg(k);
  RCHK(rr_polished_h(c, k, ev.data(), sg.data()));  // H psi of the pairs in states2
  std::vector<long> idx(k);
  for (long i = 0; i < k; i++) idx[i] = i;
  const bool win = lo <= hi;
  std::stable_sort(idx.begin(), idx.end(), [&](long a, long b) {
    if (win) {
      const bool ia = ev[a] >= lo && ev[a] <= hi, ib = ev[b] >= lo && ev[b] <= hi;
      if (ia != ib) return ia;
    }
    return sg[a] < sg[b];
  });
  const long keep = k < m ? k : m;
  idx.resize(keep);
  std::sort(idx.begin(), idx.end());  // Ritz values are ascending in index
  // gather the kept states to the front in place: idx ascends with
  // idx[j] >= j, so no source is overwritten before it is read
  for (long j = 0; j < keep; j++) {
    eval[j] = ev[idx[j]];
    sige[j] = sg[idx[j]];
    if (idx[j] != j) {
      HIPCHK(c, hipMemcpyAsync(c->states + j * N, c->states + idx[j] * N, sizeof(double) * N,
                               hipMemcpyDeviceToDevice, c->st));
      HIPCHK(c, hipMemcpyAsync(c->states2 + j * N, c->states2 + idx[j] * N, sizeof(double) * N,
                               hipMemcpyDeviceToDevice, c->st));
    }
  }
  for (long j = keep; j < m; j++) eval[j] = sige[j] = 0.0;
  HIPCHK(c, hipStreamSynchronize(c->st));
  *kept = keep;
  return SEP_OK;  // states2 rows 0..keep: H psi of the kept pairs
}

extern "C" int sep_refine(sep_ctx *c, long m, double *eval, double *sige) {
  if (!c || !eval || !sige || m < 1 || m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  long keep = 0;
  RCHK(refine_pass(c, m, eval, sige, 1.0, 0.0, &keep));
  return keep == m ? SEP_OK : fail(c, SEP_ERR_NUMERIC, "refine: basis of %ld < %ld states", keep, m);
}

// Small dense products of the window passes (dimensions up to a few
// hundred): C[i ldc + j] = sum_l A[i sai + l sal] B[l sbl + j sbj], the sum
// over l in ascending order; strides give the transposes.
__global__ void k_small_mm(const double *A, long sai, long sal, const double *B, long sbl,
                           long sbj, double *C, long ldc, long m, long n, long kk) {
  const long i = (long)blockIdx.y * 16 + threadIdx.y, j = (long)blockIdx.x * 16 + threadIdx.x;
  if (i >= m || j >= n) return;
  double acc = 0.0;
  for (long l = 0; l < kk; l++) acc = fma(A[i * sai + l * sal], B[l * sbl + j * sbj], acc);
  C[i * ldc + j] = acc;
}
static int small_mm(sep_ctx *c, const double *A, long sai, long sal, const double *B, long sbl,
                    long sbj, double *C, long ldc, long m, long n, long kk) {
  hipLaunchKernelGGL(k_small_mm, dim3((unsigned)((n + 15) / 16), (unsigned)((m + 15) / 16)),
                     dim3(16, 16), 0, c->st, A, sai, sal, B, sbl, sbj, C, ldc, m, n, kk);
  HIPCHK(c, hipGetLastError());
  return SEP_OK;
}
// H <- (H + H^T) / 2 for the r x r row-major H
__global__ void k_symmetrize(double *H, long r) {
  const long i = (long)blockIdx.y * 16 + threadIdx.y, j = (long)blockIdx.x * 16 + threadIdx.x;
  if (i >= r || j >= i) return;
  const double v = 0.5 * (H[i * r + j] + H[j * r + i]);
  H[i * r + j] = v;
  H[j * r + i] = v;
}
// s[j] = sum_i C[i][j] QC[i][j] over the nz rows of the nz x r row-major C, QC
__global__ void k_col_dots(const double *C, const double *QC, long nz, long r, double *s) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= r) return;
  double acc = 0.0;
  for (long i = 0; i < nz; i++) acc = fma(C[i * r + j], QC[i * r + j], acc);
  s[j] = acc;
}

// One correction pass of sep_refine_window on the k pairs at the front of
// the state block (X in states rows 0..k, H X in states2 rows 0..k, energies
// ev): Rayleigh-Ritz on Z = [X, T, X_prev] (T = P R, R = H X - e X, X_prev
// and H X_prev of the previous pass when prev is given) in coefficient
// space -- Gram matrices Z Z^T, Z (H Z)^T, (H Z)(H Z)^T (matrix cores), G =
// W L W^T reduced to the directions with l > floor^2 l_max, the reduced H
// diagonalised, sigma^2 = c^T (HZ HZ^T) c - e^2 -- and only the k kept pairs
// (inside [lo, hi] first, then by variance) rotated, with their H psi, both
// normalized as filter/main.c:122.  One H application per pass (on T).
// X and H X are rows base .. base + k of the two state buffers, which hold
// base + 4k rows: T and H T follow at rows k .. 2k, X_prev and H X_prev
// (the previous pass's X, written there by that pass when keep_prev) at
// 2k .. 3k, and rows 3k .. 4k are scratch.
static int window_pass(sep_ctx *c, long base, long k, double *ev, double *sg, double lo,
                       double hi, bool have_prev, bool keep_prev) {
  const long N = c->gd.N;
  const long nz = have_prev ? 3 * k : 2 * k;
  double *Z = c->states + base * N, *HZ = c->states2 + base * N;
  // R = H X - e X into rows 3k..4k; T = P R, normalized, into rows k..2k;
  // H T into HZ rows k..2k
  RCHK(ensure_scr(c, k));
  HIPCHK(c, hipMemcpyAsync(c->scr, ev, sizeof(double) * k, hipMemcpyHostToDevice, c->st));
  RCHK(diff_scaled_rows_dev(Z + 3 * k * N, N, HZ, N, Z, N, k, N, c->scr, c->st));
  RCHK(precond_states(c, Z + 3 * k * N, Z + k * N, k, SEP_REFINE_SHIFT));
  RCHK(normalize_dev(c, Z + k * N, k));
  RCHK(apply_states(c, Z + k * N, HZ + k * N, k, nullptr, nullptr));
  // device scratch: the three Gram matrices, then the coefficient-space
  // factors (r <= nz), all placed before any of them is used
  const size_t q2 = (size_t)nz * nz;
  RCHK(ensure_scr(c, 8 * q2 + (size_t)nz * k + (size_t)nz));
  double *G_d = c->scr, *P_d = G_d + q2, *Q_d = P_d + q2, *B_d = Q_d + q2, *PB_d = B_d + q2;
  double *H_d = PB_d + q2, *C_d = H_d + q2, *QC_d = C_d + q2, *M_d = QC_d + q2, *s_d = M_d + nz * k;
  RCHK(sub_gram_sym(c, Z, Z, nz, 1.0, G_d));
  RCHK(sub_gram_sym(c, Z, HZ, nz, 1.0, P_d));
  RCHK(sub_gram_sym(c, HZ, HZ, nz, 1.0, Q_d));
  std::vector<double> w(nz);
  RCHK(dev_symeig(c, nz, G_d, w.data(), "refine: Gram"));  // column j of G_d: vector j
  std::vector<double> W((size_t)nz * nz);
  HIPCHK(c, hipMemcpyAsync(W.data(), G_d, sizeof(double) * nz * nz, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  const double lmax = w[nz - 1], floor2 = SEP_FP32_SVD_FLOOR * SEP_FP32_SVD_FLOOR;
  std::vector<long> dir;
  for (long i = nz - 1; i >= 0; i--)
    if (w[i] > floor2 * lmax) dir.push_back(i);
  const long r = (long)dir.size();
  if (r < k) return fail(c, SEP_ERR_NUMERIC, "refine: basis of %ld < %ld states", r, k);
  // B = W_r L_r^-1/2 (nz x r), reduced H = B^T P B (r x r), on the device
  std::vector<double> B((size_t)nz * r);
  for (long j = 0; j < r; j++) {
    const double sc = 1.0 / sqrt(w[dir[j]]);
    for (long i = 0; i < nz; i++) B[i * r + j] = W[dir[j] * nz + i] * sc;
  }
  HIPCHK(c, hipMemcpyAsync(B_d, B.data(), sizeof(double) * nz * r, hipMemcpyHostToDevice, c->st));
  RCHK(small_mm(c, P_d, nz, 1, B_d, r, 1, PB_d, r, nz, r, nz));  // PB = P B
  RCHK(small_mm(c, B_d, 1, r, PB_d, r, 1, H_d, r, r, r, nz));    // H = B^T PB
  hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((r + 15) / 16), (unsigned)((r + 15) / 16)),
                     dim3(16, 16), 0, c->st, H_d, r);
  HIPCHK(c, hipGetLastError());
  std::vector<double> e(r);
  RCHK(dev_symeig(c, r, H_d, e.data(), "refine: reduced H"));  // column j of H_d: vector j
  // C = B V (nz x r): coefficients of the Ritz vectors over Z; their
  // variance c_j^T Q c_j - e_j^2 (V column-major: C[i][j] = sum_l B[i][l] V[j][l])
  RCHK(small_mm(c, B_d, r, 1, H_d, 1, r, C_d, r, nz, r, r));
  RCHK(small_mm(c, Q_d, nz, 1, C_d, r, 1, QC_d, r, nz, r, nz));
  hipLaunchKernelGGL(k_col_dots, dim3((unsigned)((r + 63) / 64)), dim3(64), 0, c->st, C_d, QC_d,
                     nz, r, s_d);
  HIPCHK(c, hipGetLastError());
  std::vector<double> C((size_t)nz * r), s(r);
  HIPCHK(c, hipMemcpyAsync(C.data(), C_d, sizeof(double) * nz * r, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(s.data(), s_d, sizeof(double) * r, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (long j = 0; j < r; j++) s[j] = sqrt(fabs(s[j] - e[j] * e[j]));
  std::vector<long> idx(r);
  for (long i = 0; i < r; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](long x, long y) {
    const bool ix = e[x] >= lo && e[x] <= hi, iy = e[y] >= lo && e[y] <= hi;
    if (ix != iy) return ix;
    return s[x] < s[y];
  });
  idx.resize(k);
  std::sort(idx.begin(), idx.end());
  // the kept Ritz vectors and their H psi, normalized (Ritz vectors are
  // orthonormal in the plain product: scale 1 / sqrt(dv))
  const double nrm = 1.0 / sqrt(c->dv);
  std::vector<double> M((size_t)nz * k);
  for (long i = 0; i < nz; i++)
    for (long j = 0; j < k; j++) M[i * k + j] = C[i * r + idx[j]] * nrm;
  for (long j = 0; j < k; j++) {
    ev[j] = e[idx[j]];
    sg[j] = s[idx[j]];
  }
  HIPCHK(c, hipMemcpyAsync(M_d, M.data(), sizeof(double) * nz * k, hipMemcpyHostToDevice, c->st));
  RCHK(sub_combine(c, Z, nz, M_d, k, k, Z + 3 * k * N));
  RCHK(sub_combine(c, HZ, nz, M_d, k, k, HZ + 3 * k * N));
  if (keep_prev) {  // this pass's X and H X: the next pass's X_prev
    HIPCHK(c, hipMemcpyAsync(Z + 2 * k * N, Z, sizeof(double) * k * N, hipMemcpyDeviceToDevice,
                             c->st));
    HIPCHK(c, hipMemcpyAsync(HZ + 2 * k * N, HZ, sizeof(double) * k * N,
                             hipMemcpyDeviceToDevice, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(Z, Z + 3 * k * N, sizeof(double) * k * N, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(c, hipMemcpyAsync(HZ, HZ + 3 * k * N, sizeof(double) * k * N, hipMemcpyDeviceToDevice,
                           c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));  // M outlives the copy
  return SEP_OK;
}

extern "C" int sep_refine_window(sep_ctx *c, long m, double *eval, double *sige, double lo,
                                 double hi, int passes, long *kw) {
  if (!c || !eval || !sige || !kw || m < 1 || m > c->cap1 || !(lo < hi) || passes < 1)
    return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  std::vector<long> idx;  // the window's pairs, ascending
  for (long i = 0; i < m; i++)
    if (eval[i] >= lo && eval[i] <= hi) idx.push_back(i);
  const long k = (long)idx.size();
  *kw = k;
  if (k == 0) return SEP_OK;
  // the passes work on a copy of the window's pairs in rows m .. m + 4k of
  // both state buffers (X, P R, X_prev and an output block), so the other
  // resident pairs stay as they are
  RCHK(grow_states(c, m + 4 * k, 1, m));
  RCHK(grow_states(c, m + 4 * k, 2, 0));
  double *X = c->states + m * N;
  std::vector<double> ev(k), sg(k);
  for (long j = 0; j < k; j++) {
    ev[j] = eval[idx[j]];
    HIPCHK(c, hipMemcpyAsync(X + j * N, c->states + idx[j] * N, sizeof(double) * N,
                             hipMemcpyDeviceToDevice, c->st));
  }
  RCHK(apply_states(c, X, c->states2 + m * N, k, nullptr, nullptr));  // H X
  for (int p = 0; p < passes; p++)
    RCHK(window_pass(c, m, k, ev.data(), sg.data(), lo, hi, p > 0, p + 1 < passes));
  // the refined pairs (ascending) back into the window's positions
  for (long j = 0; j < k; j++) {
    HIPCHK(c, hipMemcpyAsync(c->states + idx[j] * N, X + j * N, sizeof(double) * N,
                             hipMemcpyDeviceToDevice, c->st));
    eval[idx[j]] = ev[j];
    sige[idx[j]] = sg[j];
  }
  // a refined level that crossed a neighbour outside the window: all pairs
  // back in ascending order (states staged in the second buffer)
  if (!std::is_sorted(eval, eval + m)) {
    std::vector<long> ord(m);
    for (long i = 0; i < m; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) { return eval[a] < eval[b]; });
    for (long i = 0; i < m; i++)
      HIPCHK(c, hipMemcpyAsync(c->states2 + i * N, c->states + ord[i] * N, sizeof(double) * N,
                               hipMemcpyDeviceToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->states, c->states2, sizeof(double) * m * N,
                             hipMemcpyDeviceToDevice, c->st));
    std::vector<double> e2(m), s2(m);
    for (long i = 0; i < m; i++) {
      e2[i] = eval[ord[i]];
      s2[i] = sige[ord[i]];
    }
    std::copy(e2.begin(), e2.end(), eval);
    std::copy(s2.begin(), s2.end(), sige);
  }
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

// ------------------------------------------------- Hellmann-Feynman adjoint
extern "C" int sep_hf_adjoint(sep_ctx *c, long npairs, const long *pairs, const double *weights,
                              long natom, const double *xyz, const int *type, const double *scale,
                              int ntype, const long *npot, const double *vr, const double *pot,
                              long stride, const double *dr, double *dR, double *dscale,
                              double *dtab) {
  if (!c || npairs < 1 || !pairs || natom < 1 || !xyz || !type || ntype < 1 || !npot || !vr ||
      !pot || !dr || stride < 2)
    return SEP_ERR_ARG;
  for (long p = 0; p < 2 * npairs; p++)
    if (pairs[p] < 0 || pairs[p] >= c->cap1)
      return fail(c, SEP_ERR_ARG, "pair %ld: state %ld not resident", p / 2, pairs[p]);
  long maxnpot = 0;
  for (long a = 0; a < natom; a++) {
    if (type[a] < 0 || type[a] >= ntype || npot[type[a]] < 2 || npot[type[a]] > stride ||
        !(dr[type[a]] > 0.0))
      return fail(c, SEP_ERR_ARG, "atom %ld: bad type or table", a);
    if (npot[type[a]] > maxnpot) maxnpot = npot[type[a]];
  }
  if (hf_pairs_per_pass(maxnpot) == 0)
    return fail(c, SEP_ERR_ARG, "tables of %ld rows exceed the LDS histogram", maxnpot);
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  const long np = weights ? 1 : npairs;
  // box points per atom -> slices per atom (>= ~4096 points per block)
  double rmax = 0.0;
  for (int t = 0; t < ntype; t++)
    if (npot[t] >= 2 && (double)(npot[t] - 1) * dr[t] > rmax) rmax = (double)(npot[t] - 1) * dr[t];
  const double box = (2.0 * rmax / c->dx + 1.0) * (2.0 * rmax / c->dy + 1.0) * (2.0 * rmax / c->dz + 1.0);
  long nsplit = (long)(box / 4096.0) + 1;
  if (nsplit > 64) nsplit = 64;
  const int pb = hf_pairs_per_pass(maxnpot);
  DevBuf xyz_d, type_d, npot_d, vr_d, pot_d, dr_d, scale_d, pairs_d, w_d, rho_d, work_d, htab_d,
      dR_d, ds_d;
  RCHK(dev_upload(c, xyz_d, xyz, sizeof(double) * 3 * natom));
  RCHK(dev_upload(c, type_d, type, sizeof(int) * natom));
  RCHK(dev_upload(c, npot_d, npot, sizeof(long) * ntype));
  RCHK(dev_upload(c, vr_d, vr, sizeof(double) * ntype * stride));
  RCHK(dev_upload(c, pot_d, pot, sizeof(double) * ntype * stride));
  RCHK(dev_upload(c, dr_d, dr, sizeof(double) * ntype));
  if (scale) RCHK(dev_upload(c, scale_d, scale, sizeof(double) * natom));
  RCHK(dev_upload(c, pairs_d, pairs, sizeof(long) * 2 * npairs));
  if (weights) RCHK(dev_upload(c, w_d, weights, sizeof(double) * npairs));
  RCHK(dev_alloc(c, rho_d, sizeof(double) * np * N));
  RCHK(dev_alloc(c, work_d, sizeof(double) * natom * nsplit * pb * 4));
  RCHK(dev_alloc(c, htab_d, sizeof(double) * np * ntype * stride));
  RCHK(dev_alloc(c, dR_d, sizeof(double) * np * natom * 3));
  RCHK(dev_alloc(c, ds_d, sizeof(double) * np * natom));
  HIPCHK(c, hipMemsetAsync(htab_d.p, 0, sizeof(double) * np * ntype * stride, c->st));
  RCHK(pair_density_dev(c->states, N, pairs_d.as<long>(), w_d.as<double>(), (int)npairs,
                        rho_d.as<double>(), c->st));
  int rc = hf_adjoint_dev(c->gd, c->xmin, c->ymin, c->zmin, c->dx, c->dy, c->dz, c->dv, (int)natom,
                          xyz_d.as<double>(), type_d.as<int>(), scale_d.as<double>(),
                          npot_d.as<long>(), maxnpot, vr_d.as<double>(), pot_d.as<double>(), stride,
                          dr_d.as<double>(), ntype, rho_d.as<double>(), (int)np, (int)nsplit,
                          c->interp, work_d.as<double>(), htab_d.as<double>(), dR_d.as<double>(),
                          ds_d.as<double>(), c->st);
  if (rc != SEP_OK) return fail(c, rc, "adjoint kernel launch failed");
  if (dR)
    HIPCHK(c, hipMemcpyAsync(dR, dR_d.p, sizeof(double) * np * natom * 3, hipMemcpyDeviceToHost, c->st));
  if (dscale)
    HIPCHK(c, hipMemcpyAsync(dscale, ds_d.p, sizeof(double) * np * natom, hipMemcpyDeviceToHost, c->st));
  if (dtab)
    HIPCHK(c, hipMemcpyAsync(dtab, htab_d.p, sizeof(double) * np * ntype * stride,
                             hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (dtab)
    for (long i = 0; i < np * ntype * stride; i++) dtab[i] *= c->dv;
  return SEP_OK;
}

extern "C" int sep_filter_dev(sep_ctx *c, const void *psi0_dev, long ns, const double *an,
                              const double *zn, long nc, long ms, double vmin, double dE,
                              int precision) {
  if (!c || !psi0_dev || ns < 1 || !an || !zn || nc < 1 || ms < 1 || !(dE > 0.0))
    return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  const double *p = static_cast<const double *>(psi0_dev);
  if (precision == SEP_PREC_FP64) return filter_impl<double>(c, p, true, ns, an, zn, nc, ms, vmin, dE);
  if (precision == SEP_PREC_FP32) return filter_impl<float>(c, p, true, ns, an, zn, nc, ms, vmin, dE);
  return fail(c, SEP_ERR_ARG, "unknown precision %d", precision);
}

// ------------------------------------------------------------ measurement
__global__ void k_fill_random(double *x, long n, unsigned long seed) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    unsigned long z = (unsigned long)i * 0x9E3779B97F4A7C15UL + seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9UL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBUL;
    z ^= z >> 31;
    x[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
}

template <typename R>
static int bench_impl(sep_ctx *c, long nvec, long ms, long iters, double *kms) {
  const long N = c->gd.N;
  const long S2 = 2 * nvec, K = 64;
  // ring of K slots of 2*nvec real states (the filter's layout) + accumulators
  RCHK(ensure(c, &c->ring, &c->ring_bytes, sizeof(R) * (size_t)K * S2 * N));
  RCHK(ensure_states(c, S2 * ms, 1));
  RCHK(ensure(c, &c->tb, &c->tb_bytes, sizeof(cx<R>) * (size_t)nvec * N));
  R *ring = reinterpret_cast<R *>(c->ring);
  if (sizeof(R) == 8) {
    hipLaunchKernelGGL(k_fill_random, dim3(4096), dim3(256), 0, c->st, (double *)ring, 2 * S2 * N, 7UL);
  } else {
    hipLaunchKernelGGL(k_fill_random, dim3(4096), dim3(256), 0, c->st, c->states, S2 * N, 7UL);
    RCHK(convert_d2f(c->states, (float *)ring, S2 * N, c->st));
    RCHK(convert_d2f(c->states, (float *)ring + S2 * N, S2 * N, c->st));
  }
  HIPCHK(c, hipMemsetAsync(c->states, 0, sizeof(double) * S2 * ms * N, c->st));
  std::vector<double> an(ms * K);
  for (long i = 0; i < ms * K; i++) an[i] = 1e-3 * (double)((i * 37) % 101) / 101.0;
  double *an_d = nullptr;
  HIPCHK(c, hipMalloc(&an_d, sizeof(double) * ms * K));
  HIPCHK(c, hipMemcpy(an_d, an.data(), sizeof(double) * ms * K, hipMemcpyHostToDevice));
  float *cf_d = nullptr;  // the complex64 filter's packed coefficients (values immaterial)
  if (sizeof(R) == 4) {
    std::vector<float> cf(accum_pack_len(K, (int)ms));
    accum_pack_f32(an.data(), K, (int)ms, cf.data());
    HIPCHK(c, hipMalloc(&cf_d, sizeof(float) * cf.size()));
    HIPCHK(c, hipMemcpy(cf_d, cf.data(), sizeof(float) * cf.size(), hipMemcpyHostToDevice));
  }
  VecView<R> in{ring, ring + N, nullptr, 2 * N};
  VecView<R> out{ring + S2 * N, ring + S2 * N + N, nullptr, 2 * N};
  cx<R> *tb = reinterpret_cast<cx<R> *>(c->tb);
  // the filter loop's split step (pass B, then pass C fused with the next
  // pass A) where the grid supports it: ev[4..7] around pass B and the fused kernel
  const bool fused = hpsi_fused_supported<R>(c->gd);
  hipEvent_t ev[8];
  for (int i = 0; i < 8; i++) HIPCHK(c, hipEventCreate(&ev[i]));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int rc = SEP_OK;
  for (long it = -1; it < iters && rc == SEP_OK; it++) {  // it = -1: warm-up
    rc = hpsi_apply<R>(c->gd, in, out, tb, pot_of<R>(c), (int)nvec, 0.3, 1.7, nullptr, nullptr,
                       c->st, ev);
    if (rc == SEP_OK) {
      rc = SEP_ERR_UNSUPPORTED;
      if (cf_d)
        rc = accum_mfma_packed_dev(reinterpret_cast<const float *>(ring), S2 * N, N, (int)K, cf_d,
                                   K, 0, (int)ms, c->states, (int)S2, N, c->st);
      if (rc == SEP_ERR_UNSUPPORTED)
        rc = accum_mfma_dev<R>(ring, S2 * N, N, (int)K, an_d, 0, (int)ms, c->states, (int)S2, N,
                               c->st);
    }
    hipEventRecord(ev[4], c->st);
    if (rc == SEP_OK && fused)
      rc = hpsi_step<R>(c->gd, in, out, tb, pot_of<R>(c), (int)nvec, 0.3, 1.7, true, c->st, ev + 4);
    if (hipEventSynchronize(ev[fused ? 7 : 4]) != hipSuccess) rc = SEP_ERR_HIP;
    if (it >= 0 && rc == SEP_OK) {
      float t;
      for (int k = 0; k < 4; k++) {
        hipEventElapsedTime(&t, ev[k], ev[k + 1]);
        acc[k] += t;
      }
      hipEventElapsedTime(&t, ev[0], ev[3]);
      acc[4] += t;
      if (fused) {  // the whole split step: spectral passes + fused kernel
        hipEventElapsedTime(&t, ev[5], ev[7]);
        acc[5] += t;
      }
    }
  }
  for (int i = 0; i < 8; i++) hipEventDestroy(ev[i]);
  hipFree(an_d);
  if (cf_d) hipFree(cf_d);
  if (rc != SEP_OK) return fail(c, rc, "bench launch failed");
  for (int k = 0; k < 6; k++) kms[k] = acc[k] / (double)(iters > 0 ? iters : 1);
  return SEP_OK;
}

// Plain copy of n 16-byte words (the HBM bandwidth probe): four loads in
// flight per lane, then four stores; the unrolled loop runs only while all
// four indices are below n, the tail one word at a time.  NT: nontemporal
// (streaming) loads and stores.
typedef float word16 __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ word16 ld16(const word16 *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT>
__device__ __forceinline__ void st16(word16 v, word16 *p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
template <bool NT>
__global__ void __launch_bounds__(256) k_copy16(const word16 *__restrict__ a,
                                                word16 *__restrict__ b, long n) {
  const long stride = (long)gridDim.x * 256;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const word16 x0 = ld16<NT>(a + i), x1 = ld16<NT>(a + i + stride);
    const word16 x2 = ld16<NT>(a + i + 2 * stride), x3 = ld16<NT>(a + i + 3 * stride);
    st16<NT>(x0, b + i);
    st16<NT>(x1, b + i + stride);
    st16<NT>(x2, b + i + 2 * stride);
    st16<NT>(x3, b + i + 3 * stride);
  }
  for (; i < n; i += stride) st16<NT>(ld16<NT>(a + i), b + i);
}

extern "C" int sep_bench_copy(sep_ctx *c, long nbytes, long iters, double *gbs) {
  if (!c || !gbs || nbytes < (1L << 20) || iters < 1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long n = nbytes / 16;
  DevBuf a, b;
  RCHK(dev_alloc(c, a, 16 * (size_t)n));
  RCHK(dev_alloc(c, b, 16 * (size_t)n));
  RCHK(zero_dev(a.p, 16 * (size_t)n, c->st));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) hipEventDestroy(e0);
    return fail(c, SEP_ERR_HIP, "copy probe: event creation failed");
  }
  double best = 0.0;
  int rc = SEP_OK;
  for (int nt = 0; nt < 2; nt++)
  for (int blocks : {1024, 2048, 4096, 8192})
    for (long it = 0; it <= iters && rc == SEP_OK; it++) {  // it 0: warm-up
      hipEventRecord(e0, c->st);
      hipLaunchKernelGGL(nt ? k_copy16<true> : k_copy16<false>, dim3(blocks), dim3(256), 0,
                         c->st, static_cast<const word16 *>(a.p), static_cast<word16 *>(b.p), n);
      hipEventRecord(e1, c->st);
      float ms = 0.0f;
      if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(c, SEP_ERR_HIP, "copy probe launch failed");
      } else if (it > 0 && ms > 0.0f) {
        best = std::max(best, 2.0 * 16.0 * (double)n / (ms * 1e-3) / 1e9);
      }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *gbs = best;
  return rc;
}

extern "C" int sep_bench_hpsi(sep_ctx *c, int precision, long nvec, long ms, long iters,
                              double *kernel_ms) {
  if (!c || nvec < 1 || ms < 1 || iters < 1 || !kernel_ms) return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  RingBusy busy;  // the benchmark works in the ring
  busy.add(c);
  if (precision == SEP_PREC_FP64) return bench_impl<double>(c, nvec, ms, iters, kernel_ms);
  if (precision == SEP_PREC_FP32) return bench_impl<float>(c, nvec, ms, iters, kernel_ms);
  return fail(c, SEP_ERR_ARG, "unknown precision %d", precision);
}

// Kernel-family choice per grid and precision: time one Newton step's H.psi
// in the filter loop's split form (spectral passes + fused kernel; A + B + C
// where a family has no split form) with the plane kernels and with the row +
// strided-line kernels on nvec resident random vectors (bench_impl); keep the
// faster.
extern "C" int sep_autotune(sep_ctx *c, int precision, long nvec, double *ms_plane,
                            double *ms_lines) {
  if (!c || nvec < 1 || (precision != SEP_PREC_FP64 && precision != SEP_PREC_FP32))
    return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  double k[6];
  double tp = -1.0, tl = -1.0;
  const bool had = c->gd.force_lines[precision];
  if (plane_supported(c->gd.nx, c->gd.ny, precision)) {
    c->gd.force_lines[precision] = false;
    const int rc = precision == SEP_PREC_FP64 ? bench_impl<double>(c, nvec, 1, 5, k)
                                              : bench_impl<float>(c, nvec, 1, 5, k);
    if (rc != SEP_OK) {
      c->gd.force_lines[precision] = had;
      return rc;
    }
    // the filter loop's split step (spectral passes + fused kernel)
    tp = k[5] > 0.0 && !c->gd.no_fuse ? k[5] : k[4];
  }
  c->gd.force_lines[precision] = true;
  const int rc = precision == SEP_PREC_FP64 ? bench_impl<double>(c, nvec, 1, 5, k)
                                            : bench_impl<float>(c, nvec, 1, 5, k);
  if (rc != SEP_OK) {
    c->gd.force_lines[precision] = had;
    return rc;
  }
  tl = k[5] > 0.0 && !c->gd.no_fuse ? k[5] : k[4];
  c->gd.force_lines[precision] = !(tp > 0.0 && tp <= tl);
  if (ms_plane) *ms_plane = tp;
  if (ms_lines) *ms_lines = tl;
  return SEP_OK;
}

// One short filter (nc Newton steps of 2 nvec random states, one target) on
// the context's buffers, timed with events on its stream: the autotuner's
// measure, which sees the lane split and the ring layout of a real filter.
template <typename R>
static int tune_filter_ms(sep_ctx *c, long nvec, long nc, double *ms_out) {
  const long ns = 2 * nvec, N = c->gd.N;
  RCHK(ensure_states(c, ns, 2));
  hipLaunchKernelGGL(k_fill_random, dim3(4096), dim3(256), 0, c->st, c->states2, ns * N, 11UL);
  HIPCHK(c, hipGetLastError());
  // values immaterial to the timing; kept bounded: |beta| <= 3, dE_1 = 1e-3
  std::vector<double> an((size_t)nc), zn((size_t)nc);
  for (long j = 0; j < nc; j++) {
    an[j] = 1.0 / (double)(j + 1);
    zn[j] = std::cos(3.141592653589793 * (double)j / (double)nc);
  }
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return fail(c, SEP_ERR_HIP, "autotune: event creation failed");
  }
  int rc = SEP_OK;
  float best = -1.0f;
  for (int it = 0; it < 3 && rc == SEP_OK; it++) {  // it = 0: warm-up
    hipEventRecord(e0, c->st);
    rc = filter_impl<R>(c, c->states2, true, ns, an.data(), zn.data(), nc, 1, -2.0, 4000.0);
    hipEventRecord(e1, c->st);
    if (rc == SEP_OK && hipEventSynchronize(e1) != hipSuccess) rc = SEP_ERR_HIP;
    float t = 0.0f;
    if (rc == SEP_OK && it > 0 && hipEventElapsedTime(&t, e0, e1) == hipSuccess)
      best = best < 0.0f || t < best ? t : best;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != SEP_OK) return rc;
  *ms_out = (double)best;
  return SEP_OK;
}

extern "C" int sep_autotune_ex(sep_ctx *c, int precision, long nvec, double *ms, int *family,
                               int *lanes) {
  if (!c || nvec < 1 || (precision != SEP_PREC_FP64 && precision != SEP_PREC_FP32))
    return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  const bool f_lines = c->gd.force_lines[precision];
  const int f_form = c->gd.fused_form[precision];
  const int l0 = c->lanes;
  const bool timing = c->timing;
  c->timing = false;  // live timing turns lanes off
  const long nc = 24;
  double t_fam[SEP_HPSI_NFAMILY];
  double best = -1.0;
  int bf = -1, bl = 1, rc = SEP_OK;
  for (int f = 0; f < SEP_HPSI_NFAMILY && rc == SEP_OK; f++) {
    t_fam[f] = -1.0;
    if (sep_set_hpsi_family(c, precision, f) != SEP_OK) {
      c->err.clear();  // a family this grid does not have is no error here
      continue;
    }
    for (int L = 1; L <= SEP_MAX_LANES && L <= nvec && rc == SEP_OK; L *= 2) {
      c->lanes = L;
      double t = -1.0;
      rc = precision == SEP_PREC_FP64 ? tune_filter_ms<double>(c, nvec, nc, &t)
                                      : tune_filter_ms<float>(c, nvec, nc, &t);
      if (rc != SEP_OK || t <= 0.0) continue;
      if (t_fam[f] < 0.0 || t < t_fam[f]) t_fam[f] = t;
      if (best < 0.0 || t < best) {
        best = t;
        bf = f;
        bl = L;
      }
    }
  }
  c->timing = timing;
  if (rc != SEP_OK || bf < 0) {
    c->gd.force_lines[precision] = f_lines;
    c->gd.fused_form[precision] = f_form;
    c->lanes = l0;
    return rc != SEP_OK ? rc : fail(c, SEP_ERR_UNSUPPORTED, "autotune: no family ran");
  }
  sep_set_hpsi_family(c, precision, bf);
  c->lanes = bl;
  if (ms)
    for (int f = 0; f < SEP_HPSI_NFAMILY; f++) ms[f] = t_fam[f] > 0.0 ? t_fam[f] / nc : -1.0;
  if (family) *family = bf;
  if (lanes) *lanes = bl;
  return SEP_OK;
}

extern "C" int sep_set_hpsi_family(sep_ctx *c, int precision, int family) {
  if (!c || (precision != SEP_PREC_FP64 && precision != SEP_PREC_FP32)) return SEP_ERR_ARG;
  switch (family) {
    case SEP_HPSI_PLANE:
    case SEP_HPSI_PLANE_CHUNKED:
    case SEP_HPSI_PLANE_SWIZZLED: {
      if (!plane_supported(c->gd.nx, c->gd.ny, precision))
        return fail(c, SEP_ERR_UNSUPPORTED, "no plane kernels for %d x %d", c->gd.nx, c->gd.ny);
      const int form = family == SEP_HPSI_PLANE_CHUNKED    ? SEP_FUSED_CHUNKED
                       : family == SEP_HPSI_PLANE_SWIZZLED ? SEP_FUSED_SWIZZLED
                                                           : SEP_FUSED_PLANE;
      if (!plane_fused_form_supported(c->gd.nx, precision, form))
        return fail(c, SEP_ERR_UNSUPPORTED, "no such fused kernel for %d x %d in this precision",
                    c->gd.nx, c->gd.ny);
      c->gd.force_lines[precision] = false;
      c->gd.fused_form[precision] = form;
      return SEP_OK;
    }
    case SEP_HPSI_LINES:
      c->gd.force_lines[precision] = true;
      c->gd.fused_form[precision] = SEP_FUSED_PLANE;
      return SEP_OK;
  }
  return fail(c, SEP_ERR_ARG, "unknown H.psi family %d", family);
}

extern "C" int sep_get_hpsi_family(sep_ctx *c, int precision, int *family) {
  if (!c || !family || (precision != SEP_PREC_FP64 && precision != SEP_PREC_FP32))
    return SEP_ERR_ARG;
  const bool plane = !c->gd.force_lines[precision] &&
                     plane_supported(c->gd.nx, c->gd.ny, precision);
  const int form = c->gd.fused_form[precision];
  *family = !plane                       ? SEP_HPSI_LINES
            : form == SEP_FUSED_CHUNKED  ? SEP_HPSI_PLANE_CHUNKED
            : form == SEP_FUSED_SWIZZLED ? SEP_HPSI_PLANE_SWIZZLED
                                         : SEP_HPSI_PLANE;
  return SEP_OK;
}

extern "C" int sep_set_timing(sep_ctx *c, int on) {
  if (!c) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (on && !c->tev[0])
    for (int i = 0; i < 6; i++) HIPCHK(c, hipEventCreate(&c->tev[i]));
  c->timing = on != 0;
  for (int q = 0; q < 4; q++) c->tsum[q] = 0.0;
  c->tcount = c->tacc_count = 0;
  return SEP_OK;
}

extern "C" int sep_trim(sep_ctx *c) {
  if (!c) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->st));  // lanes have joined into st
  if (c->ring) HIPCHK(c, hipFree(c->ring));
  c->ring = nullptr;
  c->ring_bytes = 0;
  if (c->gwork) HIPCHK(c, hipFree(c->gwork));
  c->gwork = nullptr;
  c->gwork_bytes = 0;
  return SEP_OK;
}

extern "C" int sep_set_interp(sep_ctx *c, int interp) {
  if (!c || (interp != SEP_INTERP_LINEAR && interp != SEP_INTERP_CUBIC)) return SEP_ERR_ARG;
  c->interp = interp;
  return SEP_OK;
}

extern "C" int sep_set_filter_lanes(sep_ctx *c, int lanes) {
  if (!c || lanes < 0 || lanes > SEP_MAX_LANES) return SEP_ERR_ARG;
  c->lanes = lanes;
  return SEP_OK;
}

extern "C" int sep_get_timing(sep_ctx *c, double *avg_ms, long *samples) {
  if (!c || !avg_ms) return SEP_ERR_ARG;
  for (int q = 0; q < 3; q++) avg_ms[q] = c->tcount ? c->tsum[q] / (double)c->tcount : 0.0;
  avg_ms[3] = c->tacc_count ? c->tsum[3] / (double)c->tacc_count : 0.0;
  if (samples) {
    samples[0] = c->tcount;
    samples[1] = c->tacc_count;
  }
  return SEP_OK;
}

// ------------------------------------------------- dense device contractions
extern "C" int sep_gram_dev(sep_ctx *c, int precision, const void *A, long m, const void *B,
                            long mb, double *C) {
  if (!c || !A || !B || !C || m < 1 || mb < 1) return SEP_ERR_ARG;
  if (precision != SEP_PREC_FP64 && precision != SEP_PREC_FP32)
    return fail(c, SEP_ERR_ARG, "unknown precision %d", precision);
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N;
  RCHK(ensure_scr(c, m * mb));
  double *C_d = c->scr;
  if (precision == SEP_PREC_FP64) {
    RCHK(sub_gram(c, (const double *)A, m, (const double *)B, mb, C_d));
  } else {
    const long wl = gram_work_len(m, mb, N);
    RCHK(ensure(c, (void **)&c->gwork, &c->gwork_bytes, sizeof(double) * wl));
    RCHK(gram_dev<float>((const float *)A, N, m, (const float *)B, N, mb, N, false, C_d, c->gwork,
                         wl, c->st));
  }
  HIPCHK(c, hipMemcpyAsync(C, C_d, sizeof(double) * m * mb, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

extern "C" int sep_rotate_dev(sep_ctx *c, int precision, const void *A, long m, const double *M,
                              long mo, void *Out) {
  if (!c || !A || !M || !Out || m < 1 || mo < 1 || A == Out) return SEP_ERR_ARG;
  if (precision != SEP_PREC_FP64 && precision != SEP_PREC_FP32)
    return fail(c, SEP_ERR_ARG, "unknown precision %d", precision);
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(ensure_scr(c, m * mo));
  HIPCHK(c, hipMemcpyAsync(c->scr, M, sizeof(double) * m * mo, hipMemcpyHostToDevice, c->st));
  const long N = c->gd.N;
  if (precision == SEP_PREC_FP64)
    RCHK(sub_combine(c, (const double *)A, m, c->scr, mo, mo, (double *)Out));
  else
    RCHK(combine_dev<float>((const float *)A, N, m, c->scr, mo, mo, N, (float *)Out, N, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

// ================================================================ spinors
// The reference's spin-orbit filter (spinors/filterSpinors): FP64 spinors
// stored as four real rows of N (up.re, up.im, dn.re, dn.im), the kinetic
// and local parts on the two complex components through the H.psi passes,
// V_so through sep_spinor.hip.  The complex subspace stage runs on the real
// matrix-core kernels: with X the spinors flattened to 4N reals and iX the
// spinors times i, <x_s|y_t> = X_s.Y_t + i (iX_s).Y_t, and a complex
// combination sum_s M_st x_s is the real combination of [X | iX] with the
// stacked coefficients [Re M; Im M].

static int so_contrib_buf(sep_ctx *c, long nsp) {
  return ensure(c, (void **)&c->so_contrib, &c->so_contrib_bytes,
                sizeof(double) * 4 * (size_t)(c->so.nent > 0 ? c->so.nent : 1) * nsp);
}

// dst = H src on m spinors (dst != src), 32 spinors (64 complex vectors) per pass
static int apply_spinors(sep_ctx *c, const double *src, double *dst, long m) {
  const long N = c->gd.N, n4 = 4 * N;
  for (long s0 = 0; s0 < m; s0 += 32) {
    const long nb = m - s0 < 32 ? m - s0 : 32;
    double *x = const_cast<double *>(src) + s0 * n4, *y = dst + s0 * n4;
    VecView<double> in{x, x + N, nullptr, 2 * N}, out{y, y + N, nullptr, 2 * N};
    RCHK(apply_views<double>(c, in, out, (int)(2 * nb), 1.0, 0.0, nullptr));
    if (c->so.natom > 0) {
      RCHK(so_contrib_buf(c, nb));
      if (so_apply_dev(c->so, x, N, n4, y, N, n4, nb, 1.0, c->so_contrib, c->st) != SEP_OK)
        return fail(c, SEP_ERR_HIP, "spin-orbit launch failed");
    }
  }
  return SEP_OK;
}

// out[i] = x_i . y_i over 4N reals (device out, m values)
static int spinor_dots(sep_ctx *c, const double *x, const double *y, long m, double *out_d) {
  const long n4 = 4 * c->gd.N, nb = (n4 + 16383) / 16384;
  RCHK(ensure_scr(c, (size_t)nb * m + 4 * (size_t)m));
  return dots_dev(x, n4, y, n4, m, n4, out_d, c->scr, nb * m, c->st) == SEP_OK
             ? SEP_OK
             : fail(c, SEP_ERR_HIP, "dot launch failed");
}

// normalize_all for spinors (spinors/filterSpinors/norm.c:40-62)
static int normalize_spinors(sep_ctx *c, double *x, long m) {
  if (m <= 0) return SEP_OK;
  const long n4 = 4 * c->gd.N, nb = (n4 + 16383) / 16384;
  RCHK(ensure_scr(c, (size_t)nb * m + 2 * (size_t)m));
  double *out = c->scr + nb * m;
  RCHK(spinor_dots(c, x, x, m, out));
  std::vector<double> nrm(m);
  HIPCHK(c, hipMemcpyAsync(nrm.data(), out, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (long i = 0; i < m; i++) nrm[i] = nrm[i] != 0.0 ? 1.0 / sqrt(nrm[i] * c->dv) : 0.0;
  HIPCHK(c, hipMemcpyAsync(out + m, nrm.data(), sizeof(double) * m, hipMemcpyHostToDevice, c->st));
  RCHK(scale_rows_dev(x, n4, m, n4, out + m, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

// Eigen-decomposition of the Hermitian n x n host matrix A (column-major
// complex, A[r + c n]) on the device (rocSOLVER zheevd): eigenvalues
// ascending into w, eigenvectors back into A (column j = vector j).
static int dev_heev(sep_ctx *c, long n, std::vector<rocblas_double_complex> &A, double *w,
                    const char *what) {
  RCHK(dense_init(c, n));
  RCHK(ensure_scr(c, 2 * (size_t)n * n));
  rocblas_double_complex *A_d = reinterpret_cast<rocblas_double_complex *>(c->scr);
  double *D = reinterpret_cast<double *>(c->dws), *E = D + n;
  rocblas_int *info_d = reinterpret_cast<rocblas_int *>(E + n);
  HIPCHK(c, hipMemcpyAsync(A_d, A.data(), sizeof(rocblas_double_complex) * n * n,
                           hipMemcpyHostToDevice, c->st));
  rocblas_status rs = rocsolver_zheevd(c->blas, rocblas_evect_original, rocblas_fill_upper,
                                       (rocblas_int)n, A_d, (rocblas_int)n, D, E, info_d);
  if (rs != rocblas_status_success)
    return fail(c, SEP_ERR_HIP, "%s: rocsolver_zheevd status %d", what, (int)rs);
  rocblas_int info = 0;
  HIPCHK(c, hipMemcpyAsync(w, D, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(A.data(), A_d, sizeof(rocblas_double_complex) * n * n,
                           hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(&info, info_d, sizeof(info), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  if (info != 0) return fail(c, SEP_ERR_NUMERIC, "%s: eigensolver did not converge (%d)", what, (int)info);
  return SEP_OK;
}

// Complex m x m matrix A(s, t) = (Re + i Im)(s, t) * scale from two
// row-major real device matrices, Hermitian part, column-major.
static int complex_from_parts(sep_ctx *c, const double *re_d, const double *im_d, long m,
                              double scale, std::vector<rocblas_double_complex> &A) {
  std::vector<double> r((size_t)m * m), im((size_t)m * m);
  HIPCHK(c, hipMemcpyAsync(r.data(), re_d, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(im.data(), im_d, sizeof(double) * m * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  A.assign((size_t)m * m, rocblas_double_complex(0.0, 0.0));
  for (long s = 0; s < m; s++)
    for (long t = 0; t < m; t++) {
      const double ar = 0.5 * (r[s * m + t] + r[t * m + s]) * scale;
      const double ai = 0.5 * (im[s * m + t] - im[t * m + s]) * scale;
      A[s + t * m] = rocblas_double_complex(ar, ai);
    }
  return SEP_OK;
}

// Out (k spinors) = sum_s x_s Z(s, j) for the complex coefficients Z
// (column-major m x k, scaled per column by sc[j]): one real combination of
// the 2m rows [X | iX] at XX.
static int combine_spinors(sep_ctx *c, const double *XX, long m,
                           const std::vector<rocblas_double_complex> &Z, long ldz,
                           const std::vector<long> &cols, const std::vector<double> &sc,
                           double *Out) {
  const long k = (long)cols.size(), n4 = 4 * c->gd.N;
  std::vector<double> M((size_t)2 * m * k);
  for (long s = 0; s < m; s++)
    for (long j = 0; j < k; j++) {
      const rocblas_double_complex z = Z[s + cols[j] * ldz];
      M[s * k + j] = std::real(z) * sc[j];
      M[(m + s) * k + j] = std::imag(z) * sc[j];
    }
  RCHK(ensure_scr(c, (size_t)2 * m * k));
  HIPCHK(c, hipMemcpyAsync(c->scr, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice, c->st));
  RCHK(sub_combine_k(c, XX, n4, 2 * m, c->scr, k, k, n4, Out, n4));
  HIPCHK(c, hipStreamSynchronize(c->st));  // M outlives the copy
  return SEP_OK;
}

// [X | iX] of the m resident spinors into c->states2 (2m spinors)
static int stack_x_ix(sep_ctx *c, long m) {
  const long n4 = 4 * c->gd.N;
  HIPCHK(c, hipMemcpyAsync(c->states2, c->states, sizeof(double) * m * n4, hipMemcpyDeviceToDevice,
                           c->st));
  return spinor_times_i_dev(c->states, c->states2 + m * n4, c->gd.N, m, c->st) == SEP_OK
             ? SEP_OK
             : fail(c, SEP_ERR_HIP, "spinor times i launch failed");
}

extern "C" {

int sep_so_setup(sep_ctx *c, long natom, const double *xyz, const double *vso, long *nentries) {
  if (!c || natom < 0 || (natom > 0 && (!xyz || !vso))) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->st));
  SoHost h;
  const double d[3] = {c->dx, c->dy, c->dz}, x0[3] = {c->xmin, c->ymin, c->zmin};
  RCHK(so_build_host(c->gd.nx, c->gd.ny, c->gd.nz, d, x0, natom, xyz, vso, h));
  for (void *p : {(void *)c->so.off, (void *)c->so.jxyz, (void *)c->so.geo, (void *)c->so.pidx,
                  (void *)c->so.poff, (void *)c->so.pent, (void *)c->so.kmat, (void *)c->so.koff,
                  (void *)c->so.pmap})
    if (p) hipFree(p);
  c->so = SoLists();
  SoLists &s = c->so;
  // sigma = dx (spinors/filterSpinors/init.c:96-97), hamiltonian.c:98
  const double sigma = c->dx;
  s.c = 0.5 / sqrt(6.28318530717958647692 * sigma * sigma) * c->dv;
  s.s1 = sqrt(0.5) / sigma;
  s.nent = (long)h.jxyz.size();
  s.npts = (long)h.pidx.size();
  s.maxn = h.maxn;
  if (nentries) *nentries = s.nent;
  if (s.nent == 0) return SEP_OK;
  s.natom = (int)(h.off.size() - 1);
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    HIPCHK(c, hipMalloc(dst, bytes));
    HIPCHK(c, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return SEP_OK;
  };
  RCHK(up((void **)&s.off, h.off.data(), sizeof(long) * h.off.size()));
  RCHK(up((void **)&s.jxyz, h.jxyz.data(), sizeof(int) * h.jxyz.size()));
  RCHK(up((void **)&s.geo, h.geo.data(), sizeof(double) * h.geo.size()));
  RCHK(up((void **)&s.pidx, h.pidx.data(), sizeof(int) * h.pidx.size()));
  RCHK(up((void **)&s.poff, h.poff.data(), sizeof(long) * h.poff.size()));
  RCHK(up((void **)&s.pent, h.pent.data(), sizeof(int) * h.pent.size()));
  if (so_kmat_build(s, h.off, c->st) != SEP_OK) return fail(c, SEP_ERR_HIP, "so_setup: K matrices");
  if (so_pmap_build(s, c->gd.N, c->st) != SEP_OK) return fail(c, SEP_ERR_HIP, "so_setup: point map");
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

int sep_apply_h_spinor_dev(sep_ctx *c, const void *in_dev, void *out_dev, long nspinor) {
  if (!c || !in_dev || !out_dev || nspinor < 1 || in_dev == out_dev) return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(apply_spinors(c, static_cast<const double *>(in_dev), static_cast<double *>(out_dev),
                     nspinor));
  HIPCHK(c, hipStreamSynchronize(c->st));
  return SEP_OK;
}

int sep_energy_range_spinor(sep_ctx *c, double *vmin, double *dE, long *iters) {
  if (!c || !vmin || !dE) return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  const long N = c->gd.N, n4 = 4 * N;
  sep_rng st;
  sep_rng_reset(&st);
  long idum = -874917403L;  // energy.c:58
  std::vector<double> h((size_t)n4), comp((size_t)N);
  auto init = [&]() {  // init_psi on the up, then the down component (energy.c:63-65)
    std::fill(h.begin(), h.end(), 0.0);
    for (int q = 0; q < 2; q++) {
      sep_init_psi(&st, &idum, N, c->dv, comp.data());
      std::copy(comp.begin(), comp.end(), h.begin() + 2 * q * N);
    }
  };
  double *buf = nullptr;
  HIPCHK(c, hipMalloc(&buf, sizeof(double) * 2 * n4));
  double *x = buf, *hx = buf + n4;
  int rc = SEP_OK;
  double e[2] = {0.0, 0.0};
  long its[2] = {0, 0};
  const long nb = (n4 + 16383) / 16384;
  for (int pass = 0; pass < 2 && rc == SEP_OK; pass++) {
    init();
    if (hipMemcpy(x, h.data(), sizeof(double) * n4, hipMemcpyHostToDevice) != hipSuccess) {
      rc = fail(c, SEP_ERR_HIP, "energy range: upload failed");
      break;
    }
    if ((rc = apply_spinors(c, x, hx, 1)) != SEP_OK) break;
    double ene = 0.1, old = 0.0;  // energy.c:68-69, 83-84
    long i = 0;
    for (; fabs((ene - old) / ene) > 1.0e-6 && i < 500; i++) {
      // Emin: x <- x - tau H x; Emax: x <- H x; then normalized; H of the
      // normalized iterate serves the energy and the next step
      if (pass == 0) rc = axpby_dev(-0.025, hx, 1.0, x, n4, c->st);
      else if (hipMemcpyAsync(x, hx, sizeof(double) * n4, hipMemcpyDeviceToDevice, c->st) != hipSuccess)
        rc = SEP_ERR_HIP;
      if (rc != SEP_OK) break;
      if ((rc = normalize_spinors(c, x, 1)) != SEP_OK) break;
      if ((rc = apply_spinors(c, x, hx, 1)) != SEP_OK) break;
      if ((rc = ensure_scr(c, (size_t)nb + 4)) != SEP_OK) break;
      if ((rc = spinor_dots(c, x, hx, 1, c->scr + nb)) != SEP_OK) break;
      double dot = 0.0;
      if (hipMemcpy(&dot, c->scr + nb, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        rc = SEP_ERR_HIP;
        break;
      }
      old = ene;
      ene = dot * c->dv;
    }
    e[pass] = ene;
    its[pass] = i;
  }
  hipFree(buf);
  if (rc != SEP_OK) return rc;
  const double emax = 1.2 * e[1], emin = e[0] - 0.2 * fabs(e[0]);  // energy.c:106-107
  *vmin = emin;
  *dE = emax - emin;
  if (iters) {
    iters[0] = its[0];
    iters[1] = its[1];
  }
  return SEP_OK;
}

static int filter_spinor(sep_ctx *c, const double *psi0, bool on_device, long nspinor,
                         const double *an, const double *zn, long nc, long ms, double vmin,
                         double dE) {
  if (!c || !psi0 || nspinor < 1 || !an || !zn || nc < 1 || ms < 1 || !(dE > 0.0))
    return SEP_ERR_ARG;
  if (!c->has_pot) return fail(c, SEP_ERR_ARG, "potential not set");
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(filter_impl<double>(c, psi0, on_device, 4 * nspinor, an, zn, nc, ms, vmin, dE, true));
  const long N = c->gd.N, m = nspinor * ms;
  RCHK(ensure_states(c, 4 * m, 2));
  if (spinor_gather_rows_dev(c->states, c->states2, N, nspinor, ms, c->st) != SEP_OK)
    return fail(c, SEP_ERR_HIP, "spinor reorder launch failed");
  std::swap(c->states, c->states2);
  std::swap(c->cap1, c->cap2);
  return normalize_spinors(c, c->states, m);  // filter.c:29
}

int sep_filter_spinor_dev(sep_ctx *c, const void *psi0_dev, long nspinor, const double *an,
                          const double *zn, long nc, long ms, double vmin, double dE) {
  return filter_spinor(c, static_cast<const double *>(psi0_dev), true, nspinor, an, zn, nc, ms,
                       vmin, dE);
}

int sep_filter_spinor(sep_ctx *c, const double *psi0, long nspinor, const double *an,
                      const double *zn, long nc, long ms, double vmin, double dE) {
  return filter_spinor(c, psi0, false, nspinor, an, zn, nc, ms, vmin, dE);
}

int sep_energy_spinor(sep_ctx *c, long m, double *ene) {
  if (!c || !ene || m < 1 || 4 * m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(ensure_states(c, 4 * m, 2));
  RCHK(apply_spinors(c, c->states, c->states2, m));
  RCHK(ensure_scr(c, (size_t)m * ((4 * c->gd.N + 16383) / 16384) + 4 * (size_t)m));
  std::vector<double> d(m);
  double *out = c->scr + m * ((4 * c->gd.N + 16383) / 16384) + 2 * m;
  RCHK(spinor_dots(c, c->states, c->states2, m, out));
  HIPCHK(c, hipMemcpyAsync(d.data(), out, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  for (long i = 0; i < m; i++) ene[i] = d[i] * c->dv;  // energy.c:40-44
  return SEP_OK;
}

int sep_sigma_spinor(sep_ctx *c, long m, double *sige) {
  if (!c || !sige || m < 1 || 4 * m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  RCHK(ensure_states(c, 4 * m, 2));
  RCHK(apply_spinors(c, c->states, c->states2, m));
  const long nb = (4 * c->gd.N + 16383) / 16384;
  RCHK(ensure_scr(c, (size_t)m * nb + 4 * (size_t)m));
  std::vector<double> e(m), h2(m);
  double *out = c->scr + m * nb + 2 * m;
  RCHK(spinor_dots(c, c->states, c->states2, m, out));
  HIPCHK(c, hipMemcpyAsync(e.data(), out, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  RCHK(spinor_dots(c, c->states2, c->states2, m, out));
  HIPCHK(c, hipMemcpyAsync(h2.data(), out, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipStreamSynchronize(c->st));
  // <psi|H^2|psi> = |H psi|^2 for Hermitian H (energy.c:117-147)
  for (long i = 0; i < m; i++) {
    const double ev = e[i] * c->dv;
    sige[i] = sqrt(fabs(h2[i] * c->dv - ev * ev));
  }
  return SEP_OK;
}

int sep_ortho_spinor(sep_ctx *c, long m, double cutoff, double *S, long *kept) {
  if (!c || !kept || m < 1 || 4 * m > c->cap1 || !(cutoff >= 0.0)) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long n4 = 4 * c->gd.N;
  const double floor = cutoff > SEP_SPINOR_SVD_FLOOR ? cutoff : SEP_SPINOR_SVD_FLOOR;
  long k = m;
  for (int pass = 0; pass < 2; pass++) {
    // pass 0: basis of the span (portho, ortho.c:28-40); pass 1 on that
    // basis: the same map once more, to orthonormality at rounding level
    RCHK(ensure_states(c, 8 * k, 2));
    RCHK(stack_x_ix(c, k));
    double *G_d = nullptr;
    HIPCHK(c, dev_malloc(c, (void **)&G_d, sizeof(double) * 2 * k * k));
    int rc = sub_gram_k(c, c->states, n4, k, c->states, n4, k, n4, false, G_d);
    if (rc == SEP_OK)
      rc = sub_gram_k(c, c->states2 + k * n4, n4, k, c->states, n4, k, n4, false, G_d + k * k);
    std::vector<rocblas_double_complex> A;
    if (rc == SEP_OK) rc = complex_from_parts(c, G_d, G_d + k * k, k, 1.0, A);
    hipFree(G_d);
    RCHK(rc);
    std::vector<double> w(k);
    RCHK(dev_heev(c, k, A, w.data(), "spinor ortho"));
    std::vector<double> s(k);
    for (long i = 0; i < k; i++) s[i] = sqrt(w[k - 1 - i] > 0.0 ? w[k - 1 - i] : 0.0);
    if (!(s[0] > 0.0)) return fail(c, SEP_ERR_NUMERIC, "spinor ortho: all spinors are zero");
    long kk = k;
    if (pass == 0) {
      for (long i = 0; i < k; i++)
        if (s[i] / s[0] < floor) {  // ortho.c:34-39
          kk = i;
          break;
        }
      if (S) memcpy(S, s.data(), sizeof(double) * k);
    }
    std::vector<long> cols(kk);
    std::vector<double> sc(kk);
    for (long j = 0; j < kk; j++) {
      cols[j] = k - 1 - j;
      sc[j] = 1.0 / s[j];
    }
    RCHK(combine_spinors(c, c->states2, k, A, k, cols, sc, c->states));
    k = kk;
  }
  *kept = k;
  return normalize_spinors(c, c->states, k);  // main.c:137
}

int sep_rayleigh_ritz_spinor(sep_ctx *c, long m, double *eval) {
  if (!c || !eval || m < 1 || 4 * m > c->cap1) return SEP_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long n4 = 4 * c->gd.N;
  // states2: [X | iX | H X]
  RCHK(ensure_states(c, 12 * m, 2));
  RCHK(stack_x_ix(c, m));
  double *HX = c->states2 + 2 * m * n4;
  RCHK(apply_spinors(c, c->states, HX, m));
  double *P_d = nullptr;
  HIPCHK(c, dev_malloc(c, (void **)&P_d, sizeof(double) * 2 * m * m));
  // P(s, t) = <x_s|H x_t> dv (Hmat.c:30-40)
  int rc = sub_gram_k(c, c->states, n4, m, HX, n4, m, n4, false, P_d);
  if (rc == SEP_OK) rc = sub_gram_k(c, c->states2 + m * n4, n4, m, HX, n4, m, n4, false, P_d + m * m);
  std::vector<rocblas_double_complex> A;
  if (rc == SEP_OK) rc = complex_from_parts(c, P_d, P_d + m * m, m, c->dv, A);
  hipFree(P_d);
  RCHK(rc);
  RCHK(dev_heev(c, m, A, eval, "spinor Rayleigh-Ritz"));  // Hmat.c:45
  std::vector<long> cols(m);
  std::vector<double> sc(m, 1.0);
  for (long j = 0; j < m; j++) cols[j] = j;
  RCHK(combine_spinors(c, c->states2, m, A, m, cols, sc, c->states));  // Hmat.c:49-66
  return normalize_spinors(c, c->states, m);  // main.c:146
}

}  // extern "C"
