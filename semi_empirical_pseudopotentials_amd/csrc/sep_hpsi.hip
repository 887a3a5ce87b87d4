// sep_hpsi.hip -- the H.psi operator of the reference (filter/hamiltonian.c:11-44,
// filter/hnorm.c:11-39) as hand-written CDNA4 kernels.
//
// One application on a batch of complex vectors v (in the filter each complex
// vector carries two real states, re and im, because H is real) is:
//   pass A  xy-forward FFT of every z-plane              psi  -> Tbuf
//   pass B  z-forward FFT, * min(k^2/2, Ekin)/N, z-back   Tbuf -> Tbuf (in place)
//   pass C  xy-backward FFT of every z-plane, then the fused epilogue
//           out = alpha (T psi + V psi) - beta psi      Tbuf, psi, V -> out
// alpha = 1, beta = 0 is hamiltonian(); alpha = dE_1 = 4/dE,
// beta = 2 + z_{j-1} + Vmin dE_1 is the Newton step of hnorm().
// Passes A and C keep a whole plane in LDS when it fits (160 KB on gfx950);
// otherwise each is split into a row kernel and a strided-line kernel.
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "sep_internal.h"

namespace sep {

template <typename R>
struct Epi {
  const R *pot;   // V on the grid
  double alpha, beta;
  double *part;   // per-block partial sums (4 per block) or nullptr
  // spinor filter (FP64 whole-plane fused kernel only): psi_j += so_alpha *
  // (V_so psi_{j-1}) at the grid points in a spin-orbit list, so_map[g] the
  // point's row of so_g (4 doubles per spinor: up re, im, dn re, im) or -1
  const int *so_map = nullptr;
  const double *so_g = nullptr;
  long so_nsp = 0;
  double so_alpha = 0.0;
};

// block reduction of 4 doubles; writes part[blk*4 + q]
template <int T>
__device__ __forceinline__ void block_sum4(double s[4], double *part, long blk) {
  __shared__ double red[4][T / 64];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    double x = s[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double x = 0.0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++) x += red[threadIdx.x][w];
    part[blk * 4 + threadIdx.x] = x;
  }
}

template <typename R>
__device__ __forceinline__ cx<R> epilogue(const Epi<R> &e, cx<R> t, cx<R> p, R V, double s[4]) {
  R a = (R)e.alpha, b = (R)e.beta;
  cx<R> o;
  o.re = a * (t.re + V * p.re) - b * p.re;
  o.im = a * (t.im + V * p.im) - b * p.im;
  if (e.part) {  // uniform: the FP64 sums only when a reduction was asked for
    s[0] += (double)p.re * (double)o.re;
    s[1] += (double)p.im * (double)o.im;
    s[2] += (double)o.re * (double)o.re;
    s[3] += (double)o.im * (double)o.im;
  }
  return o;
}

template <typename R, int N>
__device__ __forceinline__ void load_tw(cx<R> *dst, const cx<R> *src) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
}

// ======================================================= row kernels (x lines)
// forward: view -> Tbuf (interleaved)
template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_rows_fwd(VecView<R> in, cx<R> *tb, long nrows, long N,
                                                const cx<R> *twx) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, ROWS = T / G, P = PL::P, PN = lds_line<NX>();
  __shared__ cx<R> buf[ROWS][PN];
  __shared__ cx<R> tw[NX];
  load_tw<R, NX>(tw, twx);
  __syncthreads();
  const long vec = blockIdx.y;
  const int rl = threadIdx.x / G;
  const int g = threadIdx.x % G;
  const long row = (long)blockIdx.x * ROWS + rl;
  const bool act = (rl < ROWS) && (row < nrows);
  const int rls = rl < ROWS ? rl : 0;
  TwFor<NX, -1, R> tc;
  tc.init(g, tw);
  cx<R> v[E];
  constexpr int r0 = Pass<NX, -1>::radix(0);
  if (act) {
#pragma unroll
    for (int b = 0; b < E / r0; b++)
#pragma unroll
      for (int k = 0; k < r0; k++) v[b * r0 + k] = in.ld(vec, row * NX + in_idx<NX, -1, 0>(g, b, k));
  }
  run_passes_c<NX, -1, 0, P>(v, g, act, tc, [&](int i) -> cx<R> & { return buf[rls][lds_pad(i)]; });
  constexpr int rl_ = Pass<NX, -1>::radix(P - 1);
  if (act) {
#pragma unroll
    for (int b = 0; b < E / rl_; b++)
#pragma unroll
      for (int k = 0; k < rl_; k++)
        tb[vec * N + row * NX + out_idx<NX, -1, P - 1>(g, b, k)] = v[b * rl_ + k];
  }
}

// backward + epilogue: Tbuf -> view (psi_in read from `pin`)
template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_rows_bwd(const cx<R> *tb, VecView<R> pin, VecView<R> out,
                                                Epi<R> epi, long nrows, long N, const cx<R> *twx) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, ROWS = T / G, P = PL::P, PN = lds_line<NX>();
  __shared__ cx<R> buf[ROWS][PN];
  __shared__ cx<R> tw[NX];
  load_tw<R, NX>(tw, twx);
  __syncthreads();
  const long vec = blockIdx.y;
  const int rl = threadIdx.x / G;
  const int g = threadIdx.x % G;
  const long row = (long)blockIdx.x * ROWS + rl;
  const bool act = (rl < ROWS) && (row < nrows);
  const int rls = rl < ROWS ? rl : 0;
  TwFor<NX, 1, R> tc;
  tc.init(g, tw);
  cx<R> v[E];
  constexpr int r0 = Pass<NX, 1>::radix(0);
  constexpr int rl_ = Pass<NX, 1>::radix(P - 1);
  cx<R> p[E];
  R V[E];
  if (act) {
#pragma unroll
    for (int b = 0; b < E / r0; b++)
#pragma unroll
      for (int k = 0; k < r0; k++) v[b * r0 + k] = tb[vec * N + row * NX + in_idx<NX, 1, 0>(g, b, k)];
#pragma unroll
    for (int b = 0; b < E / rl_; b++)
#pragma unroll
      for (int k = 0; k < rl_; k++) {
        const long gi = row * NX + out_idx<NX, 1, P - 1>(g, b, k);
        p[b * rl_ + k] = pin.ld(vec, gi);
        V[b * rl_ + k] = epi.pot[gi];
      }
  }
  run_passes_c<NX, 1, 0, P>(v, g, act, tc, [&](int i) -> cx<R> & { return buf[rls][lds_pad(i)]; });
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (act) {
#pragma unroll
    for (int b = 0; b < E / rl_; b++)
#pragma unroll
      for (int k = 0; k < rl_; k++) {
        const long gi = row * NX + out_idx<NX, 1, P - 1>(g, b, k);
        out.st(vec, gi, epilogue(epi, v[b * rl_ + k], p[b * rl_ + k], V[b * rl_ + k], s));
      }
  }
  if (epi.part) block_sum4<T>(s, epi.part, (long)blockIdx.y * gridDim.x + blockIdx.x);
}

// The row + line family's counterpart of k_plane_bwd_fwd_w: the inverse x
// FFT and epilogue of Newton step j (k_rows_bwd) and the forward x FFT of the
// resulting psi_j row for step j+1 (k_rows_fwd) in one kernel; the row of tb
// read at the start is rewritten at the end by the same threads.  Saves the
// forward kernel's read of psi_j.
template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_rows_bwd_fwd(cx<R> *tb, VecView<R> pin, VecView<R> out,
                                                    Epi<R> epi, long nrows, long N,
                                                    const cx<R> *twx) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, ROWS = T / G, P = PL::P, PN = lds_line<NX>();
  __shared__ cx<R> buf[ROWS][PN];
  __shared__ cx<R> tw[NX];
  load_tw<R, NX>(tw, twx);
  __syncthreads();
  const long vec = blockIdx.y;
  const int rl = threadIdx.x / G;
  const int g = threadIdx.x % G;
  const long row = (long)blockIdx.x * ROWS + rl;
  const bool act = (rl < ROWS) && (row < nrows);
  const int rls = rl < ROWS ? rl : 0;
  TwFor<NX, 1, R> tcb;
  tcb.init(g, tw);
  cx<R> v[E];
  constexpr int r0 = Pass<NX, 1>::radix(0);
  constexpr int rl_ = Pass<NX, 1>::radix(P - 1);
  constexpr int rlf = Pass<NX, -1>::radix(P - 1);
  cx<R> p[E];
  R V[E];
  if (act) {
#pragma unroll
    for (int b = 0; b < E / r0; b++)
#pragma unroll
      for (int k = 0; k < r0; k++) v[b * r0 + k] = tb[vec * N + row * NX + in_idx<NX, 1, 0>(g, b, k)];
#pragma unroll
    for (int b = 0; b < E / rl_; b++)
#pragma unroll
      for (int k = 0; k < rl_; k++) {
        const long gi = row * NX + out_idx<NX, 1, P - 1>(g, b, k);
        p[b * rl_ + k] = pin.ld(vec, gi);
        V[b * rl_ + k] = epi.pot[gi];
      }
  }
  auto at = [&](int i) -> cx<R> & { return buf[rls][lds_pad(i)]; };
  run_passes_c<NX, 1, 0, P>(v, g, act, tcb, at);
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (act) {
#pragma unroll
    for (int b = 0; b < E / rl_; b++)
#pragma unroll
      for (int k = 0; k < rl_; k++) {
        const long gi = row * NX + out_idx<NX, 1, P - 1>(g, b, k);
        v[b * rl_ + k] = epilogue(epi, v[b * rl_ + k], p[b * rl_ + k], V[b * rl_ + k], s);
        out.st(vec, gi, v[b * rl_ + k]);
      }
  }
  // psi_j's row through the row buffer into the forward input order
  __syncthreads();
  write_out<NX, 1, P - 1>(v, g, act, at);
  __syncthreads();
  read_in<NX, -1, 0>(v, g, act, at);
  TwFor<NX, -1, R> tcf;
  tcf.init(g, tw);
  run_passes_c<NX, -1, 0, P>(v, g, act, tcf, at);
  if (act) {
#pragma unroll
    for (int b = 0; b < E / rlf; b++)
#pragma unroll
      for (int k = 0; k < rlf; k++)
        tb[vec * N + row * NX + out_idx<NX, -1, P - 1>(g, b, k)] = v[b * rlf + k];
  }
}

// =============================================== strided lines (y, fallback)
// Lines of length N along an axis with element stride `es`; the block takes W
// consecutive x positions.  Line base for (x, outer) = outer*os + x.
template <typename R, int NL, int W, int S>
__global__ void __launch_bounds__(W * FftPlan<NL>::G) k_lines(cx<R> *tb, long N, int nx, long es,
                                                              long os, const cx<R> *twl) {
  using PL = FftPlan<NL>;
  constexpr int G = PL::G, E = PL::E, P = PL::P;
  __shared__ cx<R> buf[NL][W];
  __shared__ cx<R> tw[NL];
  load_tw<R, NL>(tw, twl);
  __syncthreads();
  const int p = threadIdx.x % W;
  const int g = threadIdx.x / W;
  const long x = (long)blockIdx.x * W + p;
  const long outer = blockIdx.y;
  const long vec = blockIdx.z;
  const bool act = x < nx;
  cx<R> *base = tb + vec * N + outer * os + x;
  cx<R> v[E];
  constexpr int r0 = Pass<NL, S>::radix(0);
  if (act) {
#pragma unroll
    for (int b = 0; b < E / r0; b++)
#pragma unroll
      for (int k = 0; k < r0; k++) v[b * r0 + k] = base[(long)in_idx<NL, S, 0>(g, b, k) * es];
  }
  run_passes<NL, S, 0, P>(v, g, act, tw, [&](int i) -> cx<R> & { return buf[i][p]; });
  constexpr int rl_ = Pass<NL, S>::radix(P - 1);
  if (act) {
#pragma unroll
    for (int b = 0; b < E / rl_; b++)
#pragma unroll
      for (int k = 0; k < rl_; k++) base[(long)out_idx<NL, S, P - 1>(g, b, k) * es] = v[b * rl_ + k];
  }
}

// ======================================== z pencils: forward, * k^2, backward
// PRE: the factor is kscale / (kin + pden) (kinetic preconditioner) instead
// of kin * kscale
template <typename R, int NZ, int W, bool PRE = false>
__global__ void __launch_bounds__(W * FftPlan<NZ>::G) k_pencil(cx<R> *tb, long N, int nx, int ny,
                                                               const cx<R> *twz,
                                                               const double *kxx,
                                                               const double *kyy,
                                                               const double *kzz, double ekin,
                                                               double kscale, double pden) {
  using PL = FftPlan<NZ>;
  constexpr int G = PL::G, E = PL::E, P = PL::P;
  __shared__ cx<R> buf[NZ][W];
  __shared__ cx<R> tw[NZ];
  __shared__ double kz_s[NZ];
  load_tw<R, NZ>(tw, twz);
  for (int i = threadIdx.x; i < NZ; i += blockDim.x) kz_s[i] = kzz[i];
  __syncthreads();
  const int p = threadIdx.x % W;
  const int g = threadIdx.x / W;
  const long x = (long)blockIdx.x * W + p;
  const long y = blockIdx.y;
  const long vec = blockIdx.z;
  const bool act = x < nx;
  const long es = (long)nx * ny;
  cx<R> *base = tb + vec * N + y * nx + x;
  cx<R> v[E];
  constexpr int r0 = Pass<NZ, -1>::radix(0);
  if (act) {
#pragma unroll
    for (int b = 0; b < E / r0; b++)
#pragma unroll
      for (int k = 0; k < r0; k++) v[b * r0 + k] = base[(long)in_idx<NZ, -1, 0>(g, b, k) * es];
  }
  auto at = [&](int i) -> cx<R> & { return buf[i][p]; };
  run_passes<NZ, -1, 0, P>(v, g, act, tw, at);
  // kinetic factor, filter/init.c:138-144 order: (kx + ky) + kz, clamp
  const double kxy = act ? (kxx[x] + kyy[y]) : 0.0;
  constexpr int rl_ = Pass<NZ, -1>::radix(P - 1);
#pragma unroll
  for (int b = 0; b < E / rl_; b++)
#pragma unroll
    for (int k = 0; k < rl_; k++) {
      int z = out_idx<NZ, -1, P - 1>(g, b, k);
      double kin = kxy + kz_s[z];
      if (kin > ekin) kin = ekin;
      if constexpr (PRE)
        v[b * rl_ + k] = cscale(v[b * rl_ + k], (R)(kscale / (kin + pden)));
      else
        v[b * rl_ + k] = cscale(v[b * rl_ + k], (R)(kin * kscale));
    }
  // backward: pass 0 reads exactly the registers the forward last pass wrote
  run_passes<NZ, 1, 0, P>(v, g, act, tw, at);
  constexpr int rb = Pass<NZ, 1>::radix(P - 1);
  if (act) {
#pragma unroll
    for (int b = 0; b < E / rb; b++)
#pragma unroll
      for (int k = 0; k < rb; k++) base[(long)out_idx<NZ, 1, P - 1>(g, b, k) * es] = v[b * rb + k];
  }
}

// ================================================ xy planes held in LDS
template <int NX>
struct PlaneCfg {
  static constexpr int LD = NX + 1;  // padded row pitch
};

// forward: view plane -> x FFT rows -> LDS -> y FFT columns -> Tbuf.
// All rows of the plane are loaded into registers before any compute, so the
// whole plane is in flight at once (one block per CU: the plane fills LDS).
template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_plane_fwd(VecView<R> in, cx<R> *tb, long N, int nz,
                                                 const cx<R> *twg) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, P = PL::P, LD = PlaneCfg<NX>::LD;
  constexpr int RX = T / G;                 // lines per round
  constexpr int NR = (NX + RX - 1) / RX;    // rounds
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *tw = buf + NX * LD;
  const long vec = blockIdx.x / nz;
  const long z = blockIdx.x % nz;
  const long pbase = z * (long)NX * NX;
  const int tl = threadIdx.x / G, g = threadIdx.x % G;
  constexpr int r0 = Pass<NX, -1>::radix(0);
  constexpr int rl_ = Pass<NX, -1>::radix(P - 1);
  cx<R> v[NR][E];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int row = r * RX + tl;
    if ((tl < RX) && (row < NX)) {
#pragma unroll
      for (int b = 0; b < E / r0; b++)
#pragma unroll
        for (int k = 0; k < r0; k++)
          v[r][b * r0 + k] = in.ld(vec, pbase + row * NX + in_idx<NX, -1, 0>(g, b, k));
    }
  }
  load_tw<R, NX>(tw, twg);
  __syncthreads();
  // ---- rows
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int row = r * RX + tl;
    const bool act = (tl < RX) && (row < NX);
    const int rs = act ? row : 0;
    auto at = [&](int i) -> cx<R> & { return buf[rs * LD + i]; };
    run_passes<NX, -1, 0, P>(v[r], g, act, tw, at);
    if constexpr (P > 1) __syncthreads();
    write_out<NX, -1, P - 1>(v[r], g, act, at);
  }
  __syncthreads();
  // ---- columns (consecutive threads -> consecutive columns)
  const int cl = threadIdx.x % RX, gc = threadIdx.x / RX;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int col = r * RX + cl;
    const bool act = (gc < G) && (col < NX);
    const int cs = act ? col : 0;
    const int gs = act ? gc : 0;
    cx<R> *w = v[0];
    auto at = [&](int i) -> cx<R> & { return buf[i * LD + cs]; };
    read_in<NX, -1, 0>(w, gs, act, at);
    run_passes<NX, -1, 0, P>(w, gs, act, tw, at);
    if (act) {
#pragma unroll
      for (int b = 0; b < E / rl_; b++)
#pragma unroll
        for (int k = 0; k < rl_; k++)
          tb[vec * N + pbase + (long)out_idx<NX, -1, P - 1>(gs, b, k) * NX + cs] = w[b * rl_ + k];
    }
    __syncthreads();
  }
}

// backward + epilogue: Tbuf plane -> y FFT columns -> LDS -> x FFT rows -> out.
// The column data are prefetched for all rounds; the epilogue operands of a
// row round (psi_{j-1}, V) are loaded before that round's FFT work.
template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_plane_bwd(const cx<R> *tb, VecView<R> pin, VecView<R> out,
                                                 Epi<R> epi, long N, int nz, const cx<R> *twg) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, P = PL::P, LD = PlaneCfg<NX>::LD;
  constexpr int RX = T / G;
  constexpr int NR = (NX + RX - 1) / RX;
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *tw = buf + NX * LD;
  const long vec = blockIdx.x / nz;
  const long z = blockIdx.x % nz;
  const long pbase = z * (long)NX * NX;
  constexpr int r0 = Pass<NX, 1>::radix(0);
  constexpr int rl_ = Pass<NX, 1>::radix(P - 1);
  // ---- columns
  const int cl = threadIdx.x % RX, gc = threadIdx.x / RX;
  cx<R> v[NR][E];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int col = r * RX + cl;
    if ((gc < G) && (col < NX)) {
#pragma unroll
      for (int b = 0; b < E / r0; b++)
#pragma unroll
        for (int k = 0; k < r0; k++)
          v[r][b * r0 + k] = tb[vec * N + pbase + (long)in_idx<NX, 1, 0>(gc, b, k) * NX + col];
    }
  }
  load_tw<R, NX>(tw, twg);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int col = r * RX + cl;
    const bool act = (gc < G) && (col < NX);
    const int cs = act ? col : 0;
    const int gs = act ? gc : 0;
    auto at = [&](int i) -> cx<R> & { return buf[i * LD + cs]; };
    run_passes<NX, 1, 0, P>(v[r], gs, act, tw, at);
    if constexpr (P > 1) __syncthreads();
    write_out<NX, 1, P - 1>(v[r], gs, act, at);
  }
  __syncthreads();
  // ---- rows + epilogue
  const int tl = threadIdx.x / G, g = threadIdx.x % G;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int row = r * RX + tl;
    const bool act = (tl < RX) && (row < NX);
    const int rs = act ? row : 0;
    cx<R> p[E];
    R V[E];
    if (act) {
#pragma unroll
      for (int b = 0; b < E / rl_; b++)
#pragma unroll
        for (int k = 0; k < rl_; k++) {
          const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
          p[b * rl_ + k] = pin.ld(vec, gi);
          V[b * rl_ + k] = epi.pot[gi];
        }
    }
    cx<R> *w = v[0];
    auto at = [&](int i) -> cx<R> & { return buf[rs * LD + i]; };
    read_in<NX, 1, 0>(w, g, act, at);
    run_passes<NX, 1, 0, P>(w, g, act, tw, at);
    if (act) {
#pragma unroll
      for (int b = 0; b < E / rl_; b++)
#pragma unroll
        for (int k = 0; k < rl_; k++) {
          const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
          out.st(vec, gi, epilogue(epi, w[b * rl_ + k], p[b * rl_ + k], V[b * rl_ + k], s));
        }
    }
    __syncthreads();
  }
  if (epi.part) block_sum4<T>(s, epi.part, blockIdx.x);
}

// ------------------------------------------- wave-local plane kernels
// When the threads of one line fit a wavefront (64 % G == 0), every FFT
// exchange of a row or a column stays inside one wavefront; only the
// row <-> column transpose needs block barriers.  Line positions in LDS are
// placed by PlaneLds (below) so the strided exchange patterns spread over the
// banks; FP32 threads keep their twiddles in registers (TwFor).
// If the plane does not fit LDS, the transpose goes through the buffer in
// NH chunks of NX/NH columns while the plane stays in registers.
// The spectrum of a plane is stored column-major, Tbuf[z][kx][ky], so that
// the column phase writes (and the backward column phase reads) contiguous
// ky runs.
template <typename R, int NX>
constexpr int plane_chunks() {
  // smallest power of two (dividing NX) whose chunk buffer (plus the
  // twiddle table) fits the 160 KB of LDS of a gfx950 workgroup
  int nh = 1;
  while (nh < NX && NX % (2 * nh) == 0 &&
         sizeof(cx<R>) * ((size_t)NX * (NX / nh + 1) + NX) > 160 * 1024 - 1024)
    nh *= 2;
  return nh;
}

// LDS placement of one plane in the wave-local plane kernels: row y of the
// (transposed) plane starts at rowbase(y); during the row FFTs element x of
// a row sits at rowbase(y) + xpos(x).  Generic lengths: pitch NX/NH + 1,
// rows and in-row positions xor-swizzled.  Complex64 at 128 (G = 16 threads
// per line, two lines per half-wave): pitch 150 with per-row offsets and an
// in-row permutation that make every access of the row and column phases
// free of bank conflicts for 8-byte accesses on gfx950 (64 banks of 4 B, 32
// lanes served per cycle): the two rows (columns) of a half-wave start
// 16 mod 32 elements apart, and the 16 positions of every access pattern of
// the radix-8/8/2 passes are distinct mod 16.
template <typename R, int NX, int NH>
struct PlaneLds {
  static constexpr int LDc = NX / NH + 1;
  static __device__ __forceinline__ int rowbase(int y) { return lds_swz(y) * LDc; }
  static __device__ __forceinline__ int xpos(int x) { return lds_swz(x); }
};
template <>
struct PlaneLds<float, 128, 1> {
  static constexpr int LDc = 150;
  static __device__ __forceinline__ int rowbase(int y) {
    const int d = 3 + 2 * ((y >> 5) & 3) + ((((y >> 3) ^ (y >> 4)) & 1) ^ ((y >> 6) & 1)) -
                  3 * (y & 1);
    return y * LDc + 2 * d;
  }
  static __device__ __forceinline__ int xpos(int x) {
    const int u = x >> 4;
    return x ^ (((9 * u) ^ (8 * (u >> 2))) & 15);
  }
};

template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_plane_fwd_w(VecView<R> in, cx<R> *tb, long N, int nz,
                                                   const cx<R> *twg) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, P = PL::P;
  static_assert(64 % G == 0 && NX % 8 == 0, "lines must not straddle wavefronts");
  constexpr int NH = plane_chunks<R, NX>(), W = NX / NH;
  using L = PlaneLds<R, NX, NH>;
  constexpr int LDc = L::LDc;
  constexpr int RX = T / G;                  // lines per round
  constexpr int NR = (NX + RX - 1) / RX;     // row rounds
  constexpr int NCR = (W + RX - 1) / RX;     // column rounds per chunk
  constexpr int RS = NX + 1;                 // row-exchange slice pitch (NH > 1)
  static_assert(NH == 1 || RX * RS <= NX * LDc, "row exchange slices exceed the chunk buffer");
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);  // [NX][LDc], row y at L::rowbase(y)
  cx<R> *tw = buf + NX * LDc;
  const long vec = blockIdx.x / nz;
  const long z = blockIdx.x % nz;
  const long pbase = z * (long)NX * NX;
  const int tl = threadIdx.x / G, g = threadIdx.x % G;
  constexpr int r0 = Pass<NX, -1>::radix(0);
  constexpr int rl_ = Pass<NX, -1>::radix(P - 1);
  cx<R> v[NR][E];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int row = r * RX + tl;
    if (row < NX) {
#pragma unroll
      for (int b = 0; b < E / r0; b++)
#pragma unroll
        for (int k = 0; k < r0; k++)
          v[r][b * r0 + k] = in.ld(vec, pbase + row * NX + in_idx<NX, -1, 0>(g, b, k));
    }
  }
  load_tw<R, NX>(tw, twg);
  __syncthreads();
  TwFor<NX, -1, R> tc;
  tc.init(g, tw);
  // ---- rows: x FFT; the results stay in registers.  Exchanges use the row's
  // own buffer row (NH == 1) or slice tl of the chunk buffer.
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int row = r * RX + tl;
    const bool act = row < NX;
    const int base = (NH == 1 ? L::rowbase(act ? row : 0) : tl * RS);
    auto at = [&](int i) -> cx<R> & { return buf[base + L::xpos(i)]; };
    run_passes_c<NX, -1, 0, P, WaveSync>(v[r], g, act, tc, at);
  }
  // ---- columns, chunk by chunk
#pragma unroll
  for (int c = 0; c < NH; c++) {
    __syncthreads();  // row exchanges / the previous chunk's columns are done
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int row = r * RX + tl;
      if (row < NX) {
#pragma unroll
        for (int b = 0; b < E / rl_; b++)
#pragma unroll
          for (int k = 0; k < rl_; k++) {
            const int kx = out_idx<NX, -1, P - 1>(g, b, k);
            if (kx / W == c) buf[L::rowbase(row) + (kx - c * W)] = v[r][b * rl_ + k];
          }
      }
    }
    __syncthreads();  // the transpose
#pragma unroll
    for (int cr = 0; cr < NCR; cr++) {
      const int cl = cr * RX + tl;
      const bool act = cl < W;
      const int cs = act ? cl : 0;
      cx<R> w[E];
      auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(i) + cs]; };
      read_in<NX, -1, 0>(w, g, act, at);
      run_passes_c<NX, -1, 0, P, WaveSync>(w, g, act, tc, at);
      if (act) {
        cx<R> *dst = tb + vec * N + pbase + (long)(c * W + cs) * NX;
#pragma unroll
        for (int b = 0; b < E / rl_; b++)
#pragma unroll
          for (int k = 0; k < rl_; k++) dst[out_idx<NX, -1, P - 1>(g, b, k)] = w[b * rl_ + k];
      }
    }
  }
}

template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_plane_bwd_w(const cx<R> *tb, VecView<R> pin,
                                                   VecView<R> out, Epi<R> epi, long N, int nz,
                                                   const cx<R> *twg) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, P = PL::P;
  static_assert(64 % G == 0 && NX % 8 == 0, "lines must not straddle wavefronts");
  constexpr int NH = plane_chunks<R, NX>(), W = NX / NH;
  using L = PlaneLds<R, NX, NH>;
  constexpr int LDc = L::LDc;
  constexpr int RX = T / G;
  constexpr int NR = (NX + RX - 1) / RX;
  constexpr int NCR = (W + RX - 1) / RX;
  constexpr int RS = NX + 1;
  static_assert(NH == 1 || RX * RS <= NX * LDc, "row exchange slices exceed the chunk buffer");
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *tw = buf + NX * LDc;
  const long vec = blockIdx.x / nz;
  const long z = blockIdx.x % nz;
  const long pbase = z * (long)NX * NX;
  const int tl = threadIdx.x / G, g = threadIdx.x % G;
  constexpr int r0 = Pass<NX, 1>::radix(0);
  constexpr int rl_ = Pass<NX, 1>::radix(P - 1);
  constexpr int NVR = NH == 1 ? 1 : NR;  // row inputs held in registers
  cx<R> v[NVR][E];
  // at 1024 threads the register budget is 128: twiddles come from LDS
  using TwB = typename std::conditional<(T >= 1024), TwTable<NX, 1, R>, TwFor<NX, 1, R>>::type;
  TwB tc;
#pragma unroll
  for (int c = 0; c < NH; c++) {
    // ---- columns of chunk c: load, y FFT, into the buffer
    cx<R> w[NCR][E];
#pragma unroll
    for (int cr = 0; cr < NCR; cr++) {
      const int cl = cr * RX + tl;
      if (cl < W) {
        const cx<R> *src = tb + vec * N + pbase + (long)(c * W + cl) * NX;
#pragma unroll
        for (int b = 0; b < E / r0; b++)
#pragma unroll
          for (int k = 0; k < r0; k++) w[cr][b * r0 + k] = src[in_idx<NX, 1, 0>(g, b, k)];
      }
    }
    if (c == 0) {
      load_tw<R, NX>(tw, twg);
      __syncthreads();
      tc.init(g, tw);
    } else {
      __syncthreads();  // the previous chunk's gather is done
    }
#pragma unroll
    for (int cr = 0; cr < NCR; cr++) {
      const int cl = cr * RX + tl;
      const bool act = cl < W;
      const int cs = act ? cl : 0;
      auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(i) + cs]; };
      run_passes_c<NX, 1, 0, P, WaveSync>(w[cr], g, act, tc, at);
      WaveSync::sync();
      write_out<NX, 1, P - 1>(w[cr], g, act, at);
    }
    __syncthreads();  // the transpose
    if constexpr (NH > 1) {
      // gather this chunk's x positions of every row this thread owns
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const int row = r * RX + tl;
        if (row < NX) {
#pragma unroll
          for (int b = 0; b < E / r0; b++)
#pragma unroll
            for (int k = 0; k < r0; k++) {
              const int x = in_idx<NX, 1, 0>(g, b, k);
              if (x / W == c) v[r][b * r0 + k] = buf[L::rowbase(row) + (x - c * W)];
            }
        }
      }
    }
  }
  if constexpr (NH > 1) __syncthreads();  // the buffer is free for the row exchanges
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int row = r * RX + tl;
    const bool act = row < NX;
    const int rs = act ? row : 0;
    // psi_{j-1} and V of the row are loaded before the FFT (in flight
    // during it); at 1024 threads (register budget 128) only V is, psi_{j-1}
    // after the FFT
    constexpr bool early = T < 1024;
    cx<R> p[E];
    R V[E];
    if (act) {
#pragma unroll
      for (int b = 0; b < E / rl_; b++)
#pragma unroll
        for (int k = 0; k < rl_; k++) {
          const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
          if (early) p[b * rl_ + k] = pin.ld(vec, gi);
          V[b * rl_ + k] = epi.pot[gi];
        }
    }
    cx<R> *x = v[NH == 1 ? 0 : r];
    if constexpr (NH == 1) {
      // the row is in the buffer (at L::rowbase(y), x in natural order); the
      // exchanges then run in place in that row at the permuted positions
      auto at0 = [&](int i) -> cx<R> & { return buf[L::rowbase(rs) + i]; };
      read_in<NX, 1, 0>(x, g, act, at0);
      auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(rs) + L::xpos(i)]; };
      run_passes_c<NX, 1, 0, P, WaveSync>(x, g, act, tc, at);
    } else {
      auto at = [&](int i) -> cx<R> & { return buf[tl * RS + lds_swz(i)]; };
      run_passes_c<NX, 1, 0, P, WaveSync>(x, g, act, tc, at);
    }
    if (act) {
      if constexpr (!early) {
#pragma unroll
        for (int b = 0; b < E / rl_; b++)
#pragma unroll
          for (int k = 0; k < rl_; k++) {
            const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
            p[b * rl_ + k] = pin.ld(vec, gi);
          }
      }
#pragma unroll
      for (int b = 0; b < E / rl_; b++)
#pragma unroll
        for (int k = 0; k < rl_; k++) {
          const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
          out.st(vec, gi, epilogue(epi, x[b * rl_ + k], p[b * rl_ + k], V[b * rl_ + k], s));
        }
    }
  }
  if (epi.part) block_sum4<T>(s, epi.part, blockIdx.x);
}

// An int the compiler must treat as unknown: index and LDS address math
// derived from it is recomputed where it is used instead of being kept live
// in registers across the phases of a kernel.
__device__ __forceinline__ int opaque_int(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Pass C of Newton step j fused with pass A of step j+1 (whole-plane buffer,
// NH == 1).  The epilogue's psi_j is stored (the ring slot the accumulation
// reads) and, still in registers, transformed forward -- x rows in each row's
// own buffer row, then y columns -- into plane z of the spectrum buffer,
// which this block alone read at its start.  Saves pass A's read of psi_j;
// the arithmetic is that of k_plane_bwd_w followed by k_plane_fwd_w.
template <typename R, int NX, int T>
__global__ void __launch_bounds__(T) k_plane_bwd_fwd_w(cx<R> *tb, VecView<R> pin, VecView<R> out,
                                                       Epi<R> epi, long N, int nz,
                                                       const cx<R> *twg) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, P = PL::P;
  static_assert(64 % G == 0 && NX % 8 == 0, "lines must not straddle wavefronts");
  static_assert(plane_chunks<R, NX>() == 1, "the fused kernel holds the whole plane in LDS");
  using L = PlaneLds<R, NX, 1>;
  constexpr int LDc = L::LDc;
  constexpr int RX = T / G;
  constexpr int NR = (NX + RX - 1) / RX;
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *tw = buf + NX * LDc;
  const long vec = blockIdx.x / nz;
  const long z = blockIdx.x % nz;
  const long pbase = z * (long)NX * NX;
  const int tl = threadIdx.x / G, g0 = threadIdx.x % G;
  const int g = g0;
  constexpr int r0b = Pass<NX, 1>::radix(0);
  constexpr int rlb = Pass<NX, 1>::radix(P - 1);
  constexpr int rlf = Pass<NX, -1>::radix(P - 1);
  // psi_{j-1} before the row FFT only below 1024 threads (at 1024 it fits
  // the registers but measured 5 % slower: 138 against 132 us at 128^3)
  constexpr bool early = T < 1024;
  using TwB = typename std::conditional<(T >= 1024), TwTable<NX, 1, R>, TwFor<NX, 1, R>>::type;
  using TwF = typename std::conditional<(T >= 1024), TwTable<NX, -1, R>, TwFor<NX, -1, R>>::type;
  TwB tcb;
  TwF tcf;
  // ---- inverse y FFT of the columns (pass C)
  {
    cx<R> w[NR][E];
#pragma unroll
    for (int cr = 0; cr < NR; cr++) {
      const int cl = cr * RX + tl;
      if (cl < NX) {
        const cx<R> *src = tb + vec * N + pbase + (long)cl * NX;
#pragma unroll
        for (int b = 0; b < E / r0b; b++)
#pragma unroll
          for (int k = 0; k < r0b; k++) w[cr][b * r0b + k] = src[in_idx<NX, 1, 0>(g, b, k)];
      }
    }
    load_tw<R, NX>(tw, twg);
    __syncthreads();
    tcb.init(g, tw);
    tcf.init(g, tw);
#pragma unroll
    for (int cr = 0; cr < NR; cr++) {
      const int cl = cr * RX + tl;
      const bool act = cl < NX;
      const int cs = act ? cl : 0;
      auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(i) + cs]; };
      run_passes_c<NX, 1, 0, P, WaveSync>(w[cr], g, act, tcb, at);
      WaveSync::sync();
      write_out<NX, 1, P - 1>(w[cr], g, act, at);
    }
  }
  __syncthreads();  // the transpose
  // ---- rows: inverse x FFT, epilogue, store psi_j, forward x FFT of psi_j
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int g = opaque_int(g0);  // per-round index math (see opaque_int)
    tcb.init(g, tw);
    tcf.init(g, tw);
    const int row = r * RX + tl;
    const bool act = row < NX;
    const int rs = act ? row : 0;
    cx<R> p[E];
    R V[E];
    if (act) {
#pragma unroll
      for (int b = 0; b < E / rlb; b++)
#pragma unroll
        for (int k = 0; k < rlb; k++) {
          const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
          if (early) p[b * rlb + k] = pin.ld(vec, gi);
          V[b * rlb + k] = epi.pot[gi];
        }
    }
    cx<R> x[E];
    auto at0 = [&](int i) -> cx<R> & { return buf[L::rowbase(rs) + i]; };
    read_in<NX, 1, 0>(x, g, act, at0);
    auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(rs) + L::xpos(i)]; };
    run_passes_c<NX, 1, 0, P, WaveSync>(x, g, act, tcb, at);
    if (act) {
      if constexpr (!early) {
#pragma unroll
        for (int b = 0; b < E / rlb; b++)
#pragma unroll
          for (int k = 0; k < rlb; k++) {
            const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
            p[b * rlb + k] = pin.ld(vec, gi);
          }
      }
#pragma unroll
      for (int b = 0; b < E / rlb; b++)
#pragma unroll
        for (int k = 0; k < rlb; k++) {
          const long gi = pbase + rs * NX + out_idx<NX, 1, P - 1>(g, b, k);
          x[b * rlb + k] = epilogue(epi, x[b * rlb + k], p[b * rlb + k], V[b * rlb + k], s);
          if constexpr (sizeof(R) == 8) {
            if (epi.so_map) {  // uniform: spinor filters only
              const int u = epi.so_map[gi];
              if (u >= 0) {  // vector v: spinor v / 2, up (v even) or down
                const double *gq = epi.so_g + ((long)u * epi.so_nsp + (vec >> 1)) * 4 + 2 * (vec & 1);
                x[b * rlb + k].re = fma(epi.so_alpha, gq[0], (double)x[b * rlb + k].re);
                x[b * rlb + k].im = fma(epi.so_alpha, gq[1], (double)x[b * rlb + k].im);
              }
            }
          }
          out.st(vec, gi, x[b * rlb + k]);
        }
    }
    // psi_j's row back through the row buffer into the forward input order
    WaveSync::sync();
    write_out<NX, 1, P - 1>(x, g, act, at);
    WaveSync::sync();
    read_in<NX, -1, 0>(x, g, act, at);
    run_passes_c<NX, -1, 0, P, WaveSync>(x, g, act, tcf, at);
    WaveSync::sync();
    if (act) {
#pragma unroll
      for (int b = 0; b < E / rlf; b++)
#pragma unroll
        for (int k = 0; k < rlf; k++)
          buf[L::rowbase(rs) + out_idx<NX, -1, P - 1>(g, b, k)] = x[b * rlf + k];
    }
  }
  __syncthreads();  // the transpose
  // ---- forward y FFT of the columns (pass A of step j+1) into plane z of tb
#pragma unroll
  for (int cr = 0; cr < NR; cr++) {
    const int g = opaque_int(g0);
    tcf.init(g, tw);
    const int cl = cr * RX + tl;
    const bool act = cl < NX;
    const int cs = act ? cl : 0;
    cx<R> y[E];
    auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(i) + cs]; };
    read_in<NX, -1, 0>(y, g, act, at);
    run_passes_c<NX, -1, 0, P, WaveSync>(y, g, act, tcf, at);
    if (act) {
      cx<R> *dst = tb + vec * N + pbase + (long)cs * NX;
#pragma unroll
      for (int b = 0; b < E / rlf; b++)
#pragma unroll
        for (int k = 0; k < rlf; k++) dst[out_idx<NX, -1, P - 1>(g, b, k)] = y[b * rlf + k];
    }
  }
}

// Plane placement of the separable-layout fused kernel (complex64, 128^2):
// element (y, x) at Y::f(y) + X::f(x) (8-byte words), each a weighted sum of
// the index bits, so every FFT access pattern is a per-thread base plus a
// compile-time offset (sep_plan_ok).  On gfx950 an 8-byte LDS read serves 32
// lanes a cycle (bank = word mod 32) and a write 16 (word mod 16).  The
// weights' residues mod 32 (1, 4, 8, 6, 5, 16, 2 for bits 0..6 of either
// index) make the 16 lanes of a line write 16 distinct words mod 16 in every
// pattern of the radix-8/8/2 plan except two (two ways: index bits 3..6 and
// 0, 4, 5, 6), and the two lines of a half-wave, which differ in bit 5 of
// their line index (16 mod 32 apart), read 32 distinct words mod 32.  The
// weights are the smallest superincreasing sequence with those residues
// (so the placement is one to one): 17429 words, 136 KB.
struct PlaneSep128Y {
  __host__ __device__ static constexpr int f(int y) {
    return 33 * (y & 1) + 132 * ((y >> 1) & 1) + 264 * ((y >> 2) & 1) + 4358 * ((y >> 3) & 1) +
           1093 * ((y >> 4) & 1) + 8720 * ((y >> 5) & 1) + 66 * ((y >> 6) & 1);
  }
};
struct PlaneSep128X {
  __host__ __device__ static constexpr int f(int x) {
    return (x & 1) + 4 * ((x >> 1) & 1) + 8 * ((x >> 2) & 1) + 2182 * ((x >> 3) & 1) +
           549 * ((x >> 4) & 1) + 16 * ((x >> 5) & 1) + 2 * ((x >> 6) & 1);
  }
};
// Line of thread group tl (0..63) in a round: bits 0 and 5 swapped, so that
// the two lines of a half-wave (tl = 2m, 2m + 1) differ in bit 5.
__host__ __device__ constexpr int plane_sep_line(int tl) {
  return (tl & 30) | ((tl & 1) << 5) | ((tl >> 5) & 1);
}
constexpr int kPlaneSep128Words = PlaneSep128Y::f(127) + PlaneSep128X::f(127) + 1;
constexpr bool plane_sep128_one_to_one() {
  bool used[kPlaneSep128Words] = {};
  for (int y = 0; y < 128; y++)
    for (int x = 0; x < 128; x++) {
      const int a = PlaneSep128Y::f(y) + PlaneSep128X::f(x);
      if (used[a]) return false;
      used[a] = true;
    }
  return true;
}
static_assert(plane_sep128_one_to_one(), "plane placement must be one to one");
static_assert(sep_plan_ok<128, 1, PlaneSep128Y>() && sep_plan_ok<128, -1, PlaneSep128Y>() &&
                  sep_plan_ok<128, 1, PlaneSep128X>() && sep_plan_ok<128, -1, PlaneSep128X>(),
              "plane placement must be separable for every pass of the 128 plan");

constexpr size_t plane_sep_lds() {
  return sizeof(cx<float>) *
         (size_t)(kPlaneSep128Words + TwSep<128, 1, float>::WORDS + TwSep<128, -1, float>::WORDS);
}

// Per-thread-index twiddle tables of the separable fused kernel (TwSep, the
// inverse transform's then the forward one's), built on the host.
constexpr int plane_sep_tw_words() {
  return TwSep<128, 1, float>::WORDS + TwSep<128, -1, float>::WORDS;
}
void plane_sep_tw_host(void *tab) {
  cx<float> *t = static_cast<cx<float> *>(tab);
  TwSep<128, 1, float>::host_table(t);
  TwSep<128, -1, float>::host_table(t + TwSep<128, 1, float>::WORDS);
}
long plane_sep_tw_bytes(int n, int prec) {
  return (n == 128 && prec == SEP_PREC_FP32) ? (long)sizeof(cx<float>) * plane_sep_tw_words() : 0;
}

// k_plane_bwd_fwd_w for complex64 128^2 on the separable placement: the same
// arithmetic (pass C of step j, epilogue, pass A of step j+1, plane by plane),
// every LDS access a thread base plus an immediate offset, the twiddles read
// from per-thread-index tables (TwSep, copied from tws).  Between the inverse
// and the forward x transform of a row the data stay in registers: the last
// inverse pass writes the elements in the order the first forward pass reads
// them (both g + G b + k N / R[0]).
template <int T>
__global__ void __launch_bounds__(T) k_plane_bwd_fwd_s(cx<float> *tb, VecView<float> pin,
                                                       VecView<float> out, Epi<float> epi, long N,
                                                       int nz, const cx<float> *tws) {
  constexpr int NX = 128;
  using R = float;
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, P = PL::P;
  using LY = PlaneSep128Y;
  using LX = PlaneSep128X;
  constexpr int RX = T / G;
  constexpr int NR = NX / RX;
  static_assert(NX % RX == 0, "whole rounds of lines");
  static_assert(Pass<NX, 1>::radix(P - 1) == Pass<NX, -1>::radix(0),
                "the inverse transform's last pass is the forward transform's first");
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *twb = buf + kPlaneSep128Words;
  cx<R> *twf = twb + TwSep<NX, 1, R>::WORDS;
  const long vec = blockIdx.x / nz;
  const long z = blockIdx.x % nz;
  const long pbase = z * (long)NX * NX;
  const int tl = threadIdx.x / G, g = threadIdx.x % G;
  static_assert(RX == 64, "plane_sep_line maps 64 thread groups");
  const int ln = plane_sep_line(tl);
  constexpr int r0b = Pass<NX, 1>::radix(0);
  constexpr int rlb = Pass<NX, 1>::radix(P - 1);
  constexpr int rlf = Pass<NX, -1>::radix(P - 1);
  TwSep<NX, 1, R> tcb;
  TwSep<NX, -1, R> tcf;
  // the twiddle tables' loads are issued ahead of the plane's, so that their
  // LDS stores wait for them alone
  constexpr int TWN = plane_sep_tw_words();
  static_assert(TWN <= T, "one twiddle entry per thread");
  cx<R> twv;
  if (threadIdx.x < TWN) twv = tws[threadIdx.x];
  // ---- inverse y FFT of the columns (pass C)
  {
    cx<R> w[NR][E];
#pragma unroll
    for (int cr = 0; cr < NR; cr++) {
      const cx<R> *src = tb + vec * N + pbase + (long)(cr * RX + ln) * NX;
#pragma unroll
      for (int b = 0; b < E / r0b; b++)
#pragma unroll
        for (int k = 0; k < r0b; k++) w[cr][b * r0b + k] = src[in_idx<NX, 1, 0>(g, b, k)];
    }
    if (threadIdx.x < TWN) twb[threadIdx.x] = twv;  // twf follows twb
    __syncthreads();
    tcb.init(g, twb);
    tcf.init(g, twf);
#pragma unroll
    for (int cr = 0; cr < NR; cr++) {
      cx<R> *line = buf + LX::f(cr * RX + ln);
      run_line_sep<NX, 1, LY>(w[cr], g, tcb, line);
      WaveSync::sync();
      write_out_sep<NX, 1, P - 1, LY>(w[cr], g, line);
    }
  }
  __syncthreads();  // the transpose
  // ---- rows: inverse x FFT, epilogue, store psi_j, forward x FFT of psi_j
  double s[4] = {0.0, 0.0, 0.0, 0.0};  // unused: this path has no partial sums
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int row = r * RX + ln;
    cx<R> *line = buf + LY::f(row);
    const long rbase = pbase + (long)row * NX;
    R V[E];
#pragma unroll
    for (int b = 0; b < E / rlb; b++)
#pragma unroll
      for (int k = 0; k < rlb; k++) V[b * rlb + k] = epi.pot[rbase + out_idx<NX, 1, P - 1>(g, b, k)];
    cx<R> x[E];
    read_in_sep<NX, 1, 0, LX>(x, g, line);
    run_line_sep<NX, 1, LX>(x, g, tcb, line);
    // psi_{j-1} of the row, all loads ahead of the stores (which may alias
    // them: loads issued after a store would wait for it)
    cx<R> p[E];
#pragma unroll
    for (int b = 0; b < E / rlb; b++)
#pragma unroll
      for (int k = 0; k < rlb; k++) p[b * rlb + k] = pin.ld(vec, rbase + out_idx<NX, 1, P - 1>(g, b, k));
#pragma unroll
    for (int b = 0; b < E / rlb; b++)
#pragma unroll
      for (int k = 0; k < rlb; k++) {
        const long gi = rbase + out_idx<NX, 1, P - 1>(g, b, k);
        x[b * rlb + k] = epilogue(epi, x[b * rlb + k], p[b * rlb + k], V[b * rlb + k], s);
        out.st(vec, gi, x[b * rlb + k]);
      }
    run_line_sep<NX, -1, LX>(x, g, tcf, line);
    WaveSync::sync();
    write_out_sep<NX, -1, P - 1, LX>(x, g, line);
  }
  __syncthreads();  // the transpose
  // ---- forward y FFT of the columns (pass A of step j+1) into plane z of tb
#pragma unroll
  for (int cr = 0; cr < NR; cr++) {
    const int cl = cr * RX + ln;
    cx<R> *line = buf + LX::f(cl);
    cx<R> y[E];
    read_in_sep<NX, -1, 0, LY>(y, g, line);
    run_line_sep<NX, -1, LY>(y, g, tcf, line);
    cx<R> *dst = tb + vec * N + pbase + (long)cl * NX;
#pragma unroll
    for (int b = 0; b < E / rlf; b++)
#pragma unroll
      for (int k = 0; k < rlf; k++) dst[out_idx<NX, -1, P - 1>(g, b, k)] = y[b * rlf + k];
  }
}

// Pass C fused with the next pass A, with the plane held in registers and
// transposed through a buffer of NX/NH columns: k_plane_bwd_w's and
// k_plane_fwd_w's chunked forms back to back, same arithmetic as
// k_plane_bwd_fwd_w.  The buffer is 1/NH of a plane, so NH = 2 at complex64
// 128^2 (67 KB) lets two workgroups share a CU and one's HBM phases run
// under the other's FFTs; waves_per_eu holds the register budget to what two
// resident workgroups leave (T / 128 waves per SIMD: 128 VGPRs at T = 512).
template <typename R, int NX, int T, int NH>
__global__ void __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(T / 128)))
k_plane_bwd_fwd_h(cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi, long N, int nz,
                  const cx<R> *twg) {
  using PL = FftPlan<NX>;
  constexpr int G = PL::G, E = PL::E, P = PL::P;
  static_assert(64 % G == 0 && NX % 8 == 0, "lines must not straddle wavefronts");
  static_assert(NH > 1 && NX % NH == 0, "the chunked kernel splits the plane into NH chunks");
  constexpr int W = NX / NH;
  using L = PlaneLds<R, NX, NH>;
  constexpr int LDc = L::LDc;
  constexpr int RX = T / G;
  constexpr int NR = (NX + RX - 1) / RX;
  constexpr int NCR = (W + RX - 1) / RX;
  constexpr int RS = NX + 1;
  static_assert(RX * RS <= NX * LDc, "row exchange slices exceed the chunk buffer");
  extern __shared__ char smem_raw[];
  cx<R> *buf = reinterpret_cast<cx<R> *>(smem_raw);
  cx<R> *tw = buf + NX * LDc;
  const long vec = blockIdx.x / nz;
  const long z = blockIdx.x % nz;
  const long pbase = z * (long)NX * NX;
  const int tl = threadIdx.x / G, g0 = threadIdx.x % G;
  const int g = g0;
  constexpr int r0b = Pass<NX, 1>::radix(0);
  constexpr int rlb = Pass<NX, 1>::radix(P - 1);
  constexpr int rlf = Pass<NX, -1>::radix(P - 1);
  TwTable<NX, 1, R> tcb;
  TwTable<NX, -1, R> tcf;
  cx<R> v[NR][E];  // the plane's rows, x in registers
  // ---- inverse y FFT of the columns, chunk by chunk; each thread gathers
  // the chunk's x positions of the rows it owns
#pragma unroll
  for (int c = 0; c < NH; c++) {
    cx<R> w[NCR][E];
#pragma unroll
    for (int cr = 0; cr < NCR; cr++) {
      const int cl = cr * RX + tl;
      if (cl < W) {
        const cx<R> *src = tb + vec * N + pbase + (long)(c * W + cl) * NX;
#pragma unroll
        for (int b = 0; b < E / r0b; b++)
#pragma unroll
          for (int k = 0; k < r0b; k++) w[cr][b * r0b + k] = src[in_idx<NX, 1, 0>(g, b, k)];
      }
    }
    if (c == 0) load_tw<R, NX>(tw, twg);
    __syncthreads();  // twiddles loaded / the previous chunk's gather is done
    tcb.init(g, tw);
#pragma unroll
    for (int cr = 0; cr < NCR; cr++) {
      const int cl = cr * RX + tl;
      const bool act = cl < W;
      const int cs = act ? cl : 0;
      auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(i) + cs]; };
      run_passes_c<NX, 1, 0, P, WaveSync>(w[cr], g, act, tcb, at);
      WaveSync::sync();
      write_out<NX, 1, P - 1>(w[cr], g, act, at);
    }
    __syncthreads();  // the transpose
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int row = r * RX + tl;
      if (row < NX) {
#pragma unroll
        for (int b = 0; b < E / r0b; b++)
#pragma unroll
          for (int k = 0; k < r0b; k++) {
            const int x = in_idx<NX, 1, 0>(g, b, k);
            if (x / W == c) v[r][b * r0b + k] = buf[L::rowbase(row) + (x - c * W)];
          }
      }
    }
  }
  __syncthreads();  // the buffer is free for the row exchanges
  // ---- rows: inverse x FFT, epilogue, store psi_j, forward x FFT of psi_j.
  // Exchanges use slice tl of the buffer.  V and psi_{j-1} are loaded after
  // the FFT (the other workgroup on the CU covers the latency); psi is in
  // planar form (re and im arrays), and no partial sums are formed (the
  // filter loop asks for none).
  const R alpha = (R)epi.alpha, beta = (R)epi.beta;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int g = opaque_int(g0);  // per-round index math (see opaque_int)
    tcb.init(g, tw);
    tcf.init(g, tw);
    const int row = r * RX + tl;
    const bool act = row < NX;
    const int rs = act ? row : 0;
    auto at = [&](int i) -> cx<R> & { return buf[tl * RS + lds_swz(i)]; };
    run_passes_c<NX, 1, 0, P, WaveSync>(v[r], g, act, tcb, at);
    if (act) {
      const long ro = pbase + (long)rs * NX;
      const R *pr = pin.re + vec * pin.vstride + ro, *pi = pin.im + vec * pin.vstride + ro;
      R *qr = out.re + vec * out.vstride + ro, *qi = out.im + vec * out.vstride + ro;
      const R *vp = epi.pot + ro;
      R V[E], pre[E], pim[E];
#pragma unroll
      for (int b = 0; b < E / rlb; b++)
#pragma unroll
        for (int k = 0; k < rlb; k++) {
          const int x = out_idx<NX, 1, P - 1>(g, b, k);
          V[b * rlb + k] = vp[x];
          pre[b * rlb + k] = pr[x];
          pim[b * rlb + k] = pi[x];
        }
#pragma unroll
      for (int b = 0; b < E / rlb; b++)
#pragma unroll
        for (int k = 0; k < rlb; k++) {
          const int e = b * rlb + k, x = out_idx<NX, 1, P - 1>(g, b, k);
          cx<R> o;  // the arithmetic of epilogue()
          o.re = alpha * (v[r][e].re + V[e] * pre[e]) - beta * pre[e];
          o.im = alpha * (v[r][e].im + V[e] * pim[e]) - beta * pim[e];
          qr[x] = o.re;
          qi[x] = o.im;
          v[r][e] = o;
        }
    }
    // psi_j's row back through the slice into the forward input order
    WaveSync::sync();
    write_out<NX, 1, P - 1>(v[r], g, act, at);
    WaveSync::sync();
    read_in<NX, -1, 0>(v[r], g, act, at);
    run_passes_c<NX, -1, 0, P, WaveSync>(v[r], g, act, tcf, at);
  }
  // ---- forward y FFT of the columns (pass A of step j+1), chunk by chunk
#pragma unroll
  for (int c = 0; c < NH; c++) {
    __syncthreads();  // row exchanges / the previous chunk's columns are done
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int row = r * RX + tl;
      if (row < NX) {
#pragma unroll
        for (int b = 0; b < E / rlf; b++)
#pragma unroll
          for (int k = 0; k < rlf; k++) {
            const int kx = out_idx<NX, -1, P - 1>(g, b, k);
            if (kx / W == c) buf[L::rowbase(row) + (kx - c * W)] = v[r][b * rlf + k];
          }
      }
    }
    __syncthreads();  // the transpose
#pragma unroll
    for (int cr = 0; cr < NCR; cr++) {
      const int g = opaque_int(g0);
      tcf.init(g, tw);
      const int cl = cr * RX + tl;
      const bool act = cl < W;
      const int cs = act ? cl : 0;
      cx<R> y[E];
      auto at = [&](int i) -> cx<R> & { return buf[L::rowbase(i) + cs]; };
      read_in<NX, -1, 0>(y, g, act, at);
      run_passes_c<NX, -1, 0, P, WaveSync>(y, g, act, tcf, at);
      if (act) {
        cx<R> *dst = tb + vec * N + pbase + (long)(c * W + cs) * NX;
#pragma unroll
        for (int b = 0; b < E / rlf; b++)
#pragma unroll
          for (int k = 0; k < rlf; k++) dst[out_idx<NX, -1, P - 1>(g, b, k)] = y[b * rlf + k];
      }
    }
  }
}

// ================================================================= dispatch
template <int n>
constexpr bool plane_colmajor() {
  return 64 % FftPlan<n>::G == 0 && n % 8 == 0;
}
template <typename R, int NX>
constexpr size_t plane_lds() {
  if constexpr (plane_colmajor<NX>())
    return sizeof(cx<R>) * ((size_t)NX * PlaneLds<R, NX, plane_chunks<R, NX>()>::LDc + NX);
  return sizeof(cx<R>) * ((size_t)NX * PlaneCfg<NX>::LD + NX);
}
template <typename R, int NX>
constexpr int plane_threads() {
  if constexpr (plane_colmajor<NX>()) {
    // lines per round limited so that the row-exchange slices fit the chunk buffer
    constexpr int G = FftPlan<NX>::G;
    constexpr int LDc = NX / plane_chunks<R, NX>() + 1;
    // whole-plane buffers take one workgroup per CU: 1024 threads (16 waves)
    // hide more latency and hold half a plane's points per thread of 512
    int t = sizeof(R) == 4 && NX >= 96 && plane_chunks<R, NX>() == 1 ? 1024
                                                                      : (NX >= 64 ? 512 : 256);
    if (plane_chunks<R, NX>() > 1)
      while (t > 64 && (t / G) * (NX + 1) > NX * LDc) t -= 64;
    return t;
  }
  return NX >= 64 ? 512 : 256;
}

template <typename R>
struct HOps {
  // returns false if size unsupported
  static bool rows_fwd(int nx, VecView<R> in, cx<R> *tb, long nrows, long N, int nvec,
                       const cx<R> *tw, hipStream_t st);
  static bool rows_bwd(int nx, const cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi,
                       long nrows, long N, int nvec, const cx<R> *tw, hipStream_t st,
                       long *nblk);
  static bool lines(int nl, int S, cx<R> *tb, long N, int nx, long es, long os, int nouter,
                    int nvec, const cx<R> *tw, hipStream_t st);
  static bool pencil(int nz, cx<R> *tb, long N, int nx, int ny, int nvec, const cx<R> *tw,
                     const double *kxx, const double *kyy, const double *kzz, double ekin,
                     double kscale, double pden, hipStream_t st);
  static bool plane_fwd(int nxy, VecView<R> in, cx<R> *tb, long N, int nz, int nvec,
                        const cx<R> *tw, hipStream_t st);
  static bool plane_bwd(int nxy, const cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi,
                        long N, int nz, int nvec, const cx<R> *tw, hipStream_t st, long *nblk);
  // k_rows_bwd + the next step's k_rows_fwd in one kernel
  static bool rows_bwd_fwd(int nx, cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi,
                           long nrows, long N, int nvec, const cx<R> *tw, hipStream_t st);
  // pass C fused with the next pass A; false where the plane needs chunks.
  // form (SEP_FUSED_*): the kernel variant, where the grid has it; tws: the
  // separable form's twiddle table (GridDev::twsep32)
  static bool plane_bwd_fwd(int nxy, cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi,
                            long N, int nz, int nvec, const cx<R> *tw, const cx<R> *tws,
                            hipStream_t st, int form);
};

#define SEP_LINE_SIZES(X) X(16) X(24) X(32) X(36) X(48) X(64) X(80) X(96) X(128) X(160)

template <typename R>
bool HOps<R>::rows_fwd(int nx, VecView<R> in, cx<R> *tb, long nrows, long N, int nvec,
                       const cx<R> *tw, hipStream_t st) {
  switch (nx) {
#define C(n)                                                                              \
  case n: {                                                                               \
    constexpr int T = 256, ROWS = T / FftPlan<n>::G;                                      \
    dim3 grid((unsigned)((nrows + ROWS - 1) / ROWS), (unsigned)nvec);                     \
    hipLaunchKernelGGL((k_rows_fwd<R, n, T>), grid, dim3(T), 0, st, in, tb, nrows, N, tw); \
    return true;                                                                          \
  }
    SEP_LINE_SIZES(C)
#undef C
  }
  return false;
}

template <typename R>
bool HOps<R>::rows_bwd(int nx, const cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi,
                       long nrows, long N, int nvec, const cx<R> *tw, hipStream_t st,
                       long *nblk) {
  switch (nx) {
#define C(n)                                                                                \
  case n: {                                                                                 \
    constexpr int T = 256, ROWS = T / FftPlan<n>::G;                                        \
    dim3 grid((unsigned)((nrows + ROWS - 1) / ROWS), (unsigned)nvec);                       \
    if (nblk) *nblk = (long)grid.x * grid.y;                                                \
    hipLaunchKernelGGL((k_rows_bwd<R, n, T>), grid, dim3(T), 0, st, tb, pin, out, epi, nrows, \
                       N, tw);                                                              \
    return true;                                                                            \
  }
    SEP_LINE_SIZES(C)
#undef C
  }
  return false;
}

template <typename R>
bool HOps<R>::rows_bwd_fwd(int nx, cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi,
                           long nrows, long N, int nvec, const cx<R> *tw, hipStream_t st) {
  switch (nx) {
#define C(n)                                                                                 \
  case n: {                                                                                  \
    constexpr int T = 256, ROWS = T / FftPlan<n>::G;                                         \
    dim3 grid((unsigned)((nrows + ROWS - 1) / ROWS), (unsigned)nvec);                        \
    hipLaunchKernelGGL((k_rows_bwd_fwd<R, n, T>), grid, dim3(T), 0, st, tb, pin, out, epi,    \
                       nrows, N, tw);                                                        \
    return true;                                                                             \
  }
    SEP_LINE_SIZES(C)
#undef C
  }
  return false;
}

template <typename R>
bool HOps<R>::lines(int nl, int S, cx<R> *tb, long N, int nx, long es, long os, int nouter,
                    int nvec, const cx<R> *tw, hipStream_t st) {
  constexpr int W = sizeof(R) == 8 ? 16 : 32;
  dim3 grid((unsigned)((nx + W - 1) / W), (unsigned)nouter, (unsigned)nvec);
  switch (nl) {
#define C(n)                                                                                   \
  case n:                                                                                      \
    if (S < 0)                                                                                 \
      hipLaunchKernelGGL((k_lines<R, n, W, -1>), grid, dim3(W * FftPlan<n>::G), 0, st, tb, N, nx, \
                         es, os, tw);                                                          \
    else                                                                                       \
      hipLaunchKernelGGL((k_lines<R, n, W, 1>), grid, dim3(W * FftPlan<n>::G), 0, st, tb, N, nx,  \
                         es, os, tw);                                                          \
    return true;
    SEP_LINE_SIZES(C)
#undef C
  }
  return false;
}

// pden > 0: preconditioner factor kscale / (kin + pden), FP64 only
template <typename R>
bool HOps<R>::pencil(int nz, cx<R> *tb, long N, int nx, int ny, int nvec, const cx<R> *tw,
                     const double *kxx, const double *kyy, const double *kzz, double ekin,
                     double kscale, double pden, hipStream_t st) {
  constexpr int W = sizeof(R) == 8 ? 16 : 32;
  dim3 grid((unsigned)((nx + W - 1) / W), (unsigned)ny, (unsigned)nvec);
  if (pden > 0.0 && sizeof(R) != 8) return false;
  switch (nz) {
#define C(n)                                                                                  \
  case n:                                                                                     \
    if constexpr (sizeof(R) == 8) {                                                           \
      if (pden > 0.0) {                                                                       \
        hipLaunchKernelGGL((k_pencil<R, n, W, true>), grid, dim3(W * FftPlan<n>::G), 0, st, tb, \
                           N, nx, ny, tw, kxx, kyy, kzz, ekin, kscale, pden);                 \
        return true;                                                                          \
      }                                                                                       \
    }                                                                                         \
    hipLaunchKernelGGL((k_pencil<R, n, W>), grid, dim3(W * FftPlan<n>::G), 0, st, tb, N, nx, ny, \
                       tw, kxx, kyy, kzz, ekin, kscale, 0.0);                                 \
    return true;
    SEP_LINE_SIZES(C)
#undef C
  }
  return false;
}

// Plane-kernel lengths per precision: those whose kernels fit the register
// file without spilling (80 needs 20 points per thread; FP64 128 would hold
// a half plane of 64-bit values per thread).  Other lengths take the
// row + strided-line path.
#define SEP_PLANE_SIZES_F32(X) X(16) X(24) X(32) X(36) X(48) X(64) X(96) X(128)
#define SEP_PLANE_SIZES_F64(X) X(16) X(24) X(32) X(36) X(48) X(64) X(96)

template <typename R, int n>
static void plane_fwd_launch(VecView<R> in, cx<R> *tb, long N, int nz, int nvec, const cx<R> *tw,
                             hipStream_t st) {
  constexpr int T = plane_threads<R, n>();
  constexpr size_t lds = plane_lds<R, n>();
  static_assert(lds <= 160 * 1024, "plane buffer exceeds the LDS of a gfx950 workgroup");
  static bool attr = false;
  if constexpr (plane_colmajor<n>()) {
    if (!attr) {
      hipFuncSetAttribute((const void *)k_plane_fwd_w<R, n, T>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      attr = true;
    }
    if (nvec == 0) return;  // attribute only (before graph capture)
    hipLaunchKernelGGL((k_plane_fwd_w<R, n, T>), dim3((unsigned)(nz * nvec)), dim3(T), lds, st, in,
                       tb, N, nz, tw);
  } else {
    if (!attr) {
      hipFuncSetAttribute((const void *)k_plane_fwd<R, n, T>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      attr = true;
    }
    if (nvec == 0) return;  // attribute only (before graph capture)
    hipLaunchKernelGGL((k_plane_fwd<R, n, T>), dim3((unsigned)(nz * nvec)), dim3(T), lds, st, in,
                       tb, N, nz, tw);
  }
}
template <typename R, int n>
static void plane_bwd_launch(const cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi, long N,
                             int nz, int nvec, const cx<R> *tw, hipStream_t st) {
  constexpr int T = plane_threads<R, n>();
  constexpr size_t lds = plane_lds<R, n>();
  static_assert(lds <= 160 * 1024, "plane buffer exceeds the LDS of a gfx950 workgroup");
  static bool attr = false;
  if constexpr (plane_colmajor<n>()) {
    if (!attr) {
      hipFuncSetAttribute((const void *)k_plane_bwd_w<R, n, T>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      attr = true;
    }
    if (nvec == 0) return;  // attribute only (before graph capture)
    hipLaunchKernelGGL((k_plane_bwd_w<R, n, T>), dim3((unsigned)(nz * nvec)), dim3(T), lds, st,
                       tb, pin, out, epi, N, nz, tw);
  } else {
    if (!attr) {
      hipFuncSetAttribute((const void *)k_plane_bwd<R, n, T>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      attr = true;
    }
    if (nvec == 0) return;  // attribute only (before graph capture)
    hipLaunchKernelGGL((k_plane_bwd<R, n, T>), dim3((unsigned)(nz * nvec)), dim3(T), lds, st, tb,
                       pin, out, epi, N, nz, tw);
  }
}

// Grids whose fused kernel also has the chunked two-workgroups-per-CU form
// (k_plane_bwd_fwd_h): where the whole-plane buffer takes more than half the
// LDS and half a plane fits the registers of 512 threads.
template <typename R, int n>
constexpr bool plane_fused_h() {
  return sizeof(R) == 4 && n == 128;
}

// Grids whose fused kernel has the separable-placement form (k_plane_bwd_fwd_s).
template <typename R, int n>
constexpr bool plane_fused_s() {
  return sizeof(R) == 4 && n == 128;
}

template <typename R, int n>
static bool plane_bwd_fwd_launch(cx<R> *tb, VecView<R> pin, VecView<R> out, Epi<R> epi, long N,
                                 int nz, int nvec, const cx<R> *tw, const cx<R> *tws,
                                 hipStream_t st, int form) {
  if constexpr (plane_colmajor<n>() && plane_chunks<R, n>() == 1) {
    if constexpr (plane_fused_s<R, n>()) {
      if (form == SEP_FUSED_PLANE && !epi.part && !epi.so_map && tws) {
        constexpr int T = 1024;
        constexpr size_t lds = plane_sep_lds();
        static_assert(lds <= 160 * 1024, "plane buffer exceeds the LDS of a gfx950 workgroup");
        static bool attr = false;
        if (!attr) {
          hipFuncSetAttribute((const void *)k_plane_bwd_fwd_s<T>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
          attr = true;
        }
        if (nvec == 0) return true;  // attribute only (before graph capture)
        hipLaunchKernelGGL((k_plane_bwd_fwd_s<T>), dim3((unsigned)(nz * nvec)), dim3(T), lds, st,
                           tb, pin, out, epi, N, nz, tws);
        return true;
      }
    }
    if constexpr (plane_fused_h<R, n>()) {
      if (form == SEP_FUSED_CHUNKED && !epi.part && !epi.so_map && !pin.c && pin.im && !out.c &&
          out.im) {
        constexpr int T = 512, NH = 2;
        constexpr size_t lds = sizeof(cx<R>) * ((size_t)n * PlaneLds<R, n, NH>::LDc + n);
        static_assert(2 * lds <= 160 * 1024, "two chunk buffers must share the LDS of a CU");
        static bool attr = false;
        if (!attr) {
          hipFuncSetAttribute((const void *)k_plane_bwd_fwd_h<R, n, T, NH>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
          attr = true;
        }
        if (nvec == 0) return true;  // attribute only (before graph capture)
        hipLaunchKernelGGL((k_plane_bwd_fwd_h<R, n, T, NH>), dim3((unsigned)(nz * nvec)), dim3(T),
                           lds, st, tb, pin, out, epi, N, nz, tw);
        return true;
      }
    }
    constexpr int T = plane_threads<R, n>();
    constexpr size_t lds = plane_lds<R, n>();
    static_assert(lds <= 160 * 1024, "plane buffer exceeds the LDS of a gfx950 workgroup");
    static bool attr = false;
    if (!attr) {
      hipFuncSetAttribute((const void *)k_plane_bwd_fwd_w<R, n, T>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      attr = true;
    }
    if (nvec == 0) return true;  // attribute only (before graph capture)
    hipLaunchKernelGGL((k_plane_bwd_fwd_w<R, n, T>), dim3((unsigned)(nz * nvec)), dim3(T), lds,
                       st, tb, pin, out, epi, N, nz, tw);
    return true;
  }
  return false;
}

bool plane_fused_form_supported(int n, int prec, int form) {
  if (form == SEP_FUSED_PLANE) return true;
  if (prec == SEP_PREC_FP64) {
    switch (n) {
#define C(x) \
  case x:    \
    return form == SEP_FUSED_CHUNKED ? plane_fused_h<double, x>() : plane_fused_s<double, x>();
      SEP_PLANE_SIZES_F64(C)
#undef C
    }
    return false;
  }
  switch (n) {
#define C(x) \
  case x:    \
    return form == SEP_FUSED_CHUNKED ? plane_fused_h<float, x>() : plane_fused_s<float, x>();
    SEP_PLANE_SIZES_F32(C)
#undef C
  }
  return false;
}

template <>
bool HOps<float>::plane_bwd_fwd(int nxy, cx<float> *tb, VecView<float> pin, VecView<float> out,
                                Epi<float> epi, long N, int nz, int nvec, const cx<float> *tw,
                                const cx<float> *tws, hipStream_t st, int form) {
  switch (nxy) {
#define C(n) \
  case n:    \
    return plane_bwd_fwd_launch<float, n>(tb, pin, out, epi, N, nz, nvec, tw, tws, st, form);
    SEP_PLANE_SIZES_F32(C)
#undef C
  }
  return false;
}
template <>
bool HOps<double>::plane_bwd_fwd(int nxy, cx<double> *tb, VecView<double> pin,
                                 VecView<double> out, Epi<double> epi, long N, int nz, int nvec,
                                 const cx<double> *tw, const cx<double> *tws, hipStream_t st,
                                 int form) {
  switch (nxy) {
#define C(n) \
  case n:    \
    return plane_bwd_fwd_launch<double, n>(tb, pin, out, epi, N, nz, nvec, tw, tws, st, form);
    SEP_PLANE_SIZES_F64(C)
#undef C
  }
  return false;
}

template <>
bool HOps<float>::plane_fwd(int nxy, VecView<float> in, cx<float> *tb, long N, int nz, int nvec,
                            const cx<float> *tw, hipStream_t st) {
  switch (nxy) {
#define C(n)                                                 \
  case n:                                                    \
    plane_fwd_launch<float, n>(in, tb, N, nz, nvec, tw, st); \
    return true;
    SEP_PLANE_SIZES_F32(C)
#undef C
  }
  return false;
}
template <>
bool HOps<double>::plane_fwd(int nxy, VecView<double> in, cx<double> *tb, long N, int nz,
                             int nvec, const cx<double> *tw, hipStream_t st) {
  switch (nxy) {
#define C(n)                                                  \
  case n:                                                     \
    plane_fwd_launch<double, n>(in, tb, N, nz, nvec, tw, st); \
    return true;
    SEP_PLANE_SIZES_F64(C)
#undef C
  }
  return false;
}
template <>
bool HOps<float>::plane_bwd(int nxy, const cx<float> *tb, VecView<float> pin, VecView<float> out,
                            Epi<float> epi, long N, int nz, int nvec, const cx<float> *tw,
                            hipStream_t st, long *nblk) {
  if (nblk) *nblk = (long)nz * nvec;
  switch (nxy) {
#define C(n)                                                                 \
  case n:                                                                    \
    plane_bwd_launch<float, n>(tb, pin, out, epi, N, nz, nvec, tw, st);      \
    return true;
    SEP_PLANE_SIZES_F32(C)
#undef C
  }
  return false;
}
template <>
bool HOps<double>::plane_bwd(int nxy, const cx<double> *tb, VecView<double> pin,
                             VecView<double> out, Epi<double> epi, long N, int nz, int nvec,
                             const cx<double> *tw, hipStream_t st, long *nblk) {
  if (nblk) *nblk = (long)nz * nvec;
  switch (nxy) {
#define C(n)                                                                 \
  case n:                                                                    \
    plane_bwd_launch<double, n>(tb, pin, out, epi, N, nz, nvec, tw, st);     \
    return true;
    SEP_PLANE_SIZES_F64(C)
#undef C
  }
  return false;
}

// ----------------------------------------------------------------- public
// Plane kernels of these lengths store Tbuf column-major ([kx][ky]).
static bool plane_is_colmajor(int n) {
  switch (n) {
#define C(x) \
  case x:    \
    return plane_colmajor<x>();
    SEP_LINE_SIZES(C)
#undef C
  }
  return false;
}

bool hop_supported(int n) {
  switch (n) {
#define C(x) case x:
    SEP_LINE_SIZES(C)
#undef C
    return true;
  }
  return false;
}

bool plane_supported(int nx, int ny, int prec) {
  if (nx != ny) return false;
  if (prec == SEP_PREC_FP64) {
    switch (nx) {
#define C(x) case x:
      SEP_PLANE_SIZES_F64(C)
#undef C
      return true;
    }
    return false;
  }
  switch (nx) {
#define C(x) case x:
    SEP_PLANE_SIZES_F32(C)
#undef C
    return true;
  }
  return false;
}

// One H application (see file header) on nvec complex vectors.  With ev,
// events are recorded before pass A and after passes A, B and C (ev[0..3]).
template <typename R>
int hpsi_apply(const GridDev &gd, const VecView<R> &in, const VecView<R> &out, cx<R> *tb,
               const R *pot, int nvec, double alpha, double beta, double *part, long *nblk,
               hipStream_t st, hipEvent_t *ev) {
  const int nx = gd.nx, ny = gd.ny, nz = gd.nz;
  const long N = gd.N;
  const cx<R> *twx = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[0] : gd.tw32[0]);
  const cx<R> *twy = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[1] : gd.tw32[1]);
  const cx<R> *twz = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[2] : gd.tw32[2]);
  const int prec = sizeof(R) == 8 ? SEP_PREC_FP64 : SEP_PREC_FP32;
  const bool plane = gd.force_lines[prec] ? false : plane_supported(nx, ny, prec);
  if (ev) hipEventRecord(ev[0], st);
  // pass A
  if (plane) {
    if (!HOps<R>::plane_fwd(nx, in, tb, N, nz, nvec, twx, st)) return SEP_ERR_UNSUPPORTED;
  } else {
    if (!HOps<R>::rows_fwd(nx, in, tb, (long)ny * nz, N, nvec, twx, st)) return SEP_ERR_UNSUPPORTED;
    if (!HOps<R>::lines(ny, -1, tb, N, nx, nx, (long)nx * ny, nz, nvec, twy, st))
      return SEP_ERR_UNSUPPORTED;
  }
  if (ev) hipEventRecord(ev[1], st);
  // pass B (a column-major plane spectrum swaps the roles of x and y)
  const bool cm = plane && plane_is_colmajor(nx);
  if (!HOps<R>::pencil(nz, tb, N, cm ? ny : nx, cm ? nx : ny, nvec, twz, cm ? gd.kyy : gd.kxx,
                       cm ? gd.kxx : gd.kyy, gd.kzz, gd.ekin, gd.pden > 0.0 ? gd.pnum : 1.0,
                       gd.pden, st))
    return SEP_ERR_UNSUPPORTED;
  if (ev) hipEventRecord(ev[2], st);
  // pass C
  Epi<R> epi{pot, alpha, beta, part};
  if (plane) {
    if (!HOps<R>::plane_bwd(nx, tb, in, out, epi, N, nz, nvec, twx, st, nblk))
      return SEP_ERR_UNSUPPORTED;
  } else {
    if (!HOps<R>::lines(ny, 1, tb, N, nx, nx, (long)nx * ny, nz, nvec, twy, st))
      return SEP_ERR_UNSUPPORTED;
    if (!HOps<R>::rows_bwd(nx, tb, in, out, epi, (long)ny * nz, N, nvec, twx, st, nblk))
      return SEP_ERR_UNSUPPORTED;
  }
  if (ev) hipEventRecord(ev[3], st);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// Split Newton step: the filter loop runs pass A once on psi_0 (hpsi_fwd),
// then per step pass B and pass C fused with the next step's pass A
// (hpsi_step with fwd_next), the spectrum of psi_j staying in tb between
// steps.  Plane family: the fused plane kernel, where the plane kernels hold
// a whole plane in LDS; row + line family: the y lines of pass A move to the
// start of the next step and the x rows of pass C and the next pass A run as
// one kernel.  Both planes of the vectors must exist (the fused kernels keep
// the imaginary plane that a real-only view would drop).
template <typename R>
static const cx<R> *sep_tws(const GridDev &gd) {
  return sizeof(R) == 4 ? reinterpret_cast<const cx<R> *>(gd.twsep32) : nullptr;
}

static bool split_plane(const GridDev &gd, int prec) {
  return !gd.force_lines[prec] && plane_supported(gd.nx, gd.ny, prec);
}

template <typename R>
bool hpsi_fused_supported(const GridDev &gd) {
  const int prec = sizeof(R) == 8 ? SEP_PREC_FP64 : SEP_PREC_FP32;
  if (gd.pden > 0.0) return false;
  if (!split_plane(gd, prec)) return hop_supported(gd.nx) && hop_supported(gd.ny);
  VecView<R> none{nullptr, nullptr, nullptr, 0};
  Epi<R> epi{nullptr, 0.0, 0.0, nullptr};
  return HOps<R>::plane_bwd_fwd(gd.nx, nullptr, none, none, epi, gd.N, gd.nz, 0, nullptr,
                                sep_tws<R>(gd), nullptr, gd.fused_form[prec]);
}

template <typename R>
int hpsi_fwd(const GridDev &gd, const VecView<R> &in, cx<R> *tb, int nvec, hipStream_t st) {
  const int prec = sizeof(R) == 8 ? SEP_PREC_FP64 : SEP_PREC_FP32;
  const cx<R> *twx = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[0] : gd.tw32[0]);
  if (!hpsi_fused_supported<R>(gd) || !in.im) return SEP_ERR_UNSUPPORTED;
  const bool ok = split_plane(gd, prec)
                      ? HOps<R>::plane_fwd(gd.nx, in, tb, gd.N, gd.nz, nvec, twx, st)
                      : HOps<R>::rows_fwd(gd.nx, in, tb, (long)gd.ny * gd.nz, gd.N, nvec, twx, st);
  if (!ok) return SEP_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}

// tb holds the spectrum of in (psi_{j-1}; for the row + line family its x
// transform only); out <- alpha (T + V) in - beta in and, with fwd_next, tb
// <- the same transform of out.  Events as hpsi_apply's ev[1..3]: before the
// spectral passes (the plane family's pass B; the line family's y lines, z
// pencils and y lines), after them, after the fused kernel; ev[0] is not
// recorded.
template <typename R>
int hpsi_step(const GridDev &gd, const VecView<R> &in, const VecView<R> &out, cx<R> *tb,
              const R *pot, int nvec, double alpha, double beta, bool fwd_next, hipStream_t st,
              hipEvent_t *ev, const SoEpi *so) {
  const int nx = gd.nx, ny = gd.ny, nz = gd.nz;
  const long N = gd.N;
  const int prec = sizeof(R) == 8 ? SEP_PREC_FP64 : SEP_PREC_FP32;
  const cx<R> *twx = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[0] : gd.tw32[0]);
  const cx<R> *twy = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[1] : gd.tw32[1]);
  const cx<R> *twz = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[2] : gd.tw32[2]);
  if (!hpsi_fused_supported<R>(gd) || !in.im || !out.im) return SEP_ERR_UNSUPPORTED;
  const bool plane = split_plane(gd, prec);
  // the spin-orbit add rides only in the FP64 whole-plane fused kernel
  if (so && (sizeof(R) != 8 || !plane || !fwd_next)) return SEP_ERR_UNSUPPORTED;
  if (ev) hipEventRecord(ev[1], st);
  if (!plane && !HOps<R>::lines(ny, -1, tb, N, nx, nx, (long)nx * ny, nz, nvec, twy, st))
    return SEP_ERR_UNSUPPORTED;
  const bool cm = plane && plane_is_colmajor(nx);
  if (!HOps<R>::pencil(nz, tb, N, cm ? ny : nx, cm ? nx : ny, nvec, twz, cm ? gd.kyy : gd.kxx,
                       cm ? gd.kxx : gd.kyy, gd.kzz, gd.ekin, 1.0, 0.0, st))
    return SEP_ERR_UNSUPPORTED;
  if (!plane && !HOps<R>::lines(ny, 1, tb, N, nx, nx, (long)nx * ny, nz, nvec, twy, st))
    return SEP_ERR_UNSUPPORTED;
  if (ev) hipEventRecord(ev[2], st);
  Epi<R> epi{pot, alpha, beta, nullptr};
  if (so) {
    epi.so_map = so->map;
    epi.so_g = so->g;
    epi.so_nsp = so->nsp;
    epi.so_alpha = so->alpha;
  }
  bool ok;
  if (plane)
    ok = fwd_next ? HOps<R>::plane_bwd_fwd(nx, tb, in, out, epi, N, nz, nvec, twx, sep_tws<R>(gd),
                                           st, gd.fused_form[prec])
                  : HOps<R>::plane_bwd(nx, tb, in, out, epi, N, nz, nvec, twx, st, nullptr);
  else
    ok = fwd_next
             ? HOps<R>::rows_bwd_fwd(nx, tb, in, out, epi, (long)ny * nz, N, nvec, twx, st)
             : HOps<R>::rows_bwd(nx, tb, in, out, epi, (long)ny * nz, N, nvec, twx, st, nullptr);
  if (!ok) return SEP_ERR_UNSUPPORTED;
  if (ev) hipEventRecord(ev[3], st);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
template bool hpsi_fused_supported<double>(const GridDev &);
template bool hpsi_fused_supported<float>(const GridDev &);
template int hpsi_fwd<double>(const GridDev &, const VecView<double> &, cx<double> *, int,
                              hipStream_t);
template int hpsi_fwd<float>(const GridDev &, const VecView<float> &, cx<float> *, int, hipStream_t);
template int hpsi_step<double>(const GridDev &, const VecView<double> &, const VecView<double> &,
                               cx<double> *, const double *, int, double, double, bool,
                               hipStream_t, hipEvent_t *, const SoEpi *);
template int hpsi_step<float>(const GridDev &, const VecView<float> &, const VecView<float> &,
                              cx<float> *, const float *, int, double, double, bool, hipStream_t,
                              hipEvent_t *, const SoEpi *);
bool hpsi_plane_family(const GridDev &gd, int prec) { return split_plane(gd, prec); }

// Kernel attributes of this grid's H.psi kernels, set outside any stream
// capture (hipFuncSetAttribute is not a stream operation).
template <typename R>
int hpsi_prepare(const GridDev &gd) {
  const int prec = sizeof(R) == 8 ? SEP_PREC_FP64 : SEP_PREC_FP32;
  if (gd.force_lines[prec] || !plane_supported(gd.nx, gd.ny, prec)) return SEP_OK;
  const cx<R> *tw = reinterpret_cast<const cx<R> *>(sizeof(R) == 8 ? gd.tw64[0] : gd.tw32[0]);
  VecView<R> none{nullptr, nullptr, nullptr, 0};
  Epi<R> epi{nullptr, 0.0, 0.0, nullptr};
  HOps<R>::plane_fwd(gd.nx, none, nullptr, gd.N, gd.nz, 0, tw, nullptr);
  HOps<R>::plane_bwd(gd.nx, nullptr, none, none, epi, gd.N, gd.nz, 0, tw, nullptr, nullptr);
  for (int form : {SEP_FUSED_PLANE, SEP_FUSED_CHUNKED, SEP_FUSED_SWIZZLED})
    HOps<R>::plane_bwd_fwd(gd.nx, nullptr, none, none, epi, gd.N, gd.nz, 0, tw, sep_tws<R>(gd),
                           nullptr, form);
  return hipGetLastError() == hipSuccess ? SEP_OK : SEP_ERR_HIP;
}
template int hpsi_prepare<double>(const GridDev &);
template int hpsi_prepare<float>(const GridDev &);

template int hpsi_apply<double>(const GridDev &, const VecView<double> &, const VecView<double> &,
                                cx<double> *, const double *, int, double, double, double *,
                                long *, hipStream_t, hipEvent_t *);
template int hpsi_apply<float>(const GridDev &, const VecView<float> &, const VecView<float> &,
                               cx<float> *, const float *, int, double, double, double *, long *,
                               hipStream_t, hipEvent_t *);

}  // namespace sep
