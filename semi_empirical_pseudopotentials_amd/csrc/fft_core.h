// fft_core.h -- register/LDS FFT engine for CDNA4 (gfx950).
//
// A line of N complex points is transformed by G threads, each holding E = N/G
// points in registers.  The transform is a Stockham autosort decomposition
// N = r_0 r_1 ... r_{P-1}; pass p (stride Ns = r_0...r_{p-1}) computes, for
// each butterfly j < N/r_p with t = j % Ns:
//     v_k = in[j + k N/r_p] * w_N^{k t N/(Ns r_p)}      (k < r_p)
//     out[(j/Ns) Ns r_p + t + k Ns] = DFT_{r_p}(v)_k
// Thread g owns butterflies j = g + G b (b < E/r_p) in every pass, so each
// radix must divide E.  Between passes the line is exchanged through memory
// (LDS) by the caller; pass 0 reads and the last pass writes contiguous
// index ranges across threads, which is what makes the global loads/stores
// of the first/last pass coalesced.
//
// The inverse transform uses the reversed radix order: its pass 0 reads the
// very registers the forward last pass produced (same butterfly ownership),
// which is what lets the z-pencil kernel multiply by k^2 between the two
// transforms without a memory round trip.
//
// Sign convention: S = -1 forward (exp(-2 pi i jk/N), FFTW_FORWARD),
// S = +1 backward; unnormalised (the 1/N lives in the k^2 table).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace sep {

// Aligned to its full size so that LDS and global accesses of one element
// are single 8-byte (FP32) / 16-byte (FP64) operations.
template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R re, im;
};

// A batch of complex vectors, either interleaved (c) or as two real planes
// (re, im; im == nullptr reads as 0 and drops imaginary writes).  Vector v
// starts at element v * vstride.
template <typename R>
struct VecView {
  R *re;
  R *im;
  cx<R> *c;
  long vstride;
  __device__ __forceinline__ cx<R> ld(long v, long i) const {
    if (c) return c[v * vstride + i];
    cx<R> x;
    x.re = re[v * vstride + i];
    x.im = im ? im[v * vstride + i] : (R)0;
    return x;
  }
  __device__ __forceinline__ void st(long v, long i, cx<R> x) const {
    if (c) {
      c[v * vstride + i] = x;
      return;
    }
    re[v * vstride + i] = x.re;
    if (im) im[v * vstride + i] = x.im;
  }
};

template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// multiply by S*i  (S = +-1)
template <int S, typename R>
__device__ __forceinline__ cx<R> mul_si(cx<R> a) {
  if constexpr (S > 0) return {-a.im, a.re};
  else return {a.im, -a.re};
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return {a.re * s, a.im * s}; }

// ---------------------------------------------------------------- butterflies
template <int S, typename R>
__device__ __forceinline__ void dft2(cx<R> *v) {
  cx<R> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S, typename R>
__device__ __forceinline__ void dft4(cx<R> *v) {
  cx<R> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  cx<R> s13 = cadd(v[1], v[3]), d13 = mul_si<S>(csub(v[1], v[3]));
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13);
  v[3] = csub(d02, d13);
}

template <int S, typename R>
__device__ __forceinline__ void dft3(cx<R> *v) {
  const R c3 = (R)0.86602540378443864676;  // sin(2 pi / 3)
  cx<R> t1 = cadd(v[1], v[2]);
  cx<R> t2 = {v[0].re - (R)0.5 * t1.re, v[0].im - (R)0.5 * t1.im};
  cx<R> t3 = mul_si<S>(cscale(csub(v[1], v[2]), c3));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <int S, typename R>
__device__ __forceinline__ void dft5(cx<R> *v) {
  const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
  const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
  cx<R> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  cx<R> t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  cx<R> a1 = {v[0].re + c1 * t1.re + c2 * t2.re, v[0].im + c1 * t1.im + c2 * t2.im};
  cx<R> a2 = {v[0].re + c2 * t1.re + c1 * t2.re, v[0].im + c2 * t1.im + c1 * t2.im};
  cx<R> b1 = mul_si<S>(cx<R>{s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im});
  cx<R> b2 = mul_si<S>(cx<R>{s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im});
  v[0] = cadd(v[0], cadd(t1, t2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

// radix 8 = 2 x 4: y_{2q} = DFT4(v_k + v_{k+4}), y_{2q+1} = DFT4((v_k - v_{k+4}) w8^k)
template <int S, typename R>
__device__ __forceinline__ void dft8(cx<R> *v) {
  const R h = (R)0.70710678118654752440;
  cx<R> a[4], b[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    a[k] = cadd(v[k], v[k + 4]);
    b[k] = csub(v[k], v[k + 4]);
  }
  // w8^1 = h(1 + S i), w8^2 = S i, w8^3 = h(-1 + S i)
  b[1] = cscale(cadd(b[1], mul_si<S>(b[1])), h);
  b[2] = mul_si<S>(b[2]);
  b[3] = cscale(csub(mul_si<S>(b[3]), b[3]), h);
  dft4<S>(a);
  dft4<S>(b);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    v[2 * q] = a[q];
    v[2 * q + 1] = b[q];
  }
}

// radix 16 = 2 x 8
template <int S, typename R>
__device__ __forceinline__ void dft16(cx<R> *v) {
  // w16^k = cos(2 pi k/16) + S i sin(2 pi k/16)
  const R c[8] = {(R)1.0,
                  (R)0.92387953251128675613,
                  (R)0.70710678118654752440,
                  (R)0.38268343236508977173,
                  (R)0.0,
                  (R)-0.38268343236508977173,
                  (R)-0.70710678118654752440,
                  (R)-0.92387953251128675613};
  const R s[8] = {(R)0.0,
                  (R)0.38268343236508977173,
                  (R)0.70710678118654752440,
                  (R)0.92387953251128675613,
                  (R)1.0,
                  (R)0.92387953251128675613,
                  (R)0.70710678118654752440,
                  (R)0.38268343236508977173};
  cx<R> a[8], b[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    a[k] = cadd(v[k], v[k + 8]);
    cx<R> d = csub(v[k], v[k + 8]);
    if (k == 0)
      b[k] = d;
    else if (k == 4)
      b[k] = mul_si<S>(d);
    else
      b[k] = cmul(d, cx<R>{c[k], (R)S * s[k]});
  }
  dft8<S>(a);
  dft8<S>(b);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    v[2 * q] = a[q];
    v[2 * q + 1] = b[q];
  }
}

template <int r, int S, typename R>
__device__ __forceinline__ void dft(cx<R> *v) {
  if constexpr (r == 2) dft2<S>(v);
  else if constexpr (r == 3) dft3<S>(v);
  else if constexpr (r == 4) dft4<S>(v);
  else if constexpr (r == 5) dft5<S>(v);
  else if constexpr (r == 8) dft8<S>(v);
  else if constexpr (r == 16) dft16<S>(v);
  else static_assert(r == 2, "unsupported radix");
}

// ---------------------------------------------------------------- plans
// FftPlan<N>: E points per thread, G = N/E threads per line, P passes with
// radices R[0..P-1] (forward order); every radix divides E.
template <int N>
struct FftPlan;

#define SEP_PLAN(N_, E_, P_, ...)                         \
  template <>                                             \
  struct FftPlan<N_> {                                    \
    static constexpr int N = N_, E = E_, G = N_ / E_, P = P_; \
    static constexpr int R[P_] = {__VA_ARGS__};           \
  };

SEP_PLAN(8, 8, 1, 8)
SEP_PLAN(12, 12, 2, 4, 3)
SEP_PLAN(16, 4, 2, 4, 4)
SEP_PLAN(20, 20, 2, 4, 5)
SEP_PLAN(24, 12, 3, 4, 2, 3)
SEP_PLAN(32, 8, 2, 8, 4)
SEP_PLAN(36, 12, 3, 4, 3, 3)
SEP_PLAN(40, 20, 3, 4, 2, 5)
SEP_PLAN(48, 12, 3, 4, 4, 3)
SEP_PLAN(64, 8, 2, 8, 8)
SEP_PLAN(72, 12, 4, 4, 2, 3, 3)
SEP_PLAN(80, 20, 3, 4, 4, 5)
SEP_PLAN(96, 12, 4, 4, 4, 2, 3)
SEP_PLAN(100, 20, 3, 4, 5, 5)
SEP_PLAN(108, 12, 4, 4, 3, 3, 3)
SEP_PLAN(128, 8, 3, 8, 8, 2)
SEP_PLAN(144, 12, 4, 4, 4, 3, 3)
SEP_PLAN(160, 20, 4, 4, 4, 2, 5)
SEP_PLAN(192, 12, 4, 4, 4, 4, 3)
SEP_PLAN(200, 20, 4, 4, 2, 5, 5)
SEP_PLAN(256, 8, 3, 8, 8, 4)
#undef SEP_PLAN

// Pass descriptor in a given direction.  Forward uses R[p]; backward uses the
// reversed order R[P-1-p].
template <int N, int S>
struct Pass {
  using PL = FftPlan<N>;
  static constexpr int radix(int p) { return S < 0 ? PL::R[p] : PL::R[PL::P - 1 - p]; }
  static constexpr int ns(int p) {
    int s = 1;
    for (int q = 0; q < p; q++) s *= radix(q);
    return s;
  }
};

// Index read by element (b, k) of thread g in pass p.
template <int N, int S, int p>
__host__ __device__ __forceinline__ constexpr int in_idx(int g, int b, int k) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  return g + PL::G * b + k * (N / r);
}
// Index written by element (b, k) of thread g in pass p.
template <int N, int S, int p>
__host__ __device__ __forceinline__ constexpr int out_idx(int g, int b, int k) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  constexpr int Ns = Pass<N, S>::ns(p);
  int j = g + PL::G * b;
  return (j / Ns) * Ns * r + (j % Ns) + k * Ns;
}

// Twiddle + butterflies of pass p on the E registers of thread g.
// tw[m] = exp(-2 pi i m / N) (conjugated for S > 0).
template <int N, int S, int p, typename R>
__device__ __forceinline__ void compute_pass(cx<R> *v, int g, const cx<R> *tw) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  constexpr int Ns = Pass<N, S>::ns(p);
  constexpr int nb = PL::E / r;
#pragma unroll
  for (int b = 0; b < nb; b++) {
    if constexpr (p > 0) {
      int j = g + PL::G * b;
      int t = j % Ns;
#pragma unroll
      for (int k = 1; k < r; k++) {
        cx<R> w = tw[k * t * (N / (Ns * r))];
        if constexpr (S > 0) w.im = -w.im;
        v[b * r + k] = cmul(v[b * r + k], w);
      }
    }
    dft<r, S>(v + b * r);
  }
}

// LDS position of line element i: one pad slot every 8 elements, so the
// stride-8 access patterns of the radix-8/4 exchanges spread over the banks.
__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 3); }
// The same spreading without extra space: the low three bits are xor-ed
// with the next three (a permutation inside each aligned block of 8).
__host__ __device__ constexpr int lds_swz(int i) { return i ^ ((i >> 3) & 7); }
template <int N>
constexpr int lds_line() {
  return lds_pad(N - 1) + 1;
}

// Per-thread twiddle cache: the factors w_N^{k t N/(Ns r)} a thread needs in
// passes 1..P-1 depend only on its index g, so they are read from the table
// once per kernel and then reused for every line the thread transforms.
template <int N, int S>
struct TwLayout {
  using PL = FftPlan<N>;
  static constexpr int off(int p) {
    int o = 0;
    for (int q = 1; q < p; q++) o += (PL::E / Pass<N, S>::radix(q)) * (Pass<N, S>::radix(q) - 1);
    return o;
  }
  static constexpr int total = off(PL::P) > 0 ? off(PL::P) : 1;
};

template <int N, int S, typename R>
struct TwCache {
  cx<R> w[TwLayout<N, S>::total];
  template <int p>
  __device__ __forceinline__ void fill(int g, const cx<R> *tw) {
    if constexpr (p < FftPlan<N>::P) {
      using PL = FftPlan<N>;
      constexpr int r = Pass<N, S>::radix(p);
      constexpr int Ns = Pass<N, S>::ns(p);
      constexpr int o = TwLayout<N, S>::off(p);
#pragma unroll
      for (int b = 0; b < PL::E / r; b++) {
        const int t = (g + PL::G * b) % Ns;
#pragma unroll
        for (int k = 1; k < r; k++) {
          cx<R> x = tw[k * t * (N / (Ns * r))];
          if constexpr (S > 0) x.im = -x.im;
          w[o + b * (r - 1) + (k - 1)] = x;
        }
      }
      fill<p + 1>(g, tw);
    }
  }
  __device__ __forceinline__ void init(int g, const cx<R> *tw) { fill<1>(g, tw); }
  template <int p>
  __device__ __forceinline__ cx<R> get(int b, int k) const {
    constexpr int r = Pass<N, S>::radix(p);
    return w[TwLayout<N, S>::off(p) + b * (r - 1) + (k - 1)];
  }
};

// The same interface, reading each twiddle from the (LDS) table when it is
// used: fewer registers, for the FP64 kernels whose cache would spill.
template <int N, int S, typename R>
struct TwTable {
  const cx<R> *tw;
  int g;
  __device__ __forceinline__ void init(int g_, const cx<R> *tw_) { g = g_; tw = tw_; }
  template <int p>
  __device__ __forceinline__ cx<R> get(int b, int k) const {
    constexpr int r = Pass<N, S>::radix(p);
    constexpr int Ns = Pass<N, S>::ns(p);
    const int t = (g + FftPlan<N>::G * b) % Ns;
    cx<R> x = tw[k * t * (N / (Ns * r))];
    if constexpr (S > 0) x.im = -x.im;
    return x;
  }
};

// Twiddle source of the plane and row kernels: registers for FP32, the
// table for FP64 (a cache of 64-bit twiddles costs up to 96 VGPRs).
template <int N, int S, typename R>
using TwFor = typename std::conditional<sizeof(R) == 4, TwCache<N, S, R>, TwTable<N, S, R>>::type;

// compute_pass with the twiddles taken from a twiddle source (TwCache / TwTable).
template <int N, int S, int p, typename R, typename Tw>
__device__ __forceinline__ void compute_pass_c(cx<R> *v, const Tw &tc) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  constexpr int nb = PL::E / r;
#pragma unroll
  for (int b = 0; b < nb; b++) {
    if constexpr (p > 0) {
#pragma unroll
      for (int k = 1; k < r; k++) v[b * r + k] = cmul(v[b * r + k], tc.template get<p>(b, k));
    }
    dft<r, S>(v + b * r);
  }
}

// Exchange through a line buffer: write pass-p outputs, barrier, read pass
// p+1 inputs.  `at(i)` returns a reference into the line buffer.  The caller
// guarantees a barrier before the write if the buffer was being read.
template <int N, int S, int p, typename R, typename At>
__device__ __forceinline__ void write_out(const cx<R> *v, int g, bool act, At at) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  if (act) {
#pragma unroll
    for (int b = 0; b < PL::E / r; b++)
#pragma unroll
      for (int k = 0; k < r; k++) at(out_idx<N, S, p>(g, b, k)) = v[b * r + k];
  }
}
template <int N, int S, int p, typename R, typename At>
__device__ __forceinline__ void read_in(cx<R> *v, int g, bool act, At at) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  if (act) {
#pragma unroll
    for (int b = 0; b < PL::E / r; b++)
#pragma unroll
      for (int k = 0; k < r; k++) v[b * r + k] = at(in_idx<N, S, p>(g, b, k));
  }
}

// Synchronisation between the passes of one line: a block barrier, or a
// wavefront barrier when every line lives inside one wavefront (LDS
// operations of a wavefront complete in order; the fences and the scheduling
// barrier keep the compiler from moving accesses across it).
struct BlockSync {
  static __device__ __forceinline__ void sync() { __syncthreads(); }
};
struct WaveSync {
  static __device__ __forceinline__ void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// Runs passes p..pend-1 with LDS exchanges between them.  On entry v holds
// the pass-p inputs; on exit v holds the outputs of pass pend-1.  With
// BlockSync every thread of the block must call this.
template <int N, int S, int p, int pend, typename Sync = BlockSync, typename R, typename At>
__device__ __forceinline__ void run_passes(cx<R> *v, int g, bool act, const cx<R> *tw, At at) {
  if constexpr (p < pend) {
    compute_pass<N, S, p>(v, g, tw);
    if constexpr (p + 1 < pend) {
      Sync::sync();  // readers of the buffer are done
      write_out<N, S, p>(v, g, act, at);
      Sync::sync();
      read_in<N, S, p + 1>(v, g, act, at);
    }
    run_passes<N, S, p + 1, pend, Sync>(v, g, act, tw, at);
  }
}

// run_passes with the twiddles of a twiddle source (TwCache / TwTable).
template <int N, int S, int p, int pend, typename Sync = BlockSync, typename R, typename Tw,
          typename At>
__device__ __forceinline__ void run_passes_c(cx<R> *v, int g, bool act, const Tw &tc, At at) {
  if constexpr (p < pend) {
    compute_pass_c<N, S, p>(v, tc);
    if constexpr (p + 1 < pend) {
      Sync::sync();
      write_out<N, S, p>(v, g, act, at);
      Sync::sync();
      read_in<N, S, p + 1>(v, g, act, at);
    }
    run_passes_c<N, S, p + 1, pend, Sync>(v, g, act, tc, at);
  }
}

// ------------------------------------------------- separable line layouts
// A line layout F maps a line index i to an offset F::f(i) that is additive
// over disjoint bits, f(a + b) = f(a) + f(b) when a & b == 0 (a weighted sum
// of the index bits).  In every pass of a power-of-two Stockham plan the
// index of element (b, k) of thread g is the bitwise-disjoint sum of a thread
// part (g, 0, 0) and an element part (0, b, k), so its offset is a per-thread
// base plus a compile-time constant: one address per thread and pattern, the
// element offsets folded into the LDS instructions' immediate field.
template <int N, int S, int p, bool OUT, typename F>
constexpr bool sep_pass_ok() {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  for (int g = 0; g < PL::G; g++)
    for (int b = 0; b < PL::E / r; b++)
      for (int k = 0; k < r; k++) {
        const int i = OUT ? out_idx<N, S, p>(g, b, k) : in_idx<N, S, p>(g, b, k);
        const int t = OUT ? out_idx<N, S, p>(g, 0, 0) : in_idx<N, S, p>(g, 0, 0);
        const int c = OUT ? out_idx<N, S, p>(0, b, k) : in_idx<N, S, p>(0, b, k);
        if ((t & c) != 0 || t + c != i || F::f(i) != F::f(t) + F::f(c)) return false;
      }
  return true;
}
template <int N, int S, typename F, int p = 0>
constexpr bool sep_plan_ok() {
  if constexpr (p == FftPlan<N>::P) {
    return true;
  } else {
    return sep_pass_ok<N, S, p, false, F>() && sep_pass_ok<N, S, p, true, F>() &&
           sep_plan_ok<N, S, F, p + 1>();
  }
}

// Keeps the compiler from pairing the 8-byte LDS accesses on either side
// into ds_read2_b64 / ds_write2_b64: on gfx950 a single 8-byte access serves
// a half-wave per cycle (64 banks of 4 B; 2 cycles a wave), a paired one
// measured 8 cycles a wave whatever its two offsets, twice the cost of the
// two single accesses.  No instruction is emitted.
__device__ __forceinline__ void lds_unpaired() { asm volatile("" ::: "memory"); }

// write_out / read_in on a line at `line` laid out by F.
template <int N, int S, int p, typename F, typename R>
__device__ __forceinline__ void write_out_sep(const cx<R> *v, int g, cx<R> *line) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  cx<R> *t = line + F::f(out_idx<N, S, p>(g, 0, 0));
#pragma unroll
  for (int b = 0; b < PL::E / r; b++)
#pragma unroll
    for (int k = 0; k < r; k++) {
      t[F::f(out_idx<N, S, p>(0, b, k))] = v[b * r + k];
      lds_unpaired();
    }
}
template <int N, int S, int p, typename F, typename R>
__device__ __forceinline__ void read_in_sep(cx<R> *v, int g, const cx<R> *line) {
  using PL = FftPlan<N>;
  constexpr int r = Pass<N, S>::radix(p);
  const cx<R> *t = line + F::f(in_idx<N, S, p>(g, 0, 0));
#pragma unroll
  for (int b = 0; b < PL::E / r; b++)
#pragma unroll
    for (int k = 0; k < r; k++) {
      v[b * r + k] = t[F::f(in_idx<N, S, p>(0, b, k))];
      lds_unpaired();
    }
}

// Twiddles of passes 1..P-1 in a per-thread-index LDS table: thread g's
// factors (TwCache's order) at tab[g * SLOTS + slot], so that every twiddle
// read is the thread's base plus a compile-time slot.  host_table() builds
// the table (the values of the exp(-2 pi i m / N) table rounded to R, the
// imaginary part negated for S > 0, as TwTable reads them).
template <int N, int S, typename R>
struct TwSep {
  using PL = FftPlan<N>;
  static constexpr int SLOTS = TwLayout<N, S>::total;
  static constexpr int WORDS = SLOTS * PL::G;
  const cx<R> *base;
  static void host_table(cx<R> *tab) {
    for (int g = 0; g < PL::G; g++)
      for (int p = 1; p < PL::P; p++) {
        const int r = Pass<N, S>::radix(p), Ns = Pass<N, S>::ns(p);
        for (int b = 0; b < PL::E / r; b++)
          for (int k = 1; k < r; k++) {
            const int t = (g + PL::G * b) % Ns;
            const int m = k * t * (N / (Ns * r));
            const double ang = -6.283185307179586476925286766559 * (double)m / (double)N;
            cx<R> x;
            x.re = (R)std::cos(ang);
            x.im = (R)std::sin(ang);
            if (S > 0) x.im = -x.im;
            tab[g * SLOTS + TwLayout<N, S>::off(p) + b * (r - 1) + (k - 1)] = x;
          }
      }
  }
  __device__ __forceinline__ void init(int g, const cx<R> *tab) { base = tab + g * SLOTS; }
};

// The twiddles of pass p of a TwSep source, read into registers together
// (unpaired accesses issued back to back, one wait for all) and served to
// compute_pass_c.
template <int N, int S, int p, typename R>
struct TwSepPass {
  static constexpr int r = Pass<N, S>::radix(p);
  static constexpr int NW = p > 0 ? (FftPlan<N>::E / r) * (r - 1) : 0;
  cx<R> w[NW > 0 ? NW : 1];
  __device__ __forceinline__ void load(const TwSep<N, S, R> &t) {
#pragma unroll
    for (int i = 0; i < NW; i++) {
      w[i] = t.base[TwLayout<N, S>::off(p) + i];
      lds_unpaired();
    }
  }
  template <int q>
  __device__ __forceinline__ cx<R> get(int b, int k) const {
    return w[b * (r - 1) + (k - 1)];
  }
};

// run_passes_c on a line laid out by F with TwSep twiddles: the twiddles of
// pass p + 1 are read during the exchange after pass p.  tp holds pass p's.
template <int N, int S, int p, int pend, typename F, typename R>
__device__ __forceinline__ void run_passes_sep(cx<R> *v, int g, const TwSep<N, S, R> &tc,
                                               cx<R> *line, const TwSepPass<N, S, p, R> &tp) {
  if constexpr (p < pend) {
    compute_pass_c<N, S, p>(v, tp);
    if constexpr (p + 1 < pend) {
      WaveSync::sync();
      write_out_sep<N, S, p, F>(v, g, line);
      TwSepPass<N, S, p + 1, R> tn;
      tn.load(tc);
      WaveSync::sync();
      read_in_sep<N, S, p + 1, F>(v, g, line);
      run_passes_sep<N, S, p + 1, pend, F>(v, g, tc, line, tn);
    }
  }
}
// All passes of one line.
template <int N, int S, typename F, typename R>
__device__ __forceinline__ void run_line_sep(cx<R> *v, int g, const TwSep<N, S, R> &tc,
                                             cx<R> *line) {
  TwSepPass<N, S, 0, R> t0;  // pass 0 has no twiddles
  run_passes_sep<N, S, 0, FftPlan<N>::P, F>(v, g, tc, line, t0);
}

}  // namespace sep
