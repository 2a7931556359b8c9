// pq_fft.h -- LDS-resident FFT and DST-I building blocks for gfx950, with the
// transform length a compile-time constant.
//
// One complex FFT of length M is computed by a group of TPF = M / E threads
// (E complex elements per thread).  Stages are Stockham radix-R passes
// (R in {16, 8, 4, 2}, plus a final radix 3 when M = 3 * 2^a) run in place:
// every thread loads its inputs into registers, the group synchronises, and
// the results are written back.  LDS element addresses are padded with one
// element per 16 (pad16) so that strided and thread-blocked accesses spread
// over the banks.  With the plan known at compile time every loop unrolls
// and most LDS offsets fold into immediates.
//
// Two DST-I formulations, both computing the unnormalised
//     X_k = 2 sum_{j=1..N-1} x_j sin(pi j k / N)
// of two real rows a and b at once:
//   * DstSint (FP64): FFTPACK sint's formulation (qg3p_helmholtz.F90:
//     1380-1426): pre-twiddle into a length-N real sequence, one complex FFT
//     of length N for both rows, and a prefix sum for the odd outputs.  The
//     sum accumulates ~sqrt(N) eps relative to the FFT's output scale, which
//     is harmless in FP64 but exceeds 1e-6 in FP32.
//   * DstExt (FP32): the odd extension w (w_{2N-j} = -w_j) of a + i b, one
//     complex FFT of length 2N, X_a = -Im W, X_b = Re W (the transform of a
//     real odd sequence is imaginary).  No recurrence: ~eps log N accuracy.
#pragma once

#include <type_traits>

#include "pq_common.h"

namespace pq {

// ------------------------------------------------------------ plans
constexpr int ilog2c(int x) {
  int r = 0;
  while ((1 << r) < x) ++r;
  return r;
}

// Elements per thread for an FFT of length M in type T.
template <typename T> constexpr int choose_E(int M) {
  if (M % 3 == 0) return (M >= 768 && std::is_same<T, float>::value) ? 24 : 12;
  if (std::is_same<T, float>::value && M >= 2048) return 32;
  if (M >= 512) return 16;
  if (M >= 64) return 8;
  return 4;
}

struct StageArr {
  int nst;
  int rad[12];
  int lns[12];
};

// Radices of an FFT of length M with E elements per thread: powers of two
// (16, 8, 4, 2 as E allows), and for M = 3 * 2^a a final radix 3, or 24
// (= 8 x 3 in registers, one LDS round trip fewer) when E is a multiple of 24.
constexpr StageArr make_stages(int M, int E) {
  StageArr a{};
  const bool r24 = M % 3 == 0 && E % 24 == 0 && (M / 3) % 8 == 0;
  int rem = (M % 3 == 0) ? (r24 ? M / 24 : M / 3) : M;
  int lg = 0;
  a.nst = 0;
  while (rem > 1) {
    int r = 2;
    if (16 <= E && E % 16 == 0 && rem % 16 == 0) r = 16;
    else if (8 <= E && E % 8 == 0 && rem % 8 == 0) r = 8;
    else if (4 <= E && E % 4 == 0 && rem % 4 == 0) r = 4;
    a.rad[a.nst] = r;
    a.lns[a.nst] = lg;
    ++a.nst;
    lg += ilog2c(r);
    rem /= r;
  }
  if (M % 3 == 0) {
    a.rad[a.nst] = r24 ? 24 : 3;
    a.lns[a.nst] = lg;
    ++a.nst;
  }
  return a;
}

// EF > 0 overrides the elements per thread.
template <typename T, int M, int EF = 0> struct FftPlan {
  static constexpr int E = EF > 0 ? EF : choose_E<T>(M);
  static constexpr int TPF = M / E;
  static constexpr StageArr S = make_stages(M, E);
};

// Synchronise the threads of one FFT group.  A group of at most 64 threads
// lives inside one wavefront and owns its LDS region, so a wavefront-scope
// fence + wave barrier suffices; larger groups need the workgroup barrier
// (reached by every group of the workgroup in lockstep).
template <int TPF> __device__ __forceinline__ void group_sync() {
  if constexpr (TPF <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// ---------------------------------------------------------------- radices
template <typename T> __device__ __forceinline__ void dft2(cx<T> *v) {
  cx<T> a = v[0];
  v[0] = a + v[1];
  v[1] = a - v[1];
}

template <typename T> __device__ __forceinline__ void dft4(cx<T> *v) {
  cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
  cx<T> b0 = v[1] + v[3], b1 = mul_mi(v[1] - v[3]);
  v[0] = a0 + b0;
  v[1] = a1 + b1;
  v[2] = a0 - b0;
  v[3] = a1 - b1;
}

template <typename T> __device__ __forceinline__ void dft3(cx<T> *v) {
  const T c = T(0.86602540378443864676);  // sin(2 pi / 3)
  cx<T> s = v[1] + v[2];
  cx<T> d = mul_mi(v[1] - v[2]);
  cx<T> t = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
  v[0] = v[0] + s;
  v[1] = t + scale(d, c);
  v[2] = t - scale(d, c);
}

template <typename T> __device__ __forceinline__ void dft8(cx<T> *v) {
  const T h = T(0.70710678118654752440);
  cx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4(e);
  dft4(o);
  // o_k *= exp(-2 pi i k / 8)
  o[1] = {h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
  o[2] = mul_mi(o[2]);
  o[3] = {h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = e[k] + o[k];
    v[k + 4] = e[k] - o[k];
  }
}

template <typename T> __device__ __forceinline__ void dft16(cx<T> *v) {
  cx<T> e[8], o[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    e[k] = v[2 * k];
    o[k] = v[2 * k + 1];
  }
  dft8(e);
  dft8(o);
  // exp(-2 pi i k / 16), k = 0..7
  const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);
  const T h = T(0.70710678118654752440);
  const cx<T> w[8] = {{T(1), T(0)},  {c1, -s1},  {h, -h},  {s1, -c1},
                      {T(0), T(-1)}, {-s1, -c1}, {-h, -h}, {-c1, -s1}};
#pragma unroll
  for (int k = 1; k < 8; ++k) o[k] = cmul(o[k], w[k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = e[k] + o[k];
    v[k + 8] = e[k] - o[k];
  }
}

// 24 = 3 x 8 (Cooley-Tukey): DFT-8s of the inputs n1 + 3 n2 (n1 < 3),
// twiddles exp(-2 pi i n1 k2 / 24), then DFT-3s over n1 give X[k2 + 8 k1].
template <typename T> __device__ __forceinline__ void dft24(cx<T> *v) {
  // (cos, -sin) of 2 pi m / 24, m = 0..14
  const T c15 = T(0.96592582628906828675), s15 = T(0.25881904510252076235);
  const T c30 = T(0.86602540378443864676), h = T(0.70710678118654752440);
  const cx<T> w[15] = {{T(1), T(0)},     {c15, -s15},       {c30, T(-0.5)},    {h, -h},
                       {T(0.5), -c30},   {s15, -c15},       {T(0), T(-1)},     {-s15, -c15},
                       {T(-0.5), -c30},  {-h, -h},          {-c30, T(-0.5)},   {-c15, -s15},
                       {T(-1), T(0)},    {-c15, s15},       {-c30, T(0.5)}};
  cx<T> y[3][8];
#pragma unroll
  for (int n1 = 0; n1 < 3; ++n1)
#pragma unroll
    for (int n2 = 0; n2 < 8; ++n2) y[n1][n2] = v[n1 + 3 * n2];
  dft8(y[0]);
  dft8(y[1]);
  dft8(y[2]);
#pragma unroll
  for (int k2 = 1; k2 < 8; ++k2) {
    y[1][k2] = cmul(y[1][k2], w[k2]);
    y[2][k2] = cmul(y[2][k2], w[2 * k2]);
  }
#pragma unroll
  for (int k2 = 0; k2 < 8; ++k2) {
    cx<T> z[3] = {y[0][k2], y[1][k2], y[2][k2]};
    dft3(z);
    v[k2] = z[0];
    v[k2 + 8] = z[1];
    v[k2 + 16] = z[2];
  }
}

template <typename T, int R> __device__ __forceinline__ void dftR(cx<T> *v) {
  static_assert(R == 2 || R == 3 || R == 4 || R == 8 || R == 16 || R == 24, "radix");
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3(v);
  else if constexpr (R == 4) dft4(v);
  else if constexpr (R == 8) dft8(v);
  else if constexpr (R == 24) dft24(v);
  else dft16(v);
}

// Base twiddles of a thread: for every stage with LNS > 0 and every block b
// of the thread (j = t + b TPF, k = j mod NS), w = exp(-2 pi i k / (NS R)),
// loaded once per kernel from an FP64 table (tw64[n] = exp(-2 pi i n / M)).
// The stage multiplies input r by w^r, formed by FP64 recurrence (relative
// error <= (R - 1) eps_64, then rounded once to T), so the stages issue no
// memory loads for twiddles.
template <typename T, int M, int EF = 0> struct TwBase {
  using P = FftPlan<T, M, EF>;
  static constexpr int offset(int S) {
    int c = 0;
    for (int s = 0; s < S; ++s)
      if (P::S.lns[s] > 0) c += P::E / P::S.rad[s];
    return c;
  }
  static constexpr int NW = offset(P::S.nst) > 0 ? offset(P::S.nst) : 1;
  cx<double> w[NW];

  // tw64 holds exp(-2 pi i n / (M * ts)): the table of a transform ts times
  // longer (entry n ts of it is entry n of the length-M table).
  template <int S>
  __device__ __forceinline__ void load_stage(int t, const cx<double> *__restrict__ tw64, int ts) {
    if constexpr (S < P::S.nst) {
      constexpr int R = P::S.rad[S], LNS = P::S.lns[S];
      if constexpr (LNS > 0) {
        constexpr int NB = P::E / R, NS = 1 << LNS, STEP = (M / R) >> LNS;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int k = (t + b * P::TPF) & (NS - 1);
          w[offset(S) + b] = tw64[k * STEP * ts];
        }
      }
      load_stage<S + 1>(t, tw64, ts);
    }
  }
  __device__ __forceinline__ void load(int t, const cx<double> *__restrict__ tw64, int ts = 1) {
    load_stage<0>(t, tw64, ts);
  }
};

// Padded LDS index of x + c for a per-thread x and a compile-time c that is
// a multiple of 16: pad16(x + c) = pad16(x) + c + c / 16, so the constant
// part folds into the LDS instruction's immediate offset.
__device__ __forceinline__ int padx(int x) { return x + (x >> 4); }
constexpr int padc(int c) { return c + c / 16; }

// One in-place Stockham stage of radix R: for j < M/R, k = j mod NS, the
// inputs j + r M/R are twiddled by exp(-2 pi i r k / (NS R)), transformed,
// and written to (j - k) R + k + r NS.  Split into LDS loads, compute and
// store, so that the first stage can take its inputs from registers and the
// last stage can leave its outputs there.
template <typename T, int M, int S, int EF = 0> struct Stage {
  using P = FftPlan<T, M, EF>;
  using TB = TwBase<T, M, EF>;
  static constexpr int E = P::E, TPF = P::TPF, R = P::S.rad[S], LNS = P::S.lns[S];
  static constexpr int NB = E / R, NR = M / R, NS = 1 << LNS;
  static constexpr int STEP = NR >> LNS;  // M / (NS R)
  static constexpr bool LAST = NS * R == M;
  static constexpr bool TW = LNS > 0;
  // index shapes that allow per-thread base + constant offsets
  static constexpr bool IN_FOLD = TPF % 16 == 0 && NR % 16 == 0;
  static constexpr bool OUT_LAST_FOLD = LAST && TPF % 16 == 0 && NS % 16 == 0;
  static constexpr bool OUT_MID_FOLD = !LAST && NS % 16 == 0 && TPF % NS == 0;
  static constexpr bool OUT_FIRST_FOLD = NS == 1 && 16 % R == 0 && (TPF * R) % 16 == 0;


  // element index of input r of block b of thread t, and its padded index
  __device__ __forceinline__ static int in_index(int t, int b, int r) { return t + b * TPF + r * NR; }
  __device__ __forceinline__ static int in_pad(int t, int pt, int b, int r) {
    if constexpr (IN_FOLD) return pt + padc(b * TPF + r * NR);
    else return pad16(in_index(t, b, r));
  }
  // element index of output r of block b of thread t
  __device__ __forceinline__ static int out_index(int t, int b, int r) {
    const int j = t + b * TPF, k = j & (NS - 1);
    return (j - k) * R + k + r * NS;
  }
  // padded output index; ob is the per-thread base of out_base()
  __device__ __forceinline__ static int out_base(int t) {
    if constexpr (OUT_LAST_FOLD) return padx(t);
    else if constexpr (OUT_MID_FOLD) {
      const int k = t & (NS - 1);
      return padx((t - k) * R + k);
    } else if constexpr (OUT_FIRST_FOLD) return padx(t * R);
    else return 0;
  }
  __device__ __forceinline__ static int out_pad(int t, int ob, int b, int r) {
    if constexpr (OUT_LAST_FOLD) return ob + padc(b * TPF + r * NS);
    else if constexpr (OUT_MID_FOLD) return ob + padc(b * TPF * R + r * NS);
    else if constexpr (OUT_FIRST_FOLD) return ob + padc(b * TPF * R) + r;
    else return pad16(out_index(t, b, r));
  }

  // buf holds elements of type B (B = T, or FP32 storage of an FP64
  // computation: converted on load and rounded on store).
  template <typename B>
  __device__ __forceinline__ static void load(const cx<B> *buf, int t, cx<T> (&v)[NB][R]) {
    const int pt = padx(t);
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = cx_cast<T>(buf[in_pad(t, pt, b, r)]);
  }
  __device__ __forceinline__ static void compute(const TB &tb, cx<T> (&v)[NB][R]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if constexpr (TW) {
        cx<double> w1 = tb.w[TB::offset(S) + b];
        // opaque copy: keeps the compiler from keeping every power of w
        // alive between two transforms of one kernel (register spills)
        asm volatile("" : "+v"(w1.x), "+v"(w1.y));
        cx<double> wr = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[b][r] = cmul(v[b][r], cx<T>{T(wr.x), T(wr.y)});
          if (r + 1 < R) wr = cmul(wr, w1);
        }
      }
      dftR<T, R>(v[b]);
    }
  }
  template <typename B>
  __device__ __forceinline__ static void store(cx<B> *buf, int t, const cx<T> (&v)[NB][R]) {
    const int ob = out_base(t);
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int r = 0; r < R; ++r) buf[out_pad(t, ob, b, r)] = cx_cast<B>(v[b][r]);
  }
};

template <typename T, int M, int S, int EF, typename B>
__device__ __forceinline__ void fft_stage(cx<B> *buf, int t, const TwBase<T, M, EF> &tb) {
  using St = Stage<T, M, S, EF>;
  cx<T> v[St::NB][St::R];
  St::load(buf, t, v);
  group_sync<St::TPF>();
  St::compute(tb, v);
  St::store(buf, t, v);
  group_sync<St::TPF>();
}

// Stages S .. SEND - 1 (SEND = -1: all remaining stages).
template <typename T, int M, int S, int SEND = -1, int EF = 0, typename B>
__device__ __forceinline__ void fft_stages(cx<B> *buf, int t, const TwBase<T, M, EF> &tb) {
  using P = FftPlan<T, M, EF>;
  constexpr int END = SEND < 0 ? P::S.nst : SEND;
  if constexpr (S < END) {
    fft_stage<T, M, S, EF>(buf, t, tb);
    fft_stages<T, M, S + 1, SEND, EF>(buf, t, tb);
  }
}

// Exclusive prefix sum of (a, b) over the TPF threads of one FFT group.
// tmp: per-group scratch of 2 * (TPF / 64) elements when TPF > 64 (then the
// workgroup synchronises, so every thread must reach this call).
template <typename T, int TPF>
__device__ __forceinline__ void group_excl_scan2(T &a, T &b, int t, T *tmp) {
  constexpr int W = TPF < 64 ? TPF : 64;
  const int lane = t & (W - 1);
  T ia = a, ib = b;
#pragma unroll
  for (int d = 1; d < W; d <<= 1) {
    T oa = __shfl_up(ia, d, W);
    T ob = __shfl_up(ib, d, W);
    if (lane >= d) {
      ia += oa;
      ib += ob;
    }
  }
  T ea = __shfl_up(ia, 1, W);
  T eb = __shfl_up(ib, 1, W);
  if (lane == 0) {
    ea = T(0);
    eb = T(0);
  }
  if constexpr (TPF > 64) {
    const int w = t >> 6;
    if (lane == 63) {
      tmp[2 * w] = ia;
      tmp[2 * w + 1] = ib;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPF / 64 - 1; ++q)
      if (q < w) {
        ea += tmp[2 * q];
        eb += tmp[2 * q + 1];
      }
    __syncthreads();
  }
  a = ea;
  b = eb;
}

// ------------------------------------------------------------ DST engines
// Interface: a group's LDS region of Region elements of T holds the two rows
// a = region[0 .. RS) and b = region[RS .. 2 RS) (padded slots pad16(k),
// k < N) on output: slot k = X_k, slot 0 = 0.  The loader ld(k, kp, xa, xb)
// (kp = pad16(k), for loaders that read padded LDS rows),
// called for 1 <= k <= N - 1 (and k = N/2 by DstSint), may read LDS that
// aliases the region: all loads happen before the first store.  Only
// group-level synchronisation happens inside; callers synchronise the
// workgroup around cross-group LDS traffic.
//
// Outputs go to a sink st(k, kp, xa, xb), called once for every
// 0 <= k <= N - 1 with X_k of both rows in the compute type Tc (kp =
// pad16(k)); the default sink stores them to the region's rows.  A sink
// that writes elsewhere (global memory) leaves the region's contents
// undefined.

// Sink that stores the outputs to the padded rows a and b of a region.
template <typename T> struct RowSink {
  T *rowa, *rowb;
  template <typename Tc> __device__ __forceinline__ void operator()(int, int kp, Tc xa, Tc xb) const {
    rowa[kp] = T(xa);
    rowb[kp] = T(xb);
  }
};

// Row stride (elements of T) of a padded row of N slots.
template <int N> constexpr int row_stride() { return N + N / 16 + 1; }

// EF_ > 0 sets the elements per thread (to pair it with another engine of
// the same threads per group).
template <typename T_, int N_, int EF_ = 0> struct DstSint {
  using T = T_;
  using Tc = T_;
  static constexpr int N = N_;
  static constexpr int M = N;
  static constexpr int EF = EF_;
  using P = FftPlan<T, M, EF>;
  static constexpr int TPF = P::TPF;
  static constexpr int E = P::E;
  static constexpr int RS = row_stride<N>();
  static constexpr int Region = 2 * RS;
  static constexpr int ScanTmp = TPF > 64 ? 2 * (TPF / 64) : 0;
  // outputs are thread-blocked (thread t holds k = 2 t H .. 2 t H + 2 H - 1)
  static constexpr bool kLaneOut = false;

  static constexpr int H = E / 2;  // TPF * H = N / 2

  // Per-thread constants, loaded once per kernel: base twiddles and the
  // pre-twiddle weights was[k] = 2 sin(k pi / N) of the thread's k.
  struct Ctx {
    TwBase<T, M, EF> tw;
    T was[H];
    // tw64: exp(-2 pi i n / (M ts)); was_g: 2 sin(k pi / N), k < N / 2
    __device__ __forceinline__ void load(int t, const cx<double> *__restrict__ tw64,
                                         const T *__restrict__ was_g, int ts = 1) {
      tw.load(t, tw64, ts);
#pragma unroll
      for (int e = 0; e < H; ++e) was[e] = was_g[t + e * TPF];
    }
  };

  // Raw inputs of a thread: x_k and x_{N-k} of both rows for k = t + e TPF
  // (x_{N/2} in ka/kb for k = 0).
  struct In {
    T ka[H], kb[H], ca[H], cb[H];
  };

  template <typename Loader>
  __device__ __forceinline__ static void load_in(int t, Loader ld, In &in) {
#pragma unroll
    for (int e = 0; e < H; ++e) {
      const int k = t + e * TPF;
      if (k == 0) {
        ld(N / 2, pad16(N / 2), in.ka[e], in.kb[e]);
        in.ca[e] = in.cb[e] = T(0);
      } else {
        ld(k, pad16(k), in.ka[e], in.kb[e]);
        ld(N - k, pad16(N - k), in.ca[e], in.cb[e]);
      }
    }
  }

  template <typename Loader, typename Sink>
  __device__ __forceinline__ static void run(T *region, int t, T *scan_tmp, const Ctx &cx_,
                                             Loader ld, Sink st) {
    In in;
    load_in(t, ld, in);
    run_in(region, t, scan_tmp, cx_, in, st);
  }
  template <typename Loader>
  __device__ __forceinline__ static void run(T *region, int t, T *scan_tmp, const Ctx &cx_,
                                             Loader ld) {
    run(region, t, scan_tmp, cx_, ld, RowSink<T>{region, region + RS});
  }
  // Inputs already in the region's rows.
  template <typename Sink>
  __device__ __forceinline__ static void run_rows(T *region, int t, T *scan_tmp, const Ctx &cx_,
                                                  Sink st) {
    const T *ra = region, *rb = region + RS;
    run(region, t, scan_tmp, cx_, [&](int, int kp, T &xa, T &xb) {
      xa = ra[kp];
      xb = rb[kp];
    }, st);
  }
  __device__ __forceinline__ static void run_rows(T *region, int t, T *scan_tmp, const Ctx &cx_) {
    run_rows(region, t, scan_tmp, cx_, RowSink<T>{region, region + RS});
  }

  // The transform of loaded inputs; the group's region may still be read by
  // the loader until this call (it synchronises the group first).
  template <typename Sink>
  __device__ __forceinline__ static void run_in(T *region, int t, T *scan_tmp, const Ctx &cx_,
                                                const In &in, Sink st) {
    cx<T> *buf = reinterpret_cast<cx<T> *>(region);
    cx<T> ylo[H], yhi[H];
#pragma unroll
    for (int e = 0; e < H; ++e) {
      const int k = t + e * TPF;
      if (k == 0) {
        ylo[e] = {T(0), T(0)};
        yhi[e] = {T(4) * in.ka[e], T(4) * in.kb[e]};
      } else {
        const T xka = in.ka[e], xkb = in.kb[e], xca = in.ca[e], xcb = in.cb[e];
        const T w = cx_.was[e];
        const T t1a = xka - xca, t2a = w * (xka + xca);
        const T t1b = xkb - xcb, t2b = w * (xkb + xcb);
        ylo[e] = {t1a + t2a, t1b + t2b};
        yhi[e] = {t2a - t1a, t2b - t1b};
      }
    }
    group_sync<TPF>();
#pragma unroll
    for (int e = 0; e < H; ++e) {
      const int k = t + e * TPF;
      buf[pad16(k)] = ylo[e];
      buf[pad16(k == 0 ? N / 2 : N - k)] = yhi[e];
    }
    group_sync<TPF>();
    fft_stages<T, M, 0, -1, EF>(buf, t, cx_.tw);
    T ra[H], ia[H], rb[H], ib[H];
#pragma unroll
    for (int e = 0; e < H; ++e) {
      const int m = t * H + e;
      const cx<T> ck = buf[pad16(m)];
      const cx<T> cc = buf[pad16(m == 0 ? 0 : N - m)];
      ra[e] = T(0.5) * (ck.x + cc.x);
      ia[e] = T(0.5) * (ck.y - cc.y);
      rb[e] = T(0.5) * (ck.y + cc.y);
      ib[e] = T(0.5) * (cc.x - ck.x);
    }
    if (t == 0) {
      ra[0] *= T(0.5);
      rb[0] *= T(0.5);
    }
    T sa = T(0), sb = T(0);
#pragma unroll
    for (int e = 0; e < H; ++e) {
      sa += ra[e];
      sb += rb[e];
      ra[e] = sa;
      rb[e] = sb;
    }
    group_excl_scan2<T, TPF>(sa, sb, t, scan_tmp);
    group_sync<TPF>();
#pragma unroll
    for (int e = 0; e < H; ++e) {
      const int m = t * H + e;
      st(2 * m + 1, pad16(2 * m + 1), ra[e] + sa, rb[e] + sb);
      st(2 * m, pad16(2 * m), m == 0 ? T(0) : -ia[e], m == 0 ? T(0) : -ib[e]);
    }
  }
};

// T: the rows' and the LDS buffer's type; Tc: the arithmetic type.  With
// T = float and Tc = double every butterfly, twiddle and output is computed
// in FP64 and only the stage results are rounded to FP32 in LDS (three
// roundings per transform instead of one per operation).
//
// LOW_ > 0 splits off the inputs x_1 .. x_LOW: the FFT sees them as zero and
// their part 2 sum_{j <= LOW} x_j sin(pi j k / N) is added to every output
// in FP64 (sin by the Chebyshev recurrence from the table's sin and cos of
// pi k / N); the outputs are then FP64.  For smooth fields (spectra falling
// off with wavenumber, as in an inverse transform of the solution) the
// lowest modes carry most of the amplitude, and the FFT's FP32 rounding
// error, which scales with the norm of its input, shrinks with them.
template <typename T_, int N_, typename Tc_ = T_, int LOW_ = 0> struct DstExt {
  using T = T_;
  using Tc = Tc_;
  static constexpr int N = N_;
  static constexpr int M = 2 * N;
  static constexpr int LOW = LOW_;
  static_assert(LOW >= 0 && LOW < N / 2, "DstExt: LOW must be below N / 2");
  // type of the outputs handed to the sink
  using To = typename std::conditional<(LOW > 0), double, Tc>::type;
  // FP64 arithmetic at M >= 4096: 32 elements per thread (not 16), so that
  // one group has at most 128 threads and a column workgroup of 8 layers
  // fits in 512
  static constexpr int EF = (sizeof(Tc) > sizeof(T) && M >= 4096) ? 32 : 0;
  using P = FftPlan<Tc, M, EF>;
  static constexpr int TPF = P::TPF;
  static constexpr int E = P::E;
  static constexpr int RS = row_stride<N>();
  static constexpr int Region = 2 * (M + M / 16 + 1);
  static constexpr int ScanTmp = 0;
  // the sink is called with consecutive k on consecutive lanes
  static constexpr bool kLaneOut = true;

  static constexpr int H = E / 2;  // TPF * H = N

  struct Ctx {
    TwBase<Tc, M, EF> tw;
    const cx<double> *tab;  // exp(-2 pi i n / M): (cos, -sin) of pi n / N
    __device__ __forceinline__ void load(int t, const cx<double> *__restrict__ tw64,
                                         const Tc *__restrict__, int ts = 1) {
      tw.load(t, tw64, ts);
      tab = tw64;  // ts = 1 when LOW > 0 (checked by the caller's table length)
    }
  };

  // The first stage takes its inputs straight from the rows (w_e = x_e for
  // e < N, w_N = w_0 = 0, w_e = -x_{2N-e} for e > N; the loader is called
  // once per element it feeds) and the last stage's outputs go straight to
  // the rows (only the outputs k < N are needed), which saves two LDS round
  // trips per transform.
  static_assert(P::S.nst >= 2, "DstExt needs at least two FFT stages");
  static constexpr int LAST = P::S.nst - 1;
  using S0 = Stage<Tc, M, 0, EF>;
  using SL = Stage<Tc, M, LAST, EF>;

  // Raw inputs of a thread: the first stage's elements of the extension,
  // and 2 x_j of the split-off inputs j = 1..LOW of both rows.
  struct In {
    cx<Tc> v[S0::NB][S0::R];
    double la[LOW > 0 ? LOW : 1], lb[LOW > 0 ? LOW : 1];
  };

  template <typename Loader>
  __device__ __forceinline__ static void load_in(int t, Loader ld, In &in) {
    const int pt = padx(t);
#pragma unroll
    for (int b = 0; b < S0::NB; ++b)
#pragma unroll
      for (int r = 0; r < S0::R; ++r) {
        const int e = S0::in_index(t, b, r);
        const bool lo = e < N;
        const int k = lo ? e : M - e;
        const int kp = lo ? S0::in_pad(t, pt, b, r) : pad16(k);
        T xa = T(0), xb = T(0);
        if (k > LOW && k != N) ld(k, kp, xa, xb);
        in.v[b][r] = lo ? cx<Tc>{Tc(xa), Tc(xb)} : cx<Tc>{-Tc(xa), -Tc(xb)};
      }
#pragma unroll
    for (int j = 0; j < LOW; ++j) {
      T xa, xb;
      ld(j + 1, pad16(j + 1), xa, xb);
      in.la[j] = 2.0 * double(xa);
      in.lb[j] = 2.0 * double(xb);
    }
  }

  template <typename Loader, typename Sink>
  __device__ __forceinline__ static void run(T *region, int t, T *scan_tmp, const Ctx &cx_,
                                             Loader ld, Sink st) {
    In in;
    load_in(t, ld, in);
    run_in(region, t, scan_tmp, cx_, in, st);
  }
  template <typename Loader>
  __device__ __forceinline__ static void run(T *region, int t, T *scan_tmp, const Ctx &cx_,
                                             Loader ld) {
    run(region, t, scan_tmp, cx_, ld, RowSink<T>{region, region + RS});
  }
  // Inputs already in the region's rows.
  template <typename Sink>
  __device__ __forceinline__ static void run_rows(T *region, int t, T *scan_tmp, const Ctx &cx_,
                                                  Sink st) {
    const T *ra = region, *rb = region + RS;
    run(region, t, scan_tmp, cx_, [&](int, int kp, T &xa, T &xb) {
      xa = ra[kp];
      xb = rb[kp];
    }, st);
  }
  __device__ __forceinline__ static void run_rows(T *region, int t, T *scan_tmp, const Ctx &cx_) {
    run_rows(region, t, scan_tmp, cx_, RowSink<T>{region, region + RS});
  }

  // The transform of loaded inputs; the group's region may still be read by
  // the loader until this call (it synchronises the group first).
  template <typename Sink>
  __device__ __forceinline__ static void run_in(T *region, int t, T *, const Ctx &cx_, In in,
                                                Sink st) {
    cx<T> *buf = reinterpret_cast<cx<T> *>(region);
    group_sync<TPF>();
    S0::compute(cx_.tw, in.v);
    S0::store(buf, t, in.v);
    group_sync<TPF>();
    fft_stages<Tc, M, 1, LAST, EF>(buf, t, cx_.tw);
    cx<Tc> v[SL::NB][SL::R];
    SL::load(buf, t, v);
    group_sync<TPF>();
    SL::compute(cx_.tw, v);
    const int ob = SL::out_base(t);
#pragma unroll
    for (int b = 0; b < SL::NB; ++b)
#pragma unroll
      for (int r = 0; r < SL::R; ++r) {
        const int k = SL::out_index(t, b, r);
        if (k < N) {
          To ya = k == 0 ? To(0) : To(-v[b][r].y), yb = k == 0 ? To(0) : To(v[b][r].x);
          if constexpr (LOW > 0) {
            // s_j = sin(pi j k / N): s_{j+1} = 2 cos(pi k / N) s_j - s_{j-1}
            const cx<double> w = cx_.tab[k];
            const double c2 = 2.0 * w.x;
            double s0 = 0.0, s1 = -w.y;
#pragma unroll
            for (int j = 0; j < LOW; ++j) {
              ya = fma(in.la[j], s1, ya);
              yb = fma(in.lb[j], s1, yb);
              const double s2 = fma(c2, s1, -s0);
              s0 = s1;
              s1 = s2;
            }
          }
          st(k, SL::out_pad(t, ob, b, r), ya, yb);
        }
      }
  }
};

// ------------------------------------------------------ periodic transform
// Real periodic DFT of length M = 2N (the channel's j direction, period
// 2N = 2 (G - 1)) of two real rows a and b, packed as z = a + i b into one
// complex FFT of length M:
//   A_q = (Z_q + conj Z_{M-q}) / 2,   B_q = (Z_q - conj Z_{M-q}) / (2 i),
// q = 0..N (the half spectra; A_0, A_N, B_0, B_N are real).  On the
// physical side a row holds M reals (slot pad16(j)); on the spectral side
// it holds the half spectrum as 2 (N + 1) reals, re and im of q at slots
// pad16(2q) and pad16(2q + 1).  The inverse rebuilds Z'_q = A'_q + i B'_q
// (q <= N), conj(A'_{M-q}) + i conj(B'_{M-q}) (q > N), and returns the
// real and imaginary parts of its unnormalised inverse DFT, computed as the
// conjugate of the forward FFT of conj(Z').  Forward then inverse
// multiplies by M.
template <typename T_, int N_, typename Tc_ = T_> struct PerRfft {
  using T = T_;
  using Tc = Tc_;
  static constexpr int N = N_;
  static constexpr int M = 2 * N;
  // as DstExt: 32 elements per thread for FP64 arithmetic at M >= 4096
  static constexpr int EF = (sizeof(Tc) > sizeof(T) && M >= 4096) ? 32 : 0;
  using P = FftPlan<Tc, M, EF>;
  static constexpr int TPF = P::TPF;
  static constexpr int E = P::E;
  static constexpr int RS = row_stride<2 * N + 2>();
  static constexpr int FB = 2 * (M + M / 16 + 1);  // the complex FFT buffer, in T
  static constexpr int Region = 2 * RS > FB ? 2 * RS : FB;
  static constexpr int ScanTmp = 0;
  static constexpr int KQ = (N + 1 + TPF - 1) / TPF;  // half-spectrum points per thread
  static_assert(P::S.nst >= 2, "PerRfft needs at least two FFT stages");
  static constexpr int LAST = P::S.nst - 1;
  using S0 = Stage<Tc, M, 0, EF>;
  using SL = Stage<Tc, M, LAST, EF>;

  struct Ctx {
    TwBase<Tc, M, EF> tw;
    __device__ __forceinline__ void load(int t, const cx<double> *__restrict__ tw64, int ts = 1) {
      tw.load(t, tw64, ts);
    }
  };

  // Forward: ld(j, xa, xb) for 0 <= j < M (all loads happen before the
  // first store, so the loader may read the region); the half spectra of a
  // and b are left in the region's rows.
  template <typename Loader>
  __device__ __forceinline__ static void fwd(T *region, int t, const Ctx &c, Loader ld) {
    cx<T> *buf = reinterpret_cast<cx<T> *>(region);
    cx<Tc> v[S0::NB][S0::R];
#pragma unroll
    for (int b = 0; b < S0::NB; ++b)
#pragma unroll
      for (int r = 0; r < S0::R; ++r) {
        T xa, xb;
        ld(S0::in_index(t, b, r), xa, xb);
        v[b][r] = cx<Tc>{Tc(xa), Tc(xb)};
      }
    group_sync<TPF>();
    S0::compute(c.tw, v);
    S0::store(buf, t, v);
    group_sync<TPF>();
    fft_stages<Tc, M, 1, -1, EF>(buf, t, c.tw);  // Z_k at buf[pad16(k)]
    Tc h[KQ][4];
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
      const int q = t + s * TPF;
      if (q <= N) {
        const cx<Tc> z1 = cx_cast<Tc>(buf[pad16(q)]);
        const cx<Tc> z2 = cx_cast<Tc>(buf[pad16(q == 0 ? 0 : M - q)]);
        h[s][0] = Tc(0.5) * (z1.x + z2.x);  // A_q
        h[s][1] = Tc(0.5) * (z1.y - z2.y);
        h[s][2] = Tc(0.5) * (z1.y + z2.y);  // B_q
        h[s][3] = Tc(0.5) * (z2.x - z1.x);
      }
    }
    group_sync<TPF>();
    T *ra = region, *rb = region + RS;
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
      const int q = t + s * TPF;
      if (q <= N) {
        ra[pad16(2 * q)] = T(h[s][0]);
        ra[pad16(2 * q + 1)] = T(h[s][1]);
        rb[pad16(2 * q)] = T(h[s][2]);
        rb[pad16(2 * q + 1)] = T(h[s][3]);
      }
    }
  }

  // Inverse of the half spectra in the region's rows; st(j, xa, xb) is
  // called once for every 0 <= j < M after the region has been read in full
  // (a sink may write the rows).
  template <typename Sink>
  __device__ __forceinline__ static void inv(T *region, int t, const Ctx &c, Sink st) {
    cx<T> *buf = reinterpret_cast<cx<T> *>(region);
    const T *ra = region, *rb = region + RS;
    cx<Tc> v[S0::NB][S0::R];
#pragma unroll
    for (int b = 0; b < S0::NB; ++b)
#pragma unroll
      for (int r = 0; r < S0::R; ++r) {
        const int e = S0::in_index(t, b, r);
        const bool lo = e <= N;
        const int q = lo ? e : M - e;
        const Tc ar = Tc(ra[pad16(2 * q)]), ai = Tc(ra[pad16(2 * q + 1)]);
        const Tc br = Tc(rb[pad16(2 * q)]), bi = Tc(rb[pad16(2 * q + 1)]);
        // conj(Z'_e)
        v[b][r] = lo ? cx<Tc>{ar - bi, -(ai + br)} : cx<Tc>{ar + bi, ai - br};
      }
    group_sync<TPF>();
    S0::compute(c.tw, v);
    S0::store(buf, t, v);
    group_sync<TPF>();
    fft_stages<Tc, M, 1, LAST, EF>(buf, t, c.tw);
    cx<Tc> w[SL::NB][SL::R];
    SL::load(buf, t, w);
    group_sync<TPF>();
    SL::compute(c.tw, w);
#pragma unroll
    for (int b = 0; b < SL::NB; ++b)
#pragma unroll
      for (int r = 0; r < SL::R; ++r) st(SL::out_index(t, b, r), w[b][r].x, -w[b][r].y);
  }
};

}  // namespace pq
