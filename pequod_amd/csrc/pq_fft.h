// pq_fft.h -- LDS-resident FFT and DST-I building blocks for gfx950.
//
// One complex FFT of length N is computed by a group of TPF = N / E threads
// (E complex elements per thread, E in {4, 8, 12, 16}).  Stages are Stockham
// radix-R passes (R in {16, 8, 4, 2} and a final radix 3 for N = 3 * 2^a)
// run in place: every thread loads all of its inputs into registers, the
// workgroup synchronises, then results are written back.  LDS addresses are
// padded with one element per 16 (pad16) to avoid bank conflicts for both
// strided and thread-blocked accesses.
//
// dst_pair computes the unnormalised DST-I
//     X_k = 2 sum_{j=1..n} x_j sin(pi j k / N),  N = n + 1,
// of two real rows at once, following FFTPACK sint's formulation
// (qg3p_helmholtz.F90:1380-1426): a pre-twiddle into a length-N real
// sequence y, a real DFT of y (two rows packed as the real and imaginary
// parts of one complex FFT), and a prefix sum for the odd outputs.
#pragma once

#include "pq_common.h"

namespace pq {

struct DevPlan {
  int N;      // FFT length, N = n + 1
  int tpf;    // threads per FFT group = N / E
  int lgtpf;  // log2(tpf)
  int nst;    // number of FFT stages
  int rad[8]; // radix of each stage
  int lns[8]; // log2 of the Stockham span Ns before each stage
  int R;      // rows per tile (row kernels)
  int lgR;
  int C;      // columns per tile (column kernel)
  int RS;     // LDS row stride in elements of the compute type
  const void *tw;    // cx<T>[N]: exp(-2 pi i k / N)
  const void *was;   // T[N/2]:   2 sin(k pi / N)
  const void *lam4;  // T[N]:     (2N)^2 * (-4 sin^2(p pi / (2N)))
  const double *shat;  // double[N]: sum_{i=1..n} 2 sin(pi p i / N) (0 for even p)
};

// Synchronise the threads of one FFT group.  A group of at most 64 threads
// lives inside one wavefront and owns its LDS region, so a wavefront-scope
// fence + wave barrier suffices; larger groups need the workgroup barrier.
// The branch is uniform across the workgroup.
__device__ __forceinline__ void group_sync(int tpf) {
  if (tpf <= 64) {
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
  const cx<T> w[8] = {{T(1), T(0)}, {c1, -s1}, {h, -h}, {s1, -c1},
                      {T(0), T(-1)}, {-s1, -c1}, {-h, -h}, {-c1, -s1}};
#pragma unroll
  for (int k = 1; k < 8; ++k) o[k] = cmul(o[k], w[k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = e[k] + o[k];
    v[k + 8] = e[k] - o[k];
  }
}

template <typename T, int R> __device__ __forceinline__ void dftR(cx<T> *v) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3(v);
  else if constexpr (R == 4) dft4(v);
  else if constexpr (R == 8) dft8(v);
  else dft16(v);
}

// One in-place Stockham stage of radix R over the group's buffer.
template <typename T, int E, int R>
__device__ __forceinline__ void fft_stage(cx<T> *buf, int N, int lns, int t,
                                          int tpf,
                                          const cx<T> *__restrict__ tw) {
  constexpr int NB = E / R;
  const int NR = N / R;
  cx<T> v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * tpf;
#pragma unroll
    for (int r = 0; r < R; ++r) v[b][r] = buf[pad16(j + r * NR)];
  }
  group_sync(tpf);
  const int ns = 1 << lns;
  const int step = NR >> lns;  // N / (Ns R)
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * tpf;
    const int k = j & (ns - 1);
    if (lns > 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[r * k * step]);
    }
    dftR<T, R>(v[b]);
    const int base = ((j >> lns) << lns) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[pad16(base + r * ns)] = v[b][r];
  }
  group_sync(tpf);
}

template <typename T, int E>
__device__ __forceinline__ void fft_inplace(cx<T> *buf, const DevPlan &P,
                                            int t) {
  const cx<T> *tw = reinterpret_cast<const cx<T> *>(P.tw);
  for (int s = 0; s < P.nst; ++s) {
    const int R = P.rad[s];
    if constexpr (E % 16 == 0) {
      if (R == 16) { fft_stage<T, E, 16>(buf, P.N, P.lns[s], t, P.tpf, tw); continue; }
    }
    if constexpr (E % 8 == 0) {
      if (R == 8) { fft_stage<T, E, 8>(buf, P.N, P.lns[s], t, P.tpf, tw); continue; }
    }
    if constexpr (E % 4 == 0) {
      if (R == 4) { fft_stage<T, E, 4>(buf, P.N, P.lns[s], t, P.tpf, tw); continue; }
    }
    if constexpr (E % 3 == 0) {
      if (R == 3) { fft_stage<T, E, 3>(buf, P.N, P.lns[s], t, P.tpf, tw); continue; }
    }
    if constexpr (E % 2 == 0) {
      if (R == 2) { fft_stage<T, E, 2>(buf, P.N, P.lns[s], t, P.tpf, tw); continue; }
    }
  }
}

// Exclusive prefix sum of (a, b) over the tpf threads of one FFT group.
// tmp: per-group scratch of 2 * max(1, tpf / 64) elements.  Must be reached
// by every thread of the workgroup (it may synchronise).
template <typename T>
__device__ __forceinline__ void group_excl_scan2(T &a, T &b, int t, int tpf,
                                                 T *tmp) {
  const int width = tpf < 64 ? tpf : 64;
  const int lane = t & (width - 1);
  T ia = a, ib = b;
  for (int d = 1; d < width; d <<= 1) {
    T oa = __shfl_up(ia, d, width);
    T ob = __shfl_up(ib, d, width);
    if (lane >= d) {
      ia += oa;
      ib += ob;
    }
  }
  T ea = __shfl_up(ia, 1, width);
  T eb = __shfl_up(ib, 1, width);
  if (lane == 0) {
    ea = T(0);
    eb = T(0);
  }
  if (tpf > 64) {
    const int w = t >> 6;
    if (lane == 63) {
      tmp[2 * w] = ia;
      tmp[2 * w + 1] = ib;
    }
    __syncthreads();
    for (int q = 0; q < w; ++q) {
      ea += tmp[2 * q];
      eb += tmp[2 * q + 1];
    }
    __syncthreads();
  }
  a = ea;
  b = eb;
}

// DST-I of two real rows.  ld(k, xa, xb) returns element k (1 <= k <= n) of
// rows a and b; it may read LDS that aliases buf (all loads happen before the
// first store).  The results are written to rowa / rowb at padded slots
// 0..N-1 (slot k holds X_k; slot 0 is set to 0); rowa and rowb may alias buf.
// The caller synchronises before reading the rows.  Synchronisation inside
// is per FFT group: the loader may only read LDS written before the call
// and synchronised by the caller.
template <typename T, int E, typename Loader>
__device__ __forceinline__ void dst_pair(cx<T> *buf, T *rowa, T *rowb,
                                         const DevPlan &P, int t, T *scan_tmp,
                                         Loader ld) {
  constexpr int H = E / 2;
  const int N = P.N, tpf = P.tpf;
  const T *was = reinterpret_cast<const T *>(P.was);
  cx<T> ylo[H], yhi[H];
#pragma unroll
  for (int e = 0; e < H; ++e) {
    const int k = t + e * tpf;
    if (k == 0) {
      T xa, xb;
      ld(N / 2, xa, xb);
      ylo[e] = {T(0), T(0)};
      yhi[e] = {T(4) * xa, T(4) * xb};
    } else {
      T xka, xkb, xca, xcb;
      ld(k, xka, xkb);
      ld(N - k, xca, xcb);
      const T w = was[k];
      const T t1a = xka - xca, t2a = w * (xka + xca);
      const T t1b = xkb - xcb, t2b = w * (xkb + xcb);
      ylo[e] = {t1a + t2a, t1b + t2b};
      yhi[e] = {t2a - t1a, t2b - t1b};
    }
  }
  group_sync(tpf);
#pragma unroll
  for (int e = 0; e < H; ++e) {
    const int k = t + e * tpf;
    buf[pad16(k)] = ylo[e];
    buf[pad16(k == 0 ? N / 2 : N - k)] = yhi[e];
  }
  group_sync(tpf);
  fft_inplace<T, E>(buf, P, t);
  // Separate the two real DFTs and form the outputs; thread t owns
  // m in [t H, t H + H).
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
  group_excl_scan2(sa, sb, t, tpf, scan_tmp);
  group_sync(tpf);
#pragma unroll
  for (int e = 0; e < H; ++e) {
    const int m = t * H + e;
    rowa[pad16(2 * m + 1)] = ra[e] + sa;
    rowb[pad16(2 * m + 1)] = rb[e] + sb;
    rowa[pad16(2 * m)] = m == 0 ? T(0) : -ia[e];
    rowb[pad16(2 * m)] = m == 0 ? T(0) : -ib[e];
  }
}

}  // namespace pq
