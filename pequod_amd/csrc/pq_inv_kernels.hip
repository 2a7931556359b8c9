// pq_inv_kernels.hip -- the three passes of the modal Helmholtz inversion,
// compiled once per transform length N = n + 1 (-DPQ_N=<N>).
//
//   k_rows_fwd  per (member, layer, tile of R rows): read zeta rows, DST-I
//               along i, write the tile transposed as one contiguous block
//               S[l][jb][p][r] (r fastest) of the scratch.
//   k_cols_tri  per (member, CW adjacent columns p), all layers: read the
//               columns, project layers -> modes (z_to_modes), solve each
//               mode's tridiagonal system along j (FP64), accumulate the
//               FP64 partial of the domain mean psiC_m, project modes ->
//               layers (modes_to_z), write back in place.
//   k_psic      (host file) per member: psiC from the column pass's partials,
//               kappa_l and beta_lm.
//   k_cols_chan the periodic channel's column pass (periodic transform of
//               length 2N along j, boundary sums instead of psiC).
//   k_rows_inv  per (member, layer, tile): read the tile block, subtract the mass
//               correction in the i-spectral domain (closed form of
//               helm(1)), inverse DST-I along i, subtract the constant part,
//               write psi rows including the boundary.
//
// As the reference's helm (qg3p_helmholtz.F90:614-769): a sine transform
// along i (FFTPACK sint there), a tridiagonal solve along j per sine mode
// (the Thomas sweep there), the inverse transform along i.  The modal
// projection commutes with the i-transform, so it runs in the column pass,
// where all layers of a column are at hand.
//
// Mass correction (qg3p_aux.F90:196-230, qg3p_diagnostics.F90:615-635): with
// c_m = helm_m(1) (Dirichlet 0), corr_m = a_m + b_m c_m where
// a_m = 1 / (1 + ss_m mean(c_m)) and b_m = ss_m a_m, so
//   psi_l = sum_m M(l,m) phi'_m - kappa_l - sum_{m>=2} beta_lm c_m
// with kappa_l = psiC_1 + sum_{m>=2} M(l,m) psiC_m a_m and
// beta_lm = M(l,m) psiC_m b_m.  After the i-transform c_m is separable: for
// odd p it is shat_p / (2N) * g_mp(j) with
//   g_mp(j) = (1 - (r^j + r^(N-j)) / (1 + r^N)) / mu,  mu = lambda_p - ss_m,
// r + 1/r = 2 - mu (0 < r < 1), the exact solution of the tridiagonal
// problem g(j+1) + g(j-1) + (mu - 2) g(j) = 1, g(0) = g(N) = 0; for even p it
// is zero.  So no correction field is read from HBM.
#include "pq_fft.h"
#include "pq_inv.h"
#include "pq_mfma.h"

#ifndef PQ_N
#error "compile with -DPQ_N=<transform length>"
#endif

namespace pq {

// Forward and inverse transform engines of one precision.  Both run on the
// same FFT groups (equal threads per group, row layout and arithmetic
// type); a group's LDS region holds the larger of the two, and the twiddle
// table is the longer engine's.
template <typename F, typename I> struct EngPair {
  using Fwd = F;
  using Inv = I;
  using T = typename I::T;
  using Tc = typename I::Tc;
  static_assert(std::is_same<typename F::T, T>::value && std::is_same<typename F::Tc, Tc>::value,
                "engine pair: forward and inverse engines must share their types");
  static constexpr int N = I::N, TPF = I::TPF, RS = I::RS;
  static_assert(F::N == N && F::TPF == TPF && F::RS == RS,
                "engine pair: forward and inverse engines must share their groups");
  static constexpr int Region = F::Region > I::Region ? F::Region : I::Region;
  static constexpr int ScanTmp = F::ScanTmp > I::ScanTmp ? F::ScanTmp : I::ScanTmp;
  static constexpr int M = F::M > I::M ? F::M : I::M;
};

template <typename T> struct V16;
template <> struct V16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};
template <> struct V16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};

// Row r of a tile lives in group r/2's region (rows a, b of the group).
template <typename Eng>
__device__ __forceinline__ typename Eng::T *row_ptr(typename Eng::T *lds, int r) {
  return lds + (r >> 1) * Eng::Region + (r & 1) * Eng::RS;
}

// Sum of v[0..n) over the workgroup in a fixed order (shuffle tree inside
// each wavefront, then the wavefronts in order); every thread receives the
// sums.  red: >= 64 doubles of LDS.  All threads of the workgroup call it.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], int n, double *red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nact = min(64, (int)blockDim.x - 64 * w);
#pragma unroll
  for (int m = 0; m < NV; ++m)
    if (m < n) {
      double x = v[m];
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        const double y = __shfl_down(x, d, 64);
        if (lane + d < nact) x += y;
      }
      v[m] = x;
    }
  if (lane == 0)
    for (int m = 0; m < n; ++m) red[w * 8 + m] = v[m];
  __syncthreads();
  const int nw = ((int)blockDim.x + 63) >> 6;
#pragma unroll
  for (int m = 0; m < NV; ++m)
    if (m < n) {
      double x = 0.0;
      for (int q = 0; q < nw; ++q) x += red[q * 8 + m];
      v[m] = x;
    }
  __syncthreads();
}

// Workgroups are dispatched to the 8 XCDs round-robin (block i on XCD i % 8),
// each XCD with its own L2.  Remapping block i to
// (i % 8) * (n / 8) + i / 8 gives each XCD a contiguous range of logical
// blocks, so neighbouring blocks that share cache lines meet in one L2.
__device__ __forceinline__ int xcd_block(int i, int n) {
  return (n & 7) ? i : (i & 7) * (n >> 3) + (i >> 3);
}

// ------------------------------------------------------------------ K1
// Grid: sum over members of L_b * (N / R) workgroups of (R / 2) * TPF
// threads; LDS regions of the forward engine only.
template <typename Eng, typename Tin, typename Ts>
__global__ void __launch_bounds__(512) k_rows_fwd(InvArgs a, PtrTab ptrs) {
  using T = typename Eng::T;
  using Tc = typename Eng::Tc;
  using VT = typename V16<Ts>::type;
  constexpr int N = Eng::N, TPF = Eng::TPF, V = V16<Ts>::n;
  constexpr int NVT = 2 * N / (V * TPF);  // 16-byte vectors per thread of a tile
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int R = a.R, lgR = a.lgR, NB = (a.per ? 2 * N : N) >> lgR;
  const int b = member_of(ptrs.cum_r, ptrs.nb, blockIdx.x);
  const int rem = blockIdx.x - ptrs.cum_r[b];
  const int l = rem / NB, jb = rem - l * NB, j0 = jb << lgR;
  const int f = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int ngrp = R >> 1, nt = ngrp * TPF;
  typename Eng::Fwd::Ctx ctx;
  ctx.load(t, static_cast<const cx<double> *>(a.tw64), static_cast<const Tc *>(a.was),
           a.twm / Eng::Fwd::M);
  T *region = lds + f * Eng::Fwd::Region;
  T *scan_tmp = lds + ngrp * Eng::Fwd::Region + f * Eng::Fwd::ScanTmp;
  const Tin *z = static_cast<const Tin *>(ptrs.zeta[b]) + (size_t)l * ptrs.ldz1 * ptrs.ldz2;
  const int ja = j0 + 2 * f;
  const Tin *za = z + (size_t)ja * ptrs.ldz1;
  const Tin *zb = za + ptrs.ldz1;
  // node row 0 is the closed basin's boundary (not transformed); every row
  // of the channel is a periodic interior row
  const bool va = a.per || ja > 0;
  Eng::Fwd::run(region, t, scan_tmp, ctx, [&](int k, int, T &xa, T &xb) {
    xa = va ? T(za[k]) : T(0);
    xb = T(zb[k]);
  });
  __syncthreads();
  // Block (l, jb): S[p][r] = row r at p, r fastest; one contiguous N*R run.
  Ts *out = reinterpret_cast<Ts *>(a.scr + b * a.scr_stride) + ((size_t)(l * NB + jb) * N << lgR);
#pragma unroll
  for (int s = 0; s < NVT; ++s) {
    const int e0 = (threadIdx.x + s * nt) * V, r0 = e0 & (R - 1), p = e0 >> lgR;
    VT v;
#pragma unroll
    for (int q = 0; q < V; ++q) v[q] = Ts(row_ptr<typename Eng::Fwd>(lds, r0 + q)[pad16(p)]);
    *reinterpret_cast<VT *>(out + e0) = v;
  }
}

// Spectral step of the column pass on the points (p0 + c, q), c < C,
// 1 <= q < N, rows (l, c) at LDS row l C + c, computed in the engine's
// arithmetic type Tc.  pv_inversion (LT = L, 1..8):
// phi_m = sum_l D(m,l) x_l / ((2N)^2 (lambda_p + lambda_q - ss_m)), FP64
// psiC partial acc_m += phi_m shat_p shat_q, x_l = sum_m M(l,m) phi_m.
// helm (LT = 0): x_l /= (2N)^2 (lambda_p + lambda_q - ss_l).  The
// coefficients sit in LDS as rows of 8: cD[m * 8 + l] = D(m,l),
// cM[l * 8 + m] = M(l,m), cS[m] = (2N)^2 ss_m; up to 4 layers they are held
// in registers, above that each row is read per point as 16-byte vectors.
// PER (the channel): the rows hold half spectra along the period 2N, the
// real and imaginary parts of q = 0..N at slots e = 2q and 2q + 1, with
// the periodic eigenvalue -4 sin^2(pi q / 2N) = lambda_q; instead of psiC
// the q = 0 point accumulates the boundary sums of the modal solution,
// acc_m += phi_m w_p (wall i = 1) and acc_{8+m} += phi_m (-1)^(p+1) w_p
// (wall i = G).
// 1 / x by the hardware's reciprocal estimate and Newton-Raphson steps
// r <- r + r (1 - x r) (one for FP32, two for FP64; each squares the
// relative error), in place of an IEEE division per point and mode.
__device__ __forceinline__ float recip(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return fmaf(fmaf(-x, r, 1.0f), r, r);
}
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
}

template <typename Eng, int LT, bool PER = false>
__device__ __forceinline__ void spectral_step(const InvArgs &a, typename Eng::T *lds,
                                              const typename Eng::Tc *cD, const typename Eng::Tc *cM,
                                              const typename Eng::Tc *cS, int p0, int nt, int Lm,
                                              int lgC, double (&acc)[16]) {
  using T = typename Eng::T;
  using Tc = typename Eng::Tc;
  using VC = typename V16<Tc>::type;
  constexpr int VN = V16<Tc>::n;
  constexpr int N = Eng::N;
  constexpr bool REG = LT >= 1 && LT <= 4;
  constexpr int LR = REG ? LT : 1;
  constexpr int NV = (LT + VN - 1) / VN;  // 16-byte vectors per coefficient row
  const int C = 1 << lgC;
  const Tc *lam4 = static_cast<const Tc *>(a.lam4);
  Tc d[LR * LR], mm[LR * LR], sv[LR];
  if constexpr (REG) {
#pragma unroll
    for (int m = 0; m < LT; ++m)
#pragma unroll
      for (int l = 0; l < LT; ++l) {
        d[m * LT + l] = cD[m * 8 + l];
        mm[l * LT + m] = cM[l * 8 + m];
      }
#pragma unroll
    for (int i = 0; i < LT; ++i) sv[i] = cS[i];
  }
  for (int c = 0; c < C; ++c) {
    const int p = p0 + c;
    if (p < 1) continue;
    const Tc lp = lam4[p];
    const double sp = PER ? a.ctab[p] : a.shat[p];
    const double sgn = (p & 1) ? 1.0 : -1.0;
    constexpr int E0 = PER ? 0 : 1, E1 = PER ? 2 * N + 2 : N;
    for (int e = E0 + threadIdx.x; e < E1; e += nt) {
      const int q = PER ? e >> 1 : e;
      const Tc lpq = lp + lam4[q];
      if constexpr (LT == 0) {
        for (int ll = 0; ll < Lm; ++ll) {
          T *x = row_ptr<Eng>(lds, (ll << lgC) + c) + pad16(e);
          *x = T(Tc(*x) * recip(lpq - cS[ll]));
        }
      } else {
        if constexpr (!REG) {
          // keep the coefficient rows' reads inside the loop (hoisted into
          // registers they would spill)
          asm volatile("" ::: "memory");
        }
        // the psiC weight: zero unless p and q are odd (closed basin), so
        // the even-p columns skip the FP64 accumulation
        const bool wz = !PER && !(p & 1);
        const double wpq = PER ? (e == 0 ? sp : 0.0) : (wz ? 0.0 : sp * a.shat[q]);
        Tc col[LT], phi[LT];
#pragma unroll
        for (int ll = 0; ll < LT; ++ll) col[ll] = Tc(row_ptr<Eng>(lds, (ll << lgC) + c)[pad16(e)]);
#pragma unroll
        for (int m = 0; m < LT; ++m) {
          Tc s = Tc(0);
          if constexpr (REG) {
#pragma unroll
            for (int ll = 0; ll < LT; ++ll) s += d[m * LT + ll] * col[ll];
            phi[m] = s * recip(lpq - sv[m]);
          } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              const VC dv = *reinterpret_cast<const VC *>(cD + m * 8 + v * VN);
#pragma unroll
              for (int e = 0; e < VN; ++e)
                if (v * VN + e < LT) s += dv[e] * col[v * VN + e];
            }
            phi[m] = s * recip(lpq - cS[m]);
          }
          if (!wz) acc[m] += double(phi[m]) * wpq;
          if constexpr (PER) acc[8 + m] += double(phi[m]) * (wpq * sgn);
        }
#pragma unroll
        for (int ll = 0; ll < LT; ++ll) {
          Tc s = Tc(0);
          if constexpr (REG) {
#pragma unroll
            for (int m = 0; m < LT; ++m) s += mm[ll * LT + m] * phi[m];
          } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              const VC mv = *reinterpret_cast<const VC *>(cM + ll * 8 + v * VN);
#pragma unroll
              for (int e = 0; e < VN; ++e)
                if (v * VN + e < LT) s += mv[e] * phi[v * VN + e];
            }
          }
          row_ptr<Eng>(lds, (ll << lgC) + c)[pad16(e)] = T(s);
        }
      }
    }
  }
}


// ------------------------------------------------------------------ K2
// The closed basin's column pass as the reference's own method solves it:
// a tridiagonal system along j per column p and mode m (the Thomas sweep
// of h1d / ge_local, qg3p_helmholtz.F90:1028-1170), in a form that runs in
// parallel along j.  With x_l(p, j) the row pass's unnormalised sine
// transform, the mode's right-hand side is
//   z_m(j) = (1 / 2N) sum_l D(m,l) x_l(p, j)                (z(0) = 0),
// and y_m solves y(j-1) - (2 + mu) y(j) + y(j+1) = z(j), y(0) = y(N) = 0,
// mu = 4 sin^2(p pi / 2N) + ss_m.  With r + 1/r = 2 + mu, 0 < r < 1, the free-space
// Green's function r^|k| / (r - 1/r) gives the particular solution
//   yp(j) = (F(j) + B(j) - z(j)) / (r - 1/r),
//   F(j) = r F(j-1) + z(j),  B(j) = r B(j+1) + z(j),
// two first-order recurrences damped by r, and the walls add
// alpha r^j + beta r^(N-j) with yp(0) + alpha + beta r^N = 0 and
// yp(N) + alpha r^N + beta = 0.  Each thread holds E consecutive rows of
// one column in registers; the recurrences run locally, then as an affine
// scan across the column's threads (shuffles inside a wavefront, LDS
// across wavefronts).  FP64 arithmetic throughout; the scratch keeps the
// handle's element type.  The pass also forms the psiC partials
// sum_p shat_p sum_j y_m(p, j) and writes x_l = sum_m M(l,m) y_m back in
// place.  Grid: per member N / CW workgroups of CW columns x N / E threads.

// A thread's E rows of one layer in groups of V = RunVec<Ts, E>::V
// elements (16 or 8 bytes, the widest dividing the run), each group one
// vector access.  With V dividing both E and the block height R, a group
// never crosses a block of R rows and starts at a multiple of V elements
// (runs start at multiples of E, blocks and layers at multiples of R and
// N^2 elements).  at(j) is the element index of row j.
template <typename Ts, int E> struct RunVec {
  static constexpr int B = (E * (int)sizeof(Ts)) % 16 == 0 ? 16 : ((E * (int)sizeof(Ts)) % 8 == 0 ? 8 : (int)sizeof(Ts));
  static constexpr int V = B / (int)sizeof(Ts);
  typedef Ts type __attribute__((ext_vector_type(V)));
};
template <typename Ts, int E, typename At>
__device__ __forceinline__ void load_run(const Ts *base, At at, int j0, bool grouped, double (&x)[E]) {
  using RV = RunVec<Ts, E>;
  if (RV::V > 1 && grouped) {
#pragma unroll
    for (int k = 0; k < E / RV::V; ++k) {
      const typename RV::type v = *reinterpret_cast<const typename RV::type *>(base + at(j0 + k * RV::V));
#pragma unroll
      for (int q = 0; q < RV::V; ++q) x[k * RV::V + q] = double(v[q]);
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = double(base[at(j0 + e)]);
  }
}
template <typename Ts, int E, typename At>
__device__ __forceinline__ void store_run(Ts *base, At at, int j0, bool grouped, const double (&o)[E]) {
  using RV = RunVec<Ts, E>;
  if (RV::V > 1 && grouped) {
#pragma unroll
    for (int k = 0; k < E / RV::V; ++k) {
      typename RV::type v;
#pragma unroll
      for (int q = 0; q < RV::V; ++q) v[q] = Ts(o[k * RV::V + q]);
      *reinterpret_cast<typename RV::type *>(base + at(j0 + k * RV::V)) = v;
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) base[at(j0 + e)] = Ts(o[e]);
  }
}

// x^k1 and x^k2 for 0 <= k1, k2 < KMAX by squaring, the squarings shared (a
// fixed number of steps, so that lanes with different k stay converged).
template <int KMAX>
__device__ __forceinline__ void powu2(double x, int k1, int k2, double &r1, double &r2) {
  r1 = 1.0;
  r2 = 1.0;
#pragma unroll
  for (int bit = 1; bit < KMAX; bit <<= 1) {
    if (k1 & bit) r1 *= x;
    if (k2 & bit) r2 *= x;
    x *= x;
  }
}
// Lane moves of a double inside rows of 16 lanes (DPP; CTRL 0x110 + d:
// row_shr:d, lane q reads lane q - d; 0x100 + d: row_shl:d, lane q reads
// q + d).  Lanes whose source lies outside their row read 0.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int LANE>
__device__ __forceinline__ double readlane_d(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)b, LANE);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), LANE);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Affine scans of one column over a full wavefront (64 lanes, one column,
// so r is uniform): forward v(l) = rE v(l - 1) + f(l) and backward
// v(l) = rE v(l + 1) + b(l) over lanes l, with rE = r^E.  In-row inclusive
// scans by DPP steps of 1, 2, 4 and 8 lanes; the rows' totals read back as
// uniform values and combined in order (r16 = rE^16 per row).  Returns the
// exclusive carries cf (forward value of lane l - 1; 0 at lane 0) and cb
// (backward value of lane l + 1; 0 at lane 63), the wavefront's totals ft
// (forward at lane 63) and bt (backward at lane 0), and the powers
// pl = rE^l and pr = rE^(63 - l).
__device__ __forceinline__ void wave_scan64(double f, double bb, double rE, int lw, double &cf, double &cb,
                                            double &ft, double &bt, double &pl, double &pr) {
  const int q = lw & 15, k = lw >> 4;
  double vf = f, vb = bb, md = rE;
  vf = fma(md, dpp_row<0x111>(vf), vf);
  vb = fma(md, dpp_row<0x101>(vb), vb);
  md *= md;
  vf = fma(md, dpp_row<0x112>(vf), vf);
  vb = fma(md, dpp_row<0x102>(vb), vb);
  md *= md;
  vf = fma(md, dpp_row<0x114>(vf), vf);
  vb = fma(md, dpp_row<0x104>(vb), vb);
  md *= md;
  vf = fma(md, dpp_row<0x118>(vf), vf);
  vb = fma(md, dpp_row<0x108>(vb), vb);
  md *= md;  // r16 = rE^16
  // prefixes through the rows before row k (forward) and suffixes from the
  // rows after it (backward)
  const double T0 = readlane_d<15>(vf), T1 = readlane_d<31>(vf), T2 = readlane_d<47>(vf), T3 = readlane_d<63>(vf);
  const double U0 = readlane_d<0>(vb), U1 = readlane_d<16>(vb), U2 = readlane_d<32>(vb), U3 = readlane_d<48>(vb);
  const double P2 = fma(md, T0, T1), P3 = fma(md, P2, T2);
  const double S1 = fma(md, U3, U2), S0 = fma(md, S1, U1);
  ft = fma(md, P3, T3);
  bt = fma(md, S0, U0);
  const double Pk = k == 0 ? 0.0 : (k == 1 ? T0 : (k == 2 ? P2 : P3));
  const double Sk = k == 3 ? 0.0 : (k == 2 ? U3 : (k == 1 ? S1 : S0));
  double a1, a2, b1, b2;
  powu2<16>(rE, q, 15 - q, a1, a2);  // rE^q, rE^(15 - q)
  powu2<4>(md, k, 3 - k, b1, b2);    // r16^k, r16^(3 - k)
  // lane l - 1's forward value: its in-row value plus rE^q times the rows
  // before (q = 0: the shifted value is 0 and rE^0 = 1); likewise backward
  cf = fma(a1, Pk, dpp_row<0x111>(vf));
  cb = fma(a2, Sk, dpp_row<0x101>(vb));
  pl = a1 * b1;
  pr = a2 * b2;
}

template <int W>
__device__ __forceinline__ double seg_shfl_up(double v, int d) {
  return __shfl_up(v, d, W);
}
template <int W>
__device__ __forceinline__ double seg_shfl_down(double v, int d) {
  return __shfl_down(v, d, W);
}

// The column pass of one member of L <= LT layers (LT fixes the thread
// geometry).  Mode by mode, the column's wavefronts meet once (one barrier)
// for their scan totals, which also give the wall values; each mode has its
// own LDS slots, so no barrier guards their reuse.  The layer and mode loops
// stop at the member's runtime L (the branches also bound how far the
// scheduler interleaves the unrolled modes, which keeps the registers near
// one mode's working set).
template <typename Ts, int LT, int N>
__device__ __forceinline__ void cols_tri_body(const InvArgs &a, const PtrTab &ptrs, int b, int pblk, int lgC,
                                              double *red, double *s_tot, double *s_tc, int L) {
  constexpr int E = tri_e(N, LT), TPC = N / E;
  constexpr int W = TPC < 64 ? TPC : 64;  // lanes of a column inside one wavefront
  constexpr int NWC = TPC / W;            // wavefronts per column
  // s_tot[k][m][wavefront] (k = 0 forward, 1 backward): scan totals
  auto tot = [&](int k, int m, int q) -> double & { return s_tot[(k * 8 + m) * 8 + q]; };
  const int col = threadIdx.x / TPC, lc = threadIdx.x % TPC;
  const int p = (pblk << lgC) + col;
  const int j0 = lc * E;
  const int lw = lc & (W - 1), wc = lc / W;  // lane and wavefront inside the column
  const int wv = threadIdx.x >> 6;           // wavefront of the workgroup
  const int R = a.R, lgR = a.lgR;
  const CoefLayout CL{L, N};
  const double *coef = ptrs.coef[b];
  Ts *scr = reinterpret_cast<Ts *>(a.scr + b * a.scr_stride);
  const size_t cbase = (size_t)p * R;

  // the thread's rows of layer l: vector groups when the group size
  // divides the block height, else element by element
  const bool grouped = (R % RunVec<Ts, E>::V) == 0;
  auto elem = [&](int l, int j) -> size_t {
    return (size_t)l * N * N + (size_t)(j >> lgR) * N * R + cbase + (j & (R - 1));
  };

  // The modes' constants of the workgroup's columns into LDS, their loads
  // in flight with the field's (published by the barrier after the
  // projection): s_tc[(m C + c) 6 + k] = k-th constant of mode m, column c.
  constexpr int C = tri_cols(N, LT);
  for (int t = threadIdx.x; t < L * C * 6; t += C * TPC) {
    const int k = t % 6, mc = t / 6;
    s_tc[t] = coef[CL.tri() + ((size_t)(mc / C) * N + (pblk << lgC) + mc % C) * 8 + k];
  }

  // right-hand sides of the modes
  double z[LT][E];
#pragma unroll
  for (int m = 0; m < LT; ++m)
#pragma unroll
    for (int e = 0; e < E; ++e) z[m][e] = 0.0;
  const double s2n = 1.0 / (2.0 * N);
#pragma unroll
  for (int l = 0; l < LT; ++l) {
    if (l >= L) break;
    double x[E];
    load_run<Ts, E>(scr, [&](int j) { return elem(l, j); }, j0, grouped, x);
#pragma unroll
    for (int m = 0; m < LT; ++m) {
      if (m >= L) break;
      const double d = coef[CL.D() + m + L * l] * s2n;
#pragma unroll
      for (int e = 0; e < E; ++e) z[m][e] = fma(d, x[e], z[m][e]);
    }
  }
  __syncthreads();  // s_tc

  const double sp = a.shat[p];  // 0 for even p
  double acc[LT];
#pragma unroll
  for (int m = 0; m < LT; ++m) {
    if (m >= L) break;
    const double *tc = s_tc + (m * C + col) * 6;
    const double r = tc[0], ri = tc[1], cg = tc[2], rN = tc[3], w = tc[4], rE = tc[5];
    double cf, cb, ft = 0.0, bt = 0.0;
    double pl, pr;  // r^(E lw), r^(E (W - 1 - lw))
    {
      // Scans: the thread's recurrence totals (zero
      // carry-in), inclusive affine scans inside the column's lanes of this
      // wavefront (forward from lane 0 up, backward from the last lane
      // down), and the in-wavefront carries: F just before this thread's
      // rows, B just after.  A column inside one wavefront also takes its
      // totals F(N - 1) (last lane) and B(0) (lane 0) here; wider columns
      // publish each wavefront's totals.
      // The local recurrences Fl(e) (forward from 0 at the thread's first
      // row) and Bl(e) (backward from 0 after its last row) also give the
      // local part of the particular solution, (Fl + Bl - z) / (r - 1/r) =
      // (r Fl(e - 1) + Bl(e)) / (r - 1/r), kept in z[m] in place of the
      // right-hand side: Fl overwrites z first, and the backward sweep takes
      // z(e) back as Fl(e) - r Fl(e - 1) (an error of an ulp of Fl, of the
      // size of Bl's own rounding).
      double f = 0.0, bb = 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        f = fma(r, f, z[m][e]);
        z[m][e] = f;
      }
#pragma unroll
      for (int e = E - 1; e >= 0; --e) {
        const double g = e > 0 ? r * z[m][e - 1] : 0.0;  // r Fl(e - 1)
        bb = fma(r, bb, z[m][e] - g);
        z[m][e] = (g + bb) * cg;
      }
      if constexpr (W == 64) {
        // a wavefront holds one column: DPP row steps and uniform row totals
        double wf, wb;
        wave_scan64(f, bb, rE, lw, cf, cb, wf, wb, pl, pr);
        if constexpr (NWC > 1) {
          if (lw == W - 1) tot(0, m, wv) = wf;
          if (lw == 0) tot(1, m, wv) = wb;
          __syncthreads();  // the mode's wavefront totals (a slot per mode: no reuse)
        } else {
          ft = wf;
          bt = wb;
        }
      } else {
        double vf = f, vb = bb, mul = rE;
#pragma unroll
        for (int d = 1; d < W; d <<= 1) {
          const double uf = seg_shfl_up<W>(vf, d), ub = seg_shfl_down<W>(vb, d);
          if (lw >= d) vf = fma(mul, uf, vf);
          if (lw + d < W) vb = fma(mul, ub, vb);
          mul *= mul;
        }
        const double pf = seg_shfl_up<W>(vf, 1), pb = seg_shfl_down<W>(vb, 1);
        cf = lw == 0 ? 0.0 : pf;
        cb = lw == W - 1 ? 0.0 : pb;
        if constexpr (NWC > 1) {
          if (lw == W - 1) tot(0, m, wv) = vf;
          if (lw == 0) tot(1, m, wv) = vb;
          __syncthreads();  // the mode's wavefront totals (a slot per mode: no reuse)
        } else {
          ft = __shfl(vf, W - 1, W);
          bt = __shfl(vb, 0, W);
        }
        powu2<W>(rE, lw, W - 1 - lw, pl, pr);
      }
    }
    {
      // Then the carries across the column's wavefronts
      // and the column's totals (wavefront totals combined in order), and
      // the solution y = yp + alpha r^j + beta r^(N - j), where the
      // particular solution yp(j) = (F(j) + B(j) - z(j)) / (r - 1/r) has
      // F(j0 + e) = Fl(e) + r^(e+1) F(j0 - 1) and B(j0 + e) = Bl(e) +
      // r^(E-e) B(j0 + E), and the walls satisfy yp(0) + alpha + beta r^N = 0
      // and yp(N) + alpha r^N + beta = 0, with yp(0) = B(0) / (r - 1/r)
      // (F(0) = z(0)) and yp(N) = r F(N - 1) / (r - 1/r) (z(N) = 0).  So
      // y(j0 + e) = local(e) + A r^e + B r^(E - e) with A = (r F(j0 - 1) /
      // (r - 1/r) + alpha r^j0) and B = (B(j0 + E) / (r - 1/r) +
      // beta r^(N - j0 - E)); r^j0 = (r^E)^lc and r^(N - j0 - E) =
      // (r^E)^(TPC - 1 - lc) from the carries' powers r^(E lw),
      // r^(E (W - 1 - lw)) and those of r^(E W).
      double cfm = cf, cbm = cb, fN = ft, b0 = bt;
      double ru = pl, rd = pr;
      if constexpr (NWC > 1) {
        double rW = rE;  // r^(E W)
#pragma unroll
        for (int d = 1; d < W; d <<= 1) rW *= rW;
        const int w0 = wv - wc;  // the column's first wavefront
        double cwf = 0.0, cwb = 0.0;
        fN = 0.0;
        b0 = 0.0;
#pragma unroll
        for (int q = 0; q < NWC; ++q) {
          if (q == wc) cwf = fN;
          fN = fma(rW, fN, tot(0, m, w0 + q));
        }
#pragma unroll
        for (int q = NWC - 1; q >= 0; --q) {
          if (q == wc) cwb = b0;
          b0 = fma(rW, b0, tot(1, m, w0 + q));
        }
        // the wavefront's carry reaches lane lw through r^(E lw) (forward)
        // and r^(E (W - 1 - lw)) (backward)
        cfm = fma(pl, cwf, cfm);
        cbm = fma(pr, cwb, cbm);
        double qa, qb;  // r^(E W wc), r^(E W (NWC - 1 - wc))
        powu2<NWC>(rW, wc, NWC - 1 - wc, qa, qb);
        ru *= qa;
        rd *= qb;
      }
      const double y0 = b0 * cg, yN = r * fN * cg;
      const double al = (rN * yN - y0) * w, be = (rN * y0 - yN) * w;
      const double A = fma(cg * r, cfm, al * ru), B = fma(cg, cbm, be * rd);
      double sum = 0.0, pu = 1.0, pd = rE;  // r^e, r^(E - e)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        double y = fma(A, pu, fma(B, pd, z[m][e]));
        if (e == 0 && j0 == 0) y = 0.0;  // Dirichlet row j = 0
        z[m][e] = y;
        sum += y;
        pu *= r;
        pd *= ri;
      }
      acc[m] = sp * sum;
    }
  }

  // psiC partials of this workgroup's columns (fixed-order FP64 sum)
  if (a.mode == 0) {
    block_sum(acc, L, red);
    if (threadIdx.x == 0) {
      double *part = a.part + (size_t)b * a.part_stride + (size_t)pblk * L;
      for (int m = 0; m < L; ++m) part[m] = acc[m];
    }
  }
  // back to layers, in place
#pragma unroll
  for (int l = 0; l < LT; ++l) {
    if (l >= L) break;
    double o[E];
#pragma unroll
    for (int e = 0; e < E; ++e) o[e] = 0.0;
#pragma unroll
    for (int m = 0; m < LT; ++m) {
      if (m >= L) break;
      const double mm = coef[CL.M() + l + L * m];
#pragma unroll
      for (int e = 0; e < E; ++e) o[e] = fma(mm, z[m][e], o[e]);
    }
    store_run<Ts, E>(scr, [&](int j) { return elem(l, j); }, j0, grouped, o);
  }
}

// N is a template argument although every kernel unit has one length: the
// units are linked into one library, and a kernel's symbol must differ per
// length (otherwise the units' launches all resolve to one unit's kernel).
// LT fixes the thread geometry (tri_e); members of a ragged batch may have
// fewer layers than LT.
template <typename Ts, int LT, int N>
__global__ void __launch_bounds__(512) k_cols_tri(InvArgs a, PtrTab ptrs) {
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int b = member_of(ptrs.cum_c, ptrs.nb, bid);
  const int L = ptrs.lay[b], lgC = ptrs.lgc[b];
  const int pblk = bid - ptrs.cum_c[b];
  __shared__ double red[64];
  __shared__ double s_tot[2 * 8 * 8];
  __shared__ double s_tc[LT * tri_cols(N, LT) * 6];
  cols_tri_body<Ts, LT, N>(a, ptrs, b, pblk, lgC, red, s_tot, s_tc, L);
}

// ------------------------------------------------------------ K2 channel
// The channel's column pass: as k_cols, with the periodic transform of
// length 2N along j (CE, a PerRfft engine) in place of the sine transform,
// half spectra in the rows, and the FP64 boundary sums of the modal
// solution (two per mode, the walls i = 1 and i = G) in place of the psiC
// partials.  Scratch element (p, j) of layer l at
// l N 2N + (j / R) N R + p R + j % R.
template <typename CE, typename Ts>
__global__ void __launch_bounds__(512) k_cols_chan(InvArgs a, PtrTab ptrs) {
  using T = typename CE::T;
  using Tc = typename CE::Tc;
  using VT = typename V16<Ts>::type;
  constexpr int N = CE::N, M = CE::M, TPF = CE::TPF, V = V16<Ts>::n;
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int R = a.R, lgR = a.lgR;
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  const int b = member_of(ptrs.cum_c, ptrs.nb, bid);
  const int L = ptrs.lay[b], lgC = ptrs.lgc[b], C = 1 << lgC;
  const int pblk = bid - ptrs.cum_c[b];
  const int p0 = pblk << lgC;
  const int ngrp = blockDim.x / TPF, nt = blockDim.x;
  const int f = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int nrows = L << lgC;
  typename CE::Ctx ctx;
  ctx.load(t, static_cast<const cx<double> *>(a.tw64), a.twm / M);
  T *region = lds + f * CE::Region;
  const size_t red_off = (sizeof(T) * (size_t)ngrp * CE::Region + 15) & ~size_t(15);
  double *red = reinterpret_cast<double *>(smem + red_off);
  Tc *cD = reinterpret_cast<Tc *>(red + 64);
  Tc *cM = cD + 64;
  Tc *cS = cM + 64;
  const CoefLayout CL{L, N};
  const double *coef = ptrs.coef[b];
  double cd0 = 0.0, cm0 = 0.0, cs0 = 0.0;
  {
    const int i = threadIdx.x, r = i >> 3, c = i & 7;
    if (i < 64 && r < L && c < L) {
      cd0 = coef[CL.D() + r + L * c];
      cm0 = coef[CL.M() + r + L * c];
    }
  }
  if ((int)threadIdx.x < L) cs0 = coef[CL.ssN() + threadIdx.x];

  Ts *scr = reinterpret_cast<Ts *>(a.scr + b * a.scr_stride);
  const int ra = 2 * f, rb = ra + 1;
  const int la = ra >> lgC, pa = p0 + (ra & (C - 1));
  const int lb = rb >> lgC, pb = p0 + (rb & (C - 1));
  const bool va = ra < nrows && pa >= 1;
  const bool vb = rb < nrows && pb >= 1;
  const Ts *sa = scr + (size_t)la * N * M + (size_t)pa * R;
  const Ts *sb = scr + (size_t)lb * N * M + (size_t)pb * R;
  const int blk = N << lgR;
  CE::fwd(region, t, ctx, [&](int j, T &xa, T &xb) {
    const int off = (j >> lgR) * blk + (j & (R - 1));
    xa = va ? T(sa[off]) : T(0);
    xb = vb ? T(sb[off]) : T(0);
  });
  if ((int)threadIdx.x < 64) {
    cD[threadIdx.x] = Tc(cd0);
    cM[threadIdx.x] = Tc(cm0);
  }
  if ((int)threadIdx.x < L) cS[threadIdx.x] = Tc(cs0);
  for (int i = threadIdx.x + nt; i < 64; i += nt) {
    const int r = i >> 3, c = i & 7;
    const bool in = r < L && c < L;
    cD[i] = in ? Tc(coef[CL.D() + r + L * c]) : Tc(0);
    cM[i] = in ? Tc(coef[CL.M() + r + L * c]) : Tc(0);
  }
  for (int i = threadIdx.x + nt; i < L; i += nt) cS[i] = Tc(coef[CL.ssN() + i]);
  __syncthreads();

  double acc[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) acc[m] = 0.0;
  switch (a.mode == 0 ? L : 0) {
    case 1: spectral_step<CE, 1, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    case 2: spectral_step<CE, 2, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    case 3: spectral_step<CE, 3, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    case 4: spectral_step<CE, 4, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    case 5: spectral_step<CE, 5, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    case 6: spectral_step<CE, 6, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    case 7: spectral_step<CE, 7, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    case 8: spectral_step<CE, 8, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
    default: spectral_step<CE, 0, true>(a, lds, cD, cM, cS, p0, nt, L, lgC, acc); break;
  }
  if (a.mode == 0) {
    double top[8], bot[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      top[m] = acc[m];
      bot[m] = acc[8 + m];
    }
    block_sum(top, L, red);
    block_sum(bot, L, red);
    if (threadIdx.x == 0) {
      double *part = a.part + (size_t)b * a.part_stride + (size_t)pblk * 2 * L;
      for (int m = 0; m < L; ++m) {
        part[m] = top[m];
        part[L + m] = bot[m];
      }
    }
  } else {
    __syncthreads();
  }

  {
    T *rowa = region, *rowb = region + CE::RS;
    CE::inv(region, t, ctx, [&](int j, Tc xa, Tc xb) {
      rowa[pad16(j)] = T(xa);
      rowb[pad16(j)] = T(xb);
    });
  }
  __syncthreads();
  const int per_layer = (M << lgC) / V;
  for (int ll = 0; ll < L; ++ll) {
    for (int iv = threadIdx.x; iv < per_layer; iv += nt) {
      const int e = iv * V;
      const int r0 = e & (R - 1), c = (e >> lgR) & (C - 1), jb = e >> (lgR + lgC);
      const T *src = row_ptr<CE>(lds, (ll << lgC) + c);
      VT v;
#pragma unroll
      for (int q = 0; q < V; ++q) v[q] = Ts(src[pad16((jb << lgR) + r0 + q)]);
      Ts *dst = scr + (size_t)ll * N * M + ((size_t)jb * N + p0 + c) * R + r0;
      *reinterpret_cast<VT *>(dst) = v;
    }
  }
}

// The mass correction of the closed basin for odd p (even p have none),
// subtracted from the tile's rows in the i-spectral domain:
// x(r, p) -= sum_{m>=2} beta_lm c_m(p, j0 + r), with c_m's closed form
// shat_p / 2N g_mp(j) from the coefficient table (rho = ln r, gamma
// (scaled by beta_lm here), eta, r per mode and odd p), by recurrence over
// the tile's R rows in FP64.  sub(r, p, v) subtracts v from element
// (r, p).  All threads of the workgroup call it (tid, nt).  The terms in
// r^j are dropped below e^-cut of the constant part (cut = 40 where fields
// and transforms are FP64, 20 where either is FP32: e^-20 = 2e-9, far under
// FP32 rounding).  Where that holds for every mode (jmin rmin(h) > cut, the
// slowest decay rate rmin), the correction is the constant gam_l(h) that
// the psiC reduction formed (the sum of the modes' constant parts): no mode
// loop.  The first two columns' rmin and gam_l come preloaded (pre, loaded
// by the caller with the tile, so that their latency overlaps the tile's)
// unless k0 = 2.
struct MassPre {
  double r0, r1, g0, g1;
};
__device__ __forceinline__ MassPre mass_preload(const double *rmin, const double *gam_l, int NN, int tid,
                                                int nt) {
  const int h0 = min(tid, NN / 2 - 1), h1 = min(tid + nt, NN / 2 - 1);  // clamped: surplus unused
  return MassPre{rmin[h0], rmin[h1], gam_l[h0], gam_l[h1]};
}
template <int NN, typename Sub>
__device__ __forceinline__ void mass_correction(const double *tab, const double (&beta)[8], int L,
                                                int jb, int NB, int R, int lgR, int tid, int nt,
                                                double cut, const double *rmin, const double *gam_l,
                                                MassPre pre, int k0, Sub sub) {
  const int j0 = jb << lgR;
  const double jmin0 = double(min(j0, NN - j0 - R + 1));
  int k = k0;
  for (int h = tid; h < NN / 2; h += nt, ++k) {
    const int p = 2 * h + 1;
    const double rm = k == 0 ? pre.r0 : (k == 1 ? pre.r1 : rmin[h]);
    if (jmin0 * rm > cut) {
      const double gl = k == 0 ? pre.g0 : (k == 1 ? pre.g1 : gam_l[h]);
#pragma unroll
      for (int r = 0; r < 16; ++r)
        if (r < R) sub(r, p, gl);
      continue;
    }
    double acc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0;
    // each mode's constants loaded one mode ahead (the index clamped to
    // the table)
    double2 n01 = reinterpret_cast<const double2 *>(tab + (size_t)h * 4)[0];
    double2 n23 = reinterpret_cast<const double2 *>(tab + (size_t)h * 4)[1];
    for (int m = 1; m < L; ++m) {
      const double2 c01 = n01, c23 = n23;
      {
        const double2 *e =
            reinterpret_cast<const double2 *>(tab + ((size_t)(min(m + 1, L - 1) - 1) * (NN / 2) + h) * 4);
        n01 = e[0];
        n23 = e[1];
      }
      const double rho = c01.x, gam = beta[m] * c01.y, eta = c23.x, rr = c23.y;
      // r^j for the tile's rows is below e^-cut far from both edges: then
      // only the constant part remains
      if (jmin0 * -rho > cut) {
#pragma unroll
        for (int r = 0; r < 16; ++r)
          if (r < R) acc[r] += gam;
        continue;
      }
      // r^j0 = (r^R)^jb and r^(N - j0 - R + 1) = r (r^R)^(NB - 1 - jb), by
      // squaring (jb is the workgroup's, so the loops do not diverge)
      double rR = rr;
      for (int k = 0; k < lgR; ++k) rR *= rR;
      double u = 1.0, v = rr;
      {
        double x = rR;
        for (int k = jb; k; k >>= 1) {
          if (k & 1) u *= x;
          x *= x;
        }
        x = rR;
        for (int k = NB - 1 - jb; k; k >>= 1) {
          if (k & 1) v *= x;
          x *= x;
        }
      }
      double hh[16];
#pragma unroll
      for (int r = 0; r < 16; ++r)  // r^(j0 + r), going up
        if (r < R) {
          hh[r] = u;
          u *= rr;
        }
      // r^(N - j0 - r), going down
#pragma unroll
      for (int r = 15; r >= 0; --r)
        if (r < R) {
          hh[r] += v;
          v *= rr;
        }
#pragma unroll
      for (int r = 0; r < 16; ++r)
        if (r < R) acc[r] += gam * (1.0 - hh[r] * eta);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      if (r < R) sub(r, p, acc[r]);
  }
}

// ------------------------------------------------------------------ K3
// Grid: sum over members of L_b * (N / R) workgroups of (R / 2) * TPF threads.
// At least 3 waves per SIMD (at most 168 VGPRs) up to N = 2048: without
// the hint the correction loop's FP64 temporaries take the FP32 kernel to
// 182 VGPRs and 2 waves per SIMD (-6 % time at 1023^2 x 3 x 8).  Longer
// transforms would spill under it.
constexpr int kRowsInvWaves = PQ_N <= 2048 ? 3 : 1;
template <typename Eng, typename Tout, typename Ts>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(kRowsInvWaves)))
k_rows_inv(InvArgs a, PtrTab ptrs) {
  using T = typename Eng::T;
  using Tc = typename Eng::Tc;
  using VT = typename V16<Ts>::type;
  constexpr int N = Eng::N, TPF = Eng::TPF, V = V16<Ts>::n;
  constexpr int NVT = 2 * N / (V * TPF);  // 16-byte vectors per thread of a tile
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int R = a.R, lgR = a.lgR, NB = (a.per ? 2 * N : N) >> lgR;
  const int b = member_of(ptrs.cum_r, ptrs.nb, blockIdx.x);
  const int L = ptrs.lay[b];
  const int rem = blockIdx.x - ptrs.cum_r[b];
  const int l = rem / NB, jb = rem - l * NB, j0 = jb << lgR;
  const int f = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int ngrp = R >> 1, nt = ngrp * TPF;
  T *region = lds + f * Eng::Region;
  T *scan_tmp = lds + ngrp * Eng::Region + f * Eng::ScanTmp;
  const CoefLayout CL{L, N};

  // Issue the tile block and psiC partial loads first.
  const Ts *in = reinterpret_cast<const Ts *>(a.scr + b * a.scr_stride) + ((size_t)(l * NB + jb) * N << lgR);
  VT tile[NVT];
#pragma unroll
  for (int s = 0; s < NVT; ++s) tile[s] = *reinterpret_cast<const VT *>(in + (threadIdx.x + s * nt) * V);
  // kappa_l and beta_lm of this member and layer (written by k_psic)
  double kappa = 0.0;
  double beta[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) beta[m] = 0.0;
  // channel: the correction's wall values E_l(0), E_l(N) (written by
  // k_chan_corr)
  double e0 = 0.0, eN = 0.0;
  if (a.mode == 0 && a.per) {
    const double *E = a.kb + (size_t)b * a.kb_stride + (size_t)l * (N + 1);
    e0 = E[0];
    eN = E[N];
  } else if (a.mode == 0) {
    const double *kb = a.kb + (size_t)b * a.kb_stride;
    kappa = kb[l];
#pragma unroll
    for (int m = 1; m < 8; ++m)
      if (m < L) beta[m] = kb[L + l * L + m];
  }
  // the mass correction's per-column constants (rmin, then Gamma_l);
  // preloaded except with the FP64 engine (registers)
  constexpr int kPre = sizeof(T) == 8 ? 2 : 0;
  const double *kbm = a.kb + (size_t)b * a.kb_stride + L + L * L;
  MassPre mpre{};
  if (kPre == 0 && a.mode == 0 && !a.per && L > 1)
    mpre = mass_preload(kbm, kbm + (1 + l) * (N / 2), N, threadIdx.x, nt);
  typename Eng::Inv::Ctx ctx;
  ctx.load(t, static_cast<const cx<double> *>(a.tw64), static_cast<const Tc *>(a.was),
           a.twm / Eng::Inv::M);

  // Tile block (l, jb): row r at p <- S[p * R + r].
#pragma unroll
  for (int s = 0; s < NVT; ++s) {
    const int e0 = (threadIdx.x + s * nt) * V, r0 = e0 & (R - 1), p = e0 >> lgR;
#pragma unroll
    for (int q = 0; q < V; ++q) row_ptr<Eng>(lds, r0 + q)[pad16(p)] = T(tile[s][q]);
  }

  __syncthreads();  // publishes the tile

  // Channel: the correction E_l, uniform along j, subtracted in the
  // i-spectral domain from every row of the tile.
  if (a.mode == 0 && a.per && L > 1) {
    const double *E = a.kb + (size_t)b * a.kb_stride + (size_t)l * (N + 1);
    for (int p = 1 + threadIdx.x; p < N; p += nt) {
      const double ep = E[p];
      for (int r = 0; r < R; ++r) {
        T *x = row_ptr<Eng>(lds, r) + pad16(p);
        *x = T(double(*x) - ep);
      }
    }
    __syncthreads();
  }
  // Correction field for odd p, by recurrence over the tile's rows (FP64).
  if (a.mode == 0 && !a.per && L > 1) {
    mass_correction<N>(ptrs.coef[b] + CL.tab(), beta, L, jb, NB, R, lgR, threadIdx.x, nt,
                       sizeof(T) == 4 || sizeof(Tout) == 4 ? 20.0 : 40.0, kbm, kbm + (1 + l) * (N / 2), mpre,
                       kPre, [&](int r, int p, double v) {
                         T *x = row_ptr<Eng>(lds, r) + pad16(p);
                         *x = T(double(*x) - v);
                       });
    __syncthreads();
  }

  // Output rows ja = j0 + 2f and ja + 1 of this group (ja + 1 <= N - 1).
  // pv_inversion (mode 0): psi = X - kappa, and -kappa on the boundary
  // (X_0 = 0 gives i = 0; row j = 0 is written as X = 0).  helm (mode 1),
  // the reference's boundary semantics (qg3p_helmholtz.F90:757-761,
  // 1055-1057): row 0 gets 0 at 1 <= i <= N - 1, rows 1..N-1 get X at
  // 1 <= i <= N - 1 and 0 at i = N; column i = 0 and row N are not written.
  Tout *ps = static_cast<Tout *>(ptrs.psi[b]) + (size_t)l * ptrs.ldp1 * ptrs.ldp2;
  const long long ld1 = ptrs.ldp1;
  constexpr int G = N + 1;
  const int ja = j0 + 2 * f;
  // The channel (per): every row is interior; pv_inversion writes the
  // walls i = 0 and i = N as -E_l(0) and -E_l(N) (X_0 = 0), helm keeps
  // column 0 and zeroes i = N.
  Tout *pa = ps + (size_t)ja * ld1, *pb = pa + ld1;
  const bool za = !a.per && ja == 0;
  const bool m0 = a.mode == 0;
  const double sub = m0 ? (a.per ? 0.0 : kappa) : 0.0;
  auto sink = [&](int k, int, auto xa, auto xb) {
    if (m0 || k != 0) {
      const double s0 = (a.per && k == 0) ? e0 : sub;
      pa[k] = Tout((za ? 0.0 : double(xa)) - s0);
      pb[k] = Tout(double(xb) - s0);
    }
  };
  {
    const T *rowa = region, *rowb = region + Eng::RS;
    if constexpr (Eng::Inv::kLaneOut) {
      // outputs leave the last FFT stage with consecutive k across lanes
      Eng::Inv::run_rows(region, t, scan_tmp, ctx, sink);
    } else {
      Eng::Inv::run_rows(region, t, scan_tmp, ctx);
      group_sync<TPF>();
      for (int k = t; k < N; k += TPF) sink(k, 0, Tc(rowa[pad16(k)]), Tc(rowb[pad16(k)]));
    }
  }
  if (t == 0) {
    const Tout edge = Tout(0.0 - (a.per ? (m0 ? eN : 0.0) : sub));  // +0 for helm
    if (m0 || !za) pa[N] = edge;
    pb[N] = edge;
  }
  if (m0 && !a.per && j0 + R == N) {
    Tout *pn = ps + (size_t)N * ld1;
    for (int i = threadIdx.x; i < G; i += nt) pn[i] = Tout(-kappa);
  }
}


// ------------------------------------------------------------ K1/K3 MFMA
// PQ_PREC_FP32_MFMA row passes: the sine transform of each row on the
// matrix cores (pq_mfma.h), one row per wavefront at a time, 4 wavefronts
// per workgroup.  Workgroups are persistent: each loops over scratch
// blocks (kMfmaR rows of one layer) with the grid's stride, so the
// per-lane constants are loaded once per workgroup.  A block is staged in
// LDS as kMfmaR rows of kMfmaRs floats: all of its loads are issued
// together (the whole block in flight per workgroup), the rows are
// transformed in place, and the block moves to or from the (p, r) scratch
// layout as 16-byte vectors (the row stride keeps those transposed LDS
// accesses at most 2-way bank conflicted).  FP32 fields and scratch; the
// closed basin only.
constexpr int kMfmaR = mfma::kRows;      // rows per block (the handle's R)
constexpr int kMfmaRs = mfma::kN + 4;    // LDS row stride (floats)
constexpr int kMfmaNt = mfma::kRowThreads;

template <typename Tin>
__global__ void __launch_bounds__(mfma::kRowThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_rows_fwd_mfma(InvArgs a, PtrTab ptrs) {
  constexpr int N = mfma::kN, RS = kMfmaRs, R = kMfmaR, NB = N / R, NT = kMfmaNt;
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned *a3s = reinterpret_cast<unsigned *>(smem);
  float *st = reinterpret_cast<float *>(smem) + mfma::kA3Lds;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c = lane & 31, g = lane >> 5;
  mfma::Consts K;
  mfma::load_consts(static_cast<const unsigned *>(a.mtab), lane, K);
  mfma::stage_a3(static_cast<const unsigned *>(a.mtab), a3s, tid, NT);
  const int nblk = ptrs.cum_r[ptrs.nb];
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int b = member_of(ptrs.cum_r, ptrs.nb, blk);
    const int rem = blk - ptrs.cum_r[b];
    const int l = rem / NB, jb = rem - l * NB, j0 = jb * R;
    const Tin *z = static_cast<const Tin *>(ptrs.zeta[b]) + (size_t)l * ptrs.ldz1 * ptrs.ldz2 +
                   (size_t)j0 * ptrs.ldz1;
    // the block's rows: x_i, i = 0..N-1 (the row's node i = N is a
    // boundary node and not read); row j = 0 (the boundary) as zeros
    float v[R][4];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) v[r][q] = float(z[(size_t)r * ptrs.ldz1 + 4 * tid + q]);
    __syncthreads();  // the A3 copy (first block) / the previous block's reads of st
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) st[r * RS + 4 * tid + q] = (j0 + r == 0) ? 0.0f : v[r][q];
    __syncthreads();
    for (int r = w; r < R; r += NT / 64) {
      if (j0 + r == 0) continue;  // the boundary row stays zero
      float *row = st + r * RS;
      float x[2][8];
      const float *src = row + c + 256 * g;
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int s2 = 0; s2 < 8; ++s2) x[t][s2] = src[512 * t + 32 * s2];
      if (lane == 0) x[0][0] = 0.0f;  // x_0: the boundary node
      mfma::dst_row(K, a3s, x, lane, row);
    }
    __syncthreads();
    // Block (l, jb): S[p][r] = row r at p, r fastest; one contiguous N R run.
    float *out = reinterpret_cast<float *>(a.scr + b * a.scr_stride) + (size_t)(l * NB + jb) * N * R;
#pragma unroll
    for (int k = 0; k < N * R / 4 / NT; ++k) {
      const int i = tid + NT * k, e0 = 4 * i, p = e0 / R, r0 = e0 % R;
      const float *src = st + r0 * RS + p;
      float4 o;
      o.x = src[0];
      o.y = src[RS];
      o.z = src[2 * RS];
      o.w = src[3 * RS];
      reinterpret_cast<float4 *>(out)[i] = o;
    }
  }
}

// Lowest input modes of the inverse transform taken out of the matrix-core
// transform and added to psi in FP64 (as the FP32 engines do): the
// solution's spectrum is dominated by them, and the split operands' error
// scales with the row's largest input.  Two modes take most of the gain
// (the spectrum falls off like p^-3.5; measured on the 1023^2 workload:
// 5.1e-7 with none, 1.3e-7 with two, 1.2e-7 with four).
constexpr int kMfmaLow = 2;

// cos and sin of pi k / 16, k = 0..15 (compile-time after unrolling)
__device__ __forceinline__ double cospi16(int k) {
  constexpr double t[16] = {1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
                            0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173,
                            0.19509032201612826785, 0.0, -0.19509032201612826785, -0.38268343236508977173,
                            -0.55557023301960222474, -0.70710678118654752440, -0.83146961230254523708,
                            -0.92387953251128675613, -0.98078528040323044913};
  return t[k];
}
__device__ __forceinline__ double sinpi16(int k) { return k < 8 ? cospi16(8 - k) : cospi16(k - 8); }

template <typename Tout>
__global__ void __launch_bounds__(mfma::kRowThreads) __attribute__((amdgpu_waves_per_eu(2)))
k_rows_inv_mfma(InvArgs a, PtrTab ptrs) {
  constexpr int N = mfma::kN, G = N + 1, RS = kMfmaRs, R = kMfmaR, NB = N / R, NT = kMfmaNt;
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned *a3s = reinterpret_cast<unsigned *>(smem);
  float *st = reinterpret_cast<float *>(smem) + mfma::kA3Lds;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c = lane & 31, g = lane >> 5;
  const bool m0 = a.mode == 0;
  // sin and cos of pi lane / N (FP64) for the psi rows' low modes
  const cx<double> t0 = static_cast<const cx<double> *>(a.tw64)[lane];  // exp(-i pi lane / N)
  const double sb = -t0.y, cb = t0.x;
  mfma::Consts K;
  mfma::load_consts(static_cast<const unsigned *>(a.mtab), lane, K);
  mfma::stage_a3(static_cast<const unsigned *>(a.mtab), a3s, tid, NT);
  const int nblk = ptrs.cum_r[ptrs.nb];
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int b = member_of(ptrs.cum_r, ptrs.nb, blk);
    const int L = ptrs.lay[b];
    const int rem = blk - ptrs.cum_r[b];
    const int l = rem / NB, jb = rem - l * NB, j0 = jb * R;
    const CoefLayout CL{L, N};
    // the tile block, all loads in flight, then into LDS rows r at st + r RS
    const float4 *in = reinterpret_cast<const float4 *>(a.scr + b * a.scr_stride) + (size_t)(l * NB + jb) * N * R / 4;
    float4 v[N * R / 4 / NT];
#pragma unroll
    for (int k = 0; k < N * R / 4 / NT; ++k) v[k] = in[tid + NT * k];
    double kappa = 0.0;
    double beta[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) beta[m] = 0.0;
    if (m0) {
      const double *kb = a.kb + (size_t)b * a.kb_stride;
      kappa = kb[l];
#pragma unroll
      for (int m = 1; m < 8; ++m)
        if (m < L) beta[m] = kb[L + l * L + m];
    }
    // the mass correction's per-column constants (rmin, then Gamma_l)
    const double *kbm = a.kb + (size_t)b * a.kb_stride + L + L * L;
    MassPre mpre{};
    if (m0 && L > 1) mpre = mass_preload(kbm, kbm + (1 + l) * (N / 2), N, tid, NT);
    __syncthreads();  // the previous block's reads of st
#pragma unroll
    for (int k = 0; k < N * R / 4 / NT; ++k) {
      const int e0 = 4 * (tid + NT * k), p = e0 / R, r0 = e0 % R;
      float *dst = st + r0 * RS + p;
      dst[0] = v[k].x;
      dst[RS] = v[k].y;
      dst[2 * RS] = v[k].z;
      dst[3 * RS] = v[k].w;
    }
    __syncthreads();  // publishes the tile (and the A3 copy)
    if (m0 && L > 1) {
      mass_correction<N>(ptrs.coef[b] + CL.tab(), beta, L, jb, NB, R, ilog2c(R), tid, NT, 20.0, kbm,
                         kbm + (1 + l) * (N / 2), mpre, 0, [&](int r, int p, double x) {
                           float *e = st + r * RS + p;
                           *e = float(double(*e) - x);
                         });
      __syncthreads();
    }
    // Each wavefront transforms its rows in place and writes their psi rows
    // (no barrier: a row is only touched by its wavefront).  pv_inversion
    // (mode 0): psi = X + low - kappa, -kappa on the boundary; helm (mode
    // 1) keeps the reference's boundary semantics (row 0 zero at
    // 1 <= i <= N - 1, X at 1 <= i <= N - 1 and 0 at i = N elsewhere, column
    // 0 and row N not written).  low = sum_{p <= kMfmaLow} 2 x_p
    // sin(pi p i / N) in FP64: the lowest modes leave the matrix-core
    // transform (zeroed in its operand), their sines by the Chebyshev
    // recurrence from sin and cos(pi i / N).
    Tout *ps = static_cast<Tout *>(ptrs.psi[b]) + (size_t)l * ptrs.ldp1 * ptrs.ldp2;
    const long long ld1 = ptrs.ldp1;
    const double sub = m0 ? kappa : 0.0;
    // A wavefront takes its rows in pairs (r, r + RW): both transforms, then
    // one pass over i that forms the low modes' sines once for the two rows.
    constexpr int RW = NT / 64;
    static_assert(R % (2 * RW) == 0, "row pairs of a wavefront");
    static_assert(kMfmaLow == 2, "the low-mode sum is written for 2 modes");
    for (int r = w; r < R; r += 2 * RW) {
      // low = s1 (w1 + 2 w2 c1), s1 = sin(pi i / N), c1 = cos(pi i / N)
      // (sin 2t = 2 sin t cos t), for the weights w_p = 2 x_p; row j = 0 is
      // zero (no transform)
      double p0[2], p1[2];
      bool zr[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float *row = st + (r + h * RW) * RS;
        zr[h] = j0 + r + h * RW == 0;
        p0[h] = 0.0;
        p1[h] = 0.0;
        if (zr[h]) continue;  // uniform in the wavefront
        p0[h] = 2.0 * double(row[1]);
        p1[h] = 4.0 * double(row[2]);
        float x[2][8];
        const float *src = row + c + 256 * g;
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int s2 = 0; s2 < 8; ++s2) x[t][s2] = src[512 * t + 32 * s2];
        if (lane <= kMfmaLow) x[0][0] = 0.0f;  // x_0 and the low modes (lanes 1..kMfmaLow)
        mfma::dst_row(K, a3s, x, lane, row);
      }
      // the rows' psi at i = lane + 64 k (0..N - 1), then i = N; the
      // angle pi i / N = pi lane / N + pi k / 16
#pragma unroll 4
      for (int k = 0; k < N / 64; ++k) {
        const int i = lane + 64 * k;
        const double ca = cospi16(k), sa = sinpi16(k);
        const double s1 = fma(sb, ca, cb * sa), c1 = fma(cb, ca, -sb * sa);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const float *row = st + (r + h * RW) * RS;
          const double y = (zr[h] || i == 0) ? 0.0 : fma(s1, fma(p1[h], c1, p0[h]), double(row[i]));
          if (m0 || i != 0) ps[(size_t)(j0 + r + h * RW) * ld1 + i] = Tout(y - sub);
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (m0 || !zr[h]) ps[(size_t)(j0 + r + h * RW) * ld1 + N] = Tout(0.0 - sub);
      }
    }
    if (m0 && j0 + R == N) {
      Tout *pn = ps + (size_t)N * ld1;
      for (int i = tid; i < G; i += NT) pn[i] = Tout(-kappa);
    }
  }
}

// ------------------------------------------------------------ dispatch
template <typename Eng, typename CE, typename Ts> static EngineInfo info_of() {
  static_assert(std::is_same<typename CE::T, typename Eng::T>::value &&
                    std::is_same<typename CE::Tc, typename Eng::Tc>::value,
                "channel engine: types of the row engines");
  return EngineInfo{Eng::M, Eng::TPF, Eng::Region, Eng::ScanTmp, (int)sizeof(typename Eng::T),
                    (int)sizeof(Ts), (int)sizeof(typename Eng::Tc), Eng::Fwd::Region,
                    Eng::Fwd::ScanTmp, CE::TPF, CE::Region};
}

// Engines per precision.  FP32 transforms are odd extensions (no prefix
// sum: FFTPACK sint's recurrence accumulates ~sqrt(N) eps, which reaches psi
// wherever ss_m flattens the solve's damping of high wavenumbers); the
// inverse ones sum the lowest kLow input modes in FP64 (the solution's
// spectrum is dominated by them, so the FFT's rounding error, which scales
// with its input norm, shrinks with them).
constexpr int kLow = PQ_N >= 16 ? 4 : PQ_N / 2 - 1;
using ExtF = DstExt<float, PQ_N, float>;
using ExtFLow = DstExt<float, PQ_N, float, kLow>;
using SintD = DstSint<double, PQ_N>;
using ExtWide = DstExt<float, PQ_N, double>;
using P32 = EngPair<ExtF, ExtFLow>;
using P64 = EngPair<SintD, SintD>;
using PWide = EngPair<ExtWide, ExtWide>;
// The channel's column engines (periodic transforms of length 2N), with
// the LDS and arithmetic types of the row engines of each precision.
using C32 = PerRfft<float, PQ_N, float>;
using C64 = PerRfft<double, PQ_N, double>;
using CWide = PerRfft<float, PQ_N, double>;

// Dynamic LDS above the 64 KiB default needs a per-kernel limit (gfx950 has
// 160 KiB per CU); raised once per kernel to the largest size requested.
// An error of the attribute call is cleared; a launch that still exceeds the
// limit reports its own error.
template <typename K> static void lds_limit(K kernel, size_t bytes) {
  static size_t granted = 64 * 1024;
  if (bytes <= granted) return;
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (hipGetLastError() == hipSuccess) granted = bytes;
}

// Eng: transform engine (compute type), CE: the channel's column engine,
// Tio: zeta/psi type, Ts: scratch type.
template <typename Eng, typename CE, typename Tio, typename Ts>
static hipError_t launch_pass(int pass, const InvArgs &a, const PtrTab &pt, int blocks, int threads,
                              size_t lds, hipStream_t s) {
  if (pass == 0) {
    lds_limit(k_rows_fwd<Eng, Tio, Ts>, lds);
    hipLaunchKernelGGL((k_rows_fwd<Eng, Tio, Ts>), dim3(blocks), dim3(threads), lds, s, a, pt);
  } else if (pass == 1 && a.per) {
    lds_limit(k_cols_chan<CE, Ts>, lds);
    hipLaunchKernelGGL((k_cols_chan<CE, Ts>), dim3(blocks), dim3(threads), lds, s, a, pt);
  } else if (pass == 1) {
    switch (a.L) {
      case 1: hipLaunchKernelGGL((k_cols_tri<Ts, 1, PQ_N>), dim3(blocks), dim3(threads), 0, s, a, pt); break;
      case 2: hipLaunchKernelGGL((k_cols_tri<Ts, 2, PQ_N>), dim3(blocks), dim3(threads), 0, s, a, pt); break;
      case 3: hipLaunchKernelGGL((k_cols_tri<Ts, 3, PQ_N>), dim3(blocks), dim3(threads), 0, s, a, pt); break;
      case 4: hipLaunchKernelGGL((k_cols_tri<Ts, 4, PQ_N>), dim3(blocks), dim3(threads), 0, s, a, pt); break;
      case 5: case 6: hipLaunchKernelGGL((k_cols_tri<Ts, 6, PQ_N>), dim3(blocks), dim3(threads), 0, s, a, pt); break;
      default: hipLaunchKernelGGL((k_cols_tri<Ts, 8, PQ_N>), dim3(blocks), dim3(threads), 0, s, a, pt); break;
    }
  } else {
    lds_limit(k_rows_inv<Eng, Tio, Ts>, lds);
    hipLaunchKernelGGL((k_rows_inv<Eng, Tio, Ts>), dim3(blocks), dim3(threads), lds, s, a, pt);
  }
  return hipGetLastError();
}

#define PQ_CAT2(a, b) a##b
#define PQ_CAT(a, b) PQ_CAT2(a, b)

// PQ_PREC_FP64: FP64 everywhere (FFTPACK-style sint engine).
// PQ_PREC_FP32: FP32 fields, scratch and transforms (P32 engines).
// PQ_PREC_MIXED: FP64 fields, FP32 scratch and transforms.
// PQ_PREC_FP32_STORAGE: FP32 fields and scratch, FP64 transforms.
// PQ_PREC_FP32_WIDE: FP32 fields, scratch and LDS, FP64 arithmetic
// (odd-extension engine).
EngineInfo PQ_CAT(inv_info_, PQ_N)(int prec) {
  switch (prec) {
    case PQ_PREC_FP64: return info_of<P64, C64, double>();
    case PQ_PREC_FP32_STORAGE: return info_of<P64, C64, float>();
    case PQ_PREC_FP32_WIDE: return info_of<PWide, CWide, float>();
    default: return info_of<P32, C32, float>();
  }
}

// PQ_PREC_FP32_MFMA: FP32 fields and scratch, row transforms on the matrix
// cores (lengths mfma::supported), the FP32 column pass.
static hipError_t launch_mfma(int pass, const InvArgs &a, const PtrTab &pt, int blocks, int threads,
                              size_t lds, hipStream_t s) {
  if constexpr (mfma::supported(PQ_N)) {
    if (pass == 0) {
      lds_limit(k_rows_fwd_mfma<float>, lds);
      hipLaunchKernelGGL((k_rows_fwd_mfma<float>), dim3(blocks), dim3(threads), lds, s, a, pt);
      return hipGetLastError();
    }
    if (pass == 2) {
      lds_limit(k_rows_inv_mfma<float>, lds);
      hipLaunchKernelGGL((k_rows_inv_mfma<float>), dim3(blocks), dim3(threads), lds, s, a, pt);
      return hipGetLastError();
    }
    return launch_pass<P32, C32, float, float>(pass, a, pt, blocks, threads, lds, s);
  } else {
    return hipErrorInvalidValue;
  }
}

hipError_t PQ_CAT(inv_launch_, PQ_N)(int prec, int pass, const InvArgs &a, const PtrTab &pt,
                                     int blocks, int threads, size_t lds, hipStream_t s) {
  switch (prec) {
    case PQ_PREC_FP32_MFMA:
      return launch_mfma(pass, a, pt, blocks, threads, lds, s);
    case PQ_PREC_FP64:
      return launch_pass<P64, C64, double, double>(pass, a, pt, blocks, threads, lds, s);
    case PQ_PREC_FP32:
      return launch_pass<P32, C32, float, float>(pass, a, pt, blocks, threads, lds, s);
    case PQ_PREC_FP32_STORAGE:
      return launch_pass<P64, C64, float, float>(pass, a, pt, blocks, threads, lds, s);
    case PQ_PREC_FP32_WIDE:
      return launch_pass<PWide, CWide, float, float>(pass, a, pt, blocks, threads, lds, s);
    default:
      return launch_pass<P32, C32, double, float>(pass, a, pt, blocks, threads, lds, s);
  }
}

}  // namespace pq
