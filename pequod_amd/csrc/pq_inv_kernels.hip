// pq_inv_kernels.hip -- the three passes of the modal Helmholtz inversion,
// compiled once per transform length N = n + 1 (-DPQ_N=<N>).
//
//   k_rows_fwd  per (member, layer, tile of R rows): read zeta rows, DST-I
//               along i, write the tile transposed as one contiguous block
//               S[l][jb][p][r] (r fastest) of the scratch.
//   k_cols      per (member, C adjacent columns p), all layers: read the
//               columns (C*R contiguous elements per block), DST-I along j,
//               project layers -> modes (z_to_modes), divide by
//               (2N)^2 (lambda_p + lambda_q - ss_m), accumulate the FP64
//               partial of the domain mean psiC_m, project modes -> layers
//               (modes_to_z), inverse DST-I along j, write back in place.
//   k_rows_inv  per (member, layer, tile): reduce the psiC partials in a
//               fixed order, read the tile block, subtract the mass
//               correction in the i-spectral domain (closed form of
//               helm(1)), inverse DST-I along i, subtract the constant part,
//               write psi rows including the boundary.
//
// The sine transforms along i and j replace FFTPACK sint along i plus the
// Thomas sweep along j of the reference's helm (qg3p_helmholtz.F90:614-769):
// both solve the same 5-point system (manual/manual.tex:826-832).  The modal
// projection commutes with the transforms, so it runs in spectral space where
// all layers of a column are in LDS together.
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

#ifndef PQ_N
#error "compile with -DPQ_N=<transform length>"
#endif

namespace pq {

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

// ------------------------------------------------------------------ K1
template <typename Eng, typename Tin>
__global__ void __launch_bounds__(512) k_rows_fwd(InvArgs a, PtrTab ptrs) {
  using T = typename Eng::T;
  constexpr int N = Eng::N, TPF = Eng::TPF;
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int4 w = a.work[blockIdx.x];
  const MemDesc md = a.mems[w.x];
  const int R = a.R, lgR = a.lgR;
  const int l = w.y, jb = w.z, j0 = jb << lgR;
  const int f = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int ngrp = R >> 1;
  T *region = lds + f * Eng::Region;
  T *scan_tmp = lds + ngrp * Eng::Region + f * Eng::ScanTmp;
  const Tin *z = static_cast<const Tin *>(ptrs.zeta[w.w]) + (size_t)l * ptrs.ldz1 * ptrs.ldz2;
  const int ja = j0 + 2 * f;
  const Tin *za = z + (size_t)ja * ptrs.ldz1;
  const Tin *zb = za + ptrs.ldz1;
  const bool va = ja > 0;  // node row 0 is the boundary: not transformed
  Eng::run(region, t, scan_tmp, static_cast<const cx<T> *>(a.tw), static_cast<const T *>(a.was),
           [&](int k, T &xa, T &xb) {
             xa = va ? T(za[k]) : T(0);
             xb = T(zb[k]);
           });
  __syncthreads();
  // Block (l, jb): S[p][r] = row r at p, r fastest; one contiguous N*R run.
  using VT = typename V16<T>::type;
  constexpr int V = V16<T>::n;
  T *out = static_cast<T *>(md.scr) + ((size_t)(l * (N >> lgR) + jb) * N << lgR);
  const int nvec = (N << lgR) / V;
  for (int iv = threadIdx.x; iv < nvec; iv += blockDim.x) {
    const int e0 = iv * V, r0 = e0 & (R - 1), p = e0 >> lgR;
    VT v;
#pragma unroll
    for (int s = 0; s < V; ++s) v[s] = row_ptr<Eng>(lds, r0 + s)[pad16(p)];
    *reinterpret_cast<VT *>(out + e0) = v;
  }
}

// ------------------------------------------------------------------ K2
template <typename Eng>
__global__ void __launch_bounds__(512) k_cols(InvArgs a) {
  using T = typename Eng::T;
  constexpr int N = Eng::N, TPF = Eng::TPF;
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int4 w = a.work[blockIdx.x];
  const MemDesc md = a.mems[w.x];
  const int L = md.L, C = a.C, lgC = a.lgC, R = a.R, lgR = a.lgR;
  const int p0 = w.y << lgC;
  const int ngrp = blockDim.x / TPF;
  const int f = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int nrows = L << lgC;  // row rho = l * C + c
  T *region = lds + f * Eng::Region;
  T *scan_tmp = lds + ngrp * Eng::Region + f * Eng::ScanTmp;
  const size_t red_off = (sizeof(T) * (size_t)ngrp * (Eng::Region + Eng::ScanTmp) + 15) & ~size_t(15);
  double *red = reinterpret_cast<double *>(smem + red_off);
  T *cD = reinterpret_cast<T *>(red + 512);  // D, M, ssN in T (L <= 8)
  T *cM = cD + 64;
  T *cS = cM + 64;
  const CoefLayout CL{L, N};
  for (int i = threadIdx.x; i < L * L; i += blockDim.x) {
    cD[i] = T(md.coef[CL.D() + i]);
    cM[i] = T(md.coef[CL.M() + i]);
  }
  for (int i = threadIdx.x; i < L; i += blockDim.x) cS[i] = T(md.coef[CL.ssN() + i]);

  // Element j of row (l, p) sits at ((l * NB + j / R) * N + p) * R + j % R.
  // Layer l is N * N elements: N / R blocks of N * R.
  T *scr = static_cast<T *>(md.scr);
  const int ra = 2 * f, rb = ra + 1;
  const int la = ra >> lgC, pa = p0 + (ra & (C - 1));
  const int lb = rb >> lgC, pb = p0 + (rb & (C - 1));
  const bool va = ra < nrows && pa >= 1;
  const bool vb = rb < nrows && pb >= 1;
  const T *sa = scr + (size_t)la * N * N + (size_t)pa * R;  // + (j / R) * N * R + j % R
  const T *sb = scr + (size_t)lb * N * N + (size_t)pb * R;
  const int blk = N << lgR;
  Eng::run(region, t, scan_tmp, static_cast<const cx<T> *>(a.tw), static_cast<const T *>(a.was),
           [&](int k, T &xa, T &xb) {
             const int off = (k >> lgR) * blk + (k & (R - 1));
             xa = va ? sa[off] : T(0);
             xb = vb ? sb[off] : T(0);
           });
  __syncthreads();

  // Spectral step over the points (p0 + c, q).
  const T *lam4 = static_cast<const T *>(a.lam4);
  double acc[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) acc[m] = 0.0;
  for (int c = 0; c < C; ++c) {
    const int p = p0 + c;
    if (p < 1) continue;
    const T lp = lam4[p];
    const double sp = a.shat[p];
    for (int q = 1 + threadIdx.x; q < N; q += blockDim.x) {
      const T lpq = lp + lam4[q];
      if (md.mode == 0) {
        T col[8], phi[8];
#pragma unroll
        for (int ll = 0; ll < 8; ++ll)
          if (ll < L) col[ll] = row_ptr<Eng>(lds, (ll << lgC) + c)[pad16(q)];
#pragma unroll
        for (int m = 0; m < 8; ++m)
          if (m < L) {
            T s = T(0);
#pragma unroll
            for (int ll = 0; ll < 8; ++ll)
              if (ll < L) s += cD[m + L * ll] * col[ll];
            phi[m] = s / (lpq - cS[m]);
          }
        if ((p & 1) && (q & 1)) {
          const double wpq = sp * a.shat[q];
#pragma unroll
          for (int m = 0; m < 8; ++m)
            if (m < L) acc[m] += double(phi[m]) * wpq;
        }
#pragma unroll
        for (int ll = 0; ll < 8; ++ll)
          if (ll < L) {
            T s = T(0);
#pragma unroll
            for (int m = 0; m < 8; ++m)
              if (m < L) s += cM[ll + L * m] * phi[m];
            row_ptr<Eng>(lds, (ll << lgC) + c)[pad16(q)] = s;
          }
      } else {
#pragma unroll
        for (int ll = 0; ll < 8; ++ll)
          if (ll < L) {
            T *x = row_ptr<Eng>(lds, (ll << lgC) + c) + pad16(q);
            *x = *x / (lpq - cS[ll]);
          }
      }
    }
  }
  __syncthreads();

  // Inverse DST along j, in place.
  {
    const T *rowa = region, *rowb = region + Eng::RS;
    Eng::run(region, t, scan_tmp, static_cast<const cx<T> *>(a.tw), static_cast<const T *>(a.was),
             [&](int k, T &xa, T &xb) {
               xa = rowa[pad16(k)];
               xb = rowb[pad16(k)];
             });
  }
  __syncthreads();
  // Write back: per layer and block, C*R contiguous elements.
  using VT = typename V16<T>::type;
  constexpr int V = V16<T>::n;
  const int per_layer = (N << lgC) / V;  // vectors per layer: N/R blocks x C x R
  for (int ll = 0; ll < L; ++ll) {
    for (int iv = threadIdx.x; iv < per_layer; iv += blockDim.x) {
      const int e = iv * V;
      const int r0 = e & (R - 1), c = (e >> lgR) & (C - 1), jb = e >> (lgR + lgC);
      const T *src = row_ptr<Eng>(lds, (ll << lgC) + c);
      VT v;
#pragma unroll
      for (int s = 0; s < V; ++s) v[s] = src[pad16((jb << lgR) + r0 + s)];
      T *dst = scr + (size_t)ll * N * N + ((size_t)jb * N + p0 + c) * R + r0;
      *reinterpret_cast<VT *>(dst) = v;
    }
  }
  if (md.mode != 0) return;

  // Workgroup reduction of the psiC partials (FP64) by a fixed-order tree in
  // LDS.  The row-inverse pass reduces the per-workgroup partials of a member;
  // the kernel boundary makes them visible without device-scope fences.
  const int nt = blockDim.x, tid = threadIdx.x;
  int pw = 1;
  while (pw < nt) pw <<= 1;
  for (int m = 0; m < L; ++m) {
    red[tid] = acc[m];
    __syncthreads();
    for (int st = pw >> 1; st > 0; st >>= 1) {
      if (tid < st && tid + st < nt) red[tid] += red[tid + st];
      __syncthreads();
    }
    if (tid == 0) md.part[(size_t)w.y * L + m] = red[0];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ K3
template <typename Eng, typename Tout>
__global__ void __launch_bounds__(512) k_rows_inv(InvArgs a, PtrTab ptrs) {
  using T = typename Eng::T;
  constexpr int N = Eng::N, TPF = Eng::TPF;
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int4 w = a.work[blockIdx.x];
  const MemDesc md = a.mems[w.x];
  const int L = md.L, R = a.R, lgR = a.lgR;
  const int l = w.y, jb = w.z, j0 = jb << lgR;
  const int f = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int ngrp = R >> 1;
  T *region = lds + f * Eng::Region;
  T *scan_tmp = lds + ngrp * Eng::Region + f * Eng::ScanTmp;
  const CoefLayout CL{L, N};

  // psiC_m: reduce the column pass's partials in a fixed order (every
  // workgroup of the member does the same reduction), then the correction
  // constants (FP64).
  double kappa = 0.0;
  double beta[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) beta[m] = 0.0;
  if (md.mode == 0) {
    const size_t red_off = (sizeof(T) * (size_t)ngrp * (Eng::Region + Eng::ScanTmp) + 15) & ~size_t(15);
    double *red = reinterpret_cast<double *>(smem + red_off);
    double *s_psic = red + 512;
    const int nt = blockDim.x, tid = threadIdx.x, nblk = N >> a.lgC;
    int pw = 1;
    while (pw < nt) pw <<= 1;
    for (int m = 0; m < L; ++m) {
      double part = 0.0;
      for (int b = tid; b < nblk; b += nt) part += md.part[(size_t)b * L + m];
      red[tid] = part;
      __syncthreads();
      for (int st = pw >> 1; st > 0; st >>= 1) {
        if (tid < st && tid + st < nt) red[tid] += red[tid + st];
        __syncthreads();
      }
      if (tid == 0) s_psic[m] = red[0] / (double(N) * double(N));
      __syncthreads();
    }
    if (jb == 0 && l == 0 && tid == 0) {
      int st = 0;
      for (int m = 0; m < L; ++m) {
        const double v = s_psic[m];
        md.psic[m] = v;
        if (v != v) st |= 1;
        else if (fabs(v) > 1.0e100) st |= 2;
      }
      if (st) atomicOr(md.status, st);
    }
    const double *M = md.coef + CL.M();
    const double *ca = md.coef + CL.a();
    const double *cb = md.coef + CL.b();
    kappa = s_psic[0];
    for (int m = 1; m < L; ++m) {
      const double pc = s_psic[m];
      kappa += M[l + L * m] * pc * ca[m];
      beta[m] = M[l + L * m] * pc * cb[m];
    }
    __syncthreads();
  }

  // Tile block (l, jb): row r at p <- S[p * R + r].
  {
    using VT = typename V16<T>::type;
    constexpr int V = V16<T>::n;
    const T *in = static_cast<const T *>(md.scr) + ((size_t)(l * (N >> lgR) + jb) * N << lgR);
    const int nvec = (N << lgR) / V;
    for (int iv = threadIdx.x; iv < nvec; iv += blockDim.x) {
      const int e0 = iv * V, r0 = e0 & (R - 1), p = e0 >> lgR;
      const VT v = *reinterpret_cast<const VT *>(in + e0);
#pragma unroll
      for (int s = 0; s < V; ++s) row_ptr<Eng>(lds, r0 + s)[pad16(p)] = v[s];
    }
  }
  __syncthreads();

  // Correction field for odd p, by recurrence over the tile's rows (FP64).
  if (md.mode == 0 && L > 1) {
    const double *tab = md.coef + CL.tab();
    for (int h = threadIdx.x; h < N / 2; h += blockDim.x) {
      const int p = 2 * h + 1;
      double acc[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.0;
      for (int m = 1; m < L; ++m) {
        const double *e = tab + ((size_t)(m - 1) * (N / 2) + h) * 3;
        const double rho = e[0], gam = beta[m] * e[1], eta = e[2];
        const double rr = exp(rho);
        double hh[16];
        double u = exp(double(j0) * rho);  // r^(j0 + r), going up
#pragma unroll
        for (int r = 0; r < 16; ++r)
          if (r < R) {
            hh[r] = u;
            u *= rr;
          }
        double v = exp(double(N - j0 - R + 1) * rho);  // r^(N - j0 - r), going down
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
        if (r < R) {
          T *x = row_ptr<Eng>(lds, r) + pad16(p);
          *x = T(double(*x) - acc[r]);
        }
    }
    __syncthreads();
  }

  {
    const T *rowa = region, *rowb = region + Eng::RS;
    Eng::run(region, t, scan_tmp, static_cast<const cx<T> *>(a.tw), static_cast<const T *>(a.was),
             [&](int k, T &xa, T &xb) {
               xa = rowa[pad16(k)];
               xb = rowb[pad16(k)];
             });
  }
  __syncthreads();

  Tout *ps = static_cast<Tout *>(ptrs.psi[w.w]) + (size_t)l * ptrs.ldp1 * ptrs.ldp2;
  const long long ld1 = ptrs.ldp1;
  constexpr int G = N + 1;
  const bool last = (j0 + R == N);
  const int nrow = R + (last ? 1 : 0);
  for (int idx = threadIdx.x; idx < nrow * G; idx += blockDim.x) {
    const int r = idx / G, i = idx - r * G, j = j0 + r;
    Tout *o = ps + (size_t)j * ld1 + i;
    if (md.mode == 0) {
      double val;
      if (j == 0 || j == N || i == 0 || i == N) val = -kappa;
      else val = double(row_ptr<Eng>(lds, r)[pad16(i)]) - kappa;
      *o = Tout(val);
    } else {
      // helm's boundary semantics (qg3p_helmholtz.F90:757-761, 1055-1057)
      if (j == N || i == 0) continue;
      if (j == 0) {
        if (i <= N - 1) *o = Tout(0);
      } else {
        *o = i == N ? Tout(0) : Tout(row_ptr<Eng>(lds, r)[pad16(i)]);
      }
    }
  }
}

// ------------------------------------------------------------ dispatch
template <typename Eng> static EngineInfo info_of() {
  return EngineInfo{Eng::M, Eng::TPF, Eng::Region, Eng::ScanTmp, (int)sizeof(typename Eng::T)};
}

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

template <typename Eng, typename Tio>
static hipError_t launch_pass(int pass, const InvArgs &a, const PtrTab &pt, int blocks, int threads,
                              size_t lds, hipStream_t s) {
  if (pass == 0) {
    lds_limit(k_rows_fwd<Eng, Tio>, lds);
    hipLaunchKernelGGL((k_rows_fwd<Eng, Tio>), dim3(blocks), dim3(threads), lds, s, a, pt);
  } else if (pass == 1) {
    lds_limit(k_cols<Eng>, lds);
    hipLaunchKernelGGL((k_cols<Eng>), dim3(blocks), dim3(threads), lds, s, a);
  } else {
    lds_limit(k_rows_inv<Eng, Tio>, lds);
    hipLaunchKernelGGL((k_rows_inv<Eng, Tio>), dim3(blocks), dim3(threads), lds, s, a, pt);
  }
  return hipGetLastError();
}

#define PQ_CAT2(a, b) a##b
#define PQ_CAT(a, b) PQ_CAT2(a, b)

EngineInfo PQ_CAT(inv_info_, PQ_N)(int prec) {
  return prec == PQ_PREC_FP64 ? info_of<DstSint<double, PQ_N>>() : info_of<DstExt<float, PQ_N>>();
}

hipError_t PQ_CAT(inv_launch_, PQ_N)(int prec, int pass, const InvArgs &a, const PtrTab &pt,
                                     int blocks, int threads, size_t lds, hipStream_t s) {
  switch (prec) {
    case PQ_PREC_FP64:
      return launch_pass<DstSint<double, PQ_N>, double>(pass, a, pt, blocks, threads, lds, s);
    case PQ_PREC_FP32:
      return launch_pass<DstExt<float, PQ_N>, float>(pass, a, pt, blocks, threads, lds, s);
    default:
      return launch_pass<DstExt<float, PQ_N>, double>(pass, a, pt, blocks, threads, lds, s);
  }
}

}  // namespace pq
