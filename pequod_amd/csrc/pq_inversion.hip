// pq_inversion.hip -- modal Helmholtz / PV inversion on MI355X (gfx950).
//
// Restates the reference's pv_inversion (qg3p_aux.F90:452-521) and helm
// (qg3p_helmholtz.F90:614-769) as three bandwidth-bound passes over HBM:
//
//   k_rows_fwd  per (member, layer, tile of R rows): read zeta rows, DST-I
//               along i, write the tile transposed into scratch
//               S[l][p][j] (j contiguous).
//   k_cols      per (member, tile of C columns p), all layers: DST-I along
//               j, project layers -> modes (z_to_modes), divide by
//               (2N)^2 (lambda_p + lambda_q - ss_m), accumulate the FP64
//               domain mean psiC_m, project modes -> layers (modes_to_z),
//               inverse DST-I along j, write back.  The last workgroup of a
//               member reduces the psiC partials in a fixed order.
//   k_rows_inv  per (member, layer, tile of R rows): read the transposed
//               tile, subtract the mass-correction field in the i-spectral
//               domain (closed form of helm(1), see below), inverse DST-I
//               along i, subtract the constant part, write psi rows
//               including the boundary.
//
// Both sine transforms (i and j) replace FFTPACK sint along i plus the
// Thomas sweep along j: they solve the same 5-point system
// (manual/manual.tex:826-832).  The modal projection commutes with the
// transforms, so it is applied in spectral space where all layers of a
// column are in LDS together.
//
// Mass correction (qg3p_aux.F90:196-230, qg3p_diagnostics.F90:615-635): with
// c_m = helm_m(1) (Dirichlet 0), corr_m = a_m + b_m c_m where
// a_m = 1 / (1 + ss_m mean(c_m)) and b_m = ss_m a_m, so
//   psi_l = sum_m M(l,m) phi'_m - kappa_l - sum_{m>=2} beta_lm c_m
// with kappa_l = psiC_1 + sum_{m>=2} M(l,m) psiC_m a_m and
// beta_lm = M(l,m) psiC_m b_m.  After the i-transform, c_m is separable:
// for odd p it is shat_p / (2N) * g_mp(j) with
//   g_mp(j) = (1 - (r^j + r^(N-j)) / (1 + r^N)) / mu,  mu = lambda_p - ss_m,
// r + 1/r = 2 - mu (0 < r < 1), the exact solution of the tridiagonal
// problem g(j+1) + g(j-1) + (mu - 2) g(j) = 1, g(0) = g(N) = 0; for even p it
// is zero.  So no correction field is read from HBM.
#include "pq_fft.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace pq {

// Per-member descriptor (static part, built when a handle is created).
struct MemDesc {
  const double *coef;  // see coef_layout below
  void *scr;           // L x N x N scratch (compute type)
  double *part;        // psiC partials: (N / C) x L
  double *psic;        // L (final psiC)
  int *status;         // sticky: 1 NaN psiC, 2 |psiC| > 1e100
  int L;
  int mode;            // 0 pv_inversion, 1 helm
};

// Coefficient block of one member (doubles):
//   D[L*L]   z_to_modes, D(m,l) at m + L*l
//   M[L*L]   modes_to_z, M(l,m) at l + L*m
//   ss[L], a[L], b[L], ssN[L] = (2N)^2 ss
//   tab[(L-1) * N/2 * 3]: (rho, gamma, eta) for mode m >= 1 (0-based) and
//                         odd p = 2h + 1
struct CoefLayout {
  int L, N;
  __host__ __device__ int D() const { return 0; }
  __host__ __device__ int M() const { return L * L; }
  __host__ __device__ int ss() const { return 2 * L * L; }
  __host__ __device__ int a() const { return 2 * L * L + L; }
  __host__ __device__ int b() const { return 2 * L * L + 2 * L; }
  __host__ __device__ int ssN() const { return 2 * L * L + 3 * L; }
  __host__ __device__ int tab() const { return 2 * L * L + 4 * L; }
  __host__ __device__ size_t size() const { return (size_t)tab() + (size_t)(L > 1 ? L - 1 : 0) * (N / 2) * 3; }
};

constexpr int kMaxBatchArg = 64;
struct PtrTab {
  const void *zeta[kMaxBatchArg];
  void *psi[kMaxBatchArg];
  long long ldz1, ldz2, ldp1, ldp2;
};

template <typename T> __device__ __forceinline__ T dev_exp(T x);
template <> __device__ __forceinline__ float dev_exp<float>(float x) { return __expf(x); }
template <> __device__ __forceinline__ double dev_exp<double>(double x) { return exp(x); }

// ------------------------------------------------------------------ K1
template <typename Tin, typename T, int E>
__global__ void __launch_bounds__(512)
k_rows_fwd(DevPlan P, const MemDesc *__restrict__ mems, PtrTab ptrs,
           const int4 *__restrict__ work) {
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int4 w = work[blockIdx.x];
  const MemDesc md = mems[w.x];
  const int l = w.y, j0 = w.z * P.R, N = P.N;
  const int f = threadIdx.x >> P.lgtpf, t = threadIdx.x & (P.tpf - 1);
  const int nfft = blockDim.x >> P.lgtpf;
  T *region = lds + (size_t)2 * f * P.RS;
  T *scan_tmp = lds + (size_t)P.R * P.RS + 2 * f * (P.tpf > 64 ? P.tpf / 64 : 1);
  const Tin *z = reinterpret_cast<const Tin *>(ptrs.zeta[w.w]) +
                 (size_t)l * ptrs.ldz1 * ptrs.ldz2;
  const int ja = j0 + 2 * f, jb = ja + 1;
  const Tin *za = z + (size_t)ja * ptrs.ldz1, *zb = z + (size_t)jb * ptrs.ldz1;
  const bool va = ja > 0;  // node row 0 is the boundary: not transformed
  auto ld = [&](int k, T &xa, T &xb) {
    xa = va ? T(za[k]) : T(0);
    xb = T(zb[k]);
  };
  dst_pair<T, E>(reinterpret_cast<cx<T> *>(region), region, region + P.RS, P,
                 t, scan_tmp, ld);
  (void)nfft;
  __syncthreads();
  // Transposed store: S[l][p][j0 + r], r fastest.
  T *out = reinterpret_cast<T *>(md.scr) + (size_t)l * N * N + j0;
  const int R = P.R, lgR = P.lgR;
  for (int idx = threadIdx.x; idx < N * R; idx += blockDim.x) {
    const int r = idx & (R - 1), p = idx >> lgR;
    out[(size_t)p * N + r] = lds[r * P.RS + pad16(p)];
  }
}

// ------------------------------------------------------------------ K2
template <typename T, int E>
__global__ void __launch_bounds__(512)
k_cols(DevPlan P, const MemDesc *__restrict__ mems,
       const int4 *__restrict__ work) {
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int4 w = work[blockIdx.x];
  const MemDesc md = mems[w.x];
  const int L = md.L, N = P.N, n = N - 1, C = P.C;
  const int p0 = w.y * C;
  const int nfft = blockDim.x >> P.lgtpf;
  const int f = threadIdx.x >> P.lgtpf, t = threadIdx.x & (P.tpf - 1);
  const int nrows = 2 * nfft;
  T *region = lds + (size_t)2 * f * P.RS;
  T *scan_tmp = lds + (size_t)nrows * P.RS + 2 * f * (P.tpf > 64 ? P.tpf / 64 : 1);
  const size_t red_off =
      ((sizeof(T) * ((size_t)nrows * P.RS + 2 * nfft * (P.tpf > 64 ? P.tpf / 64 : 1) + 2)) + 15) &
      ~size_t(15);
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

  const T *scr = reinterpret_cast<const T *>(md.scr);
  const int ra = 2 * f, rb = ra + 1;  // row index rho = c * L + l
  const int ca = ra / L, la = ra - ca * L, cb = rb / L, lb = rb - cb * L;
  const int pa = p0 + ca, pb = p0 + cb;
  const bool va = ra < L * C && pa >= 1 && pa <= n;
  const bool vb = rb < L * C && pb >= 1 && pb <= n;
  const T *sa = scr + ((size_t)la * N + pa) * N;
  const T *sb = scr + ((size_t)lb * N + pb) * N;
  auto ldg = [&](int k, T &xa, T &xb) {
    xa = va ? sa[k] : T(0);
    xb = vb ? sb[k] : T(0);
  };
  dst_pair<T, E>(reinterpret_cast<cx<T> *>(region), region, region + P.RS, P,
                 t, scan_tmp, ldg);
  __syncthreads();

  // Spectral step over points (c, q).
  const T *lam4 = reinterpret_cast<const T *>(P.lam4);
  double acc[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) acc[m] = 0.0;
  for (int idx = threadIdx.x; idx < C * N; idx += blockDim.x) {
    const int c = idx / N, q = idx - c * N, p = p0 + c;
    if (p < 1 || p > n || q < 1) continue;
    T *col = lds + (size_t)(c * L) * P.RS + pad16(q);
    const T lpq = lam4[p] + lam4[q];
    if (md.mode == 0) {
      T phi[8];
      for (int m = 0; m < L; ++m) {
        T s = T(0);
        for (int ll = 0; ll < L; ++ll) s += cD[m + L * ll] * col[ll * P.RS];
        phi[m] = s / (lpq - cS[m]);
      }
      if ((p & 1) && (q & 1)) {
        const double wpq = P.shat[p] * P.shat[q];
        for (int m = 0; m < L; ++m) acc[m] += double(phi[m]) * wpq;
      }
      for (int ll = 0; ll < L; ++ll) {
        T s = T(0);
        for (int m = 0; m < L; ++m) s += cM[ll + L * m] * phi[m];
        col[ll * P.RS] = s;
      }
    } else {
      for (int ll = 0; ll < L; ++ll) col[ll * P.RS] = col[ll * P.RS] / (lpq - cS[ll]);
    }
  }
  __syncthreads();

  // Inverse DST along j, in place.
  T *rowa = lds + (size_t)ra * P.RS, *rowb = lds + (size_t)rb * P.RS;
  auto lds_ld = [&](int k, T &xa, T &xb) {
    xa = rowa[pad16(k)];
    xb = rowb[pad16(k)];
  };
  dst_pair<T, E>(reinterpret_cast<cx<T> *>(region), rowa, rowb, P, t, scan_tmp,
                 lds_ld);
  __syncthreads();
  T *scw = reinterpret_cast<T *>(md.scr);
  for (int rho = 0; rho < L * C; ++rho) {
    const int c = rho / L, ll = rho - c * L, p = p0 + c;
    if (p < 1 || p > n) continue;
    T *dst = scw + ((size_t)ll * N + p) * N;
    const T *src = lds + (size_t)rho * P.RS;
    for (int k = threadIdx.x; k < N; k += blockDim.x) dst[k] = src[pad16(k)];
  }
  if (md.mode != 0) return;

  // Workgroup reduction of the psiC partials (FP64) by a fixed-order tree in
  // LDS (any workgroup size).  The row-inverse kernel reduces the per-block
  // partials of a member; the kernel boundary makes them visible without
  // device-scope fences (which would write back the XCD's L2).
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
template <typename T, typename Tout, int E>
__global__ void __launch_bounds__(512)
k_rows_inv(DevPlan P, const MemDesc *__restrict__ mems, PtrTab ptrs,
           const int4 *__restrict__ work) {
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int4 w = work[blockIdx.x];
  const MemDesc md = mems[w.x];
  const int L = md.L, l = w.y, j0 = w.z * P.R, N = P.N, n = N - 1;
  const int R = P.R, lgR = P.lgR;
  const int f = threadIdx.x >> P.lgtpf, t = threadIdx.x & (P.tpf - 1);
  T *region = lds + (size_t)2 * f * P.RS;
  T *scan_tmp = lds + (size_t)R * P.RS + 2 * f * (P.tpf > 64 ? P.tpf / 64 : 1);
  const CoefLayout CL{L, N};

  // psiC_m: reduce the column kernel's per-block partials in a fixed order
  // (every workgroup of the member does the same reduction), then the
  // correction constants (FP64).
  double kappa = 0.0;
  double beta[8];
  if (md.mode == 0) {
    double *red = reinterpret_cast<double *>(
        smem + ((sizeof(T) * ((size_t)R * P.RS + 2 * (blockDim.x >> P.lgtpf) *
                               (P.tpf > 64 ? P.tpf / 64 : 1) + 2) + 15) & ~size_t(15)));
    double *s_psic = red + 512;
    const int nt = blockDim.x, tid = threadIdx.x, nblk = N / P.C;
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
    if (w.z == 0 && l == 0 && tid == 0) {
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

  // Transposed tile load: row r <- S[l][p][j0 + r].
  const T *in = reinterpret_cast<const T *>(md.scr) + (size_t)l * N * N + j0;
  for (int idx = threadIdx.x; idx < N * R; idx += blockDim.x) {
    const int r = idx & (R - 1), p = idx >> lgR;
    lds[r * P.RS + pad16(p)] = in[(size_t)p * N + r];
  }
  __syncthreads();

  // Correction field for odd p, by recurrence over the tile's rows.
  if (md.mode == 0 && L > 1) {
    const double *tab = md.coef + CL.tab();
    for (int h = threadIdx.x; h < N / 2; h += blockDim.x) {
      const int p = 2 * h + 1;
      T acc[16];
      for (int r = 0; r < R; ++r) acc[r] = T(0);
      for (int m = 1; m < L; ++m) {
        const double *e = tab + ((size_t)(m - 1) * (N / 2) + h) * 3;
        const T rho = T(e[0]), gam = T(beta[m] * e[1]), eta = T(e[2]);
        const T rr = dev_exp<T>(rho);
        T hh[16];
        T u = dev_exp<T>(T(j0) * rho);  // r^(j0 + r), going up
        for (int r = 0; r < R; ++r) {
          hh[r] = u;
          u *= rr;
        }
        T v = dev_exp<T>(T(N - j0 - R + 1) * rho);  // r^(N - j0 - r), going down
        for (int r = R - 1; r >= 0; --r) {
          hh[r] += v;
          v *= rr;
        }
        for (int r = 0; r < R; ++r) acc[r] += gam * (T(1) - hh[r] * eta);
      }
      for (int r = 0; r < R; ++r) lds[r * P.RS + pad16(p)] -= acc[r];
    }
    __syncthreads();
  }

  T *rowa = lds + (size_t)(2 * f) * P.RS, *rowb = rowa + P.RS;
  auto lds_ld = [&](int k, T &xa, T &xb) {
    xa = rowa[pad16(k)];
    xb = rowb[pad16(k)];
  };
  dst_pair<T, E>(reinterpret_cast<cx<T> *>(region), rowa, rowb, P, t, scan_tmp,
                 lds_ld);
  __syncthreads();

  Tout *ps = reinterpret_cast<Tout *>(ptrs.psi[w.w]) + (size_t)l * ptrs.ldp1 * ptrs.ldp2;
  const long long ld1 = ptrs.ldp1;
  const int G = N + 1;
  const bool last = (j0 + R == N);
  const int nrow = R + (last ? 1 : 0);
  for (int idx = threadIdx.x; idx < nrow * G; idx += blockDim.x) {
    const int r = idx / G, i = idx - r * G, j = j0 + r;
    Tout *o = ps + (size_t)j * ld1 + i;
    if (md.mode == 0) {
      double val;
      if (j == 0 || j == N || i == 0 || i == N) val = -kappa;
      else val = double(lds[r * P.RS + pad16(i)]) - kappa;
      *o = Tout(val);
    } else {
      // helm's boundary semantics (qg3p_helmholtz.F90:757-761, 1055-1057)
      if (j == N || i == 0) continue;
      if (j == 0) {
        if (i <= n) *o = Tout(0);
      } else {
        *o = i == N ? Tout(0) : Tout(lds[r * P.RS + pad16(i)]);
      }
    }
  }
}

// ================================================================== host
struct PlanHost {
  int N = 0, E = 0, prec = 0;  // prec: 0 double compute, 1 float compute
  DevPlan dp{};
  void *tw = nullptr, *was = nullptr, *lam4 = nullptr;
  double *shat = nullptr;
  std::vector<double> shat_h, lam_h;  // lam_h[p] = -4 sin^2(p pi / 2N)
};

static int choose_E(int N) {
  if (N % 3 == 0) return 12;
  if (N >= 512) return 16;
  if (N >= 64) return 8;
  return 4;
}

static int ilog2(int x) {
  int r = 0;
  while ((1 << r) < x) ++r;
  return r;
}

template <typename T> static int upload_tables(PlanHost &ph) {
  const int N = ph.N;
  std::vector<T> tw(2 * N), was(N / 2), lam4(N);
  for (int k = 0; k < N; ++k) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * k / N;
    tw[2 * k] = T(cosl(a));
    tw[2 * k + 1] = T(sinl(a));
  }
  for (int k = 0; k < N / 2; ++k)
    was[k] = T(2.0L * sinl((long double)k * 3.141592653589793238462643383279502884L / N));
  const long double fourN2 = 4.0L * N * (long double)N;
  ph.lam_h.assign(N, 0.0);
  ph.shat_h.assign(N, 0.0);
  for (int p = 0; p < N; ++p) {
    long double s = sinl((long double)p * 3.141592653589793238462643383279502884L / (2.0L * N));
    long double lam = -4.0L * s * s;
    ph.lam_h[p] = (double)lam;
    lam4[p] = T(fourN2 * lam);
    if (p & 1) {
      long double x = (long double)p * 3.141592653589793238462643383279502884L / (2.0L * N);
      ph.shat_h[p] = (double)(2.0L * cosl(x) / sinl(x));
    }
  }
  PQ_HIP(hipMalloc(&ph.tw, sizeof(T) * 2 * N));
  PQ_HIP(hipMalloc(&ph.was, sizeof(T) * (N / 2)));
  PQ_HIP(hipMalloc(&ph.lam4, sizeof(T) * N));
  PQ_HIP(hipMalloc(&ph.shat, sizeof(double) * N));
  PQ_HIP(hipMemcpy(ph.tw, tw.data(), sizeof(T) * 2 * N, hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(ph.was, was.data(), sizeof(T) * (N / 2), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(ph.lam4, lam4.data(), sizeof(T) * N, hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(ph.shat, ph.shat_h.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  return PQ_OK;
}

static int make_plan(PlanHost &ph, int n, int compute_float, int R, int C) {
  const int N = n + 1;
  ph.N = N;
  ph.prec = compute_float;
  ph.E = choose_E(N);
  DevPlan &d = ph.dp;
  d.N = N;
  d.tpf = N / ph.E;
  d.lgtpf = ilog2(d.tpf);
  int rem = (N % 3 == 0) ? N / 3 : N;
  d.nst = 0;
  int lg = 0;
  while (rem > 1) {
    int r = 2;
    for (int cand : {16, 8, 4, 2})
      if (cand <= ph.E && ph.E % cand == 0 && rem % cand == 0) {
        r = cand;
        break;
      }
    d.rad[d.nst] = r;
    d.lns[d.nst] = lg;
    ++d.nst;
    lg += ilog2(r);
    rem /= r;
  }
  if (N % 3 == 0) {
    d.rad[d.nst] = 3;
    d.lns[d.nst] = lg;
    ++d.nst;
  }
  d.R = R;
  d.lgR = ilog2(R);
  d.C = C;
  d.RS = N + N / 16 + 1;
  int rc = compute_float ? upload_tables<float>(ph) : upload_tables<double>(ph);
  if (rc) return rc;
  d.tw = ph.tw;
  d.was = ph.was;
  d.lam4 = ph.lam4;
  d.shat = ph.shat;
  return PQ_OK;
}

static void free_plan(PlanHost &ph) {
  (void)hipFree(ph.tw);
  (void)hipFree(ph.was);
  (void)hipFree(ph.lam4);
  (void)hipFree(ph.shat);
  ph.tw = ph.was = ph.lam4 = nullptr;
  ph.shat = nullptr;
}

}  // namespace pq

using namespace pq;

// ------------------------------------------------------------------ handle
struct pq_inv {
  int g = 0, n = 0, N = 0, L = 0, prec = 0, batch_max = 0, mode = 0;
  int R = 0, C = 0;
  PlanHost plan;
  std::vector<double> coef_h;
  double *coef = nullptr;
  void *scr = nullptr;
  double *part = nullptr, *psic = nullptr;
  int *status = nullptr;
  MemDesc *mems = nullptr;
  int4 *work1 = nullptr, *work2 = nullptr, *work3 = nullptr;
  int nwork1 = 0, nwork2 = 0, nwork3 = 0;  // per member
  hipStream_t stream = nullptr;
  // staging for the host drop-in
  void *dz = nullptr, *dp = nullptr;
  size_t elem = 8;
};

struct pq_helm {
  pq_inv *inv = nullptr;
};

static int compute_coefs(pq_inv *h, const double *ss, const double *m2z,
                         const double *z2m) {
  const int L = h->L, N = h->N;
  CoefLayout CL{L, N};
  h->coef_h.assign(CL.size(), 0.0);
  double *c = h->coef_h.data();
  const double fourN2 = 4.0 * N * (double)N;
  for (int i = 0; i < L * L; ++i) {
    c[CL.D() + i] = z2m ? z2m[i] : (i % (L + 1) == 0 ? 1.0 : 0.0);
    c[CL.M() + i] = m2z ? m2z[i] : (i % (L + 1) == 0 ? 1.0 : 0.0);
  }
  const std::vector<double> &lam = h->plan.lam_h, &shat = h->plan.shat_h;
  for (int m = 0; m < L; ++m) {
    c[CL.ss() + m] = ss[m];
    c[CL.ssN() + m] = fourN2 * ss[m];
    // mean of c_m = helm_m(1): (1/N^2) sum_{p,q odd} shat_p^2 shat_q^2 /
    // ((2N)^2 (lambda_p + lambda_q - ss_m))
    long double s = 0.0L;
    for (int p = 1; p < N; p += 2)
      for (int q = 1; q < N; q += 2)
        s += (long double)shat[p] * shat[p] * shat[q] * shat[q] /
             ((long double)fourN2 * ((long double)lam[p] + lam[q] - ss[m]));
    const double meanc = (double)(s / ((long double)N * N));
    const double a = 1.0 / (1.0 + ss[m] * meanc);
    c[CL.a() + m] = a;
    c[CL.b() + m] = ss[m] * a;
  }
  for (int m = 1; m < L; ++m)
    for (int hh = 0; hh < N / 2; ++hh) {
      const int p = 2 * hh + 1;
      const double mu = lam[p] - ss[m];
      const double d = -mu;
      const double rho = -2.0 * asinh(std::sqrt(d) / 2.0);  // ln r
      const double rN = std::exp((double)N * rho);
      double *e = c + CL.tab() + ((size_t)(m - 1) * (N / 2) + hh) * 3;
      e[0] = rho;
      e[1] = shat[p] / (2.0 * N * mu);
      e[2] = 1.0 / (1.0 + rN);
    }
  return PQ_OK;
}

static int build_work(std::vector<int4> &v, int nmem, int L, int ntile,
                      bool per_layer) {
  v.clear();
  for (int b = 0; b < nmem; ++b)
    for (int l = 0; l < (per_layer ? L : 1); ++l)
      for (int t = 0; t < ntile; ++t) v.push_back(make_int4(b, per_layer ? l : t, per_layer ? t : 0, b));
  return PQ_OK;
}

static size_t lds_rows_bytes(const pq_inv *h, int rows) {
  const size_t el = h->prec ? 4 : 8;
  const int tpf = h->plan.dp.tpf;
  const int nfft = rows / 2;
  const size_t scan = (size_t)2 * nfft * (tpf > 64 ? tpf / 64 : 1);
  return el * ((size_t)rows * h->plan.dp.RS + scan + 2);
}

static size_t lds_cols_bytes(const pq_inv *h, int rows) {
  // rows + scan + 2, then the reduction buffer (512 doubles) and coefficient copies
  const size_t el = h->prec ? 4 : 8;
  return ((lds_rows_bytes(h, rows) + 15) & ~size_t(15)) + 512 * 8 + el * (64 + 64 + 8) + 16;
}

static int inv_create(pq_inv **out, int g, int L, const double *ss,
                      const double *m2z, const double *z2m, int precision,
                      int batch_max, int mode) {
  PQ_CHECK_ARG(out != nullptr, "pq_inv_create: null handle pointer");
  PQ_CHECK_ARG(g >= 9 && pq_supported_n(g - 2),
               "pq_inv_create: unsupported grid size g=%d (need g-1 = 2^a or 3*2^a, 8 <= g-1 <= 4096)", g);
  PQ_CHECK_ARG(L >= 1 && L <= 8, "pq_inv_create: layers must be in 1..8 (got %d)", L);
  PQ_CHECK_ARG(precision >= 0 && precision <= 2, "pq_inv_create: bad precision %d", precision);
  PQ_CHECK_ARG(batch_max >= 1, "pq_inv_create: batch_max must be >= 1");
  PQ_CHECK_ARG(ss != nullptr, "pq_inv_create: ss is null");
  if (mode == 0) {
    PQ_CHECK_ARG(m2z && z2m, "pq_inv_create: modes_to_z / z_to_modes are null");
    for (int l = 0; l < L; ++l)
      PQ_CHECK_ARG(m2z[l] == 1.0, "pq_inv_create: modes_to_z(:,1) must be the barotropic mode (all ones)");
    PQ_CHECK_ARG(ss[0] == 0.0, "pq_inv_create: ss(1) must be 0 (barotropic mode)");
  }
  for (int l = 0; l < L; ++l)
    PQ_CHECK_ARG(ss[l] >= 0.0 && std::isfinite(ss[l]), "pq_inv_create: ss must be finite and >= 0");
  pq_inv *h = new pq_inv();
  h->g = g;
  h->n = g - 2;
  h->N = g - 1;
  h->L = L;
  h->prec = precision;
  h->batch_max = batch_max;
  h->mode = mode;
  const bool fc = precision != PQ_PREC_FP64;
  // Tile shapes: 64-byte transposed segments (16 floats / 8 doubles); for
  // N >= 2048 fewer rows keep the workgroup within 512 threads.
  int R = fc ? 16 : 8;
  const int E = choose_E(h->N);
  while (R > 2 && (R / 2) * (h->N / E) > 512) R /= 2;
  while (R > 2 && (size_t)R * (h->N + h->N / 16 + 1) * (fc ? 4 : 8) > 96 * 1024) R /= 2;
  while (R > 2 && h->N % R != 0) R /= 2;  // tiles must cover N rows exactly
  int C = 2;
  while (C > 1 && ((L * C + 1) / 2) * (h->N / E) > 512) C /= 2;
  h->R = R;
  h->C = C;
  int rc = make_plan(h->plan, h->n, fc ? 1 : 0, R, C);
  if (rc) {
    delete h;
    return rc;
  }
  compute_coefs(h, ss, m2z, z2m);
  const size_t el = fc ? 4 : 8;
  const int N = h->N;
  const size_t scr_member = (size_t)L * N * N * el;
  auto fail = [&](int code) {
    pq_inv_destroy(h);
    return code;
  };
  if (hipMalloc(&h->coef, h->coef_h.size() * sizeof(double)) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (hipMemcpy(h->coef, h->coef_h.data(), h->coef_h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(PQ_ERR_HIP);
  if (hipMalloc(&h->scr, scr_member * batch_max) != hipSuccess) {
    pq_set_error("pq_inv_create: cannot allocate %zu bytes of scratch", scr_member * batch_max);
    return fail(PQ_ERR_NOMEM);
  }
  const int nblk2 = N / C;
  if (hipMalloc(&h->part, sizeof(double) * (size_t)nblk2 * L * batch_max) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (hipMalloc(&h->psic, sizeof(double) * (size_t)L * batch_max) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (hipMalloc(&h->status, sizeof(int) * batch_max) != hipSuccess) return fail(PQ_ERR_NOMEM);
  (void)hipMemset(h->status, 0, sizeof(int) * batch_max);
  (void)hipMemset(h->psic, 0, sizeof(double) * (size_t)L * batch_max);
  std::vector<MemDesc> md(batch_max);
  for (int b = 0; b < batch_max; ++b) {
    md[b].coef = h->coef;
    md[b].scr = (char *)h->scr + scr_member * b;
    md[b].part = h->part + (size_t)nblk2 * L * b;
    md[b].psic = h->psic + (size_t)L * b;
    md[b].status = h->status + b;
    md[b].L = L;
    md[b].mode = mode;
  }
  if (hipMalloc(&h->mems, sizeof(MemDesc) * batch_max) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (hipMemcpy(h->mems, md.data(), sizeof(MemDesc) * batch_max, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PQ_ERR_HIP);
  // Work lists: each entry is (member slot, layer|pblock, tile, pointer index);
  // built for a full chunk of kMaxBatchArg members (the pointer index w is the
  // member's position in the launch's pointer table).
  const int chunk = std::min(batch_max, kMaxBatchArg);
  std::vector<int4> w1, w2;
  h->nwork1 = L * (N / R);
  h->nwork2 = N / C;
  for (int b = 0; b < chunk; ++b)
    for (int l = 0; l < L; ++l)
      for (int t = 0; t < N / R; ++t) w1.push_back(make_int4(b, l, t, b));
  for (int b = 0; b < chunk; ++b)
    for (int pb = 0; pb < N / C; ++pb) w2.push_back(make_int4(b, pb, 0, b));
  if (hipMalloc(&h->work1, sizeof(int4) * w1.size()) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (hipMalloc(&h->work2, sizeof(int4) * w2.size()) != hipSuccess) return fail(PQ_ERR_NOMEM);
  (void)hipMemcpy(h->work1, w1.data(), sizeof(int4) * w1.size(), hipMemcpyHostToDevice);
  (void)hipMemcpy(h->work2, w2.data(), sizeof(int4) * w2.size(), hipMemcpyHostToDevice);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(PQ_ERR_HIP);
  *out = h;
  return PQ_OK;
}

// Dynamic LDS above the 64 KiB default needs an explicit per-kernel limit
// (gfx950 has 160 KiB of LDS per CU).  Any error of the attribute call is
// cleared here; a launch that still exceeds the limit reports its own error.
template <typename K> static void allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return;
  (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  (void)hipGetLastError();
}

template <typename Tin, typename T, int E>
static void launch_fwd(pq_inv *h, int nb, int mem0, const PtrTab &pt, hipStream_t s) {
  const DevPlan &P = h->plan.dp;
  const int threads = (h->R / 2) * P.tpf;
  const size_t lds = lds_rows_bytes(h, h->R);
  allow_lds(k_rows_fwd<Tin, T, E>, lds);
  hipLaunchKernelGGL((k_rows_fwd<Tin, T, E>), dim3(nb * h->nwork1), dim3(threads), lds, s, P,
                     h->mems + mem0, pt, h->work1);
}

template <typename T, int E>
static void launch_cols(pq_inv *h, int nb, int mem0, hipStream_t s) {
  const DevPlan &P = h->plan.dp;
  const int nfft = (h->L * h->C + 1) / 2;
  const int threads = nfft * P.tpf;
  const size_t lds = lds_cols_bytes(h, 2 * nfft);
  allow_lds(k_cols<T, E>, lds);
  hipLaunchKernelGGL((k_cols<T, E>), dim3(nb * h->nwork2), dim3(threads), lds, s, P, h->mems + mem0,
                     h->work2);
}

template <typename T, typename Tout, int E>
static void launch_inv(pq_inv *h, int nb, int mem0, const PtrTab &pt, hipStream_t s) {
  const DevPlan &P = h->plan.dp;
  const int threads = (h->R / 2) * P.tpf;
  // tile rows + scan scratch, then the psiC reduction buffer and psiC
  const size_t lds = ((lds_rows_bytes(h, h->R) + 15) & ~size_t(15)) + 512 * 8 + 8 * 8;
  allow_lds(k_rows_inv<T, Tout, E>, lds);
  hipLaunchKernelGGL((k_rows_inv<T, Tout, E>), dim3(nb * h->nwork1), dim3(threads), lds, s, P,
                     h->mems + mem0, pt, h->work1);
}

// pass_mask bit 0: row forward pass, bit 1: column pass, bit 2: row inverse.
template <int E>
static void launch_all_E(pq_inv *h, int nb, int mem0, const PtrTab &pt, hipStream_t s,
                         int pass_mask) {
  switch (h->prec) {
    case PQ_PREC_FP64:
      if (pass_mask & 1) launch_fwd<double, double, E>(h, nb, mem0, pt, s);
      if (pass_mask & 2) launch_cols<double, E>(h, nb, mem0, s);
      if (pass_mask & 4) launch_inv<double, double, E>(h, nb, mem0, pt, s);
      break;
    case PQ_PREC_FP32:
      if (pass_mask & 1) launch_fwd<float, float, E>(h, nb, mem0, pt, s);
      if (pass_mask & 2) launch_cols<float, E>(h, nb, mem0, s);
      if (pass_mask & 4) launch_inv<float, float, E>(h, nb, mem0, pt, s);
      break;
    default:
      if (pass_mask & 1) launch_fwd<double, float, E>(h, nb, mem0, pt, s);
      if (pass_mask & 2) launch_cols<float, E>(h, nb, mem0, s);
      if (pass_mask & 4) launch_inv<float, double, E>(h, nb, mem0, pt, s);
      break;
  }
}

static int inv_launch(pq_inv *h, int batch, const void *const *zeta, long ldz1, long ldz2,
                      void *const *psi, long ldp1, long ldp2, hipStream_t s,
                      int pass_mask = 7) {
  PQ_CHECK_ARG(batch >= 1 && batch <= h->batch_max, "pq_inv_run_dev: batch %d outside 1..%d", batch,
               h->batch_max);
  PQ_CHECK_ARG(ldz1 >= h->g && ldz2 >= h->g && ldp1 >= h->g && ldp2 >= h->g,
               "pq_inv_run_dev: leading dimensions smaller than g=%d", h->g);
  for (int b = 0; b < batch; ++b)
    PQ_CHECK_ARG(zeta[b] != nullptr && psi[b] != nullptr, "pq_inv_run_dev: null member pointer %d", b);
  const int E = h->plan.E;
  for (int b0 = 0; b0 < batch; b0 += kMaxBatchArg) {
    const int nb = std::min(kMaxBatchArg, batch - b0);
    PtrTab pt;
    std::memset(&pt, 0, sizeof(pt));
    for (int b = 0; b < nb; ++b) {
      pt.zeta[b] = zeta[b0 + b];
      pt.psi[b] = psi[b0 + b];
    }
    pt.ldz1 = ldz1;
    pt.ldz2 = ldz2;
    pt.ldp1 = ldp1;
    pt.ldp2 = ldp2;
    switch (E) {
      case 4: launch_all_E<4>(h, nb, b0, pt, s, pass_mask); break;
      case 8: launch_all_E<8>(h, nb, b0, pt, s, pass_mask); break;
      case 12: launch_all_E<12>(h, nb, b0, pt, s, pass_mask); break;
      default: launch_all_E<16>(h, nb, b0, pt, s, pass_mask); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      pq_set_error("pq_inv: kernel launch failed: %s", hipGetErrorString(e));
      return PQ_ERR_HIP;
    }
  }
  return PQ_OK;
}

extern "C" {

int pq_inv_create(pq_inv **h, int g, int layers, const double *ss, const double *modes_to_z,
                  const double *z_to_modes, int precision, int batch_max) {
  return inv_create(h, g, layers, ss, modes_to_z, z_to_modes, precision, batch_max, 0);
}

int pq_inv_destroy(pq_inv *h) {
  if (!h) return PQ_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_plan(h->plan);
  (void)hipFree(h->coef);
  (void)hipFree(h->scr);
  (void)hipFree(h->part);
  (void)hipFree(h->psic);
  (void)hipFree(h->status);
  (void)hipFree(h->mems);
  (void)hipFree(h->work1);
  (void)hipFree(h->work2);
  (void)hipFree(h->dz);
  (void)hipFree(h->dp);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PQ_OK;
}

void *pq_inv_stream(pq_inv *h) { return h ? (void *)h->stream : nullptr; }

int pq_inv_run_dev(pq_inv *h, int batch, const void *const *zeta, long ldz1, long ldz2,
                   void *const *psi, long ldp1, long ldp2, void *stream) {
  PQ_CHECK_ARG(h != nullptr, "pq_inv_run_dev: null handle");
  PQ_CHECK_ARG(zeta && psi, "pq_inv_run_dev: null pointer table");
  return inv_launch(h, batch, zeta, ldz1, ldz2, psi, ldp1, ldp2, pq_resolve_stream(stream, h->stream));
}

int pq_inv_status(pq_inv *h, double *psic) {
  PQ_CHECK_ARG(h != nullptr, "pq_inv_status: null handle");
  PQ_HIP(hipStreamSynchronize(h->stream));
  PQ_HIP(hipDeviceSynchronize());
  std::vector<int> st(h->batch_max);
  PQ_HIP(hipMemcpy(st.data(), h->status, sizeof(int) * h->batch_max, hipMemcpyDeviceToHost));
  if (psic)
    PQ_HIP(hipMemcpy(psic, h->psic, sizeof(double) * h->L * h->batch_max, hipMemcpyDeviceToHost));
  int worst = 0;
  for (int v : st) worst |= v;
  PQ_HIP(hipMemset(h->status, 0, sizeof(int) * h->batch_max));
  if (worst & 1) {
    pq_set_error("NaN psiC encountered");
    return PQ_ERR_NAN;
  }
  if (worst & 2) {
    pq_set_error("WARNING: Large psiC encountered; simulation may be blowing up");
    return PQ_WARN_LARGE;
  }
  return PQ_OK;
}

// Host-view copy helpers: g x g x L block of a (ld1, ld2) host view <-> a
// packed device array.
static int copy_view_h2d(void *dev, const double *host, int g, int L, long ld1, long ld2,
                         size_t el, hipStream_t s) {
  if (el == 8) {
    for (int l = 0; l < L; ++l)
      PQ_HIP(hipMemcpy2DAsync((char *)dev + (size_t)l * g * g * 8, (size_t)g * 8,
                              host + (size_t)l * ld1 * ld2, (size_t)ld1 * 8, (size_t)g * 8, g,
                              hipMemcpyHostToDevice, s));
    return PQ_OK;
  }
  std::vector<float> tmp((size_t)g * g * L);
  for (int l = 0; l < L; ++l)
    for (int j = 0; j < g; ++j)
      for (int i = 0; i < g; ++i)
        tmp[((size_t)l * g + j) * g + i] = (float)host[i + ld1 * (j + ld2 * (size_t)l)];
  PQ_HIP(hipMemcpyAsync(dev, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice, s));
  PQ_HIP(hipStreamSynchronize(s));
  return PQ_OK;
}

static int copy_view_d2h(double *host, const void *dev, int g, int L, long ld1, long ld2,
                         size_t el, hipStream_t s) {
  if (el == 8) {
    for (int l = 0; l < L; ++l)
      PQ_HIP(hipMemcpy2DAsync(host + (size_t)l * ld1 * ld2, (size_t)ld1 * 8,
                              (const char *)dev + (size_t)l * g * g * 8, (size_t)g * 8,
                              (size_t)g * 8, g, hipMemcpyDeviceToHost, s));
    PQ_HIP(hipStreamSynchronize(s));
    return PQ_OK;
  }
  std::vector<float> tmp((size_t)g * g * L);
  PQ_HIP(hipMemcpyAsync(tmp.data(), dev, tmp.size() * 4, hipMemcpyDeviceToHost, s));
  PQ_HIP(hipStreamSynchronize(s));
  for (int l = 0; l < L; ++l)
    for (int j = 0; j < g; ++j)
      for (int i = 0; i < g; ++i)
        host[i + ld1 * (j + ld2 * (size_t)l)] = tmp[((size_t)l * g + j) * g + i];
  return PQ_OK;
}

static int run_host(pq_inv *h, const double *zeta, int ldz1, int ldz2, double *psi, int ldp1,
                    int ldp2) {
  PQ_CHECK_ARG(zeta && psi, "pq_inv_run: null array");
  PQ_CHECK_ARG(ldz1 >= h->g && ldz2 >= h->g && ldp1 >= h->g && ldp2 >= h->g,
               "pq_inv_run: leading dimensions smaller than g=%d", h->g);
  const int g = h->g, L = h->L;
  const size_t el = h->prec == PQ_PREC_FP32 ? 4 : 8;
  void *mz = pq_mirror_device_ptr(zeta);
  void *mp = pq_mirror_device_ptr(psi);
  if (mz && mp && el == 8) {
    const void *zt[1] = {mz};
    void *pt[1] = {mp};
    int rc = inv_launch(h, 1, zt, ldz1, ldz2, pt, ldp1, ldp2, h->stream);
    if (rc) return rc;
    return pq_inv_status(h, nullptr);
  }
  const size_t bytes = (size_t)g * g * L * el;
  if (!h->dz) {
    PQ_HIP(hipMalloc(&h->dz, bytes));
    PQ_HIP(hipMalloc(&h->dp, bytes));
  }
  int rc = copy_view_h2d(h->dz, zeta, g, L, ldz1, ldz2, el, h->stream);
  if (rc) return rc;
  if (h->mode == 1) {
    // helm works in place: the output keeps input values on part of the
    // boundary, so start psi from zeta.
    PQ_HIP(hipMemcpyAsync(h->dp, h->dz, bytes, hipMemcpyDeviceToDevice, h->stream));
  }
  const void *zt[1] = {h->dz};
  void *pt[1] = {h->dp};
  rc = inv_launch(h, 1, zt, g, g, pt, g, g, h->stream);
  if (rc) return rc;
  rc = copy_view_d2h(psi, h->dp, g, L, ldp1, ldp2, el, h->stream);
  if (rc) return rc;
  return pq_inv_status(h, nullptr);
}

int pq_inv_run(pq_inv *h, const double *zeta, int ldz1, int ldz2, double *psi, int ldp1,
               int ldp2) {
  PQ_CHECK_ARG(h != nullptr, "pq_inv_run: null handle");
  return run_host(h, zeta, ldz1, ldz2, psi, ldp1, ldp2);
}

int pq_inv_pass_times(pq_inv *h, int batch, const void *const *zeta, long ldz1, long ldz2,
                      void *const *psi, long ldp1, long ldp2, int reps, float *ms3) {
  PQ_CHECK_ARG(h && zeta && psi && ms3 && reps >= 1, "pq_inv_pass_times: bad arguments");
  hipEvent_t ev[2];
  PQ_HIP(hipEventCreate(&ev[0]));
  PQ_HIP(hipEventCreate(&ev[1]));
  // full passes once so every pass reads valid intermediate data
  int rc = inv_launch(h, batch, zeta, ldz1, ldz2, psi, ldp1, ldp2, h->stream);
  for (int k = 0; k < 3 && rc == PQ_OK; ++k) {
    PQ_HIP(hipEventRecord(ev[0], h->stream));
    for (int r = 0; r < reps && rc == PQ_OK; ++r)
      rc = inv_launch(h, batch, zeta, ldz1, ldz2, psi, ldp1, ldp2, h->stream, 1 << k);
    PQ_HIP(hipEventRecord(ev[1], h->stream));
    PQ_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    PQ_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    ms3[k] = ms / reps;
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  return pq_inv_status(h, nullptr);
}

// ---- helm ---------------------------------------------------------------
int pq_helm_create(pq_helm **out, int ii, int jj, int layers, const double *ss, int precision) {
  PQ_CHECK_ARG(out != nullptr, "pq_helm_create: null handle pointer");
  PQ_CHECK_ARG(ii == jj, "pq_helm_create: only square closed basins are supported (ii=%d, jj=%d)", ii, jj);
  pq_inv *inv = nullptr;
  int rc = inv_create(&inv, ii, layers, ss, nullptr, nullptr, precision, 1, 1);
  if (rc) return rc;
  pq_helm *h = new pq_helm();
  h->inv = inv;
  *out = h;
  return PQ_OK;
}

int pq_helm_destroy(pq_helm *h) {
  if (!h) return PQ_OK;
  pq_inv_destroy(h->inv);
  delete h;
  return PQ_OK;
}

int pq_helm_solve(pq_helm *h, double *zeta, int ld1, int ld2) {
  PQ_CHECK_ARG(h != nullptr, "pq_helm_solve: null handle");
  return run_host(h->inv, zeta, ld1, ld2, zeta, ld1, ld2);
}

int pq_helm_solve_dev(pq_helm *h, void *zeta, long ld1, long ld2, void *stream) {
  PQ_CHECK_ARG(h != nullptr, "pq_helm_solve_dev: null handle");
  const void *zt[1] = {zeta};
  void *pt[1] = {zeta};
  return inv_launch(h->inv, 1, zt, ld1, ld2, pt, ld1, ld2,
                    pq_resolve_stream(stream, h->inv->stream));
}

}  // extern "C"
