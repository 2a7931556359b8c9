// pq_inversion.hip -- host side of the modal Helmholtz / PV inversion:
// handles, tables, tile choice, launches and the C ABI.
//
// Restates the reference's pv_inversion (qg3p_aux.F90:452-521) and helm
// (qg3p_helmholtz.F90:614-769) as three bandwidth-bound passes over HBM
// (kernels in pq_inv_kernels.hip, one unit per transform length):
//   row forward DST (i)  ->  column pass (DST j, modal projection, divide,
//   psiC partials, back-projection, inverse DST j)  ->  row inverse pass
//   (psiC, mass correction, inverse DST i, boundary).
// Scratch layout (per member, compute type T, N = n + 1): layer l is N x N
// elements stored as N / R blocks of N x R, block jb holding rows
// j = jb R .. jb R + R - 1 transposed (element (p, j) at
// l N^2 + jb N R + p R + j % R).  The row passes read and write whole blocks
// contiguously; the column pass moves C R contiguous elements per block.
#include "pq_internal.h"
#include "pq_inv.h"
#include "pq_mfma.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace pq {

typedef EngineInfo (*InfoFn)(int);
typedef hipError_t (*LaunchFn)(int, int, const InvArgs &, const PtrTab &, int, int, size_t,
                               hipStream_t);
struct KernelSet {
  InfoFn info = nullptr;
  LaunchFn launch = nullptr;
};

static bool kernels_for(int N, KernelSet &k) {
  switch (N) {
#define PQ_CASE(n)                                                             \
  case n:                                                                      \
    k.info = inv_info_##n;                                                     \
    k.launch = inv_launch_##n;                                                 \
    return true;
    PQ_INV_SIZES(PQ_CASE)
#undef PQ_CASE
    default:
      return false;
  }
}

static int ilog2i(int x) {
  int r = 0;
  while ((1 << r) < x) ++r;
  return r;
}

// psiC of each member from the column pass's per-workgroup partials, summed
// in a fixed order (FP64), then the constants of the row-inverse pass:
// kappa_l = psiC_1 + sum_{m>=2} M(l,m) psiC_m a_m and
// beta_lm = M(l,m) psiC_m b_m (qg3p_diagnostics.F90:615-635 rewritten for
// the closed-form correction).  Status bits: 1 NaN psiC, 2 |psiC| > 1e100
// (qg3p_aux.F90:504-517).
//
// Far from both walls the correction of layer l at odd p = 2h + 1 is the
// constant Gamma_l(h) = sum_{m>=2} beta_lm gamma_m(h) (the r^j terms have
// decayed); with rmin(h), the slowest decay rate -rho over those modes,
// the inverse pass uses it for blocks whose rows all lie where they have.
// Layout of kb after beta: rmin[N/2], then Gamma_l[N/2] for l = 0..L-1.
//
// Grid: k_psic_groups(N) workgroups of 256 threads per member, one column h
// per thread for Gamma and rmin; every workgroup forms psiC from the
// partials (a few KB), the first one publishes psiC, kappa, beta and the
// status.
__global__ void __launch_bounds__(256) k_psic(InvArgs a, PtrTab ptrs, int N) {
  __shared__ double red[4 * 8];
  __shared__ double s_psic[8];
  __shared__ double s_beta[8 * 8];
  const int ng = k_psic_groups(N);
  const int b = blockIdx.x / ng, grp = blockIdx.x % ng, L = ptrs.lay[b], tid = threadIdx.x;
  const int h = grp * 256 + tid;
  const bool lead = grp == 0;
  const int nblk = N >> ptrs.lgc[b];
  const double *part = a.part + (size_t)b * a.part_stride;
  // the mass correction's mode constants (rho, gamma) of this thread's
  // column, issued first: their latency overlaps the reduction (the
  // indices clamped to the table; the surplus unused)
  const double *tab = ptrs.coef[b] + CoefLayout{L, N}.tab();
  double2 e[7];
  if (L >= 2) {
#pragma unroll
    for (int q = 0; q < 7; ++q)
      e[q] = *reinterpret_cast<const double2 *>(tab + ((size_t)min(q, L - 2) * (N / 2) + min(h, N / 2 - 1)) * 4);
  }
  // the modal coefficients the first L threads combine with psiC (M(l, m),
  // a_m, b_m), also issued before the reduction
  const CoefLayout CL{L, N};
  const double *coef = ptrs.coef[b];
  double cm[8], ca[8], cbv[8];
  if (tid < L) {
#pragma unroll
    for (int m = 1; m < 8; ++m)
      if (m < L) {
        cm[m] = coef[CL.M() + tid + L * m];
        ca[m] = coef[CL.a() + m];
        cbv[m] = coef[CL.b() + m];
      }
  }
  double acc[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) acc[m] = 0.0;
  for (int q = tid; q < nblk; q += 256)
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (m < L) acc[m] += part[(size_t)q * L + m];
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int m = 0; m < 8; ++m)
    if (m < L) {
      double x = acc[m];
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) x += __shfl_down(x, d, 64);
      if (lane == 0) red[w * 8 + m] = x;
    }
  __syncthreads();
  if (tid == 0) {
    const double inv_n2 = 1.0 / (double(N) * double(N));
    int st = 0;
    for (int m = 0; m < L; ++m) {
      const double v = (red[m] + red[8 + m] + red[16 + m] + red[24 + m]) * inv_n2;
      s_psic[m] = v;
      if (lead) a.psic[(size_t)b * a.psic_stride + m] = v;
      if (v != v) st |= 1;
      else if (fabs(v) > 1.0e100) st |= 2;
    }
    if (st && lead) atomicOr(a.status + b, st);
  }
  __syncthreads();
  double *kb = a.kb + (size_t)b * a.kb_stride;
  if (tid < L) {
    const int l = tid;
    double kap = s_psic[0];
#pragma unroll
    for (int m = 1; m < 8; ++m) {
      if (m >= L) break;
      kap += cm[m] * s_psic[m] * ca[m];
      const double be = cm[m] * s_psic[m] * cbv[m];
      if (lead) kb[L + l * L + m] = be;
      s_beta[l * 8 + m] = be;
    }
    if (lead) kb[l] = kap;
  }
  if (L < 2) return;  // uniform in the workgroup
  __syncthreads();    // s_beta
  if (h >= N / 2) return;
  double *rmin = kb + L + L * L, *gam = rmin + N / 2;
  double rm = INFINITY;
  double g[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) g[l] = 0.0;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int m = q + 1;
    if (m >= L) break;
    rm = fmin(rm, -e[q].x);
#pragma unroll
    for (int l = 0; l < 8; ++l)
      if (l < L) g[l] += s_beta[l * 8 + m] * e[q].y;
  }
  rmin[h] = rm;
#pragma unroll
  for (int l = 0; l < 8; ++l)
    if (l < L) gam[(size_t)l * (N / 2) + h] = g[l];
}

// The channel's momentum-conserving correction (qg3p_aux.F90:524-575) per
// member, from the column pass's boundary-sum partials (FP64, fixed order):
// for every baroclinic mode m the mean boundary relative PV of the modal
// solution on the walls i = 1 and i = G, (z_b, z_d) = sum_p phi_m(p, q=0)
// (w_p, (-1)^(p+1) w_p), gives psiC_m = inv_corr_a_m (z_b, z_d).  The
// correction functions are uniform along j, c1_m(i) = 1 on i = 1, 0 on
// i = G, c2_m mirrored, with the i-spectral coefficients (in the inverse
// row pass's units) -sin(pi p / N) / (N (lambda_p - ss_m)) and
// (-1)^(p+1) times it.  So psi_l -= E_l with
//   E_l(p) = sum_{m>=2} M(l,m) (psiC_m1 + (-1)^(p+1) psiC_m2) c_m(p),
// E_l(0) = sum M(l,m) psiC_m1, E_l(N) = sum M(l,m) psiC_m2 on the walls.
// Status bits as k_psic.  Grid: one 256-thread workgroup per member.
__global__ void __launch_bounds__(256) k_chan_corr(InvArgs a, PtrTab ptrs, int N) {
  __shared__ double red[4 * 16];
  __shared__ double s_pc[16];
  const int b = blockIdx.x, L = ptrs.lay[b], tid = threadIdx.x;
  const int nblk = N >> ptrs.lgc[b];
  const double *part = a.part + (size_t)b * a.part_stride;
  double acc[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) acc[m] = 0.0;
  for (int q = tid; q < nblk; q += 256)
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (m < L) {
        acc[m] += part[(size_t)q * 2 * L + m];
        acc[8 + m] += part[(size_t)q * 2 * L + L + m];
      }
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int m = 0; m < 16; ++m)
    if ((m & 7) < L) {
      double x = acc[m];
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) x += __shfl_down(x, d, 64);
      if (lane == 0) red[w * 16 + m] = x;
    }
  __syncthreads();
  const CoefLayout CL{L, N};
  const double *coef = ptrs.coef[b];
  if (tid == 0) {
    int st = 0;
    s_pc[0] = s_pc[8] = 0.0;
    for (int m = 1; m < L; ++m) {
      const double zb = red[m] + red[16 + m] + red[32 + m] + red[48 + m];
      const double zd = red[8 + m] + red[24 + m] + red[40 + m] + red[56 + m];
      const double *ia = coef + CL.tab() + 4 * m;
      const double p1 = ia[0] * zb + ia[2] * zd, p2 = ia[1] * zb + ia[3] * zd;
      s_pc[m] = p1;
      s_pc[8 + m] = p2;
      a.psic[(size_t)b * a.psic_stride + 2 * m] = p1;
      a.psic[(size_t)b * a.psic_stride + 2 * m + 1] = p2;
      if (p1 != p1 || p2 != p2) st |= 1;
      else if (fabs(p1) > 1.0e100 || fabs(p2) > 1.0e100) st |= 2;
    }
    if (st) atomicOr(a.status + b, st);
  }
  __syncthreads();
  const double *M = coef + CL.M(), *ss = coef + CL.ss();
  const double *sinN = a.ctab + N, *lam = a.ctab + 2 * N;
  double *E = a.kb + (size_t)b * a.kb_stride;
  for (int idx = tid; idx < L * (N + 1); idx += 256) {
    const int l = idx / (N + 1), p = idx - l * (N + 1);
    double v = 0.0;
    if (p == 0 || p == N) {
      for (int m = 1; m < L; ++m) v += M[l + L * m] * s_pc[(p == 0 ? 0 : 8) + m];
    } else {
      const double sgn = (p & 1) ? 1.0 : -1.0;
      for (int m = 1; m < L; ++m)
        v += M[l + L * m] * (s_pc[m] + sgn * s_pc[8 + m]) * (-sinN[p] / (lam[p] - ss[m]));
    }
    E[idx] = v;
  }
}

constexpr size_t kLdsMax = 160 * 1024;
constexpr int kThreadsMax = 512;

}  // namespace pq

using namespace pq;

// ------------------------------------------------------------------ handle
struct pq_inv {
  int g = 0, N = 0, L = 0, prec = 0, batch_max = 0, mode = 0;
  // periodic channel: gj = 2N nodes along j (else gj = g), walls' boundary
  // condition of the momentum constraint
  int per = 0, gj = 0, bc_type = 1;
  double alpha = 0.0;
  double *ctab = nullptr;
  int twm = 0;  // twiddle table length
  int R = 0, C = 0;
  std::vector<int> lay;  // layers of each member slot (<= L)
  KernelSet ks;
  EngineInfo ei{};
  // tables
  void *tw64 = nullptr, *was = nullptr, *lam4 = nullptr;
  unsigned *mtab = nullptr;  // PQ_PREC_FP32_MFMA: per-lane constants of the row engine
  int ncu = 0;               // compute units of the handle's device (persistent row grids)
  double *shat = nullptr;
  std::vector<double> lam_h, shat_h;  // lam_h[p] = -4 sin^2(p pi / 2N)
  size_t coef_size = 0;  // doubles per member
  double *coef = nullptr;
  void *scr = nullptr;
  double *part = nullptr, *psic = nullptr, *kb = nullptr;
  int *status = nullptr;
  size_t scr_member = 0;  // scratch bytes per member
  int part_cap = 0;       // psiC partial slots per member (N / C_min)
  hipStream_t stream = nullptr;
  // staging for the host drop-in
  void *dz = nullptr, *dp = nullptr;
};

struct pq_helm {
  pq_inv *inv = nullptr;
};

template <typename T> static int upload_tables(pq_inv *h, int M) {
  const int N = h->N;
  const long double pi = 3.141592653589793238462643383279502884L;
  // the channel's column engine needs the table of length 2N
  if (h->per && M < 2 * N) M = 2 * N;
  h->twm = M;
  std::vector<double> tw(2 * (size_t)M);
  std::vector<T> was(N / 2), lam4(N + 1);
  for (int k = 0; k < M; ++k) {
    const long double a = -2.0L * pi * k / M;
    tw[2 * k] = (double)cosl(a);
    tw[2 * k + 1] = (double)sinl(a);
  }
  for (int k = 0; k < N / 2; ++k) was[k] = T(2.0L * sinl((long double)k * pi / N));
  const long double fourN2 = 4.0L * N * (long double)N;
  h->lam_h.assign(N, 0.0);
  h->shat_h.assign(N, 0.0);
  for (int p = 0; p <= N; ++p) {
    const long double x = (long double)p * pi / (2.0L * N);
    const long double s = sinl(x);
    const long double lam = -4.0L * s * s;
    lam4[p] = T(fourN2 * lam);  // p = N: the channel's q = N
    if (p == N) break;
    h->lam_h[p] = (double)lam;
    if (p & 1) h->shat_h[p] = (double)(2.0L * cosl(x) / s);
  }
  PQ_HIP(hipMalloc(&h->tw64, sizeof(double) * 2 * M));
  PQ_HIP(hipMalloc(&h->was, sizeof(T) * (N / 2)));
  PQ_HIP(hipMalloc(&h->lam4, sizeof(T) * (N + 1)));
  PQ_HIP(hipMalloc(&h->shat, sizeof(double) * N));
  PQ_HIP(hipMemcpy(h->tw64, tw.data(), sizeof(double) * 2 * M, hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(h->was, was.data(), sizeof(T) * (N / 2), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(h->lam4, lam4.data(), sizeof(T) * (N + 1), hipMemcpyHostToDevice));
  PQ_HIP(hipMemcpy(h->shat, h->shat_h.data(), sizeof(double) * N, hipMemcpyHostToDevice));
  if (h->per) {
    // boundary weights w_p of the mean boundary relative PV of a modal
    // field from its (p, q = 0) coefficients (zeta_boundary,
    // qg3p_aux.F90:288-379: the wall values vanish, the rows i = 2, 3 enter
    // through the inverse sine transform X_k = 2 sum_p x_p sin(pi p k / N)),
    // then sin(pi p / N) / N and lambda_p
    std::vector<double> ct(3 * (size_t)N, 0.0);
    const long double cff = h->bc_type == 1 ? (long double)h->alpha / (h->alpha + 3.0L) : 0.0L;
    for (int p = 1; p < N; ++p) {
      const long double s1 = sinl(pi * p / N), s2 = sinl(2.0L * pi * p / N);
      long double w = 0.0L;
      if (h->bc_type == 1) w = cff * 2.0L * (4.0L * s1 - 0.5L * s2);
      else if (h->bc_type == 2) w = 2.0L * 2.0L * s1;
      ct[p] = (double)w;
      ct[N + p] = (double)(s1 / N);
      ct[2 * N + p] = h->lam_h[p];
    }
    PQ_HIP(hipMalloc(&h->ctab, sizeof(double) * 3 * N));
    PQ_HIP(hipMemcpy(h->ctab, ct.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice));
  }
  return PQ_OK;
}

// Values at the nodes k = 0..N of the channel's correction function c1 of
// a mode with Helmholtz constant ss > 0: c1(k-1) + c1(k+1) - (2 + ss) c1(k)
// = 0, c1(0) = 1, c1(N) = 0, i.e. sinh(kappa (N - k)) / sinh(kappa N) with
// cosh kappa = 1 + ss / 2, in a form without overflow for large kappa N
// (c2(k) = c1(N - k)).
static double chan_c1(int N, double ss, int k) {
  const double kap = std::log1p(ss / 2.0 + std::sqrt(ss * (1.0 + ss / 4.0)));
  return std::exp(-kap * k) * std::expm1(-2.0 * kap * (N - k)) / std::expm1(-2.0 * kap * N);
}

// Coefficient block of one member with L layers (CoefLayout) into
// c[0 .. CL.size()).
static void compute_coefs(const pq_inv *h, int L, const double *ss, const double *m2z,
                          const double *z2m, double *c) {
  const int N = h->N;
  CoefLayout CL{L, N};
  std::fill(c, c + CL.size(), 0.0);
  const double fourN2 = 4.0 * N * (double)N;
  for (int i = 0; i < L * L; ++i) {
    c[CL.D() + i] = z2m ? z2m[i] : (i % (L + 1) == 0 ? 1.0 : 0.0);
    c[CL.M() + i] = m2z ? m2z[i] : (i % (L + 1) == 0 ? 1.0 : 0.0);
  }
  const std::vector<double> &lam = h->lam_h, &shat = h->shat_h;
  for (int m = 0; m < L; ++m) {
    c[CL.ss() + m] = ss[m];
    c[CL.ssN() + m] = fourN2 * ss[m];
    if (h->per) continue;  // the mass correction's constants: closed basin only
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
  if (h->per) {
    // inv_corr_a per baroclinic mode (correction_function_momentum,
    // qg3p_aux.F90:233-286): the inverse of zeta_boundary of the two
    // correction functions on the two walls (rows: walls i = 1, i = G;
    // columns: c1, c2 = c1 mirrored), column-major at tab + 4 m
    const double cff = h->bc_type == 1 ? h->alpha / (h->alpha + 3.0) : 0.0;
    for (int m = 1; m < L; ++m) {
      const double sm = ss[m];
      double v1[3], v2[3];  // c1 at the wall k = 0 and k = 1, 2; c1 at k = N, N-1, N-2
      for (int k = 0; k < 3; ++k) {
        v1[k] = chan_c1(N, sm, k);
        v2[k] = chan_c1(N, sm, N - k);
      }
      auto zw = [&](const double *v) {  // wall value, then the two rows inside
        double z = 0.0;
        if (h->bc_type == 1) z = cff * (-3.5 * v[0] + 4.0 * v[1] - 0.5 * v[2]);
        else if (h->bc_type == 2) z = 2.0 * (v[1] - v[0]);
        return z - sm * v[0];
      };
      // c1 seen from wall i = 1 is v1, from wall i = G is v2; c2 the reverse
      const double a11 = zw(v1), a21 = zw(v2), a12 = zw(v2), a22 = zw(v1);
      const double det = a11 * a22 - a12 * a21;
      double *ia = c + CL.tab() + 4 * m;
      ia[0] = a22 / det;
      ia[1] = -a21 / det;
      ia[2] = -a12 / det;
      ia[3] = a11 / det;
    }
  }
  // constants of the tridiagonal column pass (mode m, column p): r with
  // r + 1/r = 2 + mu, mu = 4 sin^2(p pi / 2N) + ss_m, from ln r =
  // -2 asinh(sqrt(mu) / 2); 1/r; 1 / (r - 1/r) = 1 / (2 sinh(ln r)); r^N;
  // 1 / (1 - r^2N); r^E for the E of the handle's launches.  The column
  // p = 0 is zero (r = 1/2 keeps its arithmetic finite).
  const int E = tri_e(N, tri_lt(h->L));
  for (int m = 0; m < L && !h->per; ++m)
    for (int p = 0; p < N; ++p) {
      const double mu = -lam[p] + ss[m];
      const double lr = (p == 0 || !(mu > 0.0)) ? std::log(0.5) : -2.0 * std::asinh(std::sqrt(mu) / 2.0);
      double *t = c + CL.tri() + ((size_t)m * N + p) * 8;
      const double rN = std::exp((double)N * lr);
      t[0] = std::exp(lr);
      t[1] = std::exp(-lr);
      t[2] = 1.0 / (2.0 * std::sinh(lr));
      t[3] = rN;
      t[4] = 1.0 / (1.0 - rN * rN);
      t[5] = std::exp((double)E * lr);
    }
  for (int m = 1; m < L && !h->per; ++m)
    for (int hh = 0; hh < N / 2; ++hh) {
      const int p = 2 * hh + 1;
      const double mu = lam[p] - ss[m];
      const double rho = -2.0 * asinh(std::sqrt(-mu) / 2.0);  // ln r
      const double rN = std::exp((double)N * rho);
      double *e = c + CL.tab() + ((size_t)(m - 1) * (N / 2) + hh) * 4;
      e[0] = rho;
      e[1] = shat[p] / (2.0 * N * mu);
      e[2] = 1.0 / (1.0 + rN);
      e[3] = std::exp(rho);
    }
}

// Whether rows-per-tile R and columns-per-workgroup C fit the kernels.

// The matrix-core row engine's per-lane constants (pq_mfma.h): word w of
// lane l at w * 64 + l, fp16 pairs (even slot in the low half) split as
// hi = fp16(v), lo = fp16(v - hi) from the long double value, and FP32
// twiddles.
namespace pq {
namespace mfma {
void build_table(unsigned *tab) {
  const long double pi = 3.141592653589793238462643383279502884L;
  auto put_half = [&](int word, int lane, int half, _Float16 v) {
    unsigned short bits;
    std::memcpy(&bits, &v, sizeof(bits));
    unsigned &w = tab[word * 64 + lane];
    w = half ? ((w & 0xffffu) | ((unsigned)bits << 16)) : ((w & 0xffff0000u) | bits);
  };
  auto put_split = [&](int whi, int wlo, int lane, int slot, long double v) {
    const _Float16 hi = (_Float16)(double)v;
    const _Float16 lo = (_Float16)(double)(v - (long double)(double)hi);
    put_half(whi + slot / 2, lane, slot & 1, hi);
    put_half(wlo + slot / 2, lane, slot & 1, lo);
  };
  auto put_f32 = [&](int word, int lane, long double v) {
    const float f = (float)v;
    std::memcpy(&tab[word * 64 + lane], &f, sizeof(f));
  };
  for (int i = 0; i < kWords * 64; ++i) tab[i] = 0;
  for (int lane = 0; lane < 64; ++lane) {
    const int c = lane & 31, g = lane >> 5;
    for (int s = 0; s < 8; ++s) {
      // stage 1: B operand, k = i2 = 16 t + 8 g + s, column c
      for (int t = 0; t < 2; ++t) {
        const int i2 = 16 * t + 8 * g + s;
        const long double re = cosl(pi * c * i2 / 32.0L);
        const long double im = c ? sinl(pi * c * i2 / 32.0L) : ((i2 & 1) ? -1.0L : 1.0L);
        put_split(kB1h + (2 * t + 0) * 4, kB1l + (2 * t + 0) * 4, lane, s, re);
        put_split(kB1h + (2 * t + 1) * 4, kB1l + (2 * t + 1) * 4, lane, s, im);
      }
      // stage 2 and the p = 32 m rows: A operand, row m = c, k-step t,
      // slot s <-> i1 = 8 t + 4 g + s % 4, part s / 4 (0: Re Z, 1: Im Z)
      for (int t = 0; t < 4; ++t) {
        const int i1 = 8 * t + 4 * g + (s & 3), part = s >> 2, m = c;
        long double v2;
        if (m < 16) {
          const long double a = 2.0L * pi * m * i1 / 32.0L;
          v2 = part ? cosl(a) : sinl(a);
        } else {
          const long double a = 2.0L * pi * (m - 15) * i1 / 32.0L;
          v2 = part ? -cosl(a) : sinl(a);
        }
        put_split(kA2h + 4 * t, kA2l + 4 * t, lane, s, v2);
        const long double v3 = (part == (m & 1)) ? sinl(pi * m * i1 / 32.0L) : 0.0L;
        put_split(kA3h + 4 * t, kA3l + 4 * t, lane, s, v3);
      }
    }
    // twiddles W^(c (4 g + b)) and W^(8 a c), W = exp(i pi / 1024)
    for (int b = 0; b < 4; ++b) {
      const long double a = pi * c * (4 * g + b) / (long double)kN;
      put_f32(kBre + b, lane, cosl(a));
      put_f32(kBim + b, lane, sinl(a));
    }
    for (int a = 1; a < 4; ++a) {
      const long double x = pi * 8.0L * a * c / (long double)kN;
      put_f32(kSre + a - 1, lane, cosl(x));
      put_f32(kSim + a - 1, lane, sinl(x));
    }
  }
}
}  // namespace mfma
}  // namespace pq

static int upload_mfma_table(pq_inv *h) {
  std::vector<unsigned> tab(mfma::kWords * 64);
  mfma::build_table(tab.data());
  PQ_HIP(hipMalloc(&h->mtab, sizeof(unsigned) * tab.size()));
  PQ_HIP(hipMemcpy(h->mtab, tab.data(), sizeof(unsigned) * tab.size(), hipMemcpyHostToDevice));
  return PQ_OK;
}

// Row-pass workgroup of a handle: threads and LDS bytes of pass 0 or 2.
static int row_threads(const pq_inv *h, int R) {
  return h->prec == PQ_PREC_FP32_MFMA ? mfma::kRowThreads : (R / 2) * h->ei.tpf;
}
static size_t row_lds(const pq_inv *h, int pass, int R) {
  if (h->prec == PQ_PREC_FP32_MFMA)  // the A3 copy, R rows of N + 4 floats
    return sizeof(float) * ((size_t)mfma::kA3Lds + (size_t)(h->N + 4) * R);
  return inv_lds_bytes(h->ei, pass, R / 2);
}

static bool tiles_ok(const pq_inv *h, int R, int C, bool report) {
  const EngineInfo &e = h->ei;
  const int V = 16 / e.selem;  // 16-byte vectors of the scratch type
  auto fail = [&](const char *why) {
    if (report)
      pq_set_error("pq_inv tiles R=%d C=%d unsupported for N=%d, L=%d, precision %d: %s", R, C,
                   h->N, h->L, h->prec, why);
    return false;
  };
  if (R < 2 || R > 16 || (R & (R - 1)) || h->N % R) return fail("R must be a power of two in 2..16 dividing N");
  if (R < V) return fail("R below the 16-byte vector width");
  if (h->prec == PQ_PREC_FP32_MFMA && R != mfma::kRows) return fail("the matrix-core row passes need R = 16");
  if (row_threads(h, R) > kThreadsMax) return fail("row workgroup above 512 threads");
  if (row_lds(h, 0, R) > kLdsMax || row_lds(h, 2, R) > kLdsMax)
    return fail("row workgroup above 160 KiB of LDS");
  if (C < 1 || C > 8 || (C & (C - 1)) || h->N % C) return fail("C must be a power of two in 1..8 dividing N");
  const int groups = (C * h->L + 1) / 2;
  const int ctpf = h->per ? e.ctpf : e.tpf;
  if (groups * ctpf > kThreadsMax) return fail("column workgroup above 512 threads");
  if (inv_lds_bytes(e, 1, groups, h->per) > kLdsMax) return fail("column workgroup above 160 KiB of LDS");
  if (h->N / C > h->part_cap) return fail("C below the handle's minimum");
  return true;
}

static void default_tiles(const pq_inv *h, int &R, int &C) {
  // 64-byte (FP64) / 64-byte (FP32) transposed runs per column; the largest
  // R the limits allow, halving when a tile does not fit.
  R = 16;
  while (R > 2 && !tiles_ok(h, R, 1, false)) R /= 2;
  if (h->ei.selem == 8 && R > 8) R = 8;  // FP64 scratch: 64-byte runs at R = 8
  // column workgroups of at most 4 FFT groups (C L <= 8 rows), so that two
  // or more fit a CU's LDS; 2 columns where that allows
  C = 2;
  while (C > 1 && ((C * h->L + 1) / 2 > 4 || !tiles_ok(h, R, C, false))) C /= 2;
}

// FP32-storage precision for a stratification: FP32 transforms when the
// layer <-> mode projection is not too ill conditioned (kappa_inf =
// |modes_to_z|_inf |z_to_modes|_inf <= 32), else FP64 transforms.  Measured
// max relative psi error of the FP32 transforms: 1.7e-7 on 1023^2 x 3
// layers (kappa_inf 4.1), 3.4e-7 on 2047^2 x 6 layers (14.4), <= 4e-7 over
// 64 members of 2..8 layers (up to 26.2).
extern "C" int pq_prec_fp32_choice(int L, const double *m2z, const double *z2m) {
  if (!m2z || !z2m) return PQ_PREC_FP32;
  auto norm_inf = [L](const double *a) {  // column-major L x L
    double n = 0.0;
    for (int i = 0; i < L; ++i) {
      double r = 0.0;
      for (int j = 0; j < L; ++j) r += std::fabs(a[i + L * j]);
      n = std::max(n, r);
    }
    return n;
  };
  return norm_inf(m2z) * norm_inf(z2m) <= 32.0 ? PQ_PREC_FP32 : PQ_PREC_FP32_STORAGE;
}

extern "C" int pq_prec_fp32_auto(int g, int L, const double *m2z, const double *z2m) {
  const int c = pq_prec_fp32_choice(L, m2z, z2m);
  return c == PQ_PREC_FP32 && mfma::supported(g - 1) ? PQ_PREC_FP32_MFMA : c;
}

// Per-member sizes (doubles) of the column-pass partials, psiC and the
// row-inverse constants.
static size_t part_per(const pq_inv *h) { return (size_t)h->part_cap * h->L * (h->per ? 2 : 1); }
static size_t psic_per(const pq_inv *h) { return (size_t)h->L * (h->per ? 2 : 1); }
static size_t kb_per(const pq_inv *h) {
  return h->per ? (size_t)h->L * (h->N + 1) : (size_t)(h->L + h->L * h->L + (h->L + 1) * (h->N / 2));
}

static int inv_create(pq_inv **out, int g, int L, const double *ss, const double *m2z,
                      const double *z2m, int precision, int batch_max, int mode, int per = 0,
                      int bc_type = 1, double alpha = 0.0) {
  PQ_CHECK_ARG(out != nullptr, "pq_inv_create: null handle pointer");
  PQ_CHECK_ARG(g >= 9 && pq_supported_n(g - 2),
               "pq_inv_create: unsupported grid size g=%d (need g-1 = 2^a or 3*2^a, 8 <= g-1 <= 4096)", g);
  PQ_CHECK_ARG(L >= 1 && L <= 8, "pq_inv_create: layers must be in 1..8 (got %d)", L);
  PQ_CHECK_ARG(precision >= 0 && precision <= PQ_PREC_FP32_MFMA, "pq_inv_create: bad precision %d",
               precision);
  PQ_CHECK_ARG(precision != PQ_PREC_FP32_MFMA || (mfma::supported(g - 1) && !per),
               "pq_inv_create: PQ_PREC_FP32_MFMA needs the closed basin with g - 1 = %d (got g = %d%s)",
               mfma::kN, g, per ? ", periodic channel" : "");
  if (precision == PQ_PREC_FP32_AUTO)
    precision = per ? pq_prec_fp32_choice(L, m2z, z2m) : pq_prec_fp32_auto(g, L, m2z, z2m);
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
  h->N = g - 1;
  h->L = L;
  h->prec = precision;
  h->batch_max = batch_max;
  h->mode = mode;
  h->per = per;
  h->gj = per ? 2 * (g - 1) : g;
  h->bc_type = bc_type;
  h->alpha = alpha;
  const int N = h->N;
  if (!kernels_for(N, h->ks)) {
    delete h;
    pq_set_error("pq_inv_create: no kernels for N=%d", N);
    return PQ_ERR_ARG;
  }
  h->ei = h->ks.info(precision);
  h->part_cap = N;  // C >= 1
  int R, C;
  default_tiles(h, R, C);
  if (!tiles_ok(h, R, C, true)) {
    delete h;
    return PQ_ERR_ARG;
  }
  h->R = R;
  h->C = C;
  auto fail = [&](int code) {
    pq_inv_destroy(h);
    return code;
  };
  int rc = h->ei.celem == 8 ? upload_tables<double>(h, h->ei.M) : upload_tables<float>(h, h->ei.M);
  if (rc) return fail(rc);
  if (precision == PQ_PREC_FP32_MFMA && (rc = upload_mfma_table(h))) return fail(rc);
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      pq_set_error("pq_inv_create: cannot query the device's compute units");
      return fail(PQ_ERR_HIP);
    }
  }
  h->coef_size = CoefLayout{L, N}.size();
  std::vector<double> coef_h(h->coef_size);
  compute_coefs(h, L, ss, m2z, z2m, coef_h.data());
  h->lay.assign(batch_max, L);
  h->scr_member = (size_t)L * N * (per ? 2 * N : N) * h->ei.selem;
  // every member slot starts with the creation stratification
  if (hipMalloc(&h->coef, h->coef_size * sizeof(double) * batch_max) != hipSuccess) return fail(PQ_ERR_NOMEM);
  for (int b = 0; b < batch_max; ++b)
    if (hipMemcpy(h->coef + h->coef_size * b, coef_h.data(), h->coef_size * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(PQ_ERR_HIP);
  if (hipMalloc(&h->scr, h->scr_member * batch_max) != hipSuccess) {
    pq_set_error("pq_inv_create: cannot allocate %zu bytes of scratch", h->scr_member * batch_max);
    return fail(PQ_ERR_NOMEM);
  }
  if (hipMalloc(&h->part, sizeof(double) * part_per(h) * batch_max) != hipSuccess)
    return fail(PQ_ERR_NOMEM);
  if (hipMalloc(&h->psic, sizeof(double) * psic_per(h) * batch_max) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (hipMalloc(&h->kb, sizeof(double) * kb_per(h) * batch_max) != hipSuccess)
    return fail(PQ_ERR_NOMEM);
  if (hipMalloc(&h->status, sizeof(int) * batch_max) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (hipMemset(h->status, 0, sizeof(int) * batch_max) != hipSuccess) return fail(PQ_ERR_HIP);
  if (hipMemset(h->psic, 0, sizeof(double) * psic_per(h) * batch_max) != hipSuccess) return fail(PQ_ERR_HIP);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(PQ_ERR_HIP);
  *out = h;
  return PQ_OK;
}

// pass_mask bit 0: row forward pass, bit 1: column pass, bit 2: row inverse.
static int launch_chunk(pq_inv *h, int nb, int mem0, const PtrTab &pt, hipStream_t s, int pass_mask) {
  const EngineInfo &e = h->ei;
  const int N = h->N, L = h->L;
  InvArgs a{};
  a.scr = static_cast<char *>(h->scr) + h->scr_member * mem0;
  a.scr_stride = (long long)h->scr_member;
  a.part = h->part + part_per(h) * mem0;
  a.part_stride = (int)part_per(h);
  a.psic = h->psic + psic_per(h) * mem0;
  a.psic_stride = (int)psic_per(h);
  a.status = h->status + mem0;
  a.kb = h->kb + kb_per(h) * mem0;
  a.kb_stride = (int)kb_per(h);
  a.tw64 = h->tw64;
  a.was = h->was;
  a.lam4 = h->lam4;
  a.shat = h->shat;
  a.ctab = h->ctab;
  a.L = L;
  a.mode = h->mode;
  a.per = h->per;
  a.twm = h->twm;
  a.R = h->R;
  a.lgR = ilog2i(h->R);
  a.mtab = h->mtab;
  const int rows_threads = row_threads(h, h->R);
  const int col_groups = (h->C * L + 1) / 2;
  for (int pass = 0; pass < 3; ++pass) {
    if (!(pass_mask & (1 << pass))) continue;
    int blocks, threads;
    size_t lds;
    if (pass == 1 && !h->per) {
      blocks = pt.cum_c[nb];
      threads = tri_threads(N, L);
      lds = 0;
    } else if (pass == 1) {
      blocks = pt.cum_c[nb];
      threads = col_groups * (h->per ? e.ctpf : e.tpf);
      lds = inv_lds_bytes(e, 1, col_groups, h->per);
    } else {
      blocks = pt.cum_r[nb];
      // the matrix-core row passes are persistent: two workgroups per CU
      if (h->prec == PQ_PREC_FP32_MFMA) blocks = std::min(blocks, 2 * std::max(h->ncu, 1));
      threads = rows_threads;
      lds = row_lds(h, pass, h->R);
    }
    if (pass == 2 && h->mode == 0) {
      if (h->per) hipLaunchKernelGGL(k_chan_corr, dim3(nb), dim3(256), 0, s, a, pt, N);
      else hipLaunchKernelGGL(k_psic, dim3(nb * k_psic_groups(N)), dim3(256), 0, s, a, pt, N);
      hipError_t e2 = hipGetLastError();
      if (e2 != hipSuccess) {
        pq_set_error("pq_inv: launch of the psiC reduction failed: %s", hipGetErrorString(e2));
        return PQ_ERR_HIP;
      }
    }
    hipError_t err = h->ks.launch(h->prec, pass, a, pt, blocks, threads, lds, s);
    if (err != hipSuccess) {
      pq_set_error("pq_inv: launch of pass %d failed (N=%d, R=%d, C=%d, %d threads, %zu B LDS): %s", pass,
                   N, h->R, h->C, threads, lds, hipGetErrorString(err));
      return PQ_ERR_HIP;
    }
  }
  return PQ_OK;
}

// coef / layers: per-member coefficient blocks and layer counts (NULL: the
// handle's member slots).
static int inv_launch(pq_inv *h, int batch, const void *const *zeta, long ldz1, long ldz2,
                      void *const *psi, long ldp1, long ldp2, hipStream_t s, int pass_mask = 7,
                      const double *const *coef = nullptr, const int *layers = nullptr) {
  PQ_CHECK_ARG(batch >= 1 && batch <= h->batch_max, "pq_inv_run_dev: batch %d outside 1..%d", batch,
               h->batch_max);
  PQ_CHECK_ARG(ldz1 >= h->g && ldz2 >= h->gj && ldp1 >= h->g && ldp2 >= h->gj,
               "pq_inv_run_dev: leading dimensions smaller than the view %d x %d", h->g, h->gj);
  for (int b = 0; b < batch; ++b) {
    PQ_CHECK_ARG(zeta[b] != nullptr && psi[b] != nullptr, "pq_inv_run_dev: null member pointer %d", b);
    PQ_CHECK_ARG(!layers || (layers[b] >= 1 && layers[b] <= h->L),
                 "pq_inv_run_dev: member %d has %d layers, outside 1..%d", b, layers ? layers[b] : 0, h->L);
    PQ_CHECK_ARG(!coef || coef[b], "pq_inv_run_dev: null coefficient block of member %d", b);
  }
  for (int b0 = 0; b0 < batch; b0 += kMaxBatchArg) {
    const int nb = std::min(kMaxBatchArg, batch - b0);
    PtrTab pt;
    std::memset(&pt, 0, sizeof(pt));
    // per-member layers, column tiles (the widest power of two up to 8
    // whose C L_b rows fit the column workgroup's groups) and block ranges
    const int groups = (h->C * h->L + 1) / 2;
    pt.nb = nb;
    pt.cum_r[0] = pt.cum_c[0] = 0;
    for (int b = 0; b < nb; ++b) {
      pt.zeta[b] = zeta[b0 + b];
      pt.psi[b] = psi[b0 + b];
      pt.coef[b] = coef ? coef[b0 + b] : h->coef + h->coef_size * (size_t)(b0 + b);
      const int Lb = layers ? layers[b0 + b] : h->lay[b0 + b];
      int Cb = h->C;
      if (!h->per) {
        Cb = tri_cols(h->N, h->L);  // one column geometry for the launch
      } else if (Lb != h->L) {
        Cb = 8;
        while (Cb > 1 && ((Cb * Lb + 1) / 2 > groups || h->N % Cb)) Cb /= 2;
      }
      pt.lay[b] = (unsigned char)Lb;
      pt.lgc[b] = (unsigned char)ilog2i(Cb);
      pt.cum_r[b + 1] = pt.cum_r[b] + Lb * ((h->per ? 2 * h->N : h->N) / h->R);
      pt.cum_c[b + 1] = pt.cum_c[b] + h->N / Cb;
    }
    pt.ldz1 = ldz1;
    pt.ldz2 = ldz2;
    pt.ldp1 = ldp1;
    pt.ldp2 = ldp2;
    int rc = launch_chunk(h, nb, b0, pt, s, pass_mask);
    if (rc) return rc;
  }
  return PQ_OK;
}

namespace pq {
size_t inv_coef_size(const pq_inv *h) { return h->coef_size; }

int inv_make_coefs(const pq_inv *h, int layers, const double *ss, const double *m2z,
                   const double *z2m, double *dev_block) {
  PQ_CHECK_ARG(h && dev_block && ss, "inv_make_coefs: null argument");
  PQ_CHECK_ARG(layers >= 1 && layers <= h->L, "inv_make_coefs: layers %d outside 1..%d", layers, h->L);
  std::vector<double> c(h->coef_size);
  compute_coefs(h, layers, ss, m2z, z2m, c.data());
  PQ_HIP(hipMemcpy(dev_block, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
  return PQ_OK;
}

int inv_run_members(pq_inv *h, int batch, const void *const *zeta, long ldz1, long ldz2,
                    void *const *psi, long ldp1, long ldp2, const double *const *coef,
                    const int *layers, hipStream_t s) {
  PQ_CHECK_ARG(h && coef && layers, "inv_run_members: null argument");
  return inv_launch(h, batch, zeta, ldz1, ldz2, psi, ldp1, ldp2, s, 7, coef, layers);
}

int *inv_status_dev(pq_inv *h) { return h->status; }
int inv_batch_max(const pq_inv *h) { return h->batch_max; }

size_t inv_device_bytes(const pq_inv *h) {
  const size_t N = (size_t)h->N, L = (size_t)h->L, B = (size_t)h->batch_max;
  return h->scr_member * B +
         sizeof(double) * (h->coef_size * B + part_per(h) * B + psic_per(h) * B + kb_per(h) * B +
                           2 * (size_t)h->twm + N + (h->per ? 3 * N : 0)) +
         (size_t)h->ei.celem * (N / 2 + N + 1) + sizeof(int) * B;
}
}  // namespace pq

extern "C" {

int pq_inv_create(pq_inv **h, int g, int layers, const double *ss, const double *modes_to_z,
                  const double *z_to_modes, int precision, int batch_max) {
  return inv_create(h, g, layers, ss, modes_to_z, z_to_modes, precision, batch_max, 0);
}

int pq_chan_create(pq_inv **h, int gi, int gj, int layers, const double *ss,
                   const double *modes_to_z, const double *z_to_modes, int bc_type, double alpha,
                   int precision, int batch_max) {
  PQ_CHECK_ARG(gj == 2 * (gi - 1), "pq_chan_create: the period must be gj = 2 (gi - 1) (gi=%d, gj=%d)",
               gi, gj);
  PQ_CHECK_ARG(bc_type >= 1 && bc_type <= 3, "pq_chan_create: bad bc_type %d", bc_type);
  PQ_CHECK_ARG(bc_type != 1 || (alpha >= 0.0 && std::isfinite(alpha)),
               "pq_chan_create: alpha must be finite and >= 0");
  return inv_create(h, gi, layers, ss, modes_to_z, z_to_modes, precision, batch_max, 0, 1, bc_type,
                    alpha);
}

int pq_inv_destroy(pq_inv *h) {
  if (!h) return PQ_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  (void)hipFree(h->tw64);
  (void)hipFree(h->was);
  (void)hipFree(h->lam4);
  (void)hipFree(h->mtab);
  (void)hipFree(h->shat);
  (void)hipFree(h->ctab);
  (void)hipFree(h->coef);
  (void)hipFree(h->scr);
  (void)hipFree(h->part);
  (void)hipFree(h->psic);
  (void)hipFree(h->kb);
  (void)hipFree(h->status);
  (void)hipFree(h->dz);
  (void)hipFree(h->dp);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PQ_OK;
}

void *pq_inv_stream(pq_inv *h) { return h ? (void *)h->stream : nullptr; }

int pq_inv_set_tiles(pq_inv *h, int rows, int cols) {
  PQ_CHECK_ARG(h != nullptr, "pq_inv_set_tiles: null handle");
  if (!tiles_ok(h, rows, cols, true)) return PQ_ERR_ARG;
  // the scratch layout depends on R: let running work finish first
  PQ_HIP(hipStreamSynchronize(h->stream));
  PQ_HIP(hipDeviceSynchronize());
  h->R = rows;
  h->C = cols;
  return PQ_OK;
}

int pq_inv_get_tiles(pq_inv *h, int *rows, int *cols) {
  PQ_CHECK_ARG(h && rows && cols, "pq_inv_get_tiles: null argument");
  *rows = h->R;
  *cols = h->C;
  return PQ_OK;
}

int pq_inv_set_member_layers(pq_inv *h, int slot, int layers, const double *ss,
                             const double *modes_to_z, const double *z_to_modes) {
  PQ_CHECK_ARG(h != nullptr, "pq_inv_set_member: null handle");
  PQ_CHECK_ARG(h->mode == 0, "pq_inv_set_member: only for PV inversion handles");
  PQ_CHECK_ARG(slot >= 0 && slot < h->batch_max, "pq_inv_set_member: slot %d outside 0..%d", slot,
               h->batch_max - 1);
  PQ_CHECK_ARG(layers >= 1 && layers <= h->L, "pq_inv_set_member: layers %d outside 1..%d", layers,
               h->L);
  PQ_CHECK_ARG(ss && modes_to_z && z_to_modes, "pq_inv_set_member: null stratification");
  const int L = layers;
  PQ_CHECK_ARG(ss[0] == 0.0, "pq_inv_set_member: ss(1) must be 0 (barotropic mode)");
  for (int l = 0; l < L; ++l) {
    PQ_CHECK_ARG(modes_to_z[l] == 1.0, "pq_inv_set_member: modes_to_z(:,1) must be all ones");
    PQ_CHECK_ARG(ss[l] >= 0.0 && std::isfinite(ss[l]), "pq_inv_set_member: ss must be finite and >= 0");
  }
  std::vector<double> c(h->coef_size);
  compute_coefs(h, L, ss, modes_to_z, z_to_modes, c.data());
  // ordered after the handle's earlier work on its stream
  PQ_HIP(hipMemcpyAsync(h->coef + h->coef_size * slot, c.data(), h->coef_size * sizeof(double),
                        hipMemcpyHostToDevice, h->stream));
  PQ_HIP(hipStreamSynchronize(h->stream));
  h->lay[slot] = L;
  return PQ_OK;
}

int pq_inv_set_member(pq_inv *h, int slot, const double *ss, const double *modes_to_z,
                      const double *z_to_modes) {
  PQ_CHECK_ARG(h != nullptr, "pq_inv_set_member: null handle");
  return pq_inv_set_member_layers(h, slot, h->L, ss, modes_to_z, z_to_modes);
}

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
  if (psic) PQ_HIP(hipMemcpy(psic, h->psic, sizeof(double) * psic_per(h) * h->batch_max, hipMemcpyDeviceToHost));
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

}  // extern "C"

// Host-view copy helpers: g x g x L block of a (ld1, ld2) host view <-> a
// packed device array of FP64 (el 8) or FP32 (el 4).
static int copy_view_h2d(void *dev, const double *host, int gi, int gj, int L, long ld1, long ld2,
                         size_t el, hipStream_t s) {
  if (el == 8) {
    for (int l = 0; l < L; ++l)
      PQ_HIP(hipMemcpy2DAsync((char *)dev + (size_t)l * gi * gj * 8, (size_t)gi * 8,
                              host + (size_t)l * ld1 * ld2, (size_t)ld1 * 8, (size_t)gi * 8, gj,
                              hipMemcpyHostToDevice, s));
    PQ_HIP(hipStreamSynchronize(s));
    return PQ_OK;
  }
  std::vector<float> tmp((size_t)gi * gj * L);
  for (int l = 0; l < L; ++l)
    for (int j = 0; j < gj; ++j)
      for (int i = 0; i < gi; ++i)
        tmp[((size_t)l * gj + j) * gi + i] = (float)host[i + ld1 * (j + ld2 * (size_t)l)];
  PQ_HIP(hipMemcpyAsync(dev, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice, s));
  PQ_HIP(hipStreamSynchronize(s));
  return PQ_OK;
}

static int copy_view_d2h(double *host, const void *dev, int gi, int gj, int L, long ld1, long ld2,
                         size_t el, hipStream_t s) {
  if (el == 8) {
    for (int l = 0; l < L; ++l)
      PQ_HIP(hipMemcpy2DAsync(host + (size_t)l * ld1 * ld2, (size_t)ld1 * 8,
                              (const char *)dev + (size_t)l * gi * gj * 8, (size_t)gi * 8,
                              (size_t)gi * 8, gj, hipMemcpyDeviceToHost, s));
    PQ_HIP(hipStreamSynchronize(s));
    return PQ_OK;
  }
  std::vector<float> tmp((size_t)gi * gj * L);
  PQ_HIP(hipMemcpyAsync(tmp.data(), dev, tmp.size() * 4, hipMemcpyDeviceToHost, s));
  PQ_HIP(hipStreamSynchronize(s));
  for (int l = 0; l < L; ++l)
    for (int j = 0; j < gj; ++j)
      for (int i = 0; i < gi; ++i) host[i + ld1 * (j + ld2 * (size_t)l)] = tmp[((size_t)l * gj + j) * gi + i];
  return PQ_OK;
}

static int run_host(pq_inv *h, const double *zeta, int ldz1, int ldz2, double *psi, int ldp1, int ldp2) {
  PQ_CHECK_ARG(zeta && psi, "pq_inv_run: null array");
  PQ_CHECK_ARG(ldz1 >= h->g && ldz2 >= h->gj && ldp1 >= h->g && ldp2 >= h->gj,
               "pq_inv_run: leading dimensions smaller than the view %d x %d", h->g, h->gj);
  const int g = h->g, gj = h->gj, L = h->L;
  // FP64 storage for FP64 and mixed handles, FP32 for FP32 handles
  const size_t el = (h->prec == PQ_PREC_FP64 || h->prec == PQ_PREC_MIXED) ? 8 : 4;
  void *mz = pq_mirror_device_ptr(zeta);
  void *mp = pq_mirror_device_ptr(psi);
  if (mz && mp && el == 8) {
    const void *zt[1] = {mz};
    void *pt[1] = {mp};
    int rc = inv_launch(h, 1, zt, ldz1, ldz2, pt, ldp1, ldp2, h->stream);
    if (rc) return rc;
    return pq_inv_status(h, nullptr);
  }
  const size_t bytes = (size_t)g * gj * L * el;
  if (!h->dz) {
    PQ_HIP(hipMalloc(&h->dz, bytes));
    PQ_HIP(hipMalloc(&h->dp, bytes));
  }
  int rc = copy_view_h2d(h->dz, zeta, g, gj, L, ldz1, ldz2, el, h->stream);
  if (rc) return rc;
  if (h->mode == 1) {
    // helm works in place: the output keeps input values on part of the
    // boundary, so start psi from zeta.
    PQ_HIP(hipMemcpyAsync(h->dp, h->dz, bytes, hipMemcpyDeviceToDevice, h->stream));
  }
  const void *zt[1] = {h->dz};
  void *pt[1] = {h->dp};
  rc = inv_launch(h, 1, zt, g, gj, pt, g, gj, h->stream);
  if (rc) return rc;
  rc = copy_view_d2h(psi, h->dp, g, gj, L, ldp1, ldp2, el, h->stream);
  if (rc) return rc;
  return pq_inv_status(h, nullptr);
}

extern "C" {

int pq_inv_run(pq_inv *h, const double *zeta, int ldz1, int ldz2, double *psi, int ldp1, int ldp2) {
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

int pq_helm_create_periodic(pq_helm **out, int ii, int jj, int layers, const double *ss,
                            int precision) {
  PQ_CHECK_ARG(out != nullptr, "pq_helm_create_periodic: null handle pointer");
  PQ_CHECK_ARG(jj == 2 * (ii - 1),
               "pq_helm_create_periodic: the period must be jj = 2 (ii - 1) (ii=%d, jj=%d)", ii, jj);
  pq_inv *inv = nullptr;
  int rc = inv_create(&inv, ii, layers, ss, nullptr, nullptr, precision, 1, 1, 1);
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
  return inv_launch(h->inv, 1, zt, ld1, ld2, pt, ld1, ld2, pq_resolve_stream(stream, h->inv->stream));
}

}  // extern "C"
