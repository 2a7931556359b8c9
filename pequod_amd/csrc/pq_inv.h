// pq_inv.h -- types shared by the inversion kernels (pq_inv_kernels.hip,
// compiled once per transform length) and the host side (pq_inversion.hip).
#pragma once

#include "pq_common.h"

namespace pq {

// Coefficient block of one member (doubles):
//   D[L*L]   z_to_modes, D(m,l) at m + L*l
//   M[L*L]   modes_to_z, M(l,m) at l + L*m
//   ss[L], a[L], b[L], ssN[L] = (2N)^2 ss
//   tab[(L-1) * N/2 * 4]: (rho, gamma, eta, r = e^rho) for mode m >= 1
//                         (0-based) and odd p = 2h + 1 (closed basin); the
//                         channel keeps
//                         inv_corr_a of mode m >= 1 at tab + 4 m (2 x 2,
//                         column-major)
//   tri[L * N * 8]: the tridiagonal column pass's constants of mode m and
//                   column p at (m N + p) * 8: r, 1/r, 1 / (r - 1/r), r^N,
//                   1 / (1 - r^2N), r^E (E of the handle's launches), 0, 0,
//                   with r + 1/r = 2 + mu, mu = 4 sin^2(p pi / 2N) + ss_m,
//                   0 < r < 1 (p = 0, an all-zero column: r = 1/2)
struct CoefLayout {
  int L, N;
  __host__ __device__ int D() const { return 0; }
  __host__ __device__ int M() const { return L * L; }
  __host__ __device__ int ss() const { return 2 * L * L; }
  __host__ __device__ int a() const { return 2 * L * L + L; }
  __host__ __device__ int b() const { return 2 * L * L + 2 * L; }
  __host__ __device__ int ssN() const { return 2 * L * L + 3 * L; }
  __host__ __device__ int tab() const { return 2 * L * L + 4 * L; }
  __host__ __device__ size_t tri() const {
    return (size_t)tab() + (size_t)(L > 1 ? L - 1 : 0) * (N / 2) * 4;
  }
  __host__ __device__ size_t size() const { return tri() + (size_t)L * N * 8; }
};

// Tridiagonal column pass: each thread solves E consecutive rows j of one
// column for every mode, so a column takes N / E threads.  E = 8 for
// N = 2^a; 12 for N = 3 * 2^a (6 above 4 layers, for registers).  Never
// more than 512 threads per column.
__host__ __device__ constexpr int tri_e(int N, int LT) {
  return N % 3 == 0 ? (LT <= 4 ? 12 : 6) : 8;
}
// Layer count the kernel is instantiated for: exact up to 4, then 6 (for
// 5 and 6 layers: 167 VGPRs at N = 2048, 3 waves per SIMD) and 8.
__host__ __device__ constexpr int tri_lt(int L) { return L <= 4 ? L : (L <= 6 ? 6 : 8); }
// Threads per workgroup and columns per workgroup of the tridiagonal pass.
__host__ __device__ constexpr int tri_tpc(int N, int L) { return N / tri_e(N, tri_lt(L)); }
// (a power of two dividing N: the psiC partials are kept per block of columns)
__host__ __device__ constexpr int tri_cols(int N, int L) {
  return tri_tpc(N, L) >= 256 ? 1
                              : (256 / tri_tpc(N, L) < (N & -N) ? 256 / tri_tpc(N, L) : (N & -N));
}
__host__ __device__ constexpr int tri_threads(int N, int L) { return tri_cols(N, L) * tri_tpc(N, L); }

// Member arrays of one launch (passed by value).  Members may have fewer
// layers than the handle (ragged batches): member b has lay[b] layers, its
// column-pass tiles are 2^lgc[b] columns wide, and its workgroups are the
// blocks cum_r[b] .. cum_r[b+1]-1 of the row passes and cum_c[b] ..
// cum_c[b+1]-1 of the column pass.  coef[b] is the member's coefficient
// block, for its layer count.
constexpr int kMaxBatchArg = 64;
struct PtrTab {
  const void *zeta[kMaxBatchArg];
  void *psi[kMaxBatchArg];
  const double *coef[kMaxBatchArg];  // coefficient block (CoefLayout) of member b
  long long ldz1, ldz2, ldp1, ldp2;
  int nb;
  int cum_r[kMaxBatchArg + 1];
  int cum_c[kMaxBatchArg + 1];
  unsigned char lay[kMaxBatchArg];
  unsigned char lgc[kMaxBatchArg];
};

// Member of block `blk` given the block prefix sums cum[0..nb] (binary search).
__device__ __forceinline__ int member_of(const int *cum, int nb, int blk) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cum[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Launch-wide arguments of the three passes.  Member b of a launch (b < nb)
// uses the scratch at scr + b * scr_stride bytes (L x N x J elements of the
// compute type, tile-blocked; J = N rows of j in the closed basin, 2N in
// the channel), the column pass's partials at part + b * part_stride
// ((N / C) x L psiC partials; channel: (N / C) x 2L boundary sums),
// psic + b * psic_stride (L values; channel: psiC pairs of the modes) and
// status + b.
struct InvArgs {
  char *scr;
  long long scr_stride;
  double *part;
  int part_stride;
  double *psic;      // psic_stride per member: psiC of the last run
  int psic_stride;
  double *kb;        // per member: kappa_l (L), then beta_lm at L + l L + m,
                     // then rmin[N/2] and Gamma_l[N/2] (l = 0..L-1; k_psic);
                     // channel: the correction E_l at l (N + 1) + p (p = 1..N-1
                     // spectral, p = 0 and N the wall values)
  int kb_stride;
  int *status;       // sticky per member: 1 NaN psiC, 2 |psiC| > 1e100
  const void *tw64;  // cx<double>[M]: exp(-2 pi i n / M), M the engine's FFT length
  const void *was;   // Tc[N/2]: 2 sin(k pi / N) (DstSint), Tc the arithmetic type
  const void *lam4;  // Tc[N + 1]: (2N)^2 * (-4 sin^2(p pi / (2N)))
  const double *shat;  // double[N]: sum_i 2 sin(pi p i / N) (0 for even p)
  // channel: double[3N]: boundary weights w_p, sin(pi p / N) / N, lambda_p
  const double *ctab;
  int L, mode;       // L: the handle's (largest) layer count; mode 0 pv_inversion, 1 helm
  int per;           // 0 closed basin (N rows of j), 1 periodic channel (2N rows)
  int twm;           // length of the twiddle table tw64
  int R, lgR;        // rows per tile of the row passes (scratch block height)
  const void *mtab;  // PQ_PREC_FP32_MFMA: the matrix-core engine's per-lane constants
};

// Static properties of a transform engine (per precision and length).
struct EngineInfo {
  int M;             // complex FFT length
  int tpf;           // threads per FFT group (one group = two rows)
  int region;        // LDS elements of T per group
  int scan;          // LDS elements of T of scan scratch per group
  int elem;          // sizeof(T), the LDS (row and FFT buffer) type
  int selem;         // sizeof of the scratch type
  int celem;         // sizeof of the arithmetic type (tables lam4, was)
  int fregion;       // LDS elements of T per group of the row forward pass
  int fscan;         //   (forward engine only) and its scan scratch
  int ctpf;          // channel column pass (periodic engine of length 2N):
  int cregion;       //   threads per group and LDS elements of T per group
};

// Supported transform lengths N = 2^a (8..4096) and 3 * 2^a (12..3072); one
// kernel unit (pq_inv_kernels.hip built with -DPQ_N=N) per length.
#define PQ_INV_SIZES(X)                                                        \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384)   \
  X(512) X(768) X(1024) X(1536) X(2048) X(3072) X(4096)

// Workgroups per member of the psiC reduction (k_psic): one 256-thread
// workgroup per 256 of the N/2 odd columns.
__host__ __device__ constexpr int k_psic_groups(int N) { return (N / 2 + 255) / 256; }

// Entry points of a kernel unit.  prec: PQ_PREC_*; pass: 0 row forward,
// 1 column, 2 row inverse.
#define PQ_INV_DECL(n)                                                         \
  EngineInfo inv_info_##n(int prec);                                           \
  hipError_t inv_launch_##n(int prec, int pass, const InvArgs &a,              \
                            const PtrTab &pt, int blocks, int threads,         \
                            size_t lds, hipStream_t s);
PQ_INV_SIZES(PQ_INV_DECL)
#undef PQ_INV_DECL

// LDS bytes of a pass for `groups` FFT groups (per: the channel's column
// pass).
inline size_t inv_lds_bytes(const EngineInfo &e, int pass, int groups, bool per = false) {
  const int reg = pass == 0 ? e.fregion + e.fscan : (pass == 1 && per) ? e.cregion : e.region + e.scan;
  size_t b = (size_t)e.elem * groups * reg;
  b = (b + 15) & ~size_t(15);
  if (pass == 1) b += 64 * 8 + (size_t)e.celem * (64 + 64 + 8);
  if (pass == 2) b += 16;
  return b;
}

}  // namespace pq
