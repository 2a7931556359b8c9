// pq_diag.hip -- on-device FP64 diagnostics of the stream function:
//   * energy (qg3p_diagnostics.F90:778-863): kinetic energy per layer and
//     the potential energy of the interfaces, closed basin on one domain;
//   * the time-mean accumulation psi_av += dt * psi (qg3.F90:672-681).
//
// Energy is a reduction over the g x g nodes and (g-1)^2 cells of every
// layer and reads psi once: HBM-bound.  A workgroup covers 64 rows and
// 4 x kCols columns of one tile (one wave per kCols columns, the 64 lanes
// along i so that every column read is one 512-byte segment); each thread
// walks its columns layer by layer, keeping the previous column (the cell's
// west corners) and the previous layer's values (the interface difference)
// in registers.  The i-1 neighbours are read from the same cache lines as
// the i values of the lane below.  Workgroups write their partial sums to a
// buffer owned by the handle, and a second kernel adds them in a fixed
// order: results are bit-identical from run to run for a given g and layer
// count.  They agree with the reference's loop order to FP64 rounding.
//
// Potential energy weights every node by w_i w_j with w = 1/2 at the first
// and last index and 1 elsewhere, which is the reference's interior +
// half-weight edges + quarter-weight corners; kinetic energy sums, per cell
// with corners (i-1, j-1), (i, j-1), (i-1, j), (i, j) for i, j in 2..g,
// 1/4 of the squares of the two doubled centred differences.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "pq_common.h"
#include "pq_internal.h"

namespace {

constexpr int kMaxL = 8;    // layer limit of the library
constexpr int kCols = 16;   // columns per wave
constexpr int kRows = 64;   // rows per workgroup (one lane each)
constexpr int kWaves = 4;   // waves per workgroup, side by side along j

struct EnergyArgs {
  const void *psi;         // FP64, or FP32 (read widened) for FP32-field models
  long ld1, lstride;       // (i, j, l) at psi[(i-1) + ld1 (j-1) + lstride l]
  int g, L, nblk;
  double fac[kMaxL];       // 0.5 s2(k) h(k) / H for the interfaces k = 1..L-1
  double hr[kMaxL];        // h(k) / H
  double cff;              // 1 / (g-1)^2
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Uniform select of a[k] (no dynamic indexing of the argument block).
__device__ __forceinline__ double pick(const double (&a)[kMaxL], int k) {
  double r = 0.0;
#pragma unroll
  for (int q = 0; q < kMaxL; ++q)
    if (q == k) r = a[q];
  return r;
}

template <typename T>
__global__ __launch_bounds__(kRows *kWaves) void k_energy_part(EnergyArgs a, double *part) {
  __shared__ double red[kWaves][kMaxL + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = a.g;
  const int i = blockIdx.x * kRows + lane + 1;
  const int jb = (blockIdx.y * kWaves + w) * kCols + 1;
  const bool row = i <= g;
  const bool up = row && i > 1;
  const double wi = (i == 1 || i == g) ? 0.5 : 1.0;
  double prev[kCols];
#pragma unroll
  for (int q = 0; q < kCols; ++q) prev[q] = 0.0;
  double ep = 0.0;
  for (int l = 0; l < a.L; ++l) {
    const T *p = static_cast<const T *>(a.psi) + a.lstride * l + (i - 1);
    const double f = l > 0 ? pick(a.fac, l - 1) : 0.0;
    double ek = 0.0;
    // west corners of the first cell: column jb - 1
    double ap = 0.0, bp = 0.0;
    if (row && jb > 1 && jb <= g) {
      ap = double(p[a.ld1 * (jb - 2)]);
      if (up) bp = double(p[a.ld1 * (jb - 2) - 1]);
    }
#pragma unroll
    for (int q = 0; q < kCols; ++q) {
      const int j = jb + q;
      const bool ok = row && j <= g;
      double x = 0.0, y = 0.0;  // psi(i, j), psi(i-1, j)
      if (ok) {
        x = double(p[a.ld1 * (j - 1)]);
        if (up) y = double(p[a.ld1 * (j - 1) - 1]);
      }
      if (ok && up && j > 1) {
        const double d1 = y + x - bp - ap;  // 2 d(psi)/dj on the cell
        const double d2 = ap + x - bp - y;  // 2 d(psi)/di on the cell
        ek += 0.25 * (d1 * d1 + d2 * d2);
      }
      if (ok && l > 0) {
        const double wj = (j == 1 || j == g) ? 0.5 : 1.0;
        const double dd = prev[q] - x;
        ep += f * ((wi * wj) * (dd * dd));
      }
      prev[q] = x;
      ap = x;
      bp = y;
    }
    ek = wave_sum(ek);
    if (lane == 0) red[w][l] = ek;
  }
  ep = wave_sum(ep);
  if (lane == 0) red[w][a.L] = ep;
  __syncthreads();
  if ((int)threadIdx.x <= a.L) {
    const int q = threadIdx.x;
    double s = red[0][q];
#pragma unroll
    for (int v = 1; v < kWaves; ++v) s += red[v][q];
    part[(size_t)q * a.nblk + blockIdx.y * gridDim.x + blockIdx.x] = s;
  }
}

// One workgroup per quantity q (ekin of layer q < L, epot for q = L): sums
// the nblk partials in a fixed order and applies the reference's scaling.
__global__ __launch_bounds__(256) void k_energy_final(EnergyArgs a, const double *part, double *out) {
  __shared__ double red[4];
  const int q = blockIdx.x;
  const double *src = part + (size_t)q * a.nblk;
  double s = 0.0;
  for (int b = threadIdx.x; b < a.nblk; b += 256) s += src[b];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = ((red[0] + red[1]) + red[2]) + red[3];
    out[q] = q < a.L ? 0.5 * a.cff * pick(a.hr, q) * t : a.cff * t;
  }
}

// psi_av(i, j, l) += dt * psi(i, j, l) over all g x g nodes (psi FP64, or
// FP32 read widened).
template <typename T>
__global__ __launch_bounds__(256) void k_psi_av_add(int g, double dt, const T *psi, long ld1,
                                                    long ls, double *av, long la1, long las) {
  // a product and a sum, each rounded (no fused multiply-add), as written
#pragma clang fp contract(off)
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int l = blockIdx.z;
  if (i >= g || j >= g) return;
  double &o = av[i + la1 * j + las * l];
  o = o + dt * double(psi[i + ld1 * j + ls * l]);
}

}  // namespace

struct pq_diag {
  int g = 0, L = 0;
  EnergyArgs args{};
  dim3 grid;
  hipStream_t stream = nullptr;
  double *part = nullptr;     // (L + 1) x nblk partial sums
  double *out = nullptr;      // L + 1 results (device)
  double *out_host = nullptr; // pinned copy for the synchronous call
  double *stage = nullptr;    // g x g x L copy of a host field
};

// Energy of a device field into host ekin[L], *epot on the handle's stream
// (after the work already queued there), synchronously; PQ_ERR_NAN when a
// result is NaN (the reference aborts, qg3p_diagnostics.F90:857-860).
int pq::diag_energy_sync(pq_diag *d, const void *psi, long ld1, long ld2, double *ekin, double *epot,
                         int es) {
  const int L = d->L;
  int rc = pq::diag_energy_dev_es(d, psi, ld1, ld2, d->out, d->stream, es);
  if (rc) return rc;
  PQ_HIP(hipMemcpyAsync(d->out_host, d->out, sizeof(double) * (L + 1), hipMemcpyDeviceToHost, d->stream));
  PQ_HIP(hipStreamSynchronize(d->stream));
  bool nan = false;
  for (int k = 0; k <= L; ++k) nan = nan || std::isnan(d->out_host[k]);
  std::memcpy(ekin, d->out_host, sizeof(double) * L);
  *epot = d->out_host[L];
  if (nan) {
    pq_set_error("energy: NaN value encountered");
    return PQ_ERR_NAN;
  }
  return PQ_OK;
}

extern "C" {

int pq_diag_destroy(pq_diag *d) {
  if (!d) return PQ_OK;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  (void)hipFree(d->part);
  (void)hipFree(d->out);
  (void)hipFree(d->stage);
  (void)hipHostFree(d->out_host);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return PQ_OK;
}

int pq_diag_create(pq_diag **out, int g, int layers, const double *h, const double *s2) {
  PQ_CHECK_ARG(out && h && (layers == 1 || s2), "pq_diag_create: null argument");
  PQ_CHECK_ARG(g >= 3 && layers >= 1 && layers <= kMaxL, "pq_diag_create: bad size g=%d layers=%d", g,
               layers);
  double H = 0.0;
  for (int k = 0; k < layers; ++k) {
    PQ_CHECK_ARG(h[k] > 0.0, "pq_diag_create: layer depth %d is not positive", k + 1);
    H += h[k];
  }
  pq_diag *d = new pq_diag();
  d->g = g;
  d->L = layers;
  EnergyArgs &a = d->args;
  a.g = g;
  a.L = layers;
  for (int k = 0; k < kMaxL; ++k) {
    a.hr[k] = k < layers ? h[k] / H : 0.0;
    a.fac[k] = k < layers - 1 ? 0.5 * s2[k] * h[k] / H : 0.0;
  }
  a.cff = 1.0 / ((double)(g - 1) * (double)(g - 1));
  d->grid = dim3((g + kRows - 1) / kRows, (g + kWaves * kCols - 1) / (kWaves * kCols));
  a.nblk = (int)(d->grid.x * d->grid.y);
  auto fail = [&](int rc) {
    pq_diag_destroy(d);
    return rc;
  };
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) return fail(PQ_ERR_HIP);
  if (hipMalloc((void **)&d->part, sizeof(double) * (size_t)(layers + 1) * a.nblk) != hipSuccess ||
      hipMalloc((void **)&d->out, sizeof(double) * (layers + 1)) != hipSuccess ||
      hipHostMalloc((void **)&d->out_host, sizeof(double) * (layers + 1), 0) != hipSuccess) {
    pq_set_error("pq_diag_create: out of device memory");
    return fail(PQ_ERR_NOMEM);
  }
  *out = d;
  return PQ_OK;
}

void *pq_diag_stream(pq_diag *d) { return d ? (void *)d->stream : nullptr; }

int pq_diag_energy_dev(pq_diag *d, const void *psi, long ld1, long ld2, double *out_dev, void *stream) {
  return pq::diag_energy_dev_es(d, psi, ld1, ld2, out_dev, stream, 8);
}

}  // extern "C"

int pq::diag_energy_dev_es(pq_diag *d, const void *psi, long ld1, long ld2, double *out_dev,
                           void *stream, int es) {
  PQ_CHECK_ARG(d && psi && out_dev, "pq_diag_energy_dev: null argument");
  PQ_CHECK_ARG(ld1 >= d->g && ld2 >= d->g, "pq_diag_energy_dev: leading dimensions smaller than g=%d",
               d->g);
  const hipStream_t s = pq_resolve_stream(stream, d->stream);
  EnergyArgs a = d->args;
  a.psi = psi;
  a.ld1 = ld1;
  a.lstride = ld1 * ld2;
  if (es == 4)
    hipLaunchKernelGGL(k_energy_part<float>, d->grid, dim3(kRows * kWaves), 0, s, a, d->part);
  else
    hipLaunchKernelGGL(k_energy_part<double>, d->grid, dim3(kRows * kWaves), 0, s, a, d->part);
  hipLaunchKernelGGL(k_energy_final, dim3(d->L + 1), dim3(256), 0, s, a, (const double *)d->part,
                     out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pq_set_error("pq_diag_energy_dev: kernel launch failed: %s", hipGetErrorString(e));
    return PQ_ERR_HIP;
  }
  return PQ_OK;
}

extern "C" {

int pq_diag_energy(pq_diag *d, const double *psi, int ld1, int ld2, double *ekin, double *epot) {
  PQ_CHECK_ARG(d && psi && ekin && epot, "pq_diag_energy: null argument");
  PQ_CHECK_ARG(ld1 >= d->g && ld2 >= d->g, "pq_diag_energy: leading dimensions smaller than g=%d", d->g);
  const int g = d->g, L = d->L;
  const void *dev = pq_mirror_device_ptr(psi);
  if (dev) return pq::diag_energy_sync(d, dev, ld1, ld2, ekin, epot);
  if (!d->stage) PQ_HIP(hipMalloc((void **)&d->stage, sizeof(double) * (size_t)g * g * L));
  for (int l = 0; l < L; ++l)
    PQ_HIP(hipMemcpy2DAsync(d->stage + (size_t)g * g * l, sizeof(double) * g, psi + (size_t)ld1 * ld2 * l,
                            sizeof(double) * ld1, sizeof(double) * g, g, hipMemcpyHostToDevice, d->stream));
  return pq::diag_energy_sync(d, d->stage, g, g, ekin, epot);
}

int pq_psi_av_add_dev(int g, int layers, double dt, const void *psi, long ld1, long ld2, double *psi_av,
                      long la1, long la2, void *stream) {
  return pq::psi_av_add_es(g, layers, dt, psi, ld1, ld2, psi_av, la1, la2, stream, 8);
}

}  // extern "C"

int pq::psi_av_add_es(int g, int layers, double dt, const void *psi, long ld1, long ld2, double *psi_av,
                      long la1, long la2, void *stream, int es) {
  PQ_CHECK_ARG(psi && psi_av, "pq_psi_av_add_dev: null argument");
  PQ_CHECK_ARG(g >= 1 && layers >= 1 && ld1 >= g && ld2 >= g && la1 >= g && la2 >= g,
               "pq_psi_av_add_dev: bad size g=%d layers=%d", g, layers);
  const hipStream_t s = pq_resolve_stream(stream, (hipStream_t)0);
  const dim3 grid((g + 63) / 64, (g + 3) / 4, layers);
  if (es == 4)
    hipLaunchKernelGGL(k_psi_av_add<float>, grid, dim3(256), 0, s, g, dt, (const float *)psi, ld1,
                       ld1 * ld2, psi_av, la1, la1 * la2);
  else
    hipLaunchKernelGGL(k_psi_av_add<double>, grid, dim3(256), 0, s, g, dt, (const double *)psi, ld1,
                       ld1 * ld2, psi_av, la1, la1 * la2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pq_set_error("pq_psi_av_add_dev: kernel launch failed: %s", hipGetErrorString(e));
    return PQ_ERR_HIP;
  }
  return PQ_OK;
}
