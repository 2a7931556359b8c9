// pq_cabaret.hip -- CABARET PV advection stages and the device-resident
// time step on MI355X.
//
// Each stage kernel restates one routine of qg3p_aux.F90 (cited per kernel)
// on a single, non-periodic domain, with the reference's array shapes and
// 1-based index ranges (macros below).  FP contraction is disabled in this
// file so each statement rounds exactly as written, in the order written;
// the CPU oracle is compiled the same way, which makes the stage kernels
// bit-for-bit comparable with it.
//
// Index ranges on one domain (qg3p_base.F90:887-941): owned nodes in `data`
// are 2..G+1 (fa_*), interior nodes 3..G (f_*); i-faces of FLUX_II arrays
// sit between data nodes i-1 and i.
#pragma clang fp contract(off)

#include "pq_common.h"
#include "pq_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace {

// 1-based element addresses of the reference's arrays.  Offsets are 32-bit
// (a field of the largest supported grid, (4097 + 3)^2 x 8 values, is below
// 2^28), so each address is one 32-bit multiply-add and the loads take a
// scalar base plus a 32-bit vector offset.
struct Dims {
  int g, L;
  __device__ __host__ unsigned nd(int i, int j, int l) const {
    return (unsigned)(i - 1) + (unsigned)(g + 2) * ((unsigned)(j - 1) + (unsigned)(g + 2) * (unsigned)l);
  }
  __device__ __host__ unsigned fi(int i, int j, int l) const {
    return (unsigned)(i - 1) + (unsigned)(g + 3) * ((unsigned)(j - 1) + (unsigned)(g + 2) * (unsigned)l);
  }
  __device__ __host__ unsigned fj(int i, int j, int l) const {
    return (unsigned)(i - 1) + (unsigned)(g + 2) * ((unsigned)(j - 1) + (unsigned)(g + 3) * (unsigned)l);
  }
  __device__ __host__ unsigned rs(int i, int j, int l) const {
    return (unsigned)(i - 1) + (unsigned)g * ((unsigned)(j - 1) + (unsigned)g * (unsigned)l);
  }
  size_t n_nd() const { return (size_t)(g + 2) * (g + 2) * L; }
  size_t n_fi() const { return (size_t)(g + 3) * (g + 2) * L; }
  size_t n_fj() const { return (size_t)(g + 2) * (g + 3) * L; }
  size_t n_rs() const { return (size_t)g * g * L; }
};

#ifndef PQ_CAB_BX
#define PQ_CAB_BX 64
#define PQ_CAB_BY 4
#endif
constexpr int BX = PQ_CAB_BX, BY = PQ_CAB_BY;

inline dim3 grid_for(int ni, int nj, int L) {
  return dim3((ni + BX - 1) / BX, (nj + BY - 1) / BY, L);
}

// Workgroups are dealt to the 8 XCDs in turn (dispatch id mod 8), and each
// XCD has its own L2.  The dispatch id is remapped so that every XCD takes
// one contiguous run of tiles, ordered layer fastest, then i, then j: the
// stencil rows a tile shares with the next tile in j, and the neighbouring
// layers' values, are then read through the same L2.
__device__ __forceinline__ void pq_tile(int &bx, int &by, int &bz) {
  const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const unsigned nb = gx * gy * gz;
  const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned q = nb / 8u, r = nb % 8u, x = id % 8u, k = id / 8u;
  const unsigned lid = (x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q) + k;
  bz = (int)(lid % gz);
  const unsigned rest = lid / gz;
  bx = (int)(rest % gx);
  by = (int)(rest / gx);
}

// Thread -> (i, j, l) over i in [i0, i0+ni), j in [j0, j0+nj).
#define PQ_IDX(i0, ni, j0, nj)                                                  \
  int bx_, by_, l;                                                              \
  pq_tile(bx_, by_, l);                                                         \
  const int i = (i0) + bx_ * BX + (int)threadIdx.x;                             \
  const int j = (j0) + by_ * BY + (int)threadIdx.y;                             \
  if (i >= (i0) + (ni) || j >= (j0) + (nj)) return;

// Field accessors: the model's fields are stored as TS (double, or float
// for the FP32-storage model); every read converts to double and every
// write rounds to TS, so the arithmetic is FP64 either way (and for TS =
// double the code is the plain FP64 code).
template <typename TS> struct R {
  const TS *__restrict__ p;
  __host__ __device__ R(const TS *q) : p(q) {}
  __device__ __forceinline__ double operator[](unsigned k) const { return double(p[k]); }
};
template <typename TS> struct W {
  TS *__restrict__ p;
  __host__ __device__ W(TS *q) : p(q) {}
  struct Ref {
    TS *q;
    __device__ __forceinline__ Ref &operator=(double v) {
      *q = TS(v);
      return *this;
    }
    __device__ __forceinline__ operator double() const { return double(*q); }
  };
  __device__ __forceinline__ Ref operator[](unsigned k) const { return Ref{p + k}; }
};

// cabaret_predictor (qg3p_aux.F90:993-1020): f range 3..g
template <typename TS>
__global__ void k_predictor(Dims d, double dt, R<TS> zeta_old,
                            R<TS> res, W<TS> zeta) {
  PQ_IDX(3, d.g - 2, 3, d.g - 2);
  zeta[d.nd(i, j, l)] = zeta_old[d.nd(i, j, l)] + 0.5 * dt * res[d.rs(i - 1, j - 1, l)];
}

// cabaret_beta (qg3p_aux.F90:1023-1061): l range 2..g-1 of sdata
template <typename TS>
__global__ void k_beta(Dims d, double dt, double tw_a, const double *__restrict__ beta,
                       R<TS> ui, W<TS> zeta,
                       W<TS> res_beta) {
  PQ_IDX(2, d.g - 2, 2, d.g - 2);
  const double rb = dt * (0.5 * (ui[d.fi(i + 2, j + 1, l)] + ui[d.fi(i + 1, j + 1, l)]) * beta[l]);
  const unsigned z = d.nd(i + 1, j + 1, l);
  zeta[z] = zeta[z] + (1.0 + tw_a) * rb - tw_a * res_beta[d.rs(i, j, l)];
  res_beta[d.rs(i, j, l)] = rb;
}

// cabaret_forcing with WIND_FORCING (qg3p_aux.F90:1065-1164)
template <typename TS>
__global__ void k_forcing(Dims d, double dt, const double *__restrict__ fw, double visc,
                          double visc_bot, R<TS> rel,
                          W<TS> zeta) {
  PQ_IDX(3, d.g - 2, 3, d.g - 2);
  const int L = d.L;
  const double lap = -4.0 * rel[d.nd(i, j, l)] + rel[d.nd(i - 1, j, l)] + rel[d.nd(i + 1, j, l)] +
                     rel[d.nd(i, j - 1, l)] + rel[d.nd(i, j + 1, l)];
  const double f = fw[(unsigned)(i - 2) + (unsigned)d.g * (unsigned)(j - 2)];
  double rv;
  if (L == 1)
    rv = dt * (f - visc_bot * rel[d.nd(i, j, 0)] + visc * lap);
  else if (l == 0)
    rv = dt * (f + visc * lap);
  else if (l < L - 1)
    rv = dt * visc * lap;
  else
    rv = dt * (-visc_bot * rel[d.nd(i, j, l)] + visc * lap);
  zeta[d.nd(i, j, l)] = zeta[d.nd(i, j, l)] + rv;
}

// ---- centred (leapfrog) scheme, the driver without CABARET
// zeta_centred (qg3p_aux.F90:2027-2202) and zeta_tendency /
// zeta_forward_euler (qg3p_aux.F90:1789-2024) share one form on the interior
// nodes 3..G: dst = src + T F with
//   F = 0.25 J(psi, q) + force_wind (top) + viscosity Lap(rel_p) - visc_bot rel_p (bottom),
// q = beta_y + rel + the coupling to the neighbouring layers'
// stream functions, beta_y(i) = beta ((G + 1) / 2 - (i - 1)) (qg3.F90:302-307).
// One layer keeps visc_bot inside the viscosity bracket, as the reference.
// zeta_centred: T = 2 dt, psi and rel of level n, rel_p of level n-1, src =
// dst = zeta of level n-1; forward Euler: T = dt, one level, dst the other.
template <typename TS>
__global__ void k_zeta_step(Dims d, double T, const double *__restrict__ beta,
                            const double *__restrict__ s1, const double *__restrict__ s2,
                            const double *__restrict__ fw, double visc, double visc_bot,
                            R<TS> psi, R<TS> rel,
                            R<TS> rel_p, R<TS> src, W<TS> dst) {
  PQ_IDX(3, d.g - 2, 3, d.g - 2);
  const int L = d.L;
  const double hb = ((double)d.g + 1.0) / 2.0;
  auto q = [&](int ii, int jj) {
    const double by = beta[l] * (hb - (double)(ii - 1));
    double v = by + rel[d.nd(ii, jj, l)];
    if (L > 1) {
      if (l == 0) v = v + s2[0] * psi[d.nd(ii, jj, 1)];
      else if (l == L - 1) v = v + s1[l] * psi[d.nd(ii, jj, l - 1)];
      else v = v + s1[l] * psi[d.nd(ii, jj, l - 1)] + s2[l] * psi[d.nd(ii, jj, l + 1)];
    }
    return v;
  };
  auto P = [&](int ii, int jj) { return psi[d.nd(ii, jj, l)]; };
  const double jac = q(i, j + 1) * (P(i - 1, j + 1) - P(i + 1, j + 1)) -
                     q(i, j - 1) * (P(i - 1, j - 1) - P(i + 1, j - 1)) -
                     q(i - 1, j) * (P(i - 1, j + 1) - P(i - 1, j - 1)) +
                     q(i + 1, j) * (P(i + 1, j + 1) - P(i + 1, j - 1));
  const double lap = -4.0 * rel_p[d.nd(i, j, l)] + rel_p[d.nd(i - 1, j, l)] +
                     rel_p[d.nd(i + 1, j, l)] + rel_p[d.nd(i, j - 1, l)] + rel_p[d.nd(i, j + 1, l)];
  const double f0 = fw[(unsigned)(i - 2) + (unsigned)d.g * (unsigned)(j - 2)];
  double f;
  if (L == 1)
    f = 0.25 * jac + f0 + visc * (lap - visc_bot * rel_p[d.nd(i, j, 0)]);
  else if (l == 0)
    f = 0.25 * jac + f0 + visc * lap;
  else if (l < L - 1)
    f = 0.25 * jac + visc * lap;
  else
    f = 0.25 * jac + visc * lap - visc_bot * rel_p[d.nd(i, j, l)];
  dst[d.nd(i, j, l)] = src[d.nd(i, j, l)] + T * f;
}

// robert_asselin_filter with RA_DOUBLE_FILTER (Williams' filter, nu = 0.02,
// alpha = 0.5; qg3p_aux.F90:44-62, 1716-1786) on the owned nodes 2..G+1:
// phi_l = level n+1, phi_m = level n, phi_old_l = level n-1.
template <typename TS>
__global__ void k_ra_filter(Dims d, W<TS> phi_l, W<TS> phi_m,
                            R<TS> phi_old_l, W<TS> phi_old_m) {
  PQ_IDX(2, d.g, 2, d.g);
  const double ra_nu = 0.02, ra_alpha = 0.5;
  const double w21 = 0.5 * ra_nu * ra_alpha, w22 = 1.0 - 2.0 * w21;
  const double w11 = -0.5 * ra_nu * (1.0 - ra_alpha), w12 = -2.0 * w11, w13 = 1.0 + w11;
  const unsigned k = d.nd(i, j, l);
  const double lphi_l = phi_l[k], old = phi_old_l[k], pm = phi_m[k];
  phi_l[k] = w11 * old + w12 * pm + w13 * lphi_l;
  const double nm = w21 * (old + lphi_l) + w22 * pm;
  phi_m[k] = nm;
  phi_old_m[k] = nm;
}

template <typename A>
__device__ __forceinline__ double psi_edge(const Dims &d, A psi, int i, int j, int l) {
  return 0.25 * (psi[d.nd(i, j, l)] + psi[d.nd(i, j - 1, l)] + psi[d.nd(i - 1, j, l)] +
                 psi[d.nd(i - 1, j - 1, l)]);
}

// cabaret_cell_boundaries_time_integrated (qg3p_aux.F90:1235-1313).  One
// thread per (i, j) in 3..g+1 x 3..g+1 covers both face sets.
template <typename TS>
__global__ void k_cell_boundaries(Dims d, double tw, R<TS> psi,
                                  W<TS> ui, W<TS> uj,
                                  W<TS> uimo, W<TS> ujmo) {
  PQ_IDX(3, d.g - 1, 3, d.g - 1);
  const int g = d.g;
  const double pe = psi_edge(d, psi, i, j, l);
  if (j <= g) {
    const double uim = -psi_edge(d, psi, i, j + 1, l) + pe;
    const unsigned a = d.fi(i, j, l);
    ui[a] = (1.0 + tw) * uim - tw * uimo[a];
    uimo[a] = uim;
  }
  if (i <= g) {
    const double ujm = +psi_edge(d, psi, i + 1, j, l) - pe;
    const unsigned a = d.fj(i, j, l);
    uj[a] = (1.0 + tw) * ujm - tw * ujmo[a];
    ujmo[a] = ujm;
  }
}

// cabaret_cell_boundaries_velocity (qg3p_aux.F90:1168-1231)
template <typename TS>
__global__ void k_cell_boundaries_velocity(Dims d, R<TS> psi,
                                           W<TS> ui, W<TS> uj) {
  PQ_IDX(3, d.g - 1, 3, d.g - 1);
  const int g = d.g;
  const double pe = psi_edge(d, psi, i, j, l);
  if (j <= g) ui[d.fi(i, j, l)] = -psi_edge(d, psi, i, j + 1, l) + pe;
  if (i <= g) uj[d.fj(i, j, l)] = +psi_edge(d, psi, i + 1, j, l) - pe;
}

__device__ __forceinline__ double max3(double a, double b, double c) {
  const double m = a > b ? a : b;
  return m > c ? m : c;
}
__device__ __forceinline__ double min3(double a, double b, double c) {
  const double m = a < b ? a : b;
  return m < c ? m : c;
}

// One face of cabaret_extrapolator with CABARET_LIMITER
// (qg3p_aux.F90:1362-1404 and 1450-1484).  lo/hi: is the upwind neighbour
// face beyond the boundary (fb range test on a single domain).
__device__ __forceinline__ double extrap_face(double dt, double u1, double fi1, double u_m,
                                              double f_m, double u_p, double f_p,
                                              double zn_m, double zo_m, double zn_p,
                                              double zo_p, bool bnd_lo, bool bnd_hi) {
  double fi2, psi_n, d_f;
  if (u1 >= 0.0) {
    if (!bnd_lo) {
      fi2 = f_m;
      const double v0 = (u1 + u_m) / 2.0;
      psi_n = zn_m;
      d_f = -(psi_n - zo_m) * 2.0 - dt * v0 * (fi1 - fi2);
    } else {
      fi2 = zn_m;
      psi_n = zn_m;
      d_f = -(psi_n - zo_m);
    }
  } else {
    if (!bnd_hi) {
      fi2 = f_p;
      const double v0 = (u1 + u_p) / 2.0;
      psi_n = zn_p;
      d_f = -(psi_n - zo_p) * 2.0 - dt * v0 * (fi2 - fi1);
    } else {
      fi2 = zn_p;
      psi_n = zn_p;
      d_f = -(psi_n - zo_p);
    }
  }
  double f = 2.0 * psi_n - fi2;
  const double fmx = max3(fi1, fi2, psi_n) - d_f;
  const double fmn = min3(fi1, fi2, psi_n) - d_f;
  if (f > fmx) f = fmx;
  if (f < fmn) f = fmn;
  return f;
}

// cabaret_extrapolator, new face values into fin / fjn.
template <typename TS>
__global__ void k_extrapolator(Dims d, double dt, R<TS> zeta,
                               R<TS> zeta_old, R<TS> ui,
                               R<TS> uj, R<TS> zfi,
                               R<TS> zfj, W<TS> fin,
                               W<TS> fjn) {
  PQ_IDX(3, d.g - 1, 3, d.g - 1);
  const int g = d.g;
  if (j <= g) {  // i-face (i, j), i in 3..g+1
    const unsigned a = d.fi(i, j, l);
    const bool lo = !(i > 3), hi = !(i < g + 1);
    const double u1 = ui[a], fi1 = zfi[a];
    const double um = lo ? 0.0 : ui[d.fi(i - 1, j, l)], fm = lo ? 0.0 : zfi[d.fi(i - 1, j, l)];
    const double up = hi ? 0.0 : ui[d.fi(i + 1, j, l)], fp = hi ? 0.0 : zfi[d.fi(i + 1, j, l)];
    fin[a] = extrap_face(dt, u1, fi1, um, fm, up, fp, zeta[d.nd(i - 1, j, l)],
                         zeta_old[d.nd(i - 1, j, l)], zeta[d.nd(i, j, l)],
                         zeta_old[d.nd(i, j, l)], lo, hi);
  }
  if (i <= g) {  // j-face (i, j), j in 3..g+1
    const unsigned a = d.fj(i, j, l);
    const bool lo = !(j > 3), hi = !(j < g + 1);
    const double u1 = uj[a], fi1 = zfj[a];
    const double um = lo ? 0.0 : uj[d.fj(i, j - 1, l)], fm = lo ? 0.0 : zfj[d.fj(i, j - 1, l)];
    const double up = hi ? 0.0 : uj[d.fj(i, j + 1, l)], fp = hi ? 0.0 : zfj[d.fj(i, j + 1, l)];
    fjn[a] = extrap_face(dt, u1, fi1, um, fm, up, fp, zeta[d.nd(i, j - 1, l)],
                         zeta_old[d.nd(i, j - 1, l)], zeta[d.nd(i, j, l)],
                         zeta_old[d.nd(i, j, l)], lo, hi);
  }
}

// Copy the extrapolated faces back (qg3p_aux.F90:1526-1543).
template <typename TS>
__global__ void k_face_copy(Dims d, R<TS> fin, R<TS> fjn,
                            W<TS> zfi, W<TS> zfj) {
  PQ_IDX(3, d.g - 1, 3, d.g - 1);
  if (j <= d.g) zfi[d.fi(i, j, l)] = fin[d.fi(i, j, l)];
  if (i <= d.g) zfj[d.fj(i, j, l)] = fjn[d.fj(i, j, l)];
}

// cabaret_flux (qg3p_aux.F90:1547-1594): i-faces 2..g+2 x 3..g, j-faces
// 3..g x 2..g+2.
template <typename TS>
__global__ void k_flux(Dims d, R<TS> ui, R<TS> uj,
                       R<TS> zfi, R<TS> zfj,
                       W<TS> flux_i, W<TS> flux_j) {
  PQ_IDX(2, d.g + 1, 2, d.g + 1);
  const int g = d.g;
  if (j >= 3 && j <= g) flux_i[d.fi(i, j, l)] = ui[d.fi(i, j, l)] * zfi[d.fi(i, j, l)];
  if (i >= 3 && i <= g) flux_j[d.fj(i, j, l)] = uj[d.fj(i, j, l)] * zfj[d.fj(i, j, l)];
}

// cabaret_corrector (qg3p_aux.F90:1597-1642)
template <typename TS>
__global__ void k_corrector(Dims d, double dt, R<TS> flux_i,
                            R<TS> flux_j, W<TS> res,
                            W<TS> zeta) {
  PQ_IDX(3, d.g - 2, 3, d.g - 2);
  const double r = -(flux_i[d.fi(i + 1, j, l)] - flux_i[d.fi(i, j, l)] + flux_j[d.fj(i, j + 1, l)] -
                     flux_j[d.fj(i, j, l)]);
  res[d.rs(i - 1, j - 1, l)] = r;
  zeta[d.nd(i, j, l)] = zeta[d.nd(i, j, l)] + 0.5 * dt * r;
}

// cabaret_corrector writing the corrected field into zeta_out (the step's
// zeta_old for the next step) from zeta_in; boundary values are copied.
// Covers all owned nodes 2..g+1.
template <typename TS>
__global__ void k_corrector_to(Dims d, double dt, R<TS> flux_i,
                               R<TS> flux_j, W<TS> res,
                               R<TS> zeta_in, W<TS> zeta_out) {
  PQ_IDX(1, d.g + 2, 1, d.g + 2);
  const int g = d.g;
  const unsigned a = d.nd(i, j, l);
  if (i >= 3 && i <= g && j >= 3 && j <= g) {
    const double r = -(flux_i[d.fi(i + 1, j, l)] - flux_i[d.fi(i, j, l)] +
                       flux_j[d.fj(i, j + 1, l)] - flux_j[d.fj(i, j, l)]);
    res[d.rs(i - 1, j - 1, l)] = r;
    zeta_out[a] = zeta_in[a] + 0.5 * dt * r;
  } else {
    zeta_out[a] = zeta_in[a];
  }
}

// cabaret_extrapolator, the copy of the new face values and cabaret_flux
// of one model step in one pass: thread (i, j) over 2..g+2 extrapolates
// i-face (i, j) (i in 3..g+1, j in 3..g) and j-face (i, j) into fin / fjn,
// and writes the fluxes of the cabaret_flux ranges with the new face values
// (the boundary faces i or j = 2, g+2 keep theirs).  zeta_fi / zeta_fj are
// only read here (k_face_copy copies fin / fjn back); the model runs this
// pass only in a final partial step.
template <typename TS>
__global__ void k_post_extrap_flux(Dims d, double dt, R<TS> zeta, R<TS> zeta_old, R<TS> ui,
                                   R<TS> uj, R<TS> zfi, R<TS> zfj, W<TS> fin, W<TS> fjn,
                                   W<TS> flux_i, W<TS> flux_j) {
  PQ_IDX(2, d.g + 1, 2, d.g + 1);
  const int g = d.g;
  if (j >= 3 && j <= g) {  // i-face (i, j), i in 2..g+2
    const unsigned a = d.fi(i, j, l);
    double f;
    if (i >= 3 && i <= g + 1) {
      const bool lo = !(i > 3), hi = !(i < g + 1);
      const double u1 = ui[a], fi1 = zfi[a];
      const double um = lo ? 0.0 : ui[d.fi(i - 1, j, l)], fm = lo ? 0.0 : zfi[d.fi(i - 1, j, l)];
      const double up = hi ? 0.0 : ui[d.fi(i + 1, j, l)], fp = hi ? 0.0 : zfi[d.fi(i + 1, j, l)];
      const TS v = TS(extrap_face(dt, u1, fi1, um, fm, up, fp, zeta[d.nd(i - 1, j, l)],
                                  zeta_old[d.nd(i - 1, j, l)], zeta[d.nd(i, j, l)],
                                  zeta_old[d.nd(i, j, l)], lo, hi));
      fin[a] = double(v);
      f = double(v);
    } else {
      f = zfi[a];
    }
    flux_i[a] = ui[a] * f;
  }
  if (i >= 3 && i <= g) {  // j-face (i, j), j in 2..g+2
    const unsigned a = d.fj(i, j, l);
    double f;
    if (j >= 3 && j <= g + 1) {
      const bool lo = !(j > 3), hi = !(j < g + 1);
      const double u1 = uj[a], fi1 = zfj[a];
      const double um = lo ? 0.0 : uj[d.fj(i, j - 1, l)], fm = lo ? 0.0 : zfj[d.fj(i, j - 1, l)];
      const double up = hi ? 0.0 : uj[d.fj(i, j + 1, l)], fp = hi ? 0.0 : zfj[d.fj(i, j + 1, l)];
      const TS v = TS(extrap_face(dt, u1, fi1, um, fm, up, fp, zeta[d.nd(i, j - 1, l)],
                                  zeta_old[d.nd(i, j - 1, l)], zeta[d.nd(i, j, l)],
                                  zeta_old[d.nd(i, j, l)], lo, hi));
      fjn[a] = double(v);
      f = double(v);
    } else {
      f = zfj[a];
    }
    flux_j[a] = uj[a] * f;
  }
}

// rel_from_zeta_and_psi + apply_rel_bc (qg3p_aux.F90:738-915) over owned
// nodes 2..g+1.  Boundary nodes take the boundary formula; on corners the
// left/right (j) edge formula wins, as it is applied last in the reference.
// apply_rel_bc (qg3p_aux.F90:778-915) at a boundary node (i, j) of layer l.
template <typename A>
__device__ __forceinline__ double rel_bc(const Dims &d, int bc_type, double alpha, A psi, int i,
                                         int j, int l) {
  const int g = d.g;
  const unsigned a = d.nd(i, j, l);
  const bool jb = (j == 2 || j == g + 1);
  double r;
  if (bc_type == 1) {
    const double cff = alpha / (alpha + 3.0);
    if (jb) {
      const int s = (j == 2) ? 1 : -1;
      r = cff * (-3.5 * psi[a] + 4.0 * psi[d.nd(i, j + s, l)] - 0.5 * psi[d.nd(i, j + 2 * s, l)]);
    } else {
      const int s = (i == 2) ? 1 : -1;
      r = cff * (-3.5 * psi[a] + 4.0 * psi[d.nd(i + s, j, l)] - 0.5 * psi[d.nd(i + 2 * s, j, l)]);
    }
  } else if (bc_type == 2) {
    if (jb) {
      r = (j == 2) ? 2.0 * (psi[d.nd(i, 3, l)] - psi[a]) : 2.0 * (psi[d.nd(i, g, l)] - psi[a]);
    } else {
      r = (i == 2) ? 2.0 * (psi[d.nd(3, j, l)] - psi[a]) : 2.0 * (psi[d.nd(g, j, l)] - psi[a]);
    }
  } else {
    r = 0.0;
  }
  return r;
}

// rel of owned node (i, j) of layer l.
template <typename TS>
__device__ __forceinline__ double rel_node(const Dims &d, const double *__restrict__ s1,
                                           const double *__restrict__ s2, int bc_type, double alpha,
                                           R<TS> zeta, R<TS> psi, int i, int j, int l) {
  const int g = d.g, L = d.L;
  const unsigned a = d.nd(i, j, l);
  double r;
  const bool jb = (j == 2 || j == g + 1), ib = (i == 2 || i == g + 1);
  if (jb || ib) {
    r = rel_bc(d, bc_type, alpha, psi, i, j, l);
  } else if (L == 1) {
    r = zeta[a];
  } else if (l == 0) {
    r = zeta[a] + s2[0] * (psi[a] - psi[d.nd(i, j, 1)]);
  } else if (l < L - 1) {
    r = zeta[a] + s1[l] * (psi[a] - psi[d.nd(i, j, l - 1)]) + s2[l] * (psi[a] - psi[d.nd(i, j, l + 1)]);
  } else {
    r = zeta[a] + s1[L - 1] * (psi[a] - psi[d.nd(i, j, L - 2)]);
  }
  return r;
}

template <typename TS>
__global__ void k_rel(Dims d, const double *__restrict__ s1, const double *__restrict__ s2,
                      int bc_type, double alpha, R<TS> zeta,
                      R<TS> psi, W<TS> rel) {
  PQ_IDX(2, d.g, 2, d.g);
  rel[d.nd(i, j, l)] = rel_node(d, s1, s2, bc_type, alpha, zeta, psi, i, j, l);
}

// rel_from_psi (qg3p_aux.F90:579-603): the 5-point Laplacian on the
// interior (the reference's term order), apply_rel_bc on the boundary.
template <typename TS>
__global__ void k_rel_from_psi(Dims d, int bc_type, double alpha, R<TS> psi,
                               W<TS> rel) {
  PQ_IDX(2, d.g, 2, d.g);
  const int g = d.g;
  const unsigned a = d.nd(i, j, l);
  if (i == 2 || i == g + 1 || j == 2 || j == g + 1)
    rel[a] = rel_bc(d, bc_type, alpha, psi, i, j, l);
  else
    rel[a] = -4.0 * psi[a] + psi[d.nd(i - 1, j, l)] + psi[d.nd(i, j + 1, l)] +
             psi[d.nd(i + 1, j, l)] + psi[d.nd(i, j - 1, l)];
}

// update_zeta_bc (qg3p_aux.F90:643-734) at boundary node (i, j) of layer l
// from its rel value.
template <typename TS>
__device__ __forceinline__ double zeta_bc_node(const Dims &d, const double *__restrict__ s1,
                                               const double *__restrict__ s2, double rel, R<TS> psi,
                                               int i, int j, int l) {
  const int L = d.L;
  const unsigned a = d.nd(i, j, l);
  double z;
  if (L == 1) z = rel;
  else if (l == 0) z = rel - s2[0] * (psi[a] - psi[d.nd(i, j, 1)]);
  else if (l == L - 1) z = rel - s1[L - 1] * (psi[a] - psi[d.nd(i, j, L - 2)]);
  else z = rel - s1[l] * (psi[a] - psi[d.nd(i, j, l - 1)]) - s2[l] * (psi[a] - psi[d.nd(i, j, l + 1)]);
  return z;
}

// update_zeta_bc: boundary nodes only.
template <typename TS>
__global__ void k_zeta_bc(Dims d, const double *__restrict__ s1, const double *__restrict__ s2,
                          R<TS> rel, R<TS> psi,
                          W<TS> zeta) {
  PQ_IDX(2, d.g, 2, d.g);
  const int g = d.g;
  if (!(i == 2 || i == g + 1 || j == 2 || j == g + 1)) return;
  const unsigned a = d.nd(i, j, l);
  zeta[a] = zeta_bc_node(d, s1, s2, rel[a], psi, i, j, l);
}

// rel_from_zeta_and_psi, update_zeta_bc and
// cabaret_cell_boundaries_time_integrated of one model step in one pass:
// thread (i, j) over 1..g+2 writes rel and (on the boundary) zeta of node
// (i, j) and the i- and j-face velocities of (i, j).  rel of a one-layer
// model is zeta on the halo ring as well (set(rel, zeta) copies the whole
// data array).  Same arithmetic as k_rel, k_zeta_bc and k_cell_boundaries;
// zeta is read only at the thread's own node.
template <typename TS>
__global__ void k_post_rel_cells(Dims d, const double *__restrict__ s1,
                                 const double *__restrict__ s2, int bc_type, double alpha,
                                 double tw, R<TS> psi, W<TS> zeta, W<TS> rel, W<TS> ui, W<TS> uj,
                                 W<TS> uimo, W<TS> ujmo) {
  PQ_IDX(1, d.g + 2, 1, d.g + 2);
  const int g = d.g;
  const unsigned a = d.nd(i, j, l);
  if (i >= 2 && i <= g + 1 && j >= 2 && j <= g + 1) {
    const TS r = TS(rel_node(d, s1, s2, bc_type, alpha, R<TS>(zeta.p), psi, i, j, l));
    rel[a] = double(r);
    if (i == 2 || i == g + 1 || j == 2 || j == g + 1)
      zeta[a] = zeta_bc_node(d, s1, s2, double(r), psi, i, j, l);
  } else if (d.L == 1) {
    rel[a] = zeta[a];
  }
  if (i >= 3 && i <= g + 1 && j >= 3 && j <= g + 1) {
    const double pe = psi_edge(d, psi, i, j, l);
    if (j <= g) {
      const double uim = -psi_edge(d, psi, i, j + 1, l) + pe;
      const unsigned f = d.fi(i, j, l);
      ui[f] = (1.0 + tw) * uim - tw * uimo[f];
      uimo[f] = uim;
    }
    if (i <= g) {
      const double ujm = +psi_edge(d, psi, i + 1, j, l) - pe;
      const unsigned f = d.fj(i, j, l);
      uj[f] = (1.0 + tw) * ujm - tw * ujmo[f];
      ujmo[f] = ujm;
    }
  }
}

// ---- the model's tiled step kernels
// The model keeps zeta_old and the face values zeta_fi / zeta_fj in two
// buffers each and a device word `par` that names the current one: the
// corrector pass reads the current values around each tile (a one-node /
// one-face halo) and writes the new ones into the other buffer, so that no
// workgroup overwrites what a neighbouring one still reads.  k_step_rel_force
// flips par (it reads none of these fields); the predictor and every pass
// after k_step_rel_force read it on the device, so a captured step replays
// correctly any number of times.
template <typename TS> struct Pair {
  TS *p[2];
};

// Tile of the step kernels: TX nodes along i (one per thread column) by TY
// rows of j, 64 x 4 threads, layer blockIdx.z (remapped by pq_tile).
constexpr int TX = 64, TYS = 4;
static_assert(TX == BX && BX * BY == 256, "the step kernels take one thread column per node column");
inline dim3 step_grid(int g, int L) { return dim3((g + 2 + TX - 1) / TX, (g + 2 + TYS - 1) / TYS, L); }

// cabaret_predictor followed by cabaret_beta (qg3p_aux.F90:993-1061) in
// one pass, with zeta_old read from the current buffer: the beta update of
// data node (i, j) only reads that node's predicted value.
template <typename TS>
__global__ void k_step_predictor(Dims d, double dt, double tw_a, const double *__restrict__ beta,
                                 Pair<TS> zo, const int *__restrict__ par, R<TS> res, R<TS> ui,
                                 W<TS> zeta, W<TS> res_beta) {
  PQ_IDX(3, d.g - 2, 3, d.g - 2);
  const R<TS> zeta_old(zo.p[*par]);
  double z = zeta_old[d.nd(i, j, l)] + 0.5 * dt * res[d.rs(i - 1, j - 1, l)];
  const double rb = dt * (0.5 * (ui[d.fi(i + 1, j, l)] + ui[d.fi(i, j, l)]) * beta[l]);
  z = z + (1.0 + tw_a) * rb - tw_a * res_beta[d.rs(i - 1, j - 1, l)];
  zeta[d.nd(i, j, l)] = z;
  res_beta[d.rs(i - 1, j - 1, l)] = rb;
}

// rel_from_zeta_and_psi + update_zeta_bc + cabaret_forcing +
// cabaret_cell_boundaries_time_integrated of one model step (the
// statements of k_post_rel_cells and k_forcing, in their order and
// rounding) over a TX x TYS tile of nodes 1..g+2.  psi and zeta of the tile
// and a one-node halo are staged in LDS; rel of the owned nodes among them
// is formed there (the halo's recomputed, as this pass writes neither psi
// nor zeta), so the forcing's Laplacian and the cell-boundary velocities
// read LDS only.  The forced zeta (boundary nodes: update_zeta_bc's value,
// halo ring: zeta) goes to zf; zeta itself keeps the predicted values.
template <typename TS>
__global__ void __launch_bounds__(256)
k_step_rel_force(Dims d, const double *__restrict__ s1, const double *__restrict__ s2, int bc_type,
                 double alpha, double tw, double dt, const double *__restrict__ fw, double visc,
                 double visc_bot, R<TS> psi, R<TS> zeta, W<TS> rel, W<TS> ui, W<TS> uj, W<TS> uimo,
                 W<TS> ujmo, W<TS> zf, int *__restrict__ par) {
  constexpr int HX = TX + 2, HY = TYS + 2;
  __shared__ TS s_psi[HY][HX];
  __shared__ TS s_zeta[HY][HX];
  __shared__ double rs[HY][HX];
  int bx, by, l;
  pq_tile(bx, by, l);
  const int g = d.g, L = d.L;
  const int i0 = 1 + bx * TX, j0 = 1 + by * TYS;
  const int tid = threadIdx.x + BX * threadIdx.y;
  if (tid == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) *par = 1 - *par;
  // LDS (b, a) holds node (i0 - 1 + a, j0 - 1 + b)
  for (int s = tid; s < HX * HY; s += BX * BY) {
    const int a = s % HX, b = s / HX;
    const int i = i0 - 1 + a, j = j0 - 1 + b;
    TS pv = TS(0), zv = TS(0);
    if (i >= 1 && i <= g + 2 && j >= 1 && j <= g + 2) {
      const unsigned k = d.nd(i, j, l);
      pv = psi.p[k];
      zv = zeta.p[k];
    }
    s_psi[b][a] = pv;
    s_zeta[b][a] = zv;
  }
  __syncthreads();
  // rel (rel_node's statements) of the owned nodes, rounded to TS
  for (int s = tid; s < HX * HY; s += BX * BY) {
    const int a = s % HX, b = s / HX;
    const int i = i0 - 1 + a, j = j0 - 1 + b;
    double r = 0.0;
    if (i >= 2 && i <= g + 1 && j >= 2 && j <= g + 1) {
      if (i == 2 || i == g + 1 || j == 2 || j == g + 1) {
        r = rel_bc(d, bc_type, alpha, psi, i, j, l);
      } else {
        const double z = s_zeta[b][a], pa = s_psi[b][a];
        if (L == 1)
          r = z;
        else if (l == 0)
          r = z + s2[0] * (pa - psi[d.nd(i, j, 1)]);
        else if (l < L - 1)
          r = z + s1[l] * (pa - psi[d.nd(i, j, l - 1)]) + s2[l] * (pa - psi[d.nd(i, j, l + 1)]);
        else
          r = z + s1[L - 1] * (pa - psi[d.nd(i, j, L - 2)]);
      }
      r = double(TS(r));
    }
    rs[b][a] = r;
  }
  __syncthreads();
  const int i = i0 + (int)threadIdx.x;
  if (i > g + 2) return;
  const int ca = threadIdx.x + 1;
  // psi at node (i + di, j + dj) of LDS row cb
  auto P = [&](int cb, int di, int dj) { return double(s_psi[cb + dj][ca + di]); };
  auto pedge = [&](int cb, int di, int dj) {  // psi_edge(i + di, j + dj)
    return 0.25 * (P(cb, di, dj) + P(cb, di, dj - 1) + P(cb, di - 1, dj) + P(cb, di - 1, dj - 1));
  };
  for (int b = threadIdx.y; b < TYS; b += BY) {
    const int j = j0 + b;
    if (j > g + 2) break;
    const unsigned a = d.nd(i, j, l);
    const int cb = b + 1;
    const double z = s_zeta[cb][ca];
    if (i >= 2 && i <= g + 1 && j >= 2 && j <= g + 1) {
      const double r = rs[cb][ca];
      rel[a] = r;
      if (i == 2 || i == g + 1 || j == 2 || j == g + 1) {
        zf[a] = zeta_bc_node(d, s1, s2, r, psi, i, j, l);
      } else {
        const double lap = -4.0 * rs[cb][ca] + rs[cb][ca - 1] + rs[cb][ca + 1] + rs[cb - 1][ca] +
                           rs[cb + 1][ca];
        const double f = fw[(unsigned)(i - 2) + (unsigned)g * (unsigned)(j - 2)];
        double rv;
        if (L == 1)
          rv = dt * (f - visc_bot * r + visc * lap);
        else if (l == 0)
          rv = dt * (f + visc * lap);
        else if (l < L - 1)
          rv = dt * visc * lap;
        else
          rv = dt * (-visc_bot * r + visc * lap);
        zf[a] = z + rv;
      }
    } else {
      if (L == 1) rel[a] = z;
      zf[a] = z;
    }
    if (i >= 3 && i <= g + 1 && j >= 3 && j <= g + 1) {
      const double pe = pedge(cb, 0, 0);
      if (j <= g) {
        const double uim = -pedge(cb, 0, 1) + pe;
        const unsigned f = d.fi(i, j, l);
        ui[f] = (1.0 + tw) * uim - tw * uimo[f];
        uimo[f] = uim;
      }
      if (i <= g) {
        const double ujm = +pedge(cb, 1, 0) - pe;
        const unsigned f = d.fj(i, j, l);
        uj[f] = (1.0 + tw) * ujm - tw * ujmo[f];
        ujmo[f] = ujm;
      }
    }
  }
}

// cabaret_extrapolator + cabaret_flux + cabaret_corrector + zeta_old <-
// zeta + the face copy of one model step (the statements of
// k_post_extrap_flux and k_post_correct, in their order and rounding) over
// a TX x TYS tile of nodes 1..g+2.  The forced zeta zf and the current
// zeta_old of the tile and a one-node halo, and the velocities and current
// face values the tile's faces read, are staged in LDS; the fluxes through
// the tile's faces (TX + 1 i-faces per row, TYS + 1 j-faces per column)
// are formed there.  The corrected zeta goes to zeta and the other zeta_old
// buffer, the tile's new face values (and the unchanged boundary faces the
// extrapolator reads) to the other face buffers.
template <typename TS>
__global__ void __launch_bounds__(256)
k_step_extrap_correct(Dims d, double dt, R<TS> zf, Pair<TS> zo, R<TS> ui, R<TS> uj, Pair<TS> zfi,
                      Pair<TS> zfj, W<TS> res, W<TS> zeta, const int *__restrict__ par) {
  constexpr int NX = TX + 2, NY = TYS + 2;  // nodes i0-1 .. i0+TX, j0-1 .. j0+TYS
  constexpr int IX = TX + 3;                // i-faces i0-1 .. i0+TX+1 (rows j0 .. j0+TYS-1)
  constexpr int JY = TYS + 3;               // j-faces j0-1 .. j0+TYS+1 (columns i0 .. i0+TX-1)
  __shared__ TS s_zf[NY][NX], s_zo[NY][NX];
  __shared__ TS s_ui[TYS][IX], s_fi[TYS][IX];
  __shared__ TS s_uj[JY][TX], s_fj[JY][TX];
  __shared__ TS fxi[TYS][TX + 1];  // fluxes (stored values: TS)
  __shared__ TS fxj[TYS + 1][TX];
  int bx, by, l;
  pq_tile(bx, by, l);
  const int g = d.g;
  const int i0 = 1 + bx * TX, j0 = 1 + by * TYS;
  const int tid = threadIdx.x + BX * threadIdx.y;
  const int pn = *par;  // flipped by k_step_rel_force: the buffers written here
  const TS *zold = zo.p[1 - pn], *fic = zfi.p[1 - pn], *fjc = zfj.p[1 - pn];
  const W<TS> zon(zo.p[pn]), fin(zfi.p[pn]), fjn(zfj.p[pn]);
  for (int s = tid; s < NX * NY; s += BX * BY) {
    const int a = s % NX, b = s / NX;
    const int i = i0 - 1 + a, j = j0 - 1 + b;
    TS x = TS(0), y = TS(0);
    if (i >= 1 && i <= g + 2 && j >= 1 && j <= g + 2) {
      const unsigned k = d.nd(i, j, l);
      x = zf.p[k];
      y = zold[k];
    }
    s_zf[b][a] = x;
    s_zo[b][a] = y;
  }
  for (int s = tid; s < IX * TYS; s += BX * BY) {
    const int a = s % IX, b = s / IX;
    const int i = i0 - 1 + a, j = j0 + b;
    TS x = TS(0), y = TS(0);
    if (i >= 1 && i <= g + 3 && j <= g + 2) {
      const unsigned k = d.fi(i, j, l);
      x = ui.p[k];
      y = fic[k];
    }
    s_ui[b][a] = x;
    s_fi[b][a] = y;
  }
  for (int s = tid; s < TX * JY; s += BX * BY) {
    const int a = s % TX, b = s / TX;
    const int i = i0 + a, j = j0 - 1 + b;
    TS x = TS(0), y = TS(0);
    if (i <= g + 2 && j >= 1 && j <= g + 3) {
      const unsigned k = d.fj(i, j, l);
      x = uj.p[k];
      y = fjc[k];
    }
    s_uj[b][a] = x;
    s_fj[b][a] = y;
  }
  __syncthreads();
  const int tx = threadIdx.x;
  // i-face i0 + a of row j0 + b: LDS face column a + 1, node columns a, a + 1
  auto iface = [&](int a, int b) {
    const int i = i0 + a, j = j0 + b;
    TS fl = TS(0);
    if (j >= 3 && j <= g && i >= 2 && i <= g + 2) {
      const int c = a + 1;
      double v;
      if (i >= 3 && i <= g + 1) {
        const bool lo = !(i > 3), hi = !(i < g + 1);
        const double u1 = s_ui[b][c], fi1 = s_fi[b][c];
        const double um = lo ? 0.0 : double(s_ui[b][c - 1]), fm = lo ? 0.0 : double(s_fi[b][c - 1]);
        const double up = hi ? 0.0 : double(s_ui[b][c + 1]), fp = hi ? 0.0 : double(s_fi[b][c + 1]);
        const TS t = TS(extrap_face(dt, u1, fi1, um, fm, up, fp, s_zf[b + 1][a], s_zo[b + 1][a],
                                    s_zf[b + 1][a + 1], s_zo[b + 1][a + 1], lo, hi));
        v = double(t);
      } else {
        v = s_fi[b][c];
      }
      if (a < TX) fin[d.fi(i, j, l)] = v;
      fl = TS(double(s_ui[b][c]) * v);
    }
    fxi[b][a] = fl;
  };
  // j-face j0 + b of column i0 + a: LDS face row b + 1, node rows b, b + 1
  auto jface = [&](int a, int b) {
    const int i = i0 + a, j = j0 + b;
    TS fl = TS(0);
    if (i >= 3 && i <= g && j >= 2 && j <= g + 2) {
      const int c = b + 1;
      double v;
      if (j >= 3 && j <= g + 1) {
        const bool lo = !(j > 3), hi = !(j < g + 1);
        const double u1 = s_uj[c][a], fi1 = s_fj[c][a];
        const double um = lo ? 0.0 : double(s_uj[c - 1][a]), fm = lo ? 0.0 : double(s_fj[c - 1][a]);
        const double up = hi ? 0.0 : double(s_uj[c + 1][a]), fp = hi ? 0.0 : double(s_fj[c + 1][a]);
        const TS t = TS(extrap_face(dt, u1, fi1, um, fm, up, fp, s_zf[b][a + 1], s_zo[b][a + 1],
                                    s_zf[b + 1][a + 1], s_zo[b + 1][a + 1], lo, hi));
        v = double(t);
      } else {
        v = s_fj[c][a];
      }
      if (b < TYS) fjn[d.fj(i, j, l)] = v;
      fl = TS(double(s_uj[c][a]) * v);
    }
    fxj[b][a] = fl;
  };
  for (int b = threadIdx.y; b < TYS; b += BY) {
    iface(tx, b);
    jface(tx, b);
  }
  if (tid < TYS) iface(TX, tid);                            // the tile's last i-face column
  else if (tid >= BX && tid < 2 * BX) jface(tid - BX, TYS);  // and last j-face row
  __syncthreads();
  const int i = i0 + tx;
  if (i > g + 2) return;
  for (int b = threadIdx.y; b < TYS; b += BY) {
    const int j = j0 + b;
    if (j > g + 2) break;
    const unsigned a = d.nd(i, j, l);
    const int c = threadIdx.x;
    double v;
    if (i >= 3 && i <= g && j >= 3 && j <= g) {
      const double r = -(double(fxi[b][c + 1]) - double(fxi[b][c]) + double(fxj[b + 1][c]) -
                         double(fxj[b][c]));
      res[d.rs(i - 1, j - 1, l)] = r;
      v = double(s_zf[b + 1][c + 1]) + 0.5 * dt * r;
    } else {
      v = s_zf[b + 1][c + 1];
    }
    const TS t = TS(v);
    zon[a] = double(t);
    zeta[a] = double(t);
  }
}

// max |u| over the cabaret_cfl face ranges (qg3p_aux.F90:937-940); the
// non-negative doubles are ordered like their bit patterns.
template <typename TS>
__global__ void k_maxabs(Dims d, R<TS> ui, R<TS> uj,
                         unsigned long long *__restrict__ out) {
  const int i = 3 + (int)(blockIdx.x * BX + threadIdx.x);
  const int j = 3 + (int)(blockIdx.y * BY + threadIdx.y);
  const int l = (int)blockIdx.z;
  const int g = d.g;
  double m = 0.0;
  if (i <= g + 1 && j <= g + 1) {
    if (j <= g) m = fmax(m, fabs(ui[d.fi(i, j, l)]));
    if (i <= g) m = fmax(m, fabs(uj[d.fj(i, j, l)]));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double sm[BX * BY / 64];
  const int tid = threadIdx.y * BX + threadIdx.x;
  if ((tid & 63) == 0) sm[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < BX * BY / 64; ++w) m = fmax(m, sm[w]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// Stage entry points have no handle: NULL and PQ_STREAM_NULL both mean the
// null stream.
hipStream_t S(void *s) { return pq_resolve_stream(s, (hipStream_t)0); }

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    pq_set_error("%s: kernel launch failed: %s", what, hipGetErrorString(e));
    return PQ_ERR_HIP;
  }
  return PQ_OK;
}

#define CHECK_GL(g, L, fn)                                                     \
  PQ_CHECK_ARG((g) >= 5 && (L) >= 1 && (L) <= 64, "%s: bad size g=%d L=%d", fn, g, L)

// Work buffers of the stage entry points that need one (the CFL maximum,
// and the extrapolator's new face values when the caller passes no
// scratch), one per (device, stream).  A buffer grows when a larger grid
// comes and is kept until the process ends, so a call allocates only when
// it needs more than any earlier call on its stream.  Work on one stream is
// ordered, so calls on a stream share its buffer safely; the growth frees
// the old buffer in stream order.
struct StageBuf {
  void *p = nullptr;
  size_t bytes = 0;
};
std::mutex g_stage_mu;
std::map<std::pair<int, hipStream_t>, StageBuf> g_stage_bufs;

int stage_buffer(hipStream_t s, size_t bytes, void **out) {
  int dev = 0;
  PQ_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_stage_mu);
  StageBuf &b = g_stage_bufs[std::make_pair(dev, s)];
  if (b.bytes < bytes) {
    if (b.p) PQ_HIP(hipFreeAsync(b.p, s));
    b.p = nullptr;
    b.bytes = 0;
    PQ_HIP(hipMallocAsync(&b.p, bytes, s));
    b.bytes = bytes;
  }
  *out = b.p;
  return PQ_OK;
}

// The extrapolator into a work area tmp (one i-face and one j-face array),
// then the new face values copied over zeta_fi / zeta_fj
// (cabaret_extrapolator, qg3p_aux.F90:1317-1544, whose zeta_fin / zeta_fjn
// are `save` arrays).  The copy reads exactly the faces the first kernel
// wrote, so tmp needs no clearing.
int extrapolator_launch(const Dims &d, double dt, const double *zeta, const double *zeta_old,
                        const double *ui, const double *uj, double *zfi, double *zfj, double *tmp,
                        hipStream_t s) {
  const int g = d.g;
  hipLaunchKernelGGL(k_extrapolator<double>, grid_for(g - 1, g - 1, d.L), dim3(BX, BY), 0, s, d, dt,
                     zeta, zeta_old, ui, uj, zfi, zfj, tmp, tmp + d.n_fi());
  hipLaunchKernelGGL(k_face_copy<double>, grid_for(g - 1, g - 1, d.L), dim3(BX, BY), 0, s, d, tmp,
                     tmp + d.n_fi(), zfi, zfj);
  return PQ_OK;
}

}  // namespace

// ---- work area of the entry point that allocates nothing
// pq_cabaret_extrapolator_noalloc_dev keeps the new face values here: a
// file-scope device array of kNoAllocDoubles doubles (64 MiB), which every
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in):
// it uses the device address that hipGetSymbolAddress returns, looked up
// once per device with that device current.
static constexpr size_t kNoAllocDoubles = size_t(1) << 23;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(double **out) {
  int dev = 0;
  PQ_HIP(hipGetDevice(&dev));
  PQ_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "pq_cabaret_extrapolator_noalloc_dev: device %d", dev);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      pq_set_error("pq_cabaret_extrapolator_noalloc_dev: hipGetSymbolAddress failed: %s",
                   hipGetErrorString(e));
      return PQ_ERR_HIP;
    }
    g_noalloc_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_noalloc_ptr[dev];
  return PQ_OK;
}
}  // namespace

// -------------------------------------------------------------------- C ABI
extern "C" {

int pq_cabaret_cfl_dev(int g, int L, double dt, const double *ui, const double *uj,
                       double *cfl_host, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_cfl_dev");
  PQ_CHECK_ARG(ui && uj && cfl_host, "pq_cabaret_cfl_dev: null pointer");
  Dims d{g, L};
  unsigned long long *dm = nullptr;
  int rc = stage_buffer(S(stream), sizeof(unsigned long long), (void **)&dm);
  if (rc) return rc;
  PQ_HIP(hipMemsetAsync(dm, 0, sizeof(unsigned long long), S(stream)));
  hipLaunchKernelGGL(k_maxabs<double>, grid_for(g - 1, g - 1, L), dim3(BX, BY), 0, S(stream), d, ui, uj, dm);
  unsigned long long bits = 0;
  PQ_HIP(hipMemcpyAsync(&bits, dm, sizeof(bits), hipMemcpyDeviceToHost, S(stream)));
  PQ_HIP(hipStreamSynchronize(S(stream)));
  double m;
  std::memcpy(&m, &bits, sizeof(m));
  *cfl_host = m * dt;
  return PQ_OK;
}

int pq_cabaret_predictor_dev(int g, int L, double dt, const double *zeta_old, const double *res,
                             double *zeta, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_predictor_dev");
  hipLaunchKernelGGL(k_predictor<double>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, S(stream), Dims{g, L},
                     dt, zeta_old, res, zeta);
  return check_launch("pq_cabaret_predictor_dev");
}

int pq_cabaret_beta_dev(int g, int L, double dt, double tw_a, const double *beta, const double *ui,
                        double *zeta, double *res_beta, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_beta_dev");
  hipLaunchKernelGGL(k_beta<double>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, S(stream), Dims{g, L}, dt,
                     tw_a, beta, ui, zeta, res_beta);
  return check_launch("pq_cabaret_beta_dev");
}

int pq_cabaret_forcing_dev(int g, int L, double dt, const double *fw, double visc,
                           double visc_bot, const double *rel, double *zeta, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_forcing_dev");
  hipLaunchKernelGGL(k_forcing<double>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, S(stream), Dims{g, L},
                     dt, fw, visc, visc_bot, rel, zeta);
  return check_launch("pq_cabaret_forcing_dev");
}

int pq_cabaret_cell_boundaries_dev(int g, int L, double tw, const double *psi, double *ui,
                                   double *uj, double *uimo, double *ujmo, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_cell_boundaries_dev");
  hipLaunchKernelGGL(k_cell_boundaries<double>, grid_for(g - 1, g - 1, L), dim3(BX, BY), 0, S(stream),
                     Dims{g, L}, tw, psi, ui, uj, uimo, ujmo);
  return check_launch("pq_cabaret_cell_boundaries_dev");
}

int pq_cabaret_cell_boundaries_velocity_dev(int g, int L, const double *psi, double *ui,
                                            double *uj, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_cell_boundaries_velocity_dev");
  hipLaunchKernelGGL(k_cell_boundaries_velocity<double>, grid_for(g - 1, g - 1, L), dim3(BX, BY), 0,
                     S(stream), Dims{g, L}, psi, ui, uj);
  return check_launch("pq_cabaret_cell_boundaries_velocity_dev");
}

int pq_cabaret_extrapolator_dev(int g, int L, double dt, const double *zeta,
                                const double *zeta_old, const double *ui, const double *uj,
                                double *zfi, double *zfj, double *work, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_extrapolator_dev");
  Dims d{g, L};
  double *tmp = work;
  if (!tmp) {
    int rc = stage_buffer(S(stream), sizeof(double) * (d.n_fi() + d.n_fj()), (void **)&tmp);
    if (rc) return rc;
  }
  extrapolator_launch(d, dt, zeta, zeta_old, ui, uj, zfi, zfj, tmp, S(stream));
  return check_launch("pq_cabaret_extrapolator_dev");
}

int pq_cabaret_extrapolator_noalloc_dev(int g, int L, double dt, const double *zeta,
                                        const double *zeta_old, const double *ui, const double *uj,
                                        double *zfi, double *zfj, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_extrapolator_noalloc_dev");
  PQ_CHECK_ARG(zeta && zeta_old && ui && uj && zfi && zfj,
               "pq_cabaret_extrapolator_noalloc_dev: null pointer");
  Dims d{g, L};
  const size_t need = d.n_fi() + d.n_fj();
  PQ_CHECK_ARG(need <= kNoAllocDoubles,
               "pq_cabaret_extrapolator_noalloc_dev: g=%d, %d layers need %zu doubles of work area; "
               "the library holds %zu (2 (g+3)(g+2) layers must not exceed it)",
               g, L, need, kNoAllocDoubles);
  double *tmp = nullptr;
  int rc = noalloc_scratch(&tmp);
  if (rc) return rc;
  extrapolator_launch(d, dt, zeta, zeta_old, ui, uj, zfi, zfj, tmp, S(stream));
  return check_launch("pq_cabaret_extrapolator_noalloc_dev");
}

int pq_cabaret_flux_dev(int g, int L, const double *ui, const double *uj, const double *zfi,
                        const double *zfj, double *flux_i, double *flux_j, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_flux_dev");
  hipLaunchKernelGGL(k_flux<double>, grid_for(g + 1, g + 1, L), dim3(BX, BY), 0, S(stream), Dims{g, L}, ui,
                     uj, zfi, zfj, flux_i, flux_j);
  return check_launch("pq_cabaret_flux_dev");
}

int pq_cabaret_corrector_dev(int g, int L, double dt, const double *flux_i, const double *flux_j,
                             double *res, double *zeta, void *stream) {
  CHECK_GL(g, L, "pq_cabaret_corrector_dev");
  hipLaunchKernelGGL(k_corrector<double>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, S(stream), Dims{g, L},
                     dt, flux_i, flux_j, res, zeta);
  return check_launch("pq_cabaret_corrector_dev");
}

int pq_rel_from_zeta_and_psi_dev(int g, int L, const double *s1, const double *s2, int bc_type,
                                 double alpha, const double *zeta, const double *psi, double *rel,
                                 void *stream) {
  CHECK_GL(g, L, "pq_rel_from_zeta_and_psi_dev");
  PQ_CHECK_ARG(bc_type >= 1 && bc_type <= 3, "pq_rel_from_zeta_and_psi_dev: bad bc_type %d", bc_type);
  // one layer: set(rel, zeta) copies the whole data array, halo included
  if (L == 1)
    PQ_HIP(hipMemcpyAsync(rel, zeta, sizeof(double) * Dims{g, L}.n_nd(), hipMemcpyDeviceToDevice, S(stream)));
  hipLaunchKernelGGL(k_rel<double>, grid_for(g, g, L), dim3(BX, BY), 0, S(stream), Dims{g, L}, s1, s2,
                     bc_type, alpha, zeta, psi, rel);
  return check_launch("pq_rel_from_zeta_and_psi_dev");
}

int pq_update_zeta_bc_dev(int g, int L, const double *s1, const double *s2, const double *rel,
                          const double *psi, double *zeta, void *stream) {
  CHECK_GL(g, L, "pq_update_zeta_bc_dev");
  hipLaunchKernelGGL(k_zeta_bc<double>, grid_for(g, g, L), dim3(BX, BY), 0, S(stream), Dims{g, L}, s1, s2,
                     rel, psi, zeta);
  return check_launch("pq_update_zeta_bc_dev");
}

int pq_zeta_step_dev(int g, int L, double T, const double *beta, const double *s1,
                     const double *s2, const double *force_wind, double viscosity,
                     double visc_bot, const double *psi, const double *rel, const double *rel_p,
                     const double *src, double *dst, void *stream) {
  CHECK_GL(g, L, "pq_zeta_step_dev");
  hipLaunchKernelGGL(k_zeta_step<double>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, S(stream), Dims{g, L},
                     T, beta, s1, s2, force_wind, viscosity, visc_bot, psi, rel, rel_p, src, dst);
  return check_launch("pq_zeta_step_dev");
}

int pq_ra_filter_dev(int g, int L, double *phi_l, double *phi_m, const double *phi_old_l,
                     double *phi_old_m, void *stream) {
  CHECK_GL(g, L, "pq_ra_filter_dev");
  hipLaunchKernelGGL(k_ra_filter<double>, grid_for(g, g, L), dim3(BX, BY), 0, S(stream), Dims{g, L}, phi_l,
                     phi_m, phi_old_l, phi_old_m);
  return check_launch("pq_ra_filter_dev");
}

// ------------------------------------------------------------------ model
struct pq_model {
  int g = 0, L = 0;
  Dims d{};
  pq_inv *inv = nullptr;
  hipStream_t stream = nullptr;
  // device fields
  double *psi = nullptr, *rel = nullptr, *zeta = nullptr, *zeta_old = nullptr;
  double *zfi = nullptr, *ui = nullptr, *uimo = nullptr, *flux_i = nullptr;
  double *zfj = nullptr, *uj = nullptr, *ujmo = nullptr, *flux_j = nullptr;
  double *res = nullptr, *res_beta = nullptr, *fw = nullptr;
  double *fin = nullptr, *fjn = nullptr;
  // the model's step: the second zeta_old buffer (fin / fjn are the second
  // face buffers), the forced zeta of the step, and the device word naming
  // the current buffers (0: zeta_old, zfi, zfj; 1: zeta_old_b, fin, fjn)
  double *zeta_old_b = nullptr, *zf = nullptr;
  int *par = nullptr;
  double *s1 = nullptr, *s2 = nullptr, *beta = nullptr;
  // centred (leapfrog) scheme: the second time level; lf is the level the
  // next step writes (1: psi/rel/zeta/zeta_old, 2: the *2 arrays)
  int scheme = PQ_SCHEME_CABARET, lf = 1;
  double *psi2 = nullptr, *rel2 = nullptr, *zeta2 = nullptr, *zeta_old2 = nullptr;
  size_t bytes = 0;  // device bytes of the fields above
  // bytes per value of the fields above (8, or 4 for the FP32-storage
  // precisions); fw, s1, s2 and beta are always FP64.  The stage kernels
  // compute in FP64 either way.
  int es = 8;
  unsigned long long *maxu = nullptr;
  unsigned long long *maxu_host = nullptr;
  // scalars
  double dt = 0, tw_a = 0.5, tw_u = 0.5, model_time = 0;
  double visc = 0, visc_bot = 0, alpha = 0;
  int bc_type = 1, adapt = 0;
  // flux_i / flux_j are not written by the model's step (its fluxes stay in
  // LDS): they are formed from the velocities and face values when read
  bool flux_stale = false;
  double target_cfl = 0, max_dt = 0, max_dt_inc = 0;
  // time mean of psi (qg3.F90:672-681) and energy diagnostics
  int average = 0;
  double max_spin = 0, diag_time = 0;
  double *psi_av = nullptr;          // (G, G, L)
  std::vector<double> s2_host, diag_h;
  pq_diag *diag = nullptr;
};

static double **model_field(pq_model *m, int which, size_t *count) {
  const Dims &d = m->d;
  switch (which) {
    case 0: *count = d.n_nd(); return &m->psi;
    case 1: *count = d.n_nd(); return &m->zeta;
    case 2: *count = d.n_nd(); return &m->rel;
    case 3: *count = d.n_nd(); return &m->zeta_old;
    case 4: *count = d.n_fi(); return &m->ui;
    case 5: *count = d.n_fj(); return &m->uj;
    case 6: *count = d.n_fi(); return &m->zfi;
    case 7: *count = d.n_fj(); return &m->zfj;
    case 8: *count = d.n_rs(); return &m->res;
    case 9: *count = d.n_rs(); return &m->res_beta;
    case 10: *count = d.n_fi(); return &m->uimo;
    case 11: *count = d.n_fj(); return &m->ujmo;
    case 12: *count = d.n_fi(); return &m->flux_i;
    case 13: *count = d.n_fj(); return &m->flux_j;
    case 14: *count = d.n_nd(); return &m->psi2;
    case 15: *count = d.n_nd(); return &m->zeta2;
    case 16: *count = d.n_nd(); return &m->rel2;
    case 17: *count = d.n_nd(); return &m->zeta_old2;
    default: *count = 0; return nullptr;
  }
}

// Index of the current zeta_old / face-value buffers (always 0 for the
// centred scheme), read back once the model's stream is idle.
static int model_par(pq_model *m, int *out) {
  *out = 0;
  if (!m->par) return PQ_OK;
  PQ_HIP(hipMemcpyAsync(out, m->par, sizeof(int), hipMemcpyDeviceToHost, m->stream));
  PQ_HIP(hipStreamSynchronize(m->stream));
  return PQ_OK;
}

// Makes buffer 0 current (copying buffer 1 into it when that one is), for
// host code that addresses zeta_old, zeta_fi and zeta_fj directly.
static int model_canon(pq_model *m) {
  int p = 0;
  int rc = model_par(m, &p);
  if (rc || !p) return rc;
  const Dims &d = m->d;
  const hipStream_t s = m->stream;
  PQ_HIP(hipMemcpyAsync(m->zeta_old, m->zeta_old_b, m->es * d.n_nd(), hipMemcpyDeviceToDevice, s));
  PQ_HIP(hipMemcpyAsync(m->zfi, m->fin, m->es * d.n_fi(), hipMemcpyDeviceToDevice, s));
  PQ_HIP(hipMemcpyAsync(m->zfj, m->fjn, m->es * d.n_fj(), hipMemcpyDeviceToDevice, s));
  static const int zero = 0;
  PQ_HIP(hipMemcpyAsync(m->par, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  PQ_HIP(hipStreamSynchronize(s));
  return PQ_OK;
}

// Both face buffers equal (after the host wrote buffer 0), so that the
// faces the step never writes agree whichever buffer is current.
static int model_mirror_faces(pq_model *m) {
  const Dims &d = m->d;
  PQ_HIP(hipMemcpyAsync(m->fin, m->zfi, m->es * d.n_fi(), hipMemcpyDeviceToDevice, m->stream));
  PQ_HIP(hipMemcpyAsync(m->fjn, m->zfj, m->es * d.n_fj(), hipMemcpyDeviceToDevice, m->stream));
  PQ_HIP(hipStreamSynchronize(m->stream));
  return PQ_OK;
}

int pq_model_destroy(pq_model *m) {
  if (!m) return PQ_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (double *p : {m->psi, m->rel, m->zeta, m->zeta_old, m->zfi, m->ui, m->uimo, m->flux_i,
                    m->zfj, m->uj, m->ujmo, m->flux_j, m->res, m->res_beta, m->fw, m->fin, m->fjn,
                    m->s1, m->s2, m->beta, m->psi2, m->rel2, m->zeta2, m->zeta_old2,
                    m->zeta_old_b, m->zf})
    (void)hipFree(p);
  (void)hipFree(m->par);
  (void)hipFree(m->psi_av);
  pq_diag_destroy(m->diag);
  (void)hipFree(m->maxu);
  (void)hipHostFree(m->maxu_host);
  pq_inv_destroy(m->inv);  // also destroys m->stream, which is the handle's stream
  delete m;
  return PQ_OK;
}

int pq_model_create(pq_model **out, const pq_model_params *p) {
  PQ_CHECK_ARG(out && p, "pq_model_create: null argument");
  PQ_CHECK_ARG(p->s1 && p->s2 && p->ss && p->modes_to_z && p->z_to_modes && p->beta && p->force_wind,
               "pq_model_create: null parameter array");
  PQ_CHECK_ARG(p->dt > 0.0, "pq_model_create: dt must be > 0");
  PQ_CHECK_ARG(p->scheme == PQ_SCHEME_CABARET || p->scheme == PQ_SCHEME_LEAPFROG,
               "pq_model_create: unknown scheme %d", p->scheme);
  PQ_CHECK_ARG(p->scheme == PQ_SCHEME_CABARET || !p->adapt_dt,
               "pq_model_create: adapt_dt is a CABARET option");
  PQ_CHECK_ARG(p->precision >= 0 && p->precision <= PQ_PREC_FP32_MFMA,
               "pq_model_create: bad precision %d", p->precision);
  pq_model *m = new pq_model();
  m->es = (p->precision == PQ_PREC_FP64 || p->precision == PQ_PREC_MIXED) ? 8 : 4;
  m->scheme = p->scheme;
  m->g = p->g;
  m->L = p->layers;
  m->d = Dims{p->g, p->layers};
  int rc = pq_inv_create(&m->inv, p->g, p->layers, p->ss, p->modes_to_z, p->z_to_modes, p->precision, 1);
  if (rc) {
    delete m;
    return rc;
  }
  auto fail = [&](int code) {
    pq_model_destroy(m);
    return code;
  };
  const Dims &d = m->d;
  struct A {
    double **p;
    size_t n;
  };
  std::vector<A> allocs = {{&m->psi, d.n_nd()}, {&m->rel, d.n_nd()}, {&m->zeta, d.n_nd()},
                           {&m->zeta_old, d.n_nd()}};
  if (m->scheme == PQ_SCHEME_CABARET)
    allocs.insert(allocs.end(), {{&m->zfi, d.n_fi()}, {&m->ui, d.n_fi()}, {&m->uimo, d.n_fi()},
                                 {&m->flux_i, d.n_fi()}, {&m->zfj, d.n_fj()}, {&m->uj, d.n_fj()},
                                 {&m->ujmo, d.n_fj()}, {&m->flux_j, d.n_fj()}, {&m->res, d.n_rs()},
                                 {&m->res_beta, d.n_rs()}, {&m->fin, d.n_fi()}, {&m->fjn, d.n_fj()},
                                 {&m->zeta_old_b, d.n_nd()}, {&m->zf, d.n_nd()}});
  else
    allocs.insert(allocs.end(), {{&m->psi2, d.n_nd()}, {&m->rel2, d.n_nd()}, {&m->zeta2, d.n_nd()},
                                 {&m->zeta_old2, d.n_nd()}});
  std::vector<A> fp64 = {{&m->fw, (size_t)p->g * p->g}, {&m->s1, (size_t)p->layers},
                         {&m->s2, (size_t)p->layers}, {&m->beta, (size_t)p->layers}};
  for (auto &a : allocs) {
    if (hipMalloc((void **)a.p, a.n * m->es) != hipSuccess) return fail(PQ_ERR_NOMEM);
    if (hipMemset(*a.p, 0, a.n * m->es) != hipSuccess) return fail(PQ_ERR_HIP);
    m->bytes += a.n * m->es;
  }
  for (auto &a : fp64) {
    if (hipMalloc((void **)a.p, a.n * sizeof(double)) != hipSuccess) return fail(PQ_ERR_NOMEM);
    m->bytes += a.n * sizeof(double);
  }
  if (hipMalloc((void **)&m->maxu, sizeof(unsigned long long)) != hipSuccess) return fail(PQ_ERR_NOMEM);
  if (m->scheme == PQ_SCHEME_CABARET) {
    if (hipMalloc((void **)&m->par, sizeof(int)) != hipSuccess) return fail(PQ_ERR_NOMEM);
    if (hipMemset(m->par, 0, sizeof(int)) != hipSuccess) return fail(PQ_ERR_HIP);
  }
  // [0] the maximum's bits, [1] the inversion status word read with it
  if (hipHostMalloc((void **)&m->maxu_host, 2 * sizeof(unsigned long long), 0) != hipSuccess)
    return fail(PQ_ERR_NOMEM);
  (void)hipMemcpy(m->fw, p->force_wind, sizeof(double) * p->g * p->g, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->s1, p->s1, sizeof(double) * p->layers, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->s2, p->s2, sizeof(double) * p->layers, hipMemcpyHostToDevice);
  (void)hipMemcpy(m->beta, p->beta, sizeof(double) * p->layers, hipMemcpyHostToDevice);
  m->stream = (hipStream_t)pq_inv_stream(m->inv);
  m->s2_host.assign(p->s2, p->s2 + p->layers);
  m->dt = p->dt;
  m->visc = p->viscosity;
  m->visc_bot = p->visc_bot;
  m->bc_type = p->bc_type;
  m->alpha = p->alpha;
  m->adapt = p->adapt_dt;
  m->target_cfl = p->target_cfl;
  m->max_dt = p->max_dt;
  m->max_dt_inc = p->max_dt_increase;
  *out = m;
  return PQ_OK;
}

void *pq_model_stream(pq_model *m) { return m ? (void *)m->stream : nullptr; }

}  // extern "C"

// The model's steps, templated on the field storage type; the C entry
// points below keep the C linkage of their declarations.
// A model field viewed with its storage type.
template <typename TS> static inline TS *fld(double *p) { return reinterpret_cast<TS *>(p); }
#define PQ_BY_STORAGE(m, f, ...) ((m)->es == 4 ? f<float>(__VA_ARGS__) : f<double>(__VA_ARGS__))

template <typename TS> static int invert_into(pq_model *m, double *zeta, double *psi) {
  const int ld = m->g + 2;
  const void *z[1] = {fld<TS>(zeta) + 1 + ld};
  void *p[1] = {fld<TS>(psi) + 1 + ld};
  return pq_inv_run_dev(m->inv, 1, z, ld, ld, p, ld, ld, m->stream);
}

static int model_invert(pq_model *m) { return PQ_BY_STORAGE(m, invert_into, m, m->zeta, m->psi); }

// rel_from_zeta_and_psi and update_zeta_bc (qg3p_aux.F90:606-640, 643-734)
// on the model's fields.
template <typename TS>
static void launch_rel(pq_model *m, double *zeta, double *psi, double *rel, hipStream_t s) {
  const int g = m->g, L = m->L;
  hipLaunchKernelGGL(k_rel<TS>, grid_for(g, g, L), dim3(BX, BY), 0, s, m->d, m->s1, m->s2,
                     m->bc_type, m->alpha, fld<TS>(zeta), fld<TS>(psi), fld<TS>(rel));
  hipLaunchKernelGGL(k_zeta_bc<TS>, grid_for(g, g, L), dim3(BX, BY), 0, s, m->d, m->s1, m->s2,
                     fld<TS>(rel), fld<TS>(psi), fld<TS>(zeta));
}

// Arrays of time level k (1 or 2) of a centred-scheme model.
struct Level {
  double *psi, *rel, *zeta, *zeta_old;
};
static Level level(pq_model *m, int k) {
  return k == 1 ? Level{m->psi, m->rel, m->zeta, m->zeta_old}
                : Level{m->psi2, m->rel2, m->zeta2, m->zeta_old2};
}

// pv_inversion, rel_from_zeta_and_psi and update_zeta_bc of one level
// (qg3.F90:383-391, 640-645).  A one-layer model's rel is zeta itself
// (set(rel, zeta), halo included) before the boundary values.
template <typename TS> static int lf_diagnose(pq_model *m, const Level &v) {
  int rc = invert_into<TS>(m, v.zeta, v.psi);
  if (rc) return rc;
  if (m->L == 1)
    PQ_HIP(hipMemcpyAsync(v.rel, v.zeta, sizeof(TS) * m->d.n_nd(), hipMemcpyDeviceToDevice, m->stream));
  launch_rel<TS>(m, v.zeta, v.psi, v.rel, m->stream);
  return check_launch("pq_model_step");
}

// The time-mean accumulation after the inversion (qg3.F90:672-681), from
// psi's sdata view (FP64 sum; FP32 fields read widened).
static int model_accumulate(pq_model *m, const double *psi, hipStream_t s) {
  if (m->average && m->model_time > m->max_spin) {
    const int ld = m->g + 2;
    const void *sd = reinterpret_cast<const char *>(psi) + (size_t)m->es * (1 + ld);
    int rc = pq::psi_av_add_es(m->g, m->L, m->dt, sd, ld, ld, m->psi_av, m->g, m->g, s, m->es);
    if (rc) return rc;
    m->diag_time += m->dt;
  }
  return PQ_OK;
}

// One pass of the centred scheme's DOLEAPFROG body (qg3.F90:612-660):
// zeta_centred into level lf, the Robert-Asselin-Williams filter, then
// the level's inversion, rel and zeta boundary, model_time += dt.
template <typename TS> static int lf_step(pq_model *m) {
  const Level vl = level(m, m->lf), vm = level(m, 3 - m->lf);
  const int g = m->g, L = m->L;
  const hipStream_t s = m->stream;
  hipLaunchKernelGGL(k_zeta_step<TS>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, s, m->d,
                     2.0 * m->dt, m->beta, m->s1, m->s2, m->fw, m->visc, m->visc_bot,
                     fld<TS>(vm.psi), fld<TS>(vm.rel), fld<TS>(vl.rel), fld<TS>(vl.zeta),
                     fld<TS>(vl.zeta));
  hipLaunchKernelGGL(k_ra_filter<TS>, grid_for(g, g, L), dim3(BX, BY), 0, s, m->d, fld<TS>(vl.zeta),
                     fld<TS>(vm.zeta), fld<TS>(vl.zeta_old), fld<TS>(vm.zeta_old));
  int rc = check_launch("pq_model_step");
  if (rc) return rc;
  rc = lf_diagnose<TS>(m, vl);
  if (rc) return rc;
  m->model_time += m->dt;
  rc = model_accumulate(m, vl.psi, s);
  if (rc) return rc;
  m->lf = m->lf % 2 + 1;
  return PQ_OK;
}

// Start of the centred scheme (qg3.F90:383-391, 449-452, 456-490): level 1
// from the installed zeta, zeta_old(1) = zeta(1), the forward Euler step
// into level 2 (FORWARD_EULER_START), zeta_old(2) = zeta(2), model_time = dt.
template <typename TS> static int lf_init(pq_model *m) {
  const size_t nb = sizeof(TS) * m->d.n_nd();
  const Level v1 = level(m, 1), v2 = level(m, 2);
  int rc = lf_diagnose<TS>(m, v1);
  if (rc) return rc;
  PQ_HIP(hipMemcpyAsync(v1.zeta_old, v1.zeta, nb, hipMemcpyDeviceToDevice, m->stream));
  PQ_HIP(hipMemsetAsync(v2.zeta, 0, nb, m->stream));
  hipLaunchKernelGGL(k_zeta_step<TS>, grid_for(m->g - 2, m->g - 2, m->L), dim3(BX, BY), 0,
                     m->stream, m->d, m->dt, m->beta, m->s1, m->s2, m->fw, m->visc, m->visc_bot,
                     fld<TS>(v1.psi), fld<TS>(v1.rel), fld<TS>(v1.rel), fld<TS>(v1.zeta),
                     fld<TS>(v2.zeta));
  rc = check_launch("pq_model_init");
  if (rc) return rc;
  rc = lf_diagnose<TS>(m, v2);
  if (rc) return rc;
  PQ_HIP(hipMemcpyAsync(v2.zeta_old, v2.zeta, nb, hipMemcpyDeviceToDevice, m->stream));
  m->model_time = 0.0 + m->dt;
  m->lf = 1;
  PQ_HIP(hipStreamSynchronize(m->stream));
  return pq_inv_status(m->inv, nullptr);
}

// START_PERTURB (qg3.F90:393-425) with zeta's data array already on the
// device: CABARET work arrays zeroed, psi from the inversion, rel and the
// zeta boundary, zeta_old = zeta, model_time = dt / 2.
template <typename TS> static int cabaret_init(pq_model *m) {
  const Dims &d = m->d;
  for (double *p : {m->zfi, m->ui, m->uimo, m->flux_i, m->fin})
    PQ_HIP(hipMemsetAsync(p, 0, sizeof(TS) * d.n_fi(), m->stream));
  for (double *p : {m->zfj, m->uj, m->ujmo, m->flux_j, m->fjn})
    PQ_HIP(hipMemsetAsync(p, 0, sizeof(TS) * d.n_fj(), m->stream));
  PQ_HIP(hipMemsetAsync(m->par, 0, sizeof(int), m->stream));
  m->flux_stale = false;
  PQ_HIP(hipMemsetAsync(m->res, 0, sizeof(TS) * d.n_rs(), m->stream));
  PQ_HIP(hipMemsetAsync(m->res_beta, 0, sizeof(TS) * d.n_rs(), m->stream));
  int rc = lf_diagnose<TS>(m, level(m, 1));
  if (rc) return rc;
  PQ_HIP(hipMemcpyAsync(m->zeta_old, m->zeta, sizeof(TS) * d.n_nd(), hipMemcpyDeviceToDevice,
                        m->stream));
  m->tw_a = 0.5;
  m->tw_u = 0.5;
  m->model_time = m->dt / 2.0;
  PQ_HIP(hipStreamSynchronize(m->stream));
  return pq_inv_status(m->inv, nullptr);
}

// Host values (FP64) into / out of a model field of either storage type.
static int field_upload(pq_model *m, double *dst, const double *src, size_t n, hipStream_t s) {
  if (m->es == 8) {
    PQ_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, s));
    PQ_HIP(hipStreamSynchronize(s));
    return PQ_OK;
  }
  std::vector<float> f(src, src + n);
  PQ_HIP(hipMemcpyAsync(dst, f.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
  PQ_HIP(hipStreamSynchronize(s));
  return PQ_OK;
}
static int field_download(pq_model *m, double *dst, const double *src, size_t n, hipStream_t s) {
  if (m->es == 8) {
    PQ_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    PQ_HIP(hipStreamSynchronize(s));
    return PQ_OK;
  }
  std::vector<float> f(n);
  PQ_HIP(hipMemcpyAsync(f.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  PQ_HIP(hipStreamSynchronize(s));
  std::copy(f.begin(), f.end(), dst);
  return PQ_OK;
}

int pq_model_init(pq_model *m, const double *zeta_sdata) {
  PQ_CHECK_ARG(m && zeta_sdata, "pq_model_init: null argument");
  const int g = m->g, L = m->L;
  const Dims &d = m->d;
  std::vector<double> z(d.n_nd(), 0.0);
  for (int l = 0; l < L; ++l)
    for (int j = 2; j <= g - 1; ++j)
      for (int i = 2; i <= g - 1; ++i)
        z[d.nd(i + 1, j + 1, l)] = zeta_sdata[(size_t)(i - 1) + (size_t)g * ((j - 1) + (size_t)g * l)];
  std::vector<double *> zero = {m->psi, m->rel};
  if (m->scheme == PQ_SCHEME_LEAPFROG) zero.insert(zero.end(), {m->psi2, m->rel2});
  for (double *p : zero) PQ_HIP(hipMemsetAsync(p, 0, m->es * d.n_nd(), m->stream));
  int rc = field_upload(m, m->zeta, z.data(), d.n_nd(), m->stream);
  if (rc) return rc;
  if (m->scheme == PQ_SCHEME_LEAPFROG) return PQ_BY_STORAGE(m, lf_init, m);
  return PQ_BY_STORAGE(m, cabaret_init, m);
}

// cabaret_adapt_dt (qg3p_aux.F90:946-990) with the scalar update on the host.
template <typename TS> static int model_adapt_dt(pq_model *m) {
  const Dims &d = m->d;
  PQ_HIP(hipMemsetAsync(m->maxu, 0, sizeof(unsigned long long), m->stream));
  hipLaunchKernelGGL(k_maxabs<TS>, grid_for(m->g - 1, m->g - 1, m->L), dim3(BX, BY), 0, m->stream,
                     d, fld<TS>(m->ui), fld<TS>(m->uj), m->maxu);
  PQ_HIP(hipMemcpyAsync(m->maxu_host, m->maxu, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        m->stream));
  // the status of the inversions since the last check comes back with the
  // maximum: a NaN psiC stops the model, as the reference aborts at its
  // first one (qg3p_aux.F90:504-509)
  int *st_host = reinterpret_cast<int *>(m->maxu_host + 1);
  PQ_HIP(hipMemcpyAsync(st_host, pq::inv_status_dev(m->inv), sizeof(int), hipMemcpyDeviceToHost, m->stream));
  PQ_HIP(hipStreamSynchronize(m->stream));
  if (*st_host & 1) {
    pq_set_error("NaN psiC encountered");
    return PQ_ERR_NAN;
  }
  double max_u;
  std::memcpy(&max_u, m->maxu_host, sizeof(double));
  const double old_dt = m->dt;
  const double a = std::fabs(max_u);
  const double spacing = a == 0.0 ? 2.2250738585072014e-308 : std::nextafter(a, INFINITY) - a;
  const double tol = 10.0 * std::fmax(spacing, 2.220446049250313e-16);
  if (std::fabs(max_u - 0.0) > tol) {
    double dd = m->target_cfl / max_u;
    dd = std::fmin(dd, std::fmin(old_dt * m->max_dt_inc, m->max_dt));
    m->dt = dd;
  }
  m->tw_a = 0.5 * m->dt / old_dt;
  m->tw_u = m->dt / (old_dt + m->dt);
  return PQ_OK;
}

// First half of a step: adapt_dt (host round trip when enabled), predictor
// and beta term (qg3.F90:782-790).  The PV inversion follows.
template <typename TS> static int model_pre_t(pq_model *m, hipStream_t s) {
  if (m->adapt) {
    int rc = model_adapt_dt<TS>(m);
    if (rc) return rc;
  }
  const int g = m->g, L = m->L;
  hipLaunchKernelGGL(k_step_predictor<TS>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, s, m->d,
                     m->dt, m->tw_a, m->beta, Pair<TS>{{fld<TS>(m->zeta_old), fld<TS>(m->zeta_old_b)}},
                     m->par, fld<TS>(m->res), fld<TS>(m->ui), fld<TS>(m->zeta), fld<TS>(m->res_beta));
  return PQ_OK;
}
static int model_pre(pq_model *m, hipStream_t s) { return PQ_BY_STORAGE(m, model_pre_t, m, s); }

// Second half, after psi is up to date: relative vorticity and boundary
// values, forcing, cell boundaries, extrapolator, fluxes and (unless this is
// a final partial step) the corrector with zeta_old <- zeta (qg3.F90:790-960).
template <typename TS> static int model_post_t(pq_model *m, hipStream_t s, bool last) {
  const int g = m->g, L = m->L;
  const Dims &d = m->d;
  const double dt = m->dt;
  TS *zeta = fld<TS>(m->zeta), *zeta_old = fld<TS>(m->zeta_old), *rel = fld<TS>(m->rel);
  TS *ui = fld<TS>(m->ui), *uj = fld<TS>(m->uj), *zfi = fld<TS>(m->zfi), *zfj = fld<TS>(m->zfj);
  TS *fin = fld<TS>(m->fin), *fjn = fld<TS>(m->fjn);
  TS *flux_i = fld<TS>(m->flux_i), *flux_j = fld<TS>(m->flux_j);
  if (!last) {
    // the tiled step: two passes, the fluxes in LDS
    TS *zf = fld<TS>(m->zf);
    hipLaunchKernelGGL(k_step_rel_force<TS>, step_grid(g, L), dim3(BX, BY), 0, s, d, m->s1, m->s2,
                       m->bc_type, m->alpha, m->tw_u, dt, m->fw, m->visc, m->visc_bot,
                       fld<TS>(m->psi), zeta, rel, ui, uj, fld<TS>(m->uimo), fld<TS>(m->ujmo), zf,
                       m->par);
    m->model_time += 0.5 * dt;
    int rc = model_accumulate(m, m->psi, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_step_extrap_correct<TS>, step_grid(g, L), dim3(BX, BY), 0, s, d, dt, zf,
                       Pair<TS>{{zeta_old, fld<TS>(m->zeta_old_b)}}, ui, uj, Pair<TS>{{zfi, fin}},
                       Pair<TS>{{zfj, fjn}}, fld<TS>(m->res), zeta, m->par);
    m->model_time += 0.5 * dt;
    m->flux_stale = true;
    return check_launch("pq_model_step");
  }
  // a final partial step: the stage kernels on buffer 0
  {
    int rc = model_canon(m);
    if (rc) return rc;
  }
  m->flux_stale = false;
  hipLaunchKernelGGL(k_post_rel_cells<TS>, grid_for(g + 2, g + 2, L), dim3(BX, BY), 0, s, d, m->s1,
                     m->s2, m->bc_type, m->alpha, m->tw_u, fld<TS>(m->psi), zeta, rel, ui, uj,
                     fld<TS>(m->uimo), fld<TS>(m->ujmo));
  m->model_time += 0.5 * dt;
  {
    int rc = model_accumulate(m, m->psi, s);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_forcing<TS>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, s, d, dt, m->fw,
                     m->visc, m->visc_bot, rel, zeta);
  hipLaunchKernelGGL(k_post_extrap_flux<TS>, grid_for(g + 1, g + 1, L), dim3(BX, BY), 0, s, d, dt,
                     zeta, zeta_old, ui, uj, zfi, zfj, fin, fjn, flux_i, flux_j);
  hipLaunchKernelGGL(k_face_copy<TS>, grid_for(g - 1, g - 1, L), dim3(BX, BY), 0, s, d, fin, fjn,
                     zfi, zfj);
  return check_launch("pq_model_step");
}
static int model_post(pq_model *m, hipStream_t s, bool last) {
  return PQ_BY_STORAGE(m, model_post_t, m, s, last);
}

int pq_model_step(pq_model *m, int nsteps, int final_partial) {
  PQ_CHECK_ARG(m != nullptr && nsteps >= 0, "pq_model_step: bad arguments");
  if (m->scheme == PQ_SCHEME_LEAPFROG) {
    PQ_CHECK_ARG(!final_partial, "pq_model_step: final_partial is a CABARET option");
    for (int it = 0; it < nsteps; ++it) {
      int rc = PQ_BY_STORAGE(m, lf_step, m);
      if (rc) return rc;
    }
    return PQ_OK;
  }
  for (int it = 0; it < nsteps; ++it) {
    int rc = model_pre(m, m->stream);
    if (rc) return rc;
    rc = model_invert(m);
    if (rc) return rc;
    rc = model_post(m, m->stream, (it == nsteps - 1) && final_partial);
    if (rc) return rc;
  }
  return PQ_OK;
}


// ---- hooks for the ensemble runtime (pq_ensemble.hip) --------------------
namespace pq {
int model_step_pre(pq_model *m, hipStream_t s) { return model_pre(m, s); }
int model_step_post(pq_model *m, hipStream_t s, bool last) { return model_post(m, s, last); }
int model_step_invert(pq_model *m) { return model_invert(m); }
hipStream_t model_stream(pq_model *m) { return m->stream; }
bool model_adaptive(const pq_model *m) { return m->adapt != 0; }
void model_views(pq_model *m, int *g, int *L, void **zeta_sdata, void **psi_sdata, long *ld, int *es) {
  *g = m->g;
  *L = m->L;
  *ld = m->g + 2;
  *es = m->es;
  const size_t off = (size_t)m->es * (1 + (m->g + 2));
  *zeta_sdata = (char *)m->zeta + off;
  *psi_sdata = (char *)m->psi + off;
}
ModelClock model_clock(const pq_model *m) { return {m->model_time, m->diag_time}; }
void model_set_clock(pq_model *m, ModelClock c) {
  m->model_time = c.t;
  m->diag_time = c.diag_t;
}
int model_step_kind(const pq_model *m) {
  if (!m->average) return 0;
  return m->model_time + 0.5 * m->dt > m->max_spin ? 1 : -1;
}
int *model_status_dev(pq_model *m) { return inv_status_dev(m->inv); }
size_t model_device_bytes(const pq_model *m) {
  return m->bytes + (m->psi_av ? sizeof(double) * m->d.n_rs() : 0) + inv_device_bytes(m->inv);
}
int model_scheme(const pq_model *m) { return m->scheme; }
// flux_i / flux_j of the current face values (cabaret_flux's ranges): what
// the reference's arrays hold after a step.
template <typename TS> static int model_fluxes(pq_model *m, int p) {
  const int g = m->g, L = m->L;
  hipLaunchKernelGGL(k_flux<TS>, grid_for(g + 1, g + 1, L), dim3(BX, BY), 0, m->stream, m->d,
                     fld<TS>(m->ui), fld<TS>(m->uj), fld<TS>(p ? m->fin : m->zfi),
                     fld<TS>(p ? m->fjn : m->zfj), fld<TS>(m->flux_i), fld<TS>(m->flux_j));
  m->flux_stale = false;
  return check_launch("pq_model_get");
}
ModelView model_view(pq_model *m) {
  // the view addresses buffer 0 of zeta_old and the face values; the file
  // writers and readers using it run on an idle stream
  if (model_canon(m) != PQ_OK) return ModelView{};
  if (m->flux_stale) (void)PQ_BY_STORAGE(m, model_fluxes, m, 0);
  ModelView v{};
  v.g = m->g;
  v.L = m->L;
  v.scheme = m->scheme;
  v.es = m->es;
  v.psi = m->psi;
  v.zeta = m->zeta;
  v.rel = m->rel;
  v.zeta_old = m->zeta_old;
  v.psi2 = m->psi2;
  v.zeta2 = m->zeta2;
  v.ui = m->ui;
  v.uj = m->uj;
  v.zfi = m->zfi;
  v.zfj = m->zfj;
  v.res = m->res;
  v.res_beta = m->res_beta;
  v.psi_av = m->psi_av;
  v.dt = m->dt;
  v.diag_time = m->diag_time;
  v.stream = m->stream;
  return v;
}
// START_PICKUP after the checkpoint's fields are in place (qg3.F90:316-364,
// 449-452, 456-466).  CABARET: uimo, ujmo from psi, the fluxes of the saved
// face values and the corrector's half step, zeta_old = zeta, model_time =
// dt / 2, time weights 1/2.  Centred scheme: rel = rel_from_psi of both
// levels, zeta_old = zeta, model_time = 0, level 1 next.
template <typename TS> static int pickup_t(pq_model *m, double dt) {
  const int g = m->g, L = m->L;
  const Dims &d = m->d;
  const hipStream_t s = m->stream;
  if (m->scheme == PQ_SCHEME_LEAPFROG) {
    for (int k = 1; k <= 2; ++k) {
      const Level v = level(m, k);
      hipLaunchKernelGGL(k_rel_from_psi<TS>, grid_for(g, g, L), dim3(BX, BY), 0, s, d, m->bc_type,
                         m->alpha, fld<TS>(v.psi), fld<TS>(v.rel));
      PQ_HIP(hipMemcpyAsync(v.zeta_old, v.zeta, sizeof(TS) * d.n_nd(), hipMemcpyDeviceToDevice, s));
    }
    m->model_time = 0.0;
    m->lf = 1;
  } else {
    m->dt = dt;
    TS *flux_i = fld<TS>(m->flux_i), *flux_j = fld<TS>(m->flux_j);
    hipLaunchKernelGGL(k_cell_boundaries_velocity<TS>, grid_for(g - 1, g - 1, L), dim3(BX, BY), 0, s,
                       d, fld<TS>(m->psi), fld<TS>(m->uimo), fld<TS>(m->ujmo));
    hipLaunchKernelGGL(k_flux<TS>, grid_for(g + 1, g + 1, L), dim3(BX, BY), 0, s, d, fld<TS>(m->ui),
                       fld<TS>(m->uj), fld<TS>(m->zfi), fld<TS>(m->zfj), flux_i, flux_j);
    hipLaunchKernelGGL(k_corrector<TS>, grid_for(g - 2, g - 2, L), dim3(BX, BY), 0, s, d, dt, flux_i,
                       flux_j, fld<TS>(m->res), fld<TS>(m->zeta));
    PQ_HIP(hipMemcpyAsync(m->zeta_old, m->zeta, sizeof(TS) * d.n_nd(), hipMemcpyDeviceToDevice, s));
    m->model_time = dt / 2.0;
    m->tw_a = 0.5;
    m->tw_u = 0.5;
    m->flux_stale = false;
  }
  int rc = check_launch("pickup");
  if (rc) return rc;
  PQ_HIP(hipStreamSynchronize(s));
  return m->scheme == PQ_SCHEME_CABARET ? model_mirror_faces(m) : PQ_OK;
}
int model_pickup(pq_model *m, double dt) { return PQ_BY_STORAGE(m, pickup_t, m, dt); }
// the same two half-step additions as model_post, so times round alike
void model_account_step(pq_model *m) {
  m->model_time += 0.5 * m->dt;
  if (m->average && m->model_time > m->max_spin) m->diag_time += m->dt;
  m->model_time += 0.5 * m->dt;
}
}  // namespace pq

extern "C" {

int pq_model_get(pq_model *m, int which, double *host, size_t count) {
  PQ_CHECK_ARG(m && host, "pq_model_get: null argument");
  size_t n = 0;
  double **f = model_field(m, which, &n);
  PQ_CHECK_ARG(f != nullptr, "pq_model_get: unknown field %d", which);
  PQ_CHECK_ARG(count >= n, "pq_model_get: field %d needs %zu values, buffer holds %zu", which, n, count);
  double *src = *f;
  int p = 0;
  int rc = model_par(m, &p);
  if (rc) return rc;
  if (p && which == 3) src = m->zeta_old_b;
  if (p && which == 6) src = m->fin;
  if (p && which == 7) src = m->fjn;
  if ((which == 12 || which == 13) && m->flux_stale) {
    rc = PQ_BY_STORAGE(m, pq::model_fluxes, m, p);
    if (rc) return rc;
  }
  return field_download(m, host, src, n, m->stream);
}

int pq_model_set(pq_model *m, int which, const double *host, size_t count) {
  PQ_CHECK_ARG(m && host, "pq_model_set: null argument");
  size_t n = 0;
  double **f = model_field(m, which, &n);
  PQ_CHECK_ARG(f != nullptr, "pq_model_set: unknown field %d", which);
  PQ_CHECK_ARG(count >= n, "pq_model_set: field %d needs %zu values, buffer holds %zu", which, n, count);
  PQ_HIP(hipStreamSynchronize(m->stream));
  int rc = model_canon(m);
  if (rc) return rc;
  if (which == 12 || which == 13) {
    // the other flux array takes its formed values first, so that it keeps
    // them once flux_i / flux_j are no longer formed on read
    if (m->flux_stale) {
      rc = PQ_BY_STORAGE(m, pq::model_fluxes, m, 0);
      if (rc) return rc;
    }
    m->flux_stale = false;
  }
  rc = field_upload(m, *f, host, n, m->stream);
  if (rc || !(m->par && (which == 6 || which == 7))) return rc;
  return model_mirror_faces(m);
}

int pq_model_status(pq_model *m) {
  PQ_CHECK_ARG(m != nullptr, "pq_model_status: null model");
  return pq_inv_status(m->inv, nullptr);
}

int pq_model_scalars(pq_model *m, double *out4) {
  PQ_CHECK_ARG(m && out4, "pq_model_scalars: null argument");
  out4[0] = m->dt;
  out4[1] = m->tw_a;
  out4[2] = m->tw_u;
  out4[3] = m->model_time;
  return PQ_OK;
}

int pq_model_average(pq_model *m, int enable, double max_spin) {
  PQ_CHECK_ARG(m != nullptr, "pq_model_average: null model");
  PQ_HIP(hipStreamSynchronize(m->stream));
  if (enable) {
    if (!m->psi_av && hipMalloc((void **)&m->psi_av, sizeof(double) * m->d.n_rs()) != hipSuccess) {
      m->psi_av = nullptr;
      pq_set_error("pq_model_average: out of device memory");
      return PQ_ERR_NOMEM;
    }
    PQ_HIP(hipMemset(m->psi_av, 0, sizeof(double) * m->d.n_rs()));
    m->diag_time = 0.0;
    m->max_spin = max_spin;
  }
  m->average = enable ? 1 : 0;
  return PQ_OK;
}

int pq_model_psi_av(pq_model *m, double *host, size_t count, double *diag_time) {
  PQ_CHECK_ARG(m && host, "pq_model_psi_av: null argument");
  const size_t n = m->d.n_rs();
  PQ_CHECK_ARG(count >= n, "pq_model_psi_av: needs %zu values, buffer holds %zu", n, count);
  PQ_CHECK_ARG(m->psi_av && m->diag_time > 0.0, "pq_model_psi_av: no time-mean samples yet");
  PQ_HIP(hipStreamSynchronize(m->stream));
  PQ_HIP(hipMemcpy(host, m->psi_av, n * sizeof(double), hipMemcpyDeviceToHost));
  const double cff = 1.0 / m->diag_time;
  for (size_t k = 0; k < n; ++k) host[k] *= cff;
  if (diag_time) *diag_time = m->diag_time;
  return PQ_OK;
}

int pq_model_energy(pq_model *m, const double *h, double *ekin, double *epot) {
  PQ_CHECK_ARG(m && h && ekin && epot, "pq_model_energy: null argument");
  if (!m->diag || !std::equal(m->diag_h.begin(), m->diag_h.end(), h)) {
    pq_diag_destroy(m->diag);
    m->diag = nullptr;
    m->diag_h.clear();
    int rc = pq_diag_create(&m->diag, m->g, m->L, h, m->s2_host.data());
    if (rc) return rc;
    m->diag_h.assign(h, h + m->L);
  }
  // the model's psi is read on the model's stream; the energy handle's own
  // synchronous call then copies the results
  const int ld = m->g + 2;
  PQ_HIP(hipStreamSynchronize(m->stream));
  // the level written last (centred scheme: the one before lf)
  const double *psi = m->scheme == PQ_SCHEME_LEAPFROG ? level(m, 3 - m->lf).psi : m->psi;
  const void *sd = reinterpret_cast<const char *>(psi) + (size_t)m->es * (1 + ld);
  return pq::diag_energy_sync(m->diag, sd, ld, ld, ekin, epot, m->es);
}

}  // extern "C"
