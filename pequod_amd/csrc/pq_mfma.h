// pq_mfma.h -- DST-I of one row on the matrix cores (gfx950 MFMA), N = 1024.
//
// The row passes' sine transform X_p = 2 sum_{i=1}^{N-1} x_i sin(pi p i / N)
// (FFTPACK sint's scaling, qg3p_helmholtz.F90:1370-1426) written as two
// small GEMMs.  With N = A B, A = B = 32, i = i1 + 32 i2 and
// p = p' + 64 p'' (p' < 64, p'' < 16), and W = exp(i pi / N):
//   T[p', i1] = sum_{i2} exp(i pi p' i2 / 32) x[i1 + 32 i2]       (stage 1)
//   Z[p', i1] = W^(p' i1) T[p', i1]                                (twiddle)
//   X_p / 2   = Im sum_{i1} exp(2 pi i p'' i1 / 32) Z[p', i1]      (stage 2)
// x is real, so T[64 - p'] = conj T[p'] and stage 1 needs p' = 0..32 only:
// Re T[0..31] (n-tile 0) and Im T[1..31] with Re T[32] in column 0 (n-tile
// 1), a 32 x 32 x 32 product.  Stage 2 takes Z[p'] for p' = 1..31 and gives
// the outputs p' + 64 p'' (direct rows 0..15) and, through
// Z[64 - p'][i1] = exp(i pi i1 / 16) conj Z[p'][i1], the outputs
// 64 - p' + 64 p'' (mirror rows 16..31: sin(2 pi (p''+1) i1 / 32) on Re Z,
// -cos(2 pi (p''+1) i1 / 32) on Im Z), a 32 x 64 x 32 product.  The outputs
// p = 32 m come from T[0] and T[32]:
//   X_{32m} / 2 = sum_{i1} sin(pi m i1 / 32) (m even ? T[0][i1] : T[32][i1]),
// a third product on column 0 (its twiddles are 1, so its stage-2 operand
// holds T[0] and T[32] as they are).
//
// Precision: every operand is split into an fp16 pair hi + lo (22 bits) and
// the products hi.hi + hi.lo + lo.hi are accumulated in FP32 by
// v_mfma_f32_32x32x16_f16.  Each row is scaled by a power of two so that
// |x| < 2^10 (stage-1 values stay below 2^15, inside fp16's range).
//
// Register layouts (v_mfma_f32_32x32x16_f16, lane l, c = l % 32, g = l / 32):
//   A operand: row c, k = 16 t + 8 g + s (slot s = 0..7 of k-step t)
//   B operand: column c, same k
//   D: column c, row 8 (v / 4) + 4 g + v % 4 in register v = 0..15
// Stage 1: A = x (row i1 = c, k = i2), B = the exp(i pi p' i2 / 32) table
// (k = i2, column = the p' row above).  Its D holds, in lane (c, g),
// Re/Im T[c][i1] for i1 = 8 (v / 4) + 4 g + v % 4: exactly the stage-2 B
// operand's column c when k-step t, slot s stands for
// i1 = 8 t + 4 g + s % 4 and part (Re, Im) = s / 4.  So the two stages
// chain in registers, with no transpose between them.
#pragma once

#include <hip/hip_runtime.h>

namespace pq {
namespace mfma {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kN = 1024;  // transform length of this engine
constexpr int kRowThreads = 256;  // row-pass workgroup: 4 wavefronts, one row each at a time
constexpr int kRows = 16;         // rows per scratch block (the handle's R) of the row passes
constexpr int kWords = 110;  // per-lane constant words (table: word w of lane l at w * 64 + l)
// word offsets
constexpr int kB1h = 0, kB1l = 16, kA2h = 32, kA2l = 48, kA3h = 64, kA3l = 80;
constexpr int kBre = 96, kBim = 100, kSre = 104, kSim = 107;

__host__ __device__ constexpr bool supported(int N) { return N == kN; }

// Constants held in registers; the p = 32 m rows' A operand (a third of
// the table) is read from LDS per row instead (a copy per workgroup, made by
// stage_a3), which keeps the row passes at 2 waves per SIMD without spills.
struct Consts {
  h8 b1h[2][2], b1l[2][2];  // stage-1 B operand [t][u]
  h8 a2h[4], a2l[4];        // stage-2 A operand [t]
  float bre[4], bim[4];     // W^(c (4 g + b)), b = 0..3
  float sre[3], sim[3];     // W^(8 a c), a = 1..3
};
// LDS words of the A3 copy: [t][hi, lo][lane][4]
constexpr int kA3Lds = 4 * 2 * 64 * 4;

// The A3 words into LDS as [t][hi, lo][lane][4] (all threads of the
// workgroup; a barrier must follow before dst_row reads them).
__device__ __forceinline__ void stage_a3(const unsigned *__restrict__ tab, unsigned *a3s, int tid,
                                         int nt) {
  for (int i = tid; i < kA3Lds; i += nt) {
    const int q = i & 3, lane = (i >> 2) & 63, hl = (i >> 8) & 1, t = i >> 9;
    a3s[i] = tab[((hl ? kA3l : kA3h) + 4 * t + q) * 64 + lane];
  }
}

__device__ __forceinline__ h8 ld_h8(const unsigned *__restrict__ tab, int w, int lane) {
  u32x4 v;
  v[0] = tab[(w + 0) * 64 + lane];
  v[1] = tab[(w + 1) * 64 + lane];
  v[2] = tab[(w + 2) * 64 + lane];
  v[3] = tab[(w + 3) * 64 + lane];
  return __builtin_bit_cast(h8, v);
}

__device__ __forceinline__ void load_consts(const unsigned *__restrict__ tab, int lane, Consts &K) {
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      K.b1h[t][u] = ld_h8(tab, kB1h + (2 * t + u) * 4, lane);
      K.b1l[t][u] = ld_h8(tab, kB1l + (2 * t + u) * 4, lane);
    }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    K.a2h[t] = ld_h8(tab, kA2h + 4 * t, lane);
    K.a2l[t] = ld_h8(tab, kA2l + 4 * t, lane);
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    K.bre[b] = __uint_as_float(tab[(kBre + b) * 64 + lane]);
    K.bim[b] = __uint_as_float(tab[(kBim + b) * 64 + lane]);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    K.sre[a] = __uint_as_float(tab[(kSre + a) * 64 + lane]);
    K.sim[a] = __uint_as_float(tab[(kSim + a) * 64 + lane]);
  }
}

// a - hi (hi the low or high half of h, widened): exact, one mixed-precision
// FMA.
__device__ __forceinline__ float rem_lo(unsigned h, float a) {
  float r;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(a));
  return r;
}
__device__ __forceinline__ float rem_hi(unsigned h, float a) {
  float r;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(a));
  return r;
}
// (a, b) -> packed fp16 hi and lo words, a + b.. = hi + lo to 22 bits.
__device__ __forceinline__ void split2(float a, float b, unsigned &hi, unsigned &lo) {
  const h2 h = {(_Float16)a, (_Float16)b};
  hi = __builtin_bit_cast(unsigned, h);
  const float ra = rem_lo(hi, a), rb = rem_hi(hi, b);
  const h2 l = {(_Float16)ra, (_Float16)rb};
  lo = __builtin_bit_cast(unsigned, l);
}
__device__ __forceinline__ void split8(const float (&v)[8], h8 &hi, h8 &lo) {
  u32x4 h, l;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned a, b;
    split2(v[2 * q], v[2 * q + 1], a, b);
    h[q] = a;
    l[q] = b;
  }
  hi = __builtin_bit_cast(h8, h);
  lo = __builtin_bit_cast(h8, l);
}

__device__ __forceinline__ f32x16 mma(h8 a, h8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
// c += A B from split operands: hi.hi + hi.lo + lo.hi
__device__ __forceinline__ f32x16 mma3(h8 ah, h8 al, h8 bh, h8 bl, f32x16 c) {
  c = mma(ah, bh, c);
  c = mma(ah, bl, c);
  return mma(al, bh, c);
}

// Input element of slot (t, s) in lane (c, g): x[c + 512 t + 256 g + 32 s].
__device__ __forceinline__ int in_index(int c, int g, int t, int s) { return c + 512 * t + 256 * g + 32 * s; }

// Maximum over the wavefront of a non-negative float (its bit pattern
// orders like the value), by DPP steps: quad permutations and row mirrors
// inside each row of 16 lanes, then row broadcasts; lane 63 ends with the
// maximum, which every lane reads.
__device__ __forceinline__ float wave_max_nonneg(float x) {
  int v = __float_as_int(x);
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));  // row_mirror
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(v, 63));
}

// Transform of one row held as x[t][s] (the stage-1 A operand elements,
// x[0][0] of lane 0 being x_0 = 0), any magnitude, into the row's output
// buffer out[0..N) (X_p at out[p]; LDS).  Every lane writes its 16 direct
// and mirror outputs (c != 0) or its 16 outputs p = 32 m (c = 0), with
// compile-time offsets from three per-lane bases.
__device__ __forceinline__ void dst_row(const Consts &K, const unsigned *a3s, float (&x)[2][8], int lane,
                                        float *out) {
  const int c = lane & 31, g = lane >> 5;
  // power-of-two scale: max |x| < 2^10
  float mx = 0.0f;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int s = 0; s < 8; ++s) mx = fmaxf(mx, fabsf(x[t][s]));
  mx = wave_max_nonneg(mx);
  const int e = mx > 0.0f ? ilogbf(mx) + 1 : 0;  // mx < 2^e
  h8 ah[2], al[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float v[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) v[s] = ldexpf(x[t][s], 10 - e);
    split8(v, ah[t], al[t]);
  }
  // stage 1
  f32x16 d0 = {}, d1 = {};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    d0 = mma3(ah[t], al[t], K.b1h[t][0], K.b1l[t][0], d0);
    d1 = mma3(ah[t], al[t], K.b1h[t][1], K.b1l[t][1], d1);
  }
  // twiddles W^(c i1), i1 = 8 a + 4 g + b (register v = 4 a + b); lane 0's
  // are 1, so its operand keeps T[0] (Re) and T[32] (Im)
  float zr[16], zi[16];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      float wr = K.bre[b], wi = K.bim[b];
      if (a > 0) {
        const float tr = wr * K.sre[a - 1] - wi * K.sim[a - 1];
        wi = wr * K.sim[a - 1] + wi * K.sre[a - 1];
        wr = tr;
      }
      const int v = 4 * a + b;
      zr[v] = wr * d0[v] - wi * d1[v];
      zi[v] = wr * d1[v] + wi * d0[v];
    }
  // stage 2 and the p = 32 m rows, from the same B operand
  f32x16 d2 = {}, d3 = {};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const float v[8] = {zr[4 * t], zr[4 * t + 1], zr[4 * t + 2], zr[4 * t + 3],
                        zi[4 * t], zi[4 * t + 1], zi[4 * t + 2], zi[4 * t + 3]};
    h8 bh, bl;
    split8(v, bh, bl);
    d2 = mma3(K.a2h[t], K.a2l[t], bh, bl, d2);
    const h8 a3h = __builtin_bit_cast(h8, reinterpret_cast<const u32x4 *>(a3s)[(2 * t) * 64 + lane]);
    const h8 a3l = __builtin_bit_cast(h8, reinterpret_cast<const u32x4 *>(a3s)[(2 * t + 1) * 64 + lane]);
    d3 = mma3(a3h, a3l, bh, bl, d3);
  }
  const int oe = e - 9;  // X = 2 D / 2^(10 - e)
  // register v = 4 a + b holds row m = 8 a + 4 g + b of its tile:
  // direct (a < 2) p = c + 64 m, mirror (a >= 2) p = 64 - c + 64 (m - 16),
  // the p = 32 m rows p = 32 m
  if (c != 0) {
    float *od = out + c + 256 * g, *om = out + 64 - c + 256 * g;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const float y = ldexpf(d2[4 * a + b], oe);
        if (a < 2) od[64 * (8 * a + b)] = y;
        else om[64 * (8 * (a - 2) + b)] = y;
      }
  } else {
    float *os = out + 128 * g;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) os[32 * (8 * a + b)] = ldexpf(d3[4 * a + b], oe);
  }
}

// Host: the per-lane constant table (kWords x 64 words).
void build_table(unsigned *tab);

}  // namespace mfma
}  // namespace pq
