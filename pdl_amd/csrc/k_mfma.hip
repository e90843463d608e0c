// MFMA matmult for float ndarrays on gfx950 (lib/PDL/Primitive.pd:191-245
// computes the same C[h][w] = sum_t A[h][t] B[t][w] with a scalar loop).
//
// Split-FP16 ("split16"): each float operand element x, pre-scaled by a
// power of two per row of A / per column of B so that the row's largest
// finite magnitude lies in [2^14, 2^15), is written as hi = fp16(x),
// lo = fp16(x - hi) (22 significant bits together).  The GEMM accumulates
// hi*hi + hi*lo + lo*hi in FP32 on the matrix cores
// (v_mfma_f32_32x32x16_f16) and the epilogue undoes the scaling exactly
// (ldexp by the two scales' exponents, one rounding).
//
// Range and non-finite values.  An element far below its row's maximum
// loses bits (lo, then hi, become FP16 subnormals; below 2^-40 of the max it
// flushes to zero), and an Inf/NaN element cannot be scaled at all.  The
// packing pass writes such elements' hi/lo as they fall (non-finite ones as
// zero) and records per row, in scaled units, D = sum over the row of the
// representation error beyond 2^-22 relative (+Inf for a row holding an
// Inf or NaN).  That excess perturbs output c[h][w] by at most
// D_A[h] * max|B col w| + D_B[w] * max|A row h| <= (D_A[h] + D_B[w]) * 2^15
// in scaled units; the epilogue recomputes, in PDL's own order (multiply
// then add, ascending t, from the FP32 operands), every element where that
// bound exceeds 2^-22 |c| or c is not finite.  Rows of normal(0,1) data
// flag a few dozen elements of an 8192^2 product; a row with an Inf or NaN
// gets its whole row of C from PDL's loop, so non-finite results match the
// reference bit for bit.
//
// FP8 ("fp8", opt-in): the same pre-scaling puts each row/column's max in
// [128, 256) (below e4m3's largest normal, 448); single e4m3 products on
// v_mfma_scale_f32_32x32x64_f8f6f4 (256x256 blocks) or
// v_mfma_f32_16x16x32_fp8_fp8.  Only non-finite rows are recomputed: the
// FP8 path is approximate by design.
//
// Layout: packed operands are [row][Kp] with K contiguous, rows padded to the
// block tile and K to the k step (zeros).
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

void *scratch_alloc(size_t bytes);
void scratch_free(void *p);

typedef _Float16 half_t;
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MB = 128, NB = 128, KB = 32, KPAD = KB + 8;

// |x| for finite x, 0 for Inf and NaN: row maxima ignore non-finite values
__device__ __forceinline__ float fin_abs(float x) {
  const float a = fabsf(x);
  return a < __builtin_inff() ? a : 0.0f;
}

// exponent of a power of two held as a normal float
__device__ __forceinline__ int pow2_exp(float s) { return ((__float_as_int(s) >> 23) & 0xff) - 127; }

// hi/lo split of a scaled element y (finite), and the part of its
// representation error beyond 2^-22 |y| (both differences are exact)
__device__ __forceinline__ float split_hilo(float y, half_t &h, half_t &l) {
  h = (half_t)y;
  const float r1 = y - (float)h;
  l = (half_t)r1;
  const float e = fabsf(r1 - (float)l) - 0x1p-22f * fabsf(y);
  return e > 0.0f ? e : 0.0f;
}

// Whether output c (unscaled; the accumulator was c * 2^e) must come from
// PDL's loop: the range-loss bound (D_A[h] + D_B[w]) * 2^15 (scaled units)
// exceeds 2^-22 |c|, or c is not finite.  The GEMM epilogue and the fixup
// kernel evaluate the same test on the same stored c.
__device__ __forceinline__ bool mm_needs_fix(float c, float dd, int e) {
  if (dd == 0.0f && fabsf(c) < __builtin_inff()) return false;
  return !(ldexpf(dd * 32768.0f, -e) <= 0x1p-22f * fabsf(c));
}

// the GEMM epilogue for one output: unscale the accumulator exactly and
// mark the row of an output the fixup kernel must recompute
__device__ __forceinline__ float mm_finish(float v, float da, float db, float sa, float sb, int *rowflag, int64_t h) {
  const int e = pow2_exp(sa) + pow2_exp(sb);
  const float c = ldexpf(v, -e);
  if (mm_needs_fix(c, da + db, e)) rowflag[h] = 1;   // the same value from every writer
  return c;
}

// power of two that brings max|row| into [2^(e-1), 2^e) with e = target
__device__ inline float pow2_scale(float m, int target) {
  if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
  int e;
  frexpf(m, &e);                 // m = f * 2^e, f in [0.5, 1)
  int x = target - e;
  x = x > 120 ? 120 : (x < -120 ? -120 : x);
  return ldexpf(1.0f, x);
}

// max |x| over k of each of R rows, one wave per row, when k is the
// unit-stride index
__global__ __launch_bounds__(256) void absmax_rows_wave(const float *__restrict__ src, int64_t o, int64_t inc_r,
                                                        int64_t inc_k, int64_t R, int64_t K, float *smax) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  float m = 0.0f;
  for (int64_t k = lane; k < K; k += 64) m = fmaxf(m, fin_abs(src[o + r * inc_r + k * inc_k]));
  for (int s = 32; s >= 1; s >>= 1) m = fmaxf(m, __shfl_xor(m, s, 64));
  if (lane == 0) smax[r] = m;
}

// the same with 16-byte loads, for rows that start 16-byte aligned
__global__ __launch_bounds__(256) void absmax_rows_vec(const float *__restrict__ src, int64_t o, int64_t inc_r,
                                                       int64_t R, int64_t K, float *smax) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  const float *row = src + o + r * inc_r;
  float m = 0.0f;
  const int64_t K4 = K / 4;
  for (int64_t k = lane; k < K4; k += 64) {
    const Vec<float, 4> v = *reinterpret_cast<const Vec<float, 4> *>(row + 4 * k);
    m = fmaxf(m, fmaxf(fmaxf(fin_abs(v.v[0]), fin_abs(v.v[1])), fmaxf(fin_abs(v.v[2]), fin_abs(v.v[3]))));
  }
  for (int64_t k = 4 * K4 + lane; k < K; k += 64) m = fmaxf(m, fin_abs(row[k]));
  for (int s = 32; s >= 1; s >>= 1) m = fmaxf(m, __shfl_xor(m, s, 64));
  if (lane == 0) smax[r] = m;
}

// load a 64x64 (rows x k) tile of src into t, coalesced along whichever
// index is unit-stride; rows unit-stride and 16-byte aligned with the whole
// tile in range take 16-byte loads.  fin_abs(x) when ABS.  Zeros outside R x K.
template <bool ABS>
__device__ inline void load_tile64(float (&t)[64][65], const float *__restrict__ src, int64_t o, int64_t inc_r,
                                   int64_t inc_k, int64_t R, int64_t K, int64_t r0, int64_t k0) {
  const bool kfast = (inc_k == 1 || inc_k == -1 || inc_r != 1);
  const bool rvec = !kfast && r0 + 64 <= R && k0 + 64 <= K && inc_k % 4 == 0 && ((uintptr_t)(src + o + r0) % 16 == 0);
  if (rvec) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int rr = (e % 16) * 4, kk = e / 16;
      const Vec<float, 4> v = *reinterpret_cast<const Vec<float, 4> *>(src + o + (r0 + rr) + (k0 + kk) * inc_k);
#pragma unroll
      for (int i = 0; i < 4; i++) t[rr + i][kk] = ABS ? fin_abs(v.v[i]) : v.v[i];
    }
    return;
  }
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    int rr, kk;
    if (kfast) { kk = e % 64; rr = e / 64; } else { rr = e % 64; kk = e / 64; }
    const int64_t gr = r0 + rr, gk = k0 + kk;
    const float x = (gr < R && gk < K) ? src[o + gr * inc_r + gk * inc_k] : 0.0f;
    t[rr][kk] = ABS ? fin_abs(x) : x;
  }
}

// max finite |x| per row of one 64x64 tile (read as pack_tile reads it, coalesced
// along whichever index is unit-stride), folded into smax with an atomic max
// on the bit pattern (non-negative floats order like their bits); smax starts
// at zero (zero_u32).  For operands whose k index is strided, where a thread
// per row would walk K serially.
__global__ __launch_bounds__(256) void absmax_tile(const float *__restrict__ src, int64_t o, int64_t inc_r,
                                                   int64_t inc_k, int64_t R, int64_t K, unsigned *smax,
                                                   int64_t ntk) {
  __shared__ float t[64][65];
  const int64_t r0 = (blockIdx.x / ntk) * 64, k0 = (blockIdx.x % ntk) * 64;
  load_tile64<true>(t, src, o, inc_r, inc_k, R, K, r0, k0);
  __syncthreads();
  const int rr = threadIdx.x >> 2, part = threadIdx.x & 3;
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < 16; i++) m = fmaxf(m, t[rr][part * 16 + i]);
  m = fmaxf(m, __shfl_xor(m, 1, 64));
  m = fmaxf(m, __shfl_xor(m, 2, 64));
  if (part == 0 && r0 + rr < R) atomicMax(&smax[r0 + rr], __float_as_uint(m));
}

__global__ __launch_bounds__(256) void zero_u32(unsigned *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// scale[r] from smax[r] (1 for padded rows); the rows' range-loss sums (and
// A's recompute marks) start at zero
__global__ __launch_bounds__(256) void make_scales(float *s, float *d, int *rowflag, int64_t R, int64_t Rp,
                                                   int target) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= Rp) return;
  s[r] = r < R ? pow2_scale(s[r], target) : 1.0f;
  d[r] = 0.0f;
  if (rowflag) rowflag[r] = 0;
}

// packs 8 consecutive k of one row: hi/lo FP16 (or e4m3) of x * sc with
// non-finite elements as zero; returns their range-loss excess (scaled
// units; +Inf if any element is non-finite, which FP8 tracks alone)
template <bool FP8>
__device__ __forceinline__ float pack8(const float (&x)[8], float sc, half_t *hi, half_t *lo, uint8_t *q8) {
  float ex = 0.0f;
  float y[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const bool fin = fabsf(x[i]) < __builtin_inff();
    y[i] = fin ? x[i] * sc : 0.0f;
    ex = fin ? ex : __builtin_inff();
  }
  if constexpr (FP8) {
    uint32_t w[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      uint32_t lo01 = __builtin_amdgcn_cvt_pk_fp8_f32(y[4 * h], y[4 * h + 1], 0, false);
      w[h] = __builtin_amdgcn_cvt_pk_fp8_f32(y[4 * h + 2], y[4 * h + 3], lo01, true);
    }
    *reinterpret_cast<uint2 *>(q8) = make_uint2(w[0], w[1]);
  } else {
    half8 vh, vl;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      half_t h, l;
      ex += split_hilo(y[i], h, l);
      vh[i] = h;
      vl[i] = l;
    }
    *reinterpret_cast<half8 *>(hi) = vh;
    *reinterpret_cast<half8 *>(lo) = vl;
  }
  return ex;
}

// pack a (R x K) operand into hi/lo fp16 [Rp][Kp] through a 64x64 LDS tile so
// that both the strided read and the K-contiguous write are coalesced
template <bool FP8>
__global__ __launch_bounds__(256) void pack_tile(const float *__restrict__ src, int64_t o, int64_t inc_r,
                                                 int64_t inc_k, int64_t R, int64_t K, int64_t Kp,
                                                 const float *__restrict__ scale, float *__restrict__ dsum,
                                                 half_t *__restrict__ hi, half_t *__restrict__ lo,
                                                 uint8_t *__restrict__ q8, int64_t ntk) {
  __shared__ float t[64][65];
  const int64_t r0 = (blockIdx.x / ntk) * 64, k0 = (blockIdx.x % ntk) * 64;
  load_tile64<false>(t, src, o, inc_r, inc_k, R, K, r0, k0);
  __syncthreads();
  // a thread writes 8 consecutive k of one row (Kp and the tile's k0 are
  // multiples of 32, so a group is all in or all out)
  for (int e = threadIdx.x; e < 64 * 8; e += 256) {
    const int rr = e / 8, kg = (e % 8) * 8;
    const int64_t gr = r0 + rr, gk = k0 + kg;
    if (gk >= Kp) continue;
    const float sc = scale[gr];
    float x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = t[rr][kg + i];
    const float ex = pack8<FP8>(x, sc, hi + gr * Kp + gk, lo + gr * Kp + gk, q8 + gr * Kp + gk);
    if (ex != 0.0f && gr < R) atomicAdd(&dsum[gr], ex);
  }
}

// pack an operand whose k index is unit-stride (rows 16-byte aligned): a
// thread converts 8 consecutive k of one row (two 16-byte loads) into one
// 16-byte hi and one lo store (8 bytes for FP8); zeros past R and K
template <bool FP8>
__global__ __launch_bounds__(256) void pack_rows(const float *__restrict__ src, int64_t o, int64_t inc_r, int64_t R,
                                                 int64_t K, int64_t Rp, int64_t Kp, const float *__restrict__ scale,
                                                 float *__restrict__ dsum, half_t *__restrict__ hi,
                                                 half_t *__restrict__ lo, uint8_t *__restrict__ q8) {
  const int64_t ng = Kp / 8, total = Rp * ng;
  for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (int64_t)gridDim.x * 256) {
    const int64_t row = g / ng, k0 = (g % ng) * 8;
    float x[8];
    if (row < R && k0 + 8 <= K) {
      const float *p = src + o + row * inc_r + k0;
      const Vec<float, 4> a = *reinterpret_cast<const Vec<float, 4> *>(p);
      const Vec<float, 4> b = *reinterpret_cast<const Vec<float, 4> *>(p + 4);
#pragma unroll
      for (int i = 0; i < 4; i++) { x[i] = a.v[i]; x[4 + i] = b.v[i]; }
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) x[i] = (row < R && k0 + i < K) ? src[o + row * inc_r + k0 + i] : 0.0f;
    }
    const float ex = pack8<FP8>(x, scale[row], hi + row * Kp + k0, lo + row * Kp + k0, q8 + row * Kp + k0);
    if (ex != 0.0f && row < R) atomicAdd(&dsum[row], ex);
  }
}

__device__ inline void mm_batch_offsets_m(const MmPlan &p, int64_t bidx, int64_t *o) {
  o[0] = p.off[0]; o[1] = p.off[1]; o[2] = p.off[2];
  for (int d = 0; d < p.nd; d++) {
    int64_t i = bidx % p.dims[d];
    bidx /= p.dims[d];
    o[0] += i * p.inc[0][d]; o[1] += i * p.inc[1][d]; o[2] += i * p.inc[2][d];
  }
}

// block -> C tile.  The 8 XCDs take blocks round-robin (block b runs on XCD
// b % 8), so each XCD is given one contiguous share of the tile order; within
// a share, tiles go in groups of 8 tile-rows, column by column, so the blocks
// resident on one XCD at a time reuse the same A row-panels and B
// column-panels from that XCD's L2.
__device__ inline void tile_of(int64_t b, int64_t ntm, int64_t ntn, int64_t &tm, int64_t &tn) {
  const int64_t nb = ntm * ntn;
  const int64_t t = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  constexpr int64_t G = 8;
  const int64_t g = t / (G * ntn), r = t % (G * ntn);
  const int64_t rows = ntm - g * G < G ? ntm - g * G : G;
  tm = g * G + r % rows;
  tn = r / rows;
}

// The operands stream into LDS with direct global->LDS loads
// (global_load_lds_dwordx4: lane i of a wave writes 16 bytes at the wave's
// LDS address + 16*i), no registers in between, over two stages: while the
// matrix cores work on stage k, stage k+1 is being filled.  A wave computes
// 64x64 of C as 2x2 tiles of v_mfma_f32_32x32x16_f16 (twice the work of the
// 16x16x32 form per issue slot); the block 128x128.
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <class T> __device__ __forceinline__ void dma16(const T *g, T *l) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                   (__attribute__((address_space(3))) void *)l, 16, 0, 0);
}

// BM x BN block tile, NW = WR x WC waves, each wave (BM/WR) x (BN/WC) of C as
// TI x TJ tiles of 32x32.  128x128 (4 waves, 2x2 tiles each) for smaller
// products; 256x256 (8 waves, 2x4 tiles each) halves the bytes staged per
// MFMA and gives each step enough work to cover the loads of the next.
template <int BM, int BN, int WR, int WC>
__global__ __launch_bounds__(WR * WC * 64) void split16_gemm(const half_t *__restrict__ Ahi,
                                                             const half_t *__restrict__ Alo,
                                                             const half_t *__restrict__ Bhi,
                                                             const half_t *__restrict__ Blo,
                                                             const float *__restrict__ sA,
                                                             const float *__restrict__ sB,
                                                             const float *__restrict__ dA,
                                                             const float *__restrict__ dB, int *rowflag,
                                                             float *__restrict__ c, MmPlan p, int64_t Kp,
                                                             int64_t ntm, int64_t ntn) {
  constexpr int NW = WR * WC, TI = BM / WR / 32, TJ = BN / WC / 32;
  constexpr int RA = BM / NW, RB = BN / NW;   // rows of A / B each wave stages
  static_assert(RA % 16 == 0 && RB % 16 == 0, "a wave stages whole 16-row loads");
  // [stage][hi/lo][row][k]: rows of 32 halves (four 16-byte segments) with
  // segment s of row r stored at s ^ ((r >> 1) & 3).  LDS serves 128 bytes
  // (32 banks) per cycle, i.e. 8 lanes of a 16-byte access: the fragment
  // reads (8 rows, one segment) then fall on distinct banks
  __shared__ half_t lA[2][2][BM][KB];
  __shared__ half_t lB[2][2][BN][KB];
  auto sw = [](int row, int seg) { return ((seg ^ ((row >> 1) & 3)) * 8); };
  int64_t tm, tn;
  tile_of(blockIdx.x, ntm, ntn, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w / WC, wc = w % WC;
  const half_t *ah = Ahi + tm * BM * Kp, *al = Alo + tm * BM * Kp;
  const half_t *bh = Bhi + tn * BN * Kp, *bl = Blo + tn * BN * Kp;
  f32x16 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; i++)
#pragma unroll
    for (int j = 0; j < TJ; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;
  // lane i of a 16-row load lands on row i/4, physical segment i%4, so it
  // fetches that row's logical segment
  const int drow = lane >> 2;
  auto fill = [&](int st, int64_t k0) {
#pragma unroll
    for (int s = 0; s < RA / 16; s++) {
      const int row = RA * w + 16 * s + drow;
      const int64_t off = (int64_t)row * Kp + k0 + ((lane & 3) ^ ((row >> 1) & 3)) * 8;
      dma16(ah + off, &lA[st][0][RA * w + 16 * s][0]);
      dma16(al + off, &lA[st][1][RA * w + 16 * s][0]);
    }
#pragma unroll
    for (int s = 0; s < RB / 16; s++) {
      const int row = RB * w + 16 * s + drow;
      const int64_t off = (int64_t)row * Kp + k0 + ((lane & 3) ^ ((row >> 1) & 3)) * 8;
      dma16(bh + off, &lB[st][0][RB * w + 16 * s][0]);
      dma16(bl + off, &lB[st][1][RB * w + 16 * s][0]);
    }
  };
  fill(0, 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int cur = 0;
  // 32x32x16 operand layout: lane l holds row (l % 32), k (l / 32) * 8 .. +7
  const int lr = lane & 31, lseg = lane >> 5;
  for (int64_t k0 = 0; k0 < Kp; k0 += KB) {
    // the other stage was last read in the previous step, before the barrier
    if (k0 + KB < Kp) fill(cur ^ 1, k0 + KB);
#pragma unroll
    for (int ks = 0; ks < KB / 16; ks++) {
      half8 fbh[TJ], fbl[TJ];
#pragma unroll
      for (int j = 0; j < TJ; j++) {
        const int br = wc * 32 * TJ + j * 32 + lr;
        fbh[j] = *reinterpret_cast<const half8 *>(&lB[cur][0][br][sw(br, ks * 2 + lseg)]);
        fbl[j] = *reinterpret_cast<const half8 *>(&lB[cur][1][br][sw(br, ks * 2 + lseg)]);
      }
#pragma unroll
      for (int i = 0; i < TI; i++) {
        const int ar = wr * 32 * TI + i * 32 + lr;
        const half8 fah = *reinterpret_cast<const half8 *>(&lA[cur][0][ar][sw(ar, ks * 2 + lseg)]);
        const half8 fal = *reinterpret_cast<const half8 *>(&lA[cur][1][ar][sw(ar, ks * 2 + lseg)]);
        // product by product over the row's TJ accumulators
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fal, fbh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fah, fbl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fah, fbh[j], acc[i][j], 0, 0, 0);
      }
    }
    // the next stage's loads have landed and every wave is done with this one
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    cur ^= 1;
  }
  // D layout of a 32x32 tile: lane l, reg r -> row 8*(r/4) + 4*(l/32) + r%4, col l%32
#pragma unroll
  for (int i = 0; i < TI; i++)
#pragma unroll
    for (int j = 0; j < TJ; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int64_t h = tm * BM + wr * 32 * TI + i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
        const int64_t wcol = tn * BN + wc * 32 * TJ + j * 32 + (lane & 31);
        if (h < p.H && wcol < p.W)
          c[p.off[2] + wcol * p.c_w + h * p.c_h] = mm_finish(acc[i][j][r], dA[h], dB[wcol], sA[h], sB[wcol],
                                                             rowflag, h);
      }
}

// FP8 e4m3: one product per tile, pipelined like split16_gemm
__global__ __launch_bounds__(256) void fp8_gemm(const uint8_t *__restrict__ A8, const uint8_t *__restrict__ B8,
                                               const float *__restrict__ sA, const float *__restrict__ sB,
                                               const float *__restrict__ dA, const float *__restrict__ dB,
                                               int *rowflag, float *__restrict__ c, MmPlan p, int64_t Kp,
                                               int64_t ntm, int64_t ntn) {
  __shared__ uint8_t lA[MB][KB + 16];
  __shared__ uint8_t lB[NB][KB + 16];
  int64_t tm, tn;
  tile_of(blockIdx.x, ntm, ntn, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w >> 1, wc = w & 1;
  const uint8_t *a8 = A8 + tm * MB * Kp, *b8 = B8 + tn * NB * Kp;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  // 128 rows x 32 bytes = 256 x 16 B per array: one per thread
  const int row = tid >> 1, seg = (tid & 1) * 16;
  uint4 ra = *reinterpret_cast<const uint4 *>(a8 + row * Kp + seg);
  uint4 rb = *reinterpret_cast<const uint4 *>(b8 + row * Kp + seg);
  for (int64_t k0 = 0; k0 < Kp; k0 += KB) {
    *reinterpret_cast<uint4 *>(&lA[row][seg]) = ra;
    *reinterpret_cast<uint4 *>(&lB[row][seg]) = rb;
    __syncthreads();
    if (k0 + KB < Kp) {
      ra = *reinterpret_cast<const uint4 *>(a8 + row * Kp + k0 + KB + seg);
      rb = *reinterpret_cast<const uint4 *>(b8 + row * Kp + k0 + KB + seg);
    }
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    long fa[4], fb[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      fa[i] = *reinterpret_cast<const long *>(&lA[wr * 64 + i * 16 + lr][lk]);
      fb[i] = *reinterpret_cast<const long *>(&lB[wc * 64 + i * 16 + lr][lk]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(fa[i], fb[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t h = tm * MB + wr * 64 + i * 16 + 4 * (lane >> 4) + r;
        const int64_t wcol = tn * NB + wc * 64 + j * 16 + (lane & 15);
        if (h < p.H && wcol < p.W)
          c[p.off[2] + wcol * p.c_w + h * p.c_h] = mm_finish(acc[i][j][r], dA[h], dB[wcol], sA[h], sB[wcol],
                                                             rowflag, h);
      }
}

// FP8 on v_mfma_scale_f32_32x32x64_f8f6f4 (e4m3 operands, unit block scales:
// rows and columns are pre-scaled by powers of two as for split16), the same
// DMA two-stage pipeline over 64-byte k-steps; 256x256 blocks of 8 waves,
// each 64x128 of C as 2x4 tiles of 32x32.
typedef int i32x8 __attribute__((ext_vector_type(8)));
constexpr int KB8 = 64;

template <int BM, int BN, int WR, int WC>
__global__ __launch_bounds__(WR * WC * 64) void fp8_gemm_big(const uint8_t *__restrict__ A8,
                                                             const uint8_t *__restrict__ B8,
                                                             const float *__restrict__ sA,
                                                             const float *__restrict__ sB,
                                                             const float *__restrict__ dA,
                                                             const float *__restrict__ dB, int *rowflag,
                                                             float *__restrict__ c, MmPlan p, int64_t Kp,
                                                             int64_t ntm, int64_t ntn) {
  constexpr int NW = WR * WC, TI = BM / WR / 32, TJ = BN / WC / 32;
  constexpr int RA = BM / NW, RB = BN / NW;
  static_assert(RA % 16 == 0 && RB % 16 == 0, "a wave stages whole 16-row loads");
  // [stage][row][64 bytes], 16-byte segment s of row r at s ^ ((r >> 1) & 3)
  __shared__ uint8_t lA[2][BM][KB8];
  __shared__ uint8_t lB[2][BN][KB8];
  auto sw = [](int row, int seg) { return ((seg ^ ((row >> 1) & 3)) * 16); };
  int64_t tm, tn;
  tile_of(blockIdx.x, ntm, ntn, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w / WC, wc = w % WC;
  const uint8_t *a8 = A8 + tm * BM * Kp, *b8 = B8 + tn * BN * Kp;
  f32x16 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; i++)
#pragma unroll
    for (int j = 0; j < TJ; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;
  const int drow = lane >> 2;
  auto fill = [&](int st, int64_t k0) {
#pragma unroll
    for (int s = 0; s < RA / 16; s++) {
      const int row = RA * w + 16 * s + drow;
      dma16(a8 + (int64_t)row * Kp + k0 + ((lane & 3) ^ ((row >> 1) & 3)) * 16, &lA[st][RA * w + 16 * s][0]);
    }
#pragma unroll
    for (int s = 0; s < RB / 16; s++) {
      const int row = RB * w + 16 * s + drow;
      dma16(b8 + (int64_t)row * Kp + k0 + ((lane & 3) ^ ((row >> 1) & 3)) * 16, &lB[st][RB * w + 16 * s][0]);
    }
  };
  fill(0, 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  int cur = 0;
  // 32x32x64 fp8 operand layout: lane l holds row (l % 32), k bytes (l / 32) * 32 .. +31
  const int lr = lane & 31, sg = (lane >> 5) * 2;
  auto frag = [&](const uint8_t *rowp, int row) {
    const uint4 x = *reinterpret_cast<const uint4 *>(rowp + sw(row, sg));
    const uint4 y = *reinterpret_cast<const uint4 *>(rowp + sw(row, sg + 1));
    i32x8 v;
    v[0] = (int)x.x; v[1] = (int)x.y; v[2] = (int)x.z; v[3] = (int)x.w;
    v[4] = (int)y.x; v[5] = (int)y.y; v[6] = (int)y.z; v[7] = (int)y.w;
    return v;
  };
  for (int64_t k0 = 0; k0 < Kp; k0 += KB8) {
    if (k0 + KB8 < Kp) fill(cur ^ 1, k0 + KB8);
    i32x8 fb[TJ];
#pragma unroll
    for (int j = 0; j < TJ; j++) {
      const int br = wc * 32 * TJ + j * 32 + lr;
      fb[j] = frag(&lB[cur][br][0], br);
    }
#pragma unroll
    for (int i = 0; i < TI; i++) {
      const int ar = wr * 32 * TI + i * 32 + lr;
      const i32x8 fa = frag(&lA[cur][ar][0], ar);
#pragma unroll
      for (int j = 0; j < TJ; j++)
        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa, fb[j], acc[i][j], 0, 0, 0, 127, 0, 127);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int i = 0; i < TI; i++)
#pragma unroll
    for (int j = 0; j < TJ; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int64_t h = tm * BM + wr * 32 * TI + i * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
        const int64_t wcol = tn * BN + wc * 32 * TJ + j * 32 + (lane & 31);
        if (h < p.H && wcol < p.W)
          c[p.off[2] + wcol * p.c_w + h * p.c_h] = mm_finish(acc[i][j][r], dA[h], dB[wcol], sA[h], sB[wcol],
                                                             rowflag, h);
      }
}

// PDL's loop for one output (Primitive.pd:226-242 without bad values):
// multiply, then add, in ascending t, in float; eight loads in flight
__device__ __forceinline__ float pdl_dot_f32(const float *a, int64_t at, const float *b, int64_t bt, int64_t T) {
  float cc = 0.0f;
  int64_t t = 0;
  for (; t + 8 <= T; t += 8) {
    float x[8], y[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      x[k] = a[(t + k) * at];
      y[k] = b[(t + k) * bt];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) cc = cc + x[k] * y[k];
  }
  for (; t < T; t++) cc = cc + a[t * at] * b[t * bt];
  return cc;
}

// Recomputes with PDL's loop the outputs the GEMM epilogue marked: a marked
// row's columns are re-tested with mm_needs_fix on the stored values,
// FX_S x 256 columns at a time with all of their loads in flight, and the
// marked ones listed in LDS.  Many marked (a row with an Inf or NaN): each
// thread computes its own.  Few: each wave computes one output at a time,
// its 64 lanes loading 64 products per step, FX_U steps at a time, which
// the wave adds in ascending t (constant lane indices for full steps).
// Blocks stride over the rows, so every block reaches the end.
constexpr int FX_S = 16, FX_U = 16;
__global__ __launch_bounds__(256) void mm_fixup(const float *a, const float *b, float *c, MmPlan p,
                                                const float *__restrict__ sA, const float *__restrict__ sB,
                                                const float *__restrict__ dA, const float *__restrict__ dB,
                                                const int *__restrict__ rowflag) {
  __shared__ int list[256 * FX_S];
  __shared__ int nlist;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int64_t h = blockIdx.x; h < p.H; h += gridDim.x) {
    if (!rowflag[h]) continue;
    const float *arow = a + p.off[0] + h * p.a_h;
    const int eA = pow2_exp(sA[h]);
    const float da = dA[h];
    float *crow = c + p.off[2] + h * p.c_h;
    for (int64_t w0 = 0; w0 < p.W; w0 += 256 * FX_S) {
      if (tid == 0) nlist = 0;
      __syncthreads();
      float cv[FX_S], dv[FX_S], sv[FX_S];
#pragma unroll
      for (int k = 0; k < FX_S; k++) {
        const int64_t col = w0 + k * 256 + tid;
        const bool in = col < p.W;
        cv[k] = in ? crow[col * p.c_w] : 0.0f;
        dv[k] = in ? dB[col] : 0.0f;
        sv[k] = in ? sB[col] : 1.0f;
      }
      unsigned marks = 0;
#pragma unroll
      for (int k = 0; k < FX_S; k++)
        if (w0 + k * 256 + tid < p.W && mm_needs_fix(cv[k], da + dv[k], eA + pow2_exp(sv[k]))) marks |= 1u << k;
      const int mine = __builtin_popcount(marks);
      if (mine) {
        int at = atomicAdd(&nlist, mine);
        for (unsigned m = marks; m; m &= m - 1) list[at++] = __builtin_ctz(m) * 256 + tid;
      }
      __syncthreads();
      const int n = nlist;
      if (n >= 128) {
        for (unsigned m = marks; m; m &= m - 1) {
          const int64_t col = w0 + __builtin_ctz(m) * 256 + tid;
          crow[col * p.c_w] = pdl_dot_f32(arow, p.a_t, b + p.off[1] + col * p.b_w, p.b_t, p.T);
        }
      } else {
        for (int i = w; i < n; i += 4) {
          const int64_t col = w0 + list[i];
          const float *bcol = b + p.off[1] + col * p.b_w;
          auto prod_at = [&](int64_t t0) {
            const int64_t t = t0 + lane;
            return t < p.T ? arow[t * p.a_t] * bcol[t * p.b_t] : 0.0f;
          };
          float cc = 0.0f;
          for (int64_t t0 = 0; t0 < p.T; t0 += 64 * FX_U) {
            float v[FX_U];
#pragma unroll
            for (int u = 0; u < FX_U; u++) v[u] = prod_at(t0 + 64 * u);
            if (t0 + 64 * FX_U <= p.T) {
#pragma unroll
              for (int u = 0; u < FX_U; u++)
#pragma unroll
                for (int l = 0; l < 64; l++) cc = cc + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[u]), l));
            } else {
#pragma unroll
              for (int u = 0; u < FX_U; u++) {
                const int64_t tb = t0 + 64 * u;
                const int m = p.T - tb < 64 ? (p.T - tb > 0 ? (int)(p.T - tb) : 0) : 64;
                for (int l = 0; l < m; l++) cc = cc + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[u]), l));
              }
            }
          }
          if (lane == 0) crow[col * p.c_w] = cc;
        }
      }
      __syncthreads();   // the list is refilled for the next columns
    }
  }
}

static hipError_t prep_operand(const float *src, int64_t o, int64_t inc_r, int64_t inc_k, int64_t R, int64_t K,
                               int64_t Rp, int64_t Kp, float *scale, float *dsum, int *rowflag, half_t *hi,
                               half_t *lo, uint8_t *q8, bool fp8, hipStream_t s) {
  // rows along unit-stride k that start 16-byte aligned take the vector kernels
  const bool rows16 = inc_k == 1 && ((uintptr_t)(src + o) % 16 == 0) && inc_r % 4 == 0;
  if (rows16)
    hipLaunchKernelGGL(absmax_rows_vec, dim3((unsigned)((R + 3) / 4)), dim3(256), 0, s, src, o, inc_r, R, K, scale);
  else if (inc_k == 1 || inc_k == -1)
    hipLaunchKernelGGL(absmax_rows_wave, dim3((unsigned)((R + 3) / 4)), dim3(256), 0, s, src, o, inc_r, inc_k, R, K,
                       scale);
  else {
    const int64_t nr = (R + 63) / 64, nk = (K + 63) / 64;
    hipLaunchKernelGGL(zero_u32, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, (unsigned *)scale, R);
    hipLaunchKernelGGL(absmax_tile, dim3((unsigned)(nr * nk)), dim3(256), 0, s, src, o, inc_r, inc_k, R, K,
                       (unsigned *)scale, nk);
  }
  hipLaunchKernelGGL(make_scales, dim3((unsigned)((Rp + 255) / 256)), dim3(256), 0, s, scale, dsum, rowflag, R, Rp,
                     fp8 ? 8 : 15);
  if (rows16) {
    int64_t blocks = (Rp * (Kp / 8) + 255) / 256;
    blocks = blocks > 65536 ? 65536 : blocks;
    if (fp8)
      hipLaunchKernelGGL((pack_rows<true>), dim3((unsigned)blocks), dim3(256), 0, s, src, o, inc_r, R, K, Rp, Kp, scale,
                         dsum, hi, lo, q8);
    else
      hipLaunchKernelGGL((pack_rows<false>), dim3((unsigned)blocks), dim3(256), 0, s, src, o, inc_r, R, K, Rp, Kp,
                         scale, dsum, hi, lo, q8);
    return hipGetLastError();
  }
  const int64_t ntr = (Rp + 63) / 64, ntk = (Kp + 63) / 64;
  if (fp8)
    hipLaunchKernelGGL((pack_tile<true>), dim3((unsigned)(ntr * ntk)), dim3(256), 0, s, src, o, inc_r, inc_k, R, K, Kp,
                       scale, dsum, hi, lo, q8, ntk);
  else
    hipLaunchKernelGGL((pack_tile<false>), dim3((unsigned)(ntr * ntk)), dim3(256), 0, s, src, o, inc_r, inc_k, R, K,
                       Kp, scale, dsum, hi, lo, q8, ntk);
  return hipGetLastError();
}

// Small float products (16 <= H, W, T <= 64), typically broadcast batches:
// a block of 4 waves per product (a grid capped at 2^31 - 1 blocks loops,
// fetching its next product's operands while it multiplies the current
// one; a persistent grid of resident blocks measured slower), nothing
// staged outside the kernel.  A
// (rows h, k = t) and B (rows w, k = t, transposed on the way in) are read
// once in FP32 into LDS; each row of A and each column of B gets the same
// power-of-two scale the large kernels use; every wave then owns one 32x32
// quarter of C and, per 16-k step, reads its fragments from LDS, scales and
// splits them into FP16 hi/lo in registers and issues the three
// v_mfma_f32_32x32x16_f16 products.  A's rows of 68 floats keep its 16-byte
// fragment reads aligned and on distinct banks; B's rows of 65 floats keep
// its transposing column writes on distinct banks.  C is unscaled by the
// scales' exact reciprocals.
constexpr int SMM = 64, SMP = 68, SMPB = 65;

template <bool VEC>
__global__ __launch_bounds__(256, 4) void split16_small(const float *__restrict__ a, const float *__restrict__ b,
                                                        float *__restrict__ c, MmPlan p) {
  __shared__ __align__(16) float fA[SMM][SMP];
  __shared__ __align__(16) float fB[SMM][SMPB];
  __shared__ float scA[SMM], scB[SMM], dsA[SMM], dsB[SMM];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int H = (int)p.H, W = (int)p.W, T = (int)p.T, Tp = (T + 15) & ~15;
  // staging: thread (r0, q) holds rows r0 + 4j of A(h = r, t = q) and of
  // B(t = r, w = q), loaded from indices clamped into the operand (zeroed
  // when staged) so that no load sits behind a branch
  const int q = tid & 63, r0 = tid >> 6;
  const int qa = q < T ? q : T - 1, qb = q < W ? q : W - 1;
  // a product's own extent fits 32-bit offsets (checked by the launcher)
  const int a_h = (int)p.a_h, b_t = (int)p.b_t;
  const int qoa = qa * (int)p.a_t, qob = qb * (int)p.b_w;
  float va[16], vb[16];
  auto fetch = [&](int64_t bi) {
    int64_t o[3];
    mm_batch_offsets_m(p, bi, o);
    const float *pa = a + o[0], *pb = b + o[1];
    // the row index re-read per product keeps the compiler from hoisting
    // (and holding) 32 per-load offsets across the loop
    int rr = r0;
    asm volatile("" : "+v"(rr));
    if constexpr (VEC) {
      // 64x64 unit-stride aligned operands: thread (row tid/16 + 16j, 4
      // consecutive k or w at (tid%16)*4) with 16-byte loads
      const int rv = (tid >> 4) + (rr - r0), cq = (tid & 15) * 4;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const Vec<float, 4> x = *reinterpret_cast<const Vec<float, 4> *>(pa + (rv + 16 * j) * a_h + cq);
        const Vec<float, 4> y = *reinterpret_cast<const Vec<float, 4> *>(pb + (rv + 16 * j) * b_t + cq);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          va[4 * j + e] = x.v[e];
          vb[4 * j + e] = y.v[e];
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int r = rr + 4 * j;
        const int ra = r < H ? r : H - 1, rb = r < T ? r : T - 1;
        va[j] = pa[ra * a_h + qoa];
        vb[j] = pb[rb * b_t + qob];
      }
    }
  };
  const int wr = w >> 1, wc = w & 1;
  const int ar = 32 * wr + (lane & 31), br = 32 * wc + (lane & 31), kq = (lane >> 5) * 8;
  // products bi, bi + gridDim.x, ...; the next one's operands are in flight
  // while this one is scaled, multiplied and stored
  int64_t bi = blockIdx.x;
  if (bi < p.nbatch) fetch(bi);
  for (; bi < p.nbatch; bi += gridDim.x) {
    // per-thread LDS and store addresses are recomputed each product rather
    // than held in registers across the loop
    int tl = tid;
    asm volatile("" : "+v"(tl));
    const int ql = tl & 63, rl = tl >> 6, lanel = tl & 63;
    if constexpr (VEC) {
      const int rv = tl >> 4, cq = (tl & 15) * 4;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        Vec<float, 4> x;
#pragma unroll
        for (int e = 0; e < 4; e++) x.v[e] = va[4 * j + e];
        *reinterpret_cast<Vec<float, 4> *>(&fA[rv + 16 * j][cq]) = x;
#pragma unroll
        for (int e = 0; e < 4; e++) fB[cq + e][rv + 16 * j] = vb[4 * j + e];
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int r = rl + 4 * j;
        fA[r][ql] = (r < H && ql < T) ? va[j] : 0.0f;
        fB[ql][r] = (ql < W && r < T) ? vb[j] : 0.0f;
      }
    }
    __syncthreads();
    if (bi + gridDim.x < p.nbatch) fetch(bi + gridDim.x);
    // row scales: four threads per row of fA and of fB, 16 values each
    // (16-byte LDS reads), combined within the quad; then each row's
    // range-loss excess (scaled units, +Inf for a non-finite element) over
    // the same 16 values, summed within the quad
    {
      const int r = tl >> 2, k0 = (tl & 3) * 16;
      // largest finite and smallest nonzero magnitudes (0 for a row with an
      // Inf or NaN, which takes the excess pass below)
      float ma = 0.0f, mb = 0.0f, na = __builtin_inff(), nb = __builtin_inff();
      auto low = [](float ax) { return ax < __builtin_inff() ? (ax > 0.0f ? ax : __builtin_inff()) : 0.0f; };
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const Vec<float, 4> v = *reinterpret_cast<const Vec<float, 4> *>(&fA[r][k0 + 4 * g]);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const float xa = fabsf(v.v[e]), xb = fabsf(fB[r][k0 + 4 * g + e]);
          ma = fmaxf(ma, fin_abs(xa));
          mb = fmaxf(mb, fin_abs(xb));
          na = fminf(na, low(xa));
          nb = fminf(nb, low(xb));
        }
      }
#pragma unroll
      for (int x = 1; x <= 2; x <<= 1) {
        ma = fmaxf(ma, __shfl_xor(ma, x, 64));
        mb = fmaxf(mb, __shfl_xor(mb, x, 64));
        na = fminf(na, __shfl_xor(na, x, 64));
        nb = fminf(nb, __shfl_xor(nb, x, 64));
      }
      const float s1 = pow2_scale(ma, 15), s2 = pow2_scale(mb, 15);
      // an element's hi/lo split loses bits beyond 2^-22 only when its
      // scaled magnitude is below 2^-3 (lo then falls among FP16's
      // subnormals): rows without one skip the excess pass
      float ea = 0.0f, eb = 0.0f;
      if (na * s1 < 0.125f)
        for (int e = 0; e < 16; e++) {
          const float xa = fA[r][k0 + e];
          half_t h, l;
          ea += fabsf(xa) < __builtin_inff() ? split_hilo(xa * s1, h, l) : __builtin_inff();
        }
      if (nb * s2 < 0.125f)
        for (int e = 0; e < 16; e++) {
          const float xb = fB[r][k0 + e];
          half_t h, l;
          eb += fabsf(xb) < __builtin_inff() ? split_hilo(xb * s2, h, l) : __builtin_inff();
        }
      ea += __shfl_xor(ea, 1, 64);
      eb += __shfl_xor(eb, 1, 64);
      ea += __shfl_xor(ea, 2, 64);
      eb += __shfl_xor(eb, 2, 64);
      if ((tl & 3) == 0) {
        scA[r] = s1;
        scB[r] = s2;
        dsA[r] = ea;
        dsB[r] = eb;
      }
    }
    __syncthreads();
    const float sa = scA[ar], sb = scB[br];
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    for (int k0 = 0; k0 < Tp; k0 += 16) {
      half8 ah, al, bh, bl;
      const Vec<float, 4> a0 = *reinterpret_cast<const Vec<float, 4> *>(&fA[ar][k0 + kq]);
      const Vec<float, 4> a1 = *reinterpret_cast<const Vec<float, 4> *>(&fA[ar][k0 + kq + 4]);
#pragma unroll
      for (int i = 0; i < 8; i++) {
        // non-finite elements enter as zero (their rows are recomputed)
        const float xa = i < 4 ? a0.v[i] : a1.v[i - 4], xb = fB[br][k0 + kq + i];
        const float ya = fabsf(xa) < __builtin_inff() ? xa * sa : 0.0f;
        const float yb = fabsf(xb) < __builtin_inff() ? xb * sb : 0.0f;
        const half_t ha = (half_t)ya, hb = (half_t)yb;
        ah[i] = ha;
        al[i] = (half_t)(ya - (float)ha);
        bh[i] = hb;
        bl[i] = (half_t)(yb - (float)hb);
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc, 0, 0, 0);
    }
    int64_t o[3];
    mm_batch_offsets_m(p, bi, o);
    // D layout of a 32x32 tile: lane l, reg r -> row 8*(r/4) + 4*(l/32) + r%4, col l%32.
    // Elements whose range-loss bound exceeds 2^-22 |c| (see mm_finish) are
    // noted in a mask and then recomputed with PDL's loop from the staged
    // FP32 operands, once the accumulators are no longer live
    unsigned redo = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int h = 32 * wr + 8 * (r / 4) + 4 * (lanel >> 5) + (r & 3);
      const int wcol = 32 * wc + (lanel & 31);
      if (h < H && wcol < W) {
        const float v = acc[r], dd = dsA[h] + dsB[wcol];
        if ((dd != 0.0f || !(fabsf(v) < __builtin_inff())) && !(dd * 32768.0f <= 0x1p-22f * fabsf(v)))
          redo |= 1u << r;
        c[o[2] + wcol * p.c_w + h * p.c_h] = ldexpf(v, -(pow2_exp(scA[h]) + pow2_exp(scB[wcol])));
      }
    }
    while (redo) {
      const int r = __builtin_ctz(redo);
      redo &= redo - 1;
      const int h = 32 * wr + 8 * (r / 4) + 4 * (lanel >> 5) + (r & 3);
      const int wcol = 32 * wc + (lanel & 31);
      float out = 0.0f;
      for (int t = 0; t < T; t++) out = out + fA[h][t] * fB[wcol][t];
      c[o[2] + wcol * p.c_w + h * p.c_h] = out;
    }
    __syncthreads();   // fA, fB and the scales are refilled next
  }
}

// ---------------------------------------------------------------- FP64
// Double matmult on v_mfma_f64_16x16x4_f64 (FP64 products and sums; lane l
// holds A[l%16][l/16] and B[l/16][l%16]; result register r of lane l is
// D[4r + l/16][l%16]).  128x128 blocks of 4 waves, each wave 64x64 as 4x4
// tiles; stages of 16 k go global -> registers -> LDS as [k][row] for A
// (transposed on the way in) and [k][col] for B, so that every fragment
// read is 16 consecutive doubles; the next stage's loads are in flight
// while the matrix cores work on the current one.
typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int DKB = 16;

template <int BM, int BN, int WR, int WC>
__global__ __launch_bounds__(WR * WC * 64, 2) void dgemm_mfma(const double *__restrict__ a,
                                                              const double *__restrict__ b, double *__restrict__ c,
                                                              MmPlan p, int64_t ntm, int64_t ntn) {
  constexpr int NT = WR * WC * 64, TI = BM / WR / 16, TJ = BN / WC / 16;
  constexpr int PA = BM * DKB / (2 * NT), PB = BN * DKB / (2 * NT);   // passes of 2 doubles per thread
  __shared__ __align__(16) double sA[2][DKB][BM];
  __shared__ __align__(16) double sB[2][DKB][BN];
  int64_t tm, tn;
  tile_of(blockIdx.x, ntm, ntn, tm, tn);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = w / WC, wc = w % WC;
  const int64_t m0 = tm * BM, n0 = tn * BN;
  // A: thread (row tid / 8 + rows-per-pass * pass, k pair (tid % 8) * 2), lanes along k
  // B: thread (k tid / (BN/2) + ..., col pair (tid % (BN/2)) * 2), lanes along w
  double ra[PA][2], rb[PB][2];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < PA; s++) {
      const int row = tid / 8 + (NT / 8) * s, kk = (tid % 8) * 2;
      const int64_t h = m0 + row;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int64_t t = k0 + kk + e;
        ra[s][e] = (h < p.H && t < p.T) ? a[p.off[0] + h * p.a_h + t * p.a_t] : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < PB; s++) {
      const int kr = tid / (BN / 2) + (NT / (BN / 2)) * s, col = (tid % (BN / 2)) * 2;
      const int64_t t = k0 + kr;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int64_t wv = n0 + col + e;
        rb[s][e] = (wv < p.W && t < p.T) ? b[p.off[1] + wv * p.b_w + t * p.b_t] : 0.0;
      }
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int s = 0; s < PA; s++) {
      const int row = tid / 8 + (NT / 8) * s, kk = (tid % 8) * 2;
      sA[buf][kk][row] = ra[s][0];
      sA[buf][kk + 1][row] = ra[s][1];
    }
#pragma unroll
    for (int s = 0; s < PB; s++) {
      const int kr = tid / (BN / 2) + (NT / (BN / 2)) * s, col = (tid % (BN / 2)) * 2;
      sB[buf][kr][col] = rb[s][0];
      sB[buf][kr][col + 1] = rb[s][1];
    }
  };
  d4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; i++)
#pragma unroll
    for (int j = 0; j < TJ; j++) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  load(0);
  stage(0);
  __syncthreads();
  const int fr = lane & 15, fk = lane >> 4;
  int cur = 0;
  for (int64_t k0 = 0; k0 < p.T; k0 += DKB) {
    const bool more = k0 + DKB < p.T;
    if (more) load(k0 + DKB);
#pragma unroll
    for (int kk = 0; kk < DKB; kk += 4) {
      double af[TI], bf[TJ];
#pragma unroll
      for (int i = 0; i < TI; i++) af[i] = sA[cur][kk + fk][wr * 16 * TI + i * 16 + fr];
#pragma unroll
      for (int j = 0; j < TJ; j++) bf[j] = sB[cur][kk + fk][wc * 16 * TJ + j * 16 + fr];
#pragma unroll
      for (int i = 0; i < TI; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (more) stage(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int i = 0; i < TI; i++)
#pragma unroll
    for (int j = 0; j < TJ; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t h = m0 + wr * 16 * TI + i * 16 + 4 * r + fk;
        const int64_t wv = n0 + wc * 16 * TJ + j * 16 + fr;
        if (h < p.H && wv < p.W) c[p.off[2] + wv * p.c_w + h * p.c_h] = acc[i][j][r];
      }
}

hipError_t launch_mfma_f64(const MmPlan &p, const double *a, const double *b, double *c, hipStream_t s) {
  // one product per call, where a 128x128 block is mostly useful and there
  // are at least 64 of them (fewer leave most CUs idle: the SIMT kernel's
  // smaller tiles win below about 1024x1024)
  // variant 1: the SIMT kernel; 2: the matrix cores whatever the block count
  if (p.variant == 1 || p.nbatch != 1 || p.H < 64 || p.W < 64 || p.T < 16) return hipErrorNotSupported;
  const int64_t ntm = (p.H + 127) / 128, ntn = (p.W + 127) / 128;
  if ((ntm * ntn < 64 && p.variant != 2) || ntm * ntn >= ((int64_t)1 << 31)) return hipErrorNotSupported;
  hipLaunchKernelGGL((dgemm_mfma<128, 128, 2, 2>), dim3((unsigned)(ntm * ntn)), dim3(256), 0, s, a, b, c, p, ntm, ntn);
  return hipGetLastError();
}

hipError_t launch_mfma_f32(const MmPlan &p, const float *a, const float *b, float *c, int mode, hipStream_t s) {
  const bool fp8 = (mode == PDLHIP_MM_FP8);
  // small products (broadcast batches of them in one launch)
  auto fits32 = [](int64_t n1, int64_t i1, int64_t n2, int64_t i2) {
    const int64_t lim = (int64_t)1 << 30;
    return i1 < lim && i1 > -lim && i2 < lim && i2 > -lim && (n1 - 1) * (i1 < 0 ? -i1 : i1) + (n2 - 1) * (i2 < 0 ? -i2 : i2) < lim;
  };
  if (!fp8 && p.H >= 16 && p.W >= 16 && p.T >= 16 && p.H <= SMM && p.W <= SMM && p.T <= SMM &&
      fits32(p.H, p.a_h, p.T, p.a_t) && fits32(p.W, p.b_w, p.T, p.b_t)) {
    // a block per product (the block loop then runs once); more products
    // than 2^31 - 1 loop
    const int64_t blocks = p.nbatch < ((int64_t)1 << 31) - 1 ? p.nbatch : ((int64_t)1 << 31) - 1;
    // 64x64x64 products with unit-stride k (A) and w (B), every product's
    // rows 16-byte aligned: 16-byte staging loads
    bool vec = p.H == SMM && p.W == SMM && p.T == SMM && p.a_t == 1 && p.b_w == 1 && p.a_h % 4 == 0 &&
               p.b_t % 4 == 0 && ((uintptr_t)(a + p.off[0])) % 16 == 0 && ((uintptr_t)(b + p.off[1])) % 16 == 0;
    for (int d = 0; d < p.nd && vec; d++) vec = p.inc[0][d] % 4 == 0 && p.inc[1][d] % 4 == 0;
    if (vec) hipLaunchKernelGGL(split16_small<true>, dim3((unsigned)blocks), dim3(256), 0, s, a, b, c, p);
    else hipLaunchKernelGGL(split16_small<false>, dim3((unsigned)blocks), dim3(256), 0, s, a, b, c, p);
    return hipGetLastError();
  }
  // otherwise one matrix product per call (broadcast batches use the SIMT
  // kernel), and only where a 128x128 tile is mostly useful
  if (p.nbatch != 1 || p.H < 64 || p.W < 64 || p.T < 32) return hipErrorNotSupported;
  // 256x256 blocks once both sides fill them (FP8 then steps k by 64), or as
  // the autotuner picked (variant 128 / 256)
  const bool big = p.variant ? p.variant == 256 : (p.H >= 1024 && p.W >= 1024);
  const int64_t BMr = big ? 256 : MB, BNr = big ? 256 : NB, KS = (big && fp8) ? KB8 : KB;
  const int64_t Hp = (p.H + BMr - 1) / BMr * BMr, Wp = (p.W + BNr - 1) / BNr * BNr, Kp = (p.T + KS - 1) / KS * KS;
  const size_t esz = fp8 ? 1 : 2, ncopy = fp8 ? 1 : 2;
  // [scales A | scales B | range-loss sums A | B | recompute marks per row
  // of A] then the packed panels
  const size_t head = 2 * (size_t)(Hp + Wp) * sizeof(float) + (size_t)Hp * sizeof(int);
  char *buf = (char *)scratch_alloc((size_t)(Hp + Wp) * Kp * esz * ncopy + head + 256);
  if (!buf) return hipErrorOutOfMemory;
  float *sA = (float *)buf, *sB = sA + Hp, *dA = sB + Wp, *dB = dA + Hp;
  int *rowflag = (int *)(dB + Wp);
  char *q = buf + (head + 255) / 256 * 256;
  half_t *Ahi = nullptr, *Alo = nullptr, *Bhi = nullptr, *Blo = nullptr;
  uint8_t *A8 = nullptr, *B8 = nullptr;
  if (fp8) {
    A8 = (uint8_t *)q;
    B8 = A8 + Hp * Kp;
  } else {
    Ahi = (half_t *)q;
    Alo = Ahi + Hp * Kp;
    Bhi = Alo + Hp * Kp;
    Blo = Bhi + Wp * Kp;
  }
  int64_t o[3] = {p.off[0], p.off[1], p.off[2]};
  // A rows = h (stride a_h), k = t (stride a_t); B "rows" = w (b_w), k = t (b_t)
  hipError_t e = prep_operand(a, o[0], p.a_h, p.a_t, p.H, p.T, Hp, Kp, sA, dA, rowflag, Ahi, Alo, A8, fp8, s);
  if (e == hipSuccess) e = prep_operand(b, o[1], p.b_w, p.b_t, p.W, p.T, Wp, Kp, sB, dB, nullptr, Bhi, Blo, B8, fp8, s);
  const int64_t ntm = Hp / BMr, ntn = Wp / BNr;
  if (e == hipSuccess) {
    if (fp8 && big)
      hipLaunchKernelGGL((fp8_gemm_big<256, 256, 4, 2>), dim3((unsigned)(ntm * ntn)), dim3(512), 0, s, A8, B8, sA, sB,
                         dA, dB, rowflag, c, p, Kp, ntm, ntn);
    else if (fp8)
      hipLaunchKernelGGL(fp8_gemm, dim3((unsigned)(ntm * ntn)), dim3(256), 0, s, A8, B8, sA, sB, dA, dB, rowflag, c,
                         p, Kp, ntm, ntn);
    else if (big)
      hipLaunchKernelGGL((split16_gemm<256, 256, 4, 2>), dim3((unsigned)(ntm * ntn)), dim3(512), 0, s, Ahi, Alo, Bhi,
                         Blo, sA, sB, dA, dB, rowflag, c, p, Kp, ntm, ntn);
    else
      hipLaunchKernelGGL((split16_gemm<128, 128, 2, 2>), dim3((unsigned)(ntm * ntn)), dim3(256), 0, s, Ahi, Alo, Bhi,
                         Blo, sA, sB, dA, dB, rowflag, c, p, Kp, ntm, ntn);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const int64_t fb = p.H < 2048 ? p.H : 2048;
    hipLaunchKernelGGL(mm_fixup, dim3((unsigned)fb), dim3(256), 0, s, a, b, c, p, sA, sB, dA, dB, rowflag);
    e = hipGetLastError();
  }
  scratch_free(buf);
  return e;
}

}  // namespace pdlhip
