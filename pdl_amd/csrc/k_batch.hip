// Lists of uneven matmults in one persistent launch (the reference runs
// such a list as a Perl loop of matmult calls, lib/PDL/Primitive.pd:197-209
// per item).  The host builds one work item per (product, C tile), ordered
// by decreasing inner dimension (longest-processing-time first); persistent
// blocks take items from a device counter until the queue is drained, so
// every block reaches the exit condition.  Float and double lists in the
// default modes run on the matrix cores (list_mfma_*); PDL order and the
// integer types run the SIMT tile of k_matmult.hip (accumulation order per
// output: ascending t, so the PDL-order variant stays bit-identical to PDL's
// loop).
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

template <class T, bool FUSED, int TS>
__device__ inline void list_tile(const MmItem &it, int64_t tm, int64_t tn, T (*As)[TS + 1], T (*Bs)[TS + 1]) {
  constexpr int KS = 16;
  constexpr int PT = TS / 16;          // outputs per thread per dim (16x16 threads)
  const T *a = (const T *)it.a, *b = (const T *)it.b;
  T *c = (T *)it.c;
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  const int64_t h0 = tm * TS, w0 = tn * TS;
  T acc[PT][PT];
#pragma unroll
  for (int i = 0; i < PT; i++)
#pragma unroll
    for (int j = 0; j < PT; j++) acc[i][j] = T(0);
  constexpr int NE = TS * KS / 256;   // staged elements per thread per operand
  for (int64_t k0 = 0; k0 < it.T; k0 += KS) {
    // every load of the chunk is issued before the first LDS write; indices
    // are clamped into the operand and the element zeroed afterwards, so no
    // load sits behind a branch
    T ra[NE], rb[NE];
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = tid + 256 * q;
      const int kk = e % KS, mm = e / KS, nn = e % TS, kb = e / TS;
      const int64_t gh = h0 + mm, gt = k0 + kk, gw = w0 + nn, gt2 = k0 + kb;
      const int64_t ch = gh < it.H ? gh : it.H - 1, ct = gt < it.T ? gt : it.T - 1;
      const int64_t cw = gw < it.W ? gw : it.W - 1, ct2 = gt2 < it.T ? gt2 : it.T - 1;
      ra[q] = a[it.oa + ct * it.a_t + ch * it.a_h];
      rb[q] = b[it.ob + cw * it.b_w + ct2 * it.b_t];
    }
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = tid + 256 * q;
      const int kk = e % KS, mm = e / KS, nn = e % TS, kb = e / TS;
      As[kk][mm] = (h0 + mm < it.H && k0 + kk < it.T) ? ra[q] : T(0);
      Bs[kb][nn] = (w0 + nn < it.W && k0 + kb < it.T) ? rb[q] : T(0);
    }
    __syncthreads();
    const int kmax = (it.T - k0 < KS) ? (int)(it.T - k0) : KS;
    for (int kk = 0; kk < kmax; kk++) {
      T av[PT], bv[PT];
#pragma unroll
      for (int i = 0; i < PT; i++) av[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < PT; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < PT; i++)
#pragma unroll
        for (int j = 0; j < PT; j++) {
          if constexpr (FUSED && IsFloat<T>::v) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
          else acc[i][j] = wadd(acc[i][j], wmul(av[i], bv[j]));
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < PT; i++) {
    int64_t gh = h0 + ty + 16 * i;
    if (gh >= it.H) continue;
#pragma unroll
    for (int j = 0; j < PT; j++) {
      int64_t gw = w0 + tx + 16 * j;
      if (gw < it.W) c[it.oc + gw * it.c_w + gh * it.c_h] = acc[i][j];
    }
  }
}

// The work list holds LIST_BINS bins back to back, bin x = work[bounds[x]
// .. bounds[x + 1]), one per XCD: the host spreads the products over the
// bins by their work (largest first, each to the lightest bin), so all the
// tiles of a product sit in one bin, in order.  Workgroups go to the 8 XCDs
// round-robin, so block b starts on bin b % 8 and the tiles of a product,
// taken by neighbouring blocks of one XCD at about the same time, re-read
// its A and B panels from that XCD's L2 rather than from HBM; a block whose
// bin is empty goes on to the next bins.  counter[x * LIST_CSTRIDE] is bin
// x's counter, each in a 128-byte line of its own (reset by a kernel before
// each launch); the bounds follow the counters.
constexpr int LIST_BINS = 8, LIST_CSTRIDE = 32, LIST_BOUNDS = LIST_BINS * LIST_CSTRIDE;

__global__ void list_reset(unsigned int *counter) {
  if (threadIdx.x < LIST_BINS && blockIdx.x == 0) counter[threadIdx.x * LIST_CSTRIDE] = 0u;
}

// the block's next work item (false: every bin is empty); thread 0 takes it
// and hands it on through LDS
__device__ __forceinline__ bool list_next(unsigned int *counter, const unsigned int *__restrict__ bounds,
                                          const int3 *__restrict__ work, int &bin, unsigned &done, unsigned *sh,
                                          int3 &job) {
  if (threadIdx.x == 0) {
    unsigned w = ~0u;
    while (done != (1u << LIST_BINS) - 1u) {
      const unsigned i = atomicAdd(counter + bin * LIST_CSTRIDE, 1u);
      if (bounds[bin] + i < bounds[bin + 1]) {
        w = bounds[bin] + i;
        break;
      }
      done |= 1u << bin;
      bin = (bin + 1) % LIST_BINS;
    }
    *sh = w;
  }
  __syncthreads();
  const unsigned w = *sh;
  __syncthreads();
  if (w == ~0u) return false;
  job = work[w];
  return true;
}

// persistent: take work items until every bin is empty
template <class T, bool FUSED, int TS>
__global__ __launch_bounds__(256) void list_queue(const MmItem *__restrict__ items, const int3 *__restrict__ work,
                                                  const unsigned int *__restrict__ bounds, unsigned int *counter) {
  constexpr int KS = 16;
  __shared__ T As[KS][TS + 1];
  __shared__ T Bs[KS][TS + 1];
  __shared__ unsigned int next;
  int bin = blockIdx.x % LIST_BINS;
  unsigned done = 0;
  int3 job;
  while (list_next(counter, bounds, work, bin, done, &next, job))
    list_tile<T, FUSED, TS>(items[job.x], job.y, job.z, As, Bs);
}

// one launch per item (the comparison the list path is measured against)
template <class T, bool FUSED, int TS>
__global__ __launch_bounds__(256) void list_one(const MmItem *__restrict__ items, int idx, int64_t ntn) {
  constexpr int KS = 16;
  __shared__ T As[KS][TS + 1];
  __shared__ T Bs[KS][TS + 1];
  list_tile<T, FUSED, TS>(items[idx], blockIdx.x / ntn, blockIdx.x % ntn, As, Bs);
}

// ---------------------------------------------------------------- matrix cores
// Float and double lists in the default modes run on the matrix cores with
// FP32 / FP64 products and sums (no operand scaling: the same arithmetic as
// the FP32/FP64 SIMT kernel, summed in the matrix core's order): one work
// item = a 64x64 tile of C, four waves of 32x32 each; a wave whose quadrant
// lies outside the product skips its MFMAs.  k chunks (32 floats / 16
// doubles) go global -> registers -> LDS ([k][row], [k][col]) with the next
// chunk's loads in flight while the matrix cores work on the current one
// (two LDS buffers, one barrier per chunk).
//   float:  v_mfma_f32_32x32x2f32, lane l holds A[l%32][k + l/32] and
//           B[k + l/32][l%32]; D reg r -> row 8(r/4) + 4(l/32) + r%4, col l%32
//   double: v_mfma_f64_16x16x4f64 as 2x2 tiles, lane l holds A[l%16][k + l/16]
//           and B[k + l/16][l%16]; D reg r -> row 4r + l/16, col l%16
template <class T> struct ListMfma;
template <> struct ListMfma<float> { static constexpr int KS = 32; };
template <> struct ListMfma<double> { static constexpr int KS = 16; };

typedef float lf32x16 __attribute__((ext_vector_type(16)));
typedef double ld4 __attribute__((ext_vector_type(4)));

template <class T, int KS>
__device__ __forceinline__ void list_mfma_tile(const MmItem &it, int64_t tm, int64_t tn, T (*As)[KS][65],
                                               T (*Bs)[KS][65]) {
  constexpr int NE = KS / 4;            // staged elements per thread per operand (64 x KS / 256)
  constexpr int SK = 256 / KS;          // rows apart of a thread's elements when it runs along k
  // a product's own extent fits 32-bit offsets (checked by the host)
  const T *a = (const T *)it.a + it.oa, *b = (const T *)it.b + it.ob;
  T *c = (T *)it.c;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int H = (int)it.H, W = (int)it.W, K = (int)it.T;
  const int h0 = (int)tm * 64, w0 = (int)tn * 64;
  const int a_t = (int)it.a_t, a_h = (int)it.a_h, b_w = (int)it.b_w, b_t = (int)it.b_t;
  // threads run along whichever index of each operand is unit-stride: a
  // thread keeps one "fast" index (k, or the row/column) and steps the other
  const bool a_kfast = a_t == 1 || a_h != 1;
  const bool b_wfast = b_w == 1 || b_t != 1;
  const int fa = a_kfast ? tid % KS : tid % 64, sa = a_kfast ? tid / KS : tid / 64, da = a_kfast ? SK : 4;
  const int fb = b_wfast ? tid % 64 : tid % KS, sb = b_wfast ? tid / 64 : tid / KS, db = b_wfast ? 4 : SK;
  const bool active = h0 + 32 * wr < H && w0 + 32 * wc < W;
  // a chunk wholly inside both operands loads without clamps or masks: one
  // base address per operand, the thread's elements a constant stride apart
  const bool inside_hw = h0 + 64 <= H && w0 + 64 <= W;
  const int a_step = da * (a_kfast ? a_h : a_t), b_step = db * (b_wfast ? b_t : b_w);
  T ra[NE], rb[NE];
  // edge chunks: loads clamped into the operand (the element is zeroed when
  // staged), so no load sits behind a branch
  auto load = [&](int k0) {
    if (inside_hw && k0 + KS <= K) {
      const T *pa = a + (a_kfast ? (k0 + fa) * a_t + (h0 + sa) * a_h : (k0 + sa) * a_t + (h0 + fa) * a_h);
      const T *pb = b + (b_wfast ? (w0 + fb) * b_w + (k0 + sb) * b_t : (w0 + sb) * b_w + (k0 + fb) * b_t);
#pragma unroll
      for (int q = 0; q < NE; q++) {
        ra[q] = pa[q * a_step];
        rb[q] = pb[q * b_step];
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int s1 = sa + da * q, s2 = sb + db * q;
      const int ah = a_kfast ? h0 + s1 : h0 + fa, ak = a_kfast ? k0 + fa : k0 + s1;
      const int bw = b_wfast ? w0 + fb : w0 + s2, bk = b_wfast ? k0 + s2 : k0 + fb;
      ra[q] = a[(ak < K ? ak : K - 1) * a_t + (ah < H ? ah : H - 1) * a_h];
      rb[q] = b[(bw < W ? bw : W - 1) * b_w + (bk < K ? bk : K - 1) * b_t];
    }
  };
  // interior chunks of the common layout (a's k and b's w unit-stride):
  // every staged element lands at a slot a compile-time distance from the
  // thread's first, with no bounds tests
  const bool plain = a_t == 1 && b_w == 1;
  auto stage = [&](int buf, int k0) {
    const bool inside = inside_hw && k0 + KS <= K;
    if (plain && inside) {
#pragma unroll
      for (int q = 0; q < NE; q++) {
        As[buf][fa][sa + SK * q] = ra[q];
        Bs[buf][sb + 4 * q][fb] = rb[q];
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int s1 = sa + da * q, s2 = sb + db * q;
      const int mm = a_kfast ? s1 : fa, kk = a_kfast ? fa : s1;
      const int nn = b_wfast ? fb : s2, kb = b_wfast ? s2 : fb;
      As[buf][kk][mm] = (inside || (h0 + mm < H && k0 + kk < K)) ? ra[q] : T(0);
      Bs[buf][kb][nn] = (inside || (w0 + nn < W && k0 + kb < K)) ? rb[q] : T(0);
    }
  };
  if constexpr (sizeof(T) == 4) {
    lf32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    // an empty inner dimension loads nothing (its clamp would reach index
    // -1) and stores the zero sums
    if (K > 0) load(0);
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += KS) {
      stage(buf, k0);
      __syncthreads();
      if (k0 + KS < K) load(k0 + KS);
      // every wave multiplies, also one whose quarter lies outside the
      // product (its LDS rows are zeros and it stores nothing): a branch
      // here makes the compiler copy the accumulators out of AGPRs and back
      // each chunk
#pragma unroll
      for (int kk = 0; kk < KS; kk += 2) {
        const float av = As[buf][kk + (lane >> 5)][32 * wr + (lane & 31)];
        const float bv = Bs[buf][kk + (lane >> 5)][32 * wc + (lane & 31)];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
      }
      buf ^= 1;
    }
    if (active) {
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int64_t gh = h0 + 32 * wr + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
        const int64_t gw = w0 + 32 * wc + (lane & 31);
        if (gh < H && gw < W) c[it.oc + gw * it.c_w + gh * it.c_h] = acc[r];
      }
    }
  } else {
    ld4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = ld4{0.0, 0.0, 0.0, 0.0};
    // an empty inner dimension loads nothing (its clamp would reach index
    // -1) and stores the zero sums
    if (K > 0) load(0);
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += KS) {
      stage(buf, k0);
      __syncthreads();
      if (k0 + KS < K) load(k0 + KS);
#pragma unroll
      for (int kk = 0; kk < KS; kk += 4) {
        double af[2], bf[2];
#pragma unroll
        for (int i = 0; i < 2; i++) af[i] = As[buf][kk + (lane >> 4)][32 * wr + 16 * i + (lane & 15)];
#pragma unroll
        for (int j = 0; j < 2; j++) bf[j] = Bs[buf][kk + (lane >> 4)][32 * wc + 16 * j + (lane & 15)];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
      buf ^= 1;
    }
    if (active) {
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int64_t gh = h0 + 32 * wr + 16 * i + 4 * r + (lane >> 4);
            const int64_t gw = w0 + 32 * wc + 16 * j + (lane & 15);
            if (gh < H && gw < W) c[it.oc + gw * it.c_w + gh * it.c_h] = acc[i][j][r];
          }
    }
  }
  // the LDS buffers are refilled by the next work item
  __syncthreads();
}

// KS: k per LDS stage (the type's default, or half of it: half the LDS per
// block, more resident blocks to cover load latency); the register budget
// matches what the LDS allows (4 blocks of 4 waves per CU at the default
// stage, 6 at half)
template <class T, int KS>
__global__ __launch_bounds__(256, KS == ListMfma<T>::KS ? 4 : 6) void list_mfma_queue(const MmItem *__restrict__ items,
                                                       const int3 *__restrict__ work,
                                                       const unsigned int *__restrict__ bounds,
                                                       unsigned int *counter) {
  __shared__ T As[2][KS][65];
  __shared__ T Bs[2][KS][65];
  __shared__ unsigned int next;
  int bin = blockIdx.x % LIST_BINS;
  unsigned done = 0;
  int3 job;
  while (list_next(counter, bounds, work, bin, done, &next, job))
    list_mfma_tile<T, KS>(items[job.x], job.y, job.z, As, Bs);
}

template <class T, int KS>
__global__ __launch_bounds__(256, KS == ListMfma<T>::KS ? 4 : 6) void list_mfma_one(const MmItem *__restrict__ items,
                                                                                   int idx, int64_t ntn) {
  __shared__ T As[2][KS][65];
  __shared__ T Bs[2][KS][65];
  list_mfma_tile<T, KS>(items[idx], blockIdx.x / ntn, blockIdx.x % ntn, As, Bs);
}

// ---------------------------------------------------------------- float on BF16 matrix cores
// Float lists can also run as three-way BF16 splits: x = hi + mid + lo,
// hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid), each rounded
// to nearest even; the residuals are exact and lo takes the last 8 bits, so
// hi + mid + lo == x exactly for every normal float (down to ~2^-110,
// where lo would be subnormal).  BF16 keeps FP32's exponent range: no
// operand scaling, rows of any dynamic range.  Six
// v_mfma_f32_32x32x16_bf16 products per 16-deep k step (hl, lh, mm, hm, mh,
// hh: small terms first) with FP32 accumulation; the dropped ml, lm, ll
// terms are at most 2^-23 of |x y|.  An Inf or NaN enters as hi (mid = lo =
// 0), so non-finite values propagate as in FP32 arithmetic.  Each element
// is split once, when its k chunk is staged: LDS holds
// [stage][hi/mid/lo][row][32 k] as 16-byte segments, segment s of row r at
// s ^ ((r >> 1) & 3), so the 8 lanes of a 16-byte access hit distinct bank
// groups.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t bf16_rne(uint32_t u) { return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16; }

__device__ __forceinline__ void split3_bf16(float x, uint32_t &h, uint32_t &m, uint32_t &l) {
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) {                 // Inf keeps its bits, NaN becomes a quiet NaN
    h = (u & 0x7fffffu) ? (0x7fc0u | ((u >> 16) & 0x8000u)) : (u >> 16);
    m = 0;
    l = 0;
    return;
  }
  h = bf16_rne(u);
  if ((h & 0x7f80u) == 0x7f80u) h = u >> 16;             // above BF16's largest finite: truncate instead
  const float r = x - __uint_as_float(h << 16);            // exact
  m = bf16_rne(__float_as_uint(r));
  const float r2 = r - __uint_as_float(m << 16);           // exact
  l = bf16_rne(__float_as_uint(r2));                       // exact for normal x
}

constexpr int BKS = 32;   // k per chunk: two 16-deep MFMA steps

__device__ __forceinline__ void list_bf16x3_tile(const MmItem &it, int64_t tm, int64_t tn, uint4 (*As)[3][64][4],
                                                 uint4 (*Bs)[3][64][4]) {
  const float *a = (const float *)it.a + it.oa, *b = (const float *)it.b + it.ob;
  float *c = (float *)it.c;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int H = (int)it.H, W = (int)it.W, K = (int)it.T;
  const int h0 = (int)tm * 64, w0 = (int)tn * 64;
  const int a_t = (int)it.a_t, a_h = (int)it.a_h, b_w = (int)it.b_w, b_t = (int)it.b_t;
  const bool active = h0 + 32 * wr < H && w0 + 32 * wc < W;
  // A: thread (row tid/4, 8 consecutive k of segment tid%4); B: thread
  // (column tid%64, 8 consecutive k of segment tid/64)
  const int arow = tid >> 2, aseg = tid & 3, bcol = tid & 63, bseg = tid >> 6;
  const int ah = h0 + arow < H ? h0 + arow : H - 1, bw = w0 + bcol < W ? w0 + bcol : W - 1;
  const bool avec = a_t == 1 && (((uintptr_t)(a + (int64_t)ah * a_h)) % 16 == 0);
  float ra[8], rb[8];
  auto load = [&](int k0) {
    const int ka = k0 + aseg * 8, kb = k0 + bseg * 8;
    if (avec && ka + 8 <= K) {
      const Vec<float, 4> x = *reinterpret_cast<const Vec<float, 4> *>(a + (int64_t)ah * a_h + ka);
      const Vec<float, 4> y = *reinterpret_cast<const Vec<float, 4> *>(a + (int64_t)ah * a_h + ka + 4);
#pragma unroll
      for (int e = 0; e < 4; e++) { ra[e] = x.v[e]; ra[4 + e] = y.v[e]; }
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) ra[e] = a[(int64_t)(ka + e < K ? ka + e : K - 1) * a_t + (int64_t)ah * a_h];
    }
#pragma unroll
    for (int e = 0; e < 8; e++) rb[e] = b[(int64_t)(kb + e < K ? kb + e : K - 1) * b_t + (int64_t)bw * b_w];
  };
  auto stage = [&](int buf, int k0) {
    uint32_t ph[4], pm[4], pl[4], qh[4], qm[4], ql[4];
    const bool arow_ok = h0 + arow < H, bcol_ok = w0 + bcol < W;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      uint32_t h0_, m0_, l0_, h1_, m1_, l1_;
      const float x0 = (arow_ok && k0 + aseg * 8 + e < K) ? ra[e] : 0.0f;
      const float x1 = (arow_ok && k0 + aseg * 8 + e + 1 < K) ? ra[e + 1] : 0.0f;
      split3_bf16(x0, h0_, m0_, l0_);
      split3_bf16(x1, h1_, m1_, l1_);
      ph[e / 2] = h0_ | (h1_ << 16); pm[e / 2] = m0_ | (m1_ << 16); pl[e / 2] = l0_ | (l1_ << 16);
      const float y0 = (bcol_ok && k0 + bseg * 8 + e < K) ? rb[e] : 0.0f;
      const float y1 = (bcol_ok && k0 + bseg * 8 + e + 1 < K) ? rb[e + 1] : 0.0f;
      split3_bf16(y0, h0_, m0_, l0_);
      split3_bf16(y1, h1_, m1_, l1_);
      qh[e / 2] = h0_ | (h1_ << 16); qm[e / 2] = m0_ | (m1_ << 16); ql[e / 2] = l0_ | (l1_ << 16);
    }
    const int sa = aseg ^ ((arow >> 1) & 3), sb = bseg ^ ((bcol >> 1) & 3);
    As[buf][0][arow][sa] = make_uint4(ph[0], ph[1], ph[2], ph[3]);
    As[buf][1][arow][sa] = make_uint4(pm[0], pm[1], pm[2], pm[3]);
    As[buf][2][arow][sa] = make_uint4(pl[0], pl[1], pl[2], pl[3]);
    Bs[buf][0][bcol][sb] = make_uint4(qh[0], qh[1], qh[2], qh[3]);
    Bs[buf][1][bcol][sb] = make_uint4(qm[0], qm[1], qm[2], qm[3]);
    Bs[buf][2][bcol][sb] = make_uint4(ql[0], ql[1], ql[2], ql[3]);
  };
  lf32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  if (K > 0) load(0);                  // an empty inner dimension loads nothing
  int buf = 0;
  const int fr = lane & 31, fseg = lane >> 5;
  const int ar_ = 32 * wr + fr, br_ = 32 * wc + fr;
  for (int k0 = 0; k0 < K; k0 += BKS) {
    stage(buf, k0);
    __syncthreads();
    if (k0 + BKS < K) load(k0 + BKS);
    {   // every wave multiplies (see list_mfma_tile)
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        const int sa = (ks * 2 + fseg) ^ ((ar_ >> 1) & 3), sb = (ks * 2 + fseg) ^ ((br_ >> 1) & 3);
        bf16x8 fa[3], fb[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
          fa[q] = __builtin_bit_cast(bf16x8, As[buf][q][ar_][sa]);
          fb[q] = __builtin_bit_cast(bf16x8, Bs[buf][q][br_][sb]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0], fb[2], acc, 0, 0, 0);   // hi * lo
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[2], fb[0], acc, 0, 0, 0);   // lo * hi
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1], fb[1], acc, 0, 0, 0);   // mid * mid
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0], fb[1], acc, 0, 0, 0);   // hi * mid
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[1], fb[0], acc, 0, 0, 0);   // mid * hi
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0], fb[0], acc, 0, 0, 0);   // hi * hi
      }
    }
    buf ^= 1;
  }
  if (active) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int64_t gh = h0 + 32 * wr + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
      const int64_t gw = w0 + 32 * wc + (lane & 31);
      if (gh < H && gw < W) c[it.oc + gw * it.c_w + gh * it.c_h] = acc[r];
    }
  }
  __syncthreads();   // the LDS buffers are refilled by the next work item
}

__global__ __launch_bounds__(256) void list_bf16x3_queue(const MmItem *__restrict__ items,
                                                         const int3 *__restrict__ work,
                                                         const unsigned int *__restrict__ bounds,
                                                         unsigned int *counter) {
  __shared__ uint4 As[2][3][64][4];
  __shared__ uint4 Bs[2][3][64][4];
  __shared__ unsigned int next;
  int bin = blockIdx.x % LIST_BINS;
  unsigned done = 0;
  int3 job;
  while (list_next(counter, bounds, work, bin, done, &next, job)) list_bf16x3_tile(items[job.x], job.y, job.z, As, Bs);
}

__global__ __launch_bounds__(256) void list_bf16x3_one(const MmItem *__restrict__ items, int idx, int64_t ntn) {
  __shared__ uint4 As[2][3][64][4];
  __shared__ uint4 Bs[2][3][64][4];
  list_bf16x3_tile(items[idx], blockIdx.x / ntn, blockIdx.x % ntn, As, Bs);
}

static hipError_t list_bf16x3(const MmItem *d_items, const MmItem *h_items, int nitems, const int3 *d_work,
                              unsigned int nwork, unsigned int *d_counter, bool each, int nblocks, hipStream_t s) {
  if (!each) {
    hipLaunchKernelGGL(list_reset, dim3(1), dim3(64), 0, s, d_counter);
    hipLaunchKernelGGL(list_bf16x3_queue, dim3(nblocks), dim3(256), 0, s, d_items, d_work,
                       (const unsigned int *)d_counter + LIST_BOUNDS, d_counter);
    return hipGetLastError();
  }
  for (int i = 0; i < nitems; i++) {
    const int64_t ntm = (h_items[i].H + 63) / 64, ntn = (h_items[i].W + 63) / 64;
    if (ntm * ntn == 0) continue;             // an empty k stores zeros
    hipLaunchKernelGGL(list_bf16x3_one, dim3((unsigned)(ntm * ntn)), dim3(256), 0, s, d_items, i, ntn);
  }
  return hipGetLastError();
}

template <class T, int KS>
static hipError_t list_mfma_T(const MmItem *d_items, const MmItem *h_items, int nitems, const int3 *d_work,
                              unsigned int nwork, unsigned int *d_counter, bool each, int nblocks, hipStream_t s) {
  if (!each) {
    hipLaunchKernelGGL(list_reset, dim3(1), dim3(64), 0, s, d_counter);
    hipLaunchKernelGGL((list_mfma_queue<T, KS>), dim3(nblocks), dim3(256), 0, s, d_items, d_work,
                       (const unsigned int *)d_counter + LIST_BOUNDS, d_counter);
    return hipGetLastError();
  }
  for (int i = 0; i < nitems; i++) {
    const int64_t ntm = (h_items[i].H + 63) / 64, ntn = (h_items[i].W + 63) / 64;
    if (ntm * ntn == 0) continue;             // an empty k stores zeros
    hipLaunchKernelGGL((list_mfma_one<T, KS>), dim3((unsigned)(ntm * ntn)), dim3(256), 0, s, d_items, i, ntn);
  }
  return hipGetLastError();
}

template <class T>
static hipError_t list_T(const MmItem *d_items, const MmItem *h_items, int nitems, const int3 *d_work,
                         unsigned int nwork, unsigned int *d_counter, int tile, bool fused, bool each, int nblocks,
                         hipStream_t s) {
#define LQ(F, TS) hipLaunchKernelGGL((list_queue<T, F, TS>), dim3(nblocks), dim3(256), 0, s, d_items, d_work, \
                                     (const unsigned int *)d_counter + LIST_BOUNDS + (TS == 32 ? LIST_BINS + 1 : 0), \
                                     d_counter)
#define L1(F, TS) hipLaunchKernelGGL((list_one<T, F, TS>), dim3((unsigned)(ntm * ntn)), dim3(256), 0, s, d_items, i, ntn)
  if (!each) {
    hipLaunchKernelGGL(list_reset, dim3(1), dim3(64), 0, s, d_counter);
    if (tile == 32) { if (fused) LQ(true, 32); else LQ(false, 32); }
    else { if (fused) LQ(true, 64); else LQ(false, 64); }
    return hipGetLastError();
  }
  for (int i = 0; i < nitems; i++) {
    const int64_t ntm = (h_items[i].H + tile - 1) / tile, ntn = (h_items[i].W + tile - 1) / tile;
    if (ntm * ntn == 0) continue;
    if (tile == 32) { if (fused) L1(true, 32); else L1(false, 32); }
    else { if (fused) L1(true, 64); else L1(false, 64); }
  }
#undef LQ
#undef L1
  return hipGetLastError();
}

hipError_t launch_matmult_list(int T, const MmItem *d_items, const MmItem *h_items, int nitems, const int3 *d_work,
                               unsigned int nwork, unsigned int *d_counter, int tile, bool fused, bool each,
                               int nblocks, hipStream_t s) {
  // tile 0: the matrix-core kernels (64x64 work items), float/double lists
  // in the default modes; tile 1: float on the BF16 matrix cores
  if (tile == 1 && fused && T == PDLHIP_F)
    return list_bf16x3(d_items, h_items, nitems, d_work, nwork, d_counter, each, nblocks, s);
  if (tile == 1) tile = 0;
  constexpr int KF = ListMfma<float>::KS, KD = ListMfma<double>::KS;
  if (tile == 2 && fused && T == PDLHIP_F)
    return list_mfma_T<float, KF / 2>(d_items, h_items, nitems, d_work, nwork, d_counter, each, nblocks, s);
  if (tile == 2 && fused && T == PDLHIP_D)
    return list_mfma_T<double, KD / 2>(d_items, h_items, nitems, d_work, nwork, d_counter, each, nblocks, s);
  if (tile == 2) tile = 0;
  if (tile == 0 && fused && T == PDLHIP_F)
    return list_mfma_T<float, KF>(d_items, h_items, nitems, d_work, nwork, d_counter, each, nblocks, s);
  if (tile == 0 && fused && T == PDLHIP_D)
    return list_mfma_T<double, KD>(d_items, h_items, nitems, d_work, nwork, d_counter, each, nblocks, s);
  if (tile == 0) tile = 64;
  switch (T) {
#define X(TC) case TC: return list_T<CType<TC>::T>(d_items, h_items, nitems, d_work, nwork, d_counter, tile, \
                                                   fused, each, nblocks, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

}  // namespace pdlhip
