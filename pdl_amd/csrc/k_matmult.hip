// matmult (lib/PDL/Primitive.pd:191-245): c(w,h) = sum_t a(t,h) * b(w,t),
// broadcast dims as a batch.  With dim 0 fastest this is row-major
// C[h][w] = sum_t A[h][t] B[t][w].
//
// Kernels here:
//   mm_exact  one thread per output, PDL's loop restated including the bad
//             value rules (a bad a/b element makes c bad; a running sum that
//             equals c's bad value at a 64-byte t-tile boundary stops it,
//             Primitive.pd:226-227).  Used in bad-value mode.
//   mm_simt   LDS-tiled SIMT kernel, 64x64 output tile per 256-thread block,
//             4x4 per thread, any operand strides (transposed views load
//             through the same path).  FUSED=false multiplies then adds in
//             ascending t: bit-identical to PDL's float/double loop.
//             FUSED=true uses FMA.  Integer types wrap like PDL (-fwrapv).
// The split-FP16 and FP8 MFMA kernels for float live in k_mfma.hip.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

hipError_t launch_mfma_f32(const MmPlan &p, const float *a, const float *b, float *c, int mode, hipStream_t s);
hipError_t launch_mfma_f64(const MmPlan &p, const double *a, const double *b, double *c, hipStream_t s);

__device__ inline void mm_batch_offsets(const MmPlan &p, int64_t bidx, int64_t *o) {
  o[0] = p.off[0]; o[1] = p.off[1]; o[2] = p.off[2];
  for (int d = 0; d < p.nd; d++) {
    int64_t i = bidx % p.dims[d];
    bidx /= p.dims[d];
    o[0] += i * p.inc[0][d]; o[1] += i * p.inc[1][d]; o[2] += i * p.inc[2][d];
  }
}

template <class T, bool BAD>
__global__ __launch_bounds__(256) void mm_exact(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ c,
                                                MmPlan p, BadInfo ba, BadInfo bb, BadInfo bc) {
  const int64_t nout = p.H * p.W;
  const int64_t total = nout * p.nbatch;
  const int64_t tsiz = 8 * sizeof(double) / sizeof(T);
  const T abv = bv_as<T>(ba.bv), bbv = bv_as<T>(bb.bv), cbv = bv_as<T>(bc.bv);
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t bidx = e / nout, r = e % nout;
    const int64_t h = r / p.W, w = r % p.W;
    int64_t o[3];
    mm_batch_offsets(p, bidx, o);
    T cc = T(0);
    for (int64_t to = 0; to < p.T; to += tsiz) {
      if (BAD && isbad(cc, cbv, bc.isnan)) continue;
      const int64_t te = (to + tsiz < p.T) ? to + tsiz : p.T;
      const T *ad = a + o[0] + to * p.a_t + h * p.a_h;
      const T *bd = b + o[1] + w * p.b_w + to * p.b_t;
      bool cbad = false;
      for (int64_t t = to; t < te; t++) {
        if (BAD && (isbad(*ad, abv, ba.isnan) || isbad(*bd, bbv, bb.isnan))) { cbad = true; break; }
        cc = wadd(cc, wmul(*ad, *bd));
        ad += p.atdi;
        bd += p.btdi;
      }
      if (cbad) cc = cbv;
    }
    c[o[2] + w * p.c_w + h * p.c_h] = cc;
  }
}

// ---------------------------------------------------------------- SIMT
constexpr int SB_M = 64, SB_N = 64, SB_K = 16;

template <class T, bool FUSED>
__global__ __launch_bounds__(256) void mm_simt(const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ c,
                                               MmPlan p, int64_t ntm, int64_t ntn, int a_tfast, int b_wfast) {
  __shared__ T As[SB_K][SB_M + 1];
  __shared__ T Bs[SB_K][SB_N + 1];
  int64_t blk = blockIdx.x;
  const int64_t tm = blk % ntm;
  blk /= ntm;
  const int64_t tn = blk % ntn;
  const int64_t bidx = blk / ntn;
  int64_t o[3];
  mm_batch_offsets(p, bidx, o);
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int64_t h0 = tm * SB_M, w0 = tn * SB_N;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = T(0);
  for (int64_t k0 = 0; k0 < p.T; k0 += SB_K) {
    // A tile: SB_M rows (h) x SB_K (t); consecutive threads along the unit-stride index
#pragma unroll
    for (int r = 0; r < (SB_M * SB_K) / 256; r++) {
      int e = tid + 256 * r;
      int kk, mm;
      if (a_tfast) { kk = e % SB_K; mm = e / SB_K; } else { mm = e % SB_M; kk = e / SB_M; }
      int64_t gh = h0 + mm, gt = k0 + kk;
      As[kk][mm] = (gh < p.H && gt < p.T) ? a[o[0] + gt * p.a_t + gh * p.a_h] : T(0);
    }
#pragma unroll
    for (int r = 0; r < (SB_N * SB_K) / 256; r++) {
      int e = tid + 256 * r;
      int kk, nn;
      if (b_wfast) { nn = e % SB_N; kk = e / SB_N; } else { kk = e % SB_K; nn = e / SB_K; }
      int64_t gw = w0 + nn, gt = k0 + kk;
      Bs[kk][nn] = (gw < p.W && gt < p.T) ? b[o[1] + gw * p.b_w + gt * p.b_t] : T(0);
    }
    __syncthreads();
    const int kmax = (p.T - k0 < SB_K) ? (int)(p.T - k0) : SB_K;
    for (int kk = 0; kk < kmax; kk++) {
      T av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) av[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if constexpr (FUSED && IsFloat<T>::v) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
          else acc[i][j] = wadd(acc[i][j], wmul(av[i], bv[j]));
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int64_t gh = h0 + ty + 16 * i;
    if (gh >= p.H) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int64_t gw = w0 + tx + 16 * j;
      if (gw < p.W) c[o[2] + gw * p.c_w + gh * p.c_h] = acc[i][j];
    }
  }
}

template <class T>
static hipError_t matmult_T(const MmPlan &p, const void *va, const void *vb, void *vc, bool bad, const BadInfo *bi,
                            int mode, hipStream_t s) {
  const T *a = (const T *)va, *b = (const T *)vb;
  T *c = (T *)vc;
  const int64_t total = p.H * p.W * p.nbatch;
  if (total == 0) return hipSuccess;
  if (bad) {
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((mm_exact<T, true>), dim3((unsigned)blocks), dim3(256), 0, s, a, b, c, p, bi[0], bi[1], bi[2]);
    return hipGetLastError();
  }
  if constexpr (sizeof(T) == 4 && IsFloat<T>::v) {
    if (mode == PDLHIP_MM_SPLIT16 || mode == PDLHIP_MM_FP8) {
      hipError_t e = launch_mfma_f32(p, (const float *)a, (const float *)b, (float *)c, mode, s);
      if (e != hipErrorNotSupported) return e;
    }
  }
  if constexpr (sizeof(T) == 8 && IsFloat<T>::v) {
    // the default mode runs double products on the FP64 matrix cores
    if (mode == PDLHIP_MM_SPLIT16 || mode == PDLHIP_MM_FP8) {
      hipError_t e = launch_mfma_f64(p, (const double *)a, (const double *)b, (double *)c, s);
      if (e != hipErrorNotSupported) return e;
    }
  }
  const int64_t ntm = (p.H + SB_M - 1) / SB_M, ntn = (p.W + SB_N - 1) / SB_N;
  const int64_t blocks = ntm * ntn * p.nbatch;
  if (blocks >= ((int64_t)1 << 31)) return hipErrorInvalidConfiguration;
  const int a_tfast = (p.a_t == 1 || p.a_h != 1) ? 1 : 0;
  const int b_wfast = (p.b_w == 1 || p.b_t != 1) ? 1 : 0;
  const bool fused = IsFloat<T>::v && mode != PDLHIP_MM_PDLORDER;
  if (fused)
    hipLaunchKernelGGL((mm_simt<T, true>), dim3((unsigned)blocks), dim3(256), 0, s, a, b, c, p, ntm, ntn, a_tfast,
                       b_wfast);
  else
    hipLaunchKernelGGL((mm_simt<T, false>), dim3((unsigned)blocks), dim3(256), 0, s, a, b, c, p, ntm, ntn, a_tfast,
                       b_wfast);
  return hipGetLastError();
}

#define PDLHIP_MM_TYPE(TC) \
  hipError_t launch_matmult_t##TC(const MmPlan &p, const void *a, const void *b, void *c, bool bad, \
                                  const BadInfo *bi, int mode, hipStream_t s) { \
    return matmult_T<CType<TC>::T>(p, a, b, c, bad, bi, mode, s); \
  }
PDLHIP_MM_TYPES_IN_THIS_TU

}  // namespace pdlhip
