// Lists of uneven matmults in one persistent launch (the reference runs
// such a list as a Perl loop of matmult calls, lib/PDL/Primitive.pd:197-209
// per item).  The host builds one work item per (product, C tile), ordered
// by decreasing inner dimension (longest-processing-time first); persistent
// blocks take items from a device counter until the queue is drained, so
// every block reaches the exit condition.  Each item runs the same SIMT
// tile as k_matmult.hip (accumulation order per output: ascending t, so
// the PDL-order variant stays bit-identical to PDL's loop).
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

__global__ void list_reset(unsigned int *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

// persistent: take work items until the counter passes nwork
template <class T, bool FUSED, int TS>
__global__ __launch_bounds__(256) void list_queue(const MmItem *__restrict__ items, const int3 *__restrict__ work,
                                                  unsigned int nwork, unsigned int *counter) {
  constexpr int KS = 16;
  __shared__ T As[KS][TS + 1];
  __shared__ T Bs[KS][TS + 1];
  __shared__ unsigned int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
    __syncthreads();
    const unsigned int w = next;
    __syncthreads();
    if (w >= nwork) break;
    const int3 job = work[w];
    list_tile<T, FUSED, TS>(items[job.x], job.y, job.z, As, Bs);
  }
}

// one launch per item (the comparison the list path is measured against)
template <class T, bool FUSED, int TS>
__global__ __launch_bounds__(256) void list_one(const MmItem *__restrict__ items, int idx, int64_t ntn) {
  constexpr int KS = 16;
  __shared__ T As[KS][TS + 1];
  __shared__ T Bs[KS][TS + 1];
  list_tile<T, FUSED, TS>(items[idx], blockIdx.x / ntn, blockIdx.x % ntn, As, Bs);
}

template <class T>
static hipError_t list_T(const MmItem *d_items, const MmItem *h_items, int nitems, const int3 *d_work,
                         unsigned int nwork, unsigned int *d_counter, int tile, bool fused, bool each, int nblocks,
                         hipStream_t s) {
#define LQ(F, TS) hipLaunchKernelGGL((list_queue<T, F, TS>), dim3(nblocks), dim3(256), 0, s, d_items, d_work, nwork, \
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
  switch (T) {
#define X(TC) case TC: return list_T<CType<TC>::T>(d_items, h_items, nitems, d_work, nwork, d_counter, tile, \
                                                   fused, each, nblocks, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

}  // namespace pdlhip
