// Slices.pd index / index1d / index2d (Slices.pd:142-220) on the device:
// Code gathers c() = a(n => ind), BackCode (the vtable's writebackdata, run
// when the child is written through dataflow) scatters a(n => ind) = c().
//
// One thread per (broadcast position, m) in PDL's loop order (broadcast
// dims in the plan's order, dim 0 fastest, m innermost), grid-stride.  An
// index out of range, or bad in bad mode ($ISBADVAR), is PDL's $CROAK: the
// kernel keeps the smallest such loop position (atomicMin), the host reads
// the index there and words the message; elements at other positions are
// still processed.  index1d sets c bad for a bad index (BackCode skips it).
//
// BackCode with repeated indices: PDL's sequential loop leaves the value of
// the last position that writes each element of a.  Pass 1 takes, per
// element of a, the largest loop position that writes it (atomicMax into
// an owner table spanning a's elements, cleared by a kernel); pass 2 writes
// only from that position, so the result does not depend on thread order.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

constexpr int IX_THREADS = 256;

// offsets of loop position e: a, first index, second index, c; and m
__device__ __forceinline__ void ix_offsets(const IxPlan &p, int64_t e, int64_t *o, int64_t &m) {
  m = e % p.m;
  int64_t r = e / p.m;
  o[0] = p.off[0]; o[1] = p.off[1]; o[2] = p.off[2]; o[3] = p.off[3];
  for (int d = 0; d < p.nd; d++) {
    const int64_t i = r % p.dims[d];
    r /= p.dims[d];
    o[0] += i * p.inc[0][d];
    o[1] += i * p.inc[1][d];
    o[2] += i * p.inc[2][d];
    o[3] += i * p.inc[3][d];
  }
}

// the element of a position e addresses (offset into a's data), or -1 when
// it croaks (position recorded) or, index1d, has a bad index (c set bad
// unless writing back, -2 returned)
template <class T>
__device__ __forceinline__ int64_t ix_target(const IxPlan &p, int64_t e, const int64_t *__restrict__ i1,
                                             const int64_t *__restrict__ i2, T *__restrict__ c, T cbad,
                                             unsigned long long *err, int64_t &co) {
  int64_t o[4], m;
  ix_offsets(p, e, o, m);
  if (p.op == 2) {
    const int64_t xa = i1[o[1]], xb = i2[o[2]];
    co = o[3];
    if ((p.bad && xa == p.ind_bv[0]) || xa < 0 || xa >= p.na || (p.bad && xb == p.ind_bv[1]) || xb < 0 ||
        xb >= p.nb2) {
      atomicMin(err, (unsigned long long)e);
      return -1;
    }
    return o[0] + xa * p.ia0 + xb * p.ia1;
  }
  const int64_t x = i1[o[1] + m * p.iim];
  co = o[3] + m * p.icm;
  if (p.op == 1 && p.bad && x == p.ind_bv[0]) {
    if (!p.back) c[co] = cbad;
    return -2;
  }
  if ((p.bad && x == p.ind_bv[0]) || x < 0 || x >= p.na) {
    atomicMin(err, (unsigned long long)e);
    return -1;
  }
  return o[0] + x * p.ia0;
}

template <class T>
__global__ __launch_bounds__(IX_THREADS) void k_index_gather(IxPlan p, const T *__restrict__ a,
                                                             const int64_t *__restrict__ i1,
                                                             const int64_t *__restrict__ i2, T *__restrict__ c,
                                                             T cbad, unsigned long long *err) {
  const int64_t total = p.npos * p.m;
  for (int64_t e = (int64_t)blockIdx.x * IX_THREADS + threadIdx.x; e < total; e += (int64_t)gridDim.x * IX_THREADS) {
    int64_t co;
    const int64_t t = ix_target<T>(p, e, i1, i2, c, cbad, err, co);
    if (t >= 0) c[co] = a[t];
  }
}

// pass 1 of the writeback: the last loop position writing each element of a
__global__ __launch_bounds__(IX_THREADS) void k_index_owner(IxPlan p, const int64_t *__restrict__ i1,
                                                            const int64_t *__restrict__ i2,
                                                            unsigned long long *err, long long *owner) {
  const int64_t total = p.npos * p.m;
  for (int64_t e = (int64_t)blockIdx.x * IX_THREADS + threadIdx.x; e < total; e += (int64_t)gridDim.x * IX_THREADS) {
    int64_t co;
    const int64_t t = ix_target<char>(p, e, i1, i2, nullptr, 0, err, co);
    if (t >= 0) atomicMax(owner + (t - p.a_lo), (long long)e);
  }
}

template <class T>
__global__ __launch_bounds__(IX_THREADS) void k_index_scatter(IxPlan p, T *__restrict__ a,
                                                              const int64_t *__restrict__ i1,
                                                              const int64_t *__restrict__ i2, const T *__restrict__ c,
                                                              const long long *__restrict__ owner) {
  const int64_t total = p.npos * p.m;
  for (int64_t e = (int64_t)blockIdx.x * IX_THREADS + threadIdx.x; e < total; e += (int64_t)gridDim.x * IX_THREADS) {
    int64_t co;
    unsigned long long dummy = ~0ull;
    const int64_t t = ix_target<T>(p, e, i1, i2, const_cast<T *>(c), T(0), &dummy, co);
    if (t >= 0 && owner[t - p.a_lo] == (long long)e) a[t] = c[co];
  }
}

__global__ void k_fill_i64(long long *p, int64_t n, long long v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

static unsigned ix_grid(int64_t total) {
  int64_t g = (total + IX_THREADS - 1) / IX_THREADS;
  return (unsigned)(g < 1 ? 1 : (g > 16384 ? 16384 : g));
}

template <class T>
static hipError_t index_go(const IxPlan &p, void *a, const int64_t *i1, const int64_t *i2, void *c,
                           const pdlhip_value &cbv, unsigned long long *err, long long *owner, hipStream_t s) {
  const int64_t total = p.npos * p.m;
  if (total == 0) return hipSuccess;
  if (!p.back) {
    hipLaunchKernelGGL(k_index_gather<T>, dim3(ix_grid(total)), dim3(IX_THREADS), 0, s, p, (const T *)a, i1, i2,
                       (T *)c, bv_as<T>(cbv), err);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_fill_i64, dim3(ix_grid(p.a_span)), dim3(IX_THREADS), 0, s, owner, p.a_span, -1ll);
  hipLaunchKernelGGL(k_index_owner, dim3(ix_grid(total)), dim3(IX_THREADS), 0, s, p, i1, i2, err, owner);
  hipLaunchKernelGGL(k_index_scatter<T>, dim3(ix_grid(total)), dim3(IX_THREADS), 0, s, p, (T *)a, i1, i2,
                     (const T *)c, (const long long *)owner);
  return hipGetLastError();
}

// ---------------------------------------------------------------- rangeb
// Slices.pd rangeb (EquivCPOffsCode, Slices.pd:900-1013, with PP's
// $EQUIVCPTRUNC): a child element at index-broadcast position `item`,
// in-range position iter[] and source-broadcast position st reads the
// parent at corner + iter with each dim's boundary rule (0 forbid: $CROAK;
// 1 truncate: the child gets BAD in bad mode, else 0; 2 extend: clamp; 3
// periodic; 4 mirror); dims of the index past the parent's take rule 0/1 on
// a nonzero coordinate and ignore it otherwise.  A failing position keeps
// the smallest (item, in-range position, kind) in PDL's loop order for the
// host's message.  BackCode: the parent element gets the child of the last
// position in PDL's loop order (items, in-range positions, source
// broadcast) that maps to it and is not truncated.
struct RgElem {
  int64_t poff, coff, key;   // key: the position in PDL's loop order
  int trunc;
};

__device__ __forceinline__ bool rg_element(const RgPlan &p, int64_t e, const int64_t *__restrict__ ind,
                                           unsigned long long *err, RgElem &r) {
  int64_t c[PDLHIP_MAXDIMS];
  int64_t rem = e;
  for (int d = 0; d < p.cnd; d++) {
    c[d] = rem % p.cdims[d];
    rem /= p.cdims[d];
  }
  int64_t item = 0, istride = 1, ioff = p.ioff;
  for (int d = 0; d < p.itdim; d++) {
    item += c[d] * istride;
    istride *= p.cdims[d];
    ioff += c[d] * p.iincs[d + 1];
  }
  r.coff = p.coff;
  for (int d = 0; d < p.cnd; d++) r.coff += c[d] * p.cincs[d];
  // the in-range position (zero sizes stay at 0) and its place in PDL's loop
  int64_t it[RG_MAXR], inr = 0, inr_stride = 1;
  int j = p.itdim;
  for (int k = 0; k < p.rdim; k++) {
    it[k] = p.sizes[k] ? c[j++] : 0;
    inr += it[k] * inr_stride;
    inr_stride *= p.sizes[k] ? p.sizes[k] : 1;
  }
  const int64_t pos = item * p.ninr + inr;
  int64_t poff = p.poff;
  int trunc = 0;
  const int prdim = p.rdim < p.pdim ? p.rdim : p.pdim;
  for (int k = 0; k < p.rdim; k++) {
    int64_t ck = it[k] + ind[ioff + k * p.iincs[0]];
    if (k < prdim) {
      const int64_t pd = p.pdims[k];
      if (ck < 0 || ck >= pd) {
        switch (p.bnd[k]) {
        case 0:
          atomicMin(err, (unsigned long long)(pos * 2));        // "index out-of-bounds in range"
          return false;
        case 1: trunc = 1; break;
        case 2: ck = ck >= pd ? pd - 1 : 0; break;
        case 3: ck %= pd; if (ck < 0) ck += pd; break;
        default:
          ck += pd; ck %= pd * 2; if (ck < 0) ck += pd * 2; ck -= pd;
          if (ck < 0) ck = -ck - 1;
          break;
        }
      }
      if (!trunc) poff += ck * p.pincs[k];
    } else if (ck != 0) {
      if (p.bnd[k] == 0) {
        atomicMin(err, (unsigned long long)(pos * 2 + 1));      // "Too many dims in range index"
        return false;
      }
      if (p.bnd[k] == 1) trunc = 1;
    }
  }
  int64_t st = 0, st_stride = 1;
  for (int k = 0; k < p.stdim; k++) {
    const int64_t x = c[p.itdim + p.ntsize + k];
    poff += x * p.pincs[p.rdim + k];
    st += x * st_stride;
    st_stride *= p.cdims[p.itdim + p.ntsize + k];
  }
  r.poff = poff;
  r.trunc = trunc;
  r.key = (item * p.ninr + inr) * p.nst + st;
  return true;
}

template <class T>
__global__ __launch_bounds__(IX_THREADS) void k_range_gather(RgPlan p, const T *__restrict__ a,
                                                             const int64_t *__restrict__ ind, T *__restrict__ c,
                                                             T pbv, T cbv, int pnan, unsigned long long *err) {
  for (int64_t e = (int64_t)blockIdx.x * IX_THREADS + threadIdx.x; e < p.nchild; e += (int64_t)gridDim.x * IX_THREADS) {
    RgElem r;
    if (!rg_element(p, e, ind, err, r)) continue;
    if (r.trunc) {
      c[r.coff] = p.bad ? cbv : T(0);
      continue;
    }
    const T v = a[r.poff];
    c[r.coff] = (p.bad && isbad(v, pbv, pnan)) ? cbv : v;
  }
}

__global__ __launch_bounds__(IX_THREADS) void k_range_owner(RgPlan p, const int64_t *__restrict__ ind,
                                                            unsigned long long *err, long long *owner) {
  for (int64_t e = (int64_t)blockIdx.x * IX_THREADS + threadIdx.x; e < p.nchild; e += (int64_t)gridDim.x * IX_THREADS) {
    RgElem r;
    if (!rg_element(p, e, ind, err, r) || r.trunc) continue;
    atomicMax(owner + (r.poff - p.a_lo), (long long)r.key);
  }
}

template <class T>
__global__ __launch_bounds__(IX_THREADS) void k_range_scatter(RgPlan p, T *__restrict__ a,
                                                              const int64_t *__restrict__ ind, const T *__restrict__ c,
                                                              T pbv, T cbv, int cnan, const long long *owner) {
  for (int64_t e = (int64_t)blockIdx.x * IX_THREADS + threadIdx.x; e < p.nchild; e += (int64_t)gridDim.x * IX_THREADS) {
    RgElem r;
    unsigned long long dummy = ~0ull;
    if (!rg_element(p, e, ind, &dummy, r) || r.trunc) continue;
    if (owner[r.poff - p.a_lo] != (long long)r.key) continue;
    const T v = c[r.coff];
    a[r.poff] = (p.bad && isbad(v, cbv, cnan)) ? pbv : v;
  }
}

template <class T>
static hipError_t range_go(const RgPlan &p, void *a, const int64_t *ind, void *c, const pdlhip_value &pbv,
                           const pdlhip_value &cbv, int pnan, int cnan, unsigned long long *err, long long *owner,
                           hipStream_t s) {
  if (p.nchild == 0) return hipSuccess;
  if (!p.back) {
    hipLaunchKernelGGL(k_range_gather<T>, dim3(ix_grid(p.nchild)), dim3(IX_THREADS), 0, s, p, (const T *)a, ind,
                       (T *)c, bv_as<T>(pbv), bv_as<T>(cbv), pnan, err);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_fill_i64, dim3(ix_grid(p.a_span)), dim3(IX_THREADS), 0, s, owner, p.a_span, -1ll);
  hipLaunchKernelGGL(k_range_owner, dim3(ix_grid(p.nchild)), dim3(IX_THREADS), 0, s, p, ind, err, owner);
  hipLaunchKernelGGL(k_range_scatter<T>, dim3(ix_grid(p.nchild)), dim3(IX_THREADS), 0, s, p, (T *)a, ind,
                     (const T *)c, bv_as<T>(pbv), bv_as<T>(cbv), cnan, (const long long *)owner);
  return hipGetLastError();
}

hipError_t launch_range(int T, const RgPlan &p, void *parent, const int64_t *ind, void *child,
                        const pdlhip_value &pbv, const pdlhip_value &cbv, int pbv_nan, int cbv_nan,
                        unsigned long long *err, long long *owner, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return range_go<CType<TC>::T>(p, parent, ind, child, pbv, cbv, pbv_nan, cbv_nan, err, owner, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_index(int T, const IxPlan &p, void *a, const int64_t *i1, const int64_t *i2, void *c,
                        const pdlhip_value &cbv, unsigned long long *err, long long *owner, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return index_go<CType<TC>::T>(p, a, i1, i2, c, cbv, err, owner, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

}  // namespace pdlhip
