// Data movement kernels: dtype conversion between views with PDL's bad-value
// mapping (converttypei, lib/PDL/Core/pdlconv.c:163-174: a bad source element
// becomes the target's bad value; unsigned targets convert through
// intmax_t), which is also assgn (lib/PDL/Ops.pd:382-397) when the types
// match; and counter-based fills for synthetic inputs.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

template <class D, class S> __device__ inline D conv1(S v) {
  if constexpr (!IsFloat<D>::v && (D)-1 > (D)0) return (D)(intmax_t)v;
  else return (D)v;
}

template <class D, class S, bool BAD>
__global__ __launch_bounds__(256) void k_convert(const S *__restrict__ src, D *__restrict__ dst, EwPlan p, BadInfo sbi,
                                                 D dbv) {
  const S sbv = bv_as<S>(sbi.bv);
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < p.n; e += (int64_t)gridDim.x * 256) {
    int64_t r = e, os = p.off[0], od = p.off[2];
    for (int d = 0; d < p.nd; d++) {
      int64_t i = r % p.dims[d];
      r /= p.dims[d];
      os += i * p.inc[0][d];
      od += i * p.inc[2][d];
    }
    S v = src[os];
    if (BAD && isbad(v, sbv, sbi.isnan)) dst[od] = dbv;
    else dst[od] = conv1<D, S>(v);
  }
}

// contiguous same-type copy (the common assgn of dense data); 16-byte
// vectors when both ends are 16-byte aligned, streaming loads and stores for
// copies too large for the infinity cache
template <class T, int VEC>
__global__ __launch_bounds__(256) void k_copy_flat(const T *__restrict__ src, T *__restrict__ dst, int64_t n,
                                                   bool stream) {
  const int64_t nv = n / VEC, stride = (int64_t)gridDim.x * 256;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < nv; v += stride)
    if (stream) store_nt<T, VEC>(dst + v * VEC, load_nt<T, VEC>(src + v * VEC));
    else reinterpret_cast<Vec<T, VEC> *>(dst)[v] = reinterpret_cast<const Vec<T, VEC> *>(src)[v];
  for (int64_t e = nv * VEC + (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += stride) dst[e] = src[e];
}

template <class D, class S>
static hipError_t convert_DS(const EwPlan &p, const void *src, void *dst, bool bad, const BadInfo &sbi,
                             const pdlhip_value &dbv, hipStream_t s) {
  if (p.n == 0) return hipSuccess;
  int64_t blocks = (p.n + 1023) / 1024;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  if (!bad && sizeof(D) == sizeof(S) && IsFloat<D>::v == IsFloat<S>::v && (D)-1 == (S)-1 && p.nd == 1 &&
      p.inc[0][0] == 1 && p.inc[2][0] == 1) {
    const D *sp = (const D *)src + p.off[0];
    D *dp = (D *)dst + p.off[2];
    constexpr int VEC = 16 / sizeof(D);
    if ((uintptr_t)sp % 16 == 0 && (uintptr_t)dp % 16 == 0) {
      int64_t vb = (p.n / VEC + 255) / 256;
      vb = vb < 1 ? 1 : (vb > 65536 ? 65536 : vb);
      hipLaunchKernelGGL((k_copy_flat<D, VEC>), dim3((unsigned)vb), dim3(256), 0, s, sp, dp, p.n,
                         streaming_read((size_t)p.n * sizeof(D) * 2));
    } else {
      hipLaunchKernelGGL((k_copy_flat<D, 1>), dim3((unsigned)blocks), dim3(256), 0, s, sp, dp, p.n, false);
    }
    return hipGetLastError();
  }
  if (bad)
    hipLaunchKernelGGL((k_convert<D, S, true>), dim3((unsigned)blocks), dim3(256), 0, s, (const S *)src, (D *)dst, p,
                       sbi, bv_as<D>(dbv));
  else
    hipLaunchKernelGGL((k_convert<D, S, false>), dim3((unsigned)blocks), dim3(256), 0, s, (const S *)src, (D *)dst, p,
                       sbi, bv_as<D>(dbv));
  return hipGetLastError();
}

template <class D>
static hipError_t convert_D(int stype, const EwPlan &p, const void *src, void *dst, bool bad, const BadInfo &sbi,
                            const pdlhip_value &dbv, hipStream_t s) {
  switch (stype) {
#define X(TC) case TC: return convert_DS<D, CType<TC>::T>(p, src, dst, bad, sbi, dbv, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_convert(int stype, int dtype, const EwPlan &p, const void *src, void *dst, bool bad,
                          const BadInfo &sbi, const pdlhip_value &dbv, hipStream_t s) {
  if (type_is_cplx(stype) || type_is_cplx(dtype))
    return bad ? hipErrorInvalidValue : launch_cplx_convert(stype, dtype, p, src, dst, s);
  switch (dtype) {
#define X(TC) case TC: return convert_D<CType<TC>::T>(stype, p, src, dst, bad, sbi, dbv, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------- fills
__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ inline double u01(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

template <class D>
__global__ __launch_bounds__(256) void k_fill(D *__restrict__ dst, EwPlan p, uint64_t seed, double a, double b,
                                              int normal) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < p.n; e += (int64_t)gridDim.x * 256) {
    int64_t r = e, od = p.off[2];
    for (int d = 0; d < p.nd; d++) {
      int64_t i = r % p.dims[d];
      r /= p.dims[d];
      od += i * p.inc[2][d];
    }
    double v;
    if (normal) {
      // Box-Muller on two counter-based uniforms
      double u1 = u01(splitmix64(seed ^ (2 * (uint64_t)e))), u2 = u01(splitmix64(seed ^ (2 * (uint64_t)e + 1)));
      if (u1 < 1e-300) u1 = 1e-300;
      v = a + b * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    } else {
      v = a + (b - a) * u01(splitmix64(seed + (uint64_t)e));
    }
    dst[od] = (D)v;
  }
}

hipError_t launch_fill(int dtype, const EwPlan &p, void *dst, uint64_t seed, double a, double b, int normal,
                       hipStream_t s) {
  if (p.n == 0) return hipSuccess;
  int64_t blocks = (p.n + 1023) / 1024;
  if (blocks > 65536) blocks = 65536;
  switch (dtype) {
#define X(TC) case TC: hipLaunchKernelGGL((k_fill<CType<TC>::T>), dim3((unsigned)blocks), dim3(256), 0, s, \
                                          (CType<TC>::T *)dst, p, seed, a, b, normal); return hipGetLastError();
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

}  // namespace pdlhip
