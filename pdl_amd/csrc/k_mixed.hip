// Elementwise ops whose inputs have other types than the one the op
// computes in, in one pass.  PDL first makes a converted copy of each such
// input (pdl__get_convertedpdl, converttypei: lib/PDL/Core/pdlconv.c:143-266,
// where a bad source element becomes the target type's bad value), then
// runs the op on the copies (pdlapi.c:1262-1311).  Here the kernel loads
// each input in its own type, converts it in registers exactly as
// converttypei would (unsigned targets through intmax_t; a bad source
// element, when its ndarray is bad, becomes T's default bad value), and
// applies the op and its bad-value rule to the converted values: the same
// bits as the two passes, without the copies.
//
// The source types are kernel arguments; the switch on them is uniform
// across a wave, so each wave runs one load path.  Threads walk the
// canonical plan (dims ordered by the output's strides) four elements at a
// time, grid-stride; a one-dimensional plan skips the index decomposition,
// and a one-element operand (a broadcast scalar) is converted once.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

constexpr int MX_THREADS = 256;

template <class D, class S> __device__ __forceinline__ D mx_conv(S v) {
  if constexpr (!IsFloat<D>::v && (D)-1 > (D)0) return (D)(intmax_t)v;
  else return (D)v;
}

// U input elements of source type code `ts` at offsets o[], as their
// converted copy holds them; the switch is uniform, and each case issues its
// U loads before converting any
template <class T, int U>
__device__ __forceinline__ void mx_load_n(const void *p, int ts, const int64_t *o, bool sbad, const BadInfo &sbi,
                                          T tbad, T *out) {
  switch (ts) {
#define X(TC)                                                                             \
  case TC: {                                                                              \
    typedef typename CType<TC>::T S;                                                      \
    S v[U];                                                                               \
    _Pragma("unroll") for (int j = 0; j < U; j++) v[j] = ((const S *)p)[o[j]];            \
    const S sbv = bv_as<S>(sbi.bv);                                                       \
    _Pragma("unroll") for (int j = 0; j < U; j++)                                         \
      out[j] = (sbad && isbad(v[j], sbv, sbi.isnan)) ? tbad : mx_conv<T, S>(v[j]);        \
    return;                                                                               \
  }
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default:
    _Pragma("unroll") for (int j = 0; j < U; j++) out[j] = T(0);
  }
}

constexpr int MX_U = 4;   // elements per thread per step, their loads in flight together

__device__ __forceinline__ void mx_offsets(const EwPlan &p, int64_t e, int64_t &o0, int64_t &o1, int64_t &o2) {
  o0 = p.off[0]; o1 = p.off[1]; o2 = p.off[2];
  if (p.nd == 1) {
    o0 += e * p.inc[0][0];
    o1 += e * p.inc[1][0];
    o2 += e * p.inc[2][0];
    return;
  }
  for (int d = 0; d < p.nd; d++) {
    const int64_t i = e % p.dims[d];
    e /= p.dims[d];
    o0 += i * p.inc[0][d];
    o1 += i * p.inc[1][d];
    o2 += i * p.inc[2][d];
  }
}

// a, b: inputs of types ta, tb (sbad: bad-flagged source; *_scalar: the
// operand is one element, converted once); the op's own bad test (ba/bb:
// T's default bad value, `check` as the op gates it) then runs on the
// converted values.  A thread takes elements e, e + S, e + 2S, e + 3S
// (S = the grid's thread count): coalesced across a wave, four loads of
// each operand in flight.
template <class T, int OP, bool BAD>
__global__ __launch_bounds__(MX_THREADS) void ew_mixed(const void *__restrict__ a, int ta, int sbad_a, BadInfo sba,
                                                      const void *__restrict__ b, int tb, int sbad_b, BadInfo sbb,
                                                      T *__restrict__ c, EwPlan p, BadInfo ba, BadInfo bb, T cbad,
                                                      int a_scalar, int b_scalar) {
  const T abv = bv_as<T>(ba.bv), bbv = bv_as<T>(bb.bv);
  T xs = T(0), ys = T(0);
  const int64_t a0 = p.off[0], b0 = p.off[1];
  if (a_scalar) mx_load_n<T, 1>(a, ta, &a0, sbad_a, sba, abv, &xs);
  if (b_scalar) mx_load_n<T, 1>(b, tb, &b0, sbad_b, sbb, bbv, &ys);
  const int64_t S = (int64_t)gridDim.x * MX_THREADS;
  for (int64_t e0 = (int64_t)blockIdx.x * MX_THREADS + threadIdx.x; e0 < p.n; e0 += S * MX_U) {
    int64_t oa[MX_U], ob[MX_U], oc[MX_U];
#pragma unroll
    for (int j = 0; j < MX_U; j++) {
      const int64_t e = e0 + j * S;
      mx_offsets(p, e < p.n ? e : e0, oa[j], ob[j], oc[j]);   // out of range: reload e0, not stored
    }
    T x[MX_U], y[MX_U];
    if (a_scalar) {
#pragma unroll
      for (int j = 0; j < MX_U; j++) x[j] = xs;
    } else {
      mx_load_n<T, MX_U>(a, ta, oa, sbad_a, sba, abv, x);
    }
    if (b_scalar) {
#pragma unroll
      for (int j = 0; j < MX_U; j++) y[j] = ys;
    } else {
      mx_load_n<T, MX_U>(b, tb, ob, sbad_b, sbb, bbv, y);
    }
#pragma unroll
    for (int j = 0; j < MX_U; j++) {
      if (e0 + j * S >= p.n) break;
      if constexpr (BAD) {
        if ((ba.check && isbad(x[j], abv, ba.isnan)) || (bb.check && isbad(y[j], bbv, bb.isnan))) {
          c[oc[j]] = cbad;
          continue;
        }
      }
      c[oc[j]] = binop<OP, T>(x[j], y[j]);
    }
  }
}

template <class T, int OP>
static hipError_t mixed_go(const MixedArgs &m, const EwPlan &p, void *c, bool bad, const BadInfo *bi,
                           hipStream_t s) {
  int64_t grid = (p.n + MX_THREADS * MX_U - 1) / (MX_THREADS * MX_U);
  grid = grid < 1 ? 1 : (grid > 65536 ? 65536 : grid);
  const T cbad = bv_as<T>(bi[2].bv);
  if (bad)
    hipLaunchKernelGGL((ew_mixed<T, OP, true>), dim3((unsigned)grid), dim3(MX_THREADS), 0, s, m.a, m.ta, m.sbad_a,
                       m.sba, m.b, m.tb, m.sbad_b, m.sbb, (T *)c, p, bi[0], bi[1], cbad, m.a_scalar, m.b_scalar);
  else
    hipLaunchKernelGGL((ew_mixed<T, OP, false>), dim3((unsigned)grid), dim3(MX_THREADS), 0, s, m.a, m.ta, m.sbad_a,
                       m.sba, m.b, m.tb, m.sbad_b, m.sbb, (T *)c, p, bi[0], bi[1], cbad, m.a_scalar, m.b_scalar);
  return hipGetLastError();
}

template <class T>
static hipError_t mixed_T(int op, const MixedArgs &m, const EwPlan &p, void *c, bool bad, const BadInfo *bi,
                          hipStream_t s) {
  constexpr bool F = IsFloat<T>::v;
  switch (op) {
#define CASE(O) case O: return mixed_go<T, O>(m, p, c, bad, bi, s);
    CASE(OP_PLUS) CASE(OP_MULT) CASE(OP_MINUS) CASE(OP_DIVIDE) CASE(OP_GT) CASE(OP_LT) CASE(OP_LE) CASE(OP_GE)
    CASE(OP_EQ) CASE(OP_NE) CASE(OP_MODULO) CASE(OP_SPACESHIP) CASE(OP_SQRT) CASE(OP_SIN) CASE(OP_COS)
    CASE(OP_NOT) CASE(OP_LOG10) CASE(OP_ABS)
#undef CASE
  default: break;
  }
  if constexpr (!F) {
    switch (op) {
#define CASE(O) case O: return mixed_go<T, O>(m, p, c, bad, bi, s);
      CASE(OP_SHL) CASE(OP_SHR) CASE(OP_OR) CASE(OP_AND) CASE(OP_XOR) CASE(OP_BITNOT)
#undef CASE
    default: break;
    }
  } else {
    switch (op) {
#define CASE(O) case O: return mixed_go<T, O>(m, p, c, bad, bi, s);
      CASE(OP_POWER) CASE(OP_ATAN2) CASE(OP_EXP) CASE(OP_LOG)
#undef CASE
    default: break;
    }
  }
  return hipErrorInvalidValue;
}

#define PDLHIP_MIXED_TYPE(TC) \
  hipError_t launch_mixed_t##TC(int op, const MixedArgs &m, const EwPlan &p, void *c, bool bad, const BadInfo *bi, \
                                hipStream_t s) { \
    if (p.n == 0) return hipSuccess; \
    return mixed_T<CType<TC>::T>(op, m, p, c, bad, bi, s); \
  }
PDLHIP_MIXED_TYPES_IN_THIS_TU

}  // namespace pdlhip
