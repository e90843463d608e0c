// Complex ndarrays on the device: PDL's CF (float complex) and CD (double
// complex) types (Types.pm), for the Ops.pd biops plus, minus, mult, eq and
// ne (GenericTypes $A, Ops.pd:288-301), assgn and converttypei
// (pdlconv.c:163-174), the Ufunc.pd sums and products sumover, prodover,
// cumusumover, cumuprodover (Ufunc.pd:88-140, int+ of a complex type is
// itself), and matmult (Primitive.pd:191-245, ppdefs_all).
//
// The arithmetic is C99 complex arithmetic as GCC compiles PDL's
// `$c() = $a() * $b()` and `cc += *ad * *bd` (no FMA contraction on the
// x86-64 baseline, as this build uses -ffp-contract=off): a product is
// (ac - bd) + i(ad + bc), every product and sum rounded once in the real
// type; when both parts come out NaN the infinities are recovered as C99
// Annex G (G.5.1, the _Cmultd example) specifies and libgcc's
// __mulsc3/__muldc3 do.  A complex converted to a real type keeps its real
// part (through intmax_t for an unsigned target); a real one becomes
// (x, 0).  Bad values in complex ndarrays stay on the CPU (refused by the
// host).
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

template <class R> struct alignas(2 * sizeof(R)) Cplx { R re, im; };

template <class R> __device__ inline R csign(R mag, R sgn) { return __builtin_copysign(mag, sgn); }
template <> __device__ inline float csign<float>(float mag, float sgn) { return __builtin_copysignf(mag, sgn); }

template <class R> __device__ inline Cplx<R> cmul(Cplx<R> x, Cplx<R> y) {
  R a = x.re, b = x.im, c = y.re, d = y.im;
  const R ac = a * c, bd = b * d, ad = a * d, bc = b * c;
  R re = ac - bd, im = ad + bc;
  if (__builtin_isnan(re) && __builtin_isnan(im)) {
    bool recalc = false;
    if (__builtin_isinf(a) || __builtin_isinf(b)) {          // x infinite: box it, NaNs of y to 0
      a = csign<R>(__builtin_isinf(a) ? R(1) : R(0), a);
      b = csign<R>(__builtin_isinf(b) ? R(1) : R(0), b);
      if (__builtin_isnan(c)) c = csign<R>(R(0), c);
      if (__builtin_isnan(d)) d = csign<R>(R(0), d);
      recalc = true;
    }
    if (__builtin_isinf(c) || __builtin_isinf(d)) {          // y infinite
      c = csign<R>(__builtin_isinf(c) ? R(1) : R(0), c);
      d = csign<R>(__builtin_isinf(d) ? R(1) : R(0), d);
      if (__builtin_isnan(a)) a = csign<R>(R(0), a);
      if (__builtin_isnan(b)) b = csign<R>(R(0), b);
      recalc = true;
    }
    if (!recalc && (__builtin_isinf(ac) || __builtin_isinf(bd) || __builtin_isinf(ad) || __builtin_isinf(bc))) {
      if (__builtin_isnan(a)) a = csign<R>(R(0), a);            // overflow: NaNs to 0
      if (__builtin_isnan(b)) b = csign<R>(R(0), b);
      if (__builtin_isnan(c)) c = csign<R>(R(0), c);
      if (__builtin_isnan(d)) d = csign<R>(R(0), d);
      recalc = true;
    }
    if (recalc) {
      const R inf = __builtin_huge_val();
      re = inf * (a * c - b * d);
      im = inf * (a * d + b * c);
    }
  }
  return {re, im};
}

template <class R> __device__ inline R cfabs(R v) { return v < R(0) ? -v : (v == R(0) ? R(0) : v); }

// x / y as the C99 `/` of two complex operands compiles under GCC 11: a call
// to libgcc's __divsc3 / __divdc3, Smith's method in the operands' own type
// (the ratio of y's smaller part to its larger), then C99 Annex G's recovery
// of infinities and zeros when both parts came out NaN (G.5.1)
template <class R> __device__ inline Cplx<R> cdiv(Cplx<R> x, Cplx<R> y) {
  R a = x.re, b = x.im, c = y.re, d = y.im, re, im;
  if (cfabs(c) < cfabs(d)) {
    const R ratio = c / d, denom = c * ratio + d;
    re = (a * ratio + b) / denom;
    im = (b * ratio - a) / denom;
  } else {
    const R ratio = d / c, denom = d * ratio + c;
    re = (b * ratio + a) / denom;
    im = (b - a * ratio) / denom;
  }
  if (__builtin_isnan(re) && __builtin_isnan(im)) {
    const R inf = __builtin_huge_val();
    if (c == R(0) && d == R(0) && (!__builtin_isnan(a) || !__builtin_isnan(b))) {
      re = csign<R>(inf, c) * a;                                // a nonzero x over zero
      im = csign<R>(inf, c) * b;
    } else if ((__builtin_isinf(a) || __builtin_isinf(b)) && __builtin_isfinite(c) && __builtin_isfinite(d)) {
      a = csign<R>(__builtin_isinf(a) ? R(1) : R(0), a);        // infinite x over finite y
      b = csign<R>(__builtin_isinf(b) ? R(1) : R(0), b);
      re = inf * (a * c + b * d);
      im = inf * (b * c - a * d);
    } else if ((__builtin_isinf(c) || __builtin_isinf(d)) && __builtin_isfinite(a) && __builtin_isfinite(b)) {
      c = csign<R>(__builtin_isinf(c) ? R(1) : R(0), c);        // finite x over infinite y
      d = csign<R>(__builtin_isinf(d) ? R(1) : R(0), d);
      re = R(0) * (a * c + b * d);
      im = R(0) * (b * c - a * d);
    }
  }
  return {re, im};
}

template <class R> __device__ inline Cplx<R> cbinop(int op, Cplx<R> x, Cplx<R> y) {
  switch (op) {
  case OP_PLUS: return {x.re + y.re, x.im + y.im};
  case OP_MINUS: return {x.re - y.re, x.im - y.im};
  case OP_MULT: return cmul(x, y);
  case OP_DIVIDE: return cdiv(x, y);
  case OP_EQ: return {(x.re == y.re && x.im == y.im) ? R(1) : R(0), R(0)};
  case OP_NE: return {(x.re != y.re || x.im != y.im) ? R(1) : R(0), R(0)};
  default: return x;
  }
}

// element e of a canonical plan: operand offsets (dim 0 fastest)
__device__ inline void cplx_offsets(const EwPlan &p, int64_t e, int64_t *o) {
  o[0] = p.off[0];
  o[1] = p.off[1];
  o[2] = p.off[2];
  for (int d = 0; d < p.nd; d++) {
    const int64_t i = e % p.dims[d];
    e /= p.dims[d];
    o[0] += i * p.inc[0][d];
    o[1] += i * p.inc[1][d];
    o[2] += i * p.inc[2][d];
  }
}

template <class R>
__global__ __launch_bounds__(256) void k_cplx_binop(const Cplx<R> *a, const Cplx<R> *b,
                                                    Cplx<R> *c, EwPlan p, int op) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < p.n; e += (int64_t)gridDim.x * 256) {
    int64_t o[3];
    cplx_offsets(p, e, o);
    c[o[2]] = cbinop<R>(op, a[o[0]], b[o[1]]);
  }
}

// one element of converttypei with a complex end
template <class D, class S> __device__ inline D to_real(S v) {
  if constexpr (!IsFloat<D>::v && (D)-1 > (D)0) return (D)(intmax_t)v;
  else return (D)v;
}
template <class S> struct IsCplx { static constexpr bool v = false; };
template <class R> struct IsCplx<Cplx<R>> { static constexpr bool v = true; };

template <class D, class S> __device__ inline D cconv(S v) {
  if constexpr (IsCplx<D>::v && IsCplx<S>::v) {
    typedef decltype(D().re) RD;
    return D{(RD)v.re, (RD)v.im};
  } else if constexpr (IsCplx<D>::v) {
    typedef decltype(D().re) RD;
    return D{(RD)v, RD(0)};
  } else {
    return to_real<D>(v.re);
  }
}

template <class D, class S>
__global__ __launch_bounds__(256) void k_cplx_convert(const S *src, D *dst, EwPlan p) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < p.n; e += (int64_t)gridDim.x * 256) {
    int64_t o[3];
    cplx_offsets(p, e, o);
    dst[o[2]] = cconv<D, S>(src[o[0]]);
  }
}

// matmult: one output per thread, threads along w; PDL's sum in ascending t
// from zero, each step a complex product then a complex sum
template <class R>
__global__ __launch_bounds__(256) void k_cplx_matmult(const Cplx<R> *__restrict__ a, const Cplx<R> *__restrict__ b,
                                                      Cplx<R> *__restrict__ c, MmPlan p) {
  const int64_t per = p.H * p.W, total = per * p.nbatch;
  for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (int64_t)gridDim.x * 256) {
    const int64_t w = g % p.W, h = (g / p.W) % p.H;
    int64_t r = g / per, oa = p.off[0], ob = p.off[1], oc = p.off[2];
    for (int d = 0; d < p.nd; d++) {
      const int64_t i = r % p.dims[d];
      r /= p.dims[d];
      oa += i * p.inc[0][d];
      ob += i * p.inc[1][d];
      oc += i * p.inc[2][d];
    }
    const Cplx<R> *ap = a + oa + h * p.a_h;
    const Cplx<R> *bp = b + ob + w * p.b_w;
    Cplx<R> cc = {R(0), R(0)};
    for (int64_t t = 0; t < p.T; t++) {
      const Cplx<R> m = cmul(ap[t * p.a_t], bp[t * p.b_t]);
      cc.re = cc.re + m.re;
      cc.im = cc.im + m.im;
    }
    c[oc + w * p.c_w + h * p.c_h] = cc;
  }
}

static unsigned grid_for(int64_t n) {
  int64_t blocks = (n + 255) / 256;
  return (unsigned)(blocks < 1 ? 1 : (blocks > 65536 ? 65536 : blocks));
}

hipError_t launch_cplx_binop(int T, int op, const EwPlan &p, const void *a, const void *b, void *c, hipStream_t s) {
  if (p.n == 0) return hipSuccess;
  if (T == PDLHIP_CF)
    hipLaunchKernelGGL(k_cplx_binop<float>, dim3(grid_for(p.n)), dim3(256), 0, s, (const Cplx<float> *)a,
                       (const Cplx<float> *)b, (Cplx<float> *)c, p, op);
  else if (T == PDLHIP_CD)
    hipLaunchKernelGGL(k_cplx_binop<double>, dim3(grid_for(p.n)), dim3(256), 0, s, (const Cplx<double> *)a,
                       (const Cplx<double> *)b, (Cplx<double> *)c, p, op);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <class D>
static hipError_t cconv_D(int stype, const EwPlan &p, const void *src, void *dst, hipStream_t s) {
  switch (stype) {
#define X(TC)                                                                                                \
  case TC:                                                                                                   \
    hipLaunchKernelGGL((k_cplx_convert<D, CType<TC>::T>), dim3(grid_for(p.n)), dim3(256), 0, s,             \
                       (const CType<TC>::T *)src, (D *)dst, p);                                              \
    break;
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  case PDLHIP_CF:
    hipLaunchKernelGGL((k_cplx_convert<D, Cplx<float>>), dim3(grid_for(p.n)), dim3(256), 0, s,
                       (const Cplx<float> *)src, (D *)dst, p);
    break;
  case PDLHIP_CD:
    hipLaunchKernelGGL((k_cplx_convert<D, Cplx<double>>), dim3(grid_for(p.n)), dim3(256), 0, s,
                       (const Cplx<double> *)src, (D *)dst, p);
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cplx_convert(int stype, int dtype, const EwPlan &p, const void *src, void *dst, hipStream_t s) {
  if (p.n == 0) return hipSuccess;
  switch (dtype) {
  case PDLHIP_CF: return cconv_D<Cplx<float>>(stype, p, src, dst, s);
  case PDLHIP_CD: return cconv_D<Cplx<double>>(stype, p, src, dst, s);
#define X(TC)                                                                                                \
  case TC:                                                                                                   \
    if (stype == PDLHIP_CF)                                                                                  \
      hipLaunchKernelGGL((k_cplx_convert<CType<TC>::T, Cplx<float>>), dim3(grid_for(p.n)), dim3(256), 0, s,  \
                         (const Cplx<float> *)src, (CType<TC>::T *)dst, p);                                  \
    else if (stype == PDLHIP_CD)                                                                             \
      hipLaunchKernelGGL((k_cplx_convert<CType<TC>::T, Cplx<double>>), dim3(grid_for(p.n)), dim3(256), 0, s, \
                         (const Cplx<double> *)src, (CType<TC>::T *)dst, p);                                 \
    else                                                                                                     \
      return hipErrorInvalidValue;                                                                           \
    return hipGetLastError();
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

// sumover / prodover (int+ = the complex type) and their cumulative forms
// (Ufunc.pd:88-140): one thread per output row, PDL's sequential loop;
// prodover stops once the running product is zero (both parts)
template <class R, int OP>
__global__ __launch_bounds__(256) void k_cplx_reduce(const Cplx<R> *a, Cplx<R> *b, RedPlan p) {
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < p.nout; r += (int64_t)gridDim.x * 256) {
    int64_t rest = r, oa = p.off[0], ob = p.off[1];
    for (int d = 0; d < p.nd; d++) {
      const int64_t i = rest % p.dims[d];
      rest /= p.dims[d];
      oa += i * p.inc[0][d];
      ob += i * p.inc[1][d];
    }
    const bool prod = OP == R_PROD || OP == R_CUMUPROD;
    Cplx<R> tmp = {prod ? R(1) : R(0), R(0)};
    for (int64_t t = 0; t < p.n; t++) {
      const Cplx<R> v = a[oa + t * p.inc_n[0]];
      if (prod) tmp = cmul(tmp, v);
      else tmp = {tmp.re + v.re, tmp.im + v.im};
      if (OP == R_CUMUSUM || OP == R_CUMUPROD) b[ob + t * p.inc_n[1]] = tmp;
      if (OP == R_PROD && tmp.re == R(0) && tmp.im == R(0)) break;
    }
    if (OP == R_SUM || OP == R_PROD) b[ob] = tmp;
  }
}

hipError_t launch_cplx_reduce(int T, int op, const RedPlan &p, const void *a, void *b, hipStream_t s) {
  if (p.nout == 0) return hipSuccess;
#define CR(R, OP) hipLaunchKernelGGL((k_cplx_reduce<R, OP>), dim3(grid_for(p.nout)), dim3(256), 0, s, (const Cplx<R> *)a, (Cplx<R> *)b, p)
#define CRT(R) switch (op) { case R_SUM: CR(R, R_SUM); break; case R_PROD: CR(R, R_PROD); break; \
    case R_CUMUSUM: CR(R, R_CUMUSUM); break; case R_CUMUPROD: CR(R, R_CUMUPROD); break; default: return hipErrorInvalidValue; }
  if (T == PDLHIP_CF) { CRT(float) } else if (T == PDLHIP_CD) { CRT(double) } else return hipErrorInvalidValue;
#undef CRT
#undef CR
  return hipGetLastError();
}

hipError_t launch_cplx_matmult(int T, const MmPlan &p, const void *a, const void *b, void *c, hipStream_t s) {
  const int64_t total = p.H * p.W * p.nbatch;
  if (total == 0) return hipSuccess;
  if (T == PDLHIP_CF)
    hipLaunchKernelGGL(k_cplx_matmult<float>, dim3(grid_for(total)), dim3(256), 0, s, (const Cplx<float> *)a,
                       (const Cplx<float> *)b, (Cplx<float> *)c, p);
  else if (T == PDLHIP_CD)
    hipLaunchKernelGGL(k_cplx_matmult<double>, dim3(grid_for(total)), dim3(256), 0, s, (const Cplx<double> *)a,
                       (const Cplx<double> *)b, (Cplx<double> *)c, p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace pdlhip
