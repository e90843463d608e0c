// Device helpers: bad-value tests, the Ops.pd operators with C's integer
// promotion and wrap-around (PDL is built with -fwrapv), and offset math.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "common.hpp"

namespace pdlhip {

// A pointer read from a table in device memory is generic to the compiler,
// which then emits flat loads/stores and waits for all of them before any
// use; these casts state that the data is in global memory (it always is
// here), so the accesses become global_* instructions with counted waits.
template <class T> using gptr = T __attribute__((address_space(1))) *;
template <class T> __device__ inline gptr<const T> as_global(const T *p) { return (gptr<const T>)p; }
template <class T> __device__ inline gptr<T> as_global(T *p) { return (gptr<T>)p; }


template <class T> struct Unsigned { typedef T U; };
template <> struct Unsigned<int8_t> { typedef uint8_t U; };
template <> struct Unsigned<int16_t> { typedef uint16_t U; };
template <> struct Unsigned<int32_t> { typedef uint32_t U; };
template <> struct Unsigned<int64_t> { typedef uint64_t U; };

template <class T> struct IsFloat { static constexpr bool v = false; };
template <> struct IsFloat<float> { static constexpr bool v = true; };
template <> struct IsFloat<double> { static constexpr bool v = true; };

// C integer promotion: types narrower than int compute as int
template <class T> struct Promo { typedef T P; };
template <> struct Promo<int8_t> { typedef int32_t P; };
template <> struct Promo<uint8_t> { typedef int32_t P; };
template <> struct Promo<int16_t> { typedef int32_t P; };
template <> struct Promo<uint16_t> { typedef int32_t P; };

template <class T> __device__ __host__ inline T bv_as(const pdlhip_value &v) {
  T r;
  __builtin_memcpy(&r, &v, sizeof(T));
  return r;
}

template <class T> __device__ inline bool is_nan(T v) {
  if constexpr (IsFloat<T>::v) return v != v;
  else return false;
}

// PDL_ISBAD2 (pdl.h.PL:259-262)
template <class T> __device__ inline bool isbad(T v, T bv, int bvnan) {
  return bvnan ? is_nan(v) : (v == bv);
}

// wrapping ring arithmetic in T's width: what PDL's C gives under -fwrapv
template <class T> __device__ inline T wadd(T a, T b) {
  if constexpr (IsFloat<T>::v) return a + b;
  else { typedef typename Unsigned<T>::U U; return (T)(U)((uint64_t)(U)a + (uint64_t)(U)b); }
}
template <class T> __device__ inline T wsub(T a, T b) {
  if constexpr (IsFloat<T>::v) return a - b;
  else { typedef typename Unsigned<T>::U U; return (T)(U)((uint64_t)(U)a - (uint64_t)(U)b); }
}
template <class T> __device__ inline T wmul(T a, T b) {
  if constexpr (IsFloat<T>::v) return a * b;
  else if constexpr (sizeof(T) <= 4) {
    typedef typename Unsigned<T>::U U;
    return (T)(U)((uint32_t)(U)a * (uint32_t)(U)b);
  } else {
    return (T)((uint64_t)a * (uint64_t)b);
  }
}

// Ops.pd:68-70 MOD / BU_MOD / SPACE, with the same C conversions
#define PDLHIP_ABS(A) ((A) >= 0 ? (A) : -(A))
template <class T> __device__ inline T op_mod(T X, T N) {
  if constexpr (!IsFloat<T>::v && (T)-1 > (T)0) {   // unsigned: BU_MOD
    return (T)(((N) == 0) ? 0 : ((X) - (N) * ((uint64_t)((X) / (N)))));
  } else {
    return (T)(((N) == 0) ? 0
             : ((X) - (PDLHIP_ABS(N)) * ((long long)((X) / (PDLHIP_ABS(N))) +
                 ((((N) * ((long long)((X) / (N)))) != (X)) ? ((((N) < 0) ? 1 : 0) + ((((X) < 0) ? -1 : 0))) : 0))));
  }
}

template <class T> __device__ inline T shift_count_mask(T b) {
  typedef typename Promo<T>::P P;
  return (T)((uint64_t)b & (sizeof(P) * 8 - 1));
}

template <int OP, class T> __device__ inline T binop(T a, T b) {
  typedef typename Promo<T>::P P;
  if constexpr (OP == OP_PLUS) return wadd(a, b);
  else if constexpr (OP == OP_MULT) return wmul(a, b);
  else if constexpr (OP == OP_MINUS) return wsub(a, b);
  else if constexpr (OP == OP_DIVIDE) {
    if constexpr (IsFloat<T>::v) return a / b;
    else return (T)((P)a / (P)b);
  }
  else if constexpr (OP == OP_GT) return (T)(a > b);
  else if constexpr (OP == OP_LT) return (T)(a < b);
  else if constexpr (OP == OP_LE) return (T)(a <= b);
  else if constexpr (OP == OP_GE) return (T)(a >= b);
  else if constexpr (OP == OP_EQ) return (T)(a == b);
  else if constexpr (OP == OP_NE) return (T)(a != b);
  else if constexpr (OP == OP_SHL) {
    if constexpr (IsFloat<T>::v) return a;
    else { typedef typename Unsigned<P>::U UP; return (T)(P)((UP)(P)a << (int)shift_count_mask(b)); }
  }
  else if constexpr (OP == OP_SHR) {
    if constexpr (IsFloat<T>::v) return a;
    else return (T)((P)a >> (int)shift_count_mask(b));
  }
  else if constexpr (OP == OP_OR) { if constexpr (IsFloat<T>::v) return a; else return (T)(a | b); }
  else if constexpr (OP == OP_AND) { if constexpr (IsFloat<T>::v) return a; else return (T)(a & b); }
  else if constexpr (OP == OP_XOR) { if constexpr (IsFloat<T>::v) return a; else return (T)(a ^ b); }
  else if constexpr (OP == OP_POWER) {
    // a**2 is a*a: the correctly rounded square (NaN, infinities, signed
    // zeros and overflow behave as pow's).  glibc's powf/pow(x, 2), which
    // PDL calls, is accurate to < 1 ulp but not correctly rounded: it
    // differs from a*a by one ulp on ~0.07% of float inputs
    // (powf(x,2) vs x*x over 200000 normal floats); libm pow costs ~60 VALU
    // instructions per element.
    if constexpr (IsFloat<T>::v) { if (b == T(2)) return a * a; }
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return powf(a, b);
    else if constexpr (IsFloat<T>::v) return pow(a, b);
    else return a;
  }
  else if constexpr (OP == OP_ATAN2) {
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return atan2f(a, b);
    else if constexpr (IsFloat<T>::v) return atan2(a, b);
    else return a;
  }
  else if constexpr (OP == OP_MODULO) return op_mod<T>(a, b);
  else if constexpr (OP == OP_SPACESHIP) return (T)(((a) < (b)) ? -1 : ((a) != (b)));
  // ufuncs: b = ($GENERIC(b))func(a) through <tgmath.h> (an integer
  // argument takes the double function; the cast back truncates)
  else if constexpr (OP == OP_BITNOT) { if constexpr (IsFloat<T>::v) return a; else return (T)(~(P)a); }
  else if constexpr (OP == OP_NOT) return (T)(!a);
  else if constexpr (OP == OP_ABS) {
    if constexpr (!IsFloat<T>::v && (T)-1 > (T)0) return a;     // unsigned: _rabs copies
    else return (T)PDLHIP_ABS(a);                                // PDL_ABS: -0.0 stays, NaN flips sign
  }
  else if constexpr (OP == OP_SQRT) {
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return sqrtf(a);
    else return (T)sqrt((double)a);
  }
  else if constexpr (OP == OP_SIN) {
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return sinf(a);
    else return (T)sin((double)a);
  }
  else if constexpr (OP == OP_COS) {
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return cosf(a);
    else return (T)cos((double)a);
  }
  else if constexpr (OP == OP_EXP) {
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return expf(a);
    else return (T)exp((double)a);
  }
  else if constexpr (OP == OP_LOG) {
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return logf(a);
    else return (T)log((double)a);
  }
  else if constexpr (OP == OP_LOG10) {
    if constexpr (sizeof(T) == 4 && IsFloat<T>::v) return log10f(a);
    else if constexpr (IsFloat<T>::v) return log10(a);
    else {
      // an exact power of ten gives its exponent exactly (as glibc's log10
      // does), so the truncation back to the integer type cannot land one
      // below it
      const double x = (double)a, r = log10(x);
      const double k = rint(r);
      if (k >= 0.0 && k <= 22.0 && fabs(r - k) < 1e-9) {
        double t = 1.0;
        for (int i = 0; i < (int)k; i++) t *= 10.0;   // exact up to 1e22
        if (t == x) return (T)k;
      }
      return (T)r;
    }
  }
  else return a;
}

// VEC elements loaded/stored as one 16-byte (or narrower) access
template <class T, int VEC> struct alignas(sizeof(T) * VEC) Vec { T v[VEC]; };

// 16-byte store that bypasses cache residency (streaming output)
template <class T, int VEC> __device__ inline void store_nt(T *p, const Vec<T, VEC> &v) {
  if constexpr (sizeof(T) * VEC == 16) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 w;
    __builtin_memcpy(&w, &v, 16);
    __builtin_nontemporal_store(w, reinterpret_cast<u32x4 *>(p));
  } else {
    *reinterpret_cast<Vec<T, VEC> *>(p) = v;
  }
}

// 16-byte streaming load
template <class T, int VEC> __device__ inline Vec<T, VEC> load_nt(const T *p) {
  Vec<T, VEC> v;
  if constexpr (sizeof(T) * VEC == 16) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
    __builtin_memcpy(&v, &w, 16);
  } else {
    v = *reinterpret_cast<const Vec<T, VEC> *>(p);
  }
  return v;
}

// offsets of element (i0, row) where `row` enumerates dims 1..nd-1
struct RowOffsets {
  int64_t o[3];
};

__device__ inline void row_offsets(const EwPlan &p, int64_t row, int64_t *o) {
  o[0] = p.off[0]; o[1] = p.off[1]; o[2] = p.off[2];
  for (int d = 1; d < p.nd; d++) {
    int64_t dd = p.dims[d];
    int64_t i = row % dd;
    row /= dd;
    o[0] += i * p.inc[0][d]; o[1] += i * p.inc[1][d]; o[2] += i * p.inc[2][d];
  }
}

}  // namespace pdlhip
