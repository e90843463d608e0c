// Shared reduction machinery (the per-lane/per-block partial of a row and
// how partials combine), used by the row reductions of k_reduce.hip and the
// ragged-list reductions of k_ragged.hip.  The rules restate
// lib/PDL/Ufunc.pd: sums/products in the output type (integers wrap),
// min/max with PDL's NaN and tie rules (Ufunc.pd:446-482), minmaximum
// skipping NaN (Ufunc.pd:563-613), average's integer truncation and empty
// results (Ufunc.pd:413-444).
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace pdlhip {

template <class T> struct IPlus { typedef T O; };
template <> struct IPlus<int8_t> { typedef int32_t O; };
template <> struct IPlus<uint8_t> { typedef int32_t O; };
template <> struct IPlus<int16_t> { typedef int32_t O; };
template <> struct IPlus<uint16_t> { typedef int32_t O; };

constexpr bool is_sum(int op) { return op == R_SUM || op == R_DSUM || op == R_AVERAGE || op == R_DAVERAGE; }
constexpr bool is_prod(int op) { return op == R_PROD || op == R_DPROD; }
constexpr bool is_minmax(int op) { return op == R_MIN || op == R_MAX || op == R_MIN_IND || op == R_MAX_IND; }
constexpr bool is_max(int op) { return op == R_MAX || op == R_MAX_IND; }

// ---------------------------------------------------------------- partials
// Everything one lane / block knows about its share of a row.
template <class Tin, class Tacc> struct Part {
  Tacc s, c;          // sum + compensation (floats), wrapped sum/product (ints)
  int64_t cnt;        // good elements seen
  Tin v0; int64_t i0; // extreme (min for minmaximum); i0 < 0: none
  Tin v1; int64_t i1; // max for minmaximum
};

template <int OP, class Tin, class Tacc> __device__ inline void part_init(Part<Tin, Tacc> &p) {
  p.s = is_prod(OP) ? Tacc(1) : Tacc(0);
  p.c = Tacc(0);
  p.cnt = 0;
  p.v0 = Tin(0); p.i0 = -1;
  p.v1 = Tin(0); p.i1 = -1;
}

// Neumaier step (floats) or wrapping add (ints)
template <class Tacc> __device__ inline void acc_add(Tacc &s, Tacc &c, Tacc x) {
  if constexpr (IsFloat<Tacc>::v) {
    Tacc t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  } else {
    s = wadd(s, x);
  }
}

// PDL's min/max preference as a total order on (value, index): non-NaN beats
// NaN; smaller (larger for max) value wins; equal values: earlier index; two
// NaNs: later index.
template <bool MAX, class T> __device__ inline bool pref(T v, int64_t i, T w, int64_t j) {
  bool vn = is_nan(v), wn = is_nan(w);
  if (vn != wn) return !vn;
  if (vn) return i > j;
  if (MAX ? (v > w) : (v < w)) return true;
  if (MAX ? (w > v) : (w < v)) return false;
  return i < j;
}

template <int OP, class Tin, class Tacc> __device__ inline void part_push(Part<Tin, Tacc> &p, Tin x, int64_t i) {
  if constexpr (is_sum(OP)) {
    acc_add<Tacc>(p.s, p.c, (Tacc)x);
    p.cnt++;
  } else if constexpr (is_prod(OP)) {
    p.s = wmul(p.s, (Tacc)x);
    p.cnt++;
  } else if constexpr (is_minmax(OP)) {
    // elements reach a lane in increasing index order, so PDL's sequential
    // rule (Ufunc.pd:458-460) is exact here: replace on strict improvement or
    // when the current value is NaN
    bool better = is_max(OP) ? (x > p.v0) : (x < p.v0);
    if (p.i0 < 0 || better || is_nan(p.v0)) { p.v0 = x; p.i0 = i; }
  } else if constexpr (OP == R_MINMAXIMUM) {
    if (is_nan(x)) return;
    if (p.i0 < 0 || x < p.v0) { p.v0 = x; p.i0 = i; }
    if (p.i1 < 0 || x > p.v1) { p.v1 = x; p.i1 = i; }
  }
}

// double-double add for combining compensated sums
__device__ inline void dd_add(double &hi, double &lo, double bh, double bl) {
  double s = hi + bh;
  double bb = s - hi;
  double e = (hi - (s - bb)) + (bh - bb);
  e += lo + bl;
  hi = s + e;
  lo = e - (hi - s);
}

template <int OP, class Tin, class Tacc> __device__ inline void part_combine(Part<Tin, Tacc> &p, const Part<Tin, Tacc> &q) {
  if constexpr (is_sum(OP)) {
    if constexpr (IsFloat<Tacc>::v) {
      double hi = (double)p.s, lo = (double)p.c;
      dd_add(hi, lo, (double)q.s, (double)q.c);
      // keep the pair in Tacc: for float accumulators hold hi/lo as the
      // float-rounded pair; the final value is formed in double
      p.s = (Tacc)hi;
      p.c = (Tacc)((hi - (double)p.s) + lo);
    } else {
      p.s = wadd(p.s, q.s);
    }
    p.cnt += q.cnt;
  } else if constexpr (is_prod(OP)) {
    p.s = wmul(p.s, q.s);
    p.cnt += q.cnt;
  } else if constexpr (is_minmax(OP)) {
    if (q.i0 >= 0 && (p.i0 < 0 || pref<is_max(OP)>(q.v0, q.i0, p.v0, p.i0))) { p.v0 = q.v0; p.i0 = q.i0; }
  } else if constexpr (OP == R_MINMAXIMUM) {
    if (q.i0 >= 0 && (p.i0 < 0 || q.v0 < p.v0 || (q.v0 == p.v0 && q.i0 < p.i0))) { p.v0 = q.v0; p.i0 = q.i0; }
    if (q.i1 >= 0 && (p.i1 < 0 || q.v1 > p.v1 || (q.v1 == p.v1 && q.i1 < p.i1))) { p.v1 = q.v1; p.i1 = q.i1; }
  }
}

template <class T> __device__ inline T shfl_xor(T v, int m) {
  if constexpr (sizeof(T) == 8) {
    int64_t x;
    __builtin_memcpy(&x, &v, 8);
    int lo = __shfl_xor((int)(x & 0xffffffff), m, 64), hi = __shfl_xor((int)(x >> 32), m, 64);
    x = ((int64_t)(uint32_t)lo) | ((int64_t)hi << 32);
    T r;
    __builtin_memcpy(&r, &x, 8);
    return r;
  } else if constexpr (sizeof(T) == 4) {
    int x;
    __builtin_memcpy(&x, &v, 4);
    x = __shfl_xor(x, m, 64);
    T r;
    __builtin_memcpy(&r, &x, 4);
    return r;
  } else {
    int x = (int)v;
    x = __shfl_xor(x, m, 64);
    return (T)x;
  }
}

template <class T> __device__ inline T shfl_up(T v, int d) {
  if constexpr (sizeof(T) == 8) {
    int64_t x;
    __builtin_memcpy(&x, &v, 8);
    int lo = __shfl_up((int)(x & 0xffffffff), d, 64), hi = __shfl_up((int)(x >> 32), d, 64);
    x = ((int64_t)(uint32_t)lo) | ((int64_t)hi << 32);
    T r;
    __builtin_memcpy(&r, &x, 8);
    return r;
  } else {
    int x;
    __builtin_memcpy(&x, &v, 4);
    x = __shfl_up(x, d, 64);
    T r;
    __builtin_memcpy(&r, &x, 4);
    return r;
  }
}

template <int OP, class Tin, class Tacc> __device__ inline void wave_reduce(Part<Tin, Tacc> &p) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    Part<Tin, Tacc> q;
    q.s = shfl_xor(p.s, m);
    q.c = shfl_xor(p.c, m);
    q.cnt = shfl_xor(p.cnt, m);
    q.v0 = shfl_xor(p.v0, m);
    q.i0 = shfl_xor(p.i0, m);
    q.v1 = shfl_xor(p.v1, m);
    q.i1 = shfl_xor(p.i1, m);
    part_combine<OP>(p, q);
  }
}

// ---------------------------------------------------------------- outputs
struct Outs {
  void *p[4];
  pdlhip_value bv[4];
};

template <class T> __device__ inline void put(void *base, int64_t off, T v) { as_global((T *)base)[off] = v; }

template <class Tout> __device__ inline Tout avg_div(Tout tmp, int64_t cnt) {
  if constexpr (IsFloat<Tout>::v) return tmp / (Tout)cnt;
  else if constexpr (sizeof(Tout) == 8 && (Tout)-1 > (Tout)0) return tmp / (uint64_t)cnt;
  else return (Tout)((int64_t)tmp / cnt);
}

template <int OP, class Tin, class Tout, class Tacc>
__device__ inline void part_store(const Part<Tin, Tacc> &p, const Outs &o, const int64_t *off, bool bad) {
  if constexpr (OP == R_SUM || OP == R_DSUM || OP == R_PROD || OP == R_DPROD) {
    Tout v;
    if constexpr (IsFloat<Tacc>::v && !is_prod(OP)) v = (Tout)((double)p.s + (double)p.c);
    else v = (Tout)p.s;
    if (bad && p.cnt == 0) v = bv_as<Tout>(o.bv[0]);
    put<Tout>(o.p[0], off[1], v);
  } else if constexpr (OP == R_AVERAGE || OP == R_DAVERAGE) {
    Tout v;
    if (p.cnt == 0) {
      if (bad) v = bv_as<Tout>(o.bv[0]);
      else v = IsFloat<Tin>::v ? (Tout)NAN : (Tout)0;
    } else {
      Tout tmp;
      if constexpr (IsFloat<Tacc>::v) tmp = (Tout)((double)p.s + (double)p.c);
      else tmp = (Tout)p.s;
      v = avg_div<Tout>(tmp, p.cnt);
    }
    put<Tout>(o.p[0], off[1], v);
  } else if constexpr (OP == R_MIN || OP == R_MAX) {
    put<Tin>(o.p[0], off[1], p.i0 >= 0 ? p.v0 : bv_as<Tin>(o.bv[0]));
  } else if constexpr (OP == R_MIN_IND || OP == R_MAX_IND) {
    put<int64_t>(o.p[0], off[1], p.i0 >= 0 ? p.i0 : bv_as<int64_t>(o.bv[0]));
  } else if constexpr (OP == R_MINMAXIMUM) {
    bool f = p.i0 >= 0;
    put<Tin>(o.p[0], off[1], f ? p.v0 : bv_as<Tin>(o.bv[0]));
    put<Tin>(o.p[1], off[2], f ? p.v1 : bv_as<Tin>(o.bv[1]));
    put<int64_t>(o.p[2], off[3], f ? p.i0 : bv_as<int64_t>(o.bv[2]));
    put<int64_t>(o.p[3], off[4], f ? p.i1 : bv_as<int64_t>(o.bv[3]));
  }
}

}  // namespace pdlhip
