// Bad.pd's elementwise ops over a canonical plan (the reference's loops:
// lib/PDL/Bad.pd:343-395 isbad/isgood, 584-634 setbadif, 639-676
// setvaltobad, 679-775 setnantobad/setinftobad/setnonfinitetobad, 777-806
// setbadtonan, 808-840 setbadtoval).
//
// $ISBAD(x) compiles to PDL_ISBAD2(x, x_badval, ...) in the good and the bad
// code alike (PP/PDLCode.pm:664-668); only PDL_IF_BAD(t, f) depends on the
// trans's bvalflag.  So setbadtonan and setvaltobad test values in both
// modes, isbad/isgood/setbadif/setbadtoval only in bad mode.
//
// One thread per output element, grid-stride; the plan orders dims by the
// output's strides, so consecutive threads write consecutive elements.  The
// ops that set the output's bad flag only when they found something
// (setnantobad, setinftobad, setnonfinitetobad) raise a device flag the host
// reads back.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

constexpr int BADOP_THREADS = 256;

template <class T> __device__ inline bool finite_v(T v) {
  if constexpr (IsFloat<T>::v) return isfinite(v);
  else return true;
}

template <class T, class To, int OP>
__global__ __launch_bounds__(BADOP_THREADS) void bad_ew(const T *__restrict__ a, const int32_t *__restrict__ m,
                                                        To *__restrict__ b, EwPlan p, BadInfo ba, int32_t mbv,
                                                        To bbv, T val, int bad, int *flag) {
  const T abv = bv_as<T>(ba.bv);
  bool hit = false;
  for (int64_t e = (int64_t)blockIdx.x * BADOP_THREADS + threadIdx.x; e < p.n;
       e += (int64_t)gridDim.x * BADOP_THREADS) {
    int64_t r = e, o0 = p.off[0], o1 = p.off[1], o2 = p.off[2];
    for (int d = 0; d < p.nd; d++) {
      const int64_t i = r % p.dims[d];
      r /= p.dims[d];
      o0 += i * p.inc[0][d];
      o1 += i * p.inc[1][d];
      o2 += i * p.inc[2][d];
    }
    const T x = a[o0];
    if constexpr (OP == B_ISBAD) {
      b[o2] = (To)(bad ? isbad(x, abv, ba.isnan) : 0);
    } else if constexpr (OP == B_ISGOOD) {
      b[o2] = (To)(bad ? !isbad(x, abv, ba.isnan) : 1);
    } else if constexpr (OP == B_SETBADIF) {
      const int32_t mk = m[o1];
      b[o2] = ((bad && mk == mbv) || mk) ? bbv : (To)x;
    } else if constexpr (OP == B_SETVALTOBAD) {
      b[o2] = (x == val) ? bbv : (To)x;
    } else if constexpr (OP == B_SETNANTOBAD || OP == B_SETINFTOBAD || OP == B_SETNONFINITETOBAD) {
      bool set;
      if constexpr (OP == B_SETNANTOBAD) set = is_nan(x);
      else if constexpr (OP == B_SETINFTOBAD) set = !finite_v(x) && !is_nan(x);
      else set = !finite_v(x);
      b[o2] = set ? bbv : (To)x;
      hit |= set;
    } else if constexpr (OP == B_SETBADTONAN) {
      if constexpr (IsFloat<T>::v) b[o2] = isbad(x, abv, ba.isnan) ? (To)__builtin_nan("") : (To)x;
    } else if constexpr (OP == B_SETBADTOVAL) {
      b[o2] = (bad && isbad(x, abv, ba.isnan)) ? (To)val : (To)x;
    }
  }
  if (hit) *flag = 1;   // the same value from every writer
}

template <class T, class To, int OP>
static hipError_t bad_go(const EwPlan &p, const void *a, const void *m, void *b, const BadInfo &ba, int32_t mbv,
                         const pdlhip_value &bbv, double dval, bool bad, int *flag, hipStream_t s) {
  int64_t grid = (p.n + BADOP_THREADS * 4 - 1) / (BADOP_THREADS * 4);
  grid = grid < 1 ? 1 : (grid > 65536 ? 65536 : grid);
  // $GENERIC(a) $COMP(value): the double OtherPar in the trans type, as C converts it
  const T val = (T)dval;
  hipLaunchKernelGGL((bad_ew<T, To, OP>), dim3((unsigned)grid), dim3(BADOP_THREADS), 0, s, (const T *)a,
                     (const int32_t *)m, (To *)b, p, ba, mbv, bv_as<To>(bbv), val, bad ? 1 : 0, flag);
  return hipGetLastError();
}

template <class T>
static hipError_t bad_T(int op, const EwPlan &p, const void *a, const void *m, void *b, const BadInfo &ba,
                        int32_t mbv, const pdlhip_value &bbv, double dval, bool bad, int *flag, hipStream_t s) {
  switch (op) {
  case B_ISBAD: return bad_go<T, int32_t, B_ISBAD>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
  case B_ISGOOD: return bad_go<T, int32_t, B_ISGOOD>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
  case B_SETBADIF: return bad_go<T, T, B_SETBADIF>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
  case B_SETVALTOBAD: return bad_go<T, T, B_SETVALTOBAD>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
  case B_SETBADTOVAL: return bad_go<T, T, B_SETBADTOVAL>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
  default: break;
  }
  if constexpr (IsFloat<T>::v) {
    switch (op) {
    case B_SETNANTOBAD: return bad_go<T, T, B_SETNANTOBAD>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
    case B_SETINFTOBAD: return bad_go<T, T, B_SETINFTOBAD>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
    case B_SETNONFINITETOBAD:
      return bad_go<T, T, B_SETNONFINITETOBAD>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
    case B_SETBADTONAN: return bad_go<T, T, B_SETBADTONAN>(p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
    default: break;
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_badop(int T, int op, const EwPlan &p, const void *a, const void *m, void *b, const BadInfo &ba,
                        int32_t mbv, const pdlhip_value &bbv, double dval, bool bad, int *flag, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return bad_T<CType<TC>::T>(op, p, a, m, b, ba, mbv, bbv, dval, bad, flag, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

}  // namespace pdlhip
