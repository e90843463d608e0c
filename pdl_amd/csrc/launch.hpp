// Host-side launcher entry points of the HIP kernels (one translation unit
// per element type, so the type switch compiles in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace pdlhip {

#define PDLHIP_FOR_EACH_TYPE(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)

// elementwise binops over a canonical plan (k_elementwise.hip)
#define X(TC) hipError_t launch_binop_t##TC(int op, const EwPlan &p, const void *a, const void *b, void *c, \
                                            bool bad, const BadInfo *bi, hipStream_t s);
PDLHIP_FOR_EACH_TYPE(X)
#undef X

// reductions / scans (k_reduce.hip): a of type TC, outputs per op
#define X(TC) hipError_t launch_reduce_t##TC(int op, int otype, const RedPlan &p, const void *a, void *const *outs, \
                                             bool bad, const BadInfo &ba, const pdlhip_value *obv, \
                                             int *d_flag, hipStream_t s);
PDLHIP_FOR_EACH_TYPE(X)
#undef X

// matmult (k_matmult.hip)
#define X(TC) hipError_t launch_matmult_t##TC(const MmPlan &p, const void *a, const void *b, void *c, bool bad, \
                                              const BadInfo *bi, int mode, hipStream_t s);
PDLHIP_FOR_EACH_TYPE(X)
#undef X

// a list of products in one persistent work-queue launch, or one launch per
// product (each = true) (k_batch.hip)
hipError_t launch_matmult_list(int T, const MmItem *d_items, const MmItem *h_items, int nitems, const int3 *d_work,
                               unsigned int nwork, unsigned int *d_counter, int tile, bool fused, bool each,
                               int nblocks, hipStream_t s);

// Bad.pd elementwise ops (k_badops.hip): a of type T, mask int32 (setbadif),
// b of type T (int32 for isbad/isgood); dval the double OtherPar; flag set
// where an op found values to set bad (setnantobad and its like)
hipError_t launch_badop(int T, int op, const EwPlan &p, const void *a, const void *m, void *b, const BadInfo &ba,
                        int32_t mbv, const pdlhip_value &bbv, double dval, bool bad, int *flag, hipStream_t s);

// ragged lists (k_ragged.hip)
hipError_t launch_ragged_reduce(int T, int op, const RagLaunch &L, hipStream_t s);
hipError_t launch_ragged_scan(int T, int op, const RagLaunch &L, hipStream_t s);

// conversion / assignment between views (k_misc.hip): dst = (dtype) src with
// PDL's bad-value mapping (pdlconv.c:163-174); also assgn (Ops.pd:382-397)
hipError_t launch_convert(int stype, int dtype, const EwPlan &p, const void *src, void *dst, bool bad,
                          const BadInfo &sbi, const pdlhip_value &dbv, hipStream_t s);
hipError_t launch_fill(int dtype, const EwPlan &p, void *dst, uint64_t seed, double a, double b, int normal,
                       hipStream_t s);

}  // namespace pdlhip
