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

// elementwise ops on inputs of other types, converted in registers
// (k_mixed.hip): a, b of types ta, tb; sbad_*: the source is bad-flagged
// (its bad elements become T's default bad value, as in a converted copy)
struct MixedArgs {
  const void *a, *b;
  int32_t ta, tb;
  int32_t sbad_a, sbad_b;
  BadInfo sba, sbb;
  int32_t a_scalar, b_scalar;   // every increment of the operand is 0: one element
};
#define X(TC) hipError_t launch_mixed_t##TC(int op, const MixedArgs &m, const EwPlan &p, void *c, bool bad, \
                                            const BadInfo *bi, hipStream_t s);
PDLHIP_FOR_EACH_TYPE(X)
#undef X
inline hipError_t launch_mixed(int T, int op, const MixedArgs &m, const EwPlan &p, void *c, bool bad,
                               const BadInfo *bi, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return launch_mixed_t##TC(op, m, p, c, bad, bi, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

// ragged lists (k_ragged.hip)
hipError_t launch_ragged_reduce(int T, int op, const RagLaunch &L, hipStream_t s);
hipError_t launch_ragged_scan(int T, int op, const RagLaunch &L, hipStream_t s);

// conversion / assignment between views (k_misc.hip): dst = (dtype) src with
// PDL's bad-value mapping (pdlconv.c:163-174); also assgn (Ops.pd:382-397)
hipError_t launch_cplx_binop(int T, int op, const EwPlan &p, const void *a, const void *b, void *c, hipStream_t s);
hipError_t launch_cplx_convert(int stype, int dtype, const EwPlan &p, const void *src, void *dst, hipStream_t s);
hipError_t launch_cplx_matmult(int T, const MmPlan &p, const void *a, const void *b, void *c, hipStream_t s);
hipError_t launch_cplx_reduce(int T, int op, const RedPlan &p, const void *a, void *b, hipStream_t s);
hipError_t launch_convert(int stype, int dtype, const EwPlan &p, const void *src, void *dst, bool bad,
                          const BadInfo &sbi, const pdlhip_value &dbv, hipStream_t s);
// PDL's random / randsym / grandom stream (k_rng.hip).  kind 0 random, 1
// randsym, 2 grandom; the output's logical element e (dim 0 fastest) sits at
// off + sum_d i_d * incs[d] (off + e when dense) and takes draw e
struct RngPlan {
  int32_t nd;
  int32_t dense;
  int64_t n;
  int64_t off;
  int64_t dims[PDLHIP_MAXDIMS];
  int64_t incs[PDLHIP_MAXDIMS];
};
void rng_seed(uint64_t seed, uint64_t s[4]);             // pdl_srand's first state
void rng_advance(uint64_t s[4], uint64_t n);             // s = the state n draws later
void rng_chunk_states(const uint64_t s[4], int64_t nchunks, uint64_t *out);  // every 4096 draws
const uint64_t *rng_jump_table(size_t *nwords);          // the in-chunk jumps, for the device
hipError_t launch_rng(int T, int kind, const RngPlan &p, void *out, const uint64_t *d_states, const uint64_t *d_jumps,
                      int64_t nchunks, int *d_flag, hipStream_t s);
constexpr int64_t RNG_CHUNK = 4096;

// Slices.pd index / index1d / index2d (k_index.hip): op 0/1/2, back = the
// BackCode (scatter into a).  Pointer slots: 0 a, 1 first index, 2 second
// index (index2d), 3 c; broadcast dims in PDL's loop order, m innermost.
struct IxPlan {
  int32_t op, back, nd, bad;
  int64_t npos;                         // broadcast positions
  int64_t m;                            // index1d: the m extent; 1 otherwise
  int64_t dims[PDLHIP_MAXDIMS + 2];
  int64_t inc[4][PDLHIP_MAXDIMS + 2];
  int64_t off[4];
  int64_t na, nb2;                      // a's core extents (n; na, nb)
  int64_t ia0, ia1;                     // a's core strides
  int64_t iim, icm;                     // index1d: ind's and c's strides along m
  int64_t ind_bv[2];                    // the indices' bad values
  int64_t a_lo, a_span;                 // BackCode: the owner table spans a's elements [a_lo, a_lo + a_span)
};
hipError_t launch_index(int T, const IxPlan &p, void *a, const int64_t *i1, const int64_t *i2, void *c,
                        const pdlhip_value &cbv, unsigned long long *err, long long *owner, hipStream_t s);

// Slices.pd rangeb (k_index.hip): one thread per child element; the
// child's dims are the index broadcast dims, the nonzero sizes, then the
// source's dims past rdim (RedoDims' order)
constexpr int RG_MAXR = 32;
struct RgPlan {
  int32_t rdim, pdim, itdim, ntsize, stdim, bad, back, cnd;
  int64_t nitems, nchild, ninr, nst;
  int64_t sizes[RG_MAXR];
  int32_t bnd[RG_MAXR];
  int64_t pdims[PDLHIP_MAXDIMS], pincs[PDLHIP_MAXDIMS], poff;
  int64_t iincs[PDLHIP_MAXDIMS + 1], ioff;       // the index: coordinate stride, then its broadcast dims
  int64_t cdims[PDLHIP_MAXDIMS], cincs[PDLHIP_MAXDIMS], coff;
  int64_t a_lo, a_span;                           // BackCode: the owner table spans the parent's elements
};
hipError_t launch_range(int T, const RgPlan &p, void *parent, const int64_t *ind, void *child,
                        const pdlhip_value &pbv, const pdlhip_value &cbv, int pbv_nan, int cbv_nan,
                        unsigned long long *err, long long *owner, hipStream_t s);

hipError_t launch_fill(int dtype, const EwPlan &p, void *dst, uint64_t seed, double a, double b, int normal,
                       hipStream_t s);

}  // namespace pdlhip
