// Shared host/device definitions of the pdlhip library: op codes, the
// canonical launch plans the host planner hands to kernels, and type traits.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/pdlhip.h"

namespace pdlhip {

constexpr int MAXK = 8;      // loop dims of a canonical launch plan
constexpr int MAXP = 6;      // pdls per trans on this path

// Ops.pd biops (lib/PDL/Ops.pd:288-324) and assgn (382-397)
enum BinOp : int {
  OP_PLUS, OP_MULT, OP_MINUS, OP_DIVIDE, OP_GT, OP_LT, OP_LE, OP_GE, OP_EQ, OP_NE,
  OP_SHL, OP_SHR, OP_OR, OP_AND, OP_XOR, OP_POWER, OP_ATAN2, OP_MODULO, OP_SPACESHIP,
  OP_NBINOPS,
  // Ops.pd ufuncs (Ops.pd:222-265, 327-381): one input; the elementwise
  // kernels' second operand is ignored.  Not valid in fused chains.
  OP_BITNOT = 32, OP_SQRT, OP_SIN, OP_COS, OP_NOT, OP_EXP, OP_LOG, OP_LOG10, OP_ABS
};

// Bad.pd elementwise ops (Bad.pd:343-395, 584-840)
enum BadOp : int {
  B_ISBAD, B_ISGOOD, B_SETBADIF, B_SETVALTOBAD, B_SETNANTOBAD, B_SETINFTOBAD, B_SETNONFINITETOBAD,
  B_SETBADTONAN, B_SETBADTOVAL
};

// Ufunc.pd reductions / scans (lib/PDL/Ufunc.pd:88-140, 413-613)
enum RedOp : int {
  R_SUM, R_PROD, R_DSUM, R_DPROD, R_CUMUSUM, R_CUMUPROD, R_DCUMUSUM, R_DCUMUPROD,
  R_AVERAGE, R_DAVERAGE, R_MIN, R_MAX, R_MIN_IND, R_MAX_IND, R_MINMAXIMUM,
  R_MIN_N_IND, R_MAX_N_IND, R_NREDOPS
};

// Bad-value description of one operand as a kernel sees it
// ($ISBAD / $SETBAD of PDLCode.pm:652-693; PDL_ISBAD2 pdl.h.PL:259-262)
struct BadInfo {
  int32_t check;       // this operand's values are tested against its bad value
  int32_t isnan;       // the bad value is NaN: test with isnan
  pdlhip_value bv;     // the bad value in the operand's own type
};

// Elementwise launch plan: `nd` loop dims ordered by the output's strides,
// unit dims dropped, contiguous neighbours merged.
struct EwPlan {
  int32_t nd;
  int32_t pad_;
  int64_t n;                  // total elements
  int64_t dims[MAXK];
  int64_t inc[3][MAXK];       // operand 0,1 inputs; 2 output
  int64_t off[3];
};

// Reduction / scan launch plan: core dim n (reduced or scanned) plus the
// broadcast dims of the output(s).
struct RedPlan {
  int64_t n;                  // core dim size
  int64_t inc_n[2];           // a's and (scans) b's stride along n
  int32_t nd;                 // broadcast dims after canonicalisation
  int32_t npdl;
  int64_t nout;               // product of broadcast dims
  int64_t dims[MAXK];
  int64_t inc[MAXP][MAXK];    // per pdl broadcast strides
  int64_t off[MAXP];
  int64_t m;                  // *_n_ind: m_size; scans: b's n extent
  int64_t inc_m;              // *_n_ind: c's stride along m
  void *scratch;              // caller-provided scratch for the long-row split (null: from the cache)
  int64_t scratch_bytes;      // its size
};

// Matmult launch plan (Primitive.pd:211-245): a(t,h) b(w,t) c(w,h) with
// broadcast dims as a batch.
struct MmPlan {
  int64_t T, H, W;
  int64_t a_t, a_h, b_w, b_t, c_w, c_h;  // core strides (0 where size <= 1)
  int64_t atdi, btdi;                     // PDL_REPRINCS(a)[0], (b)[1]
  int32_t nd;
  int32_t variant;                        // kernel variant the autotuner picked (0: the default rule)
  int64_t nbatch;
  int64_t dims[MAXK];
  int64_t inc[3][MAXK];
  int64_t off[3];
};

// one product of a matmult list (pdlhip_run_matmult_list): resolved
// pointers, sizes, core strides and element offsets
struct MmItem {
  const void *a, *b;
  void *c;
  int64_t H, W, T;
  int64_t a_t, a_h, b_w, b_t, c_w, c_h;
  int64_t oa, ob, oc;
};

// ragged lists (k_ragged.hip): item table, segments and a fused program
constexpr int RAG_MAXIN = 6, RAG_MAXNODE = 12, RAG_SEG = 4096;
constexpr int RAG_NOREDUCE = -1;   // final op "none": the chain's values only
struct RagNode { int32_t op, x, y, pad; };
struct RagProg {
  int32_t nin, nnodes;
  int32_t type[RAG_MAXIN];
  RagNode node[RAG_MAXNODE];     // general form: operand k >= 0 slot k, k < 0 node -k-1
  // left-deep form (lin = 1): v = slot[first]; node k: v = op_k(v, slot[lslot[k]])
  // (operands swapped when lswap[k])
  int32_t lin, first;
  int32_t lslot[RAG_MAXNODE], lswap[RAG_MAXNODE];
  // nodes computing in different types (PDL's trans type of each op):
  // the general form, each node in ntype[k] with its operands converted
  int32_t mixed;
  int32_t ntype[RAG_MAXNODE];
  int32_t store;                 // write every chain value to the item's `val`
  int32_t stream;                // streaming loads/stores (the list is larger than the infinity cache)
  pdlhip_value out_bv;
};
struct RagItem {
  const char *ptr[RAG_MAXIN];   // element 0 of the item in each input slot
  int64_t inc[RAG_MAXIN];       // stride (0 for a 0-dim scalar slot)
  char *out;                    // reduction result (scans: the scanned vector)
  int64_t out_inc;
  char *val;                    // chain values (when stored), the chain's type
  int64_t val_inc;
  int64_t len, seg0, nseg;
};
struct RagSeg { int32_t item, single; int64_t start, end; };
struct RagLaunch {
  RagProg prog;
  const RagItem *items;
  // reductions: one wave per segment of <= RAG_SEG elements; items of several
  // segments are finished by a combine over `multi`
  const RagSeg *segs;
  int64_t nseg;
  const int32_t *multi;
  int64_t nmulti;
  void *scratch;                // 64 bytes per segment
  int32_t *empty;               // minimum/maximum: per item, 1 when no good value was found
  // scans: one block per item; block b runs items order[bstart[b] .. bstart[b+1])
  const int32_t *order;
  const int32_t *bstart;
  int32_t nblk;
};

// type-code -> C type
template <int TC> struct CType;
template <> struct CType<PDLHIP_SB> { typedef int8_t T; };
template <> struct CType<PDLHIP_B> { typedef uint8_t T; };
template <> struct CType<PDLHIP_S> { typedef int16_t T; };
template <> struct CType<PDLHIP_US> { typedef uint16_t T; };
template <> struct CType<PDLHIP_L> { typedef int32_t T; };
template <> struct CType<PDLHIP_UL> { typedef uint32_t T; };
template <> struct CType<PDLHIP_IND> { typedef int64_t T; };
template <> struct CType<PDLHIP_ULL> { typedef uint64_t T; };
template <> struct CType<PDLHIP_LL> { typedef int64_t T; };
template <> struct CType<PDLHIP_F> { typedef float T; };
template <> struct CType<PDLHIP_D> { typedef double T; };

inline int type_size(int t) {
  static const int s[] = {1, 1, 2, 2, 4, 4, 8, 8, 8, 4, 8, 16, 8, 16, 32};
  return (t >= 0 && t < PDLHIP_NTYPES) ? s[t] : 0;
}
inline bool type_on_device(int t) { return t >= PDLHIP_SB && t <= PDLHIP_D; }
// float and double complex: plus, minus, mult, eq, ne, assgn, conversion and
// matmult on the device (k_complex.hip); other ops refuse them
inline bool type_is_cplx(int t) { return t == PDLHIP_CF || t == PDLHIP_CD; }

// Streaming (nontemporal) loads and stores raise the rate of a pure stream by
// about a tenth on this part, but leave nothing in the 256 MiB infinity cache
// for the next op to reuse; kernels stream only data at least that large.
inline bool streaming_read(size_t bytes) { return bytes >= ((size_t)256 << 20); }

}  // namespace pdlhip
