/*
 * pdlcpu.h -- the ORACLE: a plain-C restatement of PDL's CPU path for the
 * broadcast-loop hot path (broadcast engine, Ops binops, Ufunc reductions and
 * scans, Primitive matmult, xoshiro256+ inputs).
 *
 * This is TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker (or the
 * timed CPU baseline).  The product (pdl_amd/, include/pdlhip.h) never links
 * or calls it.
 *
 * Every function follows a reference file:line (paths relative to the PDL
 * source tree, PDL 2.106):
 *   type table / promotion      lib/PDL/Types.pm:27-255, lib/PDL/Core/pdlapi.c:1169-1327
 *   broadcast plan + errors     lib/PDL/Core/pdlbroadcast.c:211-523, pdlapi.c:829-869,
 *                               lib/PDL/Core/pdlutil.c:65-145
 *   broadcast loop              lib/PDL/Core/pdl.h.PL:640-675, pdlbroadcast.h:72-84
 *   pthread split               pdlbroadcast.c:95-209, 469-484
 *   dtype conversion            lib/PDL/Core/pdlconv.c:163-201
 *   binops                      lib/PDL/Ops.pd:66-69, 104-219, 288-324, 382-397
 *   reductions / scans          lib/PDL/Ufunc.pd:88-140, 413-613
 *   matmult                     lib/PDL/Primitive.pd:191-245
 *   RNG                         lib/PDL/Primitive-xoshiro256plus.c:28-43, 111-138
 *
 * Pinning status of each part is recorded in the header comment of pdlcpu.c.
 */
#ifndef PDLCPU_H
#define PDLCPU_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Type codes in PDL's promotion order (lib/PDL/Types.pm:27-255). */
enum {
  PDLC_SB = 0, PDLC_B, PDLC_S, PDLC_US, PDLC_L, PDLC_UL, PDLC_IND, PDLC_ULL,
  PDLC_LL, PDLC_F, PDLC_D, PDLC_LD, PDLC_CF, PDLC_CD, PDLC_CLD, PDLC_NTYPES
};

/* pdl state flags used by this path (lib/PDL/Core/pdl.h.PL:300-330). */
#define PDLC_NOMYDIMS 0x1   /* null ndarray: dims/type to be created by the op */
#define PDLC_BADVAL   0x2   /* bad-value state flag (PDL_BADVAL) */
#define PDLC_VAFFINE  0x4   /* this is a virtual-affine view (has vafftrans) */

#define PDLC_MAXDIMS 16
#define PDLC_MAXBIDS 8

typedef union {
  int8_t A; uint8_t B; int16_t S; uint16_t U; int32_t L; uint32_t K;
  int64_t N; uint64_t P; int64_t Q; float F; double D;
} pdlc_value;

typedef struct {
  int32_t type;      /* type code, or -1 for none */
  int32_t pad_;
  pdlc_value value;
} pdlc_anyval;

/* One ndarray as the broadcast engine sees it: dims, the composed
 * virtual-affine increments/offset into a root buffer (PDL_REPRINCS /
 * PDL_REPROFFS / PDL_REPRP, pdl.h.PL:511-517), bad-value state and value. */
typedef struct {
  int32_t datatype;
  int32_t state;
  int32_t ndims;
  int32_t has_badvalue;
  int64_t dims[PDLC_MAXDIMS];
  int64_t incs[PDLC_MAXDIMS];
  int64_t offs;
  void *data;
  pdlc_anyval badvalue;
  /* pdl->broadcastids / nbroadcastids (pdl.h.PL:569-570); 0 = one id, ndims */
  int32_t nbroadcastids;
  int32_t pad2_;
  int64_t broadcastids[PDLC_MAXBIDS];
} pdlc_pdl;

typedef struct {
  int32_t error;            /* 0 = none, 1 = user error, 2 = fatal */
  char message[4096];
} pdlc_error;

/* Run one op the way pdl_run_<op> does (PP.pm:1998-2014): null checks,
 * type coercion, redodims, output creation (calloc'd host memory, released
 * with pdlc_free), readdata.  otherpars: biops -> {swap}; others ignore it.
 * nthreads: the autopthread target (0 = off); the split rule of
 * pdl_autopthreadmagic with a threshold of `pthread_size_mi` Mi-elements. */
int pdlc_run(const char *op, pdlc_pdl **pdls, int npdls, const int64_t *otherpars,
             int nthreads, int pthread_size_mi, pdlc_error *err);

/* The broadcast plan of the next pdlc_run calls (after redodims, before the
 * loop), for tests that hand PDL's own plan to a readdata-level entry:
 * dims/incs (incs[d*npdls + p]) and offsets of the broadcast struct
 * (pdlbroadcast.h:18-61), the trans's core ind_sizes and inc_sizes
 * (pdlapi.c:846-864), the trans type and bvalflag.  pdlc_capture_plan(dst)
 * fills *dst on each later run until called with NULL. */
typedef struct {
  int npdls, ndims;
  int64_t dims[20];
  int64_t incs[20 * 6];
  int64_t offs[6];
  int64_t ind_sizes[3];
  int64_t inc_sizes[6][2];
  int datatype, bvalflag;
} pdlc_plan;
void pdlc_capture_plan(pdlc_plan *dst);
void pdlc_free(void *p);

/* pdl_find_max_pthread (pdlbroadcast.c:95-159) on plain dim lists, for the
 * t/pthread.t table. dims: nd dims of the single non-created pdl. */
void pdlc_find_max_pthread(const int64_t *dims, int nd, int realdims, int target,
                           int *max_pthread, int *max_dim);

/* Type names as PDL prints them (Types.pm ppforcetype). */
const char *pdlc_type_name(int t);
int pdlc_type_size(int t);
/* Default bad value of a type (Types.pm defbval), stored into *out. */
void pdlc_default_badval(int t, pdlc_anyval *out);

/* Element conversion as pdlconv.c:163-174 (unsigned targets via intmax_t). */
void pdlc_convert(const void *src, int srctype, void *dst, int dsttype, int64_t n);

/* xoshiro256+ / splitmix64 (Primitive-xoshiro256plus.c:28-43, 111-138). */
uint64_t pdlc_splitmix64_next(uint64_t *x);
uint64_t pdlc_xoshiro256plus_next(uint64_t *s);
void pdlc_srand(uint64_t *s, uint64_t seed, int nthreads);   /* s: 4*nthreads */
double pdlc_drand(uint64_t *s);
/* n draws of pdl_drand from one stream seeded with `seed` (PDL random with
 * autopthread off). */
void pdlc_random_fill(double *out, int64_t n, uint64_t seed);
void pdlc_random_fill_from(double *out, int64_t n, uint64_t seed, uint64_t skip);
void pdlc_rand_state_after(uint64_t *out, uint64_t seed, uint64_t skip);
double pdlc_ndtri(double y);
int pdlc_rangeb(pdlc_pdl *parent, const pdlc_pdl *ind, const int64_t *size, int64_t nsize, const char *boundary,
                pdlc_pdl *child, int back, pdlc_error *err);
void pdlc_ndtri_n(const double *y, double *x, int64_t n);

/* FP64 / extended-precision truth helpers (not PDL functions: the "PDL
 * computing in double" references the float results are scored against). */
void pdlc_truth_matmult_ld(const double *a, const double *b, double *c,
                           int64_t h, int64_t t, int64_t w);

#ifdef __cplusplus
}
#endif
#endif
