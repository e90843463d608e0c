/*
 * pdlcpu.c -- ORACLE (test infrastructure): plain-C restatement of PDL's CPU
 * path for the broadcast-loop hot path.  Built by oracle/Makefile with the
 * flags Perl's ccflags give PDL (-O2 -fwrapv -fno-strict-aliasing,
 * Makefile.PL:44-48), so integer arithmetic wraps exactly as in PDL.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library; the product never does.
 *
 * Pinning (what checks each part; see tests/test_oracle_*.py):
 *   type selection / promotion  t/ops.t:33,67,116-123 vectors + hand-derived
 *                               rules of pdlapi.c:1182-1236 (tests/golden)
 *   broadcast plan + errors     t/core.t:249-290 dim-compatibility table
 *                               (append rows restated with this engine's
 *                               generic signatures), t/ops.t:283 empty dims
 *   pthread split               t/pthread.t:12-33 table
 *   binops                      t/ops.t vectors (minus/swap, modulus tables,
 *                               longlong precision, bitwise, shifts), plus
 *                               numpy closed forms (IEEE single/double ops,
 *                               wrapping ints) on random inputs
 *   reductions / scans          t/ufunc.t:85-104,180-219 vectors, numpy
 *                               brute force (ints wrap, so any order pins)
 *   matmult                     t/primitive-matmult.t:21-99 vectors, numpy
 *                               sequential-order float32 products
 *   bad values                  t/bad.t:111-136,293-300, t/ufunc.t, matmult
 *                               BAD/NaN rows
 *   RNG                         oracle/_ref (the reference's own
 *                               Primitive-xoshiro256plus.c compiled here)
 *   power / atan2 (libm)        t/ops.t:33,126-127 vectors, plus
 *                               tests/test_oracle_libm.py: within 1 ulp of a
 *                               120-bit mpmath value on random operands in
 *                               every sign/quadrant, and C99 Annex F special
 *                               cases (zeros, infinities, NaN) bit-exact
 */
#include "pdlcpu.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <complex.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXP 6     /* max pdls per trans on this path (minmaximum has 5) */
#define MAXB 20    /* max broadcast dims */

/* ------------------------------------------------------------------ */
/* Type table (lib/PDL/Types.pm:27-255)                                */
/* ------------------------------------------------------------------ */
static const char *type_names[PDLC_NTYPES] = {
  "sbyte", "byte", "short", "ushort", "long", "ulong", "indx", "ulonglong",
  "longlong", "float", "double", "ldouble", "cfloat", "cdouble", "cldouble"};
static const int type_sizes[PDLC_NTYPES] = {1, 1, 2, 2, 4, 4, 8, 8, 8, 4, 8, 16, 8, 16, 32};

const char *pdlc_type_name(int t) { return (t >= 0 && t < PDLC_NTYPES) ? type_names[t] : "INVALID"; }
int pdlc_type_size(int t) { return (t >= 0 && t < PDLC_NTYPES) ? type_sizes[t] : 0; }

/* defbval column of Types.pm */
void pdlc_default_badval(int t, pdlc_anyval *out) {
  memset(out, 0, sizeof *out);
  out->type = t;
  switch (t) {
  case PDLC_SB: out->value.A = SCHAR_MIN; break;
  case PDLC_B: out->value.B = UCHAR_MAX; break;
  case PDLC_S: out->value.S = SHRT_MIN; break;
  case PDLC_US: out->value.U = USHRT_MAX; break;
  case PDLC_L: out->value.L = INT32_MIN; break;
  case PDLC_UL: out->value.K = UINT32_MAX; break;
  case PDLC_IND: out->value.N = PTRDIFF_MIN; break;
  case PDLC_ULL: out->value.P = UINT64_MAX; break;
  case PDLC_LL: out->value.Q = INT64_MIN; break;
  case PDLC_F: out->value.F = -FLT_MAX; break;
  case PDLC_D: out->value.D = -DBL_MAX; break;
  default: out->type = -1; break;
  }
}

static int type_supported(int t) { return t >= PDLC_SB && t <= PDLC_D; }
/* float and double complex (C99 _Complex, as GCC compiles PDL's PP code):
 * restated for plus, minus, mult, divide (libgcc's __divsc3/__divdc3, which
 * GCC calls for `/`), eq, ne, assgn, conversion, matmult and the sums and
 * products over a dim */
static int type_cplx(int t) { return t == PDLC_CF || t == PDLC_CD; }

/* X-macro over the types this restatement computes in:
 * X(code, ctype, ppsym, is_integer, is_unsigned) */
#define TYPES_REAL(X) \
  X(PDLC_SB, int8_t, A, 1, 0) X(PDLC_B, uint8_t, B, 1, 1) X(PDLC_S, int16_t, S, 1, 0) \
  X(PDLC_US, uint16_t, U, 1, 1) X(PDLC_L, int32_t, L, 1, 0) X(PDLC_UL, uint32_t, K, 1, 1) \
  X(PDLC_IND, int64_t, N, 1, 0) X(PDLC_ULL, uint64_t, P, 1, 1) X(PDLC_LL, int64_t, Q, 1, 0) \
  X(PDLC_F, float, F, 0, 0) X(PDLC_D, double, D, 0, 0)
/* second copy for nested expansion (a macro is not re-expanded inside itself) */
#define TYPES_REAL_(X) \
  X(PDLC_SB, int8_t, A, 1, 0) X(PDLC_B, uint8_t, B, 1, 1) X(PDLC_S, int16_t, S, 1, 0) \
  X(PDLC_US, uint16_t, U, 1, 1) X(PDLC_L, int32_t, L, 1, 0) X(PDLC_UL, uint32_t, K, 1, 1) \
  X(PDLC_IND, int64_t, N, 1, 0) X(PDLC_ULL, uint64_t, P, 1, 1) X(PDLC_LL, int64_t, Q, 1, 0) \
  X(PDLC_F, float, F, 0, 0) X(PDLC_D, double, D, 0, 0)

/* PDL_ISNAN_<ppsym> (pdl.h.PL:82-87) */
#define ISNAN_A(x) 0
#define ISNAN_B(x) 0
#define ISNAN_S(x) 0
#define ISNAN_U(x) 0
#define ISNAN_L(x) 0
#define ISNAN_K(x) 0
#define ISNAN_N(x) 0
#define ISNAN_P(x) 0
#define ISNAN_Q(x) 0
#define ISNAN_F(x) (isnan(x) ? 1 : 0)
#define ISNAN_D(x) (isnan(x) ? 1 : 0)

/* `int+` output ctype for a given trans type (PdlParObj.pm:149-159):
 * max(long, T) */
#define IPLUS_A int32_t
#define IPLUS_B int32_t
#define IPLUS_S int32_t
#define IPLUS_U int32_t
#define IPLUS_L int32_t
#define IPLUS_K uint32_t
#define IPLUS_N int64_t
#define IPLUS_P uint64_t
#define IPLUS_Q int64_t
#define IPLUS_F float
#define IPLUS_D double

/* PDL_ISBAD2 (pdl.h.PL:259-262) */
#define ISBAD2(v, bv, bvnan, ppsym) ((bvnan) ? ISNAN_##ppsym(v) : (v) == (bv))

static double anyval_as_double(const pdlc_anyval *v) {
  switch (v->type) {
#define X(code, ctype, ppsym, i, u) case code: return (double)v->value.ppsym;
    TYPES_REAL(X)
#undef X
  default: return 0;
  }
}

/* ------------------------------------------------------------------ */
/* Errors (lib/PDL/Core/pdlutil.c:27-145)                              */
/* ------------------------------------------------------------------ */
static int make_error(pdlc_error *e, int kind, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(e->message, sizeof e->message, fmt, ap);
  va_end(ap);
  e->error = kind;
  return kind;
}

/* ------------------------------------------------------------------ */
/* Op vtables: signature, types, flags (the subset of pdl_transvtable  */
/* this path uses, pdl.h.PL:381-441)                                   */
/* ------------------------------------------------------------------ */
#define P_ISOUT 0x1
#define P_ISCREAT 0x2
#define P_ISTYPED 0x4
#define P_ISTPLUS 0x8

struct trans;
typedef void (*readdata_fn)(struct trans *tr, int thr);

typedef struct vtable {
  const char *name;
  int npdls, nparents;
  const char *par_names[MAXP];
  int par_realdims[MAXP];
  int par_ind[MAXP][2];
  const char *ind_names[3];
  int par_flags[MAXP];
  int par_types[MAXP];
  const int *gentypes;           /* -1 terminated */
  readdata_fn readdata;
  int opcode;
  int badmode_biop;              /* 1: biop gating, 2: bifunc gating, 3: double OtherPar */
} vtable;

/* GenericTypes lists, in the order PP emits them (Ops.pd:1-14) */
static const int GT_ALL[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, -1};
static const int GT_REAL[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, -1};
static const int GT_INT[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, -1};
/* [@$C, @$F] with $F = F,LD,D (D moved last so non-floats default to D) */
static const int GT_POWER[] = {12, 13, 14, 9, 11, 10, -1};
static const int GT_ATAN2[] = {9, 11, 10, -1};
/* $AF of Ops.pd/Bad.pd: the non-integer types, D moved last */
static const int GT_AF[] = {9, 11, 12, 13, 14, 10, -1};

/* ------------------------------------------------------------------ */
/* Broadcast struct (pdlbroadcast.h:18-61), implicit broadcasting only */
/* ------------------------------------------------------------------ */
typedef struct {
  int npdls;
  int ndims, nimpl, nextra;
  int64_t dims[MAXB];
  int64_t incs[MAXB * MAXP];   /* incs[d*npdls + j] (PDL_BRC_INC) */
  int64_t offs[MAXP];
  int mag_nth, mag_nthpdl, mag_nthr;
  int64_t mag_stride, mag_skip;
} brc_t;

typedef struct trans {
  const vtable *vt;
  pdlc_pdl *pdls[MAXP];
  int datatype;
  int bvalflag;
  int64_t swap;
  double dpar;                 /* a double OtherPar (Bad.pd: value / newval) */
  brc_t brc;
  int64_t ind_sizes[3];
  int64_t inc_sizes[MAXP][2];
  pdlc_anyval badval[MAXP];    /* effective bad value per pdl */
  int badval_isnan[MAXP];
  volatile char setbad[MAXP];  /* $PDLSTATESETBAD requests */
  volatile char setgood[MAXP]; /* $PDLSTATESETGOOD requests */
  volatile int croaked;        /* $CROAK from the loop body (PP.pm:371) */
  char croak[512];
} trans;

/* per-thread dims/offsets (PDL_BRC_THR_OFFSET, pdlbroadcast.h:65-70;
 * split of pdlbroadcast.c:469-484). Returns 0 if the loop is empty. */
static int brc_thread_setup(const brc_t *b, int thr, int64_t *dims, int64_t *inds, int64_t *offs) {
  int d, j;
  for (d = 0; d < b->ndims; d++) { dims[d] = b->dims[d]; inds[d] = 0; }
  for (j = 0; j < b->npdls; j++) offs[j] = b->offs[j];
  if (b->mag_nthr > 0) {
    int64_t n1 = b->mag_stride + (b->mag_skip ? 1 : 0);
    dims[b->mag_nth] = (b->mag_skip && thr >= b->mag_skip) ? n1 - 1 : n1;
    int64_t off = (int64_t)thr * b->mag_stride + (thr < b->mag_skip ? thr : b->mag_skip);
    for (j = 0; j < b->npdls; j++) offs[j] += off * b->incs[b->mag_nth * b->npdls + j];
  }
  for (d = 0; d < b->ndims; d++) if (!dims[d]) return 0;   /* pdlbroadcast.c:560-561 */
  return 1;
}

/* pdl_broadcast_nd_step (pdlbroadcast.h:72-84) */
static inline int nd_step(int npdls, int64_t *offsp, int nth, int ndims, const int64_t *incs,
                          const int64_t *dims, int64_t *inds) {
  int i, j;
  for (i = nth; i < ndims; i++) {
    for (j = 0; j < npdls; j++) offsp[j] += incs[i * npdls + j];
    if (++inds[i] < dims[i]) return 1;
    inds[i] = 0;
    for (j = 0; j < npdls; j++) offsp[j] -= incs[i * npdls + j] * dims[i];
  }
  return 0;
}

/* Restatement of PDL_BROADCASTLOOP_START/END (pdl.h.PL:640-675): the two
 * innermost broadcast dims are nested loops with per-pdl increments in the
 * for-increment clause (so `continue` still advances), outer dims advance
 * by the odometer.  The body sees element offsets O0..O4. */
#define BLOOP(tr, thr, ...) do { \
  const brc_t *B_ = &(tr)->brc; const int NP_ = B_->npdls; \
  int64_t D_[MAXB], I_[MAXB], OF_[MAXP]; \
  if (!brc_thread_setup(B_, (thr), D_, I_, OF_)) break; \
  const int64_t i00_ = B_->incs[0], i01_ = B_->incs[NP_ + 0]; \
  const int64_t i10_ = NP_ > 1 ? B_->incs[1] : 0, i11_ = NP_ > 1 ? B_->incs[NP_ + 1] : 0; \
  const int64_t i20_ = NP_ > 2 ? B_->incs[2] : 0, i21_ = NP_ > 2 ? B_->incs[NP_ + 2] : 0; \
  const int64_t i30_ = NP_ > 3 ? B_->incs[3] : 0, i31_ = NP_ > 3 ? B_->incs[NP_ + 3] : 0; \
  const int64_t i40_ = NP_ > 4 ? B_->incs[4] : 0, i41_ = NP_ > 4 ? B_->incs[NP_ + 4] : 0; \
  do { \
    int64_t O0 = OF_[0], O1 = NP_ > 1 ? OF_[1] : 0, O2 = NP_ > 2 ? OF_[2] : 0; \
    int64_t O3 = NP_ > 3 ? OF_[3] : 0, O4 = NP_ > 4 ? OF_[4] : 0; \
    (void)O1; (void)O2; (void)O3; (void)O4; \
    const int64_t TD0_ = D_[0], TD1_ = D_[1]; \
    for (int64_t t1_ = 0; t1_ < TD1_; t1_++, \
         O0 += i01_ - i00_ * TD0_, O1 += i11_ - i10_ * TD0_, O2 += i21_ - i20_ * TD0_, \
         O3 += i31_ - i30_ * TD0_, O4 += i41_ - i40_ * TD0_) { \
      for (int64_t t0_ = 0; t0_ < TD0_; t0_++, O0 += i00_, O1 += i10_, O2 += i20_, \
           O3 += i30_, O4 += i40_) { \
        __VA_ARGS__ \
      } \
    } \
  } while (nd_step(NP_, OF_, 2, B_->ndims, B_->incs, D_, I_)); \
} while (0)

/* typed base pointers / bad values for pdl k */
#define DECL(k, ctype, ppsym, name) \
  ctype *name = (ctype *)tr->pdls[k]->data; \
  const ctype name##_bv = (ctype)tr->badval[k].value.ppsym; (void)name##_bv; \
  const int name##_bvnan = tr->badval_isnan[k]; (void)name##_bvnan;

/* ------------------------------------------------------------------ */
/* Ops.pd biop / bifunc / assgn                                        */
/* ------------------------------------------------------------------ */
enum {
  OP_PLUS, OP_MULT, OP_MINUS, OP_DIVIDE, OP_GT, OP_LT, OP_LE, OP_GE, OP_EQ, OP_NE,
  OP_SHL, OP_SHR, OP_OR, OP_AND, OP_XOR, OP_POWER, OP_ATAN2, OP_MODULO, OP_SPACESHIP,
  OP_ASSGN
};

/* Ops.pd:68-70 */
#define PDL_ABS(A) ((A) >= 0 ? (A) : -(A))
#define MOD(X, N) (((N) == 0) ? 0 : ((X) - (PDL_ABS(N)) * ((long long)((X) / (PDL_ABS(N))) + ((((N) * ((long long)((X) / (N)))) != (X)) ? ((((N) < 0) ? 1 : 0) + ((((X) < 0) ? -1 : 0))) : 0))))
#define BU_MOD(X, N) (((N) == 0) ? 0 : ((X) - (N) * ((uint64_t)((X) / (N)))))
#define SPACE(A, B) (((A) < (B)) ? -1 : ((A) != (B)))

/* tgmath dispatch of pow/atan2 for the float types */
static inline float tg_pow_f(float a, float b) { return powf(a, b); }
static inline double tg_pow_d(double a, double b) { return pow(a, b); }
static inline float tg_atan2_f(float a, float b) { return atan2f(a, b); }
static inline double tg_atan2_d(double a, double b) { return atan2(a, b); }

#define BIOP_LOOP(ctype, ppsym, EXPR) do { \
  DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, b) DECL(2, ctype, ppsym, c) \
  if (!tr->bvalflag) { \
    BLOOP(tr, thr, { ctype av = a[O0], bv = b[O1]; (void)av; (void)bv; c[O2] = (ctype)(EXPR); }); \
  } else { \
    const int sa = tr->pdls[0]->state & PDLC_BADVAL, sb = tr->pdls[1]->state & PDLC_BADVAL; \
    char anybad = 0; \
    BLOOP(tr, thr, { ctype av = a[O0], bv = b[O1]; \
      if ((sa && ISBAD2(av, a_bv, a_bvnan, ppsym)) || (sb && ISBAD2(bv, b_bv, b_bvnan, ppsym))) { \
        c[O2] = c_bv; anybad = 1; } else c[O2] = (ctype)(EXPR); }); \
    if (anybad) tr->setbad[2] = 1; \
  } \
} while (0)

#define BIFUNC_LOOP(ctype, ppsym, EXPR) do { \
  DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, b) DECL(2, ctype, ppsym, c) \
  if (!tr->bvalflag) { \
    BLOOP(tr, thr, { ctype av = a[O0], bv = b[O1]; c[O2] = (ctype)(EXPR); }); \
  } else { \
    char anybad = 0; \
    BLOOP(tr, thr, { ctype av = a[O0], bv = b[O1]; \
      if (ISBAD2(av, a_bv, a_bvnan, ppsym) || ISBAD2(bv, b_bv, b_bvnan, ppsym)) { \
        anybad = 1; c[O2] = c_bv; } else { c[O2] = (ctype)(EXPR); } }); \
    if (anybad) tr->setbad[2] = 1; \
  } \
} while (0)

#define ARITH_CASES(X) \
  X(PDLC_SB, int8_t, A) X(PDLC_B, uint8_t, B) X(PDLC_S, int16_t, S) X(PDLC_US, uint16_t, U) \
  X(PDLC_L, int32_t, L) X(PDLC_UL, uint32_t, K) X(PDLC_IND, int64_t, N) X(PDLC_ULL, uint64_t, P) \
  X(PDLC_LL, int64_t, Q) X(PDLC_F, float, F) X(PDLC_D, double, D)
#define INT_CASES(X) \
  X(PDLC_SB, int8_t, A) X(PDLC_B, uint8_t, B) X(PDLC_S, int16_t, S) X(PDLC_US, uint16_t, U) \
  X(PDLC_L, int32_t, L) X(PDLC_UL, uint32_t, K) X(PDLC_IND, int64_t, N) X(PDLC_ULL, uint64_t, P) \
  X(PDLC_LL, int64_t, Q)
#define SIGNED_AND_FLOAT_CASES(X) \
  X(PDLC_SB, int8_t, A) X(PDLC_S, int16_t, S) X(PDLC_L, int32_t, L) X(PDLC_IND, int64_t, N) \
  X(PDLC_LL, int64_t, Q) X(PDLC_F, float, F) X(PDLC_D, double, D)
#define UNSIGNED_CASES(X) \
  X(PDLC_B, uint8_t, B) X(PDLC_US, uint16_t, U) X(PDLC_UL, uint32_t, K) X(PDLC_ULL, uint64_t, P)

#define CPLX_BIOP_LOOP(ctype, EXPR) do { \
  ctype *a = tr->pdls[0]->data, *b = tr->pdls[1]->data, *c = tr->pdls[2]->data; \
  BLOOP(tr, thr, { ctype av = a[O0], bv = b[O1]; c[O2] = (ctype)(EXPR); }); \
} while (0)

static void rd_biop(trans *tr, int thr) {
  int op = tr->vt->opcode;
  if (type_cplx(tr->datatype)) {
#define CPLX_CASE(OPC, EXPR) \
    if (op == OPC) { \
      if (tr->datatype == PDLC_CF) { CPLX_BIOP_LOOP(float complex, EXPR); } \
      else { CPLX_BIOP_LOOP(double complex, EXPR); } \
      return; \
    }
    CPLX_CASE(OP_PLUS, av + bv)
    CPLX_CASE(OP_MULT, av * bv)
    CPLX_CASE(OP_MINUS, av - bv)
    CPLX_CASE(OP_DIVIDE, av / bv)
    CPLX_CASE(OP_EQ, av == bv)
    CPLX_CASE(OP_NE, av != bv)
#undef CPLX_CASE
    return;
  }
#define CASE_ARITH(OPC, EXPR) \
  if (op == OPC) { switch (tr->datatype) { \
    case PDLC_SB: BIOP_LOOP(int8_t, A, EXPR); break; case PDLC_B: BIOP_LOOP(uint8_t, B, EXPR); break; \
    case PDLC_S: BIOP_LOOP(int16_t, S, EXPR); break; case PDLC_US: BIOP_LOOP(uint16_t, U, EXPR); break; \
    case PDLC_L: BIOP_LOOP(int32_t, L, EXPR); break; case PDLC_UL: BIOP_LOOP(uint32_t, K, EXPR); break; \
    case PDLC_IND: BIOP_LOOP(int64_t, N, EXPR); break; case PDLC_ULL: BIOP_LOOP(uint64_t, P, EXPR); break; \
    case PDLC_LL: BIOP_LOOP(int64_t, Q, EXPR); break; case PDLC_F: BIOP_LOOP(float, F, EXPR); break; \
    case PDLC_D: BIOP_LOOP(double, D, EXPR); break; default: break; } return; }
#define CASE_INT(OPC, EXPR) \
  if (op == OPC) { switch (tr->datatype) { \
    case PDLC_SB: BIOP_LOOP(int8_t, A, EXPR); break; case PDLC_B: BIOP_LOOP(uint8_t, B, EXPR); break; \
    case PDLC_S: BIOP_LOOP(int16_t, S, EXPR); break; case PDLC_US: BIOP_LOOP(uint16_t, U, EXPR); break; \
    case PDLC_L: BIOP_LOOP(int32_t, L, EXPR); break; case PDLC_UL: BIOP_LOOP(uint32_t, K, EXPR); break; \
    case PDLC_IND: BIOP_LOOP(int64_t, N, EXPR); break; case PDLC_ULL: BIOP_LOOP(uint64_t, P, EXPR); break; \
    case PDLC_LL: BIOP_LOOP(int64_t, Q, EXPR); break; default: break; } return; }
  CASE_ARITH(OP_PLUS, av + bv)
  CASE_ARITH(OP_MULT, av * bv)
  CASE_ARITH(OP_MINUS, av - bv)
  CASE_ARITH(OP_DIVIDE, av / bv)
  CASE_ARITH(OP_GT, av > bv)
  CASE_ARITH(OP_LT, av < bv)
  CASE_ARITH(OP_LE, av <= bv)
  CASE_ARITH(OP_GE, av >= bv)
  CASE_ARITH(OP_EQ, av == bv)
  CASE_ARITH(OP_NE, av != bv)
  CASE_INT(OP_SHL, av << bv)
  CASE_INT(OP_SHR, av >> bv)
  CASE_INT(OP_OR, av | bv)
  CASE_INT(OP_AND, av & bv)
  CASE_INT(OP_XOR, av ^ bv)
#undef CASE_ARITH
#undef CASE_INT
}

static void rd_bifunc(trans *tr, int thr) {
  int op = tr->vt->opcode;
  if (op == OP_POWER) {
    if (tr->datatype == PDLC_F) BIFUNC_LOOP(float, F, tg_pow_f(av, bv));
    else if (tr->datatype == PDLC_D) BIFUNC_LOOP(double, D, tg_pow_d(av, bv));
    return;
  }
  if (op == OP_ATAN2) {
    if (tr->datatype == PDLC_F) BIFUNC_LOOP(float, F, tg_atan2_f(av, bv));
    else if (tr->datatype == PDLC_D) BIFUNC_LOOP(double, D, tg_atan2_d(av, bv));
    return;
  }
  if (op == OP_MODULO) {
    switch (tr->datatype) {
    case PDLC_SB: BIFUNC_LOOP(int8_t, A, MOD(av, bv)); break;
    case PDLC_B: BIFUNC_LOOP(uint8_t, B, BU_MOD(av, bv)); break;
    case PDLC_S: BIFUNC_LOOP(int16_t, S, MOD(av, bv)); break;
    case PDLC_US: BIFUNC_LOOP(uint16_t, U, BU_MOD(av, bv)); break;
    case PDLC_L: BIFUNC_LOOP(int32_t, L, MOD(av, bv)); break;
    case PDLC_UL: BIFUNC_LOOP(uint32_t, K, BU_MOD(av, bv)); break;
    case PDLC_IND: BIFUNC_LOOP(int64_t, N, MOD(av, bv)); break;
    case PDLC_ULL: BIFUNC_LOOP(uint64_t, P, BU_MOD(av, bv)); break;
    case PDLC_LL: BIFUNC_LOOP(int64_t, Q, MOD(av, bv)); break;
    case PDLC_F: BIFUNC_LOOP(float, F, MOD(av, bv)); break;
    case PDLC_D: BIFUNC_LOOP(double, D, MOD(av, bv)); break;
    default: break;
    }
    return;
  }
  if (op == OP_SPACESHIP) {
    switch (tr->datatype) {
    case PDLC_SB: BIFUNC_LOOP(int8_t, A, SPACE(av, bv)); break;
    case PDLC_B: BIFUNC_LOOP(uint8_t, B, SPACE(av, bv)); break;
    case PDLC_S: BIFUNC_LOOP(int16_t, S, SPACE(av, bv)); break;
    case PDLC_US: BIFUNC_LOOP(uint16_t, U, SPACE(av, bv)); break;
    case PDLC_L: BIFUNC_LOOP(int32_t, L, SPACE(av, bv)); break;
    case PDLC_UL: BIFUNC_LOOP(uint32_t, K, SPACE(av, bv)); break;
    case PDLC_IND: BIFUNC_LOOP(int64_t, N, SPACE(av, bv)); break;
    case PDLC_ULL: BIFUNC_LOOP(uint64_t, P, SPACE(av, bv)); break;
    case PDLC_LL: BIFUNC_LOOP(int64_t, Q, SPACE(av, bv)); break;
    case PDLC_F: BIFUNC_LOOP(float, F, SPACE(av, bv)); break;
    case PDLC_D: BIFUNC_LOOP(double, D, SPACE(av, bv)); break;
    default: break;
    }
  }
}

/* Ops.pd:382-397 */
#define ASSGN_LOOP(ctype, ppsym) do { \
  DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, b) \
  if (!tr->bvalflag) { BLOOP(tr, thr, { b[O1] = a[O0]; }); } \
  else { char anybad = 0; \
    BLOOP(tr, thr, { if (ISBAD2(a[O0], a_bv, a_bvnan, ppsym)) { anybad = 1; b[O1] = b_bv; continue; } \
                     b[O1] = a[O0]; }); \
    if (anybad) tr->setbad[1] = 1; } \
} while (0)

static void rd_assgn(trans *tr, int thr) {
  if (tr->datatype == PDLC_CF || tr->datatype == PDLC_CD) {
    const size_t sz = type_sizes[tr->datatype];
    char *a = tr->pdls[0]->data, *b = tr->pdls[1]->data;
    BLOOP(tr, thr, { memcpy(b + O1 * sz, a + O0 * sz, sz); });
    return;
  }
  switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: ASSGN_LOOP(ctype, ppsym); break;
    ARITH_CASES(X)
#undef X
  default: break;
  }
}

/* ------------------------------------------------------------------ */
/* Ops.pd ufunc (Ops.pd:222-265, 327-333, 356-381): b() =              */
/* ($GENERIC(b))func($a()) through <tgmath.h> (integer arguments take  */
/* the double function), bad mode: ISBAD(a) -> SETBAD(b).  _rabs       */
/* (Ops.pd:343-358): unsigned copy, PDL_ABS otherwise.                 */
/* ------------------------------------------------------------------ */
enum { U_BITNOT = 100, U_SQRT, U_SIN, U_COS, U_NOT, U_EXP, U_LOG, U_LOG10, U_ABS };

#define TG1(fn, x) _Generic((x), float: fn##f, default: fn)(x)

#define OPSUFUNC_LOOP(ctype, ppsym, EXPR) do { \
  DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, b) \
  if (!tr->bvalflag) { \
    BLOOP(tr, thr, { ctype av = a[O0]; (void)av; b[O1] = (ctype)(EXPR); }); \
  } else { \
    BLOOP(tr, thr, { ctype av = a[O0]; (void)av; \
      if (ISBAD2(av, a_bv, a_bvnan, ppsym)) b[O1] = b_bv; else b[O1] = (ctype)(EXPR); }); \
  } \
} while (0)

static void rd_opsufunc(trans *tr, int thr) {
  const int op = tr->vt->opcode;
#define X(code, ctype, ppsym) case code: \
    if (op == U_SQRT) OPSUFUNC_LOOP(ctype, ppsym, TG1(sqrt, av)); \
    else if (op == U_SIN) OPSUFUNC_LOOP(ctype, ppsym, TG1(sin, av)); \
    else if (op == U_COS) OPSUFUNC_LOOP(ctype, ppsym, TG1(cos, av)); \
    else if (op == U_NOT) OPSUFUNC_LOOP(ctype, ppsym, !av); \
    else if (op == U_EXP) OPSUFUNC_LOOP(ctype, ppsym, TG1(exp, av)); \
    else if (op == U_LOG) OPSUFUNC_LOOP(ctype, ppsym, TG1(log, av)); \
    else if (op == U_LOG10) OPSUFUNC_LOOP(ctype, ppsym, TG1(log10, av)); \
    break;
  switch (tr->datatype) {
    ARITH_CASES(X)
  default: break;
  }
#undef X
  if (op == U_BITNOT) {
    switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: OPSUFUNC_LOOP(ctype, ppsym, ~av); break;
      INT_CASES(X)
#undef X
    default: break;
    }
  }
  if (op == U_ABS) {
    switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: OPSUFUNC_LOOP(ctype, ppsym, av); break;
      UNSIGNED_CASES(X)
#undef X
#define X(code, ctype, ppsym) case code: OPSUFUNC_LOOP(ctype, ppsym, PDL_ABS(av)); break;
      SIGNED_AND_FLOAT_CASES(X)
#undef X
    default: break;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Bad.pd elementwise ops.  $ISBAD compiles to PDL_ISBAD2 in good and  */
/* bad code alike (PP/PDLCode.pm:664-668); only PDL_IF_BAD switches.   */
/*   isbad / isgood      Bad.pd:343-395: b = IF_BAD(ISBAD(a), 0) / IF_BAD(ISGOOD(a), 1) */
/*   setbadif            Bad.pd:584-634                                 */
/*   setvaltobad         Bad.pd:639-676                                 */
/*   setnantobad etc.    Bad.pd:679-775                                 */
/*   setbadtonan         Bad.pd:777-806                                 */
/*   setbadtoval         Bad.pd:808-840                                 */
/* ------------------------------------------------------------------ */
enum { B_ISBAD = 200, B_ISGOOD, B_SETBADIF, B_SETVALTOBAD, B_SETNANTOBAD, B_SETINFTOBAD, B_SETNONFINITETOBAD,
       B_SETBADTONAN, B_SETBADTOVAL };

static void rd_badop(trans *tr, int thr) {
  const int op = tr->vt->opcode;
  const int bad = tr->bvalflag;
  if (op == B_ISBAD || op == B_ISGOOD) {
    int32_t *b = (int32_t *)tr->pdls[1]->data;
    switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: { DECL(0, ctype, ppsym, a) \
      BLOOP(tr, thr, { int isb = ISBAD2(a[O0], a_bv, a_bvnan, ppsym); \
        b[O1] = op == B_ISBAD ? (bad ? isb : 0) : (bad ? !isb : 1); }); break; }
      ARITH_CASES(X)
#undef X
    default: break;
    }
    return;
  }
  if (op == B_SETBADIF) {
    int32_t *m = (int32_t *)tr->pdls[1]->data;
    int32_t m_bv;
    memcpy(&m_bv, &tr->badval[1].value, sizeof m_bv);
    switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: { DECL(0, ctype, ppsym, a) DECL(2, ctype, ppsym, b) (void)a_bv; (void)a_bvnan; \
      BLOOP(tr, thr, { if ((bad && m[O1] == m_bv) || m[O1]) b[O2] = b_bv; else b[O2] = a[O0]; }); break; }
      ARITH_CASES(X)
#undef X
    default: break;
    }
    tr->setbad[2] = 1;
    return;
  }
  char flag = 0;
  switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: { DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, b) \
    (void)a_bvnan; (void)b_bv; \
    const ctype val = (ctype)tr->dpar; (void)val; \
    BLOOP(tr, thr, { ctype av = a[O0]; \
      switch (op) { \
      case B_SETVALTOBAD: if (av == val) b[O1] = b_bv; else b[O1] = av; break; \
      case B_SETNANTOBAD: if (ISNAN_##ppsym(av)) { b[O1] = b_bv; flag = 1; } else b[O1] = av; break; \
      case B_SETINFTOBAD: if (!isfinite((double)av) && !ISNAN_##ppsym(av)) { b[O1] = b_bv; flag = 1; } \
                          else b[O1] = av; break; \
      case B_SETNONFINITETOBAD: if (!isfinite((double)av)) { b[O1] = b_bv; flag = 1; } else b[O1] = av; break; \
      case B_SETBADTONAN: b[O1] = ISBAD2(av, a_bv, a_bvnan, ppsym) ? (ctype)NAN : av; break; \
      case B_SETBADTOVAL: b[O1] = (bad && ISBAD2(av, a_bv, a_bvnan, ppsym)) ? val : av; break; \
      default: break; } }); \
    break; }
    ARITH_CASES(X)
#undef X
  default: break;
  }
  if (op == B_SETVALTOBAD || flag) tr->setbad[1] = 1;
  if (op == B_SETBADTONAN || op == B_SETBADTOVAL) tr->setgood[1] = 1;
}

/* ------------------------------------------------------------------ */
/* Ufunc.pd reductions and scans                                       */
/* ------------------------------------------------------------------ */
enum { R_SUMOVER, R_PRODOVER, R_DSUMOVER, R_DPRODOVER, R_CUMUSUMOVER, R_CUMUPRODOVER,
       R_DCUMUSUMOVER, R_DCUMUPRODOVER, R_AVERAGE, R_DAVERAGE, R_MINIMUM, R_MAXIMUM,
       R_MINIMUM_IND, R_MAXIMUM_IND, R_MINMAXIMUM };

/* Ufunc.pd:96-118: tmp in the output type, sequential, prodover stops at a
 * zero running product. */
#define OVER_LOOP(ctype, ppsym, otype, INIT, OPEQ, LEFTZERO) do { \
  DECL(0, ctype, ppsym, a) \
  otype *b = (otype *)tr->pdls[1]->data; \
  otype b_bv; memcpy(&b_bv, &tr->badval[1].value, sizeof b_bv); \
  const int64_t N_ = tr->ind_sizes[0], ia = tr->inc_sizes[0][0]; \
  if (!tr->bvalflag) { \
    BLOOP(tr, thr, { otype tmp = INIT; const ctype *ap = a + O0; \
      for (int64_t n = 0; n < N_; n++) { tmp OPEQ ap[n * ia]; if (LEFTZERO) break; } \
      b[O1] = tmp; }); \
  } else { \
    BLOOP(tr, thr, { otype tmp = INIT; int flag = 0; const ctype *ap = a + O0; \
      for (int64_t n = 0; n < N_; n++) { ctype v = ap[n * ia]; \
        if (ISBAD2(v, a_bv, a_bvnan, ppsym)) continue; flag = 1; \
        tmp OPEQ v; if (LEFTZERO) break; } \
      if (!flag) b[O1] = b_bv; else b[O1] = tmp; }); \
  } \
} while (0)

/* Ufunc.pd:120-140 */
#define CUMU_LOOP(ctype, ppsym, otype, INIT, OPEQ) do { \
  DECL(0, ctype, ppsym, a) \
  otype *b = (otype *)tr->pdls[1]->data; \
  otype b_bv; memcpy(&b_bv, &tr->badval[1].value, sizeof b_bv); \
  const int64_t N_ = tr->ind_sizes[0], ia = tr->inc_sizes[0][0], ib = tr->inc_sizes[1][0]; \
  if (!tr->bvalflag) { \
    BLOOP(tr, thr, { otype tmp = INIT; const ctype *ap = a + O0; otype *bp = b + O1; \
      for (int64_t n = 0; n < N_; n++) { tmp OPEQ ap[n * ia]; bp[n * ib] = tmp; } }); \
  } else { \
    BLOOP(tr, thr, { otype tmp = INIT; const ctype *ap = a + O0; otype *bp = b + O1; \
      for (int64_t n = 0; n < N_; n++) { ctype v = ap[n * ia]; \
        if (ISBAD2(v, a_bv, a_bvnan, ppsym)) { bp[n * ib] = b_bv; continue; } \
        tmp OPEQ v; bp[n * ib] = tmp; } }); \
  } \
} while (0)

/* Ufunc.pd:413-444: integer average truncates (PDL_Indx cnt), empty ->
 * 0 for integer generic types, NaN for float ones; BAD in bad mode. */
#define AVG_LOOP(ctype, ppsym, otype, IS_INT) do { \
  DECL(0, ctype, ppsym, a) \
  otype *b = (otype *)tr->pdls[1]->data; \
  otype b_bv; memcpy(&b_bv, &tr->badval[1].value, sizeof b_bv); \
  const int64_t N_ = tr->ind_sizes[0], ia = tr->inc_sizes[0][0]; \
  const int bad = tr->bvalflag; \
  BLOOP(tr, thr, { otype tmp = 0; int64_t cnt = 0; const ctype *ap = a + O0; \
    for (int64_t n = 0; n < N_; n++) { ctype v = ap[n * ia]; \
      if (bad && ISBAD2(v, a_bv, a_bvnan, ppsym)) continue; cnt++; tmp += v; } \
    if (!cnt) { if (bad) b[O1] = b_bv; else b[O1] = (otype)((IS_INT) ? 0 : NAN); } \
    else b[O1] = tmp / cnt; }); \
} while (0)

/* Ufunc.pd:446-482: a NaN never displaces a non-NaN, a non-NaN displaces
 * a NaN; ties keep the first. */
#define MINMAX_LOOP(ctype, ppsym, OPC, IND) do { \
  DECL(0, ctype, ppsym, a) \
  const int64_t N_ = tr->ind_sizes[0], ia = tr->inc_sizes[0][0]; \
  const int bad = tr->bvalflag; \
  if (!(IND)) { DECL(1, ctype, ppsym, c) \
    BLOOP(tr, thr, { ctype cur = 0; int flag = 0; const ctype *ap = a + O0; \
      for (int64_t n = 0; n < N_; n++) { ctype v = ap[n * ia]; \
        if (bad && ISBAD2(v, a_bv, a_bvnan, ppsym)) continue; \
        if (flag && !(v OPC cur) && !ISNAN_##ppsym(cur)) continue; \
        cur = v; flag = 1; } \
      if (flag) c[O1] = cur; else { c[O1] = c_bv; tr->setbad[1] = 1; } }); \
  } else { int64_t *c = (int64_t *)tr->pdls[1]->data; const int64_t c_bv = tr->badval[1].value.N; \
    BLOOP(tr, thr, { ctype cur = 0; int64_t curind = -1; const ctype *ap = a + O0; \
      for (int64_t n = 0; n < N_; n++) { ctype v = ap[n * ia]; \
        if (bad && ISBAD2(v, a_bv, a_bvnan, ppsym)) continue; \
        if (curind != -1 && !(v OPC cur) && !ISNAN_##ppsym(cur)) continue; \
        cur = v; curind = n; } \
      if (curind != -1) c[O1] = curind; else { c[O1] = c_bv; tr->setbad[1] = 1; } }); \
  } \
} while (0)

/* Ufunc.pd:563-613: NaNs skipped entirely */
#define MINMAXIMUM_LOOP(ctype, ppsym) do { \
  DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, cmin) DECL(2, ctype, ppsym, cmax) \
  int64_t *cmin_ind = (int64_t *)tr->pdls[3]->data, *cmax_ind = (int64_t *)tr->pdls[4]->data; \
  const int64_t N_ = tr->ind_sizes[0], ia = tr->inc_sizes[0][0]; \
  const int bad = tr->bvalflag; \
  BLOOP(tr, thr, { ctype curmin = 0, curmax = 0; int64_t mi = 0, xi = 0; int flag = 0; \
    const ctype *ap = a + O0; \
    for (int64_t n = 0; n < N_; n++) { ctype v = ap[n * ia]; \
      if (bad && ISBAD2(v, a_bv, a_bvnan, ppsym)) continue; \
      if (ISNAN_##ppsym(v)) continue; \
      if (!flag) { curmin = curmax = v; mi = xi = n; flag = 1; } \
      else { if (v < curmin) { curmin = v; mi = n; } if (v > curmax) { curmax = v; xi = n; } } } \
    if (!flag) { cmin[O1] = cmin_bv; cmax[O2] = cmax_bv; cmin_ind[O3] = tr->badval[3].value.N; \
      cmax_ind[O4] = tr->badval[4].value.N; \
      tr->setbad[1] = tr->setbad[2] = tr->setbad[3] = tr->setbad[4] = 1; } \
    else { cmin[O1] = curmin; cmax[O2] = curmax; cmin_ind[O3] = mi; cmax_ind[O4] = xi; } }); \
} while (0)

/* complex sums and products (int+ of a complex type is itself): C99
 * complex += and *= in PDL's loop, prodover's stop at a zero product */
#define CPLX_OVER_LOOP(ctype, INIT, OPEQ, LEFTZERO, CUMU) do { \
  ctype *a = tr->pdls[0]->data, *b = tr->pdls[1]->data; \
  const int64_t N_ = tr->ind_sizes[0], ia = tr->inc_sizes[0][0], ib = CUMU ? tr->inc_sizes[1][0] : 0; \
  BLOOP(tr, thr, { ctype tmp = INIT; const ctype *ap = a + O0; ctype *bp = b + O1; \
    for (int64_t n = 0; n < N_; n++) { tmp OPEQ ap[n * ia]; if (CUMU) bp[n * ib] = tmp; if (LEFTZERO) break; } \
    if (!CUMU) *bp = tmp; }); \
} while (0)

static void rd_ufunc(trans *tr, int thr) {
  int op = tr->vt->opcode;
  int dt = tr->datatype;
  if (dt == PDLC_CF || dt == PDLC_CD) {
#define CPLX_RED(ctype) \
    switch (op) { \
    case R_SUMOVER: CPLX_OVER_LOOP(ctype, 0, +=, 0, 0); break; \
    case R_PRODOVER: CPLX_OVER_LOOP(ctype, 1, *=, tmp == 0, 0); break; \
    case R_CUMUSUMOVER: CPLX_OVER_LOOP(ctype, 0, +=, 0, 1); break; \
    case R_CUMUPRODOVER: CPLX_OVER_LOOP(ctype, 1, *=, 0, 1); break; \
    default: break; }
    if (dt == PDLC_CF) { CPLX_RED(float complex) } else { CPLX_RED(double complex) }
#undef CPLX_RED
    return;
  }
  switch (op) {
  case R_SUMOVER: case R_PRODOVER:
    switch (dt) {
#define X(code, ctype, ppsym) case code: \
      if (op == R_SUMOVER) { OVER_LOOP(ctype, ppsym, IPLUS_##ppsym, 0, +=, 0); } \
      else { OVER_LOOP(ctype, ppsym, IPLUS_##ppsym, 1, *=, tmp == 0); } break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  case R_DSUMOVER: case R_DPRODOVER:
    switch (dt) {
#define X(code, ctype, ppsym) case code: \
      if (op == R_DSUMOVER) { OVER_LOOP(ctype, ppsym, double, 0, +=, 0); } \
      else { OVER_LOOP(ctype, ppsym, double, 1, *=, tmp == 0); } break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  case R_CUMUSUMOVER: case R_CUMUPRODOVER:
    switch (dt) {
#define X(code, ctype, ppsym) case code: \
      if (op == R_CUMUSUMOVER) { CUMU_LOOP(ctype, ppsym, IPLUS_##ppsym, 0, +=); } \
      else { CUMU_LOOP(ctype, ppsym, IPLUS_##ppsym, 1, *=); } break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  case R_DCUMUSUMOVER: case R_DCUMUPRODOVER:
    switch (dt) {
#define X(code, ctype, ppsym) case code: \
      if (op == R_DCUMUSUMOVER) { CUMU_LOOP(ctype, ppsym, double, 0, +=); } \
      else { CUMU_LOOP(ctype, ppsym, double, 1, *=); } break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  case R_AVERAGE:
    switch (dt) {
#define X(code, ctype, ppsym) case code: AVG_LOOP(ctype, ppsym, IPLUS_##ppsym, code < PDLC_F); break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  case R_DAVERAGE:
    switch (dt) {
#define X(code, ctype, ppsym) case code: AVG_LOOP(ctype, ppsym, double, code < PDLC_F); break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  case R_MINIMUM: case R_MAXIMUM: case R_MINIMUM_IND: case R_MAXIMUM_IND: {
    int ind = (op == R_MINIMUM_IND || op == R_MAXIMUM_IND);
    int ismin = (op == R_MINIMUM || op == R_MINIMUM_IND);
    switch (dt) {
#define X(code, ctype, ppsym) case code: \
      if (ismin) { MINMAX_LOOP(ctype, ppsym, <, ind); } else { MINMAX_LOOP(ctype, ppsym, >, ind); } break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  }
  case R_MINMAXIMUM:
    switch (dt) {
#define X(code, ctype, ppsym) case code: MINMAXIMUM_LOOP(ctype, ppsym); break;
      ARITH_CASES(X)
#undef X
    default: break;
    }
    break;
  }
}

/* ------------------------------------------------------------------ */
/* Primitive.pd:211-245 matmult: tsiz = 64/sizeof(T) tiles; the t-run   */
/* accumulates sequentially in T from c's running value.                */
/* ------------------------------------------------------------------ */
#define MATMULT_LOOP(ctype, ppsym) do { \
  DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, b) DECL(2, ctype, ppsym, c) \
  const int64_t tsiz = 8 * sizeof(double) / sizeof(ctype); \
  const int64_t atdi = tr->pdls[0]->incs[0]; \
  const int64_t btdi = tr->pdls[1]->incs[1]; \
  const int64_t T_ = tr->ind_sizes[0], H_ = tr->ind_sizes[1], W_ = tr->ind_sizes[2]; \
  const int64_t a_t = tr->inc_sizes[0][0], a_h = tr->inc_sizes[0][1]; \
  const int64_t b_w = tr->inc_sizes[1][0], b_t = tr->inc_sizes[1][1]; \
  const int64_t c_w = tr->inc_sizes[2][0], c_h = tr->inc_sizes[2][1]; \
  const int bad = tr->bvalflag; \
  BLOOP(tr, thr, { \
    for (int64_t ho = 0; ho < H_; ho += tsiz) for (int64_t wo = 0; wo < W_; wo += tsiz) { \
      int64_t he = ho + tsiz < H_ ? ho + tsiz : H_, we = wo + tsiz < W_ ? wo + tsiz : W_; \
      for (int64_t h = ho; h < he; h++) for (int64_t w = wo; w < we; w++) \
        c[O2 + w * c_w + h * c_h] = 0; \
      for (int64_t to = 0; to < T_; to += tsiz) { \
        int64_t te = to + tsiz < T_ ? to + tsiz : T_; \
        for (int64_t h = ho; h < he; h++) for (int64_t w = wo; w < we; w++) { \
          ctype *cp = c + O2 + w * c_w + h * c_h; \
          ctype cc = *cp; \
          if (bad && ISBAD2(cc, c_bv, c_bvnan, ppsym)) continue; \
          const ctype *ad = a + O0 + to * a_t + h * a_h; \
          const ctype *bd = b + O1 + w * b_w + to * b_t; \
          char c_isbad = 0; \
          for (int64_t t = to; t < te; t++) { \
            if (bad && (ISBAD2(*ad, a_bv, a_bvnan, ppsym) || ISBAD2(*bd, b_bv, b_bvnan, ppsym))) { \
              c_isbad = 1; break; } \
            cc += *ad * *bd; ad += atdi; bd += btdi; } \
          if (c_isbad) { *cp = c_bv; continue; } \
          *cp = cc; \
        } \
      } \
    } }); \
} while (0)

#define CPLX_MATMULT_LOOP(ctype) do { \
  ctype *a = tr->pdls[0]->data, *b = tr->pdls[1]->data, *c = tr->pdls[2]->data; \
  const int64_t tsiz = 8 * sizeof(double) / sizeof(ctype); \
  const int64_t atdi = tr->pdls[0]->incs[0]; \
  const int64_t btdi = tr->pdls[1]->incs[1]; \
  const int64_t T_ = tr->ind_sizes[0], H_ = tr->ind_sizes[1], W_ = tr->ind_sizes[2]; \
  const int64_t a_t = tr->inc_sizes[0][0], a_h = tr->inc_sizes[0][1]; \
  const int64_t b_w = tr->inc_sizes[1][0], b_t = tr->inc_sizes[1][1]; \
  const int64_t c_w = tr->inc_sizes[2][0], c_h = tr->inc_sizes[2][1]; \
  (void)a_t; (void)b_t; \
  BLOOP(tr, thr, { \
    for (int64_t ho = 0; ho < H_; ho += tsiz) for (int64_t wo = 0; wo < W_; wo += tsiz) { \
      int64_t he = ho + tsiz < H_ ? ho + tsiz : H_, we = wo + tsiz < W_ ? wo + tsiz : W_; \
      for (int64_t h = ho; h < he; h++) for (int64_t w = wo; w < we; w++) \
        c[O2 + w * c_w + h * c_h] = 0; \
      for (int64_t to = 0; to < T_; to += tsiz) { \
        int64_t te = to + tsiz < T_ ? to + tsiz : T_; \
        for (int64_t h = ho; h < he; h++) for (int64_t w = wo; w < we; w++) { \
          ctype *cp = c + O2 + w * c_w + h * c_h; \
          ctype cc = *cp; \
          const ctype *ad = a + O0 + to * a_t + h * a_h; \
          const ctype *bd = b + O1 + w * b_w + to * b_t; \
          for (int64_t t = to; t < te; t++) { cc += *ad * *bd; ad += atdi; bd += btdi; } \
          *cp = cc; \
        } \
      } \
    } }); \
} while (0)

static void rd_matmult(trans *tr, int thr) {
  if (tr->datatype == PDLC_CF) { CPLX_MATMULT_LOOP(float complex); return; }
  if (tr->datatype == PDLC_CD) { CPLX_MATMULT_LOOP(double complex); return; }
  switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: MATMULT_LOOP(ctype, ppsym); break;
    ARITH_CASES(X)
#undef X
  default: break;
  }
}

/* Ufunc.pd:484-540 minimum_n_ind / maximum_n_ind: the first m_size
 * extreme indices, each search skipping indices already chosen. */
#define NIND_LOOP(ctype, ppsym, OPC) do { \
  DECL(0, ctype, ppsym, a) \
  int64_t *c = (int64_t *)tr->pdls[1]->data; const int64_t c_bv = tr->badval[1].value.N; \
  const int64_t ns = tr->ind_sizes[0], ms = tr->ind_sizes[1]; \
  const int64_t ia = tr->inc_sizes[0][0], ic = tr->inc_sizes[1][0]; \
  const int bad = tr->bvalflag; \
  BLOOP(tr, thr, { ctype cur = 0; int64_t curind; \
    tr->setbad[1] = 0; tr->setgood[1] = 1; \
    for (int64_t m = 0; m < ms; m++) { curind = ns; \
      for (int64_t n = 0; n < ns; n++) { int flag = 0; \
        for (int64_t mm = 0; mm < m; mm++) if (c[O1 + mm * ic] == n) { flag = 1; break; } \
        ctype v = a[O0 + n * ia]; \
        if (!flag && (!bad || !ISBAD2(v, a_bv, a_bvnan, ppsym)) && \
            ((curind == ns) || v OPC cur || ISNAN_##ppsym(cur))) { cur = v; curind = n; } } \
      if (curind != ns) c[O1 + m * ic] = curind; \
      else { c[O1 + m * ic] = c_bv; tr->setbad[1] = 1; } } }); \
} while (0)

static void rd_nind(trans *tr, int thr) {
  int ismin = tr->vt->opcode == 0;
  switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: \
    if (ismin) { NIND_LOOP(ctype, ppsym, <); } else { NIND_LOOP(ctype, ppsym, >); } break;
    ARITH_CASES(X)
#undef X
  default: break;
  }
}

/* Primitive.pd:1351-1374 append: used to pin the broadcast engine with the
 * reference's own dim-compatibility table (t/core.t:249-290). */
#define APPEND_LOOP(ctype, ppsym) do { \
  DECL(0, ctype, ppsym, a) DECL(1, ctype, ppsym, b) DECL(2, ctype, ppsym, c) \
  const int64_t ns = tr->ind_sizes[0], nmn = tr->ind_sizes[2]; \
  const int64_t ia = tr->inc_sizes[0][0], ib = tr->inc_sizes[1][0], ic = tr->inc_sizes[2][0]; \
  BLOOP(tr, thr, { for (int64_t n = 0; n < ns; n++) c[O2 + n * ic] = a[O0 + n * ia]; \
    for (int64_t mn = ns; mn < nmn; mn++) c[O2 + mn * ic] = b[O1 + (mn - ns) * ib]; }); \
} while (0)

static void rd_append(trans *tr, int thr) {
  switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: APPEND_LOOP(ctype, ppsym); break;
    ARITH_CASES(X)
#undef X
  default: break;
  }
}

/* ------------------------------------------------------------------ */
/* Slices.pd index / index1d / index2d (Slices.pd:142-220): Code       */
/* gathers c() = a(n => ind), BackCode (the dataflow writeback,        */
/* pdl.h.PL writebackdata) scatters a(n => ind) = c().  An index out of */
/* range, or bad in bad mode, $CROAKs "Error in <op>:invalid index ..."  */
/* except index1d, which sets c bad for a bad index (and skips it when  */
/* writing back).  Opcodes: 0 index, 1 index1d, 2 index2d; +10 BackCode */
/* ------------------------------------------------------------------ */
enum { X_INDEX = 0, X_INDEX1D = 1, X_INDEX2D = 2, X_BACK = 10 };

#define CROAK_AT(...) do { \
  if (!tr->croaked) { tr->croaked = 1; snprintf(tr->croak, sizeof tr->croak, __VA_ARGS__); } \
  goto croak_out; } while (0)

#define INDEX_LOOP(ctype, ppsym, BACK) do { \
  DECL(0, ctype, ppsym, a) DECL(2, ctype, ppsym, c) \
  const int64_t *ind = (const int64_t *)tr->pdls[1]->data; const int64_t ind_bv = tr->badval[1].value.N; \
  const int64_t ns = tr->ind_sizes[0], ia = tr->inc_sizes[0][0]; const int bad = tr->bvalflag; \
  BLOOP(tr, thr, { const int64_t this_ind = ind[O1]; \
    if ((bad && this_ind == ind_bv) || this_ind < 0 || this_ind >= ns) \
      CROAK_AT("Error in %s:invalid index %lld (valid range 0..%lld)", name, (long long)this_ind, (long long)(ns - 1)); \
    if (BACK) a[O0 + this_ind * ia] = c[O2]; else c[O2] = a[O0 + this_ind * ia]; }); \
} while (0)

#define INDEX1D_LOOP(ctype, ppsym, BACK) do { \
  DECL(0, ctype, ppsym, a) DECL(2, ctype, ppsym, c) \
  const int64_t *ind = (const int64_t *)tr->pdls[1]->data; const int64_t ind_bv = tr->badval[1].value.N; \
  const int64_t ns = tr->ind_sizes[0], ms = tr->ind_sizes[1], ia = tr->inc_sizes[0][0]; \
  const int64_t ii = tr->inc_sizes[1][0], ic = tr->inc_sizes[2][0]; const int bad = tr->bvalflag; \
  BLOOP(tr, thr, { for (int64_t m = 0; m < ms; m++) { const int64_t this_ind = ind[O1 + m * ii]; \
    if (bad && this_ind == ind_bv) { if (!(BACK)) c[O2 + m * ic] = c_bv; continue; } \
    if (this_ind < 0 || this_ind >= ns) \
      CROAK_AT("Error in %s:invalid index %lld at pos %lld (valid range 0..%lld)", name, (long long)this_ind, \
               (long long)m, (long long)(ns - 1)); \
    if (BACK) a[O0 + this_ind * ia] = c[O2 + m * ic]; else c[O2 + m * ic] = a[O0 + this_ind * ia]; } }); \
} while (0)

#define INDEX2D_LOOP(ctype, ppsym, BACK) do { \
  DECL(0, ctype, ppsym, a) DECL(3, ctype, ppsym, c) \
  const int64_t *inda = (const int64_t *)tr->pdls[1]->data, *indb = (const int64_t *)tr->pdls[2]->data; \
  const int64_t bva = tr->badval[1].value.N, bvb = tr->badval[2].value.N; \
  const int64_t na = tr->ind_sizes[0], nb = tr->ind_sizes[1]; \
  const int64_t ia = tr->inc_sizes[0][0], ib = tr->inc_sizes[0][1]; const int bad = tr->bvalflag; \
  BLOOP(tr, thr, { const int64_t xa = inda[O1], xb = indb[O2]; \
    if ((bad && xa == bva) || xa < 0 || xa >= na) \
      CROAK_AT("Error in %s:invalid x-index %lld (valid range 0..%lld)", name, (long long)xa, (long long)(na - 1)); \
    if ((bad && xb == bvb) || xb < 0 || xb >= nb) \
      CROAK_AT("Error in %s:invalid y-index %lld (valid range 0..%lld)", name, (long long)xb, (long long)(nb - 1)); \
    if (BACK) a[O0 + xa * ia + xb * ib] = c[O3]; else c[O3] = a[O0 + xa * ia + xb * ib]; }); \
} while (0)

static void rd_index(trans *tr, int thr) {
  const int op = tr->vt->opcode % X_BACK, back = tr->vt->opcode >= X_BACK;
  const char *name = op == X_INDEX ? "index" : op == X_INDEX1D ? "index1d" : "index2d";
  switch (tr->datatype) {
#define X(code, ctype, ppsym) case code: \
    if (op == X_INDEX) { if (back) INDEX_LOOP(ctype, ppsym, 1); else INDEX_LOOP(ctype, ppsym, 0); } \
    else if (op == X_INDEX1D) { if (back) INDEX1D_LOOP(ctype, ppsym, 1); else INDEX1D_LOOP(ctype, ppsym, 0); } \
    else { if (back) INDEX2D_LOOP(ctype, ppsym, 1); else INDEX2D_LOOP(ctype, ppsym, 0); } \
    break;
    ARITH_CASES(X)
#undef X
  default: break;
  }
croak_out:
  return;
}

/* ------------------------------------------------------------------ */
/* The vtables                                                         */
/* ------------------------------------------------------------------ */
#define BIOP_VT(nm, opc, gt, rd, mode) \
  {nm, 3, 2, {"a", "b", "c"}, {0, 0, 0}, {{0}}, {0}, {0, 0, P_ISOUT | P_ISCREAT}, {-1, -1, -1}, gt, rd, opc, mode}
#define UFUNC_VT(nm, opc, gt) \
  {nm, 2, 1, {"a", "b"}, {0, 0}, {{0}}, {0}, {0, P_ISOUT | P_ISCREAT}, {-1, -1}, gt, rd_opsufunc, opc, 0}
/* mode 3: takes a double OtherPar */
#define BADOP_VT(nm, opc, gt, mode) \
  {nm, 2, 1, {"a", "b"}, {0, 0}, {{0}}, {0}, {0, P_ISOUT | P_ISCREAT}, {-1, -1}, gt, rd_badop, opc, mode}
#define RED_VT(nm, opc, oflags, otype) \
  {nm, 2, 1, {"a", "b"}, {1, 0}, {{0}, {0}}, {"n"}, {0, P_ISOUT | P_ISCREAT | (oflags)}, {-1, otype}, NULL, rd_ufunc, opc, 0}
#define RED_VT_C(nm, opc, oflags, otype) \
  {nm, 2, 1, {"a", "c"}, {1, 0}, {{0}, {0}}, {"n"}, {0, P_ISOUT | P_ISCREAT | (oflags)}, {-1, otype}, NULL, rd_ufunc, opc, 0}
#define SCAN_VT(nm, opc, oflags, otype) \
  {nm, 2, 1, {"a", "b"}, {1, 1}, {{0}, {0}}, {"n"}, {0, P_ISOUT | P_ISCREAT | (oflags)}, {-1, otype}, NULL, rd_ufunc, opc, 0}

static vtable vtables[] = {
  BIOP_VT("plus", OP_PLUS, GT_ALL, rd_biop, 1),
  BIOP_VT("mult", OP_MULT, GT_ALL, rd_biop, 1),
  BIOP_VT("minus", OP_MINUS, GT_ALL, rd_biop, 1),
  BIOP_VT("divide", OP_DIVIDE, GT_ALL, rd_biop, 1),
  BIOP_VT("gt", OP_GT, GT_REAL, rd_biop, 1),
  BIOP_VT("lt", OP_LT, GT_REAL, rd_biop, 1),
  BIOP_VT("le", OP_LE, GT_REAL, rd_biop, 1),
  BIOP_VT("ge", OP_GE, GT_REAL, rd_biop, 1),
  BIOP_VT("eq", OP_EQ, GT_ALL, rd_biop, 1),
  BIOP_VT("ne", OP_NE, GT_ALL, rd_biop, 1),
  BIOP_VT("shiftleft", OP_SHL, GT_INT, rd_biop, 1),
  BIOP_VT("shiftright", OP_SHR, GT_INT, rd_biop, 1),
  BIOP_VT("or2", OP_OR, GT_INT, rd_biop, 1),
  BIOP_VT("and2", OP_AND, GT_INT, rd_biop, 1),
  BIOP_VT("xor", OP_XOR, GT_INT, rd_biop, 1),
  BIOP_VT("power", OP_POWER, GT_POWER, rd_bifunc, 2),
  BIOP_VT("atan2", OP_ATAN2, GT_ATAN2, rd_bifunc, 2),
  BIOP_VT("modulo", OP_MODULO, GT_REAL, rd_bifunc, 2),
  BIOP_VT("spaceship", OP_SPACESHIP, GT_REAL, rd_bifunc, 2),
  {"assgn", 2, 1, {"a", "b"}, {0, 0}, {{0}}, {0}, {0, P_ISOUT | P_ISCREAT}, {-1, -1}, GT_ALL, rd_assgn, OP_ASSGN, 0},
  UFUNC_VT("bitnot", U_BITNOT, GT_INT),
  UFUNC_VT("sqrt", U_SQRT, GT_ALL),
  UFUNC_VT("sin", U_SIN, GT_ALL),
  UFUNC_VT("cos", U_COS, GT_ALL),
  UFUNC_VT("not", U_NOT, GT_REAL),
  UFUNC_VT("exp", U_EXP, GT_POWER),
  UFUNC_VT("log", U_LOG, GT_POWER),
  UFUNC_VT("log10", U_LOG10, GT_ALL),
  UFUNC_VT("abs", U_ABS, GT_REAL),
  {"isbad", 2, 1, {"a", "b"}, {0, 0}, {{0}}, {0}, {0, P_ISOUT | P_ISCREAT | P_ISTYPED}, {-1, PDLC_L}, GT_ALL,
   rd_badop, B_ISBAD, 0},
  {"isgood", 2, 1, {"a", "b"}, {0, 0}, {{0}}, {0}, {0, P_ISOUT | P_ISCREAT | P_ISTYPED}, {-1, PDLC_L}, GT_ALL,
   rd_badop, B_ISGOOD, 0},
  {"setbadif", 3, 2, {"a", "mask", "b"}, {0, 0, 0}, {{0}}, {0}, {0, P_ISTYPED, P_ISOUT | P_ISCREAT}, {-1, PDLC_L, -1},
   GT_ALL, rd_badop, B_SETBADIF, 0},
  BADOP_VT("setvaltobad", B_SETVALTOBAD, GT_ALL, 3),
  BADOP_VT("setnantobad", B_SETNANTOBAD, GT_AF, 0),
  BADOP_VT("setinftobad", B_SETINFTOBAD, GT_AF, 0),
  BADOP_VT("setnonfinitetobad", B_SETNONFINITETOBAD, GT_AF, 0),
  BADOP_VT("setbadtonan", B_SETBADTONAN, GT_AF, 0),
  BADOP_VT("setbadtoval", B_SETBADTOVAL, GT_ALL, 3),
  RED_VT("sumover", R_SUMOVER, P_ISTYPED | P_ISTPLUS, PDLC_L),
  RED_VT("prodover", R_PRODOVER, P_ISTYPED | P_ISTPLUS, PDLC_L),
  RED_VT("dsumover", R_DSUMOVER, P_ISTYPED, PDLC_D),
  RED_VT("dprodover", R_DPRODOVER, P_ISTYPED, PDLC_D),
  SCAN_VT("cumusumover", R_CUMUSUMOVER, P_ISTYPED | P_ISTPLUS, PDLC_L),
  SCAN_VT("cumuprodover", R_CUMUPRODOVER, P_ISTYPED | P_ISTPLUS, PDLC_L),
  SCAN_VT("dcumusumover", R_DCUMUSUMOVER, P_ISTYPED, PDLC_D),
  SCAN_VT("dcumuprodover", R_DCUMUPRODOVER, P_ISTYPED, PDLC_D),
  RED_VT("average", R_AVERAGE, P_ISTYPED | P_ISTPLUS, PDLC_L),
  RED_VT("daverage", R_DAVERAGE, P_ISTYPED, PDLC_D),
  /* min/max name their output c (Ufunc.pd:451,468) */
  RED_VT_C("minimum", R_MINIMUM, 0, -1),
  RED_VT_C("maximum", R_MAXIMUM, 0, -1),
  RED_VT_C("minimum_ind", R_MINIMUM_IND, P_ISTYPED, PDLC_IND),
  RED_VT_C("maximum_ind", R_MAXIMUM_IND, P_ISTYPED, PDLC_IND),
  {"minmaximum", 5, 1, {"a", "cmin", "cmax", "cmin_ind", "cmax_ind"}, {1, 0, 0, 0, 0}, {{0}}, {"n"},
   {0, P_ISOUT | P_ISCREAT, P_ISOUT | P_ISCREAT, P_ISOUT | P_ISCREAT | P_ISTYPED, P_ISOUT | P_ISCREAT | P_ISTYPED},
   {-1, -1, -1, PDLC_IND, PDLC_IND}, NULL, rd_ufunc, R_MINMAXIMUM, 0},
  {"matmult", 3, 2, {"a", "b", "c"}, {2, 2, 2}, {{0, 1}, {2, 0}, {2, 1}}, {"t", "h", "w"},
   {0, 0, P_ISOUT | P_ISCREAT}, {-1, -1, -1}, GT_ALL, rd_matmult, 0, 0},
  {"minimum_n_ind", 2, 1, {"a", "c"}, {1, 1}, {{0}, {1}}, {"n", "m"}, {0, P_ISOUT | P_ISCREAT | P_ISTYPED},
   {-1, PDLC_IND}, GT_REAL, rd_nind, 0, 0},
  {"maximum_n_ind", 2, 1, {"a", "c"}, {1, 1}, {{0}, {1}}, {"n", "m"}, {0, P_ISOUT | P_ISCREAT | P_ISTYPED},
   {-1, PDLC_IND}, GT_REAL, rd_nind, 1, 0},
  {"append", 3, 2, {"a", "b", "c"}, {1, 1, 1}, {{0}, {1}, {2}}, {"n", "m", "mn"},
   {0, 0, P_ISOUT | P_ISCREAT}, {-1, -1, -1}, GT_ALL, rd_append, 0, 0},
  /* Slices.pd:149-220 (GenericTypes ppdefs_all; ind typed indx) */
  {"index", 3, 2, {"a", "ind", "c"}, {1, 0, 0}, {{0}}, {"n"}, {0, P_ISTYPED, P_ISOUT | P_ISCREAT},
   {-1, PDLC_IND, -1}, GT_ALL, rd_index, X_INDEX, 0},
  {"index1d", 3, 2, {"a", "ind", "c"}, {1, 1, 1}, {{0}, {1}, {1}}, {"n", "m"}, {0, P_ISTYPED, P_ISOUT | P_ISCREAT},
   {-1, PDLC_IND, -1}, GT_ALL, rd_index, X_INDEX1D, 0},
  {"index2d", 4, 3, {"a", "inda", "indb", "c"}, {2, 0, 0, 0}, {{0, 1}}, {"na", "nb"},
   {0, P_ISTYPED, P_ISTYPED, P_ISOUT | P_ISCREAT}, {-1, PDLC_IND, PDLC_IND, -1}, GT_ALL, rd_index, X_INDEX2D, 0},
  /* their BackCode: every ndarray exists, a is written */
  {"index_back", 3, 3, {"a", "ind", "c"}, {1, 0, 0}, {{0}}, {"n"}, {0, P_ISTYPED, 0},
   {-1, PDLC_IND, -1}, GT_ALL, rd_index, X_INDEX + X_BACK, 0},
  {"index1d_back", 3, 3, {"a", "ind", "c"}, {1, 1, 1}, {{0}, {1}, {1}}, {"n", "m"}, {0, P_ISTYPED, 0},
   {-1, PDLC_IND, -1}, GT_ALL, rd_index, X_INDEX1D + X_BACK, 0},
  {"index2d_back", 4, 4, {"a", "inda", "indb", "c"}, {2, 0, 0, 0}, {{0, 1}}, {"na", "nb"},
   {0, P_ISTYPED, P_ISTYPED, 0}, {-1, PDLC_IND, PDLC_IND, -1}, GT_ALL, rd_index, X_INDEX2D + X_BACK, 0},
};

static void init_vtables(void) {
  static int done = 0;
  if (done) return;
  for (size_t i = 0; i < sizeof vtables / sizeof *vtables; i++)
    if (!vtables[i].gentypes) vtables[i].gentypes = GT_REAL;
  /* sumover/prodover/cumu*: $A; d* and average/min/max: default reals */
  for (size_t i = 0; i < sizeof vtables / sizeof *vtables; i++) {
    int op = vtables[i].opcode;
    if (vtables[i].readdata == rd_ufunc &&
        (op == R_SUMOVER || op == R_PRODOVER || op == R_CUMUSUMOVER || op == R_CUMUPRODOVER))
      vtables[i].gentypes = GT_ALL;
  }
  done = 1;
}

static const vtable *find_vtable(const char *name) {
  init_vtables();
  for (size_t i = 0; i < sizeof vtables / sizeof *vtables; i++)
    if (!strcmp(vtables[i].name, name)) return &vtables[i];
  return NULL;
}

/* ------------------------------------------------------------------ */
/* Type selection (pdlapi.c:1169-1236)                                 */
/* ------------------------------------------------------------------ */
static int type_adjust_from_trans(int ttype, int flags, int par_type) {
  if (flags & P_ISTYPED) return !(flags & P_ISTPLUS) ? par_type : (par_type > ttype ? par_type : ttype);
  return ttype;
}

static int transtype_select(const vtable *vt, pdlc_pdl **pdls) {
  int avail[PDLC_NTYPES] = {0}, last = -1, i, retval = -1;
  for (i = 0; vt->gentypes[i] >= 0; i++) avail[last = vt->gentypes[i]] = 1;
  if (vt->gentypes[0] == last) return last;
  int use_last = 0;
  for (i = vt->npdls - 1; i >= 0; i--) {
    pdlc_pdl *p = pdls[i];
    if (!(p->state & PDLC_NOMYDIMS) && !(vt->par_flags[i] & P_ISTYPED)) {
      int nt = p->datatype;
      int is_avail = (nt >= 0 && nt < PDLC_NTYPES) ? avail[nt] : 0;
      if (!is_avail && nt > last) use_last = 1;
      if (nt >= 0 && is_avail && retval < nt) retval = nt;
    }
    if (i == vt->nparents && retval >= 0) return retval;
  }
  if (use_last || retval < 0 || !avail[retval]) retval = last;
  return retval;
}

/* ------------------------------------------------------------------ */
/* pdl_param_error (pdlutil.c:65-145), exact format                    */
/* ------------------------------------------------------------------ */
static int param_error(pdlc_error *e, const vtable *vt, int pidx, pdlc_pdl **pdls, int nimpl,
                       const int *creating, const char *fmt, ...) {
  char msg[4096] = {0};
  size_t len = 0;
#define APP(...) do { if (len < sizeof msg) len += snprintf(msg + len, sizeof msg - len, __VA_ARGS__); } while (0)
  int i, j, maxrealdims;
  if (pidx < 0 || pidx >= vt->npdls) APP("ERROR: UNKNOWN PARAMETER");
  else {
    APP("%s(", vt->name);
    for (i = 0; i < vt->npdls; i++) { APP("%s", vt->par_names[i]); if (i < vt->npdls - 1) APP(","); }
    APP("): parameter '%s': ", vt->par_names[pidx]);
  }
  va_list ap;
  va_start(ap, fmt);
  if (len < sizeof msg) len += vsnprintf(msg + len, sizeof msg - len, fmt, ap);
  va_end(ap);
  APP("\nThere are %d PDLs in the expression; %d broadcast dim%s.\n", vt->npdls, nimpl, nimpl == 1 ? "" : "s");
  for (i = maxrealdims = 0; i < vt->npdls; i++)
    if (vt->par_realdims[i] > maxrealdims) maxrealdims = vt->par_realdims[i];
  APP("   PDL IN EXPR.    ");
  if (maxrealdims > 0) {
    char format[80];
    snprintf(format, sizeof format - 1, "%%%ds", 8 * maxrealdims + 3);
    APP(format, "ACTIVE DIMS | ");
  }
  APP("BROADCAST DIMS\n");
  for (i = 0; i < vt->npdls; i++) {
    APP("   #%3d (%s", i, creating[i] ? "null)\n" : "normal): ");
    if (creating[i]) continue;
    if (maxrealdims == 1) APP("    ");
    for (j = 0; j < maxrealdims - vt->par_realdims[i]; j++) APP("%8s", " ");
    for (j = 0; j < vt->par_realdims[i]; j++) APP("%8lld", (long long)pdls[i]->dims[j]);
    if (maxrealdims) APP(" | ");
    for (j = 0; j < nimpl && j + vt->par_realdims[i] < pdls[i]->ndims; j++)
      APP("%8lld", (long long)pdls[i]->dims[j + vt->par_realdims[i]]);
    APP("\n");
  }
#undef APP
  e->error = 1;
  memcpy(e->message, msg, sizeof msg);
  e->message[sizeof e->message - 1] = 0;
  return 1;
}

/* ------------------------------------------------------------------ */
/* pthread split choice (pdlbroadcast.c:95-159)                        */
/* ------------------------------------------------------------------ */
static void find_max_pthread(pdlc_pdl **pdls, int npdls, const int *realdims, const int *creating,
                             int target, int *p_max, int *p_dim, int *p_pdl) {
  int64_t max_remainder = 0;
  *p_max = 0;
  for (int j = 0; j < npdls; j++) {
    if (creating[j]) continue;
    for (int t = realdims[j]; t < pdls[j]->ndims; t++) {
      int64_t this_dim = pdls[j]->dims[t];
      int64_t rem = this_dim % target;
      if (rem == 0) { *p_max = target; *p_pdl = j; *p_dim = t; break; }
      if (this_dim > *p_max && rem > max_remainder) {
        max_remainder = rem;
        *p_max = (int)(target < this_dim ? target : this_dim);
        *p_pdl = j; *p_dim = t;
      }
    }
    if (*p_max == target) break;
  }
}

void pdlc_find_max_pthread(const int64_t *dims, int nd, int realdims, int target, int *max_pthread, int *max_dim) {
  pdlc_pdl p;
  memset(&p, 0, sizeof p);
  p.ndims = nd;
  for (int i = 0; i < nd; i++) p.dims[i] = dims[i];
  pdlc_pdl *pp[1] = {&p};
  int rd[1] = {realdims}, cr[1] = {0}, pdl = -1;
  *max_dim = -1;
  find_max_pthread(pp, 1, rd, cr, target, max_pthread, max_dim, &pdl);
}

/* ------------------------------------------------------------------ */
/* initbroadcaststruct (implicit broadcasting) + dim checks            */
/* (pdlbroadcast.c:211-327, 340-488; pdlapi.c:829-869)                 */
/* ------------------------------------------------------------------ */
static int64_t nvals_of(const pdlc_pdl *p) {
  int64_t n = 1;
  for (int i = 0; i < p->ndims; i++) n *= p->dims[i];
  return n;
}

/* pdl_dim_checks (pdlbroadcast.c:211-273) */
static int dim_checks(trans *tr, const int *creating, int load_only, int nimpl, pdlc_error *err) {
  const vtable *vt = tr->vt;
  pdlc_pdl **pdls = tr->pdls;
  int npdls = vt->npdls, j;
  for (int i = 0; i < npdls; i++) {
    pdlc_pdl *p = pdls[i];
    int ninds = vt->par_realdims[i], nd = p->ndims;
    if (!load_only && creating[i]) continue;
    int isoutput = i >= vt->nparents;
    for (j = 0; j < ninds; j++) {
      int id = vt->par_ind[i][j];
      int64_t sz = tr->ind_sizes[id];
      if (j >= nd && sz == -1) sz = tr->ind_sizes[id] = 1;
      if (load_only && creating[i]) continue;
      if (sz == -1 || (j < nd && sz == 1)) { tr->ind_sizes[id] = p->dims[j]; continue; }
      if (j >= nd && isoutput && sz != 1)
        return param_error(err, vt, i, pdls, nimpl, creating,
                           "index '%s' size %lld, can't broadcast over output ndarray with size > 1",
                           vt->ind_names[id], (long long)sz);
      if (isoutput && sz != 1 && (p->state & PDLC_VAFFINE) && j < nd && p->incs[j] == 0)
        return param_error(err, vt, i, pdls, nimpl, creating,
                           "index '%s' size %lld, can't broadcast over dummy dim with size > 1",
                           vt->ind_names[id], (long long)sz);
      if (j < nd && sz != p->dims[j] && (isoutput || p->dims[j] != 1))
        return param_error(err, vt, i, pdls, nimpl, creating, "index '%s' size %lld, but ndarray dim has size %lld",
                           vt->ind_names[id], (long long)sz, (long long)p->dims[j]);
    }
  }
  return 0;
}

/* pdl->broadcastids[k]: a plain ndarray (nbroadcastids 0 here) has the one
 * entry ndims; slots past the list are the zeroed def_broadcastids of
 * pdl_pdlnew (pdlapi.c:211-214, 226-228) */
static int64_t bid_at(const pdlc_pdl *p, int k) {
  int n = p->nbroadcastids > 0 ? p->nbroadcastids : 1;
  if (n > PDLC_MAXBIDS) n = PDLC_MAXBIDS;
  if (k >= n) return 0;
  return p->nbroadcastids > 0 ? p->broadcastids[k] : p->ndims;
}
static int nbids(const pdlc_pdl *p) {
  int n = p->nbroadcastids > 0 ? p->nbroadcastids : 1;
  return n > PDLC_MAXBIDS ? PDLC_MAXBIDS : n;
}

static int redodims(trans *tr, const int *creating, int nthreads, int pthread_size_mi, pdlc_error *err) {
  const vtable *vt = tr->vt;
  brc_t *b = &tr->brc;
  pdlc_pdl **pdls = tr->pdls;
  int npdls = vt->npdls, j, nth;
  memset(b, 0, sizeof *b);
  b->npdls = npdls;
  b->mag_nth = b->mag_nthpdl = b->mag_nthr = -1;
  /* pdl_initbroadcaststruct (pdlbroadcast.c:354-395) */
  int nimpl = 0, nids = 0;
  for (j = 0; j < npdls; j++) {
    if (creating[j]) continue;
    /* lists that do not ascend to ndims let one dim serve two broadcast
     * dims (walked past the ndarray's end): refused, as the device does */
    pdlc_pdl *p = pdls[j];
    if (p->nbroadcastids > PDLC_MAXBIDS)
      return make_error(err, 1, "pdlhip: %s: parameter '%s' has more than %d broadcast ids", vt->name,
                        vt->par_names[j], PDLC_MAXBIDS);
    for (int k = 0; k < nbids(p); k++)
      if (bid_at(p, k) < (k ? bid_at(p, k - 1) : 0) || bid_at(p, k) > p->ndims ||
          (k == nbids(p) - 1 && bid_at(p, k) != p->ndims))
        return make_error(err, 1, "pdlhip: %s: parameter '%s': broadcast ids not ascending to ndims %d", vt->name,
                          vt->par_names[j], p->ndims);
    if (nbids(pdls[j]) > nids) nids = nbids(pdls[j]);
    int nb = (int)bid_at(pdls[j], 0) - vt->par_realdims[j];
    if (nb > nimpl) nimpl = nb;
  }
  b->nimpl = nimpl;
  /* explicit dims per id: every ndarray's counts add to ndims, the last
   * ndarray's counts are the ones looped over (pdlbroadcast.c:389-392) */
  int64_t nbroadcastids[PDLC_MAXBIDS] = {0}, nexpl = 0;
  for (j = 0; j < npdls; j++) {
    if (creating[j]) continue;
    for (int i = 0; i < nids; i++) {
      int64_t d = bid_at(pdls[j], i + 1) - bid_at(pdls[j], i);
      nbroadcastids[i] = d > 0 ? d : 0;
      nexpl += nbroadcastids[i];
    }
  }
  /* pdl_autopthreadmagic (pdlbroadcast.c:168-209) */
  int mag_pdl = -1, mag_dim = -1, mag_thr = 0;
  if (nthreads > 0) {
    int64_t largest = 0;
    for (j = 0; j < npdls; j++) {
      if (creating[j]) continue;
      int64_t nv = nvals_of(pdls[j]);
      if (nv > largest) largest = nv;
    }
    if (!((largest >> 20) < pthread_size_mi)) {
      int maxp = 0;
      find_max_pthread(pdls, npdls, vt->par_realdims, creating, nthreads, &maxp, &mag_dim, &mag_pdl);
      if (maxp > 1) mag_thr = maxp;
    }
  }
  int ndims = nimpl + (int)nexpl;
  b->nextra = 2 - ndims > 0 ? 2 - ndims : 0;
  ndims += b->nextra;
  if (ndims > MAXB) return make_error(err, 2, "too many broadcast dims");
  b->ndims = ndims;
  if (mag_thr) {
    b->mag_nth = mag_dim - vt->par_realdims[mag_pdl];
    b->mag_nthpdl = mag_pdl;
    b->mag_nthr = mag_thr;
  }
  for (nth = 0; nth < ndims; nth++) b->dims[nth] = 1;
  /* pdl_broadcast_dim_checks (pdlbroadcast.c:275-327) */
  for (nth = 0; nth < nimpl; nth++) {
    for (j = 0; j < npdls; j++) {
      if (creating[j]) continue;
      int isoutput = j >= vt->nparents;
      int rd = vt->par_realdims[j];
      pdlc_pdl *p = pdls[j];
      if (nth >= bid_at(p, 0) - rd) {
        if (isoutput && b->dims[nth] != 1)
          return param_error(err, vt, nth, pdls, nimpl, creating,
                             "implicit dim %d size %lld, can't broadcast over output ndarray with size > 1",
                             nth, (long long)b->dims[nth]);
        continue;
      }
      int64_t cur = p->dims[nth + rd];
      if (isoutput && cur == 1 && cur != b->dims[nth])
        return param_error(err, vt, nth, pdls, nimpl, creating, "implicit dim %d size %lld, but dim has size %lld",
                           nth, (long long)b->dims[nth], (long long)cur);
      if (isoutput && cur != 1 && (p->state & PDLC_VAFFINE) && p->incs[nth + rd] == 0)
        return param_error(err, vt, nth, pdls, nimpl, creating, "implicit dim %d size %lld, but dim is dummy",
                           nth, (long long)b->dims[nth]);
      if (cur != 1) {
        if (b->dims[nth] != 1) {
          if (b->dims[nth] != cur)
            return param_error(err, vt, j, pdls, nimpl, creating,
                               "Mismatched implicit broadcast dimension %d: size %lld vs. %lld", nth,
                               (long long)b->dims[nth], (long long)cur);
        } else {
          b->dims[nth] = cur;
        }
        b->incs[nth * npdls + j] = p->incs[nth + rd];
      }
    }
  }
  /* the explicit ids' dims (pdlbroadcast.c:437-467) */
  nth = nimpl;
  for (int id = 0; id < nids; id++) {
    for (int64_t i = 0; i < nbroadcastids[id]; i++, nth++) {
      for (j = 0; j < npdls; j++) {
        if (creating[j]) continue;
        pdlc_pdl *p = pdls[j];
        if (nbids(p) < id) continue;
        if (bid_at(p, id + 1) - bid_at(p, id) <= i) continue;
        int64_t which = i + bid_at(p, id);
        if (which < 0 || which >= p->ndims) return make_error(err, 1, "broadcast id names a dim past the ndarray");
        int64_t mydim = p->dims[which];
        if (mydim == 1) continue;
        if (which < vt->par_realdims[j])
          return make_error(err, 1, "pdlhip: %s: parameter '%s': dim %lld is both an active dim and in broadcast id %d",
                            vt->name, vt->par_names[j], (long long)which, id + 1);
        if (b->dims[nth] == 1) {
          b->dims[nth] = mydim;
        } else if (b->dims[nth] != mydim) {
          int64_t at = i + vt->par_realdims[j];
          return param_error(err, vt, j, pdls, nimpl, creating,
                             "Mismatched implicit broadcast dimension %lld: should be %lld, is %lld", (long long)i,
                             (long long)b->dims[nth], (long long)(at < p->ndims ? p->dims[at] : 0));
        }
        b->incs[nth * npdls + j] = p->incs[which];
      }
    }
  }
  /* thread split of the magicked dim (pdlbroadcast.c:469-484) */
  if (b->mag_nthr > 0) {
    int64_t md = b->dims[b->mag_nth];
    b->mag_stride = md / b->mag_nthr;
    b->mag_skip = md % b->mag_nthr;
  }
  if (dim_checks(tr, creating, 0, nimpl, err)) return 1;
  /* creating-pdl dims promotion for dims a creating pdl alone names */
  for (int i = 0; i < npdls; i++) {
    if (!creating[i]) continue;
    for (j = 0; j < vt->par_realdims[i]; j++) {
      int id = vt->par_ind[i][j];
      if (tr->ind_sizes[id] == -1) tr->ind_sizes[id] = 1;
    }
  }
  /* pdl_broadcast_create_parameter (pdlbroadcast.c:490-523) */
  for (int i = 0; i < npdls; i++) {
    if (!creating[i]) continue;
    if (nimpl != b->ndims - b->nextra)
      return make_error(err, 1, "%s: trying to create parameter '%s' while explicitly broadcasting.\n"
                        "See the manual for why this is impossible", vt->name, vt->par_names[i]);
    pdlc_pdl *p = pdls[i];
    int rd = vt->par_realdims[i];
    if (rd + nimpl > PDLC_MAXDIMS) return make_error(err, 2, "too many dims");
    p->nbroadcastids = 0;
    p->ndims = rd + nimpl;
    for (j = 0; j < rd; j++) p->dims[j] = tr->ind_sizes[vt->par_ind[i][j]];
    for (j = 0; j < nimpl; j++) p->dims[rd + j] = b->dims[j];
    int64_t inc = 1;
    for (j = 0; j < p->ndims; j++) { p->incs[j] = inc; inc *= p->dims[j]; }
    p->offs = 0;
    for (j = 0; j < nimpl; j++) b->incs[j * npdls + i] = p->incs[j + rd];
  }
  for (int i = 0; i < npdls; i++) b->offs[i] = pdls[i]->offs;
  /* core inc_sizes (pdlapi.c:858-864) */
  for (int i = 0; i < npdls; i++) {
    pdlc_pdl *p = pdls[i];
    for (j = 0; j < vt->par_realdims[i]; j++)
      tr->inc_sizes[i][j] = (p->ndims <= j || p->dims[j] <= 1) ? 0 : p->incs[j];
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* dtype conversion (pdlconv.c:163-174)                                */
/* ------------------------------------------------------------------ */
static void get_elem_as(const void *base, int t, int64_t off, long double *ld, intmax_t *im, uintmax_t *um) {
  switch (t) {
#define X(code, ctype, ppsym, isint, isuns) case code: { ctype v = ((const ctype *)base)[off]; \
      *ld = (long double)v; *im = (intmax_t)v; *um = (uintmax_t)v; break; }
    TYPES_REAL(X)
#undef X
  default: *ld = 0; *im = 0; *um = 0;
  }
}

/* one element: src (type st) -> dst (type dt); unsigned targets go through
 * intmax_t as COPYCONVERT does */
#define CONV_CASE_DST(dcode, dctype, dppsym, disint, disuns) case dcode: \
  if (disuns) ((dctype *)dst)[doff] = (dctype)(intmax_t)sv; else ((dctype *)dst)[doff] = (dctype)sv; break;
#define CONV_CASE_SRC(scode, sctype, sppsym, sisint, sisuns) case scode: { \
    sctype sv = ((const sctype *)src)[soff]; \
    switch (dt) { TYPES_REAL_(CONV_CASE_DST) default: break; } break; }
/* with a complex end: C's conversions (a complex converted to a real type
 * keeps its real part; through intmax_t for an unsigned target, as
 * COPYCONVERT writes it) */
#define TYPES_CPLX_DST(X) TYPES_REAL_(X) X(PDLC_CF, float complex, G, 0, 0) X(PDLC_CD, double complex, C, 0, 0)
#define CONV_CASE_SRC_C(scode, sctype, sppsym, sisint, sisuns) case scode: { \
    sctype sv = ((const sctype *)src)[soff]; \
    switch (dt) { TYPES_CPLX_DST(CONV_CASE_DST) default: break; } break; }
static void convert_one(const void *src, int st, int64_t soff, void *dst, int dt, int64_t doff) {
  if (type_cplx(st) || type_cplx(dt)) {
    switch (st) {
      TYPES_REAL(CONV_CASE_SRC_C)
      CONV_CASE_SRC_C(PDLC_CF, float complex, G, 0, 0)
      CONV_CASE_SRC_C(PDLC_CD, double complex, C, 0, 0)
    default: break;
    }
    return;
  }
  switch (st) {
    TYPES_REAL(CONV_CASE_SRC)
  default: break;
  }
}

void pdlc_convert(const void *src, int srctype, void *dst, int dsttype, int64_t n) {
  for (int64_t i = 0; i < n; i++) convert_one(src, srctype, i, dst, dsttype, i);
}

/* walk every element of a view in dim-0-fastest order */
typedef void (*elem_fn)(void *ctx, int64_t view_off, int64_t dense_idx);
static void walk_view(const pdlc_pdl *p, elem_fn fn, void *ctx) {
  int64_t n = nvals_of(p);
  int64_t idx[PDLC_MAXDIMS] = {0};
  int64_t off = p->offs;
  for (int64_t k = 0; k < n; k++) {
    fn(ctx, off, k);
    for (int d = 0; d < p->ndims; d++) {
      off += p->incs[d];
      if (++idx[d] < p->dims[d]) break;
      off -= p->incs[d] * p->dims[d];
      idx[d] = 0;
    }
  }
}

typedef struct {
  const pdlc_pdl *src;
  void *dst;
  int dsttype;
  int bad;
  pdlc_anyval sbv, dbv;
  int sbv_isnan;
} conv_ctx;

static int elem_isbad(const pdlc_pdl *p, int64_t off, const pdlc_anyval *bv, int bvnan) {
  switch (p->datatype) {
#define X(code, ctype, ppsym, isint, isuns) case code: { ctype v = ((const ctype *)p->data)[off]; \
      return ISBAD2(v, bv->value.ppsym, bvnan, ppsym); }
    TYPES_REAL(X)
#undef X
  default: return 0;
  }
}

static void set_badval_elem(void *base, int t, int64_t off, const pdlc_anyval *bv) {
  switch (t) {
#define X(code, ctype, ppsym, isint, isuns) case code: ((ctype *)base)[off] = bv->value.ppsym; break;
    TYPES_REAL(X)
#undef X
  default: break;
  }
}

static void conv_in_fn(void *vctx, int64_t off, int64_t k) {
  conv_ctx *c = vctx;
  if (c->bad && elem_isbad(c->src, off, &c->sbv, c->sbv_isnan)) set_badval_elem(c->dst, c->dsttype, k, &c->dbv);
  else convert_one(c->src->data, c->src->datatype, off, c->dst, c->dsttype, k);
}

static void effective_badval(const pdlc_pdl *p, pdlc_anyval *bv, int *isnan_) {
  if (p->has_badvalue && p->badvalue.type == p->datatype) *bv = p->badvalue;
  else pdlc_default_badval(p->datatype, bv);
  *isnan_ = (p->datatype == PDLC_F) ? isnan(bv->value.F) : (p->datatype == PDLC_D) ? isnan(bv->value.D) : 0;
}

/* ------------------------------------------------------------------ */
/* threads (pdlmagic.c:222-331)                                        */
/* ------------------------------------------------------------------ */
typedef struct { trans *tr; int thr; } thr_arg;
static void *thr_main(void *v) {
  thr_arg *a = v;
  a->tr->vt->readdata(a->tr, a->thr);
  return NULL;
}

static void run_readdata(trans *tr) {
  if (tr->brc.mag_nthr > 0) {
    int n = tr->brc.mag_nthr;
    pthread_t *th = malloc(sizeof(pthread_t) * n);
    thr_arg *args = malloc(sizeof(thr_arg) * n);
    int started = 0;
    for (int i = 0; i < n; i++) {
      args[i].tr = tr; args[i].thr = i;
      if (pthread_create(&th[i], NULL, thr_main, &args[i])) break;
      started++;
    }
    for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
    for (int i = started; i < n; i++) tr->vt->readdata(tr, i);
    free(th); free(args);
  } else {
    tr->vt->readdata(tr, 0);
  }
}

/* ------------------------------------------------------------------ */
/* pdl_run_<op> (PP.pm:1998-2014 -> pdlapi.c:746-869, 1314-1327)       */
/* ------------------------------------------------------------------ */
void pdlc_free(void *p) { free(p); }

static pdlc_plan *g_capture;
void pdlc_capture_plan(pdlc_plan *dst) { g_capture = dst; }

static void capture_plan(const trans *tr) {
  if (!g_capture) return;
  pdlc_plan *c = g_capture;
  memset(c, 0, sizeof *c);
  c->npdls = tr->brc.npdls;
  c->ndims = tr->brc.ndims;
  for (int d = 0; d < tr->brc.ndims && d < 20; d++) {
    c->dims[d] = tr->brc.dims[d];
    for (int j = 0; j < tr->brc.npdls; j++) c->incs[d * tr->brc.npdls + j] = tr->brc.incs[d * tr->brc.npdls + j];
  }
  for (int j = 0; j < tr->brc.npdls && j < 6; j++) {
    c->offs[j] = tr->brc.offs[j];
    c->inc_sizes[j][0] = tr->inc_sizes[j][0];
    c->inc_sizes[j][1] = tr->inc_sizes[j][1];
  }
  for (int k = 0; k < 3; k++) c->ind_sizes[k] = tr->ind_sizes[k];
  c->datatype = tr->datatype;
  c->bvalflag = tr->bvalflag;
}

int pdlc_run(const char *op, pdlc_pdl **user_pdls, int npdls, const int64_t *otherpars, int nthreads,
             int pthread_size_mi, pdlc_error *err) {
  memset(err, 0, sizeof *err);
  const vtable *vt = find_vtable(op);
  if (!vt) return make_error(err, 1, "pdlcpu: unknown op '%s'", op);
  if (npdls != vt->npdls) return make_error(err, 1, "pdlcpu: %s takes %d ndarrays, got %d", op, vt->npdls, npdls);
  trans tr;
  memset(&tr, 0, sizeof tr);
  tr.vt = vt;
  int i, j;
  for (i = 0; i < npdls; i++) tr.pdls[i] = user_pdls[i];
  /* biop/bifunc HdrCode: swap the inputs (Ops.pd:108-114) */
  if ((vt->badmode_biop == 1 || vt->badmode_biop == 2) && otherpars && otherpars[0]) {
    pdlc_pdl *t = tr.pdls[0]; tr.pdls[0] = tr.pdls[1]; tr.pdls[1] = t;
  }
  /* bvalflag: any input with the BADVAL state (pdlapi.c:760-766) */
  int inputs_bad = 0;
  for (i = 0; i < vt->nparents; i++) if (tr.pdls[i]->state & PDLC_BADVAL) inputs_bad = 1;
  tr.bvalflag = inputs_bad;
  /* null checks (pdlapi.c:711-738) */
  for (i = 0; i < vt->nparents; i++)
    if (tr.pdls[i]->state & PDLC_NOMYDIMS)
      return make_error(err, 1, "Error in %s: input parameter '%s' is null", vt->name, vt->par_names[i]);
  for (; i < npdls; i++)
    if ((tr.pdls[i]->state & PDLC_NOMYDIMS) && !(vt->par_flags[i] & P_ISCREAT))
      return make_error(err, 1, "Error in %s: io parameter '%s' is null", vt->name, vt->par_names[i]);
  /* type coercion */
  int ttype = transtype_select(vt, tr.pdls);
  tr.datatype = ttype;
  const int cplx_vt = vt->readdata == rd_matmult || vt->readdata == rd_assgn ||
                      (vt->readdata == rd_ufunc && (vt->opcode == R_SUMOVER || vt->opcode == R_PRODOVER ||
                                                    vt->opcode == R_CUMUSUMOVER || vt->opcode == R_CUMUPRODOVER)) ||
                      (vt->readdata == rd_biop && (vt->opcode == OP_PLUS || vt->opcode == OP_MINUS ||
                                                   vt->opcode == OP_MULT || vt->opcode == OP_DIVIDE ||
                                                   vt->opcode == OP_EQ || vt->opcode == OP_NE));
  if (!type_supported(ttype) && !(type_cplx(ttype) && cplx_vt))
    return make_error(err, 1, "pdlcpu: %s: type %s is not supported by this restatement", vt->name,
                      pdlc_type_name(ttype));
  if (type_cplx(ttype) && inputs_bad)
    return make_error(err, 1, "pdlcpu: %s: bad values in %s ndarrays are not restated", vt->name,
                      pdlc_type_name(ttype));
  int creating[MAXP] = {0};
  int req_type[MAXP];
  for (i = 0; i < npdls; i++) {
    req_type[i] = type_adjust_from_trans(ttype, vt->par_flags[i], vt->par_types[i]);
    if (tr.pdls[i]->state & PDLC_NOMYDIMS) {
      creating[i] = 1;
      tr.pdls[i]->datatype = req_type[i];
      tr.pdls[i]->has_badvalue = 0;
    }
  }
  /* converted copies for mismatched params (pdlapi.c:1262-1311) */
  pdlc_pdl conv[MAXP];
  int converted[MAXP] = {0};
  pdlc_pdl *orig[MAXP];
  for (i = 0; i < npdls; i++) orig[i] = tr.pdls[i];
  for (i = npdls - 1; i >= 0; i--) {
    pdlc_pdl *p = tr.pdls[i];
    if (creating[i]) continue;
    if (p->datatype == req_type[i]) continue;
    if (!type_supported(p->datatype) && !(type_cplx(p->datatype) && cplx_vt))
      return make_error(err, 1, "pdlcpu: type %s is not supported", pdlc_type_name(p->datatype));
    /* inplace: the same pdl as input and output converts once */
    int dup = -1;
    for (j = i + 1; j < npdls; j++) if (orig[j] == p && converted[j]) dup = j;
    if (dup >= 0) { tr.pdls[i] = &conv[dup]; continue; }
    pdlc_pdl *c = &conv[i];
    memset(c, 0, sizeof *c);
    c->datatype = req_type[i];
    c->state = p->state & PDLC_BADVAL;
    c->ndims = p->ndims;
    int64_t inc = 1;
    for (j = 0; j < p->ndims; j++) { c->dims[j] = p->dims[j]; c->incs[j] = inc; inc *= p->dims[j]; }
    /* the converted child keeps the broadcast ids (pdlconv.c:156-158) */
    c->nbroadcastids = p->nbroadcastids;
    memcpy(c->broadcastids, p->broadcastids, sizeof c->broadcastids);
    c->data = calloc(inc ? inc : 1, type_sizes[c->datatype]);
    conv_ctx cx;
    cx.src = p; cx.dst = c->data; cx.dsttype = c->datatype; cx.bad = (p->state & PDLC_BADVAL) != 0;
    int sbn;
    effective_badval(p, &cx.sbv, &sbn);
    cx.sbv_isnan = sbn;
    pdlc_default_badval(c->datatype, &cx.dbv);
    walk_view(p, conv_in_fn, &cx);
    tr.pdls[i] = c;
    converted[i] = 1;
  }
  /* REDODIMS (pdlapi.c:41-75): ops with dim setters (append's
   * mn=CALC(n+m)) first load sizes with a load-only dim check */
  for (j = 0; j < 3; j++) tr.ind_sizes[j] = -1;
  int rc = 0;
  if (vt->readdata == rd_append) {
    rc = dim_checks(&tr, creating, 1, 0, err);
    if (rc) goto cleanup;
    tr.ind_sizes[2] = tr.ind_sizes[0] + tr.ind_sizes[1];
  }
  if (vt->readdata == rd_nind) {
    rc = dim_checks(&tr, creating, 1, 0, err);
    if (rc) goto cleanup;
    tr.ind_sizes[1] = otherpars ? otherpars[0] : 0;
    if (tr.ind_sizes[1] > tr.ind_sizes[0]) {
      rc = make_error(err, 1, "Error in %s:m_size > n_size", vt->name);
      goto cleanup;
    }
  }
  rc = redodims(&tr, creating, nthreads, pthread_size_mi, err);
  if (rc) goto cleanup;
  /* bad flag propagation to outputs (pdlapi.c:806-808) */
  if (inputs_bad)
    for (i = vt->nparents; i < npdls; i++) tr.pdls[i]->state |= PDLC_BADVAL;
  /* allocate created outputs, zero-filled (pdlapi.c:204) */
  for (i = 0; i < npdls; i++) {
    if (!creating[i]) continue;
    int64_t n = nvals_of(tr.pdls[i]);
    tr.pdls[i]->data = calloc(n ? n : 1, type_sizes[tr.pdls[i]->datatype]);
    tr.pdls[i]->state &= ~PDLC_NOMYDIMS;
  }
  for (i = 0; i < npdls; i++) {
    int isn;
    effective_badval(tr.pdls[i], &tr.badval[i], &isn);
    tr.badval_isnan[i] = isn;
  }
  if ((vt->badmode_biop == 1 || vt->badmode_biop == 2) && otherpars) tr.swap = otherpars[0];
  if (vt->badmode_biop == 3 && otherpars) memcpy(&tr.dpar, &otherpars[0], sizeof tr.dpar);
  capture_plan(&tr);
  /* a BackCode writes its first parameter: it must already have the trans type */
  if (vt->readdata == rd_index && vt->opcode >= X_BACK && converted[0]) {
    rc = make_error(err, 1, "pdlcpu: %s: a must have the type of c", vt->name);
    goto cleanup;
  }
  run_readdata(&tr);
  if (tr.croaked) {
    rc = make_error(err, 1, "%s", tr.croak);
    goto cleanup;
  }
  for (i = 0; i < npdls; i++) {
    if (tr.setgood[i]) tr.pdls[i]->state &= ~PDLC_BADVAL;
    if (tr.setbad[i]) tr.pdls[i]->state |= PDLC_BADVAL;
  }
  /* write converted outputs back (pdlconv.c writebackdata) */
  for (i = vt->nparents; i < npdls; i++) {
    if (!converted[i]) continue;
    pdlc_pdl *c = &conv[i], *p = orig[i];
    int64_t n = nvals_of(p);
    int64_t idx[PDLC_MAXDIMS] = {0};
    int64_t off = p->offs;
    pdlc_anyval cbv, pbv;
    pdlc_default_badval(c->datatype, &cbv);
    int pbn;
    effective_badval(p, &pbv, &pbn);
    int cbn = (c->datatype == PDLC_F) ? isnan(cbv.value.F) : (c->datatype == PDLC_D) ? isnan(cbv.value.D) : 0;
    for (int64_t k = 0; k < n; k++) {
      if ((c->state & PDLC_BADVAL) && elem_isbad(c, k, &cbv, cbn)) set_badval_elem(p->data, p->datatype, off, &pbv);
      else convert_one(c->data, c->datatype, k, p->data, p->datatype, off);
      for (int d = 0; d < p->ndims; d++) {
        off += p->incs[d];
        if (++idx[d] < p->dims[d]) break;
        off -= p->incs[d] * p->dims[d];
        idx[d] = 0;
      }
    }
    if (c->state & PDLC_BADVAL) p->state |= PDLC_BADVAL;
  }
cleanup:
  for (i = 0; i < npdls; i++) if (converted[i]) free(conv[i].data);
  (void)anyval_as_double;
  return rc;
}

/* ------------------------------------------------------------------ */
/* RNG (Primitive-xoshiro256plus.c:28-43, 111-138)                     */
/* ------------------------------------------------------------------ */
uint64_t pdlc_splitmix64_next(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t pdlc_xoshiro256plus_next(uint64_t *s) {
  uint64_t result = s[0] + s[3];
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return result;
}

void pdlc_srand(uint64_t *s, uint64_t seed, int nthreads) {
  uint64_t x = seed;
  for (int i = 0; i < 4 * nthreads; i++) s[i] = pdlc_splitmix64_next(&x);
}

double pdlc_drand(uint64_t *s) { return (double)(pdlc_xoshiro256plus_next(s) >> 11) * 0x1.0p-53; }

void pdlc_random_fill(double *out, int64_t n, uint64_t seed) {
  uint64_t s[4];
  pdlc_srand(s, seed, 1);
  for (int64_t i = 0; i < n; i++) out[i] = pdlc_drand(s);
}

/* draws [skip, skip + n) of the single stream seeded with `seed` */
void pdlc_random_fill_from(double *out, int64_t n, uint64_t seed, uint64_t skip) {
  uint64_t s[4];
  pdlc_srand(s, seed, 1);
  for (uint64_t i = 0; i < skip; i++) (void)pdlc_xoshiro256plus_next(s);
  for (int64_t i = 0; i < n; i++) out[i] = pdlc_drand(s);
}

/* the state after `skip` draws of the stream seeded with `seed` */
void pdlc_rand_state_after(uint64_t *out, uint64_t seed, uint64_t skip) {
  pdlc_srand(out, seed, 1);
  for (uint64_t i = 0; i < skip; i++) (void)pdlc_xoshiro256plus_next(out);
}

/* ------------------------------------------------------------------ */
/* ndtri: PDL::Math's inverse of the normal distribution function      */
/* (Math.pd:346-356 -> Math/ndtri.c, Cephes 2.1, S. L. Moshier), which  */
/* grandom applies to randsym draws (Primitive.pd:2256-2261).  Cephes'  */
/* published rational approximations: around the middle                */
/* x = w + w^3 P0(w^2)/Q0(w^2), w = y - 1/2, times sqrt(2 pi); in the    */
/* tails z = sqrt(-2 log y), x = z - log(z)/z - P(1/z)/(z Q(1/z)) with   */
/* (P1, Q1) for z < 8 and (P2, Q2) beyond; Q tables without their       */
/* leading 1.  polevl/p1evl are Horner's rule from the leading          */
/* coefficient (p1evl with an implicit leading 1).                      */
/* ------------------------------------------------------------------ */
static const double ND_P0[5] = {-5.99633501014107895267E1, 9.80010754185999661536E1, -5.66762857469070293439E1,
                                1.39312609387279679503E1, -1.23916583867381258016E0};
static const double ND_Q0[8] = {1.95448858338141759834E0, 4.67627912898881538453E0, 8.63602421390890590575E1,
                                -2.25462687854119370527E2, 2.00260212380060660359E2, -8.20372256168333339912E1,
                                1.59056225126211695515E1, -1.18331621121330003142E0};
static const double ND_P1[9] = {4.05544892305962419923E0, 3.15251094599893866154E1, 5.71628192246421288162E1,
                                4.40805073893200834700E1, 1.46849561928858024014E1, 2.18663306850790267539E0,
                                -1.40256079171354495875E-1, -3.50424626827848203418E-2, -8.57456785154685413611E-4};
static const double ND_Q1[8] = {1.57799883256466749731E1, 4.53907635128879210584E1, 4.13172038254672030440E1,
                                1.50425385692907503408E1, 2.50464946208309415979E0, -1.42182922854787788574E-1,
                                -3.80806407691578277194E-2, -9.33259480895457427372E-4};
static const double ND_P2[9] = {3.23774891776946035970E0, 6.91522889068984211695E0, 3.93881025292474443415E0,
                                1.33303460815807542389E0, 2.01485389549179081538E-1, 1.23716634817820021358E-2,
                                3.01581553508235416007E-4, 2.65806974686737550832E-6, 6.23974539184983293730E-9};
static const double ND_Q2[8] = {6.02427039364742014255E0, 3.67983563856160859403E0, 1.37702099489081330271E0,
                                2.16236993594496635890E-1, 1.34204006088543189037E-2, 3.28014464682127739104E-4,
                                2.89247864745380683936E-6, 6.79019408009981274425E-9};

static double nd_polevl(double x, const double *c, int n) {
  double r = c[0];
  for (int i = 1; i <= n; i++) r = r * x + c[i];
  return r;
}
static double nd_p1evl(double x, const double *c, int n) {
  double r = x + c[0];
  for (int i = 1; i < n; i++) r = r * x + c[i];
  return r;
}

double pdlc_ndtri(double y0) {
  const double expm2 = 0.13533528323661269189, s2pi = 2.50662827463100050242E0, maxnum = 1.79769313486231570815E308;
  if (y0 <= 0.0) return -maxnum;
  if (y0 >= 1.0) return maxnum;
  int neg = 1;
  double y = y0;
  if (y > 1.0 - expm2) {
    y = 1.0 - y;
    neg = 0;
  }
  if (y > expm2) {
    y = y - 0.5;
    double y2 = y * y;
    double x = y + y * (y2 * nd_polevl(y2, ND_P0, 4) / nd_p1evl(y2, ND_Q0, 8));
    return x * s2pi;
  }
  double x = sqrt(-2.0 * log(y));
  double x0 = x - log(x) / x;
  double z = 1.0 / x;
  double x1 = x < 8.0 ? z * nd_polevl(z, ND_P1, 8) / nd_p1evl(z, ND_Q1, 8) : z * nd_polevl(z, ND_P2, 8) / nd_p1evl(z, ND_Q2, 8);
  x = x0 - x1;
  return neg ? -x : x;
}

void pdlc_ndtri_n(const double *y, double *x, int64_t n) {
  for (int64_t i = 0; i < n; i++) x[i] = pdlc_ndtri(y[i]);
}

/* ------------------------------------------------------------------ */
/* Truth helper: row-major C[h][w] = sum_t A[h][t] B[t][w] in long     */
/* double (64-bit significand) for scoring double results.             */
/* ------------------------------------------------------------------ */
void pdlc_truth_matmult_ld(const double *a, const double *b, double *c, int64_t h, int64_t t, int64_t w) {
  long double *row = malloc(sizeof(long double) * (w ? w : 1));
  for (int64_t i = 0; i < h; i++) {
    for (int64_t k = 0; k < w; k++) row[k] = 0;
    for (int64_t s = 0; s < t; s++) {
      long double av = a[i * t + s];
      const double *bp = b + s * w;
      for (int64_t k = 0; k < w; k++) row[k] += av * (long double)bp[k];
    }
    for (int64_t k = 0; k < w; k++) c[i * w + k] = (double)row[k];
  }
  free(row);
}

/* ------------------------------------------------------------------ */
/* Slices.pd rangeb (Slices.pd:561-1015), the engine of range and      */
/* indexND: MakeComp (index and size conditioning, boundary letters,   */
/* the corners), RedoDims (the child's dims) and the EquivCPOffs loop   */
/* with PP's $EQUIVCPTRUNC (PP.pm:1272-1298): forward, a truncated or   */
/* (bad mode) bad source element gives the child BAD / 0; back, the    */
/* child's elements go to the parent unless truncated.  `size` is NULL */
/* for an undefined size, else nsize values; ind must be indx.         */
/* The child is created (calloc'd, dense).  back: run the BackCode into */
/* parent from an existing child of the same dims.                     */
/* ------------------------------------------------------------------ */
int pdlc_rangeb(pdlc_pdl *parent, const pdlc_pdl *ind, const int64_t *size, int64_t nsize, const char *boundary,
                pdlc_pdl *child, int back, pdlc_error *err) {
  memset(err, 0, sizeof *err);
  if (ind->datatype != PDLC_IND) return make_error(err, 1, "pdlcpu: rangeb: the index must be indx");
  if (!type_supported(parent->datatype)) return make_error(err, 1, "pdlcpu: rangeb: type not supported");
  if (ind->ndims && ind->dims[0] == 0)
    return make_error(err, 1, "Error in rangeb:can't handle Empty indices -- call range instead");
  const int64_t rdim = ind->ndims ? ind->dims[0] : 1;
  int64_t nitems = 1;
  for (int i = 1; i < ind->ndims; i++) nitems *= ind->dims[i];
  const int itdim = ind->ndims ? ind->ndims - 1 : 0;
  if (rdim > 64 || itdim > PDLC_MAXDIMS) return make_error(err, 1, "pdlcpu: rangeb: too many dims");
  /* boundary letters (MakeComp) */
  char bnd[64];
  const size_t blen = boundary ? strlen(boundary) : 0;
  for (int64_t i = 0; i < rdim; i++) {
    if (!blen) { bnd[i] = 0; continue; }
    switch (boundary[i < (int64_t)blen - 1 ? i : (int64_t)blen - 1]) {
    case '0': case 'f': case 'F': bnd[i] = 0; break;
    case '1': case 't': case 'T': bnd[i] = 1; break;
    case '2': case 'e': case 'E': bnd[i] = 2; break;
    case '3': case 'p': case 'P': bnd[i] = 3; break;
    case '4': case 'm': case 'M': bnd[i] = 4; break;
    default:
      return make_error(err, 1, "Error in rangeb:Unknown boundary condition '%c' in range", boundary[i]);
    }
  }
  /* sizes */
  int64_t sizes[64];
  for (int64_t i = 0; i < rdim; i++) sizes[i] = 0;
  if (size && nsize > 0) {
    if (nsize == 1) {
      for (int64_t i = 0; i < rdim; i++) sizes[i] = size[0];
      if (rdim > 0 && sizes[0] < 0) return make_error(err, 1, "Error in rangeb:Negative range size is not allowed\n");
    } else if (nsize <= rdim) {
      for (int64_t i = 0; i < rdim; i++) {
        sizes[i] = i < nsize ? size[i] : 0;
        if (sizes[i] < 0) return make_error(err, 1, "Error in rangeb:Negative range sizes are not allowed\n");
      }
    } else {
      return make_error(err, 1, "Error in rangeb:Size must match index's 0th dim\n");
    }
  }
  int ntsize = 0;
  for (int64_t i = 0; i < rdim; i++) ntsize += sizes[i] != 0;
  /* corners[item][k] */
  int64_t *corners = malloc(sizeof(int64_t) * (size_t)(rdim * (nitems ? nitems : 1)));
  {
    int64_t it[PDLC_MAXDIMS] = {0};
    for (int64_t item = 0; item < nitems; item++) {
      int64_t ioff = ind->offs;
      for (int k = 0; k < itdim; k++) ioff += it[k] * ind->incs[k + 1];
      for (int64_t j = 0; j < rdim; j++) corners[item * rdim + j] = ((const int64_t *)ind->data)[ioff + (ind->ndims ? ind->incs[0] : 0) * j];
      for (int k = 0; k < itdim && ++it[k] >= ind->dims[k + 1]; k++) it[k] = 0;
    }
  }
  /* RedoDims */
  const int pdim = parent->ndims;
  if (rdim > pdim + 5 && nsize != rdim) {
    free(corners);
    return make_error(err, 1,
                      "Error in rangeb:Ludicrous number of extra dims in range index; leaving child null.\n"
                      "  (%lld implicit dims is > 5; index has %lld dims; source has %d dim%s.)\n"
                      "  This often means that your index PDL is incorrect.\n"
                      "  To avoid this message, allocate dummy dims in\n"
                      "  the source or use %lld dims in range's size field.\n",
                      (long long)(rdim - pdim), (long long)rdim, pdim, pdim > 1 ? "s" : "", (long long)rdim);
  }
  const int stdim = pdim - rdim > 0 ? (int)(pdim - rdim) : 0;
  const int cnd = itdim + ntsize + stdim;
  if (cnd > PDLC_MAXDIMS) { free(corners); return make_error(err, 1, "pdlcpu: rangeb: too many child dims"); }
  /* the child's dims (RedoDims): index broadcast dims, nonzero sizes,
     source broadcast dims */
  int64_t cdims[PDLC_MAXDIMS], inc = 1;
  {
    int dim = itdim;
    for (int64_t i = 0; i < rdim; i++)
      if (sizes[i]) inc *= (cdims[dim++] = sizes[i]);
    for (dim = 0; dim < itdim; dim++) inc *= (cdims[dim] = ind->dims[dim + 1]);
    dim = itdim + ntsize;
    for (int i = 0; i < stdim; i++) inc *= (cdims[dim++] = parent->dims[i + rdim]);
  }
  if (parent->ndims && parent->dims[0] == 0)
    for (int64_t k = 0; k < rdim; k++) if (bnd[k]) bnd[k] = 1;
  const int64_t nchild = inc;
  if (!back) {
    memset(child, 0, sizeof *child);
    child->datatype = parent->datatype;
    child->ndims = cnd;
    int64_t dinc = 1;
    for (int d = 0; d < cnd; d++) { child->dims[d] = cdims[d]; child->incs[d] = dinc; dinc *= cdims[d]; }
    child->state = parent->state & PDLC_BADVAL;
    child->data = calloc(nchild ? nchild : 1, type_sizes[child->datatype]);
  }
  /* child element (coords cc) -> offset in the child as the caller holds it */
  const int bad = (parent->state & PDLC_BADVAL) != 0;
  pdlc_anyval pbv, cbv;
  int pbn, cbn;
  effective_badval(parent, &pbv, &pbn);
  effective_badval(child, &cbv, &cbn);
  /* EquivCPOffsCode */
  const int prdim = (int)(rdim < pdim ? rdim : pdim);
  int64_t szv[64 + PDLC_MAXDIMS];
  for (int64_t k = 0; k < rdim; k++) szv[k] = sizes[k];
  for (int64_t k = rdim; k < pdim; k++) szv[k] = parent->dims[k];
  int rc = 0;
  for (int64_t item = 0; item < nitems && !rc; item++) {
    int64_t iter[64 + PDLC_MAXDIMS];
    const int64_t nit = rdim > pdim ? rdim : pdim;
    for (int64_t k = 0; k < nit; k++) iter[k] = 0;
    int64_t k;
    do {
      int64_t coords[64];
      int trunc = 0;
      for (k = 0; k < prdim; k++) {
        int64_t ck = iter[k] + corners[item * rdim + k];
        const int64_t pd = parent->dims[k];
        if (ck < 0 || ck >= pd) {
          switch (bnd[k]) {
          case 0: rc = make_error(err, 1, "Error in rangeb:index out-of-bounds in range (index vector #%ld)", (long)item); break;
          case 1: trunc = 1; break;
          case 2: ck = (ck >= pd) ? pd - 1 : 0; break;
          case 3: ck %= pd; if (ck < 0) ck += pd; break;
          case 4:
            ck += pd; ck %= pd * 2; if (ck < 0) ck += pd * 2; ck -= pd;
            if (ck < 0) { ck *= -1; ck -= 1; }
            break;
          }
          if (rc) break;
        }
        coords[k] = ck;
      }
      if (rc) break;
      for (; k < rdim; k++) {
        int64_t ck = iter[k] + corners[item * rdim + k];
        switch (bnd[k]) {
        case 0:
          if (ck != 0) rc = make_error(err, 1, "Error in rangeb:Too many dims in range index (and you've forbidden boundary violations)");
          break;
        case 1: trunc |= (ck != 0); break;
        default: break;
        }
        if (rc) break;
      }
      if (rc) break;
      /* the child's coordinates: the item over the index broadcast dims,
         the in-range position over the nonzero sizes, then the source
         broadcast dims (RedoDims' order) */
      int64_t cbase = child->offs, poff = 0;
      {
        int64_t r = item;
        for (int d = 0; d < itdim; d++) { cbase += (r % cdims[d]) * child->incs[d]; r /= cdims[d]; }
        int k2 = itdim;
        for (k = 0; k < rdim; k++) {
          if (k < prdim) poff += coords[k] * parent->incs[k];
          if (sizes[k]) cbase += iter[k] * child->incs[k2++];
        }
      }
      do {
        int64_t poff1 = parent->offs + poff, cpos = cbase;
        int k2 = itdim + ntsize;
        for (k = rdim; k < pdim; k++) {
          poff1 += iter[k] * parent->incs[k];
          cpos += iter[k] * child->incs[k2++];
        }
        if (!back) {
          if (trunc || (bad && elem_isbad(parent, poff1, &pbv, pbn))) {
            if (bad) set_badval_elem(child->data, child->datatype, cpos, &cbv);
            else set_badval_elem(child->data, child->datatype, cpos, &(pdlc_anyval){child->datatype, 0, {0}});
          } else {
            convert_one(parent->data, parent->datatype, poff1, child->data, child->datatype, cpos);
          }
        } else if (!trunc) {
          if (bad && elem_isbad(child, cpos, &cbv, cbn)) set_badval_elem(parent->data, parent->datatype, poff1, &pbv);
          else convert_one(child->data, child->datatype, cpos, parent->data, parent->datatype, poff1);
        }
        for (k = rdim; k < pdim && ++iter[k] >= parent->dims[k]; k++) iter[k] = 0;
      } while (k < pdim);
      for (k = 0; k < rdim && ++iter[k] >= szv[k]; k++) iter[k] = 0;
    } while (k < rdim);
  }
  free(corners);
  if (rc && !back) { free(child->data); child->data = NULL; }
  return rc;
}
