// pdlhip C ABI: host planner + dispatch.
//
// The planner restates what PDL does before a readdata runs:
//   type selection      pdl__transtype_select, lib/PDL/Core/pdlapi.c:1182-1236
//   output types        pdl__set_output_type_badvalue, pdlapi.c:1240-1260
//   input conversion    pdl__type_convert, pdlapi.c:1262-1311 (device kernel)
//   null checks         pdl_trans_flow_null_checks, pdlapi.c:711-738
//   bad-value flag      pdl_make_trans_mutual, pdlapi.c:760-770, 806-808
//   broadcast plan      pdl_initbroadcaststruct / pdl_broadcast_dim_checks /
//                       pdl_dim_checks / pdl_broadcast_create_parameter,
//                       lib/PDL/Core/pdlbroadcast.c:211-523
//   core inc sizes      pdl_redodims_default, pdlapi.c:829-869
//   error texts         pdl_param_error, lib/PDL/Core/pdlutil.c:65-145
// then turns PDL's broadcast plan into a canonical launch plan (unit dims
// dropped, dims ordered by the output's strides, contiguous dims merged) and
// launches the HIP kernels of k_*.hip.  There is no CPU compute path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <functional>
#include <map>
#include <mutex>
#include <queue>
#include <string>
#include <tuple>
#include <vector>

#include <sys/stat.h>

#include "../../include/pdlhip.h"
#include "common.hpp"
#include "launch.hpp"

namespace pdlhip {

// ------------------------------------------------------------------ errors
static thread_local std::string g_msg;

// ------------------------------------------------------------------ allocation-free reductions
// The scratch of pdlhip_run_reduce_noalloc: the long-row reductions'
// per-chunk partials (at most 2048 + 1023 of them, under 200 KiB) live in
// this array instead of a cached hipMalloc block.  Each GPU holds its own
// copy until the process ends; host code uses the device address from
// hipGetSymbolAddress (pdlhip_init), never the array's name.
constexpr size_t NOALLOC_BYTES = (size_t)256 << 10;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static thread_local bool g_noalloc = false;   // inside pdlhip_run_reduce_noalloc
static thread_local bool g_writeback = false; // inside pdlhip_writebackdata: run the op's BackCode

static pdlhip_error ok() { return pdlhip_error{PDLHIP_ENONE, nullptr}; }

static pdlhip_error make_err(int kind, const char *fmt, ...) {
  char buf[8192];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_msg = buf;
  return pdlhip_error{kind, g_msg.c_str()};
}

static bool capturing();

// device memory that could not be had: during graph capture only cached
// blocks are handed out (no hipMalloc, no synchronisation), so an op whose
// scratch is not cached yet must run once before the capture
static pdlhip_error oom_err(const char *what) {
  if (g_noalloc)
    return make_err(PDLHIP_EUSERERROR,
                    "pdlhip_run_reduce_noalloc: %s needs more scratch than the %zu bytes this entry point reserves",
                    what, NOALLOC_BYTES);
  if (capturing())
    return make_err(PDLHIP_EUSERERROR,
                    "pdlhip: %s: needs device memory that is not cached while a graph is being captured "
                    "(run it once before pdlhip_graph_begin)",
                    what);
  return make_err(PDLHIP_EFATAL, "pdlhip: %s: out of device memory", what);
}

static pdlhip_error hip_err(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory) return oom_err(what);
  return make_err(PDLHIP_EFATAL, "pdlhip: %s: %s", what, hipGetErrorString(e));
}

#define RETURN_IF(e) do { pdlhip_error e_ = (e); if (e_.error) return e_; } while (0)

static const char *type_name(int t) {
  static const char *n[] = {"sbyte", "byte", "short", "ushort", "long", "ulong", "indx", "ulonglong",
                            "longlong", "float", "double", "ldouble", "cfloat", "cdouble", "cldouble"};
  return (t >= 0 && t < PDLHIP_NTYPES) ? n[t] : "INVALID";
}

// ------------------------------------------------------------------ state
struct Deferred {
  std::vector<MmItem> items;
  int T = -1;
  bool fused = true;
  int missed = 0;      // list items that had to run at once (bad values, conversions, another type)
};

struct Global {
  Deferred *defer = nullptr;   // set while pdlhip_run_matmult_list plans its items
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int mm_mode = PDLHIP_MM_SPLIT16;
  bool plan_only = false;
  std::mutex mu;
  // graph capture (pdlhip_graph_begin .. _end): blocks released while
  // capturing stay out of the cache until the graph that may read or write
  // them is destroyed
  bool capturing = false;
  std::vector<void *> pinned;
  void *noalloc_scratch = nullptr;   // device address of `scratch`
  // PDL's random-number stream (one state: PDL with autopthread off)
  uint64_t rng[4] = {0, 0, 0, 0};
  bool rng_seeded = false;
  uint64_t *rng_jumps = nullptr;     // the in-chunk jump matrices, on the device
};
static Global g;

// true while work enqueued on the library stream is being captured (by
// pdlhip_graph_begin or by a caller's own capture of pdlhip_stream())
static bool capturing() {
  if (g.capturing) return true;
  if (!g.stream) return false;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(g.stream, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

static pdlhip_error ensure_init() {
  if (g.plan_only) return ok();
  if (g.device >= 0) return ok();
  return make_err(PDLHIP_EFATAL, "pdlhip: not initialised (call pdlhip_init)");
}

// host -> device copy ordered on the library stream (a block from the cache
// may still be read by work queued before it was released), complete on
// return so the host buffer may go away
static hipError_t upload_now(void *dst, const void *src, size_t bytes) {
  if (!bytes) return hipSuccess;
  if (capturing()) return hipErrorStreamCaptureUnsupported;   // preparing steps run before a capture
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g.stream);
  return e == hipSuccess ? hipStreamSynchronize(g.stream) : e;
}

// ------------------------------------------------------------------ allocator
// A size-bucketed cache of HBM blocks.  Every launch of the library is
// ordered on one stream, so a block released here can be handed to the next
// op without a device sync.
struct Cache {
  std::mutex mu;
  std::multimap<size_t, void *> freelist;
  std::map<void *, size_t> owned;
  size_t cached = 0;
};
static Cache g_cache;

static size_t round_size(size_t n) {
  if (n == 0) n = 1;
  if (n <= ((size_t)1 << 20)) return (n + 511) & ~(size_t)511;
  return (n + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
}

static void release_cached() {
  for (auto &kv : g_cache.freelist) {
    g_cache.owned.erase(kv.second);
    (void)hipFree(kv.second);
  }
  g_cache.freelist.clear();
  g_cache.cached = 0;
}

void *dev_alloc(size_t n) {
  if (g_noalloc) return nullptr;
  size_t sz = round_size(n);
  std::lock_guard<std::mutex> lk(g_cache.mu);
  auto it = g_cache.freelist.lower_bound(sz);
  if (it != g_cache.freelist.end() && it->first <= sz + sz / 4) {
    void *p = it->second;
    g_cache.cached -= it->first;
    g_cache.freelist.erase(it);
    return p;
  }
  if (capturing()) return nullptr;   // no hipMalloc or synchronisation inside a capture
  void *p = nullptr;
  if (hipMalloc(&p, sz) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamSynchronize(g.stream);
    release_cached();
    p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
  }
  g_cache.owned[p] = sz;
  return p;
}

bool dev_free(void *p) {
  if (!p) return true;
  std::lock_guard<std::mutex> lk(g_cache.mu);
  auto it = g_cache.owned.find(p);
  if (it == g_cache.owned.end()) return false;
  if (capturing()) {
    // a captured kernel may use this block at every replay: it returns to
    // the cache when the graph is destroyed
    g.pinned.push_back(p);
    return true;
  }
  g_cache.freelist.insert({it->second, p});
  g_cache.cached += it->second;
  if (g_cache.cached > ((size_t)64 << 30)) {  // keep the cache bounded
    (void)hipStreamSynchronize(g.stream);
    release_cached();
  }
  return true;
}

void *scratch_alloc(size_t bytes) { return dev_alloc(bytes); }
void scratch_free(void *p) { dev_free(p); }

// ------------------------------------------------------------------ signatures
enum Family { FAM_BIOP, FAM_BIFUNC, FAM_ASSGN, FAM_RED, FAM_NIND, FAM_MM, FAM_UFUNC, FAM_BADOP, FAM_INDEX };
enum ParFlag { P_OUT = 1, P_CREAT = 2, P_TYPED = 4, P_TPLUS = 8 };

struct Sig {
  std::string name;
  int npdls, nparents;
  std::vector<std::string> par_names;
  std::vector<int> realdims;
  std::vector<std::vector<int>> ind;   // per par, ind id of each core dim
  std::vector<std::string> ind_names;
  std::vector<int> flags;
  std::vector<int> par_types;
  std::vector<int> gentypes;           // GenericTypes in PP order
  int family, op;
};

// GenericTypes lists of Ops.pd:1-14 / Ufunc.pd / Primitive.pd
static const std::vector<int> GT_ALL = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14};
static const std::vector<int> GT_REAL = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
static const std::vector<int> GT_INT = {0, 1, 2, 3, 4, 5, 6, 7, 8};
static const std::vector<int> GT_POWER = {12, 13, 14, 9, 11, 10};  // [@$C, @$F], $F = F,LD,D
static const std::vector<int> GT_ATAN2 = {9, 11, 10};
static const std::vector<int> GT_AF = {9, 11, 12, 13, 14, 10};     // $AF: non-integer, D last

static std::vector<Sig> build_sigs() {
  std::vector<Sig> v;
  auto bi = [&](const char *n, int op, const std::vector<int> &gt, int fam) {
    v.push_back(Sig{n, 3, 2, {"a", "b", "c"}, {0, 0, 0}, {{}, {}, {}}, {}, {0, 0, P_OUT | P_CREAT}, {-1, -1, -1}, gt,
                    fam, op});
  };
  bi("plus", OP_PLUS, GT_ALL, FAM_BIOP);
  bi("mult", OP_MULT, GT_ALL, FAM_BIOP);
  bi("minus", OP_MINUS, GT_ALL, FAM_BIOP);
  bi("divide", OP_DIVIDE, GT_ALL, FAM_BIOP);
  bi("gt", OP_GT, GT_REAL, FAM_BIOP);
  bi("lt", OP_LT, GT_REAL, FAM_BIOP);
  bi("le", OP_LE, GT_REAL, FAM_BIOP);
  bi("ge", OP_GE, GT_REAL, FAM_BIOP);
  bi("eq", OP_EQ, GT_ALL, FAM_BIOP);
  bi("ne", OP_NE, GT_ALL, FAM_BIOP);
  bi("shiftleft", OP_SHL, GT_INT, FAM_BIOP);
  bi("shiftright", OP_SHR, GT_INT, FAM_BIOP);
  bi("or2", OP_OR, GT_INT, FAM_BIOP);
  bi("and2", OP_AND, GT_INT, FAM_BIOP);
  bi("xor", OP_XOR, GT_INT, FAM_BIOP);
  bi("power", OP_POWER, GT_POWER, FAM_BIFUNC);
  bi("atan2", OP_ATAN2, GT_ATAN2, FAM_BIFUNC);
  bi("modulo", OP_MODULO, GT_REAL, FAM_BIFUNC);
  bi("spaceship", OP_SPACESHIP, GT_REAL, FAM_BIFUNC);
  v.push_back(Sig{"assgn", 2, 1, {"a", "b"}, {0, 0}, {{}, {}}, {}, {0, P_OUT | P_CREAT}, {-1, -1}, GT_ALL, FAM_ASSGN, 0});
  auto red = [&](const char *n, int op, int oflags, int otype, const std::vector<int> &gt, bool scan) {
    v.push_back(Sig{n, 2, 1, {"a", "b"}, {1, scan ? 1 : 0}, {{0}, scan ? std::vector<int>{0} : std::vector<int>{}},
                    {"n"}, {0, P_OUT | P_CREAT | oflags}, {-1, otype}, gt, FAM_RED, op});
  };
  red("sumover", R_SUM, P_TYPED | P_TPLUS, PDLHIP_L, GT_ALL, false);
  red("prodover", R_PROD, P_TYPED | P_TPLUS, PDLHIP_L, GT_ALL, false);
  red("dsumover", R_DSUM, P_TYPED, PDLHIP_D, GT_REAL, false);
  red("dprodover", R_DPROD, P_TYPED, PDLHIP_D, GT_REAL, false);
  red("cumusumover", R_CUMUSUM, P_TYPED | P_TPLUS, PDLHIP_L, GT_ALL, true);
  red("cumuprodover", R_CUMUPROD, P_TYPED | P_TPLUS, PDLHIP_L, GT_ALL, true);
  red("dcumusumover", R_DCUMUSUM, P_TYPED, PDLHIP_D, GT_REAL, true);
  red("dcumuprodover", R_DCUMUPROD, P_TYPED, PDLHIP_D, GT_REAL, true);
  red("average", R_AVERAGE, P_TYPED | P_TPLUS, PDLHIP_L, GT_REAL, false);
  red("daverage", R_DAVERAGE, P_TYPED, PDLHIP_D, GT_REAL, false);
  red("minimum", R_MIN, 0, -1, GT_REAL, false);
  red("maximum", R_MAX, 0, -1, GT_REAL, false);
  red("minimum_ind", R_MIN_IND, P_TYPED, PDLHIP_IND, GT_REAL, false);
  red("maximum_ind", R_MAX_IND, P_TYPED, PDLHIP_IND, GT_REAL, false);
  v.back().par_names[1] = "c";
  for (auto &s : v)
    if (s.name == "minimum" || s.name == "maximum" || s.name == "minimum_ind") s.par_names[1] = "c";
  v.push_back(Sig{"minmaximum", 5, 1, {"a", "cmin", "cmax", "cmin_ind", "cmax_ind"}, {1, 0, 0, 0, 0},
                  {{0}, {}, {}, {}, {}}, {"n"},
                  {0, P_OUT | P_CREAT, P_OUT | P_CREAT, P_OUT | P_CREAT | P_TYPED, P_OUT | P_CREAT | P_TYPED},
                  {-1, -1, -1, PDLHIP_IND, PDLHIP_IND}, GT_REAL, FAM_RED, R_MINMAXIMUM});
  auto nind = [&](const char *n, int op) {
    v.push_back(Sig{n, 2, 1, {"a", "c"}, {1, 1}, {{0}, {1}}, {"n", "m"}, {0, P_OUT | P_CREAT | P_TYPED},
                    {-1, PDLHIP_IND}, GT_REAL, FAM_NIND, op});
  };
  nind("minimum_n_ind", R_MIN_N_IND);
  nind("maximum_n_ind", R_MAX_N_IND);
  v.push_back(Sig{"matmult", 3, 2, {"a", "b", "c"}, {2, 2, 2}, {{0, 1}, {2, 0}, {2, 1}}, {"t", "h", "w"},
                  {0, 0, P_OUT | P_CREAT}, {-1, -1, -1}, GT_ALL, FAM_MM, 0});
  // Ops.pd ufuncs: Pars a(); [o]b() (Ops.pd:222-265); GenericTypes of
  // Ops.pd:327-333, 365 ($T, $A, reals, [@$C, @$F]) and _rabs (343-358)
  auto uf = [&](const char *n, int op, const std::vector<int> &gt) {
    v.push_back(Sig{n, 2, 1, {"a", "b"}, {0, 0}, {{}, {}}, {}, {0, P_OUT | P_CREAT}, {-1, -1}, gt, FAM_UFUNC, op});
  };
  uf("bitnot", OP_BITNOT, GT_INT);
  uf("sqrt", OP_SQRT, GT_ALL);
  uf("sin", OP_SIN, GT_ALL);
  uf("cos", OP_COS, GT_ALL);
  uf("not", OP_NOT, GT_REAL);
  uf("exp", OP_EXP, GT_POWER);
  uf("log", OP_LOG, GT_POWER);
  uf("log10", OP_LOG10, GT_ALL);
  uf("abs", OP_ABS, GT_REAL);
  // Bad.pd elementwise ops (Bad.pd:343-395, 584-840)
  auto bo = [&](const char *n, int op, const std::vector<int> &gt, int oflags, int otype) {
    v.push_back(Sig{n, 2, 1, {"a", "b"}, {0, 0}, {{}, {}}, {}, {0, P_OUT | P_CREAT | oflags}, {-1, otype}, gt,
                    FAM_BADOP, op});
  };
  bo("isbad", B_ISBAD, GT_ALL, P_TYPED, PDLHIP_L);
  bo("isgood", B_ISGOOD, GT_ALL, P_TYPED, PDLHIP_L);
  bo("setvaltobad", B_SETVALTOBAD, GT_ALL, 0, -1);
  bo("setnantobad", B_SETNANTOBAD, GT_AF, 0, -1);
  bo("setinftobad", B_SETINFTOBAD, GT_AF, 0, -1);
  bo("setnonfinitetobad", B_SETNONFINITETOBAD, GT_AF, 0, -1);
  bo("setbadtonan", B_SETBADTONAN, GT_AF, 0, -1);
  bo("setbadtoval", B_SETBADTOVAL, GT_ALL, 0, -1);
  v.push_back(Sig{"setbadif", 3, 2, {"a", "mask", "b"}, {0, 0, 0}, {{}, {}, {}}, {}, {0, P_TYPED, P_OUT | P_CREAT},
                  {-1, PDLHIP_L, -1}, GT_ALL, FAM_BADOP, B_SETBADIF});
  // Slices.pd:149-220: index a(n); indx ind(); [oca] c()  /  index1d a(n);
  // indx ind(m); [oca] c(m)  /  index2d a(na,nb); indx inda(); indx indb();
  // [oca] c()  (GenericTypes ppdefs_all)
  v.push_back(Sig{"index", 3, 2, {"a", "ind", "c"}, {1, 0, 0}, {{0}, {}, {}}, {"n"}, {0, P_TYPED, P_OUT | P_CREAT},
                  {-1, PDLHIP_IND, -1}, GT_ALL, FAM_INDEX, 0});
  v.push_back(Sig{"index1d", 3, 2, {"a", "ind", "c"}, {1, 1, 1}, {{0}, {1}, {1}}, {"n", "m"},
                  {0, P_TYPED, P_OUT | P_CREAT}, {-1, PDLHIP_IND, -1}, GT_ALL, FAM_INDEX, 1});
  v.push_back(Sig{"index2d", 4, 3, {"a", "inda", "indb", "c"}, {2, 0, 0, 0}, {{0, 1}, {}, {}, {}}, {"na", "nb"},
                  {0, P_TYPED, P_TYPED, P_OUT | P_CREAT}, {-1, PDLHIP_IND, PDLHIP_IND, -1}, GT_ALL, FAM_INDEX, 2});
  return v;
}

static const Sig *find_sig(const char *name) {
  static const std::vector<Sig> sigs = build_sigs();
  for (auto &s : sigs)
    if (s.name == name) return &s;
  return nullptr;
}

// ------------------------------------------------------------------ bad values
static pdlhip_value default_badval(int t) {
  pdlhip_value v;
  std::memset(&v, 0, sizeof v);
  switch (t) {
  case PDLHIP_SB: v.A = INT8_MIN; break;
  case PDLHIP_B: v.B = UINT8_MAX; break;
  case PDLHIP_S: v.S = INT16_MIN; break;
  case PDLHIP_US: v.U = UINT16_MAX; break;
  case PDLHIP_L: v.L = INT32_MIN; break;
  case PDLHIP_UL: v.K = UINT32_MAX; break;
  case PDLHIP_IND: v.N = INT64_MIN; break;
  case PDLHIP_ULL: v.P = UINT64_MAX; break;
  case PDLHIP_LL: v.Q = INT64_MIN; break;
  case PDLHIP_F: v.F = -3.40282346638528859812e+38F; break;
  case PDLHIP_D: v.D = -1.79769313486231570815e+308; break;
  default: break;
  }
  return v;
}

// per-ndarray bad value if it has one of its own type, else the type default
// (PDL_DECLARE_PARAMETER_BADVAL, lib/PDL/Core/pdlcore.h:202-205)
static BadInfo badinfo(const pdlhip_pdl *p, int check) {
  BadInfo b;
  std::memset(&b, 0, sizeof b);
  b.check = check;
  b.bv = (p->has_badvalue && p->badvalue.type == p->datatype) ? p->badvalue.value : default_badval(p->datatype);
  b.isnan = (p->datatype == PDLHIP_F) ? std::isnan(b.bv.F) : (p->datatype == PDLHIP_D) ? std::isnan(b.bv.D) : 0;
  return b;
}

// ------------------------------------------------------------------ type selection
static int adjust_from_trans(int ttype, int flags, int par_type) {
  if (flags & P_TYPED) return !(flags & P_TPLUS) ? par_type : std::max(par_type, ttype);
  return ttype;
}

static int select_transtype(const Sig &s, pdlhip_pdl **pdls) {
  bool avail[PDLHIP_NTYPES] = {false};
  int last = -1;
  for (int t : s.gentypes) avail[last = t] = true;
  if (s.gentypes.front() == last) return last;
  bool use_last = false;
  int ret = -1;
  for (int i = s.npdls - 1; i >= 0; i--) {
    const pdlhip_pdl *p = pdls[i];
    if (!(p->state & PDLHIP_NOMYDIMS) && !(s.flags[i] & P_TYPED)) {
      int t = p->datatype;
      bool is_avail = (t >= 0 && t < PDLHIP_NTYPES) && avail[t];
      if (!is_avail && t > last) use_last = true;
      if (t >= 0 && is_avail && ret < t) ret = t;
    }
    if (i == s.nparents && ret >= 0) return ret;
  }
  if (use_last || ret < 0 || !avail[ret]) ret = last;
  return ret;
}

// ------------------------------------------------------------------ broadcast plan
struct Brc {
  int npdls = 0, ndims = 0, nimpl = 0;
  int64_t dims[PDLHIP_MAXDIMS + 2];
  int64_t incs[(PDLHIP_MAXDIMS + 2) * MAXP];
  int64_t offs[MAXP];
  int64_t ind_sizes[4];
  int64_t inc_sizes[MAXP][2];
};

static pdlhip_error param_error(const Sig &s, int pidx, pdlhip_pdl **pdls, int nimpl, const bool *creating,
                                const char *fmt, ...) {
  std::string m;
  char tmp[1024];
  if (pidx < 0 || pidx >= s.npdls) {
    m += "ERROR: UNKNOWN PARAMETER";
  } else {
    m += s.name + "(";
    for (int i = 0; i < s.npdls; i++) { m += s.par_names[i]; if (i < s.npdls - 1) m += ","; }
    m += "): parameter '" + s.par_names[pidx] + "': ";
  }
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(tmp, sizeof tmp, fmt, ap);
  va_end(ap);
  m += tmp;
  snprintf(tmp, sizeof tmp, "\nThere are %d PDLs in the expression; %d broadcast dim%s.\n", s.npdls, nimpl,
           nimpl == 1 ? "" : "s");
  m += tmp;
  int maxrd = 0;
  for (int i = 0; i < s.npdls; i++) maxrd = std::max(maxrd, s.realdims[i]);
  m += "   PDL IN EXPR.    ";
  if (maxrd > 0) {
    char f[32];
    snprintf(f, sizeof f, "%%%ds", 8 * maxrd + 3);
    snprintf(tmp, sizeof tmp, f, "ACTIVE DIMS | ");
    m += tmp;
  }
  m += "BROADCAST DIMS\n";
  for (int i = 0; i < s.npdls; i++) {
    snprintf(tmp, sizeof tmp, "   #%3d (%s", i, creating[i] ? "null)\n" : "normal): ");
    m += tmp;
    if (creating[i]) continue;
    if (maxrd == 1) m += "    ";
    for (int j = 0; j < maxrd - s.realdims[i]; j++) { snprintf(tmp, sizeof tmp, "%8s", " "); m += tmp; }
    for (int j = 0; j < s.realdims[i]; j++) { snprintf(tmp, sizeof tmp, "%8lld", (long long)pdls[i]->dims[j]); m += tmp; }
    if (maxrd) m += " | ";
    for (int j = 0; j < nimpl && j + s.realdims[i] < pdls[i]->ndims; j++) {
      snprintf(tmp, sizeof tmp, "%8lld", (long long)pdls[i]->dims[j + s.realdims[i]]);
      m += tmp;
    }
    m += "\n";
  }
  return make_err(PDLHIP_EUSERERROR, "%s", m.c_str());
}

// pdl_dim_checks (pdlbroadcast.c:211-273)
static pdlhip_error dim_checks(const Sig &s, pdlhip_pdl **pdls, const bool *creating, bool load_only, int nimpl,
                               int64_t *ind_sizes) {
  for (int i = 0; i < s.npdls; i++) {
    const pdlhip_pdl *p = pdls[i];
    int nd = p->ndims;
    if (!load_only && creating[i]) continue;
    bool isout = i >= s.nparents;
    for (int j = 0; j < s.realdims[i]; j++) {
      int id = s.ind[i][j];
      int64_t sz = ind_sizes[id];
      if (j >= nd && sz == -1) sz = ind_sizes[id] = 1;
      if (load_only && creating[i]) continue;
      if (sz == -1 || (j < nd && sz == 1)) { ind_sizes[id] = p->dims[j]; continue; }
      if (j >= nd && isout && sz != 1)
        return param_error(s, i, pdls, nimpl, creating,
                           "index '%s' size %lld, can't broadcast over output ndarray with size > 1",
                           s.ind_names[id].c_str(), (long long)sz);
      if (isout && sz != 1 && (p->state & PDLHIP_VAFFINE) && j < nd && p->incs[j] == 0)
        return param_error(s, i, pdls, nimpl, creating,
                           "index '%s' size %lld, can't broadcast over dummy dim with size > 1",
                           s.ind_names[id].c_str(), (long long)sz);
      if (j < nd && sz != p->dims[j] && (isout || p->dims[j] != 1))
        return param_error(s, i, pdls, nimpl, creating, "index '%s' size %lld, but ndarray dim has size %lld",
                           s.ind_names[id].c_str(), (long long)sz, (long long)p->dims[j]);
    }
  }
  return ok();
}

// entry k of an ndarray's broadcast-id list: a plain ndarray has the one
// entry ndims; entries past the list read as 0, as PDL's zeroed
// def_broadcastids slots do (pdlapi.c:211-214)
static int64_t bcast_id(const pdlhip_pdl *p, int k) {
  const int n = p->nbroadcastids > 0 ? std::min(p->nbroadcastids, PDLHIP_MAXBIDS) : 1;
  if (k >= n) return 0;
  return p->nbroadcastids > 0 ? p->broadcastids[k] : p->ndims;
}
static int bcast_nids(const pdlhip_pdl *p) {
  return p->nbroadcastids > 0 ? std::min(p->nbroadcastids, PDLHIP_MAXBIDS) : 1;
}
// an ndarray carrying explicit broadcast ids (a broadcastI view)
static bool explicit_ids(const pdlhip_pdl *p) {
  return p && (p->nbroadcastids > 1 || (p->nbroadcastids == 1 && p->broadcastids[0] != p->ndims));
}
// the entry points with planners of their own (ranges, ragged lists, fused
// chains) take implicit broadcasting only
static pdlhip_error refuse_explicit_ids(const char *who, pdlhip_pdl *const *ps, int n) {
  for (int i = 0; i < n; i++)
    if (explicit_ids(ps[i]))
      return make_err(PDLHIP_EUSERERROR, "%s: ndarray %d has explicit broadcast ids, which this entry point does "
                      "not take", who, i);
  return ok();
}

// pdl_initbroadcaststruct (implicit broadcast dims, then the dims of the
// explicit broadcast ids, nobl = 2) + checks + output creation + core inc
// sizes.  The CPU pthread split is not planned: the whole broadcast is one
// device launch.
static pdlhip_error redodims(const Sig &s, pdlhip_pdl **pdls, const bool *creating, Brc &b) {
  const int np = s.npdls;
  b.npdls = np;
  int nimpl = 0, nids = 0;
  for (int j = 0; j < np; j++) {
    if (creating[j]) continue;
    // an id list must ascend within [0, ndims] and end at ndims (as
    // broadcastI's do), so that no dim serves two broadcast dims
    const pdlhip_pdl *p = pdls[j];
    if (p->nbroadcastids > PDLHIP_MAXBIDS)
      return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: parameter '%s' has more than %d broadcast ids", s.name.c_str(),
                      s.par_names[j].c_str(), PDLHIP_MAXBIDS);
    for (int k = 0; k < bcast_nids(p); k++)
      if (bcast_id(p, k) < (k ? bcast_id(p, k - 1) : 0) || bcast_id(p, k) > p->ndims ||
          (k == bcast_nids(p) - 1 && bcast_id(p, k) != p->ndims))
        return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: parameter '%s': broadcast ids not ascending to ndims %d",
                        s.name.c_str(), s.par_names[j].c_str(), p->ndims);
    nids = std::max(nids, bcast_nids(pdls[j]));
    nimpl = std::max<int>(nimpl, (int)(bcast_id(pdls[j], 0) - s.realdims[j]));
  }
  if (nimpl > PDLHIP_MAXDIMS) return make_err(PDLHIP_EFATAL, "pdlhip: too many broadcast dims");
  // dims per explicit id (pdlbroadcast.c:389-392): PDL sums every ndarray's
  // counts into ndims but loops over the last ndarray's counts only; the
  // other dims it adds stay size 1 and are not planned here
  int64_t nexpl_all = 0, nexpl[PDLHIP_MAXBIDS] = {0};
  for (int j = 0; j < np; j++) {
    if (creating[j]) continue;
    for (int i = 0; i < nids; i++) {
      nexpl[i] = std::max<int64_t>(0, bcast_id(pdls[j], i + 1) - bcast_id(pdls[j], i));
      nexpl_all += nexpl[i];
    }
  }
  int nplanned = nimpl;
  for (int i = 0; i < nids; i++) nplanned += (int)nexpl[i];
  if (nplanned > PDLHIP_MAXDIMS + 2) return make_err(PDLHIP_EFATAL, "pdlhip: too many broadcast dims");
  b.nimpl = nimpl;
  b.ndims = nplanned + std::max(0, 2 - nplanned);
  for (int d = 0; d < b.ndims; d++) b.dims[d] = 1;
  for (int k = 0; k < b.ndims * np; k++) b.incs[k] = 0;
  // pdl_broadcast_dim_checks (pdlbroadcast.c:275-327)
  for (int nth = 0; nth < nimpl; nth++) {
    for (int j = 0; j < np; j++) {
      if (creating[j]) continue;
      const pdlhip_pdl *p = pdls[j];
      bool isout = j >= s.nparents;
      int rd = s.realdims[j];
      if (nth >= bcast_id(p, 0) - rd) {
        if (isout && b.dims[nth] != 1)
          return param_error(s, nth, pdls, nimpl, creating,
                             "implicit dim %d size %lld, can't broadcast over output ndarray with size > 1", nth,
                             (long long)b.dims[nth]);
        continue;
      }
      int64_t cur = p->dims[nth + rd];
      if (isout && cur == 1 && cur != b.dims[nth])
        return param_error(s, nth, pdls, nimpl, creating, "implicit dim %d size %lld, but dim has size %lld", nth,
                           (long long)b.dims[nth], (long long)cur);
      if (isout && cur != 1 && (p->state & PDLHIP_VAFFINE) && p->incs[nth + rd] == 0)
        return param_error(s, nth, pdls, nimpl, creating, "implicit dim %d size %lld, but dim is dummy", nth,
                           (long long)b.dims[nth]);
      if (cur != 1) {
        if (b.dims[nth] != 1) {
          if (b.dims[nth] != cur)
            return param_error(s, j, pdls, nimpl, creating,
                               "Mismatched implicit broadcast dimension %d: size %lld vs. %lld", nth,
                               (long long)b.dims[nth], (long long)cur);
        } else {
          b.dims[nth] = cur;
        }
        b.incs[nth * np + j] = p->incs[nth + rd];
      }
    }
  }
  // the explicit ids' dims, after the implicit ones (pdlbroadcast.c:437-467)
  int nth = nimpl;
  for (int id = 0; id < nids; id++) {
    for (int i = 0; i < nexpl[id]; i++, nth++) {
      for (int j = 0; j < np; j++) {
        if (creating[j]) continue;
        const pdlhip_pdl *p = pdls[j];
        if (bcast_nids(p) < id) continue;
        if (bcast_id(p, id + 1) - bcast_id(p, id) <= i) continue;
        const int64_t which = i + bcast_id(p, id);
        if (which < 0 || which >= p->ndims) return make_err(PDLHIP_EUSERERROR, "pdlhip: broadcast id %d of "
                                                            "parameter '%s' names dim %lld of %d", id,
                                                            s.par_names[j].c_str(), (long long)which, p->ndims);
        const int64_t mydim = p->dims[which];
        if (mydim == 1) continue;
        // a dim both active and broadcast would be walked twice, past the
        // ndarray's end (PDL does not check this)
        if (which < s.realdims[j])
          return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: parameter '%s': dim %lld is both an active dim and in "
                          "broadcast id %d", s.name.c_str(), s.par_names[j].c_str(), (long long)which, id + 1);
        if (b.dims[nth] == 1) {
          b.dims[nth] = mydim;
        } else if (b.dims[nth] != mydim) {
          // PDL prints the dim at i + realdims as the size found
          const int at = i + s.realdims[j];
          return param_error(s, j, pdls, nimpl, creating,
                             "Mismatched implicit broadcast dimension %d: should be %lld, is %lld", i,
                             (long long)b.dims[nth], (long long)(at < p->ndims ? p->dims[at] : 0));
        }
        b.incs[nth * np + j] = p->incs[which];
      }
    }
  }
  RETURN_IF(dim_checks(s, pdls, creating, false, nimpl, b.ind_sizes));
  for (int i = 0; i < np; i++) {
    if (!creating[i]) continue;
    for (int j = 0; j < s.realdims[i]; j++)
      if (b.ind_sizes[s.ind[i][j]] == -1) b.ind_sizes[s.ind[i][j]] = 1;
  }
  // pdl_broadcast_create_parameter (pdlbroadcast.c:490-523)
  for (int i = 0; i < np; i++) {
    if (!creating[i]) continue;
    if (nexpl_all > 0)
      return make_err(PDLHIP_EUSERERROR,
                      "%s: trying to create parameter '%s' while explicitly broadcasting.\n"
                      "See the manual for why this is impossible",
                      s.name.c_str(), s.par_names[i].c_str());
    pdlhip_pdl *p = pdls[i];
    int rd = s.realdims[i];
    if (rd + nimpl > PDLHIP_MAXDIMS) return make_err(PDLHIP_EFATAL, "pdlhip: too many dims");
    p->ndims = rd + nimpl;
    p->nbroadcastids = 0;
    for (int j = 0; j < rd; j++) p->dims[j] = b.ind_sizes[s.ind[i][j]];
    for (int j = 0; j < nimpl; j++) p->dims[rd + j] = b.dims[j];
    int64_t inc = 1;
    for (int j = 0; j < p->ndims; j++) { p->incs[j] = inc; inc *= p->dims[j]; }
    p->offs = 0;
    p->state &= ~PDLHIP_VAFFINE;
    for (int j = 0; j < nimpl; j++) b.incs[j * np + i] = p->incs[j + rd];
  }
  for (int i = 0; i < np; i++) b.offs[i] = pdls[i]->offs;
  // core inc_sizes (pdlapi.c:858-864)
  for (int i = 0; i < np; i++)
    for (int j = 0; j < s.realdims[i]; j++)
      b.inc_sizes[i][j] = (pdls[i]->ndims <= j || pdls[i]->dims[j] <= 1) ? 0 : pdls[i]->incs[j];
  return ok();
}

// PDL's loop writes an output once per step of a broadcast dim over which
// the output does not move (inc 0): an output whose dim in an explicit
// broadcast id is 1 or a dummy (pdlbroadcast.c:437-467 checks only the
// sizes), or an output the implicit checks met before a later operand set
// the size.  The last write stays, so the plan keeps that dim's last index
// only.  Refused when outputs disagree on such a dim, or when an input
// shares an output's buffer (PDL's later steps would read earlier writes).
static pdlhip_error collapse_repeated_writes(const Sig &s, pdlhip_pdl **pdls, Brc &b) {
  const int np = s.npdls;
  for (int d = 0; d < b.ndims; d++) {
    if (b.dims[d] <= 1) continue;
    int nzero = 0;
    for (int j = s.nparents; j < np; j++) nzero += b.incs[d * np + j] == 0;
    if (nzero == 0) continue;
    if (nzero != np - s.nparents)
      return make_err(PDLHIP_EUSERERROR,
                      "pdlhip: %s: broadcast dim %d of size %lld repeats some outputs' writes but not others': "
                      "not supported on the device", s.name.c_str(), d, (long long)b.dims[d]);
    for (int i = 0; i < s.nparents; i++)
      for (int o = s.nparents; o < np; o++)
        if (pdls[i]->data && pdls[i]->data == pdls[o]->data)
          return make_err(PDLHIP_EUSERERROR,
                          "pdlhip: %s: output '%s' is written repeatedly over broadcast dim %d and shares its "
                          "buffer with input '%s': not supported on the device", s.name.c_str(),
                          s.par_names[o].c_str(), d, s.par_names[i].c_str());
    for (int i = 0; i < s.nparents; i++) {
      b.offs[i] += (b.dims[d] - 1) * b.incs[d * np + i];
      b.incs[d * np + i] = 0;
    }
    b.dims[d] = 1;
  }
  return ok();
}

// ------------------------------------------------------------------ canonical plans
struct Dimset {
  int n = 0;
  int np = 0;
  int64_t dims[PDLHIP_MAXDIMS + 2];
  int64_t inc[MAXP][PDLHIP_MAXDIMS + 2];
};

// drop unit dims, order by the key operand's |stride|, merge dims that are
// contiguous for every operand.  Returns false if the result is empty.
static bool canonicalise(Dimset &d, int key) {
  Dimset o;
  o.np = d.np;
  bool empty = false;
  for (int k = 0; k < d.n; k++) {
    if (d.dims[k] == 0) empty = true;
    if (d.dims[k] == 1) continue;
    o.dims[o.n] = d.dims[k];
    for (int p = 0; p < d.np; p++) o.inc[p][o.n] = d.inc[p][k];
    o.n++;
  }
  if (empty) { d.n = 0; return false; }
  // stable insertion sort on |inc[key]|
  for (int i = 1; i < o.n; i++) {
    for (int j = i; j > 0; j--) {
      int64_t x = std::llabs(o.inc[key][j]), y = std::llabs(o.inc[key][j - 1]);
      if (x >= y) break;
      std::swap(o.dims[j], o.dims[j - 1]);
      for (int p = 0; p < o.np; p++) std::swap(o.inc[p][j], o.inc[p][j - 1]);
    }
  }
  Dimset m;
  m.np = o.np;
  for (int k = 0; k < o.n; k++) {
    if (m.n > 0) {
      bool merge = true;
      for (int p = 0; p < o.np; p++)
        if (o.inc[p][k] != m.inc[p][m.n - 1] * m.dims[m.n - 1]) { merge = false; break; }
      if (merge) { m.dims[m.n - 1] *= o.dims[k]; continue; }
    }
    m.dims[m.n] = o.dims[k];
    for (int p = 0; p < o.np; p++) m.inc[p][m.n] = o.inc[p][k];
    m.n++;
  }
  if (m.n == 0) {
    m.n = 1;
    m.dims[0] = 1;
    for (int p = 0; p < m.np; p++) m.inc[p][0] = 0;
  }
  d = m;
  return true;
}

static Dimset brc_dimset(const Brc &b, const int *which, int np) {
  Dimset d;
  d.np = np;
  d.n = b.ndims;
  for (int k = 0; k < b.ndims; k++) {
    d.dims[k] = b.dims[k];
    for (int p = 0; p < np; p++) d.inc[p][k] = b.incs[k * b.npdls + which[p]];
  }
  return d;
}

static inline bool IsFloatCode(int T) { return T == PDLHIP_F || T == PDLHIP_D; }

// ------------------------------------------------------------------ dispatch
static hipError_t binop_dispatch(int T, int op, const EwPlan &p, const void *a, const void *b, void *c, bool bad,
                                 const BadInfo *bi, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return launch_binop_t##TC(op, p, a, b, c, bad, bi, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}
static hipError_t reduce_dispatch(int T, int op, int otype, const RedPlan &p, const void *a, void *const *outs,
                                  bool bad, const BadInfo &ba, const pdlhip_value *obv, int *flag, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return launch_reduce_t##TC(op, otype, p, a, outs, bad, ba, obv, flag, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}
static hipError_t matmult_dispatch(int T, const MmPlan &p, const void *a, const void *b, void *c, bool bad,
                                   const BadInfo *bi, int mode, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return launch_matmult_t##TC(p, a, b, c, bad, bi, mode, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

static bool fill_ew(EwPlan &e, Dimset &d) {
  std::memset(&e, 0, sizeof e);
  if (d.n > MAXK) return false;
  e.nd = d.n;
  e.n = 1;
  for (int k = 0; k < d.n; k++) {
    e.dims[k] = d.dims[k];
    e.n *= d.dims[k];
    for (int p = 0; p < 3; p++) e.inc[p][k] = p < d.np ? d.inc[p][k] : 0;
  }
  return true;
}

static int tune(const std::string &family, const std::vector<int> &cands, int dflt,
                const std::function<hipError_t(int)> &run_one);
static int ilog2(int64_t v);

// a fill kernel rather than hipMemsetAsync, so the clear is an ordinary
// kernel node wherever the work is captured
__global__ void zero_i32(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

static pdlhip_error readdata_index(const Sig &s, int T, bool bad, const Brc &b, pdlhip_pdl **pdls);

// the ops that take float and double complex on the device
static bool cplx_op(const Sig &s) {
  if (s.family == FAM_MM || s.family == FAM_ASSGN) return true;
  if (s.family == FAM_RED)
    return s.op == R_SUM || s.op == R_PROD || s.op == R_CUMUSUM || s.op == R_CUMUPROD;
  return s.family == FAM_BIOP && (s.op == OP_PLUS || s.op == OP_MINUS || s.op == OP_MULT || s.op == OP_DIVIDE ||
                                  s.op == OP_EQ || s.op == OP_NE);
}

// readdata of a complex op (k_complex.hip); every operand has type T here
static pdlhip_error readdata_cplx(const Sig &s, int T, const Brc &b, pdlhip_pdl **pdls) {
  hipError_t e = hipSuccess;
  if (s.family == FAM_BIOP || s.family == FAM_ASSGN) {
    const bool as = s.family == FAM_ASSGN;
    int which[3] = {0, as ? 0 : 1, as ? 1 : 2};
    Dimset d = brc_dimset(b, which, 3);
    if (!canonicalise(d, 2)) return ok();
    EwPlan p;
    if (!fill_ew(p, d)) return make_err(PDLHIP_EFATAL, "pdlhip: %s: too many distinct dims", s.name.c_str());
    for (int q = 0; q < 3; q++) p.off[q] = b.offs[which[q]];
    if (as) e = launch_cplx_convert(T, T, p, pdls[0]->data, pdls[1]->data, g.stream);
    else e = launch_cplx_binop(T, s.op, p, pdls[0]->data, pdls[1]->data, pdls[2]->data, g.stream);
  } else if (s.family == FAM_RED) {
    int which[2] = {0, 1};
    Dimset d = brc_dimset(b, which, 2);
    RedPlan p;
    std::memset(&p, 0, sizeof p);
    p.npdl = 2;
    p.n = b.ind_sizes[0];
    p.inc_n[0] = b.inc_sizes[0][0];
    p.inc_n[1] = (s.realdims[1] == 1) ? b.inc_sizes[1][0] : 0;
    const bool nonempty = canonicalise(d, 1);
    if (d.n > MAXK) return make_err(PDLHIP_EFATAL, "pdlhip: %s: too many distinct dims", s.name.c_str());
    p.nd = nonempty ? d.n : 0;
    p.nout = nonempty ? 1 : 0;
    for (int k = 0; k < p.nd; k++) {
      p.dims[k] = d.dims[k];
      p.nout *= d.dims[k];
      for (int q = 0; q < 2; q++) p.inc[q][k] = d.inc[q][k];
    }
    for (int q = 0; q < 2; q++) p.off[q] = b.offs[q];
    e = launch_cplx_reduce(T, s.op, p, pdls[0]->data, pdls[1]->data, g.stream);
  } else {
    int which[3] = {0, 1, 2};
    Dimset d = brc_dimset(b, which, 3);
    MmPlan p;
    std::memset(&p, 0, sizeof p);
    p.T = b.ind_sizes[0]; p.H = b.ind_sizes[1]; p.W = b.ind_sizes[2];
    p.a_t = b.inc_sizes[0][0]; p.a_h = b.inc_sizes[0][1];
    p.b_w = b.inc_sizes[1][0]; p.b_t = b.inc_sizes[1][1];
    p.c_w = b.inc_sizes[2][0]; p.c_h = b.inc_sizes[2][1];
    const bool nonempty = canonicalise(d, 2);
    if (d.n > MAXK) return make_err(PDLHIP_EFATAL, "pdlhip: matmult: too many distinct dims");
    p.nd = nonempty ? d.n : 0;
    p.nbatch = nonempty ? 1 : 0;
    for (int k = 0; k < p.nd; k++) {
      p.dims[k] = d.dims[k];
      p.nbatch *= d.dims[k];
      for (int q = 0; q < 3; q++) p.inc[q][k] = d.inc[q][k];
    }
    for (int q = 0; q < 3; q++) p.off[q] = b.offs[q];
    e = launch_cplx_matmult(T, p, pdls[0]->data, pdls[1]->data, pdls[2]->data, g.stream);
  }
  return e == hipSuccess ? ok() : hip_err(e, s.name.c_str());
}

// the readdata of one op over a planned broadcast (the pdlhip_readdata body)
static pdlhip_error readdata(const Sig &s, int T, bool bad, const Brc &b, pdlhip_pdl **pdls, const int64_t *other) {
  if (g.plan_only) return ok();
  if (type_is_cplx(T) && cplx_op(s) && !bad) return readdata_cplx(s, T, b, pdls);
  if (!type_on_device(T)) return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: type %s is not supported on the device",
                                          s.name.c_str(), type_name(T));
  if (s.family == FAM_INDEX) return readdata_index(s, T, bad, b, pdls);
  hipStream_t st = g.stream;
  hipError_t e = hipSuccess;
  const bool mixed = (s.family == FAM_BIOP || s.family == FAM_BIFUNC || s.family == FAM_UFUNC) &&
                     (pdls[0]->datatype != T || (s.family != FAM_UFUNC && pdls[1]->datatype != T));
  if (!mixed && (s.family == FAM_BIOP || s.family == FAM_BIFUNC)) {
    int which[3] = {0, 1, 2};
    Dimset d = brc_dimset(b, which, 3);
    if (!canonicalise(d, 2)) return ok();
    EwPlan p;
    if (!fill_ew(p, d)) return make_err(PDLHIP_EFATAL, "pdlhip: %s: too many distinct dims", s.name.c_str());
    for (int k = 0; k < 3; k++) p.off[k] = b.offs[k];
    BadInfo bi[3];
    for (int k = 0; k < 3; k++)
      bi[k] = badinfo(pdls[k], s.family == FAM_BIFUNC ? 1 : (pdls[k]->state & PDLHIP_BADVAL) != 0);
    e = binop_dispatch(T, s.op, p, pdls[0]->data, pdls[1]->data, pdls[2]->data, bad, bi, st);
  } else if (mixed) {
    // inputs of other types, converted in registers as their converted
    // copies would hold them (k_mixed.hip)
    const bool uf = s.family == FAM_UFUNC;
    int which[3] = {0, uf ? 0 : 1, uf ? 1 : 2};
    Dimset d = brc_dimset(b, which, 3);
    if (uf)
      for (int k = 0; k < d.n; k++) d.inc[1][k] = 0;
    if (!canonicalise(d, 2)) return ok();
    EwPlan p;
    if (!fill_ew(p, d)) return make_err(PDLHIP_EFATAL, "pdlhip: %s: too many distinct dims", s.name.c_str());
    p.off[0] = b.offs[0];
    p.off[1] = uf ? b.offs[0] : b.offs[1];
    p.off[2] = b.offs[uf ? 1 : 2];
    const pdlhip_pdl *pa = pdls[0], *pb = uf ? pdls[0] : pdls[1];
    MixedArgs m;
    std::memset(&m, 0, sizeof m);
    m.a = pa->data; m.ta = pa->datatype; m.sbad_a = (pa->state & PDLHIP_BADVAL) != 0; m.sba = badinfo(pa, 1);
    m.b = pb->data; m.tb = pb->datatype; m.sbad_b = (pb->state & PDLHIP_BADVAL) != 0; m.sbb = badinfo(pb, 1);
    m.a_scalar = m.b_scalar = 1;
    for (int k = 0; k < p.nd; k++) {
      if (p.inc[0][k] != 0 && p.dims[k] > 1) m.a_scalar = 0;
      if (p.inc[1][k] != 0 && p.dims[k] > 1) m.b_scalar = 0;
    }
    // the op's bad test on the converted values: the copies' bad value is
    // the default of T; a biop tests an input only if it is bad-flagged
    BadInfo bi[3];
    for (int k = 0; k < 2; k++) {
      const pdlhip_pdl *src = k == 0 ? pa : pb;
      std::memset(&bi[k], 0, sizeof(BadInfo));
      bi[k].bv = default_badval(T);
      bi[k].isnan = 0;
      bi[k].check = uf ? (k == 0) : (s.family == FAM_BIFUNC ? 1 : (src->state & PDLHIP_BADVAL) != 0);
    }
    bi[2] = badinfo(pdls[uf ? 1 : 2], 1);
    e = launch_mixed(T, s.op, m, p, pdls[uf ? 1 : 2]->data, bad, bi, st);
  } else if (s.family == FAM_UFUNC) {
    // the binop kernels with the second operand a zero-stride view of a's
    // first element, which the ufunc ignores (and the bad test skips)
    int which[3] = {0, 0, 1};
    Dimset d = brc_dimset(b, which, 3);
    for (int k = 0; k < d.n; k++) d.inc[1][k] = 0;
    if (!canonicalise(d, 2)) return ok();
    EwPlan p;
    if (!fill_ew(p, d)) return make_err(PDLHIP_EFATAL, "pdlhip: %s: too many distinct dims", s.name.c_str());
    p.off[0] = b.offs[0];
    p.off[1] = b.offs[0];
    p.off[2] = b.offs[1];
    BadInfo bi[3] = {badinfo(pdls[0], 1), badinfo(pdls[0], 0), badinfo(pdls[1], 1)};
    e = binop_dispatch(T, s.op, p, pdls[0]->data, pdls[0]->data, pdls[1]->data, bad, bi, st);
  } else if (s.family == FAM_BADOP) {
    const bool mask = s.op == B_SETBADIF;
    const int out = mask ? 2 : 1;
    int which[3] = {0, mask ? 1 : 0, out};
    Dimset d = brc_dimset(b, which, 3);
    if (!mask)
      for (int k = 0; k < d.n; k++) d.inc[1][k] = 0;
    if (!canonicalise(d, 2)) return ok();
    EwPlan p;
    if (!fill_ew(p, d)) return make_err(PDLHIP_EFATAL, "pdlhip: %s: too many distinct dims", s.name.c_str());
    p.off[0] = b.offs[0];
    p.off[1] = mask ? b.offs[1] : b.offs[0];
    p.off[2] = b.offs[out];
    // a double OtherPar arrives as its IEEE bit pattern
    double dval = 0.0;
    if (other && (s.op == B_SETVALTOBAD || s.op == B_SETBADTOVAL)) std::memcpy(&dval, &other[0], sizeof dval);
    const bool found = s.op == B_SETNANTOBAD || s.op == B_SETINFTOBAD || s.op == B_SETNONFINITETOBAD;
    int *flag = nullptr;
    if (found) {
      // the output's bad flag depends on the data (read back below)
      if (capturing())
        return make_err(PDLHIP_EUSERERROR,
                        "pdlhip: %s sets its output's bad flag from the data and cannot be captured in a graph",
                        s.name.c_str());
      flag = (int *)dev_alloc(sizeof(int));
      if (!flag) return oom_err(s.name.c_str());
      hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, st, flag, 1);
    }
    int32_t mbv = 0;
    if (mask) {
      const pdlhip_value mv = badinfo(pdls[1], 1).bv;
      std::memcpy(&mbv, &mv, sizeof mbv);
    }
    e = launch_badop(T, s.op, p, pdls[0]->data, mask ? pdls[1]->data : nullptr, pdls[out]->data,
                     badinfo(pdls[0], 1), mbv, badinfo(pdls[out], 1).bv, dval, bad, flag, st);
    // $PDLSTATESETBAD / $PDLSTATESETGOOD of the reference's code
    if (s.op == B_SETBADIF || s.op == B_SETVALTOBAD) pdls[out]->state |= PDLHIP_BADVAL;
    if (s.op == B_SETBADTONAN || s.op == B_SETBADTOVAL) pdls[out]->state &= ~PDLHIP_BADVAL;
    if (flag) {
      int h = 0;
      if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      dev_free(flag);
      if (h) pdls[out]->state |= PDLHIP_BADVAL;
    }
  } else if (s.family == FAM_ASSGN) {
    int which[3] = {0, 1, 1};
    Dimset d = brc_dimset(b, which, 3);
    if (!canonicalise(d, 2)) return ok();
    EwPlan p;
    if (!fill_ew(p, d)) return make_err(PDLHIP_EFATAL, "pdlhip: assgn: too many distinct dims");
    p.off[0] = b.offs[0];
    p.off[2] = b.offs[1];
    for (int k = 0; k < p.nd; k++) p.inc[1][k] = 0;
    BadInfo sb = badinfo(pdls[0], 1);
    BadInfo db = badinfo(pdls[1], 1);
    e = launch_convert(T, T, p, pdls[0]->data, pdls[1]->data, bad, sb, db.bv, st);
  } else if (s.family == FAM_RED || s.family == FAM_NIND) {
    int np = s.npdls;
    int which[MAXP] = {0, 1, 2, 3, 4, 5};
    Dimset d = brc_dimset(b, which, np);
    RedPlan p;
    std::memset(&p, 0, sizeof p);
    p.npdl = np;
    p.n = b.ind_sizes[0];
    p.inc_n[0] = b.inc_sizes[0][0];
    p.inc_n[1] = (s.realdims[1] == 1) ? b.inc_sizes[1][0] : 0;
    if (s.family == FAM_NIND) { p.m = b.ind_sizes[1]; p.inc_m = b.inc_sizes[1][0]; p.inc_n[1] = 0; }
    bool nonempty = canonicalise(d, 1);
    if (d.n > MAXK) return make_err(PDLHIP_EFATAL, "pdlhip: %s: too many distinct dims", s.name.c_str());
    p.nd = nonempty ? d.n : 0;
    p.nout = nonempty ? 1 : 0;
    for (int k = 0; k < p.nd; k++) {
      p.dims[k] = d.dims[k];
      p.nout *= d.dims[k];
      for (int q = 0; q < np; q++) p.inc[q][k] = d.inc[q][k];
    }
    for (int q = 0; q < np; q++) p.off[q] = b.offs[q];
    if (g_noalloc) {
      p.scratch = g.noalloc_scratch;
      p.scratch_bytes = (int64_t)NOALLOC_BYTES;
    }
    void *outs[4] = {nullptr, nullptr, nullptr, nullptr};
    pdlhip_value obv[4];
    std::memset(obv, 0, sizeof obv);
    for (int q = 1; q < np && q <= 4; q++) {
      outs[q - 1] = pdls[q]->data;
      obv[q - 1] = badinfo(pdls[q], 1).bv;
    }
    BadInfo ba = badinfo(pdls[0], 1);
    int *flag = nullptr;
    if (s.family == FAM_NIND && bad && p.nout > 0) {
      // the output's bad flag depends on the data (read back below)
      if (capturing())
        return make_err(PDLHIP_EUSERERROR,
                        "pdlhip: %s with bad values sets its output's bad flag from the data and cannot be "
                        "captured in a graph",
                        s.name.c_str());
      flag = (int *)dev_alloc(sizeof(int));
      if (!flag) return oom_err(s.name.c_str());
      hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, st, flag, 1);
    }
    e = reduce_dispatch(T, s.op, pdls[1]->datatype, p, pdls[0]->data, outs, bad, ba, obv, flag, st);
    // state flags the reference sets from inside readdata
    bool minmax = (s.op == R_MIN || s.op == R_MAX || s.op == R_MIN_IND || s.op == R_MAX_IND || s.op == R_MINMAXIMUM);
    if (minmax && p.n == 0 && p.nout > 0)
      for (int q = 1; q < np; q++) pdls[q]->state |= PDLHIP_BADVAL;  // Ufunc.pd:463-464
    if (s.family == FAM_NIND && p.nout > 0) {
      pdls[1]->state &= ~PDLHIP_BADVAL;  // $PDLSTATESETGOOD(c)
      int h = 0;
      if (flag) {
        if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        dev_free(flag);
      }
      if (h) pdls[1]->state |= PDLHIP_BADVAL;
    }
  } else if (s.family == FAM_MM) {
    int which[3] = {0, 1, 2};
    Dimset d = brc_dimset(b, which, 3);
    MmPlan p;
    std::memset(&p, 0, sizeof p);
    p.T = b.ind_sizes[0]; p.H = b.ind_sizes[1]; p.W = b.ind_sizes[2];
    p.a_t = b.inc_sizes[0][0]; p.a_h = b.inc_sizes[0][1];
    p.b_w = b.inc_sizes[1][0]; p.b_t = b.inc_sizes[1][1];
    p.c_w = b.inc_sizes[2][0]; p.c_h = b.inc_sizes[2][1];
    p.atdi = pdls[0]->ndims > 0 ? pdls[0]->incs[0] : 0;  // PDL_REPRINCS(a)[0]
    p.btdi = pdls[1]->ndims > 1 ? pdls[1]->incs[1] : 0;  // PDL_REPRINCS(b)[1]
    bool nonempty = canonicalise(d, 2);
    if (d.n > MAXK) return make_err(PDLHIP_EFATAL, "pdlhip: matmult: too many distinct dims");
    p.nd = nonempty ? d.n : 0;
    p.nbatch = nonempty ? 1 : 0;
    for (int k = 0; k < p.nd; k++) {
      p.dims[k] = d.dims[k];
      p.nbatch *= d.dims[k];
      for (int q = 0; q < 3; q++) p.inc[q][k] = d.inc[q][k];
    }
    for (int q = 0; q < 3; q++) p.off[q] = b.offs[q];
    BadInfo bi[3] = {badinfo(pdls[0], 1), badinfo(pdls[1], 1), badinfo(pdls[2], 1)};
    if (g.defer && !bad && (g.defer->T < 0 || g.defer->T == T)) {
      // a list item: record one product per broadcast position, launch later
      g.defer->T = T;
      g.defer->fused = IsFloatCode(T) && g.mm_mode != PDLHIP_MM_PDLORDER;
      for (int64_t bidx = 0; bidx < p.nbatch; bidx++) {
        MmItem it;
        it.a = pdls[0]->data; it.b = pdls[1]->data; it.c = pdls[2]->data;
        it.H = p.H; it.W = p.W; it.T = p.T;
        it.a_t = p.a_t; it.a_h = p.a_h; it.b_w = p.b_w; it.b_t = p.b_t; it.c_w = p.c_w; it.c_h = p.c_h;
        it.oa = p.off[0]; it.ob = p.off[1]; it.oc = p.off[2];
        int64_t r = bidx;
        for (int d = 0; d < p.nd; d++) {
          int64_t i = r % p.dims[d];
          r /= p.dims[d];
          it.oa += i * p.inc[0][d]; it.ob += i * p.inc[1][d]; it.oc += i * p.inc[2][d];
        }
        g.defer->items.push_back(it);
      }
      return ok();
    }
    if (g.defer) g.defer->missed++;
    // the autotuner's block shape for float products on the matrix cores
    // (128x128 or 256x256) and engine for mid-size double products (SIMT 1
    // or FP64 matrix cores 2), timed into a scratch output
    const bool mm16 = T == PDLHIP_F && (g.mm_mode == PDLHIP_MM_SPLIT16 || g.mm_mode == PDLHIP_MM_FP8);
    const int64_t nb128 = ((p.H + 127) / 128) * ((p.W + 127) / 128);
    const bool f64mid = T == PDLHIP_D && g.mm_mode != PDLHIP_MM_PDLORDER && g.mm_mode != PDLHIP_MM_FP32 &&
                        p.nbatch == 1 && p.H >= 64 && p.W >= 64 && p.T >= 16 && nb128 >= 16 && nb128 < 256;
    if (!bad && p.nbatch == 1 && ((mm16 && p.H >= 512 && p.W >= 512 && p.T >= 32) || f64mid)) {
      char fam[160];
      snprintf(fam, sizeof fam, "matmult:%s:%s:h2^%d:w2^%d:t2^%d", type_name(T),
               g.mm_mode == PDLHIP_MM_FP8 ? "fp8" : "default", ilog2(p.H), ilog2(p.W), ilog2(p.T));
      std::vector<int> cands = mm16 ? std::vector<int>{128, 256} : std::vector<int>{1, 2};
      const int dflt = mm16 ? ((p.H >= 1024 && p.W >= 1024) ? 256 : 128) : (nb128 >= 64 ? 2 : 1);
      void *tc = nullptr;
      auto run_one = [&](int cand) -> hipError_t {
        if (!tc) tc = dev_alloc((size_t)p.H * p.W * type_size(T));
        if (!tc) return hipErrorOutOfMemory;
        MmPlan q = p;
        q.variant = cand;
        q.off[2] = 0;
        q.c_w = 1;
        q.c_h = p.W;
        return matmult_dispatch(T, q, pdls[0]->data, pdls[1]->data, tc, false, bi, g.mm_mode, st);
      };
      p.variant = tune(fam, cands, dflt, run_one);
      if (tc) dev_free(tc);
    }
    e = matmult_dispatch(T, p, pdls[0]->data, pdls[1]->data, pdls[2]->data, bad, bi, g.mm_mode, st);
  }
  if (e != hipSuccess) return hip_err(e, s.name.c_str());
  (void)other;
  return ok();
}

__global__ void fill_u64(unsigned long long *p, unsigned long long v) { p[threadIdx.x] = v; }

// Slices.pd index / index1d / index2d: Code (gather), or BackCode (scatter
// into a) under pdlhip_writebackdata; a $CROAK position found by the kernel
// becomes PDL's message, worded from the index read back at that position
static pdlhip_error readdata_index(const Sig &s, int T, bool bad, const Brc &b, pdlhip_pdl **pdls) {
  if (capturing())
    return make_err(PDLHIP_EUSERERROR, "pdlhip: %s checks its indices on the host and cannot be captured in a graph",
                    s.name.c_str());
  IxPlan ip;
  std::memset(&ip, 0, sizeof ip);
  ip.op = s.op;
  ip.back = g_writeback ? 1 : 0;
  ip.bad = bad ? 1 : 0;
  const int ci = s.npdls - 1;
  const int slot[4] = {0, 1, s.op == 2 ? 2 : -1, ci};
  ip.npos = 1;
  for (int k = 0; k < b.ndims; k++) {
    if (b.dims[k] == 1) continue;
    ip.dims[ip.nd] = b.dims[k];
    for (int q = 0; q < 4; q++) ip.inc[q][ip.nd] = slot[q] >= 0 ? b.incs[k * b.npdls + slot[q]] : 0;
    ip.npos *= b.dims[k];
    ip.nd++;
  }
  for (int q = 0; q < 4; q++) ip.off[q] = slot[q] >= 0 ? b.offs[slot[q]] : 0;
  ip.m = s.op == 1 ? b.ind_sizes[1] : 1;
  ip.na = b.ind_sizes[0];
  ip.nb2 = s.op == 2 ? b.ind_sizes[1] : 0;
  ip.ia0 = b.inc_sizes[0][0];
  ip.ia1 = s.op == 2 ? b.inc_sizes[0][1] : 0;
  ip.iim = s.op == 1 ? b.inc_sizes[1][0] : 0;
  ip.icm = s.op == 1 ? b.inc_sizes[2][0] : 0;
  for (int k = 0; k < (s.op == 2 ? 2 : 1); k++) {
    const pdlhip_value v = badinfo(pdls[1 + k], 1).bv;
    std::memcpy(&ip.ind_bv[k], &v, sizeof(int64_t));
  }
  if (ip.npos == 0 || ip.m == 0) return ok();
  const pdlhip_pdl *a = pdls[0];
  long long *owner = nullptr;
  if (ip.back) {
    int64_t lo = a->offs, hi = a->offs;
    for (int d = 0; d < a->ndims; d++) {
      if (a->dims[d] <= 0) return ok();
      const int64_t ext = (a->dims[d] - 1) * a->incs[d];
      if (ext < 0) lo += ext; else hi += ext;
    }
    ip.a_lo = lo;
    ip.a_span = hi - lo + 1;
    owner = (long long *)dev_alloc((size_t)ip.a_span * sizeof(long long));
    if (!owner) return oom_err(s.name.c_str());
  }
  unsigned long long *err = (unsigned long long *)dev_alloc(sizeof(unsigned long long));
  if (!err) {
    dev_free(owner);
    return oom_err(s.name.c_str());
  }
  hipStream_t st = g.stream;
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, st, err, ~0ull);
  hipError_t e = launch_index(T, ip, pdls[0]->data, (const int64_t *)pdls[1]->data,
                              s.op == 2 ? (const int64_t *)pdls[2]->data : nullptr, pdls[ci]->data,
                              badinfo(pdls[ci], 1).bv, err, owner, st);
  unsigned long long pos = ~0ull;
  if (e == hipSuccess) e = hipMemcpyAsync(&pos, err, sizeof pos, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dev_free(err);
  dev_free(owner);
  if (e != hipSuccess) return hip_err(e, s.name.c_str());
  if (pos == ~0ull) return ok();
  // the loop position PDL croaks at: read its index (or indices)
  int64_t m = (int64_t)pos % ip.m, r = (int64_t)pos / ip.m, o1 = ip.off[1], o2 = ip.off[2];
  for (int d = 0; d < ip.nd; d++) {
    const int64_t i = r % ip.dims[d];
    r /= ip.dims[d];
    o1 += i * ip.inc[1][d];
    o2 += i * ip.inc[2][d];
  }
  int64_t x[2] = {0, 0};
  e = hipMemcpy(&x[0], (const int64_t *)pdls[1]->data + o1 + m * ip.iim, sizeof(int64_t), hipMemcpyDeviceToHost);
  if (e == hipSuccess && s.op == 2)
    e = hipMemcpy(&x[1], (const int64_t *)pdls[2]->data + o2, sizeof(int64_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(e, s.name.c_str());
  if (s.op == 0)
    return make_err(PDLHIP_EUSERERROR, "Error in index:invalid index %lld (valid range 0..%lld)", (long long)x[0],
                    (long long)(ip.na - 1));
  if (s.op == 1)
    return make_err(PDLHIP_EUSERERROR, "Error in index1d:invalid index %lld at pos %lld (valid range 0..%lld)",
                    (long long)x[0], (long long)m, (long long)(ip.na - 1));
  if ((bad && x[0] == ip.ind_bv[0]) || x[0] < 0 || x[0] >= ip.na)
    return make_err(PDLHIP_EUSERERROR, "Error in index2d:invalid x-index %lld (valid range 0..%lld)",
                    (long long)x[0], (long long)(ip.na - 1));
  return make_err(PDLHIP_EUSERERROR, "Error in index2d:invalid y-index %lld (valid range 0..%lld)", (long long)x[1],
                  (long long)(ip.nb2 - 1));
}

// dense temp of type `t` shaped like `p` (the converted pdl of
// pdl__get_convertedpdl_recprotect)
static pdlhip_error make_converted(const pdlhip_pdl *p, int t, pdlhip_pdl *out, bool copy_in) {
  *out = *p;
  out->datatype = t;
  out->state = p->state & PDLHIP_BADVAL;
  out->has_badvalue = 0;
  out->offs = 0;
  int64_t n = 1;
  for (int j = 0; j < p->ndims; j++) { out->incs[j] = n; n *= p->dims[j]; }
  out->data = nullptr;
  if (g.plan_only) return ok();
  out->data = dev_alloc((size_t)std::max<int64_t>(n, 1) * type_size(t));
  if (!out->data) return oom_err("converted copy");
  if (copy_in && n > 0) {
    Dimset d;
    d.np = 3;
    d.n = p->ndims;
    for (int j = 0; j < p->ndims; j++) {
      d.dims[j] = p->dims[j];
      d.inc[0][j] = p->incs[j];
      d.inc[1][j] = 0;
      d.inc[2][j] = out->incs[j];
    }
    if (!canonicalise(d, 2)) return ok();
    EwPlan e;
    if (!fill_ew(e, d)) return make_err(PDLHIP_EFATAL, "pdlhip: too many distinct dims");
    e.off[0] = p->offs;
    e.off[2] = 0;
    BadInfo sb = badinfo(p, 1);
    hipError_t he = launch_convert(p->datatype, t, e, p->data, out->data, (p->state & PDLHIP_BADVAL) != 0, sb,
                                   default_badval(t), g.stream);
    if (he != hipSuccess) return hip_err(he, "convert");
  }
  return ok();
}

static pdlhip_error write_back(const pdlhip_pdl *tmp, pdlhip_pdl *p) {
  if (g.plan_only) return ok();
  Dimset d;
  d.np = 3;
  d.n = p->ndims;
  for (int j = 0; j < p->ndims; j++) {
    d.dims[j] = p->dims[j];
    d.inc[0][j] = tmp->incs[j];
    d.inc[1][j] = 0;
    d.inc[2][j] = p->incs[j];
  }
  if (!canonicalise(d, 2)) return ok();
  EwPlan e;
  if (!fill_ew(e, d)) return make_err(PDLHIP_EFATAL, "pdlhip: too many distinct dims");
  e.off[0] = 0;
  e.off[2] = p->offs;
  BadInfo sb = badinfo(tmp, 1);
  hipError_t he = launch_convert(tmp->datatype, p->datatype, e, tmp->data, p->data, (tmp->state & PDLHIP_BADVAL) != 0,
                                 sb, badinfo(p, 1).bv, g.stream);
  if (he != hipSuccess) return hip_err(he, "convert back");
  if (tmp->state & PDLHIP_BADVAL) p->state |= PDLHIP_BADVAL;
  return ok();
}

// pdl_run_<op>: everything PDL does between the XS call and readdata
static pdlhip_error run(const Sig &s, pdlhip_pdl **user, int npdls, const int64_t *other) {
  if (npdls != s.npdls)
    return make_err(PDLHIP_EUSERERROR, "pdlhip: %s takes %d ndarrays, got %d", s.name.c_str(), s.npdls, npdls);
  RETURN_IF(ensure_init());
  pdlhip_pdl *pdls[MAXP];
  for (int i = 0; i < npdls; i++) {
    if (!user[i]) return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: ndarray %d is NULL", s.name.c_str(), i);
    if (!(user[i]->state & PDLHIP_NOMYDIMS) && (user[i]->ndims < 0 || user[i]->ndims > PDLHIP_MAXDIMS))
      return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: ndarray %d has %d dims", s.name.c_str(), i, user[i]->ndims);
    pdls[i] = user[i];
  }
  if ((s.family == FAM_BIOP || s.family == FAM_BIFUNC) && other && other[0]) std::swap(pdls[0], pdls[1]);
  bool inputs_bad = false;
  for (int i = 0; i < s.nparents; i++)
    if (pdls[i]->state & PDLHIP_BADVAL) inputs_bad = true;
  for (int i = 0; i < s.nparents; i++)
    if (pdls[i]->state & PDLHIP_NOMYDIMS)
      return make_err(PDLHIP_EUSERERROR, "Error in %s: input parameter '%s' is null", s.name.c_str(),
                      s.par_names[i].c_str());
  for (int i = s.nparents; i < npdls; i++)
    if ((pdls[i]->state & PDLHIP_NOMYDIMS) && !(s.flags[i] & P_CREAT))
      return make_err(PDLHIP_EUSERERROR, "Error in %s: io parameter '%s' is null", s.name.c_str(),
                      s.par_names[i].c_str());
  const int T = select_transtype(s, pdls);
  if (!type_on_device(T) && !(type_is_cplx(T) && cplx_op(s)))
    return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: type %s is not supported on the device", s.name.c_str(),
                    type_name(T));
  if (type_is_cplx(T) && inputs_bad)
    return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: bad values in %s ndarrays are not supported on the device",
                    s.name.c_str(), type_name(T));
  bool creating[MAXP] = {false};
  int req[MAXP];
  for (int i = 0; i < npdls; i++) {
    req[i] = adjust_from_trans(T, s.flags[i], s.par_types[i]);
    if (pdls[i]->state & PDLHIP_NOMYDIMS) {
      creating[i] = true;
      pdls[i]->datatype = req[i];
      pdls[i]->has_badvalue = 0;
    }
  }
  // converted copies (pdl__type_convert); the same ndarray used as input and
  // output (inplace) converts once
  pdlhip_pdl conv[MAXP];
  bool converted[MAXP] = {false};
  pdlhip_pdl *orig[MAXP];
  for (int i = 0; i < npdls; i++) orig[i] = pdls[i];
  pdlhip_error err = ok();
  // elementwise ops read an input of another type directly and convert it
  // in registers (k_mixed.hip): no copy, unless it is also an output; an
  // input of a few elements (a Perl scalar such as the 2 in x * 2) is
  // converted into a tiny copy instead, so the op keeps the same-type kernels
  const bool ew = s.family == FAM_BIOP || s.family == FAM_BIFUNC || s.family == FAM_UFUNC;
  for (int i = npdls - 1; i >= 0 && !err.error; i--) {
    if (creating[i] || pdls[i]->datatype == req[i]) continue;
    if (!type_on_device(pdls[i]->datatype) && !(type_is_cplx(pdls[i]->datatype) && cplx_op(s))) {
      err = make_err(PDLHIP_EUSERERROR, "pdlhip: %s: type %s is not supported on the device", s.name.c_str(),
                     type_name(pdls[i]->datatype));
      break;
    }
    if (ew && i < s.nparents && !type_is_cplx(T) && !type_is_cplx(pdls[i]->datatype)) {
      bool inout = false;
      for (int j = s.nparents; j < npdls; j++) inout = inout || orig[j] == orig[i];
      int64_t nel = 1;
      for (int d = 0; d < pdls[i]->ndims; d++) nel *= pdls[i]->dims[d];
      if (!inout && nel > 64) continue;
    }
    int dup = -1;
    for (int j = i + 1; j < npdls; j++)
      if (orig[j] == orig[i] && converted[j]) dup = j;
    if (dup >= 0) { pdls[i] = &conv[dup]; continue; }
    err = make_converted(orig[i], req[i], &conv[i], true);
    if (!err.error) { pdls[i] = &conv[i]; converted[i] = true; }
  }
  Brc b;
  for (int k = 0; k < 4; k++) b.ind_sizes[k] = -1;
  // dims setters run first with a load-only check (REDODIMS, pdlapi.c:41-75)
  if (!err.error && s.family == FAM_NIND) {
    err = dim_checks(s, pdls, creating, true, 0, b.ind_sizes);
    if (!err.error) {
      b.ind_sizes[1] = other ? other[0] : 0;
      if (b.ind_sizes[1] > b.ind_sizes[0])
        err = make_err(PDLHIP_EUSERERROR, "Error in %s:m_size > n_size", s.name.c_str());
    }
  }
  if (!err.error) err = redodims(s, pdls, creating, b);
  if (!err.error) err = collapse_repeated_writes(s, pdls, b);
  if (!err.error) {
    if (inputs_bad)
      for (int i = s.nparents; i < npdls; i++) pdls[i]->state |= PDLHIP_BADVAL;
    for (int i = 0; i < npdls && !err.error; i++) {
      if (!creating[i]) continue;
      int64_t n = 1;
      for (int j = 0; j < pdls[i]->ndims; j++) n *= pdls[i]->dims[j];
      pdls[i]->data = nullptr;
      if (!g.plan_only) {
        pdls[i]->data = dev_alloc((size_t)std::max<int64_t>(n, 1) * type_size(pdls[i]->datatype));
        if (!pdls[i]->data) err = oom_err(s.name.c_str());
      }
      pdls[i]->state &= ~PDLHIP_NOMYDIMS;
    }
  }
  // a list item whose operands needed converted copies runs now: its
  // temporaries are released below, before a deferred launch would read them
  Deferred *defer_saved = g.defer;
  bool any_conv = false;
  for (int i = 0; i < npdls; i++) any_conv = any_conv || converted[i];
  if (any_conv && g.defer) {
    g.defer->missed++;
    g.defer = nullptr;
  }
  if (!err.error) err = readdata(s, T, inputs_bad, b, pdls, other);
  g.defer = defer_saved;
  for (int i = s.nparents; i < npdls && !err.error; i++) {
    if (!converted[i]) continue;
    err = write_back(&conv[i], orig[i]);
    orig[i]->state |= conv[i].state & PDLHIP_BADVAL;
  }
  for (int i = 0; i < npdls; i++)
    if (converted[i]) dev_free(conv[i].data);
  return err;
}

// ------------------------------------------------------------------ autotuner
// Measured choices of kernel variants per shape class, kept in a text file
// per GPU architecture so that later processes reuse them: one line per
// entry, "<key> <choice>", the key naming the kernel family, element type,
// shape class and architecture (e.g. "list:float:mfma1:dim2^8:items2^12:gfx950").
// The file is $PDLHIP_TUNE_FILE, else $XDG_CACHE_HOME/pdlhip/tune-<arch>.txt,
// else $HOME/.cache/pdlhip/tune-<arch>.txt; PDLHIP_TUNE=0 turns timing off
// (each family's default rule applies), PDLHIP_TUNE_VERBOSE=1 prints each
// candidate's time to stderr.  A miss times every candidate with
// HIP events on scratch outputs (never the caller's), keeps the fastest and
// appends it to the file.  Not during a graph capture.
struct Tuner {
  std::mutex mu;
  bool loaded = false, enabled = true, verbose = false;
  std::string arch, path;
  std::map<std::string, int> choice;
  int64_t measured = 0, reused = 0;
};
static Tuner g_tuner;

static void mkdirs_for(const std::string &file) {
  for (size_t i = 1; i < file.size(); i++)
    if (file[i] == '/') (void)mkdir(file.substr(0, i).c_str(), 0755);
}

static void tuner_load() {
  if (g_tuner.loaded) return;
  g_tuner.loaded = true;
  hipDeviceProp_t prop;
  g_tuner.arch = "gfx";
  if (g.device >= 0 && hipGetDeviceProperties(&prop, g.device) == hipSuccess) {
    g_tuner.arch = prop.gcnArchName;
    const size_t colon = g_tuner.arch.find(':');
    if (colon != std::string::npos) g_tuner.arch.resize(colon);
  }
  const char *off = getenv("PDLHIP_TUNE");
  g_tuner.enabled = !(off && off[0] == '0');
  const char *verbose = getenv("PDLHIP_TUNE_VERBOSE");
  g_tuner.verbose = verbose && verbose[0] && verbose[0] != '0';
  const char *f = getenv("PDLHIP_TUNE_FILE");
  if (f && *f) {
    g_tuner.path = f;
  } else {
    const char *xdg = getenv("XDG_CACHE_HOME"), *home = getenv("HOME");
    std::string dir = (xdg && *xdg) ? std::string(xdg) : (home && *home) ? std::string(home) + "/.cache" : "";
    if (!dir.empty()) g_tuner.path = dir + "/pdlhip/tune-" + g_tuner.arch + ".txt";
  }
  if (g_tuner.path.empty()) return;
  FILE *fp = fopen(g_tuner.path.c_str(), "r");
  if (!fp) return;
  char key[512];
  int v;
  while (fscanf(fp, "%511s %d", key, &v) == 2) g_tuner.choice[key] = v;
  fclose(fp);
}

// the variant for `key` among `cands`: cached, else timed with `run_one`
// (which enqueues one run of a candidate; it is run twice, the second timed)
static int tune(const std::string &family, const std::vector<int> &cands, int dflt,
                const std::function<hipError_t(int)> &run_one) {
  std::lock_guard<std::mutex> lk(g_tuner.mu);
  tuner_load();
  const std::string key = family + ":" + g_tuner.arch;
  auto it = g_tuner.choice.find(key);
  if (it != g_tuner.choice.end() && std::find(cands.begin(), cands.end(), it->second) != cands.end()) {
    g_tuner.reused++;
    return it->second;
  }
  if (!g_tuner.enabled || g.plan_only || capturing() || cands.size() < 2) return dflt;
  int best = dflt;
  float best_ms = 1e30f;
  for (int c : cands) {
    if (run_one(c) != hipSuccess) continue;   // warm-up
    (void)hipEventRecord(g.ev0, g.stream);
    hipError_t e = run_one(c);
    (void)hipEventRecord(g.ev1, g.stream);
    float ms = 0;
    if (e == hipSuccess) e = hipEventSynchronize(g.ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, g.ev0, g.ev1);
    if (e == hipSuccess && ms < best_ms) { best_ms = ms; best = c; }
    if (g_tuner.verbose) fprintf(stderr, "pdlhip tune %s: candidate %d %.4f ms\n", key.c_str(), c, e == hipSuccess ? ms : -1.0f);
  }
  (void)hipGetLastError();
  g_tuner.measured++;
  g_tuner.choice[key] = best;
  if (!g_tuner.path.empty()) {
    mkdirs_for(g_tuner.path);
    if (FILE *fp = fopen(g_tuner.path.c_str(), "a")) {
      fprintf(fp, "%s %d\n", key.c_str(), best);
      fclose(fp);
    }
  }
  return best;
}

static int ilog2(int64_t v) {
  int r = 0;
  while (v > 1) { v >>= 1; r++; }
  return r;
}

// A planned matmult list: the items (one product per broadcast position of
// every list entry) and both tile decompositions in HBM, the tile picked by
// the autotuner.  A launch allocates nothing.
struct ListH {
  Deferred d;
  std::vector<int3> work32, work64;
  MmItem *d_items = nullptr;
  unsigned int *d_counter = nullptr;
  int3 *d_w32 = nullptr, *d_w64 = nullptr;
  int tile = 64;        // 32 / 64: SIMT tiles; 0: the matrix-core kernel (64x64 work items)
  bool mfma = false;    // float/double list in a default mode whose items fit 32-bit offsets
};

static void list_release(ListH &h) {
  dev_free(h.d_items);
  dev_free(h.d_counter);
  dev_free(h.d_w32);
  dev_free(h.d_w64);
  h.d_items = nullptr; h.d_counter = nullptr; h.d_w32 = nullptr; h.d_w64 = nullptr;
}

static hipError_t list_launch(const ListH &h, int tile, bool each) {
  const std::vector<int3> &w = tile == 32 ? h.work32 : h.work64;   // the MFMA kernel takes 64x64 items
  const int nblocks = (int)std::max<size_t>(std::min<size_t>(w.size(), 2048), 1);
  return launch_matmult_list(h.d.T, h.d_items, h.d.items.data(), (int)h.d.items.size(), tile == 32 ? h.d_w32 : h.d_w64,
                             (unsigned)w.size(), h.d_counter, tile, h.d.fused, each, nblocks, g.stream);
}

// plan every entry through pdl_run_matmult's checks (outputs created as
// PDL would), then build the work lists; `strict`: refuse entries that would
// have to run at once instead of joining the list
static pdlhip_error list_build(int n, pdlhip_pdl **a, pdlhip_pdl **b, pdlhip_pdl **c, bool strict, ListH &h) {
  const Sig *s = find_sig("matmult");
  g.defer = &h.d;
  pdlhip_error err = ok();
  int odd = 0;   // entries that cannot join (found from their types and flags, also when only planning)
  for (int i = 0; i < n && !err.error; i++) {
    pdlhip_pdl *v[3] = {a[i], b[i], c[i]};
    err = run(*s, v, 3, nullptr);
    if (!err.error && (c[i]->datatype != c[0]->datatype || a[i]->datatype != c[i]->datatype ||
                       b[i]->datatype != c[i]->datatype || ((a[i]->state | b[i]->state) & PDLHIP_BADVAL)))
      odd++;
  }
  g.defer = nullptr;
  if (err.error) return err;
  h.d.missed = std::max(h.d.missed, odd);
  if (strict && h.d.missed)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_matmult_list_prepare: %d entries need bad values, a type conversion "
                    "or another type than the first; convert them or run them one by one", h.d.missed);
  if (h.d.items.empty() || g.plan_only) return ok();
  const int nit = (int)h.d.items.size();
  int64_t maxdim = 1;
  for (auto &it : h.d.items) maxdim = std::max(maxdim, std::max(it.H, it.W));
  // work items, largest inner dimension first
  std::vector<int> order(nit);
  for (int i = 0; i < nit; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return h.d.items[x].T > h.d.items[y].T; });
  // 8 bins, one per XCD (the kernels' list_next): each product, largest
  // inner dimension first, goes whole to the bin with the least work so far
  // (tiles x k); the bins lie back to back, bounds[x] .. bounds[x + 1]
  auto build_work = [&](int tile, std::vector<int3> &work, unsigned *bounds) {
    std::vector<std::vector<int3>> bins(8);
    std::vector<double> load(8, 0.0);
    for (int i : order) {
      const MmItem &it = h.d.items[i];
      int64_t ntm = (it.H + tile - 1) / tile, ntn = (it.W + tile - 1) / tile;
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[x] += (double)(ntm * ntn) * (double)std::max<int64_t>(it.T, 1);
      for (int64_t tm = 0; tm < ntm; tm++)
        for (int64_t tn = 0; tn < ntn; tn++) bins[x].push_back(make_int3(i, (int)tm, (int)tn));
    }
    work.clear();
    for (int x = 0; x < 8; x++) {
      bounds[x] = (unsigned)work.size();
      work.insert(work.end(), bins[x].begin(), bins[x].end());
    }
    bounds[8] = (unsigned)work.size();
  };
  // the device block: 8 bin counters 128 bytes apart, then the 64-tile
  // list's bounds, then the 32-tile list's
  unsigned qinit[256 + 9 + 9] = {0};
  build_work(64, h.work64, qinit + 256);
  build_work(32, h.work32, qinit + 265);
  // the matrix-core kernel indexes within a product with 32-bit offsets
  auto span = [](int64_t n1, int64_t i1, int64_t n2, int64_t i2) {
    return (n1 > 0 ? n1 - 1 : 0) * (i1 < 0 ? -i1 : i1) + (n2 > 0 ? n2 - 1 : 0) * (i2 < 0 ? -i2 : i2);
  };
  h.mfma = h.d.fused && (h.d.T == PDLHIP_F || h.d.T == PDLHIP_D);
  for (const MmItem &it : h.d.items)
    if (span(it.T, it.a_t, it.H, it.a_h) >= ((int64_t)1 << 31) || span(it.T, it.b_t, it.W, it.b_w) >= ((int64_t)1 << 31) ||
        it.H >= ((int64_t)1 << 30) || it.W >= ((int64_t)1 << 30) || it.T >= ((int64_t)1 << 30))
      h.mfma = false;
  h.d_items = (MmItem *)dev_alloc(sizeof(MmItem) * nit);
  h.d_counter = (unsigned int *)dev_alloc(sizeof qinit);
  h.d_w32 = (int3 *)dev_alloc(sizeof(int3) * std::max<size_t>(h.work32.size(), 1));
  h.d_w64 = (int3 *)dev_alloc(sizeof(int3) * std::max<size_t>(h.work64.size(), 1));
  if (!h.d_items || !h.d_counter || !h.d_w32 || !h.d_w64) {
    list_release(h);
    return oom_err("matmult list");
  }
  hipError_t e = upload_now(h.d_items, h.d.items.data(), sizeof(MmItem) * nit);
  if (e == hipSuccess) e = upload_now(h.d_counter, qinit, sizeof qinit);
  if (e == hipSuccess) e = upload_now(h.d_w32, h.work32.data(), sizeof(int3) * h.work32.size());
  if (e == hipSuccess) e = upload_now(h.d_w64, h.work64.data(), sizeof(int3) * h.work64.size());
  // tile by the autotuner: 32 / 64 SIMT tiles, the FP32/FP64 matrix-core
  // kernel (0; 2: half the k per LDS stage) and, for float, the BF16
  // three-way split (1); timed on a copy
  // of the item table whose outputs all point into one scratch block
  if (e == hipSuccess) {
    std::vector<int> cands = {32, 64};
    if (h.mfma) { cands.push_back(0); cands.push_back(2); }
    if (h.mfma && h.d.T == PDLHIP_F) cands.push_back(1);
    const int dflt = h.mfma ? 0 : 64;
    char fam[160];
    snprintf(fam, sizeof fam, "list:%s:mfma%d:fused%d:dim2^%d:items2^%d", type_name(h.d.T), h.mfma ? 1 : 0,
             h.d.fused ? 1 : 0, ilog2(maxdim), ilog2(nit));
    MmItem *t_items = nullptr;
    void *t_c = nullptr;
    auto run_one = [&](int cand) -> hipError_t {
      if (!t_items) {
        int64_t most = 1;
        for (auto &it : h.d.items) most = std::max(most, it.H * it.W);
        std::vector<MmItem> tv = h.d.items;
        t_c = dev_alloc((size_t)most * type_size(h.d.T));
        t_items = (MmItem *)dev_alloc(sizeof(MmItem) * nit);
        if (!t_c || !t_items) return hipErrorOutOfMemory;
        for (auto &it : tv) { it.c = t_c; it.oc = 0; it.c_w = 1; it.c_h = it.W; }
        hipError_t ue = upload_now(t_items, tv.data(), sizeof(MmItem) * nit);
        if (ue != hipSuccess) return ue;
      }
      ListH t = h;
      t.d_items = t_items;
      hipError_t le = list_launch(t, cand, false);
      t.d_items = nullptr;
      return le;
    };
    h.tile = tune(fam, cands, dflt, run_one);
    if (t_items || t_c) {
      (void)hipStreamSynchronize(g.stream);
      dev_free(t_items);
      dev_free(t_c);
    }
  }
  if (e != hipSuccess) {
    list_release(h);
    return hip_err(e, "matmult list");
  }
  return ok();
}

static pdlhip_error run_list(int n, pdlhip_pdl **a, pdlhip_pdl **b, pdlhip_pdl **c, int mode) {
  ListH h;
  RETURN_IF(list_build(n, a, b, c, false, h));
  if (h.d.items.empty() || g.plan_only) return ok();
  hipError_t e = list_launch(h, h.tile, mode == PDLHIP_LIST_EACH);
  // the host item table of a per-product launch is read at launch time only
  list_release(h);
  return e == hipSuccess ? ok() : hip_err(e, "matmult list");
}

// ------------------------------------------------------------------ ragged lists
struct RaggedH {
  bool scan = false;
  int T = -1, op = -1, n = 0;
  RagLaunch L;
  void *d_items = nullptr, *d_segs = nullptr, *d_multi = nullptr, *d_order = nullptr, *d_bstart = nullptr,
       *d_scratch = nullptr, *d_empty = nullptr;
};

static const char *op_name(int op) {
  static const char *names[] = {"plus", "mult", "minus", "divide", "gt", "lt", "le", "ge", "eq", "ne", "shiftleft",
                                "shiftright", "or2", "and2", "xor", "power", "atan2", "modulo", "spaceship"};
  return (op >= 0 && op < OP_NBINOPS) ? names[op] : "?";
}

// PDL's trans type of one binop on operands of types ta, tb
static int binop_type(int op, int ta, int tb) {
  const Sig *s = find_sig(op_name(op));
  pdlhip_pdl x, y, z;
  std::memset(&x, 0, sizeof x); std::memset(&y, 0, sizeof y); std::memset(&z, 0, sizeof z);
  x.datatype = ta; y.datatype = tb; z.state = PDLHIP_NOMYDIMS; z.datatype = -1;
  pdlhip_pdl *v[3] = {&x, &y, &z};
  return select_transtype(*s, v);
}

// the op applied per item after the chain ("none": the chain's values only)
struct Fin { const char *name; int op; bool scan; };
static const Fin *find_final(const char *name) {
  static const Fin fins[] = {
      {"none", RAG_NOREDUCE, false}, {"sumover", R_SUM, false}, {"dsumover", R_DSUM, false},
      {"average", R_AVERAGE, false}, {"daverage", R_DAVERAGE, false}, {"minimum", R_MIN, false},
      {"maximum", R_MAX, false}, {"cumusumover", R_CUMUSUM, true}, {"cumuprodover", R_CUMUPROD, true}};
  if (!name) return nullptr;
  for (const Fin &f : fins) if (!strcmp(f.name, name)) return &f;
  return nullptr;
}

// The fused program of a chain over input slots of the given types: PDL's
// trans type of every node (the chain's type is the last node's; a chain
// whose nodes differ runs each node in its own type), the left-deep form
// when the program has one and one type, and the type of the final op's
// output.
static pdlhip_error chain_setup(const char *who, int nin, const int *stype, int nnodes, const pdlhip_node *nodes,
                                const Fin &f, RaggedH &h, int *otype) {
  h.scan = f.scan;
  h.op = f.op;
  std::memset(&h.L, 0, sizeof h.L);
  RagProg &pg = h.L.prog;
  pg.nin = nin;
  pg.nnodes = nnodes;
  for (int k = 0; k < nin; k++) {
    if (!type_on_device(stype[k]))
      return make_err(PDLHIP_EUSERERROR, "%s: type %s not supported", who, type_name(stype[k]));
    pg.type[k] = stype[k];
  }
  if (f.scan && (nnodes || nin != 1))
    return make_err(PDLHIP_EUSERERROR, "%s: %s takes one input and no chain", who, f.name);
  std::vector<int> ntype(nnodes);
  for (int j = 0; j < nnodes; j++) {
    const pdlhip_node &nd = nodes[j];
    if (nd.op < 0 || nd.op >= OP_NBINOPS) return make_err(PDLHIP_EUSERERROR, "%s: node %d op", who, j);
    auto otp = [&](int x) -> int {
      if (x >= 0 && x < nin) return stype[x];
      if (x < 0 && -x - 1 < j) return ntype[-x - 1];
      return -2;
    };
    const int ta = otp(nd.x), tb = otp(nd.y);
    if (ta < -1 || tb < -1) return make_err(PDLHIP_EUSERERROR, "%s: node %d operand", who, j);
    ntype[j] = binop_type(nd.op, ta, tb);
    pg.node[j].op = nd.op; pg.node[j].x = nd.x; pg.node[j].y = nd.y;
  }
  const int T = nnodes ? ntype[nnodes - 1] : stype[0];
  // left-deep form: node 0 combines two slots, every later node the previous
  // result and one slot
  pg.lin = 1;
  pg.first = 0;
  for (int j = 0; j < nnodes && pg.lin; j++) {
    const int x = nodes[j].x, y = nodes[j].y;
    if (j == 0) {
      if (x < 0 || y < 0) pg.lin = 0;
      pg.first = x;
      pg.lslot[0] = y;
      pg.lswap[0] = 0;
    } else if (x == -j && y >= 0) {
      pg.lslot[j] = y; pg.lswap[j] = 0;
    } else if (y == -j && x >= 0) {
      pg.lslot[j] = x; pg.lswap[j] = 1;
    } else {
      pg.lin = 0;
    }
  }
  // nodes in other types than the last: the general form computes each in
  // its own type
  pg.mixed = 0;
  for (int j = 0; j < nnodes; j++) {
    pg.ntype[j] = ntype[j];
    if (!type_on_device(ntype[j]))
      return make_err(PDLHIP_EUSERERROR, "%s: node %d computes in %s, not supported on the device", who, j,
                      type_name(ntype[j]));
    if (ntype[j] != T) pg.mixed = 1;
  }
  if (pg.mixed) pg.lin = 0;
  if (!f.scan && T != PDLHIP_F && T != PDLHIP_D)
    return make_err(PDLHIP_EUSERERROR, "%s: %s of a %s chain: float or double only", who, f.name, type_name(T));
  if (f.scan && T >= PDLHIP_F)
    return make_err(PDLHIP_EUSERERROR, "%s: %s of %s: integer types only", who, f.name, type_name(T));
  h.T = T;
  if (f.op == RAG_NOREDUCE) *otype = -1;
  else if (f.op == R_DSUM || f.op == R_DAVERAGE) *otype = PDLHIP_D;
  else if (f.op == R_MIN || f.op == R_MAX) *otype = T;
  else *otype = std::max<int>(PDLHIP_L, T);
  pg.out_bv = default_badval(*otype >= 0 ? *otype : T);
  return ok();
}

// a null output is created with `type` and `dims`, as PDL creates an op's
// output; an existing one must have that type and those dims
static pdlhip_error output_for(const char *who, const char *what, int64_t idx, pdlhip_pdl *o, int type, int ndims,
                               const int64_t *dims) {
  if (!o) return make_err(PDLHIP_EUSERERROR, "%s: %s %lld is missing", who, what, (long long)idx);
  if (o->state & PDLHIP_NOMYDIMS) {
    int64_t n = 1;
    o->datatype = type;
    o->has_badvalue = 0;
    o->ndims = ndims;
    for (int d = 0; d < ndims; d++) {
      o->dims[d] = dims[d];
      o->incs[d] = n;
      n *= dims[d];
    }
    o->offs = 0;
    o->data = g.plan_only ? nullptr : dev_alloc((size_t)std::max<int64_t>(n, 1) * type_size(type));
    if (!g.plan_only && !o->data) return oom_err(who);
    o->state &= ~PDLHIP_NOMYDIMS;
    return ok();
  }
  if (o->datatype != type)
    return make_err(PDLHIP_EUSERERROR, "%s: %s %lld must exist with type %s", who, what, (long long)idx,
                    type_name(type));
  bool same = o->ndims == ndims;
  for (int d = 0; same && d < ndims; d++) same = o->dims[d] == dims[d];
  if (!same) return make_err(PDLHIP_EUSERERROR, "%s: %s %lld has the wrong shape", who, what, (long long)idx);
  return ok();
}

// segments, the scans' block assignment and the device copies of an item
// table
static pdlhip_error ragged_finish(const char *who, RaggedH &h, std::vector<RagItem> &items, RaggedH **hout) {
  const int n = (int)items.size();
  h.n = n;
  // streaming access when the chain's inputs are too large to stay cached
  size_t bytes = 0;
  for (const RagItem &it : items)
    for (int k = 0; k < h.L.prog.nin; k++)
      if (it.inc[k]) bytes += (size_t)it.len * type_size(h.L.prog.type[k]);
  h.L.prog.stream = streaming_read(bytes) ? 1 : 0;
  std::vector<RagSeg> segs;
  std::vector<int32_t> multi;
  for (int i = 0; i < n; i++) {
    RagItem &it = items[i];
    it.seg0 = (int64_t)segs.size();
    it.nseg = std::max<int64_t>(1, (it.len + RAG_SEG - 1) / RAG_SEG);
    for (int64_t k = 0; k < it.nseg; k++) {
      RagSeg sg;
      sg.item = i;
      sg.single = it.nseg == 1;
      sg.start = k * RAG_SEG;
      sg.end = std::min<int64_t>(it.len, (k + 1) * RAG_SEG);
      segs.push_back(sg);
    }
    if (it.nseg > 1 && h.op != RAG_NOREDUCE) multi.push_back(i);
  }
  h.L.nseg = (int64_t)segs.size();
  h.L.nmulti = (int64_t)multi.size();
  // scans: items to blocks longest-first onto the least loaded block (LPT);
  // each block then runs its items longest-first
  std::vector<int32_t> order, bstart;
  if (h.scan) {
    const int nblk = std::min(n, 1024);
    std::vector<int32_t> byl(n);
    for (int i = 0; i < n; i++) byl[i] = i;
    std::stable_sort(byl.begin(), byl.end(), [&](int32_t x, int32_t y) { return items[x].len > items[y].len; });
    std::vector<std::vector<int32_t>> bins(nblk);
    std::priority_queue<std::pair<int64_t, int32_t>, std::vector<std::pair<int64_t, int32_t>>,
                        std::greater<std::pair<int64_t, int32_t>>> pq;
    for (int b = 0; b < nblk; b++) pq.push({0, b});
    for (int32_t i : byl) {
      auto top = pq.top();
      pq.pop();
      bins[top.second].push_back(i);
      pq.push({top.first + items[i].len + 256, top.second});   // + per-item overhead
    }
    bstart.push_back(0);
    for (auto &bn : bins) {
      order.insert(order.end(), bn.begin(), bn.end());
      bstart.push_back((int32_t)order.size());
    }
    h.L.nblk = nblk;
  }
  if (g.plan_only) { *hout = new RaggedH(h); return ok(); }
  auto up = [&](const void *src, size_t bytes, void **dst) -> bool {
    *dst = dev_alloc(std::max<size_t>(bytes, 1));
    if (!*dst) return false;
    return upload_now(*dst, src, bytes) == hipSuccess;
  };
  bool okk = up(items.data(), sizeof(RagItem) * items.size(), &h.d_items) &&
             up(segs.data(), sizeof(RagSeg) * segs.size(), &h.d_segs) &&
             up(multi.data(), sizeof(int32_t) * multi.size(), &h.d_multi) &&
             up(order.data(), sizeof(int32_t) * order.size(), &h.d_order) &&
             up(bstart.data(), sizeof(int32_t) * bstart.size(), &h.d_bstart);
  h.d_scratch = dev_alloc(std::max<size_t>(64 * segs.size(), 64));
  h.d_empty = dev_alloc(std::max<size_t>(sizeof(int32_t) * n, 4));
  if (!okk || !h.d_scratch || !h.d_empty) {
    for (void *p : {h.d_items, h.d_segs, h.d_multi, h.d_order, h.d_bstart, h.d_scratch, h.d_empty}) dev_free(p);
    return oom_err(who);
  }
  h.L.empty = (int32_t *)h.d_empty;
  h.L.items = (const RagItem *)h.d_items;
  h.L.segs = (const RagSeg *)h.d_segs;
  h.L.multi = (const int32_t *)h.d_multi;
  h.L.order = (const int32_t *)h.d_order;
  h.L.bstart = (const int32_t *)h.d_bstart;
  h.L.scratch = h.d_scratch;
  *hout = new RaggedH(h);
  return ok();
}

static char *elem_ptr(const pdlhip_pdl *p, int64_t off) {
  return (char *)p->data + off * type_size(p->datatype);
}

static pdlhip_error ragged_prepare(int n, int nin, pdlhip_pdl **in, int nnodes, const pdlhip_node *nodes,
                                   const char *final_op, pdlhip_pdl **out, pdlhip_pdl **values, RaggedH **hout) {
  const char *who = "pdlhip_ragged_prepare";
  if (n < 0 || nin < 1 || nin > RAG_MAXIN || nnodes < 0 || nnodes > RAG_MAXNODE || (n && !in))
    return make_err(PDLHIP_EUSERERROR, "%s: bad arguments", who);
  const Fin *f = find_final(final_op);
  if (!f) return make_err(PDLHIP_EUSERERROR, "%s: unsupported final op '%s'", who, final_op ? final_op : "");
  if (f->op == RAG_NOREDUCE && !values) return make_err(PDLHIP_EUSERERROR, "%s: 'none' needs value outputs", who);
  if (f->op != RAG_NOREDUCE && n && !out) return make_err(PDLHIP_EUSERERROR, "%s: bad arguments", who);
  if (f->scan && values) return make_err(PDLHIP_EUSERERROR, "%s: %s has no chain values", who, f->name);
  // slot types and shapes, uniform over the items
  std::vector<int> stype(nin, PDLHIP_D), scalar(nin, 0);
  for (int k = 0; k < nin; k++) {
    for (int i = 0; i < n; i++) {
      const pdlhip_pdl *p = in[i * nin + k];
      if (!p || (p->state & PDLHIP_NOMYDIMS))
        return make_err(PDLHIP_EUSERERROR, "%s: item %d input %d is null", who, i, k);
      if (p->state & PDLHIP_BADVAL)
        return make_err(PDLHIP_EUSERERROR, "%s: item %d input %d has bad values; run the ops per ndarray", who, i, k);
      const int sc = p->ndims == 0 ? 1 : 0;
      if (i == 0) { stype[k] = p->datatype; scalar[k] = sc; }
      else if (stype[k] != p->datatype || scalar[k] != sc)
        return make_err(PDLHIP_EUSERERROR, "%s: input %d differs in type or rank across items", who, k);
      if (!sc && p->ndims != 1) return make_err(PDLHIP_EUSERERROR, "%s: item %d input %d is not 1-D", who, i, k);
    }
  }
  RaggedH h;
  int otype = -1;
  RETURN_IF(chain_setup(who, nin, stype.data(), nnodes, nodes, *f, h, &otype));
  h.L.prog.store = values != nullptr;
  std::vector<RagItem> items(n);
  for (int i = 0; i < n; i++) {
    RagItem &it = items[i];
    std::memset(&it, 0, sizeof it);
    int64_t len = -1;
    for (int k = 0; k < nin; k++) {
      const pdlhip_pdl *p = in[i * nin + k];
      it.ptr[k] = elem_ptr(p, p->offs);
      if (scalar[k]) { it.inc[k] = 0; continue; }
      it.inc[k] = p->incs[0];
      if (len >= 0 && p->dims[0] != len)
        return make_err(PDLHIP_EUSERERROR, "%s: item %d: lengths %lld and %lld differ", who, i, (long long)len,
                        (long long)p->dims[0]);
      len = p->dims[0];
    }
    if (len < 0) len = 1;
    it.len = len;
    if (f->op != RAG_NOREDUCE) {
      RETURN_IF(output_for(who, "output", i, out[i], otype, f->scan ? 1 : 0, &len));
      it.out = elem_ptr(out[i], out[i]->offs);
      it.out_inc = f->scan ? out[i]->incs[0] : 0;
    }
    if (values) {
      RETURN_IF(output_for(who, "value output", i, values[i], h.T, 1, &len));
      it.val = elem_ptr(values[i], values[i]->offs);
      it.val_inc = values[i]->incs[0];
    }
  }
  return ragged_finish(who, h, items, hout);
}

// A chain over ndarrays that broadcast together (implicit dims, size-1 dims
// repeat), reduced along dim 0 by `final_op` into `result` (dims 1.. of the
// broadcast) and/or kept whole in `value` (all broadcast dims).  Each row
// along dim 0 is one item of a ragged launch of equal lengths; with no
// reduction, rows that are contiguous in every operand are merged first.
static pdlhip_error chain_prepare(int nin, pdlhip_pdl **in, int nnodes, const pdlhip_node *nodes,
                                  const char *final_op, pdlhip_pdl *value, pdlhip_pdl *result, RaggedH **hout) {
  const char *who = "pdlhip_chain_prepare";
  if (nin < 1 || nin > RAG_MAXIN || nnodes < 0 || nnodes > RAG_MAXNODE || !in)
    return make_err(PDLHIP_EUSERERROR, "%s: bad arguments", who);
  const Fin *f = find_final(final_op);
  if (!f) return make_err(PDLHIP_EUSERERROR, "%s: unsupported final op '%s'", who, final_op ? final_op : "");
  if (f->op == RAG_NOREDUCE && !value) return make_err(PDLHIP_EUSERERROR, "%s: 'none' needs a value output", who);
  if (f->op != RAG_NOREDUCE && !result) return make_err(PDLHIP_EUSERERROR, "%s: %s needs a result", who, f->name);
  if (f->scan && value) return make_err(PDLHIP_EUSERERROR, "%s: %s has no chain values", who, f->name);
  std::vector<int> stype(nin);
  int nd = 1;
  for (int k = 0; k < nin; k++) {
    const pdlhip_pdl *p = in[k];
    if (!p || (p->state & PDLHIP_NOMYDIMS)) return make_err(PDLHIP_EUSERERROR, "%s: input %d is null", who, k);
    if (p->state & PDLHIP_BADVAL)
      return make_err(PDLHIP_EUSERERROR, "%s: input %d has bad values; run the ops one by one", who, k);
    stype[k] = p->datatype;
    nd = std::max(nd, p->ndims);
  }
  // broadcast dims (pdl_broadcast_dim_checks: sizes must agree or be 1)
  std::vector<int64_t> bd(nd, 1);
  for (int k = 0; k < nin; k++)
    for (int d = 0; d < in[k]->ndims; d++) {
      const int64_t s = in[k]->dims[d];
      if (s == 1) continue;
      if (bd[d] != 1 && bd[d] != s)
        return make_err(PDLHIP_EUSERERROR, "%s: mismatched broadcast dim %d: size %lld vs %lld", who, d,
                        (long long)bd[d], (long long)s);
      bd[d] = s;
    }
  RaggedH h;
  int otype = -1;
  RETURN_IF(chain_setup(who, nin, stype.data(), nnodes, nodes, *f, h, &otype));
  h.L.prog.store = value != nullptr;
  if (value) RETURN_IF(output_for(who, "value output", 0, value, h.T, nd, bd.data()));
  if (f->op != RAG_NOREDUCE) {
    if (f->scan) RETURN_IF(output_for(who, "result", 0, result, otype, nd, bd.data()));
    else RETURN_IF(output_for(who, "result", 0, result, otype, nd - 1, bd.data() + 1));
  }
  // per operand (inputs, value, result) strides over the broadcast dims; 0
  // where the operand repeats
  const int nop = nin + 2;
  std::vector<std::vector<int64_t>> inc(nop, std::vector<int64_t>(nd, 0));
  std::vector<int64_t> off(nop, 0);
  for (int k = 0; k < nin; k++) {
    off[k] = in[k]->offs;
    for (int d = 0; d < in[k]->ndims; d++) inc[k][d] = in[k]->dims[d] == 1 ? 0 : in[k]->incs[d];
  }
  if (value) {
    off[nin] = value->offs;
    for (int d = 0; d < nd; d++) inc[nin][d] = bd[d] == 1 ? 0 : value->incs[d];
  }
  if (f->op != RAG_NOREDUCE) {
    off[nin + 1] = result->offs;
    if (f->scan)
      for (int d = 0; d < nd; d++) inc[nin + 1][d] = bd[d] == 1 ? 0 : result->incs[d];
    else
      for (int d = 1; d < nd; d++) inc[nin + 1][d] = bd[d] == 1 ? 0 : result->incs[d - 1];
  }
  // no reduction: fold dim d+1 into dim d where every operand is contiguous
  // across them
  std::vector<int64_t> dims = bd;
  if (f->op == RAG_NOREDUCE) {
    for (int d = 0; d + 1 < (int)dims.size();) {
      bool merge = true;
      for (int k = 0; k < nop && merge; k++) merge = inc[k][d + 1] == inc[k][d] * dims[d];
      if (!merge) { d++; continue; }
      dims[d] *= dims[d + 1];
      dims.erase(dims.begin() + d + 1);
      for (auto &v : inc) v.erase(v.begin() + d + 1);
    }
  }
  const int rd = (int)dims.size();
  int64_t rows = 1;
  for (int d = 1; d < rd; d++) rows *= dims[d];
  if (rows > INT32_MAX) return make_err(PDLHIP_EUSERERROR, "%s: too many rows", who);
  std::vector<RagItem> items((size_t)rows);
  std::vector<int64_t> idx(rd, 0);
  for (int64_t r = 0; r < rows; r++) {
    RagItem &it = items[r];
    std::memset(&it, 0, sizeof it);
    int64_t o[RAG_MAXIN + 2];
    for (int k = 0; k < nop; k++) {
      o[k] = off[k];
      for (int d = 1; d < rd; d++) o[k] += idx[d] * inc[k][d];
    }
    for (int k = 0; k < nin; k++) {
      it.ptr[k] = elem_ptr(in[k], o[k]);
      it.inc[k] = inc[k][0];
    }
    it.len = dims[0];
    if (value) { it.val = elem_ptr(value, o[nin]); it.val_inc = inc[nin][0]; }
    if (f->op != RAG_NOREDUCE) { it.out = elem_ptr(result, o[nin + 1]); it.out_inc = f->scan ? inc[nin + 1][0] : 0; }
    for (int d = 1; d < rd; d++) {
      if (++idx[d] < dims[d]) break;
      idx[d] = 0;
    }
  }
  return ragged_finish(who, h, items, hout);
}

}  // namespace pdlhip

// ==================================================================== C ABI
using namespace pdlhip;

extern "C" {

int pdlhip_abi_version(void) { return PDLHIP_ABI_VERSION; }

pdlhip_error pdlhip_init(int device) {
  std::lock_guard<std::mutex> lk(g.mu);
  if (g.device == device) return ok();
  if (g.device >= 0) return make_err(PDLHIP_EUSERERROR, "pdlhip: already initialised on device %d", g.device);
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_err(e, "hipSetDevice");
  e = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_err(e, "hipStreamCreate");
  if ((e = hipEventCreate(&g.ev0)) != hipSuccess) return hip_err(e, "hipEventCreate");
  if ((e = hipEventCreate(&g.ev1)) != hipSuccess) return hip_err(e, "hipEventCreate");
  // this GPU's copy of the allocation-free entry point's scratch
  void *ptr = nullptr;
  if ((e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch))) != hipSuccess) return hip_err(e, "hipGetSymbolAddress");
  g.noalloc_scratch = ptr;
  g.device = device;
  return ok();
}

pdlhip_error pdlhip_shutdown(void) {
  std::lock_guard<std::mutex> lk(g.mu);
  if (g.device < 0) return ok();
  (void)hipStreamSynchronize(g.stream);
  {
    std::lock_guard<std::mutex> lc(g_cache.mu);
    release_cached();
  }
  if (g.rng_jumps) (void)hipFree(g.rng_jumps);
  g.rng_jumps = nullptr;
  (void)hipEventDestroy(g.ev0);
  (void)hipEventDestroy(g.ev1);
  (void)hipStreamDestroy(g.stream);
  g.device = -1;
  g.stream = nullptr;
  return ok();
}

pdlhip_error pdlhip_set_plan_only(int enable) {
  g.plan_only = enable != 0;
  return ok();
}

pdlhip_error pdlhip_alloc(size_t nbytes, void **data) {
  RETURN_IF(ensure_init());
  if (g.plan_only) { *data = nullptr; return ok(); }
  *data = dev_alloc(nbytes);
  if (!*data) return oom_err("pdlhip_alloc");
  return ok();
}

pdlhip_error pdlhip_free(void *data) {
  if (!data) return ok();
  if (!dev_free(data)) return make_err(PDLHIP_EUSERERROR, "pdlhip_free: %p was not allocated by pdlhip_alloc", data);
  return ok();
}

pdlhip_error pdlhip_upload(void *dst, const void *src, size_t nbytes) {
  RETURN_IF(ensure_init());
  if (capturing()) return make_err(PDLHIP_EUSERERROR, "pdlhip_upload: waits for the stream; not allowed while a graph is being captured");
  if (!nbytes) return ok();
  hipError_t e = hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, g.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
  return e == hipSuccess ? ok() : hip_err(e, "upload");
}

pdlhip_error pdlhip_download(void *dst, const void *src, size_t nbytes) {
  RETURN_IF(ensure_init());
  if (capturing()) return make_err(PDLHIP_EUSERERROR, "pdlhip_download: waits for the stream; not allowed while a graph is being captured");
  if (!nbytes) return ok();
  hipError_t e = hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToHost, g.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
  return e == hipSuccess ? ok() : hip_err(e, "download");
}

pdlhip_error pdlhip_synchronize(void) {
  RETURN_IF(ensure_init());
  if (capturing()) return make_err(PDLHIP_EUSERERROR, "pdlhip_synchronize: waits for the stream; not allowed while a graph is being captured");
  hipError_t e = hipStreamSynchronize(g.stream);
  return e == hipSuccess ? ok() : hip_err(e, "synchronize");
}

void *pdlhip_stream(void) { return (void *)g.stream; }

pdlhip_error pdlhip_op_info(const char *op, int *npdls, int *nparents) {
  const Sig *s = op ? find_sig(op) : nullptr;
  if (!s) return make_err(PDLHIP_EUSERERROR, "pdlhip: unknown op '%s'", op ? op : "(null)");
  if (npdls) *npdls = s->npdls;
  if (nparents) *nparents = s->nparents;
  return ok();
}

pdlhip_error pdlhip_run(const char *op, pdlhip_pdl **pdls, int npdls, const int64_t *otherpars) {
  const Sig *s = op ? find_sig(op) : nullptr;
  if (!s) return make_err(PDLHIP_EUSERERROR, "pdlhip: unknown op '%s'", op ? op : "(null)");
  return run(*s, pdls, npdls, otherpars);
}

#define BIOP_ENTRY(NAME) \
  pdlhip_error pdlhip_run_##NAME(pdlhip_pdl *a, pdlhip_pdl *b, pdlhip_pdl *c, int swap) { \
    pdlhip_pdl *v[3] = {a, b, c}; \
    int64_t o[1] = {swap}; \
    return pdlhip_run(#NAME, v, 3, o); \
  }
BIOP_ENTRY(plus)
BIOP_ENTRY(mult)
BIOP_ENTRY(minus)
BIOP_ENTRY(divide)
BIOP_ENTRY(gt)
BIOP_ENTRY(lt)
BIOP_ENTRY(le)
BIOP_ENTRY(ge)
BIOP_ENTRY(eq)
BIOP_ENTRY(ne)
BIOP_ENTRY(shiftleft)
BIOP_ENTRY(shiftright)
BIOP_ENTRY(or2)
BIOP_ENTRY(and2)
BIOP_ENTRY(xor)
BIOP_ENTRY(power)
BIOP_ENTRY(atan2)
BIOP_ENTRY(modulo)
BIOP_ENTRY(spaceship)

#define UNARY_ENTRY(NAME) \
  pdlhip_error pdlhip_run_##NAME(pdlhip_pdl *a, pdlhip_pdl *b) { \
    pdlhip_pdl *v[2] = {a, b}; \
    return pdlhip_run(#NAME, v, 2, nullptr); \
  }
UNARY_ENTRY(assgn)
UNARY_ENTRY(sumover)
UNARY_ENTRY(prodover)
UNARY_ENTRY(dsumover)
UNARY_ENTRY(dprodover)
UNARY_ENTRY(cumusumover)
UNARY_ENTRY(cumuprodover)
UNARY_ENTRY(dcumusumover)
UNARY_ENTRY(dcumuprodover)
UNARY_ENTRY(average)
UNARY_ENTRY(daverage)
UNARY_ENTRY(minimum)
UNARY_ENTRY(maximum)
UNARY_ENTRY(minimum_ind)
UNARY_ENTRY(maximum_ind)

pdlhip_error pdlhip_run_minmaximum(pdlhip_pdl *a, pdlhip_pdl *cmin, pdlhip_pdl *cmax, pdlhip_pdl *cmin_ind,
                                   pdlhip_pdl *cmax_ind) {
  pdlhip_pdl *v[5] = {a, cmin, cmax, cmin_ind, cmax_ind};
  return pdlhip_run("minmaximum", v, 5, nullptr);
}

pdlhip_error pdlhip_run_minimum_n_ind(pdlhip_pdl *a, pdlhip_pdl *c, int64_t m_size) {
  pdlhip_pdl *v[2] = {a, c};
  int64_t o[1] = {m_size};
  return pdlhip_run("minimum_n_ind", v, 2, o);
}

pdlhip_error pdlhip_run_maximum_n_ind(pdlhip_pdl *a, pdlhip_pdl *c, int64_t m_size) {
  pdlhip_pdl *v[2] = {a, c};
  int64_t o[1] = {m_size};
  return pdlhip_run("maximum_n_ind", v, 2, o);
}

pdlhip_error pdlhip_run_matmult(pdlhip_pdl *a, pdlhip_pdl *b, pdlhip_pdl *c) {
  pdlhip_pdl *v[3] = {a, b, c};
  return pdlhip_run("matmult", v, 3, nullptr);
}

pdlhip_error pdlhip_tune_stats(int64_t *measured, int64_t *reused) {
  if (!measured || !reused) return make_err(PDLHIP_EUSERERROR, "pdlhip_tune_stats: NULL argument");
  std::lock_guard<std::mutex> lk(g_tuner.mu);
  *measured = g_tuner.measured;
  *reused = g_tuner.reused;
  return ok();
}

pdlhip_error pdlhip_run_reduce_noalloc(const char *op, pdlhip_pdl *a, pdlhip_pdl *b) {
  const Sig *s = op ? find_sig(op) : nullptr;
  if (!s || s->family != FAM_RED || s->npdls != 2 || s->realdims[1] != 0)
    return make_err(PDLHIP_EUSERERROR,
                    "pdlhip_run_reduce_noalloc: '%s' is not one of sumover, prodover, dsumover, dprodover, "
                    "average, daverage, minimum, maximum, minimum_ind, maximum_ind",
                    op ? op : "(null)");
  if (!a || !b) return make_err(PDLHIP_EUSERERROR, "pdlhip_run_reduce_noalloc: %s: NULL ndarray", op);
  if (b->state & PDLHIP_NOMYDIMS)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_run_reduce_noalloc: %s: the output must exist (none is created)", op);
  RETURN_IF(ensure_init());
  if (g.plan_only) return make_err(PDLHIP_EUSERERROR, "pdlhip_run_reduce_noalloc: not available when only planning");
  pdlhip_pdl *v[2] = {a, b};
  const int T = select_transtype(*s, v);
  const int tout = adjust_from_trans(T, s->flags[1], s->par_types[1]);
  if (a->datatype != T || b->datatype != tout)
    return make_err(PDLHIP_EUSERERROR,
                    "pdlhip_run_reduce_noalloc: %s: a must be %s and the output %s (no converted copy is made)", op,
                    type_name(T), type_name(tout));
  g_noalloc = true;
  pdlhip_error e = run(*s, v, 2, nullptr);
  g_noalloc = false;
  return e;
}

UNARY_ENTRY(bitnot)
UNARY_ENTRY(sqrt)
UNARY_ENTRY(sin)
UNARY_ENTRY(cos)
UNARY_ENTRY(not)
UNARY_ENTRY(exp)
UNARY_ENTRY(log)
UNARY_ENTRY(log10)
UNARY_ENTRY(abs)
UNARY_ENTRY(isbad)
UNARY_ENTRY(isgood)
UNARY_ENTRY(setnantobad)
UNARY_ENTRY(setinftobad)
UNARY_ENTRY(setnonfinitetobad)
UNARY_ENTRY(setbadtonan)

pdlhip_error pdlhip_run_index(pdlhip_pdl *a, pdlhip_pdl *ind, pdlhip_pdl *c) {
  pdlhip_pdl *v[3] = {a, ind, c};
  return pdlhip_run("index", v, 3, nullptr);
}
pdlhip_error pdlhip_run_index1d(pdlhip_pdl *a, pdlhip_pdl *ind, pdlhip_pdl *c) {
  pdlhip_pdl *v[3] = {a, ind, c};
  return pdlhip_run("index1d", v, 3, nullptr);
}
pdlhip_error pdlhip_run_index2d(pdlhip_pdl *a, pdlhip_pdl *inda, pdlhip_pdl *indb, pdlhip_pdl *c) {
  pdlhip_pdl *v[4] = {a, inda, indb, c};
  return pdlhip_run("index2d", v, 4, nullptr);
}

// the vtable's writebackdata for the ops with a BackCode (Slices.pd index,
// index1d, index2d): every ndarray exists, a and c already have the type the
// op computes in; a is written, the others are left as they are
pdlhip_error pdlhip_writebackdata(const char *op, pdlhip_pdl **pdls, int npdls) {
  const Sig *s = op ? find_sig(op) : nullptr;
  if (!s || s->family != FAM_INDEX)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_writebackdata: '%s' has no BackCode here (index, index1d, index2d)",
                    op ? op : "(null)");
  if (!pdls || npdls != s->npdls)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_writebackdata: %s takes %d ndarrays", op, s->npdls);
  for (int i = 0; i < npdls; i++)
    if (!pdls[i] || (pdls[i]->state & PDLHIP_NOMYDIMS))
      return make_err(PDLHIP_EUSERERROR, "pdlhip_writebackdata: %s: parameter '%s' is null", op,
                      s->par_names[i].c_str());
  const int T = select_transtype(*s, pdls);
  pdlhip_pdl *a = pdls[0], *c = pdls[npdls - 1];
  if (a->datatype != T || c->datatype != T)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_writebackdata: %s: a and c must both be %s", op, type_name(T));
  const int32_t cstate = c->state;
  g_writeback = true;
  pdlhip_error e = run(*s, pdls, npdls, nullptr);
  g_writeback = false;
  c->state = cstate;
  return e;
}

// Slices.pd rangeb (the engine of range and indexND): MakeComp's index and
// size conditioning and boundary letters, RedoDims' child dims, then the
// EquivCPOffs loop on the device (k_index.hip); back: its BackCode
static pdlhip_error rangeb(pdlhip_pdl *parent, pdlhip_pdl *ind, const int64_t *size, int64_t nsize,
                           const char *boundary, pdlhip_pdl *child, bool back) {
  RETURN_IF(ensure_init());
  if (!parent || !ind || !child) return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: NULL ndarray");
  {
    pdlhip_pdl *ps[3] = {parent, ind, child};
    RETURN_IF(refuse_explicit_ids("pdlhip: rangeb", ps, 3));
  }
  if ((parent->state & PDLHIP_NOMYDIMS) || (ind->state & PDLHIP_NOMYDIMS))
    return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: the source and the index must exist");
  if (!type_on_device(parent->datatype) || !type_on_device(ind->datatype))
    return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: type not supported on the device");
  if (capturing())
    return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb checks its indices on the host and cannot be captured");
  if (ind->ndims && ind->dims[0] == 0)
    return make_err(PDLHIP_EUSERERROR, "Error in rangeb:can't handle Empty indices -- call range instead");
  RgPlan p;
  std::memset(&p, 0, sizeof p);
  const int64_t rdim = ind->ndims ? ind->dims[0] : 1;
  if (rdim > RG_MAXR) return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: more than %d index coordinates", RG_MAXR);
  p.rdim = (int)rdim;
  p.itdim = ind->ndims ? ind->ndims - 1 : 0;
  p.nitems = 1;
  for (int i = 1; i < ind->ndims; i++) p.nitems *= ind->dims[i];
  const size_t blen = boundary ? strlen(boundary) : 0;
  for (int64_t i = 0; i < rdim; i++) {
    if (!blen) continue;
    switch (boundary[i < (int64_t)blen - 1 ? i : (int64_t)blen - 1]) {
    case '0': case 'f': case 'F': p.bnd[i] = 0; break;
    case '1': case 't': case 'T': p.bnd[i] = 1; break;
    case '2': case 'e': case 'E': p.bnd[i] = 2; break;
    case '3': case 'p': case 'P': p.bnd[i] = 3; break;
    case '4': case 'm': case 'M': p.bnd[i] = 4; break;
    default:
      return make_err(PDLHIP_EUSERERROR, "Error in rangeb:Unknown boundary condition '%c' in range",
                      i < (int64_t)blen ? boundary[i] : '\0');
    }
  }
  if (size && nsize > 0) {
    if (nsize == 1) {
      for (int64_t i = 0; i < rdim; i++) p.sizes[i] = size[0];
      if (rdim > 0 && p.sizes[0] < 0)
        return make_err(PDLHIP_EUSERERROR, "Error in rangeb:Negative range size is not allowed\n");
    } else if (nsize <= rdim) {
      for (int64_t i = 0; i < rdim; i++) {
        p.sizes[i] = i < nsize ? size[i] : 0;
        if (p.sizes[i] < 0) return make_err(PDLHIP_EUSERERROR, "Error in rangeb:Negative range sizes are not allowed\n");
      }
    } else {
      return make_err(PDLHIP_EUSERERROR, "Error in rangeb:Size must match index's 0th dim\n");
    }
  }
  for (int64_t i = 0; i < rdim; i++) p.ntsize += p.sizes[i] != 0;
  p.pdim = parent->ndims;
  if (rdim > p.pdim + 5 && nsize != rdim)
    return make_err(PDLHIP_EUSERERROR,
                    "Error in rangeb:Ludicrous number of extra dims in range index; leaving child null.\n"
                    "  (%lld implicit dims is > 5; index has %lld dims; source has %d dim%s.)\n"
                    "  This often means that your index PDL is incorrect.\n"
                    "  To avoid this message, allocate dummy dims in\n"
                    "  the source or use %lld dims in range's size field.\n",
                    (long long)(rdim - p.pdim), (long long)rdim, p.pdim, p.pdim > 1 ? "s" : "", (long long)rdim);
  p.stdim = p.pdim - (int)rdim > 0 ? p.pdim - (int)rdim : 0;
  p.cnd = p.itdim + p.ntsize + p.stdim;
  if (p.cnd > PDLHIP_MAXDIMS) return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: the child would have %d dims", p.cnd);
  int64_t cd[PDLHIP_MAXDIMS];
  {
    int dim = p.itdim;
    for (int64_t i = 0; i < rdim; i++) if (p.sizes[i]) cd[dim++] = p.sizes[i];
    for (dim = 0; dim < p.itdim; dim++) cd[dim] = ind->dims[dim + 1];
    dim = p.itdim + p.ntsize;
    for (int i = 0; i < p.stdim; i++) cd[dim++] = parent->dims[i + rdim];
  }
  // an empty source: the boundaries that do not croak truncate (RedoDims);
  // a zero extent under extend / periodic / mirror would divide by zero
  for (int64_t k = 0; k < rdim && k < p.pdim; k++)
    if (p.bnd[k] >= 2 && (parent->dims[k] == 0 || (p.pdim && parent->dims[0] == 0))) p.bnd[k] = 1;
  if (p.pdim && parent->dims[0] == 0)
    for (int64_t k = 0; k < rdim; k++) if (p.bnd[k]) p.bnd[k] = 1;
  p.nchild = 1;
  for (int d = 0; d < p.cnd; d++) p.nchild *= cd[d];
  p.ninr = 1;
  for (int64_t k = 0; k < rdim; k++) p.ninr *= p.sizes[k] ? p.sizes[k] : 1;
  p.nst = 1;
  for (int i = 0; i < p.stdim; i++) p.nst *= parent->dims[i + rdim];
  if (back) {
    if (child->state & PDLHIP_NOMYDIMS) return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: no child to write back");
    bool same = child->ndims == p.cnd && child->datatype == parent->datatype;
    for (int d = 0; same && d < p.cnd; d++) same = child->dims[d] == cd[d];
    if (!same) return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: the child does not have the range's dims and type");
  } else if (child->state & PDLHIP_NOMYDIMS) {
    child->datatype = parent->datatype;
    child->ndims = p.cnd;
    int64_t inc = 1;
    for (int d = 0; d < p.cnd; d++) { child->dims[d] = cd[d]; child->incs[d] = inc; inc *= cd[d]; }
    child->offs = 0;
    child->has_badvalue = 0;
    child->data = g.plan_only ? nullptr : dev_alloc((size_t)std::max<int64_t>(p.nchild, 1) * type_size(child->datatype));
    if (!g.plan_only && !child->data) return oom_err("rangeb");
    child->state = parent->state & PDLHIP_BADVAL;
  } else {
    bool same = child->ndims == p.cnd && child->datatype == parent->datatype;
    for (int d = 0; same && d < p.cnd; d++) same = child->dims[d] == cd[d];
    if (!same) return make_err(PDLHIP_EUSERERROR, "pdlhip: rangeb: the output does not have the range's dims and type");
  }
  if (g.plan_only || p.nchild == 0 || p.nitems == 0) return ok();
  // the index as indx (MakeComp converts it)
  pdlhip_pdl iconv;
  pdlhip_pdl *ix = ind;
  bool conv = false;
  if (ind->datatype != PDLHIP_IND) {
    RETURN_IF(make_converted(ind, PDLHIP_IND, &iconv, true));
    ix = &iconv;
    conv = true;
  }
  p.ioff = ix->offs;
  p.iincs[0] = ix->ndims ? ix->incs[0] : 0;
  for (int d = 0; d < p.itdim; d++) p.iincs[d + 1] = ix->incs[d + 1];
  for (int d = 0; d < p.pdim; d++) { p.pdims[d] = parent->dims[d]; p.pincs[d] = parent->incs[d]; }
  p.poff = parent->offs;
  for (int d = 0; d < p.cnd; d++) { p.cdims[d] = cd[d]; p.cincs[d] = child->incs[d]; }
  p.coff = child->offs;
  p.bad = (parent->state & PDLHIP_BADVAL) != 0;
  p.back = back ? 1 : 0;
  long long *owner = nullptr;
  if (back) {
    int64_t lo = parent->offs, hi = parent->offs;
    for (int d = 0; d < parent->ndims; d++) {
      const int64_t ext = (parent->dims[d] - 1) * parent->incs[d];
      if (ext < 0) lo += ext; else hi += ext;
    }
    p.a_lo = lo;
    p.a_span = hi - lo + 1;
    owner = (long long *)dev_alloc((size_t)p.a_span * sizeof(long long));
    if (!owner) {
      if (conv) dev_free(iconv.data);
      return oom_err("rangeb");
    }
  }
  unsigned long long *err = (unsigned long long *)dev_alloc(sizeof(unsigned long long));
  if (!err) {
    dev_free(owner);
    if (conv) dev_free(iconv.data);
    return oom_err("rangeb");
  }
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, g.stream, err, ~0ull);
  const BadInfo pb = badinfo(parent, 1), cb = badinfo(child, 1);
  hipError_t e = launch_range(parent->datatype, p, parent->data, (const int64_t *)ix->data, child->data, pb.bv, cb.bv,
                              pb.isnan, cb.isnan, err, owner, g.stream);
  unsigned long long key = ~0ull;
  if (e == hipSuccess) e = hipMemcpyAsync(&key, err, sizeof key, hipMemcpyDeviceToHost, g.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
  dev_free(err);
  dev_free(owner);
  if (conv) dev_free(iconv.data);
  if (e != hipSuccess) return hip_err(e, "rangeb");
  if (key == ~0ull) return ok();
  if (key & 1)
    return make_err(PDLHIP_EUSERERROR,
                    "Error in rangeb:Too many dims in range index (and you've forbidden boundary violations)");
  return make_err(PDLHIP_EUSERERROR, "Error in rangeb:index out-of-bounds in range (index vector #%ld)",
                  (long)((key >> 1) / p.ninr));
}

pdlhip_error pdlhip_run_rangeb(pdlhip_pdl *parent, pdlhip_pdl *ind, const int64_t *size, int64_t nsize,
                               const char *boundary, pdlhip_pdl *child) {
  return rangeb(parent, ind, size, nsize, boundary, child, false);
}
pdlhip_error pdlhip_writeback_rangeb(pdlhip_pdl *parent, pdlhip_pdl *ind, const int64_t *size, int64_t nsize,
                                     const char *boundary, pdlhip_pdl *child) {
  return rangeb(parent, ind, size, nsize, boundary, child, true);
}

pdlhip_error pdlhip_run_setbadif(pdlhip_pdl *a, pdlhip_pdl *mask, pdlhip_pdl *b) {
  pdlhip_pdl *v[3] = {a, mask, b};
  return pdlhip_run("setbadif", v, 3, nullptr);
}

pdlhip_error pdlhip_run_setvaltobad(pdlhip_pdl *a, pdlhip_pdl *b, double value) {
  pdlhip_pdl *v[2] = {a, b};
  int64_t o[1];
  std::memcpy(&o[0], &value, sizeof value);
  return pdlhip_run("setvaltobad", v, 2, o);
}

pdlhip_error pdlhip_run_setbadtoval(pdlhip_pdl *a, pdlhip_pdl *b, double newval) {
  pdlhip_pdl *v[2] = {a, b};
  int64_t o[1];
  std::memcpy(&o[0], &newval, sizeof newval);
  return pdlhip_run("setbadtoval", v, 2, o);
}

pdlhip_error pdlhip_run_matmult_list(int n, pdlhip_pdl **a, pdlhip_pdl **b, pdlhip_pdl **c, int mode) {
  if (n < 0 || (n > 0 && (!a || !b || !c))) return make_err(PDLHIP_EUSERERROR, "pdlhip_run_matmult_list: bad arguments");
  if (mode != PDLHIP_LIST_QUEUE && mode != PDLHIP_LIST_EACH)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_run_matmult_list: unknown mode %d", mode);
  return run_list(n, a, b, c, mode);
}

struct pdlhip_mmlist { pdlhip::ListH h; };

pdlhip_error pdlhip_matmult_list_prepare(int n, pdlhip_pdl **a, pdlhip_pdl **b, pdlhip_pdl **c,
                                         pdlhip_mmlist **handle) {
  if (!handle || n < 0 || (n > 0 && (!a || !b || !c)))
    return make_err(PDLHIP_EUSERERROR, "pdlhip_matmult_list_prepare: bad arguments");
  RETURN_IF(ensure_init());
  pdlhip_mmlist *l = new pdlhip_mmlist;
  pdlhip_error e = list_build(n, a, b, c, true, l->h);
  if (e.error) {
    delete l;
    return e;
  }
  *handle = l;
  return ok();
}

pdlhip_error pdlhip_matmult_list_launch(pdlhip_mmlist *handle, int mode) {
  if (!handle) return make_err(PDLHIP_EUSERERROR, "pdlhip_matmult_list_launch: no handle");
  if (mode != PDLHIP_LIST_QUEUE && mode != PDLHIP_LIST_EACH)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_matmult_list_launch: unknown mode %d", mode);
  RETURN_IF(ensure_init());
  if (g.plan_only || handle->h.d.items.empty()) return ok();
  hipError_t e = list_launch(handle->h, handle->h.tile, mode == PDLHIP_LIST_EACH);
  return e == hipSuccess ? ok() : hip_err(e, "matmult list");
}

pdlhip_error pdlhip_matmult_list_free(pdlhip_mmlist *handle) {
  if (!handle) return ok();
  list_release(handle->h);
  delete handle;
  return ok();
}

struct pdlhip_ragged { pdlhip::RaggedH h; };

pdlhip_error pdlhip_ragged_prepare(int nitems, int ninputs, pdlhip_pdl **inputs, int nnodes, const pdlhip_node *nodes,
                                   const char *final_op, pdlhip_pdl **outputs, pdlhip_pdl **values,
                                   pdlhip_ragged **handle) {
  if (!handle) return make_err(PDLHIP_EUSERERROR, "pdlhip_ragged_prepare: no handle");
  RETURN_IF(ensure_init());
  if (inputs && nitems > 0 && ninputs > 0)
    RETURN_IF(refuse_explicit_ids("pdlhip_ragged_prepare", inputs, nitems * ninputs));
  if (outputs && nitems > 0) RETURN_IF(refuse_explicit_ids("pdlhip_ragged_prepare", outputs, nitems));
  if (values && nitems > 0) RETURN_IF(refuse_explicit_ids("pdlhip_ragged_prepare", values, nitems));
  RaggedH *h = nullptr;
  pdlhip_error e = ragged_prepare(nitems, ninputs, inputs, nnodes, nodes, final_op, outputs, values, &h);
  if (e.error) return e;
  *handle = new pdlhip_ragged{*h};
  delete h;
  return ok();
}

pdlhip_error pdlhip_chain_prepare(int ninputs, pdlhip_pdl **inputs, int nnodes, const pdlhip_node *nodes,
                                  const char *final_op, pdlhip_pdl *value, pdlhip_pdl *result,
                                  pdlhip_ragged **handle) {
  if (!handle) return make_err(PDLHIP_EUSERERROR, "pdlhip_chain_prepare: no handle");
  RETURN_IF(ensure_init());
  if (inputs && ninputs > 0) RETURN_IF(refuse_explicit_ids("pdlhip_chain_prepare", inputs, ninputs));
  {
    pdlhip_pdl *ps[2] = {value, result};
    RETURN_IF(refuse_explicit_ids("pdlhip_chain_prepare", ps, 2));
  }
  RaggedH *h = nullptr;
  pdlhip_error e = chain_prepare(ninputs, inputs, nnodes, nodes, final_op, value, result, &h);
  if (e.error) return e;
  *handle = new pdlhip_ragged{*h};
  delete h;
  return ok();
}

pdlhip_error pdlhip_ragged_run(pdlhip_ragged *handle) {
  if (!handle) return make_err(PDLHIP_EUSERERROR, "pdlhip_ragged_run: no handle");
  RETURN_IF(ensure_init());
  if (g.plan_only) return ok();
  const RaggedH &h = handle->h;
  hipError_t e = h.scan ? launch_ragged_scan(h.T, h.op, h.L, g.stream) : launch_ragged_reduce(h.T, h.op, h.L, g.stream);
  return e == hipSuccess ? ok() : hip_err(e, "ragged run");
}

pdlhip_error pdlhip_ragged_badflags(pdlhip_ragged *handle, int32_t *flags) {
  if (!handle || !flags) return make_err(PDLHIP_EUSERERROR, "pdlhip_ragged_badflags: bad arguments");
  RETURN_IF(ensure_init());
  const RaggedH &h = handle->h;
  if (h.op != R_MIN && h.op != R_MAX) {
    std::memset(flags, 0, sizeof(int32_t) * h.n);
    return ok();
  }
  if (g.plan_only || h.n == 0) return ok();
  if (capturing())
    return make_err(PDLHIP_EUSERERROR, "pdlhip_ragged_badflags: waits for the stream; not allowed while a graph is "
                                       "being captured");
  hipError_t e = hipStreamSynchronize(g.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(flags, h.d_empty, sizeof(int32_t) * h.n, hipMemcpyDeviceToHost, g.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
  return e == hipSuccess ? ok() : hip_err(e, "ragged badflags");
}

pdlhip_error pdlhip_ragged_count(pdlhip_ragged *handle, int64_t *nitems) {
  if (!handle || !nitems) return make_err(PDLHIP_EUSERERROR, "pdlhip_ragged_count: bad arguments");
  *nitems = handle->h.n;
  return ok();
}

pdlhip_error pdlhip_ragged_free(pdlhip_ragged *handle) {
  if (!handle) return ok();
  RaggedH &h = handle->h;
  for (void *p : {h.d_items, h.d_segs, h.d_multi, h.d_order, h.d_bstart, h.d_scratch, h.d_empty}) dev_free(p);
  delete handle;
  return ok();
}

// a captured graph and the blocks its kernels use (released to the cache
// while it was captured): they go back to the cache when it is destroyed
struct GraphH {
  hipGraphExec_t exec = nullptr;
  std::vector<void *> pinned;
};

static void unpin(std::vector<void *> &blocks) {
  for (void *p : blocks) dev_free(p);
  blocks.clear();
}

pdlhip_error pdlhip_graph_begin(void) {
  RETURN_IF(ensure_init());
  if (capturing()) return make_err(PDLHIP_EUSERERROR, "pdlhip_graph_begin: a capture is already in progress");
  hipError_t e = hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_err(e, "hipStreamBeginCapture");
  g.capturing = true;
  return ok();
}

pdlhip_error pdlhip_graph_end(void **graph) {
  RETURN_IF(ensure_init());
  if (!graph) return make_err(PDLHIP_EUSERERROR, "pdlhip_graph_end: graph is NULL");
  *graph = nullptr;
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(g.stream, &gr);
  g.capturing = false;
  std::vector<void *> pinned;
  {
    std::lock_guard<std::mutex> lk(g_cache.mu);
    pinned.swap(g.pinned);
  }
  if (e != hipSuccess) {
    unpin(pinned);
    return hip_err(e, "hipStreamEndCapture");
  }
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) {
    unpin(pinned);
    return hip_err(e, "hipGraphInstantiate");
  }
  GraphH *h = new GraphH;
  h->exec = ex;
  h->pinned.swap(pinned);
  *graph = (void *)h;
  return ok();
}

pdlhip_error pdlhip_graph_launch(void *graph) {
  RETURN_IF(ensure_init());
  if (!graph) return make_err(PDLHIP_EUSERERROR, "pdlhip_graph_launch: graph is NULL");
  hipError_t e = hipGraphLaunch(((GraphH *)graph)->exec, g.stream);
  return e == hipSuccess ? ok() : hip_err(e, "hipGraphLaunch");
}

pdlhip_error pdlhip_graph_destroy(void *graph) {
  if (!graph) return ok();
  GraphH *h = (GraphH *)graph;
  hipError_t e = hipGraphExecDestroy(h->exec);
  // replays already queued on the library stream come before any later use
  // of these blocks
  unpin(h->pinned);
  delete h;
  return e == hipSuccess ? ok() : hip_err(e, "hipGraphExecDestroy");
}

pdlhip_error pdlhip_set_matmult_precision(int mode) {
  if (mode < PDLHIP_MM_SPLIT16 || mode > PDLHIP_MM_FP8)
    return make_err(PDLHIP_EUSERERROR, "pdlhip: unknown matmult precision %d", mode);
  g.mm_mode = mode;
  return ok();
}

int pdlhip_get_matmult_precision(void) { return g.mm_mode; }

pdlhip_error pdlhip_readdata(const char *op, int datatype, int bvalflag, const pdlhip_broadcast *brc,
                             pdlhip_pdl **pdls, const int64_t *otherpars) {
  const Sig *s = op ? find_sig(op) : nullptr;
  if (!s) return make_err(PDLHIP_EUSERERROR, "pdlhip: unknown op '%s'", op ? op : "(null)");
  RETURN_IF(ensure_init());
  if (!brc || brc->npdls != s->npdls || brc->ndims < 0 || brc->ndims > PDLHIP_MAXDIMS + 2)
    return make_err(PDLHIP_EUSERERROR, "pdlhip_readdata: %s: broadcast struct does not match the signature",
                    s->name.c_str());
  Brc b;
  b.npdls = brc->npdls;
  b.ndims = brc->ndims;
  b.nimpl = brc->ndims;
  for (int d = 0; d < b.ndims; d++) b.dims[d] = brc->dims[d];
  for (int k = 0; k < b.ndims * b.npdls; k++) b.incs[k] = brc->incs[k];
  for (int p = 0; p < b.npdls; p++) b.offs[p] = brc->offs[p];
  for (int k = 0; k < 4; k++) b.ind_sizes[k] = brc->ind_sizes[k];
  for (int p = 0; p < b.npdls; p++)
    for (int j = 0; j < 2; j++) b.inc_sizes[p][j] = brc->inc_sizes[p * 2 + j];
  pdlhip_pdl *v[MAXP];
  for (int i = 0; i < s->npdls; i++) v[i] = pdls[i];
  if ((s->family == FAM_BIOP || s->family == FAM_BIFUNC) && otherpars && otherpars[0]) {
    // PDL's HdrCode swaps the pdls before the trans is built, so a readdata
    // hook receives them already swapped
  }
  return readdata(*s, datatype, bvalflag != 0, b, v, otherpars);
}

pdlhip_error pdlhip_timer_start(void) {
  RETURN_IF(ensure_init());
  hipError_t e = hipEventRecord(g.ev0, g.stream);
  return e == hipSuccess ? ok() : hip_err(e, "timer");
}

pdlhip_error pdlhip_timer_stop(float *ms) {
  RETURN_IF(ensure_init());
  hipError_t e = hipEventRecord(g.ev1, g.stream);
  if (e == hipSuccess) e = hipEventSynchronize(g.ev1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, g.ev0, g.ev1);
  return e == hipSuccess ? ok() : hip_err(e, "timer");
}

pdlhip_error pdlhip_event_create(void **ev) {
  RETURN_IF(ensure_init());
  hipEvent_t e;
  hipError_t he = hipEventCreate(&e);
  if (he != hipSuccess) return hip_err(he, "hipEventCreate");
  *ev = (void *)e;
  return ok();
}

pdlhip_error pdlhip_event_record(void *ev) {
  RETURN_IF(ensure_init());
  hipError_t he = hipEventRecord((hipEvent_t)ev, g.stream);
  return he == hipSuccess ? ok() : hip_err(he, "hipEventRecord");
}

pdlhip_error pdlhip_event_elapsed(void *a, void *b, float *ms) {
  RETURN_IF(ensure_init());
  hipError_t he = hipEventSynchronize((hipEvent_t)b);
  if (he == hipSuccess) he = hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b);
  return he == hipSuccess ? ok() : hip_err(he, "hipEventElapsedTime");
}

pdlhip_error pdlhip_event_destroy(void *ev) {
  hipError_t he = hipEventDestroy((hipEvent_t)ev);
  return he == hipSuccess ? ok() : hip_err(he, "hipEventDestroy");
}

static pdlhip_error fill(pdlhip_pdl *p, uint64_t seed, double a, double b, int normal) {
  RETURN_IF(ensure_init());
  if (!p || (p->state & PDLHIP_NOMYDIMS)) return make_err(PDLHIP_EUSERERROR, "pdlhip_fill: null ndarray");
  if (!type_on_device(p->datatype)) return make_err(PDLHIP_EUSERERROR, "pdlhip_fill: type not supported");
  Dimset d;
  d.np = 3;
  d.n = p->ndims;
  for (int j = 0; j < p->ndims; j++) {
    d.dims[j] = p->dims[j];
    d.inc[0][j] = d.inc[1][j] = 0;
    d.inc[2][j] = p->incs[j];
  }
  if (!canonicalise(d, 2)) return ok();
  EwPlan e;
  if (!fill_ew(e, d)) return make_err(PDLHIP_EFATAL, "pdlhip_fill: too many distinct dims");
  e.off[2] = p->offs;
  hipError_t he = launch_fill(p->datatype, e, p->data, seed, a, b, normal, g.stream);
  return he == hipSuccess ? ok() : hip_err(he, "fill");
}

// ------------------------------------------------------------------ random numbers
// srandom / random / randsym / grandom (Primitive.pd:2113-2262) on PDL's
// stream (k_rng.hip).  PDL_MAYBE_SRAND (Primitive.pd:2095-2097) seeds from
// the clock when nothing has seeded the stream yet.
static pdlhip_error rand_fill(pdlhip_pdl *p, int kind, const char *name) {
  RETURN_IF(ensure_init());
  if (!p || (p->state & PDLHIP_NOMYDIMS))
    return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: the output ndarray must exist", name);
  if (!type_on_device(p->datatype))
    return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: type %s is not supported on the device", name,
                    type_name(p->datatype));
  if (kind == 2 && !IsFloatCode(p->datatype))
    return make_err(PDLHIP_EUSERERROR, "pdlhip: grandom: type %s is not a floating type", type_name(p->datatype));
  if (capturing())
    return make_err(PDLHIP_EUSERERROR, "pdlhip: %s: the stream's state lives on the host; it cannot be captured", name);
  if (g.plan_only) return ok();
  if (!g.rng_seeded) {
    rng_seed((uint64_t)time(nullptr), g.rng);
    g.rng_seeded = true;
  }
  RngPlan rp;
  std::memset(&rp, 0, sizeof rp);
  rp.nd = p->ndims;
  rp.off = p->offs;
  rp.n = 1;
  int64_t dense_inc = 1;
  rp.dense = 1;
  for (int d = 0; d < p->ndims; d++) {
    rp.dims[d] = p->dims[d];
    rp.incs[d] = p->incs[d];
    rp.n *= p->dims[d];
    if (p->dims[d] > 1 && p->incs[d] != dense_inc) rp.dense = 0;
    dense_inc *= p->dims[d];
  }
  if (rp.n == 0) return ok();
  if (!g.rng_jumps) {
    size_t nw = 0;
    const uint64_t *jt = rng_jump_table(&nw);
    hipError_t e = hipMalloc((void **)&g.rng_jumps, nw * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(g.rng_jumps, jt, nw * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (g.rng_jumps) (void)hipFree(g.rng_jumps);
      g.rng_jumps = nullptr;
      return hip_err(e, name);
    }
  }
  const int64_t nchunks = (rp.n + RNG_CHUNK - 1) / RNG_CHUNK;
  std::vector<uint64_t> states((size_t)nchunks * 4);
  rng_chunk_states(g.rng, nchunks, states.data());
  uint64_t *d_states = (uint64_t *)dev_alloc(states.size() * sizeof(uint64_t));
  int *d_flag = (int *)dev_alloc(sizeof(int));
  if (!d_states || !d_flag) {
    if (d_states) dev_free(d_states);
    if (d_flag) dev_free(d_flag);
    return oom_err(name);
  }
  hipError_t e = upload_now(d_states, states.data(), states.size() * sizeof(uint64_t));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, g.stream, d_flag, 1);
    e = launch_rng(p->datatype, kind, rp, p->data, d_states, g.rng_jumps, nchunks, d_flag, g.stream);
  }
  int h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, g.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g.stream);
  dev_free(d_states);
  dev_free(d_flag);
  if (e != hipSuccess) return hip_err(e, name);
  rng_advance(g.rng, (uint64_t)rp.n);
  if (h)
    return make_err(PDLHIP_EFATAL,
                    "pdlhip: %s: the stream drew an exact 0.0, which PDL draws again; the device does not", name);
  return ok();
}

pdlhip_error pdlhip_srandom(uint64_t seed) {
  rng_seed(seed, g.rng);
  g.rng_seeded = true;
  return ok();
}
pdlhip_error pdlhip_rand_skip(uint64_t n) {
  if (!g.rng_seeded) return make_err(PDLHIP_EUSERERROR, "pdlhip: the random stream has not been seeded");
  rng_advance(g.rng, n);
  return ok();
}
pdlhip_error pdlhip_rand_state(uint64_t state[4]) {
  if (!g.rng_seeded) return make_err(PDLHIP_EUSERERROR, "pdlhip: the random stream has not been seeded");
  std::memcpy(state, g.rng, sizeof g.rng);
  return ok();
}
pdlhip_error pdlhip_run_random(pdlhip_pdl *a) { return rand_fill(a, 0, "random"); }
pdlhip_error pdlhip_run_randsym(pdlhip_pdl *a) { return rand_fill(a, 1, "randsym"); }
pdlhip_error pdlhip_grandom(pdlhip_pdl *a) { return rand_fill(a, 2, "grandom"); }

pdlhip_error pdlhip_fill_uniform(pdlhip_pdl *p, uint64_t seed, double lo, double hi) { return fill(p, seed, lo, hi, 0); }
pdlhip_error pdlhip_fill_normal(pdlhip_pdl *p, uint64_t seed, double mean, double sd) {
  return fill(p, seed, mean, sd, 1);
}

}  // extern "C"
