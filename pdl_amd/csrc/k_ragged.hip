// Ragged lists: many 1-D ndarrays of uneven lengths processed in a fixed set
// of launches (the reference runs such a list as a Perl loop over ops, one
// broadcast loop per ndarray per op).  The host prepares an item table (per
// item, per input slot: pointer, stride, length; output pointer).
//
// Reductions with a fused chain: each element's value is computed in
// registers by a small program of Ops.pd binops (each node rounding in the
// chain's trans type, exactly as PDL's eager temporaries would, with no FMA
// contraction), optionally stored, then accumulated with the partial rules of
// Ufunc.pd (reduce_common.hpp) unless the final op is "none".  Items are cut into segments of at most RAG_SEG
// elements, one wave per segment; items of several segments are finished by a
// combine kernel.  A left-deep chain (each node takes the previous result
// and one input slot) runs in a kernel that works on U elements per lane per
// step with wave-uniform type and op switches; any other program shape runs
// in a general (slower) kernel.
//
// Integer scans (cumusumover/cumuprodover): one block per item, the item
// scanned in chunks of RAG_SEG elements staged through LDS with a running
// carry, so every element is read and written once.  The host assigns items
// to blocks longest-first (LPT), which bounds the makespan by the longest
// item.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"
#include "reduce_common.hpp"

namespace pdlhip {

template <class D, class S> __device__ inline D conv_to(S v) {
  if constexpr (!IsFloat<D>::v && (D)-1 > (D)0) return (D)(intmax_t)v;
  else return (D)v;
}

template <class T> __device__ inline T load_as(const char *p, int type, int64_t idx) {
  switch (type) {
#define X(TC) case TC: return conv_to<T>(as_global((const CType<TC>::T *)p)[idx]);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return T(0);
  }
}

// U elements of one slot: element e = base + j*64 (+ the lane's offset in
// base); all lanes read the same address for a 0-dim slot (inc 0)
template <class T, class S, int U, bool TAIL>
__device__ __forceinline__ void load_u(T *v, const char *p, int64_t inc, int64_t base, int64_t end, bool stream) {
  const auto q = as_global((const S *)p);
  if (inc == 0) {
    const T x = conv_to<T>(q[0]);
#pragma unroll
    for (int j = 0; j < U; j++) v[j] = x;
  } else if (inc == 1 && stream) {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = base + (int64_t)j * 64;
      v[j] = (!TAIL || e < end) ? conv_to<T>(__builtin_nontemporal_load(q + e)) : T(0);
    }
  } else {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = base + (int64_t)j * 64;
      v[j] = (!TAIL || e < end) ? conv_to<T>(q[e * inc]) : T(0);
    }
  }
}

template <class T, int U, bool TAIL>
__device__ __forceinline__ void load_slot(T *v, int type, const char *p, int64_t inc, int64_t base, int64_t end,
                                          bool stream) {
  switch (type) {
#define X(TC) case TC: load_u<T, CType<TC>::T, U, TAIL>(v, p, inc, base, end, stream); break;
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: break;
  }
}

template <int OP, class T, int U> __device__ __forceinline__ void apply_u(T *acc, const T *x, bool swap) {
  if (swap) {
#pragma unroll
    for (int j = 0; j < U; j++) acc[j] = binop<OP, T>(x[j], acc[j]);
  } else {
#pragma unroll
    for (int j = 0; j < U; j++) acc[j] = binop<OP, T>(acc[j], x[j]);
  }
}

template <class T, int U> __device__ __forceinline__ void run_op_u(int op, T *acc, const T *x, bool swap) {
  switch (op) {
  case OP_PLUS: apply_u<OP_PLUS, T, U>(acc, x, swap); break;
  case OP_MULT: apply_u<OP_MULT, T, U>(acc, x, swap); break;
  case OP_MINUS: apply_u<OP_MINUS, T, U>(acc, x, swap); break;
  case OP_DIVIDE: apply_u<OP_DIVIDE, T, U>(acc, x, swap); break;
  case OP_GT: apply_u<OP_GT, T, U>(acc, x, swap); break;
  case OP_LT: apply_u<OP_LT, T, U>(acc, x, swap); break;
  case OP_LE: apply_u<OP_LE, T, U>(acc, x, swap); break;
  case OP_GE: apply_u<OP_GE, T, U>(acc, x, swap); break;
  case OP_EQ: apply_u<OP_EQ, T, U>(acc, x, swap); break;
  case OP_NE: apply_u<OP_NE, T, U>(acc, x, swap); break;
  case OP_POWER: apply_u<OP_POWER, T, U>(acc, x, swap); break;
  case OP_ATAN2: apply_u<OP_ATAN2, T, U>(acc, x, swap); break;
  case OP_MODULO: apply_u<OP_MODULO, T, U>(acc, x, swap); break;
  case OP_SPACESHIP: apply_u<OP_SPACESHIP, T, U>(acc, x, swap); break;
  default: break;
  }
}

template <class T> __device__ inline T run_op(int op, T a, T b) {
  switch (op) {
  case OP_PLUS: return binop<OP_PLUS, T>(a, b);
  case OP_MULT: return binop<OP_MULT, T>(a, b);
  case OP_MINUS: return binop<OP_MINUS, T>(a, b);
  case OP_DIVIDE: return binop<OP_DIVIDE, T>(a, b);
  case OP_GT: return binop<OP_GT, T>(a, b);
  case OP_LT: return binop<OP_LT, T>(a, b);
  case OP_LE: return binop<OP_LE, T>(a, b);
  case OP_GE: return binop<OP_GE, T>(a, b);
  case OP_EQ: return binop<OP_EQ, T>(a, b);
  case OP_NE: return binop<OP_NE, T>(a, b);
  case OP_SHL: return binop<OP_SHL, T>(a, b);
  case OP_SHR: return binop<OP_SHR, T>(a, b);
  case OP_OR: return binop<OP_OR, T>(a, b);
  case OP_AND: return binop<OP_AND, T>(a, b);
  case OP_XOR: return binop<OP_XOR, T>(a, b);
  case OP_POWER: return binop<OP_POWER, T>(a, b);
  case OP_ATAN2: return binop<OP_ATAN2, T>(a, b);
  case OP_MODULO: return binop<OP_MODULO, T>(a, b);
  case OP_SPACESHIP: return binop<OP_SPACESHIP, T>(a, b);
  default: return a;
  }
}

// a node's value in a mixed chain, kept as 64 bits: integers sign- or
// zero-extended, float and double as their own bits
template <class S> __device__ inline uint64_t to_raw64(S v) {
  if constexpr (sizeof(S) == 4 && IsFloat<S>::v) return (uint64_t)__float_as_uint(v);
  else if constexpr (IsFloat<S>::v) return (uint64_t)__double_as_longlong(v);
  else if constexpr ((S)-1 > (S)0) return (uint64_t)v;
  else return (uint64_t)(int64_t)v;
}
template <class S> __device__ inline S from_raw64(uint64_t r) {
  if constexpr (sizeof(S) == 4 && IsFloat<S>::v) return __uint_as_float((uint32_t)r);
  else if constexpr (IsFloat<S>::v) return __longlong_as_double((long long)r);
  else return (S)r;
}
template <class T> __device__ inline T raw_as(uint64_t r, int type) {
  switch (type) {
#define X(TC) case TC: return conv_to<T>(from_raw64<CType<TC>::T>(r));
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return T(0);
  }
}

// operand x of a node, converted to the node's type S (converttypei,
// pdlconv.c:163-174): slot x >= 0, or node -x-1's value
template <class S>
__device__ inline S mixed_operand(const RagProg &pg, const RagItem &it, int64_t i, int x, const uint64_t *v) {
  return x >= 0 ? load_as<S>(it.ptr[x], pg.type[x], i * it.inc[x]) : raw_as<S>(v[-x - 1], pg.ntype[-x - 1]);
}

// value of a mixed chain at element i: each node runs in its own type, as
// PDL runs each op of the chain with its own trans type
template <class T> __device__ inline T chain_value_mixed(const RagProg &pg, const RagItem &it, int64_t i) {
  uint64_t v[RAG_MAXNODE];
  for (int n = 0; n < pg.nnodes; n++) {
    const int x = pg.node[n].x, y = pg.node[n].y, op = pg.node[n].op;
    switch (pg.ntype[n]) {
#define X(TC)                                                                                   \
  case TC: {                                                                                    \
    typedef CType<TC>::T S;                                                                     \
    v[n] = to_raw64<S>(run_op<S>(op, mixed_operand<S>(pg, it, i, x, v), mixed_operand<S>(pg, it, i, y, v))); \
    break;                                                                                      \
  }
      PDLHIP_FOR_EACH_TYPE(X)
#undef X
    default: v[n] = 0; break;
    }
  }
  return raw_as<T>(v[pg.nnodes - 1], pg.ntype[pg.nnodes - 1]);
}

// general program: value of the chain at element i of item `it`
template <class T> __device__ inline T chain_value(const RagProg &pg, const RagItem &it, int64_t i) {
  T in[RAG_MAXIN];
#pragma unroll
  for (int k = 0; k < RAG_MAXIN; k++)
    if (k < pg.nin) in[k] = load_as<T>(it.ptr[k], pg.type[k], i * it.inc[k]);
  if (pg.nnodes == 0) return in[0];
  T v[RAG_MAXNODE];
  for (int n = 0; n < pg.nnodes; n++) {
    const int x = pg.node[n].x, y = pg.node[n].y;
    const T a = x >= 0 ? in[x] : v[-x - 1];
    const T b = y >= 0 ? in[y] : v[-y - 1];
    v[n] = run_op<T>(pg.node[n].op, a, b);
  }
  return v[pg.nnodes - 1];
}

// U chain values of a left-deep program at elements base + j*64: each node
// loads its slot and applies its op (loading every slot up front doubled the
// registers and made the kernel slower)
template <class T, int U, bool TAIL>
__device__ __forceinline__ void chain_lin(T *acc, const RagProg &pg, const RagItem &it, int64_t base, int64_t end) {
  load_slot<T, U, TAIL>(acc, pg.type[pg.first], it.ptr[pg.first], it.inc[pg.first], base, end, pg.stream != 0);
  for (int n = 0; n < pg.nnodes; n++) {
    T x[U];
    const int s = pg.lslot[n];
    load_slot<T, U, TAIL>(x, pg.type[s], it.ptr[s], it.inc[s], base, end, pg.stream != 0);
    run_op_u<T, U>(pg.node[n].op, acc, x, pg.lswap[n] != 0);
  }
}

// Preloaded form of chain_lin for chains of at most RAG_PRE slots whose
// operand types are all at most 4 bytes: every slot's U elements are loaded
// as raw 32-bit words first, so a step costs one memory round trip instead
// of one per node, and converted only when their node runs.
constexpr int RAG_PRE = 5;

// raw word of an element as its load instruction leaves it (sign- or
// zero-extended), so nothing has to wait for the data before conversion
template <class S> __device__ __forceinline__ uint32_t raw_bits(S v) {
  if constexpr (IsFloat<S>::v) {
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    return u;
  } else if constexpr ((S)-1 < (S)0) {
    return (uint32_t)(int32_t)v;
  } else {
    return (uint32_t)v;
  }
}

template <class S, int U, bool TAIL>
__device__ __forceinline__ void load_raw_u(uint32_t *raw, const char *p, int64_t inc, int64_t base, int64_t end,
                                           bool stream) {
  const auto q = as_global((const S *)p);
  if (inc == 0) {
    raw[0] = raw_bits(q[0]);   // broadcast when converted
  } else if (inc == 1 && stream) {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = base + (int64_t)j * 64;
      raw[j] = (!TAIL || e < end) ? raw_bits(__builtin_nontemporal_load(q + e)) : 0u;
    }
  } else {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = base + (int64_t)j * 64;
      raw[j] = (!TAIL || e < end) ? raw_bits(q[e * inc]) : 0u;
    }
  }
}

template <int U, bool TAIL>
__device__ __forceinline__ void load_raw(uint32_t *raw, int type, const char *p, int64_t inc, int64_t base,
                                         int64_t end, bool stream) {
  switch (type) {
#define X(TC)                                                                       \
  case TC:                                                                          \
    if constexpr (sizeof(CType<TC>::T) <= 4) load_raw_u<CType<TC>::T, U, TAIL>(raw, p, inc, base, end, stream); \
    break;
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: break;
  }
}

template <class T, int U> __device__ __forceinline__ void conv_raw(T *v, int type, int64_t inc, const uint32_t *raw) {
  switch (type) {
#define X(TC)                                                      \
  case TC:                                                         \
    if constexpr (sizeof(CType<TC>::T) <= 4) {                     \
      _Pragma("unroll") for (int j = 0; j < U; j++) {              \
        CType<TC>::T s;                                            \
        __builtin_memcpy(&s, &raw[inc == 0 ? 0 : j], sizeof(s));   \
        v[j] = conv_to<T>(s);                                      \
      }                                                            \
    }                                                              \
    break;
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: break;
  }
}

template <class T, int U, bool TAIL>
__device__ __forceinline__ void chain_lin_pre(T *acc, const RagProg &pg, const RagItem &it, int64_t base,
                                              int64_t end) {
  uint32_t raw[RAG_PRE][U];
#pragma unroll
  for (int s = 0; s < RAG_PRE; s++) {
    if (s > pg.nnodes) break;
    const int slot = s == 0 ? pg.first : pg.lslot[s - 1];
    load_raw<U, TAIL>(raw[s], pg.type[slot], it.ptr[slot], it.inc[slot], base, end, pg.stream != 0);
  }
  conv_raw<T, U>(acc, pg.type[pg.first], it.inc[pg.first], raw[0]);
#pragma unroll
  for (int n = 0; n < RAG_PRE - 1; n++) {
    if (n >= pg.nnodes) break;
    T x[U];
    conv_raw<T, U>(x, pg.type[pg.lslot[n]], it.inc[pg.lslot[n]], raw[n + 1]);
    run_op_u<T, U>(pg.node[n].op, acc, x, pg.lswap[n] != 0);
  }
}

// the chain's values at elements base + j*64 into the item's value output
template <class T, int U, bool TAIL>
__device__ __forceinline__ void store_u(const T *v, char *p, int64_t inc, int64_t base, int64_t end, bool stream) {
  const auto q = as_global((T *)p);
  if (inc == 1 && stream) {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = base + (int64_t)j * 64;
      if (!TAIL || e < end) __builtin_nontemporal_store(v[j], q + e);
    }
  } else {
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = base + (int64_t)j * 64;
      if (!TAIL || e < end) q[e * inc] = v[j];
    }
  }
}

template <int OP, class T, class Tout, class Tacc>
__device__ inline void seg_finish(const Part<T, Tacc> &p, const RagProg &pg, const RagItem &it, int32_t item,
                                  int32_t *empty) {
  Outs o;
  o.p[0] = it.out;
  o.bv[0] = pg.out_bv;
  int64_t off[2] = {0, 0};
  part_store<OP, T, Tout, Tacc>(p, o, off, false);
  if constexpr (is_minmax(OP)) empty[item] = p.i0 < 0 ? 1 : 0;
}

template <class T, int OP, class Tout, class Tacc, int U, bool LIN, bool PRE = false>
__global__ __launch_bounds__(256) void rag_reduce(RagProg pg, const RagItem *__restrict__ items,
                                                  const RagSeg *__restrict__ segs, int64_t nseg,
                                                  Part<T, Tacc> *partials, int32_t *empty) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t s = (int64_t)blockIdx.x * 4 + wv; s < nseg; s += nw) {
    const RagSeg sg = segs[s];
    const RagItem &it = items[sg.item];
    constexpr bool RED = OP != RAG_NOREDUCE;
    Part<T, Tacc> p;
    if constexpr (RED) part_init<OP, T, Tacc>(p);
    if constexpr (LIN) {
      constexpr int64_t STEP = 64 * U;
      int64_t b = sg.start;
      for (; b + STEP <= sg.end; b += STEP) {
        T v[U];
        if constexpr (PRE) chain_lin_pre<T, U, false>(v, pg, it, b + lane, sg.end);
        else chain_lin<T, U, false>(v, pg, it, b + lane, sg.end);
        if (pg.store) store_u<T, U, false>(v, it.val, it.val_inc, b + lane, sg.end, pg.stream != 0);
        if constexpr (RED) {
#pragma unroll
          for (int j = 0; j < U; j++) part_push<OP, T, Tacc>(p, v[j], b + lane + (int64_t)j * 64);
        }
      }
      if (b < sg.end) {
        T v[U];
        if constexpr (PRE) chain_lin_pre<T, U, true>(v, pg, it, b + lane, sg.end);
        else chain_lin<T, U, true>(v, pg, it, b + lane, sg.end);
        if (pg.store) store_u<T, U, true>(v, it.val, it.val_inc, b + lane, sg.end, pg.stream != 0);
        if constexpr (RED) {
#pragma unroll
          for (int j = 0; j < U; j++) {
            const int64_t e = b + lane + (int64_t)j * 64;
            if (e < sg.end) part_push<OP, T, Tacc>(p, v[j], e);
          }
        }
      }
    } else {
      for (int64_t i = sg.start + lane; i < sg.end; i += 64) {
        const T x = pg.mixed ? chain_value_mixed<T>(pg, it, i) : chain_value<T>(pg, it, i);
        if (pg.store) as_global((T *)it.val)[i * it.val_inc] = x;
        if constexpr (RED) part_push<OP, T, Tacc>(p, x, i);
      }
    }
    if constexpr (RED) {
      wave_reduce<OP, T, Tacc>(p);
      if (lane == 0) {
        if (sg.single) seg_finish<OP, T, Tout, Tacc>(p, pg, it, sg.item, empty);
        else partials[s] = p;
      }
    }
  }
}

template <class T, int OP, class Tout, class Tacc>
__global__ __launch_bounds__(256) void rag_combine(RagProg pg, const RagItem *__restrict__ items,
                                                   const int32_t *__restrict__ multi, int64_t nmulti,
                                                   const Part<T, Tacc> *__restrict__ partials, int32_t *empty) {
  const int lane = threadIdx.x & 63;
  const int64_t m = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= nmulti) return;
  const RagItem &it = items[multi[m]];
  Part<T, Tacc> p;
  part_init<OP, T, Tacc>(p);
  for (int64_t k = lane; k < it.nseg; k += 64) part_combine<OP>(p, partials[it.seg0 + k]);
  wave_reduce<OP, T, Tacc>(p);
  if (lane == 0) seg_finish<OP, T, Tout, Tacc>(p, pg, it, multi[m], empty);
}

// ---------------------------------------------------------------- scans
constexpr int RS_IT = RAG_SEG / 256;
__device__ inline int rs_pad(int i) { return i + (i >> 4); }

// one chunk (<= RAG_SEG elements) of an item into registers, element
// tid + 256*k in pf[k]; ident past the item's end
template <class Tin, class Tout, int IT>
__device__ inline void scan_fetch(Tout *pf, const RagItem &it, int64_t c0, int tid, Tout ident) {
  const auto ap = as_global((const Tin *)it.ptr[0]);
  const int64_t inc = it.inc[0];
  const int n = (int)(it.len - c0 < RAG_SEG ? it.len - c0 : RAG_SEG);
  if (inc == 1 && n == RAG_SEG) {
#pragma unroll
    for (int k = 0; k < IT; k++) pf[k] = (Tout)__builtin_nontemporal_load(ap + c0 + tid + 256 * k);
  } else {
#pragma unroll
    for (int k = 0; k < IT; k++) {
      const int idx = tid + 256 * k;
      pf[k] = idx < n ? (Tout)ap[(c0 + idx) * inc] : ident;
    }
  }
}

// next non-empty chunk of the block's item list after (q, c0)
__device__ inline void scan_next(const RagItem *items, const int32_t *order, int qend, int &q, int64_t &c0) {
  c0 += RAG_SEG;
  while (q < qend && c0 >= items[order[q]].len) { q++; c0 = 0; }
}

template <class Tin, class Tout, bool PR>
__global__ __launch_bounds__(256) void rag_scan(const RagItem *__restrict__ items, const int32_t *__restrict__ order,
                                                const int32_t *__restrict__ bstart) {
  __shared__ Tout sh[RAG_SEG + RAG_SEG / 16];
  __shared__ Tout wsum[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const Tout ident = PR ? Tout(1) : Tout(0);
  const int qend = bstart[blockIdx.x + 1];
  int q = bstart[blockIdx.x];
  int64_t c0 = 0;
  while (q < qend && items[order[q]].len == 0) q++;
  Tout pf[RS_IT];
  if (q < qend) scan_fetch<Tin, Tout, RS_IT>(pf, items[order[q]], c0, tid, ident);
  Tout carry = ident;
  while (q < qend) {
    const RagItem &it = items[order[q]];
    const auto bp = as_global((Tout *)it.out);
    const int64_t binc = it.out_inc;
    const int n = (int)(it.len - c0 < RAG_SEG ? it.len - c0 : RAG_SEG);
    if (c0 == 0) carry = ident;
#pragma unroll
    for (int k = 0; k < RS_IT; k++) sh[rs_pad(tid + 256 * k)] = pf[k];
    // the next chunk's loads are in flight while this one is scanned
    int q2 = q;
    int64_t c2 = c0;
    scan_next(items, order, qend, q2, c2);
    if (q2 < qend) scan_fetch<Tin, Tout, RS_IT>(pf, items[order[q2]], c2, tid, ident);
    __syncthreads();
    // each thread scans RS_IT consecutive elements, then the wave and the
    // block combine the thread totals
    Tout run = ident;
    const int base = tid * RS_IT;
#pragma unroll
    for (int j = 0; j < RS_IT; j++) {
      run = PR ? wmul(run, sh[rs_pad(base + j)]) : wadd(run, sh[rs_pad(base + j)]);
      sh[rs_pad(base + j)] = run;
    }
    Tout incl = run;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      Tout t = shfl_up(incl, d);
      if (lane >= d) incl = PR ? wmul(t, incl) : wadd(t, incl);
    }
    Tout excl = shfl_up(incl, 1);
    if (lane == 0) excl = ident;
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    Tout pre = carry, tot = carry;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k < w) pre = PR ? wmul(pre, wsum[k]) : wadd(pre, wsum[k]);
      tot = PR ? wmul(tot, wsum[k]) : wadd(tot, wsum[k]);
    }
    pre = PR ? wmul(pre, excl) : wadd(pre, excl);
#pragma unroll
    for (int j = 0; j < RS_IT; j++) {
      Tout v = sh[rs_pad(base + j)];
      sh[rs_pad(base + j)] = PR ? wmul(pre, v) : wadd(pre, v);
    }
    __syncthreads();
    if (binc == 1 && n == RAG_SEG) {
#pragma unroll
      for (int k = 0; k < RS_IT; k++) __builtin_nontemporal_store(sh[rs_pad(tid + 256 * k)], bp + c0 + tid + 256 * k);
    } else {
#pragma unroll
      for (int k = 0; k < RS_IT; k++) {
        const int idx = tid + 256 * k;
        if (idx < n) bp[(c0 + idx) * binc] = sh[rs_pad(idx)];
      }
    }
    carry = tot;
    __syncthreads();
    q = q2;
    c0 = c2;
  }
}

// ---------------------------------------------------------------- launchers
template <class T, int OP, class Tout, class Tacc>
static hipError_t rag_red(const RagLaunch &L, hipStream_t s) {
  int64_t blocks = (L.nseg + 3) / 4;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) blocks = 1;
  Part<T, Tacc> *part = (Part<T, Tacc> *)L.scratch;
  constexpr int U = sizeof(T) == 4 ? 8 : 4;
  bool pre = L.prog.lin && L.prog.nnodes + 1 <= RAG_PRE;
  for (int k = 0; k < L.prog.nin && pre; k++) pre = type_size(L.prog.type[k]) <= 4;
  if (pre)
    hipLaunchKernelGGL((rag_reduce<T, OP, Tout, Tacc, U, true, true>), dim3((unsigned)blocks), dim3(256), 0, s,
                       L.prog, L.items, L.segs, L.nseg, part, L.empty);
  else if (L.prog.lin)
    hipLaunchKernelGGL((rag_reduce<T, OP, Tout, Tacc, U, true>), dim3((unsigned)blocks), dim3(256), 0, s, L.prog,
                       L.items, L.segs, L.nseg, part, L.empty);
  else
    hipLaunchKernelGGL((rag_reduce<T, OP, Tout, Tacc, 1, false>), dim3((unsigned)blocks), dim3(256), 0, s, L.prog,
                       L.items, L.segs, L.nseg, part, L.empty);
  if constexpr (OP != RAG_NOREDUCE) {
    if (L.nmulti > 0)
      hipLaunchKernelGGL((rag_combine<T, OP, Tout, Tacc>), dim3((unsigned)((L.nmulti + 3) / 4)), dim3(256), 0, s,
                         L.prog, L.items, L.multi, L.nmulti, (const Part<T, Tacc> *)part, L.empty);
  }
  return hipGetLastError();
}

template <class T>
static hipError_t rag_red_T(int op, const RagLaunch &L, hipStream_t s) {
  typedef typename IPlus<T>::O TI;
  switch (op) {
  case RAG_NOREDUCE: return rag_red<T, RAG_NOREDUCE, T, T>(L, s);
  case R_SUM: return rag_red<T, R_SUM, TI, TI>(L, s);
  case R_DSUM: return rag_red<T, R_DSUM, double, double>(L, s);
  case R_AVERAGE: return rag_red<T, R_AVERAGE, TI, TI>(L, s);
  case R_DAVERAGE: return rag_red<T, R_DAVERAGE, double, double>(L, s);
  case R_MIN: return rag_red<T, R_MIN, T, T>(L, s);
  case R_MAX: return rag_red<T, R_MAX, T, T>(L, s);
  default: return hipErrorInvalidValue;
  }
}

static_assert(sizeof(Part<double, double>) <= 64, "ragged partial slots are 64 bytes");

hipError_t launch_ragged_reduce(int T, int op, const RagLaunch &L, hipStream_t s) {
  if (L.nseg == 0) return hipSuccess;
  switch (T) {
  case PDLHIP_F: return rag_red_T<float>(op, L, s);
  case PDLHIP_D: return rag_red_T<double>(op, L, s);
  default: return hipErrorInvalidValue;
  }
}

template <class Tin, bool PR>
static hipError_t rag_scan_T(const RagLaunch &L, hipStream_t s) {
  typedef typename IPlus<Tin>::O TO;
  hipLaunchKernelGGL((rag_scan<Tin, TO, PR>), dim3((unsigned)L.nblk), dim3(256), 0, s, L.items, L.order, L.bstart);
  return hipGetLastError();
}

hipError_t launch_ragged_scan(int T, int op, const RagLaunch &L, hipStream_t s) {
  if (L.nblk == 0) return hipSuccess;
  const bool pr = (op == R_CUMUPROD);
  switch (T) {
#define X(TC) case TC: return pr ? rag_scan_T<CType<TC>::T, true>(L, s) : rag_scan_T<CType<TC>::T, false>(L, s);
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8)
#undef X
  default: return hipErrorInvalidValue;
  }
}

}  // namespace pdlhip
