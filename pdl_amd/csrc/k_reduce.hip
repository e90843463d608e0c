// Reductions and scans of lib/PDL/Ufunc.pd over the core dim n (dim 0 of a),
// one output element per broadcast position:
//   sumover/prodover (+d)       Ufunc.pd:88-118
//   cumusumover/cumuprodover    Ufunc.pd:120-140
//   average/daverage            Ufunc.pd:413-444
//   minimum/maximum(+_ind)      Ufunc.pd:446-482
//   minimum_n_ind/maximum_n_ind Ufunc.pd:484-540
//   minmaximum                  Ufunc.pd:563-613
//
// Kernels:
//   wave   one 64-lane wavefront per output (rows with unit stride), 16-byte
//          vector loads, cross-lane combine with shuffles.
//   split  long rows and few outputs: blocks reduce row chunks into partials
//          in scratch, a second kernel combines them.
//   thread one thread per output walking n in index order (strided rows,
//          scans, float products, *_n_ind): adjacent threads read adjacent
//          addresses when the broadcast dim 0 is the contiguous one.
//
// Semantics kept exactly: integer sums/products wrap in the output type
// (any order gives PDL's bits), min/max follow PDL's rules (a NaN never
// displaces a non-NaN, a non-NaN displaces a NaN, ties keep the first,
// all-NaN gives the last), bad values are skipped, empty/all-bad rows give
// BAD (or 0 / NaN for average in good mode).  Float sums are compensated
// (Neumaier per lane, double-double across lanes and blocks), so they are
// at least as accurate as PDL's sequential float loop; float products and
// float scans run in PDL's exact sequential order.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"
#include "reduce_common.hpp"

namespace pdlhip {

void *scratch_alloc(size_t bytes);
void scratch_free(void *p);

__device__ inline void out_offsets(const RedPlan &p, int64_t e, int64_t *off) {
  for (int k = 0; k < p.npdl; k++) off[k] = p.off[k];
  for (int d = 0; d < p.nd; d++) {
    int64_t i = e % p.dims[d];
    e /= p.dims[d];
    for (int k = 0; k < p.npdl; k++) off[k] += i * p.inc[k][d];
  }
}

// ---------------------------------------------------------------- wave
constexpr int RT = 256;  // threads per block (4 waves)

template <int OP, class Tin, class Tout, class Tacc, bool BAD, int VEC>
__device__ inline void row_segment(Part<Tin, Tacc> &pt, const Tin *__restrict__ a, int64_t o, int64_t begin,
                                   int64_t end, int64_t inc, int lane, const BadInfo &ba, bool stream) {
  const Tin abv = bv_as<Tin>(ba.bv);
  int64_t i = begin + (int64_t)lane * VEC;
  // 16-byte loads in flight per lane: 4 for 4- and 8-byte elements, 2 for
  // the narrower ones (eight or sixteen elements per load already)
  constexpr int RED_U = sizeof(Tin) >= 4 ? 4 : 2;
  if constexpr (VEC > 1) {
    // begin is a multiple of VEC and the row start is 16-byte aligned; RED_U
    // vectors in flight per lane, consumed in increasing index order
    for (; i + (RED_U - 1) * 64 * VEC + VEC <= end; i += RED_U * 64 * VEC) {
      Vec<Tin, VEC> x[RED_U];
#pragma unroll
      for (int u = 0; u < RED_U; u++)
        x[u] = stream ? load_nt<Tin, VEC>(a + o + i + u * 64 * VEC)
                      : *reinterpret_cast<const Vec<Tin, VEC> *>(a + o + i + u * 64 * VEC);
#pragma unroll
      for (int u = 0; u < RED_U; u++)
#pragma unroll
        for (int k = 0; k < VEC; k++) {
          if constexpr (BAD) { if (isbad(x[u].v[k], abv, ba.isnan)) continue; }
          part_push<OP, Tin, Tacc>(pt, x[u].v[k], i + u * 64 * VEC + k);
        }
    }
    for (; i + VEC <= end; i += 64 * VEC) {
      Vec<Tin, VEC> x = stream ? load_nt<Tin, VEC>(a + o + i) : *reinterpret_cast<const Vec<Tin, VEC> *>(a + o + i);
#pragma unroll
      for (int k = 0; k < VEC; k++) {
        if constexpr (BAD) { if (isbad(x.v[k], abv, ba.isnan)) continue; }
        part_push<OP, Tin, Tacc>(pt, x.v[k], i + k);
      }
    }
    // tail: fewer than VEC*64 elements remain past the vector loop's reach
    for (int64_t t = ((end - begin) / VEC) * VEC + begin + lane; t < end; t += 64) {
      Tin x = a[o + t];
      if constexpr (BAD) { if (isbad(x, abv, ba.isnan)) continue; }
      part_push<OP, Tin, Tacc>(pt, x, t);
    }
  } else {
    for (; i < end; i += 64) {
      Tin x = a[o + i * inc];
      if constexpr (BAD) { if (isbad(x, abv, ba.isnan)) continue; }
      part_push<OP, Tin, Tacc>(pt, x, i);
    }
  }
}

template <int OP, class Tin, class Tout, class Tacc, bool BAD, int VEC>
__global__ __launch_bounds__(RT) void red_wave(const Tin *__restrict__ a, RedPlan p, Outs o, BadInfo ba,
                                               bool stream) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = (int64_t)gridDim.x * (RT / 64);
  for (int64_t e = (int64_t)blockIdx.x * (RT / 64) + (threadIdx.x >> 6); e < p.nout; e += nw) {
    int64_t off[MAXP];
    out_offsets(p, e, off);
    Part<Tin, Tacc> pt;
    part_init<OP, Tin, Tacc>(pt);
    row_segment<OP, Tin, Tout, Tacc, BAD, VEC>(pt, a, off[0], 0, p.n, p.inc_n[0], lane, ba, stream);
    wave_reduce<OP, Tin, Tacc>(pt);
    if (lane == 0) part_store<OP, Tin, Tout, Tacc>(pt, o, off, BAD);
  }
}

// ---------------------------------------------------------------- split
// stage 1: block (e, k) reduces chunk k of row e into parts[e * nchunk + k]
template <int OP, class Tin, class Tout, class Tacc, bool BAD, int VEC>
__global__ __launch_bounds__(RT) void red_split1(const Tin *__restrict__ a, RedPlan p, int64_t nchunk, int64_t clen,
                                                 Part<Tin, Tacc> *parts, BadInfo ba, bool stream) {
  __shared__ Part<Tin, Tacc> sh[RT / 64];
  const int64_t e = blockIdx.x / nchunk, k = blockIdx.x % nchunk;
  int64_t off[MAXP];
  out_offsets(p, e, off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t cb = k * clen, ce = (cb + clen < p.n) ? cb + clen : p.n;
  Part<Tin, Tacc> pt;
  part_init<OP, Tin, Tacc>(pt);
  // the 4 waves of the block take interleaved sub-chunks of the chunk
  const int64_t sub = ((ce - cb) + 3) / 4;
  const int64_t subr = ((sub + VEC - 1) / VEC) * VEC;
  const int64_t sb = cb + w * subr, se = (sb + subr < ce) ? sb + subr : ce;
  if (sb < se) row_segment<OP, Tin, Tout, Tacc, BAD, VEC>(pt, a, off[0], sb, se, p.inc_n[0], lane, ba, stream);
  wave_reduce<OP, Tin, Tacc>(pt);
  if (lane == 0) sh[w] = pt;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 1; j < RT / 64; j++) part_combine<OP>(pt, sh[j]);
    parts[blockIdx.x] = pt;
  }
}

// stage 2: one wave per row combines its chunk partials
template <int OP, class Tin, class Tout, class Tacc, bool BAD>
__global__ __launch_bounds__(RT) void red_split2(RedPlan p, int64_t nchunk, const Part<Tin, Tacc> *parts, Outs o) {
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * (RT / 64) + (threadIdx.x >> 6);
  if (e >= p.nout) return;
  int64_t off[MAXP];
  out_offsets(p, e, off);
  Part<Tin, Tacc> pt;
  part_init<OP, Tin, Tacc>(pt);
  for (int64_t k = lane; k < nchunk; k += 64) part_combine<OP>(pt, parts[e * nchunk + k]);
  wave_reduce<OP, Tin, Tacc>(pt);
  if (lane == 0) part_store<OP, Tin, Tout, Tacc>(pt, o, off, BAD);
}

// ---------------------------------------------------------------- scan
// Integer cumusumover/cumuprodover with unit-stride rows: one block per row,
// chunks of 4096 staged through LDS (coalesced in and out), a sequential scan
// of 16 elements per thread, then wave and block prefix combines.  Wrapping
// integer + and * are associative, so the result has PDL's exact bits.  A bad
// element contributes the identity and is written as the output's bad value
// (Ufunc.pd:127-132).
constexpr int SC_IT = 16, SC_CH = RT * SC_IT;
__device__ inline int sc_pad(int i) { return i + (i >> 4); }

template <int OP, class Tin, class Tout, bool BAD>
__global__ __launch_bounds__(RT) void scan_rows(const Tin *__restrict__ a, RedPlan p, Outs o, BadInfo ba) {
  constexpr bool PR = (OP == R_CUMUPROD);
  __shared__ Tout sh[SC_CH + SC_CH / 16];
  __shared__ uint8_t badm[BAD ? SC_CH : 1];
  __shared__ Tout wsum[RT / 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const Tin abv = bv_as<Tin>(ba.bv);
  const Tout obv = bv_as<Tout>(o.bv[0]);
  const Tout ident = PR ? Tout(1) : Tout(0);
  for (int64_t e = blockIdx.x; e < p.nout; e += gridDim.x) {
    int64_t off[MAXP];
    out_offsets(p, e, off);
    const Tin *ap = a + off[0];
    Tout *bp = (Tout *)o.p[0] + off[1];
    Tout carry = ident;
    for (int64_t c0 = 0; c0 < p.n; c0 += SC_CH) {
      const int len = (p.n - c0 < SC_CH) ? (int)(p.n - c0) : SC_CH;
      // every load of the chunk is issued before the first LDS write
      Tin xv[SC_IT];
#pragma unroll
      for (int k = 0; k < SC_IT; k++) {
        const int idx = tid + RT * k;
        xv[k] = idx < len ? ap[c0 + idx] : Tin(0);
      }
#pragma unroll
      for (int k = 0; k < SC_IT; k++) {
        const int idx = tid + RT * k;
        Tout v = ident;
        if (idx < len) {
          bool isb = false;
          if constexpr (BAD) { isb = isbad(xv[k], abv, ba.isnan); badm[idx] = isb; }
          if (!isb) v = (Tout)xv[k];
        }
        sh[sc_pad(idx)] = v;
      }
      __syncthreads();
      // sequential scan of this thread's 16 consecutive elements
      Tout run = ident;
      const int base = tid * SC_IT;
#pragma unroll
      for (int j = 0; j < SC_IT; j++) {
        run = PR ? wmul(run, sh[sc_pad(base + j)]) : wadd(run, sh[sc_pad(base + j)]);
        sh[sc_pad(base + j)] = run;
      }
      // inclusive wave scan of the thread totals
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
      Tout wpre = ident, tot = ident;
#pragma unroll
      for (int k = 0; k < RT / 64; k++) {
        if (k < w) wpre = PR ? wmul(wpre, wsum[k]) : wadd(wpre, wsum[k]);
        tot = PR ? wmul(tot, wsum[k]) : wadd(tot, wsum[k]);
      }
      Tout pre = PR ? wmul(wmul(carry, wpre), excl) : wadd(wadd(carry, wpre), excl);
#pragma unroll
      for (int j = 0; j < SC_IT; j++) {
        Tout v = sh[sc_pad(base + j)];
        sh[sc_pad(base + j)] = PR ? wmul(pre, v) : wadd(pre, v);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < SC_IT; k++) {
        const int idx = tid + RT * k;
        if (idx < len) {
          Tout v = sh[sc_pad(idx)];
          if constexpr (BAD) { if (badm[idx]) v = obv; }
          bp[c0 + idx] = v;
        }
      }
      carry = PR ? wmul(carry, tot) : wadd(carry, tot);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- thread
template <int OP, class Tin, class Tout, class Tacc, bool BAD>
__global__ __launch_bounds__(RT) void red_thread(const Tin *__restrict__ a, RedPlan p, Outs o, BadInfo ba,
                                                 int *flag) {
  const Tin abv = bv_as<Tin>(ba.bv);
  for (int64_t e = (int64_t)blockIdx.x * RT + threadIdx.x; e < p.nout; e += (int64_t)gridDim.x * RT) {
    int64_t off[MAXP];
    out_offsets(p, e, off);
    const Tin *ap = a + off[0];
    const int64_t n = p.n, ia = p.inc_n[0];
    if constexpr (OP == R_CUMUSUM || OP == R_DCUMUSUM || OP == R_CUMUPROD || OP == R_DCUMUPROD) {
      // Ufunc.pd:120-140 in index order (PDL's exact rounding for floats)
      Tout *bp = (Tout *)o.p[0] + off[1];
      const int64_t ib = p.inc_n[1];
      const Tout obv = bv_as<Tout>(o.bv[0]);
      constexpr bool PR = (OP == R_CUMUPROD || OP == R_DCUMUPROD);
      Tout tmp = PR ? Tout(1) : Tout(0);
      for (int64_t i = 0; i < n; i++) {
        Tin x = ap[i * ia];
        if constexpr (BAD) { if (isbad(x, abv, ba.isnan)) { bp[i * ib] = obv; continue; } }
        tmp = PR ? wmul(tmp, (Tout)x) : wadd(tmp, (Tout)x);
        bp[i * ib] = tmp;
      }
    } else if constexpr (OP == R_MIN_N_IND || OP == R_MAX_N_IND) {
      // Ufunc.pd:484-540: m passes, each skipping indices already chosen
      int64_t *cp = (int64_t *)o.p[0] + off[1];
      const int64_t ic = p.inc_m, m = p.m;
      const int64_t cbv = bv_as<int64_t>(o.bv[0]);
      Tin cur = 0;
      for (int64_t mi = 0; mi < m; mi++) {
        int64_t curind = n;
        for (int64_t i = 0; i < n; i++) {
          bool taken = false;
          for (int64_t mm = 0; mm < mi; mm++) if (cp[mm * ic] == i) { taken = true; break; }
          if (taken) continue;
          Tin x = ap[i * ia];
          if (BAD && isbad(x, abv, ba.isnan)) continue;
          bool better = (OP == R_MIN_N_IND) ? (x < cur) : (x > cur);
          if (curind == n || better || is_nan(cur)) { cur = x; curind = i; }
        }
        if (curind != n) cp[mi * ic] = curind;
        else {
          cp[mi * ic] = cbv;
          // PDL clears c's bad flag at the start of every broadcast iteration
          // and sets it on a shortfall, so the last iteration decides
          if (flag && e == p.nout - 1) atomicOr(flag, 1);
        }
      }
    } else if constexpr (is_prod(OP) && IsFloat<Tacc>::v) {
      // Ufunc.pd:104-108: sequential, stop at a zero running product
      Tacc tmp = 1;
      bool any = false;
      for (int64_t i = 0; i < n; i++) {
        Tin x = ap[i * ia];
        if constexpr (BAD) { if (isbad(x, abv, ba.isnan)) continue; }
        any = true;
        tmp *= (Tacc)x;
        if (tmp == 0) break;
      }
      Tout v = (Tout)tmp;
      if (BAD && !any) v = bv_as<Tout>(o.bv[0]);
      put<Tout>(o.p[0], off[1], v);
    } else {
      Part<Tin, Tacc> pt;
      part_init<OP, Tin, Tacc>(pt);
      for (int64_t i = 0; i < n; i++) {
        Tin x = ap[i * ia];
        if constexpr (BAD) { if (isbad(x, abv, ba.isnan)) continue; }
        part_push<OP, Tin, Tacc>(pt, x, i);
      }
      part_store<OP, Tin, Tout, Tacc>(pt, o, off, BAD);
    }
  }
}

// ---------------------------------------------------------------- dispatch
template <int OP, class Tin, class Tout, class Tacc, bool BAD>
static hipError_t red_launch(const RedPlan &p, const Tin *a, const Outs &o, const BadInfo &ba, int *flag,
                             hipStream_t s) {
  if (p.nout == 0) return hipSuccess;
  constexpr bool SCAN = (OP == R_CUMUSUM || OP == R_DCUMUSUM || OP == R_CUMUPROD || OP == R_DCUMUPROD);
  constexpr bool SEQ = SCAN || OP == R_MIN_N_IND || OP == R_MAX_N_IND || (is_prod(OP) && IsFloat<Tacc>::v);
  if constexpr (!SEQ) {
    if (p.n >= 64 && p.inc_n[0] == 1) {
      constexpr int VEC = 16 / sizeof(Tin);
      bool aligned = ((uintptr_t)(a + p.off[0]) % 16 == 0);
      for (int d = 0; d < p.nd; d++) aligned = aligned && (p.inc[0][d] % VEC == 0);
      const bool stream = streaming_read((size_t)p.n * (size_t)p.nout * sizeof(Tin));
      if (p.nout < 1024 && p.n >= 65536) {
        // long rows, few outputs: split each row into chunks
        int64_t want = (2048 + p.nout - 1) / p.nout;
        int64_t clen = (p.n + want - 1) / want;
        if (clen < 16384) clen = 16384;
        clen = ((clen + VEC * 256 - 1) / (VEC * 256)) * (VEC * 256);
        int64_t nchunk = (p.n + clen - 1) / clen;
        // partials in the caller's scratch when it brings one, else from the cache
        const size_t need = sizeof(Part<Tin, Tacc>) * (size_t)p.nout * (size_t)nchunk;
        Part<Tin, Tacc> *parts;
        if (p.scratch) {
          if (need > (size_t)p.scratch_bytes) return hipErrorOutOfMemory;
          parts = (Part<Tin, Tacc> *)p.scratch;
        } else {
          parts = (Part<Tin, Tacc> *)scratch_alloc(need);
          if (!parts) return hipErrorOutOfMemory;
        }
        if (aligned)
          hipLaunchKernelGGL((red_split1<OP, Tin, Tout, Tacc, BAD, VEC>), dim3((unsigned)(p.nout * nchunk)), dim3(RT),
                             0, s, a, p, nchunk, clen, parts, ba, stream);
        else
          hipLaunchKernelGGL((red_split1<OP, Tin, Tout, Tacc, BAD, 1>), dim3((unsigned)(p.nout * nchunk)), dim3(RT),
                             0, s, a, p, nchunk, clen, parts, ba, stream);
        hipLaunchKernelGGL((red_split2<OP, Tin, Tout, Tacc, BAD>), dim3((unsigned)((p.nout + 3) / 4)), dim3(RT), 0, s,
                           p, nchunk, parts, o);
        hipError_t err = hipGetLastError();
        if (!p.scratch) scratch_free(parts);
        return err;
      }
      int64_t blocks = (p.nout + 3) / 4;
      if (blocks > 65536) blocks = 65536;
      if (aligned)
        hipLaunchKernelGGL((red_wave<OP, Tin, Tout, Tacc, BAD, VEC>), dim3((unsigned)blocks), dim3(RT), 0, s, a, p, o,
                           ba, stream);
      else
        hipLaunchKernelGGL((red_wave<OP, Tin, Tout, Tacc, BAD, 1>), dim3((unsigned)blocks), dim3(RT), 0, s, a, p, o,
                           ba, stream);
      return hipGetLastError();
    }
  }
  if constexpr ((OP == R_CUMUSUM || OP == R_CUMUPROD) && !IsFloat<Tout>::v) {
    if (p.inc_n[0] == 1 && p.inc_n[1] == 1 && p.n >= 256) {
      int64_t blocks = p.nout < 65536 ? p.nout : 65536;
      hipLaunchKernelGGL((scan_rows<OP, Tin, Tout, BAD>), dim3((unsigned)blocks), dim3(RT), 0, s, a, p, o, ba);
      return hipGetLastError();
    }
  }
  int64_t blocks = (p.nout + RT - 1) / RT;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((red_thread<OP, Tin, Tout, Tacc, BAD>), dim3((unsigned)blocks), dim3(RT), 0, s, a, p, o, ba,
                     flag);
  return hipGetLastError();
}

template <int OP, class Tin, class Tout, class Tacc>
static hipError_t red_bad(bool bad, const RedPlan &p, const Tin *a, const Outs &o, const BadInfo &ba, int *flag,
                          hipStream_t s) {
  return bad ? red_launch<OP, Tin, Tout, Tacc, true>(p, a, o, ba, flag, s)
             : red_launch<OP, Tin, Tout, Tacc, false>(p, a, o, ba, flag, s);
}

template <class Tin>
static hipError_t reduce_T(int op, const RedPlan &p, const void *va, void *const *outs, bool bad, const BadInfo &ba,
                           const pdlhip_value *obv, int *flag, hipStream_t s) {
  typedef typename IPlus<Tin>::O TI;
  const Tin *a = (const Tin *)va;
  Outs o;
  for (int k = 0; k < 4; k++) { o.p[k] = outs[k]; o.bv[k] = obv[k]; }
  switch (op) {
  case R_SUM: return red_bad<R_SUM, Tin, TI, TI>(bad, p, a, o, ba, flag, s);
  case R_PROD: return red_bad<R_PROD, Tin, TI, TI>(bad, p, a, o, ba, flag, s);
  case R_DSUM: return red_bad<R_DSUM, Tin, double, double>(bad, p, a, o, ba, flag, s);
  case R_DPROD: return red_bad<R_DPROD, Tin, double, double>(bad, p, a, o, ba, flag, s);
  case R_CUMUSUM: return red_bad<R_CUMUSUM, Tin, TI, TI>(bad, p, a, o, ba, flag, s);
  case R_CUMUPROD: return red_bad<R_CUMUPROD, Tin, TI, TI>(bad, p, a, o, ba, flag, s);
  case R_DCUMUSUM: return red_bad<R_DCUMUSUM, Tin, double, double>(bad, p, a, o, ba, flag, s);
  case R_DCUMUPROD: return red_bad<R_DCUMUPROD, Tin, double, double>(bad, p, a, o, ba, flag, s);
  case R_AVERAGE: return red_bad<R_AVERAGE, Tin, TI, TI>(bad, p, a, o, ba, flag, s);
  case R_DAVERAGE: return red_bad<R_DAVERAGE, Tin, double, double>(bad, p, a, o, ba, flag, s);
  case R_MIN: return red_bad<R_MIN, Tin, Tin, Tin>(bad, p, a, o, ba, flag, s);
  case R_MAX: return red_bad<R_MAX, Tin, Tin, Tin>(bad, p, a, o, ba, flag, s);
  case R_MIN_IND: return red_bad<R_MIN_IND, Tin, int64_t, Tin>(bad, p, a, o, ba, flag, s);
  case R_MAX_IND: return red_bad<R_MAX_IND, Tin, int64_t, Tin>(bad, p, a, o, ba, flag, s);
  case R_MINMAXIMUM: return red_bad<R_MINMAXIMUM, Tin, Tin, Tin>(bad, p, a, o, ba, flag, s);
  case R_MIN_N_IND: return red_bad<R_MIN_N_IND, Tin, int64_t, Tin>(bad, p, a, o, ba, flag, s);
  case R_MAX_N_IND: return red_bad<R_MAX_N_IND, Tin, int64_t, Tin>(bad, p, a, o, ba, flag, s);
  default: return hipErrorInvalidValue;
  }
}

#define PDLHIP_RED_TYPE(TC) \
  hipError_t launch_reduce_t##TC(int op, int otype, const RedPlan &p, const void *a, void *const *outs, bool bad, \
                                 const BadInfo &ba, const pdlhip_value *obv, int *flag, hipStream_t s) { \
    (void)otype; \
    return reduce_T<CType<TC>::T>(op, p, a, outs, bad, ba, obv, flag, s); \
  }
PDLHIP_RED_TYPES_IN_THIS_TU

}  // namespace pdlhip
