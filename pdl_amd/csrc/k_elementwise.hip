// Elementwise kernels for the Ops.pd binops (lib/PDL/Ops.pd:104-219),
// assgn (382-397) and dtype conversion (lib/PDL/Core/pdlconv.c:163-174),
// over a canonical EwPlan built from PDL's broadcast plan.
//
// Three shapes:
//   flat  every operand contiguous or a broadcast scalar: 16-byte vector
//         loads and stores, grid-stride.
//   tile  an input is transposed against the output (its unit stride lies on
//         another dim, e.g. x->xchg(0,1)): 64x64 tiles staged through LDS so
//         that both the transposed read and the output write are coalesced.
//   rows  anything else: one block walks a chunk of one output row (dim 0),
//         the row's offsets decomposed once per block; handles bad values.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "launch.hpp"

namespace pdlhip {

constexpr int EW_THREADS = 256;

// ---------------------------------------------------------------- flat
template <class T, int OP, int VEC>
__global__ __launch_bounds__(EW_THREADS) void ew_flat(const T *__restrict__ a, const T *__restrict__ b,
                                                     T *__restrict__ c, int64_t n, int sa, int sb, bool stream) {
  const int64_t nvec = n / VEC;
  const int64_t stride = (int64_t)gridDim.x * EW_THREADS;
  const T a0 = sa ? a[0] : T(0), b0 = sb ? b[0] : T(0);
  for (int64_t v = (int64_t)blockIdx.x * EW_THREADS + threadIdx.x; v < nvec; v += stride) {
    Vec<T, VEC> av, bv, cv;
    if (sa) {
#pragma unroll
      for (int k = 0; k < VEC; k++) av.v[k] = a0;
    } else {
      av = stream ? load_nt<T, VEC>(a + v * VEC) : reinterpret_cast<const Vec<T, VEC> *>(a)[v];
    }
    if (sb) {
#pragma unroll
      for (int k = 0; k < VEC; k++) bv.v[k] = b0;
    } else {
      bv = stream ? load_nt<T, VEC>(b + v * VEC) : reinterpret_cast<const Vec<T, VEC> *>(b)[v];
    }
#pragma unroll
    for (int k = 0; k < VEC; k++) cv.v[k] = binop<OP, T>(av.v[k], bv.v[k]);
    if (stream) store_nt<T, VEC>(c + v * VEC, cv);
    else reinterpret_cast<Vec<T, VEC> *>(c)[v] = cv;
  }
  // tail
  int64_t t = nvec * VEC + (int64_t)blockIdx.x * EW_THREADS + threadIdx.x;
  if (t < n && t < nvec * VEC + VEC) {
    T x = sa ? a0 : a[t];
    T y = sb ? b0 : b[t];
    c[t] = binop<OP, T>(x, y);
  }
}

// ---------------------------------------------------------------- rows
// Block b handles chunk (b % cpr) of row (b / cpr); `chunk` elements.
template <class T, int OP, bool BAD>
__global__ __launch_bounds__(EW_THREADS) void ew_rows(const T *__restrict__ a, const T *__restrict__ b,
                                                     T *__restrict__ c, EwPlan p, int64_t cpr, int64_t chunk,
                                                     BadInfo ba, BadInfo bb, T cbad) {
  const int64_t row = blockIdx.x / cpr;
  const int64_t i_begin = (blockIdx.x % cpr) * chunk;
  int64_t i_end = i_begin + chunk;
  if (i_end > p.dims[0]) i_end = p.dims[0];
  int64_t o[3];
  row_offsets(p, row, o);
  const int64_t ia = p.inc[0][0], ib = p.inc[1][0], ic = p.inc[2][0];
  const T abv = bv_as<T>(ba.bv), bbv = bv_as<T>(bb.bv);
  for (int64_t i = i_begin + threadIdx.x; i < i_end; i += EW_THREADS) {
    T x = a[o[0] + i * ia];
    T y = b[o[1] + i * ib];
    if constexpr (BAD) {
      if ((ba.check && isbad(x, abv, ba.isnan)) || (bb.check && isbad(y, bbv, bb.isnan))) {
        c[o[2] + i * ic] = cbad;
        continue;
      }
    }
    c[o[2] + i * ic] = binop<OP, T>(x, y);
  }
}

// small dim 0 and many rows: one thread per element, flat index decomposed
template <class T, int OP, bool BAD>
__global__ __launch_bounds__(EW_THREADS) void ew_elems(const T *__restrict__ a, const T *__restrict__ b,
                                                      T *__restrict__ c, EwPlan p, BadInfo ba, BadInfo bb,
                                                      T cbad) {
  const T abv = bv_as<T>(ba.bv), bbv = bv_as<T>(bb.bv);
  for (int64_t e = (int64_t)blockIdx.x * EW_THREADS + threadIdx.x; e < p.n; e += (int64_t)gridDim.x * EW_THREADS) {
    int64_t r = e, o0 = p.off[0], o1 = p.off[1], o2 = p.off[2];
    for (int d = 0; d < p.nd; d++) {
      int64_t i = r % p.dims[d];
      r /= p.dims[d];
      o0 += i * p.inc[0][d]; o1 += i * p.inc[1][d]; o2 += i * p.inc[2][d];
    }
    T x = a[o0], y = b[o1];
    if constexpr (BAD) {
      if ((ba.check && isbad(x, abv, ba.isnan)) || (bb.check && isbad(y, bbv, bb.isnan))) {
        c[o2] = cbad;
        continue;
      }
    }
    c[o2] = binop<OP, T>(x, y);
  }
}

// ---------------------------------------------------------------- tile
// Tiles of TI (dim 0, output-contiguous) x TJ (dim q, where a transposed
// input is contiguous).  TMASK bit k: input k is read through LDS.  For
// 4- and 8-byte types each thread moves 16-byte vectors on both sides: the
// transposed input is read along j, the output written along i.
// Shapes measured at 1 GiB with the reductions that follow (streaming):
// 4-byte types 256 (i) x 32 (j) with one transposed input (1 KB runs on
// the write, 128-byte runs on the read: 6.37 TB/s, against 5.95 for 64 x
// 128 and 6.18 for 128 x 32) and 128 x 32 with two (6.41 against 6.02 for
// 32 x 128); 8-byte types 64 x 64 (6.3 TB/s, against 5.8-5.9 for 128 x 32,
// 128 x 16 and 64 x 32); others 64 x 64.  About 33 KB of LDS per block.
template <class T, int TMASK> struct TileShape {
  static constexpr int I = sizeof(T) == 4 ? (TMASK == 3 ? 128 : 256) : 64;
  static constexpr int J = sizeof(T) == 4 ? 32 : 64;
};
constexpr int TILE_GROUP = 8;   // tile-rows (along i) visited together, column after column

template <class T, int OP, int TMASK, int V, int TI = TileShape<T, TMASK>::I, int TJ = TileShape<T, TMASK>::J>
__global__ __launch_bounds__(EW_THREADS) void ew_tile(const T *__restrict__ a, const T *__restrict__ b,
                                                     T *__restrict__ c, EwPlan p, int q, int64_t ntile_i,
                                                     int64_t ntile_j, bool stream) {
  constexpr int NL = ((TMASK & 1) ? 1 : 0) + ((TMASK & 2) ? 1 : 0);
  __shared__ T lds[NL][TI][TJ + 1];
  // blocks walk groups of TILE_GROUP tile-rows column by column: the blocks
  // in flight read and write a compact band of both operands
  const int64_t per = ntile_i * ntile_j;
  int64_t rest = (int64_t)blockIdx.x / per, r = (int64_t)blockIdx.x % per;
  const int64_t grp = r / (TILE_GROUP * ntile_j);
  const int64_t gsz = ntile_i - grp * TILE_GROUP < TILE_GROUP ? ntile_i - grp * TILE_GROUP : TILE_GROUP;
  r -= grp * TILE_GROUP * ntile_j;
  const int64_t ti = grp * TILE_GROUP + r % gsz, tj = r / gsz;
  int64_t o[3] = {p.off[0], p.off[1], p.off[2]};
  for (int d = 1; d < p.nd; d++) {
    if (d == q) continue;
    int64_t i = rest % p.dims[d];
    rest /= p.dims[d];
    o[0] += i * p.inc[0][d]; o[1] += i * p.inc[1][d]; o[2] += i * p.inc[2][d];
  }
  const int64_t i0 = ti * TI, j0 = tj * TJ;
  const int64_t ni = p.dims[0], nj = p.dims[q];
  const bool full = (i0 + TI <= ni) && (j0 + TJ <= nj);
  const T *src[2] = {a, b};
  // load: lanes along j, V elements each
  constexpr int LJ = TJ / V;                  // threads per tile row
  constexpr int LR = EW_THREADS / LJ;         // rows per pass
  {
    const int jj = (threadIdx.x % LJ) * V;
    const int ii0 = threadIdx.x / LJ;
    if (V > 1 && full) {
      // every row pass's vector load is issued before the first LDS write
      constexpr int NP = TI / LR;
      Vec<T, V> xv[2][NP];
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (!(TMASK & (1 << k))) continue;
        const int64_t inc_i = p.inc[k][0];
        const T *sp = src[k] + o[k] + (i0 + ii0) * inc_i + j0 + jj;
        if (stream) {
#pragma unroll
          for (int q = 0; q < NP; q++) xv[k][q] = load_nt<T, V>(sp + (int64_t)q * LR * inc_i);
        } else {
#pragma unroll
          for (int q = 0; q < NP; q++) xv[k][q] = *reinterpret_cast<const Vec<T, V> *>(sp + (int64_t)q * LR * inc_i);
        }
      }
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (!(TMASK & (1 << k))) continue;
        const int slot = (k == 1 && (TMASK & 1)) ? 1 : 0;
#pragma unroll
        for (int q = 0; q < NP; q++)
#pragma unroll
          for (int e = 0; e < V; e++) lds[slot][ii0 + q * LR][jj + e] = xv[k][q].v[e];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (!(TMASK & (1 << k))) continue;
        const int slot = (k == 1 && (TMASK & 1)) ? 1 : 0;
        const int64_t inc_i = p.inc[k][0];
#pragma unroll
        for (int ii = ii0; ii < TI; ii += LR) {
          const int64_t gi = i0 + ii, gj = j0 + jj;
          const T *sp = src[k] + o[k] + gi * inc_i + gj;
#pragma unroll
          for (int e = 0; e < V; e++)
            if (gi < ni && gj + e < nj) lds[slot][ii][jj + e] = sp[e];
        }
      }
    }
  }
  __syncthreads();
  // compute + store: lanes along i, V elements each.  Pointers advance by a
  // fixed step per row of the tile; a directly read input that does not vary
  // along j (a dummy dim, increment 0) is loaded once.
  constexpr int LI = TI / V;
  constexpr int LC = EW_THREADS / LI;
  const int ii = (threadIdx.x % LI) * V;
  const int jj0 = threadIdx.x / LI;
  const int64_t gi = i0 + ii;
  if (V > 1 && full) {
    const int64_t cstep = LC * p.inc[2][q];
    T *cp = c + o[2] + gi + (j0 + jj0) * p.inc[2][q];
    const T *dp[2] = {nullptr, nullptr};
    int64_t dstep[2] = {0, 0};
    T dv[2][V];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (TMASK & (1 << k)) continue;
      dp[k] = src[k] + o[k] + gi * p.inc[k][0] + (j0 + jj0) * p.inc[k][q];
      dstep[k] = LC * p.inc[k][q];
      if (dstep[k] == 0) {
        if (p.inc[k][0] == 0) {
#pragma unroll
          for (int e = 0; e < V; e++) dv[k][e] = dp[k][0];
        } else {
          const Vec<T, V> v = *reinterpret_cast<const Vec<T, V> *>(dp[k]);
#pragma unroll
          for (int e = 0; e < V; e++) dv[k][e] = v.v[e];
        }
      }
    }
#pragma unroll
    for (int jj = jj0; jj < TJ; jj += LC) {
      T x[V], y[V];
#pragma unroll
      for (int k = 0; k < 2; k++) {
        T *dst = k == 0 ? x : y;
        if (TMASK & (1 << k)) {
          const int slot = (k == 1 && (TMASK & 1)) ? 1 : 0;
#pragma unroll
          for (int e = 0; e < V; e++) dst[e] = lds[slot][ii + e][jj];
        } else if (dstep[k] == 0) {
#pragma unroll
          for (int e = 0; e < V; e++) dst[e] = dv[k][e];
        } else if (p.inc[k][0] == 0) {
#pragma unroll
          for (int e = 0; e < V; e++) dst[e] = dp[k][0];
          dp[k] += dstep[k];
        } else {
          const Vec<T, V> v = *reinterpret_cast<const Vec<T, V> *>(dp[k]);
#pragma unroll
          for (int e = 0; e < V; e++) dst[e] = v.v[e];
          dp[k] += dstep[k];
        }
      }
      Vec<T, V> out;
#pragma unroll
      for (int e = 0; e < V; e++) out.v[e] = binop<OP, T>(x[e], y[e]);
      if (stream) store_nt<T, V>(cp, out);
      else *reinterpret_cast<Vec<T, V> *>(cp) = out;
      cp += cstep;
    }
  } else {
#pragma unroll
    for (int jj = jj0; jj < TJ; jj += LC) {
      const int64_t gj = j0 + jj;
      T x[V], y[V];
#pragma unroll
      for (int e = 0; e < V; e++) {
        if (gi + e >= ni || gj >= nj) continue;
        if constexpr (TMASK & 1) x[e] = lds[0][ii + e][jj];
        else x[e] = a[o[0] + (gi + e) * p.inc[0][0] + gj * p.inc[0][q]];
        if constexpr (TMASK & 2) y[e] = lds[(TMASK & 1) ? 1 : 0][ii + e][jj];
        else y[e] = b[o[1] + (gi + e) * p.inc[1][0] + gj * p.inc[1][q]];
        c[o[2] + (gi + e) * p.inc[2][0] + gj * p.inc[2][q]] = binop<OP, T>(x[e], y[e]);
      }
    }
  }
}

// ---------------------------------------------------------------- dispatch
static int grid_for(int64_t work, int64_t per_block) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

template <class T, int OP>
static hipError_t binop_T_OP(const EwPlan &p, const void *va, const void *vb, void *vc, bool bad,
                             const BadInfo *bi, hipStream_t s) {
  const T *a = (const T *)va;
  const T *b = (const T *)vb;
  T *c = (T *)vc;
  if (p.n == 0) return hipSuccess;
  const T cbad = bv_as<T>(bi[2].bv);
  if (!bad) {
    // flat: one dim, output contiguous, inputs contiguous or scalar
    if (p.nd == 1 && p.inc[2][0] == 1 && (p.inc[0][0] == 0 || p.inc[0][0] == 1) &&
        (p.inc[1][0] == 0 || p.inc[1][0] == 1)) {
      const T *pa = a + p.off[0], *pb = b + p.off[1];
      T *pc = c + p.off[2];
      constexpr int VEC = 16 / sizeof(T);
      bool aligned = ((uintptr_t)pc % 16 == 0) && (p.inc[0][0] == 0 || (uintptr_t)pa % 16 == 0) &&
                     (p.inc[1][0] == 0 || (uintptr_t)pb % 16 == 0);
      int64_t nthreads_needed = p.n / (aligned ? VEC : 1);
      int grid = grid_for(nthreads_needed, EW_THREADS * 4);
      if (grid < 256 * 4 && nthreads_needed > EW_THREADS * 256 * 4) grid = 1024;
      const bool stream = streaming_read((size_t)p.n * sizeof(T) * 3);
      if (aligned)
        hipLaunchKernelGGL((ew_flat<T, OP, VEC>), dim3(grid), dim3(EW_THREADS), 0, s, pa, pb, pc, p.n,
                           p.inc[0][0] == 0, p.inc[1][0] == 0, stream);
      else
        hipLaunchKernelGGL((ew_flat<T, OP, 1>), dim3(grid), dim3(EW_THREADS), 0, s, pa, pb, pc, p.n,
                           p.inc[0][0] == 0, p.inc[1][0] == 0, stream);
      return hipGetLastError();
    }
    // tile: output unit stride on dim 0; a transposed input has its unit stride on dim q
    if (p.nd >= 2 && p.inc[2][0] == 1) {
      int q = -1, tmask = 0;
      for (int k = 0; k < 2; k++) {
        if (p.inc[k][0] == 1 || p.inc[k][0] == 0) continue;
        for (int d = 1; d < p.nd; d++)
          if (p.inc[k][d] == 1 && (q < 0 || q == d)) { q = d; tmask |= 1 << k; }
      }
      if (q > 0 && tmask) {
        // tile counts of the shape the launched kernel uses
        const int TI = tmask == 3 ? TileShape<T, 3>::I : TileShape<T, 1>::I;
        const int TJ = tmask == 3 ? TileShape<T, 3>::J : TileShape<T, 1>::J;
        int64_t nti = (p.dims[0] + TI - 1) / TI, ntj = (p.dims[q] + TJ - 1) / TJ;
        int64_t nrest = p.n / (p.dims[0] * p.dims[q]);
        int64_t nblk = nti * ntj * nrest;
        // 16-byte vectors need every row start of every operand 16-byte aligned
        constexpr int V4 = (sizeof(T) == 4) ? 4 : ((sizeof(T) == 8) ? 2 : 1);
        bool vec = V4 > 1;
        for (int k = 0; k < 2; k++)  // an input read directly must be unit-stride or broadcast along i
          if (!(tmask & (1 << k)) && p.inc[k][0] != 0 && p.inc[k][0] != 1) vec = false;
        for (int k = 0; k < 3 && vec; k++) {
          const void *base = k == 0 ? (const void *)a : k == 1 ? (const void *)b : (const void *)c;
          if (((uintptr_t)base + p.off[k] * sizeof(T)) % 16) vec = false;
          for (int d = 0; d < p.nd && vec; d++) {
            bool unit = (k < 2 && (tmask & (1 << k))) ? (d == q) : (d == 0);
            if (!unit && p.inc[k][d] % V4) vec = false;
          }
        }
        if (nblk < (int64_t)1 << 31) {
          const bool stream = streaming_read((size_t)p.n * sizeof(T) * 2);
#define TILE_LAUNCH(TM, VV) hipLaunchKernelGGL((ew_tile<T, OP, TM, VV>), dim3((unsigned)nblk), dim3(EW_THREADS), 0, s, \
                                               a, b, c, p, q, nti, ntj, stream)
          if (vec) {
            if (tmask == 1) TILE_LAUNCH(1, V4); else if (tmask == 2) TILE_LAUNCH(2, V4); else TILE_LAUNCH(3, V4);
          } else {
            if (tmask == 1) TILE_LAUNCH(1, 1); else if (tmask == 2) TILE_LAUNCH(2, 1); else TILE_LAUNCH(3, 1);
          }
#undef TILE_LAUNCH
          return hipGetLastError();
        }
      }
    }
  }
  const int64_t rows = p.n / p.dims[0];
  if (p.dims[0] >= 128) {
    int64_t chunk = p.dims[0] < 8192 ? p.dims[0] : 8192;
    int64_t cpr = (p.dims[0] + chunk - 1) / chunk;
    int64_t nblk = rows * cpr;
    if (nblk < (int64_t)1 << 31) {
      if (bad)
        hipLaunchKernelGGL((ew_rows<T, OP, true>), dim3((unsigned)nblk), dim3(EW_THREADS), 0, s, a, b, c, p, cpr,
                           chunk, bi[0], bi[1], cbad);
      else
        hipLaunchKernelGGL((ew_rows<T, OP, false>), dim3((unsigned)nblk), dim3(EW_THREADS), 0, s, a, b, c, p, cpr,
                           chunk, bi[0], bi[1], cbad);
      return hipGetLastError();
    }
  }
  int grid = grid_for(p.n, EW_THREADS * 4);
  if (bad)
    hipLaunchKernelGGL((ew_elems<T, OP, true>), dim3(grid), dim3(EW_THREADS), 0, s, a, b, c, p, bi[0], bi[1], cbad);
  else
    hipLaunchKernelGGL((ew_elems<T, OP, false>), dim3(grid), dim3(EW_THREADS), 0, s, a, b, c, p, bi[0], bi[1], cbad);
  return hipGetLastError();
}

template <class T>
static hipError_t binop_T(int op, const EwPlan &p, const void *a, const void *b, void *c, bool bad,
                          const BadInfo *bi, hipStream_t s) {
  constexpr bool F = IsFloat<T>::v;
  switch (op) {
  case OP_PLUS: return binop_T_OP<T, OP_PLUS>(p, a, b, c, bad, bi, s);
  case OP_MULT: return binop_T_OP<T, OP_MULT>(p, a, b, c, bad, bi, s);
  case OP_MINUS: return binop_T_OP<T, OP_MINUS>(p, a, b, c, bad, bi, s);
  case OP_DIVIDE: return binop_T_OP<T, OP_DIVIDE>(p, a, b, c, bad, bi, s);
  case OP_GT: return binop_T_OP<T, OP_GT>(p, a, b, c, bad, bi, s);
  case OP_LT: return binop_T_OP<T, OP_LT>(p, a, b, c, bad, bi, s);
  case OP_LE: return binop_T_OP<T, OP_LE>(p, a, b, c, bad, bi, s);
  case OP_GE: return binop_T_OP<T, OP_GE>(p, a, b, c, bad, bi, s);
  case OP_EQ: return binop_T_OP<T, OP_EQ>(p, a, b, c, bad, bi, s);
  case OP_NE: return binop_T_OP<T, OP_NE>(p, a, b, c, bad, bi, s);
  case OP_MODULO: return binop_T_OP<T, OP_MODULO>(p, a, b, c, bad, bi, s);
  case OP_SPACESHIP: return binop_T_OP<T, OP_SPACESHIP>(p, a, b, c, bad, bi, s);
  case OP_SQRT: return binop_T_OP<T, OP_SQRT>(p, a, b, c, bad, bi, s);
  case OP_SIN: return binop_T_OP<T, OP_SIN>(p, a, b, c, bad, bi, s);
  case OP_COS: return binop_T_OP<T, OP_COS>(p, a, b, c, bad, bi, s);
  case OP_NOT: return binop_T_OP<T, OP_NOT>(p, a, b, c, bad, bi, s);
  case OP_LOG10: return binop_T_OP<T, OP_LOG10>(p, a, b, c, bad, bi, s);
  case OP_ABS: return binop_T_OP<T, OP_ABS>(p, a, b, c, bad, bi, s);
  default: break;
  }
  if constexpr (!F) {
    switch (op) {
    case OP_SHL: return binop_T_OP<T, OP_SHL>(p, a, b, c, bad, bi, s);
    case OP_SHR: return binop_T_OP<T, OP_SHR>(p, a, b, c, bad, bi, s);
    case OP_OR: return binop_T_OP<T, OP_OR>(p, a, b, c, bad, bi, s);
    case OP_AND: return binop_T_OP<T, OP_AND>(p, a, b, c, bad, bi, s);
    case OP_XOR: return binop_T_OP<T, OP_XOR>(p, a, b, c, bad, bi, s);
    case OP_BITNOT: return binop_T_OP<T, OP_BITNOT>(p, a, b, c, bad, bi, s);
    default: break;
    }
  } else {
    switch (op) {
    case OP_POWER: return binop_T_OP<T, OP_POWER>(p, a, b, c, bad, bi, s);
    case OP_ATAN2: return binop_T_OP<T, OP_ATAN2>(p, a, b, c, bad, bi, s);
    case OP_EXP: return binop_T_OP<T, OP_EXP>(p, a, b, c, bad, bi, s);
    case OP_LOG: return binop_T_OP<T, OP_LOG>(p, a, b, c, bad, bi, s);
    default: break;
    }
  }
  return hipErrorInvalidValue;
}

#define PDLHIP_EW_TYPE(TC) \
  hipError_t launch_binop_t##TC(int op, const EwPlan &p, const void *a, const void *b, void *c, bool bad, \
                                const BadInfo *bi, hipStream_t s) { \
    return binop_T<CType<TC>::T>(op, p, a, b, c, bad, bi, s); \
  }
PDLHIP_EW_TYPES_IN_THIS_TU

}  // namespace pdlhip
