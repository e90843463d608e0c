// PDL's random, randsym and grandom (Primitive.pd:2142-2262) on the device,
// drawing PDL's own stream.
//
// PDL seeds xoshiro256+ (Primitive-xoshiro256plus.c:28-43) with four
// splitmix64 outputs of the seed (pdl_srand, :111-130) and each element of
// the output, in broadcast order (dim 0 fastest), takes the next
// pdl_drand = (x >> 11) * 2^-53 (:132-138).  With more than one CPU thread
// PDL gives each pthread its own state; the stream here is the single one
// PDL draws with autopthread off.
//
// The stream is sequential, but xoshiro256's state transition is linear over
// GF(2): 256 bits in, 256 bits out, s' = M s.  So the state k draws ahead is
// M^k s, and the host builds M^(2^i) once by repeated squaring.  A launch
// splits the n draws into chunks of 4096; the host steps the state from
// chunk to chunk (one 256x256 bit matrix-vector product, M^4096, per chunk)
// and each of a block's 256 threads jumps to its own 16 draws with the
// matrices M^(16*2^k), k < 8, held in LDS (threadIdx.x's binary digits
// select them), then runs the generator itself.  The bits are those of the
// sequential loop.
//
//   random:  a = pdl_drand(), converted to the output type
//   randsym: the same draw, which PDL redraws when it is exactly 0
//            (probability 2^-53 per element); a launch that meets one
//            reports it and the host fails the call rather than shift the
//            stream
//   grandom: ndtri(randsym(x)) in x's type (Primitive.pd:2256-2261): the
//            randsym value is rounded to the type first, then
//            PDL::Math::ndtri (Math.pd:346-356) takes it as a double and the
//            result is converted back; ndtri is Cephes' inverse of the
//            normal distribution function (Math/ndtri.c, S. L. Moshier):
//            w + w^3 R(w^2)/S(w^2) around the middle, z - log(z)/z -
//            P(1/z)/(z Q(1/z)) with z = sqrt(-2 log y) in the tails, with
//            Cephes' published coefficients.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "launch.hpp"

namespace pdlhip {

constexpr int RNG_THREADS = 256;
constexpr int RNG_PER_THREAD = 16;
constexpr int RNG_LOG_JUMPS = 8;  // 256 threads: M^(16 * 2^k), k < 8

// ---------------------------------------------------------------- host: GF(2)
namespace {

typedef uint64_t Col[4];
struct Mat { Col c[256]; };  // column j = M e_j (bit j = bit j%64 of word j/64)

void xo_step(uint64_t s[4]) {
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = (s[3] << 45) | (s[3] >> 19);
}

void matvec(const Mat &m, const uint64_t x[4], uint64_t y[4]) {
  uint64_t r[4] = {0, 0, 0, 0};
  for (int w = 0; w < 4; w++) {
    uint64_t v = x[w];
    while (v) {
      const int b = __builtin_ctzll(v);
      v &= v - 1;
      const Col &c = m.c[64 * w + b];
      r[0] ^= c[0]; r[1] ^= c[1]; r[2] ^= c[2]; r[3] ^= c[3];
    }
  }
  std::memcpy(y, r, sizeof r);
}

struct Tables {
  Mat pow2[64];  // M^(2^i)
  std::vector<uint64_t> jumps;  // M^(16 * 2^k), k < 8, as [k][j][4]
};

const Tables &tables() {
  static Tables *t = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    Tables *r = new Tables;
    for (int j = 0; j < 256; j++) {
      uint64_t s[4] = {0, 0, 0, 0};
      s[j >> 6] = 1ull << (j & 63);
      xo_step(s);
      std::memcpy(r->pow2[0].c[j], s, sizeof s);
    }
    for (int i = 1; i < 64; i++)
      for (int j = 0; j < 256; j++) matvec(r->pow2[i - 1], r->pow2[i - 1].c[j], r->pow2[i].c[j]);
    r->jumps.resize((size_t)RNG_LOG_JUMPS * 256 * 4);
    for (int k = 0; k < RNG_LOG_JUMPS; k++)
      std::memcpy(&r->jumps[(size_t)k * 1024], r->pow2[k + 4].c, sizeof(Mat));
    t = r;
  });
  return *t;
}

}  // namespace

void rng_seed(uint64_t seed, uint64_t s[4]) {
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    s[i] = z ^ (z >> 31);
  }
}

void rng_advance(uint64_t s[4], uint64_t n) {
  const Tables &t = tables();
  for (int i = 0; n; i++, n >>= 1)
    if (n & 1) matvec(t.pow2[i], s, s);
}

void rng_chunk_states(const uint64_t s[4], int64_t nchunks, uint64_t *out) {
  const Mat &step = tables().pow2[12];  // 4096 draws per chunk
  uint64_t cur[4];
  std::memcpy(cur, s, sizeof cur);
  for (int64_t c = 0; c < nchunks; c++) {
    std::memcpy(out + 4 * c, cur, sizeof cur);
    if (c + 1 < nchunks) matvec(step, cur, cur);
  }
}

const uint64_t *rng_jump_table(size_t *nwords) {
  const Tables &t = tables();
  *nwords = t.jumps.size();
  return t.jumps.data();
}

// ---------------------------------------------------------------- device
// Cephes ndtri's coefficient tables (Math/ndtri.c): P0/Q0 for
// |y - 0.5| <= 3/8, P1/Q1 for z = sqrt(-2 log y) in [2, 8), P2/Q2 for z >= 8;
// the Q tables omit their leading 1.
__constant__ double NDTRI_P0[5] = {-5.99633501014107895267E1, 9.80010754185999661536E1, -5.66762857469070293439E1,
                                   1.39312609387279679503E1, -1.23916583867381258016E0};
__constant__ double NDTRI_Q0[8] = {1.95448858338141759834E0, 4.67627912898881538453E0, 8.63602421390890590575E1,
                                   -2.25462687854119370527E2, 2.00260212380060660359E2, -8.20372256168333339912E1,
                                   1.59056225126211695515E1, -1.18331621121330003142E0};
__constant__ double NDTRI_P1[9] = {4.05544892305962419923E0, 3.15251094599893866154E1, 5.71628192246421288162E1,
                                   4.40805073893200834700E1, 1.46849561928858024014E1, 2.18663306850790267539E0,
                                   -1.40256079171354495875E-1, -3.50424626827848203418E-2, -8.57456785154685413611E-4};
__constant__ double NDTRI_Q1[8] = {1.57799883256466749731E1, 4.53907635128879210584E1, 4.13172038254672030440E1,
                                   1.50425385692907503408E1, 2.50464946208309415979E0, -1.42182922854787788574E-1,
                                   -3.80806407691578277194E-2, -9.33259480895457427372E-4};
__constant__ double NDTRI_P2[9] = {3.23774891776946035970E0, 6.91522889068984211695E0, 3.93881025292474443415E0,
                                   1.33303460815807542389E0, 2.01485389549179081538E-1, 1.23716634817820021358E-2,
                                   3.01581553508235416007E-4, 2.65806974686737550832E-6, 6.23974539184983293730E-9};
__constant__ double NDTRI_Q2[8] = {6.02427039364742014255E0, 3.67983563856160859403E0, 1.37702099489081330271E0,
                                   2.16236993594496635890E-1, 1.34204006088543189037E-2, 3.28014464682127739104E-4,
                                   2.89247864745380683936E-6, 6.79019408009981274425E-9};

// Horner from the leading coefficient (Cephes polevl), and with an implicit
// leading 1 (p1evl); no contraction, as the CPU build rounds each step
template <int N> __device__ __forceinline__ double polevl(double x, const double *c) {
  double r = c[0];
#pragma unroll
  for (int i = 1; i <= N; i++) r = r * x + c[i];
  return r;
}
template <int N> __device__ __forceinline__ double p1evl(double x, const double *c) {
  double r = x + c[0];
#pragma unroll
  for (int i = 1; i < N; i++) r = r * x + c[i];
  return r;
}

__device__ double ndtri(double y0) {
  const double EXPM2 = 0.13533528323661269189;  // exp(-2)
  const double S2PI = 2.50662827463100050242E0; // sqrt(2 pi)
  if (y0 <= 0.0) return -1.79769313486231570815E308;  // Cephes MAXNUM
  if (y0 >= 1.0) return 1.79769313486231570815E308;
  bool neg = true;
  double y = y0;
  if (y > 1.0 - EXPM2) {
    y = 1.0 - y;
    neg = false;
  }
  if (y > EXPM2) {
    y = y - 0.5;
    const double y2 = y * y;
    double x = y + y * (y2 * polevl<4>(y2, NDTRI_P0) / p1evl<8>(y2, NDTRI_Q0));
    return x * S2PI;
  }
  const double x = sqrt(-2.0 * log(y));
  const double x0 = x - log(x) / x;
  const double z = 1.0 / x;
  const double x1 = x < 8.0 ? z * polevl<8>(z, NDTRI_P1) / p1evl<8>(z, NDTRI_Q1)
                            : z * polevl<8>(z, NDTRI_P2) / p1evl<8>(z, NDTRI_Q2);
  const double r = x0 - x1;
  return neg ? -r : r;
}

// y = J x over GF(2), J in LDS as 256 columns of 4 words; branch-free, every
// lane reads the same column (an LDS broadcast)
__device__ __forceinline__ void jump(const uint64_t (*J)[4], uint64_t s[4]) {
  uint64_t y0 = 0, y1 = 0, y2 = 0, y3 = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    const uint64_t v = s[w];
#pragma unroll 8
    for (int b = 0; b < 64; b++) {
      const uint64_t m = 0 - ((v >> b) & 1);
      const uint64_t *c = J[64 * w + b];
      y0 ^= c[0] & m;
      y1 ^= c[1] & m;
      y2 ^= c[2] & m;
      y3 ^= c[3] & m;
    }
  }
  s[0] = y0; s[1] = y1; s[2] = y2; s[3] = y3;
}

// the element offset of logical element e (dim 0 fastest)
__device__ __forceinline__ int64_t rng_offset(const RngPlan &p, int64_t e) {
  if (p.dense) return p.off + e;
  int64_t o = p.off;
  for (int d = 0; d < p.nd; d++) {
    const int64_t i = e % p.dims[d];
    e /= p.dims[d];
    o += i * p.incs[d];
  }
  return o;
}

template <class T, int KIND>
__global__ __launch_bounds__(RNG_THREADS) void k_rng(T *__restrict__ out, RngPlan p,
                                                     const uint64_t *__restrict__ chunk_state,
                                                     const uint64_t *__restrict__ jumps, int64_t nchunks,
                                                     int *__restrict__ zero_draw) {
  __shared__ uint64_t J[RNG_LOG_JUMPS][256][4];  // 64 KiB
  for (int i = threadIdx.x; i < RNG_LOG_JUMPS * 256 * 4; i += RNG_THREADS) (&J[0][0][0])[i] = jumps[i];
  __syncthreads();
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    uint64_t s[4] = {chunk_state[4 * c], chunk_state[4 * c + 1], chunk_state[4 * c + 2], chunk_state[4 * c + 3]};
#pragma unroll 1
    for (int k = 0; k < RNG_LOG_JUMPS; k++) {
      uint64_t t[4] = {s[0], s[1], s[2], s[3]};
      jump(J[k], t);
      if ((threadIdx.x >> k) & 1) {
        s[0] = t[0]; s[1] = t[1]; s[2] = t[2]; s[3] = t[3];
      }
    }
    const int64_t e0 = c * (RNG_THREADS * RNG_PER_THREAD) + (int64_t)threadIdx.x * RNG_PER_THREAD;
#pragma unroll
    for (int j = 0; j < RNG_PER_THREAD; j++) {
      const uint64_t x = s[0] + s[3];
      const uint64_t t = s[1] << 17;
      s[2] ^= s[0];
      s[3] ^= s[1];
      s[1] ^= s[2];
      s[0] ^= s[3];
      s[2] ^= t;
      s[3] = (s[3] << 45) | (s[3] >> 19);
      const int64_t e = e0 + j;
      if (e >= p.n) break;
      const double v = (double)(x >> 11) * 0x1.0p-53;
      T r;
      if constexpr (KIND == 0) {
        r = (T)v;
      } else {
        if (v == 0.0) atomicOr(zero_draw, 1);
        if constexpr (KIND == 1) r = (T)v;
        else r = (T)ndtri((double)(T)v);
      }
      out[rng_offset(p, e)] = r;
    }
  }
}

template <class T>
static hipError_t rng_go(int kind, const RngPlan &p, void *out, const uint64_t *d_states, const uint64_t *d_jumps,
                         int64_t nchunks, int *d_flag, hipStream_t s) {
  int64_t grid = nchunks < 2048 ? nchunks : 2048;
  if (grid < 1) return hipSuccess;
  if (kind == 0)
    hipLaunchKernelGGL((k_rng<T, 0>), dim3((unsigned)grid), dim3(RNG_THREADS), 0, s, (T *)out, p, d_states, d_jumps,
                       nchunks, d_flag);
  else if (kind == 1)
    hipLaunchKernelGGL((k_rng<T, 1>), dim3((unsigned)grid), dim3(RNG_THREADS), 0, s, (T *)out, p, d_states, d_jumps,
                       nchunks, d_flag);
  else {
    if constexpr (std::is_floating_point<T>::value)
      hipLaunchKernelGGL((k_rng<T, 2>), dim3((unsigned)grid), dim3(RNG_THREADS), 0, s, (T *)out, p, d_states,
                         d_jumps, nchunks, d_flag);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_rng(int T, int kind, const RngPlan &p, void *out, const uint64_t *d_states, const uint64_t *d_jumps,
                      int64_t nchunks, int *d_flag, hipStream_t s) {
  switch (T) {
#define X(TC) case TC: return rng_go<CType<TC>::T>(kind, p, out, d_states, d_jumps, nchunks, d_flag, s);
    PDLHIP_FOR_EACH_TYPE(X)
#undef X
  default: return hipErrorInvalidValue;
  }
}

}  // namespace pdlhip
