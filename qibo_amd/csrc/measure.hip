// Probabilities, the exact sequential-equivalent cumulative sum, inverse-CDF
// search and sample post-processing, for gfx950.
//
// This translation unit is compiled with -ffp-contract=off: every floating
// point operation below must round exactly like numpy's, so no multiply may be
// fused into a following add behind our back. Fused multiply-adds appear only
// where numpy itself uses one (its complex abs). Division and square root are
// the compiler's correctly rounded IEEE operations (plain `/` and
// __builtin_sqrt[f]); HIP's __fsqrt_rn maps to the approximate native square
// root unless OCML_BASIC_ROUNDED_OPERATIONS is defined, so it is not used.
//
// Exact cumulative sum. numpy's cumsum (inside np.random.choice) is the
// sequential chain s_i = fl(s_{i-1} + p_i). A tree scan rounds differently and
// moves sampled indices. The chain is parallelised exactly as follows. While s
// stays inside one binade [2^e, 2^(e+1)) every double there is an integer
// multiple m of u = 2^(e-52), and for p >= 0
//     fl(s + p) = u * (m + c(p, m mod 2))
// where c is p/u rounded to nearest, a tie going to the value that makes
// m + c even. So each element is the map m -> m + c[m & 1], a pair of
// integers, and composing two such maps gives another one:
//     (g o f)[b] = f[b] + g[(b + f[b]) & 1].
// Composition is associative, so tiles of 4096 elements reduce to one map in
// parallel, groups of 64 tiles reduce to one map, a single wavefront chains
// the group maps (64 groups = 2^24 elements per step), the tile starts inside
// each group follow from a scan of its tile maps, and each tile then
// re-derives every thread's exact start value and finishes its 16 elements
// with ordinary IEEE additions, which are now exact replicas of numpy's. A tile is only handled by maps if an approximate prefix puts it well
// inside one binade, and the chain re-checks that exactly (start in the binade,
// end below 2^(e+1)); the few tiles where s crosses a power of two are summed
// element by element in order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "qsim_common.h"

namespace qsim {

// ---------------------------------------------------------------------------
// numpy's complex abs (loops_unary_complex: larger * sqrt(fma(r, r, 1)),
// r = smaller / larger, r = 0 when larger == 0 or smaller == inf), squared.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double np_abs2(c128 z) {
  double re = fabs(z.re), im = fabs(z.im);
  const double inf = __builtin_inf();
  if (re == inf) im = inf;
  if (im == inf) re = inf;
  if (re != re) im = re;
  if (im != im) re = im;
  const double l = fmax(re, im);
  const double sm = fmin(im, re);
  const double r = (l == 0.0 || sm == inf) ? 0.0 : sm / l;
  const double h = __builtin_sqrt(__builtin_fma(r, r, 1.0)) * l;
  return h * h;
}

__device__ __forceinline__ float np_abs2(c64 z) {
  float re = fabsf(z.re), im = fabsf(z.im);
  const float inf = __builtin_inff();
  if (re == inf) im = inf;
  if (im == inf) re = inf;
  if (re != re) im = re;
  if (im != im) re = im;
  const float l = fmaxf(re, im);
  const float sm = fminf(im, re);
  const float r = (l == 0.0f || sm == inf) ? 0.0f : sm / l;
  const float h = __builtin_sqrtf(__builtin_fmaf(r, r, 1.0f)) * l;
  return h * h;
}

constexpr int kT = 256;

// Block-wide sum of one double per thread in a fixed order (deterministic).
__device__ double block_sum_fixed(double v, double* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) lds[threadIdx.x] = lds[threadIdx.x] + lds[threadIdx.x + off];
    __syncthreads();
  }
  const double r = lds[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// probabilities
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(kT) void k_probs_all(const cplx<R>* __restrict__ psi, R* __restrict__ out,
                                                  uint64_t n) {
  const uint64_t i = block_id() * kT + threadIdx.x;
  if (i < n) out[i] = np_abs2(psi[i]);
}

struct PermArgs {
  uint64_t nout;
  int m;
  uint8_t ibit[kMaxFixed];  // input bit of output bit (m-1-k), k = 0..m-1
};

// All qubits measured but in another order: gather with a bit permutation.
template <typename R>
__global__ __launch_bounds__(kT) void k_probs_perm(const cplx<R>* __restrict__ psi, R* __restrict__ out,
                                                   const PermArgs a) {
  const uint64_t o = block_id() * kT + threadIdx.x;
  if (o >= a.nout) return;
  uint64_t i = 0;
  for (int k = 0; k < a.m; ++k) i |= ((o >> (a.m - 1 - k)) & 1) << a.ibit[k];
  out[o] = np_abs2(psi[i]);
}

struct MargArgs {
  uint64_t nout;      // 2^m
  uint64_t nin;       // 2^(n-m) inputs per output
  uint64_t nchunks;   // blocks per output (partial sums)
  int m;
  int nmpos;          // = m
  uint8_t ibit[kMaxFixed];   // input bit of output bit (m-1-k)
  uint8_t mpos[kMaxFixed];   // measured bit positions, ascending
};

__device__ __forceinline__ uint64_t scatter_out(uint64_t o, const MargArgs& a) {
  uint64_t i = 0;
  for (int k = 0; k < a.m; ++k) i |= ((o >> (a.m - 1 - k)) & 1) << a.ibit[k];
  return i;
}

// One thread per output: few inputs per output.
template <typename R>
__global__ __launch_bounds__(kT) void k_marg_thread(const cplx<R>* __restrict__ psi, R* __restrict__ out,
                                                    const MargArgs a) {
  const uint64_t o = block_id() * kT + threadIdx.x;
  if (o >= a.nout) return;
  const uint64_t ob = scatter_out(o, a);
  double acc = 0.0;
  for (uint64_t v = 0; v < a.nin; ++v) acc += double(np_abs2(psi[insert_zeros(v, a.mpos, a.nmpos) | ob]));
  out[o] = R(acc);
}

// One block per (output, chunk): many inputs per output.
template <typename R>
__global__ __launch_bounds__(kT) void k_marg_block(const cplx<R>* __restrict__ psi, R* __restrict__ out,
                                                   double* __restrict__ partial, const MargArgs a) {
  __shared__ double lds[kT];
  const uint64_t b = block_id();
  const uint64_t o = b / a.nchunks, c = b % a.nchunks;
  const uint64_t len = a.nin / a.nchunks;
  const uint64_t ob = scatter_out(o, a);
  double acc = 0.0;
  for (uint64_t v = c * len + threadIdx.x; v < (c + 1) * len; v += kT)
    acc += double(np_abs2(psi[insert_zeros(v, a.mpos, a.nmpos) | ob]));
  const double tot = block_sum_fixed(acc, lds);
  if (threadIdx.x == 0) {
    if (a.nchunks == 1)
      out[o] = R(tot);
    else
      partial[b] = tot;
  }
}

template <typename R>
__global__ __launch_bounds__(kT) void k_marg_final(const double* __restrict__ partial, R* __restrict__ out,
                                                   uint64_t nout, uint64_t nchunks) {
  const uint64_t o = block_id() * kT + threadIdx.x;
  if (o >= nout) return;
  double acc = 0.0;
  for (uint64_t c = 0; c < nchunks; ++c) acc += partial[o * nchunks + c];
  out[o] = R(acc);
}

// ---------------------------------------------------------------------------
// marginals in numpy's order: np.sum(probs, axis=unmeasured) on the (2,)*n
// array (calculate_probabilities, abstract.py:2739-2741)
// ---------------------------------------------------------------------------
// numpy reduces in C order after merging adjacent axes of the same kind
// (unmeasured / measured), in the probability dtype, from 0:
// * last qubit unmeasured: the innermost merged axis is a run of L = 2^t
//   unmeasured elements; each output adds, in C order of the outer unmeasured
//   qubits, the pairwise sums of its runs taken in buffers of at most 8192
//   elements (the pairwise rule of np.sum, see k_np_buffer_sums);
// * last qubit measured: each output adds its elements one by one in C order
//   of the unmeasured qubits.
// A serial chain of more than kMaxSerial additions per output would take
// seconds, so such marginals keep the parallel fixed-order FP64 sum.
constexpr uint64_t kMargBuf = 8192;
constexpr uint64_t kMaxSerial = uint64_t(1) << 24;

template <typename R>
__device__ R np_leaf_sum_psi(const cplx<R>* a, int n) {
  if (n < 8) {
    R res = R(0);
    for (int i = 0; i < n; ++i) res = res + np_abs2(a[i]);
    return res;
  }
  R r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = np_abs2(a[j]);
  int i = 8;
  for (; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = r[j] + np_abs2(a[i + j]);
  }
  R res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res = res + np_abs2(a[i]);
  return res;
}

// Pairwise sums of consecutive runs of lc probabilities (lc a power of two,
// 128 <= lc <= 8192): one 64-thread block per run, 128-element leaves
// combined as a balanced tree (the pairwise split of a power of two).
template <typename R>
__global__ __launch_bounds__(64) void k_run_sums(const cplx<R>* __restrict__ psi, uint64_t lc,
                                                 R* __restrict__ runsum) {
  __shared__ R leaf[kMargBuf / 128];
  const uint64_t run = block_id();
  const cplx<R>* a = psi + run * lc;
  const int nleaf = int(lc / 128);
  if (int(threadIdx.x) < nleaf) leaf[threadIdx.x] = np_leaf_sum_psi(a + uint64_t(threadIdx.x) * 128, 128);
  __syncthreads();
  for (int w = nleaf; w > 1; w >>= 1) {
    R v = R(0);
    if (int(threadIdx.x) < w / 2) v = leaf[2 * threadIdx.x] + leaf[2 * threadIdx.x + 1];
    __syncthreads();
    if (int(threadIdx.x) < w / 2) leaf[threadIdx.x] = v;
    __syncthreads();
  }
  if (threadIdx.x == 0) runsum[run] = leaf[0];
}

struct NpMargArgs {
  uint64_t nout;   // 2^m outputs, requested order
  uint64_t nred;   // outer unmeasured combinations (runs case) or all unmeasured combinations
  uint64_t nsub;   // run sums per outer index (runs case) / run length (short runs)
  int m;
  uint8_t obit[kMaxFixed];  // index bit of requested output bit k (k = 0 most significant)
  uint8_t mpos[kMaxFixed];  // the same bits, ascending
};

__device__ __forceinline__ uint64_t np_out_bits(uint64_t o, const NpMargArgs& a) {
  uint64_t ob = 0;
  for (int k = 0; k < a.m; ++k) ob |= ((o >> (a.m - 1 - k)) & 1) << a.obit[k];
  return ob;
}

// Runs of >= 128 elements: add the run sums of each output in C order.
template <typename R>
__global__ __launch_bounds__(kT) void k_marg_np_runs(const R* __restrict__ runsum, R* __restrict__ out,
                                                     const NpMargArgs a) {
  const uint64_t o = block_id() * kT + threadIdx.x;
  if (o >= a.nout) return;
  const uint64_t ob = np_out_bits(o, a);
  R acc = R(0);
  for (uint64_t r = 0; r < a.nred; ++r) {
    const R* sub = runsum + (insert_zeros(r, a.mpos, a.m) | ob) * a.nsub;
    for (uint64_t c = 0; c < a.nsub; ++c) acc = acc + sub[c];
  }
  out[o] = acc;
}

// Runs of fewer than 128 elements (nsub = run length, index = outer << t).
template <typename R>
__global__ __launch_bounds__(kT) void k_marg_np_short(const cplx<R>* __restrict__ psi, R* __restrict__ out,
                                                      const NpMargArgs a) {
  const uint64_t o = block_id() * kT + threadIdx.x;
  if (o >= a.nout) return;
  const uint64_t ob = np_out_bits(o, a);
  R acc = R(0);
  for (uint64_t r = 0; r < a.nred; ++r)
    acc = acc + np_leaf_sum_psi(psi + (insert_zeros(r, a.mpos, a.m) | ob) * a.nsub, int(a.nsub));
  out[o] = acc;
}

// Last qubit measured: every element in C order of the unmeasured qubits.
template <typename R>
__global__ __launch_bounds__(kT) void k_marg_np_seq(const cplx<R>* __restrict__ psi, R* __restrict__ out,
                                                    const NpMargArgs a) {
  const uint64_t o = block_id() * kT + threadIdx.x;
  if (o >= a.nout) return;
  const uint64_t ob = np_out_bits(o, a);
  R acc = R(0);
  for (uint64_t v = 0; v < a.nred; ++v) acc = acc + np_abs2(psi[insert_zeros(v, a.mpos, a.m) | ob]);
  out[o] = acc;
}

// Launches numpy's order for a marginal; returns 1 if the serial chain is too
// long (the caller then sums in parallel), 0 on launch, <0 on error.
template <typename R>
static int marg_numpy_order(const cplx<R>* psi, int n, const int* qubits, int m, R* out, void* ws, size_t ws_bytes,
                            hipStream_t s) {
  bool measured[64] = {false};
  for (int k = 0; k < m; ++k) measured[qubits[k]] = true;
  NpMargArgs a;
  std::memset(&a, 0, sizeof(a));
  a.m = m;
  a.nout = uint64_t(1) << m;
  if (measured[n - 1]) {
    a.nred = uint64_t(1) << (n - m);
    if (a.nred > kMaxSerial) return 1;
    for (int k = 0; k < m; ++k) a.obit[k] = uint8_t(n - 1 - qubits[k]);
    sorted_positions(n, qubits, m, nullptr, 0, a.mpos);
    hipLaunchKernelGGL(k_marg_np_seq<R>, launch_grid(ceil_div_u(a.nout, kT)), dim3(kT), 0, s, psi, out, a);
    QSIM_CHECK_LAUNCH();
    return 0;
  }
  int t = 0;
  while (t < n && !measured[n - 1 - t]) ++t;  // trailing unmeasured run (qubits n-t .. n-1)
  const uint64_t L = uint64_t(1) << t;
  const int outer_bits = n - t;
  a.nred = uint64_t(1) << (outer_bits - m);
  // positions in the outer index (full index >> t)
  for (int k = 0; k < m; ++k) a.obit[k] = uint8_t(n - 1 - qubits[k] - t);
  for (int k = 0; k < m; ++k) a.mpos[k] = a.obit[k];
  std::sort(a.mpos, a.mpos + m);
  if (L < 128) {
    if (a.nred * L > kMaxSerial) return 1;
    a.nsub = L;
    hipLaunchKernelGGL(k_marg_np_short<R>, launch_grid(ceil_div_u(a.nout, kT)), dim3(kT), 0, s, psi, out, a);
    QSIM_CHECK_LAUNCH();
    return 0;
  }
  const uint64_t lc = L < kMargBuf ? L : kMargBuf;
  a.nsub = L / lc;
  if (a.nred * a.nsub > kMaxSerial) return 1;
  const uint64_t nruns = (uint64_t(1) << n) / lc;
  QSIM_CHECK_ARG(ws != nullptr && ws_bytes >= nruns * sizeof(R), "qsim_probabilities: workspace too small");
  R* runsum = reinterpret_cast<R*>(ws);
  hipLaunchKernelGGL(k_run_sums<R>, launch_grid(nruns), dim3(64), 0, s, psi, lc, runsum);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_marg_np_runs<R>, launch_grid(ceil_div_u(a.nout, kT)), dim3(kT), 0, s, runsum, out, a);
  QSIM_CHECK_LAUNCH();
  return 0;
}

static void marg_plan(int n, int m, uint64_t* nchunks, bool* per_thread) {
  const uint64_t nout = uint64_t(1) << m, nin = uint64_t(1) << (n - m);
  if (nin <= 64) {
    *per_thread = true;
    *nchunks = 1;
    return;
  }
  *per_thread = false;
  uint64_t c = 1;
  while (c * 2 * nout <= 4096 && (nin / (c * 2)) >= 256) c *= 2;
  *nchunks = c;
}

template <typename R>
static int probabilities_impl(const void* state, int n, const int* qubits, int m, void* probs, void* ws,
                              size_t ws_bytes, hipStream_t s) {
  const cplx<R>* psi = reinterpret_cast<const cplx<R>*>(state);
  R* out = reinterpret_cast<R*>(probs);
  bool identity = (m == n);
  for (int k = 0; identity && k < m; ++k) identity = (qubits[k] == k);
  if (identity) {
    const uint64_t N = uint64_t(1) << n;
    hipLaunchKernelGGL(k_probs_all<R>, launch_grid(ceil_div_u(N, kT)), dim3(kT), 0, s, psi, out, N);
    QSIM_CHECK_LAUNCH();
    return QSIM_OK;
  }
  if (m == n) {
    PermArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nout = uint64_t(1) << m;
    a.m = m;
    for (int k = 0; k < m; ++k) a.ibit[k] = uint8_t(n - 1 - qubits[k]);
    hipLaunchKernelGGL(k_probs_perm<R>, launch_grid(ceil_div_u(a.nout, kT)), dim3(kT), 0, s, psi, out, a);
    QSIM_CHECK_LAUNCH();
    return QSIM_OK;
  }
  const int rc = marg_numpy_order<R>(psi, n, qubits, m, out, ws, ws_bytes, s);
  if (rc <= 0) return rc;
  MargArgs a;
  std::memset(&a, 0, sizeof(a));
  a.m = m;
  a.nmpos = m;
  a.nout = uint64_t(1) << m;
  a.nin = uint64_t(1) << (n - m);
  for (int k = 0; k < m; ++k) a.ibit[k] = uint8_t(n - 1 - qubits[k]);
  sorted_positions(n, qubits, m, nullptr, 0, a.mpos);
  bool per_thread;
  marg_plan(n, m, &a.nchunks, &per_thread);
  if (per_thread) {
    hipLaunchKernelGGL(k_marg_thread<R>, launch_grid(ceil_div_u(a.nout, kT)), dim3(kT), 0, s, psi, out, a);
    QSIM_CHECK_LAUNCH();
    return QSIM_OK;
  }
  const uint64_t need = (a.nchunks > 1) ? a.nout * a.nchunks * sizeof(double) : 0;
  QSIM_CHECK_ARG(need <= ws_bytes && (need == 0 || ws != nullptr), "qsim_probabilities: workspace too small");
  double* partial = reinterpret_cast<double*>(ws);
  hipLaunchKernelGGL(k_marg_block<R>, launch_grid(a.nout * a.nchunks), dim3(kT), 0, s, psi, out, partial, a);
  QSIM_CHECK_LAUNCH();
  if (a.nchunks > 1) {
    hipLaunchKernelGGL(k_marg_final<R>, launch_grid(ceil_div_u(a.nout, kT)), dim3(kT), 0, s, partial, out, a.nout,
                       a.nchunks);
    QSIM_CHECK_LAUNCH();
  }
  return QSIM_OK;
}

// ---------------------------------------------------------------------------
// exact cumulative sum
// ---------------------------------------------------------------------------
constexpr int kPer = 16;                 // elements per thread
constexpr int kTile = kT * kPer;         // 4096 elements per tile
constexpr int64_t kInvalid = int64_t(1) << 60;
constexpr int64_t kTwo53 = int64_t(1) << 53;

struct Map {
  int64_t c0, c1;  // m -> m + c[m & 1]
};

__device__ __forceinline__ int64_t sat_add(int64_t a, int64_t b) {
  const int64_t r = a + b;
  return r > kInvalid ? kInvalid : r;
}

// g after f
__device__ __forceinline__ Map compose(Map f, Map g) {
  Map h;
  h.c0 = sat_add(f.c0, (f.c0 & 1) ? g.c1 : g.c0);
  h.c1 = sat_add(f.c1, ((1 + f.c1) & 1) ? g.c1 : g.c0);
  return h;
}

// Biased exponent of the binade whose grid s lies on (subnormals share the
// grid of the first normal binade).
__device__ __forceinline__ int binade_of(double s) {
  const int E = int((__double_as_longlong(s) >> 52) & 0x7ff);
  return E < 1 ? 1 : E;
}

__device__ __forceinline__ bool in_binade(double s, int Eb) {
  const int E = int((__double_as_longlong(s) >> 52) & 0x7ff);
  return E == Eb || (Eb == 1 && E == 0);
}

__device__ __forceinline__ int64_t m_of(double s, int Eb) {
  return int64_t(__double_as_longlong(s)) - (int64_t(Eb - 1) << 52);
}

__device__ __forceinline__ double from_m(int64_t m, int Eb) {
  return __longlong_as_double((long long)((int64_t(Eb - 1) << 52) + m));
}

// The map of adding p (>= 0, finite) on the grid of binade Eb.
__device__ __forceinline__ Map elem_map(double p, int Eb) {
  const uint64_t bits = uint64_t(__double_as_longlong(p));
  const int Ep = int((bits >> 52) & 0x7ff);
  uint64_t mp = bits & ((uint64_t(1) << 52) - 1);
  if (Ep != 0) mp |= uint64_t(1) << 52;
  if (mp == 0) return {0, 0};
  const int sh = Eb - (Ep < 1 ? 1 : Ep);  // p / u = mp * 2^-sh
  if (sh < 0) return {kInvalid, kInvalid};
  if (sh == 0) return {int64_t(mp), int64_t(mp)};
  if (sh >= 63) return {0, 0};
  const uint64_t k = mp >> sh;
  const uint64_t rem = mp & ((uint64_t(1) << sh) - 1);
  const uint64_t half = uint64_t(1) << (sh - 1);
  if (rem < half) return {int64_t(k), int64_t(k)};
  if (rem > half) return {int64_t(k + 1), int64_t(k + 1)};
  // tie: pick the neighbour that makes m + c even
  return {int64_t(k + (k & 1)), int64_t(k + ((k + 1) & 1))};
}

// The same element map and composition with the integers held in doubles
// (exact below 2^53), for the bulk of k_tile_map: an FP64 pipe operation
// each instead of 64-bit integer sequences. x = |p| / u = |p| 2^(1075 - Eb)
// is exact whenever x >= 2^-1022, and a smaller x rounds to 0 either way;
// c0 = x rounded to nearest, ties to even (m even: m + c even), c1 the same
// with ties to odd (m odd). A p above the binade gives x >= 2^53, which the
// caller's final check (< 2^53) rejects like the integer form's kInvalid;
// sums of values >= 2^53 stay >= 2^53, and NaN fails the check.
struct DMap {
  double c0, c1;
};

__device__ __forceinline__ DMap elem_map_fp(double p, int Eb) {
  const double x = ldexp(fabs(p), 1075 - Eb);
  const double c0 = __builtin_rint(x);
  const double fl = __builtin_floor(x);
  const bool tie = (x - fl) == 0.5;
  return {c0, tie ? (fl + fl + 1.0) - c0 : c0};
}

__device__ __forceinline__ bool odd_fp(double c) {
  const double h = c * 0.5;
  return h != __builtin_floor(h);
}

// g after f
__device__ __forceinline__ DMap compose_fp(DMap f, DMap g) {
  return {f.c0 + (odd_fp(f.c0) ? g.c1 : g.c0), f.c1 + (odd_fp(1.0 + f.c1) ? g.c1 : g.c0)};
}

__device__ __forceinline__ Map to_int_map(DMap f) {
  return {f.c0 < 0x1p53 ? int64_t(f.c0) : kInvalid, f.c1 < 0x1p53 ? int64_t(f.c1) : kInvalid};
}

template <typename P>
__device__ __forceinline__ double load_p(const P* p, int64_t i) {
  return double(p[i]);
}

__device__ __forceinline__ Map shfl_up_map(Map v, int delta) {
  Map r;
  r.c0 = __shfl_up(v.c0, delta, 64);
  r.c1 = __shfl_up(v.c1, delta, 64);
  return r;
}

// Lane `lane`'s value for the whole wavefront; `lane` must be the same on
// every lane (a loop counter or a constant): one v_readlane per 32 bits,
// no LDS round trip as in a general shuffle.
__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ int64_t rl(int64_t v, int lane) {
  const uint64_t u = uint64_t(v);
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), lane));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), lane));
  return int64_t((uint64_t(hi) << 32) | lo);
}

__device__ __forceinline__ double rl(double v, int lane) {
  return __longlong_as_double((long long)rl(int64_t(__double_as_longlong(v)), lane));
}

__device__ __forceinline__ Map shfl_map(Map v, int lane) {
  Map r;
  r.c0 = rl(v.c0, lane);
  r.c1 = rl(v.c1, lane);
  return r;
}

// Inclusive scan of maps over the 64 lanes of a wavefront (lane order).
__device__ __forceinline__ Map wave_scan_incl(Map v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const Map o = shfl_up_map(v, off);
    if (lane >= off) v = compose(o, v);
  }
  return v;
}

struct TileInfo {
  int64_t c0, c1;
  int32_t Eb;
  int32_t flag;  // k_tile_map: 1 = maps usable; k_carry: 0 = apply by maps, 1 = done sequentially
};

// (A) approximate tile sums
template <typename P>
__global__ __launch_bounds__(kT) void k_tile_sum(const P* __restrict__ p, int64_t n, double* __restrict__ tsum) {
  __shared__ double lds[kT];
  const int64_t base = int64_t(blockIdx.x) * kTile;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = base + k * kT + threadIdx.x;
    if (i < n) acc += load_p(p, i);
  }
  const double tot = block_sum_fixed(acc, lds);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// (B) approximate exclusive prefix of the tile sums (one block)
__global__ __launch_bounds__(1024) void k_approx_prefix(const double* __restrict__ tsum, int64_t ntiles,
                                                        double* __restrict__ astart) {
  __shared__ double lds[1024];
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t lo = int64_t(threadIdx.x) * per;
  const int64_t hi = lo + per < ntiles ? lo + per : ntiles;
  double acc = 0.0;
  for (int64_t t = lo; t < hi; ++t) acc += tsum[t];
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const double v = threadIdx.x >= off ? lds[threadIdx.x - off] : 0.0;
    __syncthreads();
    lds[threadIdx.x] += v;
    __syncthreads();
  }
  double run = threadIdx.x > 0 ? lds[threadIdx.x - 1] : 0.0;
  for (int64_t t = lo; t < hi; ++t) {
    astart[t] = run;
    run += tsum[t];
  }
}

// Stage a tile into LDS with coalesced loads; returns the thread's 16
// consecutive elements' offset in the tile.
template <typename P>
__device__ __forceinline__ void stage_tile(const P* __restrict__ p, int64_t n, int64_t base, double* tile) {
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = base + k * kT + threadIdx.x;
    tile[k * kT + threadIdx.x] = i < n ? load_p(p, i) : 0.0;
  }
  __syncthreads();
}

// Exclusive scan of one map per thread across the block (thread order).
__device__ __forceinline__ Map block_scan_excl(Map v, Map* wave_tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const Map incl = wave_scan_incl(v);
  Map excl = shfl_up_map(incl, 1);
  if (lane == 0) excl = {0, 0};
  if (lane == 63) wave_tot[wave] = incl;
  __syncthreads();
  Map pre = {0, 0};
  for (int w = 0; w < wave; ++w) pre = compose(pre, wave_tot[w]);
  __syncthreads();
  return compose(pre, excl);
}

// The 16 consecutive elements [i0, i0 + 16) of one thread, widened to
// double (0 past n), read with 16-byte loads when the run is whole and the
// array 16-byte aligned.
__device__ __forceinline__ void load_run(const double* __restrict__ p, int64_t i0, int64_t n, double* x) {
  if (i0 + kPer <= n && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    const double2* q = reinterpret_cast<const double2*>(p + i0);
#pragma unroll
    for (int k = 0; k < kPer / 2; ++k) {
      const double2 v = q[k];
      x[2 * k] = v.x;
      x[2 * k + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kPer; ++k) x[k] = i0 + k < n ? p[i0 + k] : 0.0;
  }
}

__device__ __forceinline__ void load_run(const float* __restrict__ p, int64_t i0, int64_t n, double* x) {
  if (i0 + kPer <= n && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    const float4* q = reinterpret_cast<const float4*>(p + i0);
#pragma unroll
    for (int k = 0; k < kPer / 4; ++k) {
      const float4 v = q[k];
      x[4 * k] = double(v.x);
      x[4 * k + 1] = double(v.y);
      x[4 * k + 2] = double(v.z);
      x[4 * k + 3] = double(v.w);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kPer; ++k) x[k] = i0 + k < n ? double(p[i0 + k]) : 0.0;
  }
}

// (C) per-tile decision and map. Each thread reads its own 16 consecutive
// elements with wide loads (every byte of the tile's cache lines is used,
// without an LDS staging pass), so the kernel runs at 8 waves per SIMD.
template <typename P>
__global__ __launch_bounds__(kT) void k_tile_map(const P* __restrict__ p, int64_t n, const double* __restrict__ astart,
                                                 const double* __restrict__ tsum, TileInfo* __restrict__ info,
                                                 Map* __restrict__ xmap) {
  __shared__ Map wave_tot[kT / 64];
  const int64_t t = int64_t(xcd_block());  // each XCD walks one contiguous range of tiles
  if (t >= (n + kTile - 1) / kTile) return;
  const double A = astart[t];
  const double Z = A + tsum[t];
  const int Eb = binade_of(A);
  const double margin = 1.0 + 0x1p-20;
  const double lo = ldexp(1.0, Eb - 1023), hi = ldexp(1.0, Eb - 1022);
  const bool safe = (Eb == 1 || A >= lo * margin) && (Z * margin < hi);
  if (!safe) {
    if (threadIdx.x == 0) info[t] = {0, 0, Eb, 0};
    return;
  }
  double x[kPer];
  load_run(p, t * kTile + int64_t(threadIdx.x) * kPer, n, x);
  DMap fd = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kPer; ++k) fd = compose_fp(fd, elem_map_fp(x[k], Eb));
  const Map f = to_int_map(fd);
  const Map excl = block_scan_excl(f, wave_tot);
  xmap[t * kT + threadIdx.x] = excl;  // reused by k_tile_apply (same tile, same grid Eb)
  if (threadIdx.x == kT - 1) {
    const Map tot = compose(excl, f);
    const bool ok = tot.c0 < kTwo53 && tot.c1 < kTwo53;
    info[t] = {tot.c0, tot.c1, Eb, ok ? 1 : 0};
  }
}

// (D1) maps of groups of 64 tiles: usable when all 64 tiles are usable on
// the same grid (flag 1), else flag 0.
__global__ __launch_bounds__(64) void k_group_map(const TileInfo* __restrict__ info, int64_t ntiles,
                                                  TileInfo* __restrict__ ginfo) {
  const int lane = threadIdx.x;
  const int64_t t = int64_t(blockIdx.x) * 64 + lane;
  TileInfo ti = t < ntiles ? info[t] : TileInfo{0, 0, 0, 1};
  const int Eb0 = rl(ti.Eb, 0);
  if (t >= ntiles) ti.Eb = Eb0;
  const bool uniform = __all(ti.flag == 1 && ti.Eb == Eb0);
  const Map incl = wave_scan_incl(Map{ti.c0, ti.c1});
  if (lane == 63) {
    const bool ok = uniform && incl.c0 < kTwo53 && incl.c1 < kTwo53;
    ginfo[blockIdx.x] = {incl.c0, incl.c1, Eb0, ok ? 1 : 0};
  }
}

// Exact in-order sum of elements [lo, hi) (at most one tile) from the exact
// start s, by the whole 256-thread block, in rounds. The tile is staged in
// LDS; thread j owns the 16 consecutive elements from 16 j. A round takes the
// binade of the current exact sum, composes every remaining element's map on
// its grid (the floating-point form), scans the thread maps across the block
// and lets each thread walk its elements from its exact start: every partial
// sum that stays in the binade is final, and the first element whose sum
// leaves it (m + c >= 2^53: the sum reaches the next power of two) is found
// with an LDS minimum. That element is added with one IEEE addition and the
// next round starts after it on its new binade. Without a crossing the tile
// ends in one round; a round per binade crossed otherwise. The results are
// stored coalesced to cdf[lo, hi); returns the end value: bitwise the
// sequential chain.
struct ChainLds {
  double lt[kTile];
  Map wave_tot[kT / 64];
  double last;  // exact sum before the crossing element
  int cross;
};

template <typename P>
__device__ double chain_tile_rounds(const P* __restrict__ p, int64_t lo, int64_t hi, double s,
                                    double* __restrict__ cdf, ChainLds& sh) {
  const int tid = threadIdx.x;
  const int len = int(hi - lo);  // 1 .. 4096
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = k * kT + tid;
    sh.lt[e] = e < len ? load_p(p, lo + e) : 0.0;
  }
  if (tid == 0) sh.cross = len;
  __syncthreads();
  const int e0 = tid * kPer;
  double vals[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) vals[k] = 0.0;
  int pos = 0;  // first element not summed yet (the same on every thread)
  while (pos < len) {
    const int Eb = binade_of(s);
    const int64_t m0 = m_of(s, Eb);
    DMap fd = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int e = e0 + k;
      if (e >= pos && e < len) fd = compose_fp(fd, elem_map_fp(sh.lt[e], Eb));
    }
    const Map excl = block_scan_excl(to_int_map(fd), sh.wave_tot);
    // the thread's exact start; >= 2^53 when the sum left the binade in an
    // earlier thread (that thread reports the crossing)
    int64_t m = m0 + ((m0 & 1) ? excl.c1 : excl.c0);
    int cross = len;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int e = e0 + k;
      if (e >= pos && e < len && cross == len && m < kTwo53) {
        const DMap c = elem_map_fp(sh.lt[e], Eb);
        const double cc = (m & 1) ? c.c1 : c.c0;
        if (cc < 0x1p53 && m + int64_t(cc) < kTwo53) {
          m += int64_t(cc);
          vals[k] = from_m(m, Eb);
        } else {
          cross = e;
        }
      }
    }
    if (cross < len) atomicMin(&sh.cross, cross);
    __syncthreads();
    const int cr = sh.cross;
    if (cr == len) {
      // no crossing: the last thread with elements holds the end value
      if (tid == (len - 1) / kPer) {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          if (k == (len - 1) % kPer) sh.last = vals[k];
      }
      __syncthreads();
      if (pos < len) s = sh.last;
      pos = len;
    } else {
      // elements [pos, cr) are final; element cr is added in IEEE arithmetic
      if (cr > pos) {
        if (tid == (cr - 1) / kPer) {
#pragma unroll
          for (int k = 0; k < kPer; ++k)
            if (k == (cr - 1) % kPer) sh.last = vals[k];
        }
        __syncthreads();
        s = sh.last;
      }
      s = s + sh.lt[cr];
      if (tid == cr / kPer) {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
          if (k == cr % kPer) vals[k] = s;
      }
      pos = cr + 1;
    }
    __syncthreads();
    if (tid == 0) sh.cross = len;
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    if (e0 + k < len) sh.lt[e0 + k] = vals[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = k * kT + tid;
    if (e < len) cdf[lo + e] = sh.lt[e];
  }
  __syncthreads();
  return s;
}

// Tile-level chain of one group (64 tiles) from the exact start s; every
// lane of every wave computes the same scalar chain from the same data (the
// waves only split the work of chain_tile_rounds), and wave 0 writes. Writes
// tile starts, tile flags (0 = finish by maps, 1 = summed here) and the cdf
// of tiles summed here. Returns s.
template <typename P>
__device__ double chain_group_tiles(const P* __restrict__ p, int64_t n, int64_t ntiles, int64_t g,
                                    TileInfo* __restrict__ info, double* __restrict__ tstart,
                                    double* __restrict__ cdf, double s, ChainLds& sh) {
  const int lane = threadIdx.x & 63;
  const bool w0 = threadIdx.x < 64;
  const int64_t t = g * 64 + lane;
  TileInfo ti = t < ntiles ? info[t] : TileInfo{0, 0, 0, 1};
  // every wave holds the group's tile records before wave 0 rewrites a flag
  __syncthreads();
  const int Eb0 = rl(ti.Eb, 0);
  if (t >= ntiles) ti.Eb = Eb0;
  if (__all(ti.flag == 1 && ti.Eb == Eb0) && in_binade(s, Eb0)) {
    const Map incl = wave_scan_incl(Map{ti.c0, ti.c1});
    const Map tot = shfl_map(incl, 63);
    const int64_t m0 = m_of(s, Eb0);
    const int64_t mend = m0 + ((m0 & 1) ? tot.c1 : tot.c0);
    if (mend < kTwo53) {
      Map excl = shfl_up_map(incl, 1);
      if (lane == 0) excl = {0, 0};
      if (w0 && t < ntiles) {
        tstart[t] = from_m(m0 + ((m0 & 1) ? excl.c1 : excl.c0), Eb0);
        info[t].flag = 0;
      }
      return from_m(mend, Eb0);
    }
  }
  const int64_t gend = g * 64 + 64 < ntiles ? g * 64 + 64 : ntiles;
  for (int64_t u = g * 64; u < gend; ++u) {
    // tile u's info is held by lane u - 64 g (read once above)
    const int src = int(u - g * 64);
    TileInfo tu;
    tu.c0 = rl(ti.c0, src);
    tu.c1 = rl(ti.c1, src);
    tu.Eb = rl(ti.Eb, src);
    tu.flag = rl(ti.flag, src);
    if (threadIdx.x == 0) tstart[u] = s;
    bool ok = tu.flag == 1 && in_binade(s, tu.Eb);
    int64_t m2 = 0;
    if (ok) {
      const int64_t m = m_of(s, tu.Eb);
      m2 = m + ((m & 1) ? tu.c1 : tu.c0);
      ok = m2 < kTwo53;
    }
    if (ok) {
      s = from_m(m2, tu.Eb);
      if (threadIdx.x == 0) info[u].flag = 0;
    } else {
      const int64_t lo = u * kTile;
      const int64_t hi = lo + kTile < n ? lo + kTile : n;
      s = chain_tile_rounds(p, lo, hi, s, cdf, sh);
      if (threadIdx.x == 0) info[u].flag = 1;
    }
  }
  return s;
}

// (D2) exact chain over groups: 64 groups (2^24 elements) per step when they
// are all resolvable by their maps; group flags become 0 (start known, tiles
// finished by k_tile_starts) or 1 (chained tile by tile). One block of four
// waves: each wave runs the same scalar chain (wave 0 writes), and all four
// sum the tiles that cross a power of two (chain_tile_rounds).
template <typename P>
__global__ __launch_bounds__(kT) void k_carry(const P* __restrict__ p, int64_t n, int64_t ntiles, int64_t ngroups,
                                              TileInfo* __restrict__ info, TileInfo* __restrict__ ginfo,
                                              double* __restrict__ gstart, double* __restrict__ tstart,
                                              double* __restrict__ cdf, const int* __restrict__ valid) {
  __shared__ ChainLds sh;
  const int lane = threadIdx.x & 63;
  const bool w0 = threadIdx.x < 64;
  // probabilities with a NaN, a negative entry or an infinite sum are not
  // chained (np.random.choice rejects them; the chain would cross binades on
  // every element)
  if (valid != nullptr && *valid == 0) return;
  double s = 0.0;
  for (int64_t k = 0; k < ngroups; k += 64) {
    const int64_t gg = k + lane;
    TileInfo gi = gg < ngroups ? ginfo[gg] : TileInfo{0, 0, 0, 1};
    // every wave holds the step's group records before wave 0 rewrites a flag
    __syncthreads();
    const int Eb0 = rl(gi.Eb, 0);
    if (gg >= ngroups) gi.Eb = Eb0;
    if (__all(gi.flag == 1 && gi.Eb == Eb0) && in_binade(s, Eb0)) {
      const Map incl = wave_scan_incl(Map{gi.c0, gi.c1});
      const Map tot = shfl_map(incl, 63);
      const int64_t m0 = m_of(s, Eb0);
      const int64_t mend = m0 + ((m0 & 1) ? tot.c1 : tot.c0);
      if (mend < kTwo53) {
        Map excl = shfl_up_map(incl, 1);
        if (lane == 0) excl = {0, 0};
        if (w0 && gg < ngroups) {
          gstart[gg] = from_m(m0 + ((m0 & 1) ? excl.c1 : excl.c0), Eb0);
          ginfo[gg].flag = 0;
        }
        s = from_m(mend, Eb0);
        continue;
      }
    }
    const int64_t kend = k + 64 < ngroups ? k + 64 : ngroups;
    for (int64_t g = k; g < kend; ++g) {
      // group g's info is held by lane g - k (read once above)
      const int src = int(g - k);
      TileInfo gu;
      gu.c0 = rl(gi.c0, src);
      gu.c1 = rl(gi.c1, src);
      gu.Eb = rl(gi.Eb, src);
      gu.flag = rl(gi.flag, src);
      bool ok = gu.flag == 1 && in_binade(s, gu.Eb);
      int64_t m2 = 0;
      if (ok) {
        const int64_t m = m_of(s, gu.Eb);
        m2 = m + ((m & 1) ? gu.c1 : gu.c0);
        ok = m2 < kTwo53;
      }
      if (ok) {
        if (threadIdx.x == 0) {
          gstart[g] = s;
          ginfo[g].flag = 0;
        }
        s = from_m(m2, gu.Eb);
      } else {
        if (threadIdx.x == 0) ginfo[g].flag = 1;
        s = chain_group_tiles(p, n, ntiles, g, info, tstart, cdf, s, sh);
      }
    }
  }
}

// (D3) tile starts inside the groups the chain resolved by their maps
__global__ __launch_bounds__(64) void k_tile_starts(TileInfo* __restrict__ info, int64_t ntiles,
                                                    const TileInfo* __restrict__ ginfo, const double* __restrict__ gstart,
                                                    double* __restrict__ tstart) {
  const int64_t g = blockIdx.x;
  if (ginfo[g].flag != 0) return;
  const int lane = threadIdx.x;
  const int64_t t = g * 64 + lane;
  const TileInfo ti = t < ntiles ? info[t] : TileInfo{0, 0, 0, 1};
  const int Eb = ginfo[g].Eb;
  const Map incl = wave_scan_incl(Map{ti.c0, ti.c1});
  Map excl = shfl_up_map(incl, 1);
  if (lane == 0) excl = {0, 0};
  if (t < ntiles) {
    const int64_t m0 = m_of(gstart[g], Eb);
    tstart[t] = from_m(m0 + ((m0 & 1) ? excl.c1 : excl.c0), Eb);
    info[t].flag = 0;
  }
}

// (E) finish every tile that the chain resolved by maps
// A tile with flag 0 was resolved by its map, so k_tile_map computed its
// threads' exclusive maps on the grid ti.Eb: they are read back, not redone.
template <typename P>
__global__ __launch_bounds__(kT) void k_tile_apply(const P* __restrict__ p, int64_t n, const TileInfo* __restrict__ info,
                                                   const double* __restrict__ tstart, const Map* __restrict__ xmap,
                                                   double* __restrict__ cdf) {
  __shared__ double tile[kTile];
  const int64_t t = blockIdx.x;
  const TileInfo ti = info[t];
  if (ti.flag != 0) return;
  const int64_t base = t * kTile;
  const Map excl = xmap[t * kT + threadIdx.x];
  stage_tile(p, n, base, tile);
  const double s0 = tstart[t];
  const int64_t m0 = m_of(s0, ti.Eb);
  double s = from_m(m0 + ((m0 & 1) ? excl.c1 : excl.c0), ti.Eb);
  double out[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    s = s + tile[threadIdx.x * kPer + k];
    out[k] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) tile[threadIdx.x * kPer + k] = out[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = base + k * kT + threadIdx.x;
    if (i < n) cdf[i] = tile[k * kT + threadIdx.x];
  }
}

struct CumsumWs {
  double* tsum;
  double* astart;
  double* tstart;
  TileInfo* info;
  TileInfo* ginfo;
  double* gstart;
  Map* xmap;  // per-thread exclusive maps of every tile (16 bytes per 16 elements)
};

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static size_t cumsum_ws_bytes(int64_t n) {
  const size_t T = size_t((n + kTile - 1) / kTile);
  const size_t G = (T + 63) / 64;
  return 3 * align256(T * sizeof(double)) + align256(T * sizeof(TileInfo)) + align256(G * sizeof(TileInfo)) +
         align256(G * sizeof(double)) + align256(T * kT * sizeof(Map));
}

template <typename P>
static int cumsum_impl(const void* probs, int64_t n, double* cdf, void* ws, hipStream_t s) {
  const P* p = reinterpret_cast<const P*>(probs);
  const int64_t T = (n + kTile - 1) / kTile;
  const int64_t G = (T + 63) / 64;
  char* w = reinterpret_cast<char*>(ws);
  CumsumWs c;
  c.tsum = reinterpret_cast<double*>(w);
  w += align256(T * sizeof(double));
  c.astart = reinterpret_cast<double*>(w);
  w += align256(T * sizeof(double));
  c.tstart = reinterpret_cast<double*>(w);
  w += align256(T * sizeof(double));
  c.info = reinterpret_cast<TileInfo*>(w);
  w += align256(T * sizeof(TileInfo));
  c.ginfo = reinterpret_cast<TileInfo*>(w);
  w += align256(G * sizeof(TileInfo));
  c.gstart = reinterpret_cast<double*>(w);
  w += align256(G * sizeof(double));
  c.xmap = reinterpret_cast<Map*>(w);
  hipLaunchKernelGGL(k_tile_sum<P>, dim3(unsigned(T)), dim3(kT), 0, s, p, n, c.tsum);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_approx_prefix, dim3(1), dim3(1024), 0, s, c.tsum, T, c.astart);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_tile_map<P>, launch_grid(uint64_t(T)), dim3(kT), 0, s, p, n, c.astart, c.tsum, c.info, c.xmap);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_group_map, dim3(unsigned(G)), dim3(64), 0, s, c.info, T, c.ginfo);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_carry<P>, dim3(1), dim3(kT), 0, s, p, n, T, G, c.info, c.ginfo, c.gstart, c.tstart, cdf,
                     static_cast<const int*>(nullptr));
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_tile_starts, dim3(unsigned(G)), dim3(64), 0, s, c.info, T, c.ginfo, c.gstart, c.tstart);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_tile_apply<P>, dim3(unsigned(T)), dim3(kT), 0, s, p, n, c.info, c.tstart, c.xmap, cdf);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

// ---------------------------------------------------------------------------
// sampling without the full cumulative sum
// ---------------------------------------------------------------------------
// np.random.choice(n, size, p=p) returns, for a uniform u, the number of i
// with fl(cdf[i] / cdf[n-1]) <= u (searchsorted side="right" on the
// normalised cumsum). A draw compares only a few cdf values, and the chain
// already fixes all of them without writing the cdf: k_carry yields every
// tile's exact start, a tile it resolved by its map has every 16-element
// chunk's exact start in its threads' exclusive maps (k_tile_map, on the
// tile's grid), and a tile it summed element by element has its cdf written.
// So each draw bisects the tile ends, then the chunk ends of its tile, then
// adds at most 16 elements from the chunk's exact start; the full cdf (and
// k_tile_apply's pass over every element) is never needed. The statistics
// np.random.choice checks (sum, negative entries, NaNs) come from the same
// pass as the tile sums; invalid probabilities are not chained.

// (S0) per tile: the sum of k_tile_sum and the counts of negative and NaN
// entries
template <typename P>
__global__ __launch_bounds__(kT) void k_tile_stats(const P* __restrict__ p, int64_t n, double* __restrict__ tsum,
                                                   int2* __restrict__ tbad) {
  __shared__ double lds[kT];
  const int64_t base = int64_t(blockIdx.x) * kTile;
  double acc = 0.0;
  int neg = 0, nan = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = base + k * kT + threadIdx.x;
    if (i < n) {
      const double x = load_p(p, i);
      acc += x;
      neg += x < 0.0 ? 1 : 0;
      nan += x != x ? 1 : 0;
    }
  }
  const double tot = block_sum_fixed(acc, lds);
  const double nneg = block_sum_fixed(double(neg), lds);
  const double nnan = block_sum_fixed(double(nan), lds);
  if (threadIdx.x == 0) {
    tsum[blockIdx.x] = tot;
    tbad[blockIdx.x] = make_int2(int(nneg), int(nnan));
  }
}

// Probabilities of all qubits in order (k_probs_all) and, in the same pass,
// the tile statistics of k_tile_stats over them (same order: bitwise the
// same tile sums).
template <typename R>
__global__ __launch_bounds__(kT) void k_probs_tiles(const cplx<R>* __restrict__ psi, R* __restrict__ out, int64_t n,
                                                    double* __restrict__ tsum, int2* __restrict__ tbad) {
  __shared__ double lds[kT];
  // each XCD walks one contiguous range of tiles (see xcd_block)
  const int64_t tile = int64_t(xcd_block());
  if (tile >= (n + kTile - 1) / kTile) return;  // a 2-D grid's last row may run past the tiles
  const int64_t base = tile * kTile;
  double acc = 0.0;
  int neg = 0, nan = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int64_t i = base + k * kT + threadIdx.x;
    if (i < n) {
      // streaming loads and stores: the state is read once here and p is
      // read once more by the maps, neither stays in the caches usefully
      const R v = np_abs2(ld_nt(psi, uint64_t(i)));
      __builtin_nontemporal_store(v, out + i);
      const double x = double(v);
      acc += x;
      neg += x < 0.0 ? 1 : 0;
      nan += x != x ? 1 : 0;
    }
  }
  const double tot = block_sum_fixed(acc, lds);
  const double nneg = block_sum_fixed(double(neg), lds);
  const double nnan = block_sum_fixed(double(nan), lds);
  if (threadIdx.x == 0) {
    tsum[tile] = tot;
    tbad[tile] = make_int2(int(nneg), int(nnan));
  }
}

static size_t tile_stats_bytes(int64_t n) {
  const size_t T = size_t((n + kTile - 1) / kTile);
  return align256(T * sizeof(double)) + align256(T * sizeof(int2));
}

// Approximate exclusive prefix of the tile sums in three steps: 16 tiles
// per thread, 4096 per block (S1), the block partials in one block (S2),
// the tile starts per block (S3). Fixed order throughout (deterministic).
constexpr int kScanPer = 16;
constexpr int64_t kScanBlock = int64_t(kT) * kScanPer;

// Inclusive scan of one double per thread across the block (Hillis-Steele).
__device__ double block_scan_incl(double v, double* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kT; off <<= 1) {
    const double o = threadIdx.x >= off ? lds[threadIdx.x - off] : 0.0;
    __syncthreads();
    lds[threadIdx.x] += o;
    __syncthreads();
  }
  const double r = lds[threadIdx.x];
  __syncthreads();
  return r;
}

// (S1)
__global__ __launch_bounds__(kT) void k_scan_partials(const double* __restrict__ tsum, const int2* __restrict__ tbad,
                                                      int64_t ntiles, double* __restrict__ psum,
                                                      double2* __restrict__ pbad) {
  __shared__ double lds[kT];
  const int64_t t0 = int64_t(blockIdx.x) * kScanBlock + int64_t(threadIdx.x) * kScanPer;
  double acc = 0.0, neg = 0.0, nan = 0.0;
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t t = t0 + k;
    if (t < ntiles) {
      acc += tsum[t];
      const int2 b = tbad[t];
      neg += double(b.x);
      nan += double(b.y);
    }
  }
  acc = block_sum_fixed(acc, lds);
  neg = block_sum_fixed(neg, lds);
  nan = block_sum_fixed(nan, lds);
  if (threadIdx.x == 0) {
    psum[blockIdx.x] = acc;
    pbad[blockIdx.x] = make_double2(neg, nan);
  }
}

// (S2) one block; nparts <= 2048 (n <= 2^35)
__global__ __launch_bounds__(kT) void k_scan_top(const double* __restrict__ psum, const double2* __restrict__ pbad,
                                                 int64_t nparts, double* __restrict__ pstart,
                                                 double* __restrict__ stats, int* __restrict__ valid) {
  __shared__ double lds[kT];
  const int64_t per = (nparts + kT - 1) / kT;
  const int64_t lo = int64_t(threadIdx.x) * per;
  double acc = 0.0, neg = 0.0, nan = 0.0;
  for (int64_t b = lo; b < lo + per && b < nparts; ++b) {
    acc += psum[b];
    neg += pbad[b].x;
    nan += pbad[b].y;
  }
  const double incl = block_scan_incl(acc, lds);
  lds[threadIdx.x] = incl;
  __syncthreads();
  const double total = lds[kT - 1];
  __syncthreads();
  double run = incl - acc;
  for (int64_t b = lo; b < lo + per && b < nparts; ++b) {
    pstart[b] = run;
    run += psum[b];
  }
  neg = block_sum_fixed(neg, lds);
  nan = block_sum_fixed(nan, lds);
  if (threadIdx.x == 0) {
    stats[0] = total;
    stats[1] = neg;
    stats[2] = nan;
    *valid = (neg == 0.0 && nan == 0.0 && __builtin_isfinite(total)) ? 1 : 0;
  }
}

// (S3)
__global__ __launch_bounds__(kT) void k_scan_starts(const double* __restrict__ tsum, int64_t ntiles,
                                                    const double* __restrict__ pstart, double* __restrict__ astart) {
  __shared__ double lds[kT];
  const int64_t t0 = int64_t(blockIdx.x) * kScanBlock + int64_t(threadIdx.x) * kScanPer;
  double acc = 0.0;
  for (int k = 0; k < kScanPer; ++k)
    if (t0 + k < ntiles) acc += tsum[t0 + k];
  const double incl = block_scan_incl(acc, lds);
  double run = pstart[blockIdx.x] + (incl - acc);
  for (int k = 0; k < kScanPer; ++k) {
    const int64_t t = t0 + k;
    if (t < ntiles) {
      astart[t] = run;
      run += tsum[t];
    }
  }
}

// Exact cdf value of the start of 16-element chunk j of a tile the chain
// resolved by its map (the value before the chunk's first element).
__device__ __forceinline__ double chunk_start(const Map* __restrict__ xmap, int64_t t, int64_t j, int64_t m0, int Eb) {
  const Map x = xmap[t * kT + j];
  return from_m(m0 + ((m0 & 1) ? x.c1 : x.c0), Eb);
}

// (F) the exact total cdf[n - 1], the normaliser of the draws
template <typename P>
__global__ __launch_bounds__(64) void k_final_value(const P* __restrict__ p, int64_t n, int64_t ntiles,
                                                    const TileInfo* __restrict__ info, const double* __restrict__ tstart,
                                                    const Map* __restrict__ xmap, const double* __restrict__ cdf,
                                                    const int* __restrict__ valid, double* __restrict__ fin) {
  if (threadIdx.x != 0) return;
  if (*valid == 0) {
    *fin = 0.0;
    return;
  }
  const int64_t t = ntiles - 1, base = t * kTile;
  const TileInfo ti = info[t];
  if (ti.flag != 0) {
    *fin = cdf[n - 1];
    return;
  }
  const int64_t j = (n - 1 - base) / kPer;
  double s = chunk_start(xmap, t, j, m_of(tstart[t], ti.Eb), ti.Eb);
  for (int64_t i = base + j * kPer; i < n; ++i) s = s + load_p(p, i);
  *fin = s;
}

// (Q) one draw per thread
template <typename P>
__global__ __launch_bounds__(kT) void k_sample_search(const P* __restrict__ p, int64_t n, int64_t ntiles,
                                                      const TileInfo* __restrict__ info,
                                                      const double* __restrict__ tstart, const Map* __restrict__ xmap,
                                                      const double* __restrict__ cdf, const int* __restrict__ valid,
                                                      const double* __restrict__ fin, const double* __restrict__ u,
                                                      int64_t nshots, int64_t* __restrict__ out) {
  const int64_t k = int64_t(block_id()) * kT + threadIdx.x;
  if (k >= nshots) return;
  if (*valid == 0) {
    out[k] = 0;
    return;
  }
  const double total = *fin;
  const double v = u[k];
  // first tile whose end (the next tile's start; the total for the last
  // tile, whose normalised end is 1 > v) exceeds v after normalisation
  int64_t lo = 0, hi = ntiles - 1;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (tstart[mid + 1] / total <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  const int64_t t = lo, base = t * kTile;
  const int64_t end = base + kTile < n ? base + kTile : n;
  const TileInfo ti = info[t];
  if (ti.flag != 0) {
    // summed element by element: its cdf is written; cdf[end - 1] > v
    int64_t a = base, b = end - 1;
    while (a < b) {
      const int64_t mid = a + ((b - a) >> 1);
      if (cdf[mid] / total <= v)
        a = mid + 1;
      else
        b = mid;
    }
    out[k] = a;
    return;
  }
  // first chunk whose end (the next chunk's start; the tile end for the
  // last chunk) exceeds v, then the first element of it
  const int64_t m0 = m_of(tstart[t], ti.Eb);
  int64_t a = 0, b = (end - base - 1) / kPer;
  while (a < b) {
    const int64_t mid = a + ((b - a) >> 1);
    if (chunk_start(xmap, t, mid + 1, m0, ti.Eb) / total <= v)
      a = mid + 1;
    else
      b = mid;
  }
  double s = chunk_start(xmap, t, a, m0, ti.Eb);
  int64_t i = base + a * kPer;
  const int64_t last = (i + kPer < end ? i + kPer : end) - 1;
  for (; i < last; ++i) {
    s = s + load_p(p, i);
    if (!(s / total <= v)) break;
  }
  out[k] = i;
}

struct SampleWs {
  CumsumWs c;
  int2* tbad;
  double* psum;
  double2* pbad;
  double* pstart;
  double* fin;
  int* valid;
  double* cdf;  // written only for the tiles the chain sums element by element
};

static size_t sample_ws_bytes(int64_t n) {
  const size_t T = size_t((n + kTile - 1) / kTile);
  const size_t NP = size_t((int64_t(T) + kScanBlock - 1) / kScanBlock);
  return cumsum_ws_bytes(n) + align256(T * sizeof(int2)) + align256(NP * sizeof(double)) +
         align256(NP * sizeof(double2)) + align256(NP * sizeof(double)) + 2 * align256(sizeof(double)) +
         align256(size_t(n) * sizeof(double));
}

template <typename P>
static int sample_impl(const void* probs, int64_t n, const void* tiles, const double* u, int64_t nshots, int64_t* out,
                       double* stats, void* ws, hipStream_t s) {
  const P* p = reinterpret_cast<const P*>(probs);
  const int64_t T = (n + kTile - 1) / kTile;
  const int64_t G = (T + 63) / 64;
  const int64_t NP = (T + kScanBlock - 1) / kScanBlock;
  char* w = reinterpret_cast<char*>(ws);
  SampleWs a;
  a.c.tsum = reinterpret_cast<double*>(w);
  w += align256(T * sizeof(double));
  a.c.astart = reinterpret_cast<double*>(w);
  w += align256(T * sizeof(double));
  a.c.tstart = reinterpret_cast<double*>(w);
  w += align256(T * sizeof(double));
  a.c.info = reinterpret_cast<TileInfo*>(w);
  w += align256(T * sizeof(TileInfo));
  a.c.ginfo = reinterpret_cast<TileInfo*>(w);
  w += align256(G * sizeof(TileInfo));
  a.c.gstart = reinterpret_cast<double*>(w);
  w += align256(G * sizeof(double));
  a.c.xmap = reinterpret_cast<Map*>(w);
  w += align256(T * kT * sizeof(Map));
  a.tbad = reinterpret_cast<int2*>(w);
  w += align256(T * sizeof(int2));
  a.psum = reinterpret_cast<double*>(w);
  w += align256(NP * sizeof(double));
  a.pbad = reinterpret_cast<double2*>(w);
  w += align256(NP * sizeof(double2));
  a.pstart = reinterpret_cast<double*>(w);
  w += align256(NP * sizeof(double));
  a.fin = reinterpret_cast<double*>(w);
  w += align256(sizeof(double));
  a.valid = reinterpret_cast<int*>(w);
  w += align256(sizeof(double));
  a.cdf = reinterpret_cast<double*>(w);
  if (tiles != nullptr) {
    // computed with the probabilities (qsim_probabilities_tiles)
    a.c.tsum = reinterpret_cast<double*>(const_cast<void*>(tiles));
    a.tbad = reinterpret_cast<int2*>(reinterpret_cast<char*>(a.c.tsum) + align256(T * sizeof(double)));
  } else {
    hipLaunchKernelGGL(k_tile_stats<P>, dim3(unsigned(T)), dim3(kT), 0, s, p, n, a.c.tsum, a.tbad);
    QSIM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_scan_partials, dim3(unsigned(NP)), dim3(kT), 0, s, a.c.tsum, a.tbad, T, a.psum, a.pbad);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kT), 0, s, a.psum, a.pbad, NP, a.pstart, stats, a.valid);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_starts, dim3(unsigned(NP)), dim3(kT), 0, s, a.c.tsum, T, a.pstart, a.c.astart);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_tile_map<P>, launch_grid(uint64_t(T)), dim3(kT), 0, s, p, n, a.c.astart, a.c.tsum, a.c.info,
                     a.c.xmap);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_group_map, dim3(unsigned(G)), dim3(64), 0, s, a.c.info, T, a.c.ginfo);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_carry<P>, dim3(1), dim3(kT), 0, s, p, n, T, G, a.c.info, a.c.ginfo, a.c.gstart, a.c.tstart,
                     a.cdf, static_cast<const int*>(a.valid));
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_tile_starts, dim3(unsigned(G)), dim3(64), 0, s, a.c.info, T, a.c.ginfo, a.c.gstart,
                     a.c.tstart);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_final_value<P>, dim3(1), dim3(64), 0, s, p, n, T, a.c.info, a.c.tstart, a.c.xmap, a.cdf,
                     a.valid, a.fin);
  QSIM_CHECK_LAUNCH();
  if (nshots > 0) {
    hipLaunchKernelGGL(k_sample_search<P>, launch_grid(ceil_div_u(uint64_t(nshots), kT)), dim3(kT), 0, s, p, n, T,
                       a.c.info, a.c.tstart, a.c.xmap, a.cdf, a.valid, a.fin, u, nshots, out);
    QSIM_CHECK_LAUNCH();
  }
  return QSIM_OK;
}

// ---------------------------------------------------------------------------
// search, bits, histogram
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_search(const double* __restrict__ cdf, int64_t n,
                                               const double* __restrict__ u, int64_t nshots,
                                               int64_t* __restrict__ out) {
  const int64_t k = int64_t(block_id()) * kT + threadIdx.x;
  if (k >= nshots) return;
  const double total = cdf[n - 1];
  const double v = u[k];
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (cdf[mid] / total <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  out[k] = lo;
}

__global__ __launch_bounds__(kT) void k_to_binary(const int64_t* __restrict__ smp, int64_t nshots, int nbits,
                                                  int64_t* __restrict__ out) {
  const int64_t i = int64_t(block_id()) * kT + threadIdx.x;
  if (i >= nshots * nbits) return;
  const int64_t s = i / nbits;
  const int b = int(i % nbits);
  out[i] = (smp[s] >> (nbits - 1 - b)) & 1;
}

__global__ __launch_bounds__(kT) void k_histogram(const int64_t* __restrict__ smp, int64_t nshots,
                                                  unsigned long long* __restrict__ counts, int64_t nbins) {
  const int64_t i = int64_t(block_id()) * kT + threadIdx.x;
  if (i >= nshots) return;
  const int64_t v = smp[i];
  if (v >= 0 && v < nbins) atomicAdd(&counts[v], 1ull);
}


// ---------------------------------------------------------------------------
// numpy's np.sum of a 1-D array, bit for bit
// ---------------------------------------------------------------------------
// np.sum reduces a contiguous array buffer by buffer (8192 elements per
// buffer, numpy's default ufunc buffer size), adding each buffer's pairwise
// sum to a running total that starts at 0. The pairwise sum of a buffer
// splits it in halves (rounded down to a multiple of 8) down to blocks of at
// most 128 elements; a block of fewer than 8 elements is summed in order from
// 0, a larger block with 8 interleaved accumulators combined as
// ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7)) plus a sequential tail.
// For power-of-two lengths every split is exact, so a buffer of 2^b >= 256
// elements is a balanced tree over 128-element leaves. Sample frequencies
// normalise by this sum (Backend.sample_frequencies, abstract.py:2762).
constexpr int kNpBuf = 8192;
constexpr int kNpLeaf = 128;

template <typename F>
__device__ F np_leaf_sum(const F* a, int n) {
  if (n < 8) {
    F res = F(0);
    for (int i = 0; i < n; ++i) res = res + a[i];
    return res;
  }
  F r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  int i = 8;
  for (; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = r[j] + a[i + j];
  }
  F res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res = res + a[i];
  return res;
}

// One block of 64 threads per 8192-element buffer (power-of-two length n).
template <typename F>
__global__ __launch_bounds__(64) void k_np_buffer_sums(const F* __restrict__ p, int64_t n, F* __restrict__ bsum) {
  __shared__ F leaf[kNpBuf / kNpLeaf];
  const int64_t lo = int64_t(blockIdx.x) * kNpBuf;
  const int64_t len = (n - lo) < kNpBuf ? (n - lo) : kNpBuf;
  if (len <= kNpLeaf) {
    if (threadIdx.x == 0) bsum[blockIdx.x] = np_leaf_sum(p + lo, int(len));
    return;
  }
  const int nleaf = int(len / kNpLeaf);
  if (threadIdx.x < nleaf) leaf[threadIdx.x] = np_leaf_sum(p + lo + int64_t(threadIdx.x) * kNpLeaf, kNpLeaf);
  __syncthreads();
  for (int w = nleaf; w > 1; w >>= 1) {
    F v = F(0);
    if (threadIdx.x < w / 2) v = leaf[2 * threadIdx.x] + leaf[2 * threadIdx.x + 1];
    __syncthreads();
    if (threadIdx.x < w / 2) leaf[threadIdx.x] = v;
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = leaf[0];
}

template <typename F>
__global__ void k_np_total(const F* __restrict__ bsum, int64_t nbuf, F* __restrict__ out) {
  F acc = F(0);
  for (int64_t b = 0; b < nbuf; ++b) acc = acc + bsum[b];
  out[0] = acc;
}

template <typename F>
__global__ __launch_bounds__(kT) void k_div_by(const F* __restrict__ p, int64_t n, const F* __restrict__ d,
                                               F* __restrict__ out) {
  const int64_t i = int64_t(block_id()) * kT + threadIdx.x;
  if (i < n) out[i] = p[i] / d[0];
}

// ---------------------------------------------------------------------------
// checks of a probability vector before drawing: sum, negatives, NaNs
// ---------------------------------------------------------------------------
constexpr int kStatBlocks = 2048;

template <typename P>
__global__ __launch_bounds__(kT) void k_prob_stats(const P* __restrict__ p, int64_t n, double* __restrict__ part) {
  __shared__ double lds[kT];
  double sum = 0.0, neg = 0.0, nan = 0.0;
  const int64_t stride = int64_t(gridDim.x) * kT;
  for (int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x; i < n; i += stride) {
    const double x = double(p[i]);
    sum += x;
    neg += x < 0.0 ? 1.0 : 0.0;
    nan += x != x ? 1.0 : 0.0;
  }
  sum = block_sum_fixed(sum, lds);
  neg = block_sum_fixed(neg, lds);
  nan = block_sum_fixed(nan, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = sum;
    part[3 * blockIdx.x + 1] = neg;
    part[3 * blockIdx.x + 2] = nan;
  }
}

__global__ __launch_bounds__(kT) void k_prob_stats_final(const double* __restrict__ part, int nblocks,
                                                         double* __restrict__ out) {
  __shared__ double lds[kT];
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += kT)
    for (int k = 0; k < 3; ++k) v[k] += part[3 * b + k];
  for (int k = 0; k < 3; ++k) {
    const double r = block_sum_fixed(v[k], lds);
    if (threadIdx.x == 0) out[k] = r;
  }
}

// ---------------------------------------------------------------------------
// readout bitflips on binary samples (Backend.apply_bitflips,
// abstract.py:2279-2295): with u uniform per sample bit,
// noisy = s + (1 - s) [u < p0] - s [u < p1], in integers
// ---------------------------------------------------------------------------
struct BitflipArgs {
  int64_t n;     // nshots * nbits
  int nbits;
  double p0[kMaxFixed], p1[kMaxFixed];
};

__global__ __launch_bounds__(kT) void k_bitflips(const int64_t* __restrict__ s, const double* __restrict__ u,
                                                 const BitflipArgs a, int64_t* __restrict__ out) {
  const int64_t i = int64_t(block_id()) * kT + threadIdx.x;
  if (i >= a.n) return;
  const int col = int(i % a.nbits);
  const int64_t x = s[i];
  const double r = u[i];
  const int64_t f0 = r < a.p0[col] ? 1 : 0, f1 = r < a.p1[col] ? 1 : 0;
  out[i] = (x + (1 - x) * f0) - x * f1;
}

}  // namespace qsim

using namespace qsim;

extern "C" {

size_t qsim_probabilities_workspace(int nqubits, int nmeasured) {
  if (nqubits < 1 || nqubits > QSIM_MAX_QUBITS || nmeasured < 1 || nmeasured > nqubits) return 0;
  if (nmeasured == nqubits) return 0;
  uint64_t nchunks;
  bool per_thread;
  marg_plan(nqubits, nmeasured, &nchunks, &per_thread);
  const size_t fixed_order = (per_thread || nchunks == 1) ? 0 : size_t((uint64_t(1) << nmeasured) * nchunks * sizeof(double));
  // run sums of numpy's order: at most 2^n / 128 values (double for complex128)
  const size_t runs = nqubits >= 7 ? size_t(uint64_t(1) << (nqubits - 7)) * sizeof(double) : 0;
  return fixed_order > runs ? fixed_order : runs;
}

int qsim_probabilities(const void* state, int nqubits, int dtype, const int* qubits, int nmeasured,
                       void* probs, void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && probs != nullptr, "qsim_probabilities: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_probabilities: bad dtype");
  QSIM_CHECK_ARG(nmeasured >= 1 && nmeasured <= nqubits, "qsim_probabilities: bad number of qubits");
  if (!validate_qubits(nqubits, qubits, nmeasured, nullptr, 0, "qsim_probabilities")) return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return probabilities_impl<double>(state, nqubits, qubits, nmeasured, probs, workspace, workspace_bytes, s);
  return probabilities_impl<float>(state, nqubits, qubits, nmeasured, probs, workspace, workspace_bytes, s);
}

size_t qsim_prob_stats_workspace(int64_t n) { return n < 1 ? 0 : size_t(3 * kStatBlocks) * sizeof(double); }

int qsim_prob_stats(const void* probs, int prob_dtype, int64_t n, double* out, void* workspace, size_t workspace_bytes,
                    void* stream) {
  clear_error();
  QSIM_CHECK_ARG(probs != nullptr && out != nullptr, "qsim_prob_stats: null pointer");
  QSIM_CHECK_ARG(n >= 1 && n <= (int64_t(1) << 40), "qsim_prob_stats: bad length");
  QSIM_CHECK_ARG(prob_dtype == QSIM_F64 || prob_dtype == QSIM_F32, "qsim_prob_stats: bad dtype");
  QSIM_CHECK_ARG(workspace != nullptr && workspace_bytes >= qsim_prob_stats_workspace(n),
                 "qsim_prob_stats: workspace too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double* part = reinterpret_cast<double*>(workspace);
  const int64_t want = (n + kT - 1) / kT;
  const int nb = int(want < kStatBlocks ? want : kStatBlocks);
  if (prob_dtype == QSIM_F64)
    hipLaunchKernelGGL(k_prob_stats<double>, dim3(nb), dim3(kT), 0, s, reinterpret_cast<const double*>(probs), n, part);
  else
    hipLaunchKernelGGL(k_prob_stats<float>, dim3(nb), dim3(kT), 0, s, reinterpret_cast<const float*>(probs), n, part);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_prob_stats_final, dim3(1), dim3(kT), 0, s, part, nb, out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

size_t qsim_cumsum_workspace(int64_t n) { return n < 1 ? 0 : cumsum_ws_bytes(n); }

int qsim_cumsum_exact(const void* probs, int prob_dtype, int64_t n, double* cdf, void* workspace,
                      size_t workspace_bytes, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(probs != nullptr && cdf != nullptr, "qsim_cumsum_exact: null pointer");
  // one 1-D grid of 256-thread workgroups per 4096-element tile: n <= 2^35
  // keeps it below HIP's 2^32 work-items per grid dimension
  QSIM_CHECK_ARG(n >= 1 && n <= (int64_t(1) << 35), "qsim_cumsum_exact: length must be in [1, 2^35]");
  QSIM_CHECK_ARG(prob_dtype == QSIM_F64 || prob_dtype == QSIM_F32, "qsim_cumsum_exact: bad dtype");
  QSIM_CHECK_ARG(workspace != nullptr && workspace_bytes >= cumsum_ws_bytes(n),
                 "qsim_cumsum_exact: workspace too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (prob_dtype == QSIM_F64) return cumsum_impl<double>(probs, n, cdf, workspace, s);
  return cumsum_impl<float>(probs, n, cdf, workspace, s);
}

size_t qsim_sample_workspace(int64_t n) {
  return (n < 1 || n > (int64_t(1) << 35)) ? 0 : sample_ws_bytes(n);
}

size_t qsim_tile_stats_bytes(int64_t n) { return (n < 1 || n > (int64_t(1) << 35)) ? 0 : tile_stats_bytes(n); }

int qsim_probabilities_tiles(const void* state, int nqubits, int dtype, void* probs, void* tile_stats,
                             size_t tile_stats_bytes_, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && probs != nullptr && tile_stats != nullptr,
                 "qsim_probabilities_tiles: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_probabilities_tiles: bad dtype");
  QSIM_CHECK_ARG(nqubits >= 0 && nqubits <= 35 && nqubits <= QSIM_MAX_QUBITS,
                 "qsim_probabilities_tiles: bad number of qubits");
  const int64_t n = int64_t(1) << nqubits;
  QSIM_CHECK_ARG(tile_stats_bytes_ >= tile_stats_bytes(n), "qsim_probabilities_tiles: tile_stats too small");
  const int64_t T = (n + kTile - 1) / kTile;
  double* tsum = reinterpret_cast<double*>(tile_stats);
  int2* tbad = reinterpret_cast<int2*>(reinterpret_cast<char*>(tile_stats) + align256(T * sizeof(double)));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_probs_tiles<double>, launch_grid(uint64_t(T)), dim3(kT), 0, s,
                       reinterpret_cast<const cplx<double>*>(state), reinterpret_cast<double*>(probs), n, tsum, tbad);
  else
    hipLaunchKernelGGL(k_probs_tiles<float>, launch_grid(uint64_t(T)), dim3(kT), 0, s,
                       reinterpret_cast<const cplx<float>*>(state), reinterpret_cast<float*>(probs), n, tsum, tbad);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_sample_exact(const void* probs, int prob_dtype, int64_t n, const void* tile_stats, const double* uniforms,
                      int64_t nshots, int64_t* out, double* stats, void* workspace, size_t workspace_bytes,
                      void* stream) {
  clear_error();
  QSIM_CHECK_ARG(probs != nullptr && stats != nullptr, "qsim_sample_exact: null pointer");
  QSIM_CHECK_ARG(nshots == 0 || (uniforms != nullptr && out != nullptr), "qsim_sample_exact: null pointer");
  QSIM_CHECK_ARG(n >= 1 && n <= (int64_t(1) << 35), "qsim_sample_exact: length must be in [1, 2^35]");
  QSIM_CHECK_ARG(nshots >= 0 && nshots <= (int64_t(1) << 40), "qsim_sample_exact: bad number of shots");
  QSIM_CHECK_ARG(prob_dtype == QSIM_F64 || prob_dtype == QSIM_F32, "qsim_sample_exact: bad dtype");
  QSIM_CHECK_ARG(workspace != nullptr && workspace_bytes >= sample_ws_bytes(n),
                 "qsim_sample_exact: workspace too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (prob_dtype == QSIM_F64)
    return sample_impl<double>(probs, n, tile_stats, uniforms, nshots, out, stats, workspace, s);
  return sample_impl<float>(probs, n, tile_stats, uniforms, nshots, out, stats, workspace, s);
}

size_t qsim_np_sum_workspace(int64_t n) {
  return n < 1 ? 0 : size_t((n + kNpBuf - 1) / kNpBuf) * sizeof(double);
}

int qsim_np_sum(const void* probs, int prob_dtype, int64_t n, void* out, void* workspace, size_t workspace_bytes,
                void* stream) {
  clear_error();
  QSIM_CHECK_ARG(probs != nullptr && out != nullptr && workspace != nullptr, "qsim_np_sum: null pointer");
  QSIM_CHECK_ARG(n >= 1 && (n & (n - 1)) == 0, "qsim_np_sum: length must be a power of two");
  QSIM_CHECK_ARG(prob_dtype == QSIM_F64 || prob_dtype == QSIM_F32, "qsim_np_sum: bad dtype");
  const int64_t nbuf = (n + kNpBuf - 1) / kNpBuf;
  QSIM_CHECK_ARG(workspace_bytes >= size_t(nbuf) * sizeof(double), "qsim_np_sum: workspace too small");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (prob_dtype == QSIM_F64) {
    hipLaunchKernelGGL(k_np_buffer_sums<double>, dim3(unsigned(nbuf)), dim3(64), 0, s, (const double*)probs, n,
                       (double*)workspace);
    QSIM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_np_total<double>, dim3(1), dim3(1), 0, s, (const double*)workspace, nbuf, (double*)out);
  } else {
    hipLaunchKernelGGL(k_np_buffer_sums<float>, dim3(unsigned(nbuf)), dim3(64), 0, s, (const float*)probs, n,
                       (float*)workspace);
    QSIM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_np_total<float>, dim3(1), dim3(1), 0, s, (const float*)workspace, nbuf, (float*)out);
  }
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_divide_by(const void* probs, int prob_dtype, int64_t n, const void* divisor, void* out, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(probs != nullptr && divisor != nullptr && out != nullptr, "qsim_divide_by: null pointer");
  QSIM_CHECK_ARG(n >= 1, "qsim_divide_by: bad length");
  QSIM_CHECK_ARG(prob_dtype == QSIM_F64 || prob_dtype == QSIM_F32, "qsim_divide_by: bad dtype");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (prob_dtype == QSIM_F64)
    hipLaunchKernelGGL(k_div_by<double>, launch_grid(ceil_div_u(uint64_t(n), kT)), dim3(kT), 0, s, (const double*)probs, n,
                       (const double*)divisor, (double*)out);
  else
    hipLaunchKernelGGL(k_div_by<float>, launch_grid(ceil_div_u(uint64_t(n), kT)), dim3(kT), 0, s, (const float*)probs, n,
                       (const float*)divisor, (float*)out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_search_sorted(const double* cdf, int64_t n, const double* uniforms, int64_t nshots, int64_t* out,
                       void* stream) {
  clear_error();
  QSIM_CHECK_ARG(cdf != nullptr && uniforms != nullptr && out != nullptr, "qsim_search_sorted: null pointer");
  QSIM_CHECK_ARG(n >= 1 && nshots >= 0, "qsim_search_sorted: bad sizes");
  if (nshots == 0) return QSIM_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_search, launch_grid(ceil_div_u(uint64_t(nshots), kT)), dim3(kT), 0, s, cdf, n, uniforms, nshots, out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_samples_to_binary(const int64_t* samples, int64_t nshots, int nbits, int64_t* out, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(samples != nullptr && out != nullptr, "qsim_samples_to_binary: null pointer");
  QSIM_CHECK_ARG(nbits >= 1 && nbits <= 63 && nshots >= 0, "qsim_samples_to_binary: bad sizes");
  if (nshots == 0) return QSIM_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const uint64_t total = uint64_t(nshots) * uint64_t(nbits);
  hipLaunchKernelGGL(k_to_binary, launch_grid(ceil_div_u(total, kT)), dim3(kT), 0, s, samples, nshots, nbits, out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_histogram(const int64_t* samples, int64_t nshots, int64_t* counts, int64_t nbins, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(samples != nullptr && counts != nullptr, "qsim_histogram: null pointer");
  QSIM_CHECK_ARG(nbins >= 1 && nshots >= 0, "qsim_histogram: bad sizes");
  if (nshots == 0) return QSIM_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_histogram, launch_grid(ceil_div_u(uint64_t(nshots), kT)), dim3(kT), 0, s, samples, nshots,
                     reinterpret_cast<unsigned long long*>(counts), nbins);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_apply_bitflips(const int64_t* samples, int64_t nshots, int nbits, const double* uniforms,
                        const double* p0, const double* p1, int64_t* out, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(samples != nullptr && uniforms != nullptr && out != nullptr && p0 != nullptr && p1 != nullptr,
                 "qsim_apply_bitflips: null pointer");
  QSIM_CHECK_ARG(nbits >= 1 && nbits <= kMaxFixed && nshots >= 0 && nshots <= (int64_t(1) << 40),
                 "qsim_apply_bitflips: bad sizes");
  BitflipArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = nshots * nbits;
  a.nbits = nbits;
  for (int k = 0; k < nbits; ++k) {
    a.p0[k] = p0[k];
    a.p1[k] = p1[k];
  }
  if (a.n == 0) return QSIM_OK;
  hipLaunchKernelGGL(k_bitflips, launch_grid(ceil_div_u(uint64_t(a.n), kT)), dim3(kT), 0,
                     reinterpret_cast<hipStream_t>(stream), samples, uniforms, a, out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

}  // extern "C"
