// Comparison of two states for accuracy reports: the overlap <a|b>, both
// norms and the largest amplitude difference, accumulated in FP64 in one
// streaming read of each state (a complex64 state is widened per element, so
// no complex128 copy of it is made).
//
// Reference quantities: qibo's pure-state fidelity |conj(a) @ b|^2 and
// infidelity 1 - fidelity (src/qibo/quantum_info/metrics.py:200-291); the
// max-abs amplitude error max_i |a_i - b_i|.
//
// Stage 1: a fixed grid of workgroups walks the state with a grid stride;
// each workgroup writes five partials (Re, Im of the overlap, |a|^2, |b|^2,
// max |a - b|). Stage 2: one workgroup reduces the partials in a fixed order,
// so repeated calls give bitwise identical results.

#include <hip/hip_runtime.h>

#include "qsim_common.h"

namespace qsim {

constexpr int kCmpThreads = 256;
constexpr int kCmpBlocks = 2048;
constexpr int kCmpFields = 5;

template <typename R>
__device__ __forceinline__ void widen(const cplx<R>* p, uint64_t i, double& re, double& im) {
  const cplx<R> v = ld_nt(p, i);
  re = double(v.re);
  im = double(v.im);
}

__device__ __forceinline__ double wave_sum_d(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

__device__ __forceinline__ double wave_max_d(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
  return x;
}

// Workgroup reduction of five values; thread 0 returns the totals in v.
__device__ __forceinline__ void block_reduce5(double* v) {
  __shared__ double red[kCmpFields][kCmpThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = wave_sum_d(v[k]);
  v[4] = wave_max_d(v[4]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kCmpFields; ++k) red[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
    v[4] = fmax(fmax(red[4][0], red[4][1]), fmax(red[4][2], red[4][3]));
  }
}

template <typename RA, typename RB>
__global__ __launch_bounds__(kCmpThreads) void k_compare(const cplx<RA>* __restrict__ a,
                                                         const cplx<RB>* __restrict__ b, uint64_t n,
                                                         double* __restrict__ partials) {
  double v[kCmpFields] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const uint64_t stride = uint64_t(gridDim.x) * kCmpThreads;
  for (uint64_t i = uint64_t(blockIdx.x) * kCmpThreads + threadIdx.x; i < n; i += stride) {
    double ar, ai, br, bi;
    widen(a, i, ar, ai);
    widen(b, i, br, bi);
    // conj(a) b = (ar br + ai bi) + i (ar bi - ai br)
    v[0] = fma(ar, br, fma(ai, bi, v[0]));
    v[1] = fma(ar, bi, fma(-ai, br, v[1]));
    v[2] = fma(ar, ar, fma(ai, ai, v[2]));
    v[3] = fma(br, br, fma(bi, bi, v[3]));
    v[4] = fmax(v[4], hypot(ar - br, ai - bi));
  }
  block_reduce5(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kCmpFields; ++k) partials[uint64_t(blockIdx.x) * kCmpFields + k] = v[k];
  }
}

__global__ __launch_bounds__(kCmpThreads) void k_compare_final(const double* __restrict__ partials, int nblocks,
                                                               double* __restrict__ out) {
  double v[kCmpFields] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int blk = threadIdx.x; blk < nblocks; blk += kCmpThreads) {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] += partials[blk * kCmpFields + k];
    v[4] = fmax(v[4], partials[blk * kCmpFields + 4]);
  }
  block_reduce5(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kCmpFields; ++k) out[k] = v[k];
  }
}

static int compare_blocks(uint64_t n) {
  const uint64_t want = ceil_div_u(n, kCmpThreads);
  return int(want < uint64_t(kCmpBlocks) ? (want > 0 ? want : 1) : uint64_t(kCmpBlocks));
}

template <typename RA, typename RB>
static int launch_compare(const void* a, const void* b, uint64_t n, double* out, double* partials, hipStream_t s) {
  const int nb = compare_blocks(n);
  hipLaunchKernelGGL((k_compare<RA, RB>), dim3(nb), dim3(kCmpThreads), 0, s, reinterpret_cast<const cplx<RA>*>(a),
                     reinterpret_cast<const cplx<RB>*>(b), n, partials);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_compare_final, dim3(1), dim3(kCmpThreads), 0, s, partials, nb, out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

extern "C" size_t qsim_compare_workspace(int nqubits) {
  if (nqubits < 0 || nqubits > QSIM_MAX_QUBITS) return 0;
  return size_t(compare_blocks(uint64_t(1) << nqubits)) * kCmpFields * sizeof(double);
}

extern "C" int qsim_compare_states(const void* a, int a_dtype, const void* b, int b_dtype, int nqubits, double* out,
                                   void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(a != nullptr && b != nullptr && out != nullptr && workspace != nullptr,
                 "qsim_compare_states: null pointer");
  QSIM_CHECK_ARG(a_dtype == QSIM_C128 || a_dtype == QSIM_C64, "qsim_compare_states: bad dtype of a");
  QSIM_CHECK_ARG(b_dtype == QSIM_C128 || b_dtype == QSIM_C64, "qsim_compare_states: bad dtype of b");
  QSIM_CHECK_ARG(nqubits >= 0 && nqubits <= QSIM_MAX_QUBITS, "qsim_compare_states: nqubits out of range");
  QSIM_CHECK_ARG(workspace_bytes >= qsim_compare_workspace(nqubits), "qsim_compare_states: workspace too small");
  const uint64_t n = uint64_t(1) << nqubits;
  double* partials = reinterpret_cast<double*>(workspace);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (a_dtype == QSIM_C128)
    return b_dtype == QSIM_C128 ? launch_compare<double, double>(a, b, n, out, partials, s)
                                : launch_compare<double, float>(a, b, n, out, partials, s);
  return b_dtype == QSIM_C128 ? launch_compare<float, double>(a, b, n, out, partials, s)
                              : launch_compare<float, float>(a, b, n, out, partials, s);
}
