// Dense k-qubit blocks (k <= 5) with the matrix in device memory: the
// application of a fused gate group.
//
// psi viewed as a (2^k) x (2^(n-k)) matrix after gathering the k target bits
// is multiplied on the left by the 2^k x 2^k block (abstract.py:2350-2361
// with the matrix of matrix_fused, abstract.py:2680-2717). Each work item is
// one column: its 2^k amplitudes are loaded to registers, the block is read
// from LDS (every lane reads the same element, a broadcast), and row r of the
// product is stored as soon as it is complete (in place is safe: a column is
// owned by one lane).

#include <hip/hip_runtime.h>

#include <cstring>

#include "qsim_common.h"

namespace qsim {

constexpr int kFusedThreads = 256;

struct FusedArgs {
  uint64_t nwork;
  uint64_t off[32];
  int nfixed;
  uint8_t pos[kMaxFixed];
};

template <typename R, int K>
__global__ __launch_bounds__(kFusedThreads) void k_fused_fp(const FusedArgs a, const cplx<R>* __restrict__ mat,
                                                             cplx<R>* __restrict__ psi) {
  constexpr int D = 1 << K;
  __shared__ cplx<R> sm[D * D];
  for (int i = threadIdx.x; i < D * D; i += kFusedThreads) sm[i] = mat[i];
  __syncthreads();
  const uint64_t j = uint64_t(blockIdx.x) * kFusedThreads + threadIdx.x;
  if (j >= a.nwork) return;
  const uint64_t base = insert_zeros(j, a.pos, a.nfixed);
  cplx<R> v[D];
#pragma unroll
  for (int c = 0; c < D; ++c) v[c] = ld(psi, base | a.off[c]);
#pragma unroll 2
  for (int r = 0; r < D; ++r) {
    cplx<R> acc = cmul(sm[r * D], v[0]);
#pragma unroll
    for (int c = 1; c < D; ++c) acc = cfma(sm[r * D + c], v[c], acc);
    st(psi, base | a.off[r], acc);
  }
}

template <typename R, int K>
static int launch_fused_fp(void* state, int n, const int* targets, const void* mat, hipStream_t s) {
  FusedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nfixed = sorted_positions(n, targets, K, nullptr, 0, a.pos);
  a.nwork = uint64_t(1) << (n - K);
  for (int r = 0; r < (1 << K); ++r) {
    uint64_t off = 0;
    for (int t = 0; t < K; ++t)
      if ((r >> (K - 1 - t)) & 1) off |= uint64_t(1) << (n - 1 - targets[t]);
    a.off[r] = off;
  }
  const int blocks = ceil_div_i(a.nwork, kFusedThreads);
  hipLaunchKernelGGL((k_fused_fp<R, K>), dim3(blocks), dim3(kFusedThreads), 0, s, a,
                     reinterpret_cast<const cplx<R>*>(mat), reinterpret_cast<cplx<R>*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

template <typename R>
static int dispatch_fused(void* state, int n, const int* targets, int k, const void* mat, hipStream_t s) {
  switch (k) {
    case 1: return launch_fused_fp<R, 1>(state, n, targets, mat, s);
    case 2: return launch_fused_fp<R, 2>(state, n, targets, mat, s);
    case 3: return launch_fused_fp<R, 3>(state, n, targets, mat, s);
    case 4: return launch_fused_fp<R, 4>(state, n, targets, mat, s);
    default: return launch_fused_fp<R, 5>(state, n, targets, mat, s);
  }
}

}  // namespace qsim

using namespace qsim;

extern "C" int qsim_apply_fused(void* state, int nqubits, int dtype, const int* targets, int ntargets,
                                const void* matrix_dev, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && matrix_dev != nullptr, "qsim_apply_fused: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_fused: bad dtype");
  QSIM_CHECK_ARG(ntargets >= 1 && ntargets <= 5, "qsim_apply_fused: 1..5 targets supported");
  if (!validate_qubits(nqubits, targets, ntargets, nullptr, 0, "qsim_apply_fused")) return QSIM_EINVAL;
  QSIM_CHECK_ARG(nqubits >= ntargets, "qsim_apply_fused: more targets than qubits");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128) return dispatch_fused<double>(state, nqubits, targets, ntargets, matrix_dev, s);
  return dispatch_fused<float>(state, nqubits, targets, ntargets, matrix_dev, s);
}
