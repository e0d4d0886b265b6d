// Dense k-qubit blocks (k <= 5) with the matrix in device memory: the
// application of a fused gate group. complex128 blocks and complex64 blocks
// of fewer than 4 qubits use FMA arithmetic per column (k_fused_fp);
// complex64 blocks of 4 or 5 qubits use the MFMA kernel further below.
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
  const uint64_t j = block_id() * kFusedThreads + threadIdx.x;
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
  hipLaunchKernelGGL((k_fused_fp<R, K>), launch_grid(ceil_div_u(a.nwork, kFusedThreads)), dim3(kFusedThreads), 0, s, a,
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


// ---------------------------------------------------------------------------
// complex64 blocks of k = 4 or 5 qubits on MFMA (v_mfma_f32_32x32x2f32)
// ---------------------------------------------------------------------------
// The block acts on the 2^k-amplitude "columns" of the state (one column per
// setting of the non-target bits). With the real embedding
//     [y_re; y_im] = [[M_re, -M_im], [M_im, M_re]] [x_re; x_im]
// a batch of 32 columns is a (2D x 2D) x (2D x 32) real product, D = 2^k,
// computed by 32x32x2 FP32 MFMA steps with FP32 accumulation. A workgroup
// stages a tile of 4096 amplitudes (k target bits + 12-k column bits, the
// column bits being the lowest non-target bits so that loads are coalesced)
// in LDS, each wavefront multiplies groups of 32 columns, and the tile is
// streamed back. The embedded matrix is held in registers as MFMA A
// fragments for the whole workgroup.
//
// Operand layout of v_mfma_f32_32x32x2f32 (one 32x32 block, K = 2):
//   A: lane l supplies A[l % 32][l / 32];  B: lane l supplies B[l / 32][l % 32];
//   D: register j of lane l holds D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kMfmaTileBits = 12;

struct MfmaArgs {
  uint64_t ntiles;
  uint8_t tbits[kMfmaTileBits];   // global bit of local tile bit i, ascending
  uint8_t rowbit[5];              // local tile bit of matrix index bit (K-1-t), t = target order
  uint8_t colbit[8];              // local tile bits of the column index, ascending
};

// Local tile index of (row r, column c).
template <int K>
__device__ __forceinline__ uint32_t mfma_local(const MfmaArgs& a, uint32_t r, uint32_t c) {
  uint32_t u = 0;
#pragma clang loop unroll(full)
  for (int t = 0; t < K; ++t)
    if ((r >> (K - 1 - t)) & 1) u |= 1u << a.rowbit[t];
#pragma clang loop unroll(full)
  for (int j = 0; j < kMfmaTileBits - K; ++j)
    if ((c >> j) & 1) u |= 1u << a.colbit[j];
  return u;
}

template <int K>
__global__ __launch_bounds__(256) void k_fused_mfma_c64(const MfmaArgs a, const c64* __restrict__ mat,
                                                       c64* __restrict__ psi) {
  constexpr int D = 1 << K;             // complex rows
  constexpr int MB = (2 * D) / 32;      // 32-row blocks of the real embedding (1 or 2)
  constexpr int KS = (2 * D) / 2;       // K = 2 steps over the 2D real inner dimension
  constexpr int NCOL = 1 << (kMfmaTileBits - K);
  constexpr int NCG = NCOL / 32;        // 32-column groups (4 for K = 5, 8 for K = 4)
  __shared__ c64 tile[1 << kMfmaTileBits];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t t = xcd_block();
  const uint64_t base = insert_zeros(t, a.tbits, kMfmaTileBits);
  __shared__ uint32_t rowoff[D];  // local-index bits of matrix row r
  if (threadIdx.x < D) rowoff[threadIdx.x] = mfma_local<K>(a, threadIdx.x, 0);
  // coalesced streaming load: thread u of each 256-slice reads local index
  // u, whose global offset is the lane part (bits 0-7) | the slice part (8-11)
  uint64_t glow = 0;
#pragma clang loop unroll(full)
  for (int i = 0; i < 8; ++i)
    if ((threadIdx.x >> i) & 1) glow |= uint64_t(1) << a.tbits[i];
  glow |= base;
  c64 ld[16];
#pragma clang loop unroll(full)
  for (int k = 0; k < 16; ++k) {
    uint64_t g = glow;
#pragma clang loop unroll(full)
    for (int i = 0; i < 4; ++i)
      if ((k >> i) & 1) g |= uint64_t(1) << a.tbits[8 + i];
    ld[k] = ld_nt(psi, g);
  }
#pragma clang loop unroll(full)
  for (int k = 0; k < 16; ++k) tile[k * 256 + threadIdx.x] = ld[k];
  // A fragments of the real embedding: row i = 32 mb + l % 32, col k = 2 ks + l / 32
  float afrag[MB][KS];
#pragma clang loop unroll(full)
  for (int mb = 0; mb < MB; ++mb) {
#pragma clang loop unroll(full)
    for (int ks = 0; ks < KS; ++ks) {
      const int i = 32 * mb + (lane & 31), k = 2 * ks + (lane >> 5);
      const int ri = i % D, ci = k % D;
      const c64 m = mat[ri * D + ci];
      float v;
      if (i < D) v = (k < D) ? m.re : -m.im;
      else v = (k < D) ? m.im : m.re;
      afrag[mb][ks] = v;
    }
  }
  __syncthreads();
  // one 32-column group per wave at a time; the groups are disjoint, so each
  // wave reads and writes only its own columns of the tile
  for (int cg = wave; cg < NCG; cg += 4) {
    const uint32_t colpart = mfma_local<K>(a, 0, 32 * cg + (lane & 31));
    floatx16 acc[MB];
#pragma clang loop unroll(full)
    for (int mb = 0; mb < MB; ++mb) acc[mb] = floatx16(0.0f);
#pragma clang loop unroll(full)
    for (int ks = 0; ks < KS; ++ks) {
      // real inner index k = 2 ks + lane / 32: k < D -> re of row k, else im of row k - D
      const c64 x = tile[rowoff[(2 * ks + (lane >> 5)) % D] | colpart];
      const float b = (2 * ks + (lane >> 5) < D) ? x.re : x.im;
#pragma clang loop unroll(full)
      for (int mb = 0; mb < MB; ++mb)
        acc[mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(afrag[mb][ks], b, acc[mb], 0, 0, 0);
    }
    // D rows: real part rows [0, D), imaginary part rows [D, 2D); the wave's
    // reads of these columns are complete (they fed the accumulators above)
#pragma clang loop unroll(full)
    for (int j = 0; j < 16; ++j) {
      const int row = 8 * (j / 4) + 4 * (lane >> 5) + (j % 4);
#pragma clang loop unroll(full)
      for (int mb = 0; mb < MB; ++mb) {
        const int i = 32 * mb + row;
        if (i < D) {
          tile[rowoff[i] | colpart].re = acc[mb][j];
        } else {
          tile[rowoff[i - D] | colpart].im = acc[mb][j];
        }
      }
    }
  }
  __syncthreads();
#pragma clang loop unroll(full)
  for (int k = 0; k < 16; ++k) ld[k] = tile[k * 256 + threadIdx.x];
#pragma clang loop unroll(full)
  for (int k = 0; k < 16; ++k) {
    uint64_t g = glow;
#pragma clang loop unroll(full)
    for (int i = 0; i < 4; ++i)
      if ((k >> i) & 1) g |= uint64_t(1) << a.tbits[8 + i];
    st_nt(psi, g, ld[k]);
  }
}

template <int K>
static int launch_fused_mfma(void* state, int n, const int* targets, const void* mat, hipStream_t s) {
  MfmaArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntiles = uint64_t(1) << (n - kMfmaTileBits);
  // tile bits: the K target bits plus the lowest non-target bits
  bool is_target[QSIM_MAX_QUBITS] = {false};
  for (int t = 0; t < K; ++t) is_target[n - 1 - targets[t]] = true;
  int nb = 0;
  uint8_t bits[kMfmaTileBits];
  for (int b = 0; b < n && nb < kMfmaTileBits; ++b) {
    int need_cols = kMfmaTileBits - K;
    int cols_so_far = 0;
    for (int j = 0; j < nb; ++j) cols_so_far += is_target[bits[j]] ? 0 : 1;
    if (is_target[b] || cols_so_far < need_cols) bits[nb++] = uint8_t(b);
  }
  for (int i = 0; i < kMfmaTileBits; ++i) a.tbits[i] = bits[i];
  int nc = 0;
  for (int i = 0; i < kMfmaTileBits; ++i) {
    if (!is_target[bits[i]]) a.colbit[nc++] = uint8_t(i);
  }
  for (int t = 0; t < K; ++t) {
    const int gb = n - 1 - targets[t];
    for (int i = 0; i < kMfmaTileBits; ++i)
      if (bits[i] == gb) a.rowbit[t] = uint8_t(i);
  }
  hipLaunchKernelGGL((k_fused_mfma_c64<K>), launch_grid(a.ntiles), dim3(256), 0, s, a,
                     reinterpret_cast<const c64*>(mat), reinterpret_cast<c64*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
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
  if (nqubits >= kMfmaTileBits && ntargets == 5) return launch_fused_mfma<5>(state, nqubits, targets, matrix_dev, s);
  if (nqubits >= kMfmaTileBits && ntargets == 4) return launch_fused_mfma<4>(state, nqubits, targets, matrix_dev, s);
  return dispatch_fused<float>(state, nqubits, targets, ntargets, matrix_dev, s);
}
