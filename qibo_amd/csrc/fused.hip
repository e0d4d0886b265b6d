// Dense k-qubit blocks (k <= 5) with the matrix in device memory: the
// application of a fused gate group. Blocks of fewer than 4 qubits use FMA
// arithmetic per column (k_fused_fp); blocks of 4 or 5 qubits run on MFMA
// through an LDS tile (complex64: FP32 MFMA, k_fused_mfma_c64; complex128:
// FP64 MFMA, k_fused_tile_c128), further below.
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
// setting of the non-target bits). For k = 4, with the real embedding
//     [y_re; y_im] = [[M_re, -M_im], [M_im, M_re]] [x_re; x_im]
// a batch of 32 columns is a (2D x 2D) x (2D x 32) real product, D = 2^k,
// computed by 32x32x2 FP32 MFMA steps with FP32 accumulation; for k = 5 the
// three 32 x 32 real products of the Gauss form (see the complex128 kernel)
// take 48 MFMA steps per 32 columns instead of 64. A workgroup
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

// LDS slot of local index u: bits 0-4 ^= bits 5-9. Column groups whose 32
// columns are 32 apart (target bits at the bottom of the tile) would
// otherwise put all 32 lanes on one bank; the slices of the streaming phases
// stay conflict-free (a permutation inside each 32-slot row).
__device__ __forceinline__ uint32_t mfma_slot(uint32_t u) { return u ^ ((u >> 5) & 31u); }

// Staging moves pairs of neighbouring amplitudes (16 bytes per lane): the
// tile's lowest local bit is global bit 0 (the lowest index bit is always a
// target or a column bit), so elements 2p and 2p + 1 are adjacent in memory.
// Global index of element 2 (k * 256 + thread) of the tile whose local bits
// 1-8 (the thread) are already in glow (coalesced 4 KB slices).
__device__ __forceinline__ uint64_t mfma_addr(const MfmaArgs& a, uint64_t glow, int k) {
  uint64_t g = glow;
#pragma clang loop unroll(full)
  for (int i = 0; i < 3; ++i)
    if ((k >> i) & 1) g |= uint64_t(1) << a.tbits[9 + i];
  return g;
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 ld_pair(const c64* psi, uint64_t i) {
  return __builtin_nontemporal_load(reinterpret_cast<const floatx4*>(psi + i));
}

__device__ __forceinline__ void st_pair(c64* psi, uint64_t i, floatx4 v) {
  __builtin_nontemporal_store(v, reinterpret_cast<floatx4*>(psi + i));
}

// Persistent workgroups, software-pipelined: while the MFMA product of tile t
// runs, the loads of the workgroup's next tile are in flight, so memory
// latency is hidden inside each workgroup (tiles t = blockIdx.x + i gridDim.x).
template <int K>
__global__ __launch_bounds__(256) void k_fused_mfma_c64(const MfmaArgs a, const c64* __restrict__ mat,
                                                       c64* __restrict__ psi) {
  constexpr int D = 1 << K;             // complex rows
  constexpr int MB = (2 * D) / 32;      // 32-row blocks of the real embedding (1 or 2)
  constexpr int KS = (2 * D) / 2;       // K = 2 steps over the 2D real inner dimension
  constexpr int NCOL = 1 << (kMfmaTileBits - K);
  constexpr int NCG = NCOL / 32;        // 32-column groups (4 for K = 5, 8 for K = 4)
  __shared__ c64 tile[1 << kMfmaTileBits];
  __shared__ uint32_t rowoff[D];  // local-index bits of matrix row r
  // K = 4: the real embedding, column-major (element (i, k) at k * 2D + i):
  // the A operand of step (mb, ks) is two consecutive 32-float runs.
  // K = 5: the three real 32 x 32 matrices of the Gauss form (below) in lane
  // order: fragment (p, ks), lane l = M_p[l % 32][2 ks + l / 32] with M_0 =
  // Re M, M_1 = Im M, M_2 = Re M + Im M.
  constexpr bool GAUSS = K == 5;
  constexpr int NA = GAUSS ? 3 * (D / 2) * 64 : 4 * D * D;
  __shared__ float amat[NA];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x < D) rowoff[threadIdx.x] = mfma_local<K>(a, threadIdx.x, 0);
  for (int e = threadIdx.x; e < NA; e += 256) {
    if constexpr (GAUSS) {
      const int l = e & 63, f = e >> 6;
      const int ks = f % (D / 2), pr = f / (D / 2);
      const c64 m = mat[(l & 31) * D + 2 * ks + (l >> 5)];
      amat[e] = pr == 0 ? m.re : (pr == 1 ? m.im : m.re + m.im);
    } else {
      const int k = e / (2 * D), i = e % (2 * D);
      const c64 m = mat[(i % D) * D + (k % D)];
      amat[e] = (i < D) ? ((k < D) ? m.re : -m.im) : ((k < D) ? m.im : m.re);
    }
  }
  // lane part (local bits 1-8) of this thread's slice pair, as global bits
  uint64_t lanebits = 0;
#pragma clang loop unroll(full)
  for (int i = 0; i < 8; ++i)
    if ((threadIdx.x >> i) & 1) lanebits |= uint64_t(1) << a.tbits[i + 1];
  uint64_t t = blockIdx.x;
  floatx4 nxt[8];  // pair k: elements 2 (k * 256 + thread) and 2 (k * 256 + thread) + 1
  if (t < a.ntiles) {
    const uint64_t glow = insert_zeros(t, a.tbits, kMfmaTileBits) | lanebits;
#pragma clang loop unroll(full)
    for (int k = 0; k < 8; ++k) nxt[k] = ld_pair(psi, mfma_addr(a, glow, k));
  }
  while (t < a.ntiles) {
    const uint64_t glow = insert_zeros(t, a.tbits, kMfmaTileBits) | lanebits;
    // each thread writes (and later reads back) only its own slots here
#pragma clang loop unroll(full)
    for (int k = 0; k < 8; ++k) {
      const uint32_t e = 2u * (uint32_t(k) * 256u + threadIdx.x);
      tile[mfma_slot(e)] = c64{nxt[k].x, nxt[k].y};
      tile[mfma_slot(e + 1)] = c64{nxt[k].z, nxt[k].w};
    }
    __syncthreads();
    const uint64_t tn = t + gridDim.x;
    if (tn < a.ntiles) {  // next tile's loads overlap this tile's product
      const uint64_t gnext = insert_zeros(tn, a.tbits, kMfmaTileBits) | lanebits;
#pragma clang loop unroll(full)
      for (int k = 0; k < 8; ++k) nxt[k] = ld_pair(psi, mfma_addr(a, gnext, k));
    }
    // one 32-column group per wave at a time; the groups are disjoint, so each
    // wave reads and writes only its own columns of the tile
    for (int cg = wave; cg < NCG; cg += 4) {
      const uint32_t colpart = mfma_local<K>(a, 0, 32 * cg + (lane & 31));
      // the A operands are read from LDS at each use (an opaque lane index
      // keeps the compiler from hoisting all 64 of them into registers)
      int alane = lane;
      asm volatile("" : "+v"(alane));
      if constexpr (GAUSS) {
        // Y = M X as three real products (Gauss): P1 = Re M Re X, P2 = Im M
        // Im X, P3 = (Re M + Im M)(Re X + Im X); Re Y = P1 - P2, Im Y =
        // (P3 - P1) - P2. 48 MFMA steps per 32 columns instead of the 64 of
        // the real embedding.
        floatx16 acc0 = floatx16(0.0f), acc1 = floatx16(0.0f), acc2 = floatx16(0.0f);
#pragma clang loop unroll(full)
        for (int ks = 0; ks < D / 2; ++ks) {
          const c64 x = tile[mfma_slot(rowoff[2 * ks + (lane >> 5)] | colpart)];
          const float b0 = x.re, b1 = x.im, b2 = x.re + x.im;
          acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(amat[(0 * (D / 2) + ks) * 64 + alane], b0, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(amat[(1 * (D / 2) + ks) * 64 + alane], b1, acc1, 0, 0, 0);
          acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(amat[(2 * (D / 2) + ks) * 64 + alane], b2, acc2, 0, 0, 0);
        }
#pragma clang loop unroll(full)
        for (int j = 0; j < 16; ++j) {
          const int row = 8 * (j / 4) + 4 * (lane >> 5) + (j % 4);
          c64 y;
          y.re = acc0[j] - acc1[j];
          y.im = (acc2[j] - acc0[j]) - acc1[j];
          tile[mfma_slot(rowoff[row] | colpart)] = y;
        }
        continue;
      }
      floatx16 acc[MB];
#pragma clang loop unroll(full)
      for (int mb = 0; mb < MB; ++mb) acc[mb] = floatx16(0.0f);
#pragma clang loop unroll(full)
      for (int ks = 0; ks < KS; ++ks) {
        // real inner index k = 2 ks + lane / 32: k < D -> re of row k, else im of row k - D
        const c64 x = tile[mfma_slot(rowoff[(2 * ks + (lane >> 5)) % D] | colpart)];
        const float b = (2 * ks + (lane >> 5) < D) ? x.re : x.im;
#pragma clang loop unroll(full)
        for (int mb = 0; mb < MB; ++mb) {
          const float av = amat[(2 * ks + (alane >> 5)) * (2 * D) + 32 * mb + (alane & 31)];
          acc[mb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, b, acc[mb], 0, 0, 0);
        }
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
            tile[mfma_slot(rowoff[i] | colpart)].re = acc[mb][j];
          } else {
            tile[mfma_slot(rowoff[i - D] | colpart)].im = acc[mb][j];
          }
        }
      }
    }
    __syncthreads();
#pragma clang loop unroll(full)
    for (int k = 0; k < 8; ++k) {
      const uint32_t e = 2u * (uint32_t(k) * 256u + threadIdx.x);
      const c64 y0 = tile[mfma_slot(e)], y1 = tile[mfma_slot(e + 1)];
      st_pair(psi, mfma_addr(a, glow, k), floatx4{y0.re, y0.im, y1.re, y1.im});
    }
    t = tn;
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
  // persistent grid: as many workgroups as the device keeps resident
  static int per_cu[2] = {0, 0};  // per K = 4, 5 (the answer is fixed for the kernel)
  static int cus = 0;
  if (per_cu[K - 4] == 0) {
    int nb = 0, dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_mfma_c64<K>, 256, 0) != hipSuccess || nb < 1)
      nb = 2;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    per_cu[K - 4] = nb;
  }
  const uint64_t resident = uint64_t(cus) * uint64_t(per_cu[K - 4]);
  const unsigned grid = unsigned(a.ntiles < resident ? a.ntiles : resident);
  hipLaunchKernelGGL((k_fused_mfma_c64<K>), dim3(grid), dim3(256), 0, s, a, reinterpret_cast<const c64*>(mat),
                     reinterpret_cast<c64*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

// ---------------------------------------------------------------------------
// complex128 blocks: FP64 MFMA (v_mfma_f64_16x16x4f64)
// ---------------------------------------------------------------------------
// Y = M X for 16 columns at a time (one wavefront), as three real products
// (Gauss): P1 = Re M Re X, P2 = Im M Im X, P3 = (Re M + Im M)(Re X + Im X),
// then Re Y = P1 - P2 and Im Y = (P3 - P1) - P2: 3 D x D real products
// instead of the 4 of the real embedding. Each product is (D/16) row blocks
// x (D/4) K-steps of 16x16x4 FP64 MFMA with FP64 accumulation.
//
// Operand layout of v_mfma_f64_16x16x4f64 (wave64):
//   A (16 x 4):  lane l supplies A[l % 16][l / 16]
//   B (4 x 16):  lane l supplies B[l / 16][l % 16]
//   D (16 x 16): register r of lane l holds D[4 r + l / 16][l % 16]
// (the f64 form interleaves the rows over the four lane groups, unlike the
// f32 16x16x4 form). Lane (j = l % 16, q = l / 16) therefore reads the D/4
// rows (D/4) q .. (D/4) q + D/4 - 1 of column j (K-step s uses row
// (D/4) q + s) and writes rows 16 mb + 4 r + q of the same column. The
// matrix fragments (A operands) sit in LDS in lane order, read once per MFMA
// without bank conflicts.
typedef double doublex4 __attribute__((ext_vector_type(4)));

// k = 4 or 5 qubits, with the state staged through LDS as in the complex64
// kernel: a workgroup owns a tile of 2048 amplitudes (the k target bits plus
// the lowest 11 - k other bits), so every global load and store is a
// coalesced slice of 256 consecutive tile elements whatever the targets are.
// (Column loads straight from HBM into the B operands were measured at half
// the bandwidth when the targets include the lowest index bits: 13.5 against
// 6.9 ms per block at 30 qubits.)
// Each wavefront multiplies groups of 16 columns read from the tile and
// writes its results back to the same slots; the tile is then streamed out.
// Persistent workgroups; the next tile's loads are in flight during the
// products of the current one.
constexpr int kT128Bits = 11;

struct Tile128Args {
  uint64_t ntiles;
  uint8_t tbits[kT128Bits];   // global bit of local tile bit i, ascending
  uint8_t rowbit[5];          // local tile bit of matrix row bit (K-1-t), t = target order
  uint8_t colbit[kT128Bits];  // local tile bits of the column index, ascending (11 - K of them)
};

// LDS slot of local index u (16-byte slots): bits 0-3 ^= bits 4-7 ^ bits 8-10.
// A bijection of [0, 2048). Sixteen lanes that read 16 columns whose index
// bits are four consecutive local bits land on 16 distinct 16-byte bank
// groups wherever those bits sit; the staging slices (consecutive elements)
// stay conflict-free.
__device__ __forceinline__ uint32_t swz128(uint32_t u) { return u ^ ((u >> 4) & 15u) ^ ((u >> 8) & 7u); }

template <int K>
__device__ __forceinline__ uint32_t tile128_col(const Tile128Args& a, uint32_t c) {
  uint32_t u = 0;
#pragma clang loop unroll(full)
  for (int i = 0; i < kT128Bits - K; ++i)
    if ((c >> i) & 1) u |= 1u << a.colbit[i];
  return u;
}

// Global index of staging element k * 256 + thread of the tile whose local
// bits 0-7 (the thread) are already in glow.
__device__ __forceinline__ uint64_t tile128_addr(const Tile128Args& a, uint64_t glow, int k) {
  uint64_t g = glow;
#pragma clang loop unroll(full)
  for (int i = 0; i < kT128Bits - 8; ++i)
    if ((k >> i) & 1) g |= uint64_t(1) << a.tbits[8 + i];
  return g;
}

template <int K>
__global__ __launch_bounds__(256) void k_fused_tile_c128(const Tile128Args a, const c128* __restrict__ mat,
                                                        c128* __restrict__ psi) {
  constexpr int D = 1 << K;          // complex rows
  constexpr int RPL = D / 4;         // rows per lane = K-steps per product
  constexpr int MB = D / 16;         // 16-row output blocks per product
  constexpr int NFRAG = 3 * MB * RPL;
  constexpr int NCG = (1 << (kT128Bits - K)) / 16;  // 16-column groups per tile: 4 (K = 5) or 8 (K = 4)
  constexpr int PER = (1 << kT128Bits) / 256;       // staging elements per thread
  __shared__ c128 tile[1 << kT128Bits];
  __shared__ double frag[NFRAG * 64];
  __shared__ uint32_t rowoff[D];
  for (int e = threadIdx.x; e < NFRAG * 64; e += 256) {
    const int l = e & 63, f = e >> 6;
    const int st = f % RPL, mb = (f / RPL) % MB, p = f / (RPL * MB);
    const c128 m = mat[(16 * mb + (l & 15)) * D + RPL * (l >> 4) + st];
    frag[e] = p == 0 ? m.re : (p == 1 ? m.im : m.re + m.im);
  }
  if (threadIdx.x < D) {
    uint32_t u = 0;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t)
      if ((threadIdx.x >> (K - 1 - t)) & 1) u |= 1u << a.rowbit[t];
    rowoff[threadIdx.x] = u;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 15, q = lane >> 4;
  uint64_t lanebits = 0;  // local bits 0-7 of this thread's staging elements, as global bits
#pragma clang loop unroll(full)
  for (int i = 0; i < 8; ++i)
    if ((threadIdx.x >> i) & 1) lanebits |= uint64_t(1) << a.tbits[i];
  uint64_t t = blockIdx.x;
  c128 nxt[PER];
  if (t < a.ntiles) {
    const uint64_t glow = insert_zeros(t, a.tbits, kT128Bits) | lanebits;
#pragma clang loop unroll(full)
    for (int k = 0; k < PER; ++k) nxt[k] = ld_nt(psi, tile128_addr(a, glow, k));
  }
  while (t < a.ntiles) {
    const uint64_t glow = insert_zeros(t, a.tbits, kT128Bits) | lanebits;
    // each thread writes (and later reads back) only its own staging slots
#pragma clang loop unroll(full)
    for (int k = 0; k < PER; ++k) tile[swz128(uint32_t(k) * 256u + threadIdx.x)] = nxt[k];
    __syncthreads();  // (the first pass through also publishes frag and rowoff)
    const uint64_t tn = t + gridDim.x;
    if (tn < a.ntiles) {  // the next tile's loads overlap this tile's products
      const uint64_t gnext = insert_zeros(tn, a.tbits, kT128Bits) | lanebits;
#pragma clang loop unroll(full)
      for (int k = 0; k < PER; ++k) nxt[k] = ld_nt(psi, tile128_addr(a, gnext, k));
    }
    // one 16-column group per wave at a time; the groups are disjoint, so each
    // wave reads and writes only its own columns of the tile
    for (int cg = wave; cg < NCG; cg += 4) {
      const uint32_t colpart = tile128_col<K>(a, uint32_t(16 * cg + j));
      doublex4 acc[3][MB];
#pragma clang loop unroll(full)
      for (int p = 0; p < 3; ++p)
#pragma clang loop unroll(full)
        for (int mb = 0; mb < MB; ++mb) acc[p][mb] = doublex4(0.0);
#pragma clang loop unroll(full)
      for (int st = 0; st < RPL; ++st) {
        const c128 x = tile[swz128(rowoff[RPL * q + st] | colpart)];
        const double b0 = x.re, b1 = x.im, b2 = x.re + x.im;
#pragma clang loop unroll(full)
        for (int mb = 0; mb < MB; ++mb) {
          acc[0][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[((0 * MB + mb) * RPL + st) * 64 + lane], b0,
                                                            acc[0][mb], 0, 0, 0);
          acc[1][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[((1 * MB + mb) * RPL + st) * 64 + lane], b1,
                                                            acc[1][mb], 0, 0, 0);
          acc[2][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[((2 * MB + mb) * RPL + st) * 64 + lane], b2,
                                                            acc[2][mb], 0, 0, 0);
        }
      }
      // rows 16 mb + 4 r + q of column j; this wave's reads of these columns
      // fed the accumulators above
#pragma clang loop unroll(full)
      for (int mb = 0; mb < MB; ++mb)
#pragma clang loop unroll(full)
        for (int r = 0; r < 4; ++r) {
          c128 y;
          y.re = acc[0][mb][r] - acc[1][mb][r];
          y.im = (acc[2][mb][r] - acc[0][mb][r]) - acc[1][mb][r];
          tile[swz128(rowoff[16 * mb + 4 * r + q] | colpart)] = y;
        }
    }
    __syncthreads();
#pragma clang loop unroll(full)
    for (int k = 0; k < PER; ++k) st_nt(psi, tile128_addr(a, glow, k), tile[swz128(uint32_t(k) * 256u + threadIdx.x)]);
    t = tn;
  }
}

template <int K>
static int launch_fused_tile_c128(void* state, int n, const int* targets, const void* mat, hipStream_t s) {
  Tile128Args a;
  std::memset(&a, 0, sizeof(a));
  a.ntiles = uint64_t(1) << (n - kT128Bits);
  // tile bits: the K target bits plus the lowest non-target bits
  bool is_target[QSIM_MAX_QUBITS] = {false};
  for (int t = 0; t < K; ++t) is_target[n - 1 - targets[t]] = true;
  uint8_t bits[kT128Bits];
  int nb = 0, ncols = 0;
  for (int b = 0; b < n && nb < kT128Bits; ++b) {
    if (is_target[b]) {
      bits[nb++] = uint8_t(b);
    } else if (ncols < kT128Bits - K) {
      bits[nb++] = uint8_t(b);
      ++ncols;
    }
  }
  if (nb != kT128Bits) {
    set_error("qsim_apply_fused: tile bits do not fit the register");
    return QSIM_EINVAL;
  }
  int nc = 0;
  for (int i = 0; i < kT128Bits; ++i) {
    a.tbits[i] = bits[i];
    if (!is_target[bits[i]]) a.colbit[nc++] = uint8_t(i);
  }
  for (int t = 0; t < K; ++t) {
    const int gb = n - 1 - targets[t];
    for (int i = 0; i < kT128Bits; ++i)
      if (bits[i] == gb) a.rowbit[t] = uint8_t(i);
  }
  // persistent grid: as many workgroups as the device keeps resident
  static const unsigned resident = [] {
    int nbk = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbk, k_fused_tile_c128<K>, 256, 0) != hipSuccess || nbk < 1)
      nbk = 2;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    return unsigned(nbk) * unsigned(cus);
  }();
  const unsigned grid = unsigned(a.ntiles < resident ? a.ntiles : resident);
  hipLaunchKernelGGL((k_fused_tile_c128<K>), dim3(grid), dim3(256), 0, s, a, reinterpret_cast<const c128*>(mat),
                     reinterpret_cast<c128*>(state));
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
  if (dtype == QSIM_C128) {
    if (nqubits >= kT128Bits && ntargets == 5) return launch_fused_tile_c128<5>(state, nqubits, targets, matrix_dev, s);
    if (nqubits >= kT128Bits && ntargets == 4) return launch_fused_tile_c128<4>(state, nqubits, targets, matrix_dev, s);
    return dispatch_fused<double>(state, nqubits, targets, ntargets, matrix_dev, s);
  }
  if (nqubits >= kMfmaTileBits && ntargets == 5) return launch_fused_mfma<5>(state, nqubits, targets, matrix_dev, s);
  if (nqubits >= kMfmaTileBits && ntargets == 4) return launch_fused_mfma<4>(state, nqubits, targets, matrix_dev, s);
  return dispatch_fused<float>(state, nqubits, targets, ntargets, matrix_dev, s);
}
