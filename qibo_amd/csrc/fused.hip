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

// LDS slot of local index u: bits 0-4 ^= bits 5-9. Column groups whose 32
// columns are 32 apart (target bits at the bottom of the tile) would
// otherwise put all 32 lanes on one bank; the slices of the streaming phases
// stay conflict-free (a permutation inside each 32-slot row).
__device__ __forceinline__ uint32_t mfma_slot(uint32_t u) { return u ^ ((u >> 5) & 31u); }

// Global address of element k * 256 + thread of tile t (coalesced slices).
__device__ __forceinline__ uint64_t mfma_addr(const MfmaArgs& a, uint64_t glow, int k) {
  uint64_t g = glow;
#pragma clang loop unroll(full)
  for (int i = 0; i < 4; ++i)
    if ((k >> i) & 1) g |= uint64_t(1) << a.tbits[8 + i];
  return g;
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
  // the real embedding, column-major (element (i, k) at k * 2D + i): the A
  // operand of step (mb, ks) is two consecutive 32-float runs
  __shared__ float amat[4 * D * D];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x < D) rowoff[threadIdx.x] = mfma_local<K>(a, threadIdx.x, 0);
  for (int e = threadIdx.x; e < 4 * D * D; e += 256) {
    const int k = e / (2 * D), i = e % (2 * D);
    const c64 m = mat[(i % D) * D + (k % D)];
    amat[e] = (i < D) ? ((k < D) ? m.re : -m.im) : ((k < D) ? m.im : m.re);
  }
  // lane part (local bits 0-7) of this thread's slice element, as global bits
  uint64_t lanebits = 0;
#pragma clang loop unroll(full)
  for (int i = 0; i < 8; ++i)
    if ((threadIdx.x >> i) & 1) lanebits |= uint64_t(1) << a.tbits[i];
  uint64_t t = blockIdx.x;
  c64 nxt[16];
  if (t < a.ntiles) {
    const uint64_t glow = insert_zeros(t, a.tbits, kMfmaTileBits) | lanebits;
#pragma clang loop unroll(full)
    for (int k = 0; k < 16; ++k) nxt[k] = ld_nt(psi, mfma_addr(a, glow, k));
  }
  while (t < a.ntiles) {
    const uint64_t glow = insert_zeros(t, a.tbits, kMfmaTileBits) | lanebits;
    // each thread writes (and later reads back) only its own slots here
#pragma clang loop unroll(full)
    for (int k = 0; k < 16; ++k) tile[mfma_slot(k * 256 + threadIdx.x)] = nxt[k];
    __syncthreads();
    const uint64_t tn = t + gridDim.x;
    if (tn < a.ntiles) {  // next tile's loads overlap this tile's product
      const uint64_t gnext = insert_zeros(tn, a.tbits, kMfmaTileBits) | lanebits;
#pragma clang loop unroll(full)
      for (int k = 0; k < 16; ++k) nxt[k] = ld_nt(psi, mfma_addr(a, gnext, k));
    }
    // one 32-column group per wave at a time; the groups are disjoint, so each
    // wave reads and writes only its own columns of the tile
    for (int cg = wave; cg < NCG; cg += 4) {
      const uint32_t colpart = mfma_local<K>(a, 0, 32 * cg + (lane & 31));
      // the A operands are read from LDS at each use (an opaque lane index
      // keeps the compiler from hoisting all 64 of them into registers)
      int alane = lane;
      asm volatile("" : "+v"(alane));
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
    for (int k = 0; k < 16; ++k) st_nt(psi, mfma_addr(a, glow, k), tile[mfma_slot(k * 256 + threadIdx.x)]);
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
// complex128 blocks of k = 4 or 5 qubits on FP64 MFMA (v_mfma_f64_16x16x4f64)
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
// f32 16x16x4 form). Lane (j = l % 16, q = l / 16) therefore loads the D/4
// rows (D/4) q .. (D/4) q + D/4 - 1 of column j (K-step s uses row
// (D/4) q + s) straight from HBM into registers, and stores rows
// 16 mb + 4 r + q of the same column: no LDS staging of the state. The matrix fragments (A
// operands) sit in LDS in lane order, read once per MFMA without bank
// conflicts. A column group belongs to one wavefront, so the in-place update
// needs no synchronisation: its loads complete before its stores.
typedef double doublex4 __attribute__((ext_vector_type(4)));

struct Mfma64Args {
  uint64_t ngroups;        // groups of 16 columns: 2^(n-K) / 16
  uint64_t off[32];        // index bits of matrix row r (row bit K-1-t <-> target t)
  int nfixed;
  uint8_t pos[kMaxFixed];  // target bit positions, ascending (for insert_zeros)
};

#ifndef QSIM_FUSED_WAVES
#define QSIM_FUSED_WAVES 4
#endif
#ifndef QSIM_FUSED_PREFETCH
#define QSIM_FUSED_PREFETCH 0
#endif

template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QSIM_FUSED_WAVES)))
void k_fused_mfma_c128(const Mfma64Args a, const c128* __restrict__ mat, c128* __restrict__ psi) {
  constexpr int D = 1 << K;          // complex rows
  constexpr int RPL = D / 4;         // rows per lane = K-steps per product
  constexpr int MB = D / 16;         // 16-row output blocks per product
  constexpr int NFRAG = 3 * MB * RPL;
  __shared__ double frag[NFRAG * 64];
  // fragment (p, mb, s), lane l: M_p[16 mb + l % 16][RPL (l / 16) + s] with
  // M_0 = Re M, M_1 = Im M, M_2 = Re M + Im M
  for (int e = threadIdx.x; e < NFRAG * 64; e += 256) {
    const int l = e & 63, f = e >> 6;
    const int st = f % RPL, mb = (f / RPL) % MB, p = f / (RPL * MB);
    const c128 m = mat[(16 * mb + (l & 15)) * D + RPL * (l >> 4) + st];
    frag[e] = p == 0 ? m.re : (p == 1 ? m.im : m.re + m.im);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int j = lane & 15, q = lane >> 4;
  // row offsets are OR-linear in the row bits: off[RPL q + s] = off[RPL q] | off[s]
  const uint64_t qin = a.off[RPL * q];
  const uint64_t qout = a.off[q];
  // XCD-contiguous work split: workgroup b runs on XCD b % 8; each XCD walks one
  // contiguous eighth of the column groups, its wavefronts interleaved
  const uint32_t nxcd = gridDim.x >= 8 ? 8u : 1u;
  const uint32_t xcd = blockIdx.x % nxcd;
  const uint64_t wpx = uint64_t(gridDim.x / nxcd) * 4;  // wavefronts per XCD
  const uint64_t w = uint64_t(blockIdx.x / nxcd) * 4 + (threadIdx.x >> 6);
  const uint64_t chunk = (a.ngroups + nxcd - 1) / nxcd;
  const uint64_t gend0 = (uint64_t(xcd) + 1) * chunk;
  const uint64_t gend = gend0 < a.ngroups ? gend0 : a.ngroups;
  uint64_t g = uint64_t(xcd) * chunk + w;
  c128 x[RPL];
  if (QSIM_FUSED_PREFETCH && g < gend) {
    const uint64_t base = insert_zeros(16 * g + j, a.pos, a.nfixed) | qin;
#pragma clang loop unroll(full)
    for (int s = 0; s < RPL; ++s) x[s] = ld_nt(psi, base | a.off[s]);
  }
  while (g < gend) {
    const uint64_t gn = g + wpx;
    c128 xn[RPL];
    if (QSIM_FUSED_PREFETCH) {
      if (gn < gend) {  // the next group's loads overlap this group's products
        const uint64_t bn = insert_zeros(16 * gn + j, a.pos, a.nfixed) | qin;
#pragma clang loop unroll(full)
        for (int s = 0; s < RPL; ++s) xn[s] = ld_nt(psi, bn | a.off[s]);
      }
    } else {
      // no register double buffer: the other resident waves cover the latency
      // (addresses formed next to each load, not hoisted and held)
      const c128* pin = psi + (insert_zeros(16 * g + j, a.pos, a.nfixed) | qin);
#pragma clang loop unroll(full)
      for (int s = 0; s < RPL; ++s) {
        const c128* q = pin;
        asm volatile("" : "+v"(q));
        x[s] = ld_nt(q, a.off[s]);
      }
    }
    doublex4 acc[3][MB];
#pragma clang loop unroll(full)
    for (int p = 0; p < 3; ++p)
#pragma clang loop unroll(full)
      for (int mb = 0; mb < MB; ++mb) acc[p][mb] = doublex4(0.0);
#pragma clang loop unroll(full)
    for (int s = 0; s < RPL; ++s) {
      const double b0 = x[s].re, b1 = x[s].im, b2 = x[s].re + x[s].im;
#pragma clang loop unroll(full)
      for (int mb = 0; mb < MB; ++mb) {
        acc[0][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[((0 * MB + mb) * RPL + s) * 64 + lane], b0,
                                                          acc[0][mb], 0, 0, 0);
        acc[1][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[((1 * MB + mb) * RPL + s) * 64 + lane], b1,
                                                          acc[1][mb], 0, 0, 0);
        acc[2][mb] = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[((2 * MB + mb) * RPL + s) * 64 + lane], b2,
                                                          acc[2][mb], 0, 0, 0);
      }
    }
    c128* pout = psi + (insert_zeros(16 * g + j, a.pos, a.nfixed) | qout);
#pragma clang loop unroll(full)
    for (int mb = 0; mb < MB; ++mb)
#pragma clang loop unroll(full)
      for (int r = 0; r < 4; ++r) {
        c128 y;
        y.re = acc[0][mb][r] - acc[1][mb][r];
        y.im = (acc[2][mb][r] - acc[0][mb][r]) - acc[1][mb][r];
        c128* q = pout;
        asm volatile("" : "+v"(q));
        st_nt(q, a.off[16 * mb + 4 * r], y);
      }
    if (QSIM_FUSED_PREFETCH) {
#pragma clang loop unroll(full)
      for (int s = 0; s < RPL; ++s) x[s] = xn[s];
    }
    g = gn;
  }
}

// ---------------------------------------------------------------------------
// complex64 blocks of 5 qubits straight from HBM (v_mfma_f32_32x32x2f32)
// ---------------------------------------------------------------------------
// Y = M X for 32 columns at a time (one wavefront) as the three real 32 x 32
// products of the complex128 kernel above (Gauss), FP32 operands and FP32
// accumulation: P1 = Re M Re X, P2 = Im M Im X, P3 = (Re M + Im M)(Re X +
// Im X); Re Y = P1 - P2, Im Y = (P3 - P1) - P2. Each product is 16 K-steps of
// the 32x32x2 MFMA.
//
// Operand layout of v_mfma_f32_32x32x2f32 (wave64):
//   A (32 x 2):  lane l supplies A[l % 32][l / 32]
//   B (2 x 32):  lane l supplies B[l / 32][l % 32]
//   D (32 x 32): register j of lane l holds D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32]
// Lane (c = l % 32, h = l / 32) therefore loads rows h, h + 2, ..., h + 30 of
// column c (K-step s uses row 2 s + h) into registers and stores rows
// 8 (j / 4) + 4 h + j % 4 of the same column. No LDS holds the state: a
// column group belongs to one wavefront, whose loads complete before its
// stores. The matrix fragments sit in LDS in lane order.
template <int K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QSIM_FUSED_WAVES)))
void k_fused_mfma32_c64(const Mfma64Args a, const c64* __restrict__ mat, c64* __restrict__ psi) {
  static_assert(K == 5, "32x32 blocks");
  constexpr int D = 1 << K;
  constexpr int KS = D / 2;  // K-steps per product
  __shared__ float frag[3 * KS * 64];
  // fragment (p, s), lane l: M_p[l % 32][2 s + l / 32], M_0 = Re M, M_1 = Im M, M_2 = Re M + Im M
  for (int e = threadIdx.x; e < 3 * KS * 64; e += 256) {
    const int l = e & 63, f = e >> 6;
    const int st = f % KS, p = f / KS;
    const c64 m = mat[(l & 31) * D + 2 * st + (l >> 5)];
    frag[e] = p == 0 ? m.re : (p == 1 ? m.im : m.re + m.im);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int c = lane & 31, h = lane >> 5;
  const uint64_t hin = a.off[h];       // row h
  const uint64_t hout = a.off[4 * h];  // row 4 h
  const uint32_t nxcd = gridDim.x >= 8 ? 8u : 1u;
  const uint32_t xcd = blockIdx.x % nxcd;
  const uint64_t wpx = uint64_t(gridDim.x / nxcd) * 4;
  const uint64_t w = uint64_t(blockIdx.x / nxcd) * 4 + (threadIdx.x >> 6);
  const uint64_t chunk = (a.ngroups + nxcd - 1) / nxcd;
  const uint64_t gend0 = (uint64_t(xcd) + 1) * chunk;
  const uint64_t gend = gend0 < a.ngroups ? gend0 : a.ngroups;
  for (uint64_t g = uint64_t(xcd) * chunk + w; g < gend; g += wpx) {
    const uint64_t col = insert_zeros(32 * g + c, a.pos, a.nfixed);
    // lane pointers; the row offsets are uniform (index bits disjoint from the
    // column's, so OR = ADD), each address formed next to its access
    const c64* pin = psi + (col | hin);
    c64 x[KS];
#pragma clang loop unroll(full)
    for (int st = 0; st < KS; ++st) {
      const c64* q = pin;
      asm volatile("" : "+v"(q));  // keep the 16 addresses from being hoisted and held
      x[st] = ld_nt(q, a.off[2 * st]);
    }
    floatx16 acc0 = floatx16(0.0f), acc1 = floatx16(0.0f), acc2 = floatx16(0.0f);
#pragma clang loop unroll(full)
    for (int st = 0; st < KS; ++st) {
      const float b0 = x[st].re, b1 = x[st].im, b2 = x[st].re + x[st].im;
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(frag[(0 * KS + st) * 64 + lane], b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(frag[(1 * KS + st) * 64 + lane], b1, acc1, 0, 0, 0);
      acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(frag[(2 * KS + st) * 64 + lane], b2, acc2, 0, 0, 0);
    }
    c64* pout = psi + (col | hout);
#pragma clang loop unroll(full)
    for (int j = 0; j < 16; ++j) {
      c64 y;
      y.re = acc0[j] - acc1[j];
      y.im = (acc2[j] - acc0[j]) - acc1[j];
      c64* q = pout;
      asm volatile("" : "+v"(q));
      st_nt(q, a.off[8 * (j / 4) + (j % 4)], y);
    }
  }
}

static int launch_fused_mfma32_c64(void* state, int n, const int* targets, const void* mat, hipStream_t s) {
  constexpr int K = 5;
  Mfma64Args a;
  std::memset(&a, 0, sizeof(a));
  a.nfixed = sorted_positions(n, targets, K, nullptr, 0, a.pos);
  a.ngroups = (uint64_t(1) << (n - K)) / 32;
  for (int r = 0; r < (1 << K); ++r) {
    uint64_t off = 0;
    for (int t = 0; t < K; ++t)
      if ((r >> (K - 1 - t)) & 1) off |= uint64_t(1) << (n - 1 - targets[t]);
    a.off[r] = off;
  }
  static const unsigned resident = [] {
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_mfma32_c64<K>, 256, 0) != hipSuccess || nb < 1)
      nb = 2;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    return unsigned(nb) * unsigned(cus);
  }();
  const uint64_t want = (a.ngroups + 3) / 4;
  unsigned grid = unsigned(want < resident ? want : resident);
  if (grid >= 8) grid &= ~7u;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((k_fused_mfma32_c64<K>), dim3(grid), dim3(256), 0, s, a, reinterpret_cast<const c64*>(mat),
                     reinterpret_cast<c64*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

template <int K>
static int launch_fused_mfma_c128(void* state, int n, const int* targets, const void* mat, hipStream_t s) {
  Mfma64Args a;
  std::memset(&a, 0, sizeof(a));
  a.nfixed = sorted_positions(n, targets, K, nullptr, 0, a.pos);
  a.ngroups = (uint64_t(1) << (n - K)) / 16;
  for (int r = 0; r < (1 << K); ++r) {
    uint64_t off = 0;
    for (int t = 0; t < K; ++t)
      if ((r >> (K - 1 - t)) & 1) off |= uint64_t(1) << (n - 1 - targets[t]);
    a.off[r] = off;
  }
  // persistent grid of the workgroups the device keeps resident (a multiple of 8)
  static const unsigned resident = [] {
    int nb = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused_mfma_c128<K>, 256, 0) != hipSuccess || nb < 1)
      nb = 2;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    return unsigned(nb) * unsigned(cus);
  }();
  const uint64_t want = (a.ngroups + 3) / 4;  // one group per wavefront at least
  unsigned grid = unsigned(want < resident ? want : resident);
  if (grid >= 8) grid &= ~7u;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((k_fused_mfma_c128<K>), dim3(grid), dim3(256), 0, s, a, reinterpret_cast<const c128*>(mat),
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
    if (nqubits >= 9 && ntargets == 5) return launch_fused_mfma_c128<5>(state, nqubits, targets, matrix_dev, s);
    if (nqubits >= 8 && ntargets == 4) return launch_fused_mfma_c128<4>(state, nqubits, targets, matrix_dev, s);
    return dispatch_fused<double>(state, nqubits, targets, ntargets, matrix_dev, s);
  }
  if (nqubits >= 10 && ntargets == 5) return launch_fused_mfma32_c64(state, nqubits, targets, matrix_dev, s);
  if (nqubits >= kMfmaTileBits && ntargets == 4) return launch_fused_mfma<4>(state, nqubits, targets, matrix_dev, s);
  return dispatch_fused<float>(state, nqubits, targets, ntargets, matrix_dev, s);
}
