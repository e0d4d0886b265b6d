// Many gates per pass over the state: register-blocked tiles for gfx950.
//
// A gate pass is HBM-bound (each amplitude is read and written once per
// gate), so a circuit's gates are grouped by the host into passes that each
// touch the state once. A pass owns 12 "tile bits" of the amplitude index
// (always including the lowest bits, so loads stay coalesced); one workgroup
// loads a tile of 4096 amplitudes whose other bits are fixed, applies the
// pass's whole instruction list to it in registers, and streams it back.
//
// Each of the 256 lanes holds 16 amplitudes. The 12 local tile bits are split
// into three groups of four, and a "mapping" decides which group lives in a
// lane's registers (the other 8 bits select the lane):
//   mapping 0: registers hold local bits 8-11 (the load/store layout),
//   mapping 1: local bits 4-7,   mapping 2: local bits 0-3.
// A one-qubit gate on a local bit of the current register group is applied
// entirely in registers; diagonal gates need no exchange at all (every lane
// knows the global index of each of its amplitudes); a MAP instruction moves
// to another mapping through LDS; two- and three-qubit gates run on the tile
// in LDS. The arithmetic per gate is exactly that of the one-gate kernels
// (same fused multiply-adds in the same order), so a blocked pass gives
// bitwise the same state as applying its gates one by one.
//
// Reference semantics: repeated Backend.apply_gate (abstract.py:2322-2361,
// 3176-3197) over the gates of Circuit.queue (abstract.py:3321-3322).

#include <hip/hip_runtime.h>

#include <cstring>

#include "qsim_common.h"

namespace qsim {

constexpr int kBT = 256;              // lanes per workgroup
constexpr int kBReg = 16;             // amplitudes per lane
constexpr int kBTileBits = 12;
constexpr int kBTile = 1 << kBTileBits;

enum : uint32_t { kOpDiag = 0, kOpDense1 = 1, kOpMap = 2, kOpDenseLds = 3 };

// Instruction record (host packs the same layout: 32-byte header + 64 complex).
template <typename R>
struct BlkInstr {
  uint32_t op;
  uint32_t arg;      // Dense1: register bit 0..3; Map: mapping 0..2; DenseLds: number of targets 1..3
  uint64_t cmask;    // control bits (global amplitude-index mask), all must be 1
  uint32_t nt;       // Diag: number of target bits (0..5)
  uint8_t bits[8];   // Diag: global target bits, t0 = MSB of the diagonal index
                     // DenseLds: [0..k) local target bits in matrix order, [4..4+k) the same ascending
  uint32_t pad;
  cplx<R> m[64];     // Diag: 2^nt entries; Dense1: 2x2; DenseLds: 2^k x 2^k (row-major)
};

struct BlkArgs {
  uint64_t ntiles;
  int ninstr;
  uint8_t tbits[kBTileBits];  // global bit of local tile bit i, ascending
};

__device__ __forceinline__ int reg_base(int map) { return map == 0 ? 8 : (map == 1 ? 4 : 0); }

// Local tile index of register r of `lane` under mapping `map`.
__device__ __forceinline__ uint32_t local_index(int map, uint32_t lane, uint32_t r) {
  const int r0 = reg_base(map);
  return (lane & ((1u << r0) - 1)) | (r << r0) | ((lane >> r0) << (r0 + 4));
}

// XOR swizzle of LDS slots: spreads the strided accesses of mappings 1 and 2
// over all banks. A bijection of [0, 4096).
__device__ __forceinline__ uint32_t swz(uint32_t u) { return u ^ ((u >> 4) & 15); }

// Global-index bits contributed by local index u.
__device__ __forceinline__ uint64_t deposit(uint32_t u, const uint64_t* gbit) {
  uint64_t g = 0;
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i)
    if ((u >> i) & 1) g |= gbit[i];
  return g;
}

template <typename R>
__device__ __forceinline__ void to_lds(cplx<R>* lds, const cplx<R>* v, int map) {
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) lds[swz(local_index(map, threadIdx.x, r))] = v[r];
}

template <typename R>
__device__ __forceinline__ void from_lds(const cplx<R>* lds, cplx<R>* v, int map) {
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) v[r] = lds[swz(local_index(map, threadIdx.x, r))];
}

// One-qubit gate on register bit J of the current mapping.
template <typename R, int J>
__device__ __forceinline__ void dense1_regs(cplx<R>* v, const BlkInstr<R>& in, uint64_t lane_idx,
                                            const uint64_t* rbit) {
  const cplx<R> m0 = in.m[0], m1 = in.m[1], m2 = in.m[2], m3 = in.m[3];
  const uint64_t cmask = in.cmask;
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    if ((r >> J) & 1) continue;
    const int r1 = r | (1 << J);
    uint64_t idx = lane_idx;
#pragma clang loop unroll(full)
    for (int j = 0; j < 4; ++j)
      if ((r >> j) & 1) idx |= rbit[j];
    if ((idx & cmask) != cmask) continue;
    const cplx<R> x[2] = {v[r], v[r1]};
    cplx<R> a = cmul(m0, x[0]);
    a = cfma(m1, x[1], a);
    cplx<R> b = cmul(m2, x[0]);
    b = cfma(m3, x[1], b);
    v[r] = a;
    v[r1] = b;
  }
}

template <typename R>
__device__ __forceinline__ void diag_regs(cplx<R>* v, const BlkInstr<R>& in, uint64_t lane_idx, const uint64_t* rbit) {
  const uint64_t cmask = in.cmask;
  const uint32_t nt = in.nt < 5 ? in.nt : 5;
  if (nt == 0) {
    const cplx<R> d = in.m[0];
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      uint64_t idx = lane_idx;
#pragma clang loop unroll(full)
      for (int j = 0; j < 4; ++j)
        if ((r >> j) & 1) idx |= rbit[j];
      if ((idx & cmask) == cmask) v[r] = cmul(d, v[r]);
    }
    return;
  }
  if (nt == 1) {
    const cplx<R> d0 = in.m[0], d1 = in.m[1];
    const int tb = in.bits[0] & 63;
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      uint64_t idx = lane_idx;
#pragma clang loop unroll(full)
      for (int j = 0; j < 4; ++j)
        if ((r >> j) & 1) idx |= rbit[j];
      if ((idx & cmask) == cmask) v[r] = cmul(((idx >> tb) & 1) ? d1 : d0, v[r]);
    }
    return;
  }
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    uint64_t idx = lane_idx;
#pragma clang loop unroll(full)
    for (int j = 0; j < 4; ++j)
      if ((r >> j) & 1) idx |= rbit[j];
    if ((idx & cmask) != cmask) continue;
    uint32_t di = 0;
    for (uint32_t t = 0; t < nt; ++t) di = (di << 1) | uint32_t((idx >> (in.bits[t] & 63)) & 1);
    v[r] = cmul(in.m[di], v[r]);
  }
}

// Dense 2^k block (k = 1..3) on local bits, on the tile held in LDS.
template <typename R, int K>
__device__ __forceinline__ void dense_lds(cplx<R>* lds, const BlkInstr<R>& in, uint64_t tile_base,
                                          const uint64_t* gbit) {
  constexpr int D = 1 << K;
  constexpr int kGroups = kBTile >> K;
  uint8_t srt[K];
  uint32_t off[D];
#pragma clang loop unroll(full)
  for (int t = 0; t < K; ++t) srt[t] = in.bits[4 + t] < kBTileBits ? in.bits[4 + t] : kBTileBits - 1;
#pragma clang loop unroll(full)
  for (int r = 0; r < D; ++r) {
    uint32_t o = 0;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) {
      const uint32_t b = in.bits[t] < kBTileBits ? in.bits[t] : kBTileBits - 1;
      if ((r >> (K - 1 - t)) & 1) o |= 1u << b;
    }
    off[r] = o;
  }
  const uint64_t cmask = in.cmask;
  for (int g = threadIdx.x; g < kGroups; g += kBT) {
    uint32_t base = g;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) {
      const uint32_t low = base & ((1u << srt[t]) - 1);
      base = low | ((base ^ low) << 1);
    }
    if (((tile_base | deposit(base, gbit)) & cmask) != cmask) continue;
    cplx<R> x[D];
#pragma clang loop unroll(full)
    for (int c = 0; c < D; ++c) x[c] = lds[swz(base | off[c])];
#pragma clang loop unroll(full)
    for (int r = 0; r < D; ++r) {
      cplx<R> acc = cmul(in.m[r * D], x[0]);
#pragma clang loop unroll(full)
      for (int c = 1; c < D; ++c) acc = cfma(in.m[r * D + c], x[c], acc);
      lds[swz(base | off[r])] = acc;
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kBT) void k_blocked(const BlkArgs a, const BlkInstr<R>* __restrict__ prog,
                                                 cplx<R>* __restrict__ psi) {
  __shared__ cplx<R> lds[kBTile];
  const uint64_t tile = xcd_block();
  if (tile >= a.ntiles) return;
  uint64_t gbit[kBTileBits];
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i) gbit[i] = uint64_t(1) << a.tbits[i];
  const uint64_t tile_base = insert_zeros(tile, a.tbits, kBTileBits);

  int map = 0;
  uint64_t lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);
  uint64_t rbit[4];
#pragma clang loop unroll(full)
  for (int j = 0; j < 4; ++j) rbit[j] = gbit[reg_base(0) + j];

  cplx<R> v[kBReg];
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    uint64_t idx = lane_idx;
#pragma clang loop unroll(full)
    for (int j = 0; j < 4; ++j)
      if ((r >> j) & 1) idx |= rbit[j];
    v[r] = ld_nt(psi, idx);
  }

  for (int k = 0; k < a.ninstr; ++k) {
    const BlkInstr<R>& in = prog[k];
    const uint32_t op = in.op;
    if (op == kOpDiag) {
      diag_regs(v, in, lane_idx, rbit);
    } else if (op == kOpDense1) {
      switch (in.arg & 3) {
        case 0: dense1_regs<R, 0>(v, in, lane_idx, rbit); break;
        case 1: dense1_regs<R, 1>(v, in, lane_idx, rbit); break;
        case 2: dense1_regs<R, 2>(v, in, lane_idx, rbit); break;
        default: dense1_regs<R, 3>(v, in, lane_idx, rbit); break;
      }
    } else if (op == kOpMap) {
      const int nmap = in.arg < 3 ? int(in.arg) : 0;
      to_lds(lds, v, map);
      __syncthreads();
      map = nmap;
      from_lds(lds, v, map);
      __syncthreads();
      lane_idx = tile_base | deposit(local_index(map, threadIdx.x, 0), gbit);
#pragma clang loop unroll(full)
      for (int j = 0; j < 4; ++j) rbit[j] = gbit[reg_base(map) + j];
    } else if (op == kOpDenseLds) {
      to_lds(lds, v, map);
      __syncthreads();
      const uint32_t kk = in.arg;
      if (kk <= 1)
        dense_lds<R, 1>(lds, in, tile_base, gbit);
      else if (kk == 2)
        dense_lds<R, 2>(lds, in, tile_base, gbit);
      else
        dense_lds<R, 3>(lds, in, tile_base, gbit);
      __syncthreads();
      from_lds(lds, v, map);
      __syncthreads();
    }
  }

  if (map != 0) {
    to_lds(lds, v, map);
    __syncthreads();
    map = 0;
    from_lds(lds, v, map);
    lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);
#pragma clang loop unroll(full)
    for (int j = 0; j < 4; ++j) rbit[j] = gbit[reg_base(0) + j];
  }
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    uint64_t idx = lane_idx;
#pragma clang loop unroll(full)
    for (int j = 0; j < 4; ++j)
      if ((r >> j) & 1) idx |= rbit[j];
    st_nt(psi, idx, v[r]);
  }
}

template <typename R>
static int launch_blocked(void* state, int n, const int* tbits, const void* prog, int ninstr, hipStream_t s) {
  BlkArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntiles = uint64_t(1) << (n - kBTileBits);
  a.ninstr = ninstr;
  for (int i = 0; i < kBTileBits; ++i) a.tbits[i] = uint8_t(tbits[i]);
  hipLaunchKernelGGL((k_blocked<R>), dim3(unsigned(a.ntiles)), dim3(kBT), 0, s, a,
                     reinterpret_cast<const BlkInstr<R>*>(prog), reinterpret_cast<cplx<R>*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

extern "C" {

size_t qsim_blocked_record_bytes(int dtype) {
  if (dtype == QSIM_C128) return sizeof(BlkInstr<double>);
  if (dtype == QSIM_C64) return sizeof(BlkInstr<float>);
  return 0;
}

int qsim_apply_blocked(void* state, int nqubits, int dtype, const int* tile_bits, const void* program_dev,
                       int ninstr, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && tile_bits != nullptr, "qsim_apply_blocked: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_blocked: bad dtype");
  QSIM_CHECK_ARG(nqubits >= kBTileBits && nqubits <= QSIM_MAX_QUBITS, "qsim_apply_blocked: nqubits out of range");
  QSIM_CHECK_ARG(ninstr >= 0 && (ninstr == 0 || program_dev != nullptr), "qsim_apply_blocked: bad program");
  for (int i = 0; i < kBTileBits; ++i) {
    QSIM_CHECK_ARG(tile_bits[i] >= 0 && tile_bits[i] < nqubits, "qsim_apply_blocked: tile bit out of range");
    QSIM_CHECK_ARG(i == 0 || tile_bits[i] > tile_bits[i - 1], "qsim_apply_blocked: tile bits must ascend");
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128) return launch_blocked<double>(state, nqubits, tile_bits, program_dev, ninstr, s);
  return launch_blocked<float>(state, nqubits, tile_bits, program_dev, ninstr, s);
}

}  // extern "C"
