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
// Bits of a lane's amplitudes that are not register bits (the lane bits and
// the bits outside the tile) are fixed per lane.
//
// Instructions (variable-length records packed by the host: a 32-byte header
// then the record's complex entries, padded to 16 bytes; a pass's records are
// a few KB, so they stay in the scalar cache while every tile walks them):
//   DIAG (0)        diagonal on up to 5 bits, one complex multiply per amplitude;
//                   a target byte REG_FLAG | j names register bit j
//   DENSE1 (1)      complex 2x2 on register bit `arg`
//   MAP (2)         change to mapping `arg` (transpose through LDS)
//   DENSE_LDS (3)   2^k x 2^k (k = arg <= 3) on local bits, tile in LDS
//   DENSE1_REAL (4) 2x2 with zero imaginary parts on register bit `arg`
//   REG_PHASE (5)   v[r] *= m[r]: diagonal gates on register bits, folded by the host
//   LANE_PHASE (6)  diagonal on per-lane-fixed bits: multiplied into the lane's
//                   pending factor P, applied to the 16 amplitudes before the
//                   next MAP / DENSE_LDS / store (it commutes with everything
//                   that stays inside the lane's registers)
//   NEG (7)         v = -v (the -1 phase of CZ-like gates)
// Conditions: `cmask` = condition bits fixed per lane (DENSE_LDS: all condition
// bits); `rreq` = register-bit conditions as a 4-bit mask over register
// positions. DIAG / DENSE1 / DENSE_LDS / MAP perform exactly the fused
// multiply-adds of the one-gate kernels, so a pass built only from them is
// bitwise equal to applying its gates one by one; the other forms change the
// order of roundings.
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
constexpr uint32_t kRegFlag = 0x80;

enum : uint32_t {
  kOpDiag = 0,
  kOpDense1 = 1,
  kOpMap = 2,
  kOpDenseLds = 3,
  kOpDense1Real = 4,
  kOpRegPhase = 5,
  kOpLanePhase = 6,
  kOpNeg = 7
};

// Record header (the host packs the same layout).
struct BlkHdr {
  uint8_t op;
  uint8_t arg;       // DENSE1*: register bit 0..3; MAP: mapping 0..2; DENSE_LDS: number of targets 1..3
  uint8_t nt;        // DIAG / LANE_PHASE: number of target bits (0..5)
  uint8_t rreq;      // register-bit conditions (bit j: register bit j must be 1)
  uint32_t nent;     // complex entries following the header
  uint64_t cmask;    // condition bits fixed per lane (DENSE_LDS: all condition bits), all must be 1
  uint8_t bits[8];   // DIAG / LANE_PHASE: target bits, t0 = MSB of the diagonal index
                     // DENSE_LDS: [0..k) local target bits in matrix order, [4..4+k) the same ascending
  uint64_t reserved;
};
static_assert(sizeof(BlkHdr) == 32, "record header is 32 bytes");
// Entries: DIAG / LANE_PHASE 2^nt; DENSE1*: 2x2 row-major; REG_PHASE: 16;
// DENSE_LDS: 2^k x 2^k row-major; MAP / NEG: none.

struct BlkArgs {
  uint64_t ntiles;
  uint64_t nbytes;   // program size: records past it are never read
  int ninstr;
  uint8_t tbits[kBTileBits];  // global bit of local tile bit i, ascending
};

// Complex entries a record of this header must carry.
__device__ __forceinline__ uint32_t needed_entries(const BlkHdr& h) {
  switch (h.op) {
    case kOpDiag:
    case kOpLanePhase: return 1u << (h.nt < 5 ? h.nt : 5);
    case kOpDense1:
    case kOpDense1Real: return 4;
    case kOpRegPhase: return 16;
    case kOpDenseLds: return h.arg <= 1 ? 4u : (h.arg == 2 ? 16u : 64u);
    default: return 0;
  }
}

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

// Register r satisfies the register-bit conditions `rreq` (r is a constant
// after unrolling, rreq is uniform: a scalar test).
__device__ __forceinline__ bool reg_ok(int r, uint32_t rreq) { return (uint32_t(r) & rreq) == rreq; }

// One-qubit gate on register bit J of the current mapping: the arithmetic of
// k_dense (cmul then cfma) for DENSE1; two real multiplies per component for
// DENSE1_REAL.
template <typename R, bool REAL>
__device__ __forceinline__ void dense1_pair(cplx<R>& x0r, cplx<R>& x1r, cplx<R> m0, cplx<R> m1, cplx<R> m2,
                                            cplx<R> m3) {
  const cplx<R> x0 = x0r, x1 = x1r;
  if constexpr (REAL) {
    x0r.re = fmaR(m1.re, x1.re, m0.re * x0.re);
    x0r.im = fmaR(m1.re, x1.im, m0.re * x0.im);
    x1r.re = fmaR(m3.re, x1.re, m2.re * x0.re);
    x1r.im = fmaR(m3.re, x1.im, m2.re * x0.im);
  } else {
    x0r = cfma(m1, x1, cmul(m0, x0));
    x1r = cfma(m3, x1, cmul(m2, x0));
  }
}

template <typename R, int J, bool REAL>
__device__ __forceinline__ void dense1_regs(cplx<R>* v, const BlkHdr& h, const cplx<R>* m, bool lane_ok) {
  const cplx<R> m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
  const uint32_t rreq = h.rreq;
  if (!lane_ok) return;
  if (rreq == 0) {  // no register-bit conditions: straight-line code
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r)
      if (!((r >> J) & 1)) dense1_pair<R, REAL>(v[r], v[r | (1 << J)], m0, m1, m2, m3);
    return;
  }
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    if ((r >> J) & 1) continue;
    if (!reg_ok(r, rreq)) continue;
    dense1_pair<R, REAL>(v[r], v[r | (1 << J)], m0, m1, m2, m3);
  }
}

// Diagonal index contributions of one DIAG / LANE_PHASE record: `dl` from the
// lane's fixed bits, rm[j] from register bit j (uniform).
__device__ __forceinline__ uint32_t diag_lane_part(const BlkHdr& h, uint64_t lane_idx, uint32_t* rm) {
  const uint32_t nt = h.nt < 5 ? h.nt : 5;
  uint32_t dl = 0;
  rm[0] = rm[1] = rm[2] = rm[3] = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t b = h.bits[t];
    const uint32_t w = 1u << (nt - 1 - t);
    if (b & kRegFlag) {
      rm[0] |= (b & 3) == 0 ? w : 0u;
      rm[1] |= (b & 3) == 1 ? w : 0u;
      rm[2] |= (b & 3) == 2 ? w : 0u;
      rm[3] |= (b & 3) == 3 ? w : 0u;
    } else if ((lane_idx >> (b & 63)) & 1) {
      dl |= w;
    }
  }
  return dl;
}

template <typename R>
__device__ __forceinline__ void diag_regs(cplx<R>* v, const BlkHdr& h, const cplx<R>* m, uint64_t lane_idx,
                                          bool lane_ok) {
  uint32_t rm[4];
  const uint32_t dl = diag_lane_part(h, lane_idx, rm);
  const uint32_t rreq = h.rreq;
  if (!lane_ok) return;
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    if (!reg_ok(r, rreq)) continue;
    const uint32_t di = dl | ((r & 1) ? rm[0] : 0u) | ((r & 2) ? rm[1] : 0u) | ((r & 4) ? rm[2] : 0u) |
                        ((r & 8) ? rm[3] : 0u);
    v[r] = cmul(m[di], v[r]);
  }
}

template <typename R>
__device__ __forceinline__ void flush_lane_phase(cplx<R>* v, cplx<R>& P, bool& pdirty) {
  if (!pdirty) return;
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) v[r] = cmul(P, v[r]);
  P.re = R(1);
  P.im = R(0);
  pdirty = false;
}

// Dense 2^k block (k = 1..3) on local bits, on the tile held in LDS.
template <typename R, int K>
__device__ __forceinline__ void dense_lds(cplx<R>* lds, const BlkHdr& in, const cplx<R>* m, uint64_t tile_base,
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
      cplx<R> acc = cmul(m[r * D], x[0]);
#pragma clang loop unroll(full)
      for (int c = 1; c < D; ++c) acc = cfma(m[r * D + c], x[c], acc);
      lds[swz(base | off[r])] = acc;
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kBT) void k_blocked(const BlkArgs a, const uint8_t* __restrict__ prog,
                                                 cplx<R>* __restrict__ psi) {
  __shared__ cplx<R> lds[kBTile];
  const uint64_t tile = xcd_block();
  if (tile >= a.ntiles) return;
  uint64_t gbit[kBTileBits];
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i) gbit[i] = uint64_t(1) << a.tbits[i];
  const uint64_t tile_base = insert_zeros(tile, a.tbits, kBTileBits);

  int map = 0;
  // global index of the lane's amplitudes with the register bits cleared
  uint64_t lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);

  cplx<R> v[kBReg];
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) v[r] = ld_nt(psi, lane_idx | deposit(uint32_t(r) << 8, gbit));

  cplx<R> P;  // pending per-lane factor of LANE_PHASE records
  P.re = R(1);
  P.im = R(0);
  bool pdirty = false;

  uint64_t off = 0;
  for (int k = 0; k < a.ninstr; ++k) {
    if (off + sizeof(BlkHdr) > a.nbytes) break;
    const BlkHdr& h = *reinterpret_cast<const BlkHdr*>(prog + off);
    const cplx<R>* m = reinterpret_cast<const cplx<R>*>(prog + off + sizeof(BlkHdr));
    const uint32_t nent = h.nent;
    off += sizeof(BlkHdr) + ((uint64_t(nent) * sizeof(cplx<R>) + 15) & ~uint64_t(15));
    if (off > a.nbytes || nent < needed_entries(h)) break;  // malformed program: stop reading it
    const bool lane_ok = (lane_idx & h.cmask) == h.cmask;
    switch (h.op) {
      case kOpDense1Real:
        switch (h.arg & 3) {
          case 0: dense1_regs<R, 0, true>(v, h, m, lane_ok); break;
          case 1: dense1_regs<R, 1, true>(v, h, m, lane_ok); break;
          case 2: dense1_regs<R, 2, true>(v, h, m, lane_ok); break;
          default: dense1_regs<R, 3, true>(v, h, m, lane_ok); break;
        }
        break;
      case kOpDense1:
        switch (h.arg & 3) {
          case 0: dense1_regs<R, 0, false>(v, h, m, lane_ok); break;
          case 1: dense1_regs<R, 1, false>(v, h, m, lane_ok); break;
          case 2: dense1_regs<R, 2, false>(v, h, m, lane_ok); break;
          default: dense1_regs<R, 3, false>(v, h, m, lane_ok); break;
        }
        break;
      case kOpRegPhase:
#pragma clang loop unroll(full)
        for (int r = 0; r < kBReg; ++r) v[r] = cmul(m[r], v[r]);
        break;
      case kOpLanePhase: {
        uint32_t rm[4];
        const uint32_t di = diag_lane_part(h, lane_idx, rm);  // no register bits: rm stays 0
        if (lane_ok) P = cmul(m[di], P);
        pdirty = true;
        break;
      }
      case kOpNeg: {
        const uint32_t rreq = h.rreq;
        if (lane_ok) {
#pragma clang loop unroll(full)
          for (int r = 0; r < kBReg; ++r) {
            if (!reg_ok(r, rreq)) continue;
            v[r].re = -v[r].re;
            v[r].im = -v[r].im;
          }
        }
        break;
      }
      case kOpDiag:
        diag_regs(v, h, m, lane_idx, lane_ok);
        break;
      case kOpMap: {
        flush_lane_phase(v, P, pdirty);
        const int nmap = h.arg < 3 ? int(h.arg) : 0;
        to_lds(lds, v, map);
        __syncthreads();
        map = nmap;
        from_lds(lds, v, map);
        __syncthreads();
        lane_idx = tile_base | deposit(local_index(map, threadIdx.x, 0), gbit);
        break;
      }
      case kOpDenseLds: {
        flush_lane_phase(v, P, pdirty);
        to_lds(lds, v, map);
        __syncthreads();
        const uint32_t kk = h.arg;
        if (kk <= 1)
          dense_lds<R, 1>(lds, h, m, tile_base, gbit);
        else if (kk == 2)
          dense_lds<R, 2>(lds, h, m, tile_base, gbit);
        else
          dense_lds<R, 3>(lds, h, m, tile_base, gbit);
        __syncthreads();
        from_lds(lds, v, map);
        __syncthreads();
        break;
      }
      default:
        break;
    }
  }

  flush_lane_phase(v, P, pdirty);
  if (map != 0) {
    to_lds(lds, v, map);
    __syncthreads();
    map = 0;
    from_lds(lds, v, map);
    lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);
  }
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) st_nt(psi, lane_idx | deposit(uint32_t(r) << 8, gbit), v[r]);
}

template <typename R>
static int launch_blocked(void* state, int n, const int* tbits, const void* prog, size_t nbytes, int ninstr,
                          hipStream_t s) {
  BlkArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntiles = uint64_t(1) << (n - kBTileBits);
  a.nbytes = nbytes;
  a.ninstr = ninstr;
  for (int i = 0; i < kBTileBits; ++i) a.tbits[i] = uint8_t(tbits[i]);
  hipLaunchKernelGGL((k_blocked<R>), launch_grid(a.ntiles), dim3(kBT), 0, s, a,
                     reinterpret_cast<const uint8_t*>(prog), reinterpret_cast<cplx<R>*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

extern "C" {

size_t qsim_blocked_header_bytes(void) { return sizeof(BlkHdr); }

int qsim_apply_blocked(void* state, int nqubits, int dtype, const int* tile_bits, const void* program_dev,
                       size_t program_bytes, int ninstr, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && tile_bits != nullptr, "qsim_apply_blocked: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_blocked: bad dtype");
  QSIM_CHECK_ARG(nqubits >= kBTileBits && nqubits <= QSIM_MAX_QUBITS, "qsim_apply_blocked: nqubits out of range");
  QSIM_CHECK_ARG(ninstr >= 0 && (ninstr == 0 || (program_dev != nullptr && program_bytes >= sizeof(BlkHdr))),
                 "qsim_apply_blocked: bad program");
  QSIM_CHECK_ARG((reinterpret_cast<uintptr_t>(program_dev) & 15) == 0, "qsim_apply_blocked: program not 16-byte aligned");
  for (int i = 0; i < kBTileBits; ++i) {
    QSIM_CHECK_ARG(tile_bits[i] >= 0 && tile_bits[i] < nqubits, "qsim_apply_blocked: tile bit out of range");
    QSIM_CHECK_ARG(i == 0 || tile_bits[i] > tile_bits[i - 1], "qsim_apply_blocked: tile bits must ascend");
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return launch_blocked<double>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, s);
  return launch_blocked<float>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, s);
}

}  // extern "C"
