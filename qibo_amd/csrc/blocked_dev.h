// Device side of the blocked passes: the record layout, the register and LDS
// helpers and the per-record operations that k_blocked interprets. Nothing
// here needs the host runtime: the library also compiles this text under
// hiprtc into kernels specialised to one pass's record sequence (the same
// operations in the same order, without the interpreter's dispatch).
#pragma once

#include "qsim_device.h"

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
  kOpNeg = 7,
  kOpInner1 = 8,     // adjoint sweep only: Re <w| M |v> for a 2x2 M on a register bit
  kOpInnerDiag = 9,  // adjoint sweep only: Re <w| diag |v> on up to 5 bits
  kOpRot = 10,       // unconditional 2x2 rotation / reflection on a register bit, as three shears;
                     // no entries: t, s in the header (cmask, bits fields), float64 or float32
  kOpLaneRegPhase = 11,  // per-lane factor Q of the registers with one bit set (kept in LDS)
  kOpQFlush = 12,        // apply Q to the registers with bit arg set, reset it to 1
  kOpNop = 0xFF          // no effect (pads the programs of a batched launch to one record count)
};

// Record header (the host packs the same layout).
struct BlkHdr {
  uint8_t op;
  uint8_t arg;       // DENSE1*: register bit 0..3; ROT: bit | mode << 2; MAP: mapping 0..2; DENSE_LDS: number of targets 1..3
  uint8_t nt;        // DIAG / LANE_PHASE: number of target bits (0..5)
  uint8_t rreq;      // register-bit conditions (bit j: register bit j must be 1)
  uint32_t nent;     // complex entries following the header
  uint64_t cmask;    // lane-selecting condition bits (DENSE_LDS: all condition bits), all must be 1
  uint8_t bits[8];   // DIAG / LANE_PHASE: target bits, t0 = MSB of the diagonal index
                     // DENSE_LDS: [0..k) local target bits in matrix order, [4..4+k) the same ascending
  uint64_t tmask;    // condition bits outside the tile, all must be 1
};
static_assert(sizeof(BlkHdr) == 32, "record header is 32 bytes");
// Entries: DIAG / LANE_PHASE 2^nt; DENSE1*: 2x2 row-major; REG_PHASE: 16;
// DENSE_LDS: 2^k x 2^k row-major; MAP / NEG: none.

constexpr uint8_t kDenseTiles = 0xFF;               // BlkArgs::nfree of a pass over every tile
// From |0...0>, states of at least this many qubits are zero-filled and the
// first pass runs tile 0 alone; smaller ones run every tile in one launch.
constexpr int kFillFromZeroQubits = 22;
constexpr int kMaxOuterBits = 36 - kBTileBits;      // QSIM_MAX_QUBITS - 12

struct BlkArgs {
  uint64_t ntiles;
  uint64_t nbytes;        // program size: records past it are never read
  uint64_t state_stride;  // amplitudes between the states of a batch (grid z)
  uint64_t prog_stride;   // bytes between the programs of a batch
  int ninstr;
  uint8_t tbits[kBTileBits];  // global bit of local tile bit i, ascending
  // Tiles of the pass: every tile (nfree = kDenseTiles), or only those whose
  // bits outside the tile are zero except on the nfree bits fbits (ascending):
  // the input vanishes wherever another outer bit is 1, and so does the output.
  uint8_t nfree;
  uint8_t fbits[kMaxOuterBits];
  // Local bits outside the input's support: an amplitude whose local index
  // has one of them set is zero and is not read (its memory may hold
  // anything); 0 reads the whole tile.
  uint32_t zload;
};

// Global index of the first amplitude of tile number `tile`.
__device__ __forceinline__ uint64_t tile_origin(const BlkArgs& a, uint64_t tile) {
  if (a.nfree == kDenseTiles) return insert_zeros(tile, a.tbits, kBTileBits);
  uint64_t b = 0;
  for (int i = 0; i < int(a.nfree); ++i) b |= ((tile >> i) & 1) << a.fbits[i];
  return b;
}

// A header decoded from its eight dwords. gfx950 has no scalar byte loads, so
// the header is read as whole dwords (uniform address: scalar loads) and the
// byte fields are extracted with scalar shifts.
struct Hdr {
  uint32_t op, arg, nt, rreq, nent;
  uint64_t cmask, tmask;
  uint32_t b0, b1;  // bits[0..3], bits[4..7]
};

__device__ __forceinline__ Hdr load_hdr(const uint8_t* p) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
  const uint32_t w0 = w[0];
  Hdr h;
  h.op = w0 & 255u;
  h.arg = (w0 >> 8) & 255u;
  h.nt = (w0 >> 16) & 255u;
  h.rreq = w0 >> 24;
  h.nent = w[1];
  h.cmask = uint64_t(w[2]) | (uint64_t(w[3]) << 32);
  h.b0 = w[4];
  h.b1 = w[5];
  h.tmask = uint64_t(w[6]) | (uint64_t(w[7]) << 32);
  return h;
}

// The first 64 bytes of a record's entries, read with its header one record
// ahead by the adjoint sweep: the forms with at most 64 bytes of uniform
// entries (DENSE1*, INNER1) take them from these scalar registers instead of
// loading them when they run.
struct Ent {
  uint32_t w[16];
};

__device__ __forceinline__ Ent load_ent(const uint8_t* p) {
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
  Ent e;
#pragma clang loop unroll(full)
  for (int i = 0; i < 16; ++i) e.w[i] = q[i];
  return e;
}

// Entry i of the prefetched 64 bytes (i < 4 for complex128, i < 8 for complex64).
template <typename R>
__device__ __forceinline__ cplx<R> ent(const Ent& e, int i);

template <>
__device__ __forceinline__ cplx<double> ent<double>(const Ent& e, int i) {
  cplx<double> c;
  c.re = __hiloint2double(int(e.w[4 * i + 1]), int(e.w[4 * i]));
  c.im = __hiloint2double(int(e.w[4 * i + 3]), int(e.w[4 * i + 2]));
  return c;
}

template <>
__device__ __forceinline__ cplx<float> ent<float>(const Ent& e, int i) {
  cplx<float> c;
  c.re = __uint_as_float(e.w[2 * i]);
  c.im = __uint_as_float(e.w[2 * i + 1]);
  return c;
}

__device__ __forceinline__ uint32_t hbit(const Hdr& h, uint32_t t) {
  return ((t < 4 ? h.b0 : h.b1) >> (8 * (t & 3))) & 255u;
}

// Readable bytes the caller guarantees past the end of the record stream: the
// largest record (header + 64 complex128 entries) fits in it, so the
// read-ahead of one header and the entry reads of a record that starts inside
// the stream never leave the allocation.
constexpr uint64_t kBlockedPad = 1056;
constexpr uint32_t kMaxEntries = 64;  // largest record: an 8x8 block
constexpr size_t kMaxProgramBytes = size_t(1) << 31;  // record offsets are 32-bit

__device__ __forceinline__ int reg_base(int map) { return map == 0 ? 8 : (map == 1 ? 4 : 0); }

// Local tile index of register r of `lane` under mapping `map`.
__device__ __forceinline__ uint32_t local_index(int map, uint32_t lane, uint32_t r) {
  const int r0 = reg_base(map);
  return (lane & ((1u << r0) - 1)) | (r << r0) | ((lane >> r0) << (r0 + 4));
}

// XOR swizzle of LDS slots (8- or 4-byte components): bits 0-3 ^= bits 4-7 and
// bits 4-5 ^= bits 8-9. For every mapping the 64 lanes of a wave then hit each
// of the 64 four-byte banks at most twice with 8-byte slots and once with
// 4-byte slots (the minimum). A bijection of [0, 4096) (bits 6-11 are kept,
// bits 4-5 are recovered from bits 8-9, then bits 0-3 from bits 4-7), and
// linear over GF(2): swz(a ^ b) = swz(a) ^ swz(b), so a lane's slots are its
// own part XOR a per-register constant.
__device__ __forceinline__ uint32_t swz(uint32_t u) { return u ^ ((u >> 4) & 15u) ^ (((u >> 8) & 3u) << 4); }

// LDS slots of 8-byte components: an additive layout, slot(u) = the sum of
// kSlotW[i] over the set bits i of u (u + 1, 2, 4 for bits 5-7 and + 16, 40,
// 63, 126 for bits 8-11). The weights grow faster than the sum of the smaller
// ones, so the layout is a bijection onto a subset of [0, kSlotsF64), and
// slot(a | b) = slot(a) + slot(b) for disjoint a, b: a lane's slots are its
// own part plus a per-register constant, which the LDS instructions take as
// their immediate offset (no address arithmetic per access). Banks: 64 of 4
// bytes, two per 8-byte slot, so the slot residue mod 32 names the bank pair.
// The residues of the weights mod 32 are 1, 2, 4, 8, 16 on bits 0-4, 1, 2, 4
// on bits 5-7, 16 on bit 8 and 8 on bit 9: each half-wave (32 lanes) of
// mapping 0 (lane bits on u bits 0-4) and mapping 1 (u bits 0-3 and 8) hits 32
// distinct bank pairs; under mapping 2 (u bits 4-8) bits 4 and 8 share a
// residue and pairs of lanes share a bank pair (a two-way conflict), while
// each whole wave still spreads evenly over the 64 banks. No additive layout
// does better on all three mappings: a half-wave is conflict-free only if its
// five weights have residues of 2-adic valuation 0, 1, 2, 3, 4, so bits 4 and
// 8 would need the same valuation (they complete bits 0-3 under mappings 0
// and 1) and then collide under mapping 2.
constexpr uint32_t kSlotW[kBTileBits] = {1, 2, 4, 8, 16, 33, 66, 132, 272, 552, 1087, 2174};
constexpr uint32_t kSlotsF64 = 4348;  // 1 + the sum of the weights (34 784 bytes)

__device__ __forceinline__ constexpr uint32_t slot_f64(uint32_t u) {
  uint32_t s = 0;
  for (int i = 0; i < kBTileBits; ++i)
    if ((u >> i) & 1) s += kSlotW[i];
  return s;
}

// Slots of one LDS buffer of 4096 components of type R: the additive layout
// for 8-byte components in the kernels specialised to one record sequence
// (which define QSIM_LDS_ADDITIVE to 1: their mappings are constants), the XOR
// swizzle otherwise -- for 4-byte components, whose whole-wave accesses it
// spreads over all 64 banks (no additive layout does), and in the interpreter,
// whose run-time mappings would need a copy of every transpose per mapping
// (more scalar registers than the interpreter has left).
#ifndef QSIM_LDS_ADDITIVE
#define QSIM_LDS_ADDITIVE 0
#endif
template <typename R>
struct LdsSlots {
  static constexpr bool kAdd = QSIM_LDS_ADDITIVE && sizeof(R) == 8;
  // complex64 transposes of the specialised kernels move whole amplitudes
  // (8 bytes) in the additive layout: one exchange instead of one per
  // component; the other LDS uses keep 4-byte swizzled slots
  static constexpr bool kPair = QSIM_LDS_ADDITIVE && sizeof(R) == 4;
  static constexpr uint32_t kLen = kAdd ? kSlotsF64 : (kPair ? 2 * kSlotsF64 : uint32_t(kBTile));
  __device__ static __forceinline__ uint32_t of(uint32_t u) { return kAdd ? slot_f64(u) : swz(u); }
  // Slot of a | b for disjoint a, b from their slots.
  __device__ static __forceinline__ uint32_t join(uint32_t sa, uint32_t sb) { return kAdd ? sa + sb : sa ^ sb; }
};

// Local index of register 0 of `lane` under mapping `map` (register bits zero).
__device__ __forceinline__ uint32_t lane_local(int map, uint32_t lane) { return local_index(map, lane, 0); }

// Global-index bits contributed by local index u.
__device__ __forceinline__ uint64_t deposit(uint32_t u, const uint64_t* gbit) {
  uint64_t g = 0;
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i)
    if ((u >> i) & 1) g |= gbit[i];
  return g;
}

// The lane's amplitudes under mapping 0 (lane_idx: their global index with
// the register bits cleared), reading only those inside the support.
template <typename R>
__device__ __forceinline__ void load_lane(const BlkArgs& a, const cplx<R>* psi, uint64_t lane_idx,
                                          const uint64_t* gbit, cplx<R>* v) {
  if (a.zload == 0) {
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) v[r] = ld_nt(psi, lane_idx | deposit(uint32_t(r) << 8, gbit));
  } else {
    const uint32_t lz = local_index(0, threadIdx.x, 0) & a.zload;
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if ((lz | ((uint32_t(r) << 8) & a.zload)) == 0)
        v[r] = ld_nt(psi, lane_idx | deposit(uint32_t(r) << 8, gbit));
      else
        v[r] = cplx<R>{R(0), R(0)};
    }
  }
}

// The workgroup's LDS buffer holds one real component of the tile (4096
// values: 34 KB in complex128 with the additive layout's gaps, 16 KB in
// complex64), so a transpose moves the real parts, then the imaginary parts.
// Half the LDS of a whole complex tile lets twice as many workgroups share a
// CU.
// XOR swizzle: byte offset of register r's slot = the lane's part XOR a
// uniform constant, formed by one VALU instruction right before each access
// (volatile, so the 16 offsets are not kept live across the transpose: they
// would cost 16 registers each way).
__device__ __forceinline__ uint32_t slot_bytes(uint32_t lane_bytes, uint32_t reg_bytes) {
  uint32_t a;
  asm volatile("v_xor_b32 %0, %1, %2" : "=v"(a) : "s"(reg_bytes), "v"(lane_bytes));
  return a;
}

// Additive layout: register r's slot is the lane's base plus a constant after
// unrolling, the immediate offset of the LDS instruction. The mapping must be
// a constant for that, so a run-time mapping (the interpreter) branches to
// one of three copies.
template <int MAP, typename R>
__device__ __forceinline__ void lds_put_add(R* lds, const R* x) {
  R* base = lds + slot_f64(lane_local(MAP, threadIdx.x));
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) base[slot_f64(uint32_t(r) << reg_base(MAP))] = x[r];
}

template <int MAP, typename R>
__device__ __forceinline__ void lds_get_add(const R* lds, R* x) {
  const R* base = lds + slot_f64(lane_local(MAP, threadIdx.x));
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) x[r] = base[slot_f64(uint32_t(r) << reg_base(MAP))];
}

template <typename R>
__device__ __forceinline__ void lds_put(R* lds, const R* x, int map) {
  if constexpr (LdsSlots<R>::kAdd) {
    if (map == 0)
      lds_put_add<0>(lds, x);
    else if (map == 1)
      lds_put_add<1>(lds, x);
    else
      lds_put_add<2>(lds, x);
  } else {
    const int r0 = reg_base(map);
    const uint32_t s = swz(lane_local(map, threadIdx.x)) * uint32_t(sizeof(R));
    char* base = reinterpret_cast<char*>(lds);
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r)
      *reinterpret_cast<R*>(base + slot_bytes(s, swz(uint32_t(r) << r0) * uint32_t(sizeof(R)))) = x[r];
  }
}

template <typename R>
__device__ __forceinline__ void lds_get(const R* lds, R* x, int map) {
  if constexpr (LdsSlots<R>::kAdd) {
    if (map == 0)
      lds_get_add<0>(lds, x);
    else if (map == 1)
      lds_get_add<1>(lds, x);
    else
      lds_get_add<2>(lds, x);
  } else {
    const int r0 = reg_base(map);
    const uint32_t s = swz(lane_local(map, threadIdx.x)) * uint32_t(sizeof(R));
    const char* base = reinterpret_cast<const char*>(lds);
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r)
      x[r] = *reinterpret_cast<const R*>(base + slot_bytes(s, swz(uint32_t(r) << r0) * uint32_t(sizeof(R))));
  }
}

// Change of register mapping: amplitude (lane, r) of mapping `from` moves to
// its place under mapping `to`, through LDS, one component at a time (whole
// amplitudes for complex64 in the additive layout).
template <typename R>
__device__ __forceinline__ void remap(R* lds, cplx<R>* v, int from, int to) {
  if constexpr (LdsSlots<R>::kPair) {
    cplx<R>* l = reinterpret_cast<cplx<R>*>(lds);  // 8-byte aligned (the buffer is 16-byte aligned)
    lds_put(l, v, from);
    __syncthreads();
    lds_get(l, v, to);
    __syncthreads();
    return;
  }
  R x[kBReg];
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) x[r] = v[r].re;
  lds_put(lds, x, from);
  __syncthreads();
  lds_get(lds, x, to);
  __syncthreads();
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    v[r].re = x[r];
    x[r] = v[r].im;
  }
  lds_put(lds, x, from);
  __syncthreads();
  lds_get(lds, x, to);
  __syncthreads();
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) v[r].im = x[r];
}

// Register r satisfies the register-bit conditions `rreq` (r is a constant
// after unrolling, rreq is uniform: a scalar test).
__device__ __forceinline__ bool reg_ok(int r, uint32_t rreq) { return (uint32_t(r) & rreq) == rreq; }

// One-qubit gate on register bit J of the current mapping: the arithmetic of
// k_dense (cmul then cfma) for DENSE1; two real multiplies per component for
// DENSE1_REAL.
// In-place kernels of the two hot forms. The values are updated in their own
// registers (tied asm operands), with the roundings of the C++ forms below
// (v_mul then v_fma, exactly fmaR(b, x1, a * x0)); written this way because
// the compiler otherwise routes the results through v_fmac temporaries and
// copies the lane's amplitude array at the instruction loop's merge points.
__device__ __forceinline__ void real2_inplace(double& x0, double& x1, double a, double b, double c, double d) {
  double t0, t1;
  asm("v_mul_f64 %[t0], %[a], %[x0]\n\t"
      "v_mul_f64 %[t1], %[c], %[x0]\n\t"
      "v_fma_f64 %[x0], %[b], %[x1], %[t0]\n\t"
      "v_fma_f64 %[x1], %[d], %[x1], %[t1]"
      : [x0] "+v"(x0), [x1] "+v"(x1), [t0] "=&v"(t0), [t1] "=&v"(t1)
      : [a] "s"(a), [b] "s"(b), [c] "s"(c), [d] "s"(d));
}

__device__ __forceinline__ void real2_inplace(float& x0, float& x1, float a, float b, float c, float d) {
  float t0, t1;
  asm("v_mul_f32 %[t0], %[a], %[x0]\n\t"
      "v_mul_f32 %[t1], %[c], %[x0]\n\t"
      "v_fma_f32 %[x0], %[b], %[x1], %[t0]\n\t"
      "v_fma_f32 %[x1], %[d], %[x1], %[t1]"
      : [x0] "+v"(x0), [x1] "+v"(x1), [t0] "=&v"(t0), [t1] "=&v"(t1)
      : [a] "s"(a), [b] "s"(b), [c] "s"(c), [d] "s"(d));
}

// One shear step in place, d <- F(d, a*x) with a uniform:
// F = 0: d - a*x, 1: d + a*x, 2: a*x - d (source negation is free in VOP3).
template <int F>
__device__ __forceinline__ void shear_inplace(double& d, double a, double x) {
  if constexpr (F == 0)
    asm("v_fma_f64 %[d], -%[a], %[x], %[d]" : [d] "+v"(d) : [a] "s"(a), [x] "v"(x));
  else if constexpr (F == 1)
    asm("v_fma_f64 %[d], %[a], %[x], %[d]" : [d] "+v"(d) : [a] "s"(a), [x] "v"(x));
  else
    asm("v_fma_f64 %[d], %[a], %[x], -%[d]" : [d] "+v"(d) : [a] "s"(a), [x] "v"(x));
}

template <int F>
__device__ __forceinline__ void shear_inplace(float& d, float a, float x) {
  if constexpr (F == 0)
    asm("v_fma_f32 %[d], -%[a], %[x], %[d]" : [d] "+v"(d) : [a] "s"(a), [x] "v"(x));
  else if constexpr (F == 1)
    asm("v_fma_f32 %[d], %[a], %[x], %[d]" : [d] "+v"(d) : [a] "s"(a), [x] "v"(x));
  else
    asm("v_fma_f32 %[d], %[a], %[x], -%[d]" : [d] "+v"(d) : [a] "s"(a), [x] "v"(x));
}

// The rotation [[c, -s], [s, c]] of the real pair (x0, x1) as three shears
// with t = s / (1 + c), c >= 0 (Paeth's decomposition):
//   x0 -= t x1;  x1 += s x0;  x0 -= t x1
// -- three multiply-adds instead of two multiplies and two multiply-adds.
// In exact arithmetic 1 - t s = c and t (2 - t s) = s, so the product is the
// rotation; each step rounds once. NEG=true rotates by the opposite angle.
template <bool NEG, typename R>
__device__ __forceinline__ void rot_pair(R& x0, R& x1, R t, R s) {
  if constexpr (!NEG) {
    shear_inplace<0>(x0, t, x1);
    shear_inplace<1>(x1, s, x0);
    shear_inplace<0>(x0, t, x1);
  } else {
    shear_inplace<1>(x0, t, x1);
    shear_inplace<0>(x1, s, x0);
    shear_inplace<1>(x0, t, x1);
  }
}

// The reflection [[c, s], [s, -c]] = rotation . diag(1, -1), same shears:
//   x0 += t x1;  x1 = s x0 - x1;  x0 -= t x1
template <typename R>
__device__ __forceinline__ void refl_pair(R& x0, R& x1, R t, R s) {
  shear_inplace<1>(x0, t, x1);
  shear_inplace<2>(x1, s, x0);
  shear_inplace<0>(x0, t, x1);
}

// ROT record on register bit J, mode = arg >> 2, entry 0 = (t, s):
//   0: [[c, -s], [s, c]]   (RY)           -- each component rotated
//   1: [[c, s], [s, -c]]   (H)            -- each component reflected
//   2: [[c, -is], [-is, c]] (RX)          -- (x0.im, x1.re) rotated by the
//      angle, (x0.re, x1.im) by its opposite
template <typename R, int J>
__device__ __forceinline__ void rot_regs(cplx<R>* v, uint32_t mode, R t, R s) {
  if (mode == 0) {
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if ((r >> J) & 1) continue;
      rot_pair<false>(v[r].re, v[r | (1 << J)].re, t, s);
      rot_pair<false>(v[r].im, v[r | (1 << J)].im, t, s);
    }
  } else if (mode == 1) {
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if ((r >> J) & 1) continue;
      refl_pair(v[r].re, v[r | (1 << J)].re, t, s);
      refl_pair(v[r].im, v[r | (1 << J)].im, t, s);
    }
  } else {
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if ((r >> J) & 1) continue;
      rot_pair<false>(v[r].im, v[r | (1 << J)].re, t, s);
      rot_pair<true>(v[r].re, v[r | (1 << J)].im, t, s);
    }
  }
}

// x *= d with d uniform: exactly cmul(d, x) (re = fma(d.re, x.re, -(d.im x.im)),
// im = fma(d.re, x.im, d.im x.re)).
__device__ __forceinline__ void cmul_inplace(cplx<double>& x, cplx<double> d) {
  double t1, t2;
  asm("v_mul_f64 %[t1], %[di], %[xi]\n\t"
      "v_mul_f64 %[t2], %[di], %[xr]\n\t"
      "v_fma_f64 %[xr], %[dr], %[xr], -%[t1]\n\t"
      "v_fma_f64 %[xi], %[dr], %[xi], %[t2]"
      : [xr] "+v"(x.re), [xi] "+v"(x.im), [t1] "=&v"(t1), [t2] "=&v"(t2)
      : [dr] "s"(d.re), [di] "s"(d.im));
}

__device__ __forceinline__ void cmul_inplace(cplx<float>& x, cplx<float> d) {
  float t1, t2;
  asm("v_mul_f32 %[t1], %[di], %[xi]\n\t"
      "v_mul_f32 %[t2], %[di], %[xr]\n\t"
      "v_fma_f32 %[xr], %[dr], %[xr], -%[t1]\n\t"
      "v_fma_f32 %[xi], %[dr], %[xi], %[t2]"
      : [xr] "+v"(x.re), [xi] "+v"(x.im), [t1] "=&v"(t1), [t2] "=&v"(t2)
      : [dr] "s"(d.re), [di] "s"(d.im));
}

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

__device__ __forceinline__ void flip_sign(double& x, uint32_t sign) {
  x = __longlong_as_double(__double_as_longlong(x) ^ (uint64_t(sign) << 32));
}

__device__ __forceinline__ void flip_sign(float& x, uint32_t sign) {
  x = __uint_as_float(__float_as_uint(x) ^ sign);
}

template <typename R>
__device__ __forceinline__ cplx<R> csel(bool c, cplx<R> a, cplx<R> b) {
  cplx<R> r;
  r.re = c ? a.re : b.re;
  r.im = c ? a.im : b.im;
  return r;
}

// Lane conditions are applied with selects, not branches: a divergent branch
// around register updates makes the compiler copy the whole register array.
template <typename R, int J, bool REAL>
__device__ __forceinline__ void dense1_regs(cplx<R>* v, const Hdr& h, const cplx<R>* m, uint64_t lane_idx) {
  const cplx<R> m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
  const uint32_t rreq = h.rreq;
  if (rreq == 0 && h.cmask == 0) {  // unconditional: straight-line code
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if ((r >> J) & 1) continue;
      if constexpr (REAL) {
        real2_inplace(v[r].re, v[r | (1 << J)].re, m0.re, m1.re, m2.re, m3.re);
        real2_inplace(v[r].im, v[r | (1 << J)].im, m0.re, m1.re, m2.re, m3.re);
      } else {
        dense1_pair<R, REAL>(v[r], v[r | (1 << J)], m0, m1, m2, m3);
      }
    }
    return;
  }
  const bool lane_ok = (lane_idx & h.cmask) == h.cmask;
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    if ((r >> J) & 1) continue;
    if (!reg_ok(r, rreq)) continue;
    cplx<R> a = v[r], b = v[r | (1 << J)];
    dense1_pair<R, REAL>(a, b, m0, m1, m2, m3);
    v[r] = csel(lane_ok, a, v[r]);
    v[r | (1 << J)] = csel(lane_ok, b, v[r | (1 << J)]);
  }
}

// Diagonal index contributions of one DIAG / LANE_PHASE record: `dl` from the
// lane's fixed bits, rm[j] from register bit j (uniform).
__device__ __forceinline__ uint32_t diag_lane_part(const Hdr& h, uint64_t lane_idx, uint32_t* rm) {
  const uint32_t nt = h.nt < 5 ? h.nt : 5;
  uint32_t dl = 0;
  rm[0] = rm[1] = rm[2] = rm[3] = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t b = hbit(h, t);
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
__device__ __forceinline__ void diag_regs(cplx<R>* v, const Hdr& h, const cplx<R>* m, uint64_t lane_idx,
                                          bool lane_ok) {
  uint32_t rm[4];
  const uint32_t dl = diag_lane_part(h, lane_idx, rm);
  const uint32_t rreq = h.rreq;
  if (h.nt <= 1) {
    // Uniform entries. Two per-lane factors, chosen once: f0 for amplitudes
    // whose diagonal index is the lane part alone, f1 for those a register
    // bit sets; a lane outside the condition multiplies by exactly 1 (the
    // per-gate arithmetic, cmul(d, v), for the others).
    const cplx<R> d0 = m[0], d1 = h.nt ? m[1] : m[0];
    cplx<R> one;
    one.re = R(1);
    one.im = R(0);
    const cplx<R> f0 = csel(lane_ok, dl ? d1 : d0, one);
    const cplx<R> f1 = csel(lane_ok, d1, one);
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if (!reg_ok(r, rreq)) continue;
      const bool rbit = ((r & 1) && rm[0]) || ((r & 2) && rm[1]) || ((r & 4) && rm[2]) || ((r & 8) && rm[3]);
      v[r] = cmul(rbit ? f1 : f0, v[r]);
    }
    return;
  }
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    if (!reg_ok(r, rreq)) continue;
    const uint32_t di = dl | ((r & 1) ? rm[0] : 0u) | ((r & 2) ? rm[1] : 0u) | ((r & 4) ? rm[2] : 0u) |
                        ((r & 8) ? rm[3] : 0u);
    v[r] = csel(lane_ok, cmul(m[di], v[r]), v[r]);
  }
}

// REG_PHASE: v[r] *= m[r] on the registers r with r & D != 0 (D = 0: all
// 16); the host writes exactly 1 for the others, which are left alone.
template <typename R, int D>
__device__ __forceinline__ void reg_phase(cplx<R>* v, const cplx<R>* m) {
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    if (D != 0 && (r & D) == 0) continue;
    cmul_inplace(v[r], m[r]);
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

// Multiply the registers whose bit J is set by the per-lane factor q.
template <typename R, int J>
__device__ __forceinline__ void flush_q_bit(cplx<R>* v, cplx<R> q) {
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r)
    if ((r >> J) & 1) v[r] = cmul(q, v[r]);
}

// Dense 2^k block (k = 1..3) on local bits. Each lane owns 16 / 2^k groups of
// 2^k amplitudes (group g = lane + 256 i); the groups are gathered from the
// tile through LDS one component at a time, multiplied in registers, and
// scattered back the same way.
template <typename R, int K>
__device__ __forceinline__ void dense_lds(R* lds, cplx<R>* v, int map, const Hdr& in, const cplx<R>* m,
                                          uint64_t tile_base, const uint64_t* gbit) {
  constexpr int D = 1 << K;
  constexpr int NG = (kBTile >> K) / kBT;  // groups per lane
  uint32_t srt[K];
  uint32_t off[D];
#pragma clang loop unroll(full)
  for (int t = 0; t < K; ++t) srt[t] = hbit(in, 4 + t) < kBTileBits ? hbit(in, 4 + t) : kBTileBits - 1;
#pragma clang loop unroll(full)
  for (int r = 0; r < D; ++r) {
    uint32_t o = 0;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) {
      const uint32_t b = hbit(in, t) < kBTileBits ? hbit(in, t) : kBTileBits - 1;
      if ((r >> (K - 1 - t)) & 1) o |= 1u << b;
    }
    off[r] = o;
  }
  uint32_t base[NG];
  bool act[NG];
#pragma clang loop unroll(full)
  for (int i = 0; i < NG; ++i) {
    uint32_t b = threadIdx.x + i * kBT;
#pragma clang loop unroll(full)
    for (int t = 0; t < K; ++t) {
      const uint32_t low = b & ((1u << srt[t]) - 1);
      b = low | ((b ^ low) << 1);
    }
    base[i] = b;
    act[i] = ((tile_base | deposit(b, gbit)) & in.cmask) == in.cmask;
  }
  using L = LdsSlots<R>;
  uint32_t slot[kBReg];  // the lane's amplitudes in the current mapping
  const uint32_t s0 = L::of(lane_local(map, threadIdx.x));
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {  // per-register constants, selected by the uniform mapping
    const uint32_t rs =
        map == 0 ? L::of(uint32_t(r) << 8) : (map == 1 ? L::of(uint32_t(r) << 4) : L::of(uint32_t(r)));
    slot[r] = L::join(s0, rs);
  }
  // group i's input c sits at base[i] | off[c] (disjoint bits)
  uint32_t bslot[NG], oslot[D];
#pragma clang loop unroll(full)
  for (int i = 0; i < NG; ++i) bslot[i] = L::of(base[i]);
#pragma clang loop unroll(full)
  for (int c = 0; c < D; ++c) oslot[c] = L::of(off[c]);
  cplx<R> g[kBReg];  // g[i * D + c]: input c of group i
  // gather: real parts, then imaginary parts
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) lds[slot[r]] = v[r].re;
  __syncthreads();
#pragma clang loop unroll(full)
  for (int i = 0; i < NG; ++i)
#pragma clang loop unroll(full)
    for (int c = 0; c < D; ++c) g[i * D + c].re = lds[L::join(bslot[i], oslot[c])];
  __syncthreads();
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) lds[slot[r]] = v[r].im;
  __syncthreads();
#pragma clang loop unroll(full)
  for (int i = 0; i < NG; ++i)
#pragma clang loop unroll(full)
    for (int c = 0; c < D; ++c) g[i * D + c].im = lds[L::join(bslot[i], oslot[c])];
  __syncthreads();
  // multiply (the arithmetic of the one-gate kernels: cmul, then cfma in column order)
#pragma clang loop unroll(full)
  for (int i = 0; i < NG; ++i) {
    if (!act[i]) continue;
    cplx<R> y[D];
#pragma clang loop unroll(full)
    for (int r = 0; r < D; ++r) {
      cplx<R> acc = cmul(m[r * D], g[i * D]);
#pragma clang loop unroll(full)
      for (int c = 1; c < D; ++c) acc = cfma(m[r * D + c], g[i * D + c], acc);
      y[r] = acc;
    }
#pragma clang loop unroll(full)
    for (int r = 0; r < D; ++r) g[i * D + r] = y[r];
  }
  // scatter back and reload the lane's amplitudes in the current mapping
#pragma clang loop unroll(full)
  for (int i = 0; i < NG; ++i)
#pragma clang loop unroll(full)
    for (int c = 0; c < D; ++c) lds[L::join(bslot[i], oslot[c])] = g[i * D + c].re;
  __syncthreads();
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) v[r].re = lds[slot[r]];
  __syncthreads();
#pragma clang loop unroll(full)
  for (int i = 0; i < NG; ++i)
#pragma clang loop unroll(full)
    for (int c = 0; c < D; ++c) lds[L::join(bslot[i], oslot[c])] = g[i * D + c].im;
  __syncthreads();
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) v[r].im = lds[slot[r]];
  __syncthreads();
}

#ifndef QSIM_BLOCKED_WAVES
#define QSIM_BLOCKED_WAVES 4
#endif

// ---------------------------------------------------------------------------
// Adjoint sweep: the same records applied to two states of each batch entry
// (psi and lambda), plus INNER records that add Re <lambda| M |psi> over the
// tile into per-wave partial sums.
// ---------------------------------------------------------------------------
struct AdjArgs {
  uint64_t ntiles;
  uint64_t nbytes;
  uint64_t state_stride;
  uint64_t prog_stride;
  int ninstr;
  int nslots;  // INNER records per program
  uint8_t tbits[kBTileBits];
};

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

template <typename R>
__device__ __forceinline__ double re_conj_mul(cplx<R> w, double yr, double yi) {
  return double(w.re) * yr + double(w.im) * yi;
}

// Re sum over the lane's pairs of conj(w) . (M v) for a 2x2 M on register bit J.
template <typename R, int J>
__device__ __forceinline__ double inner1_regs(const cplx<R>* v, const cplx<R>* w, const Hdr& h,
                                              const cplx<R>* m) {
  // Re sum conj(w) M v = sum_ij Re(m_ij S_ij) with S_ij = sum over pairs of
  // conj(w_i) v_j: the lane accumulates the S_ij (4 FMAs each per pair) and
  // applies M once. Generators of X/Y rotations have m00 = m11 = 0, so only
  // S_01 and S_10 are formed for them (a uniform branch on the record).
  const uint32_t rreq = h.rreq;
  const bool offdiag = m[0].re == R(0) && m[0].im == R(0) && m[3].re == R(0) && m[3].im == R(0);
  double s00r = 0.0, s00i = 0.0, s01r = 0.0, s01i = 0.0, s10r = 0.0, s10i = 0.0, s11r = 0.0, s11i = 0.0;
  if (offdiag) {
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if ((r >> J) & 1) continue;
      if (!reg_ok(r, rreq)) continue;
      const int r1 = r | (1 << J);
      const double a0r = w[r].re, a0i = w[r].im, a1r = w[r1].re, a1i = w[r1].im;
      const double x0r = v[r].re, x0i = v[r].im, x1r = v[r1].re, x1i = v[r1].im;
      s01r = fma(a0r, x1r, fma(a0i, x1i, s01r));
      s01i = fma(a0r, x1i, fma(-a0i, x1r, s01i));
      s10r = fma(a1r, x0r, fma(a1i, x0i, s10r));
      s10i = fma(a1r, x0i, fma(-a1i, x0r, s10i));
    }
  } else {
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      if ((r >> J) & 1) continue;
      if (!reg_ok(r, rreq)) continue;
      const int r1 = r | (1 << J);
      const double a0r = w[r].re, a0i = w[r].im, a1r = w[r1].re, a1i = w[r1].im;
      const double x0r = v[r].re, x0i = v[r].im, x1r = v[r1].re, x1i = v[r1].im;
      s00r = fma(a0r, x0r, fma(a0i, x0i, s00r));
      s00i = fma(a0r, x0i, fma(-a0i, x0r, s00i));
      s01r = fma(a0r, x1r, fma(a0i, x1i, s01r));
      s01i = fma(a0r, x1i, fma(-a0i, x1r, s01i));
      s10r = fma(a1r, x0r, fma(a1i, x0i, s10r));
      s10i = fma(a1r, x0i, fma(-a1i, x0r, s10i));
      s11r = fma(a1r, x1r, fma(a1i, x1i, s11r));
      s11i = fma(a1r, x1i, fma(-a1i, x1r, s11i));
    }
  }
  const double m0r = m[0].re, m0i = m[0].im, m1r = m[1].re, m1i = m[1].im;
  const double m2r = m[2].re, m2i = m[2].im, m3r = m[3].re, m3i = m[3].im;
  return (m0r * s00r - m0i * s00i) + (m1r * s01r - m1i * s01i) + (m2r * s10r - m2i * s10i) +
         (m3r * s11r - m3i * s11i);
}

template <typename R>
__device__ __forceinline__ double inner_diag_regs(const cplx<R>* v, const cplx<R>* w, const Hdr& h,
                                                  const cplx<R>* m, uint64_t lane_idx) {
  uint32_t rm[4];
  const uint32_t dl = diag_lane_part(h, lane_idx, rm);
  const uint32_t rreq = h.rreq;
  double acc = 0.0;
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    if (!reg_ok(r, rreq)) continue;
    const uint32_t di = dl | ((r & 1) ? rm[0] : 0u) | ((r & 2) ? rm[1] : 0u) | ((r & 4) ? rm[2] : 0u) |
                        ((r & 8) ? rm[3] : 0u);
    const cplx<R> d = m[di];
    const double yr = double(d.re) * double(v[r].re) - double(d.im) * double(v[r].im);
    const double yi = double(d.re) * double(v[r].im) + double(d.im) * double(v[r].re);
    acc += re_conj_mul(w[r], yr, yi);
  }
  return acc;
}

template <typename R>
__device__ __forceinline__ void flush_lane_phase2(cplx<R>* v, cplx<R>* w, cplx<R>& P, bool& pdirty) {
  if (!pdirty) return;
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    v[r] = cmul(P, v[r]);
    w[r] = cmul(P, w[r]);
  }
  P.re = R(1);
  P.im = R(0);
  pdirty = false;
}

#ifndef QSIM_ADJOINT_WAVES
#define QSIM_ADJOINT_WAVES 3
#endif

}  // namespace qsim
