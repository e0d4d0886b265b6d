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
//   REG_PHASE (5)   v[r] *= m[r]: diagonal gates on register bits, folded by the host;
//                   arg = D != 0: only the registers with r & D != 0 (m[r] = 1 elsewhere)
//   LANE_PHASE (6)  diagonal on per-lane-fixed bits: multiplied into the lane's
//                   pending factor P, applied to the 16 amplitudes before the
//                   next MAP / DENSE_LDS / store (it commutes with everything
//                   that stays inside the lane's registers)
//   NEG (7)         v = -v (the -1 phase of CZ-like gates)
// Conditions: `tmask` = condition bits outside the tile (the same for the
// whole tile: a record whose tmask fails is skipped by the workgroup);
// `cmask` = condition bits inside the tile that select lanes (DENSE_LDS: all
// condition bits); `rreq` = register-bit conditions as a 4-bit mask over
// register positions. DIAG / DENSE1 / DENSE_LDS / MAP perform exactly the fused
// multiply-adds of the one-gate kernels, so a pass built only from them is
// bitwise equal to applying its gates one by one; the other forms change the
// order of roundings.
//
// Reference semantics: repeated Backend.apply_gate (abstract.py:2322-2361,
// 3176-3197) over the gates of Circuit.queue (abstract.py:3321-3322).

#include <hip/hip_runtime.h>

#include <cstring>

#include "blocked_dev.h"
#include "qsim_common.h"

namespace qsim {

// FROM_ZERO: the input is |0...0>, generated in registers (the state is only written).
template <typename R, bool FROM_ZERO>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(QSIM_BLOCKED_WAVES)))
void k_blocked(const BlkArgs a, const uint8_t* __restrict__ prog,
                                                 cplx<R>* __restrict__ psi) {
  __shared__ __align__(16) R lds[LdsSlots<R>::kLen];
  const uint64_t tile = xcd_block();
  if (tile >= a.ntiles) return;
  psi += blockIdx.z * a.state_stride;  // state and program of this batch entry
  prog += blockIdx.z * a.prog_stride;
  uint64_t gbit[kBTileBits];
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i) gbit[i] = uint64_t(1) << a.tbits[i];
  const uint64_t tile_base = tile_origin(a, tile);

  int map = 0;
  // global index of the lane's amplitudes with the register bits cleared
  uint64_t lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);

  cplx<R> v[kBReg];
  if constexpr (FROM_ZERO) {
#pragma clang loop unroll(full)
    for (int r = 0; r < kBReg; ++r) {
      v[r].re = (lane_idx | deposit(uint32_t(r) << 8, gbit)) == 0 ? R(1) : R(0);
      v[r].im = R(0);
    }
  } else {
    load_lane(a, psi, lane_idx, gbit, v);
  }

  cplx<R> P;  // pending per-lane factor of LANE_PHASE records
  P.re = R(1);
  P.im = R(0);
  bool pdirty = false;
  // Per-lane factor of LANE_REGPHASE records (controlled phases between one
  // register bit and lane bits), applied to the registers with that bit set
  // by a QFLUSH record; the host places the flushes. It lives in LDS so that
  // no register is held for it across the record loop.
  __shared__ cplx<R> q_lds[kBT];
  q_lds[threadIdx.x] = cplx<R>{R(1), R(0)};

  // The record stream is read ahead by one header. The buffer is readable for
  // nbytes + kBlockedPad bytes, so a header or entry read that starts inside
  // nbytes stays inside the allocation whatever the record claims.
  // 32-bit offsets (the host keeps programs below 2^31 bytes): the bound
  // checks then stay on the scalar unit (it has no 64-bit unsigned compare).
  const uint32_t nbytes = __builtin_amdgcn_readfirstlane(uint32_t(a.nbytes));
  const uint64_t tb_lo = __builtin_amdgcn_readfirstlane(uint32_t(tile_base));
  const uint64_t tb_hi = __builtin_amdgcn_readfirstlane(uint32_t(tile_base >> 32));
  const uint64_t tile_base_s = tb_lo | (tb_hi << 32);  // uniform copy (scalar registers)
  uint32_t off = 0;
  Hdr h = {};
  if (a.ninstr > 0) h = load_hdr(prog);
  for (int k = 0; k < a.ninstr && off + uint32_t(sizeof(BlkHdr)) <= nbytes; ++k) {
    const cplx<R>* m = reinterpret_cast<const cplx<R>*>(prog + off + sizeof(BlkHdr));
    const uint32_t nent = h.nent < kMaxEntries ? h.nent : kMaxEntries;  // no record is larger
    off += uint32_t(sizeof(BlkHdr)) + ((nent * uint32_t(sizeof(cplx<R>)) + 15u) & ~15u);
    const uint32_t next = off < nbytes ? off : nbytes;  // stays inside the padded buffer
    const Hdr hn = load_hdr(prog + next);  // next header, in flight while this record runs
    if ((tile_base_s & h.tmask) != h.tmask) {  // condition fails on the whole tile
      h = hn;
      continue;
    }
    // (entries read where used: reading them one record ahead as well, as the
    // adjoint kernel does, measured 7 % slower here -- SGPR pressure and the
    // copies of the read-ahead registers)
    const cplx<R>* me = m;
    // the lane condition is formed only by the forms that use it
    auto lane_ok = [&]() { return (lane_idx & h.cmask) == h.cmask; };
    switch (h.op) {
      case kOpLaneRegPhase: {
        const cplx<R> q = q_lds[threadIdx.x];
        q_lds[threadIdx.x] = csel(lane_ok(), cmul(me[0], q), q);
        break;
      }
      case kOpQFlush: {
        const cplx<R> q = q_lds[threadIdx.x];
        q_lds[threadIdx.x] = cplx<R>{R(1), R(0)};
        switch (h.arg & 3) {
          case 0: flush_q_bit<R, 0>(v, q); break;
          case 1: flush_q_bit<R, 1>(v, q); break;
          case 2: flush_q_bit<R, 2>(v, q); break;
          default: flush_q_bit<R, 3>(v, q); break;
        }
        break;
      }
      case kOpDense1Real:
        switch (h.arg & 3) {
          case 0: dense1_regs<R, 0, true>(v, h, me, lane_idx); break;
          case 1: dense1_regs<R, 1, true>(v, h, me, lane_idx); break;
          case 2: dense1_regs<R, 2, true>(v, h, me, lane_idx); break;
          default: dense1_regs<R, 3, true>(v, h, me, lane_idx); break;
        }
        break;
      case kOpDense1:
        switch (h.arg & 3) {
          case 0: dense1_regs<R, 0, false>(v, h, me, lane_idx); break;
          case 1: dense1_regs<R, 1, false>(v, h, me, lane_idx); break;
          case 2: dense1_regs<R, 2, false>(v, h, me, lane_idx); break;
          default: dense1_regs<R, 3, false>(v, h, me, lane_idx); break;
        }
        break;
      case kOpRot: {
        // (t, s) ride in the header (cmask and bits fields): no entry load
        R t, sn;
        if constexpr (sizeof(R) == 8) {
          t = __longlong_as_double(static_cast<long long>(h.cmask));
          sn = __hiloint2double(int(h.b1), int(h.b0));
        } else {
          t = __uint_as_float(uint32_t(h.cmask));
          sn = __uint_as_float(h.b0);
        }
        const uint32_t mode = h.arg >> 2;
        switch (h.arg & 3) {
          case 0: rot_regs<R, 0>(v, mode, t, sn); break;
          case 1: rot_regs<R, 1>(v, mode, t, sn); break;
          case 2: rot_regs<R, 2>(v, mode, t, sn); break;
          default: rot_regs<R, 3>(v, mode, t, sn); break;
        }
        break;
      }
      case kOpRegPhase:
        switch (h.arg & 15) {
          case 0: reg_phase<R, 0>(v, m); break;
          case 1: reg_phase<R, 1>(v, m); break;
          case 2: reg_phase<R, 2>(v, m); break;
          case 3: reg_phase<R, 3>(v, m); break;
          case 4: reg_phase<R, 4>(v, m); break;
          case 5: reg_phase<R, 5>(v, m); break;
          case 6: reg_phase<R, 6>(v, m); break;
          case 7: reg_phase<R, 7>(v, m); break;
          case 8: reg_phase<R, 8>(v, m); break;
          case 9: reg_phase<R, 9>(v, m); break;
          case 10: reg_phase<R, 10>(v, m); break;
          case 11: reg_phase<R, 11>(v, m); break;
          case 12: reg_phase<R, 12>(v, m); break;
          case 13: reg_phase<R, 13>(v, m); break;
          case 14: reg_phase<R, 14>(v, m); break;
          default: reg_phase<R, 15>(v, m); break;
        }
        break;
      case kOpLanePhase: {
        uint32_t rm[4];
        const uint32_t di = diag_lane_part(h, lane_idx, rm);  // no register bits: rm stays 0
        if (h.nt <= 1) {  // uniform entries: scalar loads and a per-lane select
          const cplx<R> d0 = me[0], d1 = h.nt ? me[1] : me[0];
          P = csel(lane_ok(), cmul(di ? d1 : d0, P), P);
        } else {
          P = csel(lane_ok(), cmul(m[di], P), P);
        }
        pdirty = true;
        break;
      }
      case kOpNeg: {
        // flip the sign bits where the lane condition holds: one XOR per component
        const uint32_t rreq = h.rreq;
        const uint32_t sign = lane_ok() ? 0x80000000u : 0u;
#pragma clang loop unroll(full)
        for (int r = 0; r < kBReg; ++r) {
          if (!reg_ok(r, rreq)) continue;
          flip_sign(v[r].re, sign);
          flip_sign(v[r].im, sign);
        }
        break;
      }
      case kOpDiag:
        diag_regs(v, h, m, lane_idx, lane_ok());
        break;
      case kOpMap: {
        flush_lane_phase(v, P, pdirty);
        const int nmap = h.arg < 3 ? int(h.arg) : 0;
        remap(lds, v, map, nmap);
        map = nmap;
        lane_idx = tile_base | deposit(local_index(map, threadIdx.x, 0), gbit);
        break;
      }
      case kOpDenseLds: {
        flush_lane_phase(v, P, pdirty);
        const uint32_t kk = h.arg;
        if (kk <= 1)
          dense_lds<R, 1>(lds, v, map, h, m, tile_base, gbit);
        else if (kk == 2)
          dense_lds<R, 2>(lds, v, map, h, m, tile_base, gbit);
        else
          dense_lds<R, 3>(lds, v, map, h, m, tile_base, gbit);
        break;
      }
      default:
        break;
    }
    h = hn;
  }

  flush_lane_phase(v, P, pdirty);
  if (map != 0) {
    remap(lds, v, map, 0);
    map = 0;
    lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);
  }
  // recompute the store addresses here instead of keeping the load addresses
  // live through the instruction loop (which spilled them to scratch)
  asm volatile("" : "+v"(lane_idx));
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) st_nt(psi, lane_idx | deposit(uint32_t(r) << 8, gbit), v[r]);
}

// The tiles of a pass with tile bits `tbits` over n qubits whose input
// vanishes wherever a bit outside `support` is 1 (all ones: every tile).
// (also used by the specialised passes, jit.hip)
void set_tiles(BlkArgs& a, int n, const int* tbits, uint64_t support) {
  uint64_t tmask = 0;
  for (int i = 0; i < kBTileBits; ++i) tmask |= uint64_t(1) << tbits[i];
  const uint64_t all = (uint64_t(1) << n) - 1;
  const uint64_t outer = all & ~tmask;
  const uint64_t freeb = outer & support;
  a.zload = 0;
  for (int i = 0; i < kBTileBits; ++i)
    if (!((support >> tbits[i]) & 1)) a.zload |= 1u << i;
  if (freeb == outer) {
    a.nfree = kDenseTiles;
    a.ntiles = uint64_t(1) << (n - kBTileBits);
    return;
  }
  a.nfree = 0;
  for (int b = 0; b < n; ++b)
    if ((freeb >> b) & 1) a.fbits[a.nfree++] = uint8_t(b);
  a.ntiles = uint64_t(1) << a.nfree;
}

template <typename R>
static int launch_blocked(void* state, int n, const int* tbits, const void* prog, size_t nbytes, int ninstr,
                          int nbatch, size_t prog_stride, hipStream_t s, bool from_zero = false,
                          uint64_t support = ~uint64_t(0)) {
  BlkArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nbytes = nbytes;
  a.state_stride = uint64_t(1) << n;
  a.prog_stride = prog_stride;
  a.ninstr = ninstr;
  for (int i = 0; i < kBTileBits; ++i) a.tbits[i] = uint8_t(tbits[i]);
  set_tiles(a, n, tbits, support);
  dim3 grid = launch_grid(a.ntiles);
  grid.z = unsigned(nbatch);
  if (from_zero)
    hipLaunchKernelGGL((k_blocked<R, true>), grid, dim3(kBT), 0, s, a, reinterpret_cast<const uint8_t*>(prog),
                       reinterpret_cast<cplx<R>*>(state));
  else
    hipLaunchKernelGGL((k_blocked<R, false>), grid, dim3(kBT), 0, s, a, reinterpret_cast<const uint8_t*>(prog),
                       reinterpret_cast<cplx<R>*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

// ---------------------------------------------------------------------------
// Adjoint sweep: the same records applied to two states of each batch entry
// (psi and lambda), plus INNER records that add Re <lambda| M |psi> over the
// tile into per-wave partial sums.
// ---------------------------------------------------------------------------
// (AdjArgs and the INNER-record helpers are in blocked_dev.h)

template <typename R>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(QSIM_ADJOINT_WAVES)))
void k_blocked_adjoint(const AdjArgs a, const uint8_t* __restrict__ prog,
                                                         cplx<R>* __restrict__ psi, cplx<R>* __restrict__ lam,
                                                         double* __restrict__ part) {
  __shared__ __align__(16) R lds[LdsSlots<R>::kLen];
  const uint64_t tile = xcd_block();
  if (tile >= a.ntiles) return;
  const uint64_t bz = blockIdx.z;
  psi += bz * a.state_stride;
  lam += bz * a.state_stride;
  prog += bz * a.prog_stride;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t gbit[kBTileBits];
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i) gbit[i] = uint64_t(1) << a.tbits[i];
  const uint64_t tile_base = insert_zeros(tile, a.tbits, kBTileBits);
  int map = 0;
  uint64_t lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);
  cplx<R> v[kBReg], w[kBReg];
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    const uint64_t g = lane_idx | deposit(uint32_t(r) << 8, gbit);
    v[r] = ld_nt(psi, g);
    w[r] = ld_nt(lam, g);
  }
  cplx<R> P;
  P.re = R(1);
  P.im = R(0);
  bool pdirty = false;
  // 32-bit offsets and a header read ahead, as in k_blocked: the record
  // decode stays on the scalar unit and overlaps the previous record
  const uint32_t nbytes = __builtin_amdgcn_readfirstlane(uint32_t(a.nbytes));
  const uint64_t tb_lo = __builtin_amdgcn_readfirstlane(uint32_t(tile_base));
  const uint64_t tb_hi = __builtin_amdgcn_readfirstlane(uint32_t(tile_base >> 32));
  const uint64_t tile_base_s = tb_lo | (tb_hi << 32);
  uint32_t off = 0;
  int slot = 0;
  Hdr h = {};
  Ent e = {};
  if (a.ninstr > 0) {
    h = load_hdr(prog);
    e = load_ent(prog + sizeof(BlkHdr));
  }
  for (int k = 0; k < a.ninstr && off + uint32_t(sizeof(BlkHdr)) <= nbytes; ++k) {
    // this record's header and first entries were read during the previous
    // one; they are awaited before the next record's reads are issued
    asm volatile("" ::"s"(h.op), "s"(e.w[0]), "s"(e.w[15]) : "memory");
    const cplx<R>* m = reinterpret_cast<const cplx<R>*>(prog + off + sizeof(BlkHdr));
    const uint32_t nent = h.nent < kMaxEntries ? h.nent : kMaxEntries;
    off += uint32_t(sizeof(BlkHdr)) + ((nent * uint32_t(sizeof(cplx<R>)) + 15u) & ~15u);
    const uint32_t next = off < nbytes ? off : nbytes;
    const Hdr hn = load_hdr(prog + next);
    const Ent en = load_ent(prog + next + sizeof(BlkHdr));
    cplx<R> me[4];
#pragma clang loop unroll(full)
    for (int i = 0; i < 4; ++i) me[i] = ent<R>(e, i);
    const bool tile_ok = (tile_base_s & h.tmask) == h.tmask;
    if (h.op == kOpInner1 || h.op == kOpInnerDiag) {
      // every INNER record writes its partial, zero where the condition fails
      double acc = 0.0;
      if (tile_ok) {
        if (h.op == kOpInnerDiag) {
          acc = inner_diag_regs(v, w, h, m, lane_idx);
        } else {
          switch (h.arg & 3) {
            case 0: acc = inner1_regs<R, 0>(v, w, h, me); break;
            case 1: acc = inner1_regs<R, 1>(v, w, h, me); break;
            case 2: acc = inner1_regs<R, 2>(v, w, h, me); break;
            default: acc = inner1_regs<R, 3>(v, w, h, me); break;
          }
        }
        // a pending lane factor P (a product of diagonal unitaries, applied
        // to both states) drops out: conj(P w) M (P v) = |P|^2 conj(w) M v
        acc = ((lane_idx & h.cmask) == h.cmask) ? acc : 0.0;
      }
      acc = wave_sum(acc);
      if (lane == 0 && slot < a.nslots)
        part[((bz * uint64_t(a.nslots) + uint64_t(slot)) * a.ntiles + tile) * 4 + wave] = acc;
      ++slot;
      h = hn;
      e = en;
      continue;
    }
    if (!tile_ok) {
      h = hn;
      e = en;
      continue;
    }
    const bool lane_ok = (lane_idx & h.cmask) == h.cmask;
    switch (h.op) {
      case kOpDense1Real:
      case kOpDense1: {
        const bool real = h.op == kOpDense1Real;
        switch ((h.arg & 3) | (real ? 4 : 0)) {
          case 0: dense1_regs<R, 0, false>(v, h, me, lane_idx); dense1_regs<R, 0, false>(w, h, me, lane_idx); break;
          case 1: dense1_regs<R, 1, false>(v, h, me, lane_idx); dense1_regs<R, 1, false>(w, h, me, lane_idx); break;
          case 2: dense1_regs<R, 2, false>(v, h, me, lane_idx); dense1_regs<R, 2, false>(w, h, me, lane_idx); break;
          case 3: dense1_regs<R, 3, false>(v, h, me, lane_idx); dense1_regs<R, 3, false>(w, h, me, lane_idx); break;
          case 4: dense1_regs<R, 0, true>(v, h, me, lane_idx); dense1_regs<R, 0, true>(w, h, me, lane_idx); break;
          case 5: dense1_regs<R, 1, true>(v, h, me, lane_idx); dense1_regs<R, 1, true>(w, h, me, lane_idx); break;
          case 6: dense1_regs<R, 2, true>(v, h, me, lane_idx); dense1_regs<R, 2, true>(w, h, me, lane_idx); break;
          default: dense1_regs<R, 3, true>(v, h, me, lane_idx); dense1_regs<R, 3, true>(w, h, me, lane_idx); break;
        }
        break;
      }
      case kOpDiag:
        diag_regs(v, h, m, lane_idx, lane_ok);
        diag_regs(w, h, m, lane_idx, lane_ok);
        break;
      case kOpNeg: {
        // -1 on the controlled subspace of both states: sign-bit flips
        const uint32_t rreq = h.rreq;
        const uint32_t sign = lane_ok ? 0x80000000u : 0u;
#pragma clang loop unroll(full)
        for (int r = 0; r < kBReg; ++r) {
          if (!reg_ok(r, rreq)) continue;
          flip_sign(v[r].re, sign);
          flip_sign(v[r].im, sign);
          flip_sign(w[r].re, sign);
          flip_sign(w[r].im, sign);
        }
        break;
      }
      case kOpLanePhase: {
        uint32_t rm[4];
        const uint32_t di = diag_lane_part(h, lane_idx, rm);
        P = csel(lane_ok, cmul(m[di], P), P);
        pdirty = true;
        break;
      }
      case kOpMap: {
        flush_lane_phase2(v, w, P, pdirty);
        const int nmap = h.arg < 3 ? int(h.arg) : 0;
        remap(lds, v, map, nmap);
        remap(lds, w, map, nmap);
        map = nmap;
        lane_idx = tile_base | deposit(local_index(map, threadIdx.x, 0), gbit);
        break;
      }
      default:
        break;
    }
    h = hn;
    e = en;
  }
  flush_lane_phase2(v, w, P, pdirty);
  if (map != 0) {
    remap(lds, v, map, 0);
    remap(lds, w, map, 0);
    map = 0;
    lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);
  }
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    const uint64_t g = lane_idx | deposit(uint32_t(r) << 8, gbit);
    st_nt(psi, g, v[r]);
    st_nt(lam, g, w[r]);
  }
}

// out[b * nparams + param_of[s]] = 2 * (sum over tiles and waves of part[b][s]),
// summed in a fixed order (one workgroup per (b, s)).
__global__ __launch_bounds__(256) void k_adjoint_reduce(const double* __restrict__ part, uint64_t per_slot,
                                                        int nslots, const int* __restrict__ param_of,
                                                        int nparams, double* __restrict__ out) {
  __shared__ double red[256];
  const uint64_t bs = blockIdx.x;  // b * nslots + s
  const double* src = part + bs * per_slot;
  double acc = 0.0;
  for (uint64_t i = threadIdx.x; i < per_slot; i += 256) acc += src[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint64_t b = bs / uint64_t(nslots);
    const int s = int(bs % uint64_t(nslots));
    const int p = param_of[s];
    if (p >= 0 && p < nparams) out[b * uint64_t(nparams) + uint64_t(p)] = 2.0 * red[0];
  }
}

template <typename R>
static int launch_adjoint(void* psi, void* lam, int n, const int* tbits, const void* prog, size_t nbytes,
                          int ninstr, int nbatch, size_t prog_stride, double* part, int nslots, hipStream_t s) {
  AdjArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntiles = uint64_t(1) << (n - kBTileBits);
  a.nbytes = nbytes;
  a.state_stride = uint64_t(1) << n;
  a.prog_stride = prog_stride;
  a.ninstr = ninstr;
  a.nslots = nslots;
  for (int i = 0; i < kBTileBits; ++i) a.tbits[i] = uint8_t(tbits[i]);
  dim3 grid = launch_grid(a.ntiles);
  grid.z = unsigned(nbatch);
  hipLaunchKernelGGL((k_blocked_adjoint<R>), grid, dim3(kBT), 0, s, a, reinterpret_cast<const uint8_t*>(prog),
                     reinterpret_cast<cplx<R>*>(psi), reinterpret_cast<cplx<R>*>(lam), part);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

namespace qsim {
// (also used by the specialised passes, jit.hip)
int check_blocked_args(int nqubits, int dtype, const int* tile_bits, const void* program_dev, size_t program_bytes,
                       int ninstr, const char* what) {
  QSIM_CHECK_ARG(tile_bits != nullptr, what);
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, what);
  QSIM_CHECK_ARG(nqubits >= kBTileBits && nqubits <= QSIM_MAX_QUBITS, what);
  QSIM_CHECK_ARG(ninstr >= 0 && (ninstr == 0 || (program_dev != nullptr && program_bytes >= sizeof(BlkHdr))), what);
  QSIM_CHECK_ARG(program_bytes < kMaxProgramBytes, what);
  QSIM_CHECK_ARG((reinterpret_cast<uintptr_t>(program_dev) & 15) == 0, what);
  for (int i = 0; i < kBTileBits; ++i) {
    QSIM_CHECK_ARG(tile_bits[i] >= 0 && tile_bits[i] < nqubits, what);
    QSIM_CHECK_ARG(i == 0 || tile_bits[i] > tile_bits[i - 1], what);
  }
  return QSIM_OK;
}
}  // namespace qsim

extern "C" {

size_t qsim_blocked_header_bytes(void) { return sizeof(BlkHdr); }

int qsim_apply_blocked_batched(void* states, int nbatch, int nqubits, int dtype, const int* tile_bits,
                               const void* programs_dev, size_t program_stride, size_t program_bytes, int ninstr,
                               void* stream) {
  clear_error();
  QSIM_CHECK_ARG(states != nullptr && nbatch >= 1 && nbatch <= 65535, "qsim_apply_blocked_batched: bad batch");
  QSIM_CHECK_ARG(program_stride % 16 == 0 && (nbatch == 1 || program_stride >= program_bytes),
                 "qsim_apply_blocked_batched: bad program stride");
  if (qsim::check_blocked_args(nqubits, dtype, tile_bits, programs_dev, program_bytes, ninstr,
                         "qsim_apply_blocked_batched: bad arguments") != QSIM_OK)
    return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return launch_blocked<double>(states, nqubits, tile_bits, programs_dev, program_bytes, ninstr, nbatch,
                                  program_stride, s);
  return launch_blocked<float>(states, nqubits, tile_bits, programs_dev, program_bytes, ninstr, nbatch,
                               program_stride, s);
}

size_t qsim_adjoint_partials_count(int nqubits, int nbatch, int nslots) {
  if (nqubits < kBTileBits || nqubits > QSIM_MAX_QUBITS || nbatch < 1 || nslots < 0) return 0;
  return (size_t(1) << (nqubits - kBTileBits)) * 4 * size_t(nbatch) * size_t(nslots);
}

int qsim_adjoint_blocked(void* psi, void* lam, int nbatch, int nqubits, int dtype, const int* tile_bits,
                         const void* programs_dev, size_t program_stride, size_t program_bytes, int ninstr,
                         int nslots, double* partials, size_t partials_count, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(psi != nullptr && lam != nullptr && psi != lam, "qsim_adjoint_blocked: bad states");
  QSIM_CHECK_ARG(nbatch >= 1 && nbatch <= 65535, "qsim_adjoint_blocked: bad batch");
  QSIM_CHECK_ARG(program_stride % 16 == 0 && (nbatch == 1 || program_stride >= program_bytes),
                 "qsim_adjoint_blocked: bad program stride");
  QSIM_CHECK_ARG(nslots >= 0 && (nslots == 0 || partials != nullptr), "qsim_adjoint_blocked: bad slots");
  QSIM_CHECK_ARG(partials_count >= qsim_adjoint_partials_count(nqubits, nbatch, nslots),
                 "qsim_adjoint_blocked: partials buffer too small");
  if (qsim::check_blocked_args(nqubits, dtype, tile_bits, programs_dev, program_bytes, ninstr,
                         "qsim_adjoint_blocked: bad arguments") != QSIM_OK)
    return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return launch_adjoint<double>(psi, lam, nqubits, tile_bits, programs_dev, program_bytes, ninstr, nbatch,
                                  program_stride, partials, nslots, s);
  return launch_adjoint<float>(psi, lam, nqubits, tile_bits, programs_dev, program_bytes, ninstr, nbatch,
                               program_stride, partials, nslots, s);
}

int qsim_adjoint_reduce(const double* partials, size_t partials_count, int nbatch, int nqubits, int nslots,
                        const int* param_of_dev, int nparams, double* grads, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(partials != nullptr && param_of_dev != nullptr && grads != nullptr, "qsim_adjoint_reduce: null");
  QSIM_CHECK_ARG(nbatch >= 1 && nslots >= 1 && nparams >= 1, "qsim_adjoint_reduce: bad sizes");
  QSIM_CHECK_ARG(nqubits >= kBTileBits && nqubits <= QSIM_MAX_QUBITS, "qsim_adjoint_reduce: bad nqubits");
  QSIM_CHECK_ARG(partials_count >= qsim_adjoint_partials_count(nqubits, nbatch, nslots),
                 "qsim_adjoint_reduce: partials buffer smaller than nbatch * nslots tiles of partial sums");
  const uint64_t per_slot = (uint64_t(1) << (nqubits - kBTileBits)) * 4;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_adjoint_reduce, dim3(unsigned(uint64_t(nbatch) * uint64_t(nslots))), dim3(256), 0, s,
                     partials, per_slot, nslots, param_of_dev, nparams, grads);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_apply_blocked_from_zero(void* state, int nqubits, int dtype, const int* tile_bits, const void* program_dev,
                                 size_t program_bytes, int ninstr, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr, "qsim_apply_blocked_from_zero: null state");
  if (qsim::check_blocked_args(nqubits, dtype, tile_bits, program_dev, program_bytes, ninstr,
                         "qsim_apply_blocked_from_zero: bad arguments") != QSIM_OK)
    return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // |0...0> has its only nonzero amplitude in tile 0 and the pass acts on the
  // tile bits only: every other tile of the output is zero. A large state is
  // filled with |0...0> (zeros at streaming speed) and the pass runs on tile
  // 0 alone; a small one (a launch costs more than its tiles) runs every tile
  // from zeros generated in registers, in one launch.
  uint64_t support = ~uint64_t(0);
  if (nqubits >= kFillFromZeroQubits) {
    const int zrc = qsim_zero_state(state, nqubits, dtype, stream);
    if (zrc != QSIM_OK) return zrc;
    support = 0;
  }
  if (dtype == QSIM_C128)
    return launch_blocked<double>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, 1, 0, s, true,
                                  support);
  return launch_blocked<float>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, 1, 0, s, true,
                               support);
}

int qsim_apply_blocked_support(void* state, int nqubits, int dtype, const int* tile_bits, uint64_t support,
                               const void* program_dev, size_t program_bytes, int ninstr, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr, "qsim_apply_blocked_support: null state");
  if (qsim::check_blocked_args(nqubits, dtype, tile_bits, program_dev, program_bytes, ninstr,
                               "qsim_apply_blocked_support: bad arguments") != QSIM_OK)
    return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return launch_blocked<double>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, 1, 0, s, false,
                                  support);
  return launch_blocked<float>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, 1, 0, s, false,
                               support);
}

int qsim_apply_blocked(void* state, int nqubits, int dtype, const int* tile_bits, const void* program_dev,
                       size_t program_bytes, int ninstr, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && tile_bits != nullptr, "qsim_apply_blocked: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_blocked: bad dtype");
  QSIM_CHECK_ARG(nqubits >= kBTileBits && nqubits <= QSIM_MAX_QUBITS, "qsim_apply_blocked: nqubits out of range");
  QSIM_CHECK_ARG(ninstr >= 0 && (ninstr == 0 || (program_dev != nullptr && program_bytes >= sizeof(BlkHdr))),
                 "qsim_apply_blocked: bad program");
  QSIM_CHECK_ARG(program_bytes < kMaxProgramBytes, "qsim_apply_blocked: program of 2^31 bytes or more");
  static_assert(kBlockedPad == QSIM_BLOCKED_PAD, "padding contract of the C ABI");
  QSIM_CHECK_ARG((reinterpret_cast<uintptr_t>(program_dev) & 15) == 0, "qsim_apply_blocked: program not 16-byte aligned");
  for (int i = 0; i < kBTileBits; ++i) {
    QSIM_CHECK_ARG(tile_bits[i] >= 0 && tile_bits[i] < nqubits, "qsim_apply_blocked: tile bit out of range");
    QSIM_CHECK_ARG(i == 0 || tile_bits[i] > tile_bits[i - 1], "qsim_apply_blocked: tile bits must ascend");
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return launch_blocked<double>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, 1, 0, s);
  return launch_blocked<float>(state, nqubits, tile_bits, program_dev, program_bytes, ninstr, 1, 0, s);
}

}  // extern "C"
