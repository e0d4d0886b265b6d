// Blocked passes specialised to their record sequence, compiled at run time.
//
// k_blocked interprets a pass's records: per record it loads the header,
// branches on the opcode and dispatches on the register bit / mapping / mask
// argument. A program's record *sequence* (opcode, argument, whether the
// record carries a tile condition) fixes all of those branches, so the
// library can emit one straight-line kernel per sequence: the same device
// operations (blocked_dev.h) in the same order, with the opcode and argument
// as constants and every other field (angles, masks, entries) still read from
// the program at run time. The result is bitwise that of k_blocked on the
// same program; the compiler can now schedule across records and hoist the
// scalar loads of later records. Programs that differ only in their values
// (rotation angles, phases, condition masks) share one kernel.
//
// The source is compiled with hiprtc for gfx950 with the flags of the
// library's own build, loaded into the current device's context, and kept
// for the life of the process in a table keyed by the sequence.

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "blocked_dev.h"
#include "qsim_common.h"

namespace qsim {

int check_blocked_args(int nqubits, int dtype, const int* tile_bits, const void* program_dev, size_t program_bytes,
                       int ninstr, const char* what);
void set_tiles(BlkArgs& a, int n, const int* tbits, uint64_t support);

namespace {

// The text of qsim_device.h and blocked_dev.h (written by the build).
#include "jit_sources.inc"

struct RecView {
  uint32_t op, arg;
  bool cond;      // nonzero tile condition: the record runs only where it holds
  uint32_t nent;  // entries (at most kMaxEntries): fixes the next record's offset
  uint32_t off;
};

// What a specialised kernel does: a pass (k_blocked), a pass whose input is
// |0...0> (k_blocked<R, true>), an adjoint sweep (k_blocked_adjoint).
constexpr int kVariantPass = 0, kVariantFromZero = 1, kVariantAdjoint = 2;

// The records k_blocked would visit: at most ninstr, while a header fits in
// the program (the interpreter's loop condition), with its record lengths.
bool decode(const uint8_t* prog, size_t nbytes, int ninstr, int dtype, std::vector<RecView>& out, std::string& why) {
  const uint32_t esize = dtype == QSIM_C128 ? 16u : 8u;
  uint32_t off = 0;
  for (int k = 0; k < ninstr && size_t(off) + sizeof(BlkHdr) <= nbytes; ++k) {
    BlkHdr h;
    std::memcpy(&h, prog + off, sizeof(BlkHdr));
    const uint32_t nent = h.nent < kMaxEntries ? h.nent : kMaxEntries;
    out.push_back({h.op, h.arg, h.tmask != 0, nent, off});
    if (h.op == kOpMap && h.tmask != 0) {
      // the mapping after a conditional MAP depends on the tile: not static
      why = "conditional mapping change";
      return false;
    }
    off += uint32_t(sizeof(BlkHdr)) + ((nent * esize + 15u) & ~15u);
  }
  return true;
}

// Estimated hiprtc compile time of a sequence, in milliseconds of host time
// (measured on this build: rotations, phase tables and sign flips compile in
// ~15 ms each, 2x2 and diagonal records in ~100-120 ms, the adjoint sweep's
// inner products in ~450 ms). Longer sequences are not specialised: a
// kernel that takes minutes to compile would not pay for itself.
constexpr uint64_t kMaxSpecCostMs = 10000;

uint64_t spec_cost_ms(const std::vector<RecView>& recs) {
  uint64_t c = 0;
  for (const RecView& r : recs) {
    switch (r.op) {
      case kOpRot: case kOpRegPhase: case kOpNeg: c += 15; break;
      case kOpLanePhase: case kOpLaneRegPhase: case kOpQFlush: c += 30; break;
      case kOpMap: case kOpDense1: case kOpDense1Real: c += 100; break;
      case kOpDiag: c += 120; break;
      case kOpDenseLds: c += 300; break;
      case kOpInner1: case kOpInnerDiag: c += 450; break;
      default: break;
    }
  }
  return c;
}

std::string structure_key(const std::vector<RecView>& recs, int dtype, int variant) {
  std::string key;
  key.reserve(recs.size() * 4 + 2);
  key.push_back(char('0' + dtype));
  key.push_back(char('0' + variant));
  for (const RecView& r : recs) {
    key.push_back(char(r.op));
    key.push_back(char(r.arg));
    key.push_back(char(r.cond ? 1 : 0));
    key.push_back(char(r.nent));  // <= 64
  }
  return key;
}

// One record's code: the body of k_blocked's switch for this opcode with the
// argument fixed; `map` is the mapping in force (every MAP is unconditional).
void emit_record(std::string& s, const RecView& r, bool dbl, int& map) {
  const std::string off = std::to_string(r.off);
  s += "  {\n    const Hdr h = load_hdr(prog + " + off + "u);\n";
  s += "    const cplx<R>* m = reinterpret_cast<const cplx<R>*>(prog + " + std::to_string(r.off + 32) + "u);\n";
  s += "    const cplx<R>* me = m;\n    (void)me;\n";
  s += "    auto lane_ok = [&]() { return (lane_idx & h.cmask) == h.cmask; };\n    (void)lane_ok;\n";
  s += r.cond ? "    if ((tile_base_s & h.tmask) == h.tmask) {\n" : "    {\n";
  const std::string J = std::to_string(r.arg & 3);
  switch (r.op) {
    case kOpLaneRegPhase:
      s += "      const cplx<R> q = q_lds[threadIdx.x];\n"
           "      q_lds[threadIdx.x] = csel(lane_ok(), cmul(me[0], q), q);\n";
      break;
    case kOpQFlush:
      s += "      const cplx<R> q = q_lds[threadIdx.x];\n"
           "      q_lds[threadIdx.x] = cplx<R>{R(1), R(0)};\n"
           "      flush_q_bit<R, " + J + ">(v, q);\n";
      break;
    case kOpDense1Real:
      s += "      dense1_regs<R, " + J + ", true>(v, h, me, lane_idx);\n";
      break;
    case kOpDense1:
      s += "      dense1_regs<R, " + J + ", false>(v, h, me, lane_idx);\n";
      break;
    case kOpRot:
      if (dbl)
        s += "      const R t = __longlong_as_double(static_cast<long long>(h.cmask));\n"
             "      const R sn = __hiloint2double(int(h.b1), int(h.b0));\n";
      else
        s += "      const R t = __uint_as_float(uint32_t(h.cmask));\n"
             "      const R sn = __uint_as_float(h.b0);\n";
      s += "      rot_regs<R, " + J + ">(v, " + std::to_string(r.arg >> 2) + "u, t, sn);\n";
      break;
    case kOpRegPhase:
      s += "      reg_phase<R, " + std::to_string(r.arg & 15) + ">(v, m);\n";
      break;
    case kOpLanePhase:
      s += "      uint32_t rm[4];\n"
           "      const uint32_t di = diag_lane_part(h, lane_idx, rm);\n"
           "      if (h.nt <= 1) {\n"
           "        const cplx<R> d0 = me[0], d1 = h.nt ? me[1] : me[0];\n"
           "        P = csel(lane_ok(), cmul(di ? d1 : d0, P), P);\n"
           "      } else {\n"
           "        P = csel(lane_ok(), cmul(m[di], P), P);\n"
           "      }\n"
           "      pdirty = true;\n";
      break;
    case kOpNeg:
      s += "      const uint32_t rreq = h.rreq;\n"
           "      const uint32_t sign = lane_ok() ? 0x80000000u : 0u;\n"
           "#pragma clang loop unroll(full)\n"
           "      for (int r = 0; r < kBReg; ++r) {\n"
           "        if (!reg_ok(r, rreq)) continue;\n"
           "        flip_sign(v[r].re, sign);\n"
           "        flip_sign(v[r].im, sign);\n"
           "      }\n";
      break;
    case kOpDiag:
      s += "      diag_regs(v, h, m, lane_idx, lane_ok());\n";
      break;
    case kOpMap: {
      const int nmap = r.arg < 3 ? int(r.arg) : 0;
      s += "      flush_lane_phase(v, P, pdirty);\n";
      s += "      remap(lds, v, " + std::to_string(map) + ", " + std::to_string(nmap) + ");\n";
      s += "      lane_idx = tile_base | deposit(local_index(" + std::to_string(nmap) + ", threadIdx.x, 0), gbit);\n";
      map = nmap;
      break;
    }
    case kOpDenseLds: {
      const int kk = r.arg <= 1 ? 1 : (r.arg == 2 ? 2 : 3);
      s += "      flush_lane_phase(v, P, pdirty);\n";
      s += "      dense_lds<R, " + std::to_string(kk) + ">(lds, v, " + std::to_string(map) +
           ", h, m, tile_base, gbit);\n";
      break;
    }
    default:  // NOP and the adjoint-only forms: no effect, as in k_blocked
      break;
  }
  s += "    }\n  }\n";
}

// The adjoint sweep's kernel (k_blocked_adjoint): every record applied to
// both states, INNER records writing their per-wave partial sums at slot
// numbers fixed here.
std::string make_adjoint_source(const std::vector<RecView>& recs, int dtype) {
  const bool dbl = dtype == QSIM_C128;
  std::string s;
  s.reserve(2048 + recs.size() * 500);
  s += "#define QSIM_LDS_ADDITIVE 1\n#include \"blocked_dev.h\"\nnamespace qsim {\n";
  s += "extern \"C\" __global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(QSIM_ADJOINT_WAVES)))\n";
  s += dbl ? "void qsim_adjoint_spec(const AdjArgs a, const uint8_t* __restrict__ prog, cplx<double>* __restrict__ psi, "
             "cplx<double>* __restrict__ lam, double* __restrict__ part) {\n  using R = double;\n"
           : "void qsim_adjoint_spec(const AdjArgs a, const uint8_t* __restrict__ prog, cplx<float>* __restrict__ psi, "
             "cplx<float>* __restrict__ lam, double* __restrict__ part) {\n  using R = float;\n";
  s += R"(  __shared__ __align__(16) R lds[LdsSlots<R>::kLen];
  const uint64_t tile = xcd_block();
  if (tile >= a.ntiles) return;
  const uint64_t bz = blockIdx.z;
  psi += bz * a.state_stride;
  lam += bz * a.state_stride;
  prog += bz * a.prog_stride;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  (void)wave;
  (void)lane;
  uint64_t gbit[kBTileBits];
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i) gbit[i] = uint64_t(1) << a.tbits[i];
  const uint64_t tile_base = insert_zeros(tile, a.tbits, kBTileBits);
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
  const uint64_t tb_lo = __builtin_amdgcn_readfirstlane(uint32_t(tile_base));
  const uint64_t tb_hi = __builtin_amdgcn_readfirstlane(uint32_t(tile_base >> 32));
  const uint64_t tile_base_s = tb_lo | (tb_hi << 32);
  (void)tile_base_s;
)";
  int map = 0, slot = 0;
  for (const RecView& r : recs) {
    s += "  {\n    const Hdr h = load_hdr(prog + " + std::to_string(r.off) + "u);\n";
    s += "    const cplx<R>* m = reinterpret_cast<const cplx<R>*>(prog + " + std::to_string(r.off + 32) + "u);\n";
    s += "    (void)m;\n";
    const std::string J = std::to_string(r.arg & 3);
    if (r.op == kOpInner1 || r.op == kOpInnerDiag) {
      s += "    double acc = 0.0;\n";
      s += r.cond ? "    if ((tile_base_s & h.tmask) == h.tmask) {\n" : "    {\n";
      if (r.op == kOpInnerDiag)
        s += "      acc = inner_diag_regs(v, w, h, m, lane_idx);\n";
      else
        s += "      acc = inner1_regs<R, " + J + ">(v, w, h, m);\n";
      s += "      acc = ((lane_idx & h.cmask) == h.cmask) ? acc : 0.0;\n    }\n";
      s += "    acc = wave_sum(acc);\n";
      s += "    if (lane == 0 && " + std::to_string(slot) + " < a.nslots)\n";
      s += "      part[((bz * uint64_t(a.nslots) + " + std::to_string(slot) +
           "u) * a.ntiles + tile) * 4 + wave] = acc;\n  }\n";
      ++slot;
      continue;
    }
    s += r.cond ? "    if ((tile_base_s & h.tmask) == h.tmask) {\n" : "    {\n";
    s += "      const bool lane_ok = (lane_idx & h.cmask) == h.cmask;\n      (void)lane_ok;\n";
    switch (r.op) {
      case kOpDense1Real:
      case kOpDense1: {
        const std::string real = r.op == kOpDense1Real ? "true" : "false";
        s += "      dense1_regs<R, " + J + ", " + real + ">(v, h, m, lane_idx);\n";
        s += "      dense1_regs<R, " + J + ", " + real + ">(w, h, m, lane_idx);\n";
        break;
      }
      case kOpDiag:
        s += "      diag_regs(v, h, m, lane_idx, lane_ok);\n      diag_regs(w, h, m, lane_idx, lane_ok);\n";
        break;
      case kOpNeg:
        s += "      const uint32_t rreq = h.rreq;\n"
             "      const uint32_t sign = lane_ok ? 0x80000000u : 0u;\n"
             "#pragma clang loop unroll(full)\n"
             "      for (int r = 0; r < kBReg; ++r) {\n"
             "        if (!reg_ok(r, rreq)) continue;\n"
             "        flip_sign(v[r].re, sign);\n"
             "        flip_sign(v[r].im, sign);\n"
             "        flip_sign(w[r].re, sign);\n"
             "        flip_sign(w[r].im, sign);\n"
             "      }\n";
        break;
      case kOpLanePhase:
        s += "      uint32_t rm[4];\n"
             "      const uint32_t di = diag_lane_part(h, lane_idx, rm);\n"
             "      P = csel(lane_ok, cmul(m[di], P), P);\n"
             "      pdirty = true;\n";
        break;
      case kOpMap: {
        const int nmap = r.arg < 3 ? int(r.arg) : 0;
        s += "      flush_lane_phase2(v, w, P, pdirty);\n";
        s += "      remap(lds, v, " + std::to_string(map) + ", " + std::to_string(nmap) + ");\n";
        s += "      remap(lds, w, " + std::to_string(map) + ", " + std::to_string(nmap) + ");\n";
        s += "      lane_idx = tile_base | deposit(local_index(" + std::to_string(nmap) +
             ", threadIdx.x, 0), gbit);\n";
        map = nmap;
        break;
      }
      default:  // forms the sweep does not use: no effect, as in k_blocked_adjoint
        break;
    }
    s += "    }\n  }\n";
  }
  s += "  flush_lane_phase2(v, w, P, pdirty);\n";
  if (map != 0) {
    s += "  remap(lds, v, " + std::to_string(map) + ", 0);\n";
    s += "  remap(lds, w, " + std::to_string(map) + ", 0);\n";
    s += "  lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);\n";
  }
  s += R"(#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    const uint64_t g = lane_idx | deposit(uint32_t(r) << 8, gbit);
    st_nt(psi, g, v[r]);
    st_nt(lam, g, w[r]);
  }
}
}  // namespace qsim
)";
  return s;
}

std::string make_source(const std::vector<RecView>& recs, int dtype, int variant) {
  if (variant == kVariantAdjoint) return make_adjoint_source(recs, dtype);
  const int from_zero = variant == kVariantFromZero ? 1 : 0;
  const bool dbl = dtype == QSIM_C128;
  std::string s;
  s.reserve(2048 + recs.size() * 400);
  s += "#define QSIM_LDS_ADDITIVE 1\n#include \"blocked_dev.h\"\nnamespace qsim {\n";
  s += "extern \"C\" __global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(QSIM_BLOCKED_WAVES)))\n";
  s += "void qsim_blocked_spec(const BlkArgs a, const uint8_t* __restrict__ prog, cplx<";
  s += dbl ? "double" : "float";
  s += ">* __restrict__ psi) {\n";
  s += dbl ? "  using R = double;\n" : "  using R = float;\n";
  s += R"(  __shared__ __align__(16) R lds[LdsSlots<R>::kLen];
  const uint64_t tile = xcd_block();
  if (tile >= a.ntiles) return;
  psi += blockIdx.z * a.state_stride;
  prog += blockIdx.z * a.prog_stride;
  uint64_t gbit[kBTileBits];
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i) gbit[i] = uint64_t(1) << a.tbits[i];
  const uint64_t tile_base = tile_origin(a, tile);
  uint64_t lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);
  cplx<R> v[kBReg];
)";
  if (from_zero)
    s += R"(#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) {
    v[r].re = (lane_idx | deposit(uint32_t(r) << 8, gbit)) == 0 ? R(1) : R(0);
    v[r].im = R(0);
  }
)";
  else
    s += "  load_lane(a, psi, lane_idx, gbit, v);\n";
  s += R"(  cplx<R> P;
  P.re = R(1);
  P.im = R(0);
  bool pdirty = false;
)";
  // the per-lane factor of LANE_REGPHASE records lives in LDS; a sequence
  // without them leaves it out, so the kernel needs only the transpose buffer
  // (32 KiB in complex128: five workgroups per CU instead of four)
  bool qfactor = false;
  for (const RecView& r : recs) qfactor |= r.op == kOpLaneRegPhase || r.op == kOpQFlush;
  if (qfactor)
    s += R"(  __shared__ cplx<R> q_lds[kBT];
  q_lds[threadIdx.x] = cplx<R>{R(1), R(0)};
)";
  s += R"(  const uint64_t tb_lo = __builtin_amdgcn_readfirstlane(uint32_t(tile_base));
  const uint64_t tb_hi = __builtin_amdgcn_readfirstlane(uint32_t(tile_base >> 32));
  const uint64_t tile_base_s = tb_lo | (tb_hi << 32);
  (void)tile_base_s;
)";
  int map = 0;
  for (const RecView& r : recs) emit_record(s, r, dbl, map);
  s += "  flush_lane_phase(v, P, pdirty);\n";
  if (map != 0) {
    s += "  remap(lds, v, " + std::to_string(map) + ", 0);\n";
    s += "  lane_idx = tile_base | deposit(local_index(0, threadIdx.x, 0), gbit);\n";
  }
  s += R"(  asm volatile("" : "+v"(lane_idx));
#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) st_nt(psi, lane_idx | deposit(uint32_t(r) << 8, gbit), v[r]);
}
}  // namespace qsim
)";
  return s;
}

// hiprtc compile with the flags of the library's build of blocked.hip.
int compile(const std::string& src, std::vector<char>& code, std::string& log) {
  hiprtcProgram prog;
  const char* hdr_src[2] = {kJitQsimDeviceH, kJitBlockedDevH};
  const char* hdr_names[2] = {"qsim_device.h", "blocked_dev.h"};
  if (hiprtcCreateProgram(&prog, src.c_str(), "qsim_blocked_spec.hip", 2, hdr_src, hdr_names) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return -1;
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        // no packed FP32 instructions, as in the library's own build
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-mllvm", "-structurizecfg-skip-uniform-regions=true"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
  size_t lsz = 0;
  hiprtcGetProgramLogSize(prog, &lsz);
  if (lsz > 1) {
    log.resize(lsz);
    hiprtcGetProgramLog(prog, &log[0]);
  }
  if (rc != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return -1;
  }
  size_t csz = 0;
  hiprtcGetCodeSize(prog, &csz);
  code.resize(csz);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return 0;
}

struct Spec {
  std::string key;
  hipModule_t module;
  hipFunction_t fn;
  int device, dtype, variant, nrec;
  size_t min_bytes;  // the last record's header ends here: a program must be at least this long
};

std::mutex g_mu;
std::vector<Spec> g_specs;                         // handle = index + 1
std::unordered_map<std::string, uint64_t> g_index;  // key -> handle

int prepare(const void* program, size_t program_bytes, int ninstr, int dtype, int variant,
            std::vector<RecView>& recs, const char* what, bool budget = true) {
  QSIM_CHECK_ARG(program != nullptr || ninstr == 0, what);
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, what);
  QSIM_CHECK_ARG(variant >= kVariantPass && variant <= kVariantAdjoint, what);
  QSIM_CHECK_ARG(ninstr >= 0 && program_bytes < kMaxProgramBytes, what);
  std::string why;
  if (!decode(reinterpret_cast<const uint8_t*>(program), program_bytes, ninstr, dtype, recs, why)) {
    set_error(std::string(what) + ": " + why);
    return QSIM_EINVAL;
  }
  if (budget && spec_cost_ms(recs) > kMaxSpecCostMs) {
    set_error(std::string(what) + ": sequence too long to specialise (estimated compile time " +
              std::to_string(spec_cost_ms(recs)) + " ms)");
    return QSIM_EINVAL;
  }
  return QSIM_OK;
}

std::string device_key(const std::vector<RecView>& recs, int dtype, int variant, int device) {
  return structure_key(recs, dtype, variant) + "@" + std::to_string(device);
}

// The kernel of a handle, checked against the launch's device, dtype, variant
// and program size.
int spec_for(uint64_t handle, int dtype, int variant, size_t program_bytes, int ninstr, Spec& sp, const char* what) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    QSIM_CHECK_ARG(handle >= 1 && handle <= g_specs.size(), what);
    sp = g_specs[handle - 1];
  }
  int device = 0;
  QSIM_CHECK_HIP(hipGetDevice(&device));
  QSIM_CHECK_ARG(device == sp.device && dtype == sp.dtype && variant == sp.variant, what);
  QSIM_CHECK_ARG(ninstr >= sp.nrec && program_bytes >= sp.min_bytes, what);
  return QSIM_OK;
}

}  // namespace
}  // namespace qsim

using namespace qsim;

extern "C" {

size_t qsim_blocked_specialized_source(const void* program, size_t program_bytes, int ninstr, int dtype,
                                       int variant, char* out, size_t cap) {
  clear_error();
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, variant, recs, "qsim_blocked_specialized_source", false) !=
      QSIM_OK)
    return 0;
  const std::string src = make_source(recs, dtype, variant);
  if (out != nullptr && cap > 0) {
    const size_t n = src.size() < cap - 1 ? src.size() : cap - 1;
    std::memcpy(out, src.data(), n);
    out[n] = '\0';
  }
  return src.size() + 1;
}

int qsim_blocked_compile_check(const void* program, size_t program_bytes, int ninstr, int dtype, int variant,
                               size_t* code_bytes) {
  clear_error();
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, variant, recs, "qsim_blocked_compile_check", false) != QSIM_OK)
    return QSIM_EINVAL;
  std::vector<char> code;
  std::string log;
  if (compile(make_source(recs, dtype, variant), code, log) != 0) {
    set_error("qsim_blocked_compile_check: hiprtc failed: " + log);
    return QSIM_EHIP;
  }
  if (code_bytes != nullptr) *code_bytes = code.size();
  return QSIM_OK;
}

int qsim_blocked_specialize(const void* program, size_t program_bytes, int ninstr, int dtype, int variant,
                            uint64_t* handle) {
  clear_error();
  QSIM_CHECK_ARG(handle != nullptr, "qsim_blocked_specialize: null handle");
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, variant, recs, "qsim_blocked_specialize") != QSIM_OK)
    return QSIM_EINVAL;
  int device = 0;
  QSIM_CHECK_HIP(hipGetDevice(&device));
  const std::string key = device_key(recs, dtype, variant, device);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_index.find(key);
    if (it != g_index.end()) {
      *handle = it->second;
      return QSIM_OK;
    }
  }
  std::vector<char> code;
  std::string log;
  if (compile(make_source(recs, dtype, variant), code, log) != 0) {
    set_error("qsim_blocked_specialize: hiprtc failed: " + log);
    return QSIM_EHIP;
  }
  hipModule_t mod;
  QSIM_CHECK_HIP(hipModuleLoadData(&mod, code.data()));
  hipFunction_t fn;
  const char* name = variant == kVariantAdjoint ? "qsim_adjoint_spec" : "qsim_blocked_spec";
  const hipError_t e = hipModuleGetFunction(&fn, mod, name);
  if (e != hipSuccess) {
    (void)hipModuleUnload(mod);
    set_error(std::string("qsim_blocked_specialize: hipModuleGetFunction: ") + hipGetErrorString(e));
    return QSIM_EHIP;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_index.find(key);
  if (it != g_index.end()) {  // another thread finished the same sequence first
    (void)hipModuleUnload(mod);
    *handle = it->second;
    return QSIM_OK;
  }
  const size_t min_bytes = recs.empty() ? 0 : size_t(recs.back().off) + sizeof(BlkHdr);
  g_specs.push_back({key, mod, fn, device, dtype, variant, int(recs.size()), min_bytes});
  *handle = uint64_t(g_specs.size());
  g_index.emplace(key, *handle);
  return QSIM_OK;
}

int qsim_blocked_specialized_lookup(const void* program, size_t program_bytes, int ninstr, int dtype, int variant,
                                    uint64_t* handle, uint64_t* key_hash) {
  clear_error();
  QSIM_CHECK_ARG(handle != nullptr, "qsim_blocked_specialized_lookup: null handle");
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, variant, recs, "qsim_blocked_specialized_lookup") != QSIM_OK)
    return QSIM_EINVAL;
  int device = 0;
  QSIM_CHECK_HIP(hipGetDevice(&device));
  const std::string key = device_key(recs, dtype, variant, device);
  if (key_hash != nullptr) {  // FNV-1a of the key
    uint64_t hsh = 1469598103934665603ull;
    for (const char c : key) hsh = (hsh ^ uint8_t(c)) * 1099511628211ull;
    *key_hash = hsh;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_index.find(key);
  *handle = it == g_index.end() ? 0 : it->second;
  return QSIM_OK;
}

int qsim_apply_blocked_specialized(uint64_t handle, void* states, int nbatch, int nqubits, int dtype,
                                   const int* tile_bits, uint64_t support, const void* programs_dev,
                                   size_t program_stride, size_t program_bytes, int ninstr, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(states != nullptr && nbatch >= 1 && nbatch <= 65535, "qsim_apply_blocked_specialized: bad batch");
  QSIM_CHECK_ARG(program_stride % 16 == 0 && (nbatch == 1 || program_stride >= program_bytes),
                 "qsim_apply_blocked_specialized: bad program stride");
  if (check_blocked_args(nqubits, dtype, tile_bits, programs_dev, program_bytes, ninstr,
                         "qsim_apply_blocked_specialized: bad arguments") != QSIM_OK)
    return QSIM_EINVAL;
  Spec sp;
  {
    // a pass kernel of either kind (the variant decides the input)
    std::lock_guard<std::mutex> lk(g_mu);
    QSIM_CHECK_ARG(handle >= 1 && handle <= g_specs.size(), "qsim_apply_blocked_specialized: unknown handle");
    sp = g_specs[handle - 1];
  }
  QSIM_CHECK_ARG(sp.variant != kVariantAdjoint, "qsim_apply_blocked_specialized: handle of an adjoint sweep");
  if (spec_for(handle, dtype, sp.variant, program_bytes, ninstr, sp,
               "qsim_apply_blocked_specialized: kernel for another device, dtype or program") != QSIM_OK)
    return QSIM_EINVAL;
  if (sp.variant == kVariantFromZero) {
    // as qsim_apply_blocked_from_zero: a large state filled with |0...0> and
    // the pass run on tile 0, a small one run on every tile
    QSIM_CHECK_ARG(nbatch == 1, "qsim_apply_blocked_specialized: a from-zero pass takes one state");
    support = ~uint64_t(0);
    if (nqubits >= kFillFromZeroQubits) {
      const int zrc = qsim_zero_state(states, nqubits, dtype, stream);
      if (zrc != QSIM_OK) return zrc;
      support = 0;
    }
  }
  BlkArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nbytes = program_bytes;
  a.state_stride = uint64_t(1) << nqubits;
  a.prog_stride = program_stride;
  a.ninstr = ninstr;
  for (int i = 0; i < kBTileBits; ++i) a.tbits[i] = uint8_t(tile_bits[i]);
  set_tiles(a, nqubits, tile_bits, support);
  const dim3 grid = launch_grid(a.ntiles);
  const uint8_t* prog = reinterpret_cast<const uint8_t*>(programs_dev);
  void* args[] = {&a, &prog, &states};
  QSIM_CHECK_HIP(hipModuleLaunchKernel(sp.fn, grid.x, grid.y, unsigned(nbatch), kBT, 1, 1, 0,
                                       reinterpret_cast<hipStream_t>(stream), args, nullptr));
  return QSIM_OK;
}

int qsim_adjoint_blocked_specialized(uint64_t handle, void* psi, void* lam, int nbatch, int nqubits, int dtype,
                                     const int* tile_bits, const void* programs_dev, size_t program_stride,
                                     size_t program_bytes, int ninstr, int nslots, double* partials,
                                     size_t partials_count, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(psi != nullptr && lam != nullptr && psi != lam, "qsim_adjoint_blocked_specialized: bad states");
  QSIM_CHECK_ARG(nbatch >= 1 && nbatch <= 65535, "qsim_adjoint_blocked_specialized: bad batch");
  QSIM_CHECK_ARG(program_stride % 16 == 0 && (nbatch == 1 || program_stride >= program_bytes),
                 "qsim_adjoint_blocked_specialized: bad program stride");
  QSIM_CHECK_ARG(nslots >= 0 && (nslots == 0 || partials != nullptr), "qsim_adjoint_blocked_specialized: bad slots");
  QSIM_CHECK_ARG(partials_count >= qsim_adjoint_partials_count(nqubits, nbatch, nslots),
                 "qsim_adjoint_blocked_specialized: partials buffer too small");
  if (check_blocked_args(nqubits, dtype, tile_bits, programs_dev, program_bytes, ninstr,
                         "qsim_adjoint_blocked_specialized: bad arguments") != QSIM_OK)
    return QSIM_EINVAL;
  Spec sp;
  if (spec_for(handle, dtype, kVariantAdjoint, program_bytes, ninstr, sp,
               "qsim_adjoint_blocked_specialized: not an adjoint kernel of this device, dtype or program") != QSIM_OK)
    return QSIM_EINVAL;
  AdjArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntiles = uint64_t(1) << (nqubits - kBTileBits);
  a.nbytes = program_bytes;
  a.state_stride = uint64_t(1) << nqubits;
  a.prog_stride = program_stride;
  a.ninstr = ninstr;
  a.nslots = nslots;
  for (int i = 0; i < kBTileBits; ++i) a.tbits[i] = uint8_t(tile_bits[i]);
  const dim3 grid = launch_grid(a.ntiles);
  const uint8_t* prog = reinterpret_cast<const uint8_t*>(programs_dev);
  void* args[] = {&a, &prog, &psi, &lam, &partials};
  QSIM_CHECK_HIP(hipModuleLaunchKernel(sp.fn, grid.x, grid.y, unsigned(nbatch), kBT, 1, 1, 0,
                                       reinterpret_cast<hipStream_t>(stream), args, nullptr));
  return QSIM_OK;
}

size_t qsim_blocked_specialized_count(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_specs.size();
}

}  // extern "C"
