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

namespace {

// The text of qsim_device.h and blocked_dev.h (written by the build).
#include "jit_sources.inc"

struct RecView {
  uint32_t op, arg;
  bool cond;      // nonzero tile condition: the record runs only where it holds
  uint32_t nent;  // entries (at most kMaxEntries): fixes the next record's offset
  uint32_t off;
};

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

std::string structure_key(const std::vector<RecView>& recs, int dtype, int from_zero) {
  std::string key;
  key.reserve(recs.size() * 4 + 2);
  key.push_back(char('0' + dtype));
  key.push_back(char('0' + from_zero));
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

std::string make_source(const std::vector<RecView>& recs, int dtype, int from_zero) {
  const bool dbl = dtype == QSIM_C128;
  std::string s;
  s.reserve(2048 + recs.size() * 400);
  s += "#include \"blocked_dev.h\"\nnamespace qsim {\n";
  s += "extern \"C\" __global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(QSIM_BLOCKED_WAVES)))\n";
  s += "void qsim_blocked_spec(const BlkArgs a, const uint8_t* __restrict__ prog, cplx<";
  s += dbl ? "double" : "float";
  s += ">* __restrict__ psi) {\n";
  s += dbl ? "  using R = double;\n" : "  using R = float;\n";
  s += R"(  __shared__ R lds[kBTile];
  const uint64_t tile = xcd_block();
  if (tile >= a.ntiles) return;
  psi += blockIdx.z * a.state_stride;
  prog += blockIdx.z * a.prog_stride;
  uint64_t gbit[kBTileBits];
#pragma clang loop unroll(full)
  for (int i = 0; i < kBTileBits; ++i) gbit[i] = uint64_t(1) << a.tbits[i];
  const uint64_t tile_base = insert_zeros(tile, a.tbits, kBTileBits);
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
    s += R"(#pragma clang loop unroll(full)
  for (int r = 0; r < kBReg; ++r) v[r] = ld_nt(psi, lane_idx | deposit(uint32_t(r) << 8, gbit));
)";
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
  int device, dtype, nrec;
  size_t min_bytes;  // the last record's header ends here: a program must be at least this long
};

std::mutex g_mu;
std::vector<Spec> g_specs;                         // handle = index + 1
std::unordered_map<std::string, uint64_t> g_index;  // key -> handle

int prepare(const void* program, size_t program_bytes, int ninstr, int dtype, int from_zero,
            std::vector<RecView>& recs, const char* what) {
  QSIM_CHECK_ARG(program != nullptr || ninstr == 0, what);
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, what);
  QSIM_CHECK_ARG(from_zero == 0 || from_zero == 1, what);
  QSIM_CHECK_ARG(ninstr >= 0 && program_bytes < kMaxProgramBytes, what);
  std::string why;
  if (!decode(reinterpret_cast<const uint8_t*>(program), program_bytes, ninstr, dtype, recs, why)) {
    set_error(std::string(what) + ": " + why);
    return QSIM_EINVAL;
  }
  return QSIM_OK;
}

}  // namespace
}  // namespace qsim

using namespace qsim;

extern "C" {

size_t qsim_blocked_specialized_source(const void* program, size_t program_bytes, int ninstr, int dtype,
                                       int from_zero, char* out, size_t cap) {
  clear_error();
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, from_zero, recs, "qsim_blocked_specialized_source") != QSIM_OK)
    return 0;
  const std::string src = make_source(recs, dtype, from_zero);
  if (out != nullptr && cap > 0) {
    const size_t n = src.size() < cap - 1 ? src.size() : cap - 1;
    std::memcpy(out, src.data(), n);
    out[n] = '\0';
  }
  return src.size() + 1;
}

int qsim_blocked_compile_check(const void* program, size_t program_bytes, int ninstr, int dtype, int from_zero,
                               size_t* code_bytes) {
  clear_error();
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, from_zero, recs, "qsim_blocked_compile_check") != QSIM_OK)
    return QSIM_EINVAL;
  std::vector<char> code;
  std::string log;
  if (compile(make_source(recs, dtype, from_zero), code, log) != 0) {
    set_error("qsim_blocked_compile_check: hiprtc failed: " + log);
    return QSIM_EHIP;
  }
  if (code_bytes != nullptr) *code_bytes = code.size();
  return QSIM_OK;
}

int qsim_blocked_specialize(const void* program, size_t program_bytes, int ninstr, int dtype, int from_zero,
                            uint64_t* handle) {
  clear_error();
  QSIM_CHECK_ARG(handle != nullptr, "qsim_blocked_specialize: null handle");
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, from_zero, recs, "qsim_blocked_specialize") != QSIM_OK)
    return QSIM_EINVAL;
  int device = 0;
  QSIM_CHECK_HIP(hipGetDevice(&device));
  const std::string key = structure_key(recs, dtype, from_zero) + "@" + std::to_string(device);
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
  if (compile(make_source(recs, dtype, from_zero), code, log) != 0) {
    set_error("qsim_blocked_specialize: hiprtc failed: " + log);
    return QSIM_EHIP;
  }
  hipModule_t mod;
  QSIM_CHECK_HIP(hipModuleLoadData(&mod, code.data()));
  hipFunction_t fn;
  const hipError_t e = hipModuleGetFunction(&fn, mod, "qsim_blocked_spec");
  if (e != hipSuccess) {
    hipModuleUnload(mod);
    set_error(std::string("qsim_blocked_specialize: hipModuleGetFunction: ") + hipGetErrorString(e));
    return QSIM_EHIP;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_index.find(key);
  if (it != g_index.end()) {  // another thread finished the same sequence first
    hipModuleUnload(mod);
    *handle = it->second;
    return QSIM_OK;
  }
  const size_t min_bytes = recs.empty() ? 0 : size_t(recs.back().off) + sizeof(BlkHdr);
  g_specs.push_back({key, mod, fn, device, dtype, int(recs.size()), min_bytes});
  *handle = uint64_t(g_specs.size());
  g_index.emplace(key, *handle);
  return QSIM_OK;
}

int qsim_apply_blocked_specialized(uint64_t handle, void* state, int nqubits, int dtype, const int* tile_bits,
                                   const void* program_dev, size_t program_bytes, int ninstr, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr, "qsim_apply_blocked_specialized: null state");
  if (check_blocked_args(nqubits, dtype, tile_bits, program_dev, program_bytes, ninstr,
                         "qsim_apply_blocked_specialized: bad arguments") != QSIM_OK)
    return QSIM_EINVAL;
  Spec sp;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    QSIM_CHECK_ARG(handle >= 1 && handle <= g_specs.size(), "qsim_apply_blocked_specialized: unknown handle");
    sp = g_specs[handle - 1];
  }
  int device = 0;
  QSIM_CHECK_HIP(hipGetDevice(&device));
  QSIM_CHECK_ARG(device == sp.device, "qsim_apply_blocked_specialized: kernel compiled for another device");
  QSIM_CHECK_ARG(dtype == sp.dtype, "qsim_apply_blocked_specialized: kernel compiled for another dtype");
  QSIM_CHECK_ARG(ninstr >= sp.nrec && program_bytes >= sp.min_bytes,
                 "qsim_apply_blocked_specialized: program shorter than the kernel's records");
  BlkArgs a;
  std::memset(&a, 0, sizeof(a));
  a.ntiles = uint64_t(1) << (nqubits - kBTileBits);
  a.nbytes = program_bytes;
  a.state_stride = uint64_t(1) << nqubits;
  a.prog_stride = 0;
  a.ninstr = ninstr;
  for (int i = 0; i < kBTileBits; ++i) a.tbits[i] = uint8_t(tile_bits[i]);
  const dim3 grid = launch_grid(a.ntiles);
  const uint8_t* prog = reinterpret_cast<const uint8_t*>(program_dev);
  void* args[] = {&a, &prog, &state};
  QSIM_CHECK_HIP(hipModuleLaunchKernel(sp.fn, grid.x, grid.y, 1, kBT, 1, 1, 0,
                                       reinterpret_cast<hipStream_t>(stream), args, nullptr));
  return QSIM_OK;
}

int qsim_blocked_specialized_lookup(const void* program, size_t program_bytes, int ninstr, int dtype, int from_zero,
                                    uint64_t* handle, uint64_t* key_hash) {
  clear_error();
  QSIM_CHECK_ARG(handle != nullptr, "qsim_blocked_specialized_lookup: null handle");
  std::vector<RecView> recs;
  if (prepare(program, program_bytes, ninstr, dtype, from_zero, recs, "qsim_blocked_specialized_lookup") != QSIM_OK)
    return QSIM_EINVAL;
  int device = 0;
  QSIM_CHECK_HIP(hipGetDevice(&device));
  const std::string key = structure_key(recs, dtype, from_zero) + "@" + std::to_string(device);
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

size_t qsim_blocked_specialized_count(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_specs.size();
}

}  // extern "C"
