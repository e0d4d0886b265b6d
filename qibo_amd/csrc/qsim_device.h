// Device-side helpers shared by the gfx950 kernels: complex types, complex
// products with explicit fused multiply-adds, bit-insertion index math,
// streaming loads and stores, the XCD-aware workgroup order. Nothing here
// needs the host runtime, so the same text also compiles under hiprtc for
// the kernels the library specialises at run time (blocked passes).
#pragma once

#ifdef __HIPCC_RTC__
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
using __hip_internal::uint8_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace qsim {

// ---------------------------------------------------------------------------
// complex element types (interleaved re, im; same layout as torch complex)
// ---------------------------------------------------------------------------
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
};

using c128 = cplx<double>;
using c64 = cplx<float>;

__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Complex products with the fused multiply-adds written out, so that every
// kernel (one gate per pass or many gates per pass) rounds identically and
// the compiler's contraction choices cannot differ between them.
template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  cplx<R> r;
  r.re = fmaR(a.re, b.re, -(a.im * b.im));
  r.im = fmaR(a.re, b.im, a.im * b.re);
  return r;
}

template <typename R>
__device__ __forceinline__ cplx<R> cfma(cplx<R> a, cplx<R> b, cplx<R> acc) {
  // acc + a*b
  acc.re = fmaR(a.re, b.re, fmaR(-a.im, b.im, acc.re));
  acc.im = fmaR(a.re, b.im, fmaR(a.im, b.re, acc.im));
  return acc;
}

// ---------------------------------------------------------------------------
// index math
// ---------------------------------------------------------------------------
constexpr int kMaxFixed = 36;  // QSIM_MAX_QUBITS (checked in qsim_common.h)

// Insert a zero bit at each position of `pos` (ascending) into j. The first
// kUnrolledFixed positions are read with constant indices (scalar loads from
// the kernel arguments, wave-uniform branches); longer lists (many controls)
// continue in a loop.
constexpr int kUnrolledFixed = 8;

__device__ __forceinline__ uint64_t insert_zeros(uint64_t j, const uint8_t* pos, int npos) {
#pragma clang loop unroll(full)
  for (int f = 0; f < kUnrolledFixed; ++f) {
    if (f < npos) {
      const uint64_t low = j & ((uint64_t(1) << pos[f]) - 1);
      j = low | ((j ^ low) << 1);
    }
  }
  for (int f = kUnrolledFixed; f < npos; ++f) {
    const uint64_t low = j & ((uint64_t(1) << pos[f]) - 1);
    j = low | ((j ^ low) << 1);
  }
  return j;
}

// Load / store one amplitude as a single 8- or 16-byte access.
template <typename R>
__device__ __forceinline__ cplx<R> ld(const cplx<R>* p, uint64_t i) {
  const R* q = reinterpret_cast<const R*>(p + i);
  cplx<R> v;
  v.re = q[0];
  v.im = q[1];
  return v;
}

template <typename R>
__device__ __forceinline__ void st(cplx<R>* p, uint64_t i, cplx<R> v) {
  R* q = reinterpret_cast<R*>(p + i);
  q[0] = v.re;
  q[1] = v.im;
}

// Streaming (nontemporal) variants: a gate pass touches each line of the
// state once, so keeping it in L2/MALL only evicts useful data. Measured on
// MI355X: an in-place streaming pass runs at 6.6 TB/s with these against
// 5.9 TB/s with ordinary loads and stores.
template <typename R>
__device__ __forceinline__ cplx<R> ld_nt(const cplx<R>* p, uint64_t i) {
  const R* q = reinterpret_cast<const R*>(p + i);
  cplx<R> v;
  v.re = __builtin_nontemporal_load(q);
  v.im = __builtin_nontemporal_load(q + 1);
  return v;
}

template <typename R>
__device__ __forceinline__ void st_nt(cplx<R>* p, uint64_t i, cplx<R> v) {
  R* q = reinterpret_cast<R*>(p + i);
  __builtin_nontemporal_store(v.re, q);
  __builtin_nontemporal_store(v.im, q + 1);
}

// Workgroup ids are dealt round-robin to the 8 XCDs of an MI355X. Remapping
// them so that each XCD walks one contiguous eighth of the grid keeps every
// XCD's requests on a compact address range (measured: 6.3 vs 5.9 TB/s for a
// streaming pass). Blocks past the last full multiple of 8 keep their id.
// HIP bounds each grid dimension at 2^32 - 1 work-items, so launches of more
// than kMaxGridX workgroups use a 2-D grid of rows of kMaxGridX workgroups;
// block_id() is the linear id blockIdx.y * gridDim.x + blockIdx.x (the
// dispatch order). kMaxGridX x 1024 threads stays below 2^32.
constexpr uint64_t kMaxGridX = uint64_t(1) << 21;
__device__ __forceinline__ uint64_t block_id() { return uint64_t(blockIdx.y) * gridDim.x + blockIdx.x; }

// Linear workgroup id with the XCD remap applied within each grid row (a row
// of kMaxGridX workgroups is a multiple of 8, so the round-robin holds per row).
__device__ __forceinline__ uint64_t xcd_block() {
  const uint32_t b = blockIdx.x, per = gridDim.x / 8;
  const uint64_t x = (b < per * 8) ? uint64_t(b % 8) * per + b / 8 : uint64_t(b);
  return uint64_t(blockIdx.y) * gridDim.x + x;
}

}  // namespace qsim
