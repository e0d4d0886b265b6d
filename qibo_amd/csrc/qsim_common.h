// Shared helpers of the gfx950 state-vector kernels: complex types, error
// reporting for the C ABI, bit-insertion index math and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/qsim_mi355x.h"

namespace qsim {

// ---------------------------------------------------------------------------
// error reporting (thread-local: every entry point is re-entrant)
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define QSIM_CHECK_ARG(cond, msg)              \
  do {                                         \
    if (!(cond)) {                             \
      ::qsim::set_error(std::string(msg));     \
      return QSIM_EINVAL;                      \
    }                                          \
  } while (0)

#define QSIM_CHECK_HIP(expr)                                                       \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      ::qsim::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
      return QSIM_EHIP;                                                            \
    }                                                                              \
  } while (0)

// Launch-error check after a kernel launch.
#define QSIM_CHECK_LAUNCH() QSIM_CHECK_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// complex element types (interleaved re, im; same layout as torch complex)
// ---------------------------------------------------------------------------
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
};

using c128 = cplx<double>;
using c64 = cplx<float>;

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename R>
__device__ __forceinline__ cplx<R> cfma(cplx<R> a, cplx<R> b, cplx<R> acc) {
  // acc + a*b
  acc.re = acc.re + a.re * b.re - a.im * b.im;
  acc.im = acc.im + a.re * b.im + a.im * b.re;
  return acc;
}

// ---------------------------------------------------------------------------
// index math
// ---------------------------------------------------------------------------
constexpr int kMaxFixed = QSIM_MAX_QUBITS;

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

// Host-side validation of a qubit list: every qubit in [0, n), no repeats
// within or across the two lists. Returns false and sets the error otherwise.
bool validate_qubits(int nqubits, const int* a, int na, const int* b, int nb, const char* what);

// Sorted bit positions (bit = n-1-q) of the union of two qubit lists.
int sorted_positions(int nqubits, const int* a, int na, const int* b, int nb, uint8_t* out);

inline int ceil_div_i(uint64_t a, uint64_t b) { return int((a + b - 1) / b); }

}  // namespace qsim
