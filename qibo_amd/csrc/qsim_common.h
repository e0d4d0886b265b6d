// Shared helpers of the gfx950 state-vector kernels: complex types, error
// reporting for the C ABI, bit-insertion index math and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/qsim_mi355x.h"
#include "qsim_device.h"

namespace qsim {

static_assert(kMaxFixed == QSIM_MAX_QUBITS, "device-side qubit bound");

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

// host side of the launch geometry: see block_id() / xcd_block() in qsim_device.h
inline dim3 launch_grid(uint64_t nblocks) {
  if (nblocks <= kMaxGridX) return dim3(unsigned(nblocks));
  return dim3(unsigned(kMaxGridX), unsigned((nblocks + kMaxGridX - 1) / kMaxGridX));
}

// Host-side validation of a qubit list: every qubit in [0, n), no repeats
// within or across the two lists. Returns false and sets the error otherwise.
bool validate_qubits(int nqubits, const int* a, int na, const int* b, int nb, const char* what);

// Sorted bit positions (bit = n-1-q) of the union of two qubit lists.
int sorted_positions(int nqubits, const int* a, int na, const int* b, int nb, uint8_t* out);

inline uint64_t ceil_div_u(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace qsim
