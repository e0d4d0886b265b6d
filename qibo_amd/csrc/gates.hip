// Gate passes on a state vector held in HBM, in place, for gfx950.
//
// A gate touching k target bits under c control bits is applied by walking
// the 2^(n-k-c) "base" indices (all target bits 0, all control bits 1) and
// updating the 2^k amplitudes of each base from registers. Base indices are
// consecutive integers with zero bits inserted at the fixed positions, so
// consecutive lanes of a wavefront touch consecutive addresses for every
// target position: for a high target the two halves of a pair are two
// contiguous streams, for a low target a lane's pair shares a 128-byte line
// with its neighbours' pairs and both loads of the line are issued back to
// back, so each line crosses HBM once; targets inside one 128-byte line go
// through an LDS-staged tile instead. Each lane handles one base index with
// streaming (nontemporal) loads and stores, and workgroups are remapped so
// that each XCD walks a contiguous part of the state.
//
// Reference semantics: qibo's Backend.apply_gate (abstract.py:2322-2361) and
// _apply_gate_controlled_by (abstract.py:3176-3197); the matrix row/column
// index has gate.qubits[0] as its most significant bit
// (einsum_utils.permutations, einsum_utils.py:94-110).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "qsim_common.h"

namespace qsim {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

bool validate_qubits(int nqubits, const int* a, int na, const int* b, int nb, const char* what) {
  if (nqubits < 1 || nqubits > QSIM_MAX_QUBITS) {
    set_error(std::string(what) + ": nqubits out of range");
    return false;
  }
  uint64_t seen = 0;
  const int* lists[2] = {a, b};
  const int counts[2] = {na, nb};
  for (int l = 0; l < 2; ++l) {
    if (counts[l] < 0 || (counts[l] > 0 && lists[l] == nullptr)) {
      set_error(std::string(what) + ": bad qubit list");
      return false;
    }
    for (int i = 0; i < counts[l]; ++i) {
      const int q = lists[l][i];
      if (q < 0 || q >= nqubits) {
        set_error(std::string(what) + ": qubit " + std::to_string(q) + " out of range for " +
                  std::to_string(nqubits) + " qubits");
        return false;
      }
      const uint64_t bit = uint64_t(1) << q;
      if (seen & bit) {
        set_error(std::string(what) + ": qubit " + std::to_string(q) + " given twice");
        return false;
      }
      seen |= bit;
    }
  }
  return true;
}

int sorted_positions(int nqubits, const int* a, int na, const int* b, int nb, uint8_t* out) {
  int k = 0;
  for (int i = 0; i < na; ++i) out[k++] = uint8_t(nqubits - 1 - a[i]);
  for (int i = 0; i < nb; ++i) out[k++] = uint8_t(nqubits - 1 - b[i]);
  std::sort(out, out + k);
  return k;
}

static inline uint64_t qmask(int nqubits, int q) { return uint64_t(1) << (nqubits - 1 - q); }

// ---------------------------------------------------------------------------
// |0...0>
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_zero_state(cplx<R>* __restrict__ psi, uint64_t n) {
  const uint64_t i = xcd_block() * 256 + threadIdx.x;
  if (i < n) {
    cplx<R> v;
    v.re = (i == 0) ? R(1) : R(0);
    v.im = R(0);
    st_nt(psi, i, v);
  }
}

// Zeros at the amplitudes whose index has a bit of `outside` set; the others
// are neither read nor written.
template <typename R>
__global__ __launch_bounds__(256) void k_zero_outside(cplx<R>* __restrict__ psi, uint64_t n, uint64_t outside) {
  const uint64_t i = xcd_block() * 256 + threadIdx.x;
  if (i < n && (i & outside) != 0) st_nt(psi, i, cplx<R>{R(0), R(0)});
}

// Collapse onto a measurement outcome: amplitudes whose measured bits equal
// `value` are divided by `norm`, the others set to zero (one pass).
template <typename R>
__global__ __launch_bounds__(256) void k_collapse(cplx<R>* __restrict__ psi, uint64_t n, uint64_t mask,
                                                  uint64_t value, R norm) {
  const uint64_t i = xcd_block() * 256 + threadIdx.x;
  if (i < n) {
    cplx<R> v = ld_nt(psi, i);
    if ((i & mask) == value) {
      v.re = v.re / norm;
      v.im = v.im / norm;
    } else {
      v.re = R(0);
      v.im = R(0);
    }
    st_nt(psi, i, v);
  }
}

constexpr int kThreads = 256;

// Amplitudes per 128-byte line: a target bit below log2 of this pairs
// amplitudes inside one line.
template <typename R>
constexpr int line_bits() {
  return sizeof(R) == 8 ? 3 : 4;
}

// ---------------------------------------------------------------------------
// dense 2^NT x 2^NT on targets, under controls: one base index per lane
// ---------------------------------------------------------------------------
template <typename R, int NT>
struct DenseArgs {
  uint64_t nwork;              // number of base indices
  uint64_t cmask;              // control bits, all set in every base
  uint64_t off[1 << NT];       // offset of matrix index r from the base
  cplx<R> m[(1 << NT) * (1 << NT)];
  int nfixed;                  // controls + targets
  int tbit;                    // bit of the single target (NT == 1, LDS variant)
  uint8_t pos[kMaxFixed];      // their bit positions, ascending
};

template <typename R, int D>
__device__ __forceinline__ cplx<R> row_times(const cplx<R>* m, const cplx<R>* v) {
  cplx<R> acc = cmul(m[0], v[0]);
#pragma clang loop unroll(full)
  for (int c = 1; c < D; ++c) acc = cfma(m[c], v[c], acc);
  return acc;
}

template <typename R, int NT>
__global__ __launch_bounds__(kThreads) void k_dense(const DenseArgs<R, NT> a, cplx<R>* __restrict__ psi) {
  constexpr int D = 1 << NT;
  const uint64_t j = xcd_block() * kThreads + threadIdx.x;
  if (j >= a.nwork) return;
  const uint64_t base = insert_zeros(j, a.pos, a.nfixed) | a.cmask;
  cplx<R> v[D];
#pragma clang loop unroll(full)
  for (int r = 0; r < D; ++r) v[r] = ld_nt(psi, base | a.off[r]);
#pragma clang loop unroll(full)
  for (int r = 0; r < D; ++r) st_nt(psi, base | a.off[r], row_times<R, D>(a.m + r * D, v));
}

// One-qubit gate whose target pairs amplitudes inside a 128-byte line: the
// block stages a contiguous tile through LDS with coalesced streaming loads,
// updates the pairs there (same arithmetic as k_dense) and streams it back.
constexpr int kLdsTileE = 8;                       // amplitudes per lane
constexpr int kLdsTile = kThreads * kLdsTileE;     // 2048 amplitudes per tile
constexpr int kLdsTileBits = 11;

template <typename R>
__global__ __launch_bounds__(kThreads) void k_dense1_lds(const DenseArgs<R, 1> a, cplx<R>* __restrict__ psi) {
  __shared__ cplx<R> t[kLdsTile];
  const uint64_t base = xcd_block() * kLdsTile;
#pragma clang loop unroll(full)
  for (int e = 0; e < kLdsTileE; ++e) t[e * kThreads + threadIdx.x] = ld_nt(psi, base + e * kThreads + threadIdx.x);
  __syncthreads();
  const uint32_t lowmask = (1u << a.tbit) - 1;
#pragma clang loop unroll(full)
  for (int e = 0; e < kLdsTileE / 2; ++e) {
    const uint32_t jj = e * kThreads + threadIdx.x;
    const uint32_t low = jj & lowmask;
    const uint32_t i0 = low | ((jj ^ low) << 1), i1 = i0 | (1u << a.tbit);
    if (((base + i0) & a.cmask) == a.cmask) {
      cplx<R> v[2] = {t[i0], t[i1]};
      t[i0] = row_times<R, 2>(a.m, v);
      t[i1] = row_times<R, 2>(a.m + 2, v);
    }
  }
  __syncthreads();
#pragma clang loop unroll(full)
  for (int e = 0; e < kLdsTileE; ++e) st_nt(psi, base + e * kThreads + threadIdx.x, t[e * kThreads + threadIdx.x]);
}

// ---------------------------------------------------------------------------
// diagonal on targets, under controls
// ---------------------------------------------------------------------------
template <typename R>
struct DiagArgs {
  uint64_t nwork;          // amplitudes walked
  uint64_t cmask;
  cplx<R> d[32];
  int nfixed;              // controls only (subspace walk)
  int nt;                  // number of targets (0..5)
  uint8_t tbit[5];         // bit positions of targets, target 0 = MSB of diag index
  uint8_t pos[kMaxFixed];
};

template <typename R>
__device__ __forceinline__ int diag_index(const DiagArgs<R>& a, uint64_t idx) {
  int r = 0;
  for (int t = 0; t < a.nt; ++t) r = (r << 1) | int((idx >> a.tbit[t]) & 1);
  return r;
}

// Walks only the controlled subspace: used when every control bit is at or
// above the line size, so the walked amplitudes form whole lines.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_diag(const DiagArgs<R> a, cplx<R>* __restrict__ psi) {
  const uint64_t j = xcd_block() * kThreads + threadIdx.x;
  if (j >= a.nwork) return;
  const uint64_t idx = insert_zeros(j, a.pos, a.nfixed) | a.cmask;
  st_nt(psi, idx, cmul(a.d[diag_index(a, idx)], ld_nt(psi, idx)));
}

// Walks every amplitude contiguously and updates the controlled ones: used
// when a control bit lies inside a line, where a subspace walk would fetch
// each line several times through strided accesses.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_diag_full(const DiagArgs<R> a, cplx<R>* __restrict__ psi) {
  const uint64_t idx = xcd_block() * kThreads + threadIdx.x;
  if (idx >= a.nwork) return;
  const cplx<R> v = ld_nt(psi, idx);
  if ((idx & a.cmask) == a.cmask) st_nt(psi, idx, cmul(a.d[diag_index(a, idx)], v));
}

// ---------------------------------------------------------------------------
// swap of two amplitudes per base, under controls
// ---------------------------------------------------------------------------
struct SwapArgs {
  uint64_t nwork;
  uint64_t cmask;
  uint64_t m0, m1;
  int nfixed;
  uint8_t pos[kMaxFixed];
};

template <typename R>
__global__ __launch_bounds__(kThreads) void k_swap(const SwapArgs a, cplx<R>* __restrict__ psi) {
  const uint64_t j = xcd_block() * kThreads + threadIdx.x;
  if (j >= a.nwork) return;
  const uint64_t base = insert_zeros(j, a.pos, a.nfixed) | a.cmask;
  const cplx<R> x = ld_nt(psi, base | a.m0), y = ld_nt(psi, base | a.m1);
  st_nt(psi, base | a.m0, y);
  st_nt(psi, base | a.m1, x);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename R, int NT>
static int launch_dense(void* state, int n, const int* targets, const int* controls, int nc,
                        const void* matrix, hipStream_t s) {
  DenseArgs<R, NT> a;
  std::memset(&a, 0, sizeof(a));
  a.nfixed = sorted_positions(n, controls, nc, targets, NT, a.pos);
  a.nwork = uint64_t(1) << (n - a.nfixed);
  a.cmask = 0;
  for (int i = 0; i < nc; ++i) a.cmask |= qmask(n, controls[i]);
  constexpr int D = 1 << NT;
  for (int r = 0; r < D; ++r) {
    uint64_t off = 0;
    for (int t = 0; t < NT; ++t)
      if ((r >> (NT - 1 - t)) & 1) off |= qmask(n, targets[t]);
    a.off[r] = off;
  }
  std::memcpy(a.m, matrix, sizeof(a.m));
  a.tbit = n - 1 - targets[0];
  if constexpr (NT == 1) {
    if (a.tbit < line_bits<R>() && n >= kLdsTileBits) {
      hipLaunchKernelGGL((k_dense1_lds<R>), launch_grid((uint64_t(1) << n) / kLdsTile), dim3(kThreads), 0, s, a,
                         reinterpret_cast<cplx<R>*>(state));
      QSIM_CHECK_LAUNCH();
      return QSIM_OK;
    }
  }
  hipLaunchKernelGGL((k_dense<R, NT>), launch_grid(ceil_div_u(a.nwork, kThreads)), dim3(kThreads), 0, s, a,
                     reinterpret_cast<cplx<R>*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

template <typename R>
static int launch_diag(void* state, int n, const int* targets, int nt, const int* controls, int nc,
                       const void* diag, hipStream_t s) {
  DiagArgs<R> a;
  std::memset(&a, 0, sizeof(a));
  a.nfixed = sorted_positions(n, controls, nc, nullptr, 0, a.pos);
  a.cmask = 0;
  for (int i = 0; i < nc; ++i) a.cmask |= qmask(n, controls[i]);
  a.nt = nt;
  for (int t = 0; t < nt; ++t) a.tbit[t] = uint8_t(n - 1 - targets[t]);
  std::memcpy(a.d, diag, sizeof(cplx<R>) * (size_t(1) << nt));
  const bool control_in_line = a.nfixed > 0 && a.pos[0] < line_bits<R>();
  if (control_in_line) {
    a.nwork = uint64_t(1) << n;
    hipLaunchKernelGGL((k_diag_full<R>), launch_grid(ceil_div_u(a.nwork, kThreads)), dim3(kThreads), 0, s, a,
                       reinterpret_cast<cplx<R>*>(state));
  } else {
    a.nwork = uint64_t(1) << (n - a.nfixed);
    hipLaunchKernelGGL((k_diag<R>), launch_grid(ceil_div_u(a.nwork, kThreads)), dim3(kThreads), 0, s, a,
                       reinterpret_cast<cplx<R>*>(state));
  }
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

template <typename R>
static int launch_swap(void* state, int n, int q0, int q1, const int* controls, int nc, hipStream_t s) {
  SwapArgs a;
  std::memset(&a, 0, sizeof(a));
  const int t[2] = {q0, q1};
  a.nfixed = sorted_positions(n, controls, nc, t, 2, a.pos);
  a.nwork = uint64_t(1) << (n - a.nfixed);
  a.cmask = 0;
  for (int i = 0; i < nc; ++i) a.cmask |= qmask(n, controls[i]);
  a.m0 = qmask(n, q1);  // |..q0=0..q1=1..>
  a.m1 = qmask(n, q0);  // |..q0=1..q1=0..>
  hipLaunchKernelGGL((k_swap<R>), launch_grid(ceil_div_u(a.nwork, kThreads)), dim3(kThreads), 0, s, a,
                     reinterpret_cast<cplx<R>*>(state));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

extern "C" {

int qsim_version(void) { return QSIM_ABI_VERSION; }

const char* qsim_last_error(void) { return g_last_error.c_str(); }

int qsim_zero_state(void* state, int nqubits, int dtype, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr, "qsim_zero_state: null state");
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= QSIM_MAX_QUBITS, "qsim_zero_state: nqubits out of range");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_zero_state: bad dtype");
  const uint64_t n = uint64_t(1) << nqubits;
  const dim3 blocks = launch_grid(ceil_div_u(n, 256));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_zero_state<double>, blocks, dim3(256), 0, s, (c128*)state, n);
  else
    hipLaunchKernelGGL(k_zero_state<float>, blocks, dim3(256), 0, s, (c64*)state, n);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_zero_outside_support(void* state, int nqubits, int dtype, uint64_t support, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr, "qsim_zero_outside_support: null state");
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= QSIM_MAX_QUBITS, "qsim_zero_outside_support: nqubits out of range");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_zero_outside_support: bad dtype");
  const uint64_t n = uint64_t(1) << nqubits;
  const uint64_t outside = (n - 1) & ~support;
  if (outside == 0) return QSIM_OK;
  const dim3 blocks = launch_grid(ceil_div_u(n, 256));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_zero_outside<double>, blocks, dim3(256), 0, s, (c128*)state, n, outside);
  else
    hipLaunchKernelGGL(k_zero_outside<float>, blocks, dim3(256), 0, s, (c64*)state, n, outside);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_collapse(void* state, int nqubits, int dtype, const int* qubits, int nmeasured, int64_t outcome,
                  double norm, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr, "qsim_collapse: null state");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_collapse: bad dtype");
  QSIM_CHECK_ARG(nmeasured >= 1 && nmeasured <= nqubits && nmeasured <= 62, "qsim_collapse: bad qubit count");
  if (!validate_qubits(nqubits, qubits, nmeasured, nullptr, 0, "qsim_collapse")) return QSIM_EINVAL;
  QSIM_CHECK_ARG(outcome >= 0 && outcome < (int64_t(1) << nmeasured), "qsim_collapse: outcome out of range");
  QSIM_CHECK_ARG(norm > 0.0 && norm < 1e300, "qsim_collapse: norm must be positive and finite");
  // outcome bit j (most significant first) belongs to qubits[j], amplitude bit n-1-q
  uint64_t mask = 0, value = 0;
  for (int j = 0; j < nmeasured; ++j) {
    const uint64_t b = uint64_t(1) << (nqubits - 1 - qubits[j]);
    mask |= b;
    if ((uint64_t(outcome) >> (nmeasured - 1 - j)) & 1) value |= b;
  }
  const uint64_t n = uint64_t(1) << nqubits;
  const dim3 blocks = launch_grid(ceil_div_u(n, 256));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_collapse<double>, blocks, dim3(256), 0, s, (c128*)state, n, mask, value, norm);
  else
    hipLaunchKernelGGL(k_collapse<float>, blocks, dim3(256), 0, s, (c64*)state, n, mask, value, float(norm));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_apply_dense(void* state, int nqubits, int dtype, const int* targets, int ntargets,
                     const int* controls, int ncontrols, const void* matrix, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && matrix != nullptr, "qsim_apply_dense: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_dense: bad dtype");
  QSIM_CHECK_ARG(ntargets >= 1 && ntargets <= 3, "qsim_apply_dense: 1..3 targets supported");
  if (!validate_qubits(nqubits, targets, ntargets, controls, ncontrols, "qsim_apply_dense"))
    return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128) {
    switch (ntargets) {
      case 1: return launch_dense<double, 1>(state, nqubits, targets, controls, ncontrols, matrix, s);
      case 2: return launch_dense<double, 2>(state, nqubits, targets, controls, ncontrols, matrix, s);
      default: return launch_dense<double, 3>(state, nqubits, targets, controls, ncontrols, matrix, s);
    }
  }
  switch (ntargets) {
    case 1: return launch_dense<float, 1>(state, nqubits, targets, controls, ncontrols, matrix, s);
    case 2: return launch_dense<float, 2>(state, nqubits, targets, controls, ncontrols, matrix, s);
    default: return launch_dense<float, 3>(state, nqubits, targets, controls, ncontrols, matrix, s);
  }
}

int qsim_apply_diag(void* state, int nqubits, int dtype, const int* targets, int ntargets,
                    const int* controls, int ncontrols, const void* diag, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && diag != nullptr, "qsim_apply_diag: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_diag: bad dtype");
  QSIM_CHECK_ARG(ntargets >= 0 && ntargets <= 5, "qsim_apply_diag: 0..5 targets supported");
  if (!validate_qubits(nqubits, targets, ntargets, controls, ncontrols, "qsim_apply_diag"))
    return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return launch_diag<double>(state, nqubits, targets, ntargets, controls, ncontrols, diag, s);
  return launch_diag<float>(state, nqubits, targets, ntargets, controls, ncontrols, diag, s);
}

int qsim_apply_swap(void* state, int nqubits, int dtype, int q0, int q1, const int* controls,
                    int ncontrols, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr, "qsim_apply_swap: null state");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_swap: bad dtype");
  const int t[2] = {q0, q1};
  if (!validate_qubits(nqubits, t, 2, controls, ncontrols, "qsim_apply_swap")) return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128) return launch_swap<double>(state, nqubits, q0, q1, controls, ncontrols, s);
  return launch_swap<float>(state, nqubits, q0, q1, controls, ncontrols, s);
}

}  // extern "C"
