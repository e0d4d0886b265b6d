// Expectation values of Pauli sums, Hamiltonian-vector products and
// <bra|M|ket> inner products for gfx950, all accumulated in FP64.
//
// A Pauli string is i^{ny} X^x Z^z with bit masks x, z over amplitude
// indices (Y = i X Z). Then
//   (P psi)_i = i^{ny} (-1)^{popcount((i ^ x) & z)} psi_{i ^ x}
// and <psi|P|psi> pairs amplitude i0 (bit h of x clear, h = top bit of x) with
// i1 = i0 ^ x: A = conj(psi_i0) psi_i1 contributes
//   Re( i^{ny} ( (-1)^{|i1 & z|} A + (-1)^{|i0 & z|} conj(A) ) ).
// Strings with equal x share one pass over the state; x = 0 strings reduce
// |psi_i|^2 with their signs in one pass. This replaces the per-term einsum
// of Backend.exp_value_observable_symbolic (abstract.py:2946-3054).
//
// Reductions use a fixed grid and fixed-order trees, so results do not
// depend on scheduling (a request gives the same value alone or in a mix).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "qsim_common.h"

namespace qsim {

constexpr int kET = 256;
constexpr int kEBlocks = 1024;  // fixed grid of every reduction pass
constexpr int kMaxGroupTerms = 32;

struct PauliGroupArgs {
  uint64_t nwork;
  uint64_t x;      // flip mask of the group (0 = diagonal)
  uint64_t hbit;   // top bit of x (unused when x = 0)
  int hpos;
  int nterms;
  uint64_t z[kMaxGroupTerms];
  double c[kMaxGroupTerms];
  int ny[kMaxGroupTerms];
};

template <typename R>
__device__ __forceinline__ double sq(cplx<R> v) {
  return double(v.re) * double(v.re) + double(v.im) * double(v.im);
}

__device__ double block_sum(double v, double* lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int off = kET / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  return lds[0];
}

template <typename R>
__global__ __launch_bounds__(kET) void k_pauli_group(const cplx<R>* __restrict__ psi, const PauliGroupArgs a,
                                                     double* __restrict__ partial) {
  __shared__ double lds[kET];
  double acc = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * kET;
  if (a.x == 0) {
    for (uint64_t i = uint64_t(blockIdx.x) * kET + threadIdx.x; i < a.nwork; i += stride) {
      const double p = sq(psi[i]);
      double w = 0.0;
      for (int t = 0; t < a.nterms; ++t) w += (__popcll(i & a.z[t]) & 1) ? -a.c[t] : a.c[t];
      acc += w * p;
    }
  } else {
    for (uint64_t j = uint64_t(blockIdx.x) * kET + threadIdx.x; j < a.nwork; j += stride) {
      const uint64_t low = j & (a.hbit - 1);
      const uint64_t i0 = low | ((j ^ low) << 1);
      const uint64_t i1 = i0 ^ a.x;
      const cplx<R> u = psi[i0], v = psi[i1];
      // A = conj(u) * v
      const double are = double(u.re) * double(v.re) + double(u.im) * double(v.im);
      const double aim = double(u.re) * double(v.im) - double(u.im) * double(v.re);
      for (int t = 0; t < a.nterms; ++t) {
        const double s1 = (__popcll(i1 & a.z[t]) & 1) ? -1.0 : 1.0;
        const double s0 = (__popcll(i0 & a.z[t]) & 1) ? -1.0 : 1.0;
        // w = s1 A + s0 conj(A)
        const double wre = (s1 + s0) * are;
        const double wim = (s1 - s0) * aim;
        // Re(i^ny w)
        double re;
        switch (a.ny[t] & 3) {
          case 0: re = wre; break;
          case 1: re = -wim; break;
          case 2: re = -wre; break;
          default: re = wim; break;
        }
        acc += a.c[t] * re;
      }
    }
  }
  const double tot = block_sum(acc, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kET) void k_sum_partials(const double* __restrict__ partial, int64_t count,
                                                      double* __restrict__ out) {
  __shared__ double lds[kET];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += kET) acc += partial[i];
  const double tot = block_sum(acc, lds);
  if (threadIdx.x == 0) out[0] = tot;
}

static int grid_for(uint64_t nwork) {
  const uint64_t b = (nwork + kET - 1) / kET;
  return int(b < kEBlocks ? (b < 1 ? 1 : b) : kEBlocks);
}

// The launches of a term table, in the order every entry point uses: the
// distinct x masks in ascending order, each split into chunks of at most
// kMaxGroupTerms terms in table order. Found by repeated minimum scans, so no
// container is allocated (the no-allocation entry point shares it).
template <typename F>
static void for_each_group(int nterms, const uint64_t* xm, const uint64_t* zm, const int* ny, const double* c,
                           F&& visit) {
  bool have = false;
  uint64_t cur = 0;
  for (int t = 0; t < nterms; ++t)
    if (!have || xm[t] < cur) cur = xm[t], have = true;
  while (have) {
    PauliGroupArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = cur;
    for (int t = 0; t < nterms; ++t) {
      if (xm[t] != cur) continue;
      a.z[a.nterms] = zm[t];
      a.c[a.nterms] = c[t];
      a.ny[a.nterms] = ny[t];
      if (++a.nterms == kMaxGroupTerms) {
        visit(a);
        a.nterms = 0;
      }
    }
    if (a.nterms > 0) visit(a);
    bool next = false;
    uint64_t nxt = 0;
    for (int t = 0; t < nterms; ++t)
      if (xm[t] > cur && (!next || xm[t] < nxt)) nxt = xm[t], next = true;
    have = next;
    cur = nxt;
  }
}

// Launches the group passes into partial[0 ..) and the final fixed-order sum;
// the partial buffer must hold ngroups * kEBlocks doubles (checked first).
template <typename R>
static int expval_impl(const void* state, int n, int nterms, const uint64_t* xm, const uint64_t* zm,
                       const int* ny, const double* c, double* out, void* ws, size_t ws_bytes, hipStream_t s,
                       const char* who) {
  size_t ngroups = 0;
  for_each_group(nterms, xm, zm, ny, c, [&](const PauliGroupArgs&) { ++ngroups; });
  if (ngroups * kEBlocks * sizeof(double) > ws_bytes) {
    set_error(std::string(who) + ": workspace too small for " + std::to_string(ngroups) +
              " term groups (distinct X masks, 32 terms per group)");
    return QSIM_EINVAL;
  }
  double* partial = reinterpret_cast<double*>(ws);
  int64_t used = 0;
  int rc = QSIM_OK;
  for_each_group(nterms, xm, zm, ny, c, [&](const PauliGroupArgs& g) {
    if (rc != QSIM_OK) return;
    PauliGroupArgs a = g;
    if (a.x == 0) {
      a.nwork = uint64_t(1) << n;
    } else {
      a.hpos = 63 - __builtin_clzll(a.x);
      a.hbit = uint64_t(1) << a.hpos;
      a.nwork = uint64_t(1) << (n - 1);
    }
    const int blocks = grid_for(a.nwork);
    hipLaunchKernelGGL(k_pauli_group<R>, dim3(blocks), dim3(kET), 0, s, reinterpret_cast<const cplx<R>*>(state), a,
                       partial + used);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string(who) + ": launch failed: " + hipGetErrorString(e));
      rc = QSIM_EHIP;
      return;
    }
    used += blocks;
  });
  if (rc != QSIM_OK) return rc;
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kET), 0, s, partial, used, out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

// ---------------------------------------------------------------------------
// out = sum_t c_t P_t psi (gather form, one output amplitude per thread)
// ---------------------------------------------------------------------------
struct PauliApplyArgs {
  uint64_t n;
  int nterms;
  uint64_t x[64];
  uint64_t z[64];
  double c[64];
  int ny[64];
};

template <typename R>
__global__ __launch_bounds__(kET) void k_pauli_apply(const cplx<R>* __restrict__ psi, cplx<R>* __restrict__ out,
                                                     const PauliApplyArgs a, int accumulate) {
  const uint64_t i = block_id() * kET + threadIdx.x;
  if (i >= a.n) return;
  double re = 0.0, im = 0.0;
  if (accumulate) {
    re = double(out[i].re);
    im = double(out[i].im);
  }
  for (int t = 0; t < a.nterms; ++t) {
    const uint64_t j = i ^ a.x[t];
    const cplx<R> v = psi[j];
    const double sgn = (__popcll(j & a.z[t]) & 1) ? -a.c[t] : a.c[t];
    double vr = sgn * double(v.re), vi = sgn * double(v.im);
    switch (a.ny[t] & 3) {  // multiply by i^ny
      case 0: break;
      case 1: { const double tr = -vi; vi = vr; vr = tr; } break;
      case 2: vr = -vr; vi = -vi; break;
      default: { const double tr = vi; vi = -vr; vr = tr; } break;
    }
    re += vr;
    im += vi;
  }
  out[i].re = R(re);
  out[i].im = R(im);
}

// ---------------------------------------------------------------------------
// Re <bra| M |ket> on targets under controls
// ---------------------------------------------------------------------------
template <typename R, int NT>
struct InnerArgs {
  uint64_t nwork;
  uint64_t cmask;
  uint64_t off[1 << NT];
  cplx<R> m[(1 << NT) * (1 << NT)];
  int nfixed;
  uint8_t pos[kMaxFixed];
};

template <typename R, int NT>
__global__ __launch_bounds__(kET) void k_inner(const cplx<R>* __restrict__ bra, const cplx<R>* __restrict__ ket,
                                               const InnerArgs<R, NT> a, double* __restrict__ partial) {
  __shared__ double lds[kET];
  constexpr int D = 1 << NT;
  double acc = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * kET;
  for (uint64_t j = uint64_t(blockIdx.x) * kET + threadIdx.x; j < a.nwork; j += stride) {
    const uint64_t base = insert_zeros(j, a.pos, a.nfixed) | a.cmask;
    cplx<R> k[D];
#pragma unroll
    for (int c = 0; c < D; ++c) k[c] = ket[base | a.off[c]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double yr = 0.0, yi = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        yr += double(a.m[r * D + c].re) * double(k[c].re) - double(a.m[r * D + c].im) * double(k[c].im);
        yi += double(a.m[r * D + c].re) * double(k[c].im) + double(a.m[r * D + c].im) * double(k[c].re);
      }
      const cplx<R> b = bra[base | a.off[r]];
      acc += double(b.re) * yr + double(b.im) * yi;  // Re(conj(b) y)
    }
  }
  const double tot = block_sum(acc, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <typename R, int NT>
static int launch_inner(const void* bra, const void* ket, int n, const int* targets, const int* controls, int nc,
                        const void* matrix, double* out, double* partial, hipStream_t s) {
  InnerArgs<R, NT> a;
  std::memset(&a, 0, sizeof(a));
  a.nfixed = sorted_positions(n, controls, nc, targets, NT, a.pos);
  a.nwork = uint64_t(1) << (n - a.nfixed);
  for (int i = 0; i < nc; ++i) a.cmask |= uint64_t(1) << (n - 1 - controls[i]);
  for (int r = 0; r < (1 << NT); ++r) {
    uint64_t off = 0;
    for (int t = 0; t < NT; ++t)
      if ((r >> (NT - 1 - t)) & 1) off |= uint64_t(1) << (n - 1 - targets[t]);
    a.off[r] = off;
  }
  std::memcpy(a.m, matrix, sizeof(a.m));
  const int blocks = grid_for(a.nwork);
  hipLaunchKernelGGL((k_inner<R, NT>), dim3(blocks), dim3(kET), 0, s, reinterpret_cast<const cplx<R>*>(bra),
                     reinterpret_cast<const cplx<R>*>(ket), a, partial);
  QSIM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kET), 0, s, partial, int64_t(blocks), out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

// ---------------------------------------------------------------------------
// Scratch for callers that cannot allocate (qsim_expval_pauli_noalloc): the
// per-block partial sums of up to kNoAllocGroups term groups in a file-scope
// device array. Each GPU holds its own copy (512 KiB) until the process ends;
// its address on a device comes from hipGetSymbolAddress, taken once per
// device with that device current. The array's name in host code is only a
// handle for that call, never a device address.
// ---------------------------------------------------------------------------
constexpr int kNoAllocGroups = 64;
constexpr int kNoAllocDevices = 64;
__device__ __align__(256) double scratch[kNoAllocGroups * kEBlocks];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kNoAllocDevices];  // per device, set once under g_scratch_mu

static int scratch_on_current_device(void* stream, double** out, const char* who) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error(std::string(who) + ": hipGetDevice: " + hipGetErrorString(e));
    return QSIM_EHIP;
  }
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error(std::string(who) + ": device ordinal out of range");
    return QSIM_EINVAL;
  }
  if (stream != nullptr) {
    hipDevice_t sdev = -1;
    e = hipStreamGetDevice(reinterpret_cast<hipStream_t>(stream), &sdev);
    if (e != hipSuccess) {
      set_error(std::string(who) + ": hipStreamGetDevice: " + hipGetErrorString(e));
      return QSIM_EHIP;
    }
    if (int(sdev) != dev) {
      set_error(std::string(who) + ": the stream belongs to another device than the current one");
      return QSIM_EINVAL;
    }
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string(who) + ": hipGetSymbolAddress(scratch): " + hipGetErrorString(e));
      return QSIM_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = reinterpret_cast<double*>(g_scratch_ptr[dev]);
  return QSIM_OK;
}

}  // namespace qsim

using namespace qsim;

extern "C" {

size_t qsim_expval_workspace(int nqubits, int nterms) {
  (void)nqubits;
  if (nterms < 1) nterms = 1;
  return size_t(nterms) * kEBlocks * sizeof(double);
}

int qsim_expval_pauli(const void* state, int nqubits, int dtype, int nterms, const uint64_t* xmasks,
                      const uint64_t* zmasks, const int* ny, const double* coeffs, double* out, void* workspace,
                      size_t workspace_bytes, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && out != nullptr && workspace != nullptr, "qsim_expval_pauli: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_expval_pauli: bad dtype");
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= QSIM_MAX_QUBITS, "qsim_expval_pauli: nqubits out of range");
  QSIM_CHECK_ARG(nterms >= 1 && xmasks && zmasks && ny && coeffs, "qsim_expval_pauli: bad term table");
  const uint64_t full = (nqubits == 64) ? ~uint64_t(0) : ((uint64_t(1) << nqubits) - 1);
  for (int t = 0; t < nterms; ++t)
    QSIM_CHECK_ARG((xmasks[t] & ~full) == 0 && (zmasks[t] & ~full) == 0, "qsim_expval_pauli: mask out of range");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (dtype == QSIM_C128)
    return expval_impl<double>(state, nqubits, nterms, xmasks, zmasks, ny, coeffs, out, workspace, workspace_bytes, s,
                               "qsim_expval_pauli");
  return expval_impl<float>(state, nqubits, nterms, xmasks, zmasks, ny, coeffs, out, workspace, workspace_bytes, s,
                            "qsim_expval_pauli");
}

int qsim_expval_pauli_noalloc(const void* state, int nqubits, int dtype, int nterms, const uint64_t* xmasks,
                              const uint64_t* zmasks, const int* ny, const double* coeffs, double* out,
                              void* stream) {
  clear_error();
  const char* who = "qsim_expval_pauli_noalloc";
  QSIM_CHECK_ARG(state != nullptr && out != nullptr, "qsim_expval_pauli_noalloc: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_expval_pauli_noalloc: bad dtype");
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= QSIM_MAX_QUBITS, "qsim_expval_pauli_noalloc: nqubits out of range");
  QSIM_CHECK_ARG(nterms >= 1 && xmasks && zmasks && ny && coeffs, "qsim_expval_pauli_noalloc: bad term table");
  const uint64_t full = (uint64_t(1) << nqubits) - 1;
  for (int t = 0; t < nterms; ++t)
    QSIM_CHECK_ARG((xmasks[t] & ~full) == 0 && (zmasks[t] & ~full) == 0,
                   "qsim_expval_pauli_noalloc: mask out of range");
  size_t ngroups = 0;
  for_each_group(nterms, xmasks, zmasks, ny, coeffs, [&](const PauliGroupArgs&) { ++ngroups; });
  if (ngroups > size_t(kNoAllocGroups)) {
    set_error("qsim_expval_pauli_noalloc: " + std::to_string(ngroups) + " term groups (distinct X masks, 32 terms " +
              "per group) exceed the maximum of " + std::to_string(kNoAllocGroups) +
              "; use qsim_expval_pauli with a workspace");
    return QSIM_EINVAL;
  }
  double* ptr = nullptr;
  const int rc = scratch_on_current_device(stream, &ptr, who);
  if (rc != QSIM_OK) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const size_t bytes = sizeof(double) * size_t(kNoAllocGroups) * size_t(kEBlocks);
  if (dtype == QSIM_C128)
    return expval_impl<double>(state, nqubits, nterms, xmasks, zmasks, ny, coeffs, out, ptr, bytes, s, who);
  return expval_impl<float>(state, nqubits, nterms, xmasks, zmasks, ny, coeffs, out, ptr, bytes, s, who);
}

int qsim_expval_noalloc_max_groups(void) { return kNoAllocGroups; }

int qsim_apply_pauli_sum(const void* state, int nqubits, int dtype, int nterms, const uint64_t* xmasks,
                         const uint64_t* zmasks, const int* ny, const double* coeffs, void* out, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(state != nullptr && out != nullptr && state != out, "qsim_apply_pauli_sum: bad pointers");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_apply_pauli_sum: bad dtype");
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= QSIM_MAX_QUBITS, "qsim_apply_pauli_sum: nqubits out of range");
  QSIM_CHECK_ARG(nterms >= 1 && xmasks && zmasks && ny && coeffs, "qsim_apply_pauli_sum: bad term table");
  const uint64_t full = (uint64_t(1) << nqubits) - 1;
  for (int t = 0; t < nterms; ++t)
    QSIM_CHECK_ARG((xmasks[t] & ~full) == 0 && (zmasks[t] & ~full) == 0, "qsim_apply_pauli_sum: mask out of range");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const uint64_t N = uint64_t(1) << nqubits;
  for (int t0 = 0; t0 < nterms; t0 += 64) {
    PauliApplyArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = N;
    a.nterms = std::min(64, nterms - t0);
    for (int t = 0; t < a.nterms; ++t) {
      a.x[t] = xmasks[t0 + t];
      a.z[t] = zmasks[t0 + t];
      a.c[t] = coeffs[t0 + t];
      a.ny[t] = ny[t0 + t];
    }
    const int acc = t0 > 0 ? 1 : 0;
    if (dtype == QSIM_C128)
      hipLaunchKernelGGL(k_pauli_apply<double>, launch_grid(ceil_div_u(N, kET)), dim3(kET), 0, s, (const c128*)state,
                         (c128*)out, a, acc);
    else
      hipLaunchKernelGGL(k_pauli_apply<float>, launch_grid(ceil_div_u(N, kET)), dim3(kET), 0, s, (const c64*)state,
                         (c64*)out, a, acc);
    QSIM_CHECK_LAUNCH();
  }
  return QSIM_OK;
}

int qsim_inner_matrix(const void* bra, const void* ket, int nqubits, int dtype, const int* targets, int ntargets,
                      const int* controls, int ncontrols, const void* matrix, double* out, void* workspace,
                      size_t workspace_bytes, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(bra && ket && matrix && out && workspace, "qsim_inner_matrix: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_inner_matrix: bad dtype");
  QSIM_CHECK_ARG(ntargets >= 1 && ntargets <= 3, "qsim_inner_matrix: 1..3 targets supported");
  QSIM_CHECK_ARG(workspace_bytes >= kEBlocks * sizeof(double), "qsim_inner_matrix: workspace too small");
  if (!validate_qubits(nqubits, targets, ntargets, controls, ncontrols, "qsim_inner_matrix")) return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double* partial = reinterpret_cast<double*>(workspace);
  if (dtype == QSIM_C128) {
    switch (ntargets) {
      case 1: return launch_inner<double, 1>(bra, ket, nqubits, targets, controls, ncontrols, matrix, out, partial, s);
      case 2: return launch_inner<double, 2>(bra, ket, nqubits, targets, controls, ncontrols, matrix, out, partial, s);
      default: return launch_inner<double, 3>(bra, ket, nqubits, targets, controls, ncontrols, matrix, out, partial, s);
    }
  }
  switch (ntargets) {
    case 1: return launch_inner<float, 1>(bra, ket, nqubits, targets, controls, ncontrols, matrix, out, partial, s);
    case 2: return launch_inner<float, 2>(bra, ket, nqubits, targets, controls, ncontrols, matrix, out, partial, s);
    default: return launch_inner<float, 3>(bra, ket, nqubits, targets, controls, ncontrols, matrix, out, partial, s);
  }
}

}  // extern "C"
