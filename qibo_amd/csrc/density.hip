// Density-matrix kernels.
//
// A density matrix rho of n qubits is stored row-major as 2^(2n) interleaved
// complex values, element (r, c) at r * 2^n + c: it is the state vector of a
// 2n-qubit register whose qubit q is row qubit q and whose qubit n + q is
// column qubit q. Unitary gates therefore need no kernel of their own (U on
// the row qubits, conj(U) on the column qubits, through the state-vector
// kernels); this file holds what is specific to matrices:
//
//   qsim_dm_axpy           y = a x, y = y + a x or y = x / a (channel sums, means)
//   qsim_dm_depolarize     (1 - lam) rho + lam Tr_T(rho) (x) I / 2^k
//   qsim_dm_reset          (1 - p0 - p1) rho + Tr_q(rho) (x) (p0 |0><0| + p1 |1><1|)
//   qsim_dm_probabilities  |sum of diagonal entries| per outcome of the measured qubits
//   qsim_dm_expval_pauli   sum_t c_t Re Tr(P_t rho)
//
// The arithmetic follows numpy's evaluation of the reference formulas
// (abstract.py:2095-2131, 2212-2232, 2297-2311, 2734-2748): each product is
// rounded before the sum it enters (built with -ffp-contract=off), numpy's
// complex abs for probabilities.

#include <hip/hip_runtime.h>

#include <cstring>

#include "qsim_common.h"

namespace qsim {
namespace {

constexpr int kT = 256;

// mode 0: y = a x; 1: y = y + a x; 2: y = x / a
template <typename R>
__global__ __launch_bounds__(kT) void k_axpy(cplx<R>* __restrict__ y, const cplx<R>* __restrict__ x, uint64_t n,
                                             R a, int mode) {
  const uint64_t i = xcd_block() * kT + threadIdx.x;
  if (i >= n) return;
  const cplx<R> v = ld_nt(x, i);
  cplx<R> o;
  if (mode == 2) {
    o.re = v.re / a;
    o.im = v.im / a;
  } else if (mode == 1) {
    const cplx<R> w = ld_nt(y, i);
    const R pr = a * v.re, pi = a * v.im;
    o.re = w.re + pr;
    o.im = w.im + pi;
  } else {
    o.re = a * v.re;
    o.im = a * v.im;
  }
  st_nt(y, i, o);
}

// Groups of the elements that share their non-target row and column bits:
// a group index is deposited into the free bits of the flat index.
struct GroupArgs {
  uint64_t ngroups;   // 2^(2n - 2k)
  uint8_t fixed[64];  // sorted positions of the 2k target bits
  uint8_t tr[8];      // row position of target j (j = 0 is the most significant target bit)
  uint8_t tc[8];
  int nfixed;
  int k;
};

__device__ __forceinline__ uint64_t spread(uint32_t t, const uint8_t* pos, int k) {
  // bit j of t (j = 0 least significant) goes to position pos[k - 1 - j]
  uint64_t out = 0;
  for (int j = 0; j < k; ++j)
    if ((t >> j) & 1u) out |= uint64_t(1) << pos[k - 1 - j];
  return out;
}

// Tr_T(rho) (x) I / 2^k mixed in with weight lam; one thread per (row, column)
// pattern of the non-target bits. The trace sums the 2^k diagonal target
// entries in increasing order of the target index (targets sorted, first =
// most significant), as numpy's einsum over the transposed matrix.
template <typename R>
__global__ __launch_bounds__(kT) void k_depolarize(cplx<R>* __restrict__ rho, GroupArgs a, R keep, R lam,
                                                   R inv_dim) {
  const uint64_t g = xcd_block() * kT + threadIdx.x;
  if (g >= a.ngroups) return;
  const uint64_t base = insert_zeros(g, a.fixed, a.nfixed);
  const uint32_t dim = 1u << a.k;
  R tre = 0, tim = 0;
  for (uint32_t t = 0; t < dim; ++t) {
    const uint64_t i = base | spread(t, a.tr, a.k) | spread(t, a.tc, a.k);
    const cplx<R> v = rho[i];
    tre = tre + v.re;
    tim = tim + v.im;
  }
  const R ire = tre * inv_dim, iim = tim * inv_dim;
  const R lre = lam * ire, lim = lam * iim;
  for (uint32_t tr = 0; tr < dim; ++tr) {
    const uint64_t ir = base | spread(tr, a.tr, a.k);
    for (uint32_t tc = 0; tc < dim; ++tc) {
      const uint64_t i = ir | spread(tc, a.tc, a.k);
      cplx<R> v = rho[i];
      const R kr = keep * v.re, ki = keep * v.im;
      if (tr == tc) {
        v.re = kr + lre;
        v.im = ki + lim;
      } else {
        v.re = kr + R(0);
        v.im = ki + R(0);
      }
      rho[i] = v;
    }
  }
}

// Reset of qubit with row bit br and column bit bc: one thread per 2x2 block.
template <typename R>
__global__ __launch_bounds__(kT) void k_reset(cplx<R>* __restrict__ rho, GroupArgs a, R keep, R p0, R p1) {
  const uint64_t g = xcd_block() * kT + threadIdx.x;
  if (g >= a.ngroups) return;
  const uint64_t base = insert_zeros(g, a.fixed, a.nfixed);
  const uint64_t br = uint64_t(1) << a.tr[0], bc = uint64_t(1) << a.tc[0];
  const cplx<R> r00 = rho[base], r01 = rho[base | bc], r10 = rho[base | br], r11 = rho[base | br | bc];
  const R tre = r00.re + r11.re, tim = r00.im + r11.im;
  // state = keep * rho + p0 * zero, then + p1 * (X zero X), with zero = T |0><0|
  cplx<R> o00, o01, o10, o11;
  o00.re = (keep * r00.re + p0 * tre) + R(0);
  o00.im = (keep * r00.im + p0 * tim) + R(0);
  o01.re = (keep * r01.re + R(0)) + R(0);
  o01.im = (keep * r01.im + R(0)) + R(0);
  o10.re = (keep * r10.re + R(0)) + R(0);
  o10.im = (keep * r10.im + R(0)) + R(0);
  o11.re = (keep * r11.re + R(0)) + p1 * tre;
  o11.im = (keep * r11.im + R(0)) + p1 * tim;
  rho[base] = o00;
  rho[base | bc] = o01;
  rho[base | br] = o10;
  rho[base | br | bc] = o11;
}

// numpy's complex abs (loops_unary_complex: larger * sqrt(fma(r, r, 1)),
// r = smaller / larger, r = 0 when larger == 0 or smaller == inf).
__device__ __forceinline__ double np_abs(double zr, double zi) {
  double re = fabs(zr), im = fabs(zi);
  const double inf = __builtin_inf();
  if (re == inf) im = inf;
  if (im == inf) re = inf;
  if (re != re) im = re;
  if (im != im) re = im;
  const double l = fmax(re, im);
  const double sm = fmin(im, re);
  const double r = (l == 0.0 || sm == inf) ? 0.0 : sm / l;
  return __builtin_sqrt(__builtin_fma(r, r, 1.0)) * l;
}
__device__ __forceinline__ float np_abs(float zr, float zi) {
  float re = fabsf(zr), im = fabsf(zi);
  const float inf = __builtin_inff();
  if (re == inf) im = inf;
  if (im == inf) re = inf;
  if (re != re) im = re;
  if (im != im) re = im;
  const float l = fmaxf(re, im);
  const float sm = fminf(im, re);
  const float r = (l == 0.0f || sm == inf) ? 0.0f : sm / l;
  return __builtin_sqrtf(__builtin_fmaf(r, r, 1.0f)) * l;
}

struct ProbArgs {
  uint64_t nout;       // 2^m
  uint64_t nsum;       // 2^(n - m)
  uint64_t stride;     // 2^n + 1: flat distance between consecutive diagonal entries
  uint8_t mbit[64];    // bit (of the row index) of output bit j, j = 0 most significant
  uint8_t upos[64];    // sorted bit positions of the measured qubits (for insert_zeros)
  int m, n;
};

// One thread per outcome: the complex sum of its diagonal entries over the
// unmeasured qubits in increasing index order, then numpy's abs.
template <typename R>
__global__ __launch_bounds__(kT) void k_dm_probs(const cplx<R>* __restrict__ rho, ProbArgs a, R* __restrict__ out) {
  const uint64_t o = xcd_block() * kT + threadIdx.x;
  if (o >= a.nout) return;
  uint64_t fixed = 0;
  for (int j = 0; j < a.m; ++j)
    if ((o >> (a.m - 1 - j)) & 1) fixed |= uint64_t(1) << a.mbit[j];
  R sre = 0, sim = 0;
  for (uint64_t b = 0; b < a.nsum; ++b) {
    const uint64_t r = insert_zeros(b, a.upos, a.m) | fixed;
    const cplx<R> v = rho[r * a.stride];
    sre = sre + v.re;
    sim = sim + v.im;
  }
  out[o] = np_abs(sre, sim);
}

// sum over i of i^ny (-1)^popc(i & z) rho[i, i ^ x] for one Pauli term per
// workgroup (up to 64 terms per launch); FP64 accumulation, one value per term.
struct DmPauliArgs {
  uint64_t x[64], z[64];
  double c[64];
  int ny[64];
  int nterms;
};

template <typename R>
__global__ __launch_bounds__(kT) void k_dm_pauli(const cplx<R>* __restrict__ rho, int n, DmPauliArgs a,
                                                 double* __restrict__ terms) {
  __shared__ double red[kT];
  const int t = blockIdx.x;
  const uint64_t x = a.x[t], z = a.z[t], dim = uint64_t(1) << n;
  double sre = 0.0, sim = 0.0;
  for (uint64_t i = threadIdx.x; i < dim; i += kT) {
    const cplx<R> v = rho[(i << n) | (i ^ x)];
    const double s = (__builtin_popcountll(i & z) & 1) ? -1.0 : 1.0;
    sre += s * double(v.re);
    sim += s * double(v.im);
  }
  // Re(i^ny * (sre + i sim))
  const int k = a.ny[t] & 3;
  const double val = k == 0 ? sre : k == 1 ? -sim : k == 2 ? -sre : sim;
  red[threadIdx.x] = val;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) terms[t] = a.c[t] * red[0];
}

__global__ void k_sum_terms(const double* __restrict__ terms, int nterms, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double acc = 0.0;
    for (int t = 0; t < nterms; ++t) acc += terms[t];
    out[0] = acc;
  }
}

int check_dm(const void* rho, int nqubits, int dtype, const char* what) {
  QSIM_CHECK_ARG(rho != nullptr, std::string(what) + ": null state");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, std::string(what) + ": bad dtype");
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= 17, std::string(what) + ": density matrices of 1..17 qubits");
  return QSIM_OK;
}

// Fills the group layout for targets (qibo numbering) of an n-qubit density matrix.
int group_args(int nqubits, const int* targets, int k, GroupArgs& a, const char* what) {
  std::memset(&a, 0, sizeof(a));
  QSIM_CHECK_ARG(targets != nullptr && k >= 1 && k <= 8 && k <= nqubits, std::string(what) + ": bad targets");
  if (!validate_qubits(nqubits, targets, k, nullptr, 0, what)) return QSIM_EINVAL;
  // targets in sorted qubit order: the first is the most significant target bit
  int sorted[8];
  for (int j = 0; j < k; ++j) sorted[j] = targets[j];
  for (int i = 1; i < k; ++i)
    for (int j = i; j > 0 && sorted[j - 1] > sorted[j]; --j) {
      const int tmp = sorted[j];
      sorted[j] = sorted[j - 1];
      sorted[j - 1] = tmp;
    }
  int npos = 0;
  uint8_t pos[16];
  for (int j = 0; j < k; ++j) {
    a.tr[j] = uint8_t(2 * nqubits - 1 - sorted[j]);
    a.tc[j] = uint8_t(nqubits - 1 - sorted[j]);
    pos[npos++] = a.tr[j];
    pos[npos++] = a.tc[j];
  }
  for (int i = 1; i < npos; ++i)
    for (int j = i; j > 0 && pos[j - 1] > pos[j]; --j) {
      const uint8_t tmp = pos[j];
      pos[j] = pos[j - 1];
      pos[j - 1] = tmp;
    }
  for (int j = 0; j < npos; ++j) a.fixed[j] = pos[j];
  a.nfixed = npos;
  a.k = k;
  a.ngroups = uint64_t(1) << (2 * nqubits - 2 * k);
  return QSIM_OK;
}

}  // namespace
}  // namespace qsim

using namespace qsim;

extern "C" {

int qsim_dm_axpy(void* y, const void* x, uint64_t nelem, int dtype, double a, int accumulate, void* stream) {
  clear_error();
  QSIM_CHECK_ARG(y != nullptr && x != nullptr, "qsim_dm_axpy: null pointer");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_dm_axpy: bad dtype");
  QSIM_CHECK_ARG(accumulate >= 0 && accumulate <= 2, "qsim_dm_axpy: mode must be 0, 1 or 2");
  QSIM_CHECK_ARG(accumulate != 2 || a != 0.0, "qsim_dm_axpy: division by zero");
  if (nelem == 0) return QSIM_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid = launch_grid(ceil_div_u(nelem, kT));
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_axpy<double>, grid, dim3(kT), 0, s, (cplx<double>*)y, (const cplx<double>*)x, nelem, a,
                       accumulate);
  else
    hipLaunchKernelGGL(k_axpy<float>, grid, dim3(kT), 0, s, (cplx<float>*)y, (const cplx<float>*)x, nelem,
                       float(a), accumulate);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_dm_depolarize(void* rho, int nqubits, int dtype, const int* targets, int ntargets, double lam,
                       void* stream) {
  clear_error();
  if (check_dm(rho, nqubits, dtype, "qsim_dm_depolarize") != QSIM_OK) return QSIM_EINVAL;
  GroupArgs a;
  if (group_args(nqubits, targets, ntargets, a, "qsim_dm_depolarize") != QSIM_OK) return QSIM_EINVAL;
  QSIM_CHECK_ARG(ntargets <= 6, "qsim_dm_depolarize: at most 6 target qubits");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid = launch_grid(ceil_div_u(a.ngroups, kT));
  const double keep = 1.0 - lam, inv = 1.0 / double(1u << ntargets);
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_depolarize<double>, grid, dim3(kT), 0, s, (cplx<double>*)rho, a, keep, lam, inv);
  else
    hipLaunchKernelGGL(k_depolarize<float>, grid, dim3(kT), 0, s, (cplx<float>*)rho, a, float(keep), float(lam),
                       float(inv));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_dm_reset(void* rho, int nqubits, int dtype, int qubit, double p0, double p1, void* stream) {
  clear_error();
  if (check_dm(rho, nqubits, dtype, "qsim_dm_reset") != QSIM_OK) return QSIM_EINVAL;
  GroupArgs a;
  if (group_args(nqubits, &qubit, 1, a, "qsim_dm_reset") != QSIM_OK) return QSIM_EINVAL;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid = launch_grid(ceil_div_u(a.ngroups, kT));
  const double keep = 1.0 - p0 - p1;
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_reset<double>, grid, dim3(kT), 0, s, (cplx<double>*)rho, a, keep, p0, p1);
  else
    hipLaunchKernelGGL(k_reset<float>, grid, dim3(kT), 0, s, (cplx<float>*)rho, a, float(keep), float(p0),
                       float(p1));
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

int qsim_dm_probabilities(const void* rho, int nqubits, int dtype, const int* qubits, int nmeasured, void* probs,
                          void* stream) {
  clear_error();
  if (check_dm(rho, nqubits, dtype, "qsim_dm_probabilities") != QSIM_OK) return QSIM_EINVAL;
  QSIM_CHECK_ARG(probs != nullptr && qubits != nullptr, "qsim_dm_probabilities: null pointer");
  QSIM_CHECK_ARG(nmeasured >= 1 && nmeasured <= nqubits, "qsim_dm_probabilities: bad qubit count");
  if (!validate_qubits(nqubits, qubits, nmeasured, nullptr, 0, "qsim_dm_probabilities")) return QSIM_EINVAL;
  ProbArgs a;
  std::memset(&a, 0, sizeof(a));
  a.m = nmeasured;
  a.n = nqubits;
  a.nout = uint64_t(1) << nmeasured;
  a.nsum = uint64_t(1) << (nqubits - nmeasured);
  a.stride = (uint64_t(1) << nqubits) + 1;
  for (int j = 0; j < nmeasured; ++j) a.mbit[j] = uint8_t(nqubits - 1 - qubits[j]);
  if (sorted_positions(nqubits, qubits, nmeasured, nullptr, 0, a.upos) != nmeasured) {
    set_error("qsim_dm_probabilities: bad qubits");
    return QSIM_EINVAL;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid = launch_grid(ceil_div_u(a.nout, kT));
  if (dtype == QSIM_C128)
    hipLaunchKernelGGL(k_dm_probs<double>, grid, dim3(kT), 0, s, (const cplx<double>*)rho, a, (double*)probs);
  else
    hipLaunchKernelGGL(k_dm_probs<float>, grid, dim3(kT), 0, s, (const cplx<float>*)rho, a, (float*)probs);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

size_t qsim_dm_expval_workspace(int nterms) { return nterms > 0 ? size_t(nterms) * sizeof(double) : 0; }

int qsim_dm_expval_pauli(const void* rho, int nqubits, int dtype, int nterms, const uint64_t* xmasks,
                         const uint64_t* zmasks, const int* ny, const double* coeffs, double* out,
                         void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  if (check_dm(rho, nqubits, dtype, "qsim_dm_expval_pauli") != QSIM_OK) return QSIM_EINVAL;
  QSIM_CHECK_ARG(nterms >= 1 && nterms <= 65535, "qsim_dm_expval_pauli: bad term count");
  QSIM_CHECK_ARG(xmasks && zmasks && ny && coeffs && out && workspace, "qsim_dm_expval_pauli: null pointer");
  QSIM_CHECK_ARG(workspace_bytes >= qsim_dm_expval_workspace(nterms), "qsim_dm_expval_pauli: workspace too small");
  const uint64_t lim = uint64_t(1) << nqubits;
  for (int t = 0; t < nterms; ++t)
    QSIM_CHECK_ARG(xmasks[t] < lim && zmasks[t] < lim, "qsim_dm_expval_pauli: mask outside the register");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double* terms = reinterpret_cast<double*>(workspace);
  for (int t0 = 0; t0 < nterms; t0 += 64) {
    DmPauliArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nterms = nterms - t0 < 64 ? nterms - t0 : 64;
    for (int t = 0; t < a.nterms; ++t) {
      a.x[t] = xmasks[t0 + t];
      a.z[t] = zmasks[t0 + t];
      a.c[t] = coeffs[t0 + t];
      a.ny[t] = ny[t0 + t];
    }
    if (dtype == QSIM_C128)
      hipLaunchKernelGGL(k_dm_pauli<double>, dim3(a.nterms), dim3(kT), 0, s, (const cplx<double>*)rho, nqubits, a,
                         terms + t0);
    else
      hipLaunchKernelGGL(k_dm_pauli<float>, dim3(a.nterms), dim3(kT), 0, s, (const cplx<float>*)rho, nqubits, a,
                         terms + t0);
    QSIM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_sum_terms, dim3(1), dim3(64), 0, s, terms, nterms, out);
  QSIM_CHECK_LAUNCH();
  return QSIM_OK;
}

}  // extern "C"
