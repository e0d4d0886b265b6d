// Host-side fusion: the neighbour-fusion planner of Circuit.fuse and the
// product of a fused group's matrices (Backend.matrix_fused), in C++.
//
// qsim_fuse_plan restates Circuit.fuse (circuit.py:954-1003) with
// _Queue.to_fused / from_fused (:55-95) and FusedGate.can_fuse / fuse /
// append / prepend (gates/special.py:57-153), step for step: every gate
// becomes a group linked to its left and right neighbour on each of its
// qubits; in queue order, each unmarked group tries, for each of its qubits
// (ascending, as at the start of its turn), to absorb its right neighbour and
// to be absorbed by its left neighbour, when the union of their qubits stays
// within max_qubits and the neighbour links allow it. The grouping decides
// the fused matrices, so it is kept exactly (the CPU tests compare it with the
// Python restatement on many circuits).
//
// qsim_matrix_fused forms the product of a group's gate matrices on the
// group's sorted qubits (abstract.py:2680-2717): starting from the identity,
// each gate (with its controlled_by controls as a block diagonal) is applied
// to every column of the running product, as a gate is applied to a state,
// instead of building the Kronecker-embedded matrix and multiplying it.

#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "qsim_common.h"

namespace {

struct Group {
  uint64_t qset = 0;
  bool marked = false;
  std::vector<int> gates;
  std::vector<int> left, right;  // per qubit: neighbour group, -1 for none
};

int popcount64(uint64_t x) { return __builtin_popcountll(x); }

struct Planner {
  int n;
  int max_qubits;
  std::vector<Group> g;

  bool can_fuse(int a, int b) const {
    if (b < 0 || g[size_t(a)].marked || g[size_t(b)].marked) return false;
    return popcount64(g[size_t(a)].qset | g[size_t(b)].qset) <= max_qubits;
  }

  // number of distinct neighbours in `links` other than `except`
  int distinct_except(const std::vector<int>& links, int except) const {
    std::vector<int> seen;
    for (int v : links)
      if (v >= 0 && v != except && std::find(seen.begin(), seen.end(), v) == seen.end()) seen.push_back(v);
    return int(seen.size());
  }

  void grow(int p, int c) { g[size_t(p)].qset |= g[size_t(c)].qset; }

  // FusedGate.fuse: `b` is the later group in time.
  void fuse(int a, int b) {
    Group& A = g[size_t(a)];
    Group& B = g[size_t(b)];
    const int bl = distinct_except(A.right, b);
    const int br = distinct_except(B.left, a);
    if (bl > 0 && br > 0) return;
    const uint64_t shared = A.qset & B.qset;
    int parent, child;
    if (bl > br) {
      parent = a, child = b;
      Group& P = g[size_t(parent)];
      bool only = true;
      for (int q = 0; q < n; ++q)
        if ((shared >> q) & 1) only = only && P.right[size_t(q)] == child;
      if (only) {
        Group& Cg = g[size_t(child)];
        Cg.marked = true;
        grow(parent, child);
        P.gates.insert(P.gates.end(), Cg.gates.begin(), Cg.gates.end());
        for (int q = 0; q < n; ++q) {
          if (!((shared >> q) & 1)) continue;
          const int nxt = Cg.right[size_t(q)];
          if (nxt >= 0) {
            P.right[size_t(q)] = nxt;
            g[size_t(nxt)].left[size_t(q)] = parent;
          } else {
            P.right[size_t(q)] = -1;
          }
        }
      }
    } else {
      parent = b, child = a;
      Group& P = g[size_t(parent)];
      bool only = true;
      for (int q = 0; q < n; ++q)
        if ((shared >> q) & 1) only = only && P.left[size_t(q)] == child;
      if (only) {
        Group& Cg = g[size_t(child)];
        Cg.marked = true;
        grow(parent, child);
        P.gates.insert(P.gates.begin(), Cg.gates.begin(), Cg.gates.end());
        for (int q = 0; q < n; ++q) {
          if (!((shared >> q) & 1)) continue;
          const int prv = Cg.left[size_t(q)];
          if (prv >= 0) {
            P.left[size_t(q)] = prv;
            g[size_t(prv)].right[size_t(q)] = parent;
          }
        }
      }
    }
    Group& Cg = g[size_t(child)];
    if (Cg.marked) {
      Group& P = g[size_t(parent)];
      const uint64_t rest = Cg.qset & ~shared;
      for (int q = 0; q < n; ++q) {
        if (!((rest >> q) & 1)) continue;
        const int nxt = Cg.right[size_t(q)];
        if (nxt >= 0) {
          P.right[size_t(q)] = nxt;
          g[size_t(nxt)].left[size_t(q)] = parent;
        }
        const int prv = Cg.left[size_t(q)];
        if (prv >= 0) {
          P.left[size_t(q)] = prv;
          g[size_t(prv)].right[size_t(q)] = parent;
        }
      }
    }
  }
};

struct Cx {
  double re, im;
};

}  // namespace

extern "C" {

int qsim_fuse_plan(int nqubits, int ngates, const uint64_t* qubit_masks, const uint8_t* kind, int max_qubits,
                   int* status, int* members, int* member_offsets) {
  qsim::clear_error();
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= 64, "qsim_fuse_plan: nqubits out of range");
  QSIM_CHECK_ARG(ngates >= 0 && (ngates == 0 || (qubit_masks && kind && status && members && member_offsets)),
                 "qsim_fuse_plan: bad arrays");
  QSIM_CHECK_ARG(max_qubits >= 1, "qsim_fuse_plan: max_qubits must be positive");
  Planner p;
  p.n = nqubits;
  p.max_qubits = max_qubits;
  p.g.resize(size_t(ngates));
  const uint64_t all = nqubits == 64 ? ~uint64_t(0) : ((uint64_t(1) << nqubits) - 1);
  std::vector<int> last(size_t(nqubits), -1);
  for (int i = 0; i < ngates; ++i) {
    QSIM_CHECK_ARG(kind[i] <= 2, "qsim_fuse_plan: kind must be 0 (gate), 1 (measurement) or 2 (special)");
    QSIM_CHECK_ARG((qubit_masks[i] & ~all) == 0, "qsim_fuse_plan: qubit out of range");
    Group& G = p.g[size_t(i)];
    G.qset = kind[i] == 2 ? all : qubit_masks[i];
    G.marked = kind[i] != 0;
    G.gates.push_back(i);
    G.left.assign(size_t(nqubits), -1);
    G.right.assign(size_t(nqubits), -1);
    for (int q = 0; q < nqubits; ++q) {
      if (!((G.qset >> q) & 1)) continue;
      if (last[size_t(q)] >= 0) {
        G.left[size_t(q)] = last[size_t(q)];
        p.g[size_t(last[size_t(q)])].right[size_t(q)] = i;
      }
      last[size_t(q)] = i;
    }
  }
  for (int i = 0; i < ngates; ++i) {
    if (p.g[size_t(i)].marked) continue;
    const uint64_t qs = p.g[size_t(i)].qset;  // its qubits at the start of its turn
    for (int q = 0; q < nqubits; ++q) {
      if (!((qs >> q) & 1)) continue;
      int nb = p.g[size_t(i)].right[size_t(q)];
      if (p.can_fuse(i, nb)) p.fuse(i, nb);
      nb = p.g[size_t(i)].left[size_t(q)];
      if (p.can_fuse(i, nb)) p.fuse(nb, i);
    }
  }
  // status: 0 dropped (absorbed), 1 the gate itself, 2 a fused group
  int off = 0;
  for (int i = 0; i < ngates; ++i) {
    const Group& G = p.g[size_t(i)];
    member_offsets[i] = off;
    if (!G.marked) {
      status[i] = G.gates.size() == 1 ? 1 : 2;
      for (int m : G.gates) members[off++] = m;
    } else {
      status[i] = kind[G.gates[0]] != 0 ? 1 : 0;
    }
  }
  member_offsets[ngates] = off;
  return QSIM_OK;
}

int qsim_matrix_fused(int rank, const int* qubits, int ngates, const int* gate_desc, const double* mats,
                      const int64_t* mat_offsets, double* out) {
  qsim::clear_error();
  QSIM_CHECK_ARG(rank >= 1 && rank <= 10, "qsim_matrix_fused: rank must be in 1..10");
  QSIM_CHECK_ARG(qubits != nullptr && out != nullptr && ngates >= 0 &&
                     (ngates == 0 || (gate_desc && mats && mat_offsets)),
                 "qsim_matrix_fused: bad arrays");
  const int dim = 1 << rank;
  std::vector<Cx> m(size_t(dim) * size_t(dim));
  std::vector<Cx> col(static_cast<size_t>(dim));
  for (int i = 0; i < dim; ++i) m[size_t(i) * size_t(dim) + size_t(i)] = {1.0, 0.0};
  const Cx* all = reinterpret_cast<const Cx*>(mats);
  for (int k = 0; k < ngates; ++k) {
    // gate_desc row: ncontrols, ntargets, then the controls and the targets
    // (qubit numbers, targets in matrix order); 2 + rank ints per gate
    const int* d = gate_desc + size_t(k) * size_t(2 + rank);
    const int nc = d[0], nt = d[1];
    QSIM_CHECK_ARG(nt >= 1 && nc >= 0 && nc + nt <= rank, "qsim_matrix_fused: bad gate size");
    int cpos[10], tpos[10];
    for (int j = 0; j < nc + nt; ++j) {
      int pos = -1;
      for (int r = 0; r < rank; ++r)
        if (qubits[r] == d[2 + j]) pos = rank - 1 - r;  // bit of the group index (first qubit = MSB)
      QSIM_CHECK_ARG(pos >= 0, "qsim_matrix_fused: a gate qubit outside the group");
      (j < nc ? cpos[j] : tpos[j - nc]) = pos;
    }
    int cmask = 0, tmask = 0;
    for (int j = 0; j < nc; ++j) cmask |= 1 << cpos[j];
    for (int j = 0; j < nt; ++j) tmask |= 1 << tpos[j];
    QSIM_CHECK_ARG(popcount64(uint64_t(cmask | tmask)) == nc + nt, "qsim_matrix_fused: a qubit given twice");
    const Cx* u = all + mat_offsets[k];
    const int tdim = 1 << nt;
    std::vector<Cx> x(static_cast<size_t>(tdim));
    // M <- (U on the targets, under the controls) M, column by column
    for (int c = 0; c < dim; ++c) {
      for (int r = 0; r < dim; ++r) col[size_t(r)] = m[size_t(r) * size_t(dim) + size_t(c)];
      for (int base = 0; base < dim; ++base) {
        if ((base & tmask) != 0 || (base & cmask) != cmask) continue;
        auto index = [&](int a) {
          int idx = base;
          for (int j = 0; j < nt; ++j)
            if ((a >> (nt - 1 - j)) & 1) idx |= 1 << tpos[j];
          return idx;
        };
        for (int a = 0; a < tdim; ++a) x[size_t(a)] = col[size_t(index(a))];
        for (int a = 0; a < tdim; ++a) {
          double re = 0.0, im = 0.0;
          for (int b = 0; b < tdim; ++b) {
            const Cx w = u[size_t(a) * size_t(tdim) + size_t(b)];
            re += w.re * x[size_t(b)].re - w.im * x[size_t(b)].im;
            im += w.re * x[size_t(b)].im + w.im * x[size_t(b)].re;
          }
          m[size_t(index(a)) * size_t(dim) + size_t(c)] = {re, im};
        }
      }
    }
  }
  std::memcpy(out, m.data(), sizeof(Cx) * m.size());
  return QSIM_OK;
}

}  // extern "C"
