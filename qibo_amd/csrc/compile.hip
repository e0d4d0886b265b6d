// Host-side record compiler of one blocked pass (qsim_compile_pass).
//
// The circuit loop (abstract.py:3321-3322) runs as blocked passes: for each
// pass the host turns its planned gates into the record stream that
// k_blocked (blocked.hip) interprets. This is that compiler in C++; its
// output is the Python restatement's (qibo_amd.blocking.compile_pass),
// record for record (the CPU test compares the two on many circuits):
//
//  1. dependency graph of the pass: per amplitude bit, a non-diagonal user
//     follows the last non-diagonal user and every diagonal user since; a
//     diagonal user follows the last non-diagonal user;
//  2. list scheduling in gate order, preferring gates that run in the current
//     register mapping; when none does, the mapping that reaches the most
//     one-qubit register gates is chosen (ties: the lower mapping);
//  3. emission: diagonal gates on lane bits become per-lane factors
//     (LANE_PHASE), diagonal gates on register bits are folded into a 16-entry
//     table (REG_PHASE) flushed before a register gate that does not commute
//     with it (after the pass's first table, the value at register 0 is taken
//     out as a global factor folded into the first table, and the record
//     multiplies only the registers whose bits the table depends on), -1 phases become sign flips (NEG), controlled phases between one
//     register bit and lane bits become per-lane factors with a flush (QFLUSH),
//     unconditional rotations / reflections become three-shear ROT records,
//     (their t and s in the header), other one-qubit gates DENSE1(_REAL),
//     two- and three-qubit gates
//     DENSE_LDS; with `ordered` only the exact per-gate forms are used;
//  4. a net -1 left by rotations with c < 0 is folded into the entries of the
//     first unconditional linear record, else a final sign flip.
//
// Complex products are formed as fma(a.re, b.re, -(a.im b.im)),
// fma(a.re, b.im, a.im b.re): numpy's complex multiply on an AVX-512 host,
// which the Python restatement uses to fold tables.

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qsim_common.h"

namespace {

constexpr int kTileBits = 12;
constexpr int kHeader = 32;
constexpr int kDescInts = 44;  // kind, ntargets, ncontrols, -, targets[5], controls[35]
constexpr int kMaxTargets = 5;
constexpr int kMaxControls = 35;
constexpr int kRegGroups[3][4] = {{8, 9, 10, 11}, {4, 5, 6, 7}, {0, 1, 2, 3}};
constexpr uint32_t kRegFlag = 0x80;

enum Kind { kNoop = 0, kDiag = 1, kSwap = 2, kDense = 3 };
enum Rec : uint8_t {
  kRDiag = 0, kRDense1 = 1, kRMap = 2, kRDenseLds = 3, kRDense1Real = 4, kRRegPhase = 5, kRLanePhase = 6,
  kRNeg = 7, kRRot = 10, kRLaneRegPhase = 11, kRQFlush = 12
};

struct C {
  double re, im;
};

inline C cmulf(C a, C b) { return {std::fma(a.re, b.re, -(a.im * b.im)), std::fma(a.re, b.im, a.im * b.re)}; }
inline bool ceq(C a, C b) { return a.re == b.re && a.im == b.im; }

struct Op {
  int kind, nt, nc;
  int t[kMaxTargets];
  int c[kMaxControls];
  const C* data;
  int ndata;
  uint64_t nd_mask, all_mask;
};

struct Writer {
  uint8_t* out;
  size_t cap, len = 0;
  int itemsize;
  int count = 0;
  bool overflow = false;
  long absorb_off = -1;  // byte offset of the entries of the first unconditional linear record
  int absorb_n = 0;

  // Overwrite the n entries at byte offset `off` (complex128 values, stored in
  // the record's precision).
  void rewrite(long off, const C* data, int n) {
    uint8_t* e = out + off;
    for (int i = 0; i < n; ++i) {
      if (itemsize == 16) {
        std::memcpy(e + 16 * i, &data[i].re, 8);
        std::memcpy(e + 16 * i + 8, &data[i].im, 8);
      } else {
        const float re = float(data[i].re), im = float(data[i].im);
        std::memcpy(e + 8 * i, &re, 4);
        std::memcpy(e + 8 * i + 4, &im, 4);
      }
    }
  }

  void emit(uint8_t op, int arg, uint64_t cmask, int nt, const int* bits, int nbits, const C* data, int ndata,
            int rreq = 0) {
    const size_t raw = size_t(ndata) * size_t(itemsize);
    const size_t rec = kHeader + (raw + 15) / 16 * 16;
    if (len + rec > cap) {
      overflow = true;
      return;
    }
    if (absorb_off < 0 && cmask == 0 && rreq == 0 && data != nullptr &&
        (op == kRDense1 || op == kRDense1Real || op == kRRegPhase || op == kRLanePhase || op == kRDiag)) {
      absorb_off = long(len + kHeader);
      absorb_n = ndata;
    }
    uint8_t* h = out + len;
    std::memset(h, 0, rec);
    h[0] = op;
    h[1] = uint8_t(arg);
    h[2] = uint8_t(nt);
    h[3] = uint8_t(rreq);
    const uint32_t nent = uint32_t(ndata);
    std::memcpy(h + 4, &nent, 4);
    std::memcpy(h + 8, &cmask, 8);
    for (int i = 0; i < nbits && i < 8; ++i) h[16 + i] = uint8_t(bits[i]);
    uint8_t* e = h + kHeader;
    for (int i = 0; i < ndata; ++i) {
      if (itemsize == 16) {
        std::memcpy(e + 16 * i, &data[i].re, 8);
        std::memcpy(e + 16 * i + 8, &data[i].im, 8);
      } else {
        const float re = float(data[i].re), im = float(data[i].im);
        std::memcpy(e + 8 * i, &re, 4);
        std::memcpy(e + 8 * i + 4, &im, 4);
      }
    }
    len += rec;
    ++count;
  }

  // ROT record: no entries, (t, s) in the header's cmask and bits fields
  // (float64 at bytes 8 and 16; complex64: float32 at bytes 8 and 16)
  void emit_rot(int arg, double t, double s) {
    if (len + kHeader > cap) {
      overflow = true;
      return;
    }
    uint8_t* h = out + len;
    std::memset(h, 0, kHeader);
    h[0] = kRRot;
    h[1] = uint8_t(arg);
    if (itemsize == 16) {
      std::memcpy(h + 8, &t, 8);
      std::memcpy(h + 16, &s, 8);
    } else {
      const float tf = float(t), sf = float(s);
      std::memcpy(h + 8, &tf, 4);
      std::memcpy(h + 16, &sf, 4);
    }
    len += kHeader;
    ++count;
  }

  void negate_absorbed() {
    for (int i = 0; i < 2 * absorb_n; ++i) {
      uint8_t* p = out + absorb_off + size_t(i) * size_t(itemsize / 2);
      if (itemsize == 16) {
        double v;
        std::memcpy(&v, p, 8);
        v = -v;
        std::memcpy(p, &v, 8);
      } else {
        float v;
        std::memcpy(&v, p, 4);
        v = -v;
        std::memcpy(p, &v, 4);
      }
    }
  }
};

// (mode, t, s, negated) of a 2x2 of the ROT forms with |c^2 + s^2 - 1| <= tol
// (1e-14 for complex128 programs; 4e-7 for complex64 ones, whose entries are
// float32 roundings of cos and sin: the rotation by atan2(s, c) differs from
// the rounded matrix by less than float32's own rounding).
bool rotation_form(const C* m, double tol, int* mode, double* t, double* s, bool* neg) {
  const C a = m[0], b = m[1], cc = m[2], d = m[3];
  const bool real = a.im == 0 && b.im == 0 && cc.im == 0 && d.im == 0;
  double c, sn;
  if (real && a.re == d.re && b.re == -cc.re) {
    *mode = 0, c = a.re, sn = cc.re;
  } else if (real && a.re == -d.re && b.re == cc.re) {
    *mode = 1, c = a.re, sn = cc.re;
  } else if (a.re == d.re && a.im == 0 && d.im == 0 && b.re == 0 && cc.re == 0 && b.im == cc.im) {
    *mode = 2, c = a.re, sn = -b.im;
  } else {
    return false;
  }
  if (std::fabs(c * c + sn * sn - 1.0) > tol) return false;
  *neg = c < 0;
  if (*neg) c = -c, sn = -sn;
  *t = sn / (1.0 + c);
  *s = sn;
  return true;
}

int compile(int n, const int* tile_bits, std::vector<Op>& items, bool ordered, Writer& w, int* switches) {
  const int N = int(items.size());
  int local[64];
  for (int b = 0; b < 64; ++b) local[b] = -1;
  for (int i = 0; i < kTileBits; ++i) local[tile_bits[i]] = i;

  // 1. dependencies
  std::vector<std::vector<int>> succ(static_cast<size_t>(N));
  std::vector<int> npred(size_t(N), 0);
  if (ordered) {
    for (int i = 0; i + 1 < N; ++i) {
      succ[size_t(i)].push_back(i + 1);
      npred[size_t(i + 1)] = 1;
    }
  } else {
    int last_nd[64];
    for (int b = 0; b < 64; ++b) last_nd[b] = -1;
    std::vector<int> diag_since[64];
    std::vector<int> preds;
    for (int j = 0; j < N; ++j) {
      const Op& it = items[size_t(j)];
      preds.clear();
      for (int b = 0; b < 64; ++b) {
        if (!((it.all_mask >> b) & 1)) continue;
        if ((it.nd_mask >> b) & 1) preds.insert(preds.end(), diag_since[b].begin(), diag_since[b].end());
        if (last_nd[b] >= 0) preds.push_back(last_nd[b]);
      }
      for (int b = 0; b < 64; ++b) {
        if (!((it.all_mask >> b) & 1)) continue;
        if ((it.nd_mask >> b) & 1) {
          last_nd[b] = j;
          diag_since[b].clear();
        } else {
          diag_since[b].push_back(j);
        }
      }
      std::sort(preds.begin(), preds.end());
      preds.erase(std::unique(preds.begin(), preds.end()), preds.end());
      for (int i : preds) succ[size_t(i)].push_back(j);
      npred[size_t(j)] = int(preds.size());
    }
  }

  auto reg_gate = [&](int i) { return items[size_t(i)].kind == kDense && items[size_t(i)].nt == 1; };
  auto runs_here = [&](int i, int m) {
    const Op& o = items[size_t(i)];
    if (o.kind == kDiag) return true;
    if (o.kind == kDense && o.nt == 1) {
      const int lb = local[n - 1 - o.t[0]];
      for (int k = 0; k < 4; ++k)
        if (kRegGroups[m][k] == lb) return true;
    }
    return false;
  };
  auto reg_index = [&](int m, int lb) {
    for (int k = 0; k < 4; ++k)
      if (kRegGroups[m][k] == lb) return k;
    return -1;
  };
  std::vector<int> ready;
  for (int i = 0; i < N; ++i)
    if (npred[size_t(i)] == 0) ready.push_back(i);

  auto reach = [&](int m) {
    std::vector<int> left(npred);
    std::vector<int> frontier;
    for (int i : ready)
      if (runs_here(i, m)) frontier.push_back(i);
    int done = 0;
    while (!frontier.empty()) {
      const int i = frontier.back();
      frontier.pop_back();
      if (reg_gate(i)) ++done;
      for (int j : succ[size_t(i)]) {
        if (--left[size_t(j)] == 0 && runs_here(j, m)) frontier.push_back(j);
      }
    }
    return done;
  };

  int cur = 0;
  bool have_pending = false;
  C pending[16];
  // the pass's first REG_PHASE table (all registers) and the global factor
  // taken out of the later ones, folded into it at the end
  long reg_first_off = -1;
  C reg_first[16];
  C gfac = {1.0, 0.0};
  bool sign = false;
  int qpend = -1;
  *switches = 0;

  auto reg_bits = [&](int m, int* rb) {
    for (int k = 0; k < 4; ++k) rb[k] = tile_bits[kRegGroups[m][k]];
  };
  auto rb_index = [&](const int* rb, int b) {
    for (int k = 0; k < 4; ++k)
      if (rb[k] == b) return k;
    return -1;
  };
  auto qflush = [&](int j) {
    if (qpend < 0 || (j >= 0 && j != qpend)) return;
    w.emit(kRQFlush, qpend, 0, 0, nullptr, 0, nullptr, 0);
    qpend = -1;
  };
  auto flush_reg = [&](int j) {
    if (!have_pending) return;
    auto depends_on = [&](int k) {
      for (int r = 0; r < 16; ++r)
        if (!ceq(pending[r], pending[r ^ (1 << k)])) return true;
      return false;
    };
    if (j >= 0 && !depends_on(j)) return;
    if (reg_first_off < 0) {
      reg_first_off = long(w.len) + kHeader;
      for (int r = 0; r < 16; ++r) reg_first[r] = pending[r];
      w.emit(kRRegPhase, 0, 0, 0, nullptr, 0, pending, 16);
    } else {
      // T = T[0] (T / T[0]): the factor goes to the first table, the record
      // multiplies only the registers where a bit T depends on is set
      const C t0 = pending[0];
      const double mag = t0.re * t0.re + t0.im * t0.im;
      if (std::fabs(mag - 1.0) <= 1e-12) {
        int dep = 0;
        for (int k = 0; k < 4; ++k)
          if (depends_on(k)) dep |= 1 << k;
        gfac = cmulf(gfac, t0);
        if (dep) {
          const C inv = {t0.re / mag, -t0.im / mag};
          C table[16];
          for (int r = 0; r < 16; ++r) table[r] = (r & dep) ? cmulf(pending[r], inv) : C{1.0, 0.0};
          w.emit(kRRegPhase, dep, 0, 0, nullptr, 0, table, 16);
        }
      } else {
        w.emit(kRRegPhase, 0, 0, 0, nullptr, 0, pending, 16);
      }
    }
    have_pending = false;
  };
  auto split_cond = [&](const Op& o, int m, uint64_t* cm, int* rreq) {
    int rb[4];
    reg_bits(m, rb);
    *cm = 0;
    *rreq = 0;
    for (int k = 0; k < o.nc; ++k) {
      const int b = n - 1 - o.c[k];
      const int ri = rb_index(rb, b);
      if (ri >= 0)
        *rreq |= 1 << ri;
      else
        *cm |= uint64_t(1) << b;
    }
  };

  auto emit_diag = [&](const Op& o) {
    int rb[4];
    reg_bits(cur, rb);
    int tb[kMaxTargets], cb[kMaxControls];
    for (int k = 0; k < o.nt; ++k) tb[k] = n - 1 - o.t[k];
    for (int k = 0; k < o.nc; ++k) cb[k] = n - 1 - o.c[k];
    const int nt = o.nt;
    if (!ordered) {
      bool none_reg = true, all_reg = true;
      for (int k = 0; k < nt; ++k) (rb_index(rb, tb[k]) >= 0 ? none_reg : all_reg) = false;
      for (int k = 0; k < o.nc; ++k) (rb_index(rb, cb[k]) >= 0 ? none_reg : all_reg) = false;
      if (none_reg) {
        uint64_t cm = 0;
        for (int k = 0; k < o.nc; ++k) cm |= uint64_t(1) << cb[k];
        w.emit(kRLanePhase, 0, cm, nt, tb, nt, o.data, o.ndata);
        return;
      }
      if (all_reg) {
        if (!have_pending) {
          for (int r = 0; r < 16; ++r) pending[r] = {1.0, 0.0};
          have_pending = true;
        }
        for (int r = 0; r < 16; ++r) {
          bool ok = true;
          for (int k = 0; k < o.nc; ++k) ok = ok && ((r >> rb_index(rb, cb[k])) & 1);
          if (!ok) continue;
          int di = 0;
          for (int k = 0; k < nt; ++k) di = (di << 1) | ((r >> rb_index(rb, tb[k])) & 1);
          pending[r] = cmulf(pending[r], o.data[di]);
        }
        return;
      }
    }
    uint64_t cm;
    int rreq;
    split_cond(o, cur, &cm, &rreq);
    int enc[kMaxTargets];
    for (int k = 0; k < nt; ++k) {
      const int ri = rb_index(rb, tb[k]);
      enc[k] = ri >= 0 ? int(kRegFlag | uint32_t(ri)) : tb[k];
    }
    if (!ordered && nt == 0 && o.data[0].re == -1.0 && o.data[0].im == 0.0) {
      w.emit(kRNeg, 0, cm, 0, nullptr, 0, nullptr, 0, rreq);
    } else if (!ordered && nt == 0 && (rreq == 1 || rreq == 2 || rreq == 4 || rreq == 8)) {
      const int j = rreq == 1 ? 0 : (rreq == 2 ? 1 : (rreq == 4 ? 2 : 3));
      if (qpend != j) {
        qflush(-1);
        qpend = j;
      }
      w.emit(kRLaneRegPhase, j, cm, 0, nullptr, 0, o.data, 1);
    } else {
      w.emit(kRDiag, 0, cm, nt, enc, nt, o.data, o.ndata, rreq);
    }
  };

  static const C kSwapMat[16] = {{1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {1, 0}, {0, 0},
                                 {0, 0}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {1, 0}};

  while (!ready.empty()) {
    std::sort(ready.begin(), ready.end());
    int pick = -1;
    for (int i : ready)
      if (runs_here(i, cur)) {
        pick = i;
        break;
      }
    if (pick < 0) {
      bool anyreg = false;
      for (int i : ready) anyreg = anyreg || reg_gate(i);
      if (anyreg) {
        int best = -1, bestc = -1;
        for (int m = 0; m < 3; ++m) {
          if (m == cur) continue;
          const int cnt = reach(m);
          if (cnt > bestc) best = m, bestc = cnt;  // ties keep the lower mapping
        }
        cur = best;
        flush_reg(-1);
        qflush(-1);
        w.emit(kRMap, cur, 0, 0, nullptr, 0, nullptr, 0);
        ++*switches;
        for (int i : ready)
          if (runs_here(i, cur)) {
            pick = i;
            break;
          }
        if (pick < 0) return -1;  // cannot happen: the chosen mapping runs a ready gate
      } else {
        pick = ready[0];
      }
    }
    ready.erase(std::find(ready.begin(), ready.end(), pick));
    for (int j : succ[size_t(pick)])
      if (--npred[size_t(j)] == 0) ready.push_back(j);

    const Op& o = items[size_t(pick)];
    if (o.kind == kDiag) {
      emit_diag(o);
    } else if (o.kind == kDense && o.nt == 1 && runs_here(pick, cur)) {
      const int lb = local[n - 1 - o.t[0]];
      const int j = reg_index(cur, lb);
      flush_reg(j);
      qflush(j);
      uint64_t cm;
      int rreq;
      split_cond(o, cur, &cm, &rreq);
      int mode;
      double t, sn;
      bool neg;
      if (!ordered && cm == 0 && rreq == 0 && rotation_form(o.data, w.itemsize == 16 ? 1e-14 : 4e-7, &mode, &t, &sn, &neg)) {
        w.emit_rot(j | (mode << 2), t, sn);
        sign ^= neg;
        continue;
      }
      bool real = !ordered;
      for (int k = 0; k < 4 && real; ++k) real = o.data[k].im == 0;
      w.emit(real ? kRDense1Real : kRDense1, j, cm, 0, nullptr, 0, o.data, 4, rreq);
    } else {
      flush_reg(-1);
      qflush(-1);
      const C* mat = o.kind == kSwap ? kSwapMat : o.data;
      const int k = o.nt;
      int bits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      int lbits[3];
      for (int q = 0; q < k; ++q) lbits[q] = local[n - 1 - o.t[q]];
      for (int q = 0; q < k; ++q) bits[q] = lbits[q];
      std::sort(lbits, lbits + k);
      for (int q = 0; q < k; ++q) bits[4 + q] = lbits[q];
      uint64_t cm = 0;
      for (int q = 0; q < o.nc; ++q) cm |= uint64_t(1) << (n - 1 - o.c[q]);
      w.emit(kRDenseLds, k, cm, 0, bits, 8, mat, (1 << k) * (1 << k));
    }
  }
  flush_reg(-1);
  qflush(-1);
  if (reg_first_off >= 0 && !(gfac.re == 1.0 && gfac.im == 0.0) && !w.overflow) {
    C table[16];
    for (int r = 0; r < 16; ++r) table[r] = cmulf(reg_first[r], gfac);
    w.rewrite(reg_first_off, table, 16);
  }
  if (sign) {
    if (w.absorb_off >= 0)
      w.negate_absorbed();
    else
      w.emit(kRNeg, 0, 0, 0, nullptr, 0, nullptr, 0);
  }
  return 0;
}

}  // namespace

extern "C" {

size_t qsim_compile_pass_capacity(int nops, int dtype) {
  if (nops < 0) return 0;
  const size_t item = dtype == QSIM_C64 ? 8 : 16;
  // per op at most a mapping change, a table flush, a factor flush and its own
  // record (<= 64 entries); then the final flushes and sign flip
  const size_t per_op = 4 * kHeader + (16 + 64) * item + 32;
  return size_t(nops) * per_op + 4 * kHeader + 16 * item;
}

int qsim_compile_pass(int nqubits, const int* tile_bits, int nops, const int* op_desc, const double* op_data,
                      const int64_t* data_offsets, int dtype, int ordered, uint8_t* out, size_t capacity,
                      size_t* out_bytes, int* count, int* switches) {
  qsim::clear_error();
  QSIM_CHECK_ARG(nqubits >= kTileBits && nqubits <= QSIM_MAX_QUBITS, "qsim_compile_pass: nqubits out of range");
  QSIM_CHECK_ARG(tile_bits != nullptr && out_bytes != nullptr && count != nullptr && switches != nullptr,
                 "qsim_compile_pass: null pointer");
  QSIM_CHECK_ARG(nops >= 0 && (nops == 0 || (op_desc && op_data && data_offsets)), "qsim_compile_pass: bad ops");
  QSIM_CHECK_ARG(dtype == QSIM_C128 || dtype == QSIM_C64, "qsim_compile_pass: bad dtype");
  QSIM_CHECK_ARG(out != nullptr && capacity >= qsim_compile_pass_capacity(nops, dtype),
                 "qsim_compile_pass: output buffer smaller than qsim_compile_pass_capacity");
  uint64_t tmask = 0;
  for (int i = 0; i < kTileBits; ++i) {
    QSIM_CHECK_ARG(tile_bits[i] >= 0 && tile_bits[i] < nqubits && (i == 0 || tile_bits[i] > tile_bits[i - 1]),
                   "qsim_compile_pass: tile bits must ascend inside the register");
    tmask |= uint64_t(1) << tile_bits[i];
  }
  std::vector<Op> items;
  items.reserve(size_t(nops));
  for (int i = 0; i < nops; ++i) {
    const int* d = op_desc + size_t(i) * kDescInts;
    Op o;
    std::memset(&o, 0, sizeof(o));
    o.kind = d[0];
    o.nt = d[1];
    o.nc = d[2];
    QSIM_CHECK_ARG(o.kind >= kNoop && o.kind <= kDense, "qsim_compile_pass: bad op kind");
    QSIM_CHECK_ARG(o.nt >= 0 && o.nt <= kMaxTargets && o.nc >= 0 && o.nc <= kMaxControls,
                   "qsim_compile_pass: bad qubit counts");
    if (o.kind == kNoop) continue;
    uint64_t seen = 0;
    for (int k = 0; k < o.nt + o.nc; ++k) {
      const int q = k < o.nt ? d[4 + k] : d[9 + (k - o.nt)];
      QSIM_CHECK_ARG(q >= 0 && q < nqubits, "qsim_compile_pass: qubit out of range");
      const uint64_t bit = uint64_t(1) << (nqubits - 1 - q);
      QSIM_CHECK_ARG(!(seen & bit), "qsim_compile_pass: qubit given twice");
      seen |= bit;
      if (k < o.nt)
        o.t[k] = q;
      else
        o.c[k - o.nt] = q;
    }
    int need = 0;
    if (o.kind == kDiag) {
      QSIM_CHECK_ARG(o.nt <= 5, "qsim_compile_pass: diagonal on more than 5 targets");
      need = 1 << o.nt;
    } else if (o.kind == kDense) {
      QSIM_CHECK_ARG(o.nt >= 1 && o.nt <= 3, "qsim_compile_pass: dense gates in a pass act on 1..3 targets");
      need = (1 << o.nt) * (1 << o.nt);
    } else {
      QSIM_CHECK_ARG(o.nt == 2, "qsim_compile_pass: SWAP acts on 2 targets");
    }
    const int64_t off = data_offsets[i];
    QSIM_CHECK_ARG(need == 0 || off >= 0, "qsim_compile_pass: missing gate data");
    o.data = need ? reinterpret_cast<const C*>(op_data) + off : nullptr;
    o.ndata = need;
    for (int k = 0; k < o.nt; ++k) {
      const uint64_t bit = uint64_t(1) << (nqubits - 1 - o.t[k]);
      o.all_mask |= bit;
      if (o.kind != kDiag) {
        QSIM_CHECK_ARG(tmask & bit, "qsim_compile_pass: a non-diagonal target outside the tile bits");
        o.nd_mask |= bit;
      }
    }
    for (int k = 0; k < o.nc; ++k) o.all_mask |= uint64_t(1) << (nqubits - 1 - o.c[k]);
    items.push_back(o);
  }
  Writer w;
  w.out = out;
  w.cap = capacity;
  w.itemsize = dtype == QSIM_C64 ? 8 : 16;
  int sw = 0;
  if (compile(nqubits, tile_bits, items, ordered != 0, w, &sw) != 0) {
    qsim::set_error("qsim_compile_pass: scheduling failed");
    return QSIM_EINVAL;
  }
  QSIM_CHECK_ARG(!w.overflow, "qsim_compile_pass: output buffer overflow");
  *out_bytes = w.len;
  *count = w.count;
  *switches = sw;
  return QSIM_OK;
}

}  // extern "C"
