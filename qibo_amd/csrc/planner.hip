// Host-side pass selection of the blocked-execution planner.
//
// The circuit loop (abstract.py:3321-3322) runs here as blocked passes:
// qibo_amd.blocking.partition repeatedly picks, among the next pending
// gates, the set that one pass can apply. Its inner loop is a greedy scan
// over the pending gates, tried from several seeds of tile bits; this file
// runs that scan natively (one call per pass instead of a Python loop over
// up to 768 gates for each of about n seeds). The rule is the Python
// restatement's, item for item (qibo_amd.blocking._scan):
//
//   a gate is taken when it is blockable, its non-diagonal targets do not
//   touch any bit of a gate deferred before it, no deferred gate's
//   non-diagonal targets touch its bits, and the tile bits plus its
//   non-diagonal targets fit in TILE_BITS; a taken gate adds its
//   non-diagonal targets to the tile. With `ordered` nothing is taken
//   after the first deferred gate.
//
// Among the seed "the low_bits lowest bits" and the seeds "low bits plus a
// window of tile_bits - low_bits consecutive bits", the first seed that takes
// the most gates wins.

#include <stdint.h>

#include <vector>

#include "qsim_common.h"

namespace {

int scan(int nitems, const uint64_t* nd, const uint64_t* all, const uint8_t* ok, uint64_t tmask, int tile_bits,
         bool ordered, uint8_t* taken, uint64_t* tile) {
  uint64_t td = 0, ad = 0;
  bool deferred = false;
  int count = 0;
  for (int i = 0; i < nitems; ++i) {
    const uint64_t m = nd[i];
    bool take = ok[i] && !(m & ad) && !(td & all[i]);
    if (take && m) take = __builtin_popcountll(tmask | m) <= tile_bits;
    if (take && !(ordered && deferred)) {
      taken[i] = 1;
      tmask |= m;
      ++count;
    } else {
      taken[i] = 0;
      deferred = true;
      td |= m;
      ad |= all[i];
    }
  }
  *tile = tmask;
  return count;
}

}  // namespace

extern "C" {

int qsim_plan_pass(int nitems, const uint64_t* nd_masks, const uint64_t* all_masks, const uint8_t* blockable,
                   int nqubits, int low_bits, int tile_bits, int ordered, uint8_t* taken, uint64_t* tile_mask) {
  qsim::clear_error();
  QSIM_CHECK_ARG(nitems >= 0 && (nitems == 0 || (nd_masks && all_masks && blockable && taken)) && tile_mask,
                 "qsim_plan_pass: bad arrays");
  QSIM_CHECK_ARG(nqubits >= 1 && nqubits <= 64 && tile_bits >= 1 && tile_bits <= 64 && low_bits >= 0 &&
                     low_bits <= tile_bits,
                 "qsim_plan_pass: bad sizes");
  const uint64_t low = low_bits >= 64 ? ~uint64_t(0) : ((uint64_t(1) << low_bits) - 1);
  uint64_t tile = 0;
  int best = scan(nitems, nd_masks, all_masks, blockable, low, tile_bits, ordered != 0, taken, &tile);
  if (!ordered) {
    const int width = tile_bits - low_bits;
    std::vector<uint8_t> cand(size_t(nitems > 0 ? nitems : 1));
    for (int start = low_bits; start + width <= nqubits; ++start) {
      const uint64_t win = (width >= 64 ? ~uint64_t(0) : ((uint64_t(1) << width) - 1)) << start;
      uint64_t t2 = 0;
      const int c = scan(nitems, nd_masks, all_masks, blockable, low | win, tile_bits, false, cand.data(), &t2);
      if (c > best) {
        best = c;
        tile = t2;
        for (int i = 0; i < nitems; ++i) taken[i] = cand[size_t(i)];
      }
    }
  }
  *tile_mask = tile;
  return best;
}

}  // extern "C"
