// Binomial reverse schedule on segments (see hr_revolve.h).
#include "hr_revolve.h"

#include <algorithm>
#include <climits>

namespace hr {

void RevolvePlan::build(int nseg, int nslot) {
  m = std::max(nseg, 1);
  s = std::max(0, std::min(nslot, m - 1));
  const size_t W = (size_t)s + 1;
  F.assign((size_t)(m + 1) * W, 0);
  T.assign((size_t)(m + 1) * W, 0);
  fsplit.assign((size_t)(m + 1) * W, 1);
  tsplit.assign((size_t)(m + 1) * W, 1);
  for (int k = 1; k <= m; ++k) {
    // no free slot: advance from the held state to every segment in turn; on the
    // forward's chain the last segment's slices are already recorded
    F[k * W] = (long long)k * (k - 1) / 2;
    T[k * W] = (long long)(k - 1) * (k - 2) / 2;
    for (int j = 1; j <= s; ++j) {
      if (k == 1) continue;
      long long bf = LLONG_MAX, bt = LLONG_MAX;
      int af = 1, at = 1;
      for (int c = 1; c < k; ++c) {
        // advance c segments, snapshot, reverse the right part with one slot less,
        // then the left part with the slot free again
        const long long vf = c + F[(size_t)(k - c) * W + j - 1] + F[(size_t)c * W + j];
        const long long vt = T[(size_t)(k - c) * W + j - 1] + F[(size_t)c * W + j];
        if (vf < bf) { bf = vf; af = c; }
        if (vt < bt) { bt = vt; at = c; }
      }
      F[k * W + j] = bf;
      T[k * W + j] = bt;
      fsplit[k * W + j] = af;
      tsplit[k * W + j] = at;
    }
  }
}

std::vector<int> RevolvePlan::chain() const {
  std::vector<int> out;
  const size_t W = (size_t)s + 1;
  int a = 0, j = s;
  while (m - a > 1 && j > 0) {
    a += tsplit[(size_t)(m - a) * W + j];
    out.push_back(a);
    --j;
  }
  return out;
}

namespace {
struct Gen {
  const RevolvePlan& P;
  bool last_recorded;
  int cur = -1;   // segment whose start state the recompute buffers hold (-1: unknown)
  std::vector<RevAction> out;

  void need(int a, int slot) {
    if (cur == a) return;
    out.push_back({slot < 0 ? REV_ZERO : REV_RESTORE, slot, a});
    cur = a;
  }
  void leaf(int c, int a, int slot_a) {
    if (c == P.m - 1 && last_recorded) {
      out.push_back({REV_CONSUME, c, 0});
      return;
    }
    need(a, slot_a);
    if (c > a) {
      out.push_back({REV_ADVANCE, a, c});
      cur = c;
    }
    out.push_back({REV_RECORD, c, 0});
    cur = c + 1;
    out.push_back({REV_CONSUME, c, 0});
  }
  void run(int a, int b, int j, int slot_a, bool top, int depth) {
    const size_t W = (size_t)P.s + 1;
    if (b - a == 1) {
      leaf(a, a, slot_a);
      return;
    }
    if (j == 0) {
      for (int c = b - 1; c >= a; --c) leaf(c, a, slot_a);
      return;
    }
    const int c = a + (top ? P.tsplit : P.fsplit)[(size_t)(b - a) * W + j];
    if (!top) {
      need(a, slot_a);
      out.push_back({REV_ADVANCE, a, c});
      cur = c;
      out.push_back({REV_SNAP, depth, c});
    }
    run(c, b, j - 1, depth, top, depth + 1);
    run(a, c, j, slot_a, false, depth);
  }
};
}  // namespace

std::vector<RevAction> RevolvePlan::actions(bool last_recorded) const {
  Gen g{*this, last_recorded};
  g.run(0, m, s, -1, true, 0);
  return g.out;
}

}  // namespace hr
