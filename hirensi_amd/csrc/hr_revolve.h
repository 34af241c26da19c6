// Binomial checkpointing of the forward wavefield for the reverse (adjoint)
// sweep: the schedule of Griewank & Walther, "Algorithm 799: revolve"
// (ACM TOMS 26(1), 2000), on segments of the time axis.
//
// The forward history is cut into m segments of L steps.  A snapshot holds the
// full RK4 state (x1, x2, x3) at a segment start; a slice buffer holds the
// displacement slices of one segment.  The reverse sweep consumes segments
// m-1, ..., 0: a segment is consumed after it has been recomputed from the
// state at its start with its slices recorded.  With s snapshot slots the
// schedule minimises the segment advances (recomputation) of the sweep; the
// first chain of snapshots is taken by the forward solve itself.
#pragma once
#include <vector>

namespace hr {

enum RevOp {
  REV_ZERO = 0,      // recompute state <- the zero state (start of segment 0)
  REV_RESTORE = 1,   // recompute state <- snapshot slot a
  REV_ADVANCE = 2,   // advance the recompute state from segment a to segment b, no recording
  REV_SNAP = 3,      // snapshot slot a <- recompute state (start of segment b)
  REV_RECORD = 4,    // advance through segment a, recording its slices
  REV_CONSUME = 5,   // adjoint steps of segment a, from its last step down to its first
};

struct RevAction {
  int op, a, b;
};

struct RevolvePlan {
  int m = 0;                     // segments
  int s = 0;                     // snapshot slots
  std::vector<long long> F, T;   // (m + 1) x (s + 1) minimal segment advances
  std::vector<int> fsplit, tsplit;

  // F(k, j): reverse k segments with j free slots, the state of the first one held.
  // T(k, j): the same when the chain of first snapshots is already taken and the
  // slices of the last segment are recorded (F(k, j) = T(k, j) + k - 1).
  void build(int nseg, int nslot);
  long long cost() const { return T[(size_t)m * (s + 1) + s]; }
  // segments the forward snapshots, slot i holds chain[i]
  std::vector<int> chain() const;
  // the reverse sweep; last_recorded: the forward kept the slices of segment m-1
  std::vector<RevAction> actions(bool last_recorded) const;
};

}  // namespace hr
