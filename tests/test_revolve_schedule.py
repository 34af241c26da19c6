"""The binomial reverse schedule of the checkpointed history (host code of
libhirensi.so, no GPU): every segment is consumed once, last to first, from a
state that the schedule has actually reconstructed, within the snapshot slots,
and the number of recomputed segments is the minimum of Griewank & Walther's
revolve: with c checkpoints (the held start state included) and m steps,
t = r m - beta(c + 1, r - 1), beta(c, r) = C(c + r, c), r the least integer with
beta(c, r) >= m.  Here the forward itself takes the first chain of snapshots and
keeps the last segment's slices, which saves exactly m - 1 advances.
"""
from math import comb

import pytest

from hirensi_amd import solver as S

ZERO, RESTORE, ADVANCE, SNAP, RECORD, CONSUME = range(6)


def griewank_t(m, c):
    if m <= 1:
        return 0
    r = 0
    while comb(c + r, c) < m:
        r += 1
    return r * m - (comb(c + 1 + r - 1, c + 1) if r >= 1 else 0)


def simulate(m, s):
    acts, adv = S.revolve_schedule(m, s)
    slots = {}                    # slot -> segment whose start state it holds
    pos = None                    # segment whose start state the recompute buffers hold
    recorded = {m - 1}            # the forward keeps the last segment's slices
    consumed = []
    advances = 0
    sweep = False
    for op, a, b in acts:
        if op == SNAP and not sweep:
            assert a not in slots and a < max(s, 1)
            slots[a] = b            # forward chain
            continue
        sweep = True
        if op == ZERO:
            pos = 0
        elif op == RESTORE:
            assert a in slots
            pos = slots[a]
        elif op == ADVANCE:
            assert pos == a and b > a
            advances += b - a
            pos = b
        elif op == SNAP:
            assert pos == b and a < s
            slots[a] = b
        elif op == RECORD:
            assert pos == a
            recorded.add(a)
            pos = a + 1
        elif op == CONSUME:
            assert a in recorded
            consumed.append(a)
            recorded.discard(a)
        # at most s snapshots alive
        assert len(slots) <= max(s, 0) or s == 0
    assert consumed == list(range(m - 1, -1, -1))
    return advances, adv


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 13, 21, 40, 63])
@pytest.mark.parametrize("s", [0, 1, 2, 3, 5, 9])
def test_schedule_is_valid_and_optimal(m, s):
    advances, adv = simulate(m, s)
    assert advances == adv
    c = min(s, max(m - 1, 0)) + 1
    assert adv + (m - 1 if m > 0 else 0) == griewank_t(m, c), (m, s, adv)


def test_enough_slots_is_uniform_checkpointing():
    m = 17
    acts, adv = S.revolve_schedule(m, m - 1)
    assert adv == 0
    chain = [b for op, a, b in acts if op == SNAP][: m - 1]
    assert chain == list(range(1, m))
