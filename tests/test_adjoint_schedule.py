"""The adjoint's reverse schedule (fsibe_adjoint_plan, host only, no GPU).

A schedule is checked by playing it: every restore reads a slot that holds
the state of the unit it names (or the initial state for unit 0), every
advance and store starts from the current state, at most `nslots` slots are
ever in use, and every unit is recorded and reversed once, last to first,
from its own start.  Its advance count is checked against the least count
of the recursion over split schedules, computed here by brute-force dynamic
programming (no closed form), and against the binomial closed form the
library reports (fsibe_adjoint_plan_cost)."""
from functools import lru_cache
from math import comb

import pytest

from fsibe_2d_amd.adjoint import plan, plan_cost

RESTORE, ADVANCE, SAVE, REVERSE = 0, 1, 2, 3


@lru_cache(None)
def dp_cost(l, s):
    """least units advanced over all split schedules: advance k units, store,
    reverse the top l - k with s - 1 slots, then the bottom k with s"""
    if l <= 1:
        return 0
    if s == 0:
        return sum(range(l))  # from unit 0 to each unit in turn
    return min(k + dp_cost(l - k, s - 1) + dp_cost(k, s) for k in range(1, l))


def play(nunits, nslots):
    acts, adv = plan(nunits, nslots)
    slots = {}
    cur, order, advanced = 0, [], 0
    lowest = nunits  # every unit from `lowest` up has been reversed
    for op, unit, arg in acts:
        if op == RESTORE:
            if arg < 0:
                assert unit == 0
            else:
                assert slots.get(arg) == unit, (op, unit, arg, slots)
            cur = unit
        elif op == ADVANCE:
            assert cur == unit and arg >= 1
            cur += arg
            advanced += arg
            assert cur < nunits
        elif op == SAVE:
            assert cur == unit and 0 <= arg < max(nslots, 1)
            # a slot is reused only once the unit it held no longer needs it:
            # every unit above the stored one has been reversed
            old = slots.get(arg)
            if old is not None:
                assert lowest <= old, (arg, old)
            slots[arg] = unit
            assert len(slots) <= nslots
        elif op == REVERSE:
            assert cur == unit and unit == lowest - 1
            order.append(unit)
            lowest = unit
            cur = unit + 1
        else:
            raise AssertionError(op)
    assert order == list(range(nunits - 1, -1, -1))
    assert advanced == adv
    return adv


@pytest.mark.parametrize("nslots", [0, 1, 2, 3, 5, 8])
def test_schedule_plays_and_is_least(nslots):
    for l in range(1, 61):
        adv = play(l, nslots)
        s = min(nslots, l - 1)
        assert adv == dp_cost(l, s), (l, nslots)
        assert adv == plan_cost(l, nslots), (l, nslots)


def test_cost_closed_form_and_limits():
    # one slot per unit start: one forward sweep, l - 1 units advanced
    for l in (1, 2, 10, 100, 1000):
        assert plan_cost(l, l) == l - 1
        assert plan_cost(l, l - 1) == l - 1
    # no slot: from unit 0 to every unit
    for l in (1, 2, 7, 50):
        assert plan_cost(l, 0) == l * (l - 1) // 2
    # the binomial form against the recursion at sizes the recursion reaches
    for s in range(1, 7):
        for l in (20, 57, 120):
            assert plan_cost(l, s) == dp_cost(l, s), (l, s)
    # the largest l reversible with r = 2 repetitions and s slots is
    # C(s + 3, 2); there t = 2 l - (s + 3)
    for s in (1, 4, 10):
        l = comb(s + 3, 2)
        assert plan_cost(l, s) == 2 * l - (s + 3)


def test_long_horizon_schedule():
    """10 000 units with 20 slots: the schedule is valid and agrees with the
    closed form, which advances each unit fewer than r = 4 times on average
    (C(21 + 3, 3) = 2024 < 10 000 <= C(21 + 4, 4) = 12 650)"""
    adv = play(10000, 20)
    assert adv == plan_cost(10000, 20) == 4 * 10000 - comb(22 + 3, 3)
    assert adv < 4 * 10000


def test_extreme_slot_counts():
    """a slot per unit start for 100 000 units (one sweep storing each unit's
    start, built without deep recursion) and a single slot for 3000 units
    (t = r l - C(r + 2, r - 1) with r ~ sqrt(2 l))"""
    assert play(100000, 99999) == 99999
    adv = play(3000, 1)
    r = next(r for r in range(1, 200) if comb(2 + r, r) >= 3000)
    assert adv == plan_cost(3000, 1) == r * 3000 - comb(3 + r - 1, r - 1)


def test_bad_sizes_are_refused():
    from fsibe_2d_amd._lib import lib

    assert lib().fsibe_adjoint_plan_cost(0, 3) == -1
    assert lib().fsibe_adjoint_plan_cost(5, -1) == -1
    assert lib().fsibe_adjoint_plan(0, 3, None, 0, None, None) != 0
