"""Binomial (revolve) reversal schedule used inside each checkpoint interval
of the adjoint pass: every step is reversed once, in decreasing order, from
the state that precedes it, with the forward-step count of Griewank's
optimal schedule."""
from functools import lru_cache
from math import comb

import pytest

import lbm


def simulate(n, slots):
    """Replay the schedule: returns (forward steps, reversed steps in order)."""
    held = {-1: 0}  # buffer -> state index held
    fwd, order = 0, []
    for op, a, b, bf, bt in lbm.checkpoint_schedule(n, slots):
        assert held.get(bf) == a, (op, a, b, bf, bt, held)
        if op == 0:
            assert b > a and bt != -1
            assert bt == -2 or 0 <= bt < slots
            held[bt] = b
            fwd += b - a
        else:
            assert op == 1 and b == a + 1
            order.append(b)
    return fwd, order


def optimal(n, s):
    """Minimal forward steps to reverse n steps with s stored states beside
    the interval's start state and one scratch state: Griewank's
    r n - C(c + r, c + 1), r least with C(c + r, c) >= n, for c = s + 1
    buffers; recompute-from-start, n (n - 1) / 2, when s = 0."""
    if n <= 1:
        return 0
    if s == 0:
        return n * (n - 1) // 2
    c = s + 1
    r = 1
    while comb(c + r, c) < n:
        r += 1
    return r * n - comb(c + r, c + 1)


@lru_cache(None)
def brute(n, s):
    """The same minimum by exhaustive split search."""
    if n <= 1:
        return 0
    if s == 0:
        return n * (n - 1) // 2
    return min(m + brute(n - m, s - 1) + brute(m, s) for m in range(1, n))


@pytest.mark.parametrize("n", list(range(1, 26)) + [64, 200])
def test_every_step_reversed_once_in_order(n):
    for s in sorted({0, 1, 2, 3, 5, n // 2, n - 1, n + 3}):
        if s < 0:
            continue
        fwd, order = simulate(n, s)
        assert order == list(range(n, 0, -1))
        k = min(s, max(n - 1, 0))
        assert fwd == optimal(n, k), (n, s, fwd)
        if n <= 30:
            assert fwd == brute(n, k)


def test_store_all_and_single_slot_costs():
    assert simulate(200, 199)[0] == 199
    assert simulate(30, 0)[0] == 30 * 29 // 2
    assert simulate(30, 1)[0] == optimal(30, 1) < 30 * 29 // 2
    assert simulate(1, 0) == (0, [1])
