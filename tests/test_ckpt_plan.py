"""The binomial checkpointing schedule, dry-run on the host (no GPU).

Pinned against a dynamic program over every split (the optimum the revolve
schedule of Griewank & Walther is proven to reach) and against the closed form
of the minimal recomputation: with s = c + 1 snapshots (the held start
included) and t the least integer such that C(s + t, t) >= l, reversing l
epochs costs t*l - C(s + t, t - 1) advances.
"""
from functools import lru_cache
from math import comb

import pytest

import sphadjoint_amd as S


@lru_cache(maxsize=None)
def dp(l, c):
    """Fewest epoch advances that reverse l epochs whose first state is held,
    with c more snapshots free (each reversed epoch is re-run once with its
    tape, which is counted apart)."""
    if l == 1:
        return 0
    if c == 0:
        return l * (l - 1) // 2
    return min(j + dp(l - j, c - 1) + dp(j, c) for j in range(1, l))


def closed_form(l, c):
    if l == 1:
        return 0
    s = c + 1
    t = 0
    while comb(s + t, t) < l:
        t += 1
    return t * l - comb(s + t, t - 1)


@pytest.mark.parametrize("c", [0, 1, 2, 3, 5, 8])
def test_schedule_is_optimal(c):
    for E in range(1, 61):
        plan = S.ckpt_plan(E, c)
        assert plan["taped"] == E
        assert plan["forward_epochs"] - E == dp(E, c), (E, c)
        assert plan["forward_epochs"] - E == closed_form(E, c), (E, c)
        assert plan["peak"] <= min(c, E)


def test_schedule_large_horizon():
    # 5333 steps at interval 4 -> 1334 epochs; 64 snapshots reach it with t = 2:
    # under two extra forward passes
    E = 1334
    plan = S.ckpt_plan(E, 64)
    assert plan["peak"] <= 64
    assert plan["forward_epochs"] - E == closed_form(E, 64)
    assert plan["forward_epochs"] < 3 * E


def test_schedule_argument_errors():
    with pytest.raises(Exception):
        S.ckpt_plan(0, 4)
    with pytest.raises(Exception):
        S.ckpt_plan(5, -1)
