"""Graph-captured forward steps (sph_simulate_graph) on the GPU.

Bar: the replayed graph leaves the state exactly where eager sph_simulate
does -- X and V bit for bit -- for the FP64 and f32c forms and a batched
state, over several calls (each replaying the cached graph, with eager work
and host reads in between), after the design field changes (a new capture),
and with ragged step counts before and after the whole cycles.
"""
import sys

import numpy as np
import pytest

import sphadjoint_amd as S
from sphadjoint_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def be():
    return S.MI355XBackend(0)


@pytest.mark.parametrize("precision,interval", [("f64", 4), ("f32c", 4), ("f32c", 3)])
def test_graph_equals_eager(be, precision, interval):
    say("graph vs eager", precision, interval)
    sc = scenes.brinkman_56x37(nsteps=0)
    pg = S.SPHParams(**sc.params)
    a = S.State(be, sc.X0, sc.V0, pg, interval=interval, precision=precision)
    b = S.State(be, sc.X0, sc.V0, pg, interval=interval, precision=precision)
    th = S.Field(be, sc.theta)
    replays = 0
    for n in (37, 16, 5, 40):
        S.simulate_(a, th, pg, be, n)
        replays += S.simulate_graph_(b, th, pg, be, n)
        np.testing.assert_array_equal(b.X, a.X)
        np.testing.assert_array_equal(b.V, a.V)
    assert replays >= 8
    # a new design field: the graph is captured again
    th2 = S.Field(be, sc.theta * 0.5)
    S.simulate_(a, th2, pg, be, 24)
    assert S.simulate_graph_(b, th2, pg, be, 24) >= 2
    np.testing.assert_array_equal(b.X, a.X)
    np.testing.assert_array_equal(b.V, a.V)


def test_graph_batched_state(be):
    say("graph on a batched state")
    scs = scenes.batch_line_search(seed=5, nscenes=3, nmin=300, nmax=3000, nsteps=0)
    args = [(S.SPHParams(**sc.params), sc.X0, sc.V0) for sc in scs]
    a = S.BatchState(be, args)
    b = S.BatchState(be, args)
    th = a.field([sc.theta for sc in scs])
    for n in (20, 33):
        S.simulate_(a, th, a.p, be, n)
        S.simulate_graph_(b, th, b.p, be, n)
        np.testing.assert_array_equal(b.X, a.X)
        np.testing.assert_array_equal(b.V, a.V)
