"""Batched scenes (one f32c state holding several independent scenes) on the GPU.

Bar: every scene of a batch evolves, differentiates and scores exactly as the
same scene run alone in its own f32c state -- X, V, J, gX, gV and gtheta bit
for bit -- since each workgroup works inside one scene with that scene's
constants and the sums keep their order.
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


def alone(be, sc, nsteps, seedV_scale):
    pg = S.SPHParams(**sc.params)
    st = S.State(be, sc.X0, sc.V0, pg, precision="f32c")
    tape = S.Tape(be, sc.N, nsteps, pg, precision="f32c")
    S.simulate_(st, sc.theta, pg, be, nsteps, tape=tape)
    J, seed = S.target_objective(st, 0.3, 0.2, 0.1)
    sv = seedV_scale * np.ones_like(seed)
    ws = S.AdjointWorkspace(be, sc.N, pg, precision="f32c")
    S.backward_(ws, tape, sc.theta, pg, be, seedX=seed, seedV=sv)
    return st.X, st.V, J, ws.gX, ws.gV, ws.gtheta


@pytest.mark.parametrize("keep", [False, True])
@pytest.mark.parametrize("nsteps", [1, 9, 30])
def test_batch_equals_scenes_alone(be, nsteps, keep):
    """A batch of four scenes equals each scene run alone (plain tapes), bit
    for bit; with keep=True the batch's tape also keeps (pterm, invrho) and
    the exact masks."""
    say("batch of 4 vs alone, steps", nsteps, "keep", keep)
    scs = scenes.batch_line_search(seed=3, nscenes=4, nmin=300, nmax=5000, nsteps=nsteps)
    bst = S.BatchState(be, [(S.SPHParams(**sc.params), sc.X0, sc.V0) for sc in scs])
    bth = bst.field([sc.theta for sc in scs])
    tape = S.Tape(be, bst.N, nsteps, bst.p, precision="f32c", keep_pi=keep, keep_lists=keep)
    S.simulate_(bst, bth, bst.p, be, nsteps, tape=tape)
    J, seed = S.batch_objective(bst, 0.3, 0.2, 0.1)
    ws = S.BatchWorkspace(be, bst)
    S.backward_(ws, tape, bth, bst.p, be, seedX=seed, seedV=0.5 * np.ones_like(seed))
    Xs, Vs, gXs, gVs = bst.split(bst.X), bst.split(bst.V), bst.split(ws.gX), bst.split(ws.gV)
    gth = ws.gtheta
    for k, sc in enumerate(scs):
        X, V, Jk, gX, gV, gt = alone(be, sc, nsteps, 0.5)
        np.testing.assert_array_equal(Xs[k], X, err_msg=f"X scene {k}")
        np.testing.assert_array_equal(Vs[k], V, err_msg=f"V scene {k}")
        assert J[k] == Jk
        np.testing.assert_array_equal(gXs[k], gX, err_msg=f"gX scene {k}")
        np.testing.assert_array_equal(gVs[k], gV, err_msg=f"gV scene {k}")
        np.testing.assert_array_equal(gth[k], gt, err_msg=f"gtheta scene {k}")


def test_batch_rejects_single_scene_only_calls(be):
    say("batch API guards")
    scs = scenes.batch_line_search(seed=1, nscenes=2, nmin=300, nmax=600, nsteps=2)
    bst = S.BatchState(be, [(S.SPHParams(**sc.params), sc.X0, sc.V0) for sc in scs])
    with pytest.raises(Exception):
        S.target_objective(bst, 0.3, 0.2, 0.1)
    with pytest.raises(Exception):
        bst.cl.arrays()
    wrong = S.Field(be, np.zeros((5, 5)))
    with pytest.raises(Exception):
        S.step_(bst, wrong)


def test_autotune_keeps_results(be):
    say("autotune: any workgroup size gives the same numbers")
    sc = scenes.brinkman_56x37(nsteps=0)
    pg = S.SPHParams(**sc.params)
    ref = S.State(be, sc.X0, sc.V0, pg, precision="f32c")
    st = S.State(be, sc.X0, sc.V0, pg, precision="f32c")
    chosen = S.autotune_(st, sc.theta, steps=4)
    assert set(chosen.values()) <= {64, 128, 256, -64, -128, -256}
    S.simulate_(ref, sc.theta, pg, be, 48)          # the tuner advanced st by 12*4 steps
    np.testing.assert_array_equal(st.X, ref.X)
    np.testing.assert_array_equal(st.V, ref.V)
    # forward + adjoint with the tuned sizes equal the default ones
    outs = []
    for s_ in (ref, st):
        tape = S.Tape(be, sc.N, 12, pg, precision="f32c")
        S.simulate_(s_, sc.theta, pg, be, 12, tape=tape)
        J, seed = S.target_objective(s_, 0.3, 0.2, 0.1)
        ws = S.AdjointWorkspace(be, sc.N, pg, precision="f32c")
        S.backward_(ws, tape, sc.theta, pg, be, seedX=seed, seedV=np.zeros_like(seed))
        outs.append((J, ws.gX, ws.gV, ws.gtheta))
    assert outs[0][0] == outs[1][0]
    for a, b in zip(outs[0][1:], outs[1][1:]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("size,keep", [(64, "plain"), (128, "plain"), (64, "pi"),
                                       (-256, "pi_lists"), (-64, "plain"),
                                       (128, "pi_lists")])
def test_small_workgroups_stage_the_whole_window(be, size, keep):
    """With 64- or 128-thread workgroups the pair passes still deal each
    256-slot window by height, so every workgroup stages its whole window's
    rows: forward + adjoint equal the 256-thread launches bit for bit, on a
    batch (scene-local windows).  At 256 threads P1 and the gtheta gather
    share one launch, at the smaller sizes they launch apart; the tapes cover
    P1 over the exact list (plain, pi_lists) and over the mask list (pi).  A
    negative size launches the passes without LDS staging (neighbours read
    from global memory): the same results again."""
    say("workgroup size", size, "tape", keep)
    scs = scenes.batch_line_search(seed=5, nscenes=3, nmin=800, nmax=6000, nsteps=16)
    args = [(S.SPHParams(**sc.params), sc.X0, sc.V0) for sc in scs]
    outs = []
    for sz in (256, size):
        bst = S.BatchState(be, args)
        names = ("nl_build", "density", "force_integrate", "adjoint_r", "adjoint_p1",
                 "adjoint_p2")
        S.tuning(bst, {n: sz for n in names})
        th = bst.field([sc.theta for sc in scs])
        tape = S.Tape(be, bst.N, 16, bst.p, precision="f32c", keep_pi=keep != "plain",
                      keep_lists=keep == "pi_lists")
        S.simulate_(bst, th, bst.p, be, 16, tape=tape)
        J, seed = S.batch_objective(bst, 0.3, 0.2, 0.1)
        ws = S.BatchWorkspace(be, bst)
        S.backward_(ws, tape, th, bst.p, be, seedX=seed, seedV=np.zeros_like(seed))
        outs.append((bst.X, J, ws.gX, ws.gV, ws.gtheta))
    a, b = outs
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(a[3], b[3])
    for x, y in zip(a[4], b[4]):
        np.testing.assert_array_equal(x, y)
