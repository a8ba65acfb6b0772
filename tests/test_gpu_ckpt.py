"""Binomial checkpointing on the GPU, through the C ABI.

Bar: the checkpointed forward ends in the same state as simulate_ and the
checkpointed adjoint gives the same gX, gV, gtheta as backward_ over a full
tape, bit for bit, in both arithmetic forms: a restored snapshot sits at a
neighbour-list rebuild, which is a pure function of the positions and ids, so
every recomputed epoch repeats the original one exactly.  The recomputation
count equals the host schedule's.
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


def full_tape(be, sc, nsteps, interval, cX, cV, precision, **keep):
    pg = S.SPHParams(**sc.params)
    st = S.State(be, sc.X0, sc.V0, pg, interval=interval, precision=precision)
    tape = S.Tape(be, sc.N, nsteps, pg, precision=precision, **keep)
    S.simulate_(st, sc.theta, pg, be, nsteps, tape=tape)
    X, V = st.X, st.V
    ws = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
    S.backward_(ws, tape, sc.theta, pg, be, seedX=cX, seedV=cV, interval=interval)
    return X, V, ws.gX, ws.gV, ws.gtheta


def checkpointed(be, sc, nsteps, interval, snapshots, seg, cX, cV, precision, repeat=1):
    pg = S.SPHParams(**sc.params)
    ck = S.Checkpointer(be, sc.N, nsteps, pg, snapshots=snapshots, seg_epochs=seg,
                        interval=interval, precision=precision)
    ws = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
    out = []
    for _ in range(repeat):
        st = ck.forward(sc.X0, sc.V0, sc.theta)
        X, V = st.X, st.V
        ck.backward(ws, seedX=cX, seedV=cV)
        out.append((X, V, ws.gX, ws.gV, ws.gtheta))
    return out, ck.stats()


CASES = [
    # (steps, interval, snapshots, epochs per segment): ragged last epoch, no
    # snapshot at all, one step per epoch, more snapshots than segments, a
    # single epoch, multi-epoch segments (ragged, exact, every segment stored)
    (37, 4, 3, 1),
    (30, 4, 0, 1),
    (25, 1, 2, 1),
    (16, 4, 10, 1),
    (3, 4, 2, 1),
    (37, 4, 2, 3),
    (40, 2, 1, 4),
    (37, 4, 20, 2),
]


@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("nsteps,interval,snapshots,seg", CASES)
def test_checkpointed_equals_full_tape(be, precision, nsteps, interval, snapshots, seg):
    say("ckpt vs full tape", precision, nsteps, interval, snapshots, seg)
    sc = scenes.brinkman_56x37(nsteps=nsteps)
    rng = np.random.default_rng(nsteps * 31 + snapshots)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    want = full_tape(be, sc, nsteps, interval, cX, cV, precision)
    got, stats = checkpointed(be, sc, nsteps, interval, snapshots, seg, cX, cV, precision)
    for a, b, name in zip(got[0], want, ("X", "V", "gX", "gV", "gtheta")):
        np.testing.assert_array_equal(a, b, err_msg=name)
    E = -(-nsteps // (interval * seg))
    plan = S.ckpt_plan(E, min(snapshots, E))
    assert stats["epochs"] == E
    assert stats["forward_epochs"] + stats["backward_forward_epochs"] == plan["forward_epochs"]
    assert stats["backward_taped_epochs"] == E - 1


@pytest.mark.parametrize("precision,keep", [("f64", dict(keep_pi=True)),
                                             ("f32c", dict(keep_pi=True)),
                                             ("f32c", dict(keep_lists=True)),
                                             ("f32c", dict(keep_pi=True, keep_lists=True))])
@pytest.mark.parametrize("nsteps,interval", [(37, 4), (25, 1)])
def test_keep_tape_equals_plain_tape(be, precision, keep, nsteps, interval):
    """A tape that keeps (pterm, invrho) and/or the exact masks of each step
    (the reverse sweep then skips the density recompute and the pair tests,
    and in the f32c form fuses each R pass into the previous P2) gives the
    plain tape's gradients bit for bit."""
    say("keep tape vs plain tape", precision, keep, nsteps, interval)
    sc = scenes.brinkman_56x37(nsteps=nsteps)
    rng = np.random.default_rng(nsteps + 7)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    want = full_tape(be, sc, nsteps, interval, cX, cV, precision)
    got = full_tape(be, sc, nsteps, interval, cX, cV, precision, **keep)
    for a, b, name in zip(got, want, ("X", "V", "gX", "gV", "gtheta")):
        np.testing.assert_array_equal(a, b, err_msg=name)


def test_keep_tape_rejects_fp64_lists(be):
    sc = scenes.gradcheck_5x5()
    with pytest.raises(Exception):
        S.Tape(be, sc.N, 4, S.SPHParams(**sc.params), precision="f64", keep_lists=True)


def test_checkpointer_reuse_and_gaussian_seed(be):
    say("ckpt reuse, gaussian target")
    sc = scenes.brinkman_56x37(nsteps=24)
    pg = S.SPHParams(**sc.params)
    ck = S.Checkpointer(be, sc.N, 24, pg, snapshots=2)
    ws = S.AdjointWorkspace(be, sc.N, pg)
    res = []
    for _ in range(2):
        st = ck.forward(sc.X0, sc.V0, sc.theta)
        J, seed = S.target_objective(st, 0.3, 0.2, 0.1)
        ck.backward(ws, seedX=seed, seedV=np.zeros_like(seed))
        res.append((J, ws.gtheta, ws.gX))
    assert res[0][0] == res[1][0]
    np.testing.assert_array_equal(res[0][1], res[1][1])
    np.testing.assert_array_equal(res[0][2], res[1][2])
    # the same through the full tape
    st = S.State(be, sc.X0, sc.V0, pg)
    tape = S.Tape(be, sc.N, 24, pg)
    S.simulate_(st, sc.theta, pg, be, 24, tape=tape)
    J, seed = S.target_objective(st, 0.3, 0.2, 0.1)
    S.backward_(ws, tape, sc.theta, pg, be, seedX=seed, seedV=np.zeros_like(seed))
    assert J == res[0][0]
    np.testing.assert_array_equal(ws.gtheta, res[0][1])


def test_checkpointer_backward_needs_forward(be):
    say("ckpt backward without forward")
    sc = scenes.random_box(200, seed=3)
    pg = S.SPHParams(**sc.params)
    ck = S.Checkpointer(be, sc.N, 8, pg, snapshots=1)
    ws = S.AdjointWorkspace(be, sc.N, pg)
    z = np.zeros((2, sc.N))
    with pytest.raises(Exception):
        ck.backward(ws, seedX=z, seedV=z, theta=sc.theta)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_torch_autograd_matches_adjoint(be, precision):
    say("torch autograd wrapper", precision)
    import torch
    from sphadjoint_amd import autograd as AG
    sc = scenes.brinkman_56x37(nsteps=20)
    pg = S.SPHParams(**sc.params)
    X0 = torch.tensor(sc.X0, dtype=torch.float64, requires_grad=True)
    V0 = torch.tensor(sc.V0, dtype=torch.float64, requires_grad=True)
    res = {}
    for snaps in (None, 2):
        sim = AG.Simulation(be, pg, sc.N, 20, snapshots=snaps, precision=precision)
        th = torch.tensor(sc.theta, dtype=torch.float64, requires_grad=True)
        XT, VT = AG.simulate(th, X0, V0, sim)
        J = AG.gaussian_target(XT, 0.3, 0.2, 0.1) + 0.5 * (VT * VT).sum()
        J.backward()
        res[snaps] = (float(J), th.grad.clone(), X0.grad.clone(), V0.grad.clone())
        X0.grad = None
        V0.grad = None
    for a, b in zip(res[None], res[2]):
        if isinstance(a, float):
            assert a == b
        else:
            assert torch.equal(a, b)
    # against the direct path, seeds formed the same way in torch
    st = S.State(be, sc.X0, sc.V0, pg, precision=precision)
    tape = S.Tape(be, sc.N, 20, pg, precision=precision)
    S.simulate_(st, sc.theta, pg, be, 20, tape=tape)
    Xt = torch.tensor(st.X, requires_grad=True)
    Vt = torch.tensor(st.V, requires_grad=True)
    (AG.gaussian_target(Xt, 0.3, 0.2, 0.1) + 0.5 * (Vt * Vt).sum()).backward()
    ws = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
    S.backward_(ws, tape, sc.theta, pg, be, seedX=Xt.grad.numpy(), seedV=Vt.grad.numpy())
    np.testing.assert_array_equal(res[None][1].numpy(), ws.gtheta)
    np.testing.assert_array_equal(res[None][2].numpy(), ws.gX)
    np.testing.assert_array_equal(res[None][3].numpy(), ws.gV)


def test_device_memory_variants(be):
    say("device-memory inputs and seeds")
    sc = scenes.brinkman_56x37(nsteps=20)
    pg = S.SPHParams(**sc.params)
    for precision in ("f64", "f32c"):
        ck = S.Checkpointer(be, sc.N, 20, pg, snapshots=2, seg_epochs=2, precision=precision)
        ws = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
        st = ck.forward(sc.X0, sc.V0, sc.theta)
        J, seed = S.target_objective(st, 0.3, 0.2, 0.1)
        ck.backward(ws, seedX=seed, seedV=np.zeros_like(seed))
        host = (J, seed, ws.gX, ws.gV, ws.gtheta)
        X0 = S.DeviceArray(be, sc.X0)
        V0 = S.DeviceArray(be, sc.V0)
        zero = S.DeviceArray(be, N=sc.N)
        np.testing.assert_array_equal(X0.numpy(), sc.X0)
        st = ck.forward(X0, V0, sc.theta)
        Jd, seedd = S.target_objective(st, 0.3, 0.2, 0.1, device=True)
        ck.backward(ws, seedX=seedd, seedV=zero)
        assert Jd == host[0]
        np.testing.assert_array_equal(seedd.numpy(), host[1])
        np.testing.assert_array_equal(ws.gX, host[2])
        np.testing.assert_array_equal(ws.gV, host[3])
        np.testing.assert_array_equal(ws.gtheta, host[4])
        # full tape with device seeds
        st = S.State(be, sc.X0, sc.V0, pg, precision=precision)
        tape = S.Tape(be, sc.N, 20, pg, precision=precision)
        S.simulate_(st, sc.theta, pg, be, 20, tape=tape)
        Jt, seedt = S.target_objective(st, 0.3, 0.2, 0.1, device=True)
        S.backward_(ws, tape, sc.theta, pg, be, seedX=seedt, seedV=zero)
        assert Jt == host[0]
        np.testing.assert_array_equal(ws.gtheta, host[4])


@pytest.mark.slow
def test_long_horizon_5333_steps(be):
    """N=2072 Brinkman scene, 5333 steps: the f32c form under a 64-snapshot
    budget gives J bit-identical over 4 repeated runs, and its gradients stay
    within the FP32-compensated form's error of a full-tape FP64 run."""
    say("long horizon 5333 steps, 64 snapshots, f32c x4 vs f64 full tape")
    n = 5333
    sc = scenes.brinkman_56x37(nsteps=n)
    pg = S.SPHParams(**sc.params)
    # FP64 full tape
    st = S.State(be, sc.X0, sc.V0, pg)
    tape = S.Tape(be, sc.N, n, pg)
    S.simulate_(st, sc.theta, pg, be, n, tape=tape)
    J64, seed = S.target_objective(st, 0.6, 0.2, 0.1)
    ws = S.AdjointWorkspace(be, sc.N, pg)
    S.backward_(ws, tape, sc.theta, pg, be, seedX=seed, seedV=np.zeros_like(seed))
    ref = (ws.gX, ws.gV, ws.gtheta)
    # the horizon's own sensitivity, for the record: FP64 again from X0 moved
    # by a relative 1e-10 (normal), and how far its gradients move
    rng = np.random.default_rng(11)
    X0p = sc.X0 * (1.0 + 1e-10 * rng.standard_normal(sc.X0.shape))
    st = S.State(be, X0p, sc.V0, pg)
    S.reset_(tape)
    S.simulate_(st, sc.theta, pg, be, n, tape=tape)
    Jp, seedp = S.target_objective(st, 0.6, 0.2, 0.1)
    S.backward_(ws, tape, sc.theta, pg, be, seedX=seedp, seedV=np.zeros_like(seedp))
    dp = [float(np.abs(a - b).max() / np.abs(b).max())
          for a, b in zip((ws.gX, ws.gV, ws.gtheta), ref)]
    print("   FP64 from X0 (1 + 1e-10 N(0,1)): J %.3e  gX %.3e  gV %.3e  gtheta %.3e"
          % (abs(Jp - J64) / abs(J64), *dp))
    del tape
    ck = S.Checkpointer(be, sc.N, n, pg, snapshots=64, precision="f32c")
    wsf = S.AdjointWorkspace(be, sc.N, pg, precision="f32c")
    runs = []
    for _ in range(4):
        stf = ck.forward(sc.X0, sc.V0, sc.theta)
        J, sd = S.target_objective(stf, 0.6, 0.2, 0.1)
        ck.backward(wsf, seedX=sd, seedV=np.zeros_like(sd))
        runs.append((J, wsf.gX, wsf.gV, wsf.gtheta))
    for r in runs[1:]:
        assert r[0] == runs[0][0]
        for a, b in zip(r[1:], runs[0][1:]):
            np.testing.assert_array_equal(a, b)
    eJ = abs(runs[0][0] - J64) / abs(J64)
    e = [float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(runs[0][1:], ref)]
    print("   f32c vs FP64: J %.3e  gX %.3e  gV %.3e  gtheta %.3e" % (eJ, *e))
    assert eJ < 1e-4
    assert e[2] < 1e-2


@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("nsteps,interval,snapshots,seg,tape_steps", [
    (37, 4, 3, 1, 12),     # two tail segments after the schedule's last
    (37, 4, 0, 1, 40),     # the whole horizon on the tape: no recomputation
    (40, 2, 1, 4, 16),     # multi-epoch segments, one tail segment
    (25, 1, 2, 1, 7),      # one step per epoch
    (37, 4, 2, 3, 5),      # a budget below one segment: one segment
])
def test_tape_budget_equals_full_tape(be, precision, nsteps, interval, snapshots, seg,
                                      tape_steps):
    """With a tape budget the last segments of the horizon are taped in the
    forward and never recomputed; the gradients are the full tape's bit for
    bit, and the reverse sweep recomputes fewer segments."""
    say("ckpt tape budget vs full tape", precision, nsteps, interval, snapshots, seg, tape_steps)
    sc = scenes.brinkman_56x37(nsteps=nsteps)
    rng = np.random.default_rng(nsteps * 13 + tape_steps)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    want = full_tape(be, sc, nsteps, interval, cX, cV, precision)
    pg = S.SPHParams(**sc.params)
    runs = {}
    for budget in (None, tape_steps):
        ck = S.Checkpointer(be, sc.N, nsteps, pg, snapshots=snapshots, seg_epochs=seg,
                            interval=interval, precision=precision, tape_steps=budget)
        ws = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
        st = ck.forward(sc.X0, sc.V0, sc.theta)
        X, V = st.X, st.V
        ck.backward(ws, seedX=cX, seedV=cV)
        runs[budget] = ((X, V, ws.gX, ws.gV, ws.gtheta), ck.stats())
    got, stats = runs[tape_steps]
    for a, b, name in zip(got, want, ("X", "V", "gX", "gV", "gtheta")):
        np.testing.assert_array_equal(a, b, err_msg=name)
    E = -(-nsteps // (interval * seg))
    units = min(max(tape_steps // (interval * seg), 1), E)
    base = runs[None][1]
    print(f"   E={E}, taped units {units}: backward recomputed {stats['backward_forward_epochs']} "
          f"segments (no budget: {base['backward_forward_epochs']})")
    assert stats["epochs"] == E
    assert stats["forward_epochs"] == base["forward_epochs"] or units > 1
    if units > 1:
        assert stats["backward_forward_epochs"] < base["backward_forward_epochs"]
    if units == E:
        assert stats["backward_forward_epochs"] == 0
