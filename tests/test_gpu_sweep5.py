"""The 5x5 half-cell sweep of the f32c list builds (sph_state_set_sweep), on the GPU.

With sweep 5 each rebuild orders every cell's slots by quadrant, then id, and
builds the mask list from the 5x5 half cells around each slot's quadrant
instead of all slots of the 3x3 cells.  Bars:
  * the exported cell list and neighbour list equal the oracle's bit for bit
    (the exports restore the reference's order);
  * the list is the same: a tape that kept the forward's exact masks (built
    by the half-cell sweep) and a plain tape, whose reverse sweep rebuilds the
    masks of each epoch by the 3x3 sweep on the same layout, give the same
    gradients bit for bit;
  * checkpointed gradients equal the full tape's bit for bit (restored
    snapshots keep their quadrant order);
  * state and gradients against FP64 within the f32c form's bars, on a dam
    break, the Brinkman scene and the 1M-particle block (the summation order
    inside a cell changes, so the f32c results change at rounding level).
"""
import sys
import warnings

import numpy as np
import pytest

import sphadjoint_amd as S
from oracle import oracle as O
from sphadjoint_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def be():
    return S.MI355XBackend(0)


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def cloud(case):
    rng = np.random.default_rng(11)
    if case == "cloud50k":
        N, Lx, Ly, h = 50000, 2.0, 1.2, 0.004
        X = np.vstack([rng.uniform(0, Lx, N), rng.uniform(0, Ly, N)])
    elif case == "outside":
        N, Lx, Ly, h = 3000, 1.0, 0.6, 0.02
        X = np.vstack([rng.uniform(0, Lx, N), rng.uniform(0, Ly, N)])
        X[:, :300] = rng.uniform(-0.3, 1.5, (2, 300)) * np.array([[Lx], [Ly]])
    else:  # one_cell_20k: a pile-up ranked by the big-cell path, with quadrant keys
        N, Lx, Ly, h = 20000, 2.0, 1.2, 0.004
        cs = 2 * h * 1.2
        X = np.vstack([0.7 + cs * rng.uniform(0.01, 0.99, N),
                       0.5 + cs * rng.uniform(0.01, 0.99, N)])
    p = dict(h=h, m=1e-3, rho0=1.0, c=10.0, mu=0.01, dt=1e-5, Lx=Lx, Ly=Ly, kw=1e4,
             ngx=9, ngy=7, gx=0.0, gy=-9.81)
    return X, p


@pytest.mark.parametrize("case", ["dambreak", "cloud50k", "outside", "one_cell_20k"])
def test_sweep5_exports_equal_the_oracle(be, case):
    say("sweep 5 exports", case)
    if case == "dambreak":
        sc = scenes.dambreak_40x40()
        X, p, maxnb = sc.X0, sc.params, 64
    else:
        X, p = cloud(case)
        maxnb = 64
    g = S.State(be, X, np.zeros_like(X), S.SPHParams(**p), precision="f32c", maxnb=maxnb)
    S.set_sweep_(g, 5)
    o = O.State(X, np.zeros_like(X), O.SPHParams(**p), layout="slot", maxnb=maxnb)
    S.build_neighbors_(g.nl)
    if case == "one_cell_20k":
        O.build_(o)    # (the cell list; a list of 20k neighbours each is not the point here)
    else:
        O.build_neighbors_(o)
    for a, b, name in zip(g.cl.arrays(), o.cell_list(), ("cellof", "counts", "starts", "order")):
        np.testing.assert_array_equal(a, b, err_msg=name)
    if case == "one_cell_20k":
        return
    gi, gc = g.nl.arrays()
    oi, oc, of = o.neighbors()
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)


def test_sweep5_pile_up_steps_like_sweep3(be):
    """20k particles in one cell: the rebuild ranks the cell by quadrant keys
    (the big-cell bitmap path), and a few steps track the 3x3 build."""
    say("sweep 5 pile-up steps")
    X, p = cloud("one_cell_20k")
    p = dict(p, dt=1e-7)
    out = []
    for sweep in (3, 5):
        g = S.State(be, X, np.zeros_like(X), S.SPHParams(**p), precision="f32c", interval=1)
        S.set_sweep_(g, sweep)
        S.simulate_(g, np.zeros((7, 9)), S.SPHParams(**p), be, 3)
        out.append((g.X, g.rho))
    ex = np.abs(out[0][0] - out[1][0]).max()
    er = relerr(out[1][1], out[0][1])
    print(f"   X abs {ex:.3e}  rho rel {er:.3e}")
    assert ex < 1e-9 and er < 1e-5


@pytest.mark.parametrize("nsteps", [25, 200])
def test_sweep5_forward_tracks_fp64(be, nsteps):
    say("sweep 5 forward dambreak", nsteps)
    sc = scenes.dambreak_40x40(nsteps=nsteps)
    pg = S.SPHParams(**sc.params)
    po = O.SPHParams(**sc.params)
    g = S.State(be, sc.X0, sc.V0, pg, precision="f32c")
    S.set_sweep_(g, 5)
    o = O.State(sc.X0, sc.V0, po)
    S.simulate_(g, sc.theta, pg, be, nsteps)
    O.simulate_(o, sc.theta, po, nsteps)
    ex = np.abs(g.X - o.X).max()
    ev = relerr(g.V, o.V)
    er = relerr(g.rho, o.rho)
    print(f"   X abs {ex:.3e}  V rel {ev:.3e}  rho rel {er:.3e}")
    # the bars of the 3x3 build (tests/test_gpu_f32c.py)
    assert ex < (1e-8 if nsteps <= 25 else 1e-7)
    assert er < 1e-6
    assert ev < 1e-3


def tape_run(be, sc, nsteps, interval, cX, cV, sweep, **keep):
    pg = S.SPHParams(**sc.params)
    st = S.State(be, sc.X0, sc.V0, pg, interval=interval, precision="f32c")
    S.set_sweep_(st, sweep)
    tape = S.Tape(be, sc.N, nsteps, pg, precision="f32c", **keep)
    S.simulate_(st, sc.theta, pg, be, nsteps, tape=tape)
    X, V = st.X, st.V
    ws = S.AdjointWorkspace(be, sc.N, pg, precision="f32c")
    S.backward_(ws, tape, sc.theta, pg, be, seedX=cX, seedV=cV, interval=interval)
    J = float((X * cX).sum() + (V * cV).sum())
    return J, X, V, ws.gX, ws.gV, ws.gtheta


@pytest.mark.parametrize("nsteps,interval", [(37, 4), (25, 1)])
def test_sweep5_masks_equal_the_3x3_masks(be, nsteps, interval):
    say("sweep 5 kept masks vs rebuilt 3x3 masks", nsteps, interval)
    sc = scenes.brinkman_56x37(nsteps=nsteps)
    rng = np.random.default_rng(nsteps + 3)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    want = tape_run(be, sc, nsteps, interval, cX, cV, 5)
    got = tape_run(be, sc, nsteps, interval, cX, cV, 5, keep_lists=True)
    for a, b, name in zip(got[1:], want[1:], ("X", "V", "gX", "gV", "gtheta")):
        np.testing.assert_array_equal(a, b, err_msg=name)


@pytest.mark.parametrize("nsteps,interval,snapshots,seg", [(37, 4, 3, 1), (37, 4, 2, 3),
                                                           (25, 1, 2, 1)])
def test_sweep5_checkpointed_equals_full_tape(be, nsteps, interval, snapshots, seg):
    say("sweep 5 ckpt vs full tape", nsteps, interval, snapshots, seg)
    sc = scenes.brinkman_56x37(nsteps=nsteps)
    rng = np.random.default_rng(nsteps * 7 + snapshots)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    want = tape_run(be, sc, nsteps, interval, cX, cV, 5)
    pg = S.SPHParams(**sc.params)
    ck = S.Checkpointer(be, sc.N, nsteps, pg, snapshots=snapshots, seg_epochs=seg,
                        interval=interval, precision="f32c")
    S.set_sweep_(ck, 5)
    ws = S.AdjointWorkspace(be, sc.N, pg, precision="f32c")
    st = ck.forward(sc.X0, sc.V0, sc.theta)
    X, V = st.X, st.V
    ck.backward(ws, seedX=cX, seedV=cV)
    for a, b, name in zip((X, V, ws.gX, ws.gV, ws.gtheta), want[1:],
                          ("X", "V", "gX", "gV", "gtheta")):
        np.testing.assert_array_equal(a, b, err_msg=name)


def test_sweep5_adjoint_brinkman_200_vs_fp64(be):
    say("sweep 5 adjoint, brinkman 56x37, 200 steps, vs FP64 GPU")
    sc = scenes.brinkman_56x37(nsteps=200)
    rng = np.random.default_rng(5)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    pg = S.SPHParams(**sc.params)
    st = S.State(be, sc.X0, sc.V0, pg, interval=4)
    tape = S.Tape(be, sc.N, 200, pg)
    S.simulate_(st, sc.theta, pg, be, 200, tape=tape)
    ws = S.AdjointWorkspace(be, sc.N, pg)
    S.backward_(ws, tape, sc.theta, pg, be, seedX=cX, seedV=cV, interval=4)
    J64 = float((st.X * cX).sum() + (st.V * cV).sum())
    J32, _, _, gX, gV, gth = tape_run(be, sc, 200, 4, cX, cV, 5)
    eJ = abs(J32 - J64) / abs(J64)
    e = [relerr(gX, ws.gX), relerr(gV, ws.gV), relerr(gth, ws.gtheta)]
    print("   J %.3e  gX %.3e  gV %.3e  gtheta %.3e" % (eJ, *e))
    assert eJ < 1e-5
    assert e[2] < 1e-5
    assert max(e[:2]) < 2e-5


def test_sweep5_block_1m_checkpointed_vs_fp64(be):
    """The 1M-particle block, 100 steps forward + adjoint: f32c checkpointed
    with the half-cell sweep against the FP64 full tape, and against the 3x3
    build (the same bars as the 500-step configuration test)."""
    say("sweep 5, 1M block, 100 steps, f32c ckpt vs FP64")
    nsteps = 100
    sc = scenes.block_1m(nsteps=nsteps)
    pg = S.SPHParams(**sc.params)
    seed = lambda st: S.target_objective(st, 0.5, 0.5, 0.1)
    st = S.State(be, sc.X0, sc.V0, pg)
    tape = S.Tape(be, sc.N, nsteps, pg)
    with warnings.catch_warnings():
        # the block's surface particles raise the reference's displacement warning
        warnings.simplefilter("ignore")
        S.simulate_(st, sc.theta, pg, be, nsteps, tape=tape)
    J64, sX = seed(st)
    ws64 = S.AdjointWorkspace(be, sc.N, pg)
    S.backward_(ws64, tape, sc.theta, pg, be, seedX=sX, seedV=np.zeros_like(sX))
    g64 = (ws64.gX, ws64.gV, ws64.gtheta)
    del tape
    res = {}
    for sweep in (3, 5):
        ck = S.Checkpointer(be, sc.N, nsteps, pg, snapshots=8, precision="f32c")
        S.set_sweep_(ck, sweep)
        ws = S.AdjointWorkspace(be, sc.N, pg, precision="f32c")
        st = ck.forward(sc.X0, sc.V0, sc.theta, check_flags=False)
        J, sX = seed(st)
        ck.backward(ws, seedX=sX, seedV=np.zeros_like(sX))
        res[sweep] = (J, ws.gX, ws.gV, ws.gtheta)
        del ck, ws
    J5, gX5, gV5, gth5 = res[5]
    eJ = abs(J5 - J64) / abs(J64)
    e = [relerr(gX5, g64[0]), relerr(gV5, g64[1]), relerr(gth5, g64[2])]
    d = [relerr(a, b) for a, b in zip(res[5][1:], res[3][1:])]
    print("   vs FP64: J %.3e  gX %.3e  gV %.3e  gtheta %.3e" % (eJ, *e))
    print("   vs 3x3:  J %.3e  gX %.3e  gV %.3e  gtheta %.3e"
          % (abs(J5 - res[3][0]) / abs(res[3][0]), *d))
    assert eJ < 1e-9
    assert e[2] < 1e-5
    assert max(e[:2]) < 1e-4


def test_set_sweep_rejections(be):
    sc = scenes.brinkman_56x37(nsteps=4)
    pg = S.SPHParams(**sc.params)
    st64 = S.State(be, sc.X0, sc.V0, pg)
    with pytest.raises(Exception, match="f32c"):
        S.set_sweep_(st64, 5)
    S.set_sweep_(st64, 3)
    st = S.State(be, sc.X0, sc.V0, pg, precision="f32c")
    with pytest.raises(Exception, match="3 or 5"):
        S.set_sweep_(st, 4)
    b = S.BatchState(be, [(pg, sc.X0, sc.V0), (pg, sc.X0, sc.V0)])
    with pytest.raises(Exception, match="one scene"):
        S.set_sweep_(b, 5)
