"""GPU parity of the plain FP64 path against the CPU oracle, through the C ABI.

Bar: integer outputs (cell list, neighbour list, flags) bit-exact; FP64
forward state bit-exact (same neighbour order, same operations, no
contraction); adjoint gX/gV bit-exact when the reverse rebuild schedule is the
reference's (interval=1) and within 1e-9 relative otherwise; gtheta, whose
scatter uses float atomics as the reference's does, within 1e-12 relative.
"""
import sys

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


def pair(sc, be, **kw):
    pg = S.SPHParams(**sc.params)
    po = O.SPHParams(**sc.params)
    g = S.State(be, sc.X0, sc.V0, pg, skin=sc.skin, interval=kw.get("interval", sc.interval),
                maxnb=sc.maxnb)
    o = O.State(sc.X0, sc.V0, po, skin=sc.skin, interval=kw.get("interval", sc.interval),
                maxnb=sc.maxnb, layout="slot")
    return g, o, pg, po


def random_cloud(N, Lx, Ly, h, seed, outside=0):
    rng = np.random.default_rng(seed)
    X = np.vstack([rng.uniform(0, Lx, N), rng.uniform(0, Ly, N)])
    if outside:
        X[:, :outside] = rng.uniform(-0.3, 1.5, (2, outside)) * np.array([[Lx], [Ly]])
    p = dict(h=h, m=1e-3, rho0=1.0, c=10.0, mu=0.01, dt=1e-5, Lx=Lx, Ly=Ly, kw=1e4,
             ngx=9, ngy=7, gx=0.0, gy=-9.81)
    return scenes.Scene("cloud", p, X, np.zeros_like(X), np.zeros((7, 9)), 0)


@pytest.mark.parametrize("case", ["random_box", "dambreak", "cloud50k", "outside"])
def test_cell_list_bit_exact(be, case):
    say("cell list", case)
    if case == "random_box":
        sc = scenes.random_box(400, seed=42)
    elif case == "dambreak":
        sc = scenes.dambreak_40x40()
    elif case == "cloud50k":
        sc = random_cloud(50000, 2.0, 1.2, 0.004, seed=1)
    else:
        sc = random_cloud(3000, 1.0, 0.6, 0.02, seed=2, outside=300)
    g, o, _, _ = pair(sc, be)
    S.build_(g.cl)
    O.build_(o)
    got = g.cl.arrays()
    want = o.cell_list()
    for a, b, name in zip(got, want, ("cellof", "counts", "starts", "order")):
        np.testing.assert_array_equal(a, b, err_msg=name)


@pytest.mark.parametrize("case", ["random_box", "dambreak", "cloud20k"])
def test_neighbor_list_bit_exact(be, case):
    say("neighbour list", case)
    if case == "random_box":
        sc = scenes.random_box(400, seed=42)
    elif case == "dambreak":
        sc = scenes.dambreak_40x40()
    else:
        sc = random_cloud(20000, 1.0, 0.6, 0.004, seed=3)
    g, o, _, _ = pair(sc, be)
    S.build_neighbors_(g.nl)
    O.build_neighbors_(o)
    gi, gc = g.nl.arrays()
    oi, oc, of = o.neighbors()
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)          # same slot-major layout, same order
    np.testing.assert_array_equal(g.nl.flags, of)
    # the cell list after build_neighbors is the one of that build
    for a, b in zip(g.cl.arrays(), o.cell_list()):
        np.testing.assert_array_equal(a, b)


def test_neighbor_overflow_flag(be):
    say("overflow flag")
    sc = scenes.random_box(400, seed=42)
    sc.maxnb = 8
    g, o, _, _ = pair(sc, be)
    S.build_neighbors_(g.nl)
    O.build_neighbors_(o)
    gi, gc = g.nl.arrays()
    oi, oc, of = o.neighbors()
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(gi, oi)
    assert g.nl.flags[0] == 1 and of[0] == 1


@pytest.mark.parametrize("nsteps,interval", [(1, 4), (25, 4), (25, 1), (60, 8)])
def test_forward_bit_exact_dambreak(be, nsteps, interval):
    say("forward dambreak", nsteps, interval)
    sc = scenes.dambreak_40x40(nsteps=nsteps)
    g, o, pg, po = pair(sc, be, interval=interval)
    S.simulate_(g, sc.theta, pg, be, nsteps)
    O.simulate_(o, sc.theta, po, nsteps)
    np.testing.assert_array_equal(g.X, o.X)
    np.testing.assert_array_equal(g.V, o.V)
    np.testing.assert_array_equal(g.rho, o.rho)
    np.testing.assert_array_equal(g.pterm, o.pterm)
    np.testing.assert_array_equal(g.invrho, o.invrho)
    np.testing.assert_array_equal(g.a, o.a)
    np.testing.assert_array_equal(g.alpha, o.alpha)
    assert g.nl.age == o.age


def test_forward_bit_exact_brinkman(be):
    say("forward brinkman 56x37, 200 steps")
    sc = scenes.brinkman_56x37(nsteps=200)
    g, o, pg, po = pair(sc, be)
    S.simulate_(g, sc.theta, pg, be, 100)
    S.simulate_(g, sc.theta, pg, be, 100)        # two calls continue the same schedule
    O.simulate_(o, sc.theta, po, 200)
    np.testing.assert_array_equal(g.X, o.X)
    np.testing.assert_array_equal(g.V, o.V)
    np.testing.assert_array_equal(g.alpha, o.alpha)


def test_step_and_build_neighbors_api(be):
    say("step_ / build_neighbors_ interleaving")
    sc = scenes.reuse_block()
    g, o, pg, po = pair(sc, be, interval=3)
    th = np.zeros((pg.ngy, pg.ngx))
    for k in range(7):
        S.step_(g, th, pg, be)
        O.step_(o, th, po)
        if k == 3:
            a_before = g.a
            S.build_neighbors_(g.nl)
            O.build_neighbors_(o)
            np.testing.assert_array_equal(g.a, a_before)     # st.a survives a rebuild
    np.testing.assert_array_equal(g.X, o.X)
    np.testing.assert_array_equal(g.a, o.a)
    assert g.nl.age == o.age


def test_displacement_flag(be):
    say("displacement flag")
    # fast particles and a long interval break the skin (test/runtests.jl:244-245)
    p = dict(h=0.06, m=0.004, rho0=1.0, c=10.0, mu=0.02, dt=1e-4, Lx=1.0, Ly=0.6, kw=1e4,
             ngx=9, ngy=7, gx=0.0, gy=-9.81)
    N = 64
    X0 = np.zeros((2, N))
    for k in range(N):
        X0[0, k] = 0.2 + 0.02 * (k % 8)
        X0[1, k] = 0.3 + 0.02 * (k // 8)
    V0 = np.full((2, N), 30.0)
    sc = scenes.Scene("fast", p, X0, V0, np.zeros((7, 9)), 20, interval=8)
    g, o, pg, po = pair(sc, be)
    with pytest.warns(UserWarning, match="interval"):
        S.simulate_(g, sc.theta, pg, be, 20)
    with pytest.warns(UserWarning, match="interval"):
        O.simulate_(o, sc.theta, po, 20)
    assert g.nl.flags.tolist() == [0, 0]        # reported and cleared


def test_set_state_and_objective(be):
    say("set X / target_objective")
    sc = scenes.brinkman_56x37()
    g, o, pg, po = pair(sc, be)
    S.simulate_(g, sc.theta, pg, be, 10)
    O.simulate_(o, sc.theta, po, 10)
    J, seed = S.target_objective(g, 0.3, 0.2, 0.1)
    Jo, so = O.target_objective(o.X, 0.3, 0.2, 0.1)
    assert J == pytest.approx(Jo, rel=1e-13)
    np.testing.assert_allclose(seed, so, rtol=1e-14, atol=1e-300)   # device exp vs libm exp
    X = g.X.copy()
    X[0] += 1e-5
    g.X = X
    np.testing.assert_array_equal(g.X, X)


def test_tape_entries_and_capacity(be):
    say("tape")
    sc = scenes.dambreak_40x40(nsteps=9)
    g, o, pg, po = pair(sc, be)
    tg = S.Tape(be, g.N, 9, pg)
    to = O.Tape(o.N, 9, po)
    S.simulate_(g, sc.theta, pg, be, 5, tape=tg)
    S.simulate_(g, sc.theta, pg, be, 4, tape=tg)
    O.simulate_(o, sc.theta, po, 9, to)
    assert len(tg) == 9 and tg.capacity() == 9
    for k in (1, 4, 5, 9):
        xg, vg = tg.get(k)
        xo, vo = to.get(k)
        np.testing.assert_array_equal(xg, xo)
        np.testing.assert_array_equal(vg, vo)
    with pytest.raises(ValueError):
        S.simulate_(g, sc.theta, pg, be, 1, tape=tg)
    S.reset_(tg)
    assert len(tg) == 0


def _grad_pair(sc, be, nsteps, interval, seedX, seedV):
    g, o, pg, po = pair(sc, be, interval=interval)
    tg = S.Tape(be, g.N, nsteps, pg)
    to = O.Tape(o.N, nsteps, po)
    S.simulate_(g, sc.theta, pg, be, nsteps, tape=tg)
    O.simulate_(o, sc.theta, po, nsteps, to)
    wg = S.AdjointWorkspace(be, g.N, pg)
    wo = O.AdjointWorkspace(o.N, po)
    S.backward_(wg, tg, sc.theta, pg, be, seedX=seedX, seedV=seedV, interval=interval)
    O.backward_(wo, to, sc.theta, po, seedX=seedX, seedV=seedV, interval=interval, layout="slot")
    return g, o, wg, wo


@pytest.mark.parametrize("theta_scale", [3.0, 3000.0])
def test_adjoint_gradcheck_scene_bit_exact(be, theta_scale):
    say("adjoint gradcheck 5x5, theta scale", theta_scale)
    sc = scenes.gradcheck_5x5(nsteps=12)
    sc.theta = sc.theta * (theta_scale / 3.0)
    g, o, wg, wo = _grad_pair(sc, be, 12, 1, sc.extra["cX"], sc.extra["cV"])
    np.testing.assert_array_equal(g.X, o.X)
    np.testing.assert_array_equal(wg.gX, wo.gX)
    np.testing.assert_array_equal(wg.gV, wo.gV)
    np.testing.assert_allclose(wg.gtheta, wo.gtheta, rtol=1e-12,
                               atol=1e-13 * np.abs(wo.gtheta).max())


def test_adjoint_brinkman_200_steps(be):
    say("adjoint brinkman 56x37, 200 steps, interval 4")
    sc = scenes.brinkman_56x37(nsteps=200)
    rng = np.random.default_rng(11)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    g, o, wg, wo = _grad_pair(sc, be, 200, 4, cX, cV)
    np.testing.assert_array_equal(g.X, o.X)
    for a, b in ((wg.gX, wo.gX), (wg.gV, wo.gV), (wg.gtheta, wo.gtheta)):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9 * np.abs(b).max())


def test_adjoint_vs_central_differences_on_gpu(be):
    say("adjoint vs central differences (GPU forward)")
    sc = scenes.gradcheck_5x5(nsteps=12)
    pg = S.SPHParams(**sc.params)
    cX, cV = sc.extra["cX"], sc.extra["cV"]

    def J(X0, V0, th, tape=None):
        st = S.State(be, X0, V0, pg, interval=1)
        S.simulate_(st, th, pg, be, 12, tape=tape)
        return float((st.X * cX).sum() + (st.V * cV).sum()), st
    tape = S.Tape(be, sc.N, 12, pg)
    _, st = J(sc.X0, sc.V0, sc.theta, tape)
    ws = S.AdjointWorkspace(be, sc.N, pg)
    S.backward_(ws, tape, sc.theta, pg, be, seedX=cX, seedV=cV, interval=1)
    h = 1e-6

    def fd(f, x, idx):
        xp = x.copy(); xp[idx] += h
        xm = x.copy(); xm[idx] -= h
        return (f(xp) - f(xm)) / (2 * h)
    for idx in [(0, 0), (3, 2), (4, 5)]:
        gfd = fd(lambda t: J(sc.X0, sc.V0, t)[0], sc.theta, idx)
        assert ws.gtheta[idx] == pytest.approx(gfd, rel=1e-5, abs=1e-9)
    for idx in [(0, 0), (1, 8), (0, 19)]:
        gfd = fd(lambda x: J(x, sc.V0, sc.theta)[0], sc.X0, idx)
        assert ws.gX[idx] == pytest.approx(gfd, rel=1e-5, abs=1e-8)
    for idx in [(0, 3), (1, 14)]:
        gfd = fd(lambda v: J(sc.X0, v, sc.theta)[0], sc.V0, idx)
        assert ws.gV[idx] == pytest.approx(gfd, rel=1e-5, abs=1e-8)


def test_repeat_runs_bit_identical(be):
    say("determinism over repeated runs")
    sc = scenes.brinkman_56x37(nsteps=200)
    pg = S.SPHParams(**sc.params)
    outs = []
    for _ in range(3):
        st = S.State(be, sc.X0, sc.V0, pg)
        S.simulate_(st, sc.theta, pg, be, 200)
        J, _ = S.target_objective(st, 0.3, 0.2, 0.1)
        outs.append((J, st.X))
    for J, X in outs[1:]:
        assert J == outs[0][0]
        np.testing.assert_array_equal(X, outs[0][1])


def test_single_particle_and_boundaries(be):
    say("N=1 and particles beyond the walls")
    p = dict(h=0.05, m=0.0025, rho0=1.0, c=10.0, mu=0.02, dt=1e-4, Lx=1.0, Ly=0.6, kw=1e4,
             ngx=5, ngy=4, gx=0.0, gy=-9.81)
    for X0 in (np.array([[0.5], [0.3]]), np.array([[-0.01, 1.02, 0.5], [0.3, 0.61, -0.02]])):
        sc = scenes.Scene("edge", p, X0, np.zeros_like(X0), np.ones((4, 5)), 10)
        g, o, pg, po = pair(sc, be)
        S.simulate_(g, sc.theta, pg, be, 10)
        O.simulate_(o, sc.theta, po, 10)
        np.testing.assert_array_equal(g.X, o.X)
        np.testing.assert_array_equal(g.V, o.V)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_gtheta_bit_identical_over_runs(be, precision):
    """gtheta is reduced by a fixed-order gather (no float atomics), so the
    whole gradient, not only J, repeats bit for bit."""
    say("gtheta determinism", precision)
    sc = scenes.brinkman_56x37(nsteps=60)
    pg = S.SPHParams(**sc.params)
    rng = np.random.default_rng(2)
    cX = rng.standard_normal((2, sc.N))
    cV = rng.standard_normal((2, sc.N))
    outs = []
    for _ in range(2):
        st = S.State(be, sc.X0, sc.V0, pg, precision=precision)
        tape = S.Tape(be, sc.N, 60, pg, precision=precision)
        S.simulate_(st, sc.theta, pg, be, 60, tape=tape)
        ws = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
        S.backward_(ws, tape, sc.theta, pg, be, seedX=cX, seedV=cV)
        outs.append((ws.gtheta, ws.gX, ws.gV))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_gtheta_reaches_particles_several_cells_from_their_binning_cell(be, precision):
    """The gtheta gather sweeps, for each theta cell, the SPH cells around it
    with a margin for the displacement since the last rebuild.  Isolated fast
    particles (no neighbours, so the reverse rebuild schedule does not matter)
    moving 2-3 cells per 20-step epoch need a wider margin than one cell: the
    R pass raises the step's sweep level and every particle's design term is
    still gathered.  gtheta against the oracle's atomic scatter (which never
    misses one): 1e-12 in FP64, the FP32 form's 1e-5."""
    say("gtheta with particles several cells from their binning cell", precision)
    rng = np.random.default_rng(17)
    p = dict(h=0.02, m=4e-4, rho0=1.0, c=10.0, mu=0.02, dt=1e-3, Lx=1.0, Ly=0.6, kw=1e4,
             ngx=9, ngy=7, gx=0.0, gy=-9.81)
    # a drifting lattice: every particle moves at (5, 3) m/s, so the spacing
    # (0.12) stays above the cutoff 2h(1 + skin) = 0.048 and no pair ever forms
    xs, ys = np.meshgrid(0.2 + 0.12 * np.arange(5), 0.15 + 0.12 * np.arange(3))
    X0 = np.vstack([xs.ravel(), ys.ravel()]) + rng.uniform(-0.01, 0.01, (2, xs.size))
    V0 = np.vstack([np.full(xs.size, 5.0), np.full(xs.size, 3.0)])
    theta = 5.0 * rng.uniform(size=(7, 9))
    n, interval = 40, 20
    sc = scenes.Scene("fast_isolated", p, X0, V0, theta, n, interval=interval)
    pg, po = S.SPHParams(**p), O.SPHParams(**p)
    g = S.State(be, X0, V0, pg, interval=interval, precision=precision)
    tg = S.Tape(be, sc.N, n, pg, precision=precision)
    o = O.State(X0, V0, po, interval=interval, layout="slot")
    to = O.Tape(sc.N, n, po)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        S.simulate_(g, theta, pg, be, n, tape=tg)
        O.simulate_(o, theta, po, n, to)
    moved = np.abs(g.X - X0).max() / (2 * p["h"] * 1.2)
    print(f"   displacement over the horizon: {moved:.2f} cells; X vs oracle "
          f"{np.abs(g.X - o.X).max():.2e}")
    assert (o.neighbors()[1] == 0).all()          # no pair in the last list either
    J, seed = S.target_objective(g, 0.5, 0.3, 0.2)
    Jo, so = O.target_objective(o.X, 0.5, 0.3, 0.2)
    wg = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
    wo = O.AdjointWorkspace(sc.N, po)
    z = np.zeros_like(seed)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        S.backward_(wg, tg, theta, pg, be, seedX=seed, seedV=z, interval=interval)
        O.backward_(wo, to, theta, po, seedX=so, seedV=z, interval=interval, layout="slot")
    e = float(np.abs(wg.gtheta - wo.gtheta).max() / np.abs(wo.gtheta).max())
    print(f"   gtheta max-norm rel err {e:.3e}; nodes carrying gradient "
          f"{int((np.abs(wo.gtheta) > 0).sum())}")
    assert moved > 4
    assert e < (1e-12 if precision == "f64" else 1e-5)
