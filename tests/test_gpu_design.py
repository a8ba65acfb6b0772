"""The design-field chain and line search on the GPU (sph_design_*).

Bars: the filter, the Borrvall-Petersson map and the chain-rule/transpose
kernels evaluate the oracle's (the reference script's) Float64 expressions in
the same order, so they match bit for bit; the volume projection sums in a
fixed order of its own, so it is held to 1e-14 and to feasibility; the design
gradient matches the oracle's to 1e-10 and central differences to 1e-5 (the
bar of scripts/10_design_gradcheck.jl); the batched line search takes the same
decisions as the sequential one.
"""
import sys

import numpy as np
import pytest

import sphadjoint_amd as S
from oracle import design as D
from oracle import oracle as O
from sphadjoint_amd import scenes
from sphadjoint_amd.design import DesignChain, Optimizer

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def be():
    return S.MI355XBackend(0)


@pytest.mark.parametrize("shape,r,amax,qp", [((17, 33), 2, 2.0e3, 0.1), ((32, 64), 2, 100.0, 10.0),
                                             ((7, 9), 3, 5.0, 1.0)])
def test_filter_and_alpha_bit_exact(be, shape, r, amax, qp):
    say("filter/alpha", shape, r)
    rng = np.random.default_rng(11)
    rd = rng.uniform(-0.1, 1.1, shape)
    ch = DesignChain(be, *shape, alpha_max=amax, qp=qp, radius=r)
    rdf, th = ch.theta(rd)
    W = D.build_filter(*shape, r)
    rdf_o = D.apply_filter(W, rd)
    assert np.array_equal(np.asarray(rdf), rdf_o)
    assert np.array_equal(np.asarray(th), D.alpha_of(rdf_o, amax, qp))


def _design_scene(nsteps=12):
    sc = scenes.gradcheck_5x5(seed=0, nsteps=nsteps)
    rng = np.random.default_rng(7)
    rd = 0.90 + 0.10 * (rng.uniform(size=(7, 7)) - 0.5)
    return sc, rd, dict(qp=10.0, alpha_max=100.0, radius=2), ((0.32, 0.32), 0.1)


def test_design_gradient_matches_oracle(be):
    say("design gradient vs oracle, gradcheck scene")
    sc, rd, cfg, (tgt, sig) = _design_scene()
    p = S.SPHParams(**sc.params)
    opt = Optimizer(be, p, sc.X0, sc.V0, sc.nsteps, tgt, sig, interval=1, **cfg)
    J, g = opt.objective_and_grad(rd)
    # oracle: the same chain around the oracle's SPH
    W = D.build_filter(7, 7, cfg["radius"])
    po = O.SPHParams(**sc.params)
    rdf = D.apply_filter(W, rd)
    th = D.alpha_of(rdf, cfg["alpha_max"], cfg["qp"])
    st = O.State(sc.X0, sc.V0, po, interval=1, layout="slot")
    tape = O.Tape(st.N, sc.nsteps, po)
    O.simulate_(st, th, po, sc.nsteps, tape)
    Jo, seed = O.target_objective(st.X, *tgt, sig)
    ws = O.AdjointWorkspace(st.N, po)
    O.backward_(ws, tape, th, po, seedX=seed, seedV=np.zeros_like(seed), interval=1,
                layout="slot")
    assert J == pytest.approx(Jo, rel=1e-13)
    # the chain itself on the GPU's own gtheta: bit for bit
    gth = opt.ws.gtheta
    g_chain = D.filter_adjoint(W, gth * D.dalpha_of(np.asarray(opt.rdf), cfg["alpha_max"],
                                                     cfg["qp"]))
    assert np.array_equal(np.asarray(g), g_chain)
    # against the oracle's whole chain
    go = D.filter_adjoint(W, ws.gtheta * D.dalpha_of(rdf, cfg["alpha_max"], cfg["qp"]))
    assert np.allclose(np.asarray(g), go, rtol=1e-10, atol=1e-14 * np.abs(go).max())


def test_design_gradient_vs_central_differences(be):
    """scripts/10_design_gradcheck.jl on the GPU: top-3 components, h = 1e-5."""
    say("design gradient vs central differences")
    sc, rd, cfg, (tgt, sig) = _design_scene()
    p = S.SPHParams(**sc.params)
    opt = Optimizer(be, p, sc.X0, sc.V0, sc.nsteps, tgt, sig, interval=1, **cfg)
    J, g = opt.objective_and_grad(rd)
    g = np.asarray(g)
    assert abs(J) > 1e-3
    hfd = 1e-5
    for li in np.argsort(-np.abs(g).ravel())[:3]:
        idx = np.unravel_index(li, g.shape)
        xp = rd.copy(); xp[idx] += hfd
        xm = rd.copy(); xm[idx] -= hfd
        fd = (opt.objective_only(xp) - opt.objective_only(xm)) / (2 * hfd)
        assert g[idx] == pytest.approx(fd, rel=1e-5)


def test_projection_and_candidate(be):
    say("projection and candidate")
    rng = np.random.default_rng(13)
    shape = (32, 64)
    ch = DesignChain(be, *shape, volfrac=0.75)
    x = rng.uniform(-0.2, 1.0, shape)
    y = np.asarray(ch.project(x))
    yo, t = D.project_volume(x, 0.75)
    assert t is not None
    # the device sums in its own fixed order (per thread, then a tree): the
    # bisection's last decisions, taken where the mean is within rounding of
    # volfrac, may differ from the sequential sum's, which moves the shift by a
    # few ulps of its value (1.3e-15 for this input); y is clamp(x + shift)
    assert np.allclose(y, yo, rtol=0, atol=1e-14)
    assert y.mean() >= 0.75 - 1e-15
    feas = rng.uniform(0.8, 1.2, shape)
    assert np.array_equal(np.asarray(ch.project(feas)), np.clip(feas, 0, 1))
    rd = ch.field(rng.uniform(0.85, 0.95, shape))
    g = ch.field(rng.standard_normal(shape))
    c, gmax = ch.candidate(rd, g, 0.05)
    co, gno = D.candidate(np.asarray(rd), np.asarray(g), 0.05, 0.75)
    assert gmax == gno
    assert np.allclose(np.asarray(c), co, rtol=0, atol=1e-14)


def test_line_search_batched_equals_sequential(be):
    """Batched candidate evaluation takes the sequential search's decisions."""
    say("line search, f32c, sequential vs batched")
    sc, rd, cfg, (tgt, sig) = _design_scene(nsteps=20)
    p = S.SPHParams(**sc.params)
    runs = []
    for batched in (False, True):
        opt = Optimizer(be, p, sc.X0, sc.V0, sc.nsteps, tgt, sig, interval=1,
                        precision="f32c", **cfg)
        rd_out, hist, it = opt.optimize(rd, maxit=3, step0=0.05, batched=batched)
        runs.append((rd_out, hist))
    (r0, h0), (r1, h1) = runs
    assert len(h0) >= 2 and all(b < a for a, b in zip(h0, h0[1:]))
    assert h0 == h1
    assert np.array_equal(r0, r1)


def test_line_search_f64_improves(be):
    say("line search, f64")
    sc, rd, cfg, (tgt, sig) = _design_scene(nsteps=20)
    p = S.SPHParams(**sc.params)
    opt = Optimizer(be, p, sc.X0, sc.V0, sc.nsteps, tgt, sig, interval=1, **cfg)
    # batched candidates exist in the f32c form only: an FP64 optimizer refuses them
    with pytest.raises(ValueError, match="f32c"):
        opt.optimize(rd, maxit=1, batched=True)
    rd_out, hist, it = opt.optimize(rd, maxit=4)
    assert len(hist) >= 2 and all(b < a for a, b in zip(hist, hist[1:]))
    assert rd_out.mean() >= 0.75 - 1e-15 and rd_out.min() >= 0 and rd_out.max() <= 1


def test_sensitivity_overlay(be):
    """scripts/04_interactive.jl:196-215 without the GUI: J and the improvement
    map against the oracle, and the predicted improvement of painting the most
    sensitive node against central differences in theta."""
    say("sensitivity overlay")
    sc = scenes.gradcheck_5x5(seed=0, nsteps=12)
    p = S.SPHParams(**sc.params)
    amax = 1500.0
    sens = S.Sensitivity(be, p, sc.N, 12, interval=1, alpha_max=amax)
    tgt, sig = (0.32, 0.32), 0.1
    J, overlay, improve, valid = sens.compute(sc.X0, sc.V0, sc.theta, tgt, sig)
    assert valid and overlay.shape == (7, 7) and np.abs(overlay).max() == 1.0
    po = O.SPHParams(**sc.params)
    st = O.State(sc.X0, sc.V0, po, interval=1, layout="slot")
    tape = O.Tape(st.N, 12, po)
    O.simulate_(st, sc.theta, po, 12, tape)
    Jo, seed = O.target_objective(st.X, *tgt, sig)
    ws = O.AdjointWorkspace(st.N, po)
    O.backward_(ws, tape, sc.theta, po, seedX=seed, seedV=np.zeros_like(seed), interval=1,
                layout="slot")
    assert J == pytest.approx(Jo, rel=1e-13)
    assert np.allclose(improve, -amax * ws.gtheta.T, rtol=1e-10,
                       atol=1e-12 * np.abs(improve).max())
    # painting node (j, i) by delta raises theta there by amax*delta
    i, j = np.unravel_index(np.argmax(np.abs(improve)), improve.shape)
    d = 1e-6

    def J_of(th):
        s2 = S.State(be, sc.X0, sc.V0, p, interval=1)
        S.simulate_(s2, th, p, be, 12)
        return S.target_objective(s2, *tgt, sig)[0]
    tp = sc.theta.copy(); tp[j, i] += amax * d
    tm = sc.theta.copy(); tm[j, i] -= amax * d
    assert -(J_of(tp) - J_of(tm)) / (2 * d) == pytest.approx(improve[i, j], rel=1e-5)
