"""Pins for the CPU restatement of the design-field chain (no GPU).

The reference checks this chain with one script: the design-variable gradient
against central differences (scripts/10_design_gradcheck.jl: top-3 components,
relative error below 1e-5, at an interior design near 0.9 with qp = 10 and
alpha_max = 100).  That check is restated here on the oracle's SPH, together
with what the mathematics fixes for each piece: the filter's weights sum to
one, it reproduces constant and (away from the edges) linear fields, its
adjoint is its exact transpose, alpha(0) = alpha_max and alpha(1) = 0 with the
stated derivative, and the volume projection is feasible, minimal and
idempotent.  A planted mistake in any of them (a dropped normalisation, an
adjoint that reuses the forward's weights, a wrong sign in dalpha, a shift in
the wrong direction) fails one of these.
"""
import numpy as np
import pytest

from oracle import design as D
from oracle import oracle as O
from sphadjoint_amd import scenes


@pytest.mark.parametrize("shape,r", [((17, 33), 2), ((7, 9), 3), ((5, 5), 1)])
def test_filter_weights_sum_to_one(shape, r):
    W = D.build_filter(*shape, r)
    for lst in W.values():
        assert sum(w for _, _, w in lst) == pytest.approx(1.0, abs=2e-15)
        assert all(w > 0 for _, _, w in lst)
    # interior node: (2r+1)^2 box minus the corners at distance >= r+1
    j, i = shape[0] // 2, shape[1] // 2
    if r <= j < shape[0] - r and r <= i < shape[1] - r:
        n_in = sum(1 for a in range(-r, r + 1) for b in range(-r, r + 1)
                   if a * a + b * b < (r + 1) ** 2)
        assert len(W[(j, i)]) == n_in


def test_filter_reproduces_constant_and_linear_fields():
    ngy, ngx, r = 17, 33, 2
    W = D.build_filter(ngy, ngx, r)
    one = D.apply_filter(W, np.ones((ngy, ngx)))
    assert np.all(one <= 1.0) and np.allclose(one, 1.0, atol=2e-15)
    jj, ii = np.meshgrid(np.arange(ngy), np.arange(ngx), indexing="ij")
    lin = 0.1 + 0.01 * ii + 0.02 * jj          # inside [0, 1]
    y = D.apply_filter(W, lin)
    inner = (slice(r, ngy - r), slice(r, ngx - r))
    assert np.allclose(y[inner], lin[inner], rtol=0, atol=2e-15)
    # at the edges the one-sided weights pull towards the interior
    assert not np.allclose(y[0, :], lin[0, :], atol=1e-6)


def test_filter_adjoint_is_the_transpose():
    rng = np.random.default_rng(3)
    ngy, ngx = 17, 33
    for r in (1, 2, 3):
        W = D.build_filter(ngy, ngx, r)
        x = rng.uniform(0.0, 1.0, (ngy, ngx))
        y = rng.standard_normal((ngy, ngx))
        lhs = float((D.apply_filter(W, x) * y).sum())
        rhs = float((x * D.filter_adjoint(W, y)).sum())
        assert lhs == pytest.approx(rhs, rel=1e-13)
        # the filter is not symmetric at the edges: its transpose differs from it
        assert not np.allclose(D.filter_adjoint(W, y), D.apply_filter(W, y), atol=1e-3)


def test_alpha_map_and_its_derivative():
    for amax, qp in [(2.0e3, 0.1), (100.0, 10.0), (5.0, 1.0)]:
        assert D.alpha_of(1.0, amax, qp) == 0.0
        assert D.alpha_of(0.0, amax, qp) == pytest.approx(amax, rel=1e-15)
        for rd in (0.05, 0.3, 0.75, 0.97):
            e = 1e-6
            fd = (D.alpha_of(rd + e, amax, qp) - D.alpha_of(rd - e, amax, qp)) / (2 * e)
            assert D.dalpha_of(rd, amax, qp) == pytest.approx(fd, rel=1e-7)
            assert D.dalpha_of(rd, amax, qp) < 0


def test_volume_projection_is_feasible_minimal_and_idempotent():
    rng = np.random.default_rng(5)
    for volfrac in (0.75, 0.5, 0.9):
        x = rng.uniform(-0.2, 1.1, (17, 33))
        y, t = D.project_volume(x, volfrac)
        assert t is not None and 0.0 < t < 1.0
        assert y.mean() >= volfrac - 1e-15
        # a shift smaller by 1e-12 would miss the bound: the shift is minimal
        assert np.clip(np.clip(x, 0, 1) + t - 1e-12, 0, 1).mean() < volfrac
        # projecting again moves nothing but rounding: the bisection returns the
        # midpoint of its last bracket, whose mean may sit an ulp below volfrac
        z, t2 = D.project_volume(y, volfrac)
        assert (t2 is None or t2 < 1e-15) and np.allclose(z, y, rtol=0, atol=1e-15)
    # a feasible design is only clamped
    x = rng.uniform(0.8, 1.3, (17, 33))
    y, t = D.project_volume(x, 0.75)
    assert t is None and np.array_equal(y, np.clip(x, 0, 1))


def test_candidate_moves_against_the_gradient():
    rng = np.random.default_rng(9)
    rd = rng.uniform(0.85, 0.95, (17, 33))
    g = rng.standard_normal((17, 33))
    c, gn = D.candidate(rd, g, 0.05, 0.75)
    assert gn == np.abs(g).max()
    assert np.allclose(c, np.clip(rd - 0.05 * g / gn, 0, 1))


def _design_scene(qp=10.0, amax=100.0, nsteps=12):
    """The reference's gradcheck scene (test/runtests.jl:163-180) with its 7x7
    theta produced by the design chain from rd ~ 0.9 +- 0.05 (an interior point,
    as scripts/10_design_gradcheck.jl:26-36 requires), and a Gaussian target on
    the block so that J carries signal."""
    sc = scenes.gradcheck_5x5(seed=0, nsteps=nsteps)
    rng = np.random.default_rng(7)
    rd = 0.90 + 0.10 * (rng.uniform(size=(7, 7)) - 0.5)
    return sc, rd, dict(qp=qp, amax=amax, r=2, target=(0.32, 0.32), sigma=0.1)


def _oracle_J(sc, W, rd, cfg, grad=False):
    p = O.SPHParams(**sc.params)
    rdf = D.apply_filter(W, rd)
    th = D.alpha_of(rdf, cfg["amax"], cfg["qp"])
    st = O.State(sc.X0, sc.V0, p, interval=1)
    tape = O.Tape(st.N, sc.nsteps, p) if grad else None
    O.simulate_(st, th, p, sc.nsteps, tape)
    J, seed = O.target_objective(st.X, *cfg["target"], cfg["sigma"])
    if not grad:
        return J
    ws = O.AdjointWorkspace(st.N, p)
    O.backward_(ws, tape, th, p, seedX=seed, seedV=np.zeros_like(seed), interval=1)
    grdf = ws.gtheta * D.dalpha_of(rdf, cfg["amax"], cfg["qp"])
    return J, D.filter_adjoint(W, grdf)


def test_design_gradient_vs_central_differences():
    """scripts/10_design_gradcheck.jl restated on the oracle: the three largest
    components of dJ/drd against central differences (h = 1e-5), below 1e-5."""
    sc, rd, cfg = _design_scene()
    W = D.build_filter(7, 7, cfg["r"])
    J, g = _oracle_J(sc, W, rd, cfg, grad=True)
    assert abs(J) > 1e-3                      # the target sees the block
    hfd = 1e-5
    for li in np.argsort(-np.abs(g).ravel())[:3]:
        idx = np.unravel_index(li, g.shape)
        xp = rd.copy(); xp[idx] += hfd
        xm = rd.copy(); xm[idx] -= hfd
        fd = (_oracle_J(sc, W, xp, cfg) - _oracle_J(sc, W, xm, cfg)) / (2 * hfd)
        assert g[idx] == pytest.approx(fd, rel=1e-5)
