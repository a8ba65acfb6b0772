"""Regression pins against the committed fixtures of tests/golden/ (written by
tests/golden/make_golden.py from the CPU oracle).

CPU: the oracle reproduces every fixture bit for bit.
GPU (FP64 form): cell list, neighbour list and forward state bit-exact; J
bit-exact; gradients within the tolerances the parity tests use (the GPU's
gtheta is summed in a fixed order of its own, and with interval > 1 its
reverse sweep reuses the forward's neighbour epochs).  GPU (f32c form): within
the FP32-compensated form's stated errors.
"""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PKEYS = ["h", "m", "rho0", "c", "mu", "dt", "Lx", "Ly", "kw", "ngx", "ngy", "gx", "gy"]
FIXTURES = ["gradcheck_5x5", "random_box_64"]


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    pr = dict(zip(PKEYS, d["params"].tolist()))
    pr["ngx"], pr["ngy"] = int(pr["ngx"]), int(pr["ngy"])
    d["p"] = pr
    return d


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def linear_or_gauss(name, X, V, d, target):
    if name == "gradcheck_5x5":
        return float((X * d["seedX"]).sum() + (V * d["seedV"]).sum()), d["seedX"], d["seedV"]
    return target(X)


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_fixture(name):
    from oracle import oracle as O
    d = load(name)
    p = O.SPHParams(**d["p"])
    kw = dict(skin=float(d["skin"]), interval=int(d["interval"]), maxnb=int(d["maxnb"]),
              layout="slot")
    st = O.State(d["X0"], d["V0"], p, **kw)
    O.build_neighbors_(st)
    for a, k in zip(st.cell_list(), ("cellof0", "counts0", "starts0", "order0")):
        np.testing.assert_array_equal(a, d[k], err_msg=k)
    idx, cnt, _ = st.neighbors()
    np.testing.assert_array_equal(cnt, d["nl_cnt0"])
    np.testing.assert_array_equal(idx, d["nl_idx0"])
    st = O.State(d["X0"], d["V0"], p, **kw)
    n = int(d["nsteps"])
    tape = O.Tape(st.N, n, p)
    O.simulate_(st, d["theta"], p, n, tape)
    np.testing.assert_array_equal(st.X, d["X"])
    np.testing.assert_array_equal(st.V, d["V"])
    np.testing.assert_array_equal(st.rho, d["rho"])
    ws = O.AdjointWorkspace(st.N, p)
    O.backward_(ws, tape, d["theta"], p, seedX=d["seedX"], seedV=d["seedV"],
                interval=int(d["interval"]), layout="slot")
    for k in ("gX", "gV", "gtheta"):
        np.testing.assert_array_equal(getattr(ws, k), d[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_matches_fixture(name, precision):
    import sphadjoint_amd as S
    print("[case] golden", name, precision, flush=True)
    d = load(name)
    be = S.MI355XBackend(0)
    p = S.SPHParams(**d["p"])
    kw = dict(skin=float(d["skin"]), interval=int(d["interval"]), maxnb=int(d["maxnb"]))
    st = S.State(be, d["X0"], d["V0"], p, precision=precision, **kw)
    S.build_neighbors_(st.nl)
    got = st.cl.arrays()
    for a, k in zip(got, ("cellof0", "counts0", "starts0", "order0")):
        np.testing.assert_array_equal(a, d[k], err_msg=k)
    if precision == "f64":
        idx, cnt = st.nl.arrays()
        np.testing.assert_array_equal(cnt, d["nl_cnt0"])
        np.testing.assert_array_equal(idx, d["nl_idx0"])
    st = S.State(be, d["X0"], d["V0"], p, precision=precision, **kw)
    n = int(d["nsteps"])
    tape = S.Tape(be, st.N, n, p, precision=precision)
    S.simulate_(st, d["theta"], p, be, n, tape=tape)
    ws = S.AdjointWorkspace(be, st.N, p, precision=precision)
    S.backward_(ws, tape, d["theta"], p, be, seedX=d["seedX"], seedV=d["seedV"],
                interval=int(d["interval"]))
    if precision == "f64":
        np.testing.assert_array_equal(st.X, d["X"])
        np.testing.assert_array_equal(st.V, d["V"])
        np.testing.assert_allclose(st.rho, d["rho"], rtol=1e-14)
        tol = 1e-12 if name == "gradcheck_5x5" else 1e-9
        for k in ("gX", "gV", "gtheta"):
            assert relerr(getattr(ws, k), d[k]) < tol, k
    else:
        assert relerr(st.X, d["X"]) < 1e-7
        assert relerr(st.V, d["V"]) < 1e-4
        for k in ("gX", "gV", "gtheta"):
            assert relerr(getattr(ws, k), d[k]) < 1e-4, k
