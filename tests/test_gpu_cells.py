"""Cell lists of pile-ups and of diverged states, on the GPU against the oracle.

Bar: bit-exact cellof / counts / starts / order (the reference's counting
sort with _cl_sort_cells!, src/neighbors.jl:44-114).  Cells with more than 128
entries are ranked by a bitmap of their ids rather than by each entry counting
the smaller ids of its cell, so a pile-up costs O(entries) per cell, not
O(entries^2); the order must not change.  A position that is not finite bins
into a clamped cell, as the reference's unsafe_trunc does, and raises a flag.
"""
import sys
import time
import warnings

import numpy as np
import pytest

import sphadjoint_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def be():
    return S.MI355XBackend(0)


def params(h):
    return dict(h=h, m=1e-3, rho0=1.0, c=10.0, mu=0.01, dt=1e-5, Lx=2.0, Ly=1.2, kw=1e4,
                ngx=9, ngy=7, gx=0.0, gy=-9.81)


def cluster(case):
    rng = np.random.default_rng(23)
    h = 0.004
    cs = 2 * h * 1.2
    if case == "one_cell_50k":
        N = 50000
        X = np.vstack([0.7 + cs * rng.uniform(0.01, 0.99, N), 0.5 + cs * rng.uniform(0.01, 0.99, N)])
    else:
        # a uniform background of 70k particles, 30k piled into one cell and
        # 10 cells of 300: big cells, and ids spanning more than one bitmap
        # window (65,536 ids) inside one cell
        X = np.vstack([rng.uniform(0, 2.0, 70000), rng.uniform(0, 1.2, 70000)])
        pile = np.vstack([1.3 + cs * rng.uniform(0.01, 0.99, 30000),
                          0.4 + cs * rng.uniform(0.01, 0.99, 30000)])
        small = [np.vstack([0.1 + 0.15 * k + cs * rng.uniform(0.01, 0.99, 300),
                            0.9 + cs * rng.uniform(0.01, 0.99, 300)]) for k in range(10)]
        X = np.hstack([X, pile] + small)
        X = X[:, rng.permutation(X.shape[1])]   # pile-up ids spread over the whole range
    return X, params(h)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("case", ["one_cell_50k", "mixed_100k"])
def test_pile_up_cell_list_bit_exact(be, case, precision):
    say("pile-up cell list", case, precision)
    X, p = cluster(case)
    g = S.State(be, X, np.zeros_like(X), S.SPHParams(**p), precision=precision)
    o = O.State(X, np.zeros_like(X), O.SPHParams(**p), layout="slot")
    be.synchronize()
    t0 = time.perf_counter()
    S.build_(g.cl)
    be.synchronize()
    dt = time.perf_counter() - t0
    O.build_(o)
    got, want = g.cl.arrays(), o.cell_list()
    print(f"   N={X.shape[1]}, largest cell {int(want[1].max())} entries, GPU build {dt * 1e3:.1f} ms")
    for a, b, name in zip(got, want, ("cellof", "counts", "starts", "order")):
        np.testing.assert_array_equal(a, b, err_msg=name)
    assert dt < 1.0


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_non_finite_position_flag(be, precision):
    """A NaN position bins into the lowest cell (the reference's clamp of
    unsafe_trunc's typemin) and the rebuild raises the non-finite flag, which
    the flag check reports and clears."""
    say("non-finite position flag", precision)
    rng = np.random.default_rng(4)
    p = params(0.02)
    X = np.vstack([rng.uniform(0.1, 1.9, 300), rng.uniform(0.1, 1.1, 300)])
    X[:, 17] = np.nan
    X[0, 40] = np.inf
    g = S.State(be, X, np.zeros_like(X), S.SPHParams(**p), precision=precision)
    o = O.State(X, np.zeros_like(X), O.SPHParams(**p), layout="slot")
    S.build_(g.cl)
    O.build_(o)
    for a, b in zip(g.cl.arrays(), o.cell_list()):
        np.testing.assert_array_equal(a, b)
    with pytest.warns(UserWarning, match="non-finite"):
        S.check_neighbor_flags(g.nl)
    f3 = np.zeros(3, np.int32)
    S._lib.call("sph_flags_get_all", g._h, f3.ctypes.data_as(S.api.C.c_void_p))
    assert f3.tolist() == [0, 0, 0]
    # a finite state raises nothing
    h = S.State(be, X[:, :10] * 0 + 0.5 + 0.01 * np.arange(10), np.zeros((2, 10)),
                S.SPHParams(**p), precision=precision)
    S.build_(h.cl)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        S.check_neighbor_flags(h.nl)
