"""The allocation-free objective (sph_target_objective_noalloc) on the GPU.

Bar: its J and seed equal sph_target_objective_dev's bit for bit, on two calls
in a row that share the GPU's scratch array, the second on a smaller state
(fewer partials than the first call left behind), and it refuses a state
beyond the scratch array's capacity.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import sphadjoint_amd as S
from sphadjoint_amd import _lib as L
from sphadjoint_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def noalloc(st, xc, yc, sigma, with_seed=True):
    J = S.DeviceArray(st.backend, N=1)          # 16 bytes of device memory; J is the first double
    seed = S.DeviceArray(st.backend, N=st.N) if with_seed else None
    L.call("sph_target_objective_noalloc", st._h, float(xc), float(yc), float(sigma), J.ptr,
           seed.ptr if seed is not None else None)
    st.backend.synchronize()
    return float(J.numpy()[0, 0]), (seed.numpy() if seed is not None else None)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_noalloc_objective_equals_the_allocating_one(precision):
    say("allocation-free objective", precision)
    be = S.MI355XBackend(0)
    big = scenes.brinkman_56x37(nsteps=30)
    small = scenes.dambreak_40x40(nsteps=30)
    for sc, (xc, yc, sig) in ((big, (0.3, 0.2, 0.1)), (small, (0.2, 0.25, 0.05))):
        pg = S.SPHParams(**sc.params)
        st = S.State(be, sc.X0, sc.V0, pg, precision=precision)
        S.simulate_(st, sc.theta, pg, be, 30)
        J, seed = S.target_objective(st, xc, yc, sig, device=True)
        Jn, seedn = noalloc(st, xc, yc, sig)
        print(f"   N={sc.N}: J {J:.17g} / {Jn:.17g}")
        assert Jn == J
        np.testing.assert_array_equal(seedn, seed.numpy())
        Jn2, none = noalloc(st, xc, yc, sig, with_seed=False)
        assert Jn2 == J and none is None


def test_noalloc_objective_rejects_states_beyond_its_scratch():
    say("allocation-free objective, size bound")
    be = S.MI355XBackend(0)
    N = 16384 * 256 + 1
    X = np.vstack([np.linspace(0.01, 1.99, N), np.full(N, 0.5)])
    p = S.SPHParams(h=0.001, m=1e-6, rho0=1.0, c=10.0, mu=0.0, dt=1e-6, Lx=2.0, Ly=1.2,
                    kw=1e4, ngx=9, ngy=7)
    st = S.State(be, X, np.zeros_like(X), p)
    J = S.DeviceArray(be, N=1)
    with pytest.raises(L.SPHError, match="scratch array holds"):
        L.call("sph_target_objective_noalloc", st._h, 0.5, 0.5, 0.1, J.ptr, None)
