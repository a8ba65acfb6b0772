#!/usr/bin/env python3
"""Writes the small regression fixtures of tests/golden/ from the CPU oracle.

The reference stores no golden vectors (its tests are invariants, brute force
and finite differences), so these fixtures are the oracle's own outputs, kept
as regression pins once the oracle has passed its checks (tests/test_oracle.py).
Each fixture holds the scene's inputs and the oracle's outputs:

  gradcheck_5x5   N=25, 12 steps, interval 1: final X, V, rho, the cell list,
                  the neighbour list, J of a random linear functional and its
                  gradients gX, gV, gtheta (test/runtests.jl:163-228's scene)
  random_box_64   N=64 in the 1.0x0.6 box, 25 steps, interval 4: the same
                  outputs with the Gaussian target as the objective

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O            # noqa: E402
from sphadjoint_amd import scenes         # noqa: E402


def run(sc, nsteps, interval, objective):
    p = O.SPHParams(**sc.params)
    st = O.State(sc.X0, sc.V0, p, skin=sc.skin, interval=interval, maxnb=sc.maxnb, layout="slot")
    O.build_neighbors_(st)
    cellof, counts, starts, order = st.cell_list()
    idx, cnt, _ = st.neighbors()
    st = O.State(sc.X0, sc.V0, p, skin=sc.skin, interval=interval, maxnb=sc.maxnb, layout="slot")
    tape = O.Tape(st.N, nsteps, p)
    O.simulate_(st, sc.theta, p, nsteps, tape)
    X, V = st.X, st.V
    if objective == "linear":
        cX, cV = sc.extra["cX"], sc.extra["cV"]
        J = float((X * cX).sum() + (V * cV).sum())
        seedX, seedV = cX, cV
    else:
        J, seedX = O.target_objective(X, 0.5, 0.3, 0.15)
        seedV = np.zeros_like(seedX)
    ws = O.AdjointWorkspace(st.N, p)
    O.backward_(ws, tape, sc.theta, p, seedX=seedX, seedV=seedV, interval=interval, layout="slot")
    return dict(params=np.array([sc.params[k] for k in PKEYS]), X0=sc.X0, V0=sc.V0,
                theta=sc.theta, nsteps=nsteps, interval=interval, skin=sc.skin, maxnb=sc.maxnb,
                cellof0=cellof, counts0=counts, starts0=starts, order0=order, nl_idx0=idx,
                nl_cnt0=cnt, X=X, V=V, rho=st.rho, J=J, seedX=seedX, seedV=seedV, gX=ws.gX,
                gV=ws.gV, gtheta=ws.gtheta)


PKEYS = ["h", "m", "rho0", "c", "mu", "dt", "Lx", "Ly", "kw", "ngx", "ngy", "gx", "gy"]


def main():
    sc = scenes.gradcheck_5x5(nsteps=12)
    np.savez(os.path.join(HERE, "gradcheck_5x5.npz"), **run(sc, 12, 1, "linear"))
    sc = scenes.random_box(64, seed=5)
    sc.theta = np.random.default_rng(9).uniform(size=sc.theta.shape)
    sc.V0 = 0.1 * np.random.default_rng(10).standard_normal(sc.X0.shape)
    np.savez(os.path.join(HERE, "random_box_64.npz"), **run(sc, 25, 4, "gaussian"))


if __name__ == "__main__":
    main()
