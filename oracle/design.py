"""CPU restatement of the reference's design-field chain (TEST INFRASTRUCTURE).

Loaded only by tests/ as the checker of the device chain (sph_design_*).  The
reference keeps this chain in its optimisation script, not in src/; every
function below follows scripts/03_optimize.jl operation for operation in
Float64 (Julia's TW), with plain Python floats so that each rounding is the
reference's:

    build_filter(ngy, ngx, r)          scripts/03_optimize.jl:156-164
    apply_filter(W, x)                 scripts/03_optimize.jl:166-183
    filter_adjoint(W, g)               scripts/03_optimize.jl:186-194
    alpha_of(rd), dalpha_of(rd)        scripts/03_optimize.jl:150-151
    project_volume(rd, volfrac)        scripts/03_optimize.jl:205-219
    candidate(rd, g, step, volfrac)    scripts/03_optimize.jl:311

Arrays are numpy (ngy, ngx); indices are 0-based here.  The one place where
bits cannot follow the reference is the sum inside project_volume (Julia's
`sum` may reassociate): it is taken here in index order, and the checks pin
the projection by what the mathematics fixes (feasibility, minimality of the
shift, idempotence) instead of bit identity.
"""
from __future__ import annotations

import math

import numpy as np


def build_filter(ngy, ngx, r):
    """Weight lists of the linear density filter: node (j, i) -> [(jj, ii, w/s)],
    jj outer, ii inner, w = (r+1) - dist kept when positive, s = sum of the kept w."""
    W = {}
    for j in range(ngy):
        for i in range(ngx):
            lst = []
            s = 0.0
            for jj in range(max(0, j - r), min(ngy - 1, j + r) + 1):
                for ii in range(max(0, i - r), min(ngx - 1, i + r) + 1):
                    w = float(r + 1) - math.sqrt(float((ii - i) ** 2 + (jj - j) ** 2))
                    if w > 0:
                        lst.append((jj, ii, w))
                        s += w
            W[(j, i)] = [(jj, ii, w / s) for (jj, ii, w) in lst]
    return W


def apply_filter(W, x):
    x = np.asarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    for j in range(x.shape[0]):
        for i in range(x.shape[1]):
            acc = 0.0
            for jj, ii, w in W[(j, i)]:
                acc += w * float(x[jj, ii])
            y[j, i] = min(max(acc, 0.0), 1.0)
    return y


def filter_adjoint(W, g):
    g = np.asarray(g, dtype=np.float64)
    out = np.zeros_like(g)
    for j in range(g.shape[0]):
        for i in range(g.shape[1]):
            for jj, ii, w in W[(j, i)]:
                out[jj, ii] += w * float(g[j, i])
    return out


def alpha_of(rd, alpha_max, qp):
    rd = np.asarray(rd, dtype=np.float64)
    return alpha_max * qp * (1 - rd) / (qp + rd)


def dalpha_of(rd, alpha_max, qp):
    rd = np.asarray(rd, dtype=np.float64)
    q = qp + rd
    return -alpha_max * qp * (qp + 1) / (q * q)


def _mean(x):
    s = 0.0
    for v in np.asarray(x, dtype=np.float64).ravel(order="F"):
        s += float(v)
    return s / x.size


def project_volume(rd, volfrac):
    """Clamp to [0,1]; if the mean is below volfrac, shift every node by the
    bisected amount t in [0,1] (60 steps) and clamp again."""
    x = np.clip(np.asarray(rd, dtype=np.float64), 0.0, 1.0)
    if _mean(x) >= volfrac:
        return x, None
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = (lo + hi) / 2
        if _mean(np.clip(x + mid, 0.0, 1.0)) < volfrac:
            lo = mid
        else:
            hi = mid
    t = (lo + hi) / 2
    return np.clip(x + t, 0.0, 1.0), t


def candidate(rd, g, step, volfrac):
    gn = float(np.max(np.abs(g)))
    return project_volume(np.asarray(rd) - step * np.asarray(g) / gn, volfrac)[0], gn
