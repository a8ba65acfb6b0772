"""The topology optimiser's design chain and line search on the device.

The reference keeps these in its script scripts/03_optimize.jl (lines
142-244 for the chain, 287-341 for the line search); here they run on the GPU
through the sph_design_* entry points of include/sphadj.h:

    DesignChain.theta(rd)        rdf = apply_filter(rd); theta = alpha_of(rdf)
    DesignChain.grad(rdf, ws)    grd = filter_adjoint(gtheta .* dalpha_of(rdf))
    DesignChain.project(x)       project_volume(x, volfrac)
    DesignChain.candidate(...)   project_volume(rd - step*g/max|g|, volfrac)
    Optimizer.objective_and_grad / objective_only / optimize
                                 the script's functions and its backtracking
                                 search (accept only improving candidates,
                                 halve the step up to 6 times, grow it by 1.3
                                 up to 0.3 after an acceptance)

Sensitivity is the look-ahead overlay of scripts/04_interactive.jl:169-215
(forward + adjoint over a short horizon from the current state, mapped to the
improvement of painting each cell solid) without the GUI.

Optimizer.optimize(batched=True) evaluates the up to six candidates of one
backtracking round together, as the scenes of one batched f32c state, and
accepts the first improving one in the order the sequential search would try
them, so it takes the same decisions as the sequential search run on the same
arithmetic.  Design variables stay FP64 on the device; they are read back only
to report them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .api import (AdjointWorkspace, BatchState, DeviceArray, Field, MI355XBackend, SPHParams,
                  State, Tape, backward_, batch_objective, reset_, simulate_, target_objective)


class DesignChain:
    """build_filter (scripts/03_optimize.jl:156-164) and the Borrvall-Petersson
    map alpha(rd) = alpha_max*qp*(1-rd)/(qp+rd) for an (ngy, ngx) grid."""

    def __init__(self, backend: MI355XBackend, ngy, ngx, *, alpha_max=2.0e3, qp=0.1,
                 volfrac=0.75, radius=2):
        self.backend = backend
        self.shape = (int(ngy), int(ngx))
        self.alpha_max, self.qp, self.volfrac, self.radius = (float(alpha_max), float(qp),
                                                               float(volfrac), int(radius))
        h = C.c_void_p()
        L.call("sph_design_create", backend._h, self.shape[0], self.shape[1], self.radius,
               self.alpha_max, self.qp, C.byref(h))
        self._h = h

    def field(self, x=None):
        return Field(self.backend, np.zeros(self.shape) if x is None else x)

    def _as_field(self, x):
        return x if isinstance(x, Field) else self.field(x)

    def theta(self, rd, rdf: Field = None, theta: Field = None, offset=0):
        """(rdf, theta) with rdf = apply_filter(rd), theta = alpha_of(rdf); theta may
        be a batched field, written at node ``offset``."""
        rd = self._as_field(rd)
        rdf = rdf or self.field()
        theta = theta or self.field()
        L.call("sph_design_theta", self._h, rd._h, rdf._h, theta._h, int(offset))
        return rdf, theta

    def grad(self, rdf: Field, ws: AdjointWorkspace, grd: Field = None, offset=0):
        """filter_adjoint(gtheta .* dalpha_of(rdf)) from the workspace's gtheta."""
        grd = grd or self.field()
        L.call("sph_design_grad", self._h, rdf._h, ws._h, int(offset), grd._h)
        return grd

    def project(self, x, out: Field = None):
        x = self._as_field(x)
        out = out or self.field()
        L.call("sph_design_project", self._h, x._h, self.volfrac, out._h)
        return out

    def candidate(self, rd: Field, g: Field, step, out: Field = None):
        """(project_volume(rd - step*g/max|g|), max|g|)."""
        out = out or self.field()
        gmax = C.c_double()
        L.call("sph_design_candidate", self._h, rd._h, g._h, float(step), self.volfrac, out._h,
               C.byref(gmax))
        return out, gmax.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.load().sph_design_destroy(h)
            self._h = None


class Optimizer:
    """objective_and_grad / objective_only / the backtracking projected-gradient
    loop of scripts/03_optimize.jl:198-341 for one scene (X0, V0, nsteps) and a
    Gaussian target (xc, yc, sigma); the neighbour-list interval defaults to the
    script's NL_INTERVAL = 2."""

    def __init__(self, backend: MI355XBackend, p: SPHParams, X0, V0, nsteps, target, sigma, *,
                 chain: DesignChain = None, interval=2, precision="f64", **chain_kw):
        self.backend, self.p = backend, p
        self.X0, self.V0 = np.asarray(X0, dtype=np.float64), np.asarray(V0, dtype=np.float64)
        self.N = self.X0.shape[1]
        self.nsteps, self.target, self.sigma = int(nsteps), tuple(target), float(sigma)
        self.interval, self.precision = int(interval), precision
        self.chain = chain or DesignChain(backend, p.ngy, p.ngx, **chain_kw)
        # allocated once and reused, as the script does with its tape
        self.tape = Tape(backend, self.N, self.nsteps, p, precision=precision)
        self.ws = AdjointWorkspace(backend, self.N, p, precision=precision)
        self.rdf, self.th = self.chain.field(), self.chain.field()
        self.evaluations = 0

    def _forward(self, rd, tape=None):
        self.chain.theta(rd, self.rdf, self.th)
        st = State(self.backend, self.X0, self.V0, self.p, interval=self.interval,
                   precision=self.precision)
        simulate_(st, self.th, self.p, self.backend, self.nsteps, tape=tape)
        self.evaluations += 1
        return st

    def objective_and_grad(self, rd):
        """(J, grd Field) for design rd (array or Field)."""
        reset_(self.tape)
        st = self._forward(rd, self.tape)
        J, seed = target_objective(st, *self.target, self.sigma, device=True)
        backward_(self.ws, self.tape, self.th, self.p, self.backend, seedX=seed,
                  seedV=DeviceArray(self.backend, N=self.N), interval=self.interval)
        return J, self.chain.grad(self.rdf, self.ws)

    def objective_only(self, rd):
        st = self._forward(rd)
        return target_objective(st, *self.target, self.sigma)[0]

    def objectives_batched(self, cands):
        """J of each candidate design (Fields), all scenes of one batched f32c state."""
        k = len(cands)
        scenes = [(self.p, self.X0, self.V0)] * k
        bst = BatchState(self.backend, scenes, interval=self.interval)
        n = self.p.ngy * self.p.ngx
        th = bst.field([np.zeros((self.p.ngy, self.p.ngx))] * k)
        rdf = self.chain.field()
        for q, c in enumerate(cands):
            self.chain.theta(c, rdf, th, offset=q * n)
        simulate_(bst, th, bst.p, self.backend, self.nsteps)
        self.evaluations += k
        return batch_objective(bst, *self.target, self.sigma)[0]

    def optimize(self, rd0=None, *, maxit=40, step0=0.05, batched=False, tries=6):
        """Projected gradient with backtracking (scripts/03_optimize.jl:287-341).
        Returns (rd as numpy, J history, iterations run).  batched=True scores
        each iteration's candidates in one batched state, which exists in the
        f32c form only: it takes the sequential search's decisions when the
        optimizer's own precision is f32c, and is refused otherwise."""
        if batched and self.precision != "f32c":
            raise ValueError("batched candidate evaluation runs in the f32c form; an "
                             f"optimizer of precision {self.precision!r} would compare its "
                             "incumbent J with f32c candidates (use precision='f32c')")
        ch = self.chain
        rd = ch.field(np.ones(ch.shape) if rd0 is None else rd0)
        step = float(step0)
        J, g = self.objective_and_grad(rd)
        hist = [J]
        it = 0
        for it in range(1, maxit + 1):
            gn = float(np.max(np.abs(np.asarray(g))))
            if gn < 1e-30:
                break
            accepted = False
            if batched:
                steps = [step * 0.5 ** t for t in range(tries)]
                cands = [ch.candidate(rd, g, s)[0] for s in steps]
                Js = self.objectives_batched(cands)
                for t, (s, Jc) in enumerate(zip(steps, Js)):
                    if Jc < J:
                        rd, J, step, accepted = cands[t], float(Jc), min(s * 1.3, 0.30), True
                        break
                if not accepted:
                    step = steps[-1] * 0.5
            else:
                for _ in range(tries):
                    cand, _gn = ch.candidate(rd, g, step)
                    Jc = self.objective_only(cand)
                    if Jc < J:
                        rd, J, accepted = cand, Jc, True
                        step = min(step * 1.3, 0.30)
                        break
                    step *= 0.5
            if not accepted:
                break
            hist.append(J)
            _, g = self.objective_and_grad(rd)
        return np.asarray(rd), hist, it



class Sensitivity:
    """The look-ahead sensitivity overlay of the interactive script
    (scripts/04_interactive.jl:169-215) without its GUI: from a state (X, V),
    run ``nsens`` steps on a tape, take J = target_objective at (xc, yc) with
    width sigma, run the adjoint, and map dJ/dtheta to the improvement of
    painting a cell solid.  In that script theta = alpha_max * (1 - canvas^T),
    so painting (canvas 1 -> 0) raises theta by alpha_max and the improvement
    is -alpha_max * gtheta^T, shown on an (ngx, ngy) canvas, normalised by the
    95th percentile of its nonzero magnitudes and clamped to [-1, 1].  A
    look-ahead whose |J| is below ``min_signal`` (0.5 in the script) carries no
    signal and is reported as invalid."""

    def __init__(self, backend: MI355XBackend, p: SPHParams, N, nsens, *, interval=2,
                 precision="f64", alpha_max=1500.0, min_signal=0.5):
        self.backend, self.p, self.N, self.nsens = backend, p, int(N), int(nsens)
        self.interval, self.precision = int(interval), precision
        self.alpha_max, self.min_signal = float(alpha_max), float(min_signal)
        # allocated once and reused at every request, as the script does
        self.tape = Tape(backend, self.N, self.nsens, p, precision=precision)
        self.ws = AdjointWorkspace(backend, self.N, p, precision=precision)

    def compute(self, X, V, theta, target, sigma):
        """(J, overlay (ngx, ngy) in [-1, 1], raw improvement (ngx, ngy), valid)."""
        st = State(self.backend, X, V, self.p, interval=self.interval,
                   precision=self.precision)      # teleport!: a fresh list, no displacement check
        reset_(self.tape)
        simulate_(st, theta, self.p, self.backend, self.nsens, tape=self.tape)
        J, seed = target_objective(st, target[0], target[1], sigma, device=True)
        backward_(self.ws, self.tape, theta, self.p, self.backend, seedX=seed,
                  seedV=DeviceArray(self.backend, N=self.N), interval=self.interval)
        improve = -self.alpha_max * self.ws.gtheta.T
        nz = np.abs(improve)[np.abs(improve) > 0.0]
        scale = float(np.quantile(nz, 0.95)) if nz.size else 1.0
        overlay = np.clip(improve / max(scale, np.finfo(np.float64).eps), -1.0, 1.0)
        valid = abs(J) >= self.min_signal
        if not valid:
            overlay = np.zeros_like(overlay)
        return J, overlay, improve, valid
