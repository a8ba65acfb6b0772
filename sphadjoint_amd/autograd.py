"""torch.autograd binding of simulate!/backward! (the rrule of the time loop).

``simulate(theta, X0, V0, sim)`` runs ``sim.nsteps`` steps of the device
simulation and returns the final state (X_T, V_T) as float64 tensors of shape
(2, N); its backward is the discrete adjoint of libsphadj (src/driver.jl:141-207
and src/adjoint.jl), which returns dJ/dtheta, dJ/dX0 and dJ/dV0 from
dJ/dX_T and dJ/dV_T.  Any torch objective of (X_T, V_T) composes with it, for
instance :func:`gaussian_target` (src/driver.jl:209-234).

With ``snapshots=None`` a :class:`Simulation` keeps the whole horizon on a
Tape, as the reference does; with an integer it runs the binomial
checkpointing schedule of :class:`sphadjoint_amd.Checkpointer` under that
snapshot budget.  Both give the same gradients bit for bit.

Tensors cross the boundary through host memory and must live on the CPU:
libsphadj drives the GPU through the system HIP runtime, while PyTorch wheels
carry their own copy of it, and the two cannot share the device in one process.
"""
from __future__ import annotations

import numpy as np
import torch

from . import api as A


class Simulation:
    """Device buffers for repeated forward/adjoint passes over one scene size."""

    def __init__(self, backend: A.MI355XBackend, p: A.SPHParams, N, nsteps, *, snapshots=None,
                 seg_epochs=1, skin=0.2, interval=4, maxnb=64, precision="f64"):
        self.backend, self.p, self.N, self.nsteps = backend, p, int(N), int(nsteps)
        self.skin, self.interval, self.maxnb, self.precision = skin, interval, maxnb, precision
        self.snapshots = snapshots
        self.ws = A.AdjointWorkspace(backend, self.N, p, precision=precision)
        if snapshots is None:
            self.ck = None
            self.tape = A.Tape(backend, self.N, self.nsteps, p, precision=precision)
        else:
            self.ck = A.Checkpointer(backend, self.N, self.nsteps, p, snapshots=snapshots,
                                     seg_epochs=seg_epochs, skin=skin, interval=interval,
                                     maxnb=maxnb, precision=precision)
            self.tape = None
        self._serial = 0       # forward passes run
        self._ready = -1       # serial of the forward the buffers hold
        self._theta = None

    def forward(self, theta, X0, V0):
        """(X_T, V_T) as (2, N) float64 arrays; keeps what the adjoint needs."""
        self._serial += 1
        self._theta = A._field(theta, self.p, self.backend)
        if self.ck is not None:
            st = self.ck.forward(X0, V0, self._theta)
        else:
            A.reset_(self.tape)
            st = A.State(self.backend, X0, V0, self.p, skin=self.skin, interval=self.interval,
                         maxnb=self.maxnb, precision=self.precision)
            A.simulate_(st, self._theta, self.p, self.backend, self.nsteps, tape=self.tape)
        self._ready = self._serial
        return st.X, st.V

    def adjoint(self, gXT, gVT, serial=None):
        """(dJ/dtheta, dJ/dX0, dJ/dV0) from (dJ/dX_T, dJ/dV_T) of the last forward."""
        if self._ready < 0 or (serial is not None and serial != self._ready):
            raise RuntimeError("the adjoint needs the forward pass it belongs to; this "
                               "Simulation has run another forward (or a backward) since")
        if self.ck is not None:
            self.ck.backward(self.ws, seedX=gXT, seedV=gVT, theta=self._theta)
        else:
            A.backward_(self.ws, self.tape, self._theta, self.p, self.backend, seedX=gXT,
                        seedV=gVT, interval=self.interval, skin=self.skin, maxnb=self.maxnb)
        self._ready = -1       # one backward per forward: the buffers are spent
        return self.ws.gtheta, self.ws.gX, self.ws.gV


class _SimulateFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, X0, V0, sim):
        XT, VT = sim.forward(theta.detach().cpu().numpy(), X0.detach().cpu().numpy(),
                             V0.detach().cpu().numpy())
        ctx.sim = sim
        ctx.serial = sim._serial
        return torch.from_numpy(XT), torch.from_numpy(VT)

    @staticmethod
    def backward(ctx, gXT, gVT):
        sim = ctx.sim
        gx = np.zeros((2, sim.N)) if gXT is None else gXT.detach().cpu().double().numpy()
        gv = np.zeros((2, sim.N)) if gVT is None else gVT.detach().cpu().double().numpy()
        gth, gX0, gV0 = sim.adjoint(gx, gv, serial=ctx.serial)
        return torch.from_numpy(gth), torch.from_numpy(gX0), torch.from_numpy(gV0), None


def simulate(theta: torch.Tensor, X0: torch.Tensor, V0: torch.Tensor, sim: Simulation):
    """(X_T, V_T) = simulate!(State(X0, V0), theta, nsteps), differentiable in
    theta, X0 and V0 through the discrete adjoint.  float64 tensors: theta
    (ngy, ngx), X0/V0 (2, N)."""
    for t, name in ((theta, "theta"), (X0, "X0"), (V0, "V0")):
        if t.dtype != torch.float64:
            raise TypeError(f"{name} must be float64")
        if t.device.type != "cpu":
            raise TypeError(f"{name} must be a CPU tensor (the simulation's own device memory "
                            "holds the state)")
    return _SimulateFn.apply(theta, X0, V0, sim)


def gaussian_target(X: torch.Tensor, xc, yc, sigma):
    """J = -sum_i exp(-|x_i - x_c|^2 / (2 sigma^2)) (src/driver.jl:219-234), in torch."""
    s2 = float(sigma) ** 2
    dx = X[0] - xc
    dy = X[1] - yc
    return -torch.exp(-(dx * dx + dy * dy) / (2 * s2)).sum()
