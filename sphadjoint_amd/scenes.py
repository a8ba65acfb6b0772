"""Synthetic scene builders for the dam-break / design-field workloads.

Each builder returns a ``Scene`` with plain numpy inputs: the SPHParams keywords,
X0 and V0 of shape (2, N) (the reference's (2,N) state layout), and theta of
shape (ngy, ngx).  Random numbers come from numpy's PCG64 with recorded seeds;
the reference's Julia MersenneTwister streams cannot be reproduced here.

Scene geometry follows the reference scripts: a lattice block starting one
spacing away from the walls, filled row by row (x inner, y outer) as in
scripts/02_dambreak.jl:46-56, and the test scenes of test/runtests.jl.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

G = 9.81


@dataclass
class Scene:
    name: str
    params: dict
    X0: np.ndarray
    V0: np.ndarray
    theta: np.ndarray
    nsteps: int
    skin: float = 0.2
    interval: int = 4
    maxnb: int = 64
    extra: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.X0.shape[1]


def lattice(nx, ny, dx, x0=None, y0=None):
    """nx*ny lattice points at spacing dx starting at (x0, y0), x fastest."""
    x0 = dx if x0 is None else x0
    y0 = dx if y0 is None else y0
    xs = x0 + dx * np.arange(nx)
    ys = y0 + dx * np.arange(ny)
    X = np.empty((2, nx * ny))
    X[0] = np.tile(xs, ny)
    X[1] = np.repeat(ys, nx)
    return X


def wcsph_params(dx, Lx, Ly, ngx, ngy, *, rho0=1000.0, hfac=1.2, H=0.4, cfl=0.25,
                 nu=0.05, kw=5.0e4):
    """Parameter set of the dam-break workloads: h = hfac*dx, m = rho0*dx^2,
    c0 = 10*sqrt(2 g H), dt = cfl*h/c0, mu = rho0*nu (Morris term m*2mu/(rho_i rho_j)
    scales like nu = mu/rho), wall stiffness kw (an acceleration)."""
    h = hfac * dx
    c = 10.0 * math.sqrt(2.0 * G * H)
    return dict(h=h, m=rho0 * dx * dx, rho0=rho0, c=c, mu=rho0 * nu, dt=cfl * h / c,
                Lx=Lx, Ly=Ly, kw=kw, ngx=ngx, ngy=ngy, gx=0.0, gy=-G)


def dambreak_40x40(seed=0, nsteps=1000):
    """Fluid block 40x40 = 1600 particles, dx = 0.01, tank 1.0 x 0.6, theta = 0
    on a 33x21 grid, seeded U(-1e-4, 1e-4) position jitter, V0 = 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 0.01
    p = wcsph_params(dx, 1.0, 0.6, 33, 21)
    X0 = lattice(40, 40, dx) + rng.uniform(-1e-4, 1e-4, size=(2, 1600))
    V0 = np.zeros_like(X0)
    theta = np.zeros((p["ngy"], p["ngx"]))
    return Scene("dambreak_40x40", p, X0, V0, theta, nsteps)


def brinkman_56x37(seed=0, nsteps=200, u0=0.5):
    """N = 2072 (56x37 lattice, dx = 0.01) in a 1.0 x 0.6 tank, moving at
    V0 = (u0, 0) into a Brinkman design field theta ~ U(0,1) on a 64x32 grid."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 0.01
    p = wcsph_params(dx, 1.0, 0.6, 64, 32)
    X0 = lattice(56, 37, dx) + rng.uniform(-1e-4, 1e-4, size=(2, 56 * 37))
    V0 = np.zeros_like(X0)
    V0[0] = u0
    theta = rng.uniform(0.0, 1.0, size=(p["ngy"], p["ngx"]))
    return Scene("brinkman_56x37", p, X0, V0, theta, nsteps)


def box_filter5(a):
    """5x5 box mean with clamp-to-edge (in-bounds mean)."""
    ny, nx = a.shape
    out = np.empty_like(a)
    for j in range(ny):
        j0, j1 = max(0, j - 2), min(ny, j + 3)
        for i in range(nx):
            i0, i1 = max(0, i - 2), min(nx, i + 3)
            out[j, i] = a[j0:j1, i0:i1].mean()
    return out


def block_1m(seed=0, nsteps=500, n_side=1000, nu=0.005):
    """n_side x n_side fluid block (1M particles at n_side=1000), dx = 1e-3,
    tank 2.0 x 1.2, theta on a 512x256 grid seeded U(0,1) and 5x5 box-filtered.

    The kinematic viscosity is set so that the explicit viscous diffusion
    number nu*dt/h^2 (0.037) matches the reference's dam-break script
    (scripts/02_dambreak.jl:31-44: 0.031); the nu = 0.05 of the dx = 0.01
    scenes would give 0.37 at this resolution, past the explicit limit, and the
    block blows up within 20 steps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = 1.0e-3
    p = wcsph_params(dx, 2.0, 1.2, 512, 256, nu=nu)
    N = n_side * n_side
    X0 = lattice(n_side, n_side, dx) + rng.uniform(-1e-2 * dx, 1e-2 * dx, size=(2, N))
    V0 = np.zeros_like(X0)
    th = rng.uniform(0.0, 1.0, size=(p["ngy"], p["ngx"]))
    # vectorised 5x5 in-bounds mean (same result as box_filter5)
    pad = np.pad(th, 2, mode="constant")
    one = np.pad(np.ones_like(th), 2, mode="constant")
    s = np.zeros_like(th)
    cnt = np.zeros_like(th)
    for dj in range(5):
        for di in range(5):
            s += pad[dj:dj + th.shape[0], di:di + th.shape[1]]
            cnt += one[dj:dj + th.shape[0], di:di + th.shape[1]]
    theta = s / cnt
    return Scene(f"block_{n_side}x{n_side}", p, X0, V0, theta, nsteps)


def gradcheck_5x5(seed=0, nsteps=12):
    """The reference's adjoint-vs-central-differences scene (test/runtests.jl:163-180):
    5x5 jittered lattice, h=0.10, m=0.010, rho0=1, c=8, mu=0.02, dt=1e-3,
    1.0x1.0 tank, kw=2e4, theta = 3*U(0,1) on 7x7, random linear functional."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = dict(h=0.10, m=0.010, rho0=1.0, c=8.0, mu=0.02, dt=1.0e-3, Lx=1.0, Ly=1.0,
             kw=2.0e4, ngx=7, ngy=7, gx=0.0, gy=-G)
    nx = ny = 5
    X0 = np.zeros((2, nx * ny))
    k = 0
    for jy in range(ny):
        for ix in range(nx):
            X0[0, k] = 0.20 + 0.06 * ix + 0.004 * rng.standard_normal()
            X0[1, k] = 0.20 + 0.06 * jy + 0.004 * rng.standard_normal()
            k += 1
    V0 = 0.05 * rng.standard_normal((2, nx * ny))
    theta = 3.0 * rng.uniform(size=(7, 7))
    cX = rng.standard_normal((2, nx * ny))
    cV = rng.standard_normal((2, nx * ny))
    return Scene("gradcheck_5x5", p, X0, V0, theta, nsteps, interval=1,
                 extra=dict(cX=cX, cV=cV))


def random_box(N=400, seed=42):
    """N particles uniform in the 1.0x0.6 box (test/runtests.jl:44-49), h=0.06."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = dict(h=0.06, m=0.004, rho0=1.0, c=10.0, mu=0.02, dt=1e-4, Lx=1.0, Ly=0.6,
             kw=1e4, ngx=9, ngy=7, gx=0.0, gy=-G)
    X0 = np.vstack([rng.uniform(size=(1, N)) * p["Lx"], rng.uniform(size=(1, N)) * p["Ly"]])
    return Scene("random_box", p, X0, np.zeros_like(X0), np.zeros((7, 9)), 0)


def reuse_block(seed=7):
    """Jittered 12x10 block with dp=0.02, h=1.3dp (test/runtests.jl:111-131)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dp = 0.02
    p = dict(h=1.3 * dp, m=dp * dp, rho0=1.0, c=10.0, mu=0.02, dt=1e-4, Lx=1.0, Ly=0.6,
             kw=1e4, ngx=9, ngy=7, gx=0.0, gy=-G)
    xs = dp * np.arange(1, 13)
    ys = dp * np.arange(1, 11)
    N = len(xs) * len(ys)
    X0 = np.zeros((2, N))
    k = 0
    for y in ys:
        for x in xs:
            X0[0, k] = x + 0.1 * dp * rng.standard_normal()
            X0[1, k] = y + 0.1 * dp * rng.standard_normal()
            k += 1
    V0 = 0.05 * rng.standard_normal((2, N))
    return Scene("reuse_block", p, X0, V0, np.zeros((7, 9)), 40)


def batch_line_search(seed=0, nscenes=64, nmin=1000, nmax=64000, nsteps=300):
    """64 independent scenes, N log-uniform in [nmin, nmax], each a rectangular
    lattice block (aspect 3:2) with dx scaled so the tank is fixed at 1.0 x 0.6.
    Scene k has theta_k = theta0 + alpha_k * d on a 64x32 grid, alpha_k
    log-spaced in [1e-3, 1], d ~ N(0,1), theta0 ~ U(0,1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Ns = np.round(np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=nscenes))).astype(int)
    theta0 = rng.uniform(0.0, 1.0, size=(32, 64))
    d = rng.standard_normal((32, 64))
    alphas = np.logspace(-3, 0, nscenes)
    scenes = []
    for k, n in enumerate(Ns):
        ny = max(2, int(round(math.sqrt(n / 1.5))))
        nx = max(2, int(round(n / ny)))
        # the block spans half the tank width: nx*dx = 0.5
        dx = 0.5 / nx
        # viscosity: 0.05 (the dam-break scenes), capped so that the explicit
        # viscous diffusion number nu*dt/h^2 stays at 0.037 or below
        h, c = 1.2 * dx, 10.0 * math.sqrt(2.0 * G * 0.4)
        nu = min(0.05, 0.037 * h * h / (0.25 * h / c))
        p = wcsph_params(dx, 1.0, 0.6, 64, 32, nu=nu)
        X0 = lattice(nx, ny, dx) + rng.uniform(-1e-2 * dx, 1e-2 * dx, size=(2, nx * ny))
        V0 = np.zeros_like(X0)
        scenes.append(Scene(f"batch_{k}", p, X0, V0, theta0 + alphas[k] * d, nsteps))
    return scenes
