"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE).

Loaded only by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
bench.py, as the checker.  The names mirror the reference's Julia API
(src/SPHAdjoint.jl:39-40) with ``!`` spelled as a trailing underscore:

    State(X0, V0, p; skin, interval, maxnb, layout)   src/driver.jl:24-36
    Tape(N, nsteps, p), reset_(tape)                  src/driver.jl:62-71
    simulate_(st, theta, p, nsteps, tape=None)        src/driver.jl:92-115
    step_(st, theta, p)                               src/forward.jl:161-173
    build_(st), build_neighbors_(st)                  src/neighbors.jl:97-114,252-263
    AdjointWorkspace(N, p), backward_(...)            src/driver.jl:128-207
    target_objective(X, xc, yc, sigma)                src/driver.jl:219-234

Arrays use numpy shape (2, N) for X/V (Julia's (2,N) column-major memory is
interleaved x,y, which is what the C side stores) and (ngy, ngx) for theta.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import warnings

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

X, V, A, RHO, PTERM, INVRHO, ALPHA, XREF = range(8)
GX, GV, GXS, GVS, ABAR, GRHO, GALPHA, GTHETA = range(8)


class orc_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("h", "m", "rho0", "c", "mu", "gx", "gy", "dt", "Lx", "Ly", "kw")] + \
               [("ngx", C.c_int64), ("ngy", C.c_int64)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    vp = C.c_void_p
    i64 = C.c_int64
    for sfx, real in (("_f64", C.c_double), ("_f32", C.c_float)):
        rp = C.POINTER(real)

        def f(name, res, *args):
            fn = getattr(L, name + sfx)
            fn.restype = res
            fn.argtypes = list(args)
        f("orc_state_create", vp, C.POINTER(orc_params), i64, vp, vp, C.c_double, i64, i64, C.c_int)
        f("orc_state_destroy", None, vp)
        f("orc_state_n", i64, vp)
        f("orc_state_get", None, vp, C.c_int, vp)
        f("orc_state_set", None, vp, C.c_int, vp)
        f("orc_state_info", None, vp, vp)
        f("orc_cell_list_get", None, vp, vp, vp, vp, vp)
        f("orc_neighbors_get", None, vp, vp, vp, vp)
        f("orc_flags_clear", None, vp)
        for n in ("orc_build", "orc_build_neighbors", "orc_maybe_rebuild", "orc_density",
                  "orc_eos", "orc_integrate"):
            f(n, None, vp)
        f("orc_accel", None, vp, vp)
        f("orc_step", None, vp, vp)
        f("orc_tape_create", vp, i64, i64)
        f("orc_tape_destroy", None, vp)
        f("orc_tape_reset", None, vp)
        f("orc_tape_length", i64, vp)
        f("orc_tape_capacity", i64, vp)
        f("orc_tape_get", None, vp, i64, vp, vp)
        f("orc_simulate", C.c_int, vp, vp, i64, vp)
        f("orc_ws_create", vp, i64, C.POINTER(orc_params))
        f("orc_ws_destroy", None, vp)
        f("orc_ws_get", None, vp, C.c_int, vp)
        f("orc_backward", C.c_int, vp, vp, vp, C.POINTER(orc_params), vp, vp, C.c_double, i64,
          i64, C.c_int, vp)
        f("orc_target_objective", real, vp, i64, C.c_double, C.c_double, C.c_double, vp)
        f("orc_w_kern", real, real, real)
        f("orc_f_kern", real, real, real, real)
        f("orc_g_kern", real, real, real, real)
        f("orc_wnorm", real, real)
        f("orc_interp_alpha", None, vp, C.POINTER(orc_params), real, real, vp)
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = None
    L.orc_get_threads.argtypes = []
    L.orc_get_threads.restype = C.c_int


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def _sfx(T):
    T = np.dtype(T)
    if T == np.float64:
        return "_f64"
    if T == np.float32:
        return "_f32"
    raise TypeError(f"oracle supports float64/float32, got {T}")


def _fn(name, T):
    return getattr(lib(), name + _sfx(T))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class SPHParams:
    """SPHParams{T} (src/params.jl:14-28); T is a numpy float dtype."""

    def __init__(self, *, h, m, rho0, c, mu, dt, Lx, Ly, kw, ngx, ngy, gx=0.0, gy=-9.81,
                 T=np.float64):
        self.T = np.dtype(T)
        self.h, self.m, self.rho0, self.c, self.mu = h, m, rho0, c, mu
        self.gx, self.gy, self.dt, self.Lx, self.Ly, self.kw = gx, gy, dt, Lx, Ly, kw
        self.ngx, self.ngy = int(ngx), int(ngy)

    def c_struct(self):
        return orc_params(self.h, self.m, self.rho0, self.c, self.mu, self.gx, self.gy,
                          self.dt, self.Lx, self.Ly, self.kw, self.ngx, self.ngy)


def _as_2n(A, T, N=None):
    A = np.asarray(A, dtype=T)
    if A.ndim != 2 or A.shape[0] != 2:
        raise ValueError("expected an array of shape (2, N)")
    if N is not None and A.shape[1] != N:
        raise ValueError("size mismatch")
    return np.ascontiguousarray(A.T)          # interleaved x,y per particle


def _theta(theta, p):
    th = np.asarray(theta, dtype=p.T)
    if th.shape != (p.ngy, p.ngx):
        raise ValueError(f"theta must be (ngy, ngx) = {(p.ngy, p.ngx)}")
    return np.asfortranarray(th)              # (ngy, ngx) column-major


class State:
    """State (src/driver.jl:5-36) with its CellList and NeighborList."""

    def __init__(self, X0, V0, p: SPHParams, *, skin=0.2, interval=4, maxnb=64, layout="auto"):
        if layout == "auto":
            layout = "particle"                # CPU backend picks particle-major
        if layout not in ("slot", "particle"):
            raise ValueError("layout must be :slot / :particle / :auto")
        self.p = p
        self.T = p.T
        x0 = _as_2n(X0, self.T)
        self.N = x0.shape[0]
        v0 = _as_2n(V0, self.T, self.N)
        cp = p.c_struct()
        self._h = _fn("orc_state_create", self.T)(C.byref(cp), self.N, _ptr(x0), _ptr(v0),
                                                 float(skin), int(interval), int(maxnb),
                                                 0 if layout == "slot" else 1)
        info = np.zeros(8, np.int64)
        _fn("orc_state_info", self.T)(self._h, _ptr(info))
        self.nx, self.ny, self.ncell, self.maxnb, self.sm, self.si = (int(v) for v in info[:6])
        self.interval = int(info[6])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _fn("orc_state_destroy", self.T)(h)
            self._h = None

    def _get(self, which, two):
        out = np.empty((self.N, 2) if two else self.N, dtype=self.T)
        _fn("orc_state_get", self.T)(self._h, which, _ptr(out))
        return np.ascontiguousarray(out.T) if two else out

    def _set(self, which, arr, two):
        a = _as_2n(arr, self.T, self.N) if two else np.ascontiguousarray(arr, dtype=self.T)
        _fn("orc_state_set", self.T)(self._h, which, _ptr(a))

    X = property(lambda s: s._get(X, True), lambda s, v: s._set(X, v, True))
    V = property(lambda s: s._get(V, True), lambda s, v: s._set(V, v, True))
    a = property(lambda s: s._get(A, True))
    rho = property(lambda s: s._get(RHO, False))
    pterm = property(lambda s: s._get(PTERM, False))
    invrho = property(lambda s: s._get(INVRHO, False))
    alpha = property(lambda s: s._get(ALPHA, False))

    @property
    def age(self):
        info = np.zeros(8, np.int64)
        _fn("orc_state_info", self.T)(self._h, _ptr(info))
        return int(info[7])

    def cell_list(self):
        """(cellof, counts, starts, order) in the reference's 1-based convention."""
        cellof = np.zeros(self.N, np.int32)
        counts = np.zeros(self.ncell, np.int32)
        starts = np.zeros(self.ncell, np.int32)
        order = np.zeros(self.N, np.int32)
        _fn("orc_cell_list_get", self.T)(self._h, _ptr(cellof), _ptr(counts), _ptr(starts),
                                          _ptr(order))
        return cellof, counts, starts, order

    def neighbors(self):
        """(indices (N*maxnb,), counts (N,), flags (2,)) with the state's layout strides."""
        idx = np.zeros(self.N * self.maxnb, np.int32)
        cnt = np.zeros(self.N, np.int32)
        flags = np.zeros(2, np.int32)
        _fn("orc_neighbors_get", self.T)(self._h, _ptr(idx), _ptr(cnt), _ptr(flags))
        return idx, cnt, flags

    def neighbor_rows(self):
        """List of per-particle neighbour id lists (1-based), in list order."""
        idx, cnt, _ = self.neighbors()
        rows = []
        for i in range(self.N):
            rows.append([int(idx[m * self.sm + i * self.si]) for m in range(int(cnt[i]))])
        return rows

    flags = property(lambda s: s._flags())

    def _flags(self):
        f = np.zeros(2, np.int32)
        _fn("orc_neighbors_get", self.T)(self._h, None, None, _ptr(f))
        return f


def check_neighbor_flags(st: State):
    """check_neighbor_flags (src/neighbors.jl:282-290): warn and reset."""
    f = st._flags()
    if f[0] != 0:
        warnings.warn(f"neighbour list row overflow (maxnb={st.maxnb}); neighbours were dropped")
    if f[1] != 0:
        warnings.warn("neighbour list rebuild interval too long (displacement exceeded skin/2)")
    _fn("orc_flags_clear", st.T)(st._h)
    return f


def build_(st: State):
    _fn("orc_build", st.T)(st._h)
    return st


def build_neighbors_(st: State):
    _fn("orc_build_neighbors", st.T)(st._h)
    return st


def maybe_rebuild_(st: State):
    _fn("orc_maybe_rebuild", st.T)(st._h)
    return st


def density_(st: State):
    _fn("orc_density", st.T)(st._h)


def eos_(st: State):
    _fn("orc_eos", st.T)(st._h)


def accel_(st: State, theta):
    th = _theta(theta, st.p)
    _fn("orc_accel", st.T)(st._h, _ptr(th))


def integrate_(st: State):
    _fn("orc_integrate", st.T)(st._h)


def step_(st: State, theta, p=None):
    th = _theta(theta, st.p)
    _fn("orc_step", st.T)(st._h, _ptr(th))
    return st


class Tape:
    """Tape (src/driver.jl:45-71): (2, N, capacity) X and V before each step."""

    def __init__(self, N, nsteps, p: SPHParams):
        self.T = p.T
        self.N = int(N)
        self._h = _fn("orc_tape_create", self.T)(self.N, int(nsteps))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _fn("orc_tape_destroy", self.T)(h)
            self._h = None

    def __len__(self):
        return int(_fn("orc_tape_length", self.T)(self._h))

    def capacity(self):
        return int(_fn("orc_tape_capacity", self.T)(self._h))

    def get(self, k):
        x = np.empty((self.N, 2), self.T)
        v = np.empty((self.N, 2), self.T)
        _fn("orc_tape_get", self.T)(self._h, int(k), _ptr(x), _ptr(v))
        return np.ascontiguousarray(x.T), np.ascontiguousarray(v.T)


def reset_(tape: Tape):
    _fn("orc_tape_reset", tape.T)(tape._h)
    return tape


def simulate_(st: State, theta, p, nsteps, tape: Tape | None = None):
    th = _theta(theta, st.p)
    rc = _fn("orc_simulate", st.T)(st._h, _ptr(th), int(nsteps), tape._h if tape is not None else None)
    if rc == 1:
        raise ValueError(f"tape capacity {tape.capacity()} too small for {len(tape) + nsteps} steps")
    check_neighbor_flags(st)
    return st


class AdjointWorkspace:
    """AdjointWorkspace (src/driver.jl:117-133)."""

    def __init__(self, N, p: SPHParams):
        self.T = p.T
        self.N = int(N)
        self.p = p
        cp = p.c_struct()
        self._h = _fn("orc_ws_create", self.T)(self.N, C.byref(cp))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _fn("orc_ws_destroy", self.T)(h)
            self._h = None

    def _get(self, which):
        if which == GTHETA:
            out = np.empty(self.p.ngx * self.p.ngy, self.T)
            _fn("orc_ws_get", self.T)(self._h, which, _ptr(out))
            return out.reshape(self.p.ngx, self.p.ngy).T.copy()
        two = which in (GX, GV, GXS, GVS, ABAR)
        out = np.empty((self.N, 2) if two else self.N, self.T)
        _fn("orc_ws_get", self.T)(self._h, which, _ptr(out))
        return np.ascontiguousarray(out.T) if two else out

    gX = property(lambda s: s._get(GX))
    gV = property(lambda s: s._get(GV))
    gtheta = property(lambda s: s._get(GTHETA))
    grho = property(lambda s: s._get(GRHO))
    galpha = property(lambda s: s._get(GALPHA))


def backward_(ws: AdjointWorkspace, tape: Tape, theta, p, *, seedX, seedV, skin=0.2,
              interval=4, maxnb=64, layout="auto"):
    if layout == "auto":
        layout = "particle"
    th = _theta(theta, ws.p)
    sx = _as_2n(seedX, ws.T, ws.N)
    sv = _as_2n(seedV, ws.T, ws.N)
    cp = ws.p.c_struct()
    flags = np.zeros(2, np.int32)
    _fn("orc_backward", ws.T)(ws._h, tape._h, _ptr(th), C.byref(cp), _ptr(sx), _ptr(sv),
                              float(skin), int(interval), int(maxnb),
                              0 if layout == "slot" else 1, _ptr(flags))
    if flags[0]:
        warnings.warn("neighbour list row overflow during backward")
    if flags[1]:
        warnings.warn("displacement exceeded skin/2 during backward")
    return ws


def target_objective(X, xc, yc, sigma, T=np.float64):
    x = _as_2n(X, T)
    seed = np.zeros_like(x)
    J = _fn("orc_target_objective", T)(_ptr(x), x.shape[0], float(xc), float(yc), float(sigma),
                                       _ptr(seed))
    return J, np.ascontiguousarray(seed.T)


def w_kern(q, A, T=np.float64):
    return _fn("orc_w_kern", T)(q, A)


def f_kern(q, A, h, T=np.float64):
    return _fn("orc_f_kern", T)(q, A, h)


def g_kern(q, A, h, T=np.float64):
    return _fn("orc_g_kern", T)(q, A, h)


def wnorm(h, T=np.float64):
    return _fn("orc_wnorm", T)(h)


def interp_alpha(theta, p: SPHParams, x, y):
    th = _theta(theta, p)
    out = np.zeros(3, p.T)
    cp = p.c_struct()
    _fn("orc_interp_alpha", p.T)(_ptr(th), C.byref(cp), x, y, _ptr(out))
    return out
