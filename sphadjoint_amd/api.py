"""Python mirror of the reference's public API over libsphadj.so.

Names, argument order and meaning follow src/SPHAdjoint.jl:39-40 with Julia's
``!`` spelled as a trailing underscore (``simulate!`` -> ``simulate_``):

    SPHParams                       src/params.jl:14-28
    State(backend, X0, V0, p; ...)  src/driver.jl:24-36
    CellList / NeighborList views   src/neighbors.jl:9-36, 150-190
    build_(cl, X, p, backend)       src/neighbors.jl:97-114
    build_neighbors_(nl, cl, X, p, backend)   src/neighbors.jl:252-263
    step_(st, theta, p, backend)    src/forward.jl:161-173
    simulate_(st, theta, p, backend, nsteps; tape)   src/driver.jl:92-115
    Tape(backend, N, nsteps, p), reset_(tape)        src/driver.jl:62-71
    AdjointWorkspace(backend, N, p)                  src/driver.jl:128-133
    backward_(ws, tape, theta, p, backend; seedX, seedV, ...)  src/driver.jl:141-207
    target_objective(X, xc, yc, sigma)               src/driver.jl:219-234

``backend`` is an :class:`MI355XBackend` (the KernelAbstractions backend
object's counterpart).  Arrays are numpy: X/V/a and seeds have shape (2, N),
theta has shape (ngy, ngx).  Reading ``st.X`` copies device data to the host,
as ``Array(st.X)`` does in Julia.
"""
from __future__ import annotations

import ctypes as C
import warnings
import weakref

import numpy as np

from . import _lib as L

_PREC = {"f64": L.SPH_F64, "float64": L.SPH_F64, "f32c": L.SPH_F32C}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class MI355XBackend:
    """One HIP device and one in-order stream (the KA backend's role).  With
    ``stream`` (a hipStream_t of the library's HIP runtime, as an int or
    c_void_p) the backend issues into that caller-owned stream instead of a
    stream of its own."""

    def __init__(self, device: int = 0, stream=None):
        h = C.c_void_p()
        if stream is None:
            L.call("sph_ctx_create", int(device), C.byref(h))
        else:
            L.call("sph_ctx_create_stream", int(device), C.c_void_p(int(getattr(stream, "value", stream))),
                   C.byref(h))
        self._h = h
        self.device = device

    def synchronize(self):
        L.call("sph_sync", self._h)

    def mem_info(self):
        """(free, total) bytes of device memory."""
        f, t = C.c_int64(), C.c_int64()
        L.call("sph_mem_info", self._h, C.byref(f), C.byref(t))
        return int(f.value), int(t.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.load().sph_ctx_destroy(h)
            self._h = None


class SPHParams:
    """SPHParams{T} (src/params.jl:14-28).  ``T`` is kept for interface parity;
    the arithmetic form is chosen per State (``precision``)."""

    def __init__(self, *, h, m, rho0, c, mu, dt, Lx, Ly, kw, ngx, ngy, gx=0.0, gy=-9.81,
                 T=np.float64):
        self.h, self.m, self.rho0, self.c, self.mu = float(h), float(m), float(rho0), float(c), float(mu)
        self.gx, self.gy, self.dt = float(gx), float(gy), float(dt)
        self.Lx, self.Ly, self.kw = float(Lx), float(Ly), float(kw)
        self.ngx, self.ngy = int(ngx), int(ngy)
        self.T = np.dtype(T)

    def c_struct(self):
        return L.sph_params(self.h, self.m, self.rho0, self.c, self.mu, self.gx, self.gy, self.dt,
                            self.Lx, self.Ly, self.kw, self.ngx, self.ngy)

    def support(self):
        return 2 * self.h

    def dxg(self):
        return self.Lx / (self.ngx - 1)

    def dyg(self):
        return self.Ly / (self.ngy - 1)


class Field:
    """Device-resident design field theta of shape (ngy, ngx)."""

    def __init__(self, backend: MI355XBackend, theta):
        th = np.asfortranarray(np.asarray(theta, dtype=np.float64))
        if th.ndim != 2:
            raise ValueError("theta must be 2-D (ngy, ngx)")
        self.shape = th.shape
        self.backend = backend
        h = C.c_void_p()
        L.call("sph_field_create", backend._h, th.shape[0], th.shape[1], _ptr(th), C.byref(h))
        self._h = h

    def set(self, theta):
        th = np.asfortranarray(np.asarray(theta, dtype=np.float64))
        if th.shape != self.shape:
            raise ValueError("shape mismatch")
        L.call("sph_field_set", self._h, _ptr(th))

    def __array__(self, dtype=None, copy=None):
        out = np.empty(self.shape[::-1], np.float64)
        L.call("sph_field_get", self._h, _ptr(out))
        out = out.T
        return out.astype(dtype) if dtype is not None else out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.load().sph_field_destroy(h)
            self._h = None


def _field(theta, p: SPHParams, backend: MI355XBackend) -> Field:
    if isinstance(theta, Field) and getattr(theta, "batched", False):
        return theta                      # the library checks the node count
    if isinstance(theta, Field):
        f = theta
    else:
        f = Field(backend, theta)
    if f.shape != (p.ngy, p.ngx):
        raise ValueError(f"theta must be (ngy, ngx) = {(p.ngy, p.ngx)}, got {f.shape}")
    return f


class DeviceArray:
    """A (2, N) float64 array in GPU memory owned by the library (sph_buffer),
    stored interleaved (x, y per particle) as the ABI expects.  Inputs and
    outputs of the device-memory calls: initial states, adjoint seeds."""

    def __init__(self, backend: MI355XBackend, A=None, N=None):
        if A is not None:
            a = _as_2n(A, N)
            self.N = a.shape[0]
            src = _ptr(a)
        else:
            self.N = int(N)
            src = None
        self.backend = backend
        h = C.c_void_p()
        L.call("sph_buffer_create", backend._h, 16 * self.N, src, C.byref(h))
        self._h = h
        p = C.c_void_p()
        L.call("sph_buffer_ptr", self._h, C.byref(p))
        self.ptr = p

    def numpy(self):
        out = np.empty((self.N, 2))
        L.call("sph_buffer_read", self._h, _ptr(out), 16 * self.N)
        return np.ascontiguousarray(out.T)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.load().sph_buffer_destroy(h)
            self._h = None


def _is_device(A):
    return isinstance(A, DeviceArray)


def _dev_2n(A, N):
    if A.N != N:
        raise ValueError(f"expected a device array of N={N}, got {A.N}")
    return A, A.ptr


def _as_2n(A, N=None):
    A = np.asarray(A, dtype=np.float64)
    if A.ndim != 2 or A.shape[0] != 2:
        raise ValueError("expected an array of shape (2, N)")
    if N is not None and A.shape[1] != N:
        raise ValueError(f"expected (2, {N}), got {A.shape}")
    return np.ascontiguousarray(A.T)       # interleaved x,y per particle


class CellList:
    """View of a state's cell list (src/neighbors.jl:9-19)."""

    def __init__(self, st):
        self._ref = weakref.ref(st)
        info = st._info()
        self.nx, self.ny = int(info[0]), int(info[1])
        self.cs = st.rc

    def arrays(self):
        st = self._st
        n, nc = st.N, self.nx * self.ny
        cellof = np.zeros(n, np.int32)
        counts = np.zeros(nc, np.int32)
        starts = np.zeros(nc, np.int32)
        order = np.zeros(n, np.int32)
        L.call("sph_cell_list_get", st._h, _ptr(cellof), _ptr(counts), _ptr(starts), _ptr(order))
        return cellof, counts, starts, order

    _st = property(lambda s: s._ref())
    cellof = property(lambda s: s.arrays()[0])
    counts = property(lambda s: s.arrays()[1])
    starts = property(lambda s: s.arrays()[2])
    order = property(lambda s: s.arrays()[3])


class NeighborList:
    """View of a state's neighbour list (src/neighbors.jl:150-162).  The list
    is materialised from the positions of the last build on request."""

    def __init__(self, st):
        self._ref = weakref.ref(st)
        info = st._info()
        self.maxnb, self.sm, self.si = int(info[3]), int(info[4]), int(info[5])
        self.rc = st.rc
        self.dmax = st.p.h * st.skin

    _st = property(lambda s: s._ref())

    @property
    def interval(self):
        return int(self._st._info()[6])

    @property
    def age(self):
        return int(self._st._info()[7])

    def arrays(self):
        st = self._st
        idx = np.zeros(st.N * self.maxnb, np.int32)
        cnt = np.zeros(st.N, np.int32)
        L.call("sph_neighbors_get", st._h, _ptr(idx), _ptr(cnt))
        return idx, cnt

    indices = property(lambda s: s.arrays()[0])
    counts = property(lambda s: s.arrays()[1])

    @property
    def flags(self):
        f = np.zeros(2, np.int32)
        L.call("sph_flags_get", self._st._h, _ptr(f))
        return f

    def rows(self):
        idx, cnt = self.arrays()
        return [[int(idx[m * self.sm + i * self.si]) for m in range(int(cnt[i]))]
                for i in range(self._st.N)]


class State:
    """State (src/driver.jl:5-36).  ``precision`` selects the arithmetic form:
    "f64" (the reference's Float64 arithmetic) or "f32c"."""

    def __init__(self, backend: MI355XBackend, X0, V0, p: SPHParams, *, skin=0.2, interval=4,
                 maxnb=64, layout="auto", precision="f64"):
        if layout == "auto":
            layout = "slot"                        # GPU backends pick slot-major
        if layout not in ("slot", "particle"):
            raise ValueError("layout must be :slot / :particle / :auto")
        self.backend, self.p = backend, p
        self.skin = float(skin)
        self.rc = (2 * p.h) * (1 + self.skin)
        self.precision = _PREC[precision]
        x0 = _as_2n(X0)
        self.N = x0.shape[0]
        v0 = _as_2n(V0, self.N)
        cp = p.c_struct()
        h = C.c_void_p()
        L.call("sph_state_create", backend._h, C.byref(cp), self.precision, self.N, _ptr(x0),
               _ptr(v0), self.skin, int(interval), int(maxnb),
               L.SPH_LAYOUT_SLOT if layout == "slot" else L.SPH_LAYOUT_PARTICLE, C.byref(h))
        self._h = h
        self.cl = CellList(self)
        self.nl = NeighborList(self)

    @classmethod
    def _borrow(cls, h, owner, backend, p, N, skin, precision):
        # a state owned by another object (a Checkpointer): never destroyed here
        st = cls.__new__(cls)
        st.backend, st.p, st.skin, st.N = backend, p, float(skin), int(N)
        st.rc = (2 * p.h) * (1 + st.skin)
        st.precision = precision
        st._h, st._owner = h, owner
        st.cl = CellList(st)
        st.nl = NeighborList(st)
        return st

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_owner", None) is None:
            L.load().sph_state_destroy(h)
        self._h = None

    def _info(self):
        o = np.zeros(8, np.int64)
        L.call("sph_state_info", self._h, _ptr(o))
        return o

    def _get(self, which, two):
        out = np.empty((self.N, 2) if two else self.N, np.float64)
        L.call("sph_state_get", self._h, which, _ptr(out))
        return np.ascontiguousarray(out.T) if two else out

    def _set(self, which, A):
        a = _as_2n(A, self.N)
        L.call("sph_state_set", self._h, which, _ptr(a))

    X = property(lambda s: s._get(L.SPH_X, True), lambda s, v: s._set(L.SPH_X, v))
    V = property(lambda s: s._get(L.SPH_V, True), lambda s, v: s._set(L.SPH_V, v))
    a = property(lambda s: s._get(L.SPH_A, True))
    rho = property(lambda s: s._get(L.SPH_RHO, False))
    pterm = property(lambda s: s._get(L.SPH_PTERM, False))
    invrho = property(lambda s: s._get(L.SPH_INVRHO, False))
    alpha = property(lambda s: s._get(L.SPH_ALPHA, False))


def check_neighbor_flags(nl: NeighborList):
    """check_neighbor_flags (src/neighbors.jl:282-290): warn, then reset."""
    f = nl.flags
    f3 = np.zeros(3, np.int32)
    L.call("sph_flags_get_all", nl._st._h, _ptr(f3))
    if f[0] != 0:
        # (the f32c form never sets it: its passes test the slots past the masks)
        warnings.warn(f"neighbour list row overflow (maxnb={nl.maxnb}); increase maxnb")
    if f[1] != 0:
        warnings.warn("neighbour-list rebuild interval too long (a displacement exceeded "
                      "skin*h); reduce interval or increase skin")
    if f3[2] != 0:
        warnings.warn("non-finite particle positions at a neighbour-list rebuild (the "
                      "simulation has diverged); they were binned into an edge cell")
    L.call("sph_flags_clear", nl._st._h)
    return f


def _state_of(obj):
    if isinstance(obj, State):
        return obj
    if isinstance(obj, (CellList, NeighborList)):
        return obj._st
    raise TypeError("expected a State, CellList or NeighborList of this backend")


def build_(cl, X=None, p=None, backend=None):
    """build!(cl, X, p, backend): rebuild the cell list from the state's X."""
    st = _state_of(cl)
    L.call("sph_build", st._h)
    return cl


def build_neighbors_(nl, cl=None, X=None, p=None, backend=None):
    """build_neighbors!(nl, cl, X, p, backend): cell list + neighbour list; age = 0."""
    st = _state_of(nl)
    L.call("sph_build_neighbors", st._h)
    return nl


def step_(st: State, theta, p: SPHParams = None, backend: MI355XBackend = None):
    f = _field(theta, st.p, st.backend)
    L.call("sph_step", st._h, f._h)
    return st


class Tape:
    """Tape (src/driver.jl:45-83): the state before every step, on the device.
    keep_pi / keep_lists (the latter f32c only) also keep each step's
    (pterm, invrho) and exact neighbour masks, so that backward_ skips the
    density recompute and the pair tests (sph_tape_create_keep); same
    gradients, more device memory."""

    def __init__(self, backend: MI355XBackend, N, nsteps, p: SPHParams, precision="f64", *,
                 keep_pi=False, keep_lists=False):
        self.backend, self.N = backend, int(N)
        h = C.c_void_p()
        keep = (1 if keep_pi else 0) | (2 if keep_lists else 0)
        if keep:
            L.call("sph_tape_create_keep", backend._h, self.N, int(nsteps), _PREC[precision],
                   keep, C.byref(h))
        else:
            L.call("sph_tape_create", backend._h, self.N, int(nsteps), _PREC[precision],
                   C.byref(h))
        self._h = h

    def __len__(self):
        n = C.c_int64()
        L.call("sph_tape_length", self._h, C.byref(n))
        return int(n.value)

    def capacity(self):
        n = C.c_int64()
        L.call("sph_tape_capacity", self._h, C.byref(n))
        return int(n.value)

    def get(self, k):
        """(tape.X[:, :, k], tape.V[:, :, k]) for 1-based k."""
        x = np.empty((self.N, 2))
        v = np.empty((self.N, 2))
        L.call("sph_tape_get", self._h, int(k), _ptr(x), _ptr(v))
        return np.ascontiguousarray(x.T), np.ascontiguousarray(v.T)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.load().sph_tape_destroy(h)
            self._h = None


def reset_(tape: Tape):
    L.call("sph_tape_reset", tape._h)
    return tape


def capacity(tape: Tape):
    return tape.capacity()


def simulate_(st: State, theta, p: SPHParams, backend: MI355XBackend, nsteps, tape: Tape = None):
    """simulate!: nsteps steps; with a tape, the state before each step is kept."""
    f = _field(theta, st.p, st.backend)
    L.call("sph_simulate", st._h, f._h, int(nsteps), tape._h if tape is not None else None)
    check_neighbor_flags(st.nl)
    return st


def simulate_graph_(st: State, theta, p: SPHParams, backend: MI355XBackend, nsteps):
    """simulate! without a tape, its steps replayed from a captured HIP graph
    (same results as simulate_).  Returns the number of graph launches."""
    f = _field(theta, st.p, st.backend)
    n = C.c_int64()
    L.call("sph_simulate_graph", st._h, f._h, int(nsteps), C.byref(n))
    check_neighbor_flags(st.nl)
    return int(n.value)


def autotune_(st: State, theta, steps=4):
    """Picks the launch configuration of the f32c pair kernels on this state
    (advances it by 12*steps steps): the workgroup size (64, 128, 256),
    negated for a kernel that is faster reading its neighbours from global
    memory than from LDS-staged rows.  Returns {kernel: configuration}."""
    f = _field(theta, st.p, st.backend)
    out = np.zeros(6, np.int32)
    L.call("sph_autotune", st._h, f._h, int(steps), _ptr(out))
    names = ("nl_build", "density", "force_integrate", "adjoint_r", "adjoint_p1", "adjoint_p2")
    return dict(zip(names, out.tolist()))


def tuning(st: State, sizes=None):
    """The tuned workgroup sizes of an f32c state; with ``sizes`` (as returned
    by autotune_ or tuning), sets them."""
    names = ("nl_build", "density", "force_integrate", "adjoint_r", "adjoint_p1", "adjoint_p2")
    if sizes is not None:
        v = np.array([int(sizes[n]) for n in names], np.int32)
        L.call("sph_tuning_set", st._h, _ptr(v))
    out = np.zeros(6, np.int32)
    L.call("sph_tuning_get", st._h, _ptr(out))
    return dict(zip(names, out.tolist()))


def set_sweep_(st, sweep):
    """The candidate sweep of an f32c state's neighbour-list builds
    (sph_state_set_sweep): 5 orders each cell's slots by quadrant and builds
    the list from the 5x5 half cells around each slot, 3 (the default) is the
    reference's 3x3 sweep.  Takes a State or a Checkpointer (its work state);
    it applies from the next rebuild."""
    if isinstance(st, Checkpointer):
        st = st.state
    L.call("sph_state_set_sweep", st._h, int(sweep))


class AdjointWorkspace:
    """AdjointWorkspace (src/driver.jl:117-133): gX, gV, gtheta after backward_."""

    def __init__(self, backend: MI355XBackend, N, p: SPHParams, precision="f64"):
        self.backend, self.N, self.p = backend, int(N), p
        cp = p.c_struct()
        h = C.c_void_p()
        L.call("sph_workspace_create", backend._h, self.N, C.byref(cp), _PREC[precision], C.byref(h))
        self._h = h

    def _get(self, which):
        if which == L.SPH_GTHETA:
            out = np.empty((self.p.ngx, self.p.ngy))
            L.call("sph_workspace_get", self._h, which, _ptr(out))
            return np.ascontiguousarray(out.T)
        out = np.empty((self.N, 2))
        L.call("sph_workspace_get", self._h, which, _ptr(out))
        return np.ascontiguousarray(out.T)

    gX = property(lambda s: s._get(L.SPH_GX))
    gV = property(lambda s: s._get(L.SPH_GV))
    gtheta = property(lambda s: s._get(L.SPH_GTHETA))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.load().sph_workspace_destroy(h)
            self._h = None


def backward_(ws: AdjointWorkspace, tape: Tape, theta, p: SPHParams, backend: MI355XBackend, *,
              seedX, seedV, scratch=None, skin=0.2, interval=4, maxnb=64, layout="auto"):
    """backward!: reverse sweep over the tape; fills ws.gX, ws.gV, ws.gtheta."""
    if layout == "auto":
        layout = "slot"
    f = _field(theta, ws.p, ws.backend)
    lay = L.SPH_LAYOUT_SLOT if layout == "slot" else L.SPH_LAYOUT_PARTICLE
    if _is_device(seedX) and _is_device(seedV):
        bx, px = _dev_2n(seedX, ws.N)
        bv, pv = _dev_2n(seedV, ws.N)
        L.call("sph_backward_dev", ws._h, tape._h, f._h, px, pv, float(skin), int(interval),
               int(maxnb), lay)
        return ws
    sx = _as_2n(seedX, ws.N)
    sv = _as_2n(seedV, ws.N)
    L.call("sph_backward", ws._h, tape._h, f._h, _ptr(sx), _ptr(sv), float(skin), int(interval),
           int(maxnb), lay)
    return ws


class Checkpointer:
    """Binomial (revolve) checkpointing of simulate!/backward! over ``nsteps``.

    The reference's Tape keeps the state before every step (src/driver.jl:38-83);
    this cuts the horizon into segments of ``seg_epochs`` neighbour-list epochs
    (``seg_epochs * interval`` steps), keeps at most ``snapshots`` segment-start
    states at the positions of Griewank & Walther's binomial schedule, and
    recomputes each segment onto a one-segment tape right before its reverse
    sweep.  Gradients are bit-identical to a full-tape ``backward_``.  For E
    segments, recomputation costs about ``t`` extra forward passes where
    C(snapshots + 1 + t, t) >= E; with snapshots >= E - 1 every segment runs
    twice in all (the minimum for any scheme that does not keep every step).

    ``forward(X0, V0, theta)`` runs the whole horizon and returns the final
    state (owned by the checkpointer; ``backward`` reuses its buffers, so read
    the final state before calling ``backward``).

    ``tape_steps`` lets the segment tape hold that many steps: the last
    segments that fit are recorded in the forward and never recomputed, and
    the schedule covers the segments before them (same gradients).  The tape
    costs 80 B (f32c) or 48 B (FP64) of device memory per particle-step.
    """

    def __init__(self, backend: MI355XBackend, N, nsteps, p: SPHParams, *, snapshots=64,
                 seg_epochs=1, skin=0.2, interval=4, maxnb=64, precision="f64",
                 tape_steps=None):
        self.backend, self.p, self.N = backend, p, int(N)
        self.nsteps, self.skin, self.interval = int(nsteps), float(skin), int(interval)
        self.seg_epochs = int(seg_epochs)
        self.precision = _PREC[precision]
        cp = p.c_struct()
        h = C.c_void_p()
        L.call("sph_ckpt_create", backend._h, C.byref(cp), self.precision, self.N, self.nsteps,
               int(snapshots), self.seg_epochs, self.skin, self.interval, int(maxnb), C.byref(h))
        self._h = h
        self._stref = None
        self._theta = None
        if tape_steps is not None:
            # the last segments that fit on a tape of tape_steps steps are
            # recorded in the forward, not recomputed (sph_ckpt_set_tape_steps)
            L.call("sph_ckpt_set_tape_steps", h, int(tape_steps))

    @property
    def state(self):
        """The work state (borrowed: it keeps the checkpointer alive, and the
        checkpointer refers to it only weakly, so that no reference cycle
        delays the release of the device memory)."""
        st = self._stref() if self._stref is not None else None
        if st is None:
            sh = C.c_void_p()
            L.call("sph_ckpt_state", self._h, C.byref(sh))
            st = State._borrow(sh, self, self.backend, self.p, self.N, self.skin, self.precision)
            self._stref = weakref.ref(st)
        return st

    def forward(self, X0, V0, theta, check_flags=True):
        """X0, V0: (2, N) numpy arrays, or DeviceArrays."""
        self._theta = _field(theta, self.p, self.backend)
        if _is_device(X0) and _is_device(V0):
            bx, px = _dev_2n(X0, self.N)
            bv, pv = _dev_2n(V0, self.N)
            L.call("sph_ckpt_forward_dev", self._h, px, pv, self._theta._h)
        else:
            x0 = _as_2n(X0, self.N)
            v0 = _as_2n(V0, self.N)
            L.call("sph_ckpt_forward", self._h, _ptr(x0), _ptr(v0), self._theta._h)
        st = self.state
        if check_flags:
            check_neighbor_flags(st.nl)
        return st

    def backward(self, ws: AdjointWorkspace, *, seedX, seedV, theta=None):
        """seedX, seedV: (2, N) numpy arrays, or DeviceArrays."""
        f = self._theta if theta is None else _field(theta, self.p, self.backend)
        if f is None:
            raise RuntimeError("Checkpointer.backward needs a preceding forward")
        if _is_device(seedX) and _is_device(seedV):
            bx, px = _dev_2n(seedX, ws.N)
            bv, pv = _dev_2n(seedV, ws.N)
            L.call("sph_ckpt_backward_dev", self._h, ws._h, f._h, px, pv)
            return ws
        sx = _as_2n(seedX, ws.N)
        sv = _as_2n(seedV, ws.N)
        L.call("sph_ckpt_backward", self._h, ws._h, f._h, _ptr(sx), _ptr(sv))
        return ws

    def stats(self):
        o = np.zeros(5, np.int64)
        L.call("sph_ckpt_stats", self._h, _ptr(o))
        return dict(epochs=int(o[0]), snapshots=int(o[1]), forward_epochs=int(o[2]),
                    backward_forward_epochs=int(o[3]), backward_taped_epochs=int(o[4]))

    def __del__(self):
        # no borrowed state is alive here: each holds the checkpointer
        h = getattr(self, "_h", None)
        if h:
            L.load().sph_ckpt_destroy(h)
            self._h = None


def ckpt_plan(epochs, snapshots):
    """The binomial schedule on the host: forward epochs run in total (the
    taped ones included), epochs taped, most snapshots alive at once."""
    o = np.zeros(3, np.int64)
    L.call("sph_ckpt_plan", int(epochs), int(snapshots), _ptr(o))
    return dict(forward_epochs=int(o[0]), taped=int(o[1]), peak=int(o[2]))


class BatchField(Field):
    """The design fields of a batch's scenes, one after the other (each (ngy, ngx)
    column-major), as one device field."""
    batched = True

    def __init__(self, backend: MI355XBackend, thetas):
        flat = np.concatenate([np.asarray(t, dtype=np.float64).ravel(order="F") for t in thetas])
        super().__init__(backend, flat.reshape(-1, 1))
        self.shapes = [np.shape(t) for t in thetas]


class BatchState(State):
    """Several independent scenes in one f32c state (sph_batch_create): every
    kernel launch covers all of them.  ``scenes`` is a list of (SPHParams, X0,
    V0); X/V of the state are the scenes' particles in order."""

    def __init__(self, backend: MI355XBackend, scenes, *, skin=0.2, interval=4, maxnb=64):
        params = (L.sph_params * len(scenes))(*[sp.c_struct() for sp, _, _ in scenes])
        Ns = np.array([np.shape(x0)[1] for _, x0, _ in scenes], np.int64)
        x0 = np.ascontiguousarray(np.concatenate([_as_2n(x) for _, x, _ in scenes]))
        v0 = np.ascontiguousarray(np.concatenate([_as_2n(v) for _, _, v in scenes]))
        self.backend, self.p = backend, scenes[0][0]
        self.skin = float(skin)
        self.rc = (2 * self.p.h) * (1 + self.skin)
        self.precision = L.SPH_F32C
        self.N = int(Ns.sum())
        h = C.c_void_p()
        L.call("sph_batch_create", backend._h, len(scenes), params, _ptr(Ns), _ptr(x0), _ptr(v0),
               self.skin, int(interval), int(maxnb), C.byref(h))
        self._h = h
        self.cl = CellList(self)
        self.nl = NeighborList(self)
        self.params = [sp for sp, _, _ in scenes]
        n = C.c_int64()
        info = np.zeros(4 * len(scenes), np.int64)
        L.call("sph_batch_info", self._h, C.byref(n), _ptr(info))
        self.layout_info = info.reshape(-1, 4)
        self.nscenes = int(n.value)

    def split(self, A):
        """(2, N) array -> list of the scenes' (2, N_k) arrays."""
        return [A[:, o:o + n] for o, n, _, _ in self.layout_info]

    def field(self, thetas):
        return BatchField(self.backend, thetas)


class BatchWorkspace(AdjointWorkspace):
    """AdjointWorkspace of a BatchState; gtheta is the list of the scenes' fields."""

    def __init__(self, backend: MI355XBackend, st: BatchState):
        self.backend, self.N, self.p = backend, st.N, st.p
        self._st_info = st.layout_info
        self._shapes = [(sp.ngy, sp.ngx) for sp in st.params]
        h = C.c_void_p()
        L.call("sph_batch_workspace_create", backend._h, st._h, C.byref(h))
        self._h = h

    def _get(self, which):
        if which == L.SPH_GTHETA:
            tot = int(sum(a * b for a, b in self._shapes))
            out = np.empty(tot)
            L.call("sph_workspace_get", self._h, which, _ptr(out))
            res, o = [], 0
            for ny, nx in self._shapes:
                res.append(out[o:o + ny * nx].reshape((ny, nx), order="F").copy())
                o += ny * nx
            return res
        return super()._get(which)


def batch_objective(st: BatchState, xc, yc, sigma, device=False):
    """target_objective per scene: (J of shape (nscenes,), dJ/dX_T of shape (2, N))."""
    J = np.zeros(st.nscenes)
    if device:
        buf = device if isinstance(device, DeviceArray) else DeviceArray(st.backend, N=st.N)
        L.call("sph_batch_objective_dev", st._h, float(xc), float(yc), float(sigma), _ptr(J),
               buf.ptr)
        return J, buf
    seed = np.empty((st.N, 2))
    L.call("sph_batch_objective", st._h, float(xc), float(yc), float(sigma), _ptr(J), _ptr(seed))
    return J, np.ascontiguousarray(seed.T)


def target_objective(st: State, xc, yc, sigma, device=False):
    """target_objective(st.X, xc, yc, sigma) -> (J, dJ/dX_T of shape (2, N)), evaluated
    on the device.  With ``device=True`` (or a DeviceArray to fill) the seed stays
    in GPU memory as a :class:`DeviceArray`."""
    if not isinstance(st, State):
        raise TypeError("target_objective takes the State whose X is evaluated on the device")
    J = C.c_double()
    if device:
        buf = device if isinstance(device, DeviceArray) else DeviceArray(st.backend, N=st.N)
        L.call("sph_target_objective_dev", st._h, float(xc), float(yc), float(sigma), C.byref(J),
               buf.ptr)
        return J.value, buf
    seed = np.empty((st.N, 2))
    L.call("sph_target_objective", st._h, float(xc), float(yc), float(sigma), C.byref(J), _ptr(seed))
    return J.value, np.ascontiguousarray(seed.T)


def profile_enable(backend: MI355XBackend, on: bool = True, only: str = None):
    """Per-kernel GPU timing with events on the backend's stream (resets counters).
    With ``only`` (a kernel name of profile_stats), that kernel alone is timed."""
    mode = 1 if on else 0
    if on and only is not None:
        names = kernel_names()
        if only not in names:
            raise ValueError(f"unknown kernel {only!r}; known: {names}")
        mode = 2 + names.index(only)
    L.call("sph_profile_enable", backend._h, mode)


def kernel_names():
    lib = L.load()
    out, k = [], 0
    while True:
        name = lib.sph_profile_name(k)
        if name is None:
            return out
        out.append(name.decode())
        k += 1


def profile_stats(backend: MI355XBackend):
    """{kernel name: (launches, total ms)} since profile_enable."""
    out = {}
    k = 0
    lib = L.load()
    while True:
        name = lib.sph_profile_name(k)
        if name is None:
            break
        n = C.c_int64()
        ms = C.c_double()
        L.call("sph_profile_get", backend._h, k, C.byref(n), C.byref(ms))
        out[name.decode()] = (int(n.value), float(ms.value))
        k += 1
    return out
