"""The boundary's own contract on the GPU: a context on a caller-owned stream.

A Julia (AMDGPU) caller hands its own HIP queue to the library
(sph_ctx_create_stream); every call then issues into that stream, and the
results are those of a context with a stream of its own, bit for bit.  The
stream comes from the HIP runtime the library itself links (libamdhip64.so.7),
created at the default priority; the library never destroys it.
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


def hip_runtime():
    L.load()                        # brings in the library's HIP runtime first
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
    hip.hipStreamCreateWithFlags.restype = C.c_int
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.restype = C.c_int
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.restype = C.c_int
    hip.hipSetDevice.argtypes = [C.c_int]
    hip.hipSetDevice.restype = C.c_int
    return hip


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_context_on_caller_stream(precision):
    say("context on a caller-owned stream", precision)
    hip = hip_runtime()
    assert hip.hipSetDevice(0) == 0
    stream = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(stream), 1) == 0     # hipStreamNonBlocking
    try:
        sc = scenes.brinkman_56x37(nsteps=40)
        pg = S.SPHParams(**sc.params)
        out = []
        for be in (S.MI355XBackend(0), S.MI355XBackend(0, stream=stream)):
            st = S.State(be, sc.X0, sc.V0, pg, precision=precision)
            tape = S.Tape(be, sc.N, 40, pg, precision=precision)
            S.simulate_(st, sc.theta, pg, be, 40, tape=tape)
            J, seed = S.target_objective(st, 0.3, 0.2, 0.1)
            ws = S.AdjointWorkspace(be, sc.N, pg, precision=precision)
            S.backward_(ws, tape, sc.theta, pg, be, seedX=seed, seedV=np.zeros_like(seed))
            out.append((st.X, J, ws.gX, ws.gtheta))
            del st, tape, ws, be
        assert out[0][1] == out[1][1]
        for a, b in zip(out[0], out[1]):
            np.testing.assert_array_equal(a, b)
        assert hip.hipStreamSynchronize(stream) == 0    # still alive after the context is gone
    finally:
        assert hip.hipStreamDestroy(stream) == 0


def test_context_rejects_null_stream():
    h = C.c_void_p()
    with pytest.raises(L.SPHError, match="stream is NULL"):
        L.call("sph_ctx_create_stream", 0, None, C.byref(h))


def test_mem_info():
    be = S.MI355XBackend(0)
    free, total = be.mem_info()
    assert 0 < free <= total
