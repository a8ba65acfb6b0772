"""ctypes binding of libsphadj.so (the C ABI declared in include/sphadj.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (make -C
sphadjoint_amd/csrc).  There is no fallback: if the library is missing or the
GPU is absent, loading fails with an error that says so.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SPHADJ_LIB names another build of the same library (e.g. a variant under test)
LIB_PATH = os.environ.get("SPHADJ_LIB") or os.path.join(_HERE, "libsphadj.so")

SPH_OK, SPH_E_ARG, SPH_E_CAPACITY, SPH_E_HIP, SPH_E_STATE, SPH_E_NOMEM = range(6)
SPH_F64, SPH_F32C = 0, 1
SPH_LAYOUT_SLOT, SPH_LAYOUT_PARTICLE = 0, 1
SPH_X, SPH_V, SPH_A, SPH_RHO, SPH_PTERM, SPH_INVRHO, SPH_ALPHA = range(7)
SPH_GX, SPH_GV, SPH_GTHETA = range(3)

# every entry point of include/sphadj.h: name -> (restype, argtypes)
_vp = C.c_void_p
_i64 = C.c_int64
_d = C.c_double
_pp = C.POINTER(C.c_void_p)


class sph_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("h", "m", "rho0", "c", "mu", "gx", "gy", "dt", "Lx", "Ly", "kw")] + \
               [("ngx", C.c_int64), ("ngy", C.c_int64)]


_PP = C.POINTER(sph_params)
SIGNATURES = {
    "sph_last_error": (C.c_char_p, []),
    "sph_ctx_create": (C.c_int, [C.c_int, _pp]),
    "sph_ctx_create_stream": (C.c_int, [C.c_int, _vp, _pp]),
    "sph_ctx_destroy": (C.c_int, [_vp]),
    "sph_mem_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "sph_sync": (C.c_int, [_vp]),
    "sph_field_create": (C.c_int, [_vp, _i64, _i64, _vp, _pp]),
    "sph_field_set": (C.c_int, [_vp, _vp]),
    "sph_field_get": (C.c_int, [_vp, _vp]),
    "sph_field_destroy": (C.c_int, [_vp]),
    "sph_state_create": (C.c_int, [_vp, _PP, C.c_int, _i64, _vp, _vp, _d, _i64, _i64, C.c_int, _pp]),
    "sph_state_destroy": (C.c_int, [_vp]),
    "sph_state_get": (C.c_int, [_vp, C.c_int, _vp]),
    "sph_state_set": (C.c_int, [_vp, C.c_int, _vp]),
    "sph_state_info": (C.c_int, [_vp, _vp]),
    "sph_build": (C.c_int, [_vp]),
    "sph_cell_list_get": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "sph_build_neighbors": (C.c_int, [_vp]),
    "sph_neighbors_get": (C.c_int, [_vp, _vp, _vp]),
    "sph_flags_get": (C.c_int, [_vp, _vp]),
    "sph_flags_get_all": (C.c_int, [_vp, _vp]),
    "sph_flags_clear": (C.c_int, [_vp]),
    "sph_step": (C.c_int, [_vp, _vp]),
    "sph_simulate": (C.c_int, [_vp, _vp, _i64, _vp]),
    "sph_simulate_graph": (C.c_int, [_vp, _vp, _i64, C.POINTER(_i64)]),
    "sph_tape_create": (C.c_int, [_vp, _i64, _i64, C.c_int, _pp]),
    "sph_tape_create_keep": (C.c_int, [_vp, _i64, _i64, C.c_int, C.c_int, _pp]),
    "sph_tape_destroy": (C.c_int, [_vp]),
    "sph_tape_reset": (C.c_int, [_vp]),
    "sph_tape_length": (C.c_int, [_vp, C.POINTER(_i64)]),
    "sph_tape_capacity": (C.c_int, [_vp, C.POINTER(_i64)]),
    "sph_tape_get": (C.c_int, [_vp, _i64, _vp, _vp]),
    "sph_workspace_create": (C.c_int, [_vp, _i64, _PP, C.c_int, _pp]),
    "sph_workspace_destroy": (C.c_int, [_vp]),
    "sph_workspace_get": (C.c_int, [_vp, C.c_int, _vp]),
    "sph_backward": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _d, _i64, _i64, C.c_int]),
    "sph_target_objective": (C.c_int, [_vp, _d, _d, _d, C.POINTER(_d), _vp]),
    "sph_buffer_create": (C.c_int, [_vp, _i64, _vp, _pp]),
    "sph_buffer_destroy": (C.c_int, [_vp]),
    "sph_buffer_ptr": (C.c_int, [_vp, _pp]),
    "sph_buffer_read": (C.c_int, [_vp, _vp, _i64]),
    "sph_buffer_write": (C.c_int, [_vp, _vp, _i64]),
    "sph_backward_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _d, _i64, _i64, C.c_int]),
    "sph_target_objective_dev": (C.c_int, [_vp, _d, _d, _d, C.POINTER(_d), _vp]),
    "sph_target_objective_noalloc": (C.c_int, [_vp, _d, _d, _d, _vp, _vp]),
    "sph_ckpt_forward_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "sph_ckpt_backward_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "sph_batch_create": (C.c_int, [_vp, _i64, _PP, _vp, _vp, _vp, _d, _i64, _i64, _pp]),
    "sph_batch_workspace_create": (C.c_int, [_vp, _vp, _pp]),
    "sph_batch_info": (C.c_int, [_vp, C.POINTER(_i64), _vp]),
    "sph_batch_objective": (C.c_int, [_vp, _d, _d, _d, _vp, _vp]),
    "sph_batch_objective_dev": (C.c_int, [_vp, _d, _d, _d, _vp, _vp]),
    "sph_autotune": (C.c_int, [_vp, _vp, _i64, _vp]),
    "sph_tuning_get": (C.c_int, [_vp, _vp]),
    "sph_tuning_set": (C.c_int, [_vp, _vp]),
    "sph_state_set_sweep": (C.c_int, [_vp, C.c_int]),
    "sph_ckpt_create": (C.c_int, [_vp, _PP, C.c_int, _i64, _i64, _i64, _i64, _d, _i64, _i64,
                                  _pp]),
    "sph_ckpt_destroy": (C.c_int, [_vp]),
    "sph_ckpt_forward": (C.c_int, [_vp, _vp, _vp, _vp]),
    "sph_ckpt_state": (C.c_int, [_vp, _pp]),
    "sph_ckpt_backward": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "sph_ckpt_stats": (C.c_int, [_vp, _vp]),
    "sph_ckpt_set_tape_steps": (C.c_int, [_vp, _i64]),
    "sph_ckpt_plan": (C.c_int, [_i64, _i64, _vp]),
    "sph_design_create": (C.c_int, [_vp, _i64, _i64, _i64, _d, _d, _pp]),
    "sph_design_destroy": (C.c_int, [_vp]),
    "sph_design_theta": (C.c_int, [_vp, _vp, _vp, _vp, _i64]),
    "sph_design_grad": (C.c_int, [_vp, _vp, _vp, _i64, _vp]),
    "sph_design_project": (C.c_int, [_vp, _vp, _d, _vp]),
    "sph_design_candidate": (C.c_int, [_vp, _vp, _vp, _d, _d, _vp, C.POINTER(_d)]),
    "sph_profile_enable": (C.c_int, [_vp, C.c_int]),
    "sph_profile_get": (C.c_int, [_vp, C.c_int, C.POINTER(_i64), C.POINTER(_d)]),
    "sph_profile_name": (C.c_char_p, [C.c_int]),
}


class SPHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[sph status {code}] {msg}")
        self.code = code


class CapacityError(SPHError, ValueError):
    """Tape too small (the reference raises ArgumentError, src/driver.jl:96-98)."""


_lib = None


def load():
    """Load libsphadj.so and declare every entry point; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (make -C sphadjoint_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if os.environ.get("SPHADJ_LIB") and not hasattr(L, name):
                continue          # an older build under comparison lacks newer entry points
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc != SPH_OK:
        msg = load().sph_last_error().decode(errors="replace")
        if rc == SPH_E_CAPACITY:
            raise CapacityError(rc, msg)
        if rc == SPH_E_ARG:
            raise SPHError(rc, msg)
        raise SPHError(rc, msg)
    return rc
