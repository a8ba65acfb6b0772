"""sphadjoint_amd: MI355X (gfx950) implementation of SPHAdjoint's differentiable
2D WCSPH step behind the C ABI of include/sphadj.h.

The public names mirror the reference's exports (src/SPHAdjoint.jl:39-40),
with Julia's ``!`` written as a trailing underscore.
"""
from .api import (AdjointWorkspace, BatchField, BatchState, BatchWorkspace, CellList,
                  Checkpointer, DeviceArray, Field, MI355XBackend, NeighborList, SPHParams,
                  State, Tape, autotune_, backward_, batch_objective, build_, build_neighbors_,
                  capacity, check_neighbor_flags, ckpt_plan, reset_, set_sweep_, simulate_,
                  simulate_graph_, step_, target_objective, tuning)
from .design import DesignChain, Optimizer, Sensitivity

__all__ = ["SPHParams", "State", "CellList", "NeighborList", "AdjointWorkspace", "Tape",
           "simulate_", "step_", "build_", "build_neighbors_", "backward_", "target_objective",
           "reset_", "capacity", "check_neighbor_flags", "MI355XBackend", "Field", "Checkpointer",
           "ckpt_plan", "DeviceArray", "BatchState", "BatchWorkspace", "BatchField",
           "batch_objective", "simulate_graph_", "autotune_",
           "tuning", "set_sweep_", "DesignChain", "Optimizer", "Sensitivity"]
