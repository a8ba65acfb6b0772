#!/usr/bin/env python3
"""Forward+adjoint particle-steps/s of the differentiable 2D WCSPH step on one MI355X.

Workload (the 1M-particle configuration): a 1000x1000 fluid block, dx = 1e-3,
h = 1.2 dx, tank 2.0 x 1.2, theta on a 512x256 grid (U(0,1), 5x5 box filter);
one bench step is one evaluation of that configuration: a horizon of H = 500
physical steps forward, the Gaussian target objective J on the device, then
the discrete adjoint back over the horizon (dJ/dX0, dJ/dV0, dJ/dtheta).
particle-steps/s = N * H / wall time of a step; recomputation under
checkpointing is not credited.  Inputs (X0, V0, theta) are resident in HBM
before any timed region and the adjoint seeds stay on the device.

Forms (each evaluates the same horizon):
  f64        plain FP64 (the reference's arithmetic) over a full tape of X, V
  f64_ckpt   FP64 under the same checkpointing as f32c_ckpt (like-for-like)
  f32c_ckpt  FP32 state with compensated sums, FP64 positions / J / gtheta,
             binomial checkpointing under a snapshot budget       -> `value`
  f32c       the f32c arithmetic over a full tape that also keeps each step's
             (pterm, invrho) and exact neighbour masks
  f64_keep   FP64 over a full tape that also keeps (pterm, invrho) (the tape
             kind of f32c, for a like-for-like full-tape ratio)

Timing: W warm-up steps, then K trials; in each trial every form runs one
step, bracketed by a device synchronisation on both sides, the order of the
forms alternating from trial to trial.  `value` = N*H*K / the K f32c_ckpt
steps' summed time; ratios are medians of per-trial ratios.  No events are
recorded in the timed trials: the per-kernel table and the roofline come
from two further passes per form (one with events around every kernel, one
with events around the dominant kernel's launches only).
"""
import argparse
import gc
import json
import math
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+adjoint particle-steps/s, N=1M FP32 vs FP64 plain; batched scenes/s"
XC, YC, SIGMA = 0.5, 0.5, 0.1          # Gaussian target inside the block

# Algorithmic HBM bytes per particle of each neighbour pass: the minimum
# compulsory traffic (each per-particle array read or written once per pass,
# neighbour-side reads served on chip), Sx = bytes of one position component,
# S = bytes of one component of any other per-particle quantity:
#   density (pass A)       read X, write rho                         2Sx + S
#   force + integrate (B)  read X, V, rho; write X', V'              4Sx + 5S
#   adjoint R              read X, write rho (density recompute)     2Sx + S
#   adjoint P1             read X, V, rho, gX', gV', V'; write gXs,  2Sx + 14S
#                          gV, rho-bar
#   adjoint P2             read X, rho-bar, gXs; write gX            2Sx + 5S
# FP64: Sx = S = 8.  The FP32 form: Sx = 8 (FP64 positions), S = 4.
def _alg(sx, s):
    return {"density": 2 * sx + s, "force_integrate": 4 * sx + 5 * s, "adjoint_r": 2 * sx + s,
            "adjoint_p1": 2 * sx + 14 * s, "adjoint_p2": 2 * sx + 5 * s}


ALG_BYTES = {"f64": _alg(8, 8), "f32c": _alg(8, 4)}
# per particle-step of the whole path: forward 6Sx + 6S, forward+adjoint 12Sx + 26S
ALG_STEP_BYTES = {"f64": {"fwd": 96, "fwd_adj": 304}, "f32c": {"fwd": 72, "fwd_adj": 200}}
# The implementation's own per-particle arrays of each kernel (what it reads
# and writes once per slot, besides the neighbour lists), for comparison:
IMPL_BYTES = {
    "f64": {"density": 16 + 16, "force_integrate": 48 + 32, "adjoint_r": 64 + 56,
            "adjoint_p1": 88 + 40, "adjoint_p2": 56 + 16},
    # f32c: X double2; V, xr, PI, abar, gX, gV float2; galpha, grho float; cell int.
    # Its R pass runs inside P2, or at an epoch's last step inside the remap
    # gather (ids, inverse, gathered gX, gV, V'; X; writes gX, gV, xr, abar,
    # galpha, key): the latter's arrays
    "f32c": {"density": 8 + 8, "force_integrate": 44 + 32, "adjoint_r": 48 + 40,
             "adjoint_p1": 64 + 20, "adjoint_p2": 32 + 8},
}
# measured HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by
# a plain copy of known size), keyed "<form>:<kernel>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_block_1000x1000.json")
# VALU instructions per wave of each kernel (rocprofv3 SQ_INSTS_VALU / SQ_WAVES)
VALU_FILE = os.path.join(ROOT, "profiles", "valu_busy_block_1000x1000.json")
SYMBOL_F32C = {"density": "k_density32", "force_integrate": "k_force32",
               "adjoint_r": "k_remap_rpw32<true>", "adjoint_p1": "k_adj_p1_theta32<true>",
               "adjoint_p2": "k_adj_p232<true>"}
SYMBOL = {"f64": {"density": "k_density", "force_integrate": "k_force", "adjoint_r": "k_adj_r",
                  "adjoint_p1": "k_adj_p1", "adjoint_p2": "k_adj_p2"},
          "f32c": SYMBOL_F32C, "f32c_ckpt": SYMBOL_F32C}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpu():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def _oracle_run(O, sc, T, steps, adjoint):
    p = O.SPHParams(T=T, **sc.params)
    st = O.State(sc.X0, sc.V0, p)
    tape = O.Tape(st.N, steps, p) if adjoint else None
    t0 = time.perf_counter()
    O.simulate_(st, sc.theta, p, steps, tape)
    if adjoint:
        J, seed = O.target_objective(st.X, XC, YC, SIGMA, T=T)
        ws = O.AdjointWorkspace(st.N, p)
        O.backward_(ws, tape, sc.theta, p, seedX=seed, seedV=np.zeros_like(seed))
    return time.perf_counter() - t0


def cpu_baseline(sc4, threads, c4_steps, trials):
    """The oracle (C restatement of the reference's CPU path, OpenMP over
    particles) on the host cores: C1 (the 1600-particle dam break, 1000 steps
    forward, Float64) in full, C2 (N = 2072, 200 steps forward + adjoint) in
    full, and a bounded sample of the 1M configuration (its first c4_steps
    steps forward + adjoint) in Float64 (`value`) and Float32; best and median
    of `trials` runs each, A/B alternated."""
    import warnings
    from oracle import oracle as O
    from sphadjoint_amd import scenes
    if threads:
        O.set_threads(threads)
    warnings.simplefilter("ignore")
    c1 = scenes.dambreak_40x40(nsteps=1000)
    c2 = scenes.brinkman_56x37(nsteps=200)
    cases = [("C1 fwd f64", c1, np.float64, 1000, False),
             ("C2 fwd+adj f64", c2, np.float64, 200, True),
             ("C2 fwd+adj f32", c2, np.float32, 200, True),
             ("C4 fwd+adj f64", sc4, np.float64, c4_steps, True),
             ("C4 fwd+adj f32", sc4, np.float32, c4_steps, True)]
    times = {c[0]: [] for c in cases}
    for t in range(trials):
        order = cases if t % 2 == 0 else cases[::-1]
        for name, sc, T, steps, adj in order:
            times[name].append(_oracle_run(O, sc, T, steps, adj))
    out = {}
    for name, sc, T, steps, adj in cases:
        ts = times[name]
        out[name] = {"N": sc.N, "steps": steps, "best_s": round(min(ts), 4),
                     "median_s": round(statistics.median(ts), 4),
                     "median_pss": round(sc.N * steps / statistics.median(ts), 1)}
    v = out["C4 fwd+adj f64"]["median_pss"]
    model, ncpu = host_cpu()
    return {
        "value": v, "unit": "particle-steps/s", "cores": O.get_threads(), "kind": "port",
        "sample": f"oracle (the reference's CPU algorithm restated in C, OpenMP over particles) "
                  f"on {O.get_threads()} threads of '{model}' ({ncpu} logical CPUs): the 1M block's "
                  f"first {c4_steps} steps forward + adjoint, Float64, median of {trials}; "
                  f"C1 and C2 in full beside it (cases)",
        "float32_value": out["C4 fwd+adj f32"]["median_pss"],
        "cases": out,
    }


class Form:
    """One form's objects, allocated once; run() evaluates one horizon."""

    def __init__(self, R, name, H):
        S, be = R.S, R.be
        self.R, self.name, self.H = R, name, H
        self.precision = name.split("_")[0]
        self.ckpt = "_ckpt" in name
        # "_s5": the list builds sweep 5x5 half cells over quadrant-ordered cells
        self.sweep = 5 if name.endswith("_s5") else 3
        self.ws = S.AdjointWorkspace(be, R.N, R.p, precision=self.precision)
        if self.ckpt:
            E = math.ceil(H / R.args.interval)
            seg = R.args.seg_epochs or max(1, math.ceil(E / (R.args.snapshots + 1)))
            self.ck = S.Checkpointer(be, R.N, H, R.p, snapshots=R.args.snapshots, seg_epochs=seg,
                                     interval=R.args.interval, precision=self.precision)
            S.set_sweep_(self.ck, self.sweep)
        else:
            keep = {}
            if name in ("f32c", "f32c_s5"):
                keep = dict(keep_pi=True, keep_lists=True)
            elif name == "f64_keep":
                keep = dict(keep_pi=True)
            self.tape = S.Tape(be, R.N, H, R.p, precision=self.precision, **keep)
        self.st = None

    def run(self, profile=None):
        """One horizon; returns (seconds, J, per-kernel events or None, extra)."""
        R, S, be = self.R, self.R.S, self.R.be
        if not self.ckpt:
            # a fresh state at X0, V0 (allocation and upload outside the timed region)
            self.st = None
            gc.collect()
            self.st = S.State(be, R.sc.X0, R.sc.V0, R.p, interval=R.args.interval,
                              precision=self.precision)
            S.set_sweep_(self.st, self.sweep)
            S.reset_(self.tape)
        be.synchronize()
        if profile:
            S.api.profile_enable(be, True, only=None if profile == "all" else profile)
        t0 = time.perf_counter()
        if self.ckpt:
            st = self.ck.forward(R.X0, R.V0, R.theta, check_flags=False)
            J, seed = S.target_objective(st, XC, YC, SIGMA, device=R.seed)
            self.ck.backward(self.ws, seedX=seed, seedV=R.zero)
        else:
            st = self.st
            L = S._lib
            f = S.api._field(R.theta, R.p, be)
            L.call("sph_simulate", st._h, f._h, self.H, self.tape._h)   # flags read after timing
            J, seed = S.target_objective(st, XC, YC, SIGMA, device=R.seed)
            S.backward_(self.ws, self.tape, R.theta, R.p, be, seedX=seed, seedV=R.zero,
                        interval=R.args.interval)
        be.synchronize()
        dt = time.perf_counter() - t0
        kern = None
        if profile:
            stats = S.api.profile_stats(be)
            S.api.profile_enable(be, False)
            kern = {k: {"launches": n, "avg_us": 1e3 * ms / n, "total_ms": ms}
                    for k, (n, ms) in stats.items() if n}
        extra = dict(self.ck.stats()) if self.ckpt else {}
        return dt, J, kern, extra

    def forward_only(self):
        """H steps forward without a tape (the forward rate of the form)."""
        R, S, be = self.R, self.R.S, self.R.be
        st = S.State(be, R.sc.X0, R.sc.V0, R.p, interval=R.args.interval,
                     precision=self.precision)
        S.set_sweep_(st, self.sweep)
        be.synchronize()
        t0 = time.perf_counter()
        f = S.api._field(R.theta, R.p, be)
        S._lib.call("sph_simulate", st._h, f._h, self.H, None)
        be.synchronize()
        return time.perf_counter() - t0


class Runner:
    def __init__(self, sc, args):
        import sphadjoint_amd as S
        self.S, self.sc, self.args = S, sc, args
        self.be = S.MI355XBackend(0)
        self.p = S.SPHParams(**sc.params)
        self.N = sc.N
        self.theta = S.Field(self.be, sc.theta)
        self.X0 = S.DeviceArray(self.be, sc.X0)
        self.V0 = S.DeviceArray(self.be, sc.V0)
        self.zero = S.DeviceArray(self.be, N=sc.N)
        self.seed = S.DeviceArray(self.be, N=sc.N)


def run_batched(nscenes, steps, seed=0):
    """Batched line search (nscenes scenes, N log-uniform in [1k, 64k],
    theta_k = theta0 + alpha_k d): one f32c state holds every scene.  After
    the autotuner has picked the workgroup sizes (on a scratch state), times
    forward + adjoint over a full tape, and forward-only candidate evaluation
    (J per scene) run eagerly and replayed from a captured HIP graph; best of
    three runs each."""
    import warnings
    import sphadjoint_amd as S
    from sphadjoint_amd import scenes
    warnings.simplefilter("ignore")
    scs = scenes.batch_line_search(seed=seed, nscenes=nscenes, nsteps=steps)
    be = S.MI355XBackend(0)
    args = [(S.SPHParams(**sc.params), sc.X0, sc.V0) for sc in scs]
    scratch = S.BatchState(be, args)
    th = scratch.field([sc.theta for sc in scs])
    sizes = S.autotune_(scratch, th, steps=4)
    del scratch
    Ntot = sum(sc.N for sc in scs)
    out = {"scenes": nscenes, "N_total": int(Ntot), "N_min": int(min(sc.N for sc in scs)),
           "N_max": int(max(sc.N for sc in scs)), "steps": steps, "workgroup_sizes": sizes}

    def fresh():
        st = S.BatchState(be, args)
        S.tuning(st, sizes)
        return st

    tape = S.Tape(be, Ntot, steps, args[0][0], precision="f32c", keep_pi=True, keep_lists=True)
    zero = S.DeviceArray(be, N=Ntot)
    ts = []
    for _ in range(4):
        st = fresh()
        ws = S.BatchWorkspace(be, st)
        S.reset_(tape)
        be.synchronize()
        t0 = time.perf_counter()
        S.simulate_(st, th, st.p, be, steps, tape=tape)
        J, seed_ = S.batch_objective(st, 0.3, 0.2, 0.1, device=True)
        S.backward_(ws, tape, th, st.p, be, seedX=seed_, seedV=zero)
        be.synchronize()
        ts.append(time.perf_counter() - t0)
    dt = min(ts[1:])
    del tape
    out["fwd_adj_s"] = round(dt, 5)
    out["fwd_adj_scenes_per_s"] = round(nscenes / dt, 2)
    out["fwd_adj_particle_steps_per_s"] = round(Ntot * steps / dt, 1)
    for mode in ("eager", "graph"):
        ts = []
        for _ in range(4):
            st = fresh()
            be.synchronize()
            t0 = time.perf_counter()
            if mode == "eager":
                S.simulate_(st, th, st.p, be, steps)
            else:
                out["graph_launches"] = S.simulate_graph_(st, th, st.p, be, steps)
            J, _ = S.batch_objective(st, 0.3, 0.2, 0.1)
            be.synchronize()
            ts.append(time.perf_counter() - t0)
        dt = min(ts[1:])
        out[f"fwd_{mode}_scenes_per_s"] = round(nscenes / dt, 2)
        out[f"fwd_{mode}_particle_steps_per_s"] = round(Ntot * steps / dt, 1)
    return out


def run_long_horizon(R, steps, snapshots, runs=2, tape_frac=0.0):
    """The long horizon at N = 1M: `steps` steps of the f32c form forward +
    adjoint under a `snapshots` budget (one-epoch segments); p-s/s, device
    memory of the checkpointer, recomputed segments, J over repeated runs.
    With tape_frac > 0 the checkpointer's tape may take that share of the free
    device memory (80 B per particle-step and a layout record per epoch): the
    last segments that fit are taped in the forward, not recomputed."""
    S, be = R.S, R.be
    gc.collect()
    be.synchronize()
    free0, total = be.mem_info()
    tape_steps = None
    if tape_frac > 0:
        per_step = 80 * R.N + (12 * R.N + 4 * 300000) // R.args.interval
        tape_steps = max(1, min(steps, int(tape_frac * free0 / per_step)))
    ck = S.Checkpointer(be, R.N, steps, R.p, snapshots=snapshots, seg_epochs=1,
                        interval=R.args.interval, precision="f32c", tape_steps=tape_steps)
    ws = S.AdjointWorkspace(be, R.N, R.p, precision="f32c")
    be.synchronize()
    free1, _ = be.mem_info()
    Js, ts, gts = [], [], []
    for _ in range(runs):
        be.synchronize()
        t0 = time.perf_counter()
        st = ck.forward(R.X0, R.V0, R.theta, check_flags=False)
        J, seed = S.target_objective(st, XC, YC, SIGMA, device=R.seed)
        ck.backward(ws, seedX=seed, seedV=R.zero)
        be.synchronize()
        ts.append(time.perf_counter() - t0)
        Js.append(J)
        gts.append(ws.gtheta)
    stats = ck.stats()
    same = all(j == Js[0] for j in Js) and all(np.array_equal(g, gts[0]) for g in gts)
    out = {"steps": steps, "N": R.N, "snapshots": snapshots, "seg_epochs": 1,
           "tape_steps": tape_steps,
           "fwd_adj_particle_steps_per_s": round(R.N * steps / min(ts), 1),
           "seconds": [round(t, 3) for t in ts],
           "checkpointer_device_bytes": int(free0 - free1),
           "device_total_bytes": int(total),
           "epochs": stats["epochs"], "forward_epochs_run": stats["forward_epochs"],
           "backward_epochs_run": stats["backward_forward_epochs"],
           "recomputed_epochs": stats["forward_epochs"] + stats["backward_forward_epochs"]
           - stats["epochs"], "J": Js[0], "J_gtheta_bit_identical": bool(same)}
    del ck, ws
    gc.collect()
    return out


def dominant(form, kern):
    table = ALG_BYTES[form.split("_")[0]]
    return max((k for k in kern if k in table), key=lambda k: kern[k]["total_ms"])


def _traffic(form, kname, N):
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        rec = tr.get("kernels", {}).get(f"{form}:{kname}")
        if rec is not None and rec.get("N") == N:
            return rec["bytes_per_launch"], os.path.relpath(TRAFFIC_FILE, ROOT)
    except (OSError, ValueError):
        pass
    return None, None


def _valu(form, kname, avg_us, N):
    try:
        with open(VALU_FILE) as f:
            vk = json.load(f)
        rec = vk["kernels"].get(SYMBOL.get(form, {}).get(kname, ""))
        if rec is not None and vk.get("N") == N:
            cyc = rec["valu_per_wave"] * 4 * rec["waves"] / vk["simds"]
            return {"valu_per_wave": rec["valu_per_wave"],
                    "busy": round(cyc / (avg_us * 1e-6 * vk["clock_GHz"] * 1e9), 3),
                    "source": os.path.relpath(VALU_FILE, ROOT)}
    except (OSError, ValueError, KeyError):
        pass
    return None


def roofline_of(form, kern, timed, N):
    """HBM roofline of the dominant kernel of `form` (from the events of the
    dominant-kernel pass), plus the same figures for every pair pass (from the
    all-kernel pass)."""
    prec = form.split("_")[0]
    table = ALG_BYTES[prec]
    dom = dominant(form, kern)
    avg = timed[dom]["avg_us"]
    alg = table[dom] * N
    achieved = alg / (avg * 1e-6) / 1e9
    traffic, source = _traffic(form, dom, N)
    per = {}
    for k in table:
        if k in kern:
            a = table[k] * N / (kern[k]["avg_us"] * 1e-6) / 1e9
            per[k] = {"alg_bytes_per_launch": table[k] * N, "avg_us": round(kern[k]["avg_us"], 3),
                      "achieved_GBs": round(a, 1), "frac": round(a / HBM_PEAK_GBS, 4),
                      "impl_bytes_per_launch": IMPL_BYTES[prec][k] * N,
                      "traffic": _traffic(form, k, N)[0]}
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": dom, "alg_bytes_per_launch": alg,
            "impl_bytes_per_launch": IMPL_BYTES[prec][dom] * N,
            "avg_launch_us": round(avg, 3), "launches": timed[dom]["launches"],
            "traffic_source": source, "valu": _valu(form, dom, avg, N), "pair_passes": per}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed trials K (one horizon of every form each)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed warm-up trials W")
    ap.add_argument("--horizon", type=int, default=500, help="physical steps H of one evaluation")
    ap.add_argument("--n-side", type=int, default=1000, help="particles per side of the block")
    ap.add_argument("--interval", type=int, default=4, help="neighbour-list rebuild interval")
    ap.add_argument("--snapshots", type=int, default=64, help="checkpoint snapshot budget")
    ap.add_argument("--seg-epochs", type=int, default=0,
                    help="epochs per checkpoint segment (0: fewest that fit the budget)")
    ap.add_argument("--forms", default="f64,f64_ckpt,f32c_ckpt,f32c,f64_keep")
    ap.add_argument("--cpu-steps", type=int, default=30, help="steps of the 1M CPU sample")
    ap.add_argument("--cpu-trials", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched leg (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--long-steps", type=int, default=5333,
                    help="steps of the long-horizon leg at N=1M (0: skip)")
    ap.add_argument("--long-tape-frac", type=float, default=0.5,
                    help="share of free device memory the long-horizon checkpointer may tape "
                         "(a second long-horizon run; 0: skip)")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the per-kernel and roofline passes")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4,
                          "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the path's sharding is a "
                                   "design, not code", "n_gpus": args.gpus}))
        sys.exit(2)
    import warnings
    warnings.simplefilter("ignore")      # the block raises the reference's displacement warning
    from sphadjoint_amd import scenes
    H, K, W = args.horizon, args.steps, args.warmup
    sc = scenes.block_1m(nsteps=H, n_side=args.n_side)
    N = sc.N
    log(f"workload {sc.name}: N={N}, horizon H={H}, trials K={K}, warm-up W={W}")
    R = Runner(sc, args)
    names = args.forms.split(",")
    forms = {n: Form(R, n, H) for n in names}
    times = {n: [] for n in names}
    Js, extras = {}, {}
    for t in range(W + K):
        order = names if t % 2 == 0 else names[::-1]
        for n in order:
            dt, J, _, extra = forms[n].run()
            if t >= W:
                times[n].append(dt)
            Js.setdefault(n, J)
            extras[n] = extra
        if t >= W:
            log(f"trial {t - W + 1}/{K}: " + ", ".join(f"{n} {times[n][-1] * 1e3:.1f} ms"
                                                       for n in names))
    kerns = {}
    if not args.no_profile:
        for n in names:
            _, _, kern, _ = forms[n].run(profile="all")
            dom = dominant(n, kern)
            _, _, kdom, _ = forms[n].run(profile=dom)
            kerns[n] = (kern, kdom)
            log(f"{n}: kernels (events on every launch, one horizon)")
            for k, v in kern.items():
                log(f"   {k:16s} {v['launches']:6d} x {v['avg_us']:9.2f} us = {v['total_ms']:9.2f} ms")
    fwd_only = {}
    for n in names:
        if n in ("f64", "f32c"):
            fwd_only[n] = min(forms[n].forward_only() for _ in range(2))
    del forms
    gc.collect()

    best = "f32c_ckpt" if "f32c_ckpt" in names else names[0]
    tb = times[best]
    value = N * H * K / sum(tb)

    def ratios(a, b):
        return [x / y for x, y in zip(times[b], times[a])]     # rate(a) / rate(b)

    speed = {}
    for a in names:
        for b in ("f64", "f64_ckpt", "f64_keep"):
            if a != b and b in names and not a.startswith("f64"):
                r = ratios(a, b)
                speed[f"{a}/{b}"] = {"median": round(statistics.median(r), 3),
                                     "min": round(min(r), 3), "max": round(max(r), 3)}
    for a in names:
        b = a[:-3]
        if a.endswith("_s5") and b in names:     # the half-cell sweep against the 3x3 one
            r = ratios(a, b)
            speed[f"{a}/{b}"] = {"median": round(statistics.median(r), 3),
                                 "min": round(min(r), 3), "max": round(max(r), 3)}
    roofline = None
    if kerns:
        roofline = roofline_of(best, *kerns[best], N)

    long_h = long_t = None
    if args.long_steps > 0 and best == "f32c_ckpt":
        log(f"long horizon: {args.long_steps} steps, f32c, {args.snapshots} snapshots")
        long_h = run_long_horizon(R, args.long_steps, args.snapshots)
        log(f"long horizon: {long_h}")
        if args.long_tape_frac > 0:
            long_t = run_long_horizon(R, args.long_steps, args.snapshots,
                                      tape_frac=args.long_tape_frac)
            log(f"long horizon with a tape budget: {long_t}")
    del R
    gc.collect()

    batched = None
    if args.batch_scenes > 0:
        log(f"batched line search: {args.batch_scenes} scenes, {args.batch_steps} steps")
        batched = run_batched(args.batch_scenes, args.batch_steps)
        log(f"batched: {batched}")

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: C1, C2 and {args.cpu_steps} steps of the 1M block on the oracle")
        cpu = cpu_baseline(sc, args.cpu_threads, args.cpu_steps, args.cpu_trials)
        log(f"cpu baseline {cpu['value']:.4g} p-s/s on {cpu['cores']} threads: {cpu['cases']}")

    ck_desc = ""
    if best.endswith("_ckpt"):
        ck_desc = f", binomial checkpointing under {args.snapshots} snapshots"
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(sum(tb) * 1e3 / K, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32c": "f32 (FP64 positions, J, gtheta)", "f64": "f64"}[best.split("_")[0]],
        "data": "synthetic",
        "config": {"workload": f"block_{args.n_side}x{args.n_side} fluid block, N={N}, dx=1e-3, "
                               f"h=1.2dx, tank 2.0x1.2, theta 512x256 box-filtered; one step = "
                               f"{H} physical steps forward + Gaussian-target adjoint" + ck_desc,
                   "N": N, "horizon_steps": H, "form": best},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": {
            n: {"fwd_adj": round(N * H / statistics.median(times[n]), 1),
                "fwd_adj_best": round(N * H / min(times[n]), 1),
                "ms_median": round(statistics.median(times[n]) * 1e3, 3),
                "ms_best": round(min(times[n]) * 1e3, 3),
                "fwd": round(N * H / fwd_only[n], 1) if n in fwd_only else None,
                "step_alg_GBs": round(ALG_STEP_BYTES[n.split("_")[0]]["fwd_adj"] * N * H
                                      / statistics.median(times[n]) / 1e9, 1),
                "J": Js[n], **extras[n]}
            for n in names},
        "speedup": speed,
        "kernels_us": ({k: round(v["avg_us"], 3) for k, v in kerns[best][0].items()}
                       if kerns else None),
        "long_horizon": long_h,
        "long_horizon_tape_budget": long_t,
        "batched_scenes_per_s": batched,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
