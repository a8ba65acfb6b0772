#!/usr/bin/env python3
"""Forward+adjoint particle-steps/s of the differentiable 2D WCSPH step on one MI355X.

Workload: 1000x1000 = 1M particle fluid block, dx = 1e-3, h = 1.2 dx, tank
2.0 x 1.2, theta on a 512x256 grid (U(0,1), 5x5 box filter), K physical steps
forward, the Gaussian target objective J on the device, then the discrete
adjoint back over the K steps (dJ/dX0, dJ/dV0, dJ/dtheta).
particle-steps/s = N * K / wall time; recomputation under checkpointing is not
credited.  Inputs (X0, V0, theta) are resident in HBM before the timed region
and the adjoint seeds stay on the device.

Forms, each timed over the same K steps:
  f64        the plain FP64 path (reference arithmetic, full tape)
  f32c_ckpt  FP32 state with compensated sums, FP64 positions / J / gtheta,
             binomial checkpointing under a snapshot budget  -> `value`
  f32c       the same arithmetic over a full tape that also keeps each step's
             (pterm, invrho) and exact neighbour masks (80 B per particle-step)
One JSON line is printed.  Each form runs a warm-up pass, a pass with HIP
events around every kernel (the per-kernel table, and the choice of the
dominant kernel), and the timed pass, in which events bracket only the
dominant kernel's launches on the library's stream: `value` comes from the
timed pass's wall clock and the roofline from its events.
"""
import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+adjoint particle-steps/s, N=1M FP32 vs FP64 plain; batched scenes/s"
XC, YC, SIGMA = 0.5, 0.5, 0.1          # Gaussian target inside the block

# Algorithmic HBM bytes per particle (one launch processes N particles): every
# own-particle array the kernel's function must read or write, once, in the
# form's precision; neighbour-side reads count as served from cache.
# FP64 form: X, V, PI=(pterm, invrho) double2, abar, gX, gV double2, galpha, grho double.
ALG_BYTES_F64 = {
    "density": 16 + 16,                 # read X; write (pterm, invrho)
    "force_integrate": 48 + 32,         # read X, V, PI; write X', V'
    "adjoint_r": 64 + 56,               # read X, V', gX, gV; write PI, abar, galpha, gV
    "adjoint_p1": 88 + 40,              # read X, V, PI, abar, galpha, gV; write gXs, grho, gV
    "adjoint_p2": 56 + 16,              # read X, grho, gXs, gX; write gX
}
# f32c form: X double2; V, xr (cell-relative position), PI, abar, gX, gV float2;
# galpha, grho float; 4-byte binning cell per slot.
ALG_BYTES_F32C = {
    "density": 8 + 8,                   # read xr; write PI
    "force_integrate": 44 + 32,         # read X, xr, V, PI, cell; write X', xr', V'
    "adjoint_r": 44 + 28,               # read X, cell, V', gX, gV; write xr, abar, galpha, gV
    "adjoint_p1": 64 + 20,              # read X, xr, V, PI, abar, galpha, gV, cell; write gXs, grho, gV
    "adjoint_p2": 32 + 8,               # read xr, grho, gXs, gX, cell; write gX
}
ALG_BYTES = {"f64": ALG_BYTES_F64, "f32c": ALG_BYTES_F32C}
# minimum HBM bytes per particle-step of the whole path (forward / forward+adjoint)
ALG_STEP_BYTES = {"f64": {"fwd": 96, "fwd_adj": 304}, "f32c": {"fwd": 72, "fwd_adj": 200}}
# measured HBM traffic per launch of a kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE,
# scaled by a plain copy of known size), keyed "<form>:<kernel>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_block_1000x1000.json")
# VALU instructions per wave of each kernel (rocprofv3 SQ_INSTS_VALU / SQ_WAVES);
# with the live launch time they give the fraction of SIMD issue cycles spent
# on VALU work (a wave64 instruction holds a 16-lane SIMD for 4 cycles)
VALU_FILE = os.path.join(ROOT, "profiles", "valu_busy_block_1000x1000.json")
SYMBOL = {"f64": {"density": "k_density", "force_integrate": "k_force", "adjoint_r": "k_adj_r",
                  "adjoint_p1": "k_adj_p1", "adjoint_p2": "k_adj_p2"},
          # the f32c tapes keep (pterm, invrho) and the masks: most P2 launches
          # also run the next step's pointwise R pass, and P1 launches carry
          # the step's gtheta gather
          "f32c": {"density": "k_density32", "force_integrate": "k_force32",
                   "adjoint_r": "k_adj_rpw32", "adjoint_p1": "k_adj_p1_theta32<true>",
                   "adjoint_p2": "k_adj_p232<true>"},
          "f32c_ckpt": {"density": "k_density32", "force_integrate": "k_force32",
                        "adjoint_r": "k_adj_rpw32", "adjoint_p1": "k_adj_p1_theta32<true>",
                        "adjoint_p2": "k_adj_p232<true>"}}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(sc, steps, threads):
    """The oracle (C restatement of the reference's CPU path, OpenMP over
    particles) on a bounded sample of the same workload: `steps` steps forward
    with a tape, then the adjoint over them, in Float64 (`value`) and Float32."""
    from oracle import oracle as O
    if threads:
        O.set_threads(threads)
    rates = {}
    for T in (np.float64, np.float32):
        p = O.SPHParams(T=T, **sc.params)
        st = O.State(sc.X0, sc.V0, p)
        tape = O.Tape(st.N, steps, p)
        t0 = time.perf_counter()
        O.simulate_(st, sc.theta, p, steps, tape)
        t1 = time.perf_counter()
        J, seed = O.target_objective(st.X, XC, YC, SIGMA, T=T)
        ws = O.AdjointWorkspace(st.N, p)
        O.backward_(ws, tape, sc.theta, p, seedX=seed, seedV=np.zeros_like(seed))
        t2 = time.perf_counter()
        rates[T] = (st.N * steps / (t2 - t0), st.N * steps / (t1 - t0), t2 - t0)
    v64, f64, s64 = rates[np.float64]
    v32, f32, s32 = rates[np.float32]
    return {
        "value": v64, "unit": "particle-steps/s", "cores": O.get_threads(), "kind": "port",
        "sample": f"oracle (the reference's CPU algorithm restated in C, OpenMP over particles), "
                  f"the same 1M scene, first {steps} steps forward + adjoint: Float64 "
                  f"{v64:.4g} p-s/s ({s64:.1f} s; forward alone {f64:.4g}), Float32 {v32:.4g} "
                  f"p-s/s ({s32:.1f} s; forward alone {f32:.4g})",
        "float32_value": v32,
    }


class Runner:
    def __init__(self, sc, args):
        import sphadjoint_amd as S
        self.S, self.sc, self.args = S, sc, args
        self.be = S.MI355XBackend(0)
        self.p = S.SPHParams(**sc.params)
        self.N = sc.N
        # inputs resident in HBM before any timed region
        self.theta = S.Field(self.be, sc.theta)
        self.X0 = S.DeviceArray(self.be, sc.X0)
        self.V0 = S.DeviceArray(self.be, sc.V0)
        self.zero = S.DeviceArray(self.be, N=sc.N)
        self.seed = S.DeviceArray(self.be, N=sc.N)

    def seg_epochs(self, K):
        a = self.args
        if a.seg_epochs > 0:
            return a.seg_epochs
        E = math.ceil(K / a.interval)
        return max(1, math.ceil(E / (a.snapshots + 1)))

    def forward_only(self, precision, K, W):
        S = self.S
        gc.collect()
        st = S.State(self.be, self.sc.X0, self.sc.V0, self.p, precision=precision)
        S.simulate_(st, self.theta, self.p, self.be, W)
        self.be.synchronize()
        t0 = time.perf_counter()
        S.simulate_(st, self.theta, self.p, self.be, K)
        self.be.synchronize()
        return time.perf_counter() - t0

    def fwd_adj(self, form, K, profile):
        """One forward+adjoint pass over K steps; returns (seconds, J, kernels, extra).
        profile: None (no events), "all", or the one kernel name to time."""
        S, be = self.S, self.be
        precision = form.split("_")[0]
        gc.collect()   # release earlier passes' device memory outside the timed region
        ws = S.AdjointWorkspace(be, self.N, self.p, precision=precision)
        if form.endswith("_ckpt"):
            ck = S.Checkpointer(be, self.N, K, self.p, snapshots=self.args.snapshots,
                                seg_epochs=self.seg_epochs(K), interval=self.args.interval,
                                precision=precision)
            be.synchronize()
            if profile:
                S.api.profile_enable(be, True, only=None if profile == "all" else profile)
            t0 = time.perf_counter()
            st = ck.forward(self.X0, self.V0, self.theta, check_flags=False)
            t1 = time.perf_counter()
            J, seed = S.target_objective(st, XC, YC, SIGMA, device=self.seed)
            ck.backward(ws, seedX=seed, seedV=self.zero)
            be.synchronize()
            t2 = time.perf_counter()
            extra = dict(ck.stats(), fwd_s=t1 - t0, seg_epochs=ck.seg_epochs,
                         snapshots_budget=self.args.snapshots)
        else:
            st = S.State(be, self.sc.X0, self.sc.V0, self.p, interval=self.args.interval,
                         precision=precision)
            # the chosen form's tape keeps (pterm, invrho) and the exact masks;
            # the plain FP64 form keeps X and V, as the reference's tape does
            keep = dict(keep_pi=True, keep_lists=True) if precision == "f32c" else {}
            tape = S.Tape(be, self.N, K, self.p, precision=precision, **keep)
            be.synchronize()
            if profile:
                S.api.profile_enable(be, True, only=None if profile == "all" else profile)
            t0 = time.perf_counter()
            S.simulate_(st, self.theta, self.p, be, K, tape=tape)
            t1 = time.perf_counter()
            J, seed = S.target_objective(st, XC, YC, SIGMA, device=self.seed)
            S.backward_(ws, tape, self.theta, self.p, be, seedX=seed, seedV=self.zero,
                        interval=self.args.interval)
            be.synchronize()
            t2 = time.perf_counter()
            extra = dict(fwd_s=t1 - t0)
        kern = None
        if profile:
            stats = S.api.profile_stats(be)
            S.api.profile_enable(be, False)
            kern = {k: {"launches": n, "avg_us": 1e3 * ms / n, "total_ms": ms}
                    for k, (n, ms) in stats.items() if n}
        return t2 - t0, J, kern, extra


def run_batched(nscenes, steps, seed=0):
    """Batched line search (nscenes scenes, N log-uniform in [1k, 64k],
    theta_k = theta0 + alpha_k d): one f32c state holds every scene.  After
    the autotuner has picked the workgroup sizes (on a scratch state), times
    forward + adjoint over a full tape, and forward-only candidate evaluation
    (J per scene) run eagerly and replayed from a captured HIP graph."""
    import sphadjoint_amd as S
    from sphadjoint_amd import scenes
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

    # forward + adjoint over a full tape (a warm-up pass, then the timed one)
    for _ in range(2):
        st = fresh()
        gc.collect()
        tape = S.Tape(be, Ntot, steps, st.p, precision="f32c", keep_pi=True, keep_lists=True)
        ws = S.BatchWorkspace(be, st)
        zero = S.DeviceArray(be, N=Ntot)
        be.synchronize()
        t0 = time.perf_counter()
        S.simulate_(st, th, st.p, be, steps, tape=tape)
        J, seed_ = S.batch_objective(st, 0.3, 0.2, 0.1, device=True)
        S.backward_(ws, tape, th, st.p, be, seedX=seed_, seedV=zero)
        be.synchronize()
        dt = time.perf_counter() - t0
        del tape, ws
    out["fwd_adj_s"] = round(dt, 5)
    out["fwd_adj_scenes_per_s"] = round(nscenes / dt, 2)
    out["fwd_adj_particle_steps_per_s"] = round(Ntot * steps / dt, 1)
    # forward-only candidate evaluation: eager steps, then graph replays
    for mode in ("eager", "graph"):
        for _ in range(2):
            st = fresh()
            be.synchronize()
            t0 = time.perf_counter()
            if mode == "eager":
                S.simulate_(st, th, st.p, be, steps)
            else:
                out["graph_launches"] = S.simulate_graph_(st, th, st.p, be, steps)
            J, _ = S.batch_objective(st, 0.3, 0.2, 0.1)
            be.synchronize()
            dt = time.perf_counter() - t0
        out[f"fwd_{mode}_scenes_per_s"] = round(nscenes / dt, 2)
        out[f"fwd_{mode}_particle_steps_per_s"] = round(Ntot * steps / dt, 1)
    return out


def dominant(form, kern):
    table = ALG_BYTES[form.split("_")[0]]
    return max((k for k in kern if k in table), key=lambda k: kern[k]["total_ms"])


def roofline_of(form, kern, N):
    table = ALG_BYTES[form.split("_")[0]]
    dom = dominant(form, kern)
    alg = table[dom] * N
    achieved = alg / (kern[dom]["avg_us"] * 1e-6) / 1e9
    traffic, source = None, None
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        rec = tr.get("kernels", {}).get(f"{form}:{dom}")
        if rec is not None and rec.get("N") == N:
            traffic = rec["bytes_per_launch"]
            source = os.path.relpath(TRAFFIC_FILE, ROOT)
    except (OSError, ValueError):
        pass
    valu = None
    try:
        with open(VALU_FILE) as f:
            vk = json.load(f)
        rec = vk["kernels"].get(SYMBOL.get(form, {}).get(dom, ""))
        if rec is not None and vk.get("N") == N:
            cyc = rec["valu_per_wave"] * 4 * rec["waves"] / vk["simds"]
            valu = {"valu_per_wave": rec["valu_per_wave"],
                    "busy": round(cyc / (kern[dom]["avg_us"] * 1e-6 * vk["clock_GHz"] * 1e9), 3),
                    "source": os.path.relpath(VALU_FILE, ROOT)}
    except (OSError, ValueError, KeyError):
        pass
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": dom, "alg_bytes_per_launch": alg, "avg_launch_us": round(kern[dom]["avg_us"], 3),
            "traffic_source": source, "valu": valu}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="timed physical steps K")
    ap.add_argument("--warmup", type=int, default=16, help="steps of the untimed warm-up pass")
    ap.add_argument("--n-side", type=int, default=1000, help="particles per side of the block")
    ap.add_argument("--interval", type=int, default=4, help="neighbour-list rebuild interval")
    ap.add_argument("--snapshots", type=int, default=64, help="checkpoint snapshot budget")
    ap.add_argument("--seg-epochs", type=int, default=0,
                    help="epochs per checkpoint segment (0: fewest that fit the budget)")
    ap.add_argument("--cpu-steps", type=int, default=100, help="steps of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="f64,f32c_ckpt,f32c")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched leg (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=300)
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
    from sphadjoint_amd import scenes
    sc = scenes.block_1m(nsteps=args.steps, n_side=args.n_side)
    N, K, W = sc.N, args.steps, args.warmup
    log(f"workload {sc.name}: N={N}, K={K}, W={W}")
    R = Runner(sc, args)

    forms = {}
    for form in args.forms.split(","):
        R.fwd_adj(form, W, profile=None)                        # warm-up pass
        _, _, kern, _ = R.fwd_adj(form, K, profile="all")       # per-kernel table
        dom = dominant(form, kern)
        secs, J, kdom, extra = R.fwd_adj(form, K, profile=dom)  # the timed pass
        r = {"s": secs, "value": N * K / secs, "J": J, "kernels": kern, "timed": kdom, **extra}
        plain = form.split("_")[0]
        if not form.endswith("_ckpt"):
            r["fwd_only_s"] = R.forward_only(plain, K, W)
        forms[form] = r
        log(f"{form}: fwd+adj {r['value']:.4g} p-s/s ({secs * 1e3 / K:.3f} ms/step), "
            f"J={J:.17g}" + (f", ckpt {extra}" if form.endswith("_ckpt") else ""))
        for k, v in kern.items():
            log(f"   {k:16s} {v['launches']:6d} x {v['avg_us']:9.2f} us = {v['total_ms']:9.2f} ms")

    best = "f32c_ckpt" if "f32c_ckpt" in forms else next(iter(forms))
    b = forms[best]
    roofline = roofline_of(best, b["timed"], N)

    batched = None
    if args.batch_scenes > 0:
        log(f"batched line search: {args.batch_scenes} scenes, {args.batch_steps} steps")
        batched = run_batched(args.batch_scenes, args.batch_steps)
        log(f"batched: {batched}")

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: {args.cpu_steps} steps forward+adjoint on the oracle")
        cpu = cpu_baseline(sc, args.cpu_steps, args.cpu_threads)
        log(f"cpu baseline {cpu['value']:.4g} p-s/s on {cpu['cores']} threads")

    ck_desc = ""
    if best.endswith("_ckpt"):
        ck_desc = (f", binomial checkpointing: {b['snapshots_budget']} snapshots, "
                   f"{b['seg_epochs']}-epoch segments")
    line = {
        "metric": METRIC, "value": round(b["value"], 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(b["s"] * 1e3 / K, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32c": "f32 (FP64 positions, J, gtheta)", "f64": "f64"}[best.split("_")[0]],
        "data": "synthetic",
        "config": {"workload": f"block_{args.n_side}x{args.n_side} fluid block, N={N}, dx=1e-3, "
                               f"h=1.2dx, tank 2.0x1.2, theta 512x256 box-filtered, {K} steps "
                               f"forward + Gaussian-target adjoint" + ck_desc,
                   "N": N, "steps": K, "form": best},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": {
            f: {"fwd_adj": round(r["value"], 1),
                "fwd": round(N * K / r["fwd_only_s"], 1) if "fwd_only_s" in r else None,
                "ms_per_step": round(r["s"] * 1e3 / K, 4),
                "step_alg_GBs": round(ALG_STEP_BYTES[f.split("_")[0]]["fwd_adj"] * N * K
                                      / r["s"] / 1e9, 1),
                "J": r["J"],
                **({k: r[k] for k in ("epochs", "forward_epochs", "backward_forward_epochs",
                                      "seg_epochs")} if f.endswith("_ckpt") else {})}
            for f, r in forms.items()},
        "speedup_vs_f64": ({f: round(r["value"] / forms["f64"]["value"], 3)
                            for f, r in forms.items() if f != "f64"}
                           if "f64" in forms else None),
        "kernels_us": {k: round(v["avg_us"], 3) for k, v in b["kernels"].items()},
        "batched_scenes_per_s": batched,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
