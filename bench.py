#!/usr/bin/env python3
"""Forward and forward+adjoint particle-steps/s of the differentiable 2D WCSPH
step on one MI355X.

Workload: 1000x1000 = 1M particle fluid block, dx = 1e-3, h = 1.2 dx, tank
2.0 x 1.2, theta on a 512x256 grid (U(0,1), 5x5 box filter), K physical steps
of simulate_ with a tape followed by backward_ over the same K steps.
particle-steps/s = N * K / wall time (all inputs resident in HBM; the seeds of
the random linear functional J = <cX, X_T> + <cV, V_T> are uploaded by
backward_ as in the reference's API).

One JSON line is printed.  `value` is the forward+adjoint rate of the best
available form; the other form's numbers, the forward-only rate, the per-kernel
roofline of the dominant kernel and the CPU baseline (the oracle, a C
restatement of the reference's CPU algorithm, on a bounded sample) are extra
fields.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+adjoint particle-steps/s, N=1M FP32 vs FP64 plain; batched scenes/s"

# Compulsory own-particle HBM bytes per launch-unit (one particle) of each
# kernel in the FP64 form with a tape: every per-particle array the kernel must
# read or write, once; neighbour-side reads are served from cache.
ALG_BYTES_F64 = {
    "density": 16 + 16 + 16,            # read X; write (pterm, invrho), tape X
    "force_integrate": 48 + 32 + 16,    # read X, V, (pterm, invrho); write X', V', tape V
    "adjoint_r": 64 + 56,               # read X, V', gX, gV; write PI, abar, galpha, gV
    "adjoint_p1": 88 + 40,              # read X, V, PI, abar, galpha, gV; write gXs, grho, gV
    "adjoint_p2": 56 + 16,              # read X, grho, gXs, gX; write gX
}
ALG_BYTES_F32C = {
    "density": 8 + 4 + 8,               # read xr, cell; write (pterm, invrho)
    "force_integrate": 44 + 56,         # read X64, xr, V, PI, cell; write X64', xr', V', tape X, V
    "adjoint_r": 28 + 72,               # xr from tape X64; read X, xr, V', gX, gV, cell; write PI, abar, galpha, gV
    "adjoint_p1": 64 + 20,              # read X, xr, V, PI, abar, galpha, gV, cell; write gXs, grho, gV
    "adjoint_p2": 32 + 8,               # read xr, grho, gXs, gX, cell; write gX
}
ALG_BYTES = {"f64": ALG_BYTES_F64, "f32c": ALG_BYTES_F32C}
# the reference-level minimum per particle-step (survey of the path):
# FP64 state 96 / 304 B; FP64 positions with FP32 rest 72 / 200 B
ALG_STEP_BYTES = {"f64": {"fwd": 96, "fwd_adj": 304}, "f32c": {"fwd": 72, "fwd_adj": 200}}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(sc, steps, threads):
    """The oracle (C restatement of the reference's CPU path, OpenMP over
    particles) on a bounded sample of the same workload."""
    from oracle import oracle as O
    if threads:
        O.set_threads(threads)
    p = O.SPHParams(**sc.params)
    rng = np.random.default_rng(1)
    cX = rng.standard_normal(sc.X0.shape)
    cV = rng.standard_normal(sc.X0.shape)
    st = O.State(sc.X0, sc.V0, p)
    tape = O.Tape(st.N, steps, p)
    t0 = time.perf_counter()
    O.simulate_(st, sc.theta, p, steps, tape)
    t1 = time.perf_counter()
    ws = O.AdjointWorkspace(st.N, p)
    O.backward_(ws, tape, sc.theta, p, seedX=cX, seedV=cV)
    t2 = time.perf_counter()
    N = st.N
    return {
        "value": N * steps / (t2 - t0), "unit": "particle-steps/s", "cores": O.get_threads(),
        "kind": "port",
        "sample": f"oracle FP64 (reference CPU algorithm restated in C, OpenMP), same 1M scene, "
                  f"{steps} steps forward+adjoint; forward alone {N * steps / (t1 - t0):.4g} p-s/s",
        "fwd_value": N * steps / (t1 - t0),
    }


def run_gpu(sc, K, W, precision):
    import sphadjoint_amd as S
    from sphadjoint_amd import api
    be = S.MI355XBackend(0)
    p = S.SPHParams(**sc.params)
    N = sc.N
    theta = S.Field(be, sc.theta)
    rng = np.random.default_rng(1)
    cX = rng.standard_normal(sc.X0.shape)
    cV = rng.standard_normal(sc.X0.shape)
    out = {}

    # forward only
    st = S.State(be, sc.X0, sc.V0, p, precision=precision)
    S.simulate_(st, theta, p, be, W)
    be.synchronize()
    t0 = time.perf_counter()
    S.simulate_(st, theta, p, be, K)
    be.synchronize()
    out["fwd_s"] = time.perf_counter() - t0
    out["fwd_value"] = N * K / out["fwd_s"]
    log(f"{precision} forward: {out['fwd_value']:.4g} p-s/s ({out['fwd_s'] * 1e3 / K:.3f} ms/step)")
    del st

    # forward + adjoint over a full tape
    tape = S.Tape(be, N, max(K, W), p, precision=precision)
    ws = S.AdjointWorkspace(be, N, p, precision=precision)
    st = S.State(be, sc.X0, sc.V0, p, precision=precision)
    S.simulate_(st, theta, p, be, W, tape=tape)
    S.backward_(ws, tape, theta, p, be, seedX=cX, seedV=cV)
    S.reset_(tape)
    del st
    st = S.State(be, sc.X0, sc.V0, p, precision=precision)
    be.synchronize()
    api.profile_enable(be, True)
    t0 = time.perf_counter()
    S.simulate_(st, theta, p, be, K, tape=tape)
    S.backward_(ws, tape, theta, p, be, seedX=cX, seedV=cV)
    be.synchronize()
    out["fwdadj_s"] = time.perf_counter() - t0
    stats = api.profile_stats(be)
    api.profile_enable(be, False)
    out["value"] = N * K / out["fwdadj_s"]
    out["kernels"] = {k: {"launches": n, "avg_us": 1e3 * ms / n, "total_ms": ms}
                      for k, (n, ms) in stats.items() if n}
    log(f"{precision} forward+adjoint: {out['value']:.4g} p-s/s "
        f"({out['fwdadj_s'] * 1e3 / K:.3f} ms/step)")
    for k, v in out["kernels"].items():
        log(f"   {k:16s} {v['launches']:6d} x {v['avg_us']:9.2f} us = {v['total_ms']:9.2f} ms")
    J, _ = S.target_objective(st, 0.5, 0.5, 0.1)
    out["J"] = J
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="timed physical steps K")
    ap.add_argument("--warmup", type=int, default=8, help="untimed warm-up steps W")
    ap.add_argument("--n-side", type=int, default=1000, help="particles per side of the block")
    ap.add_argument("--cpu-steps", type=int, default=32, help="steps of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="f64,f32c", help="comma list of f64,f32c")
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
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in "
                                   "DESIGN.md but not part of it", "n_gpus": args.gpus}))
        sys.exit(2)
    from sphadjoint_amd import scenes
    sc = scenes.block_1m(nsteps=args.steps, n_side=args.n_side)
    N = sc.N
    log(f"workload {sc.name}: N={N}, K={args.steps}, W={args.warmup}")

    forms = {}
    for form in args.forms.split(","):
        forms[form] = run_gpu(sc, args.steps, args.warmup, form)
    best_dtype = "f32c" if "f32c" in forms else "f64"
    best = forms[best_dtype]

    kern = best["kernels"]
    table = ALG_BYTES[best_dtype]
    dom = max((k for k in kern if k in table), key=lambda k: kern[k]["total_ms"])
    alg = table[dom] * N
    achieved = alg / (kern[dom]["avg_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom, "alg_bytes_per_launch": alg,
                "avg_launch_us": round(kern[dom]["avg_us"], 3)}

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: {args.cpu_steps} steps forward+adjoint on the oracle")
        cpu = cpu_baseline(sc, args.cpu_steps, args.cpu_threads)
        log(f"cpu baseline {cpu['value']:.4g} p-s/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC, "value": round(best["value"], 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(best["fwdadj_s"] * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"f32c": "f32 (FP64 positions, J, gtheta)", "f64": "f64"}[best_dtype],
        "data": "synthetic",
        "config": {"workload": f"block_{args.n_side}x{args.n_side} dam-break block, N={N}, "
                               f"h=1.2dx, tank 2.0x1.2, theta 512x256, {args.steps} steps "
                               f"forward+adjoint (full tape)",
                   "N": N, "steps": args.steps, "form": best_dtype},
        "roofline": roofline,
        "cpu_baseline": None if cpu is None else {k: v for k, v in cpu.items() if k != "fwd_value"},
        "forms": {
            f: {"fwd_adj": round(r["value"], 1), "fwd": round(r["fwd_value"], 1),
                "step_alg_GBs": round(ALG_STEP_BYTES[f]["fwd_adj"] * N * args.steps
                                      / r["fwdadj_s"] / 1e9, 1),
                "J": r["J"]}
            for f, r in forms.items()},
        "f32c_over_f64": (round(forms["f32c"]["value"] / forms["f64"]["value"], 3)
                          if "f32c" in forms and "f64" in forms else None),
        "kernels_us": {k: round(v["avg_us"], 3) for k, v in best["kernels"].items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
