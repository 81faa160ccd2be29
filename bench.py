"""Benchmark: MCWF RK try-steps/s of the batched FP32-derivative ensemble on one MI355X.

Workload: driven Jaynes-Cummings (qubit (x) mode cutoff 30, dim 60: g=3, deltaC=-1,
deltaA=0.5, eta=1, kappa=1, gamma=0.2, two jump channels), 16384 trajectories from
ground (x) vacuum, evolved to T=20 with a stream point every Dt=0.1, epsRel=1e-6,
epsAbs=1e-12 (reference semantics), dpLimit=0.01.  One "step" is one full T=20
evolution of the whole ensemble in one persistent launch (state reset to t=0 before
each step, device-side, outside the timed region is not needed: the reset is a few
device copies and is timed too).

value = RK try-steps (accepted + rejected RKCK tries, summed over trajectories) per
second over the K timed steps.  Beside it: the FP64 batched form, the plain
one-trajectory-per-launch FP64 form (a bounded sample), the roofline of the batched
kernel and the CPU oracle timed on the host (a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

# algorithmic work per RKCK try for JC-60 (Schroedinger picture):
# M = 354 complex multiply-adds per H_nH/i application (stored diagonals of the mode (176),
# qubit (120) and JC (58) terms over their slices), F = 6*8*M + 124*d
JC_M, JC_D = 354, 60
F_DERIV = 6 * 8 * JC_M          # 16992 flop per try in the derivative evaluations
F_CTRL = 124 * JC_D             # 7440 flop per try of RK algebra + error norm (FP64)
P_FP32, P_FP64 = 157.3e12, 78.6e12   # MI355X vector peaks (FLOP/s)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ntraj", type=int, default=16384)
    ap.add_argument("--T", type=float, default=20.0)
    ap.add_argument("--plain-traj", type=int, default=4)
    ap.add_argument("--cpu-traj", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(trajectories would shard by index range, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import cppqed_amd as Q
    from sysdefs import JC60, ground

    sysq = Q.System(JC60)
    psi0 = ground(sysq.dim)
    iv = Q.run_intervals(T=args.T, Dt=0.1)

    def run_batched(precision, steps, warmup):
        ens = Q.Ensemble(sysq, args.ntraj, psi0, Q.mcwf_pars(seed=1001, precision=precision, jump_log_cap=64))
        for _ in range(warmup):
            ens.reset()
            ens.advance(iv)
        ms0, n0 = ens.kernel_ms()
        tries = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            ens.reset()
            ens.advance(iv)
            tries += int(ens.counters()[1][4])
        el = time.perf_counter() - t0
        ms1, n1 = ens.kernel_ms()
        kernel_s = (ms1 - ms0) / 1e3
        _, tot = ens.counters()
        return dict(tries=tries, seconds=el, kernel_s=kernel_s, launches=n1 - n0, last_totals=tot.tolist())

    b32 = run_batched(Q.QT_FP32_DERIV, args.steps, args.warmup)
    b64 = run_batched(Q.QT_FP64, max(1, args.steps // 2), 1)
    value = b32["tries"] / b32["seconds"]
    v64 = b64["tries"] / b64["seconds"]

    plain = None
    if not args.no_plain and args.plain_traj > 0:
        t0 = time.perf_counter()
        ptries = 0
        for k in range(args.plain_traj):
            r = Q.mcwf_plain_run(sysq, psi0, iv, Q.mcwf_pars(seed=1001), traj=k)
            ptries += int(r["counters"][4])
        pel = time.perf_counter() - t0
        plain = dict(tries=ptries, seconds=pel, value=ptries / pel,
                     sample=f"{args.plain_traj} trajectories to T={args.T}")

    # roofline of the batched FP32-derivative kernel: algorithmic flops / kernel time
    f32_flops = F_DERIV * b32["tries"]
    f64_flops = F_CTRL * b32["tries"]
    achieved = (f32_flops + f64_flops) / b32["kernel_s"] / 1e12
    t_min = f32_flops / P_FP32 + f64_flops / P_FP64
    peak = (f32_flops + f64_flops) / t_min / 1e12

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "mcwf_batch_jc60_16k_traffic.json")
    if os.path.exists(tpath) and args.ntraj == 16384 and args.T == 20.0:
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]

    cpu = None
    if not args.no_cpu and args.cpu_traj > 0:
        import oracle_lib as O
        so = O.System(JC60)
        t0 = time.perf_counter()
        _, ctries = O.ensemble_run(so, psi0, O.params(rng="philox", seed=1001, stream=0),
                                   O.run_spec("Dt", Dt=0.1, T=args.T), args.cpu_traj)
        cel = time.perf_counter() - t0
        cpu = {"value": ctries / cel, "unit": "RK try-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_traj} JC-60 trajectories to T={args.T}, FP64, 1 thread ({ctries} tries, {cel:.1f} s)"}

    out = {
        "metric": "MCWF RK steps/s (JC dim 60, 16k traj) batched vs plain; p99 latency",
        "value": value,
        "unit": "RK try-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * b32["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-deriv/fp64-control",
        "data": "synthetic",
        "config": {"workload": f"driven Jaynes-Cummings dim 60, {args.ntraj} MCWF trajectories to T={args.T}, "
                               "Dt=0.1 streams, FP32 derivatives + FP64 control",
                   "n_traj": args.ntraj, "T": args.T, "dim": 60, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "note": "F = 6*8*354 FP32 + 124*60 FP64 flop per RK try; peak = mix of 157.3 FP32 / 78.6 FP64 vector"},
        "batched_fp64": {"value": v64, "seconds": b64["seconds"], "tries": b64["tries"]},
        "plain_fp64": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "batched_fp64_over_plain": (v64 / plain["value"]) if plain else None,
        "kernel": {"launches": b32["launches"], "avg_ms": 1e3 * b32["kernel_s"] / max(1, b32["launches"])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
