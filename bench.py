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
    ap.add_argument("--no-master", action="store_true")
    ap.add_argument("--master-points", type=int, default=512)
    ap.add_argument("--master-T", type=int, default=60)
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

    master = None
    if not args.no_master:
        from sysdefs import jc_master_spec
        etas = 0.5 + np.arange(args.master_points) * 14.5 / max(1, args.master_points - 1)
        msys = [Q.System(jc_master_spec(40, eta=float(e), g=10.0)) for e in etas]
        rho0 = np.zeros((80, 80), complex)
        rho0[0, 0] = 1
        sweep = Q.MasterSweep(msys, rho0, epsRel=1e-12, epsAbs=1e-12, ss_tol=1e-8)
        t0 = time.perf_counter()
        chunks = 0
        while chunks < args.master_T // 5:
            _, dn = sweep.advance(np.full(5, 1.0))
            chunks += 1
            if sweep.status()[0].all():
                break
        mel = time.perf_counter() - t0
        done, tss, tm, mcnt = sweep.status()
        mtries = int(mcnt[:, 4].sum())
        mms, _ = sweep.kernel_ms()
        d = 80
        f_gemm = 6 * 8 * d ** 3                  # dense complex H rho per RHS, on FP64 MFMA
        f_try = 6 * (8 * d ** 3 + 2 * 8 * d * d + 4 * d * d) + 124 * d * d
        m_ach = f_gemm * mtries / (mms / 1e3) / 1e12
        master = {"workload": f"{args.master_points} driven Jaynes-Cummings points, mode cutoff 40 (d=80), g=10, kappa=1, "
                              "eta in [0.5, 15], epsRel=epsAbs=1e-12, stop at ||drho/dt||_F < 1e-8",
                  "tries": mtries, "seconds": mel, "tries_per_s": mtries / mel,
                  "converged": int(done.sum()), "t_end": float(tm.max()),
                  "residual_median": float(np.median(dn[:, -1])), "residual_max": float(dn[:, -1].max()),
                  "t_steady_median": float(np.median(tss[tss > 0])) if (tss > 0).any() else None,
                  "roofline": {"bound": "mfma", "achieved": m_ach, "peak": 78.6, "unit": "TFLOP/s", "frac": m_ach / 78.6,
                               "traffic": None,
                               "note": "8 d^3 flop of dense complex (H_nH/i) rho per derivative, 6 per try; "
                                       f"{f_try / 1e6:.1f} MFLOP per try in all, sparse-equivalent 1.82 MFLOP"}}

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
        "master_sweep": master,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
