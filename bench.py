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

# Algorithmic work per RKCK try: F = 6*8*M + 124*d flop, M = complex multiply-adds of one H_nH/i
# application counted the reference's way (stored entries of every term times its slices, the
# all-zero Kerr diagonal excluded; SysInfo.m_terms).  JC-60: mode 88*2 + qubit 2*30 + JC 58 = 294,
# so 14112 flop in the derivative evaluations (FP32 in the FP32-derivative run) + 7440 FP64.
P_FP32, P_FP64 = 157.3e12, 78.6e12   # MI355X vector peaks (FLOP/s)


def try_flops(system):
    """(derivative flops, control flops) of one RKCK try of `system`."""
    return 6 * 8 * system.info.m_terms, 124 * system.dim


def valu_roofline(f32_flops, f64_flops, seconds):
    """Achieved vs the time-weighted FP32/FP64 vector peak for this flop mix (TFLOP/s)."""
    achieved = (f32_flops + f64_flops) / seconds / 1e12
    peak = (f32_flops + f64_flops) / (f32_flops / P_FP32 + f64_flops / P_FP64) / 1e12
    return achieved, peak


def copy_calibration():
    """A plain device-to-device copy of known size (3 x 256 MiB hipMemcpy): rocprofv3's FETCH_SIZE
    and WRITE_SIZE of these dispatches give the factors that turn the counters into bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("copy calibration: 3 device-to-device copies of %d bytes" % n, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ntraj", type=int, default=16384)
    ap.add_argument("--T", type=float, default=20.0)
    ap.add_argument("--Dt", type=float, default=0.1,
                    help="stream spacing of the JC-60 ensemble (counter runs at other spacings separate the "
                         "cost of a stream point from that of an RK try)")
    ap.add_argument("--plain-traj", type=int, default=4)
    ap.add_argument("--cpu-traj", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-master", action="store_true")
    ap.add_argument("--master-points", type=int, default=512)
    ap.add_argument("--master-T", type=int, default=60)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-T", type=float, default=5.0)
    ap.add_argument("--mixed-layout", default=None,
                    help="variant[,team,...] of the mixed batch instead of the autotuner's choice")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--serve-chunk", type=int, default=20,
                    help="stream intervals a request advances per loop step (20: a whole T=2 request at Dt=0.1)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the copy of known size that calibrates FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; --gpus %d is not supported "
              "(trajectories would shard by index range, one process per GPU; not built here)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import cppqed_amd as Q
    from sysdefs import JC60, ground

    sysq = Q.System(JC60)
    psi0 = ground(sysq.dim)
    iv = Q.run_intervals(T=args.T, Dt=args.Dt)

    def run_batched(precision, steps, warmup):
        ens = Q.Ensemble(sysq, args.ntraj, psi0, Q.mcwf_pars(seed=1001, precision=precision, jump_log_cap=64))
        ens.set_graph(True)  # every step replays the captured launch sequence
        per_step = None
        for _ in range(max(warmup, 1)):
            ens.reset()
            ens.advance(iv)
            per_step = int(ens.counters()[1][4])
        ms0, n0 = ens.kernel_ms()
        t0 = time.perf_counter()
        for _ in range(steps):
            ens.reset()
            ens.advance(iv)
        el = time.perf_counter() - t0
        ms1, n1 = ens.kernel_ms()
        kernel_s = (ms1 - ms0) / 1e3
        _, tot = ens.counters()
        # every step starts from the same reset state and RNG ordinals, so it makes the same
        # tries; the counters are read outside the timed region and checked against a warm-up
        if int(tot[4]) != per_step:
            raise RuntimeError("a step's try count differs from the warm-up's: %d vs %d" % (tot[4], per_step))
        tries = per_step * steps
        captures, replays = ens.set_graph(None)
        return dict(tries=tries, seconds=el, kernel_s=kernel_s, launches=n1 - n0, last_totals=tot.tolist(),
                    graph={"captures": captures, "replays": replays})

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
    fd, fc = try_flops(sysq)
    achieved, peak = valu_roofline(fd * b32["tries"], fc * b32["tries"], b32["kernel_s"])

    traffic = None
    issue = None
    tpath = os.path.join(ROOT, "profiles", "mcwf_batch_jc60_16k_traffic.json")
    if os.path.exists(tpath) and args.ntraj == 16384 and args.T == 20.0 and args.Dt == 0.1:
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj["traffic_bytes_per_launch"]
        # VALU issue: the counted instructions per try x 4 cycles per wave64 instruction over the
        # SIMDs' cycles in the measured time (256 CUs x 4 SIMDs at 2.4 GHz)
        if "sq_insts_valu_per_try" in tj:
            issue = tj["sq_insts_valu_per_try"] * b32["tries"] * 4 / (b32["kernel_s"] * 2.4e9 * 1024)

    cpu = None if args.no_cpu or args.cpu_traj <= 0 else cpu_baseline(JC60, psi0, args)

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
        # the MFMA kernel multiplies only the non-zero 16 x 4 fragments of H_nH/i: four real
        # 16x16x4 MFMAs (8192 flop) per such fragment and 16-column tile, six derivatives per try
        nb = (d + 15) // 16
        hp = np.zeros((16 * nb, 16 * nb), complex)
        hp[:d, :d] = msys[0].dense()
        nzf = sum(bool(np.any(hp[16 * r:16 * r + 16, 4 * k:4 * k + 4] != 0)) for r in range(nb) for k in range(4 * nb))
        f_gemm = 6 * nzf * nb * 8192
        f_dense = 6 * 8 * d ** 3                 # the dense complex H rho per derivative
        f_try = f_gemm + 6 * (2 * 8 * d * d + 4 * d * d) + 124 * d * d
        m_ach = f_gemm * mtries / (mms / 1e3) / 1e12
        m_dense = f_dense * mtries / (mms / 1e3) / 1e12
        # the sparse-equivalent work of the reference's column-by-column application: 6 derivatives
        # of 8 flop per complex multiply-add, M per column, d columns
        f_sparse = 6 * 8 * msys[0].info.m_terms * d
        m_sparse = f_sparse * mtries / (mms / 1e3) / 1e12
        master = {"workload": f"{args.master_points} driven Jaynes-Cummings points, mode cutoff 40 (d=80), g=10, kappa=1, "
                              "eta in [0.5, 15], epsRel=epsAbs=1e-12, stop at ||drho/dt||_F < 1e-8",
                  "tries": mtries, "seconds": mel, "tries_per_s": mtries / mel,
                  "converged": int(done.sum()), "t_end": float(tm.max()),
                  "residual_median": float(np.median(dn[:, -1])), "residual_max": float(dn[:, -1].max()),
                  "t_steady_median": float(np.median(tss[tss > 0])) if (tss > 0).any() else None,
                  "roofline": {"bound": "mfma", "achieved": m_ach, "peak": 78.6, "unit": "TFLOP/s", "frac": m_ach / 78.6,
                               "traffic": master_traffic(args, mtries, chunks),
                               "dense_equivalent_achieved": m_dense,
                               "sparse_equivalent_achieved": m_sparse,
                               "note": f"executed FP64 MFMA flop: {nzf} of {nb * 4 * nb} 16x4 fragments of H_nH/i are "
                                       f"non-zero, 8192 flop per fragment and column tile, 6 derivatives per try = "
                                       f"{f_gemm / 1e6:.2f} MFLOP per try ({f_try / 1e6:.1f} in all); "
                                       f"dense_equivalent_achieved counts 8 d^3 per derivative, sparse_equivalent_achieved "
                                       f"6*8*M*d = {f_sparse / 1e6:.2f} MFLOP per try (M = {msys[0].info.m_terms})"}}

    mixed = None if args.no_mixed else mixed_batch(Q, args)
    serving = None if args.no_serve else serve(Q, args)

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
                               f"Dt={args.Dt} streams, FP32 derivatives + FP64 control",
                   "n_traj": args.ntraj, "T": args.T, "dim": 60, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic, "valu_issue_frac": issue,
                     "note": f"F = 6*8*{sysq.info.m_terms} FP32 + 124*60 FP64 flop per RK try; "
                             "peak = time-weighted mix of 157.3 FP32 / 78.6 FP64 vector; valu_issue_frac = "
                             "counted VALU instructions x 4 cycles / SIMD cycles of the measured time"},
        "batched_fp64": {"value": v64, "seconds": b64["seconds"], "tries": b64["tries"]},
        "plain_fp64": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "batched_fp64_over_plain": (v64 / plain["value"]) if plain else None,
        "kernel": {"launches": b32["launches"], "avg_ms": 1e3 * b32["kernel_s"] / max(1, b32["launches"]),
                   "graph": b32["graph"],
                   # one step's totals: MCWF steps, RK tries, stream points (n_traj x intervals)
                   "per_step": {"mcwf_steps": int(b32["last_totals"][0]),
                                "tries": b32["tries"] // max(1, args.steps),
                                "stream_points": args.ntraj * len(iv)},
                   "note": "each step replays one captured hipGraph (jump draws, queue reset, batched kernel); "
                           "avg_ms is the graph's duration from HIP events on its stream"},
        "cpu_baseline": cpu,
        "master_sweep": master,
        "mixed_batch": mixed,
        "serving": serving,
    }
    print(json.dumps(out))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, psi0, args):
    """The oracle's ensemble (FP64, the reference's step logic, -O3) on the host: one thread over
    args.cpu_traj trajectories (the reference's serial Ensemble::advance), then one thread per
    host core over disjoint trajectory ranges (cores available to this process, at most 16, the
    CPU share of one GPU on the box).  ctypes releases the GIL inside the oracle's calls."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle_lib as O
    so = O.System(spec)
    run = O.run_spec("Dt", Dt=args.Dt, T=args.T)

    def chunk(first, n):
        return O.ensemble_run(so, psi0, O.params(rng="philox", seed=1001, stream=first), run, n)[1]

    t0 = time.perf_counter()
    t1tries = chunk(0, args.cpu_traj)
    t1 = time.perf_counter() - t0
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    per = max(1, args.cpu_traj // 2)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        tn_tries = sum(ex.map(lambda i: chunk(i * per, per), range(cores)))
    tn = time.perf_counter() - t0
    return {"value": tn_tries / tn, "unit": "RK try-steps/s", "cores": cores, "kind": "port",
            "sample": f"{cores} threads x {per} JC-60 trajectories to T={args.T}, FP64 ({tn_tries} tries, {tn:.1f} s)",
            "cpu_model": cpu_model(),
            "single_core": {"value": t1tries / t1, "cores": 1,
                            "sample": f"{args.cpu_traj} JC-60 trajectories to T={args.T}, FP64, 1 thread "
                                      f"({t1tries} tries, {t1:.1f} s)"},
            "note": "reported baseline only: the oracle (a C++ restatement of the reference's MCWF step, "
                    "compiled -O3), not the upstream binary, which cannot be built here"}


def master_traffic(args, tries, launches):
    """HBM bytes per Master launch from the committed FETCH_SIZE / WRITE_SIZE runs of this workload
    (bytes per RKCK try x this run's tries per launch), or None for another workload."""
    path = os.path.join(ROOT, "profiles", "master_sweep_jc80_512_traffic.json")
    if args.master_points != 512 or not os.path.exists(path):
        return None
    with open(path) as f:
        per_try = json.load(f)["traffic_bytes_per_try"]
    return per_try * tries / max(1, launches)


def mixed_batch(Q, args):
    """Uneven batch: 8192 trajectories drawn uniformly from four builds (dims 1024, 4096, 64, 60),
    FP32 derivatives, one batched launch (two persistent kernels on two streams: work-group teams,
    and quarter plus wavefront teams), layout chosen by the autotuner."""
    from sysdefs import MIXED, basis, mixed_counts
    counts = mixed_counts()
    iv = Q.run_intervals(T=args.mixed_T, Dt=0.1)
    systems = [Q.System(spec) for _, spec, _ in MIXED]

    def members():
        return [Q.Ensemble(s, int(n), basis(s.dim, idx), Q.mcwf_pars(seed=1000 + i, precision=Q.QT_FP32_DERIV))
                for i, (s, n, (_, _, idx)) in enumerate(zip(systems, counts, MIXED))]

    mem = members()
    t0 = time.perf_counter()
    if args.mixed_layout is None:
        b = Q.Batch(mem)
        variant, probe_ms = b.autotune(probe_span=0.3, probe_traj=2048)
    else:
        # a fixed layout (e.g. for counter runs, whose serialised timing could sway the autotuner)
        lay = [int(x) for x in args.mixed_layout.split(",")]
        b = Q.Batch(mem, variant=lay[0], team=lay[1:] or None)
        variant, probe_ms = lay[0], []
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.advance(iv)
    el = time.perf_counter() - t0
    kms, _ = b.kernel_ms()
    tries = [int(m.counters()[1][4]) for m in mem]
    f32 = sum(try_flops(s)[0] * t for s, t in zip(systems, tries))
    f64 = sum(try_flops(s)[1] * t for s, t in zip(systems, tries))
    ach, peak = valu_roofline(f32, f64, kms / 1e3)
    # the same members one after another, each in its own launch with its default layout
    seq = 0.0
    for i, (s, n, (_, _, idx)) in enumerate(zip(systems, counts, MIXED)):
        e = Q.Ensemble(s, int(n), basis(s.dim, idx), Q.mcwf_pars(seed=1000 + i, precision=Q.QT_FP32_DERIV))
        t1 = time.perf_counter()
        e.advance(iv)
        seq += time.perf_counter() - t1
    return {"workload": f"{int(counts.sum())} trajectories drawn uniformly from 4 builds "
                        f"({', '.join(f'{n} {c}' for (n, _, _), c in zip(MIXED, counts))}), T={args.mixed_T}, "
                        "Dt=0.1, FP32 derivatives, one batched launch (two persistent kernels)",
            "tries": sum(tries), "tries_per_member": tries, "seconds": el, "tries_per_s": sum(tries) / el,
            "kernel_ms": kms, "variant": Q.batch_variants()[variant], "teams": b.layout()[1],
            "autotune_s": tune_s, "autotune_probe_ms": [float(x) for x in probe_ms],
            "sequential_seconds": seq, "sequential_tries_per_s": sum(tries) / seq,
            "roofline": {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                         "note": "sum over members of (6*8*M FP32 + 124*d FP64) flop per try"}}


def serve(Q, args):
    """Serving load: Poisson arrivals at a fixed rate, each an MCWF ensemble (256-4096 trajectories,
    log-uniform) of a system drawn from the PLM-20, JC-60, mixed-batch and JC-80 builds, or a
    Master-equation parameter sweep (4-16 JC-80 points), T=2, Dt=0.1, FP32 derivatives for the
    ensembles; continuous batching with memory-based admission."""
    from sysdefs import JC60, MIXED, PLM20, basis, coherent, ground, jc_master_spec
    rng = np.random.default_rng(4242)
    pool = [("plm20", Q.System(PLM20), coherent(0.5, 20)), ("jc60", Q.System(JC60), ground(60))]
    pool += [(n, Q.System(spec), None) for n, spec, _ in MIXED[:3]]
    idx = {n: i for n, _, i in MIXED}
    jc80 = [Q.System(jc_master_spec(40, eta=float(e), g=10.0)) for e in np.linspace(0.5, 15, 8)]
    rho80 = np.zeros((80, 80), complex)
    rho80[0, 0] = 1
    srv = Q.Server(hbm_budget=int(args.serve_budget_gib * (1 << 30)), max_active=16, chunk_intervals=args.serve_chunk,
                   pars=Q.mcwf_pars(precision=Q.QT_FP32_DERIV))
    t_start = time.perf_counter()
    t_next = t_start + rng.exponential(1.0 / args.serve_rate)
    pending, lat, waits, tries, kinds = {}, [], [], 0, {}
    submitted = 0
    t_report = t_start + 60.0
    while True:
        now = time.perf_counter()
        if now >= t_report:  # progress of a long load, once a minute
            print(f"serving: {now - t_start:.0f} s, {submitted} submitted, {len(lat)} finished", file=sys.stderr,
                  flush=True)
            t_report += 60.0
        while now >= t_next and now - t_start < args.serve_seconds:
            k = int(rng.integers(0, len(pool) + 2))
            if k == len(pool) + 1:
                # a parameter-sweep solve: 4-16 JC-80 Master-equation points (eta values), T=2
                npts = int(rng.integers(4, 17))
                pts = [jc80[int(i)] for i in rng.integers(0, len(jc80), npts)]
                rid = srv.submit_sweep(pts, rho80, T=2.0, Dt=0.1)
                pending[rid] = "sweep_jc80"
                submitted += 1
                t_next += rng.exponential(1.0 / args.serve_rate)
                continue
            if k == len(pool):
                name, s = "jc80", jc80[int(rng.integers(0, len(jc80)))]
                psi = ground(80)
            else:
                name, s, psi = pool[k]
                if psi is None:
                    psi = basis(s.dim, idx[name])
            n = int(round(np.exp(rng.uniform(np.log(256), np.log(4096)))))
            rid = srv.submit(s, n, psi, T=2.0, Dt=0.1, seed=int(rng.integers(1, 1 << 40)))
            pending[rid] = name
            submitted += 1
            t_next += rng.exponential(1.0 / args.serve_rate)
        a, w = srv.step()
        for rid in list(pending):
            r = srv.result(rid)
            if r is not None:
                lat.append(r[1]); waits.append(r[2]); tries += r[3]
                kinds[pending[rid]] = kinds.get(pending[rid], 0) + 1
                srv.release(rid)
                del pending[rid]
        if a == 0 and w == 0:
            if time.perf_counter() - t_start >= args.serve_seconds:
                break
            time.sleep(max(0.0, min(t_next - time.perf_counter(), 0.05)))
    el = time.perf_counter() - t_start
    st = srv.stats()
    lat = np.array(lat)
    return {"workload": f"Poisson arrivals at {args.serve_rate}/s for {args.serve_seconds} s; ensembles of 256-4096 "
                        "trajectories (log-uniform) of PLM-20, JC-60, particle-cavity 1024, four modes 4096, six qubits 64 "
                        "or JC-80 (eta in [0.5, 15]), or Master-equation sweeps of 4-16 JC-80 points; T=2, Dt=0.1; "
                        f"FP32 derivatives; up to 16 requests per step, {args.serve_chunk} intervals per step, "
                        f"{args.serve_budget_gib} GiB admission budget",
            "requests": int(len(lat)), "submitted": submitted, "seconds": el, "tries": tries, "tries_per_s": tries / el,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max()), "queue_wait_mean_s": float(np.mean(waits)),
            "launches": st["steps"], "kernel_s": st["kernel_ms"] / 1e3, "by_system": kinds}


if __name__ == "__main__":
    main()
