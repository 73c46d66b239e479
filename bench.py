#!/usr/bin/env python3
"""Benchmark of the Lawson-RK4 propagation step on one MI355X.

Default workload (one GPU): the batched workload -- 256 independent GRIN fibers
(configs.batched_fiber(0..255): M in {3, 6, 10}, NT in {4096, 8192, 16384}, lengths
U(0.2, 1) m, pulses T0 U(0.05, 0.3) ps with peak powers log-uniform in [10 W, 10 kW],
symmetric GRIN coupling, RK4, Raman on) in nine batches of one (M, NT) class each, in
mixed precision (FP32 transforms, pair products, Raman convolutions and Kerr sums; FP64
spectral state, Lawson phases and RK accumulation). A "step" is one Lawson-RK4 step of
every fiber. The timed region starts with every fiber at z = 0 and its initial field
resident in HBM and ends after the last step's kernels complete; it covers exactly
--steps steps (HIP graphs are built when the solvers are created, outside it).

Prints one JSON line: value in M*NT*steps/s over the fibers whose fields stay finite,
the plain FP64 single-fiber solver's rate over the same fibers and the ratio, the
per-class rel-L2 of the lowest-power fiber against the FP64 oracle, the CPU baseline
(the oracle on the host cores, the reference's decomposition), the dominant kernel's
roofline from dispatch-recorded HIP events, and the per-step kernel time summed over
the classes beside ms_per_step.

Other workloads (--workload): grin6 (the 6-mode GRIN fiber, M = 6, NT = 16384, plain FP64),
gaussian2 (the 2-mode Gaussian, FP64, compared with the oracle over all 5000 steps),
grin10_mixed (M = 10, NT = 65536, mixed precision) and serving (Poisson arrivals through
the continuous-batching engine).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "M·NT·n_steps/s at stated rel-L2 vs FP64 oracle; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}  # MI355X vector peaks (FP32 with packed instructions)
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "hbm_traffic_grin6_fp64.md")
BATCHED_TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "hbm_traffic_batched.md")


def measured_traffic(path=TRAFFIC_PROFILE):
    """Bytes per launch of each kernel from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (copy-calibrated), or {} when it is absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return json.loads(line.split(":", 1)[1])
    except OSError:
        pass
    return {}


def single_workloads():
    """Single-fiber workloads: (parameters, precision, pair decode, oracle prefix steps, description)."""
    from mm_propagation_amd import _capi, configs
    return {
        "grin6": (lambda: configs.grin_fiber(M=6, nt=16384, n_steps=20000, symmetric=True),
                  _capi.MMP_PRECISION_FP64, _capi.MMP_PAIRS_PHYSICAL, 200,
                  "grin6_fp64: M=6, NT=16384, T_w=20 ps, L=1 m, n_steps=20000, RK4, Raman on, "
                  "symmetric GRIN coupling nnz=648, physical pair decode"),
        "gaussian2": (lambda: configs.two_mode_gaussian(),
                      _capi.MMP_PRECISION_FP64, _capi.MMP_PAIRS_REFERENCE_CPU, 5000,
                      "gaussian2_fp64: M=2, NT=4096, T_w=10 ps, L=0.5 m, n_steps=5000, RK4, Raman on, "
                      "3 nonzeros per mode, the reference CPU pair decode"),
        # peak powers 1-100 W: at 10-1000 W this 10-mode fiber is not resolved by 20000 steps (the
        # oracle's field, like the device's, is no longer finite by step 325 at NT = 16384)
        "grin10_mixed": (lambda: configs.grin_fiber(M=10, nt=65536, n_steps=20000, symmetric=True,
                                                    power_range=(1.0, 100.0)),
                         _capi.MMP_PRECISION_MIXED, _capi.MMP_PAIRS_PHYSICAL, 20,
                         "grin10_mixed: M=10, NT=65536, T_w=20 ps, L=1 m, n_steps=20000, RK4, Raman on, "
                         "symmetric GRIN coupling nnz=5000, peak powers log-uniform in [1, 100] W, physical "
                         "pair decode, FP32 transforms and sums with FP64 state, phases and accumulation"),
    }


def cpu_baseline(sp, steps: int, decode: int, name: str):
    """Time the FP64 C restatement of the reference CPU path on a bounded sample: the first
    `steps` steps of the solve from z = 0 (computeLawsonRK(steps), operator set-up included),
    with the reference's decomposition (OpenMP over modes, the Raman pair loop serial,
    src/MultimodePropagation_CPU.cpp:196-205). Returns the baseline record and the oracle's
    field at z = steps * h."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_binding as ob
    threads = min(16, os.cpu_count() or 1)
    o = ob.Oracle(sp, pair_decode=decode, n_threads=threads, parallel_pairs=False)
    t0 = time.perf_counter()
    o.compute_lawson_rk(steps)
    dt = time.perf_counter() - t0
    fields = o.get_result()
    o.close()
    rec = {"value": sp.n_modes * sp.nt * steps / dt, "unit": "M·NT·steps/s", "cores": threads,
           "kind": "port",
           "sample": f"{name} M={sp.n_modes} NT={sp.nt}: first {steps} of {sp.n_steps} RK4 steps in {dt:.2f} s, "
                     f"OpenMP {threads} threads over modes, the Raman pair loop serial as in the reference, "
                     "radix-2 FFT (no FFTW here)"}
    return rec, fields


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's n_steps)")
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--profile-steps", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="oracle prefix for the CPU baseline and the accuracy check (default per workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels one by one instead of replaying HIP graphs (for kernel tracers)")
    ap.add_argument("--autotune", dest="autotune", action="store_true", default=None,
                    help="MMP_FLAG_AUTOTUNE: each solver times its column-stage variants at creation and keeps the "
                         "fastest (default for the batched workload, off for the others)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--matrix-cores", choices=["default", "on", "off"], default="default",
                    help="mixed precision: the gather's coupling sums on the matrix cores by default at 4-10 and 13-16 modes "
                         "(on: MMP_FLAG_MATRIX_CORES, wherever that gather exists; off: MMP_FLAG_NO_MATRIX_CORES)")
    ap.add_argument("--separate-columns", action="store_true",
                    help="every column pass a kernel of its own (7 per stage instead of 4)")
    ap.add_argument("--mixed", action="store_true",
                    help="mixed precision (FP32 transforms and sums, FP64 state); the bench line is FP64 by default")
    ap.add_argument("--workload", choices=["grin6", "gaussian2", "grin10_mixed", "batched", "serving", "gather"],
                    default="batched",
                    help="batched (default): 256 independent GRIN fibers "
                         "(M in {3, 6, 10}, NT in {4096, 8192, 16384}) in batches per (M, NT), mixed precision; "
                         "grin6: the 6-mode GRIN fiber in FP64; gaussian2: the 2-mode Gaussian in FP64; "
                         "grin10_mixed: the 10-mode GRIN fiber at NT=65536 in mixed precision; serving: Poisson "
                         "arrivals of solves (M in {1, 3, 6, 10, 15}, NT 2^12..2^16, 500..20000 steps) through "
                         "the continuous-batching engine; gather: one batch of --fibers GRIN fibers of --modes "
                         "modes at --nt samples (1-100 W), mixed precision, the column-stage gather on the vector "
                         "ALUs and on the matrix cores, dispatch-recorded time per launch of each")
    ap.add_argument("--modes", type=int, default=10, help="gather workload: modes")
    ap.add_argument("--nt", type=int, default=16384, help="gather workload: samples")
    ap.add_argument("--fibers", type=int, default=256, help="batched workload: number of fibers")
    ap.add_argument("--fp64-batch", action="store_true", help="batched workload in FP64 (default: mixed)")
    ap.add_argument("--batch-streams", choices=["concurrent", "one", "per-class"], default="concurrent",
                    help="batched workload: concurrent (default): one graph of every class batch, each class a "
                         "parallel branch on its own stream; one: one graph, the classes in order on one stream; "
                         "per-class: each class's own graphs on its own stream, no group")
    ap.add_argument("--requests", type=int, default=32, help="serving workload: number of requests")
    ap.add_argument("--rate", type=float, default=2.0, help="serving workload: Poisson arrival rate, requests/s")
    ap.add_argument("--serving-seed", type=int, default=0, help="serving workload: arrival and shape seed")
    ap.add_argument("--slots", type=int, default=16, help="serving workload: fiber slots per batch group")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the bench, copy this many MiB device to device once with the L2 evicted first "
                         "(the reference dispatch for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(independent fibers shard across GPUs with no exchange)"}))
        return 2
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "serving":
        return run_serving(args)
    if args.workload == "gather":
        return run_gather(args)

    from mm_propagation_amd import _capi, configs
    from mm_propagation_amd.solver import Context, MultimodePropagationGPU_RamanON

    make, prec, decode, cpu_steps, desc = single_workloads()[args.workload]
    sp = make()
    if args.mixed:
        prec = _capi.MMP_PRECISION_MIXED
    steps = args.steps if args.steps is not None else sp.n_steps
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else cpu_steps
    mixed = prec == _capi.MMP_PRECISION_MIXED
    M, NT = sp.n_modes, sp.nt
    ctx = Context(0)
    flags = ((_capi.MMP_FLAG_NO_GRAPH if args.no_graph else 0) |
             (_capi.MMP_FLAG_SEPARATE_COLUMNS if args.separate_columns else 0) |
             (_capi.MMP_FLAG_AUTOTUNE if args.autotune else 0))
    s = MultimodePropagationGPU_RamanON(sp, pair_decode=decode, flags=flags, ctx=ctx, precision=prec)
    tuned = {k: s.info()[k] for k in ("gather_waves", "chunk_lines", "tuned", "pair_columns", "pair_groups",
                                      "gather_hi_occupancy")}
    psi0, _ = s.get_state()
    if args.warmup:
        s.advance(args.warmup)
    ctx.synchronize()
    s.set_state(psi0)  # back to z = 0 with the initial field
    ctx.synchronize()
    t0 = time.perf_counter()
    s.advance(steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = M * NT * steps / elapsed
    fields = s.getResult()
    energy = s.energy()
    e0 = configs.energy_per_mode(sp.initial_fields, sp.time_window)
    accuracy = {"energy_drift_rel": float(abs(energy.sum() - e0.sum()) / e0.sum()),
                "energy_note": "total energy at z = steps * h against the initial field's (the Raman term moves "
                               "energy between frequencies, so it is not an invariant of this physics)"}
    s.close()
    if mixed and steps == sp.n_steps:
        # the plain FP64 solver over the same steps: the mixed field and energy at z = L against it
        f64 = MultimodePropagationGPU_RamanON(sp, pair_decode=decode, ctx=ctx)
        f64.computeLawsonRK(steps)
        ref64 = f64.getResult()
        en64 = f64.energy()
        f64.close()
        accuracy["rel_l2_vs_fp64_gpu_at_L_max"] = float(configs.rel_l2_per_mode(fields, ref64).max())
        accuracy["energy_drift_vs_fp64_gpu_at_L"] = float((energy.sum() - en64.sum()) / en64.sum())

    # per-kernel HIP-event timings on the solver's stream over a profiled run of the same workload
    ps = MultimodePropagationGPU_RamanON(sp, pair_decode=decode,
                                         flags=_capi.MMP_FLAG_PROFILE |
                                         (flags & (_capi.MMP_FLAG_SEPARATE_COLUMNS | _capi.MMP_FLAG_AUTOTUNE)),
                                         ctx=ctx, precision=prec)
    ps.advance(8)
    ps.kernel_times()
    ps.advance(args.profile_steps)
    kt = ps.kernel_times()
    ps.close()
    total_ms = sum(k["total_ms"] for k in kt) or 1.0
    dom = max(kt, key=lambda k: k["total_ms"])
    avg_s = dom["total_ms"] / 1e3 / max(dom["launches"], 1)
    achieved = dom["bytes_per_launch"] / avg_s / 1e9
    traffic = (measured_traffic().get(dom["name"])
               if args.workload == "grin6" and not (mixed or args.separate_columns) else None)
    roofline = {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": "profiles/hbm_traffic_grin6_fp64.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, "
                                  "copy-calibrated)" if traffic else None,
                "avg_launch_us": round(avg_s * 1e6, 3), "bytes_per_launch": dom["bytes_per_launch"],
                "share_of_kernel_time": round(dom["total_ms"] / total_ms, 4),
                "kernels": {k["name"]: {"avg_us": round(k["total_ms"] * 1e3 / max(k["launches"], 1), 3),
                                        "launches": k["launches"],
                                        "gbps": round(k["bytes_per_launch"] / max(k["total_ms"] / 1e3 /
                                                                                  max(k["launches"], 1), 1e-12) / 1e9, 1)}
                            for k in kt}}

    cpu = None
    if not args.no_cpu_baseline and cpu_steps > 0:
        cpu, ref = cpu_baseline(sp, cpu_steps, decode, args.workload)
        # the same solve prefix on the device, compared field for field with the oracle
        vs = MultimodePropagationGPU_RamanON(sp, pair_decode=decode, flags=flags, ctx=ctx, precision=prec)
        vs.computeLawsonRK(cpu_steps)
        got = vs.getResult()
        en = vs.energy()
        vs.close()
        err = configs.rel_l2_per_mode(got, ref)
        e_ref = configs.energy_per_mode(ref, sp.time_window)
        accuracy["rel_l2_vs_oracle_at_steps"] = cpu_steps
        accuracy["rel_l2_vs_oracle_max"] = float(err.max())
        accuracy["rel_l2_vs_oracle_per_mode"] = [float(x) for x in err]
        accuracy["energy_drift_vs_oracle_at_steps"] = float((en.sum() - e_ref.sum()) / e_ref.sum())
        if args.workload != "gaussian2":
            accuracy["note"] = ("fields compared after the first cpu_steps steps: this GRIN workload is chaotic, "
                                "two FP64 solvers that differ only in FFT rounding separate to 1e-4 by step 500")

    out = {
        "metric": METRIC, "value": value, "unit": "M·NT·steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64", "data": "synthetic",
        "config": {"workload": desc + (" (run in mixed precision)" if args.mixed else ""),
                   "fibers": 1, "nt": NT, "devices": "one GPU"},
        "accuracy": accuracy, "roofline": roofline, "cpu_baseline": cpu, "gather_plan": tuned,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
    return 0


def colstage_gather_flops(M: int, NT: int, P: int, L: int, parity: bool) -> float:
    """Arithmetic of one fiber's column-stage Kerr + Raman gather launch (k_colstage_raman_gather),
    per the algorithm it runs, in floating-point operations: the inverse column DFTs of the L Raman
    lines (64-point DFTs: 5 * 64 * log2 64 per column, 30 per sample, plus the 6-flop four-step
    twiddle), the slot mixes kv U_a conj(U_b) + kr V_R (12 per slot and sample), the products
    X_s U_k (6 per slot, mode k and sample; ceil(M/2) modes per slot with a parity-structured
    coupling, M otherwise), the coupling sums sigma_p += Q X_s U_k (4 per slot, mode pair and
    sample) and the forward column DFT of sigma (36 per mode and sample)."""
    K = (M + 1) // 2 if parity else M
    per_sample = 36 * L + 12 * P + 6 * P * K + 4 * P * M * K + 36 * M
    return float(NT) * per_sample


def colstage_gather_mfma_flops(M: int, NT: int, P: int, L: int, parity: bool, chunks: int) -> float:
    """Arithmetic of the matrix-core form of that launch, in floating-point operations: the same
    line DFTs (30 per sample and line, plus the 6-flop twiddle), slot mixes (12 per slot) and
    forward DFT of sigma (36 per mode), the coupling products B[p][k] += Q X_s on the matrix cores
    (4 per slot, mode pair and sample, the zero padding of the 32-row tiles not counted), and
    sigma_p += U_k B[p][k] (8 per mode pair and sample, once per chunk and parity class)."""
    K = (M + 1) // 2 if parity else M
    classes = 2 if parity else 1
    per_sample = 36 * L + 12 * P + 4 * P * M * K + 8 * M * K * chunks * classes + 36 * M
    return float(NT) * per_sample


def coupling_parity(sp) -> bool:
    """Whether each pair slot (l, m) of the coupling couples only modes p, k of one parity of
    p + k (zero values aside): the structure under which the gather keeps ceil(M/2) modes per slot."""
    seen = {}
    for p, arr in enumerate(sp.coupling_coefficients):
        for nz in arr:
            (k, l, m), q = nz.index, nz.value
            if q == 0:
                continue
            par = (p + k) % 2
            if seen.setdefault((l, m), par) != par:
                return False
    return True


def run_batched(args) -> int:
    """The batched workload: 256 independent fibers (configs.batched_fiber(i), i < --fibers) in
    batches of one (M, NT) class each, mixed precision; a step is one RK4 step of every fiber.
    Timed: exactly --steps steps of every fiber, fields resident, after --warmup steps and a
    restart of every fiber at z = 0 from its initial field."""
    from mm_propagation_amd import _capi, configs
    from mm_propagation_amd.batch import MultimodePropagationBatch, SolverGroup
    from mm_propagation_amd.solver import Context, MultimodePropagationGPU_RamanON

    prec = _capi.MMP_PRECISION_FP64 if args.fp64_batch else _capi.MMP_PRECISION_MIXED
    fibers = [configs.batched_fiber(i) for i in range(args.fibers)]
    classes = {}
    for i, sp in enumerate(fibers):
        classes.setdefault((sp.n_modes, sp.nt), []).append(i)
    autotune = True if args.autotune is None else args.autotune
    bflags = ((_capi.MMP_FLAG_AUTOTUNE if autotune else 0) | (_capi.MMP_FLAG_NO_GRAPH if args.no_graph else 0) |
              {"default": 0, "on": _capi.MMP_FLAG_MATRIX_CORES, "off": _capi.MMP_FLAG_NO_MATRIX_CORES}[args.matrix_cores])
    one_stream = args.batch_streams == "one"
    use_group = args.batch_streams in ("one", "concurrent") and not args.no_graph
    shared = Context(0) if one_stream else None
    batches = []
    t_setup = time.perf_counter()
    for key in sorted(classes, key=lambda k: -k[0] * k[1] * len(classes[k])):  # largest classes first
        ctx = shared or Context(0)
        b = MultimodePropagationBatch([fibers[i] for i in classes[key]], precision=prec, ctx=ctx, flags=bflags)
        batches.append((key, classes[key], ctx, b))
    t_setup = time.perf_counter() - t_setup
    plans = {f"M{k[0]}_NT{k[1]}": {"fibers": len(idx), "gather_waves": b.info(0)["gather_waves"],
                                   "chunk_lines": b.info(0)["chunk_lines"], "pair_columns": b.info(0)["pair_columns"],
                                   "pair_groups": b.info(0)["pair_groups"],
                                   "gather_hi_occupancy": b.info(0)["gather_hi_occupancy"],
                                   "gather_matrix_cores": b.info(0)["gather_matrix_cores"]}
             for k, idx, _, b in batches}
    contexts = [shared] if one_stream else [ctx for _, _, ctx, _ in batches]
    psi0 = {(bi, f): b.state(f)[0] for bi, (_, idx, _, b) in enumerate(batches) for f in range(len(idx))}

    # every class in one graph (mmp_group: one launch per 16 steps of all 256 fibers); with
    # per-class contexts each class is a parallel branch of that graph, forked from and joined
    # into the group context's stream
    group = None
    if use_group:
        gctx = shared or Context(0)
        group = SolverGroup([b for _, _, _, b in batches], gctx)
        if gctx is not shared:
            contexts = contexts + [gctx]

    def run(n):
        """n steps of every fiber: the group's graph, or the classes' 16-step graphs interleaved."""
        if group is not None:
            group.advance(n)
        else:
            left = n
            while left > 0:
                k = min(16, left)
                for _, _, _, b in batches:
                    b.advance(k)
                left -= k
        for c in contexts:
            c.synchronize()

    def restart_all():
        for bi, (_, idx, _, b) in enumerate(batches):
            for f in range(len(idx)):
                b.set_state(f, psi0[(bi, f)])  # that fiber back to z = 0 with its initial field
        for c in contexts:
            c.synchronize()

    steps = args.steps if args.steps is not None else 512
    if args.warmup:
        run(args.warmup)
        restart_all()
    t0 = time.perf_counter()
    run(steps)
    elapsed = time.perf_counter() - t0

    # only the fibers whose field stays finite count: the device parks a fiber in the step its
    # field stops being finite (its later steps skip it), and none of its steps is throughput
    finite = set()
    halted_at = []
    for bi, (_, idx, _, b) in enumerate(batches):
        for f in range(len(idx)):
            info = b.info(f)
            if info["halted"]:
                halted_at.append(int(info["steps_taken"]))
            else:
                finite.add((bi, f))
    finite_fibers = [idx[f] for bi, (_, idx, _, _) in enumerate(batches) for f in range(len(idx)) if (bi, f) in finite]

    # the same warm-up and timed steps with the classes in order on one stream (one graph, no
    # branches; a second set of batches, the tuner's cached plans): its wall time per step is what
    # the classes' kernel times add up to, and the difference to ms_per_step is the branches' overlap
    in_order_ms = None
    if group is not None and args.batch_streams == "concurrent":
        sctx = Context(0)
        sb = [MultimodePropagationBatch([fibers[i] for i in idx], precision=prec, ctx=sctx, flags=bflags)
              for _, idx, _, _ in batches]
        sg = SolverGroup(sb, sctx)
        if args.warmup:
            sg.advance(args.warmup)
            for bi, (b2, (_, idx, _, _)) in enumerate(zip(sb, batches)):
                for f in range(len(idx)):
                    b2.set_state(f, psi0[(bi, f)])
            sctx.synchronize()
        t1 = time.perf_counter()
        sg.advance(steps)
        sctx.synchronize()
        in_order_ms = (time.perf_counter() - t1) * 1e3 / steps
        sg.close()
        for b2 in sb:
            b2.close()
        sctx.close()
    for bi, (key, idx, _, _) in enumerate(batches):
        plans[f"M{key[0]}_NT{key[1]}"]["finite_after_timed_steps"] = sum(1 for f in range(len(idx)) if (bi, f) in finite)
    units = sum(fibers[i].n_modes * fibers[i].nt for i in finite_fibers)
    value = units * steps / elapsed

    # accuracy and the CPU baseline: in every class the lowest-power fiber, after the timed
    # steps (at most the first --cpu-steps of them), against the FP64 oracle run on the host with
    # the reference's decomposition (OpenMP over modes, the Raman pair loop serial)
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_binding as ob
    acc_steps = min(steps, args.cpu_steps if args.cpu_steps is not None else 16)
    threads = min(16, os.cpu_count() or 1)
    acc, cpu_time, cpu_units, cpu_per_step = {}, 0.0, 0, {}
    if not args.no_cpu_baseline:
        restart_all()
        run(acc_steps)
        for bi, (key, idx, _, b) in enumerate(batches):
            f = min(range(len(idx)), key=lambda j: float(np.max(np.abs(fibers[idx[j]].initial_fields) ** 2)))
            sp = fibers[idx[f]]
            o = ob.Oracle(sp, pair_decode=_capi.MMP_PAIRS_PHYSICAL, n_threads=threads, parallel_pairs=False)
            tc = time.perf_counter()
            o.compute_lawson_rk(acc_steps)
            dt = time.perf_counter() - tc
            ref = o.get_result()
            o.close()
            got = b.result(f)
            err = configs.rel_l2_per_mode(got, ref)
            e_ref = configs.energy_per_mode(ref, sp.time_window).sum()
            e_got = configs.energy_per_mode(got, sp.time_window).sum()
            acc[f"M{key[0]}_NT{key[1]}_fiber{idx[f]}"] = {"rel_l2_max": float(err.max()),
                                                         "energy_drift": float((e_got - e_ref) / e_ref)}
            cpu_per_step[f"M{key[0]}_NT{key[1]}"] = round(dt / acc_steps * 1e3, 2)
            n_fin = sum(1 for f2 in range(len(idx)) if (bi, f2) in finite)
            cpu_time += dt / acc_steps * n_fin  # the class's finite fibers one after another, per step
            cpu_units += sp.n_modes * sp.nt * n_fin
    cpu = None
    if cpu_time > 0:
        cpu = {"value": cpu_units / cpu_time, "unit": "M·NT·steps/s", "cores": threads, "kind": "port",
               "sample": f"the lowest-power fiber of each of the {len(batches)} classes, {acc_steps} RK4 steps each "
                         f"(ms per step: {cpu_per_step}), extrapolated to the {len(finite)} finite fibers; the oracle "
                         f"(FP64 C restatement of the reference CPU path) with {threads} OpenMP threads over "
                         "modes and the Raman pair loop serial as in the reference; radix-2 FFT (no FFTW here)"}

    # the plain FP64 single-fiber solver on one fiber of each class, the batch's fibers one after
    # another at those rates
    plain_time = 0.0
    plain = {}
    pctx = shared or batches[0][2]
    def power(i):
        return float(np.max(np.abs(fibers[i].initial_fields) ** 2))

    for bi, (key, idx, _, _) in enumerate(batches):
        fin = sorted((idx[f] for f in range(len(idx)) if (bi, f) in finite), key=power)
        per_step = None
        # a fiber that stays finite over the timing window (a parked fiber skips its steps); the
        # window is the timed run's own length at most, over which these fibers stayed finite
        n_time = max(1, min(128, steps - 4))
        for i in fin:
            s1 = MultimodePropagationGPU_RamanON(fibers[i], ctx=pctx,
                                                 flags=_capi.MMP_FLAG_NO_GRAPH if args.no_graph else 0)
            s1.advance(4)
            pctx.synchronize()
            t1 = time.perf_counter()
            s1.advance(n_time)
            pctx.synchronize()
            dt = (time.perf_counter() - t1) / n_time
            halted = s1.info()["halted"]
            s1.close()
            if not halted:
                per_step = dt
                break
        if per_step is None:
            continue
        plain[f"M{key[0]}_NT{key[1]}"] = round(per_step * 1e6, 2)
        plain_time += per_step * sum(1 for f in range(len(idx)) if (bi, f) in finite)
    plain_value = units / plain_time if plain_time > 0 else None

    # per-kernel dispatch-recorded times (eager launches of the same kernels): every class replays
    # the timed run (from the initial fields, the same steps, the same fibers parked on the way),
    # and their summed kernel time per step stands beside ms_per_step
    kernel_us_per_step = 0.0
    per_class = {}
    per_class_kernel = {}
    for key, idx, ctx, _ in batches:
        pb = MultimodePropagationBatch([fibers[i] for i in idx], precision=prec, ctx=ctx,
                                       flags=_capi.MMP_FLAG_PROFILE | (bflags & _capi.MMP_FLAG_AUTOTUNE))
        pb.kernel_times()
        pb.advance(steps)
        kt = pb.kernel_times()
        pb.close()
        cls_us = sum(k["total_ms"] for k in kt) * 1e3 / steps
        kernel_us_per_step += cls_us
        per_class[f"M{key[0]}_NT{key[1]}"] = round(cls_us, 1)
        per_class_kernel[f"M{key[0]}_NT{key[1]}"] = {k["name"]: round(k["total_ms"] * 1e3 / steps, 1) for k in kt}
    # the dominant kernel: the class with the most finite work, its finite fibers only (every
    # fiber of each launch live)
    def finite_units(t):
        bi, (key, idx, _, _) = t
        return key[0] * key[1] * sum(1 for f in range(len(idx)) if (bi, f) in finite)

    bi, (key, idx, ctx, _) = max(enumerate(batches), key=finite_units)
    fin = [idx[f] for f in range(len(idx)) if (bi, f) in finite] or idx
    # untuned: a tuning run would step these fibers past the steps over which they stayed finite
    pb = MultimodePropagationBatch([fibers[i] for i in fin], precision=prec, ctx=ctx, flags=_capi.MMP_FLAG_PROFILE)
    pb.advance(min(4, steps))
    pb.kernel_times()
    pb.advance(max(1, min(16, steps - 4)))  # within the steps over which these fibers stayed finite
    kt = pb.kernel_times()
    info = pb.info(0)
    pb.close()
    # every pass of that class launch: algorithmic bytes over the dispatch-recorded time (the
    # transforms' HBM rates beside the dominant kernel's)
    passes = {p["name"]: {"avg_us": round(p["total_ms"] * 1e3 / max(p["launches"], 1), 3),
                          "gbps": round(p["bytes_per_launch"] / (p["total_ms"] / 1e3 / max(p["launches"], 1)) / 1e9, 1),
                          "hbm_frac": round(p["bytes_per_launch"] / (p["total_ms"] / 1e3 / max(p["launches"], 1))
                                            / 1e9 / HBM_PEAK_GBS, 4)}
              for p in kt if p["launches"] > 0}
    dom = (max(kt, key=lambda k: k["total_ms"]), key, fin, info)
    k, key, idx, info = dom
    avg_s = k["total_ms"] / 1e3 / max(k["launches"], 1)
    gbps = k["bytes_per_launch"] / avg_s / 1e9
    if k["name"] == "col_raman_gather":
        M, NT = key
        par = coupling_parity(fibers[idx[0]])
        matrix = bool(info["gather_matrix_cores"])
        if matrix:
            flops = colstage_gather_mfma_flops(M, NT, int(info["n_pairs"]), int(info["n_conv_lines"]), par,
                                               int(info["gather_chunks"])) * len(idx)
        else:
            flops = colstage_gather_flops(M, NT, int(info["n_pairs"]), int(info["n_conv_lines"]), par) * len(idx)
        tflops = flops / avg_s / 1e12
        # FP32: the matrix cores' dense FP32 peak equals the packed vector peak (157.3 TFLOP/s)
        peak = VALU_PEAK_TFLOPS["f64" if args.fp64_batch else "f32"]
        arith = "mfma" if matrix else "valu"
        if tflops / peak >= gbps / HBM_PEAK_GBS:
            roofline = {"bound": arith, "kernel": k["name"], "achieved": round(tflops, 3), "peak": peak,
                        "unit": "TFLOP/s", "frac": round(tflops / peak, 5), "traffic": None,
                        "hbm_gbps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBS, 4)}
        else:
            roofline = {"bound": "hbm", "kernel": k["name"], "achieved": round(gbps, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBS, 5), "traffic": None,
                        f"{arith}_tflops": round(tflops, 3), f"{arith}_frac": round(tflops / peak, 4)}
        roofline["matrix_cores"] = matrix
        roofline.update({"flops_per_launch": flops, "bytes_per_launch": k["bytes_per_launch"]})
    else:
        roofline = {"bound": "hbm", "kernel": k["name"], "achieved": round(gbps, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBS, 5), "traffic": None,
                    "bytes_per_launch": k["bytes_per_launch"]}
    cls = f"M={key[0]} NT={key[1]} x {len(idx)} fibers"
    traffic = measured_traffic(BATCHED_TRAFFIC_PROFILE).get(f"{k['name']} {cls}") if not args.fp64_batch else None
    if traffic:
        roofline.update({"traffic": traffic, "traffic_source": "profiles/hbm_traffic_batched.md (rocprofv3 FETCH_SIZE "
                                                               "x2.0 + WRITE_SIZE x1.0, copy-calibrated)"})
    roofline.update({"class": cls, "avg_launch_us": round(avg_s * 1e6, 3), "passes": passes,
                     "peak_note": "FP32 peak 157.3 TFLOP/s: the matrix cores' dense FP32 rate, and the vector rate "
                                  "with packed FP32 instructions, which this build does not emit (78.6 TFLOP/s "
                                  "without them, as FP64)"})
    out = {
        "metric": METRIC, "value": value, "unit": "M·NT·steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.fp64_batch else "f32+f64", "data": "synthetic",
        "config": {"workload": f"batched{args.fibers}: fibers configs.batched_fiber(0..{args.fibers - 1}) "
                               "(M in {3,6,10}, NT in {4096,8192,16384}, L U(0.2,1) m, T0 U(0.05,0.3) ps, "
                               "P log-uniform [10 W, 10 kW], symmetric GRIN coupling, RK4, Raman on, "
                               + ("FP64" if args.fp64_batch else "FP32 transforms and sums, FP64 state") +
                               f"), {len(batches)} batches by (M, NT), "
                               + ({"concurrent": "one graph for all classes, each class a parallel branch on its "
                                                 "own stream",
                                   "one": "one graph for all classes, in order on one stream",
                                   "per-class": "each class's own graphs on its own stream"}[args.batch_streams]
                                  if group is not None or args.batch_streams == "per-class"
                                  else "one stream, eager launches")
                               + (", autotuned" if autotune else ""),
                   "fibers": args.fibers, "nt": "4096-16384", "devices": "one GPU"},
        "fibers_finite": len(finite),
        "fibers_parked": {"count": len(halted_at), "steps_before_parking_median":
                          float(np.median(halted_at)) if halted_at else None},
        "setup_s": round(t_setup, 2),
        "plain_fp64_single_fiber": {"value": plain_value, "per_step_us": plain,
                                    "ratio": value / plain_value if plain_value else None,
                                    "note": "the plain FP64 solver's per-step time for one fiber of each class, "
                                            "over the same finite fibers solved one after another"},
        "cpu_ratio": (value / cpu["value"]) if cpu else None,
        "accuracy": {"rel_l2_vs_fp64_oracle": acc, "steps": acc_steps},
        "kernel_sum_ms_per_step": round(kernel_us_per_step / 1e3, 4),
        "kernel_sum_note": "each class's kernels timed alone with dispatch events, summed over classes; where "
                           "classes run on streams of their own their kernels overlap on the device, and "
                           "ms_per_step may then be below this sum",
        "ms_per_step_classes_in_order": round(in_order_ms, 4) if in_order_ms is not None else None,
        "ms_per_step_classes_in_order_note": "the same warm-up and timed steps of the same fibers with the nine "
                                             "classes in order on one stream (one graph without branches)",
        "kernel_us_per_step_by_class": per_class,
        "kernel_us_per_step_by_class_and_kernel": per_class_kernel,
        "roofline": roofline, "cpu_baseline": cpu,
        "plans": plans,
    }
    print(json.dumps(out), flush=True)
    if group is not None:
        group.close()
    for _, _, ctx, b in batches:
        b.close()
    for c in contexts:
        c.close()
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
    return 0


def run_gather(args) -> int:
    """The column-stage gather of one batch (`--fibers` GRIN fibers, `--modes` modes, `--nt`
    samples, 1-100 W so that every fiber stays finite, mixed precision, untuned) in its two forms:
    the coupling sums on the vector ALUs (MMP_FLAG_NO_MATRIX_CORES) and on the matrix cores
    (MMP_FLAG_MATRIX_CORES, where that form exists). Eager launches with dispatch-recorded events:
    4 warm-up steps, then exactly --steps steps (default 16) timed per kernel."""
    from mm_propagation_amd import _capi, configs
    from mm_propagation_amd.batch import MultimodePropagationBatch

    F = args.fibers if args.fibers != 256 else 16
    steps = args.steps if args.steps is not None else 16
    sps = [configs.grin_fiber(M=args.modes, nt=args.nt, n_steps=20000, fiber_length=0.5, seed=200 + i,
                              symmetric=True, power_range=(1.0, 100.0)) for i in range(F)]
    forms = {}
    for name, flag in (("vector", _capi.MMP_FLAG_NO_MATRIX_CORES), ("matrix", _capi.MMP_FLAG_MATRIX_CORES)):
        b = MultimodePropagationBatch(sps, pair_decode=_capi.MMP_PAIRS_PHYSICAL, precision=_capi.MMP_PRECISION_MIXED,
                                      flags=flag | _capi.MMP_FLAG_PROFILE)
        info = b.info(0)
        if name == "matrix" and not info["gather_matrix_cores"]:
            b.close()
            continue
        b.advance(4)
        b.kernel_times()
        b.advance(steps)
        kt = b.kernel_times()
        b.close()
        g = next(k for k in kt if k["name"] == "col_raman_gather")
        avg_s = g["total_ms"] / 1e3 / g["launches"]
        flops = (colstage_gather_mfma_flops(args.modes, args.nt, int(info["n_pairs"]), int(info["n_conv_lines"]), True,
                                            int(info["gather_chunks"])) if name == "matrix" else
                 colstage_gather_flops(args.modes, args.nt, int(info["n_pairs"]), int(info["n_conv_lines"]), True)) * F
        forms[name] = {"avg_launch_us": round(avg_s * 1e6, 2), "tflops": round(flops / avg_s / 1e12, 2),
                       "frac_of_157.3": round(flops / avg_s / 1e12 / VALU_PEAK_TFLOPS["f32"], 4),
                       "hbm_gbps": round(g["bytes_per_launch"] / avg_s / 1e9, 1),
                       "step_kernel_us": round(sum(k["total_ms"] for k in kt) * 1e3 / steps, 1),
                       "gather_waves": info["gather_waves"], "gather_hi_occupancy": info["gather_hi_occupancy"]}
    best = min(forms.values(), key=lambda f: f["avg_launch_us"])
    units = args.modes * args.nt * F
    out = {"metric": "column-stage gather launches per second x M·NT (one RK stage of every fiber)",
           "value": units / (best["avg_launch_us"] * 1e-6), "unit": "M·NT/s", "n_gpus": 1, "steps": steps,
           "warmup": 4, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"gather: {F} GRIN fibers, M = {args.modes}, NT = {args.nt}, 1-100 W, mixed, untuned"},
           "forms": forms}
    print(json.dumps(out), flush=True)
    return 0


def rk4_bytes_per_step(M: int, NT: int, P: int, raman: bool, c_w: int) -> float:
    """Algorithmic HBM bytes of one RK4 step: every logical array read or written once per stage,
    field state at 16 B per sample, work arrays at c_w B (16 FP64, 8 mixed):
    NT * [6 * 16 * M + c_w * (19 M + 4 R M + 8 R P)]."""
    R = 1 if raman else 0
    return NT * (6 * 16 * M + c_w * (19 * M + 4 * R * M + 8 * R * P))


def run_serving(args) -> int:
    """Continuous batching of propagation solves: `--requests` requests arrive as a Poisson process
    at `--rate` per second (serving_requests(seed)), with M in {1, 3, 6, 10, 15}, NT in
    2^12..2^16, n_steps log-uniform in [500, 20000], pulses as in the batched workload (10 W to
    10 kW, T0 U(0.05, 0.3) ps); mixed precision, physical pair decode. The engine admits them by
    device footprint into batch slots (grouped by structure) or solo solvers, each on its own
    stream. The timed region is the engine's run from the first arrival to the last result."""
    from mm_propagation_amd import _capi, configs
    from mm_propagation_amd.serving import ServingEngine, serving_requests
    from mm_propagation_amd.solver import Context, MultimodePropagationGPU_RamanON

    # warm-up: one small solve per engine path (module load, first graph instantiation)
    warm = serving_requests(4, 1e6, seed=1000 + args.serving_seed, modes=(3,), nts=(4096, 2048), steps=(32, 64),
                            power_range=(10.0, 100.0))
    eng = ServingEngine(slots=args.slots)
    eng.run(warm)
    eng.close()

    reqs = serving_requests(args.requests, args.rate, seed=args.serving_seed)
    eng = ServingEngine(slots=args.slots)
    stats = eng.run(reqs)
    eng.close()
    makespan = stats["makespan_s"]
    finite_reqs = [r for r in reqs if not r.nonfinite]
    # capacity: the same requests all submitted at once (the server's rate when the arrivals do
    # not hold it back), counted over the same finite requests
    from mm_propagation_amd.serving import Request
    burst = [Request(rid=r.rid, params=r.params, n_steps=r.n_steps, arrival=0.0, raman=r.raman) for r in reqs]
    eng = ServingEngine(slots=args.slots)
    burst_stats = eng.run(burst)
    eng.close()
    burst_finite = {r.rid for r in burst if not r.nonfinite}
    # throughput counts the requests whose field stays finite (the others are parked by the device
    # when they overflow and end at the next chunk boundary)
    units = sum(r.units for r in finite_reqs)
    value = units / makespan
    steps = sum(r.n_steps for r in reqs)

    # the same finite requests solved one after another by the plain FP64 single-fiber solver: its
    # per-step time for each shape, on a request of that shape that stays finite over the window,
    # times each request's steps
    ctx = Context(0)
    per_step = {}
    plain_time = 0.0
    for r in finite_reqs:
        key = (r.params.n_modes, r.params.nt)
        if key not in per_step:
            s1 = MultimodePropagationGPU_RamanON(r.params, pair_decode=_capi.MMP_PAIRS_PHYSICAL, ctx=ctx)
            s1.advance(16)
            ctx.synchronize()
            t1 = time.perf_counter()
            s1.advance(48)
            ctx.synchronize()
            per_step[key] = (time.perf_counter() - t1) / 48
            s1.close()
        plain_time += per_step[key] * r.n_steps

    # accuracy: the three cheapest finite requests (M * NT * steps) against the FP64 oracle on the
    # host, with the growth an FP64 rounding-level perturbation gets over the same steps (the
    # oracle with its other FFT ordering) for scale: these are whole solves of up to thousands of
    # strongly nonlinear steps
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_binding as ob
    acc = {}
    threads = min(16, os.cpu_count() or 1)
    for r in sorted(finite_reqs, key=lambda r: r.units)[:3]:
        refs = []
        for variant in (0, 1):
            o = ob.Oracle(r.params, pair_decode=_capi.MMP_PAIRS_PHYSICAL, n_threads=threads, fft_variant=variant)
            o.compute_lawson_rk(r.n_steps)
            refs.append(o.get_result())
            o.close()
        err = configs.rel_l2_per_mode(r.result, refs[0])
        e_ref = float(np.sum(np.abs(refs[0]) ** 2))
        e_got = float(np.sum(np.abs(r.result) ** 2))
        acc[f"req{r.rid}_M{r.params.n_modes}_NT{r.params.nt}_steps{r.n_steps}"] = {
            "rel_l2_max": float(err.max()), "energy_drift": (e_got - e_ref) / e_ref,
            "fp64_dit_vs_dif": float(configs.rel_l2_per_mode(refs[1], refs[0]).max())}
    ctx.close()

    # aggregate HBM rate: algorithmic bytes of every request's steps over the makespan
    bytes_total = 0.0
    for r in reqs:
        info = eng.plan_for(r)
        bytes_total += rk4_bytes_per_step(r.params.n_modes, r.params.nt, int(info["n_pairs"]), r.raman, 8) * r.n_steps
    achieved = bytes_total / makespan / 1e9
    placements = {}
    for r in reqs:
        kind = r.placement.split()[0]
        placements[kind] = placements.get(kind, 0) + 1
    out = {
        "metric": METRIC, "value": value, "unit": "M·NT·steps/s", "n_gpus": 1, "steps": steps,
        "warmup": len(warm), "ms_per_step": makespan * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"serving{args.requests}: Poisson arrivals at {args.rate}/s (seed {args.serving_seed}), "
                               "M in {1,3,6,10,15}, NT in {4096..65536}, n_steps log-uniform [500, 20000], "
                               "L U(0.2,1) m, T0 U(0.05,0.3) ps, P log-uniform [10 W, 10 kW], symmetric GRIN "
                               f"coupling, RK4, Raman on, mixed precision, {args.slots} slots per batch group",
                   "requests": args.requests, "nt": "4096-65536", "devices": "one GPU"},
        "latency_s": {"p50": stats["p50_latency_s"], "p99": stats["p99_latency_s"],
                      "max": max(r.latency for r in reqs)},
        "makespan_s": makespan,
        "requests_nonfinite": stats["nonfinite"], "requests_finite": len(finite_reqs),
        "placements": placements, "batch_groups": stats["groups"],
        "peak_device_bytes": stats["peak_device_bytes"],
        "plain_fp64_sequential": {"value": units / plain_time, "seconds": plain_time,
                                  "ratio": value / (units / plain_time),
                                  "note": "the finite requests solved one after another by the plain FP64 solver, "
                                          "with no wait for their arrivals"},
        "capacity_all_at_once": {"value": units / burst_stats["makespan_s"], "makespan_s": burst_stats["makespan_s"],
                                 "ratio_to_plain_fp64_sequential": plain_time / burst_stats["makespan_s"],
                                 "same_finite_requests": burst_finite == {r.rid for r in finite_reqs},
                                 "note": "the same requests submitted together at t = 0"},
        "accuracy_vs_fp64_oracle": acc,
        "roofline": {"bound": "hbm", "kernel": "whole run (all kernels)", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None},
    }
    print(json.dumps(out), flush=True)
    return 0


def copy_calibration(mib: int) -> None:
    """One device-to-device copy of `mib` MiB after writing 512 MiB elsewhere (evicting the
    L2s), so FETCH_SIZE and WRITE_SIZE of that dispatch should each read `mib` MiB."""
    import torch
    n = mib * 1024 * 1024 // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    evict = torch.empty(512 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    evict.fill_(2.0)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
