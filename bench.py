"""Benchmark of the BaryTree interaction path on one MI355X.

Workload (the headline configuration): Coulomb kernel with singularity
subtraction (kappa = 0.5), N = 16M particles uniform in [-1,1]^3 (sources =
targets, separate arrays), charges uniform in [-1,1], weights 1,
cluster-cluster treecode, theta = 0.8, degree 10, max leaf 8000.

A step is one BaryTree_SolveLearnedDevice call with the particles resident
in HBM: device arrays in, potential (caller order) out on the device --
trees built on the GPU, interaction lists on the host, the per-pair degrees
the shipped degree model picks, the whole GPU path in mixed precision.
`value` is seconds per call (the model is loaded once, outside the timed
region).  Beside it:
  * the same call on host arrays (PCIe in and out, BaryTree_SolveLearned);
  * the plain FP64 time-to-solution (one call on HBM-resident particles,
    and BaryTree_Solve on host arrays, the drop-in path) and the speedups
    over them;
  * the GPU path alone with its inputs resident in HBM (one plan): mixed +
    learned, mixed at the fixed degree, plain FP64 with the symmetric near
    field, plain FP64 with one evaluation per ordered near-field pair (the
    reference's form);
  * every result's relative L2 error against the FP64 direct sum at 1000
    stride-sampled targets;
  * the roofline of the dominant kernel (the symmetric near field);
  * CPU baselines: the reference's own sources on one core (a sample at
    500k, extrapolated phase by phase) and the oracle on all host cores at
    the full size (interactions sampled);
  * `configs`: the other configurations (C1 Coulomb PC 100k against the
    reference on one core, C2 Yukawa CC 1M, C3 256 batched systems, C4
    learned degrees on a 4M Plummer sphere), each with its own timings,
    errors and baselines.

Prints one JSON line on stdout.  --no-configs skips the other
configurations, --no-cpu the CPU baselines.

--workload learned measures the learned degree selection on its own
configuration: regularized Coulomb (epsilon 0.01), N = 4M Plummer sphere
(scale 0.1, truncated at r = 1, seed 4), cluster-particle, theta 0.7, leaf
2000: fixed degree 10 against the per-pair degrees of the shipped model
(trained on other Plummer realizations), in FP64 and mixed.

--workload yukawa: Yukawa (kappa 0.5), N = 1M uniform, cluster-cluster,
theta 0.8, degree 8, leaf 4000, plain FP64 (symmetric and one-directional
near field) against mixed precision.

--workload cpu-case: Coulomb, N = 100k uniform, particle-cluster, theta 0.7,
degree 8, leaf 2000, skipping, FP64, against the reference's own sources
compiled for one CPU core and the oracle on all host cores.

--workload hermite: Coulomb with the Hermite approximation, particle-cluster
and cluster-particle, degree 5, against the reference's own sources on one
CPU core at N = 100k, and at N = 2M on the GPU.

--workload kernels: the seven kernels beside the Coulomb family at N = 1M on
the GPU and at N = 20k against the reference's own sources on one CPU core.

--workload batched measures the batching extension instead: 256 independent
Coulomb systems (N log-uniform in [1k, 64k], 8-Gaussian mixtures, sigma 0.05,
sources = targets), particle-cluster, theta 0.7, degree 6: the autotuned
batched solve (one launch per phase for all systems) against 256 sequential
plain single-system solves with the same leaf sizes.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_EVAL = 17  # Coulomb-SS, operations as written in coulomb_ss_pp.c:40-47
PEAK = {"fp32": 157.3, "fp64": 78.6}  # MI355X vector TFLOP/s
# FP64 operations of one pair as the reference's formulas write them (sqrt,
# division counted as one each): Coulomb near field, coulomb_pp.c (3
# differences, 5 for r^2, sqrt, q/r, the sum); Hermite particle-cluster,
# coulomb_pc.c:111-126 (42); Hermite cluster-particle, coulomb_cp.c:127-146
# (36)
FLOPS_PP_COULOMB, FLOPS_HERM_PC, FLOPS_HERM_CP = 11, 42, 36


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_uniform(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-1.0, 1.0, n) for _ in range(3))
    q = rng.uniform(-1.0, 1.0, n)
    w = np.ones(n)
    return x, y, z, q, w


def rel_l2(ref, val):
    return float(np.sqrt(np.sum((ref - val) ** 2) / np.sum(ref ** 2)))


def time_steps(plan, precision, steps, warmup):
    import torch
    plan.set_precision(precision)
    for _ in range(warmup):
        plan.execute(download=False)
    torch.cuda.synchronize()
    times, near, sym, far = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute(download=False)  # synchronizes on its own stream
        st = plan.stats()
        times.append(st["gpu_total_s"])
        near.append(st["gpu_near_s"])
        sym.append(st["gpu_near_sym_s"])
        far.append(st["gpu_far_s"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st["gpu_near_sym_mean_s"] = float(np.mean(sym))
    st["gpu_far_mean_s"] = float(np.mean(far))
    return wall / steps, float(np.mean(times)), float(np.mean(near)), st


PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_k_near_sym_coulomb_ss_cc_n16m.md")


def near_traffic():
    """HBM bytes per launch of the symmetric near-field kernel at the
    headline size, from the committed rocprofv3 summary (FETCH_SIZE and
    WRITE_SIZE, each scaled by the same counter on a copy of known size);
    None if the summary is absent."""
    path = PMC_SUMMARY
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def near_valu_mix():
    """Per-kind VALU wave-instruction counts of one launch of the symmetric
    near-field kernel at the headline size, from the committed rocprofv3
    summary (SQ_INSTS_VALU_* counters); None if the summary is absent."""
    path = PMC_SUMMARY
    try:
        for line in open(path):
            if line.startswith("valu_mix_per_launch:"):
                return json.loads(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


# SQ_INSTS_VALU_* counter -> issue-cost probe kind (barytree_amd.probe)
MIX_COST = {"SQ_INSTS_VALU_FMA_F64": "fma_f64", "SQ_INSTS_VALU_ADD_F64": "add_f64",
            "SQ_INSTS_VALU_MUL_F64": "fma_f64", "SQ_INSTS_VALU_FMA_F32": "fma_f32",
            "SQ_INSTS_VALU_ADD_F32": "mul_f32", "SQ_INSTS_VALU_MUL_F32": "mul_f32",
            "SQ_INSTS_VALU_TRANS_F32": "trans_f32", "SQ_INSTS_VALU_CVT": "cvt_f64_f32",
            "other": "fma_f32"}


def issue_model(mix, probe, simds):
    """Seconds the VALU needs to issue a launch's instruction mix, each kind
    at its cost measured alone: the sum over kinds of wave instructions x ns
    per wave instruction and SIMD, over the SIMDs of the GPU.
    Transcendentals are priced at the mean of exp2 and rsq; instructions
    outside the counted kinds at an FP32 fma.  Kinds issued side by side can
    overlap, so a kernel may run below this."""
    cost = dict(probe)
    cost["trans_f32"] = 0.5 * (probe["exp_f32"] + probe["rsq_f32"])
    ns = sum(mix[k] * cost[MIX_COST[k]] for k in MIX_COST if k in mix)
    return ns * 1e-9 / simds


def bench_batched(args):
    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt
    from barytree_amd.workloads import batched_sizes, gaussian_mixture

    nsys = args.systems
    sizes = batched_sizes(nsys, 3, nmin=1000, nmax=64000)
    systems = []
    for i, n in enumerate(sizes):
        x, y, z, q, w = gaussian_mixture(int(n), 3000 + i)
        systems.append(((x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w)))
    total = int(sizes.sum())
    log(f"[bench] batched: {nsys} systems, {total} particles")
    kargs = (bt.Kernel.COULOMB, [0.0], bt.Singularity.SKIPPING, bt.Approximation.LAGRANGE,
             bt.ComputeType.PARTICLE_CLUSTER, 0.7, 6)
    t0 = time.perf_counter()
    sc = bt.autotune(systems, *kargs)
    tune_s = time.perf_counter() - t0
    log(f"[bench] autotune {tune_s:.2f}s: {sc.as_dict()}")
    leaf_of = [sc.leaf[int(np.floor(np.log2(n)))] for n in sizes]

    # one-shot batched solve, end to end (host trees and lists, upload, GPU,
    # download) on the concatenated arrays BaryTree_SolveBatched takes
    batch = bt.Batch(systems)

    def batched_once(precision):
        t0 = time.perf_counter()
        out = bt.solve_batched(batch, *kargs, schedule=sc, precision=precision)
        return time.perf_counter() - t0, out, bt.last_stats()
    batched_once(bt.PRECISION_FP64)  # warm-up
    e2e = [batched_once(bt.PRECISION_FP64) for _ in range(args.steps)]
    e2e_fp64 = min(r[0] for r in e2e)
    out_fp64, st_b = e2e[-1][1], e2e[-1][2]
    e2e_mixed_runs = [batched_once(bt.PRECISION_MIXED) for _ in range(args.steps)]
    e2e_mixed = min(r[0] for r in e2e_mixed_runs)
    out_mixed = e2e_mixed_runs[-1][1]
    phases = {k: round(st_b[k], 4) for k in ("host_copy_s", "host_tree_s", "host_lists_s", "host_work_s",
                                             "h2d_s", "gpu_total_s")}
    log(f"[bench] batched solve fp64 {e2e_fp64:.3f}s mixed {e2e_mixed:.3f}s (phases {phases})")

    # 256 sequential plain solves, same leaf sizes
    def sequential_once():
        t0 = time.perf_counter()
        outs, gpu = [], 0.0
        for (t, s), lf in zip(systems, leaf_of):
            outs.append(bt.solve(t, s, *kargs, lf, lf))
            gpu += bt.last_stats()["gpu_total_s"]
        return time.perf_counter() - t0, outs, gpu
    sequential_once()  # warm-up
    seq_runs = [sequential_once() for _ in range(max(1, args.steps - 1))]
    seq = min(r[0] for r in seq_runs)
    seq_gpu = min(r[2] for r in seq_runs)
    outs_seq = seq_runs[-1][1]
    # single-system solves take the symmetric near field (targets are the
    # sources), batched plans do not: the sums' order differs
    diff_seq = rel_l2(np.concatenate(outs_seq), np.concatenate(out_fp64))
    log(f"[bench] sequential {seq:.3f}s (sum of gpu phases {seq_gpu:.4f}s), batched vs sequential "
        f"relL2 {diff_seq:.2e}")

    # resident-input GPU time: batched plan vs per-system plans
    plan = bt.Plan.batched(systems, *kargs, schedule=sc)
    wall_b, gpu_b, near_b, st = time_steps(plan, bt.PRECISION_FP64, args.steps, args.warmup)
    off = plan.offsets
    idx = np.linspace(0, total - 1, 1000).astype(np.int64)
    direct = plan.direct_sample(idx)
    _, gpu_bm, _, _ = time_steps(plan, bt.PRECISION_MIXED, args.steps, args.warmup)
    plan.close()
    err64 = rel_l2(direct, np.concatenate(out_fp64)[idx])
    err32 = rel_l2(direct, np.concatenate(out_mixed)[idx])
    log(f"[bench] resident batched plan gpu fp64 {gpu_b:.4f}s mixed {gpu_bm:.4f}s; "
        f"relL2 fp64 {err64:.2e} mixed {err32:.2e}")
    evals = st["pp_ptwise"] + st["pc_ptwise"]
    line = {
        "metric": "batched time-to-solution, 256 systems (vs 256 sequential plain solves)",
        "value": round(e2e_fp64, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(e2e_fp64 * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{nsys} Coulomb systems, N log-uniform [1k,64k] (total {total}), "
                         "8-Gaussian mixtures sigma=0.05, particle-cluster theta=0.7 degree 6"),
            "systems": nsys, "particles": total, "compute_type": "PC", "theta": 0.7, "degree": 6,
        },
        "schedule": sc.as_dict(),
        "autotune_s": round(tune_s, 3),
        "sequential_s": round(seq, 4),
        "speedup_vs_sequential": round(seq / e2e_fp64, 3),
        "batched_mixed_s": round(e2e_mixed, 4),
        "time_to_solution_phases_s": phases,
        "rel_l2_batched_vs_sequential": diff_seq,
        "gpu_batched_fp64_s": round(gpu_b, 5),
        "gpu_batched_mixed_s": round(gpu_bm, 5),
        "gpu_sequential_sum_s": round(seq_gpu, 5),
        "rel_l2_err_fp64": err64,
        "rel_l2_err_mixed": err32,
        "interactions": evals,
        "interactions_per_s_gpu_fp64": evals / gpu_b,
    }
    return line


def bench_learned(args):
    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt
    from barytree_amd import learned

    cfg = learned.MODELS["regularized_coulomb_cp"]
    n = args.n if args.n != 16_000_000 else 4_000_000
    log(f"[bench] learned: Plummer N={n}")
    pts = learned.particles("plummer", n, 4)
    t0 = time.perf_counter()
    plan = learned.plan_for(cfg, *pts)
    host_setup = time.perf_counter() - t0
    model = learned.load_model("regularized_coulomb_cp")
    idx = np.arange(0, n, n // 1000)[:1000]
    direct = plan.direct_sample(idx)
    res = {}
    for form, prec, use in (("fixed_fp64", bt.PRECISION_FP64, False),
                            ("fixed_mixed", bt.PRECISION_MIXED, False),
                            ("learned_fp64", bt.PRECISION_FP64, True),
                            ("learned_mixed", bt.PRECISION_MIXED, True)):
        t0 = time.perf_counter()
        if use:
            plan.use_model(model, args.delta)
        else:
            plan.set_pair_degrees(None)
        setup = time.perf_counter() - t0
        wall, gpu, near, st = time_steps(plan, prec, args.steps, args.warmup)
        out = plan.execute()
        err = rel_l2(direct, out[idx])
        far = st["far_ptwise_variable"] if use else st["cp_ptwise"]
        res[form] = {"s": round(wall, 5), "gpu_s": round(gpu, 5), "rel_l2_err": err,
                     "far_evals": far, "degree_setup_s": round(setup, 4),
                     "mean_pair_degree": st["mean_pair_degree"] if use else 10.0,
                     "phases_s": {k: round(st[k], 5) for k in ("gpu_upward_s", "gpu_near_s",
                                                               "gpu_far_s", "gpu_down_s")},
                     "near_symmetric_launch_s": round(st["gpu_near_sym_mean_s"], 5)}
        log(f"[bench] {form}: {wall:.4f} s/step, relL2 {err:.2e}, far evals {far:.3e}, "
            f"phases {res[form]['phases_s']}, symmetric near launch {st['gpu_near_sym_mean_s']:.4f}")
    plan.close()
    line = {
        "metric": "time-to-solution, learned per-pair degree vs fixed degree 10 (C4)",
        "value": res["learned_mixed"]["s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["learned_mixed"]["s"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 pair kernel, f64 charges/accumulation",
        "data": "synthetic",
        "config": {"workload": (f"regularized Coulomb eps=0.01, N={n} Plummer (a=0.1, r<=1), "
                                "cluster-particle theta=0.7, degree 10, leaf 2000"),
                   "n": n, "compute_type": "CP", "theta": 0.7, "degree": 10, "max_leaf": 2000},
        "forms": res,
        "phases_note": ("the far field, upward pass and downpass run on a second stream beside the symmetric "
                        "near-field launch: their times are elapsed times while sharing the GPU with it"),
        "delta": args.delta,
        "speedup_learned_vs_fixed_fp64": round(res["fixed_fp64"]["s"] / res["learned_fp64"]["s"], 3),
        "speedup_learned_mixed_vs_fixed_fp64": round(res["fixed_fp64"]["s"] / res["learned_mixed"]["s"], 3),
        "host_setup_s": round(host_setup, 3),
    }
    return line


def bench_yukawa(args):
    """Yukawa kappa = 0.5, N = 1M uniform in [-1,1]^3 (seed 2), cluster-cluster,
    theta 0.8, degree 8, max leaf 4000: the plain FP64 form against the mixed
    form (FP32 pair kernel, FP64 charges and accumulation), each checked
    against the FP64 direct sum at 1000 stride-sampled targets."""
    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt

    n = args.n if args.n != 16_000_000 else 1_000_000
    x, y, z, q, w = gen_uniform(n, 2)
    targets = (x.copy(), y.copy(), z.copy(), q.copy())
    kargs = (bt.Kernel.YUKAWA, [0.5], bt.Singularity.SKIPPING, bt.Approximation.LAGRANGE,
             bt.ComputeType.CLUSTER_CLUSTER, 0.8, 8, 4000, 4000)
    t0 = time.perf_counter()
    plan = bt.Plan(targets, (x, y, z, q, w), *kargs, precision=bt.PRECISION_MIXED)
    host_setup = time.perf_counter() - t0
    idx = np.arange(0, n, n // 1000)[:1000]
    direct = plan.direct_sample(idx)
    res = {}
    for form, prec, sym in (("fp64_one_directional", bt.PRECISION_FP64, False),
                            ("fp64", bt.PRECISION_FP64, True), ("mixed", bt.PRECISION_MIXED, True)):
        plan.set_symmetric(sym)
        wall, gpu, near, st = time_steps(plan, prec, args.steps, args.warmup)
        out = plan.execute()
        err = rel_l2(direct, out[idx])
        res[form] = {"s": round(wall, 5), "gpu_s": round(gpu, 5), "near_s": round(near, 5),
                     "rel_l2_err": err,
                     "phases_s": {k: round(st[k], 5) for k in ("gpu_upward_s", "gpu_near_s",
                                                               "gpu_far_s", "gpu_down_s")}}
        log(f"[bench] yukawa {form}: {wall:.4f} s/step (near {near:.4f}), relL2 {err:.2e}")
    plan.close()
    near_evals = st["pp_ptwise"] + st["pc_ptwise"]
    evals = near_evals + st["cp_ptwise"] + st["cc_ptwise"]
    # dominant kernel: the symmetric near field (targets are the sources);
    # 14 flops per ordered Yukawa pair evaluation it delivers (yukawa_pp.c)
    sym_s = st["gpu_near_sym_mean_s"]
    achieved = 14 * st["sym_ptwise"] / sym_s / 1e12 if sym_s > 0 else 0.0
    line = {
        "metric": "time per solve, Yukawa CC N=1M, plain FP64 vs mixed precision",
        "value": res["mixed"]["s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["mixed"]["s"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 pair kernel, f64 charges/accumulation",
        "data": "synthetic",
        "config": {"workload": (f"Yukawa kappa=0.5, N={n} uniform [-1,1]^3, cluster-cluster "
                                "theta=0.8 degree 8 max leaf 4000"),
                   "n": n, "compute_type": "CC", "theta": 0.8, "degree": 8, "max_leaf": 4000},
        "forms": res,
        "speedup_mixed_vs_fp64": round(res["fp64"]["s"] / res["mixed"]["s"], 3),
        "speedup_mixed_vs_fp64_one_directional": round(res["fp64_one_directional"]["s"] / res["mixed"]["s"], 3),
        "interactions": evals,
        "roofline": {"bound": "valu", "kernel": "k_near_sym<Yukawa, skipping> (symmetric near field, mixed)",
                     "achieved": round(achieved, 2), "peak": PEAK["fp32"], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK["fp32"], 4),
                     "frac_executed": round(achieved * st["sym_evals"] / max(st["sym_ptwise"], 1) / PEAK["fp32"], 4),
                     "traffic": None, "kernel_s": round(sym_s, 5), "ordered_pairs": st["sym_ptwise"],
                     "executed_pair_evaluations": st["sym_evals"]},
        "host_setup_s": round(host_setup, 3),
    }
    return line


def bench_cpu_case(args):
    """Coulomb, N = 100k uniform in [-1,1]^3 (seed 1), particle-cluster,
    theta 0.7, degree 8, max leaf 2000, skipping, FP64: the GPU plain form
    against the reference's own sources compiled for the CPU (oracle/_ref,
    one core; the oracle's restatement when that library is absent), on the
    same particles, with their difference and the error against the FP64
    direct sum at 1000 stride-sampled targets."""
    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt
    from oracle import oracle as O

    n = args.n if args.n != 16_000_000 else 100_000
    x, y, z, q, w = gen_uniform(n, 1)
    targets = (x.copy(), y.copy(), z.copy(), q.copy())
    kargs = (bt.Kernel.COULOMB, [0.0], bt.Singularity.SKIPPING, bt.Approximation.LAGRANGE,
             bt.ComputeType.PARTICLE_CLUSTER, 0.7, 8, 2000, 2000)
    bt.solve(targets, (x, y, z, q, w), *kargs)  # warm-up (module load, first allocations)
    t0 = time.perf_counter()
    out_gpu = bt.solve(targets, (x, y, z, q, w), *kargs)
    e2e = time.perf_counter() - t0
    plan = bt.Plan(targets, (x, y, z, q, w), *kargs)
    wall, gpu, near, st = time_steps(plan, bt.PRECISION_FP64, args.steps, args.warmup)
    idx = np.arange(0, n, n // 1000)[:1000]
    direct = plan.direct_sample(idx)
    plan.close()
    kind = "reference" if O.ref_available() else "port"
    fn = O.ref_treecode if kind == "reference" else O.treecode
    t0 = time.perf_counter()
    out_cpu = fn(*targets, x, y, z, q, w, 1, [0.0], 1, 1, 1, 0.7, 8, 2000, 2000)
    cpu_s = time.perf_counter() - t0
    diff = rel_l2(out_cpu, out_gpu)
    log(f"[bench] C1 gpu {wall:.4f} s/step (end to end {e2e:.3f} s), cpu {kind} {cpu_s:.2f} s, "
        f"gpu vs cpu relL2 {diff:.2e}, vs direct {rel_l2(direct, out_gpu[idx]):.2e}")
    threads = host_threads()
    allc = cpu_run("port", (x, y, z, q, w), (1, [0.0], 1, 1, 1, 0.7, 8, 2000, 2000), 1, threads)
    log(f"[bench] C1 oracle on {threads} threads {allc['wall']:.2f} s")
    # the beta auto-tune (beta = 0.5) with the leaf sizes of the reference's
    # CPU build and of its GPU build, one call each after a warm-up
    beta = {}
    bkargs = (bt.Kernel.COULOMB, [0.0], bt.Singularity.SKIPPING, bt.Approximation.LAGRANGE,
              bt.ComputeType.PARTICLE_CLUSTER, -1.0, -1, 1, 1)
    try:
        for name, which in (("cpu_build_leaves", bt.BETA_LEAVES_CPU), ("gpu_build_leaves", bt.BETA_LEAVES_GPU)):
            bt.set_beta_leaf_sizes(which)
            bt.solve(targets, (x, y, z, q, w), *bkargs, beta=0.5)
            t0 = time.perf_counter()
            out_b = bt.solve(targets, (x, y, z, q, w), *bkargs, beta=0.5)
            dt = time.perf_counter() - t0
            sb = bt.last_stats()
            beta[name] = {"s": round(dt, 4), "theta": sb["theta"], "degree": sb["interp_degree"],
                          "leaves": [sb["max_per_source_leaf"], sb["max_per_target_leaf"]],
                          "rel_l2_err": rel_l2(direct, out_b[idx])}
            log(f"[bench] C1 beta 0.5 {name}: {dt:.4f} s, {beta[name]}")
    finally:
        bt.set_beta_leaf_sizes(bt.BETA_LEAVES_CPU)
    line = {
        "metric": "time per solve, Coulomb PC N=100k FP64, GPU vs the reference on one CPU core",
        "value": round(wall, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"Coulomb, N={n} uniform [-1,1]^3, particle-cluster theta=0.7 "
                                "degree 8 max leaf 2000, skipping"),
                   "n": n, "compute_type": "PC", "theta": 0.7, "degree": 8, "max_leaf": 2000},
        "end_to_end_s": round(e2e, 4),
        "rel_l2_err": rel_l2(direct, out_gpu[idx]),
        "rel_l2_gpu_vs_cpu": diff,
        "cpu_baseline": {"value": round(cpu_s, 3), "unit": "s", "cores": 1, "kind": kind,
                         "sample": f"the whole treecode on the same {n} particles, one call",
                         "all_cores": {"value": round(allc["wall"], 3), "unit": "s", "cores": threads,
                                       "kind": "port", "sample": "the oracle on the same particles, one call",
                                       "phases_s": {k: round(v, 4) for k, v in allc["phases"].items()}},
                         "host": host_cpu()},
        "speedup_vs_cpu": round(cpu_s / wall, 1),
        "speedup_end_to_end_vs_cpu": round(cpu_s / e2e, 1),
        "beta_0_5_one_call": beta,
    }
    return line


def bench_hermite(args):
    """Hermite approximation, Coulomb with skipping, uniform in [-1,1]^3
    (seed 7), theta 0.7, degree 5, leaf 2000, FP64: particle-cluster and
    cluster-particle at N = 100k against the reference's own sources on one
    CPU core (oracle/_ref; particle-cluster), and at N = 2M on the GPU alone,
    each checked against the FP64 direct sum at 1000 stride-sampled targets."""
    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt
    from oracle import oracle as O

    res = {}
    for n in (100_000, 2_000_000):
        x, y, z, q, w = gen_uniform(n, 7)
        targets = (x.copy(), y.copy(), z.copy(), q.copy())
        idx = np.arange(0, n, n // 1000)[:1000]
        for name, ct in (("pc", bt.ComputeType.PARTICLE_CLUSTER), ("cp", bt.ComputeType.CLUSTER_PARTICLE)):
            kargs = (bt.Kernel.COULOMB, [0.5], bt.Singularity.SKIPPING, bt.Approximation.HERMITE, ct,
                     0.7, 5, 2000, 2000)
            plan = bt.Plan(targets, (x, y, z, q, w), *kargs)
            wall, gpu, near, st = time_steps(plan, bt.PRECISION_FP64, args.steps, args.warmup)
            out = plan.execute()
            err = rel_l2(plan.direct_sample(idx), out[idx])
            plan.close()
            res[f"{name}_n{n}"] = {"s": round(wall, 5), "rel_l2_err": err, "out": out,
                                   "near_s": round(near, 6), "far_s": round(st["gpu_far_mean_s"], 6),
                                   "pp_ptwise": st["pp_ptwise"], "far_ptwise": st[f"{name}_ptwise"]}
            log(f"[bench] hermite {name} N={n}: {wall:.4f} s/step (near launch {near:.4f} s, "
                f"Hermite far-field launch {st['gpu_far_mean_s']:.4f} s), relL2 {err:.2e}")
    n = 100_000
    x, y, z, q, w = gen_uniform(n, 7)
    kind = "reference" if O.ref_available() else "port"
    fn = O.ref_treecode if kind == "reference" else O.treecode
    t0 = time.perf_counter()
    ref = fn(x.copy(), y.copy(), z.copy(), q.copy(), x, y, z, q, w, 1, [0.5], 1, 2, 1, 0.7, 5, 2000, 2000)
    cpu_s = time.perf_counter() - t0
    diff = rel_l2(ref, res["pc_n100000"]["out"])
    log(f"[bench] hermite pc N={n}: cpu {kind} {cpu_s:.2f} s, gpu vs cpu relL2 {diff:.2e}")
    for v in res.values():
        v.pop("out")
    gpu_s = res["pc_n100000"]["s"]

    def roof(name, kernel, flops, pairs, secs):
        achieved = flops * pairs / secs / 1e12
        return {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 2), "peak": PEAK["fp64"],
                "unit": "TFLOP/s", "frac": round(achieved / PEAK["fp64"], 4), "traffic": None,
                "kernel_s": round(secs, 6), "pairs": pairs, "flops_per_pair": flops, "run": name}

    pc, cp = res["pc_n100000"], res["cp_n2000000"]
    # the metric's workload spends most of its time in the FP64 near-field
    # launch; the Hermite far-field launches follow at N = 2M
    roofline = roof("pc_n100000", "k_interact<Coulomb, skipping, particle targets, FP64> (near field)",
                    FLOPS_PP_COULOMB, pc["pp_ptwise"], pc["near_s"])
    herm = [roof("pc_n2000000", "k_herm_pc<Coulomb, skipping>", FLOPS_HERM_PC,
                 res["pc_n2000000"]["far_ptwise"], res["pc_n2000000"]["far_s"]),
            roof("cp_n2000000", "k_herm_cp<Coulomb>", FLOPS_HERM_CP, cp["far_ptwise"], cp["far_s"])]
    line = {
        "metric": "time per solve, Coulomb Hermite particle-cluster N=100k FP64, GPU vs the reference on one CPU core",
        "value": gpu_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(gpu_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Coulomb skipping, Hermite, N=100k (and 2M) uniform [-1,1]^3, theta=0.7 degree 5 leaf 2000",
                   "n": n, "compute_type": "PC", "theta": 0.7, "degree": 5, "max_leaf": 2000},
        "forms": res,
        "rel_l2_gpu_vs_cpu": diff,
        "roofline": roofline,
        "roofline_hermite": herm,
        "cpu_baseline": {"value": round(cpu_s, 3), "unit": "s", "cores": 1, "kind": kind,
                         "sample": f"the whole Hermite treecode (particle-cluster) on the same {n} particles, one call"},
        "speedup_vs_cpu": round(cpu_s / gpu_s, 1),
    }
    return line


# the kernels beside the Coulomb family: id -> (name, kernel parameters)
OTHER_KERNELS = {4: ("regularized_yukawa", [0.5, 0.1]), 5: ("atan", [2.0, 0.3]), 8: ("sin_over_r", [1.5]),
                 9: ("mq", [2.0, 0.3]), 10: ("rbs_u", [0.1]), 11: ("rbs_v", [0.1]), 12: ("user", [0.5])}


def bench_kernels(args):
    """The kernels beside the Coulomb family, particle-cluster with skipping,
    uniform in [-1,1]^3 (seed 8), theta 0.7, degree 5, leaf 2000, FP64: each
    at N = 1M on the GPU against the FP64 direct sum at 1000 stride-sampled
    targets, and at N = 20k against the reference's own sources on one CPU
    core (oracle/_ref)."""
    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt
    from oracle import oracle as O

    kind = "reference" if O.ref_available() else "port"
    fn = O.ref_treecode if kind == "reference" else O.treecode
    res = {}
    big, small = 1_000_000, 20_000
    xb = gen_uniform(big, 8)
    xs = gen_uniform(small, 8)
    idx = np.arange(0, big, big // 1000)[:1000]
    cpu_total = 0.0
    for k, (name, kp) in OTHER_KERNELS.items():
        kargs = (k, kp, 1, 1, 1, 0.7, 5, 2000, 2000)
        x, y, z, q, w = xb
        plan = bt.Plan((x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w), *kargs)
        wall, gpu, near, st = time_steps(plan, bt.PRECISION_FP64, args.steps, args.warmup)
        out = plan.execute()
        err = rel_l2(plan.direct_sample(idx), out[idx])
        plan.close()
        x, y, z, q, w = xs
        t = (x.copy(), y.copy(), z.copy(), q.copy())
        got = bt.solve(t, (x, y, z, q, w), *kargs)
        t0 = time.perf_counter()
        ref = fn(*t, x, y, z, q, w, *kargs)
        cpu_s = time.perf_counter() - t0
        cpu_total += cpu_s
        res[name] = {"s_n1m": round(wall, 5), "rel_l2_err_n1m": err, "cpu_s_n20k": round(cpu_s, 3),
                     "rel_l2_gpu_vs_cpu_n20k": rel_l2(ref, got)}
        log(f"[bench] {name}: N=1M {wall:.4f} s/step relL2 {err:.2e}; N=20k cpu {kind} {cpu_s:.2f} s, "
            f"gpu vs cpu {res[name]['rel_l2_gpu_vs_cpu_n20k']:.2e}")
    total = round(sum(v["s_n1m"] for v in res.values()), 5)
    line = {
        "metric": "time per solve of the seven kernels beside the Coulomb family, N=1M PC FP64 (sum)",
        "value": total,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "regularized Yukawa, atan, sin-over-r, MQ, RBS-u, RBS-v, user example; "
                               "N=1M uniform [-1,1]^3, particle-cluster theta=0.7 degree 5 leaf 2000, skipping",
                   "n": big, "compute_type": "PC", "theta": 0.7, "degree": 5, "max_leaf": 2000},
        "kernels": res,
        "cpu_baseline": {"value": round(cpu_total, 3), "unit": "s", "cores": 1, "kind": kind,
                         "sample": f"the whole treecode of each kernel on {small} particles, one call each (sum)"},
    }
    return line


def host_threads():
    """Threads for the all-core CPU runs: OMP_NUM_THREADS when set (the
    host's share of cores), else the core count, at most 16."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return max(1, min(16, n if n > 0 else (os.cpu_count() or 1)))


def host_cpu():
    """Core count and CPU model of this host (nproc, lscpu)."""
    info = {"nproc": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        pass
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return info


def cpu_run(kind, pts, kargs, stride=1, threads=1):
    """One CPU treecode call on (x, y, z, q, w), targets = sources: kind
    "reference" runs the reference's own sources (oracle/_ref), "port" the
    oracle.  With stride k only every k-th key's interaction lists are
    computed (the trees, clusters, lists and downpass run in full).  Returns
    the wall time, the phase times and the point-wise evaluation counts
    (computed and full)."""
    from oracle import oracle as O
    x, y, z, q, w = pts
    t = (x.copy(), y.copy(), z.copy(), q.copy())
    O.set_threads(threads)
    t0 = time.perf_counter()
    if kind == "reference":
        O.ref_set_sample_stride(stride)
        O.ref_treecode(*t, x, y, z, q, w, *kargs)
        wall = time.perf_counter() - t0
        O.ref_set_sample_stride(1)
        r = O.ref_last_timing()
        done, total = r["near_done"] + r["far_done"], r["near_total"] + r["far_total"]
        nodes = None
    else:
        O.set_sample_stride(stride)
        O.treecode(*t, x, y, z, q, w, *kargs)
        wall = time.perf_counter() - t0
        O.set_sample_stride(1)
        r = O.last_stats()
        done = r["near_done"] + r["far_done"]
        total = r["pp_ptwise"] + r["pc_ptwise"] + r["cp_ptwise"] + r["cc_ptwise"]
        nodes = r["source_nodes"] + r["target_nodes"]
    O.set_threads(O.max_threads())
    phases = {k: r[k] for k in ("t_setup", "t_upward", "t_lists", "t_compute", "t_downpass", "t_finish")}
    return {"wall": wall, "phases": phases, "evals_done": done, "evals_total": total, "nodes": nodes}


def same_n_estimate(r):
    """Seconds of the whole treecode from a sampled run at the same size: every
    phase as measured, the interaction phase scaled by evaluations."""
    ph = r["phases"]
    rest = sum(v for k, v in ph.items() if k != "t_compute")
    return rest + ph["t_compute"] * r["evals_total"] / max(r["evals_done"], 1)


def cpu_baseline_large(n, kargs, st16, threads):
    """CPU baselines of the headline workload.
    reference: the reference's sources on one core at N_s = 250k (same
    parameters and generator), the interactions of one target leaf computed
    (the last node in preorder), the other phases in full; extrapolated to
    N = 16M phase by phase (interactions by the point-wise counts of the 16M
    plan, upward pass and downpass by N P + nodes P^2, trees by N log N,
    lists by nodes).
    port: the oracle on all host cores at the full N = 16M, the interactions
    of every (nodes/3)-th target node computed (the rest of the treecode in
    full), interactions scaled by evaluations.  calibration: the oracle over
    the reference, both on one core on the 250k sample."""
    import barytree_amd as bt
    from oracle import oracle as O
    ns = 250_000
    pts = gen_uniform(ns, 5)
    st_s = bt.analyze(pts[:4], pts, *kargs)
    stride = max(1, st_s["target_nodes"] - 1)  # the root (no lists) and the last leaf
    P = (kargs[6] + 1) ** 3
    ref = cpu_run("reference" if O.ref_available() else "port", pts, kargs, stride, 1)
    port1 = cpu_run("port", pts, kargs, stride, 1)
    nodes_s = port1["nodes"]
    nodes16 = st16["source_nodes"] + st16["target_nodes"]
    evals16 = st16["pp_ptwise"] + st16["pc_ptwise"] + st16["cp_ptwise"] + st16["cc_ptwise"]

    def extrapolate(r):
        ph = r["phases"]
        work = lambda nn, nodes: nn * P + nodes * P * P
        est = {
            "t_setup": ph["t_setup"] * (n * np.log(n)) / (ns * np.log(ns)),
            "t_upward": ph["t_upward"] * work(n, nodes16 / 2) / work(ns, nodes_s / 2),
            "t_lists": ph["t_lists"] * nodes16 / nodes_s,
            "t_compute": ph["t_compute"] * evals16 / max(r["evals_done"], 1),
            "t_downpass": ph["t_downpass"] * work(n, nodes16 / 2) / work(ns, nodes_s / 2),
            "t_finish": ph["t_finish"] * n / ns,
        }
        return est, sum(est.values())

    ref_est, ref_total = extrapolate(ref)
    port1_est, port1_total = extrapolate(port1)
    calib = port1["wall"] / ref["wall"]
    log(f"[bench] cpu: reference 1 core on N={ns} (stride {stride}) {ref['wall']:.1f} s -> "
        f"{ref_total:.0f} s at N={n} (extrapolated); port/reference {calib:.3f}")
    allc = None
    if n == 16_000_000 and threads > 1:
        big = gen_uniform(n, 5)
        stride16 = max(1, st16["target_nodes"] // 3)
        allc = cpu_run("port", big, kargs, stride16, threads)
        allc_total = same_n_estimate(allc)
        log(f"[bench] cpu: oracle on {threads} threads at N={n} (stride {stride16}) {allc['wall']:.1f} s -> "
            f"{allc_total:.0f} s (interactions extrapolated)")
        del big
    rate_ref = ref["evals_done"] / max(ref["phases"]["t_compute"], 1e-12)
    out = {
        "value": round(ref_total, 1),
        "unit": "s",
        "cores": 1,
        "kind": "reference" if O.ref_available() else "port",
        "sample": (f"the reference's own sources (oracle/_ref) on one core, N={ns} with the headline's "
                   f"parameters, interactions of the last of its {st_s['target_nodes']} target "
                   f"nodes ({ref['evals_done']:.3e} "
                   f"evaluations at {rate_ref:.3e}/s), the other phases in full ({ref['wall']:.1f} s in all); "
                   f"extrapolated to N={n}: interactions by the plan's {evals16:.3e} point-wise "
                   f"evaluations, upward pass and downpass by N P + nodes P^2, trees by N log N"),
        "extrapolated": True,
        "phases_s": {k: round(v, 2) for k, v in ref_est.items()},
        "sample_phases_s": {k: round(v, 3) for k, v in ref["phases"].items()},
        "calibration_port_over_reference": round(calib, 4),
        "port_one_core_extrapolated_s": round(port1_total, 1),
        "host": host_cpu(),
    }
    if allc is not None:
        out["all_cores"] = {
            "value": round(allc_total, 1), "unit": "s", "cores": threads, "kind": "port",
            "sample": (f"the oracle (OpenMP over targets and nodes, per-target sums in serial order) at "
                       f"N={n}, interactions of every {stride16}th target node ({allc['evals_done']:.3e} of "
                       f"{allc['evals_total']:.3e} evaluations), the other phases in full "
                       f"({allc['wall']:.1f} s in all)"),
            "phases_s": {k: round(v, 2) for k, v in allc["phases"].items()},
            "reference_equivalent_s": round(allc_total / calib, 1),
        }
    return out


def summary(line):
    """A workload's line without the keys every line repeats."""
    drop = ("metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline", "data",
            "ms_per_step")
    return {k: v for k, v in line.items() if k not in drop}


def bench_large(args):
    """The headline workload (docstring at the top of the file)."""
    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt
    from barytree_amd import learned

    n = args.n
    log(f"[bench] generating N={n}")
    x, y, z, q, w = gen_uniform(n, 5)
    targets = (x.copy(), y.copy(), z.copy(), q.copy())
    sources = (x, y, z, q, w)
    kargs = (bt.Kernel.COULOMB, [0.5], bt.Singularity.SUBTRACTION,
             bt.Approximation.LAGRANGE, bt.ComputeType.CLUSTER_CLUSTER, 0.8, 10, 8000, 8000)
    model = learned.load_model("coulomb_cc")  # loaded once, outside the timed region

    # 1. time-to-solution: one BaryTree_SolveLearnedDevice call per step with
    # the particles resident in HBM (device arrays in, potential out on the
    # device: trees on the GPU, lists on the host, degrees from the model,
    # the GPU path); the same call on host arrays (PCIe-inclusive) beside it
    dt = [torch.from_numpy(a).cuda() for a in targets]
    ds = [torch.from_numpy(a).cuda() for a in sources]
    dout = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def tts_device(model_, precision):
        bt.solve_device(dt, ds, dout, *kargs, model=model_, delta=args.delta, precision=precision)
        return bt.last_stats()
    for _ in range(args.warmup):
        tts_device(model, bt.PRECISION_MIXED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st_tts = tts_device(model, bt.PRECISION_MIXED)
    torch.cuda.synchronize()
    tts = (time.perf_counter() - t0) / args.steps
    out_tts = dout.cpu().numpy()
    host_s = sum(st_tts[k] for k in ("host_copy_s", "host_tree_s", "host_lists_s", "host_work_s", "h2d_s"))
    log(f"[bench] time-to-solution (mixed + learned, one call, inputs in HBM): {tts:.3f} s "
        f"(trees {st_tts['host_tree_s']:.3f} lists {st_tts['host_lists_s']:.3f} work {st_tts['host_work_s']:.3f} "
        f"copies {st_tts['h2d_s']:.3f} gpu {st_tts['gpu_total_s']:.3f})")
    t0 = time.perf_counter()
    out_host = bt.solve_learned(targets, sources, *kargs, model=model, delta=args.delta,
                                precision=bt.PRECISION_MIXED)
    tts_host = time.perf_counter() - t0
    host_equal = bool(np.array_equal(out_host, out_tts))
    log(f"[bench] the same call on host arrays (PCIe in and out): {tts_host:.3f} s (results equal: {host_equal})")

    # plain FP64 time-to-solution (the drop-in path's arithmetic)
    plain_tts = plain_tts_host = err_plain_tts = None
    if args.plain_steps > 0:
        ts = []
        for _ in range(args.plain_steps):
            t0 = time.perf_counter()
            tts_device(None, bt.PRECISION_FP64)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        plain_tts = float(np.mean(ts))
        out_plain_tts = dout.cpu().numpy()
        t0 = time.perf_counter()
        bt.solve(targets, sources, *kargs)
        plain_tts_host = time.perf_counter() - t0
        log(f"[bench] time-to-solution plain FP64 (one call, inputs in HBM): {plain_tts:.3f} s; "
            f"BaryTree_Solve on host arrays {plain_tts_host:.3f} s")
    del dt, ds, dout

    # 2. the GPU path alone, inputs resident in HBM (one plan, several forms)
    t0 = time.perf_counter()
    plan = bt.Plan(targets, sources, *kargs, precision=bt.PRECISION_MIXED)
    plan_setup = time.perf_counter() - t0
    st0 = plan.stats()
    log(f"[bench] plan {plan_setup:.2f}s (tree {st0['host_tree_s']:.2f} lists "
        f"{st0['host_lists_s']:.2f} work {st0['host_work_s']:.2f} h2d {st0['h2d_s']:.2f})")
    idx = np.arange(0, n, n // 1000)[:1000]
    direct = plan.direct_sample(idx)
    gsteps = max(1, min(args.steps, 3))
    wall_fast, gpu_fast, near_fast, st = time_steps(plan, bt.PRECISION_MIXED, gsteps, 1)
    err_fast = rel_l2(direct, plan.execute()[idx])
    log(f"[bench] mixed: {wall_fast:.3f} s/step (gpu {gpu_fast:.3f}, near {near_fast:.3f}), relL2 {err_fast:.3e}")
    t0 = time.perf_counter()
    plan.use_model(model, args.delta)
    deg_setup = time.perf_counter() - t0
    wall_ml, gpu_ml, near_ml, st_ml = time_steps(plan, bt.PRECISION_MIXED, gsteps, 1)
    err_ml = rel_l2(direct, plan.execute()[idx])
    log(f"[bench] mixed+learned: {wall_ml:.3f} s/step (gpu {gpu_ml:.3f}, far {st_ml['gpu_far_s']:.3f}, "
        f"mean degree {st_ml['mean_pair_degree']:.2f}), relL2 {err_ml:.3e}, degree setup {deg_setup:.2f}s")
    plan.set_pair_degrees(None)
    wall_plain = wall_plain1 = err_plain = None
    if args.plain_steps > 0:
        wall_plain, gpu_plain, _, st_p = time_steps(plan, bt.PRECISION_FP64, args.plain_steps, 1)
        err_plain = rel_l2(direct, plan.execute()[idx])
        log(f"[bench] fp64 (symmetric near field): {wall_plain:.3f} s/step (symmetric launch "
            f"{st_p['gpu_near_sym_mean_s']:.3f}), relL2 {err_plain:.3e}")
        # the reference's form: one evaluation per ordered near-field pair
        plan.set_symmetric(False)
        wall_plain1, _, _, _ = time_steps(plan, bt.PRECISION_FP64, 1, 0)
        plan.set_symmetric(True)
        log(f"[bench] fp64 one-directional: {wall_plain1:.3f} s/step")
    plan.close()
    err_tts = rel_l2(direct, out_tts[idx])
    if plain_tts:
        err_plain_tts = rel_l2(direct, out_plain_tts[idx])

    evals = st["pp_ptwise"] + st["pc_ptwise"] + st["cp_ptwise"] + st["cc_ptwise"]
    evals_learned = st["pp_ptwise"] + st_ml["far_ptwise_variable"]
    # dominant kernel: the symmetric near field (each leaf pair once for both
    # particles); its work counted as the reference's, 17 flops per ordered
    # pair evaluation it delivers
    sym_evals = st["sym_ptwise"]
    sym_s = st_ml["gpu_near_sym_mean_s"]
    achieved = FLOPS_PER_EVAL * sym_evals / sym_s / 1e12 if sym_s > 0 else 0.0
    roofline = {
        "bound": "valu",
        "kernel": "k_near_sym<Coulomb, subtraction> (symmetric near field, mixed precision)",
        "achieved": round(achieved, 2),
        "peak": PEAK["fp32"],
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK["fp32"], 4),
        "frac_executed": round(achieved * st["sym_evals"] / max(sym_evals, 1) / PEAK["fp32"], 4),
        "traffic": near_traffic() if n == 16_000_000 else None,
        "kernel_s": round(sym_s, 4),
        "ordered_pairs": sym_evals,
        "executed_pair_evaluations": st["sym_evals"],
        "flops_per_ordered_pair": FLOPS_PER_EVAL,
    }
    mix = near_valu_mix() if n == 16_000_000 else None
    if mix:
        from barytree_amd.probe import probe_valu
        probe = probe_valu()
        simds = 4 * torch.cuda.get_device_properties(0).multi_processor_count
        ceil_s = issue_model(mix, probe, simds)
        roofline["valu_issue_model_s"] = round(ceil_s, 4)
        roofline["valu_issue_model_over_measured"] = round(ceil_s / sym_s, 4)
        log(f"[bench] symmetric near kernel {sym_s:.3f} s, VALU issue time of its instruction mix "
            f"(kinds priced alone) {ceil_s:.3f} s; probe ns/instr {probe}")
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baselines")
        cpu = cpu_baseline_large(n, kargs, st, host_threads())

    line = {
        "metric": "time-to-solution N=16M at rel-L2 err <=1e-6 vs FP64 direct; speedup vs plain",
        "value": round(tts, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tts * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 pair kernel, f64 charges/accumulation",
        "data": "synthetic",
        "config": {
            "workload": (f"Coulomb singularity subtraction kappa=0.5, N={n} uniform [-1,1]^3, "
                         "cluster-cluster theta=0.8 degree 10 max leaf 8000"),
            "n": n, "compute_type": "CC", "theta": 0.8, "degree": 10, "max_leaf": 8000,
        },
        "form": ("one BaryTree_SolveLearnedDevice call per step, particles resident in HBM, potential "
                 "written to HBM: trees, lists, per-pair degrees (shipped coulomb_cc model, loaded once) and "
                 "the GPU path in mixed precision"),
        "rel_l2_err": err_tts,
        "time_to_solution_phases_s": {k: round(st_tts[k], 4) for k in (
            "host_copy_s", "host_tree_s", "host_lists_s", "host_work_s", "h2d_s", "gpu_total_s")},
        "host_setup_s": round(host_s, 4),
        "time_to_solution_host_arrays_s": round(tts_host, 4),
        "host_arrays_result_equal": host_equal,
        "plain_fp64_time_to_solution_s": round(plain_tts, 4) if plain_tts else None,
        "plain_fp64_time_to_solution_host_arrays_s": round(plain_tts_host, 4) if plain_tts_host else None,
        "plain_fp64_time_to_solution_rel_l2_err": err_plain_tts,
        "speedup_vs_plain": round(plain_tts / tts, 3) if plain_tts else None,
        "speedup_vs_plain_host_arrays": round(plain_tts_host / tts_host, 3) if plain_tts_host else None,
        "gpu_step_s": {
            "mixed_learned": round(wall_ml, 4),
            "mixed_fixed_degree": round(wall_fast, 4),
            "plain_fp64_symmetric": round(wall_plain, 4) if wall_plain else None,
            "plain_fp64_one_directional": round(wall_plain1, 4) if wall_plain1 else None,
        },
        "gpu_step_speedup_vs_plain_fp64_symmetric": round(wall_plain / wall_ml, 3) if wall_plain else None,
        "gpu_step_speedup_vs_plain_fp64_one_directional": round(wall_plain1 / wall_ml, 3) if wall_plain1 else None,
        "rel_l2_err_gpu_step": {"mixed_learned": err_ml, "mixed_fixed_degree": err_fast,
                                "plain_fp64": err_plain},
        "mean_pair_degree": st_ml["mean_pair_degree"],
        "far_evals_fixed": st["cc_ptwise"] + st["pc_ptwise"] + st["cp_ptwise"],
        "far_evals_learned": st_ml["far_ptwise_variable"],
        "degree_setup_s": round(deg_setup, 3),
        "phases_s": {k: round(st_ml[k], 4) for k in ("gpu_upward_s", "gpu_near_s", "gpu_near_sym_s",
                                                      "gpu_far_s", "gpu_down_s", "gpu_finish_s")},
        "phases_note": ("upward pass, far field and downpass run on a second stream beside the symmetric "
                        "near-field launch: their times are elapsed times while sharing the GPU with it"),
        "sym_pairs": sym_evals,
        "interactions": evals_learned,
        "interactions_per_s": evals_learned / tts,
        "interactions_per_s_gpu_step": evals_learned / wall_ml,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_reference_one_core"] = round(cpu["value"] / tts, 1)
        if "all_cores" in cpu:
            line["speedup_vs_cpu_all_cores"] = round(cpu["all_cores"]["value"] / tts, 1)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16_000_000)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-configs", action="store_true",
                    help="headline only: skip the other configurations' sub-objects")
    ap.add_argument("--workload", choices=["large", "batched", "learned", "yukawa", "cpu-case", "hermite",
                                           "kernels"],
                    default="large")
    ap.add_argument("--delta", type=float, default=0.5,
                    help="degree model threshold (probability that a lower degree is not enough)")
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 2 GiB (HBM-traffic counter calibration)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only print the measured issue cost of the VALU instruction kinds")
    args = ap.parse_args()
    if args.valu_probe:
        from barytree_amd.probe import probe_valu
        print(json.dumps({"valu_ns_per_wave_instruction_per_simd": probe_valu()}))
        return
    if args.calibrate_copy:
        import torch
        a = torch.empty(2 ** 28, dtype=torch.float64, device="cuda").uniform_()
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": 2 ** 31, "bytes_written": 2 ** 31}))
        return
    if args.gpus != 1:
        print("bench.py: this library runs on one GPU per process; multi-GPU runs "
              "(a locally essential tree exchange between ranks) are not part of it.", file=sys.stderr)
        sys.exit(2)

    single = {"batched": bench_batched, "learned": bench_learned, "yukawa": bench_yukawa,
              "cpu-case": bench_cpu_case, "hermite": bench_hermite, "kernels": bench_kernels}
    if args.workload in single:
        print(json.dumps(single[args.workload](args)), flush=True)
        return

    line = bench_large(args)
    if not args.no_configs:
        # the other configurations, each at its own size with a few steps
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.n = 3, 1, 16_000_000
        configs = {}
        for key, fn in (("C1_cpu_case", bench_cpu_case), ("C2_yukawa", bench_yukawa),
                        ("C3_batched", bench_batched), ("C4_learned", bench_learned)):
            log(f"[bench] configuration {key}")
            configs[key] = summary(fn(sub))
        line["configs"] = configs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
