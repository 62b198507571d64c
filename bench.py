"""Benchmark of the BaryTree interaction path on one MI355X.

Workload (the headline configuration): Coulomb kernel with singularity
subtraction (kappa = 0.5), N = 16M particles uniform in [-1,1]^3 (sources =
targets, separate arrays), charges uniform in [-1,1], weights 1,
cluster-cluster treecode, theta = 0.8, degree 10, max leaf 8000.

A step is one execution of the GPU path (upward pass, near and far field,
downpass, correction, scatter to caller order) with its inputs resident in
HBM; the host work (trees, interaction lists, upload) is done once by the
plan and reported separately.  Three forms are timed on the same plan: the
plain FP64 form; the mixed form (FP32 pair evaluation with FP64 charges and
accumulation); and the mixed, learned-degree form (each far-field pair at the
degree the shipped degree model picks) -- the one the metric is quoted on.
Each is checked against an FP64 direct sum at 1000 stride-sampled targets.

Prints one JSON line on stdout.

--workload learned measures the learned degree selection on its own
configuration: regularized Coulomb (epsilon 0.01), N = 4M Plummer sphere
(scale 0.1, truncated at r = 1, seed 4), cluster-particle, theta 0.7, leaf
2000: fixed degree 10 against the per-pair degrees of the shipped model
(trained on other Plummer realizations), in FP64 and mixed.

--workload yukawa: Yukawa (kappa 0.5), N = 1M uniform, cluster-cluster,
theta 0.8, degree 8, leaf 4000, plain FP64 against mixed precision.

--workload cpu-case: Coulomb, N = 100k uniform, particle-cluster, theta 0.7,
degree 8, leaf 2000, skipping, FP64, against the reference's own sources
compiled for one CPU core.

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


def cpu_baseline(n_evals):
    """Times the oracle (FP64 restatement of the reference kernels, 1 core)
    on a bounded sample and scales to the workload's evaluation count."""
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    m, n = 1000, 2_000_000
    x, y, z, q, w = gen_uniform(n, 99)
    t0 = time.perf_counter()
    O.direct(x[:m], y[:m], z[:m], q[:m], x, y, z, q, w, 1, [0.5], 2)
    dt = time.perf_counter() - t0
    rate = m * n / dt
    return {
        "value": n_evals / rate,
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle FP64 Coulomb-SS pair kernel (coulomb_ss_pp.c arithmetic), "
                   f"{m} targets x {n} sources = {m * n:.1e} evaluations in {dt:.1f} s "
                   f"({rate:.3e} eval/s, 1 core); workload time = its {n_evals:.3e} "
                   f"point-wise evaluations / that rate (upward/downward passes excluded)"),
    }


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

    # one-shot batched solve, end to end (host trees and lists, upload, GPU, download)
    def batched_once(precision):
        t0 = time.perf_counter()
        out = bt.solve_batched(systems, *kargs, schedule=sc, precision=precision)
        return time.perf_counter() - t0, out, bt.last_stats()
    batched_once(bt.PRECISION_FP64)  # warm-up
    e2e = [batched_once(bt.PRECISION_FP64) for _ in range(args.steps)]
    e2e_fp64 = min(r[0] for r in e2e)
    out_fp64, st_b = e2e[-1][1], e2e[-1][2]
    e2e_mixed_runs = [batched_once(bt.PRECISION_MIXED) for _ in range(args.steps)]
    e2e_mixed = min(r[0] for r in e2e_mixed_runs)
    out_mixed = e2e_mixed_runs[-1][1]
    log(f"[bench] batched solve fp64 {e2e_fp64:.3f}s mixed {e2e_mixed:.3f}s "
        f"(host tree {st_b['host_tree_s']:.3f} lists {st_b['host_lists_s']:.3f} "
        f"gpu {st_b['gpu_total_s']:.4f})")

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
    exact = all(np.array_equal(a, b) for a, b in zip(outs_seq, out_fp64))
    log(f"[bench] sequential {seq:.3f}s (sum of gpu phases {seq_gpu:.4f}s), batched==sequential: {exact}")

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
        "batched_equals_sequential_bitwise": exact,
        "gpu_batched_fp64_s": round(gpu_b, 5),
        "gpu_batched_mixed_s": round(gpu_bm, 5),
        "gpu_sequential_sum_s": round(seq_gpu, 5),
        "rel_l2_err_fp64": err64,
        "rel_l2_err_mixed": err32,
        "interactions": evals,
        "interactions_per_s_gpu_fp64": evals / gpu_b,
    }
    print(json.dumps(line), flush=True)


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
                                                               "gpu_far_s", "gpu_down_s")}}
        log(f"[bench] {form}: {wall:.4f} s/step, relL2 {err:.2e}, far evals {far:.3e}")
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
        "delta": args.delta,
        "speedup_learned_vs_fixed_fp64": round(res["fixed_fp64"]["s"] / res["learned_fp64"]["s"], 3),
        "speedup_learned_mixed_vs_fixed_fp64": round(res["fixed_fp64"]["s"] / res["learned_mixed"]["s"], 3),
        "host_setup_s": round(host_setup, 3),
    }
    print(json.dumps(line), flush=True)


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
    for form, prec in (("fp64", bt.PRECISION_FP64), ("mixed", bt.PRECISION_MIXED)):
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
        "interactions": evals,
        "roofline": {"bound": "valu", "kernel": "k_near_sym<Yukawa, skipping> (symmetric near field, mixed)",
                     "achieved": round(achieved, 2), "peak": PEAK["fp32"], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK["fp32"], 4), "traffic": None,
                     "kernel_s": round(sym_s, 5), "ordered_pairs": st["sym_ptwise"]},
        "host_setup_s": round(host_setup, 3),
    }
    print(json.dumps(line), flush=True)


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
                         "sample": f"the whole treecode on the same {n} particles, one call"},
        "speedup_vs_cpu": round(cpu_s / wall, 1),
        "speedup_end_to_end_vs_cpu": round(cpu_s / e2e, 1),
    }
    print(json.dumps(line), flush=True)


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
    print(json.dumps(line), flush=True)


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
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16_000_000)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
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

    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "learned":
        bench_learned(args)
        return
    if args.workload == "yukawa":
        bench_yukawa(args)
        return
    if args.workload == "cpu-case":
        bench_cpu_case(args)
        return
    if args.workload == "hermite":
        bench_hermite(args)
        return
    if args.workload == "kernels":
        bench_kernels(args)
        return

    import torch  # device setup and synchronisation only
    assert torch.cuda.is_available(), "no GPU"
    import barytree_amd as bt

    n = args.n
    log(f"[bench] generating N={n}")
    x, y, z, q, w = gen_uniform(n, 5)
    targets = (x.copy(), y.copy(), z.copy(), q.copy())
    sources = (x, y, z, q, w)
    kernel_args = (bt.Kernel.COULOMB, [0.5], bt.Singularity.SUBTRACTION,
                   bt.Approximation.LAGRANGE, bt.ComputeType.CLUSTER_CLUSTER, 0.8, 10, 8000, 8000)
    t0 = time.perf_counter()
    plan = bt.Plan(targets, sources, *kernel_args, precision=bt.PRECISION_MIXED)
    host_setup = time.perf_counter() - t0
    st0 = plan.stats()
    log(f"[bench] plan {host_setup:.2f}s (tree {st0['host_tree_s']:.2f} lists "
        f"{st0['host_lists_s']:.2f} work {st0['host_work_s']:.2f} h2d {st0['h2d_s']:.2f})")

    # fast form (the metric)
    wall_fast, gpu_fast, near_fast, st = time_steps(plan, bt.PRECISION_MIXED, args.steps, args.warmup)
    log(f"[bench] mixed: {wall_fast:.3f} s/step (gpu {gpu_fast:.3f}, near {near_fast:.3f})")
    idx = np.arange(0, n, n // 1000)[:1000]
    out_fast = plan.execute()
    direct = plan.direct_sample(idx)
    err_fast = rel_l2(direct, out_fast[idx])
    log(f"[bench] mixed relL2 vs FP64 direct (1000 targets): {err_fast:.3e}")

    # mixed, learned-degree form (the metric)
    from barytree_amd import learned
    t0 = time.perf_counter()
    plan.use_model(learned.load_model("coulomb_cc"), args.delta)
    deg_setup = time.perf_counter() - t0
    wall_ml, gpu_ml, near_ml, st_ml = time_steps(plan, bt.PRECISION_MIXED, args.steps, args.warmup)
    out_ml = plan.execute()
    err_ml = rel_l2(direct, out_ml[idx])
    log(f"[bench] mixed+learned: {wall_ml:.3f} s/step (gpu {gpu_ml:.3f}, far {st_ml['gpu_far_s']:.3f}, "
        f"mean degree {st_ml['mean_pair_degree']:.2f}), relL2 {err_ml:.3e}, degree setup {deg_setup:.2f}s")
    plan.set_pair_degrees(None)

    # plain FP64 form on the same plan
    wall_plain, err_plain = None, None
    if args.plain_steps > 0:
        wall_plain, gpu_plain, near_plain, _ = time_steps(plan, bt.PRECISION_FP64, args.plain_steps, 1)
        out_plain = plan.execute()
        err_plain = rel_l2(direct, out_plain[idx])
        log(f"[bench] fp64: {wall_plain:.3f} s/step (gpu {gpu_plain:.3f}), relL2 {err_plain:.3e}")
    plan.close()

    evals = st["pp_ptwise"] + st["pc_ptwise"] + st["cp_ptwise"] + st["cc_ptwise"]
    # dominant kernel: the symmetric near field (each leaf pair once for both
    # particles); its work counted as the reference's, 17 flops per ordered
    # pair evaluation it delivers
    sym_evals = st["sym_ptwise"]
    sym_s = st["gpu_near_sym_mean_s"]
    achieved = FLOPS_PER_EVAL * sym_evals / sym_s / 1e12 if sym_s > 0 else 0.0
    roofline = {
        "bound": "valu",
        "kernel": "k_near_sym<Coulomb> (symmetric near field, mixed precision)",
        "achieved": round(achieved, 2),
        "peak": PEAK["fp32"],
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK["fp32"], 4),
        "traffic": near_traffic() if n == 16_000_000 else None,
        "kernel_s": round(sym_s, 4),
        "ordered_pairs": sym_evals,
        "flops_per_ordered_pair": FLOPS_PER_EVAL,
    }
    # the VALU issue time of the kernel's instruction mix (kinds priced alone)
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
        log("[bench] timing the CPU oracle sample")
        cpu = cpu_baseline(evals)
        log(f"[bench] cpu baseline {cpu['value']:.1f} s (extrapolated)")

    line = {
        "metric": "time-to-solution N=16M at rel-L2 err <=1e-6 vs FP64 direct; speedup vs plain",
        "value": round(wall_ml, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ml * 1e3, 2),
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
        "form": "mixed precision + learned per-pair degree (shipped coulomb_cc model)",
        "rel_l2_err": err_ml,
        "speedup_vs_plain": round(wall_plain / wall_ml, 3) if wall_plain else None,
        "mixed_fixed_degree_s": round(wall_fast, 4),
        "mixed_fixed_degree_rel_l2_err": err_fast,
        "mean_pair_degree": st_ml["mean_pair_degree"],
        "far_evals_fixed": st["cc_ptwise"] + st["pc_ptwise"] + st["cp_ptwise"],
        "far_evals_learned": st_ml["far_ptwise_variable"],
        "degree_setup_s": round(deg_setup, 3),
        "phases_s": {k: round(st_ml[k], 4) for k in ("gpu_upward_s", "gpu_near_s", "gpu_near_sym_s",
                                                      "gpu_far_s", "gpu_down_s", "gpu_finish_s")},
        "plain_fp64_s": round(wall_plain, 4) if wall_plain else None,
        "plain_fp64_rel_l2_err": err_plain,
        "gpu_time_s": round(gpu_fast, 4),
        "near_phase_s": round(near_fast, 4),
        "near_sym_kernel_s": round(sym_s, 4),
        "sym_pairs": sym_evals,
        "host_setup_s": round(host_setup, 3),
        "end_to_end_s": round(host_setup + deg_setup + wall_ml, 3),
        "interactions": evals,
        "interactions_per_s": evals / wall_fast,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
