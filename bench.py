"""Benchmark of the two hot paths on one MI355X.

Primary line (value): FP32-scaled Felsenstein pruning, pattern.node updates/s,
on the 512-taxon Yule tree / 200,000 unique GTR+G4 patterns workload
(branch lengths ~ Exp(0.05), weights Poisson(3)+1).  One step = one full
lnL evaluation (P matrices, the fused post-order pass, the compensated
log-sum).  Alongside: the FP64 forms (plain = every partial stored; fused),
the derivative pass, and Fitch parsimony site.node/s on the 1,024-taxon /
10^6-site workload (down-pass and up-pass), each with its own roofline.

Inputs are generated on the host, copied to HBM before timing; timing uses
HIP events on the stream the library launches on.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TF = 157.3
FP64_PEAK_TF = 78.6

# algorithmic cost per unit (per internal pattern.node / site.node)
PRUNE_F32_BYTES = 136.0   # 2 * (R*K*4 + 4)
PRUNE_F64_BYTES = 264.0   # 2 * (R*K*8 + 4)
PRUNE_FLOPS = 272.0       # 2 children * 4 rates * 16 MAC * 2 + 16
FITCH_DOWN_BYTES = 1.0044  # per internal site.node (leaves read, prelims written, costs)
FITCH_UP_BYTES = 1.0       # per site.node (every set read once, every final written once)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(prefix):
    """HBM bytes per step of the kernels whose name starts with `prefix`, from
    the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in
    profiles/hbm_traffic.json (already scaled by the copy calibration); None
    when that kernel was not profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    hits = [v for name, v in k.items() if name.startswith(prefix)]
    if not hits:
        return None
    return sum(v["read_bytes"] + v["write_bytes"] for v in hits)


def cpu_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps  # ms per step


def bench_spr(torch, ctx, cpu_baseline: bool):
    """Exact scores of SPR neighbours (the learned-proposal workload: 256-taxon
    tree, 50,000 GTR+G4 patterns): GPU exact evaluations per second over an
    even sample of 4,096 of the round's moves, FP64 and FP32-scaled; beside
    it the FP64 oracle on 8 of the moved trees, which must agree."""
    from phylocaml_amd import spr, spr_search as SS

    inst = SS.make_instance(7, 256, 50_000, 10)
    c = SS.Context(ctx, inst)
    mv = spr.moves(inst.N, inst.children, inst.root)
    sample = mv[:: max(1, len(mv) // 4096)][:4096]
    rates, scores = {}, {}
    for name, fp64 in (("f64", True), ("f32", False)):
        c.exact(inst.children, inst.root, inst.blen, sample[:256], fp64=fp64)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scores[name] = c.exact(inst.children, inst.root, inst.blen, sample, fp64=fp64)
        torch.cuda.synchronize()
        rates[name] = len(sample) / (time.perf_counter() - t0)
    res = {"workload": "256-taxon Yule tree, 50,000 unique GTR+G4 patterns; exact lnL of "
                       f"{len(sample)} of the {len(mv)} rooted SPR neighbours (even sample)",
           "moves_per_round": int(len(mv)), "exact_evaluations_per_s": rates}
    log(f"  spr: {len(mv)} moves; exact evaluations/s f64 {rates['f64']:.0f}, "
        f"f32 {rates['f32']:.0f}")
    if cpu_baseline:
        from oracle import oracle as O

        threads = min(16, os.cpu_count() or 1)
        tips = D_unpack_dna(inst)
        m = inst.model
        # one untimed evaluation (thread pool start-up), then 8 timed ones
        O.lnl(inst.N, inst.children, inst.root, inst.blen, m.priors, m.q, m.rates, m.probs, tips,
              inst.weights, threads=threads)
        t_cpu, worst, n_cpu = 0.0, 0.0, 8
        for k in range(0, len(sample), len(sample) // n_cpu)[:n_cpu]:
            s_, v_ = (int(x) for x in sample[k])
            ch, r, bl = spr.apply(inst.N, inst.children, inst.root, inst.blen, s_, v_)
            c0 = time.perf_counter()
            ov, _ = O.lnl(inst.N, ch, r, bl, m.priors, m.q, m.rates, m.probs, tips,
                          inst.weights, threads=threads)
            t_cpu += time.perf_counter() - c0
            worst = max(worst, abs(scores["f64"][k] - ov) / abs(ov))
        res["rel_err_f64_vs_oracle"] = worst
        res["cpu_baseline"] = {"value": n_cpu / t_cpu, "unit": "exact evaluations/s",
                               "cores": threads, "kind": "port",
                               "sample": f"oracle FP64 lnL of {n_cpu} moved trees ({t_cpu:.2f}s, "
                                         f"after one untimed evaluation), {threads} OpenMP "
                                         "threads"}
        log(f"  spr cpu: {n_cpu / t_cpu:.2f} exact evaluations/s ({threads} thr), "
            f"max rel err {worst:.1e}")
    return res


def D_unpack_dna(inst):
    """uint32 state bitmasks [N, P] of an instance's packed 4-bit tips."""
    from phylocaml_amd import device as D

    return D.unpack_nibbles(inst.tips_np, inst.P).astype(np.uint32)


def bench_aa(torch, ctx, n_trees, steps):
    """Batched 20-state pruning: n_trees trees, taxa ~ U{16..256}, patterns
    ~ U{1000..50000}, one persistent BF16-MFMA launch per size bucket with
    the autotuned waves per workgroup."""
    from oracle import oracle as O
    from phylocaml_amd import aa

    log(f"aa workload: generating {n_trees} trees")
    t0 = time.time()
    trees = aa.synth_workload(n_trees, seed=3, n_sim=1000)
    log(f"  generated in {time.time() - t0:.1f}s")
    tune = aa.autotune(ctx, trees)
    log(f"  autotune: { {k: v['waves'] for k, v in tune.items()} }")
    tb = aa.TunedBatch(ctx, trees, tune)
    ms = time_steps(torch, tb.run, steps, 1)
    got = tb.lnl.cpu().numpy()
    # error against the FP64 oracle on the 8 smallest trees and, below, the 3
    # nearest the median size (CPU time bounded)
    order = np.argsort([t.n_leaves * t.n_patterns for t in trees])[:8]
    errs = []
    threads = min(16, os.cpu_count() or 1)
    for i in order:
        t = trees[i]
        ov, _ = O.lnl(t.n_leaves, t.children, t.root, t.blen, t.pi, t.q, t.rates, t.probs,
                      aa.tips_bitmask(t.tips), t.weights, threads=threads)
        errs.append(abs(got[i] - ov) / abs(ov))
    # CPU baseline: the oracle on the 3 trees nearest the median size
    size = np.array([t.n_patterns * (t.n_leaves - 1) for t in trees])
    mid = np.argsort(np.abs(size - np.median(size)))[:3]
    t_cpu = 0.0
    for i in mid:
        t = trees[i]
        tips = aa.tips_bitmask(t.tips)
        c0 = time.perf_counter()
        ov, _ = O.lnl(t.n_leaves, t.children, t.root, t.blen, t.pi, t.q, t.rates, t.probs,
                      tips, t.weights, threads=threads)
        t_cpu += time.perf_counter() - c0
        errs.append(abs(got[i] - ov) / abs(ov))
    pn_cpu = int(size[mid].sum())
    pn = tb.pattern_nodes
    useful_tf = 6400.0 * pn / (ms / 1e3) / 1e12
    executed_tf = 16384.0 * pn / (ms / 1e3) / 1e12
    res = {"workload": f"{n_trees} trees, taxa U{{16..256}}, patterns U{{1000..50000}}, "
                       "GTR-20 (190 exch ~ Dir(1), pi ~ Dir(2)), G4 alpha ~ U(0.3,2), "
                       "bl ~ Exp(0.1)",
           "pattern_nodes": pn, "ms": ms, "pattern_node_per_s": pn / (ms / 1e3),
           "autotune": tune,
           "rel_err_vs_fp64_oracle": {"max": float(max(errs)), "mean": float(np.mean(errs)),
                                      "trees": [int(i) for i in order] + [int(i) for i in mid]},
           "cpu_baseline": {"value": pn_cpu / t_cpu, "unit": "pattern.node updates/s",
                            "cores": threads, "kind": "port",
                            "sample": f"oracle FP64 20-state pruning of the 3 trees nearest "
                                      f"the batch's median size ({pn_cpu:.3e} pattern.nodes, "
                                      f"{t_cpu:.2f}s), {threads} OpenMP threads"},
           "roofline": {"bound": "mfma", "traffic": None, "achieved": useful_tf, "peak": 2500.0,
                        "unit": "TFLOP/s", "frac": useful_tf / 2500.0,
                        "executed_tflops": executed_tf,
                        "note": "6,400 useful BF16 MFMA flop per internal pattern.node "
                                "(2 children x 4 rates x 20x20 MAC); 16,384 executed "
                                "(states padded 20 -> 32 on both sides)"}}
    log(f"  aa: {ms:.1f} ms for {pn:.3e} pattern.nodes = {pn / (ms / 1e3):.3e}/s, "
        f"err max {max(errs):.2e}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["all", "prune", "fitch", "aa", "spr"], default="all")
    ap.add_argument("--aa-trees", type=int, default=2048)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): every path "
              "here runs on a single GPU; sharding across GPUs is not implemented",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0").fill_(1)
        b = torch.empty_like(a)
        ms = time_steps(torch, lambda: b.copy_(a), 10, 2)
        print(json.dumps({"copy_bytes": 1 << 30, "ms": ms,
                          "read_plus_write_GBs": 2 * (1 << 30) / (ms / 1e3) / 1e9}), flush=True)
        return

    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM
    from phylocaml_amd import nonadditive as NA, synth

    ctx = D.Context(0)
    out = {}
    steps, warm = args.steps, args.warmup

    # ------------------------------------------------------------- pruning
    N, P = 512, 200_000
    log("pruning workload: generating")
    t0 = time.time()
    tree = synth.yule(N, 2, 0.05)
    pi, co = synth.gtr_params(4, 5.0, 2)
    model = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi),
                              MM.PHY_GAMMA_MEAN))
    st = synth.simulate(tree, model.priors, model.q, model.rates, P, 2, unique=True)
    w = synth.poisson_weights(P, 3.0, 2)
    tips_np = synth.pack_dna(st)
    log(f"  generated in {time.time() - t0:.1f}s")
    topo = D.Topology(N, tree.children, tree.root, ctx)
    dm = LK.DeviceModel(ctx, model)
    tips = D.to_device(tips_np)
    wd = D.to_device(w)
    bl = D.to_device(tree.blen)
    stride = tips_np.shape[1]
    units = P * (N - 1)
    forms = {"f32_fused": 1 | 2, "f64_plain": 0, "f64_fused": 2, "f32_plain": 1}
    res = {}
    lnls = {}
    for name, form in forms.items():
        ev = LK.Evaluator(ctx, topo, dm, P, form)
        ms = time_steps(torch, lambda: ev.run(tips, stride, wd, bl), steps, warm)
        lnls[name] = float(ev.lnl.cpu()[0])
        res[name] = ms
        log(f"  {name}: {ms:.3f} ms  lnL {lnls[name]!r}")
        del ev
        torch.cuda.empty_cache()
    evd = LK.Evaluator(ctx, topo, dm, P, 0, derivs=True)
    ms_d64 = time_steps(torch, lambda: evd.run(tips, stride, wd, bl), max(3, steps // 4), 2)
    del evd
    evd32 = LK.Evaluator(ctx, topo, dm, P, 1, derivs=True)
    ms_d32 = time_steps(torch, lambda: evd32.run(tips, stride, wd, bl), max(3, steps // 4), 2)
    del evd32
    torch.cuda.empty_cache()
    log(f"  derivs f64 {ms_d64:.3f} ms, f32 {ms_d32:.3f} ms")
    ref = lnls["f64_plain"]
    rel32 = abs(lnls["f32_fused"] - ref) / abs(ref)
    ms_main = res["f32_fused"]
    value = units / (ms_main / 1e3)

    # ------------------------------------------------------------- Fitch
    fitch = None
    if args.only in ("all", "fitch"):
        NF, S = 1024, 1_000_000
        log("fitch workload: generating")
        ftree = synth.yule(NF, 1, 0.1)
        lv = synth.fitch_leaves(NF, S, 1)
        ftopo = D.Topology(NF, ftree.children, ftree.root, ctx)
        ft = NA.alloc_tree(ctx, ftopo, S, lv, with_final=True)
        ms_down = time_steps(torch, lambda: NA.downpass(ctx, ftopo, ft), steps, warm)
        ms_up = time_steps(torch, lambda: NA.uppass(ctx, ftopo, ft), steps, warm)
        total = int(ft.total.cpu()[0])
        down_units, up_units = S * (NF - 1), S * (2 * NF - 1)
        gbs_down = FITCH_DOWN_BYTES * down_units / (ms_down / 1e3) / 1e9
        gbs_up = FITCH_UP_BYTES * up_units / (ms_up / 1e3) / 1e9
        fitch = {
            "workload": "1024-taxon Yule tree, 1e6 IUPAC sites, 16 sites/uint64",
            "downpass_site_node_per_s": down_units / (ms_down / 1e3),
            "uppass_site_node_per_s": up_units / (ms_up / 1e3),
            "ms_downpass": ms_down, "ms_uppass": ms_up, "total_cost": total,
            "roofline_downpass": {"bound": "hbm", "achieved": gbs_down, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": gbs_down / HBM_PEAK_GBS,
                                  "traffic": measured_traffic("k_fitch_down<")},
            "roofline_uppass": {"bound": "hbm", "achieved": gbs_up, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": gbs_up / HBM_PEAK_GBS,
                                "traffic": measured_traffic("k_fitch_up<")},
        }
        log(f"  fitch down {ms_down:.3f} ms ({gbs_down:.0f} GB/s alg), "
            f"up {ms_up:.3f} ms ({gbs_up:.0f} GB/s alg), total {total}")
        if not args.no_cpu_baseline:
            from oracle import oracle as O

            # the same leaves through the oracle
            S_cpu = S
            thr = min(16, os.cpu_count() or 1)
            leaves = D.unpack_nibbles(lv[:, : D.words(S_cpu)], S_cpu)
            c0 = time.perf_counter()
            osets, _, _, _ = O.fitch_downpass(NF, ftree.children, ftree.root, leaves, threads=thr)
            t_down = time.perf_counter() - c0
            c0 = time.perf_counter()
            O.fitch_uppass(NF, ftree.children, ftree.root, osets)
            t_up = time.perf_counter() - c0
            fitch["cpu_baseline"] = {
                "value": S_cpu * (NF - 1) / t_down, "unit": "site.node/s (down-pass)",
                "cores": thr, "kind": "port",
                "uppass_value": S_cpu * (2 * NF - 1) / t_up, "uppass_cores": 1,
                "sample": f"oracle Fitch on the whole workload ({S_cpu:,} sites): "
                          f"down-pass {t_down:.2f}s on {thr} OpenMP threads, up-pass "
                          f"{t_up:.2f}s on 1 thread"}
            log(f"  fitch cpu: down {fitch['cpu_baseline']['value']:.3e}/s ({thr} thr), "
                f"up {fitch['cpu_baseline']['uppass_value']:.3e}/s (1 thr)")
        del ft
        torch.cuda.empty_cache()

    # ------------------------------------------------------------- batched 20-state
    aa_res = None
    if args.only in ("all", "aa"):
        aa_res = bench_aa(torch, ctx, args.aa_trees, max(2, steps // 5))

    # ------------------------------------------------------------- exact SPR scores
    spr_res = None
    if args.only in ("all", "spr"):
        spr_res = bench_spr(torch, ctx, not args.no_cpu_baseline)

    # ------------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        model_name, ncpu = cpu_info()
        threads = min(16, ncpu)
        # bounded samples of the same pruning workload (a few seconds of CPU work)
        tips_all = np.where(st < 4, 1 << np.minimum(st, 3), 15).astype(np.uint32)
        reps = 5
        t0 = time.time()
        for _ in range(reps):
            O.lnl(N, tree.children, tree.root, tree.blen, model.priors, model.q, model.rates,
                  model.probs, tips_all, w, threads=threads)
        dt = (time.time() - t0) / reps
        n_1 = 50_000
        t0 = time.time()
        O.lnl(N, tree.children, tree.root, tree.blen, model.priors, model.q, model.rates,
              model.probs, tips_all[:, :n_1], w[:n_1], threads=1)
        dt1 = time.time() - t0
        cpu = {"value": P * (N - 1) / dt, "unit": "pattern.node updates/s", "cores": threads,
               "kind": "port",
               "sample": f"oracle FP64 pruning (lnL) of the full 200,000-pattern, 512-taxon "
                         f"workload, {threads} OpenMP threads, mean of {reps} runs "
                         f"({dt:.2f}s each); host CPU: {model_name} ({ncpu} logical CPUs "
                         f"visible)",
               "value_1thread": n_1 * (N - 1) / dt1,
               "sample_1thread": f"first {n_1} patterns, 1 thread, {dt1:.2f}s"}
        log(f"  cpu baseline: {cpu['value']:.3e} ({threads} thr), "
            f"{cpu['value_1thread']:.3e} (1 thr)")

    alg_gbs = PRUNE_F32_BYTES * units / (ms_main / 1e3) / 1e9
    tflops = PRUNE_FLOPS * units / (ms_main / 1e3) / 1e12
    line = {
        "metric": "FP32-scaled vs FP64 pruning updates/s at rel ΔlnL≤1e-6; "
                  "Fitch site-node/s",
        "value": value,
        "unit": "pattern.node updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms_main,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Yule tree, simulated GTR+G4 sites)",
        "config": {"workload": "DNA pruning, 512-taxon Yule tree (bl~Exp(0.05)), 200,000 "
                               "unique GTR+G4 patterns (alpha 0.5, Yang mean-rate "
                               "categories), weights Poisson(3)+1; FP32 partials with "
                               "int32 exponents per (pattern, rate), fused post-order pass",
                   "n_taxa": N, "n_patterns": P, "n_rates": 4, "form": "f32_fused"},
        "roofline": {"bound": "hbm", "achieved": alg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg_gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic("k_prune_down<float, 4, 1, true,"),
                     "compute_tflops": tflops, "compute_peak_tflops": FP32_PEAK_TF,
                     "note": "achieved = 136 B per internal pattern.node (the unfused "
                             "algorithmic traffic) x 511 nodes x 200,000 patterns / step "
                             "time; the fused kernel keeps interior partials on chip, so "
                             "frac > 1 is the fusion gain and the kernel is bound by FP32 "
                             "VALU issue (compute_tflops); traffic = HBM bytes per launch "
                             "measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (copy-"
                             "calibrated, profiles/hbm_traffic.json)"},
        "cpu_baseline": cpu,
        "prune": {
            "ms": res, "ms_derivs_f64": ms_d64, "ms_derivs_f32": ms_d32,
            "updates_per_s": {k: units / (v / 1e3) for k, v in res.items()},
            "fp32_vs_fp64_plain": res["f64_plain"] / ms_main,
            "fp32_vs_fp64_fused": res["f64_fused"] / ms_main,
            "rel_err_fp32_vs_fp64": rel32,
            "lnl": lnls,
        },
        "fitch": fitch,
        "aa": aa_res,
        "spr": spr_res,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
