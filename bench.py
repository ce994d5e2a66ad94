"""Benchmark of the two hot paths on one MI355X.

Primary line (value): FP32-scaled Felsenstein pruning, pattern.node updates/s,
on the 512-taxon Yule tree / 200,000 unique GTR+G4 patterns workload
(branch lengths ~ Exp(0.05), weights Poisson(3)+1).  One step = one full
lnL evaluation (P matrices, the fused post-order pass, the compensated
log-sum); value = pattern.nodes / step time.  The roofline divides the
algorithmic work of one launch of the dominant kernel by that kernel's mean
duration, measured live with HIP events on the library's stream
(phy_ctx_kernel_timer) over the timed steps.

Alongside: the FP64 forms and the derivative pass (with their errors against
the FP64 oracle), Fitch parsimony on the 1,024-taxon / 10^6-site workload
beside a STREAM-style copy kernel, the 2,048-tree 20-state batch, exact SPR
neighbour scores, the learned-proposal search against exhaustive search,
the MLP kernels, and the CPU oracle on the host cores (the 32-taxon /
2,000-site case timed as the reference's single-threaded execution).

Inputs are generated on the host and copied to HBM before timing.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TF = 157.3      # MI355X vector FP32 (packed); 78.6 without packed FP32
FP32_NOPACK_TF = 78.6
FP64_PEAK_TF = 78.6
BF16_PEAK_TF = 2500.0     # dense BF16 MFMA

# algorithmic cost per unit (per internal pattern.node / site.node / move)
PRUNE_F32_BYTES = 136.0   # 2 * (R*K*4 + 4), R = 4 rate classes, K = 4 states
PRUNE_F64_BYTES = 264.0   # 2 * (R*K*8 + 4)
PRUNE_FLOPS = 272.0       # 2 children * 4 rates * 16 MAC * 2 + 16 (products)
# derivative pass, per (pattern, rate class, internal node): 4 eigen-basis
# 4x4 transforms of the two children (U^-1 D, U m) and 2 of their messages
# (U^T A), 2 message returns (U^-T), 16 FMA of L' and L'', ~40 products and
# sums: 320 flop, i.e. 1,280 per pattern.node at R = 4
DERIV_FLOPS = 1280.0
FITCH_DOWN_BYTES = 1.0044  # per internal site.node (leaves read, prelims written, costs)
FITCH_UP_BYTES = 1.0       # per site.node (every set read once, every final written once)
MLP_FWD_FLOPS = 41216.0    # 2 (32*128 + 128*128 + 128) per move
MLP_BWD_FLOPS = 73984.0    # 2 (2*128*128 + 32*128 + 128): the three weight gradients + dH1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(prefix, grid=None, suffix=""):
    """HBM bytes per launch of the kernels whose name starts with `prefix`
    and ends with `suffix` (and whose launch has `grid` work-items, when
    given), from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes summarised in profiles/hbm_traffic.json
    (already scaled by the copy calibration); None when that kernel was not
    profiled."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    hits = [v for name, v in k.items()
            if name.startswith(prefix) and name.endswith(suffix)
            and (grid is None or v.get("grid") == grid)]
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


class Clock:
    """Step time (HIP events around the timed steps) and the mean duration
    of the dominant kernel's launches within them (phy_ctx_kernel_timer:
    event pairs on the library's stream around each launch)."""

    def __init__(self, torch, ctx):
        from phylocaml_amd._lib import lib

        self.torch, self.ctx, self.lib = torch, ctx, lib

    def run(self, fn, steps, warmup, label=None):
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        self.lib.phy_ctx_kernel_timer(self.ctx.handle, 1 if label else 0)
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(steps):
            fn()
        e.record()
        torch.cuda.synchronize()
        step_ms = s.elapsed_time(e) / steps
        k_ms, n = None, 0
        if label:
            tot, cnt = C.c_double(), C.c_int64()
            self.ctx.check(self.lib.phy_ctx_kernel_time(self.ctx.handle, label.encode(),
                                                        C.byref(tot), C.byref(cnt)), "kernel_time")
            n = cnt.value
            k_ms = tot.value / max(n, 1)
        self.lib.phy_ctx_kernel_timer(self.ctx.handle, 0)
        return step_ms, k_ms, n


def kernel_totals(ctx, labels):
    """{label: total ms} of the kernels timed since phy_ctx_kernel_timer(1)."""
    from phylocaml_amd._lib import lib

    out = {}
    for label in labels:
        tot, cnt = C.c_double(), C.c_int64()
        ctx.check(lib.phy_ctx_kernel_time(ctx.handle, label.encode(), C.byref(tot), C.byref(cnt)),
                  "kernel_time")
        out[label] = round(tot.value, 3)
    return out


def stream_copy(torch, ctx, clock, gib=2):
    """The project's own STREAM-style copy kernel on `gib` GiB."""
    from phylocaml_amd import device as D
    from phylocaml_amd._lib import lib

    nbytes = gib << 30
    a = torch.empty(nbytes, dtype=torch.uint8, device=f"cuda:{ctx.device}").fill_(1)
    b = torch.empty_like(a)
    step, k_ms, _ = clock.run(lambda: ctx.check(lib.phy_copy_stream(ctx.handle, D.dptr(b),
                                                                    D.dptr(a), nbytes), "copy"),
                              10, 2, "copy_stream")
    assert torch.equal(a[:: 1 << 20], b[:: 1 << 20])
    gbs = 2 * nbytes / (k_ms / 1e3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    log(f"  copy_stream: {gib} GiB in {k_ms:.3f} ms = {gbs:.0f} GB/s (read + write)")
    return {"bytes": nbytes, "kernel_ms": k_ms, "read_plus_write_GBs": gbs,
            "frac_of_peak": gbs / HBM_PEAK_GBS,
            "kernel": "k_copy_stream: one 16-byte non-temporal load and store per thread, "
                      "a grid of 256-thread workgroups covering the buffer"}


def bench_prune(torch, ctx, clock, steps, warm, cpu_baseline):
    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM, synth

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
    tips, wd, bl = D.to_device(tips_np), D.to_device(w), D.to_device(tree.blen)
    stride = tips_np.shape[1]
    units = P * (N - 1)
    forms = {"f32_fused": 1 | 2, "f64_plain": 0, "f64_fused": 2, "f32_plain": 1}
    step_ms, kern_ms, lnls = {}, {}, {}
    for name, form in forms.items():
        ev = LK.Evaluator(ctx, topo, dm, P, form)
        s_ms, k_ms, _ = clock.run(lambda: ev.run(tips, stride, wd, bl), steps, warm, "prune_down")
        lnls[name] = float(ev.lnl.cpu()[0])
        step_ms[name], kern_ms[name] = s_ms, k_ms
        log(f"  {name}: step {s_ms:.3f} ms, prune_down kernel {k_ms:.3f} ms, lnL {lnls[name]!r}")
        del ev
        torch.cuda.empty_cache()
    deriv = {}
    for name, form in (("f64", 0), ("f32", 1)):
        ev = LK.Evaluator(ctx, topo, dm, P, form, derivs=True)
        s_ms, k_ms, _ = clock.run(lambda: ev.run(tips, stride, wd, bl), max(3, steps // 4), 2,
                                  "prune_deriv")
        deriv[name] = {"step_ms": s_ms, "kernel_ms": k_ms,
                       "d1": ev.d1.cpu().numpy(), "d2": ev.d2.cpu().numpy()}
        log(f"  derivatives {name}: lnL + d1/d2 step {s_ms:.3f} ms, prune_deriv kernel "
            f"{k_ms:.3f} ms")
        del ev
        torch.cuda.empty_cache()
    errors = None
    if cpu_baseline:
        from oracle import oracle as O

        thr = min(16, os.cpu_count() or 1)
        tips_all = np.where(st < 4, 1 << np.minimum(st, 3), 15).astype(np.uint32)
        t0 = time.time()
        ov, _, od1, od2, a1, a2 = O.lnl(N, tree.children, tree.root, tree.blen, model.priors,
                                        model.q, model.rates, model.probs, tips_all, w,
                                        derivs=True, scales=True, threads=thr)
        log(f"  oracle lnL + derivatives in {time.time() - t0:.1f}s")
        mask = np.arange(2 * N - 1) != tree.root
        errors = {"lnl_rel": {k: abs(v - ov) / abs(ov) for k, v in lnls.items()},
                  "oracle_lnl": ov,
                  "note": "derivative errors: |d - oracle| / sum_p w_p|term_p| per branch "
                          "(term = g for d1, |L''/L| + g^2 for d2), max and mean over the "
                          "1,022 branches"}
        for name in ("f64", "f32"):
            e1 = np.abs(deriv[name]["d1"] - od1)[mask] / a1[mask]
            e2 = np.abs(deriv[name]["d2"] - od2)[mask] / a2[mask]
            errors[f"d1_{name}"] = {"max": float(e1.max()), "mean": float(e1.mean())}
            errors[f"d2_{name}"] = {"max": float(e2.max()), "mean": float(e2.mean())}
        log(f"  errors vs oracle: lnL f32 fused {errors['lnl_rel']['f32_fused']:.1e}; d1/d2 "
            f"f32 {errors['d1_f32']['max']:.1e}/{errors['d2_f32']['max']:.1e}, f64 "
            f"{errors['d1_f64']['max']:.1e}/{errors['d2_f64']['max']:.1e}")
    for d in deriv.values():
        d.pop("d1"), d.pop("d2")
    return {"N": N, "P": P, "units": units, "step_ms": step_ms, "kernel_ms": kern_ms,
            "lnl": lnls, "deriv": deriv, "errors": errors,
            "oracle_inputs": (tree, model, st, w)}


def deriv_roofline(k_ms, units, peak, nbytes_prefix, grid=None):
    tf = DERIV_FLOPS * units / (k_ms / 1e3) / 1e12
    tr = measured_traffic(nbytes_prefix, grid)
    return {"bound": "valu", "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak,
            "traffic": tr,
            "hbm_GBs": None if tr is None else tr / (k_ms / 1e3) / 1e9,
            "note": "1,280 flop per internal pattern.node (320 per rate class: the eigen-"
                    "basis transforms of both children and their messages, L' and L'') / "
                    "mean k_prune_deriv duration"}


def bench_fitch(torch, ctx, clock, steps, warm, copy_gbs, cpu_baseline):
    from phylocaml_amd import device as D, nonadditive as NA, synth

    NF, S = 1024, 1_000_000
    log("fitch workload: generating")
    ftree = synth.yule(NF, 1, 0.1)
    lv = synth.fitch_leaves(NF, S, 1)
    ftopo = D.Topology(NF, ftree.children, ftree.root, ctx)
    ft = NA.alloc_tree(ctx, ftopo, S, lv, with_final=True)
    s_down, k_down, _ = clock.run(lambda: NA.downpass(ctx, ftopo, ft), steps, warm, "fitch_down")
    s_up, k_up, _ = clock.run(lambda: NA.uppass(ctx, ftopo, ft), steps, warm, "fitch_up")
    total = int(ft.total.cpu()[0])
    down_units, up_units = S * (NF - 1), S * (2 * NF - 1)

    def roof(bytes_per_unit, units, k_ms, prefix):
        gbs = bytes_per_unit * units / (k_ms / 1e3) / 1e9
        tr = measured_traffic(prefix)
        return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "frac_of_copy": gbs / copy_gbs if copy_gbs else None,
                "traffic": tr, "kernel_ms": k_ms,
                "measured_GBs": None if tr is None else tr / (k_ms / 1e3) / 1e9}

    fitch = {
        "workload": "1024-taxon Yule tree, 1e6 IUPAC sites, 16 sites/uint64",
        "downpass_site_node_per_s": down_units / (s_down / 1e3),
        "uppass_site_node_per_s": up_units / (s_up / 1e3),
        "ms_downpass": s_down, "ms_uppass": s_up, "total_cost": total,
        "roofline_downpass": roof(FITCH_DOWN_BYTES, down_units, k_down, "k_fitch_down<"),
        "roofline_uppass": roof(FITCH_UP_BYTES, up_units, k_up, "k_fitch_up<"),
    }
    log(f"  fitch down {k_down:.3f} ms ({fitch['roofline_downpass']['achieved']:.0f} GB/s alg), "
        f"up {k_up:.3f} ms ({fitch['roofline_uppass']['achieved']:.0f} GB/s alg), "
        f"total {total}")
    if cpu_baseline:
        from oracle import oracle as O

        thr = min(16, os.cpu_count() or 1)
        leaves = D.unpack_nibbles(lv[:, : D.words(S)], S)
        c0 = time.perf_counter()
        osets, _, _, otot = O.fitch_downpass(NF, ftree.children, ftree.root, leaves, threads=thr)
        t_down = time.perf_counter() - c0
        c0 = time.perf_counter()
        O.fitch_downpass(NF, ftree.children, ftree.root, leaves[:, :100_000], threads=1)
        t_down1 = time.perf_counter() - c0
        c0 = time.perf_counter()
        O.fitch_uppass(NF, ftree.children, ftree.root, osets)
        t_up = time.perf_counter() - c0
        assert otot == total, "Fitch total differs from the oracle"
        fitch["cpu_baseline"] = {
            "value": S * (NF - 1) / t_down, "unit": "site.node/s (down-pass)",
            "cores": thr, "kind": "port",
            "value_1thread": 100_000 * (NF - 1) / t_down1,
            "uppass_value": S * (2 * NF - 1) / t_up, "uppass_cores": 1,
            "sample": f"oracle Fitch down-pass on the whole workload ({S:,} sites, "
                      f"{t_down:.2f}s, {thr} OpenMP threads; equal total), on the first "
                      f"100,000 sites with 1 thread ({t_down1:.2f}s); up-pass on the whole "
                      f"workload, 1 thread ({t_up:.2f}s)"}
        log(f"  fitch cpu: down {fitch['cpu_baseline']['value']:.3e}/s ({thr} thr), "
            f"{fitch['cpu_baseline']['value_1thread']:.3e}/s (1 thr), up "
            f"{fitch['cpu_baseline']['uppass_value']:.3e}/s (1 thr)")
    del ft
    torch.cuda.empty_cache()
    return fitch


def bench_aa(torch, ctx, clock, n_trees, steps, cpu_baseline):
    """Batched 20-state pruning: n_trees trees, taxa ~ U{16..256}, patterns
    ~ U{1000..50000}: the autotuned tile per size bucket, one persistent
    launch when the buckets agree."""
    from phylocaml_amd import aa

    log(f"aa workload: generating {n_trees} trees")
    t0 = time.time()
    trees = aa.synth_workload(n_trees, seed=3, n_sim=1000)
    log(f"  generated in {time.time() - t0:.1f}s")
    tune = aa.autotune(ctx, trees)
    log(f"  autotune: { {k: v['waves'] for k, v in tune.items()} }")
    tb = aa.TunedBatch(ctx, trees, tune)
    ms, k_total, n = clock.run(tb.run, steps, 1, "aa_prune")
    k_ms = k_total * n / steps  # every launch of one evaluation
    got = tb.lnl.cpu().numpy()
    pn = tb.pattern_nodes
    useful_tf = 6400.0 * pn / (k_ms / 1e3) / 1e12
    executed_tf = 16384.0 * pn / (k_ms / 1e3) / 1e12
    res = {"workload": f"{n_trees} trees, taxa U{{16..256}}, patterns U{{1000..50000}}, "
                       "GTR-20 (190 exch ~ Dir(1), pi ~ Dir(2)), G4 alpha ~ U(0.3,2), "
                       "bl ~ Exp(0.1)",
           "pattern_nodes": pn, "ms": ms, "kernel_ms": k_ms, "launches": tb.launches,
           "pattern_node_per_s": pn / (ms / 1e3), "autotune": tune,
           "finite_and_nonpositive": bool(np.all(np.isfinite(got)) and np.all(got <= 0)),
           "roofline": {"bound": "mfma",
                        "traffic": sum(measured_traffic(f"k_aa_prune<{b.waves}, ") or 0.0
                                       for _, b in tb.parts) or None,
                        "achieved": useful_tf, "peak": BF16_PEAK_TF, "unit": "TFLOP/s",
                        "frac": useful_tf / BF16_PEAK_TF, "executed_tflops": executed_tf,
                        "note": "6,400 useful BF16 MFMA flop per internal pattern.node "
                                "(2 children x 4 rates x 20x20 MAC); 16,384 executed "
                                "(states padded 20 -> 32 on both sides)"}}
    log(f"  aa: step {ms:.1f} ms, kernel {k_ms:.1f} ms ({tb.launches} launch(es)) for "
        f"{pn:.3e} pattern.nodes = {pn / (ms / 1e3):.3e}/s")
    if cpu_baseline:
        from oracle import oracle as O

        threads = min(16, os.cpu_count() or 1)
        size = np.array([t.n_patterns * (t.n_leaves - 1) for t in trees])
        # error on the 8 smallest trees, the 3 nearest the median and the largest;
        # CPU time on the 3 nearest the median
        small = list(np.argsort(size)[:8])
        mid = list(np.argsort(np.abs(size - np.median(size)))[:3])
        big = [int(np.argmax(size))]
        errs, t_cpu = [], 0.0
        for i in small + mid + big:
            t = trees[i]
            c0 = time.perf_counter()
            ov, _ = O.lnl(t.n_leaves, t.children, t.root, t.blen, t.pi, t.q, t.rates, t.probs,
                          aa.tips_bitmask(t.tips), t.weights, threads=threads)
            if i in mid:
                t_cpu += time.perf_counter() - c0
            errs.append(abs(got[i] - ov) / abs(ov))
        pn_cpu = int(size[mid].sum())
        res["rel_err_vs_fp64_oracle"] = {"max": float(max(errs)), "mean": float(np.mean(errs)),
                                         "trees": [int(i) for i in small + mid + big]}
        res["cpu_baseline"] = {"value": pn_cpu / t_cpu, "unit": "pattern.node updates/s",
                               "cores": threads, "kind": "port",
                               "sample": f"oracle FP64 20-state pruning of the 3 trees nearest "
                                         f"the batch's median size ({pn_cpu:.3e} pattern.nodes, "
                                         f"{t_cpu:.2f}s), {threads} OpenMP threads"}
        log(f"  aa errors vs oracle: max {max(errs):.2e}; cpu {pn_cpu / t_cpu:.3e}/s")
    return res


def bench_spr(torch, ctx, cpu_baseline: bool):
    """Exact scores of SPR neighbours (the learned-proposal workload: 256-taxon
    tree, 50,000 GTR+G4 patterns): GPU exact evaluations per second over an
    even sample of 4,096 of the round's moves, FP64 and FP32-scaled; beside
    it the FP64 oracle on 8 of the moved trees, which must agree."""
    from phylocaml_amd import device as D, spr, spr_search as SS
    from phylocaml_amd._lib import lib

    inst = SS.make_instance(7, 256, 50_000, 10)
    c = SS.Context(ctx, inst)
    mv = spr.moves(inst.N, inst.children, inst.root)
    sample = mv[:: max(1, len(mv) // 4096)][:4096]
    rates, scores = {}, {}
    for name, fp64 in (("f64", True), ("f32", False)):
        c.exact(inst.children, inst.root, inst.blen, sample[:256], fp64=fp64,
                incremental=False)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scores[name] = c.exact(inst.children, inst.root, inst.blen, sample, fp64=fp64,
                               incremental=False)
        torch.cuda.synchronize()
        rates[name] = len(sample) / (time.perf_counter() - t0)
    res = {"workload": "256-taxon Yule tree, 50,000 unique GTR+G4 patterns; exact lnL of "
                       f"{len(sample)} of the {len(mv)} rooted SPR neighbours (even sample)",
           "moves_per_round": int(len(mv)), "exact_evaluations_per_s": rates}
    log(f"  spr: {len(mv)} moves; exact evaluations/s f64 {rates['f64']:.0f}, "
        f"f32 {rates['f32']:.0f}")
    # the whole round through the incremental scorer (repair + insertion
    # pass per pruned subtree), checked against the per-candidate scores
    idx = np.arange(0, len(mv), max(1, len(mv) // 4096))[:4096]
    inc = {}
    for name, form in (("f64", 0), ("f32", 1)):
        sc = spr.IncrementalScorer(ctx, c.dm, inst.N, c.tips, c.stride, c.wd, inst.P, form=form,
                                   s_batch=32)
        sc.score(inst.children, inst.root, inst.blen, mv[:2048])  # warm-up, workspace
        torch.cuda.synchronize()
        lib.phy_ctx_kernel_timer(ctx.handle, 0)
        t0 = time.perf_counter()
        got = sc.score(inst.children, inst.root, inst.blen, mv)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        # the same round again with per-launch event pairs, for the split
        lib.phy_ctx_kernel_timer(ctx.handle, 1)
        sc.score(inst.children, inst.root, inst.blen, mv)
        torch.cuda.synchronize()
        spans = kernel_totals(ctx, ("spr_insert", "spr_repair", "prune_down"))
        lib.phy_ctx_kernel_timer(ctx.handle, 0)
        got = got.cpu().numpy()
        diff = float(np.max(np.abs(got[idx] - scores["f64"]) / np.abs(scores["f64"])))
        inc[name] = {"round_s": dt, "moves_per_s": len(mv) / dt,
                     "max_rel_diff_vs_per_candidate_f64": diff,
                     "kernel_ms": spans,
                     "workspace_gib": sc.work.numel() / 2**30}
        log(f"  spr incremental {name}: whole round {len(mv)} moves in {dt * 1e3:.1f} ms = "
            f"{len(mv) / dt:.0f} moves/s (x{len(mv) / dt / rates[name]:.0f} per-candidate), "
            f"max rel diff {diff:.1e}; kernels {inc[name]['kernel_ms']}")
        del sc
    res["incremental"] = inc
    if cpu_baseline:
        from oracle import oracle as O

        threads = min(16, os.cpu_count() or 1)
        tips = D.unpack_nibbles(inst.tips_np, inst.P).astype(np.uint32)
        m = inst.model
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


def bench_tbr(torch, ctx):
    """Exact scores of TBR neighbours on the same instance (256-taxon tree,
    50,000 GTR+G4 patterns): the neighbourhood within join and reroot
    distance 3 through the per-candidate scorer (an even sample of 2,048
    moves) and whole through the incremental scorer, which must agree; then
    the neighbourhood with unlimited join distance (reroot <= 3) through the
    incremental scorer alone."""
    from phylocaml_amd import neighborhood as NB, spr_search as SS
    from phylocaml_amd._lib import lib

    inst = SS.make_instance(7, 256, 50_000, 10)
    c = SS.Context(ctx, inst)
    mv = NB.tbr_moves(inst.N, inst.children, inst.root, 3, 3)
    idx = np.arange(0, len(mv), max(1, len(mv) // 2048))[:2048]
    sample = mv[idx]
    rates, ref = {}, {}
    for name, form in (("f64", 0), ("f32", 1)):
        sc = NB.Scorer(ctx, c.dm, inst.N, c.tips, c.stride, c.wd, inst.P, form=form, chunk=1024,
                       incremental=False)
        sc.score(inst.children, inst.root, inst.blen, sample[:256])  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ref[name] = sc.score(inst.children, inst.root, inst.blen, sample).cpu().numpy()
        rates[name] = len(sample) / (time.perf_counter() - t0)
        del sc
    res = {"workload": "256-taxon Yule tree, 50,000 unique GTR+G4 patterns; exact lnL of the "
                       f"{len(mv)} TBR neighbours within join and reroot distance 3",
           "moves": int(len(mv)), "per_candidate_evaluations_per_s": rates}
    inc = {}
    for name, form in (("f64", 0), ("f32", 1)):
        sc = NB.Scorer(ctx, c.dm, inst.N, c.tips, c.stride, c.wd, inst.P, form=form,
                       incremental=True, s_batch=32)
        sc.score(inst.children, inst.root, inst.blen, mv[:4096])  # warm-up, workspace
        sc.score(inst.children, inst.root, inst.blen, mv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        got = sc.score(inst.children, inst.root, inst.blen, mv)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        lib.phy_ctx_kernel_timer(ctx.handle, 1)
        sc.score(inst.children, inst.root, inst.blen, mv)
        torch.cuda.synchronize()
        spans = kernel_totals(ctx, ("tbr_insert", "tbr_reroot", "spr_repair", "prune_down"))
        lib.phy_ctx_kernel_timer(ctx.handle, 0)
        got = got.cpu().numpy()
        diff = float(np.max(np.abs(got[idx] - ref["f64"]) / np.abs(ref["f64"])))
        inc[name] = {"round_s": dt, "moves_per_s": len(mv) / dt,
                     "vs_per_candidate": len(mv) / dt / rates[name],
                     "max_rel_diff_vs_per_candidate_f64": diff, "kernel_ms": spans,
                     "workspace_gib": sc.iwork.numel() / 2**30}
        log(f"  tbr incremental {name}: {len(mv)} moves in {dt * 1e3:.1f} ms = "
            f"{len(mv) / dt:.0f} moves/s (x{len(mv) / dt / rates[name]:.0f} per-candidate "
            f"{rates[name]:.0f}/s), max rel diff {diff:.1e}; kernels {spans}")
        del sc
    res["incremental"] = inc
    # unlimited join distance: every regraft of every rerooting within 3
    big = NB.tbr_moves(inst.N, inst.children, inst.root, -1, 3)
    sc = NB.Scorer(ctx, c.dm, inst.N, c.tips, c.stride, c.wd, inst.P, form=1, incremental=True,
                   s_batch=32)
    sc.score(inst.children, inst.root, inst.blen, big)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got = sc.score(inst.children, inst.root, inst.blen, big).cpu().numpy()
    dt = time.perf_counter() - t0
    res["incremental_unlimited_join_f32"] = {
        "moves": int(len(big)), "round_s": dt, "moves_per_s": len(big) / dt,
        "finite": bool(np.all(np.isfinite(got))), "workspace_gib": sc.iwork.numel() / 2**30}
    log(f"  tbr incremental f32, join unlimited: {len(big)} moves in {dt * 1e3:.1f} ms = "
        f"{len(big) / dt:.0f} moves/s")
    del sc
    return res


def bench_search(ctx):
    """Learned proposal against exhaustive exact scoring, both to convergence,
    on the bounded instance (64 taxa, 5,000 patterns, start = true tree after
    6 random SPR moves; MLP trained on 4 other instances)."""
    from phylocaml_amd import spr_search as SS

    t0 = time.time()
    out = SS.compare(ctx, 64, 5_000, 6, train_instances=4, top_k=64, fp64=True)
    out["total_s"] = time.time() - t0
    log(f"  search: exhaustive {out['exhaustive']['exact_evaluations']} evaluations "
        f"{out['exhaustive']['wall_s']:.1f}s, guided {out['guided']['exact_evaluations']} "
        f"{out['guided']['wall_s']:.1f}s (x{out['exact_evaluation_ratio']:.1f}), "
        f"lnL gap {out['final_lnl_gap']:.3g}")
    return out


def bench_newton(torch):
    """Branch-length Newton on a 64-taxon / 2,000-pattern GTR+G4 case, where
    launches and host round trips weigh: time per iteration of the host loop
    (optimize_branches: derivative pass, three copies to the host, host
    step), of the device loop issued eagerly (phy_newton_propose, phy_lk_tree,
    phy_newton_accept), and of the same iterations replayed from a HIP graph."""
    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM, optimize as OPT, synth

    s = torch.cuda.Stream()
    ctx = D.Context(0, stream=s.cuda_stream)
    N, P, K = 64, 2000, 20
    with torch.cuda.stream(s):
        tree = synth.yule(N, 4, 0.08)
        pi, co = synth.gtr_params(4, 5.0, 4)
        model = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi)))
        st = synth.simulate(tree, model.priors, model.q, model.rates, P, 4)
        topo = D.Topology(N, tree.children, tree.root, ctx)
        dm = LK.DeviceModel(ctx, model)
        tips, wd = D.to_device(synth.pack_dna(st)), D.to_device(synth.poisson_weights(P, 3.0, 4))
        start = np.full(2 * N - 1, 0.1)
        start[tree.root] = 0.0
        obj = OPT.DeviceObjective(ctx, topo, tips, D.words(P), wd, P)
        OPT.optimize_branches(obj, model, start, tree.children, N, tree.root, max_iter=2)
        t0 = time.perf_counter()
        host = OPT.optimize_branches(obj, model, start, tree.children, N, tree.root, max_iter=K)
        host_ms = (time.perf_counter() - t0) * 1e3 / max(host.derivative_passes, 1)
        dn = OPT.DeviceNewton(ctx, topo, dm, tips, D.words(P), wd, P, start)
        dn.run(2)
        ctx.sync()
        t0 = time.perf_counter()
        dn.run(K)
        ctx.sync()
        eager_ms = (time.perf_counter() - t0) * 1e3 / K
        dg = OPT.DeviceNewton(ctx, topo, dm, tips, D.words(P), wd, P, start)
        g = dg.capture(K)
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - t0) * 1e3 / K
        _, lg, _, _ = dg.result()
    out = {"workload": f"{N}-taxon Yule tree, {P:,} GTR+G4 patterns, FP64, every branch from 0.1",
           "host_loop_ms_per_pass": host_ms, "device_eager_ms_per_iteration": eager_ms,
           "graph_ms_per_iteration": graph_ms, "host_final_lnl": host.lnl,
           "graph_lnl_after_two_replays": lg}
    log(f"  newton: host loop {host_ms:.3f} ms/pass, device eager {eager_ms:.3f} ms/iteration, "
        f"graph replay {graph_ms:.3f} ms/iteration")
    return out


def bench_mlp(torch, ctx, clock):
    """MLP 32 -> 128 -> 128 -> 1 on 65,536 moves: forward and backward
    kernels (BF16 MFMA GEMMs, FP32 accumulation)."""
    from phylocaml_amd import device as D
    from phylocaml_amd._lib import lib
    from phylocaml_amd.mlp import MLP

    B = 65_536
    rng = np.random.default_rng(0)
    net = MLP(ctx, 32, 128, 0)
    x = D.to_device(rng.standard_normal((B, 32)).astype(np.float32))
    t = D.to_device(rng.standard_normal(B).astype(np.float32))
    y = torch.empty(B, dtype=torch.float32, device=f"cuda:{ctx.device}")
    def fwd():
        ctx.check(lib.phy_mlp_forward(ctx.handle, net.handle, D.dptr(x), B, D.dptr(y)), "mlp")

    def bwd():
        ctx.check(lib.phy_mlp_backward(ctx.handle, net.handle, D.dptr(x), D.dptr(t), B,
                                       D.dptr(net._grads), D.dptr(net._loss)), "mlp")

    _, f_ms, _ = clock.run(fwd, 20, 3, "mlp_forward")
    _, b_ms, _ = clock.run(bwd, 20, 3, "mlp_backward")
    f_tf = MLP_FWD_FLOPS * B / (f_ms / 1e3) / 1e12
    b_tf = MLP_BWD_FLOPS * B / (b_ms / 1e3) / 1e12
    log(f"  mlp: forward {f_ms * 1e3:.1f} us ({B / (f_ms / 1e3):.3e} moves/s, {f_tf:.1f} TF/s), "
        f"backward {b_ms * 1e3:.1f} us ({b_tf:.1f} TF/s)")
    net.close()
    return {"batch_moves": B,
            "forward": {"kernel_ms": f_ms, "moves_per_s": B / (f_ms / 1e3), "tflops": f_tf,
                        "mfma_frac": f_tf / BF16_PEAK_TF},
            "backward": {"kernel_ms": b_ms, "moves_per_s": B / (b_ms / 1e3), "tflops": b_tf,
                         "mfma_frac": b_tf / BF16_PEAK_TF},
            "note": "forward 41,216 flop/move, backward 73,984 (weight and input gradients, "
                    "after its own forward); the GEMMs are 16x16 BF16 MFMA tiles of one wave, "
                    "so at 128 hidden units the launches are latency-bound"}


def bench_config0(ctx):
    """The reference's own CPU-runnable case: 32-taxon Yule tree (mean bl
    0.1), 2,000 GTR+G4 sites (alpha 0.5, pi ~ Dir(5,5,5,5)) compressed to
    patterns; the oracle's FP64 lnL and Fitch per-site/total costs on 1
    thread (the reference's execution model), median of 10 after 5 warm-ups,
    against the same on the GPU."""
    from oracle import oracle as O
    from phylocaml_amd import device as D, likelihood as LK, mlmodel as MM
    from phylocaml_amd import nonadditive as NA, synth

    tree = synth.yule(32, 0, 0.1)
    pi, co = synth.gtr_params(4, 5.0, 0)
    model = MM.create(MM.Spec(4, ("GTR", co), ("DiscreteGamma", 4, 0.5), ("Empirical", pi)))
    st = synth.simulate(tree, model.priors, model.q, model.rates, 2000, 0)
    pats, w = synth.compress(st)
    P = pats.shape[1]
    tips = (1 << pats).astype(np.uint32)
    sets = (1 << st).astype(np.uint8)

    def cpu():
        v, _ = O.lnl(32, tree.children, tree.root, tree.blen, model.priors, model.q,
                     model.rates, model.probs, tips, w, threads=1)
        _, sc, _, tot = O.fitch_downpass(32, tree.children, tree.root, sets, threads=1)
        return v, tot

    for _ in range(5):
        cpu()
    ts = []
    for _ in range(10):
        c0 = time.perf_counter()
        ov, otot = cpu()
        ts.append(time.perf_counter() - c0)
    topo = D.Topology(32, tree.children, tree.root, ctx)
    dm = LK.DeviceModel(ctx, model)
    ev = LK.Evaluator(ctx, topo, dm, P, 0)
    ft = NA.alloc_tree(ctx, topo, 2000, D.pack_nibbles(sets))
    args = (D.to_device(D.pack_nibbles(tips.astype(np.uint8))), D.words(P), D.to_device(w),
            D.to_device(tree.blen))
    gts = []
    for k in range(15):
        ctx.sync()
        c0 = time.perf_counter()
        ev.run(*args)
        NA.downpass(ctx, topo, ft)
        ctx.sync()
        if k >= 5:
            gts.append(time.perf_counter() - c0)
    gl, gt = float(ev.lnl.cpu()[0]), int(ft.total.cpu()[0])
    res = {"workload": f"32-taxon Yule tree, 2,000 GTR+G4 sites ({P} patterns): FP64 lnL and "
                       "Fitch per-site and total costs",
           "cpu_ms_median": 1e3 * float(np.median(ts)), "cpu_threads": 1,
           "gpu_ms_median": 1e3 * float(np.median(gts)),
           "lnl_oracle": ov, "lnl_gpu": gl, "lnl_rel_err": abs(gl - ov) / abs(ov),
           "fitch_total_oracle": int(otot), "fitch_total_gpu": gt,
           "note": "wall time per evaluation including launches and one synchronisation "
                   "(GPU), median of 10 after 5 warm-ups"}
    log(f"  config 0: cpu {res['cpu_ms_median']:.2f} ms (1 thread), gpu "
        f"{res['gpu_ms_median']:.2f} ms, lnL err {res['lnl_rel_err']:.1e}, Fitch "
        f"{gt} vs {otot}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only", choices=["all", "prune", "fitch", "aa", "spr", "tbr", "search",
                                       "mlp", "copy", "newton"], default="all",
                    help="one part only (the pruning line is always measured unless the part "
                         "is copy or mlp)")
    ap.add_argument("--aa-trees", type=int, default=2048)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): every path "
              "here runs on a single GPU; sharding across GPUs is not implemented",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from phylocaml_amd import device as D

    ctx = D.Context(0)
    clock = Clock(torch, ctx)
    steps, warm = args.steps, args.warmup
    cpu_on = not args.no_cpu_baseline
    everything = args.only == "all"

    copy = stream_copy(torch, ctx, clock)
    if args.only in ("copy", "mlp"):
        print(json.dumps({"copy_stream": copy,
                          "mlp": bench_mlp(torch, ctx, clock) if args.only == "mlp" else None}),
              flush=True)
        return
    pr = bench_prune(torch, ctx, clock, steps, warm, cpu_on)
    units = pr["units"]
    fitch = bench_fitch(torch, ctx, clock, steps, warm, copy["read_plus_write_GBs"], cpu_on) \
        if args.only in ("all", "fitch") else None
    aa_res = bench_aa(torch, ctx, clock, args.aa_trees, max(2, steps // 5), cpu_on) \
        if args.only in ("all", "aa") else None
    spr_res = bench_spr(torch, ctx, cpu_on) if args.only in ("all", "spr") else None
    tbr_res = bench_tbr(torch, ctx) if args.only in ("all", "tbr") else None
    search = bench_search(ctx) if args.only in ("all", "search") else None
    mlp = bench_mlp(torch, ctx, clock) if everything else None
    newton = bench_newton(torch) if args.only in ("all", "newton") else None
    config0 = bench_config0(ctx) if everything and cpu_on else None

    cpu = None
    if cpu_on:
        from oracle import oracle as O

        tree, model, st, w = pr["oracle_inputs"]
        N, P = pr["N"], pr["P"]
        model_name, ncpu = cpu_info()
        threads = min(16, ncpu)
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

    ms_main, k_main = pr["step_ms"]["f32_fused"], pr["kernel_ms"]["f32_fused"]
    value = units / (ms_main / 1e3)
    tflops = PRUNE_FLOPS * units / (k_main / 1e3) / 1e12
    # the 512 x 200k pass in split form: the parts side by side (one launch),
    # then the top program; both grids are padded for the XCD-aware tile
    # order, and the traffic file keeps each kernel's largest launches
    grid = ((pr["P"] + 255) // 256) * 256 * 4
    # (the fused pass's kernels: STORE = false; the stored form's end in true)
    t_parts = measured_traffic("k_prune_part<float, 4,", suffix="false>")
    t_top = measured_traffic("k_prune_down<float, 4, 1, true, 0, 2, false>")
    traffic = None if t_parts is None or t_top is None else t_parts + t_top
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
                   "n_taxa": pr["N"], "n_patterns": pr["P"], "n_rates": 4, "form": "f32_fused"},
        "roofline": {"bound": "valu", "achieved": tflops, "peak": FP32_PEAK_TF,
                     "unit": "TFLOP/s", "frac": tflops / FP32_PEAK_TF, "traffic": traffic,
                     "kernel_ms": k_main,
                     "frac_of_nonpacked_fp32": tflops / FP32_NOPACK_TF,
                     "hbm_GBs": None if traffic is None else traffic / (k_main / 1e3) / 1e9,
                     "hbm_frac": None if traffic is None
                     else traffic / (k_main / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "note": "achieved = 272 flop per internal pattern.node x 511 nodes x "
                             "200,000 patterns / mean duration of the fused pass (HIP events "
                             "around its two launches: k_prune_part, the tree's 4 parts "
                             "side by side, then the top program k_prune_down<float, 4, 1, "
                             "true, 0, 2>); packed FP32 is kept out of the build, so "
                             "78.6 TFLOP/s is the reachable vector peak; traffic = HBM bytes "
                             "per pass from rocprofv3 FETCH_SIZE + WRITE_SIZE "
                             "(copy-calibrated) of both kernels: interior partials stay "
                             "on chip, the parts' root partials go through HBM"},
        "cpu_baseline": cpu,
        "prune": {
            "step_ms": pr["step_ms"], "kernel_ms": pr["kernel_ms"],
            "updates_per_s": {k: units / (v / 1e3) for k, v in pr["step_ms"].items()},
            "fp32_vs_fp64_plain": pr["step_ms"]["f64_plain"] / ms_main,
            "fp32_vs_fp64_fused": pr["step_ms"]["f64_fused"] / ms_main,
            "lnl": pr["lnl"], "errors_vs_fp64_oracle": pr["errors"],
            "derivatives": {
                name: {**d, "roofline": deriv_roofline(
                    d["kernel_ms"], units, FP32_PEAK_TF if name == "f32" else FP64_PEAK_TF,
                    f"k_prune_deriv<{'float' if name == 'f32' else 'double'},", grid)}
                for name, d in pr["deriv"].items()},
        },
        "copy_stream": copy,
        "fitch": fitch,
        "aa": aa_res,
        "spr": spr_res,
        "tbr": tbr_res,
        "search": search,
        "mlp": mlp,
        "newton": newton,
        "config0": config0,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
