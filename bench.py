"""Benchmark: pixel-pairs/s of the xi(r_par, r_perp) pair histograms on one MI355X.

Workload (default): 100k synthetic forests over 2,500 deg^2, z_q ~ U[2.1, 3.5],
HEALPix nside 16, 50x50 bins of 4 Mpc/h, w and delta stored as float32
(recipe in lya2pcf_amd/synth.py).  A step is one replay of the captured graph
(zero outputs + persistent kernel) over the whole pair list, with every input
already in HBM.  value = P / (seconds per step), P = sum of n1*n2 over the
forest-pair list (pruned pairs included).

The JSON line carries the headline (mixed-precision) kernel and, alongside, the
plain FP64 kernel on the same data, their bin-count agreement and the
xi error, the roofline of the mixed kernel (VALU-bound: 10 algorithmic FLOP per
pixel pair against the FP32 vector peak) and a CPU baseline: the oracle's
restatement of the reference CPU loop, timed on a bounded sample on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixel-pairs/s for ξ(r∥,r⊥); mixed vs plain FP64 kernel at ≤1e-6 rel err"
FLOP_PER_PAIR = 10.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_graph(g, eng, steps, warmup, torch):
    """Wall time per step over exactly `steps` replays and per-launch event times."""
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        g.launch(stream)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        g.launch(stream)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    ev = [s.elapsed_time(e) / 1e3 for s, e in zip(starts, ends)]
    return wall, ev


def lds_atomic_rates(torch):
    """LDS atomic adds per CU per clock for uint32, float and double (lya_debug_lds_atomic):
    why FP32 LDS partials do not pay on this part."""
    from lya2pcf_amd import _lib
    props = torch.cuda.get_device_properties(0)
    cus = props.multi_processor_count
    clock_hz = 2.4e9  # MI355X peak engine clock
    blocks, iters = cus * 4, 4096
    out = {}
    for t, name, dt in ((0, "u32", torch.int32), (1, "f32", torch.float32), (2, "f64", torch.float64)):
        buf = torch.empty(blocks, dtype=dt, device="cuda")
        for rep in range(2):  # the first launch loads the kernel
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            _lib.check(_lib.lib().lya_debug_lds_atomic(t, blocks, iters, _lib.ptr(buf), _lib.stream_ptr()))
            e1.record()
            torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3
        out[name] = blocks * 512.0 * iters / sec / cus / clock_hz
    log("LDS atomic adds per CU per clock: %s" % {k: round(v, 2) for k, v in out.items()})
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_c0(threads):
    """The restated reference CPU loop on the 2,000-forest 10 x 10 deg workload (the
    reference's own CPU-runnable case): all of it, on `threads` cores and on one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    from parameters import chiquito
    from lya2pcf_amd import forests, synth, xi
    cfg = synth.CONFIGS["patch2k"]
    fs = forests.ForestSet.from_survey(synth.generate(cfg))
    am = forests.angmax_from(fs, 200.0)
    off, idx = xi.build_pairs(fs, am)
    n = np.diff(fs.pix_off)
    f1 = np.repeat(np.arange(fs.nf), np.diff(off))
    P = float((n[f1] * n[idx]).sum())
    out = {"workload": cfg.name, "forests": fs.nf, "pixel_pairs": P}
    for key, nt in (("all_cores", threads), ("one_core", 1)):
        t0 = time.perf_counter()
        oracle.xi_cpu_form(fs, off, idx, am, chiquito, threads=nt)
        dt = time.perf_counter() - t0
        out[key] = {"value": P / dt, "seconds": dt, "cores": nt}
        log("cpu baseline c0 (%d threads): %.3g pixel-pairs/s, %.1f s" % (nt, P / dt, dt))
    return out


def cpu_baseline(fs, eng, angmax, budget_pairs, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    from parameters import chiquito
    n = np.diff(fs.pix_off)
    nb_cnt = np.diff(eng.nb_off)
    # cost per forest 1 (sum of n1*n2 over its neighbours)
    f1 = np.repeat(np.arange(fs.nf), nb_cnt)
    cost = np.bincount(f1, weights=(n[f1] * n[eng.nb_idx]).astype(np.float64), minlength=fs.nf)
    # a spread-out sample: every k-th forest until the budget is reached
    order = np.random.default_rng(0).permutation(fs.nf)
    csum = np.cumsum(cost[order])

    def run(budget, nthreads):
        m = int(np.searchsorted(csum, budget)) + 1
        sample = np.sort(order[:m]).astype(np.int32)
        P = float(cost[sample].sum())
        t0 = time.perf_counter()
        oracle.xi_cpu_form(fs, eng.nb_off, eng.nb_idx, angmax, chiquito, f1_list=sample, threads=nthreads)
        return m, P, time.perf_counter() - t0

    m, P, dt = run(budget_pairs, threads)
    m1, P1, dt1 = run(budget_pairs / 40, 1)  # the reference's numba loop runs on one core per rank
    return {"value": P / dt, "unit": "pixel-pairs/s", "cores": threads, "kind": "port",
            "sample": "%d of %d forest-1 lines (%.3g pixel pairs, %.1f s), reference CPU formula "
                      "(correlation_procedures_cpu.py:46-84) restated in C, FP64, OpenMP over forest 1"
                      % (m, fs.nf, P, dt),
            "single_core": {"value": P1 / dt1, "cores": 1,
                            "sample": "%d forest-1 lines (%.3g pixel pairs, %.1f s)" % (m1, P1, dt1)}}


def bench_distortion(args, emit=True):
    """Distortion matrix (5% subsample, FP64 accumulation) and one chi^2 gradient on
    20k forests: wall times of the device work (reported, not the headline).  With
    emit=False the figures are returned instead of printed as a JSON line."""
    import random
    import numpy as np
    import torch
    from lya2pcf_amd import adjoint, forests, synth
    from lya2pcf_amd import distortion as dist
    cfg = synth.CONFIGS["dist20k"]
    if args.forests:
        cfg = synth.scaled(cfg, args.forests)
    s = synth.generate(cfg)
    fs = forests.ForestSet.from_survey(s, float32_inputs=True)
    am = forests.angmax_from(fs, 200.0)
    eng = dist.DistortionEngine(fs, am)
    pairs = dist.subsample(fs, eng.nb_off, eng.nb_idx, range(fs.nf), 0.95, rng=random.Random(1))
    n = np.diff(fs.pix_off)
    P = float((n[pairs[:, 0]] * n[pairs[:, 1]]).sum())
    times = []
    for k in range(args.warmup + args.steps):
        eng.zero()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.accumulate(pairs)
        torch.cuda.synchronize()
        if k >= args.warmup:
            times.append(time.perf_counter() - t0)
    D, wA = eng.result()
    td = float(np.median(times))
    log("distortion: %d subsampled forest pairs, %.4g pixel pairs, %.3f s" % (len(pairs), P, td))
    # chi^2 gradient: forward pair sums + two adjoint walks, FP64 geometry (on MI355X it
    # runs at the FP32 rate and needs no edge recheck; both modes give the same bins)
    ae = adjoint.AdjointEngine(fs, am, mode="fp64", pairs=(eng.nb_off, eng.nb_idx))
    dev = "cuda"
    wm = adjoint.WeightModel(s.z, s.sigma2_pip, cfg.z_nodes, dev)
    pr = adjoint.Projection(s.pix_off, s.loglam, dev)
    m = torch.from_numpy(adjoint.model_xi()).to(dev)
    delta = torch.from_numpy(s.delta.astype(np.float32).astype(np.float64)).to(dev)
    tg = []
    for k in range(args.warmup + args.steps):
        eta = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
        s2 = torch.full((10,), 0.05, dtype=torch.float64, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        w = wm(eta, s2)
        W, DW = adjoint.pair_sums(pr(delta, w), w, ae)
        adjoint.chi2(W, DW, m, 1e-4).backward()
        torch.cuda.synchronize()
        if k >= args.warmup:
            tg.append(time.perf_counter() - t0)
    tgm = float(np.median(tg))
    P2 = ae.fwd.pixel_pairs
    log("chi2 gradient: %.3f s for %.4g pixel pairs (forward + 2 adjoint walks)" % (tgm, P2))
    if not emit:
        return {"workload": cfg.name, "forests": fs.nf, "subsampled_pairs": int(len(pairs)), "pixel_pairs": P,
                "distortion_s": td, "distortion_pixel_pairs_per_s": P / td,
                "chi2_gradient_s": tgm, "gradient_pixel_pairs": P2, "gradient_mode": "fp64"}
    print(json.dumps({"metric": "distortion matrix wall time (5% subsample, FP64)", "value": td, "unit": "s",
                      "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "data": "synthetic", "dtype": "f64",
                      "config": {"workload": cfg.name, "forests": fs.nf, "subsampled_pairs": int(len(pairs)),
                                 "pixel_pairs": P},
                      "pixel_pairs_per_s": P / td,
                      "rows_with_weight": int((wA > 0).sum()),
                      "chi2_gradient": {"seconds": tgm, "pixel_pairs": P2, "pixel_pairs_per_s": 3 * P2 / tgm,
                                        "mode": "fp64",
                                        "note": "one forward and two adjoint walks of the full pair list"}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="sky100k", choices=["patch2k", "sky100k", "sky300k", "dist20k"],
                    help="sky100k (default): the xi headline; patch2k/sky300k: xi on those surveys; dist20k: distortion + gradient")
    ap.add_argument("--forests", type=int, default=0, help="override the number of forests (same density)")
    ap.add_argument("--cpu-pairs", type=float, default=1.5e11, help="CPU baseline sample size in pixel pairs")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-distortion", action="store_true",
                    help="skip the 20k-forest distortion and gradient figures of the default line")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--counts", action="store_true", help="time with per-bin pair counts on")
    ap.add_argument("--only", choices=["mixed", "fp64", "fp64-walk", "mixed32"], default=None,
                    help="time one kernel only (profiling)")
    ap.add_argument("--ppl", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--replicas", type=int, default=0)
    ap.add_argument("--row-pad", type=int, default=-1)
    ap.add_argument("--tail-max", type=int, default=0,
                    help="longest forest-1 tail walked against forest 2 in lanes (0: tuned/default, -1: whole "
                         "slabs of 64*px_per_lane pixels only)")
    ap.add_argument("--copy-calibrate", type=int, default=0,
                    help="after timing, run one plain device copy of this many bytes (calibrates the "
                         "profiler's HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (xi would shard by output "
              "HEALPix pixel with no exchange step, the distortion with one 50 MB sum-reduce); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from lya2pcf_amd import autotune, forests, synth, xi

    torch.cuda.init()
    if args.config == "dist20k":
        return bench_distortion(args)
    cfgname = args.config
    cfg = synth.CONFIGS[cfgname]
    if args.forests:
        cfg = synth.scaled(cfg, args.forests)
    t0 = time.perf_counter()
    survey = synth.generate(cfg)
    fs = forests.ForestSet.from_survey(survey, float32_inputs=True)
    del survey
    angmax = forests.angmax_from(fs, 200.0)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    eng = xi.XiEngine(fs, angmax, pair_builder="device")
    t_setup = time.perf_counter() - t0
    P = eng.pixel_pairs
    log("data: %d forests, %d pixels, %d forest pairs, P=%.4g pixel pairs (gen %.1fs, upload+pairs %.1fs)"
        % (fs.nf, fs.n_pix, eng.n_pairs, P, t_gen, t_setup))
    # set-up outside the timed region: pair list and work items on the GPU, and the
    # host builders (OpenMP) for comparison
    setup = {"pair_list_gpu_s": eng.setup_seconds["pairs"], "work_items_gpu_s": eng.setup_seconds["items"]}
    if not args.no_cpu:
        t0 = time.perf_counter()
        off_h, idx_h = xi.build_pairs(fs, angmax)
        setup["pair_list_host_s"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        xi.build_work(fs, off_h, idx_h, eng.f1_list, eng.item_cost, 1 << 62)
        setup["work_items_host_s"] = time.perf_counter() - t0
        setup["host_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))
        setup["pair_lists_identical"] = bool(np.array_equal(off_h, eng.nb_off) and np.array_equal(idx_h, eng.nb_idx))
        del off_h, idx_h
    log("setup: %s" % setup)

    def shape_for(mode):
        walk = 1 if mode == "fp64-walk" else (0 if mode == "fp64" else None)
        base = "fp64" if mode.startswith("fp64") else mode
        shape = dict(autotune.load_best(base, cfgname, walk=walk) or {})
        if mode == "fp64-walk":
            shape = {"walk": 1, "waves": shape.get("waves", 16)}
        for k, a in (("px_per_lane", args.ppl), ("waves", args.waves), ("replicas", args.replicas)):
            if a and not shape.get("walk"):
                shape[k] = a
        if args.row_pad >= 0:
            shape["row_pad"] = args.row_pad
        if args.tail_max and not shape.get("walk"):
            shape["tail_max"] = args.tail_max
        return base, shape

    # Timed launches compute the reference's outputs (sum w1 w2, sum dw1 dw2); one
    # further untimed launch per kernel adds the per-bin pair counts for the
    # bit-exact comparison between kernels.
    out = {}
    modes = [args.only] if args.only else (["mixed"] if args.no_fp64 else ["mixed", "fp64", "fp64-walk"])
    for mode in modes:
        base, shape = shape_for(mode)
        g = eng.graph(xi.default_config(base, counts=args.counts, **shape))
        wall, ev = time_graph(g, eng, args.steps, args.warmup, torch)
        res = eng.results()
        g.close()
        t_cnt = None
        if not args.counts:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            eng.run(xi.default_config(base, counts=True, **shape))
            e1.record()
            torch.cuda.synchronize()
            t_cnt = e0.elapsed_time(e1) / 1e3
            res = eng.results()
        work = None if mode == "fp64-walk" else eng.work_stats(xi.default_config(base, **shape))
        out[mode] = {"wall": wall, "ev": ev, "res": res, "shape": shape, "with_counts_s": t_cnt, "work": work,
                     "binned": float(res[0].sum())}
        log("%s %s: %.4f s/step (events median %.4f s) -> %.4g pixel-pairs/s; binned %.4g, evaluated %s%s"
            % (mode, shape, wall, float(np.median(ev)), P / wall, out[mode]["binned"],
               "%.4g" % work["evaluated_pairs"] if work else "n/a",
               "; tasks %d, non-empty %d" % (work["tasks"], work["tasks_nonempty"]) if work else ""))

    if args.copy_calibrate:
        from lya2pcf_amd import _lib
        nbytes = (args.copy_calibrate // 16) * 16
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        _lib.check(_lib.lib().lya_copy_bytes(_lib.ptr(src), _lib.ptr(dst), nbytes, _lib.stream_ptr()))
        torch.cuda.synchronize()
        log("copy calibration: %d bytes" % nbytes)

    main_mode = "mixed" if "mixed" in out else modes[0]
    m = out[main_mode]
    value = P / m["wall"]
    ev_med = float(np.median(m["ev"]))
    achieved = P * FLOP_PER_PAIR / ev_med / 1e12
    peak = FP64_PEAK_TFLOPS if main_mode.startswith("fp64") else FP32_PEAK_TFLOPS
    B = m["binned"]
    achieved_binned = B * FLOP_PER_PAIR / ev_med / 1e12
    line = {
        "metric": METRIC, "value": value, "unit": "pixel-pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": m["wall"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if main_mode.startswith("fp64") else "f32+f64", "data": "synthetic",
        "config": {"workload": cfg.name, "forests": fs.nf, "pixels": fs.n_pix, "forest_pairs": int(eng.n_pairs),
                   "pixel_pairs": P, "bins": "50x50 of 4 Mpc/h", "kernel": main_mode, "shape": m["shape"],
                   "outputs": "sum w1w2, sum dw1dw2" + (" and pair counts" if args.counts else ""),
                   "parallelism": "single GPU, persistent work queue"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     # HBM bytes per launch measured with rocprofv3 counters on the 100k-forest
                     # workload only; other workloads report none
                     "traffic": autotune.measured_traffic(main_mode) if args.config == "sky100k" and
                     not args.forests else None,
                     # VALU issue and LDS figures of the same kernel and workload (rocprofv3 counters)
                     "counters": autotune.measured_counters(main_mode) if args.config == "sky100k" and
                     not args.forests else None,
                     "pixel_pairs": P, "binned_pairs": B,
                     "evaluated_pairs": m["work"]["evaluated_pairs"] if m["work"] else None,
                     "binned_per_pixel_pair": B / P,
                     "evaluated_per_pixel_pair": m["work"]["evaluated_pairs"] / P if m["work"] else None,
                     "achieved_binned": achieved_binned, "frac_binned": achieved_binned / peak,
                     "note": "achieved/frac: 10 FLOP per pixel pair of P (algorithmic, pruned pairs included) / "
                             "median per-launch hipEvent time; achieved_binned/frac_binned: 10 FLOP per binned "
                             "pair (sum of the per-bin counts); evaluated_pairs: pixel-pair slots the kernel steps "
                             "through after pruning (lya_xi_work_stats); no MFMA or HBM bound applies"},
        "kernels": {k: {"value": P / v["wall"], "ms_per_step": v["wall"] * 1e3,
                        "event_median_ms": float(np.median(v["ev"])) * 1e3, "shape": v["shape"],
                        "with_counts_ms": None if v["with_counts_s"] is None else v["with_counts_s"] * 1e3,
                        "binned_pairs": v["binned"],
                        "evaluated_pairs": v["work"]["evaluated_pairs"] if v["work"] else None,
                        "tasks": v["work"]["tasks"] if v["work"] else None,
                        "tasks_nonempty": v["work"]["tasks_nonempty"] if v["work"] else None,
                        "frac_binned": v["binned"] * FLOP_PER_PAIR / float(np.median(v["ev"])) / 1e12 /
                        (FP64_PEAK_TFLOPS if k.startswith("fp64") else FP32_PEAK_TFLOPS)}
                    for k, v in out.items()},
        "setup": setup,
    }
    if line["roofline"]["traffic"]:
        # achieved HBM rate: the measured bytes per launch over this run's launch time
        line["roofline"]["hbm_gbps"] = line["roofline"]["traffic"] / ev_med / 1e9
    fp64s = [k for k in ("fp64", "fp64-walk") if k in out]
    if fp64s and "mixed" in out:
        best64 = min(fp64s, key=lambda k: out[k]["wall"])
        c64, w64, d64 = out[best64]["res"]
        cm, wm, dm = out["mixed"]["res"]
        with np.errstate(divide="ignore", invalid="ignore"):
            x64 = np.where(w64 > 0, d64 / w64, 0.0)
            xm = np.where(wm > 0, dm / wm, 0.0)
        W64, Wm = w64.sum(0), wm.sum(0)
        X64 = np.where(W64 > 0, d64.sum(0) / np.where(W64 > 0, W64, 1), 0.0)
        Xm = np.where(Wm > 0, dm.sum(0) / np.where(Wm > 0, Wm, 1), 0.0)
        sel = np.abs(X64) > 0
        line["fp64"] = {"kernel": best64, "value": P / out[best64]["wall"],
                        "ms_per_step": out[best64]["wall"] * 1e3, "shape": out[best64]["shape"],
                        "speedup_mixed_over_fp64": out[best64]["wall"] / out["mixed"]["wall"],
                        "counts_bit_exact": bool(np.array_equal(c64, cm)),
                        "counts_bit_exact_all_kernels": bool(all(np.array_equal(out[k]["res"][0], cm)
                                                                 for k in fp64s)),
                        "xi_total_max_rel_err": float(np.max(np.abs(Xm - X64)[sel] / np.abs(X64[sel]))),
                        "xi_per_pixel_max_abs_err": float(np.max(np.abs(xm - x64)))}
    if "mixed32" in out and "fp64" in out:
        c64, w64, d64 = out["fp64"]["res"]
        c3, w3, d3 = out["mixed32"]["res"]
        W64, W3 = w64.sum(0), w3.sum(0)
        X64 = np.where(W64 > 0, d64.sum(0) / np.where(W64 > 0, W64, 1), 0.0)
        X3 = np.where(W3 > 0, d3.sum(0) / np.where(W3 > 0, W3, 1), 0.0)
        sel = np.abs(X64) > 0
        line["mixed32"] = {"value": P / out["mixed32"]["wall"], "ms_per_step": out["mixed32"]["wall"] * 1e3,
                           "counts_bit_exact": bool(np.array_equal(c64, c3)),
                           "xi_total_max_rel_err": float(np.max(np.abs(X3 - X64)[sel] / np.abs(X64[sel])))}
    line["lds_atomic_adds_per_cu_clock"] = lds_atomic_rates(torch)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("cpu baseline: %d threads, ~%.2g pixel pairs" % (threads, args.cpu_pairs))
        line["cpu_baseline"] = cpu_baseline(fs, eng, angmax, args.cpu_pairs, threads)
        line["cpu_baseline"]["cpu_model"] = cpu_model()
        line["cpu_baseline"]["os_cpu_count"] = os.cpu_count()
        line["cpu_baseline"]["c0"] = cpu_baseline_c0(threads)
    if args.config == "sky100k" and not args.forests and not args.no_distortion and not args.only:
        del eng
        torch.cuda.empty_cache()
        args.forests = 0
        line["distortion"] = bench_distortion(args, emit=False)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
