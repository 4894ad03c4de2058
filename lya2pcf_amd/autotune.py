"""Autotuner for the persistent pair-binning kernel.

Search space (all shapes give identical bins):
  walk 0 (lanes hold forest-1 pixels, every mode):
    px_per_lane  1, 2, 4   -> forest-1 tile of 64, 128, 256 pixels per wave
    replicas     1, 2      -> LDS histogram copies (waves w, w+R, ... share one)
    row_pad      0, 1      -> LDS row stride np_rt or np_rt + 1 (bank padding)
    waves        4, 8, 16  -> waves per workgroup
    lane_map     0, 1, 2   -> lane order of forest-1 pixels (consecutive, by 2, by 4)
    merge        0, 1      -> sum runs of equal bins per block before the LDS atomics
    tail_max     -1, 32, 48, 63 -> longest forest-1 tail walked against forest 2 in lanes
                             (-1: slabs of 64*px_per_lane pixels only)
  walk 1 (lanes walk diagonals, FP64 only):
    waves        8, 16
Two replicas exist for 8-wave workgroups only; four would not leave room in the
CU's 160 KiB of LDS for a second workgroup.  The default sweep fixes replicas=1 and
row_pad=1 and skips lane_map 0 and 4-wave workgroups; ``--full`` sweeps the whole
space.  In profiles/autotune_full_sweep_sky100k.md a second LDS replica (half the
workgroups per CU) costs 20-22 % and 4-wave workgroups 18-20 %, while padding changes
at most 1 % and lane order up to 8 % in the merged shapes; the default sweep holds the
fastest shapes.
Each candidate is timed as hipGraph replays of the whole problem, with or
without the per-bin pair counts; the best shape per workload and mode is
written to ``tuned.json`` next to this file and used by the benchmark and the
drop-in procedures.

    python -m lya2pcf_amd.autotune [--config sky100k] [--forests N] [--modes mixed,fp64] [--counts] [--full]
"""
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuned.json")
PROFILES = os.path.join(os.path.dirname(HERE), "profiles")

SPACE = {"px_per_lane": (1, 2, 4), "replicas": (1,), "row_pad": (1,), "waves": (8, 16), "lane_map": (1, 2),
         "merge": (0, 1), "tail_max": (32, 48)}
FULL_SPACE = {"px_per_lane": (1, 2, 4), "replicas": (1, 2), "row_pad": (0, 1), "waves": (4, 8, 16),
              "lane_map": (0, 1, 2), "merge": (0, 1), "tail_max": (-1, 32, 48, 63)}
WALK_SPACE = {"walk": (1,), "waves": (8, 16)}
SHAPE_KEYS = ("px_per_lane", "waves", "replicas", "row_pad", "lane_map", "walk", "merge", "tail_max")


def load_best(mode="mixed", config="sky100k", walk=None):
    """Tuned kernel shape of ``mode`` for workload ``config`` (falls back to sky100k);
    with ``walk`` = 0 or 1, the best shape of that kernel family."""
    try:
        with open(CACHE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    entry = (t.get(config) or t.get("sky100k") or {}).get(mode, {})
    if walk is None:
        best = entry.get("shape")
    else:
        best = next((r["shape"] for r in entry.get("all", []) if r["shape"].get("walk", 0) == walk), None)
    return {k: v for k, v in best.items() if k in SHAPE_KEYS} if best else None


def measured_traffic(mode="mixed"):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(PROFILES, "traffic_xi_%s_sky100k.json" % mode)
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def measured_counters(mode="mixed"):
    """VALU issue, LDS bank-conflict and wait fractions of the tile kernel from the
    committed rocprofv3 counter summary, if any."""
    path = os.path.join(PROFILES, "counters_xi_%s_sky100k.json" % mode)
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    keys = ("valu_issue", "lds_bank_conflict_per_cu_cycle", "lds_wait_per_wave_cycle",
            "valu_per_evaluated_wave_step", "cycles_per_xcd", "source")
    return {k: d[k] for k in keys if k in d}


def candidates(mode, full=False):
    """Shapes to time: the tile space (default or full) and, for fp64, the walk."""
    spaces = [FULL_SPACE if full else SPACE] + ([WALK_SPACE] if mode == "fp64" else [])
    for sp in spaces:
        keys = list(sp)
        for vals in itertools.product(*sp.values()):
            c = dict(zip(keys, vals))
            if c.get("replicas", 1) == 2 and c.get("waves", 8) != 8:  # built for 8 waves only
                continue
            yield c


def tune(eng, mode, reps=3, log=print, counts=False, full=False):
    import torch
    from . import xi
    results = []
    for shape in candidates(mode, full):
        try:
            g = eng.graph(xi.default_config(mode, counts=counts, **shape))
        except Exception as e:  # shape not supported (LDS size)
            log("skip %s: %s" % (shape, e))
            continue
        g.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            g.launch()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        g.close()
        results.append((dt, shape))
        log("%s %s: %.4f s -> %.4g pixel-pairs/s" % (mode, shape, dt, eng.pixel_pairs / dt))
    results.sort(key=lambda r: r[0])
    return results


def main():
    import argparse
    sys.path.insert(0, os.path.dirname(HERE))
    from lya2pcf_amd import forests, synth, xi
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="sky100k")
    ap.add_argument("--forests", type=int, default=30000)
    ap.add_argument("--modes", default="mixed,fp64")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--counts", action="store_true", help="tune with per-bin pair counts on")
    ap.add_argument("--full", action="store_true", help="sweep the full shape space")
    a = ap.parse_args()
    cfg = synth.scaled(synth.CONFIGS[a.config], a.forests)
    fs = forests.ForestSet.from_survey(synth.generate(cfg), float32_inputs=True)
    eng = xi.XiEngine(fs, forests.angmax_from(fs, 200.0))
    print("tuning on %d forests, %.3g pixel pairs" % (fs.nf, eng.pixel_pairs), flush=True)
    out = {}
    for mode in a.modes.split(","):
        res = tune(eng, mode, a.reps, log=lambda s: print(s, flush=True), counts=a.counts, full=a.full)
        out[mode] = {"shape": res[0][1], "seconds": res[0][0], "pixel_pairs": eng.pixel_pairs,
                     "counts": a.counts, "workload": cfg.name + " (%d forests)" % fs.nf,
                     "all": [{"shape": s, "seconds": t} for t, s in res]}
    path = os.environ.get("LYA_TUNE_OUT", CACHE)
    try:
        with open(path) as f:
            table = json.load(f)
        if "mixed" in table or "fp64" in table:  # older single-workload table
            table = {}
    except (OSError, ValueError):
        table = {}
    table.setdefault(a.config, {}).update(out)
    with open(path, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
