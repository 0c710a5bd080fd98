"""Benchmark of the 4_fbe estimator hot path on one MI355X.

A step is one slice of the closed triangles of the workload: ring gather +
c2r of every (ring, channel) plane the slice needs, the sums over reference
planes (channel spectra and their products in the lag-Fourier form, or the
per-p1 MFMA contraction and lag reduction), the counts, and the copy of
every triangle's W x W accumulator and count arrays to host memory.  The cube is
resident in HBM before the timed region starts.

Default workload (the metric's configuration): 768 channels x 512^2 grid,
24 log-spaced annular bins 1..200 cells, all 506 closed triangles, all lags
(nc10 = 767), the lag-Fourier form (the sum over reference planes as a
channel-axis triple correlation in the DFT domain, FP32 products, FP64
sums).  The split-BF16 MFMA contraction is timed over the same slices
afterwards and reported beside it (--compare).

    python bench.py [--gpus 1] [--steps 20] [--warmup 2] [--config cfg3]
                    [--precision lagft|bf16x3|fp16x3|fp32|fp8]
                    [--compare bf16x3|...|none] [--no-cpu-baseline]

A pass over the workload is split into --steps slices of equal contraction
cost (contiguous in the triangle list), so the timed steps together cover
every triangle exactly once; warm-up steps are slices too.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    "cfg1": dict(N=128, nc=64, nflag=6, rings=(2, 60, 8), nc10=63, seed=1001,
                 desc="plain small: 64 ch x 128^2, 6 flagged, 8 log bins 2-60, lags |dnu|<=63"),
    "cfg2": dict(N=512, nc=384, nflag=31, rings=(1, 200, 20), nc10=383, seed=2001,
                 desc="plain scaled: 384 ch x 512^2, 31 flagged, 20 log bins 1-200, all lags"),
    "cfg3": dict(N=512, nc=768, nflag=0, rings=(1, 200, 24), nc10=767, seed=3001,
                 desc="precision: 768 ch x 512^2, 24 log bins 1-200, all lags"),
}

# published dense MFMA peaks of the instruction each form issues; one
# algorithmic BF16x3 flop costs three BF16 MFMA flops (mfma_passes)
PEAK_TFLOPS = {"bf16x3": 2500.0, "fp16x3": 2500.0, "fp32": 157.3, "fp8": 5000.0, "lagft": 157.3}
PASSES = {"bf16x3": 3, "fp16x3": 3, "fp32": 1, "fp8": 1, "lagft": 1}
# the lag-Fourier form's products run on the vector ALUs in FP32: the
# published FP32 vector peak (157.3 TFLOP/s) counts packed FP32 pairs, which
# this build does not issue; one FP32 FMA per lane and clock is half of it
VALU_FP32_UNPACKED_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# launch-configuration table tuned on this workload on an MI355X (loaded
# by default when present)
TUNE_TABLE = os.path.join(ROOT, "amber_amd", "tune_gfx950.txt")
# HBM bytes of all contraction launches of one step of the default workload
# (grouped waves, pipelined BF16x3 form): FETCH_SIZE x 2.00 + WRITE_SIZE x
# 1.00 summed over the step's dispatches from two PMC runs of this workload,
# the factors being what the same counters read for a plain 1 GiB device
# copy.  Counters cannot be read inside a timed run; `traffic` is this total
# over the step's launches.
MEASURED_TRAFFIC_STEP = {("cfg3", "bf16x3", "grouped"): 1.909e13,
                         # lag-Fourier form (compute grids, symmetric tilings, 1024-pixel items):
                         # k_lagft_tiles, FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00
                         ("cfg3", "lagft", "grouped"): 1.468e12}


def traffic_per_launch(args, st, nsteps):
    tot = MEASURED_TRAFFIC_STEP.get((args.config, args.precision, args.batch_mode))
    if tot is None or args.tri_limit or args.ref_grid:
        return None
    # the measured total covers one pass over every triangle, as the timed steps do
    return round(tot / max(st["n_contract_launch"], 1), -6)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def tri_grids(A, tris, hi, precision="bf16x3", ref_grid=False):
    """The grid each triangle is summed on: its compute grid in the
    lag-Fourier form (the smallest alias-free grid of its rings), Ng(r1)
    otherwise"""
    if precision == "lagft" and not ref_grid:
        return np.array([A.compute_grid(hi[t[0]], hi[t[1]], hi[t[2]]) for t in tris], np.float64)
    return np.array([A.ng(hi[t[0]]) for t in tris], np.float64)


def tri_costs(A, tris, hi, precision="bf16x3", ref_grid=False):
    """Relative contraction cost of each triangle, used to cut the list into
    slices of near-equal work: its grid's pixels times the triplets it
    computes (half of them where r1 == r2 or r2 == r3 lets the kernel mirror
    the other half); in the lag-Fourier form the pixels times the share of
    the (k2, k3) half plane its tiling sums: half for r1 == r2 (sheared
    tiles) and for r2 == r3 (the tiles below the diagonal of [0, L/2)^2 and
    past the anti-diagonal left out), about a quarter for r1 == r2 == r3
    (the principal domain's tiles)."""
    ng = tri_grids(A, tris, hi, precision, ref_grid)
    sym = np.array([(t[0] == t[1]) or (t[1] == t[2]) for t in tris])
    if precision == "lagft":
        r12 = np.array([t[0] == t[1] for t in tris])
        r23 = np.array([t[1] == t[2] for t in tris])
        return ng * ng * np.where(r12 & r23, 0.25, np.where(r12 | r23, 0.5, 1.0))
    return ng * ng * np.where(sym, 0.5, 1.0)


def cost_slices(costs, nslices):
    """Contiguous slices of the triangle list (kept in its (r1, r2, r3) order,
    so triangles that share a reference ring's planes mostly share a slice)
    of near-equal total cost; every triangle lies in exactly one slice."""
    n = len(costs)
    nslices = max(1, min(int(nslices), n))
    cum = np.concatenate([[0.0], np.cumsum(costs)])
    bounds = [0]
    for k in range(1, nslices):
        target = cum[-1] * k / nslices
        i = int(np.argmin(np.abs(cum - target)))  # cut before triangle i
        i = min(max(i, bounds[-1] + 1), n - (nslices - k))
        bounds.append(i)
    bounds.append(n)
    return [np.arange(bounds[k], bounds[k + 1]) for k in range(nslices)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps; the workload's triangles are split into this many slices of equal cost, "
                         "so the timed steps together make exactly one pass over every triangle")
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed steps, each one slice (those that size the buffers first, then the heaviest)")
    ap.add_argument("--budget-s", type=float, default=540.0,
                    help="wall-clock budget of the whole run: warm-up steps, the FP64 check and the CPU "
                         "baseline are cut to fit it")
    ap.add_argument("--config", default="cfg3", choices=sorted(CONFIGS) + ["cfg4", "cfg5", "post", "maps", "grid"])
    ap.add_argument("--qps", type=float, default=0.0, help="cfg5: mean arrival rate (0: 60%% of the measured capacity)")
    ap.add_argument("--duration", type=float, default=90.0, help="cfg5: seconds of arrivals")
    ap.add_argument("--probe-s", type=float, default=20.0, help="cfg5: seconds of the one-at-a-time capacity probe")
    ap.add_argument("--upload-frac", type=float, default=0.1, help="cfg5: fraction of requests that carry a cube")
    ap.add_argument("--precision", default="lagft", choices=["bf16x3", "fp16x3", "fp32", "fp8", "lagft"],
                    help="lagft (default): the lag-Fourier form; bf16x3 / fp16x3 / fp32 / fp8: the MFMA contraction")
    ap.add_argument("--compare", default="bf16x3", choices=["none", "bf16x3", "fp16x3", "fp32", "fp8", "lagft"],
                    help="cfg1-3: a second form timed over the same slices after the main line's pass and reported "
                         "beside it (none: skip)")
    ap.add_argument("--tri-limit", type=int, default=0,
                    help="development only: K triangles (the line is then not the metric workload)")
    ap.add_argument("--tri-pick", default="first", choices=["first", "largest"])
    ap.add_argument("--ref-grid", action="store_true",
                    help="lag-Fourier form on every triangle's reference grid Ng(r1) instead of its compute grid")
    ap.add_argument("--batch-mode", default="grouped", choices=["grouped", "per_triangle"])
    ap.add_argument("--estimator", default="v1", choices=["v1", "v2"],
                    help="v1: amber (data contraction); v2: amber_v2 (data plus the normalisation contraction)")
    ap.add_argument("--tune", default=None, metavar="PATH",
                    help="launch-configuration table: loaded when PATH exists, else tuned on a sample of the "
                         "workload and saved there; 'none': the built-in defaults; default: the committed "
                         "table when present, else the defaults")
    ap.add_argument("--tune-seconds", type=float, default=150.0, help="bound of a tuning run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (counter calibration)")
    args = ap.parse_args()
    t_run = time.time()
    if args.copy_calib:
        copy_calib(args.copy_calib)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the path would shard by "
              "triangle with no collective, see the design notes)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import amber_amd as A
    if args.config == "cfg4":
        return bench_cfg4(A, args)
    if args.config == "cfg5":
        return bench_cfg5(A, args)
    if args.config == "post":
        return bench_post(A, args)
    if args.config == "maps":
        return bench_maps(A, args)
    if args.config == "grid":
        return bench_grid(A, args)
    cfg = CONFIGS[args.config]
    N, nc, nc10 = cfg["N"], cfg["nc"], cfg["nc10"]
    lo, hi = A.log_rings(*cfg["rings"])
    tris = A.triangles(lo, hi)
    if args.tri_limit:
        if args.tri_pick == "largest":
            ngs = np.array([A.ng(hi[t[0]]) for t in tris])
            tris = tris[np.sort(np.argsort(-ngs, kind="stable")[: args.tri_limit])]
        else:
            tris = tris[: args.tri_limit]
    if args.precision == "lagft" and not args.ref_grid:
        # triangles grouped by compute grid (then in (r1, r2, r3) order), so
        # the contiguous slices share more (ring, grid) sets; every step
        # still builds all the sets it reads
        g = tri_grids(A, tris, hi, args.precision)
        tris = tris[np.lexsort((tris[:, 2], tris[:, 1], tris[:, 0], g))]
    t0 = time.time()
    gv, samp, uf = A.generate(N, nc, nflag=cfg["nflag"], seed=cfg["seed"])
    fk = 1 if cfg["nflag"] else 0
    log("generated %s cube in %.1f s" % (args.config, time.time() - t0))

    ctx = A.Context(device=0, precision=args.precision)
    ctx.set_batch_mode(args.batch_mode)
    t0 = time.time()
    cube = ctx.load_cube(N, nc, 1.0, fk, gv, samp)
    upload_s = time.time() - t0
    ctx.set_channels(cube, uf if fk else None, nc10)
    ctx.set_rings(lo, hi)
    log("uploaded cube (%.2f GB of FP64 file data) in %.2f s" % ((gv.nbytes + samp.nbytes) / 1e9, upload_s))

    W = 2 * nc10 + 1
    if args.ref_grid:
        ctx.set_grid(False)
    costs = tri_costs(A, tris, hi, args.precision, args.ref_grid)
    grids = tri_grids(A, tris, hi, args.precision, args.ref_grid)
    # the sample a missing table is tuned on: up to 24 triangles evenly
    # spread over the list ordered by cost, so every Ng bucket is present
    by_cost = np.argsort(costs, kind="stable")
    sample = [(cube, *map(int, tris[k])) for k in by_cost[np.linspace(0, len(tris) - 1, min(24, len(tris))).astype(int)]]
    tuning = use_tuning(ctx, args, sample)
    slices = cost_slices(costs, max(args.steps, 1))
    nmax = max(len(sl) for sl in slices)
    out = (np.zeros((nmax, W, W), np.float32), np.zeros((nmax, W, W), np.int32),
           np.zeros((nmax, W, W), np.float32) if args.estimator == "v2" else None)
    for o in out:  # the caller's result buffers, reused by every step: their pages are mapped up front
        if o is not None:
            o.fill(0)
    log("%d triangles in %d slices of %d-%d triangles (cost spread %.3f-%.3f of a pass)"
        % (len(tris), len(slices), min(len(x) for x in slices), nmax,
           min(costs[x].sum() for x in slices) / costs.sum(), max(costs[x].sum() for x in slices) / costs.sum()))

    def step(sl):
        # every step rebuilds the filtered planes its slice needs: nothing carries over
        ctx.drop_planes(cube)
        n = len(sl)
        ctx.bispec_batch(cube, tris[sl], out=tuple(None if o is None else o[:n] for o in out))

    def footprint(sl):
        """plane sets a slice builds: n^2 per distinct (ring, grid n)"""
        sets = {(int(r), int(grids[k])) for k in sl for r in tris[k]}
        return sum(g * g for _, g in sets)

    def warm_order(slc):
        """slices for warm-up steps: the one whose plane sets take the most
        device memory, the one with the most triangles (result buffers and
        staging), then by cost; so the timed steps grow no buffer"""
        order = [int(np.argmax([footprint(x) for x in slc])), int(np.argmax([len(x) for x in slc]))]
        order += sorted(range(len(slc)), key=lambda k: -costs[slc[k]].sum())
        seen = []
        for k in order:
            if k not in seen:
                seen.append(k)
        return seen

    def timed_pass(slc):
        """warm-up on the slices that size the buffers, then every slice
        once; returns (seconds per slice, stats, per-triangle latencies)"""
        for k in warm_order(slc)[:2]:
            step(slc[k])
        ctx.reset_stats()
        secs, lt = [], []
        for sl in slc:
            t0 = time.time()
            step(sl)
            secs.append(time.time() - t0)
            lt.extend(ctx.batch_times().tolist())
        return secs, ctx.stats(), lt

    # warm-up: the slices that size the plane pool and the result buffers
    # first, then the heaviest, so no buffer grows in the timed steps
    heavy = warm_order(slices)
    nwarm = 0
    for w in range(args.warmup):
        sl = slices[heavy[w % len(slices)]]
        if w > 0 and time.time() - t_run + 2 * costs[sl].sum() / costs.sum() * 60.0 > 0.25 * args.budget_s:
            log("warm-up cut at %d steps to keep the run within its budget" % w)
            break
        t0 = time.time()
        step(sl)
        nwarm += 1
        log("warmup step %d (%d triangles): %.2f s (plane pool %.1f GB)"
            % (w, len(sl), time.time() - t0, ctx.stats()["plane_pool_bytes"] / 1e9))
    ctx.reset_stats()
    per_step = []
    lat = []
    t_start = time.time()
    for s_ in range(len(slices)):
        t0 = time.time()
        step(slices[s_])
        per_step.append(time.time() - t0)
        lat.extend(ctx.batch_times().tolist())
        log("timed step %d (%d triangles): %.2f s (plane pool %.1f GB)"
            % (s_, len(slices[s_]), per_step[-1], ctx.stats()["plane_pool_bytes"] / 1e9))
    elapsed = time.time() - t_start
    nsteps = len(slices)
    st = ctx.stats()
    value = st["evals"] / elapsed

    # roofline of the dominant kernel (the contraction), from HIP events on
    # the library's stream; work is counted once per distinct sum (a pair
    # of triplets that r1 == r2 or r2 == r3 makes equal counts once)
    ms_c = st["ms_contract"]
    lagft = args.precision == "lagft"
    work = st["flop_lagft"] if lagft else st["flop_contract_distinct"]
    ach = work / (ms_c * 1e-3) / 1e12 if ms_c > 0 else 0.0
    peak = PEAK_TFLOPS[args.precision]
    roof = {"bound": "valu" if lagft else "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
            "frac": round(ach / peak, 4),
            "traffic": traffic_per_launch(args, st, nsteps),
            "kernel": "k_lagft_tiles" if lagft else "k_contract<%s>" % args.precision,
            "work": ("14 flop per distinct (k2, k3, pixel) of the half spectrum k3 <= L/2 (one complex triple "
                     "product conj(F) G H), L = channel transform length, pixels of the grid the triangle is "
                     "summed on; half of the half spectrum where r1 == r2 "
                     "or r2 == r3 makes sums pairwise equal, a third where r1 == r2 == r3" if lagft else
                     "2 flop per (distinct triplet, pixel): 2*V*Ng^2 per triangle, halved (plus the self-mirrored "
                     "triplets) where r1 == r2 or r2 == r3"),
            "mfma_passes": PASSES[args.precision],
            "frac_of_pass_peak": round(ach * PASSES[args.precision] / peak, 4),
            "avg_launch_ms": round(ms_c / max(st["n_contract_launch"], 1), 3),
            "launches": int(st["n_contract_launch"]),
            "algorithmic_all_triplets_tflops": round(st["flop_contract"] / (ms_c * 1e-3) / 1e12, 2) if ms_c > 0 else None,
            # MFMA work actually issued (whole tiles, padded pitch, all passes,
            # skipped tiles excluded) against the dense peak of the instruction
            "issued_tflops": round(st["flop_contract_exec"] / (ms_c * 1e-3) / 1e12, 2) if ms_c > 0 else None,
            "issued_frac_of_instr_peak": round(st["flop_contract_exec"] / (ms_c * 1e-3) / 1e12 /
                                               (VALU_FP32_UNPACKED_TFLOPS if lagft else peak), 4) if ms_c > 0 else None}
    if lagft:
        roof["frac_of_unpacked_fp32_valu"] = round(ach / VALU_FP32_UNPACKED_TFLOPS, 4)
        roof["direct_sum_equivalent_tflops"] = round(st["flop_contract_distinct"] / (ms_c * 1e-3) / 1e12, 2) \
            if ms_c > 0 else None
    stages = {"fft_ms": round(st["ms_fft"], 2), "contract_ms": round(st["ms_contract"], 2),
              "norm_ms": round(st["ms_norm"], 2), "reduce_ms": round(st["ms_reduce"], 2),
              "fft_gbs": round(st["bytes_fft"] / (st["ms_fft"] * 1e-3) / 1e9, 1) if st["ms_fft"] > 0 else None}

    err = None
    ocache = {}
    if not args.no_error:
        err = fp64_error(A, ctx, cube, cfg, gv, samp, uf, fk, lo, hi, tris,
                         deadline=t_run + args.budget_s - args.cpu_seconds - 30.0, cache=ocache)
    # the second form over the same slices, when the run budget allows one
    # more pass of it (estimated from its flop rate on this workload)
    compare = None
    if args.compare not in ("none", args.precision) and not args.tri_limit:
        est = {"bf16x3": 4.6, "fp16x3": 4.9, "fp8": 2.5, "fp32": 16.0, "lagft": 0.25}[args.compare] * elapsed \
            if args.precision == "lagft" else elapsed
        if time.time() - t_run + 1.2 * est + args.cpu_seconds + 40.0 > args.budget_s:
            compare = {"precision": args.compare, "skipped": "run budget (estimated %.0f s)" % est}
        else:
            ctx.set_precision(args.compare)
            secs, st2, _ = timed_pass(slices)
            e2 = sum(secs)
            ms2 = st2["ms_contract"]
            lag2 = args.compare == "lagft"
            w2 = st2["flop_lagft"] if lag2 else st2["flop_contract_distinct"]
            a2 = w2 / (ms2 * 1e-3) / 1e12 if ms2 > 0 else 0.0
            compare = {"precision": args.compare, "value": round(st2["evals"] / e2, 1), "unit": "evals/s",
                       "pass_s": round(e2, 3), "ms_per_step": round(1e3 * e2 / len(slices), 2),
                       "roofline": {"bound": "valu" if lag2 else "mfma", "achieved": round(a2, 2),
                                    "peak": PEAK_TFLOPS[args.compare], "unit": "TFLOP/s",
                                    "frac": round(a2 / PEAK_TFLOPS[args.compare], 4),
                                    "kernel": "k_lagft_tiles" if lag2 else "k_contract<%s>" % args.compare,
                                    "issued_tflops": round(st2["flop_contract_exec"] / (ms2 * 1e-3) / 1e12, 2)
                                    if ms2 > 0 else None,
                                    "avg_launch_ms": round(ms2 / max(st2["n_contract_launch"], 1), 3)},
                       "stages": {"fft_ms": round(st2["ms_fft"], 2), "contract_ms": round(st2["ms_contract"], 2),
                                  "reduce_ms": round(st2["ms_reduce"], 2)},
                       "speedup_of_main_form": round(e2 / elapsed, 3)}
            if not args.no_error:
                compare["rel_err_vs_fp64"] = fp64_error(A, ctx, cube, cfg, gv, samp, uf, fk, lo, hi, tris,
                                                        deadline=t_run + args.budget_s - args.cpu_seconds - 30.0,
                                                        cache=ocache)
            ctx.set_precision(args.precision)
            log("compare %s: %.3g evals/s, pass %.1f s (main form %.1f s)" % (args.compare, compare["value"], e2,
                                                                            elapsed))
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(A, cfg, gv, samp, uf, fk, lo, hi, tris,
                           min(args.cpu_seconds, max(5.0, t_run + args.budget_s - time.time() - 20.0)))

    workload = cfg["desc"] + ", %d triangles" % len(tris)
    if args.estimator == "v2":
        workload += ", amber_v2 estimator (data and normalisation contractions)"
    if args.tri_limit:
        workload += " (DEVELOPMENT SUBSET, not the metric workload)"
    line = {
        "metric": "triangle-bin x lag-pair evals/sec at 768ch x 512^2; rel err vs FP64; p50/p99 latency",
        "value": round(value, 1), "unit": "evals/s", "n_gpus": 1, "steps": nsteps, "warmup": nwarm,
        "ms_per_step": round(1e3 * elapsed / nsteps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"bf16x3": "bf16", "fp16x3": "fp16", "fp32": "f32", "fp8": "fp8_e4m3", "lagft": "f32"}[args.precision],
        "precision": args.precision, "data": "synthetic",
        "config": {"workload": workload, "estimator": args.estimator, "channels": nc, "grid": N,
                   "bins": cfg["rings"][2],
                   "triangles": int(len(tris)), "nc10": nc10, "flagged": cfg["nflag"],
                   "evals_per_pass": int(st["evals"]),
                   "sum_grid": ("per triangle the smallest alias-free grid of its three rings, sums scaled by "
                                "(Ng/n)^2" if lagft and not args.ref_grid else "Ng(r1) of the reference"),
                   "step": "one slice of the triangle list%s (%d slices of equal contraction cost); the %d timed "
                           "steps cover every triangle once and each rebuilds every plane set it reads"
                           % (" grouped by compute grid" if lagft and not args.ref_grid else "", nsteps, nsteps)},
        "pass_s": round(elapsed, 3),
        "roofline": roof, "stages": stages,
        "triangle_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                "p99": round(float(np.percentile(lat, 99)), 3),
                                "max": round(float(np.max(lat)), 3),
                                "note": "time from the start of a step to each triangle's results in host memory "
                                        "(the triangles of one grouped wave complete together); request latency "
                                        "under load is --config cfg5"},
        "rel_err_vs_fp64": err, "cpu_baseline": cpu, "compare_form": compare,
        "launch_config": tuning,
        "upload_s_pcie_incl": round(upload_s, 3),
        "run_s": round(time.time() - t_run, 1),
    }
    if args.estimator == "v2":
        # waves whose normalisation sums were past 1/32 of their 1/8 grid
        line["norm_certificate"] = {"uncertified_waves": int(st["n_uncertified"]),
                                    "worst_grid_distance": round(float(st["cert_margin_max"]), 4)}
    ctx.close()
    print(json.dumps(line), flush=True)


def mixed_workload(A, ctx, shapes, ntri, seed, max_r_frac=1.0 / 3.0, keep=None):
    """Resident cubes of the given (N, nc) shapes (no flags, all lags) and
    ntri random closed triangles over them: ring widths uniform in 1..16
    cells, radii below N/3, each triangle's three rings its own.  Returns
    (cubes by shape, [(cube, r1, r2, r3)], ring lo, ring hi, triangles by cube);
    keep (a dict): receives each shape's host arrays (gv, samp)."""
    rng = np.random.default_rng(seed)
    cubes = {}
    for k, (N, nc) in enumerate(shapes):
        gv, samp, _ = A.generate(N, nc, nflag=0, seed=seed + 17 * k)
        cid = ctx.load_cube(N, nc, 1.0, 0, gv, samp)
        ctx.set_channels(cid, None, nc - 1)
        cubes[(N, nc)] = cid
        if keep is not None:
            keep[(N, nc)] = (gv, samp)
        del gv, samp
    lo, hi, reqs = [], [], []
    by_cube = {c: [] for c in cubes.values()}
    while len(reqs) < ntri:
        N, nc = shapes[rng.integers(len(shapes))]
        umax = N * max_r_frac - 1
        w = rng.integers(1, 17, 3)
        h1 = rng.uniform(8 + w[0], umax)
        h2 = min(h1, h1 * rng.uniform(0.5, 1.0) + w[1] / 2)
        if h2 <= w[2] + 1:
            continue
        h3 = rng.uniform(w[2] + 1, h2)
        rh = np.array([h3, h2, h1])
        rl = np.maximum(rh - w[[2, 1, 0]], 0.5)
        if not np.all(rh[:-1] <= rh[1:]):
            continue
        if not any((t == [2, 1, 0]).all() for t in A.triangles(rl, rh)):
            continue
        base = len(lo)
        lo.extend(rl.tolist())
        hi.extend(rh.tolist())
        cid = cubes[(N, nc)]
        reqs.append((cid, base + 2, base + 1, base))
        by_cube[cid].append((cid, base + 2, base + 1, base))
    lo, hi = np.array(lo), np.array(hi)
    ctx.set_rings(lo, hi)
    return cubes, reqs, lo, hi, by_cube


def bench_cfg4(A, args):
    """200 uneven triangles over 9 cubes {256,512,1024}^2 x {128,384,768}:
    per-triangle launches of the plain form against grouped waves and one
    captured graph, each under the built-in launch configuration and under
    the table the autotuner makes from a sample of the mix."""
    shapes = [(N, nc) for N in (256, 512, 1024) for nc in (128, 384, 768)]
    ctx = A.Context(device=0, precision=args.precision)
    t0 = time.time()
    cubes, reqs, lo, hi, _ = mixed_workload(A, ctx, shapes, 200, 4001)
    log("cfg4: 9 cubes resident, 200 triangles, %d rings, setup %.1f s" % (len(lo), time.time() - t0))

    def timed(label, fn, reps, keys=("ms_contract", "ms_norm")):
        ts, cs = [], []
        for _ in range(reps):
            ctx.reset_stats()
            t = time.time()
            out = fn()
            ts.append(time.time() - t)
            st = ctx.stats()
            cs.append(sum(st[k] for k in keys))
        log("cfg4 %s: %s s (device %s ms)" % (label, ", ".join("%.2f" % x for x in ts),
                                              ", ".join("%.0f" % x for x in cs)))
        return out, min(ts), st, min(cs)

    def same(a, b):
        return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))

    def eager():
        drop_all(ctx, cubes)
        return ctx.bispec_multi(reqs)

    def graph_run(label, reps):
        t = time.time()
        g = ctx.graph(reqs)
        cap = time.time() - t
        out, tg, _, cg = timed(label, g.run, reps, keys=("ms_graph",))
        g.close()
        return out, tg, cg, cap

    reps = 2
    ctx.tune_clear()
    ctx.set_batch_mode("per_triangle")
    per, t_per, st, c_per = timed("per-triangle launches", eager, 1)
    evals = st["evals"]
    ctx.set_batch_mode("grouped")
    fg, t_fg, _, c_fg = timed("grouped, fixed configuration", eager, reps)
    fgr, t_fgr, c_fgr, cap_f = graph_run("graph, fixed configuration", reps)
    # the autotuner on every fifth triangle of the mix
    sample = reqs[::5]
    t = time.time()
    table, _ = ctx.autotune(sample, max_seconds=args.tune_seconds,
                            path=None if args.tune in (None, "none") else args.tune)
    t_tune = time.time() - t
    log("cfg4: tuned %d buckets on %d triangles in %.1f s" % (len(table), len(sample), t_tune))
    tg, t_tg, _, c_tg = timed("grouped, tuned", eager, reps)
    tgr, t_tgr, c_tgr, cap_t = graph_run("graph, tuned", reps)

    def rate(s_):
        return {"s": round(s_, 3), "evals_per_s": round(evals / s_, 1)}

    line = {"metric": "cfg4 mixed-batch evals/sec (200 uneven triangles, 9 cubes)", "unit": "evals/s",
            "value": round(evals / t_tgr, 1), "precision": args.precision, "evals": int(evals),
            "per_triangle": dict(rate(t_per), contraction_ms=round(c_per, 1)),
            "fixed": {"grouped": dict(rate(t_fg), contraction_ms=round(c_fg, 1)),
                      "graph": dict(rate(t_fgr), device_ms=round(c_fgr, 1), capture_s=round(cap_f, 3))},
            "tuned": {"grouped": dict(rate(t_tg), contraction_ms=round(c_tg, 1)),
                      "graph": dict(rate(t_tgr), device_ms=round(c_tgr, 1), capture_s=round(cap_t, 3)),
                      "tune_s": round(t_tune, 1), "sample_triangles": len(sample), "table": table},
            "speedup_vs_per_triangle": {"fixed_grouped": round(t_per / t_fg, 3), "fixed_graph": round(t_per / t_fgr, 3),
                                        "tuned_grouped": round(t_per / t_tg, 3), "tuned_graph": round(t_per / t_tgr, 3)},
            "device_speedup_tuned_vs_fixed": {"grouped_contraction": round(c_fg / c_tg, 3),
                                              "graph_replay": round(c_fgr / c_tgr, 3)},
            "bit_identical_to_per_triangle": {"fixed_grouped": same(per, fg), "fixed_graph": same(per, fgr),
                                              "tuned_grouped": same(per, tg), "tuned_graph": same(per, tgr)},
            "note": "each timed pass rebuilds every filtered plane (eager: planes dropped first; graph: the plane "
                    "builds are nodes of the graph); value is the tuned graph; contraction_ms: device time of the "
                    "contraction kernels; device_ms of a graph: the whole replay (builds included)"}
    ctx.close()
    print(json.dumps(line), flush=True)


def use_tuning(ctx, args, sample):
    """Installs the launch-configuration table: --tune PATH is loaded when it
    exists (a table of another device is reported and not used), else the
    sample is tuned and saved there.  Returns what was done, for the
    result line."""
    path = args.tune
    if path is None and os.path.exists(TUNE_TABLE):
        path = TUNE_TABLE
    if not path or path == "none":
        return {"table": None, "note": "built-in defaults"}
    if os.path.exists(path):
        try:
            ctx.tune_load(path)
            return {"table": os.path.relpath(path, ROOT) if path.startswith(ROOT) else path,
                    "entries": len(ctx.tune_table()), "tuned_now": False}
        except Exception as e:  # a table of another device: the defaults
            log("tuning table %s not used: %s" % (path, e))
            return {"table": None, "note": "table not used: %s" % e}
    t = time.time()
    table, _ = ctx.autotune(sample, max_seconds=args.tune_seconds, path=path)
    log("tuned %d buckets on %d triangles in %.1f s -> %s" % (len(table), len(sample), time.time() - t, path))
    return {"table": path, "entries": len(table), "tuned_now": True, "tune_s": round(time.time() - t, 1),
            "sample_triangles": len(sample)}


def drop_all(ctx, cubes):
    for c in cubes.values():
        ctx.drop_planes(c)


def bench_cfg5(A, args):
    """Serving: Poisson arrivals of requests, each of 1-64 triangle bins of
    the cfg4 distribution on one cube, which is a resident cube of one of
    the nine shapes {256,512,1024}^2 x {128,384,768} or (a fraction
    --upload-frac of requests) a new cube of one of those shapes carried by
    the request and uploaded by the serving queue; continuous batching at
    work-item granularity; p50/p99 latency and throughput."""
    import threading
    shapes = [(N, nc) for N in (256, 512, 1024) for nc in (128, 384, 768)]
    ctx = A.Context(device=0, precision=args.precision)
    t0 = time.time()
    host = {}
    cubes, reqs, lo, hi, by_cube = mixed_workload(A, ctx, shapes, 600, 5001, keep=host)
    log("cfg5: 9 cubes resident, %d triangle bins, setup %.1f s" % (len(reqs), time.time() - t0))
    rng = np.random.default_rng(5002)
    shape_of = {cid: sh for sh, cid in cubes.items()}
    cube_ids = sorted(by_cube)

    def make_request():
        cid = cube_ids[rng.integers(len(cube_ids))]
        pool = by_cube[cid]
        k = int(rng.integers(1, 65))
        pick = [pool[i] for i in rng.choice(len(pool), size=min(k, len(pool)), replace=False)]
        if rng.random() < args.upload_frac:  # the same bins on a carried cube of this shape
            return shape_of[cid], [t[1:] for t in pick]
        return None, pick

    def submit(r):
        shape, tris = r
        if shape is None:
            return ctx.submit(tris)
        gv, samp = host[shape]
        return ctx.submit_cube(shape[0], shape[1], 1.0, 0, gv, samp, tris, nc10=shape[1] - 1)[0]

    ctx.serve_start(max_active=256)
    # capacity: a closed loop of requests, one at a time
    t = time.time()
    nprobe = 0
    while time.time() - t < args.probe_s or nprobe < 3:
        ctx.wait(submit(make_request()))
        nprobe += 1
    cap_rps = nprobe / (time.time() - t)
    qps = args.qps if args.qps > 0 else 0.6 * cap_rps
    log("cfg5: capacity %.3f requests/s one at a time (%d requests); offered load %.3f requests/s for %.0f s"
        % (cap_rps, nprobe, qps, args.duration))
    ctx.reset_stats()
    tickets = []
    lock = threading.Lock()
    done = []

    def waiter():
        i = 0
        while True:
            with lock:
                if i >= len(tickets):
                    if stop[0]:
                        return
                    tk = None
                else:
                    tk = tickets[i]
            if tk is None:
                time.sleep(0.001)
                continue
            outs, lat = ctx.wait(tk[0])
            done.append((lat, tk[1], tk[2]))
            i += 1
            if i % 5 == 0:
                log("cfg5: %d requests done" % i)

    stop = [False]
    th = threading.Thread(target=waiter)
    th.start()
    t_start = time.time()
    next_t = t_start
    while next_t - t_start < args.duration:
        now = time.time()
        if now < next_t:
            time.sleep(next_t - now)
        r = make_request()
        tk = submit(r)
        with lock:
            tickets.append((tk, len(r[1]), r[0] is not None))
        next_t += rng.exponential(1.0 / qps)
    stop[0] = True
    th.join()
    wall = time.time() - t_start
    ctx.serve_stop()
    st = ctx.stats()
    lat = np.array([d[0] for d in done])
    up = np.array([d[2] for d in done])
    ntri = sum(d[1] for d in done)

    def pct(x):
        return ({"p50": round(float(np.percentile(x, 50)), 1), "p99": round(float(np.percentile(x, 99)), 1),
                 "max": round(float(np.max(x)), 1), "n": int(len(x))} if len(x) else None)

    line = {"metric": "cfg5 serving latency and throughput", "precision": args.precision,
            "offered_qps": round(qps, 4), "capacity_rps_one_at_a_time": round(cap_rps, 4),
            "requests": len(done), "triangles": int(ntri), "wall_s": round(wall, 2),
            "throughput": {"requests_per_s": round(len(done) / wall, 4), "triangles_per_s": round(ntri / wall, 2),
                           "evals_per_s": round(st["evals"] / wall, 1)},
            "latency_ms": pct(lat), "latency_ms_resident_cube": pct(lat[~up]), "latency_ms_carried_cube": pct(lat[up]),
            "cube_uploads": int(st["n_cube_uploads"]), "evicted_cubes": int(st["n_evict_cubes"]),
            "evicted_plane_sets": int(st["n_evict_planes"]), "iterations": int(st["n_serve_iters"]),
            "device_ms": {"contraction": round(st["ms_contract"] + st["ms_norm"], 1), "fft": round(st["ms_fft"], 1),
                          "reduce": round(st["ms_reduce"], 1)},
            "workload": "9 resident cubes {256,512,1024}^2 x {128,384,768} ch, 600 triangle bins (widths 1-16, "
                        "radii < N/3); requests of 1-64 bins of one cube, %.0f%% of them carrying a new cube of that "
                        "shape (up to 1024^2 x 768 complex64, uploaded by the queue); Poisson arrivals for %.0f s; "
                        "HBM budget 0.9 of the free memory" % (100 * args.upload_frac, args.duration)}
    ctx.close()
    print(json.dumps(line), flush=True)


def bench_post(A, args):
    """Lag-space post-processing (get_cyl_BS) of cfg3-sized result arrays
    (W = 1535, N = 1519 after the 8-lag cut): triangles/s, and the HBM rate
    of the two FP64 FFT passes of the 2-D inverse transform (algorithmic
    bytes: the M x N real input read and the N x N output written)."""
    nc10, n = 767, 32
    W = 2 * nc10 + 1
    rng = np.random.default_rng(6001)
    bl = rng.standard_normal((n, W, W)).astype(np.float32)
    cnt = rng.integers(1, 768, (n, W, W)).astype(np.int32)
    ctx = A.Context(device=0)
    ctx.cyl_bispec(bl[:2], cnt[:2])
    ctx.reset_stats()
    steps = max(args.steps, 1)
    t = time.time()
    for _ in range(steps):
        ctx.cyl_bispec(bl, cnt)
    dt = (time.time() - t) / steps
    st = ctx.stats()
    N = W - 2 * A.ANALYSIS["ncut"]
    M = nc10 + 1 - A.ANALYSIS["ncut"]
    byt = (8.0 * M * N + 8.0 * N * N) * n * steps
    gbs = byt / (st["ms_post"] * 1e-3) / 1e9
    line = {"metric": "lag-space post-processing (fold, normalise, window, lag->k_par transform)",
            "value": round(n / dt, 2), "unit": "triangles/s", "triangles": n, "W": W, "N": N,
            "s_per_batch": round(dt, 4),
            "roofline": {"bound": "hbm", "kernel": "k_cyl_rows + k_cyl_cols (FP64 FFT passes)",
                         "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "work": "8 B per input lag cell (M x N) read + 8 B per output cell (N x N) written",
                         "nominal_fft_gflops": round(st["flop_post"] / (st["ms_post"] * 1e-3) / 1e9, 1)},
            "ms_transform_per_triangle": round(st["ms_post"] / steps / n, 4),
            "note": "host->device copy of the W x W arrays and device->host copy of the N x N results included"}
    ctx.close()
    print(json.dumps(line), flush=True)


def bench_maps(A, args):
    """MAPS C_l(dnu) (the maps program over cl_dnu) on the 768-channel 512^2
    cube of the precision workload, 15 log bins of |U| in [1, 500): every
    channel pair n1 <= n2 over every binned cell in FP64.  The host->device
    copy of the FP64 file arrays is inside each call and reported apart from
    the device time; value is channel-pair cells per second of device time."""
    import oracle as O
    cfg = CONFIGS["cfg3"]
    N, nc, Du = cfg["N"], cfg["nc"], 1.0
    gv, samp, _ = A.generate(N, nc, nflag=0, seed=cfg["seed"])
    ctx = A.Context(device=0)
    ctx.maps_cldnu(N, nc, Du, gv, samp)
    ctx.reset_stats()
    steps = max(args.steps, 1)
    t = time.time()
    for _ in range(steps):
        ctx.maps_cldnu(N, nc, Du, gv, samp)
    wall = (time.time() - t) / steps
    st = ctx.stats()
    ms_pair, ms_all = st["ms_maps"] / steps, st["ms_maps_total"] / steps
    flop = st["flop_maps"] / steps
    pair_cells = flop / 4.0
    line = {"metric": "MAPS C_l(dnu): channel-pair cells per second", "value": round(pair_cells / (ms_all * 1e-3)),
            "unit": "pair-cells/s", "steps": steps, "ms_per_step": round(ms_all, 3), "dtype": "f64",
            "data": "synthetic", "config": {"workload": "maps 768 ch x 512^2, 15 log |U| bins 1-500"},
            "roofline": {"bound": "fp64", "kernel": "k_maps_pairs", "achieved": round(flop / (ms_pair * 1e-3) / 1e12, 2),
                         "peak": 78.6, "unit": "TFLOP/s", "frac": round(flop / (ms_pair * 1e-3) / 1e12 / 78.6, 4)},
            "ms_pair_kernel": round(ms_pair, 3), "s_per_call_with_upload": round(wall, 3)}
    ctx.close()
    if not args.no_cpu_baseline:
        ncs = 48
        sub = np.ascontiguousarray(gv.reshape(-1, nc, 2)[:, :ncs]), np.ascontiguousarray(samp[:, :ncs])
        t = time.time()
        O.cldnu(N, ncs, Du, sub[0], sub[1])
        dt = time.time() - t
        line["cpu_baseline"] = {"value": round(pair_cells / (nc * (nc + 1) / 2) * ncs * (ncs + 1) / 2 / dt),
                                "unit": "pair-cells/s", "cores": 1, "kind": "port",
                                "sample": "the FP64 oracle's pair loop over the first %d channels of the same cube" % ncs}
    print(json.dumps(line), flush=True)


def bench_grid(A, args):
    """The upstream gridder (vcg) on the archived grid of the shipped
    inputparameters (Umax 602, FWHM 1480.73 -> NN 1176, dU 1.0254), 768
    channels of 2 stokes, 100k synthetic groups with radii uniform in
    [0, 1.05 Umax) (denser at short baselines, 5% beyond Umax), 20% of
    weights flagged, gridded with flags into a resident cube.  value is
    gridded (group, channel) visibilities per second of kernel time; the
    host->device copy of the group data and the cube conversion are outside
    the timed kernels and reported apart."""
    import oracle as O
    nu0, Umax, FWHM, nc, nst, ng = np.float32(1.54e8), 602.0, 1480.73, 768, 2, 100000
    dU, NN = A.grid_setup(Umax, FWHM)
    rng = np.random.default_rng(7001)
    r = Umax * 1.05 * rng.random(ng)
    th = 2 * np.pi * rng.random(ng)
    ranpar = np.zeros((ng, 5), np.float32)
    ranpar[:, 0] = r * np.cos(th) / nu0
    ranpar[:, 1] = r * np.sin(th) / nu0
    data = rng.standard_normal((ng, nc, nst, 3), dtype=np.float32)
    data[..., 2] = (rng.random((ng, nc, nst), dtype=np.float32) >= 0.2)
    ctx = A.Context(device=0)
    cube, n = ctx.grid_cube(NN, nc, ranpar, data, nu0, dU, Umax)
    ctx.release_cube(cube)
    ctx.reset_stats()
    steps = max(args.steps, 1)
    t = time.time()
    for _ in range(steps):
        cube, n = ctx.grid_cube(NN, nc, ranpar, data, nu0, dU, Umax)
        ctx.release_cube(cube)
    wall = (time.time() - t) / steps
    st = ctx.stats()
    ms = st["ms_grid"] / steps
    cells, _ = A.grid_cells(NN, ranpar, nu0, dU, Umax)
    nonempty = len(np.unique(cells[cells >= 0]))
    vis = n * nc
    byt = vis * 12.0 + nonempty * nc * 24.0
    line = {"metric": "gridded visibilities per second (vcg nearest-cell gridding)", "value": round(vis / (ms * 1e-3)),
            "unit": "visibilities/s", "steps": steps, "ms_per_step": round(ms, 3), "dtype": "f32 in, f64 sums",
            "data": "synthetic", "config": {"workload": "vcg grid NN 1176 (Umax 602, FWHM 1480.73), 768 ch x 2 stokes, "
                                                        "100k groups, flagged"},
            "groups_gridded": n, "cells_nonempty": nonempty,
            "roofline": {"bound": "hbm", "kernel": "k_grid", "achieved": round(byt / (ms * 1e-3) / 1e9, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "bytes_per_launch": byt},
            "s_per_call_with_upload_and_cube": round(wall, 3)}
    ctx.close()
    if not args.no_cpu_baseline:
        ncs = 32
        sub = np.ascontiguousarray(data[:, :ncs])
        t = time.time()
        _, _, on = O.grid(NN, ncs, ranpar, sub, nu0, dU, Umax)
        dt = time.time() - t
        line["cpu_baseline"] = {"value": round(on * ncs / dt), "unit": "visibilities/s", "cores": 1, "kind": "port",
                                "sample": "the oracle's sequential gridding loop over the first %d channels of the "
                                          "same groups (grid allocation included)" % ncs}
    print(json.dumps(line), flush=True)


def copy_calib(mib):
    """A plain HBM copy of known size, for scaling the FETCH_SIZE/WRITE_SIZE
    counters of the contraction: each dispatch reads and writes mib MiB."""
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calib_bytes": n, "dispatches": 3}), flush=True)


def canonical(tris):
    """the triangle list in the reference's (r1, r2, r3) order, whatever
    order the timed slices use"""
    return tris[np.lexsort((tris[:, 2], tris[:, 1], tris[:, 0]))]


def fp64_error(A, ctx, cube, cfg, gv, samp, uf, fk, lo, hi, tris, deadline=None, cache=None):
    """Relative error vs the FP64 oracle on sampled lag cells of four
    triangles (Ng quantiles, the largest first); a triangle whose oracle run
    would pass the deadline is left out and said so.  cache (a dict): the
    oracle's cells by triangle, filled on the first call and reused by later
    ones (another precision form on the same cube)."""
    import oracle as O
    N, nc, nc10 = cfg["N"], cfg["nc"], cfg["nc10"]
    mask = O.chan_mask(nc, fk, uf)
    tris = canonical(tris)
    ngs = np.array([A.ng(hi[t[0]]) for t in tris])
    # triangles at the 25/50/75 % quantiles of Ng and the largest; a bin whose
    # three rings cannot close a triangle on the grid has B == 0 exactly and
    # is skipped (its relative error is undefined)
    order = np.argsort(ngs, kind="stable")
    picks = sorted(set(int(order[min(len(order) - 1, int(f * len(order)))]) for f in (0.25, 0.5, 0.75, 1.0)),
                   key=lambda k: -ngs[k])
    rng = np.random.default_rng(7)
    out = []
    last_s = 0.0
    for k in picks:
        if deadline is not None and time.time() + 1.5 * last_s > deadline:
            out.append({"triangle": [int(x) for x in tris[k]], "skipped": "run budget"})
            continue
        r1, r2, r3 = map(int, tris[k])
        res = ctx.bispec_ex(cube, r1, r2, r3)
        t0 = time.time()
        if cache is not None and (r1, r2, r3) in cache:
            sel, b64, c64, babs = cache[(r1, r2, r3)]
        else:
            cells = np.argwhere(res["cnt"] > 0)
            sel = cells[rng.choice(len(cells), size=min(48, len(cells)), replace=False)]
            sel = np.vstack([sel, [[nc10, nc10]]])
            d2 = sel[:, 0] - nc10
            d3 = sel[:, 1] - nc10
            b64, c64, _, babs = O.bispec_cells(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10, d2, d3,
                                               nthreads=min(16, os.cpu_count() or 1), want_abs=True)
            if cache is not None:
                cache[(r1, r2, r3)] = (sel, b64, c64, babs)
        got = res["bl64"][sel[:, 0], sel[:, 1]]
        if np.linalg.norm(b64) <= 1e-12 * np.linalg.norm(babs):
            log("triangle %s: B == 0 up to FP64 rounding (rings cannot close), skipped" % ((r1, r2, r3),))
            continue
        # noise cubes cancel, so a lag cell's |B| can sit near zero: per-cell
        # errors are scaled by the triangle's RMS |B| over the sampled cells
        scale = float(np.sqrt(np.mean(b64 ** 2)))
        rel = np.abs(got - b64) / max(scale, 1e-300)
        out.append({"triangle": [r1, r2, r3], "ng": int(res["ng"]), "cells": int(len(sel)),
                    "counts_exact": bool(np.array_equal(res["cnt"][sel[:, 0], sel[:, 1]], c64)),
                    "cell_err_over_rms": {"median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
                                          "max": float(rel.max())},
                    "normwise": float(np.linalg.norm(got - b64) / np.linalg.norm(b64)),
                    "max_err_over_sum_abs_terms": float(np.max(np.abs(got - b64) / np.maximum(babs, 1e-300))),
                    "oracle_s": round(time.time() - t0, 1)})
        last_s = time.time() - t0
        log("fp64 check triangle %s Ng=%d: normwise %.3g" % ((r1, r2, r3), res["ng"], out[-1]["normwise"]))
    return out


def cpu_baseline(A, cfg, gv, samp, uf, fk, lo, hi, tris, budget_s):
    """Reference-algorithm CPU path (FP64 c2r + FP32 triplet loop, OpenMP),
    timed on a bounded sample and extrapolated linearly in V*Ng^2."""
    import oracle as O
    tris = canonical(tris)
    N, nc, nc10 = cfg["N"], cfg["nc"], cfg["nc10"]
    mask = O.chan_mask(nc, fk, uf)
    threads = min(16, os.cpu_count() or 1)
    ngs = np.array([A.ng(hi[t[0]]) for t in tris])
    k = int(np.argsort(ngs)[len(ngs) // 2])
    r1, r2, r3 = map(int, tris[k])
    P = int(ngs[k]) ** 2
    act = np.flatnonzero(mask)
    # probe one reference plane, then size the sample to the budget
    bl, _, cnt, n1, tf, tl = O.refalgo(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10,
                                       p1_lo=int(act[0]), p1_hi=int(act[0]) + 1, nthreads=threads)
    per_trip_px = tl / max(n1 * P, 1)
    nplanes_filled = 3 * min(nc, 1 + nc10 + 1)
    fill_per_plane_px = tf / max(nplanes_filled * P, 1)
    kp = int(max(1, min(nc, (budget_s - tf) / max(tl, 1e-3))))
    bl, _, cnt, n2, tf2, tl2 = O.refalgo(N, nc, 1.0, gv, samp, lo, hi, r1, r2, r3, mask, nc10,
                                         p1_lo=0, p1_hi=kp, nthreads=threads)
    per_trip_px = tl2 / max(n2 * P, 1)
    # whole workload: every triangle fills 3*nc planes and loops over V triplets
    V = int(np.sum(mask)) ** 3 if nc10 >= nc - 1 else None
    if V is None:
        m = mask.astype(bool)
        V = 0
        for p1 in np.flatnonzero(m):
            w = m[max(0, p1 - nc10): p1 + nc10 + 1].sum()
            V += int(w) * int(w)
    total = 0.0
    evals = 0
    W = 2 * nc10 + 1
    ev_tri = None
    for t, g in zip(tris, ngs):
        Pt = int(g) ** 2
        total += per_trip_px * V * Pt + fill_per_plane_px * 3 * nc * Pt
    ev_tri = int(np.count_nonzero(cnt_full(mask, nc10)))
    evals = ev_tri * len(tris)
    return {"value": round(evals / total, 2), "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": "reference-algorithm CPU path (oracle/orc_refalgo.c: FP64 c2r fill + FP32 triplet "
                      "loop, OpenMP dynamic) on triangle (%d,%d,%d) Ng=%d, p1 in [0,%d) = %d triplets, "
                      "%.1f s loop + %.1f s fill; extrapolated linearly in V*Ng^2 to all %d triangles"
                      % (r1, r2, r3, int(ngs[k]), kp, n2, tl2, tf2, len(tris)),
            "est_total_s": round(total, 1)}


def cnt_full(mask, nc10):
    nc = len(mask)
    W = 2 * nc10 + 1
    m = mask.astype(np.int64)
    out = np.zeros((W, W), np.int64)
    for i2 in range(W):
        d2 = i2 - nc10
        for p1 in range(max(0, -d2), min(nc, nc - d2)):
            if m[p1] and m[p1 + d2]:
                lo = max(0, -p1) if False else max(-nc10, -p1)
                hi_ = min(nc10, nc - 1 - p1)
                out[i2, lo + nc10: hi_ + nc10 + 1] += m[p1 + lo: p1 + hi_ + 1]
    return out


if __name__ == "__main__":
    main()
