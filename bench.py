"""Benchmark: exact sequential sampling on one MI355X.

Workload (one step = one exact sample): the n = 70 banded random graph state with band 30 (seed 2, natural order,
local Cliffords + RZ(alpha) RX(beta) tails), planned into an in-place arena of 2^30 entries; complex64 arena with FP64
arithmetic and reductions.  The timed region covers K samples with the plan and arena resident on the device.

The JSON line also carries:
  roofline      -- the dominant sweep kernel (the template instance with the most device time): algorithmic bytes
                   of its launches / their summed HIP-event durations, against the 8 TB/s HBM peak;
  cpu_baseline  -- one sample of the same plan by the reference executor compiled from its sources (oracle/_ref), or
                   this project's CPU restatement (oracle/gs_oracle) when that is absent, on the host's cores;
  c64_vs_c128   -- the same samples' seeds in a complex128 arena: speedup and max |d log2 Pr| (c64 bits re-scored
                   in c128 with prob mode);
  ibm_w30       -- exact samples of the reference's own width-30 plan of the 70-qubit IBM circuit
                   (tests/golden/ibm/nq70_inplace_plan.txt, Sum_k 2^r_k = 2^33), beside the reference's published
                   11.3 s per sample for that plan (README.md:24, laptop CPU).
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
SWEEP_NAMES = ("gram", "direct", "write", "write+gram", "write+direct", "write+gram of a one-slot next step",
               "pattern sums of a run of diagonal steps", "product pass of a run of diagonal steps")
ROUTES = ("m1", "split", "mixA", "mixB")


def per_step_table(steps, ms, nl, geo, title):
    """Markdown table of every step's sweeps from a profiled pass: average launch time (HIP events) and bandwidth on the
    algorithmic bytes of the launch."""
    out = [title, "", "| step | nk | slots | outputs | sweep | blocks | VG | us | GB | TB/s | of 8 TB/s |",
           "|---|---|---|---|---|---|---|---|---|---|---|"]
    for k, s in enumerate(steps):
        for sw in range(nl.shape[1]):
            if nl[k, sw] == 0:
                continue
            g = geo[k]
            us = ms[k, sw] / nl[k, sw] * 1e3
            b = sweep_bytes(s, sw, 8)
            fused = sw in (3, 4)
            out.append(f"| {k} | {s['nk']} | {g['ntq']} | {s['m']} | {SWEEP_NAMES[sw]} | "
                       f"{g['fgrid'] if fused else g['grid']} | {g['fvg'] if fused else g['vg']} | {us:.1f} | "
                       f"{b / 1e9:.3f} | {b / us / 1e6:.2f} | {b / us / 1e6 / 8:.2f} |")
    return "\n".join(out)


def sweep_bytes(step, sweep, esz):
    """Algorithmic bytes of one sweep launch on a step: norm sweeps read the live slots; write sweeps (also the ones
    that fold in the next step's norms, which read nothing more) read the live slots and write the outputs."""
    ntq = len({t[0] for t in step["terms"]})
    read = (1 << step["nk"]) * ntq * esz
    if sweep == 5:                      # pattern sums of a run: one read of its table
        return read
    if sweep == 6:                      # product pass of a run: read and write
        return 2 * read
    if sweep >= 2:
        return read + (1 << step["nk"]) * (2 if step["m"] == 2 else 1) * esz
    return read


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=("c2", "c3", "c4", "c5", "ibm", "tt"),
                    help="c2: exact sampling at width 30 (the headline line); c3 batched scoring; c4 gradients; "
                         "c5 width-34 sampling; ibm: samples of the reference's width-30 IBM plan; tt: the two-table "
                         "engine on the reference's IBM two-table plans (each prints its own JSON line)")
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--profile-steps", type=int, default=32, help="samples in the profiled (per-kernel timing) pass")
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--c128-steps", type=int, default=256)
    ap.add_argument("--ibm-steps", type=int, default=8, help="samples of the IBM width-30 plan (0: skip)")
    ap.add_argument("--tt-steps", type=int, default=2, help="samples of the two-table width-31 plan (0: skip)")
    ap.add_argument("--c3-strings", type=int, default=8192)
    ap.add_argument("--c3-lanes", type=int, default=512)
    ap.add_argument("--c3-single", type=int, default=256, help="strings timed one per launch sequence")
    ap.add_argument("--c3-budget", type=float, default=120.0, help="GB of arena per batch in the mixed schedule")
    ap.add_argument("--c3-mixed", type=int, default=1, help="run the 96-instance mixed schedule (0: skip)")
    ap.add_argument("--c5-samples", type=int, default=32)
    ap.add_argument("--c5-strings", type=int, default=256)
    ap.add_argument("--c5-c128", type=int, default=4, help="c5 samples/strings re-scored in a complex128 arena")
    ap.add_argument("--no-fusion", action="store_true", help="run every norm pass as its own sweep")
    ap.add_argument("--fused-run", type=int, default=None, metavar="LR",
                    help="log2 run length of the fused layout (complex64: 1 or 2; default: engine default)")
    ap.add_argument("--plain-pairs", type=int, default=None, metavar="NTQ",
                    help="plain sweeps of steps with at most NTQ slots per coset take two runs per thread (0: off)")
    ap.add_argument("--no-xcd-blocks", action="store_true",
                    help="width-34 workload: deal blocks round-robin to the XCDs instead of in contiguous eighths")
    ap.add_argument("--per-step", action="store_true",
                    help="width-26 and IBM workloads: print every step's sweep times and bandwidth (markdown, stderr)")
    ap.add_argument("--tt-grid-log2", type=int, default=0, metavar="LG",
                    help="two-table engine: cap its gather sweeps at 2^LG blocks (0: engine default)")
    ap.add_argument("--no-fused4", action="store_true",
                    help="norm passes of four-slot steps as their own sweeps (the rest stay folded)")
    ap.add_argument("--fold-reduce", type=int, default=0, metavar="R",
                    help="reduce a sweep's block partials in the next sweep instead of in a one-block kernel (1: every "
                         "sweep; R >= 2: next sweeps of steps with at most 2^R live slots; 0: never)")
    ap.add_argument("--grid-log2", type=int, default=0, help="cap sweeps at 2^N blocks (default: engine default)")
    ap.add_argument("--grid-log2-one-slot", type=int, default=0,
                    help="cap the plain sweeps of one-slot steps and the passes of runs of diagonal steps at 2^N blocks "
                         "(default: engine default)")
    ap.add_argument("--small-steps", type=int, default=None, metavar="NK",
                    help="runs of steps with at most 2^NK cosets go to one single-block launch (-1: never; default: "
                         "engine default)")
    ap.add_argument("--legs", default="c3,c4,c5",
                    help="short legs of the other workloads appended to the headline line (comma list; '' for none)")
    ap.add_argument("--cpu-reps", type=int, default=3, help="repetitions of each CPU baseline timing (median)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="first run one plain device copy of this many MiB (byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the in-place arena of one instance lives on one device (asked for "
              f"--gpus {args.gpus}); sharding across GPUs is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    from _pkg import egs
    gs = egs.gsamp
    if args.calibrate_copy:
        ms = gs.copy_probe(args.calibrate_copy << 20)
        print(f"copy probe: {args.calibrate_copy} MiB in {ms:.3f} ms = {2 * (args.calibrate_copy << 20) / ms / 1e6:.1f} GB/s "
              f"(read + write)", file=sys.stderr, flush=True)
    if args.workload != "c2":
        out = {"c3": workload_c3, "c4": workload_c4, "c5": workload_c5,
               "ibm": lambda gs, egs, args: ibm_w30(gs, egs, args.ibm_steps, args),
               "tt": lambda gs, egs, args: ibm_two_table(gs, egs, args.tt_steps, args.tt_grid_log2)}[args.workload](
            gs, egs, args)
        print(json.dumps(out))
        return
    inst = egs.instances.config_instance("c2")
    tmp = tempfile.mkdtemp(prefix="gs_bench_")
    plan_path = os.path.join(tmp, "c2_plan.txt")
    inst.write_plan(plan_path)
    _, R, steps = egs.planio.read_inplace_plan(plan_path)
    traffic = egs.planio.plan_traffic(steps, 8)
    print(f"workload: n=70 band 30 seed 2, width {inst.width}, R {R}, log2 sum 2^r = {inst.log2_sum_live():.3f}, "
          f"{traffic['sample'] / 1e9:.2f} GB algorithmic per sample (c64)", file=sys.stderr, flush=True)

    K, W = args.steps, args.warmup
    KP = min(K, args.profile_steps)
    with gs.Engine(plan_path, gs.C64) as e:
        if args.no_fusion:
            e.set_fusion(False)
        if args.no_fused4:
            e.set_fused4(False)
        if args.fused_run is not None:
            e.set_fused_run(args.fused_run)
        if args.plain_pairs is not None:
            e.set_plain_pairs(args.plain_pairs)
        if args.fold_reduce:
            e.set_fold_reduce(args.fold_reduce)
        if args.grid_log2:
            e.set_grid_log2(args.grid_log2)
        if args.small_steps is not None:
            e.set_small_steps(args.small_steps)
        if W:
            e.sample(W, args.seed + 10 ** 6)
        # timed region: K samples, no instrumentation.  Engine calls are synchronous (each returns after its stream
        # has drained), so the wall clock brackets exactly the K samples' device work.  torch is not loaded: it
        # bundles a second HIP runtime.
        t0 = time.perf_counter()
        bits, lp = e.sample(K, args.seed)
        wall = time.perf_counter() - t0
        dev_ms = e.last_device_ms
        # profiled pass over the first KP of the same samples: HIP events around every sweep launch on the engine's
        # stream give each kernel's duration (the events add a little time between launches, not inside them)
        e.profile(True)
        t0 = time.perf_counter()
        bits_p, _ = e.sample(KP, args.seed)
        wall_p = time.perf_counter() - t0
        ms, nl = e.profile_read()
        e.profile(False)
        assert np.array_equal(bits_p, bits[:KP])
        geo = [e.geometry(k) for k in range(e.ns)]
        if args.per_step:
            log(per_step_table(steps, ms, nl, geo, f"per-step sweeps, {KP} samples of the width-26 plan (HIP events)"))
    s_per_sample = wall / K

    # dominant kernel: aggregate per template instance <storage, VR, VG, ntq, route, sweep, next route>
    agg = {}
    for k, s in enumerate(steps):
        gk = geo[k]
        for sw in range(7):
            if nl[k, sw] == 0:
                continue
            if sw >= 5:                                                           # a run of diagonal steps
                key = (2, 1, 1, 0, sw + 1, 0)
            elif sw >= 3 and gk["fvg"] in (2, 4):
                key = (1 << gk["flr"], gk["fvg"], gk["ntq"], gk["route"], sw, gk["route2"])   # fused layouts
            elif sw >= 3:
                key = (1 << gk["flr"], 1, gk["ntq"], gk["route"], 5, 0)           # plain layout folding a 1-slot Gram
            else:
                key = (1 << gk["lr"], gk["vg"], gk["ntq"], gk["route"], sw, 0)
            a = agg.setdefault(key, dict(ms=0.0, n=0, bytes=0.0, steps=[]))
            a["ms"] += ms[k, sw]
            a["n"] += int(nl[k, sw])
            a["bytes"] += sweep_bytes(s, sw, 8) * nl[k, sw]
            a["steps"].append(k)
    moved = sum(a["bytes"] for a in agg.values()) / KP          # bytes the sweeps actually stream per sample
    key, dom = max(agg.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["bytes"] / (dom["ms"] / 1e3) / 1e9
    vr, vg, ntq, route, sw, route2 = key
    kname = (f"gs::sweep_kernel<float2, {vr}, {vg}, {ntq}, {route}, {sw}, {route2}>" if sw < 6 else
             "gs::run_hist_kernel<float2>" if sw == 6 else "gs::run_apply_kernel<float2>")
    kdesc = (f"{vr * vg} cosets/thread ({vg} run(s) of {vr}), {ntq} slots/coset, route {ROUTES[route]}, "
             f"{SWEEP_NAMES[sw]} sweep" + (f", next route {ROUTES[route2]}" if sw in (3, 4) else ""))
    traffic_meas = profiled_traffic(kname)
    total_sweep_ms = float(ms.sum())
    print(f"dominant kernel {kname} ({kdesc}): {dom['n']} launches, {dom['ms'] / dom['n']:.4f} ms avg, {achieved:.1f} GB/s; "
          f"all sweeps {total_sweep_ms / KP:.3f} ms/sample (profiled pass: {1e3 * wall_p / KP:.3f} ms/sample wall); "
          f"timed: {1e3 * s_per_sample:.3f} ms/sample", file=sys.stderr)

    # complex128 comparison (same seeds) and c64 bits re-scored in c128
    K2 = min(args.c128_steps, K)
    with gs.Engine(plan_path, gs.C128) as e2:
        e2.sample(1, args.seed + 10 ** 6)
        t0 = time.perf_counter()
        bits2, lp2 = e2.sample(K2, args.seed)
        wall2 = (time.perf_counter() - t0) / K2
        rescored = e2.prob(bits[:K2])
    dl = np.abs(rescored - lp[:K2])
    dlog = float(np.max(dl))
    same_bits = int(sum(int(np.array_equal(a, b)) for a, b in zip(bits[:K2], bits2)))

    ibm = ibm_w30(gs, egs, args.ibm_steps, args) if args.ibm_steps else None
    tt2 = ibm_two_table(gs, egs, args.tt_steps, args.tt_grid_log2) if args.tt_steps else None

    legs = {}
    for name in [x for x in args.legs.split(",") if x]:
        t0 = time.perf_counter()
        legs[name] = {"c3": leg_c3, "c4": leg_c4, "c5": leg_c5}[name](gs, egs)
        legs[name]["leg_wall_s"] = time.perf_counter() - t0
        log(f"leg {name}: {json.dumps(legs[name])}")

    cpu = cpu_detail = None
    if not args.no_cpu:
        cpu, cpu_detail = cpu_baseline(plan_path, args.seed, args.cpu_reps)

    out = {
        "metric": "s/exact sample (n=70 band-30 banded graph state, 2^30 in-place arena, complex64 + FP64 sums)",
        "value": s_per_sample,
        "unit": "s/sample",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * s_per_sample,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage": "c64",
        "data": "synthetic",
        "config": {"workload": "exact sequential sampling, n=70 banded G(n,1/2) band 30 (seed 2), width "
                               f"{inst.width}, R=30, 1 sample per step",
                   "n": 70, "band": 30, "R": R, "width": inst.width, "log2_sum_2r": round(inst.log2_sum_live(), 3)},
        "hbm_gbps_algorithmic": traffic["sample"] / s_per_sample / 1e9,
        "gb_per_sample_algorithmic": traffic["sample"] / 1e9,
        "gb_per_sample_moved": moved / 1e9,
        "hbm_gbps_moved": moved / s_per_sample / 1e9,
        "device_ms_per_sample": dev_ms / K,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_meas, "kernel": kname, "kernel_desc": kdesc,
                     "traffic_source": TRAFFIC_PROFILE if traffic_meas else None,
                     "launches": dom["n"], "avg_ms": dom["ms"] / dom["n"],
                     "bytes_per_launch": dom["bytes"] / dom["n"],
                     "timing": f"HIP events around each launch, profiled pass over the first {KP} of the timed samples",
                     "sweeps_ms_per_sample": total_sweep_ms / KP},
        "c64_vs_c128": {"c128_s_per_sample": wall2, "speedup": wall2 / s_per_sample,
                        "max_abs_dlog2p": dlog, "mean_abs_dlog2p": float(np.mean(dl)), "samples_compared": K2,
                        "identical_bits": same_bits,
                        "method": "the same seeds sampled in a complex128 arena (bits compared) and the complex64 "
                                  "samples re-scored there (prob mode)"},
        "cpu_baseline": cpu,
        "cpu_baseline_detail": cpu_detail,
        "ibm_w30": ibm,
        "ibm_two_table": tt2,
    }
    if "c4" in legs:                          # the metric's third part: loss gradients per second (c64)
        out["grads_per_s"] = legs["c4"]["c64"]["grads_per_s"]
        out["string_grads_per_s"] = legs["c4"]["c64"]["string_grads_per_s"]
    out.update(legs)
    print(json.dumps(out))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload_c3(gs, egs, args):
    """Batched scoring.  (a) 8192 strings on the n = 70, band-26 instance (seed 3): half uniform, half exact samples
    drawn by the sequential sampler; batched chain-mode scoring (arena of 2^w x B lanes, tiles from the autotuner)
    against one string per launch sequence (the single-string engine's chain mode).  (b) 96 instances, n in
    {40, 56, 70} with bands {16, 22, 26} (32 each, seeds 300..395), 64 strings each, scheduled into batches under a
    memory budget by the autotuner; p50 / p99 per-instance latency."""
    at = egs.autotune
    tmp = tempfile.mkdtemp(prefix="gs_c3_")
    cache = os.path.join(tmp, "tiles.json")
    inst = egs.instances.config_instance("c3")
    path = os.path.join(tmp, "c3.txt")
    inst.write_plan(path)
    ns = len(inst.steps)
    N, B = args.c3_strings, args.c3_lanes
    rng = np.random.default_rng(3)
    strings = rng.integers(0, 2, size=(N, ns)).astype(np.int8)
    res = {"workload": "c3 batched scoring", "width": inst.width, "strings": N, "lanes": B}
    with gs.Engine(path, gs.C64) as e:
        t0 = time.perf_counter()
        strings[N // 2:], _ = e.sample(N - N // 2, 33)
        log(f"c3: {N - N // 2} exact samples drawn in {time.perf_counter() - t0:.1f} s")
        n1 = min(N, args.c3_single)
        e.chain(strings[:4])
        t0 = time.perf_counter()
        lp1, _, _ = e.chain(strings[:n1])
        t1 = (time.perf_counter() - t0) / n1
    log(f"c3: one string per launch sequence: {1 / t1:.0f} amplitudes/s")
    with gs.BatchEngine([path], [B], storage=gs.C64) as bt:
        t0 = time.perf_counter()
        tt = at.tune_tiles(bt, [strings[:B]], cache=cache)
        log(f"c3: tiles tuned in {time.perf_counter() - t0:.1f} s ({tt.get('choice', 'cached')}, unroll {tt['lu']})")
        bt.use_graph(True)
        bt.score([strings[:B]], want_cond=False, want_amp=False)
        lps = []
        t0 = time.perf_counter()
        for i in range(0, N, B):
            lp, _, _ = bt.score([strings[i:i + B]], want_cond=False, want_amp=False)
            lps.append(lp[0])
            log(f"c3 batched: {i + B} strings, {(i + B) / (time.perf_counter() - t0):.0f} amplitudes/s")
        tb = time.perf_counter() - t0
    lpb = np.concatenate(lps)
    res["single"] = {"amplitudes_per_s": 1 / t1, "strings_timed": n1}
    res["batched"] = {"amplitudes_per_s": N / tb, "tiles": {"lg": tt["lg"], "lu": tt["lu"], "choice": tt.get("choice")},
                      "speedup_vs_single": (N / tb) * t1,
                      "max_abs_dlog2p_vs_single": float(np.max(np.abs(lpb[:n1] - lp1)))}
    log(f"c3 batched: {N / tb:.0f} amplitudes/s")
    if not args.c3_mixed:
        return res
    # mixed instances
    plans, widths, bits = [], [], []
    for j, (n, band) in enumerate([(40, 16)] * 32 + [(56, 22)] * 32 + [(70, 26)] * 32):
        mi = egs.instances.Instance(n, band, 300 + j)
        pth = os.path.join(tmp, f"m{j}.txt")
        mi.write_plan(pth)
        plans.append(gs.Plan(path=pth))
        widths.append(mi.R)
        bits.append(np.random.default_rng(1000 + j).integers(0, 2, size=(64, len(mi.steps))).astype(np.int8))
    log(f"c3 mixed: {len(plans)} instances, widths {min(widths)}..{max(widths)}")
    best, sched, _, tiles = at.tune_schedule(gs, plans, widths, 64, bits, args.c3_budget * 1e9, gs.C64,
                                             tile_cache=cache, log=log)
    res["mixed"] = {"instances": len(plans), "lanes": 64, "widths": [min(widths), max(widths)],
                    "budget_gb": args.c3_budget, "best_schedule": best, "schedules": sched, "unroll": tiles["lu"]}
    log(f"c3 mixed: {best} {sched[best]}")
    return res


def workload_c4(gs, egs, args):
    """Gradients: n = 60, band 24 (seed 4), 120 trainable angles (alpha, beta per qubit); loss = mean -ln Pr(b) over
    1024 seeded uniform strings; forward = batched chain scoring, backward = the adjoint with checkpoint-and-recompute."""
    inst = egs.instances.config_instance("c4")
    model = egs.adjoint.AngleModel(inst)
    path = os.path.join(tempfile.mkdtemp(prefix="gs_c4_"), "c4.txt")
    inst.write_plan(path)
    B = 1024
    bits = np.random.default_rng(4).integers(0, 2, size=(B, len(inst.steps))).astype(np.int8)
    log(f"c4: n=60 band 24 width {inst.width}, {B} strings")
    res = {"workload": "c4 gradients, n=60 band 24 (seed 4), 120 angles, loss over 1024 strings", "width": inst.width}
    for storage, tag in ((gs.C64, "c64"), (gs.C128, "c128")):
        with gs.BatchEngine([path], [B], storage=storage) as bt:
            bt.use_graph(True)
            M, dA, dB = model.matrices(model.alpha0, model.beta0)
            bt.set_matrices(0, M)
            bt.score([bits], want_cond=False, want_amp=False)
            t0 = time.perf_counter()
            reps = 5
            for _ in range(reps):
                lp, _, _ = bt.score([bits], want_cond=False, want_amp=False)
            fwd_wall = (time.perf_counter() - t0) / reps
            fwd_dev = bt.last_ms()[0]
            bt.grad([bits])
            t0 = time.perf_counter()
            for _ in range(reps):
                (r,), seg = bt.grad([bits])
            grad_wall = (time.perf_counter() - t0) / reps
            tot_dev, gfwd_dev = bt.grad_ms()
            ga, gb = model.grad_ln_pr(bits, r[1], r[2], dA, dB)
            loss = float(-np.mean(lp[0]) * np.log(2))
        res[tag] = {"loss": loss, "forward_ms": 1e3 * fwd_wall, "forward_device_ms": fwd_dev,
                    "grad_ms": 1e3 * grad_wall, "grad_device_ms": tot_dev, "grad_forward_part_device_ms": gfwd_dev,
                    "backward_over_forward": (tot_dev - gfwd_dev) / fwd_dev, "grad_over_forward": tot_dev / fwd_dev,
                    "grads_per_s": 1.0 / grad_wall, "string_grads_per_s": B / grad_wall, "checkpoint_segment": seg,
                    "grad_norm": float(np.sqrt(np.sum((-gb.mean(0)) ** 2) + np.sum((-ga.mean(0)) ** 2)))}
        log(f"c4 {tag}: {res[tag]}")
    return res


def workload_c5(gs, egs, args):
    """Width 34: n = 80, band 34 (seed 5) plus 40 phase-gadget ancillas projected on 0, complex64 arena of 2^34
    entries (137 GB).  Exact samples of the data qubits (hybrid mode: ancillas forced, data sampled) and chain-mode
    log2 Pr of random strings; a complex128 rerun (2^34 x 16 B = 275 GB) on a subset gives |d log2 Pr|."""
    t0 = time.perf_counter()
    inst = egs.instances.config_instance("c5")
    path = os.path.join(tempfile.mkdtemp(prefix="gs_c5_"), "c5.txt")
    inst.write_plan(path)
    _, R, steps = egs.planio.read_inplace_plan(path)
    traffic = egs.planio.plan_traffic(steps, 8)
    nanc = inst.N - inst.n
    ns = len(steps)
    log(f"c5: N={inst.N} width {inst.width} R={R} ({time.perf_counter() - t0:.0f} s to plan), "
        f"{traffic['sample'] / 1e12:.2f} TB algorithmic per sample")
    nsamp, nstr, nsub = args.c5_samples, args.c5_strings, args.c5_c128
    hy = np.full((nsamp, ns), -1, np.int8)
    hy[:, :nanc] = 0                                            # ancillas are the first nanc steps
    rnd = np.random.default_rng(55).integers(0, 2, size=(nstr, ns)).astype(np.int8)
    rnd[:, :nanc] = 0
    res = {"workload": "c5 width 34: n=80 band 34 seed 5 + 40 gadget ancillas", "N": inst.N, "width": inst.width,
           "R": R, "tb_per_sample_algorithmic": traffic["sample"] / 1e12}
    with gs.Engine(path, gs.C64) as e:
        if args.grid_log2_one_slot:
            e.set_grid_log2_one_slot(args.grid_log2_one_slot)
        if args.no_xcd_blocks:
            e.set_xcd_blocks(False)
        obs, lpss, lpcs = [], [], []
        ts = tc = 0.0
        for i in range(0, nsamp, 8):                            # chunks, so progress shows every few seconds
            t0 = time.perf_counter()
            ob, lp, _ = e.hybrid(hy[i:i + 8], 5, i)
            ts += time.perf_counter() - t0
            obs.append(ob)
            lpss.append(lp)
            log(f"c5 c64: samples {i}..{i + len(lp) - 1} done, {ts / (i + len(lp)):.3f} s/sample")
        for i in range(0, nstr, 16):
            t0 = time.perf_counter()
            lp, _, _ = e.chain(rnd[i:i + 16])
            tc += time.perf_counter() - t0
            lpcs.append(lp)
            log(f"c5 c64: strings {i}..{i + len(lp) - 1} done, {tc / (i + len(lp)):.3f} s/string")
        ob, lps, lpc = np.concatenate(obs), np.concatenate(lpss), np.concatenate(lpcs)
        ts /= nsamp
        tc /= nstr
    res["c64"] = {"s_per_sample": ts, "samples": nsamp, "hbm_gbps_algorithmic": traffic["sample"] / ts / 1e9,
                  "s_per_string_chain": tc, "strings": nstr,
                  "chain_gbps_algorithmic": traffic["chain1"] / tc / 1e9}
    if nsub:
        try:
            with gs.Engine(path, gs.C128) as e2:
                t0 = time.perf_counter()
                lp2 = e2.prob(ob[:nsub])
                t2 = (time.perf_counter() - t0) / nsub
                lpc2, _, _ = e2.chain(rnd[:nsub])
            res["c128"] = {"strings": nsub, "s_per_string_prob": t2,
                           "max_abs_dlog2p_samples": float(np.max(np.abs(lp2 - lps[:nsub]))),
                           "max_abs_dlog2p_strings": float(np.max(np.abs(lpc2 - lpc[:nsub])))}
        except gs.GsError as ex:
            res["c128"] = {"error": str(ex)}
    return res


def leg_c4(gs, egs):
    """Gradients, short form of the c4 workload: n = 60, band 24 (seed 4), 120 angles, loss over 1024 strings; 3 timed
    forward passes and 3 timed gradients per storage, plus a complex64 gradient with checkpoints every 10 steps, which
    recomputes the table inside every segment (the default complex64 budget keeps every step's input)."""
    inst = egs.instances.config_instance("c4")
    model = egs.adjoint.AngleModel(inst)
    path = os.path.join(tempfile.mkdtemp(prefix="gs_c4_"), "c4.txt")
    inst.write_plan(path)
    B, reps = 1024, 3
    bits = np.random.default_rng(4).integers(0, 2, size=(B, len(inst.steps))).astype(np.int8)
    res = {"workload": "c4 gradients, n=60 band 24 (seed 4), 120 angles (fp32), loss = mean -ln Pr over 1024 strings",
           "width": inst.width}
    for storage, tag in ((gs.C64, "c64"), (gs.C128, "c128")):
        with gs.BatchEngine([path], [B], storage=storage) as bt:
            bt.use_graph(True)
            M, dA, dB = model.matrices(model.alpha0, model.beta0)
            bt.set_matrices(0, M)
            bt.score([bits], want_cond=False, want_amp=False)
            t0 = time.perf_counter()
            for _ in range(reps):
                lp, _, _ = bt.score([bits], want_cond=False, want_amp=False)
            fwd_wall = (time.perf_counter() - t0) / reps
            fwd_dev = bt.last_ms()[0]
            runs = [("", 0)] + ([("recompute_seg10_", 10)] if storage == gs.C64 else [])
            for prefix, seg in runs:
                bt.grad([bits], seg=seg)
                t0 = time.perf_counter()
                for _ in range(reps):
                    (r,), used = bt.grad([bits], seg=seg)
                grad_wall = (time.perf_counter() - t0) / reps
                tot_dev, gfwd_dev = bt.grad_ms()
                res.setdefault(tag, {}).update({
                    prefix + "grad_ms": 1e3 * grad_wall, prefix + "grad_device_ms": tot_dev,
                    prefix + "grad_over_forward": tot_dev / fwd_dev,
                    prefix + "backward_over_forward": (tot_dev - gfwd_dev) / fwd_dev,
                    prefix + "grads_per_s": 1.0 / grad_wall, prefix + "string_grads_per_s": B / grad_wall,
                    prefix + "checkpoint_segment": used})
            res[tag].update({"loss": float(-np.mean(lp[0]) * np.log(2)), "forward_ms": 1e3 * fwd_wall,
                             "forward_device_ms": fwd_dev})
        log(f"c4 {tag}: {res[tag]}")
    return res


def leg_c3(gs, egs):
    """Batched scoring, short form of the c3 workload: 1024 strings (512 uniform, 512 exact samples) of the n = 70,
    band-26 instance in two 512-lane batches (tiles tuned over 2^10 / 2^12 cosets and unroll 1 / 2, graph replays)
    against one string per launch sequence; and a reduced mix of 12 instances (4 of each (n, band) pair of the full
    mix, 64 lanes each) through the schedule tuner, with p50 / p99 per-instance latency."""
    at = egs.autotune
    tmp = tempfile.mkdtemp(prefix="gs_c3_")
    inst = egs.instances.config_instance("c3")
    path = os.path.join(tmp, "c3.txt")
    inst.write_plan(path)
    ns, N, L = len(inst.steps), 1024, 512
    strings = np.random.default_rng(3).integers(0, 2, size=(N, ns)).astype(np.int8)
    res = {"workload": "c3 batched scoring, n=70 band 26 (seed 3) + a 12-instance mix", "width": inst.width}
    with gs.Engine(path, gs.C64) as e:
        strings[N // 2:], _ = e.sample(N - N // 2, 33)
        e.chain(strings[:4])
        t0 = time.perf_counter()
        lp1, _, _ = e.chain(strings[:128])
        t1 = (time.perf_counter() - t0) / 128
    with gs.BatchEngine([path], [L], storage=gs.C64) as bt:
        tt = at.tune_tiles(bt, [strings[:L]], candidates=(10, 12), unrolls=(1, 2))
        bt.use_graph(True)
        bt.score([strings[:L]], want_cond=False, want_amp=False)
        lps = []
        t0 = time.perf_counter()
        for i in range(0, N, L):
            lp, _, _ = bt.score([strings[i:i + L]], want_cond=False, want_amp=False)
            lps.append(lp[0])
        tb = time.perf_counter() - t0
    lpb = np.concatenate(lps)
    res["single_amplitudes_per_s"] = 1 / t1
    res["batched_amplitudes_per_s"] = N / tb
    res["batched_over_single"] = (N / tb) * t1
    res["tiles"] = {"lg": tt["lg"], "lu": tt["lu"], "choice": tt.get("choice")}
    res["max_abs_dlog2p_batched_vs_single"] = float(np.max(np.abs(lpb[:128] - lp1)))
    plans, widths, bits = [], [], []
    for j, (n, band) in enumerate([(40, 16)] * 4 + [(56, 22)] * 4 + [(70, 26)] * 4):
        seed = 300 + 32 * (j // 4) + j % 4                      # the first four seeds of each pair of the full mix
        mi = egs.instances.Instance(n, band, seed)
        pth = os.path.join(tmp, f"m{j}.txt")
        mi.write_plan(pth)
        plans.append(gs.Plan(path=pth))
        widths.append(mi.R)
        bits.append(np.random.default_rng(1000 + seed - 300).integers(0, 2, size=(64, len(mi.steps))).astype(np.int8))
    best, sched, _, _ = at.tune_schedule(gs, plans, widths, 64, bits, 120e9, gs.C64)
    res["mix"] = {"instances": len(plans), "lanes": 64, "widths": [min(widths), max(widths)], "best_schedule": best,
                  "amplitudes_per_s": sched[best]["amplitudes_per_s"], "p50_s": sched[best]["p50_s"],
                  "p99_s": sched[best]["p99_s"], "schedules": sched}
    return res


def leg_c5(gs, egs):
    """Width 34, short form of the c5 workload: n = 80, band 34 (seed 5) + 40 phase-gadget ancillas in a complex64
    arena of 2^34 entries (137 GB): 2 exact hybrid samples (ancillas forced 0) and chain-mode log2 Pr of 2 random
    strings; the 2 samples re-scored in a complex128 arena (275 GB) give |d log2 Pr|."""
    inst = egs.instances.config_instance("c5")
    path = os.path.join(tempfile.mkdtemp(prefix="gs_c5_"), "c5.txt")
    inst.write_plan(path)
    _, R, steps = egs.planio.read_inplace_plan(path)
    traffic = egs.planio.plan_traffic(steps, 8)
    ns, na = len(steps), inst.N - inst.n
    hy = np.full((2, ns), -1, np.int8)
    hy[:, :na] = 0
    rnd = np.random.default_rng(55).integers(0, 2, size=(2, ns)).astype(np.int8)
    rnd[:, :na] = 0
    with gs.Engine(path, gs.C64) as e:
        t0 = time.perf_counter()
        ob, lps, _ = e.hybrid(hy, 5)
        ts = (time.perf_counter() - t0) / 2
        t0 = time.perf_counter()
        e.chain(rnd)
        tc = (time.perf_counter() - t0) / 2
        e.profile(True)                                         # one more sample with per-kernel events
        e.hybrid(hy[:1], 5)
        ms, nl = e.profile_read()
        e.profile(False)
    moved = sum(sweep_bytes(steps[k], sw, 8) * nl[k, sw] for k in range(ns) for sw in range(nl.shape[1]))
    run_ms = float(ms[:, 5:].sum())
    run_bytes = sum(sweep_bytes(steps[k], sw, 8) * nl[k, sw] for k in range(ns) for sw in (5, 6))
    res = {"workload": "c5 width 34: n=80 band 34 seed 5 + 40 gadget ancillas, R=34, complex64", "R": R,
           "s_per_sample": ts, "hbm_gbps_algorithmic": traffic["sample"] / ts / 1e9,
           "gb_per_sample_moved": moved / 1e9, "hbm_gbps_moved": moved / ts / 1e9,
           "runs_of_diagonal_steps": {"ms_per_sample": run_ms, "gb_per_sample": run_bytes / 1e9,
                                      "gbps": run_bytes / (run_ms / 1e3) / 1e9 if run_ms else None},
           "s_per_string_chain": tc, "chain_gbps_algorithmic": traffic["chain1"] / tc / 1e9, "samples": 2,
           "strings": 2}
    try:
        with gs.Engine(path, gs.C128) as e2:
            lp2 = e2.prob(ob)
        res["max_abs_dlog2p_vs_c128"] = float(np.max(np.abs(lp2 - lps)))
    except gs.GsError as ex:
        res["c128_error"] = str(ex)
    return res


IBM_PLAN = os.path.join("tests", "golden", "ibm", "nq70_inplace_plan.txt")
IBM_PUBLISHED_S = 11.3      # README.md:24: 1,012 samples in 190 min, gsamp2 sample2, Ryzen 9 9955HX laptop


def ibm_w30(gs, egs, K, args):
    """K exact samples (seeds 2000000..) of the reference's width-30 in-place plan of the IBM #151 circuit."""
    path = os.path.join(ROOT, IBM_PLAN)
    _, R, steps = egs.planio.read_inplace_plan(path)
    traffic = egs.planio.plan_traffic(steps, 8)
    with gs.Engine(path, gs.C64) as e:
        if args.no_fusion:
            e.set_fusion(False)
        if args.no_fused4:
            e.set_fused4(False)
        if args.grid_log2:
            e.set_grid_log2(args.grid_log2)
        if args.small_steps is not None:
            e.set_small_steps(args.small_steps)
        e.sample(1, 7)
        t0 = time.perf_counter()
        bits, lp = e.sample(K, 2000000)
        wall = (time.perf_counter() - t0) / K
        if args.per_step:
            e.profile(True)
            e.sample(K, 2000000)
            ms, nl = e.profile_read()
            e.profile(False)
            log(per_step_table(steps, ms, nl, [e.geometry(k) for k in range(e.ns)],
                               f"per-step sweeps, {K} samples of the IBM width-30 plan (HIP events)"))
    return {"plan": IBM_PLAN, "R": R, "samples": K, "s_per_sample": wall,
            "gb_per_sample_algorithmic": traffic["sample"] / 1e9,
            "hbm_gbps_algorithmic": traffic["sample"] / wall / 1e9,
            "reference_published_s_per_sample": IBM_PUBLISHED_S,
            "reference_published_hardware": "laptop, AMD Ryzen 9 9955HX (README.md:24)",
            "speedup_vs_published": IBM_PUBLISHED_S / wall,
            "first_sample": [egs.planio.format_bits(bits[0]), float(lp[0])]}


def ibm_two_table(gs, egs, K, grid_log2=0):
    """The two-table engine on the reference's committed two-table plans of the IBM circuit: exact samples of the
    width-31 plan (published 13.0 s per sample, README.md:24, 2 x 2^31 complex64 tables) and prob of the two gate
    strings on the width-30 plan (published 6.9 s per string, README.md:138-140)."""
    d = os.path.join(ROOT, "tests", "golden", "ibm")
    out = {}
    with gs.TwoTableEngine(os.path.join(d, "nq70_w31_plan.txt")) as e:
        if grid_log2:
            e.set_grid_log2(grid_log2)
        e.sample(1, 7)
        t0 = time.perf_counter()
        e.sample(K, 1000)
        ts = (time.perf_counter() - t0) / K
    out["w31_sample"] = {"s_per_sample": ts, "samples": K, "reference_published_s": 13.0,
                         "speedup_vs_published": 13.0 / ts}
    with gs.TwoTableEngine(os.path.join(d, "nq70_w30_plan.txt")) as e:
        if grid_log2:
            e.set_grid_log2(grid_log2)
        strings = egs.planio.read_strings(os.path.join(d, "gate_strings.txt"), e.ns)
        e.prob(strings[:1])
        t0 = time.perf_counter()
        e.prob(strings)
        tp = (time.perf_counter() - t0) / len(strings)
        t0 = time.perf_counter()
        e.chain(strings)
        tc = (time.perf_counter() - t0) / len(strings)
    out["w30_prob"] = {"s_per_string": tp, "reference_published_s": 6.9, "speedup_vs_published": 6.9 / tp}
    out["w30_chain"] = {"s_per_string": tc}
    return out


TRAFFIC_PROFILE = "profiles/sweep_c2_hbm_traffic.md"


def profiled_traffic(kname):
    """HBM bytes per dispatch of kernel kname (read + write) from the committed rocprofv3 counter summary, or None."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4 and cells[0].strip("`") == kname:
                return (float(cells[2]) + float(cells[3])) * 1e6
    return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        return ""


def _ref_sample_s(exe, plan, seed, reps, env, n2=2):
    """Seconds per exact sample of the reference executor, excluding plan parsing, arena allocation and first touch:
    (wall time of ``sample n2`` - wall time of ``sample 1``) / (n2 - 1), both from ``seed``, median of ``reps``
    pairs (n2 > 2 for plans whose samples take less than the process start-up)."""
    diffs = []
    for _ in range(reps):
        t = []
        for n in (1, n2):
            t0 = time.perf_counter()
            subprocess.run([exe, plan, "sample", str(n), str(seed)], capture_output=True, text=True, env=env,
                           check=True, timeout=600)
            t.append(time.perf_counter() - t0)
        diffs.append((t[1] - t[0]) / (n2 - 1))
    return float(np.median(diffs)), [float(x) for x in diffs]


def _port_sample_s(exe, plan, seed, n, env, c128):
    """Seconds per exact sample of this project's CPU restatement (oracle/gs_oracle), from its own per-line timing
    (GS_ORACLE_TIMING: compute only, after parsing and allocation), median over n samples."""
    cmd = [exe] + (["--c128"] if c128 else []) + [plan, "sample", str(n), str(seed)]
    res = subprocess.run(cmd, capture_output=True, text=True, env=dict(env, GS_ORACLE_TIMING="1"), check=True,
                         timeout=600)
    t = [float(l.split()[1]) for l in res.stderr.splitlines() if l.startswith("time ")]
    return float(np.median(t)), t


def cpu_baseline(plan_path, seed, reps):
    """The same workload on the host cores.  Headline: the reference executor compiled from its own sources
    (oracle/_ref/gsamp2, complex64 arena), seconds per exact sample excluding parse, allocation and first touch,
    median of ``reps`` pairs.  Detail: the same for its complex128 build and for the n = 24 config-1 plan, and this
    project's restatement (oracle/gs_oracle) in both storages, with the restatement/reference ratios."""
    ref64 = os.path.join(ROOT, "oracle", "_ref", "gsamp2")
    ref128 = os.path.join(ROOT, "oracle", "_ref", "gsamp2_c128")
    port = os.path.join(ROOT, "oracle", "gs_oracle")
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    c1 = os.path.join(ROOT, "tests", "golden", "c1", "plan.txt")
    detail = {"cores": cores, "cpu": _cpu_model(), "method": "reference: wall(sample 2) - wall(sample 1) (n = 70 "
              "plan) or (wall(sample 2049) - wall(sample 1)) / 2048 (n = 24 plan), median of "
              f"{reps} pairs; port: its per-sample compute time, median of 3 samples"}
    head = None
    try:
        for tag, plan, n2 in (("c2", plan_path, 2), ("c1", c1, 2049)):
            for storage, exe in (("c64", ref64), ("c128", ref128)):
                if os.path.exists(exe):
                    v, all_ = _ref_sample_s(exe, plan, seed, reps, env, n2)
                    detail[f"reference_{tag}_{storage}_s"] = v
                    detail[f"reference_{tag}_{storage}_pairs_s"] = all_
                if os.path.exists(port):
                    v, all_ = _port_sample_s(port, plan, seed, 3, env, storage == "c128")
                    detail[f"port_{tag}_{storage}_s"] = v
                    if f"reference_{tag}_{storage}_s" in detail:
                        detail[f"port_over_reference_{tag}_{storage}"] = v / detail[f"reference_{tag}_{storage}_s"]
                log(f"cpu baseline {tag} {storage}: " + ", ".join(f"{k}={v}" for k, v in detail.items()
                                                                    if k.endswith(f"{tag}_{storage}_s")))
    except (subprocess.SubprocessError, OSError) as ex:
        detail["error"] = str(ex)
    if "reference_c2_c64_s" in detail:
        head = {"value": detail["reference_c2_c64_s"], "unit": "s/sample", "cores": cores, "kind": "reference",
                "sample": f"exact samples of the same plan (seed {seed}) by the reference executor built from its "
                          f"sources, complex64 arena: wall(sample 2) - wall(sample 1), median of {reps} pairs "
                          "(excludes plan parsing, allocation and first touch)",
                "cpu": detail["cpu"]}
    elif "port_c2_c64_s" in detail:
        head = {"value": detail["port_c2_c64_s"], "unit": "s/sample", "cores": cores, "kind": "port",
                "sample": "3 exact samples of the same plan by oracle/gs_oracle, per-sample compute time, median",
                "cpu": detail["cpu"]}
    return head, detail


if __name__ == "__main__":
    main()
