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
                   in c128 with prob mode).
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
SWEEP_NAMES = ("gram", "direct", "write", "write+gram", "write+direct")
ROUTES = ("m1", "split", "mixA", "mixB")


def sweep_bytes(step, sweep, esz):
    """Algorithmic bytes of one sweep launch on a step: norm sweeps read the live slots; write sweeps (also the ones
    that fold in the next step's norms, which read nothing more) read the live slots and write the outputs."""
    ntq = len({t[0] for t in step["terms"]})
    read = (1 << step["nk"]) * ntq * esz
    if sweep >= 2:
        return read + (1 << step["nk"]) * (2 if step["m"] == 2 else 1) * esz
    return read


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--c128-steps", type=int, default=8)
    ap.add_argument("--no-fusion", action="store_true", help="run every norm pass as its own sweep")
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
    inst = egs.instances.config_instance("c2")
    tmp = tempfile.mkdtemp(prefix="gs_bench_")
    plan_path = os.path.join(tmp, "c2_plan.txt")
    inst.write_plan(plan_path)
    _, R, steps = egs.planio.read_inplace_plan(plan_path)
    traffic = egs.planio.plan_traffic(steps, 8)
    print(f"workload: n=70 band 30 seed 2, width {inst.width}, R {R}, log2 sum 2^r = {inst.log2_sum_live():.3f}, "
          f"{traffic['sample'] / 1e9:.2f} GB algorithmic per sample (c64)", file=sys.stderr, flush=True)

    K, W = args.steps, args.warmup
    with gs.Engine(plan_path, gs.C64) as e:
        if args.no_fusion:
            e.set_fusion(False)
        if W:
            e.sample(W, args.seed + 10 ** 6)
        e.profile(True)
        # engine calls are synchronous: each returns after its stream has drained, so the wall clock brackets
        # exactly the K samples' device work (torch is not loaded: it bundles a second HIP runtime)
        t0 = time.perf_counter()
        bits, lp = e.sample(K, args.seed)
        wall = time.perf_counter() - t0
        dev_ms = e.last_device_ms
        ms, nl = e.profile_read()
        e.profile(False)
        geo = [e.geometry(k) for k in range(e.ns)]
    s_per_sample = wall / K

    # dominant kernel: aggregate per template instance <storage, VR, VG, ntq, route, sweep, next route>
    agg = {}
    for k, s in enumerate(steps):
        gk = geo[k]
        for sw in range(5):
            if nl[k, sw] == 0:
                continue
            if sw >= 3:
                key = (1 << 1, 2, gk["ntq"], gk["route"], sw, gk["route2"])   # fused layout: runs of 2 x 2
            else:
                key = (1 << gk["lr"], 1, gk["ntq"], gk["route"], sw, 0)
            a = agg.setdefault(key, dict(ms=0.0, n=0, bytes=0.0, steps=[]))
            a["ms"] += ms[k, sw]
            a["n"] += int(nl[k, sw])
            a["bytes"] += sweep_bytes(s, sw, 8) * nl[k, sw]
            a["steps"].append(k)
    key, dom = max(agg.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom["bytes"] / (dom["ms"] / 1e3) / 1e9
    vr, vg, ntq, route, sw, route2 = key
    kname = f"gs::sweep_kernel<float2, {vr}, {vg}, {ntq}, {route}, {sw}, {route2}>"
    kdesc = (f"{vr * vg} cosets/thread ({vg} run(s) of {vr}), {ntq} slots/coset, route {ROUTES[route]}, "
             f"{SWEEP_NAMES[sw]} sweep" + (f", next route {ROUTES[route2]}" if sw >= 3 else ""))
    traffic_meas = profiled_traffic(kname)
    total_sweep_ms = float(ms.sum())
    print(f"dominant kernel {kname} ({kdesc}): {dom['n']} launches, {dom['ms'] / dom['n']:.4f} ms avg, {achieved:.1f} GB/s; "
          f"all sweeps {total_sweep_ms / K:.3f} ms/sample of {1e3 * s_per_sample:.3f} ms wall", file=sys.stderr)

    # complex128 comparison (same seeds) and c64 bits re-scored in c128
    K2 = min(args.c128_steps, K)
    with gs.Engine(plan_path, gs.C128) as e2:
        e2.sample(1, args.seed + 10 ** 6)
        t0 = time.perf_counter()
        bits2, lp2 = e2.sample(K2, args.seed)
        wall2 = (time.perf_counter() - t0) / K2
        rescored = e2.prob(bits[:K2])
    dlog = float(np.max(np.abs(rescored - lp[:K2])))
    same_bits = int(sum(int(np.array_equal(a, b)) for a, b in zip(bits[:K2], bits2)))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(plan_path, args.seed)

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
        "device_ms_per_sample": dev_ms / K,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_meas, "kernel": kname, "kernel_desc": kdesc,
                     "traffic_source": TRAFFIC_PROFILE if traffic_meas else None,
                     "launches": dom["n"], "avg_ms": dom["ms"] / dom["n"],
                     "bytes_per_launch": dom["bytes"] / dom["n"]},
        "c64_vs_c128": {"c128_s_per_sample": wall2, "speedup": wall2 / s_per_sample,
                        "max_abs_dlog2p": dlog, "samples_compared": K2, "identical_bits": same_bits},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


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


def cpu_baseline(plan_path, seed):
    """One exact sample of the same plan on the host cores; reference executor if it was built, else the port."""
    ref = os.path.join(ROOT, "oracle", "_ref", "gsamp2")
    port = os.path.join(ROOT, "oracle", "gs_oracle")
    exe, kind = (ref, "reference") if os.path.exists(ref) else (port, "port")
    if not os.path.exists(exe):
        return None
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores))
    t0 = time.perf_counter()
    res = subprocess.run([exe, plan_path, "sample", "1", str(seed)], capture_output=True, text=True, env=env,
                         timeout=600)
    wall = time.perf_counter() - t0
    if res.returncode:
        return {"error": f"exit {res.returncode}"}
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": wall, "unit": "s/sample", "cores": cores, "kind": kind,
            "sample": "1 exact sample (seed %d) of the same plan, whole process wall time incl. plan parse" % seed,
            "cpu": cpu_model, "line": res.stdout.strip()[:200]}


if __name__ == "__main__":
    main()
