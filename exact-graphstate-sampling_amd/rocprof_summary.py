"""Summaries of rocprofv3 output for the committed profiles.

  python rocprof_summary.py stats TRACE_DIR          -> top kernels of a --kernel-trace --stats run (markdown)
  python rocprof_summary.py launches TRACE_DIR REGEX PERIOD [--range A:B] [BYTES ...]
                                                     -> the dispatches of the kernels whose name matches REGEX, in
                                                        start order (only dispatches A .. B-1 of them with --range),
                                                        grouped by position modulo PERIOD (position p: dispatches p,
                                                        p + PERIOD, ...; e.g. the steps of one pass); with BYTES (one
                                                        algorithmic byte count per position), the bandwidth and
                                                        fraction of the 8 TB/s HBM peak per position
  python rocprof_summary.py traffic FETCH_DIR WRITE_DIR MIB
                                                     -> HBM bytes per dispatch per kernel from two --pmc runs
                                                        (FETCH_SIZE, WRITE_SIZE), each scaled by what the counter
                                                        reads for the copy probe of MIB MiB (bench.py --calibrate-copy)
  python rocprof_summary.py gather TRACE_DIR FETCH_DIR WRITE_DIR PLAN MIB SAMPLES
                                                     -> the two-table engine's exact samples of PLAN (a gsamp plan:
                                                        one norm pass and one write pass per step): samples 1 ..
                                                        SAMPLES-1 of the trace (sample 0 is the warm-up), per step and
                                                        pass the kernel time, the algorithmic bytes (each old-table
                                                        entry a pass gathers read once, the writes) and the HBM bytes
                                                        of the counter runs, calibrated as in traffic

Input directories hold the CSV files rocprofv3 writes with ``-f csv -o run``.
"""
import collections
import csv
import os
import re
import sys


def _short(name):
    """Kernel name without its parameter list, with the vector types spelled float2 / double2."""
    n = name.replace("(anonymous namespace)::", "").replace("HIP_vector_type<float, 2u>", "float2") \
        .replace("HIP_vector_type<double, 2u>", "double2").replace("void ", "")
    return n.split(">(")[0] + ">" if ">(" in n else n.split("(")[0]


def stats(trace_dir, top=14):
    rows = list(csv.DictReader(open(os.path.join(trace_dir, "run_kernel_stats.csv"))))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| `{_short(r['Name'])[:100]}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    return "\n".join(out)


def launches(trace_dir, regex, period, nbytes=None, first=0, stop=None):
    rows = [r for r in csv.DictReader(open(os.path.join(trace_dir, "run_kernel_trace.csv")))
            if re.search(regex, r["Kernel_Name"])]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    rows = rows[first:stop]
    pos = collections.defaultdict(list)
    names = {}
    for i, r in enumerate(rows):
        pos[i % period].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        names.setdefault(i % period, _short(r["Kernel_Name"]))
    head = "| position | kernel | dispatches | avg us | min us | max us |"
    out = [head + (" TB/s | of 8 TB/s |" if nbytes else ""), "|---" * (6 + (2 if nbytes else 0)) + "|"]
    for p in range(period):
        d = pos.get(p, [])
        if not d:
            continue
        avg = sum(d) / len(d)
        line = f"| {p} | `{names[p][:80]}` | {len(d)} | {avg:.1f} | {min(d):.1f} | {max(d):.1f} |"
        if nbytes:
            tbs = nbytes[p] / (avg * 1e-6) / 1e12
            line += f" {tbs:.2f} | {tbs / 8:.2f} |"
        out.append(line)
    if len(rows) % period:
        out.append(f"\n({len(rows)} dispatches: not a multiple of {period})")
    return "\n".join(out)


def _counter(path, cname):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == cname:
            d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return d


def traffic(fetch_dir, write_dir, mib):
    f = _counter(os.path.join(fetch_dir, "run_counter_collection.csv"), "FETCH_SIZE")
    w = _counter(os.path.join(write_dir, "run_counter_collection.csv"), "WRITE_SIZE")
    probe = next(k for k in f if "copy_probe" in k)
    kib = mib * 1024.0
    ff = kib / f[probe][0]
    wf = kib / w[probe][0]
    out = [f"calibration (copy probe, {mib} MiB read + {mib} MiB written): FETCH_SIZE reads {f[probe][0]:.0f} KiB -> "
           f"factor {ff:.4f}; WRITE_SIZE reads {w[probe][0]:.0f} KiB -> factor {wf:.4f}",
           "", "| kernel | dispatches | HBM read MB/dispatch | HBM write MB/dispatch |", "|---|---|---|---|"]
    for k in f:
        if k == probe:
            continue
        rd = sum(f[k]) / len(f[k]) * ff * 1024 / 1e6
        wr = sum(w.get(k, [0])) / max(1, len(w.get(k, []))) * wf * 1024 / 1e6
        out.append(f"| `{_short(k)[:100]}` | {len(f[k])} | {rd:.2f} | {wr:.2f} |")
    return "\n".join(out)


def _gather_rows(path, key):
    rows = [r for r in csv.DictReader(open(path)) if "gather_kernel" in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return rows


def gather(trace_dir, fetch_dir, write_dir, plan, mib, samples):
    toks = open(plan).read().split()
    ns, i, pos = int(toks[1]), 3, []
    for k in range(ns):
        r, rn, nk = int(toks[i + 1]), int(toks[i + 2]), int(toks[i + 5])
        i += 6 + nk + 8
        rd = min(1 << r, 1 << (rn + nk)) * 8                  # c64 entries gathered, each read once
        pos += [(k, "norm", r, rn, nk, rd), (k, "write", r, rn, nk, rd + (1 << rn) * 8)]
    P = len(pos)
    tr = _gather_rows(os.path.join(trace_dir, "run_kernel_trace.csv"), "Dispatch_Id")
    fr = _gather_rows(os.path.join(fetch_dir, "run_counter_collection.csv"), "Dispatch_Id")
    wr = _gather_rows(os.path.join(write_dir, "run_counter_collection.csv"), "Dispatch_Id")
    f = _counter(os.path.join(fetch_dir, "run_counter_collection.csv"), "FETCH_SIZE")
    w = _counter(os.path.join(write_dir, "run_counter_collection.csv"), "WRITE_SIZE")
    probe = next(k for k in f if "copy_probe" in k)
    ff, wf = mib * 1024.0 / f[probe][0], mib * 1024.0 / w[probe][0]
    n = samples - 1
    dur, hr, hw = [0.0] * P, [0.0] * P, [0.0] * P
    for s in range(1, samples):
        for p in range(P):
            j = s * P + p
            want = "1" if pos[p][1] == "norm" else "2"           # gather_kernel<S, kDirect = 1 / kWrite = 2, NKS>
            for rows in (tr, fr, wr):
                assert rows[j]["Kernel_Name"].split(",")[2].strip() == want, (j, rows[j]["Kernel_Name"])
            dur[p] += (int(tr[j]["End_Timestamp"]) - int(tr[j]["Start_Timestamp"])) / 1e3 / n
            hr[p] += float(fr[j]["Counter_Value"]) * ff * 1024 / n
            hw[p] += float(wr[j]["Counter_Value"]) * wf * 1024 / n
    T, A, H = sum(dur), sum(x[5] for x in pos), sum(hr) + sum(hw)
    out = [f"per sample ({n} samples): gather kernels {T / 1e3:.1f} ms; algorithmic {A / 1e9:.1f} GB "
           f"({A / T / 1e6:.2f} TB/s); HBM {sum(hr) / 1e9:.1f} GB read + {sum(hw) / 1e9:.1f} GB written "
           f"({H / T / 1e6:.2f} TB/s, {H / A:.3f} x algorithmic)", "",
           "| step | pass | r | rn | nker | us | algorithmic GB | HBM GB read + written | TB/s | of 8 TB/s |",
           "|---|---|---|---|---|---|---|---|---|---|"]
    for p in range(P):
        k, kind, r, rn, nk, b = pos[p]
        if dur[p] < 100:
            continue
        out.append(f"| {k} | {kind} | {r} | {rn} | {nk} | {dur[p]:.0f} | {b / 1e9:.2f} | {hr[p] / 1e9:.2f} + "
                   f"{hw[p] / 1e9:.2f} | {b / dur[p] / 1e6:.2f} | {b / dur[p] / 8e6:.2f} |")
    return "\n".join(out)


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        print(stats(sys.argv[2]))
    elif sys.argv[1] == "launches":
        rest, first, stop = sys.argv[5:], 0, None
        if rest[:1] == ["--range"]:
            a, b = rest[1].split(":")
            first, stop, rest = int(a), int(b), rest[2:]
        print(launches(sys.argv[2], sys.argv[3], int(sys.argv[4]), [float(x) for x in rest] or None, first, stop))
    elif sys.argv[1] == "traffic":
        print(traffic(sys.argv[2], sys.argv[3], int(sys.argv[4])))
    elif sys.argv[1] == "gather":
        print(gather(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], int(sys.argv[6]), int(sys.argv[7])))
    else:
        raise SystemExit(__doc__)
