"""Autotuner of the batched engine: work-item tiles per step, and the schedule of instances into launches.

Tiles.  ``gs_batch`` launches one kernel per step index covering every instance; a work item is (instance, lane group
of 64 strings, a Gray-ordered range of 2^lg cosets) and one block takes one item.  Small lg gives many items (the
chip fills even for one lane group) and more partial sums; large lg gives few, long items.  ``tune_tiles`` times the
whole step sequence once per candidate lg (per-step device times from the events the engine records after every
step), keeps for every step index the fastest candidate, applies the choice with ``set_tile`` and stores it in a JSON
cache keyed by the plans' content digests, the lane counts and the storage, so a later process with the same batch
skips the measurement.  The final choice is made on graph replays of the whole batch: the per-step choice against
every uniform tile, each at every unroll depth.

Schedule.  A mixed set of instances may not fit the device at once (an instance of width w with B lanes holds
2^w * B entries).  ``schedules`` proposes partitions into batches under a memory budget (one batch per instance;
first-fit decreasing by arena size; size classes kept together), ``tune_schedule`` runs each once and keeps the one
with the best throughput, reporting per-instance latency (the end of the instance's last step, counted from the start
of the whole schedule).
"""
import json
import os
import time

import numpy as np

CANDIDATES = (8, 10, 12, 14)
UNROLLS = (0, 1, 2)


def _key(bt):
    return "|".join([str(bt.storage), ",".join(map(str, bt.lanes))] + [p.digest for p in bt.plans])


def _load(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _store(path, table):
    if not path:
        return
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f)
    os.replace(tmp, path)


def step_times(bt, bits):
    """Device ms of every step index of one scoring run (kernel-by-kernel launches, events after each step)."""
    bt.use_graph(False)
    bt.score(bits, want_cond=False, want_amp=False)
    _, ends = bt.last_ms()
    out = np.zeros(len(ends))
    prev = 0.0
    for k, e in enumerate(ends):
        if e >= 0:
            out[k] = e - prev
            prev = e
    return out


def tune_tiles(bt, bits, candidates=CANDIDATES, unrolls=UNROLLS, cache=None, repeats=2):
    """Pick lg per step index (the same for every instance, clamped to each step's coset count) and then the unroll
    depth, apply them and return ``dict(lg=[...], lu=..., cached=bool, ms_default=..., ms_tuned=...)``."""
    table = _load(cache)
    key = _key(bt)
    if key in table:
        ent = table[key]
        _apply(bt, ent["lg"])
        bt.set_unroll(ent["lu"])
        return dict(lg=ent["lg"], lu=ent["lu"], cached=True)
    times = {}
    bt.set_unroll(0)
    for lg in candidates:
        _apply(bt, [lg] * bt.maxns)
        t = np.full(bt.maxns, np.inf)
        for _ in range(repeats):
            t = np.minimum(t, step_times(bt, bits))
        times[lg] = t
    per_step = [int(min(candidates, key=lambda c: times[c][k])) for k in range(bt.maxns)]
    # The per-step choice comes from runs with an event after every step; the batch is then replayed as one graph,
    # where a choice that won a step by noise can lose.  The final choice (tiles and unroll) is the fastest graph
    # replay of the whole step sequence among the per-step choice and every uniform tile.
    plans = {"per_step": per_step}
    for lg in candidates:
        plans[f"uniform_{lg}"] = [int(lg)] * bt.maxns
    bt.use_graph(True)
    tot = {}
    for name, lgs in plans.items():
        _apply(bt, lgs)
        for lu in unrolls:
            bt.set_unroll(lu)
            bt.score(bits, want_cond=False, want_amp=False)          # graph capture
            best_t = np.inf
            for _ in range(repeats):
                t0 = time.perf_counter()
                bt.score(bits, want_cond=False, want_amp=False)
                best_t = min(best_t, time.perf_counter() - t0)
            tot[(name, lu)] = 1e3 * best_t
    name, lu = min(tot, key=tot.get)
    best = plans[name]
    _apply(bt, best)
    bt.set_unroll(lu)
    bt.use_graph(False)
    table[key] = {"lg": best, "lu": int(lu)}
    _store(cache, table)
    return dict(lg=best, lu=int(lu), cached=False, choice=name, ms_default=tot.get(("uniform_12", 0)),
                ms_tuned=tot[(name, lu)])


def _apply(bt, lgs):
    for i, p in enumerate(bt.plans):
        for k in range(p.ns):
            bt.set_tile(i, lgs[k], k)


def arena_bytes(width, lanes, esz):
    return (1 << width) * lanes * esz


def schedules(sizes, budget):
    """Candidate partitions of instances (their arena bytes ``sizes``) into batches of total size <= budget."""
    n = len(sizes)
    out = {"one_per_batch": [[i] for i in range(n)]}
    order = sorted(range(n), key=lambda i: -sizes[i])
    bins = []
    for i in order:                                             # first-fit decreasing
        for b in bins:
            if b[0] + sizes[i] <= budget:
                b[0] += sizes[i]
                b[1].append(i)
                break
        else:
            bins.append([sizes[i], [i]])
    out["first_fit_decreasing"] = [b[1] for b in bins]
    classes = {}
    for i in range(n):                                          # size classes kept together, split by the budget
        classes.setdefault(sizes[i], []).append(i)
    groups = []
    for sz in sorted(classes, reverse=True):
        cur, tot = [], 0
        for i in classes[sz]:
            if cur and tot + sz > budget:
                groups.append(cur)
                cur, tot = [], 0
            cur.append(i)
            tot += sz
        groups.append(cur)
    out["size_classes"] = groups
    return out


def apply_tiles(bt, tiles):
    """Apply a tuned ``dict(lg=[per step index], lu=...)`` to every instance of a batch."""
    _apply(bt, [tiles["lg"][k] if k < len(tiles["lg"]) else tiles["lg"][-1] for k in range(bt.maxns)])
    bt.set_unroll(tiles["lu"])


def run_schedule(gs, plans, lanes, bits, groups, storage, tiles, graph=True, log=None):
    """Score every instance once following ``groups`` (batches run one after another) with the tiles ``tiles``.
    Returns dict(score_s: summed scoring wall time, setup_s: engine creation and graph capture, latency:
    per-instance seconds from the start of the first batch's scoring to the end of the instance's last step, counting
    only scoring time), and log2p per instance."""
    lat = [0.0] * len(plans)
    lps = [None] * len(plans)
    score_s = setup_s = 0.0
    for gi, grp in enumerate(groups):
        t0 = time.perf_counter()
        with gs.BatchEngine([plans[i] for i in grp], [lanes] * len(grp), storage=storage) as bt:
            sub = [bits[i] for i in grp]
            apply_tiles(bt, tiles)
            bt.use_graph(graph)
            bt.score(sub, want_cond=False, want_amp=False)          # warm (graph capture)
            setup_s += time.perf_counter() - t0
            t1 = time.perf_counter()
            lp, _, _ = bt.score(sub, want_cond=False, want_amp=False)
            dt = time.perf_counter() - t1
            bt.use_graph(False)                                       # per-step events for the latencies
            bt.score(sub, want_cond=False, want_amp=False)
            total_ms, ends = bt.last_ms()
            for j, i in enumerate(grp):
                frac = max(ends[plans[i].ns - 1], 0.0) / max(total_ms, 1e-9)
                lat[i] = score_s + min(1.0, frac) * dt
                lps[i] = lp[j]
            score_s += dt
        if log:
            log(f"  batch {gi + 1}/{len(groups)}: {len(grp)} instances, {dt * 1e3:.1f} ms")
    return dict(score_s=score_s, setup_s=setup_s, latency=lat), lps


def tune_schedule(gs, plans, widths, lanes, bits, budget, storage, tile_cache=None, log=None):
    """Tune the tiles once on the largest batch of the first-fit schedule, then run every candidate schedule with
    them; returns (best name, {name: figures}, the best schedule's log2p per instance, tiles)."""
    esz = 8 if storage == gs.C64 else 16
    sizes = [arena_bytes(w, lanes, esz) for w in widths]
    cands = schedules(sizes, budget)
    big = max(cands["first_fit_decreasing"], key=len)
    with gs.BatchEngine([plans[i] for i in big], [lanes] * len(big), storage=storage) as bt:
        tiles = tune_tiles(bt, [bits[i] for i in big], cache=tile_cache)
    if log:
        log(f"tiles tuned on a batch of {len(big)} instances: lu {tiles['lu']}")
    res = {}
    best = None
    for name, groups in cands.items():
        if log:
            log(f"schedule {name}: {len(groups)} batches")
        r, lps = run_schedule(gs, plans, lanes, bits, groups, storage, tiles, log=log)
        lat = r["latency"]
        res[name] = dict(score_s=r["score_s"], setup_s=r["setup_s"], batches=len(groups),
                         amplitudes_per_s=len(plans) * lanes / r["score_s"],
                         p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)))
        if best is None or r["score_s"] < res[best]["score_s"]:
            best, best_lps = name, lps
    return best, res, best_lps, tiles
