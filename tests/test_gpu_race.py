"""Device race check (gs_race_check): the reference executor asserts in-place race freedom while it runs
(inplace.py:126-137: no slot read by two cosets, no slot written twice, no write onto another coset's input).  The
device check applies the same asserts to the layouts the kernels actually walk -- the plain sweeps' relabelled coset
coordinates, the fused sweeps' pairing layouts, the runs of diagonal steps -- using the arena as an owner map; a plan
with two equal kept columns is the negative control."""
import os
import time

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu


def corrupt_plan(text, k):
    """The plan text with step k's second kept column replaced by its first (two cosets share every slot)."""
    toks = text.split()
    ns = int(toks[1])
    pos = 3
    for j in range(ns):
        nk, nt = int(toks[pos + 1]), int(toks[pos + 5])
        if j == k:
            assert nk >= 2
            toks[pos + 7] = toks[pos + 6]
        pos += 6 + nk + 4 * nt + 8
    return " ".join(toks) + "\n"


@pytest.mark.parametrize("case", ["c1", "c4", "c2", "g48"])
def test_race_free_layouts(golden, tmp_path, case):
    """Every step of the plan, every layout its kernels use (plain, fused, runs of diagonal steps on the gadget
    instance): no slot read twice, written twice or written over another coset's input; every slot of each step
    checked."""
    if case == "g48":
        inst = egs.instances.Instance(48, 18, 0, gadgets_count=24)
        path = str(tmp_path / "g48.txt")
        inst.write_plan(path)
    else:
        path = os.path.join(golden, case, "plan.txt")
    _, R, steps = egs.planio.read_inplace_plan(path)
    print(f"race check {case} (R={R})", flush=True)
    with gs.Engine(path, gs.C64) as e:
        if case == "g48":
            e.set_diag_runs(8, 10)
            assert e.run_of(26) is not None
        tot = e.race_check_all()
        for k in (0, len(steps) // 2, len(steps) - 1):
            r = e.race_check(k, "logical")
            assert r["read_twice"] == r["written_twice"] == r["clobbered"] == 0, (k, r)
    want = sum((1 << s["nk"]) * len({t[0] for t in s["terms"]}) for s in steps)
    assert tot["checked"] >= want, (tot, want)


def test_race_check_catches_a_corrupt_layout(golden):
    """Negative control: step 12 of the config-1 plan with its second kept column equal to its first -- two cosets
    share every slot, and the check counts the collisions in the plan's own and the relabelled layouts."""
    text = open(os.path.join(golden, "c1", "plan.txt")).read()
    bad = corrupt_plan(text, 12)
    print("race check negative control", flush=True)
    with gs.Engine(gs.Plan(text=bad), gs.C64) as e:
        for lay in ("plain", "logical"):
            r = e.race_check(12, lay)
            assert r["read_twice"] > 0 and r["written_twice"] > 0, (lay, r)
        assert e.race_check(11, "plain")["read_twice"] == 0


def test_race_free_width34_runs(tmp_path):
    """The width-34 instance's runs of diagonal steps (2^32-2^34 cosets): every slot of each run's passes visited
    once and inside its step's table."""
    inst = egs.instances.config_instance("c5")
    path = str(tmp_path / "c5.txt")
    inst.write_plan(path)
    print("race check c5 runs", flush=True)
    with gs.Engine(path, gs.C64) as e:
        starts = sorted({r[0] for r in (e.run_of(k) for k in range(e.ns)) if r is not None})
        assert len(starts) >= 4
        for k in starts:
            t0 = time.perf_counter()
            r = e.race_check(k, "run")
            print(f"  run at step {k}: {r} ({time.perf_counter() - t0:.2f} s)", flush=True)
            assert r["written_twice"] == 0 and r["clobbered"] == 0 and r["checked"] > 0, (k, r)
