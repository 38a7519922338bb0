"""Batched chain-mode scoring (many strings, many plans, one launch per step) against the reference's chain outputs
and the single-string engine; hipGraph replay and tile choices must not change results."""
import json
import os

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu


def lines(path):
    with open(path) as f:
        return [l.split() for l in f if l.strip()]


def test_batch_c1_matches_reference_chain(golden):
    d = os.path.join(golden, "c1")
    meta = json.load(open(os.path.join(d, "meta.json")))
    strings = egs.planio.read_strings(os.path.join(d, "strings.txt"), meta["ns"])
    ref = lines(os.path.join(d, "chain_c64.txt"))
    print("batch c1 x64", flush=True)
    with gs.BatchEngine([os.path.join(d, "plan.txt")], [64]) as bt:
        lp, cond, amp = bt.score([strings])
    assert np.max(np.abs(lp[0] - np.array([float(l[0]) for l in ref]))) < 1e-5
    assert np.max(np.abs(cond[0] - np.array([[float(x) for x in l[1:]] for l in ref]))) < 1e-5
    assert np.max(np.abs(np.log2(np.abs(amp[0]) ** 2) - meta["ns"] - lp[0])) < 1e-5


def test_batch_mixed_instances_match_single_engine(golden):
    """c1 plus the 12 small circuits in one batch (uneven ns and widths), 64 random strings each, against the
    single-string engine's chain mode; graph replay gives bitwise the same numbers."""
    d = os.path.join(golden, "small")
    cases = json.load(open(os.path.join(d, "cases.json")))
    plans = [os.path.join(golden, "c1", "plan.txt")] + [os.path.join(d, f"s{c['i']}_plan.txt") for c in cases]
    rng = np.random.default_rng(4)
    pl = [gs.Plan(path=p) for p in plans]
    bits = [rng.integers(0, 2, size=(64, p.ns)).astype(np.int8) for p in pl]
    print("batch mixed", len(plans), flush=True)
    with gs.BatchEngine(pl, [64] * len(pl)) as bt:
        lp, cond, amp = bt.score(bits)
        bt.use_graph(True)
        lp2, cond2, amp2 = bt.score(bits)
        lp3, cond3, amp3 = bt.score(bits)
    for i, p in enumerate(pl):
        assert np.array_equal(lp[i], lp2[i]) and np.array_equal(cond[i], cond2[i]) and np.array_equal(amp[i], amp3[i])
        with gs.Engine(p, gs.C64) as e:
            lpe, conde, ampe = e.chain(bits[i])
        live = lpe > -40
        assert np.max(np.abs(lp[i][live] - lpe[live])) < 1e-5, i
        assert np.max(np.abs(cond[i] - conde)) < 1e-5, i
        assert np.allclose(amp[i], ampe, rtol=1e-5, atol=1e-9), i


def test_batch_tiles_do_not_change_results(tmp_path):
    """n = 60 band 24 instance (width 20: 2^20 x 64 lanes = 537 MB arena): tiles, unroll depths and the order of
    the work lists (round-robin, or consecutive per XCD in group-major or coset-major order) change only the
    summation order of the per-lane norms."""
    inst = egs.instances.config_instance("c4")
    plan = str(tmp_path / "c4.txt")
    inst.write_plan(plan)
    rng = np.random.default_rng(8)
    with gs.BatchEngine([plan], [64]) as bt:
        bits = [rng.integers(0, 2, size=(64, len(inst.steps))).astype(np.int8)]
        print("batch c4 x64 default tiles", flush=True)
        lp, cond, amp = bt.score(bits)
        runs = []
        for lg, lu, xcd in ((14, 2, 0), (10, 0, 1), (12, 1, 2), (12, 3, 1), (9, 1, 2)):
            bt.set_tile(0, lg)
            bt.set_unroll(lu)
            bt.set_xcd_items(xcd)
            print(f"batch c4 x64 tiles 2^{lg} unroll {lu} xcd items {xcd}", flush=True)
            runs.append(bt.score(bits))
    for lp2, cond2, amp2 in runs:
        assert np.max(np.abs(lp[0] - lp2[0])) < 1e-9
        assert np.array_equal(amp[0], amp2[0])      # per-lane arithmetic does not depend on the tile or unroll
