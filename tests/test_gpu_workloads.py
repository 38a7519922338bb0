"""Parity at the measured workloads' full sizes (through the C ABI).

* the reference's own width-34 doped plan (cloud_doped/doped_m20_inplace_plan.txt, 538 vertices, R = 34): IBM's first
  and last doped shots scored unsliced in a complex64 arena (2^34 x 8 B = 128 GiB) and a complex128 arena (256 GiB)
  against the reference's published values (results/doped/run15_c128/README.md:10-13, scored by
  cloud_doped/run_c128.sh:24-29 with --qiskit);
* the n = 70, band-30 sampling instance (R = 30): samples in both storages against the reference executor's (bits
  identical);
* batched scoring of the n = 70, band-26 instance at 512 lanes, and one instance of each (n, band) pair of the mixed
  batch, against the oracle's chain mode;
* the width-34 instance (n = 80, band 34, 40 gadget ancillas): complex64 against complex128 on the same seeds,
  chain against prob, and a two-vertex sliced route with twist against the unsliced prob in complex128.

Only one arena of 2^34 entries exists at a time: each engine is closed before the next is made.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "gs_oracle")


def oracle(plan, *args, c128=True):
    cmd = [ORACLE] + (["--c128"] if c128 else []) + [plan] + [str(a) for a in args]
    out = subprocess.run(cmd, capture_output=True, text=True, check=True, timeout=900).stdout
    return [l.split() for l in out.splitlines() if l.strip()]


def maxdiff(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


# ---------------------------------------------------------------------------------------------------------------
# width 34: the reference's doped plan and its published amplitudes

@pytest.fixture(scope="module")
def doped(golden):
    d = os.path.join(golden, "doped")
    with open(os.path.join(d, "doped_m20_step_vertices.json")) as f:
        sv = json.load(f)
    with open(os.path.join(d, "ibm_doped_shots_values.json")) as f:
        vals = json.load(f)
    shots = [l.strip() for l in open(os.path.join(d, "ibm_doped_shots_first_last.txt")) if l.strip()]
    return os.path.join(d, "doped_m20_inplace_plan.txt"), sv, vals, shots


@pytest.mark.parametrize("storage,tol", [(gs.C64, 3e-6), (gs.C128, 1e-8)])
def test_doped_width34_published_shots(doped, storage, tol):
    """log2 Pr_doped of IBM's first and last doped shots on the unsliced R = 34 plan: within 1e-8 of the reference's
    complex128 values in a complex128 arena, within 3e-6 of them in a complex64 arena (the reference's own complex64
    values sit 2.5e-7 and 1.4e-6 away).  The plan's one-slot steps of 2^30 cosets or more run at 2^14 blocks."""
    plan, sv, vals, shots = doped
    print(f"doped R=34 storage={'c128' if storage == gs.C128 else 'c64'}: 2 published shots", flush=True)
    with gs.Engine(plan, storage) as e:
        geo = [e.geometry(k) for k in range(e.ns)]
        _, _, steps = egs.planio.read_inplace_plan(plan)
        big = [k for k, s in enumerate(steps) if s["nk"] >= 30 and geo[k]["ntq"] == 1]
        assert big and all(geo[k]["grid"] == 1 << 14 for k in big)
        ds = egs.doped.DopedSampler(e, sv, vals["nd"], vals["na"])
        got = ds.score(shots, qiskit=True)
    print(f"  got {got}, published c128 {vals['c128']}, c64 {vals['c64']}", flush=True)
    assert maxdiff(got, vals["c128"]) <= tol
    if storage == gs.C64:
        assert maxdiff(got, vals["c64"]) <= tol


# ---------------------------------------------------------------------------------------------------------------
# width 30: the sampling instance

@pytest.mark.parametrize("storage,tol", [(gs.C64, 1e-5), (gs.C128, 1e-12)])
def test_c2_samples_vs_reference(golden, storage, tol):
    """3 exact samples (seed 4242) of the n = 70, band-30 instance in a 2^30 arena: bits identical to the reference
    executor's in the same storage (tests/golden/c2), log2 Pr within 1e-12 (complex128) / 1e-5 (complex64); prob of
    the sampled strings agrees."""
    d = os.path.join(golden, "c2")
    tag = "c128" if storage == gs.C128 else "c64"
    ref = [l.split() for l in open(os.path.join(d, f"sample_{tag}.txt")) if l.strip()]
    print(f"c2 sample 3 4242 storage={tag}", flush=True)
    with gs.Engine(os.path.join(d, "plan.txt"), storage) as e:
        bits, lp = e.sample(3, 4242)
        lpp = e.prob(bits)
    assert [egs.planio.format_bits(b) for b in bits] == [r[0] for r in ref]
    assert maxdiff(lp, [float(r[1]) for r in ref]) <= tol
    assert maxdiff(lpp, lp) <= tol


# ---------------------------------------------------------------------------------------------------------------
# batched scoring

def test_c3_batched_512_lanes_vs_oracle(tmp_path):
    """The n = 70, band-26 instance, 512 lanes in one batched launch sequence (half uniform strings, half exact
    samples): log2 Pr and conditionals of 16 of the lanes against the oracle's chain mode (complex128) within 1e-5;
    every lane's log2 Pr against the single-string engine's chain mode within 1e-5."""
    inst = egs.instances.config_instance("c3")
    path = str(tmp_path / "c3.txt")
    inst.write_plan(path)
    ns = len(inst.steps)
    B = 512
    bits = np.random.default_rng(3).integers(0, 2, size=(B, ns)).astype(np.int8)
    print("c3 n=70 band 26, 512 lanes", flush=True)
    with gs.Engine(path, gs.C64) as e:
        bits[B // 2:], _ = e.sample(B // 2, 33)
        lp1, cond1, _ = e.chain(bits)
    with gs.BatchEngine([path], [B], storage=gs.C64) as bt:
        egs.autotune.tune_tiles(bt, [bits])
        bt.use_graph(True)
        lp, cond, amp = bt.score([bits])
    pick = list(range(0, B // 2, 32)) + list(range(B // 2, B, 32))
    egs.planio.write_strings(str(tmp_path / "pick.txt"), bits[pick])
    ref = oracle(path, "chain", str(tmp_path / "pick.txt"))
    assert maxdiff(lp[0][pick], [float(r[0]) for r in ref]) <= 1e-5
    assert maxdiff(cond[0][pick], [[float(x) for x in r[1:]] for r in ref]) <= 1e-5
    assert maxdiff(lp[0], lp1) <= 1e-5
    assert maxdiff(np.log2(np.abs(amp[0]) ** 2) - ns, lp[0]) <= 1e-5


def test_c3_mixed_instances_vs_oracle(tmp_path):
    """One instance of each (n, band) pair of the mixed batch ((40, 16), (56, 22), (70, 26); the mix's seeds 300,
    332, 364), 64 lanes each, scored together: 8 lanes of each against the oracle's chain mode within 1e-5."""
    plans, bits = [], []
    for j, (n, band) in ((0, (40, 16)), (32, (56, 22)), (64, (70, 26))):
        mi = egs.instances.Instance(n, band, 300 + j)
        p = str(tmp_path / f"m{j}.txt")
        mi.write_plan(p)
        plans.append(p)
        bits.append(np.random.default_rng(1000 + j).integers(0, 2, size=(64, len(mi.steps))).astype(np.int8))
    print("c3 mixed: 3 instances x 64 lanes", flush=True)
    with gs.BatchEngine(plans, [64] * 3, storage=gs.C64) as bt:
        bt.use_graph(True)
        lp, cond, _ = bt.score(bits)
    for i, p in enumerate(plans):
        egs.planio.write_strings(str(tmp_path / f"s{i}.txt"), bits[i][::8])
        ref = oracle(p, "chain", str(tmp_path / f"s{i}.txt"))
        assert maxdiff(lp[i][::8], [float(r[0]) for r in ref]) <= 1e-5, i
        assert maxdiff(cond[i][::8], [[float(x) for x in r[1:]] for r in ref]) <= 1e-5, i


# ---------------------------------------------------------------------------------------------------------------
# width 34: the gadget instance

@pytest.fixture(scope="module")
def c5(tmp_path_factory):
    d = tmp_path_factory.mktemp("c5")
    inst = egs.instances.config_instance("c5")
    path = str(d / "plan.txt")
    inst.write_plan(path)
    return d, inst, path


def test_c5_width34_c64_vs_c128(c5):
    """n = 80, band 34 + 40 gadget ancillas, R = 34.  Two hybrid samples (ancillas forced 0, data sampled) in a
    complex64 arena and in a complex128 arena from the same seeds: identical bits, |d log2 Pr| <= 1e-5.  In the
    complex64 arena: the product of the chain conditionals is the chain's log2 Pr, chain equals prob (1e-5), and the
    hybrid log2 Pr of each sample equals its prob."""
    d, inst, path = c5
    ns, na = inst.N, inst.N - inst.n
    hy = np.full((2, ns), -1, np.int8)
    hy[:, :na] = 0
    rnd = np.random.default_rng(55).integers(0, 2, size=(2, ns)).astype(np.int8)
    rnd[:, :na] = 0
    print("c5 R=34 c64: 2 hybrid samples, chain and prob of 4 strings", flush=True)
    with gs.Engine(path, gs.C64) as e:
        ob, lph, _ = e.hybrid(hy, 5)
        strings = np.concatenate([ob, rnd])
        lpc, cond, _ = e.chain(strings)
        lpp = e.prob(strings)
    assert maxdiff(np.sum(np.log2(cond), axis=1), lpc) <= 1e-6
    assert maxdiff(lpc, lpp) <= 1e-5
    assert maxdiff(lph, lpp[:2]) <= 1e-5
    print("c5 R=34 c128: the same 2 hybrid samples, prob of the 2 random strings", flush=True)
    with gs.Engine(path, gs.C128) as e:
        ob2, lph2, _ = e.hybrid(hy, 5)
        lpp2 = e.prob(rnd)
    assert np.array_equal(ob2, ob)
    assert maxdiff(lph2, lph) <= 1e-5
    assert maxdiff(lpp2, lpp[2:]) <= 1e-5


def test_c5_width34_sliced_twist_vs_unsliced(c5):
    """Two data vertices of the width-34 instance (75 and 76, whose neighbours include the last three data steps)
    sliced out: the four slice plans' step-K tables (K = ns - 3, gs_prefix in a complex128 arena) summed with the
    twist by gs_resume give the unsliced log2 Pr (gs_prob, complex128) within 1e-10."""
    d, inst, path = c5
    ns, na = inst.N, inst.N - inst.n
    K = ns - 3
    info = egs.sliced.make_sliced_plans(inst.am.A, inst.am.M, inst.order, [75, 76], str(d / "sl"), K=K)
    assert info["twist"] and info["K"] == K
    b = np.random.default_rng(56).integers(0, 2, size=ns).astype(np.int8)
    b[:na] = 0
    print(f"c5 R=34 c128 sliced [75, 76] K={K}: unsliced prob", flush=True)
    with gs.Engine(path, gs.C128) as e:
        want = float(e.prob(b[None, :])[0])
    tabs, lss = [], []
    for p in info["plans"]:
        print(f"  prefix of {os.path.basename(p)}", flush=True)
        with gs.Engine(p, gs.C128) as e:
            tab, nk, ntq, ls, _ = e.prefix(b, K)
        tabs.append(tab)
        lss.append(ls)
    cs = [c for c, _ in info["twists"]]
    ts = [t for _, t in info["twists"]]
    with gs.Engine(info["plans"][0], gs.C128) as e:
        _, lpf, _, _ = e.resume(b, K, tabs, lss, 5, 0, cs, ts)
    print(f"  sliced {lpf:.12f} unsliced {want:.12f}", flush=True)
    assert abs(lpf - want) <= 1e-10


def test_tw70_twist_gate(golden, tmp_path):
    """The reference's twist gate at 70 qubits (cloud_nq80/run_hybrid.sh:32-43): IBM #151 on its width-30 plan, the 3
    proposals of cloud_nq80/tw70_props.txt (20 sampled trailing bits) drawn by unsliced hybrid (seed 17) and by the
    two slice plans of vertex 21 (prefix to K = 50 on each, resume with the twist file): identical bits, log2 Pr
    within 1e-5 (the gate's bound) in complex64 and within 1e-10 in complex128.  The slice plans and twist are the
    ones this package writes, byte-identical to the reference's committed files (tests/test_sliced.py)."""
    d = os.path.join(golden, "sliced")
    ibm = os.path.join(golden, "ibm")
    meta = json.load(open(os.path.join(d, "reference_sliced_plans.json")))["tw70"]
    n, T, tail = egs.tableau.clifford_body(os.path.join(ibm, "nq70.qasm"))
    order = json.load(open(os.path.join(ibm, "order_w30.json")))
    am = egs.amp.Amplitudes(n, T, tail, order)
    info = egs.sliced.make_sliced_plans(am.A, am.M, order, meta["slices"], str(tmp_path / "tw70"), K=meta["K"])
    K, seed = meta["K"], meta["hybrid_seed"]
    props = egs.planio.read_strings(os.path.join(d, "tw70_props.txt"), 70, allow_q=True)
    cs = [c for c, _ in info["twists"]]
    ts = [t for _, t in info["twists"]]
    for storage, tol in ((gs.C64, 1e-5), (gs.C128, 1e-10)):
        print(f"tw70 twist gate storage={'c128' if storage == gs.C128 else 'c64'}: unsliced hybrid", flush=True)
        with gs.Engine(os.path.join(ibm, "nq70_inplace_plan.txt"), storage) as e:
            hb, hlp, _ = e.hybrid(props, seed)
        tabs = [[None] * len(info["plans"]) for _ in props]
        lss = [[None] * len(info["plans"]) for _ in props]
        for j, p in enumerate(info["plans"]):
            with gs.Engine(p, storage) as e:
                for i, row in enumerate(props):
                    tabs[i][j], _, _, lss[i][j], _ = e.prefix(row, K)
        print("  twisted two-slice resume", flush=True)
        with gs.Engine(info["plans"][0], storage) as e:
            for i, row in enumerate(props):
                rb, lpf, _, _ = e.resume(row, K, tabs[i], lss[i], seed, i, cs, ts)
                assert np.array_equal(rb, hb[i]), i
                assert abs(lpf - hlp[i]) <= tol, (i, lpf, hlp[i])


def test_c5_width34_diag_runs_vs_steps(c5):
    """The width-34 instance's 30 diagonal steps (runs of 8, 8, 6, 5 and 3 steps of 2^32-2^34 cosets) against the same
    steps one sweep each, complex128 arena: prob of 2 strings within 1e-10; the runs are on by default."""
    d, inst, path = c5
    ns, na = inst.N, inst.N - inst.n
    rnd = np.random.default_rng(57).integers(0, 2, size=(2, ns)).astype(np.int8)
    rnd[:, :na] = 0
    print("c5 R=34 c128: runs of diagonal steps against one sweep per step", flush=True)
    with gs.Engine(path, gs.C128) as e:
        runs = sorted({e.run_of(k) for k in range(ns)} - {None})
        assert sum(r[1] for r in runs) >= 25, runs
        lp_runs = e.prob(rnd)
        e.set_diag_runs(0)
        assert all(e.run_of(k) is None for k in range(ns))
        lp_steps = e.prob(rnd)
    print(f"  runs {runs}; max |d| {maxdiff(lp_runs, lp_steps):.2e}", flush=True)
    assert maxdiff(lp_runs, lp_steps) <= 1e-10
