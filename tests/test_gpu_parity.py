"""Parity of the HIP engine (through the C ABI) with the reference executor's outputs and with brute force.

Golden outputs come from the reference executor compiled from its own sources (tests/golden/make_golden.py);
tolerances: complex128 arena 1e-12 on log2 Pr and conditionals (the FP64 reference path), complex64 arena 1e-5
(the reference's own GPU-vs-CPU bound, test_cuda.py:84; c64 rounding of two independent engines), sampled bits
identical.
"""
import json
import os

import numpy as np
import pytest

from _pkg import egs
import dense

gs = egs.gsamp
pytestmark = pytest.mark.gpu

TOL = {gs.C64: 1e-5, gs.C128: 1e-12}
TAG = {gs.C64: "c64", gs.C128: "c128"}


def lines(path):
    with open(path) as f:
        return [l.split() for l in f if l.strip()]


def close(a, b, tol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    d = np.where(both_inf, 0.0, np.abs(a - b))
    return float(np.max(d)) if d.size else 0.0


PATHS = ["small-runs", "fused", "fused-run4", "plain"]


def configure(e, path):
    """small-runs: the default engine (runs of small steps in one block, folded norms elsewhere); fused: every step
    its own sweeps, norms folded into the previous write; plain: one norm sweep per sampled step."""
    e.set_small_steps(10 if path == "small-runs" else -1)
    e.set_fusion(path != "plain")
    if path == "fused-run4" and e.storage == gs.C64:
        e.set_fused_run(2)


@pytest.fixture(scope="module")
def c1(golden):
    d = os.path.join(golden, "c1")
    with open(os.path.join(d, "meta.json")) as f:
        meta = json.load(f)
    ns = meta["ns"]
    strings = egs.planio.read_strings(os.path.join(d, "strings.txt"), ns)
    hy = egs.planio.read_strings(os.path.join(d, "hybrid.txt"), ns, allow_q=True)
    return d, meta, strings, hy


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_c1_prob_chain(c1, storage, path):
    d, meta, strings, _ = c1
    print(f"c1 prob/chain storage={TAG[storage]} path={path}", flush=True)
    with gs.Engine(os.path.join(d, "plan.txt"), storage) as e:
        configure(e, path)
        lp = e.prob(strings)
        ref = np.array([float(l[0]) for l in lines(os.path.join(d, f"prob_{TAG[storage]}.txt"))])
        assert close(lp, ref, 0) <= TOL[storage], close(lp, ref, 0)
        lpc, cond, amp = e.chain(strings)
        refc = lines(os.path.join(d, f"chain_{TAG[storage]}.txt"))
        assert close(lpc, [float(l[0]) for l in refc], 0) <= TOL[storage]
        assert close(cond, [[float(x) for x in l[1:]] for l in refc], 0) <= max(TOL[storage], 1e-9)  # printed %.9f
        # the amplitude: log2 |Z|^2 - ns is log2 Pr
        assert close(np.log2(np.abs(amp) ** 2) - meta["ns"], lpc, 0) <= 10 * TOL[storage]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
@pytest.mark.parametrize("mode", ["sample", "sample2"])
def test_c1_sample_bits_identical(c1, storage, mode, path):
    """Sampled bits identical to the reference on every execution path: the whole chain in one block (the plan's
    steps are all small), separate sweeps with the next step's norms folded into the previous write, and one norm
    sweep per sampled step."""
    d, meta, _, _ = c1
    print(f"c1 {mode} storage={TAG[storage]} path={path}", flush=True)
    with gs.Engine(os.path.join(d, "plan.txt"), storage) as e:
        configure(e, path)
        if path == "fused":
            assert any(e.geometry(k)["fusable"] for k in range(e.ns))
        bits, lp = e.sample(16, meta["sample_seed"], seeding=gs.SEED_PLAIN if mode == "sample" else gs.SEED_SEQ)
    ref = lines(os.path.join(d, f"{mode}_{TAG[storage]}.txt"))
    assert [egs.planio.format_bits(b) for b in bits] == [l[0] for l in ref]
    assert close(lp, [float(l[1]) for l in ref], 0) <= TOL[storage]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_c1_hybrid(c1, storage, path):
    d, meta, _, hy = c1
    print(f"c1 hybrid storage={TAG[storage]} path={path}", flush=True)
    with gs.Engine(os.path.join(d, "plan.txt"), storage) as e:
        configure(e, path)
        bits, lp, cond = e.hybrid(hy, meta["hybrid_seed"])
    ref = lines(os.path.join(d, f"hybrid_{TAG[storage]}.txt"))
    assert [egs.planio.format_bits(b) for b in bits] == [l[0] for l in ref]
    assert close(lp, [float(l[1]) for l in ref], 0) <= TOL[storage]
    assert close(cond, [[float(x) for x in l[2:]] for l in ref], 0) <= max(TOL[storage], 1e-9)


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
@pytest.mark.parametrize("path", ["small-runs", "fused"])
def test_c1_prefix_resume(c1, storage, path):
    d, meta, _, hy = c1
    K = meta["K"]
    c128 = storage == gs.C128
    print(f"c1 prefix/resume storage={TAG[storage]} path={path}", flush=True)
    ref_pre = lines(os.path.join(d, f"prefix_{TAG[storage]}.txt"))
    ref_res = lines(os.path.join(d, f"resume_{TAG[storage]}.txt"))
    with gs.Engine(os.path.join(d, "plan.txt"), storage) as e:
        configure(e, path)
        for li, row in enumerate(hy):
            tab, nk, ntq, ls, lp = e.prefix(row, K)
            rtab, rnk, rntq, rls = egs.planio.read_dump(os.path.join(d, f"dump_{TAG[storage]}_{li}.bin"), c128)
            assert (nk, ntq) == (rnk, rntq)
            assert abs(ls - rls) <= 1e-6 and abs(lp - float(ref_pre[li][1])) <= TOL[storage]
            scale = np.max(np.abs(rtab))
            assert np.max(np.abs(tab - rtab)) <= TOL[storage] * scale * 10
            # resume from the reference's own dump: bits and probabilities as the reference's resume
            bits, lpf, lpp, cond = e.resume(row, K, [rtab], [rls], seed=meta["resume_seed"], line=li)
            assert egs.planio.format_bits(bits) == ref_res[li][0]
            assert abs(lpf - float(ref_res[li][1])) <= TOL[storage] and abs(lpp - float(ref_res[li][2])) <= TOL[storage]
            assert close(cond[K:], [float(x) for x in ref_res[li][3:]], 0) <= max(TOL[storage], 1e-9)


@pytest.fixture(scope="module")
def small_cases(golden):
    d = os.path.join(golden, "small")
    with open(os.path.join(d, "cases.json")) as f:
        return d, json.load(f)


def _small_ops(path):
    n, ops = egs.tableau.parse(path)
    return n, ops


def test_small_circuits_all_strings_vs_statevector(small_cases):
    """Every string of 12 random circuits: log2 Pr (prob and chain, c128) equals the dense statevector at 1e-12
    relative, and equals the reference executor's outputs.  Negative control: the same plans with the rows of every
    M swapped must disagree."""
    d, cases = small_cases
    worst = worst_ref = ctrl = 0.0
    for c in cases:
        i, n = c["i"], c["n"]
        print(f"small circuit {i} n={n}", flush=True)
        plan = os.path.join(d, f"s{i}_plan.txt")
        _, ops = _small_ops(os.path.join(d, f"s{i}.qasm"))
        p = dense.probabilities(n, ops)
        _, _, steps = egs.planio.read_inplace_plan(plan)
        sv = [s["v"] for s in steps]
        strings = egs.planio.read_strings(os.path.join(d, f"s{i}_strings.txt"), n)
        idx = [dense.step_order_index(b, sv, n) for b in strings]
        with gs.Engine(plan, gs.C128) as e:
            lp = e.prob(strings)
            lpc, cond, _ = e.chain(strings)
        got = np.exp2(lp)
        worst = max(worst, float(np.max(np.abs(got - p[idx]))))
        assert np.allclose(np.exp2(lpc), got, rtol=1e-12, atol=1e-15)
        ref = np.array([float(l[0]) for l in lines(os.path.join(d, f"s{i}_prob_c128.txt"))])
        # log2 Pr is compared where the probability is not a rounding-noise zero (the reference tests do the same,
        # test_chain.py:122-124); zero-probability strings give unrelated tiny values in any two FP64 orders
        live = p[idx] > 1e-10
        worst_ref = max(worst_ref, close(lp[live], ref[live], 0))
        # negative control: M rows swapped
        text = open(plan).read()
        bad = _swap_rows(text)
        with gs.Engine(gs.Plan(text=bad), gs.C128) as e:
            ctrl = max(ctrl, float(np.max(np.abs(np.exp2(e.prob(strings)) - p[idx]))))
    print(f"max |p - p_sv| = {worst:.2e}; max |log2 - reference| = {worst_ref:.2e}; control = {ctrl:.3f}")
    assert worst < 1e-12 and worst_ref < 1e-10 and ctrl > 1e-2


def _swap_rows(text):
    toks = text.split()
    n, ns, R = int(toks[0]), int(toks[1]), int(toks[2])
    pos = 3
    out = toks[:3]
    for _ in range(ns):
        head = toks[pos:pos + 6]
        nk, nt = int(head[1]), int(head[5])
        pos2 = pos + 6 + nk + 4 * nt
        m = toks[pos2:pos2 + 8]
        out += toks[pos:pos2] + m[4:8] + m[0:4]
        pos = pos2 + 8
    return " ".join(out) + "\n"


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_ibm_gate_strings(golden, storage):
    """IBM #151, width-30 in-place plan: the two gate strings of cloud_ibmxeb/run.sh:16-25 within 1e-5."""
    d = os.path.join(golden, "ibm")
    print(f"ibm gate storage={TAG[storage]}", flush=True)
    want = [float(l[0]) for l in lines(os.path.join(d, "gate_values.txt"))]
    with gs.Engine(os.path.join(d, "nq70_inplace_plan.txt"), storage) as e:
        strings = egs.planio.read_strings(os.path.join(d, "gate_strings.txt"), e.ns)
        lp = e.prob(strings)
        lpc, cond, _ = e.chain(strings)
    assert close(lp, want, 0) < 1e-5, lp
    assert close(lpc, want, 0) < 1e-5, lpc
    assert close(np.sum(np.log2(cond), axis=1), lpc, 0) < 1e-6


def test_ibm_laptop_samples_bit_exact(golden):
    """KAT: gsamp2 ... sample 3 2000000 (results/laptop_samples_w30.txt): identical bits, log2 Pr within 1e-5."""
    d = os.path.join(golden, "ibm")
    ref = lines(os.path.join(d, "laptop_samples_w30.txt"))
    print("ibm laptop KAT sample 3 2000000", flush=True)
    with gs.Engine(os.path.join(d, "nq70_inplace_plan.txt"), gs.C64) as e:
        bits, lp = e.sample(3, 2000000)
    assert [egs.planio.format_bits(b) for b in bits] == [l[0] for l in ref]
    assert close(lp, [float(l[1]) for l in ref], 0) < 1e-5


def test_plan_errors_from_engine(golden):
    """Malformed plans give the reference's exit codes."""
    good = open(os.path.join(golden, "c1", "plan.txt")).read()
    toks = good.split()
    with pytest.raises(gs.GsError) as ex:
        gs.Plan(text="")
    assert ex.value.code == 2
    with pytest.raises(gs.GsError) as ex:
        gs.Plan(text=" ".join(toks[:3] + toks[3:5]))
    assert ex.value.code == 3


@pytest.fixture(scope="module")
def c4_plan(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("c4") / "plan.txt")
    egs.instances.config_instance("c4").write_plan(path)
    return path


@pytest.mark.parametrize("grid_log2,walk,run,xcd", [(14, 0, -1, 0), (9, 0, -1, 1), (9, 1, -1, 0), (6, 2, -1, 1),
                                                    (4, 0, -1, 0), (9, 0, 2, 1), (12, 0, -1, 1)])
def test_c4_samples_vs_oracle_any_grid(c4_plan, grid_log2, walk, run, xcd):
    """n=60 band-24 instance (2^20 arena, up to 2^10 blocks per sweep): samples from the FP64 arena equal the
    oracle's bits and log2 Pr (1e-12) whatever the block cap -- one tile per block with a separate reduce kernel,
    512 looping blocks with the reduce kernel, 64 or 16 looping blocks reducing in their last block -- and whatever the
    tile order (forward, reversed on odd steps, reversed everywhere) and whatever the block-to-XCD mapping; a second
    run gives bitwise the same samples."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    import subprocess
    out = subprocess.run([os.path.join(root, "oracle", "gs_oracle"), "--c128", c4_plan, "sample", "3", "777"],
                         capture_output=True, text=True, check=True).stdout.split("\n")
    ref = [l.split() for l in out if l.strip()]
    print(f"c4 sample 3 777 grid_log2={grid_log2} walk={walk} run={run} xcd={xcd}", flush=True)
    storage = gs.C64 if run == 2 else gs.C128        # the 4-run fused layout is a complex64 layout
    with gs.Engine(c4_plan, storage) as e:
        e.set_grid_log2(grid_log2)
        e.set_walk(walk)
        e.set_fused_run(run)
        e.set_xcd_blocks(xcd)
        assert max(e.geometry(k)["grid"] for k in range(e.ns)) <= 1 << grid_log2
        if run == 2:
            assert any(e.geometry(k)["flr"] == 2 for k in range(e.ns))
        bits, lp = e.sample(3, 777)
        lpp = e.prob(bits)
        bits2, lp2 = e.sample(3, 777)                # every coset once per sweep, fixed-order sums: bitwise repeatable
    tol = TOL[storage]
    assert [egs.planio.format_bits(b) for b in bits] == [r[0] for r in ref]
    assert close(lp, [float(r[1]) for r in ref], 0) < tol
    assert close(lpp, lp, 0) < tol
    assert np.array_equal(bits2, bits) and np.array_equal(lp2, lp)


@pytest.mark.parametrize("pair_ntq,far,fusion", [(0, 0, True), (1, 1, True), (1, 0, False), (2, 1, False),
                                                 (2, 0, True)])
def test_c4_plain_pairs_vs_oracle(c4_plan, tmp_path, pair_ntq, far, fusion):
    """n=60 band-24 instance in the FP64 arena: samples, prob and chain equal the oracle (bits identical, 1e-12)
    whatever the plain sweeps' layout -- one run per thread, or a second run at the next unit bit or at the top
    coordinate for one-slot (and two-slot) steps -- with the norm passes folded into the writes or all plain."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    import subprocess
    ora = os.path.join(root, "oracle", "gs_oracle")
    out = subprocess.run([ora, "--c128", c4_plan, "sample", "3", "778"], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    ref = [l.split() for l in out if l.strip()]
    print(f"c4 sample 3 778 plain pairs ntq<={pair_ntq} far={far} fusion={fusion}", flush=True)
    with gs.Engine(c4_plan, gs.C128) as e:
        e.set_plain_pairs(pair_ntq, far)
        e.set_fusion(fusion)
        vgs = [e.geometry(k)["vg"] for k in range(e.ns)]
        ntqs = [e.geometry(k)["ntq"] for k in range(e.ns)]
        if pair_ntq:
            assert any(v == 2 for v in vgs)
        assert all(v == 1 or n <= pair_ntq for v, n in zip(vgs, ntqs))
        bits, lp = e.sample(3, 778)
        lpp = e.prob(bits)
        lpc, cond, _ = e.chain(bits)
    assert [egs.planio.format_bits(b) for b in bits] == [r[0] for r in ref]
    assert close(lp, [float(r[1]) for r in ref], 0) < TOL[gs.C128]
    assert close(lpp, lp, 0) < TOL[gs.C128]
    assert close(lpc, lp, 0) < TOL[gs.C128]
    strings = tmp_path / "strings.txt"
    strings.write_text("\n".join(r[0] for r in ref) + "\n")
    chain_ref = subprocess.run([ora, "--c128", c4_plan, "chain", str(strings)], capture_output=True, text=True,
                               check=True).stdout.split("\n")
    cref = [l.split() for l in chain_ref if l.strip()]
    assert close(lpc, [float(r[0]) for r in cref], 0) < TOL[gs.C128]
    assert close(cond, [[float(x) for x in r[1:]] for r in cref], 0) < 2e-9      # printed with %.9f


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_sliced_prefix_resume_with_twist(tmp_path, storage):
    """Two slice vertices with suffix neighbours on the n=24 band-10 instance: the four slice tables of step K=18
    (gs_prefix on each slice plan), summed by gs_resume with the twist, give the unsliced log2 Pr (gs_prob)."""
    inst = egs.instances.Instance(24, 10, 1)
    full = str(tmp_path / "full.txt")
    inst.write_plan(full)
    info = egs.sliced.make_sliced_plans(inst.am.A, inst.am.M, inst.order, [13, 15], str(tmp_path / "sl"), K=18)
    assert info["twist"]
    bits = np.random.default_rng(7).integers(0, 2, size=(6, 24)).astype(np.int8)
    print(f"sliced [13, 15] K=18 storage={TAG[storage]}", flush=True)
    with gs.Engine(full, storage) as e:
        want = e.prob(bits)
    cs = [c for c, _ in info["twists"]]
    ts = [t for _, t in info["twists"]]
    got = []
    for b in bits:
        tabs, lss = [], []
        for p in info["plans"]:
            with gs.Engine(p, storage) as e:
                tab, nk, ntq, ls, _ = e.prefix(b, 18)
            tabs.append(tab)
            lss.append(ls)
        with gs.Engine(info["plans"][0], storage) as e:
            _, lpf, _, _ = e.resume(b, 18, tabs, lss, 5, 0, cs, ts)
        got.append(lpf)
    assert close(got, want, 0) < TOL[storage] * 10


@pytest.fixture(scope="module")
def gadget_plan(tmp_path_factory):
    """n = 48 data vertices, band 18, 24 phase-gadget ancillas (width 18): the kept columns of 43 steps are GF(2)
    combinations whose span holds the low physical bits without being them, so the engine relabels those steps'
    coset coordinates."""
    d = tmp_path_factory.mktemp("g48")
    inst = egs.instances.Instance(48, 18, 0, gadgets_count=24)
    path = str(d / "plan.txt")
    inst.write_plan(path)
    return path, inst


def test_relabelled_steps_vs_oracle(gadget_plan, tmp_path):
    """Hybrid samples (ancillas forced 0, data sampled), prob of random strings and a prefix dump, FP64 arena, against
    the oracle: bits identical, log2 Pr within 1e-12, the step-K table in the plan's own coset order within 1e-12."""
    import subprocess
    path, inst = gadget_plan
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    oracle = os.path.join(root, "oracle", "gs_oracle")
    ns, na = inst.N, inst.N - inst.n
    hy = np.full((4, ns), -1, np.int8)
    hy[:, :na] = 0
    rnd = np.random.default_rng(3).integers(0, 2, size=(4, ns)).astype(np.int8)
    rnd[:, :na] = 0
    egs.planio.write_strings(str(tmp_path / "hy.txt"), hy)
    egs.planio.write_strings(str(tmp_path / "rnd.txt"), rnd)
    run = lambda *a: [l.split() for l in subprocess.run([oracle, "--c128", path, *map(str, a)], capture_output=True,
                                                        text=True, check=True).stdout.splitlines() if l.strip()]
    ref_h = run("hybrid", str(tmp_path / "hy.txt"), 11)
    ref_p = [float(r[0]) for r in run("prob", str(tmp_path / "rnd.txt"))]
    K = na + 20
    run("prefix", str(tmp_path / "rnd.txt"), K, str(tmp_path / "d"))
    print("gadget n=48 band 18 + 24 ancillas, relabelled steps", flush=True)
    with gs.Engine(path, gs.C128) as e:
        assert sum(e.geometry(k)["lr"] > 0 for k in range(ns)) > 40
        ob, lph, _ = e.hybrid(hy, 11)
        lpp = e.prob(rnd)
        tab, nk, ntq, ls, _ = e.prefix(rnd[0], K)
    assert [egs.planio.format_bits(b) for b in ob] == [r[0] for r in ref_h]
    assert close(lph, [float(r[1]) for r in ref_h], 0) < 1e-12
    assert close(lpp, ref_p, 0) < 1e-12
    dump, dnk, dntq, dls = egs.planio.read_dump(str(tmp_path / "d_0.bin"), c128=True)
    assert (dnk, dntq) == (nk, ntq) and abs(dls - ls) < 1e-12
    assert np.max(np.abs(np.asarray(dump) - tab)) < 1e-12 * max(1.0, np.max(np.abs(tab)))


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_one_slot_folded_norms_vs_oracle(gadget_plan, tmp_path, storage):
    """14 of the gadget instance's data steps (from step 24 on) are one-slot, one-output steps whose slot space is
    the whole table the previous step writes: that write also sums their Gram matrix.  Hybrid samples with those folded norms and with
    every norm pass as its own sweep equal the oracle (bits identical; log2 Pr and conditionals 1e-12 in complex128,
    1e-5 in complex64) and each other."""
    import subprocess
    path, inst = gadget_plan
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ns, na = inst.N, inst.N - inst.n
    hy = np.full((6, ns), -1, np.int8)
    hy[:, :na] = 0
    egs.planio.write_strings(str(tmp_path / "hy.txt"), hy)
    cmd = [os.path.join(root, "oracle", "gs_oracle")] + (["--c128"] if storage == gs.C128 else [])
    ref = [l.split() for l in subprocess.run(cmd + [path, "hybrid", str(tmp_path / "hy.txt"), "23"],
                                             capture_output=True, text=True, check=True).stdout.splitlines() if l.strip()]
    print(f"gadget n=48 band 18 + 24 ancillas, one-slot folded norms, storage={TAG[storage]}", flush=True)
    got = {}
    with gs.Engine(path, storage) as e:
        geo = [e.geometry(k) for k in range(ns)]
        one_slot = [k for k in range(ns - 1) if geo[k]["fusable"] and geo[k + 1]["ntq"] == 1]
        assert len(one_slot) >= 10, one_slot
        for fusion in (True, False):
            e.set_fusion(fusion)
            got[fusion] = e.hybrid(hy, 23)
    for fusion, (ob, lp, cond) in got.items():
        assert [egs.planio.format_bits(b) for b in ob] == [r[0] for r in ref], fusion
        assert close(lp, [float(r[1]) for r in ref], 0) <= TOL[storage]
        assert close(cond, [[float(x) for x in r[2:]] for r in ref], 0) <= max(TOL[storage], 1e-9)
    assert np.array_equal(got[True][0], got[False][0])
    assert close(got[True][1], got[False][1], 0) <= TOL[storage]


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
@pytest.mark.parametrize("grid_log2", [0, 4])
def test_diag_runs_vs_oracle(gadget_plan, tmp_path, storage, grid_log2):
    """Steps 26..37 of the gadget instance are 12 consecutive diagonal steps (one slot, one output, slot offset 0,
    2^18 cosets).  Taken as runs (8 + 4 steps: a pattern-sum pass, the decisions in one block, a product pass): hybrid
    samples, prob and chain of random strings, and a prefix dump at K = 30 (inside a run, so those steps go one by one)
    equal the oracle -- bits identical, complex128 within 1e-12, complex64 within 1e-5 -- with 2^9 blocks and one tile
    each, and with 2^4 blocks looping over 32 tiles (then the second run's product pass also sums the Gram matrix of
    step 38, a two-slot step, and decides it)."""
    import subprocess
    path, inst = gadget_plan
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ns, na = inst.N, inst.N - inst.n
    hy = np.full((4, ns), -1, np.int8)
    hy[:, :na] = 0
    rnd = np.random.default_rng(31).integers(0, 2, size=(4, ns)).astype(np.int8)
    rnd[:, :na] = 0
    egs.planio.write_strings(str(tmp_path / "hy.txt"), hy)
    egs.planio.write_strings(str(tmp_path / "rnd.txt"), rnd)
    cmd = [os.path.join(root, "oracle", "gs_oracle")] + (["--c128"] if storage == gs.C128 else []) + [path]
    run = lambda *a: [l.split() for l in subprocess.run(cmd + [str(x) for x in a], capture_output=True, text=True,
                                                        check=True).stdout.splitlines() if l.strip()]
    ref_h = run("hybrid", str(tmp_path / "hy.txt"), 13)
    ref_p = [float(r[0]) for r in run("prob", str(tmp_path / "rnd.txt"))]
    ref_c = run("chain", str(tmp_path / "rnd.txt"))
    run("prefix", str(tmp_path / "rnd.txt"), 30, str(tmp_path / "d"))
    tol = TOL[storage]
    print(f"gadget n=48: runs of diagonal steps, storage={TAG[storage]} grid_log2={grid_log2}", flush=True)
    with gs.Engine(path, storage) as e:
        e.set_diag_runs(8, 10)
        if grid_log2:
            e.set_grid_log2(grid_log2)
        runs = {e.run_of(k) for k in range(ns)} - {None}
        assert {(r[0], r[1]) for r in runs} == {(26, 8), (34, 4)}, runs
        assert all(r[3] == (5 if grid_log2 else 0) for r in runs)
        # the second run is followed by a two-slot step: with tile coordinates to pair (2^4 blocks), its product pass
        # also sums that step's Gram matrix for the sampled strings
        assert {r[0]: r[4] for r in runs} == {26: 0, 34: 1 if grid_log2 else 0}, runs
        ob, lph, condh = e.hybrid(hy, 13)
        lpp = e.prob(rnd)
        lpc, cond, _ = e.chain(rnd)
        tab, nk, ntq, ls, _ = e.prefix(rnd[0], 30)
        e.set_diag_runs(0)
        ob0, lph0, _ = e.hybrid(hy, 13)
    assert [egs.planio.format_bits(b) for b in ob] == [r[0] for r in ref_h]
    assert close(lph, [float(r[1]) for r in ref_h], 0) <= tol
    assert close(condh, [[float(x) for x in r[2:]] for r in ref_h], 0) <= max(tol, 1e-9)
    assert close(lpp, ref_p, 0) <= tol
    assert close(lpc, [float(r[0]) for r in ref_c], 0) <= tol
    assert close(cond, [[float(x) for x in r[1:]] for r in ref_c], 0) <= max(tol, 1e-9)
    dump, dnk, dntq, dls = egs.planio.read_dump(str(tmp_path / "d_0.bin"), c128=storage == gs.C128)
    assert (dnk, dntq) == (nk, ntq) and abs(dls - ls) <= max(tol, 1e-12)
    assert np.max(np.abs(np.asarray(dump) - tab)) <= 10 * tol * max(1.0, np.max(np.abs(tab)))
    assert np.array_equal(ob0, ob) and close(lph0, lph, 0) <= tol


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_fold_reduce_bitwise_equal(golden, storage):
    """IBM #151 width-30 plan: sweeps that reduce the previous sweep's block partials themselves against a one-block
    reduction kernel between them (default) -- sample, prob, chain and hybrid outputs bitwise equal (the same
    fixed-order sums and chain step either way)."""
    d = os.path.join(golden, "ibm")
    strings = egs.planio.read_strings(os.path.join(d, "gate_strings.txt"), 70)
    hy = strings.copy()
    hy[:, 50:] = -1
    got = {}
    print(f"fold reduce on/off, IBM width 30, storage={TAG[storage]}", flush=True)
    with gs.Engine(os.path.join(d, "nq70_inplace_plan.txt"), storage) as e:
        for fold in (True, False):
            e.set_fold_reduce(fold)
            got[fold] = (e.sample(2, 2000000), e.prob(strings), e.chain(strings), e.hybrid(hy, 17))
    (b1, l1), p1, (c1, q1, a1), (h1, hl1, hc1) = got[True]
    (b0, l0), p0, (c0, q0, a0), (h0, hl0, hc0) = got[False]
    for x, y in ((b1, b0), (l1, l0), (p1, p0), (c1, c0), (q1, q0), (a1, a0), (h1, h0), (hl1, hl0), (hc1, hc0)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
@pytest.mark.parametrize("plan", ["c4", "ibm"])
def test_fused_four_slot_next_step(golden, storage, plan):
    """Writes followed by a four-slot step fold that step's norms in (two pairing coordinates per thread: the thread's
    four cosets hold whole next-step cosets).  The n = 60 plan (nk up to 18 at such steps) and IBM's width-30 plan:
    sample, prob, chain and hybrid with the folding on equal the same with those norm passes as their own sweeps --
    bits identical, complex128 within 1e-12, complex64 within 1e-5 -- and the n = 60 plan's chain-mode log2 Pr equal the
    reference executor's (golden)."""
    tol = TOL[storage]
    if plan == "c4":
        d = os.path.join(golden, "c4")
        path = os.path.join(d, "plan.txt")
        strings = egs.planio.read_strings(os.path.join(d, "strings.txt"), 60)
    else:
        d = os.path.join(golden, "ibm")
        path = os.path.join(d, "nq70_inplace_plan.txt")
        strings = egs.planio.read_strings(os.path.join(d, "gate_strings.txt"), 70)
    hy = strings.copy()
    hy[:, hy.shape[1] // 3:] = -1
    got = {}
    print(f"four-slot folded norms, plan={plan} storage={TAG[storage]}", flush=True)
    with gs.Engine(path, storage) as e:
        four = [k for k in range(e.ns - 1) if e.geometry(k)["fvg"] == 4]
        assert len(four) >= 5, four
        for on in (True, False):
            e.set_fused4(on)
            assert any(e.geometry(k)["fvg"] == 4 for k in range(e.ns - 1)) == on
            got[on] = (e.sample(4, 31), e.prob(strings), e.chain(strings), e.hybrid(hy, 9))
    (b1, l1), p1, (c1_, q1, a1), (h1, hl1, hc1) = got[True]
    (b0, l0), p0, (c0, q0, a0), (h0, hl0, hc0) = got[False]
    assert np.array_equal(b1, b0) and np.array_equal(h1, h0)
    for x, y in ((l1, l0), (p1, p0), (c1_, c0), (hl1, hl0)):
        assert close(x, y, 0) <= tol, (x, y)
    for x, y in ((q1, q0), (hc1, hc0)):
        assert close(x, y, 0) <= max(tol, 1e-9)
    if plan == "c4":
        ref = lines(os.path.join(d, f"chain_{TAG[storage]}.txt"))
        assert close(c1_, [float(r[0]) for r in ref], 0) <= tol


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_tiny_plans_and_empty_batches(tmp_path, storage):
    """Edge sizes: plans of 1, 2 and 3 vertices (a single step; nk = 0 tables), every string of each scored with prob
    and chain and 5 samples drawn, against the oracle (bits identical; complex128 1e-12, complex64 1e-5) and summing to
    probability 1; and empty batches (no strings, no samples) return empty results without touching the device."""
    import itertools
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    tol = TOL[storage]
    for n, band in ((1, 1), (2, 1), (3, 2)):
        inst = egs.instances.Instance(n, band, 0)
        path = str(tmp_path / f"tiny{n}.txt")
        inst.write_plan(path)
        strings = np.array(list(itertools.product([0, 1], repeat=n)), np.int8)
        egs.planio.write_strings(str(tmp_path / f"s{n}.txt"), strings)
        cmd = [os.path.join(root, "oracle", "gs_oracle")] + (["--c128"] if storage == gs.C128 else []) + [path]
        run = lambda *a: [l.split() for l in subprocess.run(cmd + [str(x) for x in a], capture_output=True, text=True,
                                                            check=True).stdout.splitlines() if l.strip()]
        ref_s = run("sample", 5, 5)
        ref_p = [float(r[0]) for r in run("prob", str(tmp_path / f"s{n}.txt"))]
        ref_c = run("chain", str(tmp_path / f"s{n}.txt"))
        print(f"tiny plan n={n} storage={TAG[storage]}", flush=True)
        with gs.Engine(path, storage) as e:
            assert e.ns == n
            bits, lp = e.sample(5, 5)
            lpp = e.prob(strings)
            lpc, cond, _ = e.chain(strings)
            eb, el = e.sample(0, 5)
            assert eb.shape == (0, n) and el.shape == (0,)
            assert e.prob(strings[:0]).shape == (0,)
            c0, q0, _ = e.chain(strings[:0])
            assert c0.shape == (0,) and q0.shape[0] == 0
        assert [egs.planio.format_bits(b) for b in bits] == [r[0] for r in ref_s]
        assert close(lp, [float(r[1]) for r in ref_s], 0) <= tol
        assert close(lpp, ref_p, 0) <= tol
        assert close(lpc, [float(r[0]) for r in ref_c], 0) <= tol
        assert abs(np.sum(2.0 ** lpp) - 1.0) <= 10 * tol
