"""Sliced plans (sliced_plans.py:26-86): the step-K tables of all slice assignments, summed by resume (with the
twist when a slice vertex has suffix neighbours), give the unsliced log2 Pr.  Checked with the FP64 oracle
(prefix/resume modes, gsamp2.cpp:162-208) on the n = 24, band-10 instance; gf2_solve against brute force."""
import os
import subprocess

import numpy as np
import pytest

from _pkg import egs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "gs_oracle")
sl = egs.sliced


def run(*args):
    out = subprocess.run([ORACLE, "--c128", *map(str, args)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    return [l.split() for l in out.stdout.splitlines() if l.strip()]


def test_gf2_solve_brute_force():
    rng = np.random.default_rng(0)
    for _ in range(300):
        basis = [int(x) for x in rng.integers(0, 1 << 9, size=rng.integers(1, 7))]
        s = int(rng.integers(0, 1 << 9))
        span = {0: 0}
        for i, b in enumerate(basis):
            span.update({x ^ b: c | (1 << i) for x, c in list(span.items()) if x ^ b not in span})
        a = sl.gf2_solve(basis, s)
        assert (a is not None) == (s in span)
        if a is not None:
            acc = 0
            for i, b in enumerate(basis):
                if (a >> i) & 1:
                    acc ^= b
            assert acc == s


@pytest.mark.parametrize("S,K", [([0], 18), ([14], 18), ([13, 15], 18), ([5, 9], 24)])
def test_sliced_sum_equals_unsliced(tmp_path, S, K):
    if not os.path.exists(ORACLE):
        pytest.skip("oracle not built")
    inst = egs.instances.Instance(24, 10, 1)
    ns = 24
    full = str(tmp_path / "full.txt")
    inst.write_plan(full)
    strings = str(tmp_path / "s.txt")
    bits = np.random.default_rng(7).integers(0, 2, size=(6, ns)).astype(np.int8)
    egs.planio.write_strings(strings, bits)
    want = [float(r[0]) for r in run(full, "prob", strings)]
    out = str(tmp_path / "sl")
    info = sl.make_sliced_plans(inst.am.A, inst.am.M, inst.order, S, out, K=K)
    assert info["R"] <= inst.width and info["K"] == K
    A = np.asarray(inst.am.A)
    assert info["twist"] == any(A[v, u] for v in S for u in inst.order[K:])      # K = ns: no suffix, no twist
    dumps = []
    for i, p in enumerate(info["plans"]):
        d = str(tmp_path / f"d{i}")
        run(p, "prefix", strings, K, d)
        dumps.append(d)
    extra = ["--twist", out + "_twist.txt"] if info["twist"] else []
    got = [float(r[1]) for r in run(info["plans"][0], "resume", strings, K, 5, *dumps, *extra)]
    assert np.max(np.abs(np.array(got) - np.array(want))) < 1e-10, (got, want)


def _digest(path):
    import hashlib
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


@pytest.fixture(scope="module")
def ref_sliced(golden):
    import json
    d = os.path.join(golden, "sliced")
    return d, json.load(open(os.path.join(d, "reference_sliced_plans.json")))


def test_tw70_sliced_plans_identical_to_reference(golden, ref_sliced, tmp_path):
    """IBM #151 (cloud_nq76/nq70.qasm, width-30 order) sliced at vertex 21 with K = 50: both slice plans and the twist
    file are byte-identical to the reference's committed cloud_nq80/tw70_s0.txt, tw70_s1.txt, tw70_twist.txt (the
    default, physical keep order; the logical order gives other files)."""
    import json
    d, ref = ref_sliced
    ibm = os.path.join(golden, "ibm")
    n, T, tail = egs.tableau.clifford_body(os.path.join(ibm, "nq70.qasm"))
    order = json.load(open(os.path.join(ibm, "order_w30.json")))
    am = egs.amp.Amplitudes(n, T, tail, order)
    out = str(tmp_path / "tw70")
    info = sl.make_sliced_plans(am.A, am.M, order, ref["tw70"]["slices"], out, K=ref["tw70"]["K"])
    assert info["twist"] and info["R"] == 30
    for t in ("0", "1"):
        assert _digest(f"{out}_s{t}.txt") == ref["sha256"][f"cloud_nq80/tw70_s{t}.txt"], t
    assert _digest(f"{out}_twist.txt") == ref["sha256"]["cloud_nq80/tw70_twist.txt"]
    assert open(f"{out}_twist.txt").read() == open(os.path.join(d, "tw70_twist.txt")).read()
    out2 = str(tmp_path / "tw70l")
    sl.make_sliced_plans(am.A, am.M, order, [21], out2, K=50, keep_order="logical")
    assert _digest(f"{out2}_s0.txt") != ref["sha256"]["cloud_nq80/tw70_s0.txt"]


def test_doped_sliced_plans_identical_to_reference(golden, ref_sliced, tmp_path):
    """The 538-vertex doped instance (T-gadget conversion of data_doped/nq70_depth70_checks27_doped.qasm, the width-34
    step order) sliced at ancillas 135 and 104: R = 33, K = 518 (the first step of the all-data suffix) and the four
    slice plans byte-identical to the reference's committed cloud_doped/doped_m20sl_s00..s11.txt."""
    import json
    _, ref = ref_sliced
    dd = os.path.join(golden, "doped")
    nd, ops = egs.tableau.parse(os.path.join(dd, "nq70_depth70_checks27_doped.qasm"))
    N, body, tail, anc = egs.gadgets.convert(nd, ops)
    q = tmp_path / "doped.qasm"
    egs.gadgets.write_qasm(str(q), N, body + tail)
    NN, TT, tl = egs.tableau.clifford_body(str(q))
    sv = json.load(open(os.path.join(dd, "doped_m20_step_vertices.json")))
    am = egs.amp.Amplitudes(NN, TT, tl, sv)
    out = str(tmp_path / "m20sl")
    info = sl.make_sliced_plans(am.A, am.M, sv, ref["doped_m20sl"]["slices"], out, nd=nd)
    assert (info["R"], info["K"], info["twist"]) == (ref["doped_m20sl"]["R"], ref["doped_m20sl"]["K"], False)
    for t in ("00", "01", "10", "11"):
        assert _digest(f"{out}_s{t}.txt") == ref["sha256"][f"cloud_doped/doped_m20sl_s{t}.txt"], t


def test_twist_line_reads_the_q_part_off_the_pivots():
    """twist_line splits a logical translation into kept coordinates and Q through the fully reduced Q basis; checked
    against a search over every element of Q on random steps of random plans (the Q part is the one element that
    agrees with the translation on every non-kept coordinate)."""
    rng = np.random.default_rng(11)
    checked = 0
    for seed in range(6):
        inst = egs.instances.Instance(14, 6, seed)
        for s in inst.steps:
            if not s["Qb"]:
                continue
            keepmask = sum(1 << i for i in s["keep"])
            for _ in range(4):
                a = int(rng.integers(0, 1 << s["r"]))
                cs, ts = sl.twist_line(s, a)
                aq = [q for q in s["Qall"] if (a ^ q) & ~keepmask == 0]
                assert len(aq) == 1
                ak = a ^ aq[0]
                assert ts == s["Tq"][aq[0]]
                assert cs == sum(((ak >> i) & 1) << j for j, i in enumerate(s["keep"]))
                checked += 1
    assert checked > 50
