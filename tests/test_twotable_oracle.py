"""The two-table restatement (oracle/twotable_ref.py) against the reference gsamp / dp outputs committed as fixtures
(tests/golden/tt, written by the reference binaries compiled from their sources), and the independent cross-check the
reference makes (results/xcheck70_independent.txt): dp's sliced complex amplitude equals the in-place planner's
Z(w_b), whose |Z|^2 2^-ns is the reference gsamp2's probability."""
import importlib.util
import os

import numpy as np
import pytest

from _pkg import egs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location("twotable_ref", os.path.join(ROOT, "oracle", "twotable_ref.py"))
tt = importlib.util.module_from_spec(spec)
spec.loader.exec_module(tt)


@pytest.fixture(scope="module")
def tdir(golden):
    return os.path.join(golden, "tt")


def lines(path):
    return [l.split() for l in open(path) if l.strip()]


def test_prob_and_chain_match_reference_gsamp(tdir, golden):
    _, _, steps = tt.load_plan(os.path.join(tdir, "c1_plan.txt"))
    strings = egs.planio.read_strings(os.path.join(golden, "c1", "strings.txt"), len(steps))[:16]
    ref_p = [float(l[0]) for l in lines(os.path.join(tdir, "c1_prob.txt"))]
    ref_c = lines(os.path.join(tdir, "c1_chain.txt"))
    for i, b in enumerate(strings):
        assert abs(tt.prob(steps, b) - ref_p[i]) < 1e-9
        lp, cond = tt.chain1(steps, b)
        assert abs(lp - float(ref_c[i][0])) < 1e-9
        assert np.max(np.abs(cond - np.array([float(x) for x in ref_c[i][1:]]))) < 2e-9


def test_two_table_equals_in_place(tdir, golden):
    """Two algorithms, one answer: the two-table chain of the config-1 plan equals the in-place executor's (reference
    gsamp2 outputs) within complex64 rounding."""
    ref_tt = [float(l[0]) for l in lines(os.path.join(tdir, "c1_prob.txt"))]
    ref_ip = [float(l[0]) for l in lines(os.path.join(golden, "c1", "prob_c128.txt"))]
    assert np.max(np.abs(np.array(ref_tt) - np.array(ref_ip))) < 1e-5


def test_dp_matches_reference_and_in_place_amplitude(tdir, golden):
    m, _, steps = tt.load_plan(os.path.join(tdir, "dp_plan.txt"), with_m=False)
    jobs = tt.read_jobs(os.path.join(tdir, "dp_jobs.txt"), m)
    want = np.array([complex(float(a), float(b)) for a, b in lines(os.path.join(tdir, "dp_out_c128.txt"))])
    got = tt.dp(steps, jobs)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    want32 = np.array([complex(float(a), float(b)) for a, b in lines(os.path.join(tdir, "dp_out_c64.txt"))])
    assert np.max(np.abs(tt.dp(steps, jobs, c64=True) - want32)) < 1e-5 * np.max(np.abs(want))
    # |Z|^2 2^-ns = Pr(b) from the in-place executor (reference gsamp2, complex128) for the same six strings
    ns = 24
    ref_ip = np.array([float(l[0]) for l in lines(os.path.join(golden, "c1", "prob_c128.txt"))][:6])
    assert np.max(np.abs(np.log2(np.abs(want) ** 2) - ns - ref_ip)) < 1e-9


def test_prefix_and_resume_match_reference_gsamp(tdir, golden):
    """prefix at K = 18 against the reference's dump files and printed lines; resume with the bits the reference drew
    (its '?' outcomes taken from its output) against its printed probabilities and conditionals."""
    import json
    meta = json.load(open(os.path.join(golden, "c1", "meta.json")))
    K = meta["K"]
    _, _, steps = tt.load_plan(os.path.join(tdir, "c1_plan.txt"))
    ref_pre = lines(os.path.join(tdir, "c1_prefix.txt"))
    ref_res = lines(os.path.join(tdir, "c1_resume.txt"))
    hy = egs.planio.read_strings(os.path.join(golden, "c1", "hybrid.txt"), len(steps), allow_q=True)
    assert len(hy) == len(ref_pre) == len(ref_res) == 8
    for li, row in enumerate(hy):
        tab, ls, lp = tt.prefix(steps, row, K)
        rtab, r, one, rls = egs.planio.read_dump(os.path.join(tdir, f"c1_dump_{li}.bin"))
        assert (one, len(tab)) == (1, 1 << r) and r == steps[K]["r"]
        assert abs(ls - rls) < 1e-6 and abs(lp - float(ref_pre[li][1])) < 1e-9
        assert np.max(np.abs(tab - rtab)) < 1e-6 * np.max(np.abs(rtab))
        bits = np.array([int(c) for c in ref_res[li][0]])
        assert all(bits[k] == row[k] for k in range(len(steps)) if row[k] >= 0)
        lpf, lpp, cond = tt.resume(steps, bits, K, [rtab], [rls])
        assert abs(lpf - float(ref_res[li][1])) < 1e-9 and abs(lpp - float(ref_res[li][2])) < 1e-9
        assert np.max(np.abs(cond[K:] - [float(x) for x in ref_res[li][3:]])) < 2e-9


def test_resume_of_two_slices_is_linear(tdir, golden):
    """Resume of a table split into two weighted slices equals resume of the whole table (the slice sum is linear;
    logscales shift the weights)."""
    import json
    meta = json.load(open(os.path.join(golden, "c1", "meta.json")))
    K = meta["K"]
    _, _, steps = tt.load_plan(os.path.join(tdir, "c1_plan.txt"))
    ref_res = lines(os.path.join(tdir, "c1_resume.txt"))
    rtab, _, _, rls = egs.planio.read_dump(os.path.join(tdir, "c1_dump_0.bin"))
    bits = np.array([int(c) for c in ref_res[0][0]])
    rng = np.random.default_rng(3)
    part = rtab * rng.uniform(0.2, 0.8, rtab.size)
    whole = tt.resume(steps, bits, K, [rtab], [rls], c64=False)
    split = tt.resume(steps, bits, K, [part * 2.0 ** -3, (rtab - part) * 2.0 ** 1], [rls + 3, rls - 1], c64=False)
    assert abs(whole[0] - split[0]) < 1e-9 and np.max(np.abs(whole[2] - split[2])) < 1e-9


def test_sliced_jobs_reproduce_the_committed_jobs(tdir, golden):
    """engine.Sliced + write_jobs rebuild the committed dp job file (vertices 3 and 11 of the config-1 instance sliced,
    weights of six strings) on which the reference dp binary's outputs were taken: every number equal."""
    inst = egs.instances.Instance(24, 10, 1)
    sl = egs.engine.Sliced(inst.am.A, inst.order, [3, 11])
    bits = egs.planio.read_strings(os.path.join(golden, "c1", "strings.txt"), len(inst.steps))[:6]
    ws = []
    for b in bits:
        bq = egs.planio.step_to_qubit(b, inst.step_vertices, inst.N)
        ws.append([(complex(inst.am.M[q][bq[q], 0]), complex(inst.am.M[q][bq[q], 1])) for q in range(inst.N)])
    import tempfile
    with tempfile.TemporaryDirectory() as t:
        path = os.path.join(t, "jobs.txt")
        egs.engine.write_jobs(path, sl, ws)
        got = [float(x) for x in open(path).read().split()]
    want = [float(x) for x in open(os.path.join(tdir, "dp_jobs.txt")).read().split()]
    assert got == want
