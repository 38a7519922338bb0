"""The two-table (gather-form) engine against the reference gsamp / dp outputs (tests/golden/tt, complex64 tables as
the reference's; dp complex128 and -DDP_FLOAT), and the reference's committed two-table chain known answers on IBM's
width-30 plans (results/chain/gate_fwd_out.txt, gate_rev_out.txt, ideal_fwd_chain.txt)."""
import importlib.util
import os
import subprocess

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location("twotable_ref", os.path.join(ROOT, "oracle", "twotable_ref.py"))
tt = importlib.util.module_from_spec(spec)
spec.loader.exec_module(tt)


def lines(path):
    return [l.split() for l in open(path) if l.strip()]


@pytest.fixture(scope="module")
def tdir(golden):
    return os.path.join(golden, "tt")


def test_c1_modes_vs_reference_gsamp(tdir, golden):
    plan = os.path.join(tdir, "c1_plan.txt")
    print("two-table c1: prob, chain, sample, hybrid", flush=True)
    with gs.TwoTableEngine(plan) as e:
        strings = egs.planio.read_strings(os.path.join(golden, "c1", "strings.txt"), e.ns)
        hy = egs.planio.read_strings(os.path.join(golden, "c1", "hybrid.txt"), e.ns, allow_q=True)
        lp = e.prob(strings)
        lpc, cond, amp = e.chain(strings)
        bits, lps = e.sample(16, 1234)
        ob, lph, condh = e.hybrid(hy, 99)
    assert np.max(np.abs(lp - [float(l[0]) for l in lines(os.path.join(tdir, "c1_prob.txt"))])) < 1e-5
    refc = lines(os.path.join(tdir, "c1_chain.txt"))
    assert np.max(np.abs(lpc - [float(l[0]) for l in refc])) < 1e-5
    assert np.max(np.abs(cond - [[float(x) for x in l[1:]] for l in refc])) < 1e-5
    assert np.max(np.abs(np.log2(np.abs(amp) ** 2) - e.ns - lpc)) < 1e-5
    refs = lines(os.path.join(tdir, "c1_sample.txt"))
    assert [egs.planio.format_bits(b) for b in bits] == [l[0] for l in refs]
    assert np.max(np.abs(lps - [float(l[1]) for l in refs])) < 1e-5
    refh = lines(os.path.join(tdir, "c1_hybrid.txt"))
    assert [egs.planio.format_bits(b) for b in ob] == [l[0] for l in refh]
    assert np.max(np.abs(lph - [float(l[1]) for l in refh])) < 1e-5


@pytest.mark.parametrize("storage,tag,tol", [(gs.C128, "c128", 1e-12), (gs.C64, "c64", 1e-5)])
def test_dp_vs_reference(tdir, storage, tag, tol):
    m, _, _ = tt.load_plan(os.path.join(tdir, "dp_plan.txt"), with_m=False)
    jobs = tt.read_jobs(os.path.join(tdir, "dp_jobs.txt"), m)
    want = np.array([complex(float(a), float(b)) for a, b in lines(os.path.join(tdir, f"dp_out_{tag}.txt"))])
    print(f"dp {tag}", flush=True)
    with gs.TwoTableEngine(os.path.join(tdir, "dp_plan.txt"), with_m=False, storage=storage) as e:
        got = e.dp(jobs)
    assert np.max(np.abs(got - want)) < tol * np.max(np.abs(want))


def test_ibm_chain_known_answers(tdir, golden):
    """gsamp chain on cloud_ibmxeb/nq70_w30_plan.txt: the gate strings (results/chain/gate_fwd_out.txt) and 4 of the
    1,012 ideal samples with all 70 conditionals (results/chain/ideal_fwd_chain.txt); the reversed-order plan
    cloud_chain/plan_rev.txt with its gate strings (results/chain/gate_rev_out.txt)."""
    print("two-table IBM w30 chain KATs", flush=True)
    with gs.TwoTableEngine(os.path.join(golden, "ibm", "nq70_w30_plan.txt")) as e:
        gate = egs.planio.read_strings(os.path.join(golden, "ibm", "gate_strings.txt"), e.ns)
        lp, _, _ = e.chain(gate)
        ideal = egs.planio.read_strings(os.path.join(tdir, "ibm_ideal_fwd_strings.txt"), e.ns)
        lpi, condi, _ = e.chain(ideal)
    want = [float(l[0]) for l in lines(os.path.join(tdir, "ibm_gate_fwd_chain.txt"))]
    assert np.max(np.abs(lp - want)) < 1e-5, (lp, want)
    ref = lines(os.path.join(tdir, "ibm_ideal_fwd_chain.txt"))
    assert np.max(np.abs(lpi - [float(l[0]) for l in ref])) < 1e-5
    assert np.max(np.abs(condi - [[float(x) for x in l[1:]] for l in ref])) < 1e-6
    with gs.TwoTableEngine(os.path.join(tdir, "ibm_plan_rev.txt")) as e:
        gate_rev = egs.planio.read_strings(os.path.join(tdir, "ibm_gate_rev.txt"), e.ns)
        lpr, _, _ = e.chain(gate_rev)
    want = [float(l[0]) for l in lines(os.path.join(tdir, "ibm_gate_rev_chain.txt"))]
    assert np.max(np.abs(lpr - want)) < 1e-5, (lpr, want)


@pytest.mark.parametrize("storage,tol", [(gs.C64, 1e-5), (gs.C128, 1e-12)])
def test_grid_caps_agree(golden, storage, tol):
    """The gather sweeps compute every output the same way whatever the launch shape; the block count changes only the
    order of the FP64 norm sums, hence the last bits of each step's scale and, in complex64, the rounding of the
    scaled table.  On IBM's width-30 plan, prob and chain of the gate strings and two samples at 2^9, 2^12 and the
    default 2^14 blocks: identical bits, log2 Pr within 1e-5 (complex64) / 1e-12 (complex128); a cap beyond 2^14 is
    rejected."""
    print(f"two-table IBM w30: grid caps 2^9, 2^12, default, storage {storage}", flush=True)
    got = {}
    with gs.TwoTableEngine(os.path.join(golden, "ibm", "nq70_w30_plan.txt"), storage=storage) as e:
        gate = egs.planio.read_strings(os.path.join(golden, "ibm", "gate_strings.txt"), e.ns)
        for lg in (9, 12, 0):
            e.set_grid_log2(lg)
            got[lg] = (e.prob(gate), e.chain(gate)[0], e.sample(2, 77))
        with pytest.raises(gs.GsError):
            e.set_grid_log2(15)
    p0, c0, (b0, l0) = got[0]
    for lg in (9, 12):
        p, c, (b, l) = got[lg]
        assert np.array_equal(b, b0)
        for x, y in ((p, p0), (c, c0), (l, l0)):
            assert np.max(np.abs(np.asarray(x) - np.asarray(y))) < tol, (lg, x, y)


def test_cli_front_ends(tdir, golden, tmp_path):
    """gsamp_amd and dp_amd print the reference's lines."""
    bindir = os.path.join(ROOT, "exact-graphstate-sampling_amd", "bin")
    out = subprocess.run([os.path.join(bindir, "gsamp_amd"), os.path.join(tdir, "c1_plan.txt"), "sample", "3", "1234"],
                         capture_output=True, text=True, check=True).stdout.splitlines()
    ref = [l for l in open(os.path.join(tdir, "c1_sample.txt")).read().splitlines()][:3]
    assert [l.split()[0] for l in out] == [l.split()[0] for l in ref]
    assert [l.split()[2] for l in out] == [l.split()[2] for l in ref]
    out = subprocess.run([os.path.join(bindir, "dp_amd"), os.path.join(tdir, "dp_plan.txt"),
                          os.path.join(tdir, "dp_jobs.txt")], capture_output=True, text=True, check=True).stdout.split()
    want = [float(x) for x in open(os.path.join(tdir, "dp_out_c128.txt")).read().split()]
    assert np.max(np.abs(np.array([float(x) for x in out]) - want)) < 1e-12 * np.max(np.abs(want))


def test_c1_prefix_resume_vs_reference_gsamp(tdir, golden):
    """gsamp prefix / resume at K = 18 on the hybrid lines: the dumped tables (the reference's own dump files) and the
    resumed samples."""
    import json
    meta = json.load(open(os.path.join(golden, "c1", "meta.json")))
    K = meta["K"]
    ref_pre = lines(os.path.join(tdir, "c1_prefix.txt"))
    ref_res = lines(os.path.join(tdir, "c1_resume.txt"))
    print("two-table prefix/resume", flush=True)
    with gs.TwoTableEngine(os.path.join(tdir, "c1_plan.txt")) as e:
        hy = egs.planio.read_strings(os.path.join(golden, "c1", "hybrid.txt"), e.ns, allow_q=True)
        assert len(hy) == len(ref_pre) == len(ref_res)
        for li, row in enumerate(hy):
            tab, r, ls, lp = e.prefix(row, K)
            rtab, rr, rone, rls = egs.planio.read_dump(os.path.join(tdir, f"c1_dump_{li}.bin"))
            assert (r, 1) == (rr, rone)
            assert abs(ls - rls) < 1e-6 and abs(lp - float(ref_pre[li][1])) < 1e-5
            assert np.max(np.abs(tab - rtab)) < 1e-5 * np.max(np.abs(rtab))
            bits, lpf, lpp, cond = e.resume(row, K, [rtab], [rls], meta["resume_seed"], li)
            assert egs.planio.format_bits(bits) == ref_res[li][0]
            assert abs(lpf - float(ref_res[li][1])) < 1e-5 and abs(lpp - float(ref_res[li][2])) < 1e-5
            assert np.max(np.abs(cond[K:] - [float(x) for x in ref_res[li][3:]])) < 1e-6


def test_prefix_resume_cli_and_two_slices(tdir, golden, tmp_path):
    """gsamp_amd prefix writes the reference's dump files; resume over two dump prefixes (the table split into two
    weighted slices) reproduces the reference's single-slice resume lines."""
    import json
    meta = json.load(open(os.path.join(golden, "c1", "meta.json")))
    K = meta["K"]
    exe = os.path.join(ROOT, "exact-graphstate-sampling_amd", "bin", "gsamp_amd")
    plan = os.path.join(tdir, "c1_plan.txt")
    hyf = os.path.join(golden, "c1", "hybrid.txt")
    out = subprocess.run([exe, plan, "prefix", hyf, str(K), str(tmp_path / "d")], capture_output=True, text=True,
                         check=True).stdout.splitlines()
    ref_pre = lines(os.path.join(tdir, "c1_prefix.txt"))
    assert len(out) == len(ref_pre)
    for li, l in enumerate(out):
        assert l.split()[0] == ref_pre[li][0] and abs(float(l.split()[1]) - float(ref_pre[li][1])) < 1e-5
        got, r, one, ls = egs.planio.read_dump(str(tmp_path / f"d_{li}.bin"))
        want, rr, _, rls = egs.planio.read_dump(os.path.join(tdir, f"c1_dump_{li}.bin"))
        assert (r, one) == (rr, 1) and abs(ls - rls) < 1e-6
        assert np.max(np.abs(got - want)) < 1e-5 * np.max(np.abs(want))
        # split the reference's table into two slices: a * 2^-2 at logscale ls + 2 and the rest at ls - 1
        a = want * np.linspace(0.25, 0.75, want.size)
        for tag, t, s in (("a", a * 0.25, rls + 2), ("b", (want - a) * 2.0, rls - 1)):
            with open(tmp_path / f"{tag}_{li}.bin", "wb") as f:
                f.write(np.array([r, 1], np.int32).tobytes() + np.array([s]).tobytes())
                f.write(np.stack([t.real, t.imag], 1).astype(np.float32).tobytes())
    out = subprocess.run([exe, plan, "resume", hyf, str(K), str(meta["resume_seed"]), str(tmp_path / "a"),
                          str(tmp_path / "b")], capture_output=True, text=True, check=True).stdout.splitlines()
    ref_res = lines(os.path.join(tdir, "c1_resume.txt"))
    assert len(out) == len(ref_res)
    for l, want in zip(out, ref_res):
        got = l.split()
        assert got[0] == want[0]
        assert np.max(np.abs(np.array([float(x) for x in got[1:]]) - [float(x) for x in want[1:]])) < 1e-5
