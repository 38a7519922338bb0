"""Known answers the reference's own runs hold, and the reference executor's outputs on two more fixture sets, against
the HIP engines (through the C ABI).

* results/ibmxeb/uniform_out.txt: 64 uniform strings (cloud_ibmxeb/uniform64.txt) scored by the two-table sampler on
  the width-30 plan; the two-table engine here, and the in-place engine on the in-place plan of the same order;
* results/samples_nq70.txt: the two-table sampler's first 4 exact samples on the width-31 plan, seeds 1000000..;
* cloud/check55_expected.txt: the 55-qubit bit-exact gate of cloud/run.sh:19-27 (gsamp sample 3 1, 1e-6);
* zero-probability branches (a plan with a zeroed M row): -inf, zero conditionals, the sampled bits set to 0 after
  the failing step, as the reference prints them (gsamp2.cpp:48, :67, :94, :116);
* the n = 60, band-24 plan (nk up to 20): every mode against the reference executor's outputs.
"""
import json
import os

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu
TOL = {gs.C64: 1e-5, gs.C128: 1e-12}
TAG = {gs.C64: "c64", gs.C128: "c128"}


def lines(path):
    with open(path) as f:
        return [l.split() for l in f if l.strip()]


def close(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    same_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    assert not np.any(np.isinf(a) ^ np.isinf(b)), (a, b)
    with np.errstate(invalid="ignore"):
        d = np.where(same_inf, 0.0, np.abs(a - b))
    return float(np.max(d)) if d.size else 0.0


def test_ibm_uniform64_known_answers(golden):
    """64 uniform strings of the IBM #151 circuit (width-30 order) against results/ibmxeb/uniform_out.txt, the
    reference's two-table sampler on cloud_ibmxeb/nq70_w30_plan.txt in complex64 (log2 Pr down to -77.6): the in-place
    engine on cloud_ibmxeb/nq70_inplace_plan.txt within 1e-5 in complex128 (the reference's own complex64 rounding,
    up to 5.9e-6 here), and both engines within 2e-5 in complex64 (two complex64 executions, each rounding its tables
    on its own: 1.02e-5 measured, the same for the same plan and algorithm)."""
    d = os.path.join(golden, "ibm")
    want = [float(l[0]) for l in lines(os.path.join(d, "uniform64_out.txt"))]
    got = {}
    for storage in (gs.C64, gs.C128):
        print(f"ibm uniform64: in-place engine storage={TAG[storage]}", flush=True)
        with gs.Engine(os.path.join(d, "nq70_inplace_plan.txt"), storage) as e:
            strings = egs.planio.read_strings(os.path.join(d, "uniform64.txt"), e.ns)
            assert len(strings) == 64
            got[storage] = e.prob(strings)
    print("ibm uniform64: two-table engine", flush=True)
    with gs.TwoTableEngine(os.path.join(d, "nq70_w30_plan.txt")) as e:
        lpt = e.prob(strings)
    print(f"  max |d| in-place c64 {close(got[gs.C64], want):.2e}, in-place c128 {close(got[gs.C128], want):.2e}, "
          f"two-table c64 {close(lpt, want):.2e}", flush=True)
    assert close(lpt, want) < 2e-5
    assert close(got[gs.C64], want) < 2e-5
    assert close(got[gs.C128], want) < 1e-5


def test_ibm_two_table_samples_known_answers(golden):
    """The first 4 lines of results/samples_nq70.txt (gsamp cloud/nq70_plan.txt sample N 1000000): identical bits and
    seeds, log2 Pr within 1e-5."""
    d = os.path.join(golden, "ibm")
    ref = lines(os.path.join(d, "samples_nq70_first4.txt"))
    print("ibm two-table width-31 samples 1000000..1000003", flush=True)
    with gs.TwoTableEngine(os.path.join(d, "nq70_w31_plan.txt")) as e:
        bits, lp = e.sample(4, 1000000)
    assert [egs.planio.format_bits(b) for b in bits] == [r[0] for r in ref]
    assert close(lp, [float(r[1]) for r in ref]) < 1e-5


def test_check55_gate(golden):
    """cloud/run.sh:19-27: gsamp check55_plan.txt sample 3 1 must reproduce check55_expected.txt -- bits identical,
    log2 Pr within 1e-6."""
    d = os.path.join(golden, "tt")
    ref = lines(os.path.join(d, "check55_expected.txt"))
    print("check55 gate: two-table sample 3 1", flush=True)
    with gs.TwoTableEngine(os.path.join(d, "check55_plan.txt")) as e:
        bits, lp = e.sample(3, 1)
    assert [egs.planio.format_bits(b) for b in bits] == [r[0] for r in ref]
    assert [r[2] for r in ref] == ["1", "2", "3"]
    assert close(lp, [float(r[1]) for r in ref]) < 1e-6


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
@pytest.mark.parametrize("path", ["default", "plain"])
def test_zero_probability_branches(golden, storage, path):
    """Row b = 1 of step 5's M is zero: prob and chain give -inf for strings forcing b_5 = 1 (chain's conditionals 0
    from step 5 on), hybrid gives -inf, conditionals 0 and bits 0 at the sampled steps after it, sampling never takes
    the zero branch -- each line as the reference executor printed it (default engine, and every norm pass as its
    own sweep with the small-step runs off)."""
    d = os.path.join(golden, "zero")
    tag = TAG[storage]
    tol = TOL[storage]
    print(f"zero-probability branches storage={tag} path={path}", flush=True)
    with gs.Engine(os.path.join(d, "plan.txt"), storage) as e:
        if path == "plain":
            e.set_fusion(False)
            e.set_small_steps(-1)
        strings = egs.planio.read_strings(os.path.join(d, "strings.txt"), e.ns)
        hy = egs.planio.read_strings(os.path.join(d, "hybrid.txt"), e.ns, allow_q=True)
        lp = e.prob(strings)
        lpc, cond, amp = e.chain(strings)
        sb, slp = e.sample(8, 77)
        hb, hlp, hcond = e.hybrid(hy, 41)
    ref = lines(os.path.join(d, f"prob_{tag}.txt"))
    assert np.isneginf(lp[:4]).all() and close(lp, [float(r[0]) for r in ref]) <= tol
    ref = lines(os.path.join(d, f"chain_{tag}.txt"))
    assert np.isneginf(lpc[:4]).all() and np.all(cond[:4, 5:] == 0)
    assert close(lpc, [float(r[0]) for r in ref]) <= tol
    assert close(cond, [[float(x) for x in r[1:]] for r in ref]) <= max(tol, 1e-9)
    assert np.all(amp[:4] == 0)
    ref = lines(os.path.join(d, f"sample_{tag}.txt"))
    assert [egs.planio.format_bits(b) for b in sb] == [r[0] for r in ref] and np.all(sb[:, 5] == 0)
    assert close(slp, [float(r[1]) for r in ref]) <= tol
    ref = lines(os.path.join(d, f"hybrid_{tag}.txt"))
    assert [egs.planio.format_bits(b) for b in hb] == [r[0] for r in ref]
    assert np.isneginf(hlp[:3]).all()
    assert close(hlp, [float(r[1]) for r in ref]) <= tol
    assert close(hcond, [[float(x) for x in r[2:]] for r in ref]) <= max(tol, 1e-9)


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_c4_every_mode_vs_reference(golden, storage):
    """The n = 60, band-24 plan: prob, chain, sample, sample2, hybrid, prefix dumps and resume against the reference
    executor's outputs (bits identical; 1e-12 complex128, 1e-5 complex64)."""
    d = os.path.join(golden, "c4")
    meta = json.load(open(os.path.join(d, "meta.json")))
    tag, tol = TAG[storage], TOL[storage]
    c128 = storage == gs.C128
    print(f"c4 every mode storage={tag}", flush=True)
    with gs.Engine(os.path.join(d, "plan.txt"), storage) as e:
        strings = egs.planio.read_strings(os.path.join(d, "strings.txt"), e.ns)
        hy = egs.planio.read_strings(os.path.join(d, "hybrid.txt"), e.ns, allow_q=True)
        ref = lines(os.path.join(d, f"prob_{tag}.txt"))
        assert close(e.prob(strings), [float(r[0]) for r in ref]) <= tol
        ref = lines(os.path.join(d, f"chain_{tag}.txt"))
        lpc, cond, _ = e.chain(strings)
        assert close(lpc, [float(r[0]) for r in ref]) <= tol
        assert close(cond, [[float(x) for x in r[1:]] for r in ref]) <= max(tol, 1e-9)
        for mode, seeding in (("sample", gs.SEED_PLAIN), ("sample2", gs.SEED_SEQ)):
            ref = lines(os.path.join(d, f"{mode}_{tag}.txt"))
            bits, lp = e.sample(len(ref), meta["sample_seed"], seeding=seeding)
            assert [egs.planio.format_bits(b) for b in bits] == [r[0] for r in ref], mode
            assert close(lp, [float(r[1]) for r in ref]) <= tol, mode
        ref = lines(os.path.join(d, f"hybrid_{tag}.txt"))
        hb, hlp, hcond = e.hybrid(hy, meta["hybrid_seed"])
        assert [egs.planio.format_bits(b) for b in hb] == [r[0] for r in ref]
        assert close(hlp, [float(r[1]) for r in ref]) <= tol
        assert close(hcond, [[float(x) for x in r[2:]] for r in ref]) <= max(tol, 1e-9)
        K = meta["K"]
        ref_pre = lines(os.path.join(d, f"prefix_{tag}.txt"))
        ref_res = lines(os.path.join(d, f"resume_{tag}.txt"))
        for li, row in enumerate(hy):
            tab, nk, ntq, ls, lpp = e.prefix(row, K)
            rtab, rnk, rntq, rls = egs.planio.read_dump(os.path.join(d, f"dump_{tag}_{li}.bin"), c128)
            assert (nk, ntq) == (rnk, rntq) and abs(ls - rls) <= 1e-6
            assert abs(lpp - float(ref_pre[li][1])) <= tol
            assert np.max(np.abs(tab - rtab)) <= 10 * tol * np.max(np.abs(rtab))
            bits, lpf, lpp2, rc = e.resume(row, K, [rtab], [rls], seed=meta["resume_seed"], line=li)
            assert egs.planio.format_bits(bits) == ref_res[li][0]
            assert abs(lpf - float(ref_res[li][1])) <= tol and abs(lpp2 - float(ref_res[li][2])) <= tol
            assert close(rc[K:], [float(x) for x in ref_res[li][3:]]) <= max(tol, 1e-9)
