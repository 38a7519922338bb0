"""The CPU oracle (oracle/gs_oracle, this project's restatement of gsamp2.cpp) is pinned before it is trusted:
identical printed outputs to the reference executor on every mode of the config-1 plan (golden files written by the
reference compiled from its own sources), probabilities equal to dense statevectors on small circuits, the
reference's exit codes on malformed plans, and -- where the reference build exists here -- the reference's
base-address unit test (test_baseof.cpp)."""
import json
import os
import subprocess

import numpy as np
import pytest

from _pkg import egs
import dense

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "gs_oracle")


def oracle(args, c128=False, cwd=None):
    cmd = [ORACLE] + (["--c128"] if c128 else []) + args
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, env=dict(os.environ, OMP_NUM_THREADS="4"))
    return res.returncode, res.stdout


def _close_lines(got, want, tol, name):
    got, want = [l.split() for l in got.splitlines()], [l.split() for l in want.splitlines()]
    assert len(got) == len(want), name
    for g, w in zip(got, want):
        assert g[0] == w[0] or abs(float(g[0]) - float(w[0])) < tol, name
        assert all(abs(float(a) - float(b)) < max(tol, 2e-9) for a, b in zip(g[1:], w[1:]) if a != b), name


@pytest.mark.parametrize("case", ["c1", "c4"])
@pytest.mark.parametrize("c128", [False, True])
def test_oracle_equals_reference_outputs(golden, case, c128, tmp_path):
    """Every mode on the config-1 plan (nk <= 9) and on the n = 60, band-24 plan (nk up to 20, so two of the
    oracle's four 10-bit lookup chunks are used): printed outputs equal the reference executor's -- complex128 within
    1e-10, complex64 within 1e-6 (the two executors sum the branch norms in different orders, so a complex64 store
    can round differently), sampled bits identical; prefix dumps entry by entry; resume from the dumps."""
    d = os.path.join(golden, case)
    meta = json.load(open(os.path.join(d, "meta.json")))
    tag = "c128" if c128 else "c64"
    tol = 1e-10 if c128 or case == "c1" else 1e-6
    plan = os.path.join(d, "plan.txt")
    nsamp = str(sum(1 for l in open(os.path.join(d, f"sample_{tag}.txt")) if l.strip()))
    cases = {
        "prob": ["prob", os.path.join(d, "strings.txt")],
        "chain": ["chain", os.path.join(d, "strings.txt")],
        "sample": ["sample", nsamp, str(meta["sample_seed"])],
        "sample2": ["sample2", nsamp, str(meta["sample_seed"])],
        "hybrid": ["hybrid", os.path.join(d, "hybrid.txt"), str(meta["hybrid_seed"])],
    }
    for name, args in cases.items():
        rc, out = oracle([plan] + args, c128)
        assert rc == 0
        _close_lines(out, open(os.path.join(d, f"{name}_{tag}.txt")).read(), tol, name)
    dump = str(tmp_path / "dump")
    rc, out = oracle([plan, "prefix", os.path.join(d, "hybrid.txt"), str(meta["K"]), dump], c128)
    assert rc == 0
    ref = [l.split() for l in open(os.path.join(d, f"prefix_{tag}.txt"))]
    got = np.array([[float(x) for x in l.split()] for l in out.splitlines()])
    assert np.allclose(got, np.array([[float(x) for x in l] for l in ref]), atol=tol)
    for li in range(len(ref)):
        a = egs.planio.read_dump(f"{dump}_{li}.bin", c128)
        b = egs.planio.read_dump(os.path.join(d, f"dump_{tag}_{li}.bin"), c128)
        # logscale: sums of the stored table's norms (complex64 stores may round differently, see above)
        assert a[1:3] == b[1:3] and abs(a[3] - b[3]) < (1e-12 if tol < 1e-9 else tol)
        assert np.allclose(a[0], b[0], rtol=0, atol=1e-6 * max(1, np.abs(b[0]).max()))
    rc, out = oracle([plan, "resume", os.path.join(d, "hybrid.txt"), str(meta["K"]), str(meta["resume_seed"]), dump],
                     c128)
    assert rc == 0
    ref = [l.split() for l in open(os.path.join(d, f"resume_{tag}.txt"))]
    for g, w in zip([l.split() for l in out.splitlines()], ref):
        assert g[0] == w[0] and np.allclose([float(x) for x in g[1:]], [float(x) for x in w[1:]], atol=max(tol, 1e-9))


@pytest.mark.parametrize("c128", [False, True])
def test_oracle_width30_sample_equals_reference(golden, c128, tmp_path):
    """The n = 70, band-30 instance (R = 30, nk up to 29: three of the oracle's 10-bit lookup chunks): the first
    sample of seed 4242 equals the reference executor's (tests/golden/c2) -- bits identical, log2 Pr within 1e-10
    (complex128) / 1e-6 (complex64).  The fixture's plan is the instance generator's."""
    d = os.path.join(golden, "c2")
    tag = "c128" if c128 else "c64"
    inst = egs.instances.config_instance("c2")
    plan_text = open(os.path.join(d, "plan.txt")).read()
    assert inst.R == 30 and max(s["nk"] for s in egs.planio.read_inplace_plan(os.path.join(d, "plan.txt"))[2]) >= 21
    inst.write_plan(os.path.join(str(tmp_path), "c2.txt"))
    assert open(os.path.join(str(tmp_path), "c2.txt")).read() == plan_text
    res = subprocess.run([ORACLE] + (["--c128"] if c128 else []) + [os.path.join(d, "plan.txt"), "sample", "1", "4242"],
                         capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 4)))
    assert res.returncode == 0
    got = res.stdout.split()
    want = open(os.path.join(d, f"sample_{tag}.txt")).readline().split()
    assert got[0] == want[0] and got[2] == want[2]
    assert abs(float(got[1]) - float(want[1])) < (1e-10 if c128 else 1e-6)


@pytest.mark.parametrize("c128", [False, True])
def test_oracle_zero_probability_semantics(golden, c128):
    """A branch of probability exactly 0 (row b = 1 of step 5's M zeroed): prob and chain print -inf, chain's
    conditionals are 0 from the failing step on, hybrid prints -inf with conditionals 0 and sampled bits 0 from there
    (gsamp2.cpp:48, :67, :94, :116) -- identical lines to the reference executor's."""
    d = os.path.join(golden, "zero")
    tag = "c128" if c128 else "c64"
    plan = os.path.join(d, "plan.txt")
    for name, args in (("prob", ["prob", os.path.join(d, "strings.txt")]),
                       ("chain", ["chain", os.path.join(d, "strings.txt")]),
                       ("sample", ["sample", "8", "77"]),
                       ("hybrid", ["hybrid", os.path.join(d, "hybrid.txt"), "41"])):
        rc, out = oracle([plan] + args, c128)
        assert rc == 0
        want = open(os.path.join(d, f"{name}_{tag}.txt")).read()
        assert sum("-inf" in l for l in want.splitlines()) >= (0 if name == "sample" else 3)
        _close_lines(out, want, 1e-10, name)


def test_oracle_small_circuits_vs_statevector(golden):
    d = os.path.join(golden, "small")
    cases = json.load(open(os.path.join(d, "cases.json")))
    worst = 0.0
    for c in cases:
        i, n = c["i"], c["n"]
        _, ops = egs.tableau.parse(os.path.join(d, f"s{i}.qasm"))
        p = dense.probabilities(n, ops)
        _, _, steps = egs.planio.read_inplace_plan(os.path.join(d, f"s{i}_plan.txt"))
        sv = [s["v"] for s in steps]
        strings = egs.planio.read_strings(os.path.join(d, f"s{i}_strings.txt"), n)
        rc, out = oracle([os.path.join(d, f"s{i}_plan.txt"), "prob", os.path.join(d, f"s{i}_strings.txt")], True)
        assert rc == 0
        got = np.exp2([float(l) for l in out.split()])
        want = np.array([p[dense.step_order_index(b, sv, n)] for b in strings])
        worst = max(worst, float(np.max(np.abs(got - want))))
    assert worst < 1e-12


def test_oracle_exit_codes(golden, tmp_path):
    toks = open(os.path.join(golden, "c1", "plan.txt")).read().split()
    s = tmp_path / "strings.txt"
    s.write_text("0" * 24 + "\n")
    nk41 = "1 1 1 " + "0 41 2 1 0 2 " + " ".join(["1"] * 41) + " 0 0 0 0 0 1 1 0 " + " ".join(["0.5"] * 8)
    cases = {2: "", 3: " ".join(toks[:3]), 7: nk41}
    for code, text in cases.items():
        p = tmp_path / f"bad{code}.txt"
        p.write_text(text)
        rc, _ = oracle([str(p), "prob", str(s)])
        assert rc == code, (code, rc)
    rc, _ = oracle([os.path.join(golden, "c1", "plan.txt")])
    assert rc == 1
    q = tmp_path / "q.txt"
    q.write_text("?" + "0" * 23 + "\n")
    rc, _ = oracle([os.path.join(golden, "c1", "plan.txt"), "prefix", str(q), "4", str(tmp_path / "x")])
    assert rc == 8


def test_reference_base_of_unit_test():
    exe = os.path.join(ROOT, "oracle", "_ref", "test_baseof")
    if not os.path.exists(exe):
        pytest.skip("reference executor not built here")
    res = subprocess.run([exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout
