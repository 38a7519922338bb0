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
