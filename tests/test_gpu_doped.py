"""Hybrid rejection driver on a small gadget instance (n = 10 data qubits, band 4, 3 phase-gadget ancillas), checked
against a dense statevector: every proposal's log2 Pr_doped and z (the prefix marginal in units of the uniform law)
are exact, the accept/reject decisions follow the reference's random.Random stream, and scoring matches."""
import os
import random

import numpy as np
import pytest

from _pkg import egs
import dense

gs = egs.gsamp
pytestmark = pytest.mark.gpu


def test_doped_sampler_vs_dense(tmp_path):
    inst = egs.instances.Instance(10, 4, 21, gadgets_count=3)
    nd, na = inst.n, inst.N - inst.n
    plan = str(tmp_path / "p.txt")
    inst.write_plan(plan)
    p = dense.probabilities(inst.N, inst.body + inst.tail).reshape([2] * inst.N)
    pd = p[(Ellipsis,) + (0,) * na] * 2.0 ** na                # doped law of the data qubits (ancillas last axes)
    assert abs(pd.sum() - 1) < 1e-12                            # Pr(all ancillas 0) = 2^-na exactly
    sv = inst.step_vertices
    m = 4
    suffix_q = sv[len(sv) - m:]
    out = str(tmp_path / "samples.txt")
    print("doped sampler n=10 band 4 + 3 gadgets", flush=True)
    with gs.Engine(plan, gs.C128) as e:
        ds = egs.doped.DopedSampler(e, sv, nd, na, m=m)
        acc, zmax = ds.sample(24, 3, 2.0, out)
        strings = ["".join(str(x) for x in np.random.default_rng(j).integers(0, 2, nd)) for j in range(5)]
        sc = ds.score(strings)
    rng = random.Random(3)
    props = [[rng.randrange(2) for _ in range(nd)] for _ in range(24)]
    u = [rng.random() for _ in range(24)]
    rows = [l.split() for l in open(out) if l.startswith(("ACC", "rej"))]
    assert len(rows) == 24 and sum(r[0] == "ACC" for r in rows) == acc
    for i, r in enumerate(rows):
        b = [int(c) for c in r[1]]
        assert all(b[q] == props[i][q] for q in range(nd) if q not in suffix_q)
        assert abs(float(r[2]) - np.log2(pd[tuple(b)])) < 1e-9
        idx = tuple(slice(None) if q in suffix_q else b[q] for q in range(nd))
        z = 2.0 ** (nd - m) * pd[idx].sum()
        zr = float(r[3][2:])
        assert abs(zr - z) < 1e-6 * max(1, z)
        assert (r[0] == "ACC") == (u[i] < zr / 2.0)
    for s, v in zip(strings, sc):
        assert abs(v - np.log2(pd[tuple(int(c) for c in s)])) < 1e-9
