"""Adjoint of log Pr(b) (hand-written backward sweep with checkpoint-and-recompute) against an FP64 dense-state
torch-autograd reference (n = 20, band 8; tests/golden/make_dense_grad.py) and Richardson-extrapolated central
differences (h = 1e-3, FP64 arena) on the n = 60, band 24 instance; segmented and unsegmented runs agree."""
import os
import subprocess
import sys

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_adjoint_matches_dense_autograd(golden, tmp_path):
    ref = np.load(os.path.join(golden, "c4dense", "grad_n20.npz"))
    inst = egs.instances.Instance(int(ref["n"]), int(ref["band"]), int(ref["seed"]))
    model = egs.adjoint.AngleModel(inst)
    assert np.array_equal(model.alpha0, ref["alpha"]) and np.array_equal(model.beta0, ref["beta"])
    plan = str(tmp_path / "p.txt")
    inst.write_plan(plan)
    bits = ref["bits"]
    print("adjoint n=20 band 8 vs dense autograd", flush=True)
    with gs.BatchEngine([plan], [64], storage=gs.C128) as bt:
        M, dA, dB = model.matrices(model.alpha0, model.beta0)
        bt.set_matrices(0, M)
        (res,), seg = bt.grad([bits])
        lp, Z, G = res
        ga, gb = model.grad_ln_pr(bits, Z, G, dA, dB)
    assert np.max(np.abs(lp * np.log(2) - ref["lnp"])) < 1e-9
    assert np.max(np.abs(-ga.mean(0) - ref["grad_alpha"])) < 1e-9
    assert np.max(np.abs(-gb.mean(0) - ref["grad_beta"])) < 1e-9
    assert np.max(np.abs(ref["grad_beta"])) > 1e-2                  # a non-trivial gradient


def test_adjoint_live_alpha_matches_dense_autograd(golden, tmp_path):
    """The tail rz(beta) h rz(alpha) h (AngleModel(last_h=True)): alpha acts before the measurement, so both angle
    sets carry gradients; the adjoint's chain rule through dM/dalpha and dM/dbeta against the FP64 dense-state torch
    autograd gradient (tests/golden/c4dense/grad_n20_live.npz) within 1e-9."""
    ref = np.load(os.path.join(golden, "c4dense", "grad_n20_live.npz"))
    inst = egs.instances.Instance(int(ref["n"]), int(ref["band"]), int(ref["seed"]))
    model = egs.adjoint.AngleModel(inst, last_h=True)
    plan = str(tmp_path / "p.txt")
    inst.write_plan(plan)
    bits = ref["bits"]
    print("adjoint n=20 band 8, live alpha, vs dense autograd", flush=True)
    with gs.BatchEngine([plan], [64], storage=gs.C128) as bt:
        M, dA, dB = model.matrices(model.alpha0, model.beta0)
        bt.set_matrices(0, M)
        (res,), seg = bt.grad([bits])
        lp, Z, G = res
        ga, gb = model.grad_ln_pr(bits, Z, G, dA, dB)
    assert np.max(np.abs(lp * np.log(2) - ref["lnp"])) < 1e-9
    assert np.max(np.abs(-ga.mean(0) - ref["grad_alpha"])) < 1e-9
    assert np.max(np.abs(-gb.mean(0) - ref["grad_beta"])) < 1e-9
    assert np.max(np.abs(ref["grad_alpha"])) > 1e-2 and np.max(np.abs(ref["grad_beta"])) > 1e-2


def test_adjoint_finite_differences_and_segments(tmp_path):
    inst = egs.instances.config_instance("c4")
    model = egs.adjoint.AngleModel(inst)
    plan = str(tmp_path / "c4.txt")
    inst.write_plan(plan)
    rng = np.random.default_rng(12)
    bits = rng.integers(0, 2, size=(64, len(inst.steps))).astype(np.int8)
    print(f"adjoint n=60 band 24 (width {inst.width}) FD check", flush=True)
    with gs.BatchEngine([plan], [64], storage=gs.C128) as bt:
        M, dA, dB = model.matrices(model.alpha0, model.beta0)
        bt.set_matrices(0, M)
        (res,), seg_full = bt.grad([bits])
        (res4,), seg4 = bt.grad([bits], seg=4)
        assert seg4 == 4 and seg_full == len(inst.steps)
        lp, Z, G = res
        assert np.array_equal(G, res4[2]) and np.array_equal(lp, res4[0])
        ga, gb = model.grad_ln_pr(bits, Z, G, dA, dB)
        loss_grad = -gb.mean(0)
        def loss_at(q, d):
            beta = model.beta0.astype(np.float64)          # the step in FP64 (the angles themselves are float32)
            beta[q] += d
            bt.set_matrices(0, model.matrices(model.alpha0, beta)[0])
            l, _, _ = bt.score([bits], want_cond=False, want_amp=False)
            return -np.mean(l[0]) * np.log(2)

        h = 1e-3
        worst = 0.0
        for q in (3, 17, 30, 44, 59):
            d1 = (loss_at(q, h) - loss_at(q, -h)) / (2 * h)
            d2 = (loss_at(q, h / 2) - loss_at(q, -h / 2)) / h
            fd = (4 * d2 - d1) / 3                                   # Richardson: O(h^4) truncation
            worst = max(worst, abs(fd - loss_grad[q]) / max(1e-3, abs(fd)))
        print(f"max relative |FD - adjoint| = {worst:.2e}")
    assert worst < 1e-7
    assert np.max(np.abs(ga)) < 1e-9


def test_torch_autograd_function(tmp_path):
    """LogProb.apply through torch autograd (in a child process: torch bundles its own HIP runtime) equals the
    direct adjoint contraction."""
    inst = egs.instances.Instance(16, 6, 5)
    plan = str(tmp_path / "p.txt")
    inst.write_plan(plan)
    code = f"""
import sys, numpy as np, torch
sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from _pkg import egs
gs = egs.gsamp
inst = egs.instances.Instance(16, 6, 5)
model = egs.adjoint.AngleModel(inst)
bits = np.random.default_rng(1).integers(0, 2, size=(64, 16)).astype(np.int8)
with gs.BatchEngine([{plan!r}], [64], storage=gs.C128) as bt:
    LogProb = egs.adjoint.make_log_prob(bt, model)
    a = torch.tensor(model.alpha0, requires_grad=True)
    b = torch.tensor(model.beta0, requires_grad=True)
    loss = -LogProb.apply(a, b, torch.tensor(bits)).mean()
    loss.backward()
    M, dA, dB = model.matrices(model.alpha0, model.beta0)
    bt.set_matrices(0, M)
    (res,), _ = bt.grad([bits])
    ga, gb = model.grad_ln_pr(bits, res[1], res[2], dA, dB)
    print(float(np.max(np.abs(b.grad.numpy() + gb.mean(0)))), float(np.max(np.abs(a.grad.numpy()))))
"""
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    d_beta, d_alpha = (float(x) for x in res.stdout.split()[-2:])
    assert d_beta < 1e-12 and d_alpha < 1e-9
