"""gs_prob_noalloc (the prob entry point that allocates no device memory: every buffer lives in the library's fixed
64 MiB device array) against gs_prob on an engine of the same plan and storage, through the C ABI: bitwise equal
log2 Pr on two calls in a row, the second with a smaller input (fewer strings, a smaller plan), and again after a
zero-probability plan; a plan whose arena does not fit is rejected with code 16."""
import os

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("storage", [gs.C64, gs.C128])
def test_prob_noalloc_equals_prob(golden, storage):
    tag = "c128" if storage == gs.C128 else "c64"
    big = os.path.join(golden, "c4", "plan.txt")           # n = 60, R = 20
    small = os.path.join(golden, "c1", "plan.txt")         # n = 24, R = 10
    zero = os.path.join(golden, "zero", "plan.txt")        # a branch of probability 0 at step 5
    rng = np.random.default_rng(7)
    b_big = rng.integers(0, 2, size=(70, 60)).astype(np.int8)      # more than one chunk of 64 strings
    b_small = rng.integers(0, 2, size=(5, 24)).astype(np.int8)
    b_zero = egs.planio.read_strings(os.path.join(golden, "zero", "strings.txt"), 24)
    print(f"no-alloc prob storage={tag}: n=60 R=20 x 70 strings, then n=24 R=10 x 5, then the zero plan", flush=True)
    got = [gs.prob_noalloc(big, b_big, storage), gs.prob_noalloc(small, b_small, storage),
           gs.prob_noalloc(zero, b_zero, storage)]
    want = []
    for plan, b in ((big, b_big), (small, b_small), (zero, b_zero)):
        with gs.Engine(plan, storage) as e:
            want.append(e.prob(b))
    for g, w in zip(got, want):
        assert same(g, w), (g, w)
    assert np.isneginf(got[2][:4]).all()
    assert same(gs.prob_noalloc(small, b_small[:2], storage), want[1][:2])      # and once more, smaller still


def test_prob_noalloc_rejects_plans_beyond_the_buffer(golden):
    with pytest.raises(gs.GsError) as ex:
        gs.prob_noalloc(os.path.join(golden, "c2", "plan.txt"), np.zeros((1, 70), np.int8), gs.C64)
    assert ex.value.code == 16
