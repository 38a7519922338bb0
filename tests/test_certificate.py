"""The hybrid sampler's certificate (hybrid_certificate.py:1-30): on exact small distributions the accepted-prefix
law's TV distance is at most eps = E_u[(z - c)^+] and its classical fidelity at least m = 1 - eps; the empirical
Bernstein bound covers the true mean in all but about delta of repetitions; no mass above the cap gives TV ~ 0."""
import math
import random

from _pkg import egs

cert = egs.doped


def _law(D, seed):
    rnd = random.Random(seed)
    z = [rnd.expovariate(1.0) for _ in range(D)]
    s = sum(z)
    return [t * D / s for t in z]                                # E_u z = 1


def test_tv_and_fidelity_inequalities():
    for c in (1.03, 1.5, 3.0):
        z = _law(4096, 1)
        w = [min(t, c) for t in z]
        m = sum(w) / len(z)
        tv = 0.5 * sum(abs(wi / m - zi) for wi, zi in zip(w, z)) / len(z)
        F = (sum(math.sqrt(wi / m * zi) for wi, zi in zip(w, z)) / len(z)) ** 2
        assert tv <= (1 - m) + 1e-12
        assert F >= m - 1e-12


def test_no_mass_above_cap():
    r = cert.certificate([1.0] * 100, 1.03, 0.05)
    assert r["tv"] < 0.2 and r["above_cap"] == 0


def test_bernstein_coverage():
    z = _law(1 << 14, 2)
    c = 1.5
    m = sum(min(t, c) for t in z) / len(z)
    rnd = random.Random(5)
    fails = sum(cert.certificate_bound([min(z[rnd.randrange(len(z))], c) for _ in range(102)], c, 0.05)[0] > m
                for _ in range(400))
    assert fails <= 30
