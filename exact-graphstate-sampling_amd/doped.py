"""Hybrid rejection sampling and scoring of data strings on a gadget (doped) instance (doped_sampler.py:1-77), plus
the sampler's statistical certificate (hybrid_certificate.py:1-61), on the HIP engine.

Conventions (doped_sampler.py:1-16).  Step order is the plan's; ``step_vertices`` maps steps to vertices; vertices
< nd are data qubits, the others ancillas, always projected on 0.  With na ancillas Pr(all ancillas 0) = 2^-na
exactly, so log2 Pr_doped(b) = log2 Pr_gadget(b, anc = 0) + na.  The trailing m data steps form the sampled suffix
(m defaults to every trailing data step).  A proposal draws the other nd - m data bits uniformly; one hybrid pass
gives the exact marginal Pr(prefix, anc = 0), from the product of the conditionals of the first K = ns - m steps,
and an exact draw of the suffix.  The proposal is accepted with probability z / cap, with
z = 2^(nd - m) * 2^na * Pr(prefix, anc = 0).  Proposals and acceptance uniforms come from Python's
``random.Random(seed)`` in the reference's order: all proposal bits first, then one uniform per proposal.  Proposal i
seeds its hybrid pass with seed * 1000003 + i.  The output lines match the reference:
``ACC|rej <bits by qubit> <log2 Pr_doped> z=<z> <seconds>s [Z_ABOVE_CAP]``.
With ``slices`` (sliced plans sharing the structure, summed at step K, with an optional twist), every evaluation runs
prefix on each slice engine and resume on the sum (doped_sampler.py:35-40).
"""
import math
import random
import time

import numpy as np


class DopedSampler:
    def __init__(self, engine, step_vertices, nd, na, m=None, slice_engines=None, K=None, twist=None):
        self.e = engine
        self.sv = list(step_vertices)
        self.ns = len(self.sv)
        self.nd, self.na = nd, na
        self.pos = {v: i for i, v in enumerate(self.sv)}
        msuf = 0
        while msuf < self.ns and self.sv[self.ns - 1 - msuf] < nd:
            msuf += 1
        self.m = msuf if m is None else min(msuf, int(m))
        self.slices = slice_engines or []
        self.K = K
        self.twist = twist                                      # (cshift list, tqshift list) or None

    def step_string(self, bits_q, free_suffix):
        s = np.zeros(self.ns, np.int8)
        for q in range(self.nd):
            s[self.pos[q]] = bits_q[q]
        if free_suffix:
            s[self.ns - self.m:] = -1
        return s

    def _sliced(self, s, seed):
        tabs, lss = [], []
        for e in self.slices:
            tab, _, _, ls, _ = e.prefix(s, self.K)
            tabs.append(tab)
            lss.append(ls)
        cs, ts = self.twist if self.twist else (None, None)
        return self.slices[0].resume(s, self.K, tabs, lss, seed, 0, cs, ts)

    def sample(self, nprop, seed, cap, out):
        """Append ``nprop`` proposals to the file ``out``; returns (accepted, max z)."""
        rng = random.Random(seed)
        props = [[rng.randrange(2) for _ in range(self.nd)] for _ in range(nprop)]
        u = [rng.random() for _ in range(nprop)]
        K = self.ns - self.m
        zmax, acc = 0.0, 0
        with open(out, "a") as f:
            f.write(f"# seed {seed} cap {cap} m_suffix {self.m} K {K} nprop {nprop}\n")
            for i in range(nprop):
                t0 = time.time()
                s = self.step_string(props[i], True)
                hseed = seed * 1000003 + i
                if self.slices:
                    bits, lpf, lpp, _ = self._sliced(s, hseed)
                    lp_full, lp_prefix = lpf + self.na, lpp + self.na
                else:
                    ob, lp, cond = self.e.hybrid(s[None, :], hseed, 0)
                    bits, lp_full = ob[0], float(lp[0]) + self.na
                    c = cond[0, :K]
                    lp_prefix = float(np.sum(np.log2(c))) + self.na if np.all(c > 0) else -math.inf
                z = 2.0 ** (lp_prefix + (self.nd - self.m))
                zmax = max(zmax, z)
                ok = u[i] < z / cap
                b = "".join(str(int(bits[self.pos[q]])) for q in range(self.nd))
                f.write(f"{'ACC' if ok else 'rej'} {b} {lp_full:.9f} z={z:.6f} {time.time() - t0:.1f}s"
                        f"{' Z_ABOVE_CAP' if z > cap else ''}\n")
                f.flush()
                acc += ok
            f.write(f"# done: {acc} accepted of {nprop}; max z {zmax:.6f}; cap {cap}\n")
        return acc, zmax

    def score(self, strings, qiskit=False):
        """log2 Pr_doped of data strings (characters by qubit; ``qiskit``: character k is qubit nd-1-k)."""
        out = []
        for s in strings:
            b = [int(s[self.nd - 1 - q]) for q in range(self.nd)] if qiskit else [int(c) for c in s]
            st = self.step_string(b, False)
            if self.slices:
                out.append(self._sliced(st, 0)[1] + self.na)
            else:
                out.append(float(self.e.prob(st[None, :])[0]) + self.na)
        return out


def certificate_bound(ws, c, delta):
    """Lower confidence bounds on m = E_u[min(z, c)] from i.i.d. draws w = min(z, c) (hybrid_certificate.py:25-30):
    (empirical Bernstein, Maurer & Pontil 2009 Thm 4; Hoeffding; the sample mean), all in units of z."""
    n = len(ws)
    x = [w / c for w in ws]
    mu = sum(x) / n
    var = sum((t - mu) ** 2 for t in x) / (n - 1)
    L = math.log(2 / delta)
    bern = mu - math.sqrt(2 * var * L / n) - 7 * L / (3 * (n - 1))
    hoeff = mu - math.sqrt(math.log(1 / delta) / (2 * n))
    return c * bern, c * hoeff, c * mu


def certificate(zs, c, delta=0.05):
    """TV(sampled, ideal) <= 1 - m and classical fidelity >= m at confidence 1 - delta, with m bounded below from the
    proposals' z values (hybrid_certificate.py:1-16).  Returns dict(tv, fidelity, m_lower, m_hoeffding, m_mean)."""
    ws = [min(z, c) for z in zs]
    lo, hoeff, mean = certificate_bound(ws, c, delta)
    return dict(tv=max(0.0, 1 - lo), fidelity=min(1.0, max(0.0, lo)), m_lower=lo, m_hoeffding=hoeff, m_mean=mean,
                n=len(zs), above_cap=sum(z > c for z in zs), max_z=max(zs))


def read_z(path):
    """z values of every ACC/rej line of a sampler output file."""
    with open(path) as f:
        return [float(l.split("z=")[1].split()[0]) for l in f if l.startswith(("ACC", "rej"))]
