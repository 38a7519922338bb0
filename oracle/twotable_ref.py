"""CPU restatement of the reference's two-table (gather-form) executors, for small plans (test infrastructure only:
the tests compare the HIP engine with it; nothing in the product imports it).

gsamp.cpp:21-29  gather: output cp of the new table sums the inputs old[c], u = pdep(cp, pivmask) ^ kspan[t],
                 c = u & ((1 << r) - 1); x = bit r of u; S0 collects x = 0 terms, S1 the x = 1 terms with sign
                 (-1)^{parity(c & rho)}.
gsamp.cpp:39-55  chain (prob mode, forced bits): per step both branch norms, p1 = n1 / (n0 + n1), log2 p_b, write the
                 chosen branch divided by sqrt(n_b).
gsamp.cpp:59-70  chain1 (chain mode): one pass per step, cond = nb / (2 prev), log2 Pr = log2(prev) - ns.
gsamp.cpp:73-91  run_steps (forced bits): per step scale by 1/sqrt(prev), cond = nb / 2, logscale += log2(prev) / 2.
gsamp.cpp:124-148 prefix / resume: the table read by step K with its logscale; resume sums slice tables with weights
                 2^(ls - max ls) and continues from prev = norm, logscale = max ls.
dp.cpp:26-62     dp: per slice, every step with per-vertex weights (W0, W1); the job sums pre * table[0] over slices.

Arithmetic is complex128 (numpy); ``store`` rounds each written table to complex64 as the reference's cf tables do
(dp's default tables are complex128).
"""
import numpy as np


def load_plan(path, with_m=True):
    """gsamp plan (with M, gsamp.cpp:96-107) or dp plan (dp.cpp:16-25): (n, maxr, steps)."""
    toks = open(path).read().split()
    n, ns, maxr = int(toks[0]), int(toks[1]), int(toks[2])
    i = 3
    steps = []
    for _ in range(ns):
        v, r, rn, rho, piv, nk = (int(x) for x in toks[i:i + 6])
        i += 6
        ker = [int(x) for x in toks[i:i + nk]]
        i += nk
        kspan = [0]
        for k in ker:
            kspan += [x ^ k for x in kspan]
        M = None
        if with_m:
            vals = [float(x) for x in toks[i:i + 8]]
            i += 8
            M = np.array([[vals[0] + 1j * vals[1], vals[2] + 1j * vals[3]], [vals[4] + 1j * vals[5], vals[6] + 1j * vals[7]]])
        steps.append(dict(v=v, r=r, rn=rn, rho=rho, piv=piv, kspan=kspan, M=M))
    return n, maxr, steps


def _pdep(cp, mask):
    """Vectorised pdep: bit j of cp goes to the j-th lowest set bit of mask."""
    out = np.zeros_like(cp)
    j = 0
    for b in range(64):
        if (mask >> b) & 1:
            out |= ((cp >> j) & 1) << b
            j += 1
    return out


def _parity(x):
    x = x.copy()
    p = np.zeros_like(x)
    while np.any(x):
        p ^= x & 1
        x >>= 1
    return p


def gather(s, old):
    """(S0, S1) over the 2^rn outputs of step s from the table old (gsamp.cpp:21-29)."""
    cp = np.arange(1 << s["rn"], dtype=np.int64)
    u0 = _pdep(cp, s["piv"])
    S0 = np.zeros(cp.size, complex)
    S1 = np.zeros(cp.size, complex)
    low = (1 << s["r"]) - 1
    for k in s["kspan"]:
        u = u0 ^ k
        c = u & low
        o = old[c].astype(complex)
        x1 = ((u >> s["r"]) & 1).astype(bool)
        sg = np.where(_parity(c & s["rho"]) == 1, -1.0, 1.0)
        S1 += np.where(x1, sg * o, 0)
        S0 += np.where(x1, 0, o)
    return S0, S1


def _store(v, c64):
    return v.astype(np.complex64) if c64 else v


def prob(steps, bits, c64=True):
    """log2 Pr of a forced string by the two-pass chain (gsamp.cpp:39-55)."""
    old = np.ones(1, complex)
    logp = 0.0
    for s, b in zip(steps, bits):
        S0, S1 = gather(s, old)
        M = s["M"]
        n0 = float(np.sum(np.abs(M[0, 0] * S0 + M[0, 1] * S1) ** 2))
        n1 = float(np.sum(np.abs(M[1, 0] * S0 + M[1, 1] * S1) ** 2))
        p1 = n1 / (n0 + n1)
        pb = p1 if b else 1 - p1
        if pb <= 0:
            return -np.inf
        logp += np.log2(pb)
        sc = 1 / np.sqrt(n1 if b else n0)
        old = _store((M[b, 0] * sc) * S0 + (M[b, 1] * sc) * S1, c64)
    return logp


def chain1(steps, bits, c64=True):
    """(log2 Pr, conditionals) of a forced string, one pass per step (gsamp.cpp:59-70)."""
    old = np.ones(1, complex)
    prev = 1.0
    cond = np.zeros(len(steps))
    for k, (s, b) in enumerate(zip(steps, bits)):
        S0, S1 = gather(s, old)
        o = s["M"][b, 0] * S0 + s["M"][b, 1] * S1
        nb = float(np.sum(np.abs(o) ** 2))
        old = _store(o, c64)
        cond[k] = nb / (2 * prev)
        prev = nb
        if nb <= 0:
            return -np.inf, cond
    return np.log2(prev) - len(steps), cond


def run_steps(steps, bits, k0, k1, old, prev, logscale, c64=True):
    """Forced steps k0..k1-1 from table ``old`` of norm ``prev`` (gsamp.cpp:73-91):
    (table, prev, logscale, log2 Pr of the steps, cond)."""
    logp = 0.0
    cond = np.zeros(len(steps))
    for k in range(k0, k1):
        s, b = steps[k], int(bits[k])
        sc = 1 / np.sqrt(prev)
        S0, S1 = gather(s, old)
        o = (s["M"][b, 0] * sc) * S0 + (s["M"][b, 1] * sc) * S1
        nb = float(np.sum(np.abs(o) ** 2))
        old = _store(o, c64)
        logscale += 0.5 * np.log2(prev)
        cond[k] = nb / 2
        prev = nb
        if nb <= 0:
            return old, prev, logscale, -np.inf, cond
        logp += np.log2(cond[k])
    return old, prev, logscale, logp, cond


def prefix(steps, bits, K, c64=True):
    """(table read by step K, logscale, log2 Pr(prefix)) after forced steps 0..K-1 (gsamp.cpp:124-135)."""
    tab, _, ls, lp, _ = run_steps(steps, bits, 0, K, np.ones(1, complex), 1.0, 0.0, c64)
    return tab.astype(complex), ls, lp


def resume(steps, bits, K, tables, logscales, c64=True):
    """(log2 Pr(full), log2 Pr(prefix), cond) of the forced string continued from the summed slice tables
    (gsamp.cpp:136-148)."""
    L = max(logscales)
    acc = sum(np.asarray(t, complex) * 2.0 ** (ls - L) for t, ls in zip(tables, logscales))
    norm = float(np.sum(np.abs(acc) ** 2))
    lpp = -K + np.log2(norm) + 2 * L
    _, _, _, lps, cond = run_steps(steps, bits, K, len(steps), _store(acc, c64), norm, L, c64)
    return lpp + lps, lpp, cond


def dp(steps, jobs, c64=False):
    """sum over each job's slices of pre * Z(w) (dp.cpp:26-62); jobs: [[(pre, [(w0, w1)] per vertex)] per job]."""
    out = []
    for job in jobs:
        tot = 0j
        for pre, w in job:
            old = np.ones(1, complex)
            for s in steps:
                S0, S1 = gather(s, old)
                W0, W1 = w[s["v"]]
                old = _store(W0 * S0 + W1 * S1, c64)
            tot += pre * complex(old[0])
        out.append(tot)
    return np.array(out)


def read_jobs(path, m):
    """dp jobs file (dp.cpp:3, engine.py:38-45)."""
    toks = open(path).read().split()
    i = 0
    nj = int(toks[i]); i += 1
    jobs = []
    for _ in range(nj):
        nsl = int(toks[i]); i += 1
        job = []
        for _ in range(nsl):
            pre = float(toks[i]) + 1j * float(toks[i + 1]); i += 2
            w = []
            for _ in range(m):
                a = [float(x) for x in toks[i:i + 4]]; i += 4
                w.append((a[0] + 1j * a[1], a[2] + 1j * a[3]))
            job.append((pre, w))
        jobs.append(job)
    return jobs
