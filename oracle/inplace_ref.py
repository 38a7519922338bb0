"""Python in-place executor (TEST INFRASTRUCTURE ONLY: a checker used by tests/, never by the product).

Restates ``sample_chain_probs`` of the reference (inplace.py:102-141): the exact conditionals Pr(b_k | b_<k) along a
forced path on the planner's in-place layout, computing both branch norms explicitly, with the table renormalised after
every step.  Like the reference it asserts, per step, that no two cosets write the same slot and that no write lands on
another coset's input (inplace.py:126-137).  Vectorised over cosets with numpy; FP64 throughout.

``steps`` are the dicts of ``inplace.plan_inplace``; ``M[v]`` the 2x2 measurement matrices.
"""
import numpy as np


def _bases(cols, nk):
    """Physical base address of every coset c < 2^nk (XOR of cols over the bits of c)."""
    b = np.zeros(1 << nk, dtype=np.uint64)
    for j in range(nk):
        half = 1 << j
        b[half:2 * half] = b[:half] ^ np.uint64(cols[j])
    return b


def _parity(x):
    x = x.copy()
    out = np.zeros(x.shape, dtype=np.int64)
    while np.any(x):
        out ^= (x & np.uint64(1)).astype(np.int64)
        x >>= np.uint64(1)
    return out


def sample_chain_probs(steps, R, M, bits, check_races=True):
    mem = np.zeros(1 << R, complex)
    mem[0] = 1.0
    out = []
    for st, bv in zip(steps, bits):
        nk = len(st["keep"])
        base = _bases(st["Tkeep"], nk)
        c = np.arange(1 << nk, dtype=np.uint64)
        cl = np.zeros(1 << nk, dtype=np.uint64)        # logical old coordinates of the representative
        for j, i in enumerate(st["keep"]):
            cl |= ((c >> np.uint64(j)) & np.uint64(1)) << np.uint64(i)
        Mv = M[st["v"]]
        m = st["m"]
        vals = np.zeros((2, m, 1 << nk), complex)        # [b, tau, c]
        for qv, x, tau in st["terms"]:
            full = cl ^ np.uint64(qv)
            sgn = np.where((x == 1) & (_parity(full & np.uint64(st["rho"])) == 1), -1.0, 1.0)
            slot = mem[(base ^ np.uint64(st["Tq"][qv])).astype(np.int64)]
            for b in (0, 1):
                vals[b, tau] += Mv[b, x] * sgn * slot
        nrm = np.array([np.sum(np.abs(vals[b]) ** 2) for b in (0, 1)])
        p1 = nrm[1] / (nrm[0] + nrm[1])
        out.append(p1 if bv else 1 - p1)
        tgt = [base] + ([base ^ np.uint64(st["g"])] if m == 2 else [])
        if check_races:
            written = np.concatenate(tgt)
            if np.unique(written).size != written.size:
                raise AssertionError("two cosets write the same address")
            owner = {}
            for qv in st["Qall"]:
                for ci, a in enumerate(base ^ np.uint64(st["Tq"][qv])):
                    owner[int(a)] = ci
            for t in tgt:
                for ci, a in enumerate(t):
                    o = owner.get(int(a))
                    if o is not None and o != ci:
                        raise AssertionError("write would clobber another coset's input")
        scale = 1.0 / np.sqrt(nrm[bv]) if nrm[bv] > 0 else 0.0
        for tau, t in enumerate(tgt):
            mem[t.astype(np.int64)] = vals[bv, tau] * scale
    return out
