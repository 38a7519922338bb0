"""Sliced in-place plans (sliced_plans.py:26-86): widths beyond one arena, summed at a step K by prefix/resume.

Fixing x_v of a slice vertex v splits the Ising sum in two; each part is a sum on the graph with v's edges removed.
The weight column w_v(1 - x_v) is zeroed, and the edge phases (-1)^{x_v x_u} go into M_u[:, 1] of every prefix
neighbour u (steps < K).  All 2^|S| assignments therefore share one plan structure (the graph with S isolated) and
differ only in M, so their step-K tables add up (gsamp2.cpp:176-208).  A suffix neighbour u (step >= K) is summed
once over the added tables and cannot carry the phase.  Its phase (-1)^{x_v <s_v, x_suffix>} multiplies the partial
sum F(x_suffix) = sum_u T(u) (-1)^{<phi(u), x_suffix>}, which translates the step-K table by a with
sum_i a_i B_i = s_v (B = the cut basis entering step K).  A slice's translation is the XOR of the a_v of its set
slice vertices.  Split into the coset coordinate part (cshift) and the Q part (tqshift = the offset of the Q vector),
it forms one line of the twist file read by resume.
"""
import itertools
import json

import numpy as np

from .inplace import pext, plan_inplace, write_inplace_plan
from .plan import plan, width


def gf2_solve(basis, s):
    """Coefficients a (bit i for basis[i]) with XOR_i a_i basis[i] = s, or None when s is outside the span (Python
    ints as bit vectors, any length)."""
    piv = {}                                                    # pivot bit -> (vector, combination)
    for i, b in enumerate(basis):
        v, comb = int(b), 1 << i
        while v:
            p = v.bit_length() - 1
            if p not in piv:
                piv[p] = (v, comb)
                break
            v ^= piv[p][0]
            comb ^= piv[p][1]
    s, comb = int(s), 0
    while s:
        p = s.bit_length() - 1
        if p not in piv:
            return None
        s ^= piv[p][0]
        comb ^= piv[p][1]
    return comb


def isolate(A, S):
    A2 = np.array(A, dtype=np.uint8, copy=True)
    for v in S:
        A2[v, :] = 0
        A2[:, v] = 0
    return A2


def make_sliced_plans(A, M, order, S, out, K=None, nd=None, keep_order="physical"):
    """Write ``{out}_s{bits}.txt`` for every assignment of the slice vertices S (bits in the order of S), plus
    ``{out}_twist.txt`` when a slice vertex has suffix neighbours and ``{out}_slices.json``.  K defaults to the first
    step after which only data vertices (< nd) follow.  Returns dict(R, K, twist, step_vertices, plans, twists)."""
    A = np.asarray(A, dtype=np.uint8)
    N = len(A)
    nd = N if nd is None else nd
    S = [int(v) for v in S]
    A2 = isolate(A, S)
    R = max(1, width(plan(A2, order)))
    steps = plan_inplace(A2, order, R, keep_order=keep_order)
    sv = [s["v"] for s in steps]
    if K is None:
        K = next(i for i in range(len(sv)) if all(v < nd for v in sv[i:]))
        if K == 0:
            K = len(sv)
    prefix, suffix = set(sv[:K]), set(sv[K:])
    for v in S:
        if v not in prefix:
            raise ValueError(f"slice vertex {v} is not among the first K = {K} steps")
    avec = {}
    for v in S:
        s_v = sum(1 << int(u) for u in np.nonzero(A[v])[0] if int(u) in suffix)
        if s_v:
            a = gf2_solve(steps[K]["B"], s_v)
            if a is None:
                raise ValueError(f"slice vertex {v}: its suffix neighbourhood is outside the cut space at step {K}")
            avec[v] = a
    Sset = set(S)
    plans, twists = [], []
    for bits in itertools.product((0, 1), repeat=len(S)):
        Ms = [np.array(m, dtype=complex, copy=True) for m in M]
        a = 0
        for v, xv in zip(S, bits):
            Ms[v][:, 1 - xv] = 0
            if xv:
                for u in np.nonzero(A[v])[0]:
                    if int(u) in prefix and int(u) not in Sset:
                        Ms[int(u)][:, 1] *= -1
                a ^= avec.get(v, 0)
        for i in range(len(S)):                                 # an edge inside S: its phase once
            for j in range(i + 1, len(S)):
                if A[S[i], S[j]] and bits[i] and bits[j]:
                    Ms[S[j]][:, 1] *= -1
        path = f"{out}_s{''.join(map(str, bits))}.txt"
        write_inplace_plan(path, steps, R, Ms, N)
        plans.append(path)
        if avec:
            st = steps[K]
            keepmask = sum(1 << i for i in st["keep"])
            aq = next(q for q in st["Qall"] if (a ^ q) & ~keepmask == 0)
            ak = a ^ aq
            cshift = sum(((ak >> i) & 1) << j for j, i in enumerate(st["keep"]))
            twists.append((cshift, st["Tq"][aq]))
    if avec:
        with open(f"{out}_twist.txt", "w") as f:
            for cs, ts in twists:
                f.write(f"{cs} {ts}\n")
    with open(f"{out}_slices.json", "w") as f:
        json.dump({"slices": S, "R": R, "K_suffix_start": K, "twist": bool(avec), "step_vertices": sv}, f)
    return dict(R=R, K=K, twist=bool(avec), step_vertices=sv, plans=plans, twists=twists)
