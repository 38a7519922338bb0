"""Sliced in-place plans: a dynamic program too wide for one arena, split over the assignments of a few vertices and
summed at a step K by prefix/resume (the reference's sliced_plans.py:26-86; gsamp2.cpp:162-208 sums the tables).

Slicing a set S fixes x_v for v in S.  Each edge (v, u) of the graph contributes (-1)^{x_v x_u} to the Ising sum; with
x_v fixed the edge becomes a local factor on u, so every assignment of S is a sum over the graph with the rows and
columns of S cleared, and all 2^|S| assignments share one plan structure (the same addresses, steps and sizes).  They
differ only in the measurement matrices, which carry the fixed values:

* v in S keeps only the live column x_v of M_v (the weight w_v(1 - x_v) is zero);
* an edge (v, u) with x_v = 1 and u processed before K (a prefix vertex outside S) negates column 1 of M_u, once per
  such edge;
* an edge inside S with both ends 1 is a constant sign, carried by the live column of its later end (in S order);
* an edge (v, u) with x_v = 1 and u at or after K cannot be folded into u: the suffix steps run once, on the summed
  table.  Its sign (-1)^{<s_v, x_suffix>} (s_v = v's suffix neighbourhood) multiplies the partial sum
  F(x_suffix) = sum_u T(u) (-1)^{<phi(u), x_suffix>} of the step-K table T, whose logical index u has the field
  phi(u) = sum_i u_i B_i over the cut basis B entering step K.  Multiplying by (-1)^{<s_v, .>} therefore translates
  the table by a_v, the coordinates of s_v in B; an assignment translates it by the XOR a of the a_v with x_v = 1.

The translation acts on the dump's two indices through step K's coordinate map: the old coordinates split into the
kept ones (the coset index c, bit j = coordinate keep[j]) and the span Q, whose basis is fully reduced on its pivot
coordinates.  The Q part of a is therefore read off a's pivot bits (the XOR of the basis vectors whose pivot is set),
and the rest of a lies on the kept coordinates: cshift = a's kept bits in keep order, tqshift = the physical offset
of the Q part.  One line ``cshift tqshift`` per assignment forms the twist file resume reads.
"""
import itertools
import json

import numpy as np

from .inplace import plan_inplace, write_inplace_plan
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
    """The graph with every edge at a vertex of S removed."""
    A2 = np.array(A, dtype=np.uint8, copy=True)
    idx = list(S)
    A2[idx, :] = 0
    A2[:, idx] = 0
    return A2


def assignments(S):
    """Assignments of the slice vertices, first vertex of S as the most significant digit (the file tag order)."""
    return list(itertools.product((0, 1), repeat=len(S)))


def fold_counts(A, S, bits):
    """For one assignment: ``(negations, inner)`` with negations[u] = the number of slice neighbours v of u (u not in
    S) with x_v = 1, and inner[j] = the number of slice vertices S[i], i < j, adjacent to S[j] with x_i = x_j = 1."""
    A = np.asarray(A)
    on = [v for v, x in zip(S, bits) if x]
    sset = set(S)
    neg = {}
    for v in on:
        for u in np.flatnonzero(A[v]):
            u = int(u)
            if u not in sset:
                neg[u] = neg.get(u, 0) + 1
    inner = [sum(1 for i in range(j) if bits[i] and bits[j] and A[S[i], S[j]]) for j in range(len(S))]
    return neg, inner


def _negate_column1(m, times):
    """Column 1 of a 2x2 complex matrix multiplied by -1, ``times`` times (complex multiplications, so that the
    signs of zero parts come out as the plan file prints them)."""
    for _ in range(times):
        m[:, 1] *= -1


def slice_matrices(A, M, S, bits, prefix):
    """Measurement matrices of one slice assignment (see the module docstring); ``prefix`` is the set of vertices
    processed before K."""
    Ms = [np.array(m, dtype=complex, copy=True) for m in M]
    for v, x in zip(S, bits):
        Ms[v][:, 1 - x] = 0
    neg, inner = fold_counts(A, S, bits)
    for u, times in neg.items():
        if u in prefix:
            _negate_column1(Ms[u], times)
    for j, times in enumerate(inner):
        _negate_column1(Ms[S[j]], times)
    return Ms


def twist_line(step, a):
    """(cshift, tqshift) of a translation by the logical vector a of step ``step``'s old coordinates."""
    aq = 0
    for b, p in step["Qb"]:
        if (a >> p) & 1:
            aq ^= b
    ak = a ^ aq
    keep = step["keep"]
    if ak & ~sum(1 << i for i in keep):
        raise AssertionError("translation outside the kept coordinates and Q")
    cshift = sum(((ak >> i) & 1) << j for j, i in enumerate(keep))
    return cshift, step["Tq"][aq]


def make_sliced_plans(A, M, order, S, out, K=None, nd=None, keep_order="physical"):
    """Write ``{out}_s{bits}.txt`` for every assignment of the slice vertices S (bits in the order of S), plus
    ``{out}_twist.txt`` when a slice vertex has neighbours at or after step K, and ``{out}_slices.json``.  K defaults
    to the first step after which only data vertices (< nd) follow.  Returns dict(R, K, twist, step_vertices, plans,
    twists)."""
    A = np.asarray(A, dtype=np.uint8)
    N = len(A)
    nd = N if nd is None else nd
    S = [int(v) for v in S]
    A2 = isolate(A, S)
    R = max(1, width(plan(A2, order)))
    steps = plan_inplace(A2, order, R, keep_order=keep_order)
    sv = [s["v"] for s in steps]
    if K is None:
        K = next((i for i in range(len(sv)) if all(v < nd for v in sv[i:])), len(sv))
        K = K or len(sv)
    prefix, suffix = set(sv[:K]), set(sv[K:])
    missing = [v for v in S if v not in prefix]
    if missing:
        raise ValueError(f"slice vertices {missing} are not among the first K = {K} steps")
    # per slice vertex with suffix neighbours: its translation a_v in the cut basis entering step K
    shift = {}
    for v in S:
        s_v = 0
        for u in np.flatnonzero(A[v]):
            if int(u) in suffix:
                s_v |= 1 << int(u)
        if s_v:
            a = gf2_solve(steps[K]["B"], s_v)
            if a is None:
                raise ValueError(f"slice vertex {v}: its suffix neighbourhood is outside the cut space at step {K}")
            shift[v] = a
    plans, twists = [], []
    for bits in assignments(S):
        path = f"{out}_s{''.join(map(str, bits))}.txt"
        write_inplace_plan(path, steps, R, slice_matrices(A, M, S, bits, prefix), N)
        plans.append(path)
        if shift:
            a = 0
            for v, x in zip(S, bits):
                if x:
                    a ^= shift.get(v, 0)
            twists.append(twist_line(steps[K], a))
    if shift:
        with open(f"{out}_twist.txt", "w") as f:
            f.writelines(f"{c} {t}\n" for c, t in twists)
    with open(f"{out}_slices.json", "w") as f:
        json.dump({"slices": S, "R": R, "K_suffix_start": K, "twist": bool(shift), "step_vertices": sv}, f)
    return dict(R=R, K=K, twist=bool(shift), step_vertices=sv, plans=plans, twists=twists)
