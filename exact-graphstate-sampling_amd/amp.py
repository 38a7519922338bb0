"""Per-vertex measurement matrices M_j of a circuit = Clifford body + single-qubit tail.

Follows amp.py:7-29 of the reference.  <b| (W_0 x ... x W_{n-1}) |S> with |S> = e^{-i phi} (x U_q^dagger)|G_A>
gives, per qubit, the 2x2 matrix M_q = W_q U_q^dagger, indexed M_q[b, x] (b = measured bit, x = graph-state
basis bit).  W_q multiplies the tail gates of qubit q with later gates on the left (amp.py:16); the tail may
hold h, t, tdg and rz(theta) = diag(e^{-i theta/2}, e^{i theta/2}) (amp.py:10-15).
"""
import re

import numpy as np

from .graphform import graph_form
from .plan import plan

_RZ = re.compile(r"rz\((.*)\)")
# e^{i pi/4} = (1 + i)/sqrt(2), both parts correctly rounded (0.7071067811865476).  This is the value the
# reference's committed plans carry; glibc's complex exp(i pi/4) returns an imaginary part one ulp lower, which
# would change the last bits of every M entry of a qubit with a t/tdg tail.
_EIGHTH_TURN = complex(np.sqrt(0.5), np.sqrt(0.5))


def gate_matrix(name):
    """2x2 matrix of one tail gate."""
    if name == "h":
        return np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    if name == "tdg":
        return np.diag([1, _EIGHTH_TURN.conjugate()])
    if name == "t":
        return np.diag([1, _EIGHTH_TURN])
    m = _RZ.match(name)
    if m:
        th = float(m.group(1))
        return np.diag([np.exp(-1j * th / 2), np.exp(1j * th / 2)])
    raise ValueError(f"unsupported tail gate {name}")


def tail_matrix(ops):
    """Product of the tail gates ``ops`` (time order), later gates on the left."""
    W = np.eye(2, dtype=complex)
    for name, _ in ops:
        W = gate_matrix(name) @ W
    return W


class Amplitudes:
    """Graph form, measurement matrices and the two-table plan of a circuit (amp.py:19-29).

    Attributes: ``n``, ``A`` (adjacency), ``W`` (tail matrices), ``M`` (measurement matrices, M[q][b, x]),
    ``order`` (vertex order, default 0..n-1) and ``steps`` (``plan(A, order)``)."""

    def __init__(self, n, T, tail, order=None):
        self.n = n
        self.A, U = graph_form(T)
        per_qubit = [[] for _ in range(n)]
        for name, qs in tail:
            if len(qs) != 1:
                raise AssertionError("tail must be single-qubit")
            per_qubit[qs[0]].append((name, qs))
        self.W = [tail_matrix(per_qubit[q]) for q in range(n)]
        self.M = [self.W[q] @ U[q].conj().T for q in range(n)]
        self.order = list(order) if order else list(range(n))
        self.steps = plan(self.A, self.order)

    def weights(self, b):
        """Per-vertex weights (M_q[b_q, 0], M_q[b_q, 1]) of the measured string b (qubit order)."""
        return [(self.M[q][b[q], 0], self.M[q][b[q], 1]) for q in range(self.n)]
