"""Stabilizer tableau of the Clifford body of a QASM circuit.

Follows the reference's ``tableau.py``: ``parse`` (tableau.py:9-19), ``Tableau`` with the gates
h, s, sdg, sx, sxdg, x, z, cz and exact sign tracking (tableau.py:21-37), ``clifford_body``
(tableau.py:39-49), ``gf2_rank`` (tableau.py:51-60) and ``cut_rank`` (tableau.py:62-66).

Row i of the tableau is the i-th stabilizer generator of C|0...0>: ``xs[i]``, ``zs[i]`` are its X and Z bit
rows (uint8, n x n) and ``r[i]`` its sign bit.  Gates act by conjugation, P -> G P G^dagger, with the
Aaronson-Gottesman column update rules; sx and sxdg are H S H and H S^dagger H (equal to sqrt(X) and its
inverse up to a global phase, which does not change stabilizers).
"""
import re

import numpy as np

CLIFFORD_GATES = ("h", "s", "sdg", "sx", "sxdg", "x", "z", "cz")
_QREG = re.compile(r"qreg\s+q\[(\d+)\]")
_QUBIT = re.compile(r"q\[(\d+)\]")


def parse(path):
    """Read a flat OpenQASM 2 file with one register ``q``.

    Returns ``(n, ops)`` with ``ops`` a list of ``(name, [qubits])``.  The gate name keeps its parameter text,
    e.g. ``"rz(0.955)"``, exactly as the reference parser does (tableau.py:9-19): a statement is split at its
    first space, the trailing ';' is dropped, header lines are skipped.
    """
    n = None
    ops = []
    with open(path) as fh:
        for raw in fh:
            stmt = raw.strip().rstrip(";")
            if not stmt or stmt.startswith("OPENQASM") or stmt.startswith("include"):
                continue
            reg = _QREG.match(stmt)
            if reg:
                n = int(reg.group(1))
                continue
            name, rest = stmt.split(" ", 1)
            ops.append((name, [int(q) for q in _QUBIT.findall(rest)]))
    return n, ops


class Tableau:
    """Stabilizer generators of an n-qubit state, starting from |0...0> (generators Z_i)."""

    def __init__(self, n):
        self.n = n
        self.xs = np.zeros((n, n), np.uint8)
        self.zs = np.eye(n, dtype=np.uint8)
        self.r = np.zeros(n, np.uint8)

    # single-qubit Clifford generators -------------------------------------------------------------
    def h(self, q):
        x = self.xs[:, q].copy()
        z = self.zs[:, q].copy()
        self.r ^= x & z
        self.xs[:, q] = z
        self.zs[:, q] = x

    def s_(self, q):
        x = self.xs[:, q]
        self.r ^= x & self.zs[:, q]
        self.zs[:, q] ^= x

    def z(self, q):
        self.r ^= self.xs[:, q]

    def x(self, q):
        self.r ^= self.zs[:, q]

    def sdg(self, q):
        # S^dagger = S Z: conjugate by Z first, then by S
        self.z(q)
        self.s_(q)

    def sx(self, q):
        self.h(q)
        self.s_(q)
        self.h(q)

    def sxdg(self, q):
        self.h(q)
        self.sdg(q)
        self.h(q)

    def cz(self, a, b):
        xa, xb = self.xs[:, a].copy(), self.xs[:, b].copy()
        self.r ^= xa & xb & (self.zs[:, a] ^ self.zs[:, b])
        self.zs[:, a] ^= xb
        self.zs[:, b] ^= xa

    def apply(self, name, qs):
        if name == "cz":
            self.cz(qs[0], qs[1])
            return
        fn = {"h": self.h, "s": self.s_, "sdg": self.sdg, "sx": self.sx, "sxdg": self.sxdg,
              "x": self.x, "z": self.z}.get(name)
        if fn is None:
            raise ValueError(f"not a Clifford gate: {name}")
        fn(qs[0])

    def copy(self):
        t = Tableau(self.n)
        t.xs, t.zs, t.r = self.xs.copy(), self.zs.copy(), self.r.copy()
        return t


def clifford_body(path):
    """``(n, tableau of the leading Clifford gates, remaining ops)``; the tail starts at the first gate that is
    not in the Clifford set (tableau.py:39-49)."""
    n, ops = parse(path)
    t = Tableau(n)
    cut = len(ops)
    for i, (name, qs) in enumerate(ops):
        if name not in CLIFFORD_GATES:
            cut = i
            break
        t.apply(name, qs)
    return n, t, ops[cut:]


def gf2_rank(mat):
    """Rank over GF(2) of a 0/1 matrix (tableau.py:51-60 computes the same number)."""
    return len(gf2_basis(bit_rows(mat)))


def bit_rows(mat):
    """Rows of a 0/1 matrix as Python ints (bit j = column j)."""
    packed = np.packbits(np.asarray(mat, dtype=np.uint8) & 1, axis=1, bitorder="little")
    return [int.from_bytes(row.tobytes(), "little") for row in packed]


def gf2_basis(vectors):
    """Echelon basis {top bit: vector} of the span of Python-int bit vectors."""
    pivots = {}
    for v in vectors:
        while v:
            top = v.bit_length() - 1
            p = pivots.get(top)
            if p is None:
                pivots[top] = v
                break
            v ^= p
    return pivots


def cut_rank(t, part):
    """Entanglement across ``part`` | rest of the stabilizer state: rank of the generator columns restricted to
    the complement B, minus |B| (tableau.py:62-66)."""
    comp = np.array(sorted(set(range(t.n)) - set(part)), dtype=int)
    g = np.concatenate([t.xs[:, comp], t.zs[:, comp]], axis=1)
    return gf2_rank(g) - len(comp)
