"""Graph form of a stabilizer state under local Cliffords.

``graph_form(t)`` returns ``(A, U)``: A is the symmetric 0/1 adjacency matrix and U a list of per-qubit 2x2
unitaries with (U_0 x ... x U_{n-1}) |S> = e^{i phi} |G_A>, where
|G_A> = 2^{-n/2} sum_x (-1)^{sum_{i<j} A_ij x_i x_j} |x>.  Follows graphform.py:21-50 of the reference:

1. greedy Hadamards until the X block of the tableau is invertible (graphform.py:27-32): up to three passes
   over the qubits; a Hadamard that does not raise the rank of the X block is undone at once, and both the
   Hadamard and its undo are kept in the gate list (they enter U as H @ H, which is not bit-exactly I);
2. row reduction of the X block to the identity with exact Pauli-product signs (graphform.py:34-40);
3. S^dagger on every qubit with a Y on the diagonal, then Z on every qubit whose sign is -1
   (graphform.py:42-45).  The Z block is then A (graphform.py:47) and U_q is the product of the gates applied
   to q, later gates on the left (graphform.py:48-49).

Only the greedy first stage has free choices; the canonical form after stage 2 (X block = I) is unique for
the stabilizer group, so stages 2-3 are fixed by the mathematics.  The matrix products go through numpy's
matmul with the same operand types as the reference (float H, complex S^dagger, integer Z), so that U, and
with it the plan's M doubles, agree to the last bit on the same numpy build.
"""
import numpy as np

from .tableau import gf2_rank

GATE_MATRICES = {
    "h": np.array([[1, 1], [1, -1]]) / np.sqrt(2),
    "sdg": np.diag([1, -1j]),
    "z": np.diag([1, -1]),
}


def _pauli_product_sign(xk, zk, rk, xi, zi, ri):
    """Sign bit of P_k * P_i for commuting Pauli rows given as Python-int bit masks.

    The phase exponent (in powers of i) of the single-qubit products is summed with the Aaronson-Gottesman
    function g; for commuting rows it is even, so the result is a sign."""
    y1 = xk & zk
    x1 = xk & ~zk
    z1 = ~xk & zk
    plus = (y1 & zi & ~xi) | (x1 & zi & xi) | (z1 & xi & ~zi)
    minus = (y1 & xi & ~zi) | (x1 & zi & ~xi) | (z1 & xi & zi)
    ph = (2 * rk + 2 * ri + bin(plus).count("1") - bin(minus).count("1")) % 4
    if ph not in (0, 2):
        raise AssertionError("anticommuting generators in a stabilizer tableau")
    return ph // 2


def _rows_as_ints(mat):
    packed = np.packbits(mat & 1, axis=1, bitorder="little")
    return [int.from_bytes(row.tobytes(), "little") for row in packed]


def _ints_as_rows(vals, n):
    out = np.zeros((len(vals), n), np.uint8)
    for i, v in enumerate(vals):
        for j in range(n):
            out[i, j] = (v >> j) & 1
    return out


def graph_form(t0):
    n = t0.n
    t = t0.copy()
    gates = []

    def do(name, q):
        gates.append((name, q))
        {"h": t.h, "sdg": t.sdg, "z": t.z}[name](q)

    # 1. greedy Hadamards
    for _ in range(3):
        for j in range(n):
            rank_now = gf2_rank(t.xs)
            if rank_now == n:
                break
            do("h", j)
            if gf2_rank(t.xs) <= rank_now:
                do("h", j)
    if gf2_rank(t.xs) != n:
        raise AssertionError("could not make the X block invertible")

    # 2. reduce the X block to the identity (rows as bit masks; signs exact)
    xs = _rows_as_ints(t.xs)
    zs = _rows_as_ints(t.zs)
    r = [int(b) for b in t.r]
    for c in range(n):
        bit = 1 << c
        p = next(i for i in range(c, n) if xs[i] & bit)
        if p != c:
            xs[c], xs[p] = xs[p], xs[c]
            zs[c], zs[p] = zs[p], zs[c]
            r[c], r[p] = r[p], r[c]
        for i in range(n):
            if i != c and xs[i] & bit:
                r[i] = _pauli_product_sign(xs[c], zs[c], r[c], xs[i], zs[i], r[i])
                xs[i] ^= xs[c]
                zs[i] ^= zs[c]
    t.xs = _ints_as_rows(xs, n)
    t.zs = _ints_as_rows(zs, n)
    t.r = np.array(r, np.uint8)

    # 3. clear the diagonal of the Z block, then the signs
    for i in range(n):
        if t.zs[i, i]:
            do("sdg", i)
    for i in range(n):
        if t.r[i]:
            do("z", i)
    if not ((t.xs == np.eye(n, dtype=np.uint8)).all() and not t.r.any() and not np.diag(t.zs).any()):
        raise AssertionError("graph form reduction failed")
    A = t.zs.astype(np.uint8)
    if not (A == A.T).all():
        raise AssertionError("graph form adjacency is not symmetric")
    U = [np.eye(2, dtype=complex) for _ in range(n)]
    for name, q in gates:
        U[q] = GATE_MATRICES[name] @ U[q]
    return A, U
