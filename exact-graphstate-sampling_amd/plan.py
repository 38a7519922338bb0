"""Cut-rank plan of Z(w) = sum_x prod_j w_j(x_j) (-1)^{sum_{i<j} A_ij x_i x_j} along a vertex order.

Follows plan.py:12-35 of the reference (the linear rank-width dynamic program of Van den Nest, Duer, Vidal
and Briegel).  After processing the prefix P of ``order`` the summed-out vertices act on the future vertices
F only through the field s = A[F, P] x_P, which lives in an r-dimensional GF(2) space (r = cut rank).  The
table is indexed by coordinates c in a basis B of that space; processing vertex v is the linear map
Lam: (c, x_v) -> c' whose generators are the old basis vectors with v's bit removed, followed by v's row
restricted to the unprocessed vertices (index r stands for x_v).

Per step the plan records ``v``, ``r`` (old rank), ``rn`` (new rank), ``rho`` (bit j set when old basis
vector j contains v: the sign (-1)^{x_v <rho, c>}), ``piv`` (generator indices that become the new basis)
and ``ker`` (kernel vectors of Lam as masks over the r+1 generators).

Vectors are Python ints (bit j = vertex j), so instances beyond 64 vertices work unchanged.
"""


def adjacency_rows(A):
    """Row i of A as a Python-int bit mask."""
    n = len(A)
    return [sum(int(A[i][j]) << j for j in range(n)) for i in range(n)]


class Eliminator:
    """Incremental GF(2) elimination with pivots at the top bit, tracking combinations.

    ``add(vec, comb)`` reduces ``vec`` against the stored vectors (largest pivot first) and either stores the
    remainder (returns None) or, when ``vec`` is already in the span, returns the combination mask of inputs
    that XOR to zero.  The stored remainder of a vector, and therefore every combination mask, is the unique
    normal form with respect to the pivot set, so the result depends only on the input sequence."""

    def __init__(self):
        self.rows = {}          # pivot bit -> (vector, combination)

    def add(self, vec, comb):
        for p in sorted(self.rows, reverse=True):
            if (vec >> p) & 1:
                pv, pc = self.rows[p]
                vec ^= pv
                comb ^= pc
        if vec:
            self.rows[vec.bit_length() - 1] = (vec, comb)
            return None
        return comb


def cut_generators(B, v, nbr_v, done):
    """Generators of Lam at vertex v: old basis vectors without v, then v's row on the unprocessed vertices."""
    vb = 1 << v
    return [b & ~vb for b in B] + [nbr_v & ~done]


def plan(A, order):
    nbr = adjacency_rows(A)
    B = []
    done = 0
    steps = []
    for v in order:
        r = len(B)
        rho = sum(1 << j for j, b in enumerate(B) if (b >> v) & 1)
        done |= 1 << v
        gens = cut_generators(B, v, nbr[v], done)
        elim = Eliminator()
        piv, ker = [], []
        for i, g in enumerate(gens):
            k = elim.add(g, 1 << i)
            if k is None:
                piv.append(i)
            else:
                ker.append(k)
        steps.append(dict(v=v, r=r, rn=len(piv), rho=rho, piv=piv, ker=ker))
        B = [gens[i] for i in piv]
    if B:
        raise AssertionError("state space must be empty at the end")
    return steps


def width(steps):
    """Largest table rank max_k max(r_k, rn_k) of a plan."""
    return max(max(s["r"], s["rn"]) for s in steps) if steps else 0
