"""In-place planner of the cut-rank dynamic program: one arena of 2^R slots (R >= largest cut rank).

Follows inplace.py:31-100 (``plan_inplace``) and inplace.py:143-159 (``write_inplace_plan``) of the reference.

Each step maps old coordinates (c, x_v) to new ones by the linear map Lam with kernel K.  Q is the span of the
old-coordinate parts of K; coset representatives have zero Q-pivot coordinates.  Every old coset c + Q
(2^q live slots) feeds m = 2^{q+1-dim K} new slots and nothing else feeds them, so each coset is a closed
butterfly that may overwrite its own slots.  Physical addresses are a GF(2)-linear map T of the logical
coordinates (the columns of T are R-bit integers).  The new coordinates are the kept (non-Q-pivot) old
coordinates, sorted by their physical column, then tau (which of the m outputs).  T_new keeps the old columns of
the kept coordinates and maps tau to g: T of the first Q basis vector when q >= 1, else the lowest physical bit
outside span(T) (a growth step).

The planner's choices (kernel discovery order, the Q basis and its order, keep order, representative/tau order,
g) are made exactly as the reference makes them, so the plan file is byte-identical; each choice is noted at
the line that makes it.  Every choice is a deterministic function of (A, order, R): kernel vectors and Q basis
vectors are unique normal forms with respect to their pivot sets, so any correct elimination that processes the
generators in the same order yields the same integers.
"""
from .plan import Eliminator, adjacency_rows, cut_generators


def rref(vectors):
    """Fully reduced echelon basis of span(vectors), as [(vector, pivot bit)] in the order the pivots first
    appear (inplace.py:14-24): each new vector is reduced by the basis, its top bit becomes a pivot and is
    cleared from the earlier basis vectors."""
    basis = []
    for v in vectors:
        for b, p in basis:
            if (v >> p) & 1:
                v ^= b
        if v:
            p = v.bit_length() - 1
            basis = [((b ^ v) if (b >> p) & 1 else b, bp) for b, bp in basis]
            basis.append((v, p))
    return basis


def in_span(v, basis):
    """Membership of v in the span of an ``rref`` basis (inplace.py:26-29)."""
    for b, p in basis:
        if (v >> p) & 1:
            v ^= b
    return v == 0


def span_list(basis_vectors):
    """All 2^k elements; element i is the XOR of the basis vectors selected by the bits of i."""
    out = [0]
    for b in basis_vectors:
        out += [a ^ b for a in out]
    return out


def apply_linear(cols, u):
    """XOR of cols[i] over the set bits i of u (the physical address of logical index u)."""
    a = 0
    i = 0
    while u:
        if u & 1:
            a ^= cols[i]
        u >>= 1
        i += 1
    return a


def plan_inplace(A, order, R, keep_order="physical"):
    """In-place plan of graph A along ``order`` in an arena of 2^R slots.

    ``keep_order="physical"`` sorts the kept coordinates by their physical column (inplace.py:50), so that
    consecutive coset indices c land on nearby addresses.  ``"logical"`` keeps them in coordinate order, which
    is what the planner did when ``cloud_ibmxeb/nq70_inplace_plan.txt`` was written: that committed plan is
    reproduced byte for byte only with ``keep_order="logical"``.  Both are valid layouts with identical
    results."""
    nbr = adjacency_rows(A)
    B, T = [], []
    done = 0
    steps = []
    for v in order:
        r = len(B)
        done |= 1 << v
        rho = sum(1 << j for j, b in enumerate(B) if (b >> v) & 1)
        gens = cut_generators(B, v, nbr[v], done)            # index r is x_v
        elim = Eliminator()
        K = [k for k in (elim.add(g, 1 << i) for i, g in enumerate(gens)) if k is not None]
        low = (1 << r) - 1
        Qb = rref([k & low for k in K if k & low])             # Q basis, in pivot-discovery order
        q = len(Qb)
        qpiv = {p for _, p in Qb}
        keep = [i for i in range(r) if i not in qpiv]          # coset-representative coordinates
        if keep_order == "physical":
            keep.sort(key=lambda i: T[i])                       # low physical columns first (inplace.py:50)
        elif keep_order != "logical":
            raise ValueError(f"keep_order must be 'physical' or 'logical', not {keep_order!r}")
        Qall = span_list([b for b, _ in Qb])
        Kset = set(span_list(K))
        reps, terms = [], []                                    # terms (qv, x, tau) in (qv, x) enumeration order
        for qv in Qall:
            for x in (0, 1):
                u = qv | (x << r)
                tau = next((ti for ti, rep in enumerate(reps) if (u ^ rep) in Kset), None)
                if tau is None:
                    reps.append(u)
                    tau = len(reps) - 1
                terms.append((qv, x, tau))
        m = len(reps)
        if m not in (1, 2) or reps[0] != 0:
            raise AssertionError(f"step of vertex {v}: {m} outputs per coset")

        def field(u):
            return apply_linear(gens, u)

        newB = [gens[i] for i in keep] + ([field(reps[1])] if m == 2 else [])
        if m == 2:
            if q >= 1:
                g = apply_linear(T, Qb[0][0])                   # second output: another slot of the same coset
            else:
                tb = rref(list(T))                              # growth: lowest physical bit outside span(T)
                g = next((1 << j for j in range(R) if not in_span(1 << j, tb)), None)
                if g is None:
                    raise AssertionError(f"arena 2^{R} too small at vertex {v}")
            newT = [T[i] for i in keep] + [g]
        else:
            g = None
            newT = [T[i] for i in keep]
        if not all(t < (1 << R) for t in newT) or len(rref(newT)) != len(newT):
            raise AssertionError("layout not injective")
        nbasis = rref(newB)
        if len(nbasis) != len(newB) or not all(in_span(field(1 << i), nbasis) for i in range(r + 1)):
            raise AssertionError("new basis does not span the field space")
        Tq = {qv: apply_linear(T, qv) for qv in Qall}
        steps.append(dict(v=v, r=r, rn=len(newB), keep=keep, Tkeep=[T[i] for i in keep], rho=rho, terms=terms,
                          Tq=Tq, m=m, g=g, B=list(B), Qall=Qall, Qb=list(Qb)))
        B, T = newB, newT
    if B:
        raise AssertionError("state space must be empty at the end")
    return steps


def pext(x, mask):
    """Bits of x at the set positions of mask, packed to the low end."""
    out = 0
    j = 0
    while mask:
        low = mask & -mask
        if x & low:
            out |= 1 << j
        j += 1
        mask ^= low
    return out


def step_record(s, Mv):
    """One plan-file record: ``v nk m g rhoc nterms``, Tkeep, (Tq x tau par) per term, M[b][x] as re im.

    rhoc carries the kept coordinates of rho (bit j of c is coordinate keep[j]); par = parity(qv & rho) is the
    Q part of the sign (inplace.py:154-158)."""
    rhoc = sum(((s["rho"] >> i) & 1) << j for j, i in enumerate(s["keep"]))
    head = f"{s['v']} {len(s['keep'])} {s['m']} {s['g'] or 0} {rhoc} {len(s['terms'])} "
    tk = " ".join(str(t) for t in s["Tkeep"]) + " "
    tt = " ".join(f"{s['Tq'][qv]} {x} {tau} {bin(qv & s['rho']).count('1') & 1}" for qv, x, tau in s["terms"]) + " "
    mm = " ".join("%.17g %.17g" % (complex(Mv[b, x]).real, complex(Mv[b, x]).imag) for b in (0, 1) for x in (0, 1))
    return head + tk + tt + mm + "\n"


def write_inplace_plan(path, steps, R, M, n):
    """Write the in-place plan file read by the engine (format: ``n ns R`` then one record per step)."""
    with open(path, "w") as f:
        f.write(f"{n} {len(steps)} {R}\n")
        for s in steps:
            f.write(step_record(s, M[s["v"]]))


def check_race_free(steps):
    """Host-side owner map of every step: asserts that no two cosets write the same slot and that no coset
    writes a slot another coset reads (the invariants inplace.py:126-137 asserts while executing).  Pure
    integer work over 2^nk cosets; meant for small plans.  Returns the number of slots checked."""
    checked = 0
    for s in steps:
        nk = len(s["keep"])
        offs = sorted(set(s["Tq"].values()))
        owner = {}
        for c in range(1 << nk):
            base = apply_linear(s["Tkeep"], c)
            for o in offs:
                a = base ^ o
                if a in owner:
                    raise AssertionError(f"step of vertex {s['v']}: slot {a} read by two cosets")
                owner[a] = c
        written = set()
        for c in range(1 << nk):
            base = apply_linear(s["Tkeep"], c)
            for a in ([base, base ^ s["g"]] if s["m"] == 2 else [base]):
                if a in written:
                    raise AssertionError(f"step of vertex {s['v']}: two cosets write slot {a}")
                written.add(a)
                if owner.get(a, c) != c:
                    raise AssertionError(f"step of vertex {s['v']}: write would clobber another coset's input")
                checked += 1
    return checked
