"""T-gadget form of a doped Clifford circuit (doped_to_qasm.py:22-34 of the reference).

A mid-circuit rz(pi/4) (or t) on qubit q becomes a CNOT from q onto a fresh ancilla a -- ``h a; cz q,a; h a`` --
in the Clifford body, and the ancilla's tail is ``t a; h a``: projecting a onto outcome 0 after T then H applies
T to q (up to normalisation and a global phase).  Ancillas are numbered n, n+1, ... in gate order.  Every gate
after the last Clifford or T gate stays in the data tail.
"""
CLIFFORD = {"h", "s", "sdg", "sx", "sxdg", "x", "z", "cz"}
T_GATES = ("rz(pi/4)", "t")


def convert(n, ops):
    """Return ``(N, body ops, tail ops, ancillas [(a, q, gate index)])``."""
    last = max(i for i, (g, _) in enumerate(ops) if g in CLIFFORD or g in T_GATES)
    body, anc = [], []
    for i, (g, qs) in enumerate(ops[:last + 1]):
        if g in T_GATES:
            a = n + len(anc)
            anc.append((a, qs[0], i))
            body += [("h", [a]), ("cz", [qs[0], a]), ("h", [a])]
        elif g in CLIFFORD:
            body.append((g, qs))
        else:
            raise ValueError(f"unexpected mid-circuit gate {g}")
    tail = list(ops[last + 1:])
    for a, _, _ in anc:
        tail += [("t", [a]), ("h", [a])]
    return n + len(anc), body, tail, anc


def write_qasm(path, N, ops):
    """Flat OpenQASM 2 file with one register q[N]."""
    with open(path, "w") as f:
        f.write(f"OPENQASM 2.0;\ninclude \"qelib1.inc\";\nqreg q[{N}];\n")
        for g, qs in ops:
            f.write(g + " " + ",".join(f"q[{q}]" for q in qs) + ";\n")
