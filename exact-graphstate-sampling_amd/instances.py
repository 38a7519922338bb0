"""Seeded banded random graph-state instances (the workloads the engine is measured on).

An instance is a circuit on n qubits:

* Clifford body: ``h`` on every qubit and ``cz`` on every edge (i, j), 0 < j - i <= band, kept with probability
  1/2 -- the banded random graph state |G> -- then one seeded element of the 24-element single-qubit Clifford
  group per qubit (as an h/s word), then ``h`` on every qubit;
* tail per qubit: ``rz(beta) h rz(alpha)``, so that with the body's last ``h`` the qubit is rotated by
  RZ(alpha) RX(beta) (RX(beta) = H RZ(beta) H), alpha and beta uniform in [0, 2 pi).

The last body layer of ``h`` makes every tail start with a non-Clifford gate, so ``clifford_body`` of the written
QASM splits body and tail exactly where this module does.  Optional phase-gadget ancillas (a T gate on the parity
of three random vertices, realised as an ancilla a with ``h a; cz q,a (x3); h a`` in the body and ``t a; h a`` in
its tail) model doped circuits: projecting every ancilla onto 0 applies the gadgets.  Gadgets attach to vertices
among the first ``band`` data vertices and the ancillas come first in the order, so every projected vertex precedes
the sampled data steps (the condition under which hybrid sampling of the data suffix is exact, gsamp2.cpp:71-74) and
the width stays at most ``band``: after the ancillas, every cut's future side touched by processed vertices lies in a
window of ``band`` data vertices.
"""
import numpy as np

from . import gadgets
from .amp import Amplitudes
from .inplace import plan_inplace, write_inplace_plan
from .plan import width
from .tableau import Tableau

_H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
_S = np.diag([1, 1j])


def _canon(u):
    """Key of a 2x2 unitary modulo global phase."""
    k = np.flatnonzero(np.abs(u.ravel()) > 1e-9)[0]
    v = u.ravel() * (abs(u.ravel()[k]) / u.ravel()[k])
    return tuple(np.round(v.real, 6)) + tuple(np.round(v.imag, 6))


def clifford_words():
    """The 24 single-qubit Cliffords modulo phase as shortest h/s words, in breadth-first order (h before s)."""
    seen = {_canon(np.eye(2)): []}
    frontier = [([], np.eye(2, dtype=complex))]
    words = [[]]
    while frontier:
        nxt = []
        for word, u in frontier:
            for g, mat in (("h", _H), ("s", _S)):
                w = mat @ u
                key = _canon(w)
                if key not in seen:
                    seen[key] = word + [g]
                    words.append(word + [g])
                    nxt.append((word + [g], w))
        frontier = nxt
    assert len(words) == 24
    return words


def banded_ops(n, band, seed, gadget_count=0, gadget_arity=3, fp32_angles=False):
    """Clifford body and tail of a seeded banded instance; returns ``(N, body, tail, ancillas)``.  With
    ``fp32_angles`` the rotation angles are rounded to single precision (the same draws, each an exact float32
    value), for workloads whose angles are float32 parameters."""
    rng = np.random.default_rng(seed)
    words = clifford_words()
    body = [("h", [q]) for q in range(n)]
    for i in range(n):
        for j in range(i + 1, min(n, i + band + 1)):
            if rng.random() < 0.5:
                body.append(("cz", [i, j]))
    cliff = rng.integers(0, 24, size=n)
    for q in range(n):
        body += [(g, [q]) for g in words[int(cliff[q])]]
    alpha = rng.uniform(0, 2 * np.pi, size=n)
    beta = rng.uniform(0, 2 * np.pi, size=n)
    if fp32_angles:
        alpha = alpha.astype(np.float32).astype(np.float64)
        beta = beta.astype(np.float32).astype(np.float64)
    anc = []
    for a in range(gadget_count):
        qs = sorted(int(q) for q in rng.choice(min(n, band), size=gadget_arity, replace=False))
        av = n + a
        body += [("h", [av])] + [("cz", [q, av]) for q in qs] + [("h", [av])]
        anc.append((av, qs))
    body += [("h", [q]) for q in range(n)]
    tail = []
    for q in range(n):
        tail += [(f"rz({float(beta[q])!r})", [q]), ("h", [q]), (f"rz({float(alpha[q])!r})", [q])]
    for av, _ in anc:
        tail += [("t", [av]), ("h", [av])]
    return n + gadget_count, body, tail, anc


def tableau_of(N, body):
    t = Tableau(N)
    for name, qs in body:
        t.apply(name, qs)
    return t


class Instance:
    """A banded instance with its graph form, measurement matrices and in-place plan.

    ``order`` defaults to the natural order (ancillas, if any, first: projected vertices must precede the sampled
    ones).  ``R`` defaults to the plan's width; a larger R sizes the arena as a workload asks."""

    def __init__(self, n, band, seed, R=None, gadgets_count=0, order=None, keep_order="physical", fp32_angles=False):
        self.n, self.band, self.seed = n, band, seed
        self.fp32_angles = fp32_angles
        self.N, self.body, self.tail, self.ancillas = banded_ops(n, band, seed, gadgets_count, fp32_angles=fp32_angles)
        self.order = list(order) if order is not None else (list(range(n, self.N)) + list(range(n)))
        self.am = Amplitudes(self.N, tableau_of(self.N, self.body), self.tail, self.order)
        self.width = width(self.am.steps)
        self.R = max(1, self.width if R is None else R)
        if self.R < self.width:
            raise ValueError(f"R = {self.R} is below the width {self.width}")
        self.steps = plan_inplace(self.am.A, self.order, self.R, keep_order=keep_order)

    @property
    def step_vertices(self):
        return [s["v"] for s in self.steps]

    def write_plan(self, path):
        write_inplace_plan(path, self.steps, self.R, self.am.M, self.N)

    def write_qasm(self, path):
        gadgets.write_qasm(path, self.N, self.body + self.tail)

    def log2_sum_live(self):
        """log2 of sum_k 2^{r_k}: live slots read by one pass over the plan."""
        return float(np.log2(sum(2.0 ** s["r"] for s in self.steps)))


# The measured workloads (seeds fixed here so every tool builds the same instance).
CONFIGS = {
    "c1": dict(n=24, band=10, seed=1, R=10),
    "c2": dict(n=70, band=30, seed=2, R=30),
    "c3": dict(n=70, band=26, seed=3),
    "c4": dict(n=60, band=24, seed=4, fp32_angles=True),
    "c5": dict(n=80, band=34, seed=5, R=34, gadgets_count=40),
}


def config_instance(name, **over):
    kw = dict(CONFIGS[name])
    kw.update(over)
    return Instance(**kw)
