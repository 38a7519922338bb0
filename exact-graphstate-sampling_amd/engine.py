"""Writers for the two-table plan formats and vertex slicing (engine.py:5-64 of the reference).

* ``write_plan`` -- plan of the complex-amplitude engine: header ``m ns maxr``, per step
  ``v r rn rho pivmask nker ker...`` (engine.py:5-12).
* ``write_sampler_plan`` -- the same records followed by the 8 doubles of M_v (two-table sampler plan,
  engine.py:55-64).
* ``Sliced`` -- Z(w) with the vertices S summed outside the dynamic program: the plan covers the graph without S
  and ``slices(w)`` yields, per assignment of S, the prefactor prod_{v in S} w_v(x_v) (-1)^{edges inside S} and
  the kept vertices' weights with w_q(1) negated when q has an odd number of sliced neighbours set to 1
  (engine.py:14-36).
* ``write_jobs`` -- job file of weight sets per slice (engine.py:38-45).

The reference's ``run_dp`` (engine.py:47-53) runs a binary inside a container; it has no counterpart here.
"""
import numpy as np

from .plan import plan, width
from .sliced import assignments, fold_counts


def _pivmask(s):
    return sum(1 << p for p in s["piv"])


def _record(s):
    return f"{s['v']} {s['r']} {s['rn']} {s['rho']} {_pivmask(s)} {len(s['ker'])} " + " ".join(map(str, s["ker"]))


def _mtext(M):
    return " ".join("%.17g %.17g" % (complex(M[b, x]).real, complex(M[b, x]).imag) for b in (0, 1) for x in (0, 1))


def write_plan(path, steps, m):
    """Plan file of the amplitude engine; returns maxr."""
    maxr = width(steps)
    with open(path, "w") as f:
        f.write(f"{m} {len(steps)} {maxr}\n")
        for s in steps:
            f.write(_record(s) + "\n")
    return maxr


def write_sampler_plan(path, am):
    """Two-table sampler plan with per-step measurement matrices M_v; returns maxr."""
    maxr = width(am.steps)
    with open(path, "w") as f:
        f.write(f"{am.n} {len(am.steps)} {maxr}\n")
        for s in am.steps:
            f.write(_record(s) + " " + _mtext(am.M[s["v"]]) + "\n")
    return maxr


class Sliced:
    """Z(w) for graph A with the vertices in S summed outside the dynamic program (graph A - S inside)."""

    def __init__(self, A, order, S):
        self.S = list(S)
        sset = set(self.S)
        self.keep = [q for q in order if q not in sset]
        self.idx = {q: i for i, q in enumerate(self.keep)}
        self.m = len(self.keep)
        self.A = A
        sub = np.asarray(A)[np.ix_(self.keep, self.keep)].astype(np.uint8)
        self.steps = plan(sub, list(range(self.m)))

    def slices(self, w):
        """Yield ``(prefactor, weights of the kept vertices)`` for every assignment of the sliced vertices: the
        prefactor is prod_{v in S} w_v(x_v) times the sign of the edges inside S with both ends 1, and a kept vertex
        with an odd number of sliced neighbours set to 1 has w_q(1) negated (the edge phases (-1)^{x_v x_q})."""
        for xs in assignments(self.S):
            neg, inner = fold_counts(self.A, self.S, xs)
            pre = 1 + 0j
            for v, xv in zip(self.S, xs):
                pre *= w[v][xv]
            if sum(inner) & 1:
                pre = -pre
            yield pre, [(w[q][0], -w[q][1] if neg.get(q, 0) & 1 else w[q][1]) for q in self.keep]


def write_jobs(path, sl, ws):
    """Job file: ``njobs``; per job ``nslices``; per slice ``pre_re pre_im`` then one weight line per kept vertex."""
    with open(path, "w") as f:
        f.write(f"{len(ws)}\n")
        for w in ws:
            sls = list(sl.slices(w))
            f.write(f"{len(sls)}\n")
            for pre, ww in sls:
                f.write("%.17g %.17g\n" % (pre.real, pre.imag))
                for a, b in ww:
                    f.write("%.17g %.17g %.17g %.17g\n" % (a.real, a.imag, b.real, b.imag))
