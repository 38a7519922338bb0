"""Writes tests/golden/refplans/: the reference planner's own in-place plans of the measured instances.

Run in the build container only, where the reference tree exists.  It imports the reference's host planner
(tableau.py, graphform.py, amp.py, inplace.py) from that tree -- nothing of it is copied into the repository -- and
feeds it each instance exactly as a reference user would: the instance's QASM (written by this package) through
clifford_body, Amplitudes(n, tableau, tail, order), plan_inplace(A, order, R) and write_inplace_plan.  The plan files
are data: the current reference planner's output (its kept coordinates sorted by physical column, inplace.py:50),
which no committed reference plan pins (those predate that line).  tests/test_planner.py compares this package's
default planner with them byte for byte.

Instances: c1 (n = 24, band 10, R = 10), c2 (n = 70, band 30, R = 30), c3 (n = 70, band 26), c4 (n = 60, band 24),
c5 (n = 80, band 34 + 40 phase-gadget ancillas, R = 34), and the n = 48, band-18 instance with 24 gadget ancillas.
"""
import importlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("GS_REFERENCE", "/root/reference")
sys.path.insert(0, os.path.join(ROOT, "tests"))
from _pkg import egs  # noqa: E402

CASES = {
    "c1": dict(config="c1"),
    "c2": dict(config="c2"),
    "c3": dict(config="c3"),
    "c4": dict(config="c4"),
    "c5": dict(config="c5"),
    "g48": dict(n=48, band=18, seed=0, gadgets_count=24),
}


def reference_modules():
    sys.path.insert(0, REF)
    try:
        tab = importlib.import_module("tableau")
        amp = importlib.import_module("amp")
        inp = importlib.import_module("inplace")
    finally:
        sys.path.pop(0)
    return tab, amp, inp


def main():
    tab, amp, inp = reference_modules()
    out = os.path.join(HERE, "refplans")
    os.makedirs(out, exist_ok=True)
    meta = {}
    for name, spec in CASES.items():
        inst = (egs.instances.config_instance(spec["config"]) if "config" in spec
                else egs.instances.Instance(**spec))
        qasm = os.path.join(out, f"{name}.qasm")
        inst.write_qasm(qasm)
        n, T, tail = tab.clifford_body(qasm)
        am = amp.Amplitudes(n, T, tail, list(inst.order))
        steps = inp.plan_inplace(am.A, list(inst.order), inst.R)
        inp.write_inplace_plan(os.path.join(out, f"{name}_plan.txt"), steps, inst.R, am.M, n)
        os.remove(qasm)
        meta[name] = dict(spec, R=inst.R, order=list(inst.order))
        print(f"{name}: N={n} R={inst.R} steps={len(steps)}")
    with open(os.path.join(out, "cases.json"), "w") as f:
        json.dump(meta, f)


if __name__ == "__main__":
    main()
