"""The C ABI library loads and exports every entry point include/gsamp_amd.h declares; plan parsing (host only)
returns the reference's exit codes.  No device calls."""
import os
import re
import subprocess

import pytest

from _pkg import egs

gs = egs.gsamp
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "gsamp_amd.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(gs_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_python_binding():
    assert sorted(gs.ABI_SYMBOLS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = gs.library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", gs.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (gs_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_library_has_gfx950_code_object():
    with open(gs.LIB_PATH, "rb") as f:
        assert b"gfx950" in f.read()


def test_plan_parse_matches_reference_exit_codes(golden):
    good = open(os.path.join(golden, "c1", "plan.txt")).read()
    p = gs.Plan(text=good)
    assert (p.n, p.ns, p.R) == (24, 24, 10)
    toks = good.split()
    # header only -> step head missing (3); head + truncated Tkeep (4); truncated term (5); truncated matrix (6)
    nk1 = int(toks[3 + 1])
    cases = {2: "", 3: " ".join(toks[:3])}
    for code, text in cases.items():
        with pytest.raises(gs.GsError) as ex:
            gs.Plan(text=text)
        assert ex.value.code == code
    # a complete step with nk = 41 -> 7
    nk41 = "1 1 1 " + "0 41 2 1 0 2 " + " ".join(["1"] * 41) + " 0 0 0 0 0 1 1 0 " + " ".join(["0.5"] * 8)
    with pytest.raises(gs.GsError) as ex:
        gs.Plan(text=nk41)
    assert ex.value.code == 7


def test_plan_parse_truncations_give_codes_4_5_6(golden):
    text = open(os.path.join(golden, "ibm", "nq70_inplace_plan.txt")).read()
    toks = text.split()
    # step 1 of the IBM plan: "53 1 2 2 1 2 | 1 | 0 0 0 0 0 1 1 0 | 8 doubles"; the first step starts at token 3
    first_len = 6 + int(toks[4]) + 4 * int(toks[8]) + 8
    s1 = 3 + first_len
    head = toks[:s1 + 6]
    with pytest.raises(gs.GsError) as ex:
        gs.Plan(text=" ".join(toks[:3]).replace(toks[1], toks[1]) + " " + " ".join(toks[3:s1 + 6]))
    assert ex.value.code == 4
    with pytest.raises(gs.GsError) as ex:
        gs.Plan(text=" ".join(toks[:s1 + 6 + 1 + 2]))
    assert ex.value.code == 5
    with pytest.raises(gs.GsError) as ex:
        gs.Plan(text=" ".join(toks[:s1 + 6 + 1 + 8 + 3]))
    assert ex.value.code == 6
    assert head


def test_ibm_and_doped_plans_parse(golden):
    p = gs.Plan(path=os.path.join(golden, "ibm", "nq70_inplace_plan.txt"))
    assert (p.n, p.ns, p.R) == (70, 70, 30)
    d = gs.Plan(path=os.path.join(golden, "doped", "doped_m20_inplace_plan.txt"))
    assert (d.n, d.ns, d.R) == (538, 538, 34)


def test_engine_without_device_fails_loudly(golden):
    """No HIP device in this container: creating an engine must raise, not fall back to the CPU."""
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a HIP device may be present")
    with pytest.raises(gs.GsError) as ex:
        gs.Engine(os.path.join(golden, "c1", "plan.txt"))
    assert ex.value.code in (16, 17)


def test_bindings_validate_shapes_before_any_call():
    """Mis-sized host arrays are rejected before the C ABI reads them (the ABI takes no lengths for them):
    BatchEngine.score/grad check one [lanes_i, ns_i] array per plan, Engine.resume the string length and K."""
    from types import SimpleNamespace
    import numpy as np
    bt = gs.BatchEngine.__new__(gs.BatchEngine)
    bt.plans, bt.lanes, bt._h = [SimpleNamespace(ns=5), SimpleNamespace(ns=7)], [64, 128], None
    good = [np.zeros((64, 5), np.int8), np.zeros((128, 7), np.int8)]
    for bad in ([good[0]], [good[0], np.zeros((7, 128), np.int8)], [np.zeros((64, 4), np.int8), good[1]]):
        with pytest.raises(ValueError):
            bt.grad(bad)
        with pytest.raises(ValueError):
            bt.score(bad)
    e = gs.Engine.__new__(gs.Engine)
    e.ns, e._h = 6, None
    with pytest.raises(ValueError):
        e.resume(np.zeros(5, np.int8), 2, [np.zeros(4)], [0.0], 1)
    with pytest.raises(ValueError):
        e.resume(np.zeros(6, np.int8), 7, [np.zeros(4)], [0.0], 1)
