"""The gsamp2-compatible front end (bin/gsamp2_amd): same arguments, stdout lines, dump files and exit codes as the
reference executor (gsamp2.cpp:149-225), checked line by line against the reference's golden outputs."""
import json
import os
import subprocess

import numpy as np
import pytest

from _pkg import egs

gs = egs.gsamp
pytestmark = pytest.mark.gpu


def cli(args):
    res = subprocess.run([gs.CLI_PATH] + args, capture_output=True, text=True, timeout=120)
    return res.returncode, res.stdout, res.stderr


def compare(out, ref, tol):
    got = [l.split() for l in out.splitlines() if l.strip()]
    want = [l.split() for l in ref.splitlines() if l.strip()]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert len(g) == len(w)
        for a, b in zip(g, w):
            if set(b) <= set("01") and len(b) > 3:
                assert a == b
            else:
                assert abs(float(a) - float(b)) <= tol or a == b, (a, b)


@pytest.mark.parametrize("c128", [False, True])
def test_cli_modes_match_reference_lines(golden, c128, tmp_path):
    d = os.path.join(golden, "c1")
    meta = json.load(open(os.path.join(d, "meta.json")))
    tag = "c128" if c128 else "c64"
    tol = 1e-12 if c128 else 1e-5
    pre = ["--c128"] if c128 else []
    plan = os.path.join(d, "plan.txt")
    print(f"cli modes {tag}", flush=True)
    for mode, args in {"prob": [os.path.join(d, "strings.txt")], "chain": [os.path.join(d, "strings.txt")],
                       "sample": ["16", str(meta["sample_seed"])], "sample2": ["16", str(meta["sample_seed"])],
                       "hybrid": [os.path.join(d, "hybrid.txt"), str(meta["hybrid_seed"])]}.items():
        rc, out, err = cli(pre + [plan, mode] + args)
        assert rc == 0, err
        compare(out, open(os.path.join(d, f"{mode}_{tag}.txt")).read(), max(tol, 1e-9))
    dump = str(tmp_path / "dump")
    rc, out, err = cli(pre + [plan, "prefix", os.path.join(d, "hybrid.txt"), str(meta["K"]), dump])
    assert rc == 0, err
    compare(out, open(os.path.join(d, f"prefix_{tag}.txt")).read(), 1e-6)
    for li in range(8):
        a = egs.planio.read_dump(f"{dump}_{li}.bin", c128)
        b = egs.planio.read_dump(os.path.join(d, f"dump_{tag}_{li}.bin"), c128)
        assert os.path.getsize(f"{dump}_{li}.bin") == os.path.getsize(os.path.join(d, f"dump_{tag}_{li}.bin"))
        assert a[1:3] == b[1:3] and np.allclose(a[0], b[0], atol=10 * tol * np.abs(b[0]).max())
    rc, out, err = cli(pre + [plan, "resume", os.path.join(d, "hybrid.txt"), str(meta["K"]), str(meta["resume_seed"]), dump])
    assert rc == 0, err
    compare(out, open(os.path.join(d, f"resume_{tag}.txt")).read(), max(tol, 1e-9))


def test_cli_exit_codes(golden, tmp_path):
    s = tmp_path / "s.txt"
    s.write_text("0" * 24 + "\n")
    toks = open(os.path.join(golden, "c1", "plan.txt")).read().split()
    for code, text in {2: "", 3: " ".join(toks[:3])}.items():
        p = tmp_path / f"p{code}.txt"
        p.write_text(text)
        assert cli([str(p), "prob", str(s)])[0] == code
    assert cli([os.path.join(golden, "c1", "plan.txt")])[0] == 1
    q = tmp_path / "q.txt"
    q.write_text("?" + "0" * 23 + "\n")
    assert cli([os.path.join(golden, "c1", "plan.txt"), "prefix", str(q), "4", str(tmp_path / "x")])[0] == 8
    assert cli([os.path.join(golden, "c1", "plan.txt"), "resume", str(s), "4", "1", str(tmp_path / "missing")])[0] == 9
