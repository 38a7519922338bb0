"""Readers and writers of the reference's file formats around the in-place engine.

* strings files: one string per line in STEP order, characters '0'/'1' ('?' for sampled steps); lines shorter
  than ns characters are skipped (gsamp2.cpp:219-221);
* step-vertex JSON: ``[v_0, v_1, ...]``, the vertex of each step; it maps qubit-order strings to step order
  (ibm_xeb_prep.py:25-28, doped_sampler.py:35-40);
* prefix dumps: int32 nk, int32 ntq, float64 logscale, then 2^nk * ntq complex entries of the arena type in
  (c, sorted offset) order (gsamp2.cpp:170-175);
* twist files: one ``cshift tqshift`` line per slice (sliced_plans.py:83-84);
* in-place plan files, parsed into Python dicts (the engine parses them itself; this reader serves host-side
  analysis and tests).
"""
import json

import numpy as np


def read_strings(path, ns, allow_q=False):
    """Strings file -> int8 array [nstr, ns] (-1 for '?')."""
    rows = []
    with open(path) as f:
        for line in f:
            if len(line) < ns:
                continue
            row = []
            for ch in line[:ns]:
                if ch in "01":
                    row.append(ord(ch) - 48)
                elif ch == "?" and allow_q:
                    row.append(-1)
                else:
                    raise ValueError(f"not a bit: {ch!r}")
            rows.append(row)
    return np.array(rows, dtype=np.int8).reshape(-1, ns)


def format_bits(row):
    return "".join("?" if b < 0 else chr(48 + int(b)) for b in row)


def write_strings(path, bits):
    with open(path, "w") as f:
        for row in np.atleast_2d(bits):
            f.write(format_bits(row) + "\n")


def read_step_vertices(path):
    with open(path) as f:
        data = json.load(f)
    if isinstance(data, dict):
        data = data["step_vertices"]
    return [int(v) for v in data]


def qubit_to_step(bits_q, step_vertices):
    """Qubit-order strings [.., n] -> step order [.., ns]."""
    b = np.asarray(bits_q)
    return b[..., np.asarray(step_vertices)]


def step_to_qubit(bits_s, step_vertices, n):
    b = np.asarray(bits_s)
    out = np.zeros(b.shape[:-1] + (n,), dtype=b.dtype)
    out[..., np.asarray(step_vertices)] = b
    return out


def read_dump(path, c128=False):
    """Prefix dump -> ``(table complex [2^nk * ntq], nk, ntq, logscale)``."""
    with open(path, "rb") as f:
        nk, ntq = np.frombuffer(f.read(8), dtype=np.int32)
        ls = float(np.frombuffer(f.read(8), dtype=np.float64)[0])
        n = (1 << int(nk)) * int(ntq)
        raw = np.frombuffer(f.read(), dtype=np.float64 if c128 else np.float32)
    if raw.size != 2 * n:
        raise ValueError(f"bad dump {path}")
    return raw[0::2].astype(np.float64) + 1j * raw[1::2].astype(np.float64), int(nk), int(ntq), ls


def write_dump(path, table, nk, ntq, logscale, c128=False):
    t = np.asarray(table)
    inter = np.empty(2 * t.size, dtype=np.float64 if c128 else np.float32)
    inter[0::2] = t.real
    inter[1::2] = t.imag
    with open(path, "wb") as f:
        f.write(np.array([nk, ntq], dtype=np.int32).tobytes())
        f.write(np.array([logscale], dtype=np.float64).tobytes())
        f.write(inter.tobytes())


def read_twist(path, nslices):
    out = []
    with open(path) as f:
        toks = f.read().split()
    for s in range(nslices):
        out.append((int(toks[2 * s]), int(toks[2 * s + 1])))
    return out


def read_inplace_plan(path):
    """In-place plan file -> ``(n, R, steps)``; each step a dict with v, nk, m, g, rhoc, Tkeep, terms
    [(tq, x, tau, par)], M (2x2 complex array, M[b, x])."""
    with open(path) as f:
        toks = f.read().split()
    pos = 0

    def take(k):
        nonlocal pos
        out = toks[pos:pos + k]
        pos += k
        return out

    n, ns, R = (int(t) for t in take(3))
    steps = []
    for _ in range(ns):
        v, nk, m, g, rhoc, nt = (int(t) for t in take(6))
        tk = [int(t) for t in take(nk)]
        terms = []
        for _ in range(nt):
            tq, x, tau, par = (int(t) for t in take(4))
            terms.append((tq, x, tau, par))
        vals = [float(t) for t in take(8)]
        M = np.array([[complex(vals[0], vals[1]), complex(vals[2], vals[3])],
                      [complex(vals[4], vals[5]), complex(vals[6], vals[7])]])
        steps.append(dict(v=v, nk=nk, m=m, g=g, rhoc=rhoc, Tkeep=tk, terms=terms, M=M))
    return n, R, steps


def plan_traffic(steps, entry_bytes):
    """Algorithmic arena bytes of one pass, per mode.  r_k = nk + log2 ntq (live slots read), rn_k = nk + [m == 2]
    (slots written).  Returns dict(chain1=sum(2^r + 2^rn), sample=sum(2*2^r + 2^rn)) times entry_bytes."""
    c1 = 0
    sm = 0
    for s in steps:
        ntq = len({t[0] for t in s["terms"]})
        r = (1 << s["nk"]) * ntq
        rn = (1 << s["nk"]) * (2 if s["m"] == 2 else 1)
        c1 += r + rn
        sm += 2 * r + rn
    return dict(chain1=c1 * entry_bytes, sample=sm * entry_bytes)
