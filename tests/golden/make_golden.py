"""Writes the golden fixtures under tests/golden/ (run in the build container, where the reference tree and the
reference executor compiled from its own sources by oracle/Makefile exist; the fixtures are data and travel with the
repository, the reference does not).

* ibm/    -- IBM's 70-qubit instance and its committed plans and known answers, copied as data:
             cloud_nq76/nq70.qasm, orders/order_w30.json, cloud_ibmxeb/nq70_inplace_plan.txt, cloud_ibmxeb/nq70_w30_plan.txt,
             cloud/nq70_plan.txt + cloud/nq70_step_qubits.json, results/nq70_graph_adjacency.txt,
             cloud_ibmxeb/gate_strings.txt with the gate values of cloud_ibmxeb/run.sh:19, and the first 3 samples of
             results/laptop_samples_w30.txt (gsamp2 sample 3 2000000);
* doped/  -- data_doped/nq70_depth70_checks27_doped.qasm, cloud_doped/doped_m20_step_vertices.json and
             cloud_doped/doped_m20_inplace_plan.txt (538 vertices, R = 34, all five step kinds);
* c1/     -- the n = 24, band 10 instance planned by this package (plan.txt), 64 uniform strings, 8 hybrid lines,
             and the reference executor's outputs for every mode in complex64 and complex128;
* small/  -- 12 random small circuits (n <= 9) planned by this package with every string of each, and the
             reference executor's prob / chain outputs (complex128);
* tt/     -- the two-table engine: the config-1 instance's gsamp plan (engine.write_sampler_plan) with the reference
             gsamp's prob / chain / sample / hybrid outputs; a dp plan and sliced jobs (engine.write_plan,
             engine.Sliced, engine.write_jobs) with the reference dp's outputs (complex128 and -DDP_FLOAT); and the
             committed chain known answers: results/chain/gate_fwd_out.txt and the first 4 lines of
             results/chain/ideal_fwd_chain.txt (strings from cloud_chain/ideal_fwd.txt) on cloud_ibmxeb/nq70_w30_plan.txt,
             cloud_chain/plan_rev.txt with cloud_chain/gate_rev.txt and results/chain/gate_rev_out.txt.
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("GS_REFERENCE", "/root/reference")
REFBIN = os.path.join(ROOT, "oracle", "_ref")
sys.path.insert(0, os.path.join(ROOT, "tests"))
from _pkg import egs  # noqa: E402

GATE_VALUES = [-69.508530214929, -68.876165103181]   # cloud_ibmxeb/run.sh:19
DOPED_C128 = [-68.295331132, -69.167144551]          # results/doped/run15_c128/README.md:10-13 (first, last shot)
DOPED_C64 = [-68.295330880, -69.167143120]


def run(binary, args, c128=False):
    exe = os.path.join(REFBIN, "gsamp2_c128" if c128 else "gsamp2")
    res = subprocess.run([exe] + args, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS=os.environ.get("GS_GOLDEN_THREADS", "4")))
    if res.returncode:
        raise RuntimeError(f"{exe} {args}: exit {res.returncode} {res.stderr}")
    return res.stdout


def copy(src, dst):
    shutil.copyfile(os.path.join(REF, src), dst)


def ibm():
    d = os.path.join(HERE, "ibm")
    os.makedirs(d, exist_ok=True)
    copy("cloud_nq76/nq70.qasm", os.path.join(d, "nq70.qasm"))
    copy("orders/order_w30.json", os.path.join(d, "order_w30.json"))
    copy("cloud_ibmxeb/nq70_inplace_plan.txt", os.path.join(d, "nq70_inplace_plan.txt"))
    copy("cloud_ibmxeb/nq70_w30_plan.txt", os.path.join(d, "nq70_w30_plan.txt"))
    copy("cloud/nq70_plan.txt", os.path.join(d, "nq70_w31_plan.txt"))
    copy("cloud/nq70_step_qubits.json", os.path.join(d, "nq70_w31_step_qubits.json"))
    copy("results/nq70_graph_adjacency.txt", os.path.join(d, "nq70_graph_adjacency.txt"))
    copy("cloud_ibmxeb/gate_strings.txt", os.path.join(d, "gate_strings.txt"))
    with open(os.path.join(d, "gate_values.txt"), "w") as f:
        f.write("\n".join("%.12f" % v for v in GATE_VALUES) + "\n")
    with open(os.path.join(REF, "results/laptop_samples_w30.txt")) as f:
        lines = [l for l in f if len(l.split()) == 3 and set(l.split()[0]) <= set("01")][:3]
    with open(os.path.join(d, "laptop_samples_w30.txt"), "w") as f:
        f.writelines(lines)


def doped():
    d = os.path.join(HERE, "doped")
    os.makedirs(d, exist_ok=True)
    copy("data_doped/nq70_depth70_checks27_doped.qasm", os.path.join(d, "nq70_depth70_checks27_doped.qasm"))
    copy("cloud_doped/doped_m20_step_vertices.json", os.path.join(d, "doped_m20_step_vertices.json"))
    copy("cloud_doped/doped_m20_inplace_plan.txt", os.path.join(d, "doped_m20_inplace_plan.txt"))
    # IBM's first and last doped shots (Qiskit order, cloud_doped/run_c128.sh:24: head -1 / grep . | tail -1) with the
    # reference's published log2 Pr_doped of each in complex128 and complex64 (results/doped/run15_c128/README.md:10-13)
    with open(os.path.join(REF, "cloud_doped/ibm_doped_shots_qiskit.txt")) as f:
        shots = [l.strip() for l in f if l.strip()]
    with open(os.path.join(d, "ibm_doped_shots_first_last.txt"), "w") as f:
        f.write(shots[0] + "\n" + shots[-1] + "\n")
    with open(os.path.join(d, "ibm_doped_shots_values.json"), "w") as f:
        json.dump({"c128": DOPED_C128, "c64": DOPED_C64, "nd": 70, "na": 468}, f)


def c1():
    d = os.path.join(HERE, "c1")
    os.makedirs(d, exist_ok=True)
    inst = egs.instances.config_instance("c1")
    plan = os.path.join(d, "plan.txt")
    inst.write_plan(plan)
    inst.write_qasm(os.path.join(d, "c1.qasm"))
    with open(os.path.join(d, "step_vertices.json"), "w") as f:
        json.dump(inst.step_vertices, f)
    ns = len(inst.steps)
    rng = np.random.default_rng(101)
    uni = rng.integers(0, 2, size=(64, ns))
    egs.planio.write_strings(os.path.join(d, "strings.txt"), uni)
    hy = rng.integers(0, 2, size=(8, ns)).astype(np.int8)
    hy[:, ns - 6:] = -1
    egs.planio.write_strings(os.path.join(d, "hybrid.txt"), hy)
    K = ns - 6
    for c128 in (False, True):
        tag = "c128" if c128 else "c64"
        outs = {
            f"prob_{tag}.txt": ["prob", plan_rel("strings.txt", d)],
            f"chain_{tag}.txt": ["chain", plan_rel("strings.txt", d)],
            f"sample_{tag}.txt": ["sample", "16", "1234"],
            f"sample2_{tag}.txt": ["sample2", "16", "1234"],
            f"hybrid_{tag}.txt": ["hybrid", plan_rel("hybrid.txt", d), "99"],
        }
        for name, args in outs.items():
            with open(os.path.join(d, name), "w") as f:
                f.write(run("gsamp2", [plan] + args, c128))
        # prefix then resume (one slice) on the hybrid lines
        dump = os.path.join(d, f"dump_{tag}")
        with open(os.path.join(d, f"prefix_{tag}.txt"), "w") as f:
            f.write(run("gsamp2", [plan, "prefix", plan_rel("hybrid.txt", d), str(K), dump], c128))
        with open(os.path.join(d, f"resume_{tag}.txt"), "w") as f:
            f.write(run("gsamp2", [plan, "resume", plan_rel("hybrid.txt", d), str(K), "5", dump], c128))
    with open(os.path.join(d, "meta.json"), "w") as f:
        json.dump(dict(K=K, sample_seed=1234, hybrid_seed=99, resume_seed=5, n=inst.N, ns=ns, R=inst.R), f)


def plan_rel(name, d):
    return os.path.join(d, name)


def small():
    """Random Clifford circuits with the reference tests' tail (tdg h rz(TH) h, test_gsamp.py:57-68)."""
    d = os.path.join(HERE, "small")
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(31)
    th = 0.9553166181245092
    cases = []
    for i in range(12):
        n = int(rng.integers(2, 10))
        ops = []
        for _ in range(int(rng.integers(10, 80))):
            g = str(rng.choice(["h", "s", "sx", "x", "z", "cz", "cz", "cz"]))
            ops.append((g, [int(q) for q in rng.choice(n, 2, replace=False)]) if g == "cz" else (g, [int(rng.integers(n))]))
        tail = []
        for q in range(n):
            tail += [("tdg", [q]), ("h", [q]), (f"rz({th})", [q]), ("h", [q])]
        order = [int(v) for v in rng.permutation(n)]
        t = egs.tableau.Tableau(n)
        for name, qs in ops:
            t.apply(name, qs)
        am = egs.amp.Amplitudes(n, t, tail, order)
        R = max(1, egs.plan.width(am.steps))
        st = egs.inplace.plan_inplace(am.A, order, R)
        plan = os.path.join(d, f"s{i}_plan.txt")
        egs.inplace.write_inplace_plan(plan, st, R, am.M, n)
        egs.gadgets.write_qasm(os.path.join(d, f"s{i}.qasm"), n, ops + tail)
        with open(os.path.join(d, f"s{i}_order.json"), "w") as f:
            json.dump(order, f)
        sv = [s["v"] for s in st]
        allq = np.array([[(idx >> (n - 1 - q)) & 1 for q in range(n)] for idx in range(2 ** n)])
        egs.planio.write_strings(os.path.join(d, f"s{i}_strings.txt"), allq[:, sv])
        for mode in ("prob", "chain"):
            with open(os.path.join(d, f"s{i}_{mode}_c128.txt"), "w") as f:
                f.write(run("gsamp2", [plan, mode, os.path.join(d, f"s{i}_strings.txt")], True))
        cases.append(dict(i=i, n=n, R=R, order=order))
    with open(os.path.join(d, "cases.json"), "w") as f:
        json.dump(cases, f)


def run_bin(name, args):
    exe = os.path.join(REFBIN, name)
    res = subprocess.run([exe] + args, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS=os.environ.get("GS_GOLDEN_THREADS", "4")))
    if res.returncode:
        raise RuntimeError(f"{exe} {args}: exit {res.returncode} {res.stderr}")
    return res.stdout


def twotable():
    d = os.path.join(HERE, "tt")
    os.makedirs(d, exist_ok=True)
    inst = egs.instances.Instance(24, 10, 1)
    plan = os.path.join(d, "c1_plan.txt")
    egs.engine.write_sampler_plan(plan, inst.am)
    strings = os.path.join(HERE, "c1", "strings.txt")
    hybrid = os.path.join(HERE, "c1", "hybrid.txt")
    for mode, args in (("prob", ["prob", strings]), ("chain", ["chain", strings]), ("sample", ["sample", "16", "1234"]),
                       ("hybrid", ["hybrid", hybrid, "99"])):
        with open(os.path.join(d, f"c1_{mode}.txt"), "w") as f:
            f.write(run_bin("gsamp", [plan] + args))
    # prefix / resume (gsamp.cpp:124-148) of the hybrid lines at K = 18: dumps and lines
    meta = json.load(open(os.path.join(HERE, "c1", "meta.json")))
    K = meta["K"]
    with open(os.path.join(d, "c1_prefix.txt"), "w") as f:
        f.write(run_bin("gsamp", [plan, "prefix", hybrid, str(K), os.path.join(d, "c1_dump")]))
    with open(os.path.join(d, "c1_resume.txt"), "w") as f:
        f.write(run_bin("gsamp", [plan, "resume", hybrid, str(K), str(meta["resume_seed"]), os.path.join(d, "c1_dump")]))
    # dp: the graph with vertices 3 and 11 sliced, weights of 6 of the strings (w_q = M_q[b_q, :])
    A, M = inst.am.A, inst.am.M
    sl = egs.engine.Sliced(A, inst.order, [3, 11])
    egs.engine.write_plan(os.path.join(d, "dp_plan.txt"), sl.steps, sl.m)
    bits = egs.planio.read_strings(strings, len(inst.steps))[:6]
    ws = []
    for b in bits:
        bq = egs.planio.step_to_qubit(b, inst.step_vertices, inst.N)
        ws.append([(complex(M[q][bq[q], 0]), complex(M[q][bq[q], 1])) for q in range(inst.N)])
    egs.engine.write_jobs(os.path.join(d, "dp_jobs.txt"), sl, ws)
    for name, tag in (("dp", "c128"), ("dp_float", "c64")):
        with open(os.path.join(d, f"dp_out_{tag}.txt"), "w") as f:
            f.write(run_bin(name, [os.path.join(d, "dp_plan.txt"), os.path.join(d, "dp_jobs.txt")]))
    # committed known answers of the reference's two-table chain mode
    with open(os.path.join(REF, "results/chain/gate_fwd_out.txt")) as f:
        open(os.path.join(d, "ibm_gate_fwd_chain.txt"), "w").write(f.read())
    with open(os.path.join(REF, "cloud_chain/ideal_fwd.txt")) as f:
        open(os.path.join(d, "ibm_ideal_fwd_strings.txt"), "w").write("".join(f.readlines()[:4]))
    with open(os.path.join(REF, "results/chain/ideal_fwd_chain.txt")) as f:
        open(os.path.join(d, "ibm_ideal_fwd_chain.txt"), "w").write("".join(f.readlines()[:4]))
    copy("cloud_chain/plan_rev.txt", os.path.join(d, "ibm_plan_rev.txt"))
    copy("cloud_chain/gate_rev.txt", os.path.join(d, "ibm_gate_rev.txt"))
    copy("results/chain/gate_rev_out.txt", os.path.join(d, "ibm_gate_rev_chain.txt"))


def zero_row_plan(text, k):
    """The plan text with row b = 1 of step k's M set to zero: x_v = 1 at that step has probability exactly 0."""
    toks = text.split()
    ns = int(toks[1])
    pos = 3
    for j in range(ns):
        nk, nt = int(toks[pos + 1]), int(toks[pos + 5])
        pos += 6 + nk + 4 * nt
        if j == k:
            toks[pos + 4:pos + 8] = ["0", "0", "0", "0"]
        pos += 8
    lines = [" ".join(toks[:3])]
    pos = 3
    for j in range(ns):
        nk, nt = int(toks[pos + 1]), int(toks[pos + 5])
        end = pos + 6 + nk + 4 * nt + 8
        lines.append(" ".join(toks[pos:end]))
        pos = end
    return "\n".join(lines) + "\n"


def zero():
    """Zero-probability branches (gsamp2.cpp:48, :67, :94, :116): the config-1 plan with row b = 1 of step 5's M zeroed.
    Strings forcing b_5 = 1 have probability 0: prob and chain give -inf, chain's conditionals are 0 from step 5 on,
    hybrid gives -inf with conditionals 0 and sampled bits 0 from step 5 on.  The reference executor's outputs in
    complex64 and complex128."""
    d = os.path.join(HERE, "zero")
    os.makedirs(d, exist_ok=True)
    plan = os.path.join(d, "plan.txt")
    with open(os.path.join(HERE, "c1", "plan.txt")) as f:
        open(plan, "w").write(zero_row_plan(f.read(), 5))
    ns = 24
    rng = np.random.default_rng(505)
    s = rng.integers(0, 2, size=(8, ns)).astype(np.int8)
    s[:4, 5] = 1
    s[4:, 5] = 0
    egs.planio.write_strings(os.path.join(d, "strings.txt"), s)
    hy = rng.integers(0, 2, size=(6, ns)).astype(np.int8)
    hy[:, ns - 8:] = -1
    hy[:3, 5] = 1
    hy[3:, 5] = 0
    hy[5, 2:] = -1                      # a sampled step 5 (never picks the zero branch) and a sampled prefix
    egs.planio.write_strings(os.path.join(d, "hybrid.txt"), hy)
    for c128 in (False, True):
        tag = "c128" if c128 else "c64"
        for mode, args in (("prob", ["prob", os.path.join(d, "strings.txt")]),
                           ("chain", ["chain", os.path.join(d, "strings.txt")]),
                           ("sample", ["sample", "8", "77"]),
                           ("hybrid", ["hybrid", os.path.join(d, "hybrid.txt"), "41"])):
            with open(os.path.join(d, f"{mode}_{tag}.txt"), "w") as f:
                f.write(run("gsamp2", [plan] + args, c128))


def c4():
    """The n = 60, band-24 instance's plan (nk up to 20: two of the oracle's 10-bit lookup chunks)
    with the reference executor's outputs of every mode in complex64 and complex128, so that the oracle is pinned
    at the sizes where the GPU tests use it as the checker."""
    d = os.path.join(HERE, "c4")
    os.makedirs(d, exist_ok=True)
    inst = egs.instances.config_instance("c4")
    plan = os.path.join(d, "plan.txt")
    inst.write_plan(plan)
    ns = len(inst.steps)
    rng = np.random.default_rng(404)
    egs.planio.write_strings(os.path.join(d, "strings.txt"), rng.integers(0, 2, size=(6, ns)))
    hy = rng.integers(0, 2, size=(3, ns)).astype(np.int8)
    hy[:, ns - 12:] = -1
    egs.planio.write_strings(os.path.join(d, "hybrid.txt"), hy)
    K = ns - 12
    for c128 in (False, True):
        tag = "c128" if c128 else "c64"
        for mode, args in (("prob", ["prob", os.path.join(d, "strings.txt")]),
                           ("chain", ["chain", os.path.join(d, "strings.txt")]),
                           ("sample", ["sample", "3", "4040"]),
                           ("sample2", ["sample2", "3", "4040"]),
                           ("hybrid", ["hybrid", os.path.join(d, "hybrid.txt"), "44"])):
            with open(os.path.join(d, f"{mode}_{tag}.txt"), "w") as f:
                f.write(run("gsamp2", [plan] + args, c128))
        dump = os.path.join(d, f"dump_{tag}")
        with open(os.path.join(d, f"prefix_{tag}.txt"), "w") as f:
            f.write(run("gsamp2", [plan, "prefix", os.path.join(d, "hybrid.txt"), str(K), dump], c128))
        with open(os.path.join(d, f"resume_{tag}.txt"), "w") as f:
            f.write(run("gsamp2", [plan, "resume", os.path.join(d, "hybrid.txt"), str(K), "45", dump], c128))
    with open(os.path.join(d, "meta.json"), "w") as f:
        json.dump(dict(K=K, sample_seed=4040, hybrid_seed=44, resume_seed=45, ns=ns, R=inst.R), f)


def c2():
    """The n = 70, band-30 sampling instance's plan (R = 30, nk up to 29: three of the oracle's 10-bit lookup chunks)
    with the reference executor's first 3 samples from seed 4242 in complex64 and complex128."""
    d = os.path.join(HERE, "c2")
    os.makedirs(d, exist_ok=True)
    plan = os.path.join(d, "plan.txt")
    egs.instances.config_instance("c2").write_plan(plan)
    for c128 in (False, True):
        with open(os.path.join(d, f"sample_{'c128' if c128 else 'c64'}.txt"), "w") as f:
            f.write(run("gsamp2", [plan, "sample", "3", "4242"], c128))


def sliced():
    """The reference's committed sliced plans, as SHA-256 digests (the files are 22-195 KB each), with the small
    twist files and the twist gate's proposals copied as data:
    * cloud_nq80/tw70_s0.txt, tw70_s1.txt, tw70_twist.txt: IBM #151 (width-30 order) sliced at vertex 21, K = 50;
      tw70_props.txt holds the 3 hybrid proposals of the gate cloud_nq80/run_hybrid.sh:32-43 (seed 17);
    * cloud_doped/doped_m20sl_s00..s11.txt: the 538-vertex doped instance sliced at ancillas 135 and 104, K = 518,
      R = 33 (cloud_doped/doped_m20sl_slices.json)."""
    import hashlib
    d = os.path.join(HERE, "sliced")
    os.makedirs(d, exist_ok=True)
    digests = {}
    for name in ["cloud_nq80/tw70_s0.txt", "cloud_nq80/tw70_s1.txt", "cloud_nq80/tw70_twist.txt"] + \
            [f"cloud_doped/doped_m20sl_s{t}.txt" for t in ("00", "01", "10", "11")]:
        with open(os.path.join(REF, name), "rb") as f:
            digests[name] = hashlib.sha256(f.read()).hexdigest()
    meta = json.load(open(os.path.join(REF, "cloud_doped/doped_m20sl_slices.json")))
    with open(os.path.join(d, "reference_sliced_plans.json"), "w") as f:
        json.dump({"sha256": digests,
                   "tw70": {"slices": [21], "K": 50, "hybrid_seed": 17},
                   "doped_m20sl": {"slices": meta["slices"], "K": meta["K_suffix_start"], "R": meta["R"]}}, f,
                  indent=1)
    copy("cloud_nq80/tw70_twist.txt", os.path.join(d, "tw70_twist.txt"))
    copy("cloud_nq80/tw70_props.txt", os.path.join(d, "tw70_props.txt"))


def kats():
    """Larger known-answer sets the reference's cloud runs produced, copied as data:
    * ibm/uniform64.txt + uniform64_out.txt -- 64 uniform strings (cloud_ibmxeb/uniform64.txt, step order) and their
      log2 Pr from the two-table sampler on cloud_ibmxeb/nq70_w30_plan.txt (results/ibmxeb/uniform_out.txt);
    * ibm/samples_nq70_first4.txt -- the first 4 exact samples of the two-table sampler on cloud/nq70_plan.txt, seeds
      1000000.. (results/samples_nq70.txt);
    * tt/check55_plan.txt + check55_expected.txt -- the 55-qubit bit-exact gate of cloud/run.sh:19-27
      (gsamp check55_plan.txt sample 3 1)."""
    d = os.path.join(HERE, "ibm")
    copy("cloud_ibmxeb/uniform64.txt", os.path.join(d, "uniform64.txt"))
    copy("results/ibmxeb/uniform_out.txt", os.path.join(d, "uniform64_out.txt"))
    with open(os.path.join(REF, "results/samples_nq70.txt")) as f:
        open(os.path.join(d, "samples_nq70_first4.txt"), "w").write("".join(f.readlines()[:4]))
    t = os.path.join(HERE, "tt")
    copy("cloud/check55_plan.txt", os.path.join(t, "check55_plan.txt"))
    copy("cloud/check55_expected.txt", os.path.join(t, "check55_expected.txt"))


if __name__ == "__main__":
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            globals()[name]()
        sys.exit(0)
    ibm()
    doped()
    c1()
    small()
    twotable()
    print("fixtures written under", HERE)
