"""Build of the native parts: the HIP engine library, its command-line front end, and the CPU oracle.

``build()`` always recompiles.  Everything lands in ``bin/`` next to this file (git-ignored, shipped with the tree).
The HIP sources are compiled for gfx950 only, with packed-FP32 instruction selection disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "bin")
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(BIN, "libgsamp_amd.so")
CLI = os.path.join(BIN, "gsamp2_amd")
TT_CLIS = (os.path.join(BIN, "gsamp_amd"), os.path.join(BIN, "dp_amd"))
ARCH = os.environ.get("GS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, quiet_warn=True):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return res


def build_library():
    os.makedirs(BIN, exist_ok=True)
    cmds, objs = [], []
    for src in ("gs_kernels.hip", "gs_engine.hip", "gs_batch.hip", "gs_twotable.hip", "gs_plan.cpp"):
        obj = os.path.join(BIN, src.split(".")[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:     # one hipcc process per source file
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)


def build_cli():
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-o", tmp, os.path.join(CSRC, "gs_cli.cpp"),
          f"-L{BIN}", "-lgsamp_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    for out in TT_CLIS:                      # one source, the file name selects the gsamp or dp contract
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-o", tmp, os.path.join(CSRC, "gs_tt_cli.cpp"),
              f"-L{BIN}", "-lgsamp_amd", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, out)


def build_oracle():
    """The CPU oracle (test infrastructure) and, where the reference tree is present, the reference executor."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])


def build():
    build_library()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build()
    print("built", LIB, CLI)
