"""Builds libfvmhip.so (HIP for gfx950 + host C++) in-tree.

    python xcalibre_tw.jl_amd/build.py [-j N]

Every HIP translation unit is compiled with -ffp-contract=off: the FP64 path
is bit-identical to the CPU restatement only when no multiply-add is fused.
Packed FP32 instructions are disabled (-packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "fvmhip")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libfvmhip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-munsafe-fp-atomics",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _compile(src, kind):
    obj = os.path.join(OBJ_DIR, src + ".o")
    path = os.path.join(CSRC, src)
    if kind == "hip":
        cmd = [HIPCC, *HIP_FLAGS, "-c", path, "-o", obj]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(jobs=None, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    hip, cpp = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs, failed = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, "hip") for s in hip] + [ex.submit(_compile, s, "cpp") for s in cpp]
        for f in futs:
            src, obj, rc, log = f.result()
            if rc != 0:
                failed.append((src, log))
            else:
                objs.append(obj)
                if verbose and log.strip():
                    print(f"--- {src}\n{log}")
    if failed:
        for src, log in failed:
            sys.stderr.write(f"=== {src} failed\n{log}\n")
        raise RuntimeError("libfvmhip build failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
