"""Build recipe for libcfdx.so (hipcc, gfx950) — run in-tree, no GPU needed.

    python build.py            # builds libcfdx.so next to this file

Every HIP translation unit is compiled with packed FP32 disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcfdx.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CFDX_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-address-of-packed-member"]
HIPFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    elif verbose and (r.stderr.strip()):
        # hipcc's host pass warns about the packed-fp32 feature; keep other output visible
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
        if lines:
            print("\n".join(lines))


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".cpp", ".hip")):
            out.append(os.path.join(CSRC, f))
    return out


def build(verbose: bool = False) -> str:
    """Compiles every source under csrc/ into libcfdx.so (always recompiles)."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    t0 = time.time()
    procs = []
    for src in sources():
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        flags = list(COMMON)
        if src.endswith(".hip"):
            cmd = [HIPCC, "-x", "hip"] + HIPFLAGS + flags + ["-c", src, "-o", obj]
        else:
            cmd = [HIPCC, "-x", "c++"] + flags + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", src, "-o", obj]
        if verbose:
            print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"FAILED: {' '.join(cmd)}\n{out}{err}\n")
        elif verbose:
            lines = [l for l in err.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
            if lines:
                print("\n".join(lines))
    if failed:
        raise RuntimeError("libcfdx compile failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
