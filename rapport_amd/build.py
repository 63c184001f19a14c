"""Build the gfx950 HIP library in-tree: rapport_amd/librapport_amd.so.

    python -m rapport_amd.build          # or rapport_amd.build.build()

hipcc cross-compiles for gfx950 without a GPU. Packed FP32 instructions are
disabled for all device code (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "rapport_amd.hip")
OUT = os.path.join(HERE, "librapport_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def sources():
    d = os.path.join(HERE, "csrc")
    return [os.path.join(d, f) for f in sorted(os.listdir(d))
            if f.endswith((".hip", ".hpp", ".h"))] + [os.path.join(ROOT, "include", "rapport_amd.h")]


def build(verbose=True, extra=()):
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, *extra, SRC, "-o", tmp]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building librapport_amd.so")
    if verbose and r.stderr.strip():
        # host-pass warning about the packed-fp32 feature is expected
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            sys.stderr.write("\n".join(lines[:40]) + "\n")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(extra=sys.argv[1:])
