"""Builds the gfx950 library and the drop-in CLI in-tree.

    python -m golaypair_amd.build        (or golaypair_amd.build.build())

Outputs (git-ignored, shipped with the tree):
    golaypair_amd/lib/libgolaypair_gpu.so   C ABI declared in include/golay_gpu.h
    golaypair_amd/bin/golaypair             CLI with the reference's `main` interface
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib", "libgolaypair_gpu.so")
BIN = os.path.join(PKG, "bin", "golaypair")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs (no accvgpr moves around the PSD reduction)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
CXX_FLAGS = ["-O2", "-std=c++20", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    hips = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(OBJ, f[:-4] + ".o") for f in hips]

    def comp(i):
        return _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, hips[i]), "-o", objs[i]])

    with ThreadPoolExecutor(max_workers=min(8, len(hips))) as ex:
        for out in ex.map(comp, range(len(hips))):
            if verbose and out.strip():
                print(out)
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs)
    os.replace(LIB + ".tmp", LIB)
    cli = os.path.join(CSRC, "cli")
    if os.path.isdir(cli):
        srcs = sorted(os.path.join(cli, f) for f in os.listdir(cli) if f.endswith(".cpp"))
        _run(["g++"] + CXX_FLAGS + srcs + ["-o", BIN + ".tmp", "-L", os.path.dirname(LIB),
                                           "-lgolaypair_gpu", "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(BIN + ".tmp", BIN)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
