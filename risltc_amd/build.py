"""Builds risltc_amd/librisltc.so for gfx950 with hipcc (in-tree, no JIT).

Every HIP source is compiled with -ffp-contract=off (the kernels' operation
order is part of their parity contract) and with packed FP32 instructions
disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librisltc.so")
SOURCES = ["capi.cpp", "formats.cpp", "shade.hip", "lbvh.hip", "sah.hip", "mlp.hip"]
ARCH = os.environ.get("RISLTC_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-value", "-Wno-unused-result"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def build(verbose=False, jobs=4):
    """Compiles every source to an object in a private build folder, then links."""
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(bdir, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmd = [hipcc()] + lang + COMMON + DEVICE + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode())
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, out.decode()))
    text = out.decode().strip()
    if text:
        print(text, file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
