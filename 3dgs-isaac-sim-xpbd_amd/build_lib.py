"""Build recipe for libxpbd_isaac.so (HIP, gfx950) and the CPU oracle.

Run ``python 3dgs-isaac-sim-xpbd_amd/build_lib.py`` (or ``__graft_entry__.build()``).
Objects are compiled in parallel; the shared library lands next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libxpbd_isaac.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle_xpbd.so")
OBJ = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off keeps every a*b+c unfused so per-node passes reproduce the
# reference arithmetic bit for bit; packed fp32 ops are disabled for gfx950.
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function",
]
ORACLE_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-I" + INCLUDE]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def product_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def build_product(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = product_sources()
    objs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in srcs]

    def one(pair):
        src, obj = pair
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", src, "-o", obj])

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    srcs = sorted(os.path.join(ORACLE_DIR, f) for f in os.listdir(ORACLE_DIR) if f.endswith(".cpp"))
    tmp = ORACLE_LIB + ".tmp"
    _run(["g++"] + ORACLE_FLAGS + ["-shared", "-o", tmp] + srcs + ["-fopenmp"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


ASAN_DIR = os.path.join(ORACLE_DIR, "_asan")
ASAN_BIN = os.path.join(ASAN_DIR, "oracle_sanitized")


def build_oracle_asan():
    """The oracle and oracle/sanitize/driver.cpp in one host executable with
    AddressSanitizer and UndefinedBehaviorSanitizer (no GPU code involved)."""
    os.makedirs(ASAN_DIR, exist_ok=True)
    srcs = sorted(os.path.join(ORACLE_DIR, f) for f in os.listdir(ORACLE_DIR) if f.endswith(".cpp"))
    srcs.append(os.path.join(ORACLE_DIR, "sanitize", "driver.cpp"))
    tmp = ASAN_BIN + ".tmp"
    _run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fno-omit-frame-pointer",
          "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I" + INCLUDE, "-o", tmp] + srcs)
    os.replace(tmp, ASAN_BIN)
    return ASAN_BIN


if __name__ == "__main__":
    what = sys.argv[1:] or ["product", "oracle"]
    if "oracle" in what:
        print(build_oracle())
    if "oracle-asan" in what:
        print(build_oracle_asan())
    if "product" in what:
        print(build_product())
