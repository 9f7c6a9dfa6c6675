"""Build recipe of libsqe.so (hand-written HIP for gfx950) -- in-tree.

    python -m sparqc_amd.build        # or __graft_entry__.build()

Each translation unit is compiled with hipcc for gfx950 in parallel, then linked
into sparqc_amd/_build/libsqe.so.  Flags:
  -ffp-contract=off   FP64 gate arithmetic keeps the reference's operation
                      order (no fused multiply-add), so results match the CPU
                      oracle bit for bit; FP32 kernels fuse explicitly.
  -target-feature -packed-fp32-ops
                      no packed FP32 instructions.
  -simplifycfg-sink-common=false
                      keeps the per-bucket amplitude arrays in registers: merging
                      the per-target-bit variants of a gate otherwise leaves
                      pointer phis into the array and forces it to scratch.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libsqe.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-simplifycfg-sink-common=false",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]

SOURCES = ["sqe_kernels_f64.hip", "sqe_kernels_f32.hip", "sqe_kernels_common.hip", "sqe_engine.cpp",
           "sqe_batch.cpp", "sqe_noalloc.hip"]


def _deps(path, seen=None):
    """The file and every local header it includes, recursively."""
    import re
    seen = set() if seen is None else seen
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for inc in re.findall(r'^\s*#\s*include\s+"([^"]+)"', f.read(), re.M):
            for base in (os.path.dirname(path), CSRC, os.path.join(ROOT, "include")):
                cand = os.path.join(base, inc)
                if os.path.exists(cand):
                    _deps(cand, seen)
                    break
    return seen


def _stale(src, obj):
    """The object is older than its source or a header the source includes."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in _deps(os.path.join(CSRC, src)))


def _compile(src, incremental=False):
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC, "--offload-arch=" + ARCH] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=False, incremental=False):
    """Compile every translation unit (incremental: only those whose object is
    older than a source or header) and link libsqe.so."""
    os.makedirs(OUT_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "circuit_check.cpp")
CHECK_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "circuit_check")


def build_circuit_check(verbose=False):
    """Host-only C++ driver of include/sparqc_circuit.hpp linked to libsqe.so
    (g++; it finds the library through an rpath relative to itself)."""
    os.makedirs(os.path.dirname(CHECK_BIN), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), CHECK_SRC,
           "-L", OUT_DIR, "-lsqe", "-Wl,-rpath,$ORIGIN/../../../sparqc_amd/_build", "-o", CHECK_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for circuit_check:\n%s" % r.stderr[-4000:])
    if verbose:
        print("built", CHECK_BIN)
    return CHECK_BIN


ENC_SRC = os.path.join(ROOT, "tests", "cpp", "encode_check.cpp")
ENC_BIN = os.path.join(ROOT, "tests", "cpp", "_build", "encode_check")


def build_encode_check(verbose=False):
    """Host-only check of the sweep-program encoder (hipcc host pass: the
    planner headers include the HIP runtime header; no device code is built)."""
    os.makedirs(os.path.dirname(ENC_BIN), exist_ok=True)
    cmd = [HIPCC, "--offload-host-only", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           ENC_SRC, "-o", ENC_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for encode_check:\n%s" % r.stderr[-4000:])
    if verbose:
        print("built", ENC_BIN)
    return ENC_BIN


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    build_circuit_check(verbose=True)
    build_encode_check(verbose=True)
    sys.exit(0)
