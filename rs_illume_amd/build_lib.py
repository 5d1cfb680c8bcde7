"""Builds librs_illume_amd.so (HIP kernels + C ABI) for gfx950 with hipcc.

Run as ``python -m rs_illume_amd.build_lib``; ``__graft_entry__.build()`` calls
:func:`build`. Objects go to ``rs_illume_amd/_build/``; the shared library is
written in-tree next to this file.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librs_illume_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every float expression rounds as written (bit-exact hit
# arithmetic). The packed-FP32 target feature is disabled for device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(REPO, "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


# Per-file device flags. guide.hip: MFMA results in ordinary VGPRs (the
# heuristic put them in AGPRs, and every bias, ReLU and conversion then paid
# an accumulator read or write).
FILE_FLAGS = {"guide.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    # Every translation unit is compiled as HIP for gfx950: hipcc would give a
    # .cpp file its default device target (not gfx950), so a kernel defined
    # there would have no code object on the device. Host-side files must not
    # define kernels at all; check_host_sources enforces that.
    cmd = [HIPCC, "-x", "hip"] + DEVICE + FILE_FLAGS.get(os.path.basename(src), []) + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def check_host_sources(srcs):
    """Kernels live in .hip files only; a __global__ in a .cpp file is an error."""
    for src in srcs:
        if src.endswith(".cpp") and "__global__" in open(src).read():
            raise RuntimeError("%s defines a kernel: move it to a .hip file" % src)


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    check_host_sources(srcs)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
