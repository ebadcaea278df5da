"""Build the fastmi HIP library (libfastmi.so) in-tree for gfx950.

Floating-point contraction is disabled (``-ffp-contract=off``) so that the kernels reproduce
the reference FastS kernels bit for bit; packed FP32 instructions are disabled as required on
this hardware pool.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfastmi.so")
SOURCES = ["fm_kernels.hip", "fm_lu.hip", "fm_adjoint.hip", "fm_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FM_OFFLOAD_ARCH", "gfx950")


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wno-unused-result"]


def build(verbose=False, jobs=None, out=None, extra=(), objdir=None):
    """Compile every source to an object in parallel, then link libfastmi.so (or ``out``);
    ``extra`` adds compiler flags (e.g. tuning macros such as -DFM_RES_WAVES=3)."""
    lib = out or LIB
    objdir = objdir or os.path.join(CSRC, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        o = os.path.join(objdir, s.replace(".hip", ".o"))
        cmd = [HIPCC] + flags() + list(extra) + ["-c", os.path.join(CSRC, s), "-o", o]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(o)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out.decode(errors="replace"))
        elif verbose and out:
            sys.stderr.write(out.decode(errors="replace"))
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
