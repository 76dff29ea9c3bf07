"""Builds the HIP library genesis6_amd/lib/libg6gpu.so for gfx950 and the
Fortran MD driver genesis6_amd/lib/g6md linked against it (in-tree).

    python -m genesis6_amd.build
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libg6gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FSRC = os.path.join(HERE, "fortran")
DRIVER = os.path.join(LIBDIR, "g6md")
MC_DRIVER = os.path.join(LIBDIR, "g6mc")
RX_DRIVER = os.path.join(LIBDIR, "g6rx")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # no packed FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose and r.stderr:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    """The ISO_C_BINDING module and the deck-driven MD driver (flang)."""
    if not os.path.exists(FLANG):
        raise RuntimeError("flang not found at %s" % FLANG)
    objdir = os.path.join(LIBDIR, "fortran_obj")
    os.makedirs(objdir, exist_ok=True)
    objs = {}
    for src in ("g6gpu.f90", "g6md.f90", "g6mc.f90", "g6rx.f90"):
        obj = os.path.join(objdir, src.replace(".f90", ".o"))
        cmd = [FLANG, "-O2", "-J", objdir, "-c", os.path.join(FSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            raise RuntimeError("flang failed on %s" % src)
        objs[src] = obj
    # the MD driver (Mixture decks), the MC driver (mc2 inputs) and the
    # replica-exchange MD driver
    for prog, exe in (("g6md.f90", DRIVER), ("g6mc.f90", MC_DRIVER), ("g6rx.f90", RX_DRIVER)):
        cmd = [FLANG, "-O2", "-o", exe + ".tmp", objs["g6gpu.f90"], objs[prog],
               "-L" + LIBDIR, "-lg6gpu", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stderr)
            raise RuntimeError("flang link of %s failed" % os.path.basename(exe))
        os.replace(exe + ".tmp", exe)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
