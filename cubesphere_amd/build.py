"""Build libcubesphere_hip.so (HIP for gfx950) in-tree.

The host metric (cs_mesh.cpp) is compiled by g++ with -ffp-contract=off so the
vertex lattice is bit-identical to the reference's Fortran arithmetic; the
kernels are compiled by hipcc for gfx950 only.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcubesphere_hip.so")
BUILD = os.path.join(HERE, "_build")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CS_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    mesh_o = os.path.join(BUILD, "cs_mesh.o")
    kern_o = os.path.join(BUILD, "cs_kernels.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-c",
          os.path.join(CSRC, "cs_mesh.cpp"), "-o", mesh_o])
    p1h_o = os.path.join(BUILD, "cs_p1s_host.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-c",
          os.path.join(CSRC, "cs_p1s_host.cpp"), "-o", p1h_o])
    dump_o = os.path.join(BUILD, "cs_dump.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, "cs_dump.cpp"), "-o", dump_o])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "cs_kernels.hip"), "-o", kern_o])
    geom_o = os.path.join(BUILD, "cs_geometry.o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "cs_geometry.hip"), "-o", geom_o])
    p1_o = os.path.join(BUILD, "cs_slopes_p1.o")
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, "cs_slopes_p1.hip"),
          "-o", p1_o])
    p1s_o = os.path.join(BUILD, "cs_p1_stage.o")
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, "cs_p1_stage.hip"),
          "-o", p1s_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", mesh_o, p1h_o, dump_o, kern_o, geom_o, p1_o, p1s_o,
          "-pthread", "-o",
          tmp])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    # hardware-counter probe: one kernel + a calibration copy, no Python in the process
    _run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "cs_probe.hip"), "-L", HERE, "-lcubesphere_hip",
          "-Wl,-rpath,$ORIGIN", "-o", PROBE])
    if verbose:
        print("built", PROBE)
    build_fortran(verbose)
    return LIB


PROBE = os.path.join(HERE, "cs_probe")


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
DEMO = os.path.join(HERE, "cs_dropin_demo")
P1DEMO = os.path.join(HERE, "cs_p1_dropin")


def build_fortran(verbose=False):
    """The ISO_C_BINDING module and the Fortran drop-in demo (the reference's
    own language); skipped when no Fortran compiler is installed."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran interface not built")
        return None
    fdir = os.path.join(HERE, "fortran")
    obj = os.path.join(BUILD, "cs_hip_iface.o")
    _run([FLANG, "-c", "-O2", "-J", BUILD, os.path.join(fdir, "cs_hip_iface.f90"), "-o", obj])
    _run([FLANG, "-O2", "-I", BUILD, os.path.join(fdir, "cs_dropin_demo.f90"), obj, "-L", HERE,
          "-lcubesphere_hip", "-Wl,-rpath,$ORIGIN", "-o", DEMO])
    _run([FLANG, "-O2", "-I", BUILD, os.path.join(fdir, "cs_p1_dropin.f90"), obj, "-L", HERE,
          "-lcubesphere_hip", "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", P1DEMO])
    if verbose:
        print("built", DEMO, P1DEMO)
    return DEMO


if __name__ == "__main__":
    build(verbose=True)
