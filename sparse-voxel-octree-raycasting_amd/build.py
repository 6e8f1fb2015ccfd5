"""Build libsvo.so (gfx950 HIP kernels + C-ABI + host builder) in-tree.

    python sparse-voxel-octree-raycasting_amd/build.py

Writes sparse-voxel-octree-raycasting_amd/svo_amd/libsvo.so.  Host C++ is
compiled by g++, device code by hipcc for gfx950; both without FMA
contraction (the ray arithmetic must match the CPU restatement bit for bit)
and with packed-fp32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "svo_amd")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SVO_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-Wall", "-pthread"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    host_o = os.path.join(BUILD, "svo_host.o")
    cmds = [["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "svo_host.cpp"), "-o", host_o]]
    objs = [host_o]
    for src in ("svo_device.hip", "svo_cnn.hip", "svo_build.hip", "svo_scene.hip"):
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    # the translation units are independent: compile them concurrently
    procs = []
    for cmd in cmds:
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    failed = [cmd for cmd, pr in zip(cmds, procs) if pr.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    out = os.path.join(OUT_DIR, "libsvo.so")
    tmp = out + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-pthread"])
    os.replace(tmp, out)
    build_example()
    return out


def build_example():
    """The C host-loop example, linked against the in-tree libsvo.so."""
    src = os.path.join(HERE, "examples", "render_flythrough.c")
    exe = os.path.join(HERE, "examples", "render_flythrough")
    run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-pedantic", "-I", os.path.join(os.path.dirname(HERE), "include"),
         src, "-L", OUT_DIR, "-lsvo", "-Wl,-rpath,$ORIGIN/../svo_amd", "-o", exe])
    return exe


if __name__ == "__main__":
    print(build())
    sys.exit(0)
