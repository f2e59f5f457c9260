"""Build libccs.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc, the
ccs_search campaign CLI on top of it, and the LDS throughput probe (lds_probe)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libccs.so")
CLI = os.path.join(HERE, "ccs_search")
PROBE = os.path.join(HERE, "lds_probe")
SOURCES = ["ccs_kernels.hip", "ccs_host.cpp", "ccs_server.cpp", "ccs_confirm.cpp", "ccs_campaign.cpp",
           "ccs_family.cpp"]
HEADERS = ["ccs_internal.hpp", "ccs_arith.hpp", os.path.join("..", "..", "include", "ccs.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CCS_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose=False):
    """Compile every source into libccs.so (always recompiles)."""
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-c", "-x", "hip", os.path.join(CSRC, src), "-o", obj] + \
            [f for f in FLAGS if f != "-shared"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    # the campaign CLI: host code on the C ABI only, so the host compiler builds it
    cli_tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(CSRC, "ccs_search.cpp"), "-L" + HERE, "-lccs", "-Wl,-rpath,$ORIGIN", "-o", cli_tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(cli_tmp, CLI)
    # the LDS throughput probe bench.py takes the dense stage's ceiling from
    probe_tmp = PROBE + ".tmp"
    cmd = [HIPCC, "-O3", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           os.path.join(CSRC, "lds_probe.hip"), "-o", probe_tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(probe_tmp, PROBE)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
