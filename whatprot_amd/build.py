"""Builds libwhatprot_amd.so for gfx950 with hipcc (in tree, no JIT cache) and
the whatprot-amd command line over its C ABI."""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libwhatprot_amd.so")
CLI_OUT = os.path.join(HERE, "whatprot-amd")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off (their forms are unreliable on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build(verbose=False):
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp"] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    # host-only executable: plain g++ against the library's C ABI
    cli = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, os.path.join(HERE, "cli", "main.cpp"),
           "-L" + HERE, "-lwhatprot_amd", "-Wl,-rpath,$ORIGIN", "-o", CLI_OUT + ".tmp"]
    if verbose:
        print(" ".join(cli), flush=True)
    subprocess.run(cli, check=True)
    os.replace(CLI_OUT + ".tmp", CLI_OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
