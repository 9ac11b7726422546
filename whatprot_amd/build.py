"""Builds libwhatprot_amd.so for gfx950 with hipcc (in tree, no JIT cache) and
the whatprot-amd command line over its C ABI.

Every source compiles to its own object in parallel (each .hip file carries its
own gfx950 code object), then one hipcc line links the shared library."""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libwhatprot_amd.so")
CLI_OUT = os.path.join(HERE, "whatprot-amd")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off (their forms are unreliable on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def _jobs():
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(16, n or os.cpu_count() or 4))


def _newest_header():
    hs = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src, verbose, incremental=False):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if incremental and os.path.exists(obj) and \
            os.path.getmtime(obj) > max(os.path.getmtime(src), _newest_header()):
        return obj
    cmd = [HIPCC] + FLAGS + ["-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.stderr:
        print(r.stderr, file=sys.stderr, end="", flush=True)
    if r.returncode != 0:
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return obj


def build(verbose=False, incremental=False):
    """Compiles every source (incremental: only those older than their object
    or than any header) and links the library and the command line."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with concurrent.futures.ThreadPoolExecutor(_jobs()) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, incremental), srcs))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
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
    build(verbose="-v" in sys.argv, incremental="-i" in sys.argv)
    print(OUT)
