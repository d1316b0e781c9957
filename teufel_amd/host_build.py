"""Build recipe for the C++ host library teufel_amd/lib/libteufel_host.so
(the reference-API mirror in teufel_amd/host/, linked against the HIP
library libteufel_gpu.so)."""
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libteufel_host.so")
INCLUDE = os.path.join(HERE, "host")
CXX = os.environ.get("CXX", "g++")


def sources():
    return sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [CXX, "-std=c++14", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", "-o", tmp] + sources() + [
        "-L" + LIB_DIR, "-lteufel_gpu", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host library build failed:\n" + r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_program(source, out, verbose=False):
    """Compile a C++ program against the host library (tests, examples)."""
    cmd = [CXX, "-std=c++14", "-O2", "-Wall", "-I" + INCLUDE, "-o", out, source, "-L" + LIB_DIR,
           "-lteufel_host", "-lteufel_gpu", "-Wl,-rpath," + LIB_DIR]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
