"""Builds the C++ host-layer test programs (tests/cpp/test_*.cpp -> tests/cpp/bin/).

They include the host headers (nrgljubljana_amd/host/nrg), link libnrgcore.so and the HIP runtime,
and the GPU program also links the oracle (oracle/liboracle.so) to check its results. Run paths are
relative to the binaries so that the tree can be moved."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

PROGRAMS = {"test_host_cpu": False, "test_host_gpu": True}   # name -> links the oracle


def build(verbose=False):
    os.makedirs(BIN, exist_ok=True)
    for name, with_oracle in PROGRAMS.items():
        src = os.path.join(HERE, name + ".cpp")
        out = os.path.join(BIN, name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
               "-I" + HERE, "-I" + os.path.join(ROOT, "nrgljubljana_amd", "host"), "-I" + os.path.join(ROCM, "include"),
               src, "-o", out + ".tmp",
               "-L" + os.path.join(ROOT, "nrgljubljana_amd"), "-lnrgcore",
               "-Wl,-rpath,$ORIGIN/../../../nrgljubljana_amd"]
        if with_oracle:
            cmd += ["-L" + os.path.join(ROOT, "oracle"), "-loracle", "-Wl,-rpath,$ORIGIN/../../../oracle"]
        cmd += ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lquadmath"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"building {name} failed")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose=True)
