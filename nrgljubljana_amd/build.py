"""Build libnrgcore.so (HIP for gfx950) in-tree with hipcc."""
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnrgcore.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
         "-Wno-unused-result", "-Wno-unused-value",
         # no packed FP32 instructions anywhere in this library
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def version():
    try:
        return subprocess.check_output(["git", "-C", HERE, "describe", "--always", "--dirty"],
                                       stderr=subprocess.DEVNULL).decode().strip()
    except Exception:
        return "dev"


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + [f'-DNRGCORE_VERSION="{version()}"'] + sources() + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
    os.replace(OUT + ".tmp", OUT)
    build_run(verbose)
    return OUT


RUN_OUT = os.path.join(HERE, "libnrgrun.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build_run(verbose=False):
    """libnrgrun.so: the whole-calculation driver (host/nrg_run.cpp, C++ on the host) over libnrgcore.so."""
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROCM, "include"), os.path.join(HERE, "host", "nrg_run.cpp"), "-o", RUN_OUT + ".tmp",
           "-L" + HERE, "-lnrgcore", "-Wl,-rpath,$ORIGIN", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lquadmath", "-ldl"]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("building libnrgrun.so failed:\n" + r.stdout + r.stderr)
    os.replace(RUN_OUT + ".tmp", RUN_OUT)
    return RUN_OUT


if __name__ == "__main__":
    print(build(verbose=True))
