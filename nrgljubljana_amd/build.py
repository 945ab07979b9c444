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
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
