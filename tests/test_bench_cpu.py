"""CPU checks of the benchmark's contract and of the models behind its roofline figures."""
import os
import re
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402
from nrgljubljana_amd import workloads  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_panel_constants_match_the_library():
    """bench.py counts the reduction's launches and bytes with the library's panel width and tail
    order; both are restated there, so they must agree with the kernels' constants."""
    src = open(os.path.join(ROOT, "nrgljubljana_amd", "csrc", "syev.hip")).read()
    nb = int(re.search(r"constexpr int NB = (\d+);", src).group(1))
    tail = int(re.search(r"constexpr int TAIL = (\d+);", src).group(1))
    assert tail == bench.TAIL
    assert bench.main_panels.__defaults__ == (nb, tail)
    assert bench.reduction_bytes.__defaults__[0] == nb


def test_launch_model_matches_the_trace():
    """Launches per call of the two column kernels on the 256-block workload, as the kernel trace in
    profiles/syev_fp64_256blk_kernels.md counts them (22176 preparation launches, over three workgroup
    sizes, and 21504 product launches over 8 calls with three chains; an earlier two-chain trace had
    14784 and 14336): the roofline's bytes per launch divide by these counts."""
    sizes = workloads.eig_sizes()
    three = bench.reduction_bytes(sizes, groups=3)
    assert three["k_col_prep"][1] == 22176 // 8
    assert three["k_col_symv"][1] == 21504 // 8
    two = bench.reduction_bytes(sizes, groups=2)
    assert two["k_col_prep"][1] == 14784 // 8
    assert two["k_col_symv"][1] == 14336 // 8
    # the bytes do not depend on how the blocks are dealt into chains
    one = bench.reduction_bytes(sizes, groups=1)
    for k in one:
        assert one[k][0] == two[k][0]


def test_launches_of_one_block():
    """One block of order 200 has three column-kernel panels (200 - 32 p > 128 for p < 3), each with
    33 preparation and 32 product launches; blocks at or below the tail order never reach them."""
    got = bench.reduction_bytes([200])
    assert bench.main_panels(200) == 3
    assert got["k_col_prep"][1] == 3 * 33
    assert got["k_col_symv"][1] == 3 * 32
    assert bench.reduction_bytes([128, 16]) == {"k_col_prep": (0.0, 0), "k_col_symv": (0.0, 0)}


def test_workload_is_reproducible():
    a, b = workloads.eig_sizes(), workloads.eig_sizes()
    assert np.array_equal(a, b)
    assert len(a) == 256 and a.min() >= 16 and a.max() <= 1024
    assert workloads.eig_flops([3]) == 90.0   # (10/3) n^3
