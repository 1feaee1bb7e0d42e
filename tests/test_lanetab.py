"""The register coil stage's adjoint lane table (disorder_amd/csrc/lanetab.hpp)
on the CPU: tests/native/lanetab_harness.cpp (built here with g++) builds the
table for random states (empty, one to ten samples, 768 or 1,536, any count
up to 2,000, half of the samples in one row; one block per chunk of 768
row-sorted samples) and runs the kernel's use of each block on integer
sample values; every row must equal the direct sum of its chunk's samples
and be written once, every sample dealt once.  A flipped close bit must be
caught in every block it is planted in."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("lanetab") / "lanetab_harness")
    subprocess.run([gxx, "-std=c++17", "-O2", "-o", out, os.path.join(HERE, "native", "lanetab_harness.cpp")],
                   check=True)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_lane_table_rows_equal_direct_sums(harness, seed):
    r = subprocess.run([harness, "5000", str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok ") and int(r.stdout.split()[1]) > 100000


def test_lane_table_check_catches_a_flipped_close_bit(harness):
    r = subprocess.run([harness, "400", "3", "plant"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    last = r.stdout.strip().splitlines()[-1].split()  # "planted P caught C"
    assert last[0] == "planted" and int(last[1]) > 100 and last[3] == last[1]
