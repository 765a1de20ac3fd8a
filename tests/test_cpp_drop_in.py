"""The C++ drop-in classes (include/sph_amd.hpp) compile against the C-ABI
and, on a GPU, run the reference's call pattern (tests/cpp/drop_in_test.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "drop_in_test.cpp")
LIBDIR = os.path.join(ROOT, "splishsplash_amd")


def build(tmp_path):
    exe = str(tmp_path / "drop_in_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), SRC,
                    "-L", LIBDIR, "-lsph_amd", "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def test_cpp_drop_in_compiles_and_links(tmp_path):
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_cpp_drop_in_runs(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
