"""The search kernel's binary searches (paracosm_amd/csrc/csm_search.hpp),
compiled for the host and checked against exhaustive searches: lower bound,
membership with version windows (per-lane and wave-shared segment forms) and
the LDS membership test, on random ascending arrays with repeated values."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_search_primitives_match_exhaustive_search(tmp_path):
    src = os.path.join(HERE, "search_primitives_check.cpp")
    exe = str(tmp_path / "search_primitives_check")
    try:
        subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe], check=True,
                       capture_output=True, text=True)
    except FileNotFoundError:
        pytest.skip("no host C++ compiler")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("bad 0")
