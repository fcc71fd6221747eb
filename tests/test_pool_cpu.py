"""CPU: the device-block cache's stream ordering (pptk_amd/csrc/pool.hpp).

tests/pool_check.cpp drives pk::BlockPool with a fake driver: blocks freed
inside a call are held behind a fence on the call's stream, so another stream
(another serving slot) never receives a block an unfinished kernel may still
use, including when an error unwinds the call between launch and
synchronisation. Built with g++ under AddressSanitizer and
UndefinedBehaviorSanitizer.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_block_pool_stream_ordering(tmp_path):
    exe = str(tmp_path / "pool_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Werror",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "pptk_amd", "csrc"),
           os.path.join(ROOT, "tests", "pool_check.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "pool_check: ok" in r.stdout
