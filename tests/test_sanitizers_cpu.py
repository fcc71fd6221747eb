"""CPU: the host-side C++ of the library and its checker, built with g++
under AddressSanitizer and UndefinedBehaviorSanitizer
(-fsanitize=address,undefined -fno-sanitize-recover=all) and run:

* tests/oracle_check.cpp: every entry point of the oracle
  (oracle/pptk_oracle.cpp) on edge-case clouds, results against brute force;
* tests/persist_check.cpp: the persisted-tree parser behind
  pptk_tree_deserialize (pptk_amd/csrc/persist_format.hpp) on intact,
  damaged, truncated, misaligned and randomly corrupted buffers;
* tests/pyutil_check.cpp: the CPython/NumPy marshalling of the drop-in
  modules (pptk_amd/csrc/pyext/pyutil.hpp) in an embedded interpreter.

The device allocator's stream ordering has its own sanitizer test
(test_pool_cpu.py).
"""
import os
import shutil
import subprocess
import sysconfig

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-std=c++17", "-O1", "-g", "-Wall", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]

needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _build_and_run(tmp_path, name, srcs, extra=(), env=None, timeout=240):
    exe = str(tmp_path / name)
    cmd = ["g++"] + SAN + list(srcs) + ["-o", exe] + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe], capture_output=True, text=True, timeout=timeout, env=e)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "%s: ok" % name in r.stdout, r.stdout[-2000:]
    return r.stdout


@needs_gxx
def test_oracle_under_sanitizers(tmp_path):
    _build_and_run(tmp_path, "oracle_check",
                   [os.path.join(ROOT, "oracle", "pptk_oracle.cpp"),
                    os.path.join(ROOT, "tests", "oracle_check.cpp"),
                    "-Wno-sign-compare", "-Wno-unknown-pragmas"])


@needs_gxx
def test_persisted_tree_parser_under_sanitizers(tmp_path):
    out = _build_and_run(tmp_path, "persist_check",
                         ["-Werror", "-I", os.path.join(ROOT, "pptk_amd", "csrc"),
                          os.path.join(ROOT, "tests", "persist_check.cpp")])
    assert "fuzz:" in out


@needs_gxx
def test_python_marshalling_under_sanitizers(tmp_path):
    np = pytest.importorskip("numpy")
    cfg = shutil.which("python3-config")
    if cfg is None:
        pytest.skip("needs python3-config to embed the interpreter")
    ld = subprocess.run([cfg, "--ldflags", "--embed"], capture_output=True,
                        text=True).stdout.split()
    inc = ["-I" + sysconfig.get_paths()["include"], "-I" + np.get_include(),
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "pptk_amd", "csrc")]
    _build_and_run(tmp_path, "pyutil_check",
                   inc + [os.path.join(ROOT, "tests", "pyutil_check.cpp")], extra=ld,
                   env={"PYTHONMALLOC": "malloc", "ASAN_OPTIONS": "detect_leaks=0"})
