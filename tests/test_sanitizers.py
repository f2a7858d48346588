"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code:

* the oracle (oracle/Makefile target `asan`), with the oracle's test files run
  against it in a child process (libasan preloaded into the interpreter);
* the product's host graph algorithms (matlab_ufvm_amd/csrc/hostalg.cpp:
  agglomeration, Galerkin gather lists, colourings, level schedules, DILU lower
  lists) in tests/native/hostalg_check.cpp, which also compares the
  agglomeration with the oracle's.

Any sanitizer report aborts the child with a nonzero status."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]


def _libasan():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    if p.returncode or not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("no libasan for this compiler")
    return path


def test_oracle_tests_under_sanitizers():
    asan = _libasan()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               UFVM_ORACLE_LIB=os.path.join(ROOT, "oracle", "libufvm_oracle_asan.so"))
    files = ["tests/test_oracle.py", "tests/test_simple_oracle.py", "tests/test_tutorials.py",
             "tests/test_fieldio.py", "tests/test_mesh_geometry.py"]
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", *files], cwd=ROOT, env=env, capture_output=True,
                       text=True)
    assert p.returncode == 0, (p.stdout[-3000:], p.stderr[-3000:])


def test_host_graph_algorithms_under_sanitizers(tmp_path):
    _libasan()
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "hostalg_check")
    subprocess.run(["gcc", "-std=c11", *SAN, "-ffp-contract=off", "-c",
                    os.path.join(ROOT, "oracle", "ufvm_oracle.c"), "-o", obj], check=True)
    subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "matlab_ufvm_amd", "csrc"),
                    "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "native", "hostalg_check.cpp"),
                    os.path.join(ROOT, "matlab_ufvm_amd", "csrc", "hostalg.cpp"), obj, "-lm",
                    "-o", exe], check=True)
    p = subprocess.run([exe], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"),
                       capture_output=True, text=True)
    assert p.returncode == 0, (p.stdout, p.stderr[-3000:])
    assert "0 failures" in p.stdout
