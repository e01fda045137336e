"""The C ABI from C: tests/c/abi_host.c includes include/cfdx.h, is compiled by gcc as C (GNU
C99, -Wall -Wextra -Werror), links libcfdx.so and runs its host-side entry points (parameters,
struct layouts, C-grid geometry and metrics, block sizing and error reporting, tile shapes,
restart files). No device call."""
import os
import shutil
import subprocess

import pytest

import cfdx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_header_compiles_as_c_and_host_calls_work(tmp_path):
    lib_dir = os.path.dirname(cfdx.LIB_PATH)
    exe = str(tmp_path / "abi_host")
    cmd = ["gcc", "-std=gnu99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "abi_host.c"), "-L", lib_dir, "-lcfdx", "-lm", f"-Wl,-rpath,{lib_dir}",
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ALL OK"), out.stdout
