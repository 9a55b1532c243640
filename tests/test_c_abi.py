"""The C ABI from C: tests/c_client/abi_client.c is compiled as strict C99 (gcc -std=c99 -pedantic
-Werror) against include/bhrt.h and include/bhrt_host.h, linked against libbhrt.so and run (no device
needed). The struct sizes the C compiler lays out equal the library's own (bhh_abi_sizes), and the
argument errors come back as status codes."""
import os
import shutil
import subprocess

import pytest

import bhrt_amd as bh
from helpers import ROOT

CLIENT = os.path.join(ROOT, "tests", "c_client", "abi_client.c")
PKG = os.path.join(ROOT, "spacetime-optics_amd")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_c_client_builds_links_and_agrees(tmp_path):
    exe = str(tmp_path / "abi_client")
    cmd = ["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           CLIENT, "-L", PKG, "-lbhrt", f"-Wl,-rpath,{PKG}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    lines = run.stdout.split()
    assert run.returncode == 0, run.stdout + run.stderr
    L = bh.load()
    import ctypes as C
    n = L.bhh_abi_sizes(None, 0)
    sizes = (C.c_size_t * n)()
    assert L.bhh_abi_sizes(sizes, n) == n
    assert [int(v) for v in lines[:n]] == list(sizes)
