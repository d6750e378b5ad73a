"""Runs the C++ test program of the qlten host API (tests/cpp/test_host_api.cc,
built by __graft_entry__.build()) on the GPU."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "bin" / "test_host_api"


def test_host_api_binary_is_built():
    assert EXE.exists(), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_host_api_on_device():
    print("case: tests/cpp/test_host_api", flush=True)
    proc = subprocess.run([str(EXE)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=110,
                          env=dict(os.environ))
    print(proc.stdout[-4000:])
    assert proc.returncode == 0, proc.stdout[-4000:]
