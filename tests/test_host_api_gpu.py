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


@pytest.mark.gpu
def test_qlten_files_of_the_host_api_match_the_python_restatement(tmp_path):
    """The C++ host layer (operator<<) writes a U(1)xU(1) complex and a U(1)
    real tensor; the Python reader reads them and the Python writer
    reproduces the files byte for byte."""
    import io
    from tensortoolkit_amd import qlten_io as Q
    print("case: .qlten files of tests/cpp/test_host_api vs tensortoolkit_amd.qlten_io", flush=True)
    path = tmp_path / "u1u1.qlten"
    env = dict(os.environ, QLTEN_IO_OUT=str(path))
    proc = subprocess.run([str(EXE)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=110,
                          env=env)
    assert proc.returncode == 0, proc.stdout[-4000:]
    for p, dtype, qn_len in ((path, "c128", 2), (path.with_name(path.name + ".real"), "f64", 1)):
        raw = p.read_bytes()
        t = Q.read(io.BytesIO(raw), dtype, qn_len)
        assert t.n_blks > 0 and t.raw_size > 0
        buf = io.BytesIO()
        Q.write(t, buf)
        assert buf.getvalue() == raw, p.name
