"""The reference's C++ smoke test (tdctest/tdctest.cpp:21-59) run through the
drop-in C++ class, no Python binding in between: tests/host/smoke_cpp.cpp,
built by __graft_entry__.build() (make -C tests/host)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "host", "smoke_cpp")


def test_cpp_smoke_runs_on_the_gpu():
    assert os.path.exists(EXE), "tests/host/smoke_cpp is not built: run __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=100)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, (r.returncode, r.stdout[-500:], r.stderr[-500:])
    assert "smoke_cpp ok" in r.stdout
    assert "Completed logging setup" in r.stdout  # the reference's first log line
