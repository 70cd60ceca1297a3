"""A C++ host program on the C ABI alone (tests/host/host_example.cpp, built
by __graft_entry__.build()): plan forward and exact adjoint equal the
drop-in entry points bit for bit, the matched dot test, graph-replayed
CGLS, and an invalid geometry reported as a status.  It runs as a child
process and must print "host example: OK"."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_host_program(gpu):
    exe = os.path.join(HERE, "host", "host_example")
    assert os.path.exists(exe), "tests/host/host_example is not built (run __graft_entry__.build())"
    print("case: C++ host program on the C ABI", flush=True)
    sys.stdout.flush()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host example: OK" in r.stdout
