"""The C ABI from C++: tests/capi/capi_check (built with the library) runs the twins of the
reference's include/cl_marshal.h - its 3-stage route and the inverse, the AoS/SoA -> ASTA
wrappers, padding - and of include/marshal.h on hipMalloc'd buffers, each checked against a
host loop, with elapsed times accumulated."""
import os
import subprocess
import sys

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "capi", "capi_check")


def test_c_caller_on_the_device():
    assert os.path.exists(EXE), "tests/capi/capi_check is built by __graft_entry__.build()"
    print("capi_check gpu", flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE, "gpu"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "capi gpu ok" in r.stdout
