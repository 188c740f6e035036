"""The reference self test's TLM sections through the C++ TlmSolver facade
(tests/selftest/tlm_selftest.cpp): on a GPU every section must pass; without
one the program reports SKIP and succeeds."""
import os
import subprocess

import pytest

from emsolver_amd import tlm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "selftest", "tlm_selftest")


def _run():
    assert os.path.exists(EXE), "build the self test first (__graft_entry__.build())"
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    return p


@pytest.mark.skipif(tlm.device_count() > 0, reason="a HIP device is present")
def test_selftest_without_device_skips():
    p = _run()
    assert p.returncode == 0 and "SKIP" in p.stdout


@pytest.mark.gpu
def test_selftest_on_gpu():
    p = _run()
    assert "FAIL" not in p.stdout, p.stdout
    assert p.returncode == 0 and "SELFTEST PASSED" in p.stdout
    assert "bit-identical" in p.stdout
