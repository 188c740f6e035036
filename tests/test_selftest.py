"""The reference self test's TLM sections through the C++ TlmSolver facade
(tests/selftest/tlm_selftest.cpp) and the cancellation program
(tests/selftest/tlm_stop.cpp): on a GPU every section must pass; without one
the programs report SKIP and succeed."""
import os
import subprocess

import pytest

from emsolver_amd import tlm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "selftest", "tlm_selftest")
STOP = os.path.join(ROOT, "tests", "selftest", "tlm_stop")


def _run(exe=EXE):
    assert os.path.exists(exe), "build the test programs first (__graft_entry__.build())"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    return p


@pytest.mark.skipif(tlm.device_count() > 0, reason="a HIP device is present")
@pytest.mark.parametrize("exe", [EXE, STOP], ids=["selftest", "stop"])
def test_programs_without_device_skip(exe):
    p = _run(exe)
    assert p.returncode == 0 and "SKIP" in p.stdout


@pytest.mark.gpu
def test_selftest_on_gpu():
    p = _run()
    assert "FAIL" not in p.stdout, p.stdout
    assert p.returncode == 0 and "SELFTEST PASSED" in p.stdout
    assert "bit-identical" in p.stdout


@pytest.mark.gpu
def test_stop_mid_run_on_gpu():
    """requestStop() from a second thread (tests/selftest/tlm_stop.cpp): the
    run ends between blocks, run() finalises the DFTs and marks the solver
    finished, and every result equals a run configured to end at that step;
    a second run() on the stopped solver (the drop-in entered at curStep > 0,
    the stop request still standing) changes nothing."""
    p = _run(STOP)
    assert "FAIL" not in p.stdout, p.stdout
    assert p.returncode == 0 and "STOPTEST PASSED" in p.stdout
