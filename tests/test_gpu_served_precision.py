"""The served stream's precision contract: a request that tolerates 1e-3 of
probe error is stored in FP16 only where the storage policy's error model
(serve.plan_storage) says FP16 meets it, and then it does. A seeded sample
of served-stream requests small enough for the FP64 oracle (axes <= 96),
every one of the sample's FP16-served requests and FP32 fallbacks, each run
on the GPU at its full step count: max over the run of |x - x64| / max |x64|
per probe <= 1e-3 for the FP16-served requests against the FP64 oracle
(TlmSolver.cpp:705-713 is the probe), and the FP32-served ones at FP32
accuracy."""
import numpy as np
import pytest

from emsolver_amd import serve, tlm
from tests import helpers

pytestmark = pytest.mark.gpu


def _sample():
    reqs = serve.served_stream(count=200, rate=50.0, lo=32, hi=96, seed=5)
    f16, fallback = [], []
    for r in reqs:
        if r.tol < serve.F16_TOL:
            continue
        sc = r.build()
        (f16 if sc.pulse == "f16" else fallback).append(sc)
    return f16, fallback[:4]


F16, FALLBACK = _sample()


def _probe_err(sc):
    g = tlm.run_scene(sc, state=False)
    r64 = helpers.oracle_run(sc, fp64=True)
    e = np.abs(g.probes - r64.probes).max(axis=(0, 2)) / np.abs(r64.probes).max(axis=(0, 2))
    ports = np.abs(g.ports - r64.ports).max(axis=0)[0] / np.abs(r64.ports).max(axis=0)[0]
    return e, ports


def test_sample_holds_fp16_requests():
    assert len(F16) >= 3 and len(F16) + len(FALLBACK) >= 8


@pytest.mark.parametrize("sc", F16, ids=lambda s: f"{s.name}_{s.nx}x{s.ny}x{s.nz}_{s.total_steps}")
def test_fp16_served_request_within_1e3(sc):
    assert sc.pulse == "f16"
    print(f"[served] {sc.name} {sc.nx}x{sc.ny}x{sc.nz} steps {sc.total_steps} "
          f"FP32 halo {sc.fp32_halo}", flush=True)
    e, ports = _probe_err(sc)
    print(f"  probe max rel err {np.round(e, 6)}  port V,I {np.round(ports, 7)}")
    assert e.max() <= 1e-3
    assert ports.max() <= 1e-3


@pytest.mark.parametrize("sc", FALLBACK, ids=lambda s: f"{s.name}_{s.nx}x{s.ny}x{s.nz}_{s.total_steps}")
def test_fp32_fallback_request(sc):
    assert sc.pulse == "f32"
    print(f"[served] {sc.name} {sc.nx}x{sc.ny}x{sc.nz} steps {sc.total_steps} (FP32 fallback)",
          flush=True)
    e, ports = _probe_err(sc)
    print(f"  probe max rel err {np.round(e, 8)}")
    assert e.max() <= 1e-4 and ports.max() <= 1e-4
