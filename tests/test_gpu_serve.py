"""Serving: requests of different sizes, step counts and storage formats
admitted into one running batch as others finish. Each result is the same
bits in the mix as alone on the GPU, and as the oracle's."""
import numpy as np
import pytest

from emsolver_amd import serve, tlm
from tests import helpers

pytestmark = pytest.mark.gpu


def _stream(n=10):
    # FP16 for every 1e-3 request (policy off): the mix holds both formats
    reqs = serve.served_stream(count=n, rate=1e9, lo=12, hi=40, steps=(30, 120), seed=21,
                               policy=False)
    for r in reqs:
        r.arrival = 0.0
    return reqs


def test_served_results_equal_alone_and_oracle():
    reqs = _stream()
    srv = serve.TlmServer(block_ticks=16, max_members=4, hbm_reserve=0,
                          tuning={"planes": 2, "block_waves": 4, "grid_blocks": 0,
                                  "graph_ticks": 8, "planes16": 4})
    done = srv.serve(reqs, realtime=False)
    print(f"[serve] {len(done)} done, mean members {srv.stats['mean_members']:.2f}", flush=True)
    assert sorted(c.rid for c in done) == list(range(len(reqs)))
    assert srv.stats["mean_members"] > 2  # they really overlapped
    assert {c.pulse for c in done} == {"f16", "f32"}
    for c in done:
        sc = reqs[c.rid].build()
        alone = tlm.run_scene(sc)
        o = helpers.oracle_run(sc)
        assert np.abs(o.probes).max() > 0 and np.abs(o.ports).max() > 0
        helpers.assert_bitexact(f"req{c.rid} probes (mix vs alone)", c.probes, alone.probes)
        helpers.assert_bitexact(f"req{c.rid} ports (mix vs alone)", c.ports, alone.ports)
        helpers.assert_bitexact(f"req{c.rid} probes (mix vs oracle)", c.probes, o.probes)
        helpers.assert_bitexact(f"req{c.rid} ports (mix vs oracle)", c.ports, o.ports)


def test_admission_waits_for_memory():
    """With a reserve that leaves room for one request at a time, requests
    are served one after another and still all complete."""
    reqs = _stream(4)
    probe = serve.TlmServer()
    smallest = min(probe._need_bytes(r.build()) for r in reqs)
    free, _ = tlm.device_memory()
    # room for one request beyond the reserve: the next waits until it leaves
    srv = serve.TlmServer(block_ticks=32, hbm_reserve=free - int(1.2 * smallest))
    done = srv.serve(reqs, realtime=False)
    assert len(done) == 4
    assert srv.stats["mean_members"] <= 1.0 + 1e-9


def test_served_device_setup_equals_host_setup():
    """Requests voxelised on the device at admission give the bits of the
    same requests uploaded as host maps."""
    def stream(dev_setup):
        reqs = serve.served_stream(count=6, rate=1e9, lo=12, hi=36, steps=(30, 90), seed=33,
                                   device_setup=dev_setup, policy=False)
        for r in reqs:
            r.arrival = 0.0
        return reqs
    out = {}
    for dev_setup in (False, True):
        srv = serve.TlmServer(block_ticks=16, max_members=3, hbm_reserve=0)
        out[dev_setup] = {c.rid: c for c in srv.serve(stream(dev_setup), realtime=False)}
    assert sorted(out[True]) == list(range(6))
    for rid, c in out[True].items():
        h = out[False][rid]
        assert np.abs(h.ports).max() > 0
        helpers.assert_bitexact(f"req{rid} probes", c.probes, h.probes)
        helpers.assert_bitexact(f"req{rid} ports", c.ports, h.ports)
