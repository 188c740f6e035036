"""Scene builders on the CPU: a scene built as device set-up operations is
the same grid as the host-map scene (its operations replayed by the
oracle's voxelisers and feed carve), the served stream is reproducible, and
its summary arithmetic is right."""
import numpy as np
import pytest

from emsolver_amd import scenes, serve
from tests import helpers


@pytest.mark.parametrize("dims,frac,boxes,seed", [((36, 30, 44), 0.5, 5, 9),
                                                    ((32, 41, 57), 0.3, 8, 2),
                                                    ((64, 33, 38), 0.7, 0, 4)])
def test_dipole_ops_equal_host_dipole(dims, frac, boxes, seed):
    kw = dict(length_frac=frac, steps=100, background_boxes=boxes, seed=seed)
    d = scenes.dipole_scene(*dims, device_setup=True, **kw)
    h = scenes.dipole_scene(*dims, **kw)
    r = helpers.oracle_build(d)
    assert np.array_equal(r.mat, h.mat)
    assert r.ports == [(list(h.ports[0][0]), 2, 1.0)]
    assert np.array_equal(d.epsr, h.epsr) and np.array_equal(d.sigma, h.sigma)
    assert d.has_diel == h.has_diel == (boxes > 0)


def test_request_scene_host_and_ops_agree():
    for dims, st, pulse, sd in (((40, 35, 50), 300, "f16", 3), ((33, 60, 41), 500, "f32", 8)):
        tol = 1e-3 if pulse == "f16" else 1e-5
        a = serve.request_scene(dims, 0.5, st, tol, sd, policy=False)
        b = serve.request_scene(dims, 0.5, st, tol, sd, device_setup=True, policy=False)
        assert np.array_equal(a.probes, b.probes)
        rb = helpers.oracle_build(b)
        assert np.array_equal(rb.mat, a.mat) and rb.ports == [(list(a.ports[0][0]), 2, 1.0)]
        assert a.pulse == b.pulse == pulse and a.total_steps == b.total_steps == st
        # probes keep off the dipole column, the only PEC cells
        assert (a.mat[a.probes] != scenes.MAT_PEC).all()


def test_served_stream_is_reproducible():
    a = serve.served_stream(count=50, seed=5)
    b = serve.served_stream(count=50, seed=5)
    assert [(r.rid, r.arrival, r.tol) for r in a] == [(r.rid, r.arrival, r.tol) for r in b]
    t = np.array([r.arrival for r in a])
    # 50 Poisson arrivals at 50 per second: about one second
    assert (np.diff(t) > 0).all() and 0.5 < t[-1] < 2.0
    assert {r.tol for r in a} == {1e-3, 1e-5}
    sa, sb = a[7].build(), b[7].build()
    assert sa.name == sb.name and np.array_equal(sa.mat, sb.mat)
    assert 32 <= min(sa.nx, sa.ny, sa.nz) and max(sa.nx, sa.ny, sa.nz) <= 384


def test_summarize():
    done = [serve.Completion(i, float(i), float(i), float(i) + 1.0 + i, 1000, 10, "f32")
            for i in range(5)]
    s = serve.summarize(done)
    lat = np.array([1.0 + i for i in range(5)])
    assert s["requests"] == 5 and s["p50_latency_s"] == np.percentile(lat, 50)
    assert s["cell_updates"] == 5 * 1000 * 10
    assert abs(s["gcell_s"] - 5e4 / (9.0 - 0.0) / 1e9) < 1e-15


def test_plane_wave_sources_are_the_air_cells_of_the_plane():
    """add_plane_wave follows TlmSolver::addPlaneWave (TlmSolver.cpp:394-407):
    the AIR cells with co[prop_axis] = plane_index, in cell order."""
    sc = scenes.dielectric_boxes(n=20, nboxes=3, steps=10, nprobes=2, nfreq=1, seed=2)
    for axis in range(3):
        sc.sources = []
        n = scenes.add_plane_wave(sc, axis, 7, (axis + 1) % 3, 0.5)
        want = [sc.cell(*co) for co in
                [(i, j, k) for k in range(sc.nz) for j in range(sc.ny) for i in range(sc.nx)]
                if co[axis] == 7 and sc.mat[sc.cell(*co)] == scenes.MAT_AIR]
        assert n == len(want) > 0
        assert [c for c, _, _ in sc.sources] == want
        assert all(p == (axis + 1) % 3 and a == 0.5 for _, p, a in sc.sources)


def test_storage_policy_rules():
    """serve.plan_storage: FP32 below the 1e-3 tolerance; otherwise an FP32
    band of the source slabs +- halo that holds every slab to keep with
    F16_MARGIN to spare, FP16 only when the grid is within F16_MAX_AXIS and
    at least F16_MIN_SHARE of the slabs stay 16-bit."""
    assert serve.plan_storage(1e-5, 100, (100, 100, 100), [50], [50])[0] == "f32"
    pulse, halo, share = serve.plan_storage(1e-3, 100, (100, 100, 100), [50], [45, 55, 48])
    assert pulse == "f16" and halo == 5 + serve.F16_MARGIN
    assert abs(share - (1 - (2 * halo + 1) / 100)) < 1e-12
    # a probe far from the source widens the band until FP16 no longer pays
    pulse, halo, share = serve.plan_storage(1e-3, 100, (100, 100, 100), [50], [45, 55, 90])
    assert pulse == "f32" and share < serve.F16_MIN_SHARE
    # beyond the sizes the error model was checked on
    big = serve.F16_MAX_AXIS + 1
    assert serve.plan_storage(1e-3, 100, (big, 100, 100), [50], [50])[0] == "f32"


def test_served_storage_same_for_host_and_device_setup():
    """A request gets the same storage and FP32 band whether its scene is a
    host map or set-up operations; the stream holds FP16 requests only where
    the policy allows, FP32 otherwise."""
    host = serve.served_stream(count=120, lo=32, hi=96, seed=5)
    dev = serve.served_stream(count=120, lo=32, hi=96, seed=5, device_setup=True)
    n16 = 0
    for a, b in zip(host, dev):
        sa, sb = a.build(), b.build()
        assert (sa.pulse, sa.fp32_halo) == (sb.pulse, sb.fp32_halo)
        if sa.pulse == "f16":
            n16 += 1
            assert a.tol >= serve.F16_TOL
            assert max(sa.nx, sa.ny, sa.nz) <= serve.F16_MAX_AXIS
            k0, k1, g0, gap = sa.dipole
            slab = sa.nx * sa.ny
            for k in [k0, k1] + [int(c) // slab for c in sa.probes]:
                assert g0 - sa.fp32_halo + serve.F16_MARGIN <= k <= g0 + gap - 1 + sa.fp32_halo - serve.F16_MARGIN
    assert n16 >= 1


class _FakeDev:
    def __init__(self, sc):
        self.sc, self.nprobe, self.nport, self.closed = sc, 0, 0, False

    def close(self):
        self.closed = True


class _FakeBatch:
    """Stands in for tlm.TlmBatch: members finish after their steps; the
    first FP16 member to leave reports a 16-bit range error, as
    tlm_batch_remove does for a member whose pulses outgrew a slab scale."""
    failed = []

    def __init__(self, device):
        self.left = {}

    def set_tuning(self, **kw):
        pass

    def get_tuning(self):
        return {}

    def add(self, dev, steps):
        self.left[dev] = steps

    def run(self, ticks):
        for d in self.left:
            self.left[d] = max(0, self.left[d] - ticks)

    def sync(self):
        pass

    def remaining(self):
        return dict(self.left)

    def remove(self, dev):
        self.left.pop(dev)
        if dev.sc.pulse == "f16" and not _FakeBatch.failed:
            _FakeBatch.failed.append(dev)
            raise serve.tlm.TlmError("tlm_batch_remove failed (-7)", serve.tlm.ERR_RANGE)

    def close(self):
        pass


def test_server_reruns_a_range_failed_member_in_fp32(monkeypatch):
    """A member whose 16-bit storage fails leaves the batch alone; the server
    solves that request again in FP32 and every request completes once."""
    _FakeBatch.failed = []
    monkeypatch.setattr(serve.tlm, "TlmBatch", _FakeBatch)
    monkeypatch.setattr(serve.tlm, "open_scene", lambda sc, device: _FakeDev(sc))
    monkeypatch.setattr(serve.tlm, "device_memory", lambda device: (1 << 40, 1 << 40))
    reqs = []
    for rid in range(6):
        sc = scenes.Scene(8, 8, 8, np.zeros(512, np.uint8), name=f"r{rid}",
                          total_steps=40 + 10 * rid, pulse="f16" if rid % 2 else "f32")
        reqs.append(serve.Request(rid, sc, 0.0))
    srv = serve.TlmServer(block_ticks=16)
    done = srv.serve(reqs, realtime=False)
    assert sorted(c.rid for c in done) == list(range(6))
    bad = _FakeBatch.failed[0]
    assert bad.closed and srv.stats["retried_f32"] == [1]
    byrid = {c.rid: c for c in done}
    assert byrid[1].pulse == "f32" and byrid[3].pulse == "f16" and byrid[0].pulse == "f32"
