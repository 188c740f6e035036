"""Scene -> voxel set-up on the device: the material maps written by the
HIP voxelisers equal the oracle's (VoxelizeMesh / VoxelizeMeshSolid /
VoxelizeWire / the feed carve) bit for bit, the device-built cell codes,
faces and Huygens lists equal the host-built ones, and a scene set up on the
device runs to the same bits as the same scene uploaded from the host and as
the CPU oracle."""
import numpy as np
import pytest

from emsolver_amd import scenes, tlm
from oracle import oracle_tlm as ot
from tests import helpers

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[voxel] {msg}", flush=True)


def _random_mesh_ops(rng, dims, origin, dl):
    """A few objects partly outside the grid: boxes, rotated boxes, spheres,
    plates (solid and surface), wires of several radii."""
    L = np.array(dims) * dl
    o = np.array(origin)
    ops = []
    for q in range(8):
        kind = q % 4
        mid = int(rng.integers(1, 6))
        if kind == 0:
            lo = o + rng.uniform(-0.2, 0.7, 3) * L
            hi = lo + rng.uniform(0.1, 0.6, 3) * L
            v, ix = scenes.box_mesh(lo, hi)
            ops.append(("mesh", v, ix, (0.0, 0.0, 0.0), mid, bool(rng.integers(0, 2))))
        elif kind == 1:
            half = rng.uniform(0.05, 0.3, 3) * L
            v, ix = scenes.box_mesh(-half, half)
            v = (v.astype(np.float64) @ scenes.rotation(*rng.uniform(-1.5, 1.5, 3)).T)
            cen = o + rng.uniform(0.0, 1.0, 3) * L
            ops.append(("mesh", v.astype(np.float32), ix, tuple(np.float32(cen)), mid, True))
        elif kind == 2:
            v, ix = scenes.icosphere(float(rng.uniform(0.1, 0.35)) * L.min(), int(rng.integers(1, 4)))
            cen = o + rng.uniform(0.1, 0.9, 3) * L
            ops.append(("mesh", v, ix, tuple(np.float32(cen)), mid, bool(rng.integers(0, 2))))
        else:
            pts = (o + rng.uniform(-0.1, 1.1, (int(rng.integers(2, 5)), 3)) * L).astype(np.float32)
            r = float(np.float32(rng.choice([0.0, 0.3, 0.8, 1.6]) * dl))
            ops.append(("wire", pts, r, (0.0, 0.0, 0.0), mid))
    # a surface plate and a feed
    c = o + 0.5 * L
    pv, pi = scenes.plate_mesh(c, (0.0, 0.7 * L[1], 0.1 * L[2]), (0.2 * L[0], 0.0, 0.6 * L[2]))
    ops.append(("mesh", pv, pi, (0.0, 0.0, 0.0), 1, False))
    a = (c + np.array([0.0, 0.0, -1.3 * dl])).astype(np.float32)
    b = (c + np.array([0.2 * dl, 0.1 * dl, 1.3 * dl])).astype(np.float32)
    ops.append(("feed", tuple(a), tuple(b), (0.0, 0.0, 0.0), 1.0))
    return ops


@pytest.mark.parametrize("seed", range(6))
def test_device_material_map_equals_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    dims = tuple(int(x) for x in rng.integers(17, 61, 3))
    origin = tuple(float(np.float32(x)) for x in rng.uniform(-0.5, 0.5, 3))
    dl = float(np.float32(rng.uniform(0.01, 0.03)))
    sc = scenes.Scene(*dims, None, dl=dl, origin=origin, name=f"rand{seed}")
    sc.ops = _random_mesh_ops(rng, dims, origin, dl)
    _say(f"seed {seed} dims {dims} ops {len(sc.ops)}")
    host = helpers.oracle_build(sc)
    with tlm.TlmDevice(*dims, dl) as dev:
        feeds = tlm.apply_ops(dev, sc)
        got = dev.mat_read()
    assert len(np.unique(host.mat)) >= 4, "degenerate case"
    bad = np.flatnonzero(got != host.mat)
    assert bad.size == 0, f"{bad.size} cells differ, first {bad[:5]}: device {got[bad[:5]]} " \
                          f"oracle {host.mat[bad[:5]]}"
    assert [(list(c), p, a) for c, p, a in feeds] == host.ports


def test_each_voxeliser_alone():
    """One call of each kind on an empty map, including a mesh with unused
    vertices (the solid's bounding box includes them, Geometry.cpp:559-561)
    and a mesh wholly outside the grid."""
    dims, origin, dl = (40, 33, 29), (-0.2, 0.1, 0.05), 0.02
    L = np.array(dims) * dl
    o = np.array(origin)
    sv, si = scenes.icosphere(0.25 * L.min(), 3)
    extra = np.vstack([sv, (o + 0.95 * L).astype(np.float32)[None]])
    # ten slabs stacked along x in one mesh: rows cross it 20 times, more
    # than the one-pass fill keeps per row (the two-pass fill takes over)
    lv, li = [], []
    for q in range(10):
        v, ix = scenes.box_mesh(o + L * np.array([0.05 + 0.09 * q, 0.1, 0.1]),
                                o + L * np.array([0.09 + 0.09 * q, 0.9, 0.8]))
        li.append(ix + 8 * q)
        lv.append(v)
    layered = np.vstack(lv).astype(np.float32), np.vstack(li).astype(np.int32)
    cases = {
        "layered solid": ("mesh", layered[0], layered[1], (0.0, 0.0, 0.0), 8, True),
        "sphere solid": ("mesh", sv, si, tuple(np.float32(o + 0.5 * L)), 3, True),
        "sphere surface": ("mesh", sv, si, tuple(np.float32(o + 0.5 * L)), 3, False),
        "unused vertex": ("mesh", extra, si, tuple(np.float32(o + 0.4 * L)), 4, True),
        "outside": ("mesh", sv, si, tuple(np.float32(o + 3.0 * L)), 5, True),
        "thick wire": ("wire", (o + np.array([[0.1, 0.2, 0.3], [0.8, 0.7, 0.5]]) * L
                                ).astype(np.float32), 2.3 * dl, (0.0, 0.0, 0.0), 6),
        "thin polyline": ("wire", (o + np.array([[0.1, 0.1, 0.1], [0.9, 0.2, 0.3],
                                                  [0.5, 0.9, 0.9]]) * L).astype(np.float32),
                          0.0, (0.0, 0.0, 0.0), 7),
    }
    for name, op in cases.items():
        sc = scenes.Scene(*dims, None, dl=dl, origin=origin)
        sc.ops = [op]
        host = helpers.oracle_build(sc)
        with tlm.TlmDevice(*dims, dl) as dev:
            tlm.apply_ops(dev, sc)
            got = dev.mat_read()
        _say(f"{name}: {int((host.mat != 0).sum())} cells")
        if name == "outside":
            assert not host.mat.any()
        else:
            assert host.mat.any()
        assert np.array_equal(got, host.mat), name


def test_device_codes_faces_and_huygens_equal_host_build():
    sc = scenes.mesh_scene(48, steps=4)
    host = helpers.oracle_build(sc)
    yh, gh = tlm.stub_admittances(sc.epsr, sc.sigma, sc.dl)
    with tlm.TlmDevice(sc.nx, sc.ny, sc.nz, sc.dl) as dev:
        tlm.apply_ops(dev, sc)
        dev.upload_mat(yh, gh)
        nf = dev.pec_faces_mat()
        nh = dev.huygens_mat(sc.huy_offset)
        fc, fa = dev.read_faces()
        hc, ha = dev.read_huygens_faces()
    c, a, _ = tlm.pec_faces(sc.nx, sc.ny, sc.nz, host.mat)
    h, b, _ = tlm.huygens_faces(sc.nx, sc.ny, sc.nz, host.mat, sc.huy_offset)
    _say(f"faces {nf} huygens {nh}")
    assert nf == c.size > 0 and nh == h.size > 0
    assert np.array_equal(fc, c) and np.array_equal(fa, a)
    assert np.array_equal(hc, h) and np.array_equal(ha, b)


def test_mesh_scene_runs_like_host_upload_and_oracle():
    """mesh_scene set up on the device == the oracle-built map uploaded from
    the host == the CPU oracle: state, probes, port, Js, Huygens, energy."""
    sc = scenes.mesh_scene(40, steps=160, huy_offset=4)
    host = helpers.oracle_build(sc)
    assert (host.mat >= 2).any() and (host.mat == 1).any() and len(host.ports) == 1
    dev_res = tlm.run_scene(sc)
    host_res = tlm.run_scene(host)
    o = helpers.oracle_run(host)
    _say(f"mesh_scene 40^3: {len(host.ports[0][0])} feed cells, probes max "
         f"{np.abs(o.probes).max():.3e}")
    helpers.compare_results(dev_res, o, host)
    helpers.compare_results(host_res, o, host)


@pytest.mark.parametrize("pulse", ["f32", "f16"])
def test_device_built_dipole_equals_host_dipole(pulse):
    """dipole_scene(device_setup=True) (cell boxes + feed carve on the
    device) is the host dipole_scene: same map, same results."""
    kw = dict(length_frac=0.5, steps=200, background_boxes=5, seed=9)
    d = scenes.dipole_scene(36, 30, 44, device_setup=True, **kw)
    h = scenes.dipole_scene(36, 30, 44, **kw)
    for s in (d, h):
        s.pulse = pulse
        s.probes = np.array([h.cell(5, 7, 9), h.cell(30, 20, 30)], np.int64)
    with tlm.TlmDevice(d.nx, d.ny, d.nz, d.dl) as dev:
        feeds = tlm.apply_ops(dev, d)
        assert np.array_equal(dev.mat_read(), h.mat)
    assert [(c, p, a) for c, p, a in feeds] == [(list(h.ports[0][0]), 2, 1.0)]
    rd = tlm.run_scene(d)
    rh = tlm.run_scene(h)
    _say(f"dipole {pulse}: ports max {np.abs(rh.ports).max():.3e}")
    assert np.abs(rh.ports).max() > 0
    for name in ("probes", "ports", "V"):
        helpers.assert_bitexact(f"{pulse} {name}", getattr(rd, name), getattr(rh, name))
    assert np.array_equal(rd.js_inst, rh.js_inst)


def test_set_up_errors():
    with tlm.TlmDevice(8, 8, 8, 0.01) as dev:
        v, ix = scenes.box_mesh((0.01, 0.01, 0.01), (0.05, 0.05, 0.05))
        bad = ix.copy()
        bad[3, 1] = 8
        with pytest.raises(tlm.TlmError, match="out of range"):
            dev.voxelize_mesh((0, 0, 0), 0.01, v, bad, (0, 0, 0), 2, True)
        with pytest.raises(tlm.TlmError, match="coordinates out of range"):
            dev.voxelize_mesh((0, 0, 0), 0.01, v * np.float32(1e12), ix, (0, 0, 0), 2, True)
        dev.mat_cells_box((2, 2, 2), (4, 4, 4), 5)
        with pytest.raises(tlm.TlmError, match="yhat/ghat"):
            dev.upload_mat([1.0], [0.0])  # id 5 needs four entries
        dev.upload_mat([0.0, 0.0, 0.0, 4.0], [0.0] * 4)
        m = dev.mat_read()
        assert (m == 5).sum() == 8 and set(np.unique(m)) == {0, 5}


def test_degenerate_inputs_match_oracle():
    """Empty meshes and wires, a single triangle asked for a solid (surface
    only: no ray pair), a feed outside the grid, a box clipped away."""
    dims, origin, dl = (20, 18, 16), (0.0, 0.0, 0.0), 0.05
    tri = np.float32([[0.2, 0.2, 0.3], [0.7, 0.3, 0.35], [0.4, 0.75, 0.4]])
    sc = scenes.Scene(*dims, None, dl=dl, origin=origin)
    sc.ops = [("mesh", np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32), (0, 0, 0), 3, True),
              ("mesh", tri, np.int32([[0, 1, 2]]), (0, 0, 0), 4, True),
              ("wire", np.float32([[0.3, 0.3, 0.3]]), 0.1, (0, 0, 0), 5),
              ("wire", np.zeros((0, 3), np.float32), 0.0, (0, 0, 0), 5),
              ("cells_box", (30, 0, 0), (40, 5, 5), 6),
              ("cells_box", (-5, -5, -5), (2, 2, 2), 7),
              ("feed", (5.0, 5.0, 5.0), (5.0, 5.0, 5.2), (0, 0, 0), 1.0)]
    host = helpers.oracle_build(sc)
    with tlm.TlmDevice(*dims, dl) as dev:
        feeds = tlm.apply_ops(dev, sc)
        got = dev.mat_read()
    assert feeds == [] and host.ports == []
    assert set(np.unique(host.mat)) == {0, 4, 7}
    assert np.array_equal(got, host.mat)
