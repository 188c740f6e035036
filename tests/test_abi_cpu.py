"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/GpuTlmHip.h, validates arguments before touching a
device, reports a missing device the way RunGpuTlm's caller expects
(GpuTlm.cpp:535-548: an error code plus a message, never a crash), and its
host helpers agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from emsolver_amd import _native, scenes, tlm
from oracle import oracle_tlm as ot
from tests import helpers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "GpuTlmHip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tlm_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = _native.load()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # and the ctypes table covers exactly the declared set
    assert sorted(_native.SIGNATURES) == names


def test_library_is_the_in_tree_build():
    assert os.path.exists(_native.LIB_PATH)
    assert _native.LIB_PATH.startswith(os.path.join(ROOT, "emsolver_amd", "lib"))


def _create(nx, ny, nz, pulse=0, dl=0.01):
    lib = _native.load()
    d = _native.GridDesc(nx, ny, nz, dl, 0.0, pulse, 0, 0)
    h = C.c_void_p()
    rc = lib.tlm_create(C.byref(d), C.byref(h))
    return rc, h, lib.tlm_last_error(None).decode()


def test_argument_checks_before_device():
    assert _create(0, 4, 4)[0] == -1
    assert _create(4, 4, 4, dl=0.0)[0] == -1
    rc, _, msg = _create(4, 4, 4, pulse=7)
    assert rc == -1 and "storage type" in msg
    rc, _, msg = _create(1100, 1100, 1000)
    assert rc == -6 and "32-bit" in msg
    lib = _native.load()
    assert lib.tlm_create(None, None) == -1


@pytest.mark.skipif(tlm.device_count() > 0, reason="a HIP device is present")
def test_no_device_reported_not_crashing():
    assert tlm.device_count() == 0
    rc, h, msg = _create(8, 8, 8)
    assert rc == -2 and "no HIP device" in msg and not h.value
    with pytest.raises(tlm.TlmError, match="no HIP device"):
        tlm.TlmDevice(8, 8, 8, 0.01)
    assert tlm.scatter_unitarity_error(1.0) < 0
    lib = _native.load()
    g = C.c_float(0)
    assert lib.tlm_copy_bandwidth(0, 1 << 20, 1, C.byref(g), None) == -2


def test_host_face_lists_match_oracle():
    sc = scenes.selftest_dipole(scenes.WF_GAUSS_SINE, diel_sphere=True)
    o = helpers.oracle_open(sc)
    oc, oa, os_, _ = o.faces()
    c, a, s = tlm.pec_faces(sc.nx, sc.ny, sc.nz, sc.mat)
    assert oc.size > 0
    assert np.array_equal(c, oc) and np.array_equal(a, oa) and np.array_equal(s, os_)
    hc, ha, hs = o.huygens_faces()
    c, a, s = tlm.huygens_faces(sc.nx, sc.ny, sc.nz, sc.mat, sc.huy_offset)
    assert hc.size > 0
    assert np.array_equal(c, hc) and np.array_equal(a, ha) and np.array_equal(s, hs)


@pytest.mark.parametrize("wf", [0, 1, 2, 3])
def test_schedule_tables_match_oracle(wf):
    dt = tlm.timestep(scenes.DL)
    sv, cph, sph = tlm.schedule_tables(wf, 1e9, dt, 3000)
    ref = np.array([ot.waveform_value(wf, 1e9, dt, n) for n in range(3000)], np.float32)
    assert np.array_equal(sv, ref)
    # phase tables: float cos/sin of 2*pi*f0*(n*dt) as the monitor computes them
    ph = (np.float32(2.0) * np.float32(np.pi) * np.float32(1e9)) * (np.arange(3000, dtype=np.float32) * dt)
    assert np.allclose(cph, np.cos(ph.astype(np.float64)), atol=2e-6)
    assert np.allclose(sph, np.sin(ph.astype(np.float64)), atol=2e-6)


def test_dft_tables():
    f = np.array([5e7, 1e9, 1.6e9])
    dt = float(tlm.timestep(scenes.DL))
    ct, st = tlm.dft_tables(f, dt, 500)
    ph = 2 * np.pi * f[None, :] * (np.arange(500)[:, None] * dt)
    assert np.allclose(ct, np.cos(ph), atol=1e-15) and np.allclose(st, np.sin(ph), atol=1e-15)


def test_stub_admittances_float_rule():
    epsr = np.array([0.5, 1.0, 3.0, 9.7], np.float32)
    sigma = np.array([-0.1, 0.0, 0.02, 0.05], np.float32)
    yh, gh = tlm.stub_admittances(epsr, sigma, scenes.DL)
    assert yh[0] == 0 and yh[1] == 0 and yh[2] == np.float32(8.0)
    assert gh[0] == 0 and gh[1] == 0
    assert abs(gh[3] - 0.05 * scenes.DL * 376.730313) < 1e-6


def test_voxel_setup_rejects_null_context():
    lib = _native.load()
    g = _native.VoxelGrid()
    g.dl = 0.01
    z = np.zeros(3, np.float32)
    zp = z.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.tlm_mat_fill(None, 0) == -1
    assert lib.tlm_mat_cells_box(None, None, None, 1) == -1
    assert lib.tlm_voxelize_mesh(None, C.byref(g), zp, 1, None, 0, zp, 2, 1) == -1
    assert lib.tlm_voxelize_wire(None, C.byref(g), zp, 1, 0.0, zp, 1) == -1
    assert lib.tlm_carve_feed(None, C.byref(g), zp, zp, zp, None, 0, None) == -1
    assert lib.tlm_mat_read(None, None) == -1
    assert lib.tlm_upload_mat(None, None, None, 0) == -1
    assert lib.tlm_pec_faces_mat(None) == -1
    assert lib.tlm_huygens_mat(None, 4) == -1
    assert lib.tlm_read_faces(None, None, None, 0) == -1
    assert "null context" in lib.tlm_last_error(None).decode()
