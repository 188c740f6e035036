"""GPU parity: the HIP step (through the C ABI) against the CPU restatement
of TlmSolver's FP32 path, on identical inputs.

Bar: bit-exact for pulses, stubs, probes, ports, Js and Huygens
accumulators and the FP64 probe DFT (same operation order on both sides,
no FMA contraction); FP64 energy within 1e-12 relative (the sum order of
the reduction differs from the reference's chunked sum).
"""
import dataclasses

import numpy as np
import pytest

from emsolver_amd import scenes, tlm
from tests import helpers

pytestmark = pytest.mark.gpu


def _random_scene(nx, ny, nz, seed, rho=0.0, pec_frac=0.05, diel=True, steps=40, random_state=True):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    mat = np.zeros(n, np.uint8)
    nd = 4 if diel else 0
    u = rng.random(n)
    mat[u < pec_frac] = scenes.MAT_PEC
    if diel:
        ids = rng.integers(scenes.MAT_DIEL0, scenes.MAT_DIEL0 + nd, n).astype(np.uint8)
        sel = (u >= pec_frac) & (u < pec_frac + 0.4)
        mat[sel] = ids[sel]
    epsr = rng.uniform(1.0, 10.0, nd).astype(np.float32)
    sigma = rng.uniform(0.0, 0.05, nd).astype(np.float32)
    sc = scenes.Scene(nx, ny, nz, mat, epsr, sigma, rho=rho, total_steps=steps, settle_steps=5,
                      name=f"random_{nx}x{ny}x{nz}_s{seed}")
    air = np.flatnonzero(mat == 0)
    if air.size == 0:  # a grid of a few cells: put the sources in dielectric cells
        air = np.flatnonzero(mat != scenes.MAT_PEC)
    sc.sources = [(int(air[len(air) // 2]), 2, 1.0), (int(air[len(air) // 3]), 0, -0.7),
                  (int(air[len(air) // 2]), 1, 0.25)]
    sc.probes = np.array(rng.choice(n, min(5, n), replace=False), np.int64)
    sc.dft_freqs = np.array([1e8, 5e8, 1e9])
    if random_state:
        V0 = rng.standard_normal(12 * n).astype(np.float32)
        V0.reshape(n, 12)[mat == scenes.MAT_PEC] = 0.0
        sc.V0 = V0
        if diel:
            Vs0 = rng.standard_normal(3 * n).astype(np.float32)
            Vs0.reshape(n, 3)[mat < scenes.MAT_DIEL0] = 0.0
            sc.Vs0 = Vs0
    return sc


def _check(sc, steps=None):
    print(f"[parity] {sc.name} rho={sc.rho} steps={steps or sc.total_steps}", flush=True)
    g = tlm.run_scene(sc, steps=steps)
    o = helpers.oracle_run(sc, steps=steps)
    helpers.compare_results(g, o, sc)
    return g, o


def test_device_present():
    assert tlm.device_count() >= 1
    print(tlm.device_name(0))


@pytest.mark.parametrize("shape", [(8, 8, 8), (13, 7, 9), (33, 5, 6), (64, 3, 17)])
@pytest.mark.parametrize("rho", [0.0, -1.0, 0.37])
def test_random_materials_bitexact(shape, rho):
    sc = _random_scene(*shape, seed=sum(shape), rho=rho, steps=24)
    _check(sc)


# Grids one or two cells thick: every cell of a 1-wide axis touches both
# outer faces of it, rows shorter than a wave, strips holding many rows.
THIN = [(1, 1, 1), (1, 4, 5), (5, 1, 4), (4, 5, 1), (2, 2, 2), (65, 1, 1), (1, 1, 70),
        (1, 70, 1), (3, 1, 66)]


@pytest.mark.parametrize("shape", THIN, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("rho", [-1.0, 0.37])
def test_thin_grids_bitexact(shape, rho):
    n = shape[0] * shape[1] * shape[2]
    sc = _random_scene(*shape, seed=n + 11, rho=rho, pec_frac=0.0 if n < 8 else 0.1, steps=30)
    _check(sc)


@pytest.mark.parametrize("prop_axis", [0, 2])
def test_plane_wave_sources_bitexact(prop_axis):
    """A plane-wave source (TlmSolver::addPlaneWave): one soft source on every
    AIR cell of a plane, over dielectric boxes and the ground plane."""
    sc = scenes.dielectric_boxes(n=40, nboxes=4, steps=50, nprobes=6, nfreq=3, seed=5)
    sc.sources = []
    nsrc = scenes.add_plane_wave(sc, prop_axis, 12, 1 if prop_axis == 0 else 0, 0.8)
    assert nsrc > 500
    _check(sc)


def test_multi_cell_ports_bitexact():
    """Ports over several cells (TlmSolver::addPort): a 2 x 3 patch fed along
    z and a row of 4 fed along y that shares a cell with a soft source, so
    the port sums run over many cells, the loop current is taken round the
    middle cell, and one cell takes a source's and a port's additions in the
    reference's order (sources, then ports)."""
    sc = _random_scene(24, 20, 18, seed=31, rho=0.37, steps=60, random_state=False)
    m = sc.mat
    patch = [sc.cell(i, j, 9) for j in (8, 9) for i in (10, 11, 12)]
    row = [sc.cell(i, 14, 5) for i in range(4, 8)]
    m[patch + row] = scenes.MAT_AIR
    sc.sources = [(row[2], 1, 0.6), (sc.cell(3, 3, 3), 2, 1.0)]
    m[sc.cell(3, 3, 3)] = scenes.MAT_AIR
    sc.ports = [(patch, 2, 1.0), (row, 1, 0.5)]
    g, o = _check(sc)
    assert np.abs(o.ports).max() > 0


def test_no_dielectric_bitexact():
    sc = _random_scene(20, 12, 16, seed=5, diel=False, steps=30)
    _check(sc)


def test_free_space_cube_short():
    sc = scenes.free_space_cube(n=32, steps=120, probe_offset=6)
    _check(sc)


def test_dielectric_boxes_small():
    sc = scenes.dielectric_boxes(n=40, nboxes=5, steps=150, nprobes=6, nfreq=8, seed=2)
    _check(sc)


def test_selftest_dipole_ports_huygens():
    sc = scenes.selftest_dipole(scenes.WF_GAUSS_SINE, extra_steps=0)
    sc = sc.with_steps(min(sc.total_steps, sc.settle_steps + 80))
    _check(sc)


def test_selftest_dipole_with_sphere():
    sc = scenes.selftest_dipole(scenes.WF_CW_RAMPED, extra_steps=0, diel_sphere=True)
    sc = sc.with_steps(min(sc.total_steps, sc.settle_steps + 40))
    _check(sc)


def test_wide_rows_two_waves():
    # nx = 300: one row spans 75 quads; exercises multi-wave rows
    sc = _random_scene(300, 6, 5, seed=11, rho=-1.0, steps=12)
    _check(sc)


def test_free_space_64_full_run_tolerances():
    """The FP32 tolerances on the 64^3 free-space case: probe max rel err
    <= 1e-5 and L2-normalised snapshot err <= 1e-6 (bit-exact in practice)."""
    sc = scenes.free_space_cube(n=64, steps=1000)
    g = tlm.run_scene(sc)
    o = helpers.oracle_run(sc)
    pe = np.abs(g.probes - o.probes).max() / np.abs(o.probes).max()
    se = np.linalg.norm(g.V - o.V) / np.linalg.norm(o.V)
    print(f"probe rel err {pe:.3e}  snapshot rel L2 {se:.3e}")
    assert pe <= 1e-5 and se <= 1e-6
    helpers.compare_results(g, o, sc)


TUNINGS = [
    {"kernel": 0, "planes": 1, "block_waves": 4, "nontemporal": 0, "grid_blocks": 0, "fold": 0,
     "occupancy": 0},
    {"kernel": 0, "planes": 0, "block_waves": 0, "nontemporal": 1, "grid_blocks": 0, "fold": 1,
     "occupancy": 8},
    {"kernel": 0, "planes": 3, "block_waves": 1, "nontemporal": 0, "grid_blocks": 5, "occupancy": 3},
    {"kernel": 0, "planes": 1, "block_waves": 2, "nontemporal": 1, "grid_blocks": 17},
    {"kernel": 0, "planes": 64, "block_waves": 4, "nontemporal": 0, "grid_blocks": 8},
    {"kernel": 0, "planes": 2, "block_waves": 4, "nontemporal": 2, "grid_blocks": 0, "occupancy": 5},
]


def _run_tuned(sc, tuning, steps=None):
    steps = sc.total_steps if steps is None else steps
    with tlm.open_scene(sc) as dev:
        dev.set_tuning(**tuning)
        dev.run(0, steps)
        dev.sync()
        return tlm.collect(dev, sc, steps)


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: "z{planes}_w{block_waves}_nt{nontemporal}_g{grid_blocks}_f{}_o{}".format(t.get("fold", 2), t.get("occupancy", -1), **t))
@pytest.mark.parametrize("shape", [(12, 9, 11), (300, 5, 7), (63, 7, 5), (2000, 2, 3)])
def test_tunings_bitexact(shape, tuning):
    """Every launch shape gives the oracle's bits, including rows that start
    mid-strip (nx = 12, 63, 300) and rows split across many strips
    (nx = 300, 2000) where the seam cells push."""
    sc = _random_scene(*shape, seed=shape[0] + 7, rho=0.37, steps=10)
    print(f"[parity] {sc.name} tuning {tuning}", flush=True)
    g = _run_tuned(sc, tuning)
    o = helpers.oracle_run(sc)
    helpers.compare_results(g, o, sc)


def test_plane_and_frames_bitexact():
    """|E| cut plane (every even step) and playback frames (|Js| + latest
    plane every record_every steps) match the reference semantics."""
    sc = scenes.dielectric_boxes(n=30, nboxes=4, steps=47, nprobes=3, nfreq=2, seed=9)
    sc.plane_axis, sc.plane_index, sc.record_every = 1, 13, 5
    _check(sc)
    sc.plane_axis, sc.plane_index = 0, 99   # clamped to the last x plane
    _check(sc.with_steps(70))


@pytest.mark.parametrize("fold", [0, 1])
def test_folded_and_launched_injection_agree(fold):
    """Soft sources and a port added inside the step (fold 1) or by the
    injection launch (fold 0): the oracle's bits either way, with probes
    and a port on the source cells themselves, PEC neighbours and stubs."""
    sc = _random_scene(30, 11, 9, seed=44, rho=-1.0, steps=30)
    port = [sc.cell(15, 5, 4), sc.cell(15, 5, 5)]   # interior: the current loop stays in the grid
    sc.mat[port] = scenes.MAT_AIR
    sc.ports = [(port, 2, 0.9)]
    sc.probes = np.array([c for c, _, _ in sc.sources] + [port[0]], np.int64)
    print(f"[parity] fold {fold}", flush=True)
    g = _run_tuned(sc, {"fold": fold})
    helpers.compare_results(g, helpers.oracle_run(sc), sc)


def test_fold_tuning_rules():
    """tuning.fold: 0..2 accepted and read back (default 2), anything else
    refused; tuning.occupancy: -1 (automatic, the default), 0 or 3..8."""
    sc = _random_scene(20, 9, 7, seed=5, diel=False, steps=4)
    with tlm.open_scene(sc) as dev:
        assert dev.get_tuning()["fold"] == 2 and dev.get_tuning()["occupancy"] == -1
        for f in (0, 1, 2):
            dev.set_tuning(fold=f)
            assert dev.get_tuning()["fold"] == f
        with pytest.raises(tlm.TlmError, match="tlm_set_tuning"):
            dev.set_tuning(fold=3)
        for o in (-1, 0, 3, 8):
            dev.set_tuning(occupancy=o)
            assert dev.get_tuning()["occupancy"] == o
        for o in (-2, 1, 2, 9):
            with pytest.raises(tlm.TlmError, match="tlm_set_tuning"):
                dev.set_tuning(occupancy=o)


@pytest.mark.parametrize("every", [1, 3])
def test_timed_steps_equal_untimed(every):
    """tlm_time_steps (events on every `every`-th step launch) runs the same
    steps as tlm_run: the oracle's bits, and a kernel average that is > 0."""
    sc = _random_scene(40, 9, 11, seed=17, rho=0.37, steps=23)
    print(f"[parity] timed steps, events every {every}", flush=True)
    with tlm.open_scene(sc) as dev:
        tot, ker = dev.time_steps(0, 10, kernel_every=every)
        assert tot > 0 and ker is not None and 0 < ker <= tot
        _, none = dev.time_steps(10, 13, kernel_every=0)
        assert none is None
        with pytest.raises(tlm.TlmError, match="kernel_every"):
            dev.time_steps(23, 0, kernel_every=-1)
        dev.sync()
        g = tlm.collect(dev, sc, sc.total_steps)
    helpers.compare_results(g, helpers.oracle_run(sc), sc)
