"""Checks at the sizes the benchmark runs.

* The 256^3 dielectric-box scene (16.8 M cells, 16 probes with a
  32-frequency FP64 DFT, the ground plane's Js faces, energy) and the 512^3
  64-box scene (134 M cells; FP32, FP16 and BF16 storage) against the oracle
  bit for bit over their first steps, at the sizes where offsets, launch grids
  and the XCD block ranges take their full-scale values.
* The batched configuration's 256 grids and requests of the served workload
  at its full range of sizes: the same bits in the mix as alone.
* Over a full 3000-step run at 256^3, properties that hold at any size: a
  lossless closed box (rho = -1, sigma = 0, PEC ground plane) keeps its
  weighted energy sum(V^2) + sum(yh * Vs^2) (the quantity the stub node's
  scattering matrix is unitary in, TlmSolver.cpp:116-169, and that connect
  permutes with signs), and each probe's running FP64 DFT equals the DFT of
  its recorded series (the convention of SelfTest.cpp:182-195).
"""
import dataclasses

import numpy as np
import pytest

from emsolver_amd import scenes, serve, tlm
from tests import helpers
from tests.test_gpu_parity import _check
from tests.test_gpu_precision import _check16

pytestmark = pytest.mark.gpu


def _near_probes(sc, offsets):
    """The scene's probes plus probes at these offsets from the centre source
    cell, which the field reaches within the few steps run here."""
    c = sc.nx // 2
    near = [sc.cell(c + a, c + b, c + d) for a, b, d in offsets]
    return dataclasses.replace(sc, probes=np.concatenate([sc.probes, np.int64(near)]))


def test_256_scene_bitexact_first_steps():
    sc = scenes.dielectric_boxes(n=256, nboxes=8, steps=40, nprobes=16, nfreq=32, seed=1)
    _check(_near_probes(sc, [(5, 0, 0), (0, 7, 3), (-9, -4, 0), (2, 2, -11)]))


def test_256_scene_bitexact_first_steps_folded_sources():
    """The same with the source added inside the step (tuning fold = 1),
    which by default only grids of at most 4 M cells do."""
    from tests.test_gpu_parity import _run_tuned
    sc = _near_probes(scenes.dielectric_boxes(n=256, nboxes=8, steps=40, nprobes=16, nfreq=32,
                                              seed=1), [(5, 0, 0), (0, 7, 3), (0, 0, 0)])
    print(f"[full] {sc.name} with folded sources", flush=True)
    g = _run_tuned(sc, {"fold": 1})
    helpers.compare_results(g, helpers.oracle_run(sc), sc)


def test_512_scene_bitexact_first_steps_fp32():
    sc = scenes.dielectric_boxes(n=512, nboxes=64, steps=6, nprobes=16, nfreq=4, seed=1)
    _check(_near_probes(sc, [(2, 0, 0), (0, 3, 0), (-1, -2, 1)]))


@pytest.mark.parametrize("pulse", ["f16", "bf16"])
def test_512_scene_bitexact_first_steps_16bit(pulse):
    """The 512^3 64-box scene in both 16-bit storage formats (the two runs of
    the mixed-precision configuration) against the oracle's restatement of
    the 16-bit scheme, bit for bit."""
    sc = scenes.dielectric_boxes(n=512, nboxes=64, steps=6, nprobes=16, nfreq=4, seed=1)
    _check16(dataclasses.replace(_near_probes(sc, [(2, 0, 0), (0, 3, 0), (-1, -2, 1)]),
                                 pulse=pulse))


def _weighted_energy(sc, V, Vs):
    """sum(V^2) + sum over stub cells of yh * (Vs_x^2 + Vs_y^2 + Vs_z^2), in
    double, with the stub admittances the solver uses."""
    n = sc.ncell
    yh_id, _ = tlm.stub_admittances(sc.epsr, sc.sigma, sc.dl)
    yh = np.zeros(256, np.float64)
    yh[scenes.MAT_DIEL0:scenes.MAT_DIEL0 + len(yh_id)] = yh_id
    e = np.dot(V.astype(np.float64), V.astype(np.float64))
    s2 = (Vs.reshape(n, 3).astype(np.float64) ** 2).sum(axis=1)
    return e + np.dot(yh[sc.mat], s2)


def test_lossless_box_keeps_weighted_energy_3000_steps():
    sc = scenes.dielectric_boxes(n=256, nboxes=8, steps=3000, nprobes=4, nfreq=0, seed=2)
    sc = dataclasses.replace(sc, rho=-1.0, sigma=np.zeros_like(sc.sigma), sources=[],
                             name="lossless_boxes_256")
    rng = np.random.default_rng(8)
    n = sc.ncell
    V0 = rng.standard_normal(12 * n, dtype=np.float32)
    V0.reshape(n, 12)[sc.mat == scenes.MAT_PEC] = 0.0
    Vs0 = rng.standard_normal(3 * n, dtype=np.float32)
    Vs0.reshape(n, 3)[sc.mat < scenes.MAT_DIEL0] = 0.0
    sc.V0, sc.Vs0 = V0, Vs0
    print(f"[full] {sc.name}: {n} cells, {sc.total_steps} steps", flush=True)
    g = tlm.run_scene(sc)
    e0 = _weighted_energy(sc, V0, Vs0)
    e1 = _weighted_energy(sc, g.V, g.Vs)
    drift = e1 / e0 - 1.0
    print(f"[full] weighted energy {e0:.9e} -> {e1:.9e}, relative change {drift:.3e}", flush=True)
    assert np.isfinite(e1) and abs(drift) < 1e-6  # 3e-9 measured over the 3000 steps
    # the state moved: a run that did nothing would pass the energy check
    assert not np.array_equal(g.V, V0)


def test_probe_dft_equals_dft_of_series_3000_steps():
    sc = scenes.dielectric_boxes(n=256, nboxes=8, steps=3000, nprobes=16, nfreq=32, seed=1)
    print(f"[full] {sc.name}: probe DFT over {sc.total_steps} steps", flush=True)
    g = tlm.run_scene(sc, state=False)
    x = g.probes.astype(np.float64)                       # [step][probe][component]
    assert np.abs(x).max() > 0
    dt = float(tlm.timestep(sc.dl))
    ph = 2.0 * np.pi * np.outer(np.arange(sc.total_steps) * dt, sc.dft_freqs)  # [step][freq]
    re = np.einsum("npc,nf->pcf", x, np.cos(ph))
    im = -np.einsum("npc,nf->pcf", x, np.sin(ph))
    gre, gim = g.probe_dft
    scale = np.abs(x).sum(axis=0)[:, :, None]             # bound on any partial sum
    err = max(np.abs(gre - re).max() / scale.max(), np.abs(gim - im).max() / scale.max())
    print(f"[full] running DFT against the series' DFT: {err:.2e} of sum|x|", flush=True)
    assert err < 1e-12


def test_batched_256_grids_equal_alone():
    """The batched configuration's 256 dipole grids (each axis log-uniform in
    [32, 128], dielectric background) in one batch with their steps cut to
    U[60, 160]: every grid's probes, ports and final state equal the same grid
    run alone, and four of them equal the oracle."""
    scs = scenes.batch_dipoles(count=256, steps=(60, 160), seed=3)
    print(f"[full] {len(scs)} grids, {sum(sc.ncell for sc in scs)} cells in one batch", flush=True)
    res = tlm.run_scenes_batched(scs)
    for i, (g, sc) in enumerate(zip(res, scs)):
        alone = tlm.run_scene(sc)
        helpers.assert_bitexact(f"grid {i} V (batch vs alone)", g.V, alone.V)
        helpers.assert_bitexact(f"grid {i} probes (batch vs alone)", g.probes, alone.probes)
        helpers.assert_bitexact(f"grid {i} ports (batch vs alone)", g.ports, alone.ports)
    for i in (0, 85, 170, 255):
        o = helpers.oracle_run(scs[i])
        assert np.abs(o.ports).max() > 0
        helpers.assert_bitexact(f"grid {i} V (batch vs oracle)", res[i].V, o.V)
        helpers.assert_bitexact(f"grid {i} ports (batch vs oracle)", res[i].ports, o.ports)


def test_served_stream_full_sizes_equal_alone():
    """Twelve requests of the served workload (each axis log-uniform in
    [32, 384], FP16 for the 1e-3 requests and FP32 for the others, steps cut
    to U[40, 120]) all arriving at once: served together, each one's probes
    and ports equal the same request run alone."""
    reqs = serve.served_stream(count=12, rate=1e9, steps=(40, 120), seed=5, policy=False)
    for r in reqs:
        r.arrival = 0.0
    srv = serve.TlmServer(block_ticks=16)
    done = srv.serve(reqs, realtime=False)
    print(f"[full] served {len(done)} requests, largest {max(r.build().ncell for r in reqs)} "
          f"cells, mean members {srv.stats['mean_members']:.2f}", flush=True)
    assert sorted(c.rid for c in done) == list(range(len(reqs)))
    assert srv.stats["mean_members"] > 2
    for c in done:
        alone = tlm.run_scene(reqs[c.rid].build())
        assert np.abs(alone.ports).max() > 0
        helpers.assert_bitexact(f"req{c.rid} probes (mix vs alone)", c.probes, alone.probes)
        helpers.assert_bitexact(f"req{c.rid} ports (mix vs alone)", c.ports, alone.ports)
