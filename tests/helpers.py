"""Test helpers: run a Scene on the oracle (CPU restatement) and compare.

The oracle result mirrors emsolver_amd.tlm.SceneResult so GPU and CPU
results compare field by field.
"""
from __future__ import annotations

import numpy as np

from emsolver_amd.tlm import SceneResult
from oracle.oracle_tlm import OracleConfig, OracleTlm


PULSE_FORMATS = {"f32": 0, "f16": 1, "bf16": 2}


def oracle_open(scene, fp64=False, threads=0) -> OracleTlm:
    """The oracle for a scene: FP32 (parity), FP64 (precision reference) or,
    for a 16-bit scene, the restatement of its 16-bit storage scheme."""
    cfg = OracleConfig(scene.nx, scene.ny, scene.nz, dl=scene.dl, origin=scene.origin,
                       f0=scene.f0, waveform=scene.waveform, total_steps=scene.total_steps,
                       settle_steps=scene.settle_steps, rho=scene.rho,
                       huy_offset=scene.huy_offset, record_every=scene.record_every,
                       gd_width=scene.gd_width, gd_delay=scene.gd_delay, threads=threads,
                       fp64=fp64, pulse_format=0 if fp64 else PULSE_FORMATS[scene.pulse],
                       fp32_halo=scene.fp32_halo)
    o = OracleTlm(cfg, scene.mat, scene.epsr, scene.sigma)
    if scene.plane_axis >= 0:
        o.set_field_plane(scene.plane_axis, scene.plane_index)
    for c, p, a in scene.sources:
        o.add_source(c, p, a)
    for cells, p, a in scene.ports:
        o.add_port(cells, p, a)
    for c in scene.probes:
        o.add_probe_cell(int(c))
    if len(scene.probes) and len(scene.dft_freqs):
        o.set_probe_dft(scene.dft_freqs)
    if scene.V0 is not None:
        o.set_state(scene.V0, scene.Vs0)
    return o


def oracle_collect(o: OracleTlm, scene) -> SceneResult:
    steps = o.cur_step
    res = SceneResult(steps=steps)
    if len(scene.probes):
        res.probes = np.stack([o.probe_series(p) for p in range(len(scene.probes))], axis=1)
        if len(scene.dft_freqs):
            re, im = zip(*[o.probe_dft(p) for p in range(len(scene.probes))])
            res.probe_dft = (np.stack(re), np.stack(im))
    if scene.ports:
        recs = [o.port_records(p) for p in range(len(scene.ports))]
        res.ports = np.stack([np.stack([v, i], axis=1) for v, i in recs], axis=1)
    inst, re, im, mag, ns = o.js()
    res.js_inst = inst
    res.js_dft4 = np.stack([re[0::2], im[0::2], re[1::2], im[1::2]], axis=1) if inst.size else \
        np.zeros((0, 4))
    res.dft_samples = ns
    if scene.huy_offset >= 2:
        d, hs = o.huygens()
        res.huygens = d.reshape(-1, 8)
        res.huy_samples = hs
    st, _, ed = o.energy()
    res.energy_steps, res.energy = st, ed
    if scene.plane_axis >= 0:
        res.plane = o.plane()
    if scene.record_every > 0:
        fr = o.frames_full()
        res.frames = fr
    res.V, res.Vs = o.state()
    return res


def oracle_run(scene, steps=None, fp64=False, threads=0) -> SceneResult:
    o = oracle_open(scene, fp64=fp64, threads=threads)
    if steps is None:
        o.run_steps(scene.total_steps)
    else:
        o.run_steps(steps)
    o.finalize()
    res = oracle_collect(o, scene)
    o.close()
    return res


def assert_bitexact(name, gpu, ref):
    """FP32 GPU values against oracle values (held as float64 copies of
    float32 numbers): equal as float32, with a helpful message."""
    g = np.asarray(gpu, np.float32)
    r = np.asarray(ref).astype(np.float32)
    assert g.shape == r.shape, f"{name}: shape {g.shape} vs {r.shape}"
    if not np.array_equal(g, r):
        bad = np.flatnonzero(g.reshape(-1) != r.reshape(-1))
        i = int(bad[0])
        raise AssertionError(f"{name}: {bad.size} of {g.size} values differ; first at flat "
                             f"{i}: gpu {g.reshape(-1)[i]!r} oracle {r.reshape(-1)[i]!r}")


def compare_results(gpu: SceneResult, ref: SceneResult, scene, state=True, energy_rtol=1e-12):
    # the compared data must be non-trivial, or a bit-exact match proves nothing
    assert ref.steps > 0 and gpu.steps == ref.steps, "step counts differ"
    if state:
        assert np.abs(ref.V).max() > 0, "oracle state is all zero"
    if len(scene.probes):
        assert np.abs(ref.probes).max() > 0, "oracle probes are all zero"
    if len(ref.energy):
        assert ref.energy.max() > 0, "oracle energy is zero"
    if state:
        assert_bitexact("V", gpu.V, ref.V)
        if ref.Vs is not None:
            assert_bitexact("Vs", gpu.Vs, ref.Vs)
    if len(scene.probes):
        assert_bitexact("probes", gpu.probes, ref.probes)
        if len(scene.dft_freqs):
            assert np.array_equal(gpu.probe_dft[0], ref.probe_dft[0]), "probe DFT re"
            assert np.array_equal(gpu.probe_dft[1], ref.probe_dft[1]), "probe DFT im"
    if scene.ports:
        assert_bitexact("ports", gpu.ports, ref.ports)
    if gpu.js_inst is not None:
        assert_bitexact("js_inst", gpu.js_inst, ref.js_inst)
        assert_bitexact("js_dft", gpu.js_dft4, ref.js_dft4)
        assert gpu.dft_samples == ref.dft_samples
    if scene.huy_offset >= 2 and gpu.huygens is not None:
        assert_bitexact("huygens", gpu.huygens, ref.huygens)
    if scene.plane_axis >= 0:
        assert_bitexact("plane", gpu.plane, ref.plane)
    if scene.record_every > 0:
        gs, gj, gp = gpu.frames
        rs, rj, rp = ref.frames
        assert np.array_equal(gs, rs), "frame steps"
        assert_bitexact("frame js", gj, rj)
        if gp is not None:
            assert_bitexact("frame planes", gp, rp)
    assert np.array_equal(gpu.energy_steps, ref.energy_steps), "energy sample steps"
    if len(ref.energy):
        rel = np.abs(gpu.energy - ref.energy) / np.maximum(np.abs(ref.energy), 1e-300)
        assert rel.max() <= energy_rtol, f"energy rel err {rel.max()}"
        # the float history the reference keeps
        assert np.array_equal(gpu.energy.astype(np.float32), ref.energy.astype(np.float32)) or \
            rel.max() <= 1e-15


def oracle_build(scene):
    """The host-map scene equal to a device-built one: its set-up operations
    run by the oracle's voxelisers and feed carve (cell boxes and fills are
    plain array stores), the feeds as ports."""
    import dataclasses

    from emsolver_amd import scenes as _scenes
    from oracle import oracle_tlm as ot
    dims = (scene.nx, scene.ny, scene.nz)
    mat = np.zeros(scene.ncell, np.uint8)
    for op in scene.ops:
        if op[0] == "fill":
            mat[:] = op[1]
        elif op[0] == "cells_box":
            _scenes.stamp_box(mat, *dims, op[1], op[2], op[3])
        elif op[0] == "mesh":
            ot.voxelize_mesh(mat, dims, scene.origin, scene.dl, op[1], op[2], op[3], op[4], op[5])
        elif op[0] == "wire":
            ot.voxelize_wire(mat, dims, scene.origin, scene.dl, op[1], op[2], op[3], op[4])
    ports = []
    for op in scene.ops:
        if op[0] == "feed":
            cells, pol = ot.carve_feed(mat, dims, scene.origin, scene.dl, op[1], op[2], op[3])
            if cells.size:
                ports.append(([int(c) for c in cells], pol, float(op[4])))
    return dataclasses.replace(scene, mat=mat, ops=None, ports=ports)
