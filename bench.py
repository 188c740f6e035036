#!/usr/bin/env python3
"""Throughput of the SCN TLM time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step" = one TLM time step over the whole grid): the 512^3
grid with the dielectric layout of the 256^3 configuration scaled up to 64
seeded boxes (eps_r in [1,10], sigma in [0,0.05] S/m), a PEC ground plane at
z = 0, a centre Ez Gaussian-derivative soft source, 16 seeded E probes and a
running FP64 probe DFT at 32 frequencies, with the PEC surface-current and
energy monitors on, as the reference runs them; link pulses in FP32. Inputs
are resident in HBM before the timed region.

Prints ONE JSON line. `value` is Gcell-updates/s of the timed K steps
(wall clock between two stream synchronisations). `roofline` prices the
fused scatter/connect kernel: algorithmic bytes per launch
(cells * (97 + 24 * f_diel) B for FP32 pulses) over its average launch
time, measured with HIP events carried by the step launches of the timed
region on the context stream, against the 8 TB/s HBM peak. `cpu_baseline`
times the CPU restatement of the reference algorithm (oracle/, the
reference's threading rule min(16, hw-1)) over a 100-step window of the same
scene. `extra` holds the other configurations: the 256^3 scene, the 512^3
scene in FP16 / BF16 storage and in free space, the 16-bit storage error
against the FP64 oracle, the 64^3 free-space case, the batched and served
workloads, scene set-up and the far-field transform.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "TLM Gcell-updates/s & % HBM roofline; served mixed-size p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PULSE_BYTES = {"f32": 4, "f16": 2, "bf16": 2}


def alg_bytes_per_cell(scene, stub_bytes=4):
    """B_alg = 2*12*s_p + 1 + 2*3*s_s*f_diel (material id 1 B). For the 16-bit
    formats s_p is counted per slab: 4 B in the FP32 slabs near sources, 2 B
    elsewhere."""
    sp = PULSE_BYTES[scene.pulse]
    if scene.pulse != "f32":
        sp = 2 + 2 * fp32_slab_fraction(scene)
    return 2 * 12 * sp + 1 + 2 * 3 * stub_bytes * scene.diel_fraction()


def fp32_slab_fraction(scene):
    """Fraction of z-slabs a 16-bit run keeps in FP32 (within fp32_halo of a
    soft-source or port slab)."""
    if scene.pulse == "f32":
        return 1.0
    if scene.fp32_halo < 0:
        return 0.0
    slab = scene.nx * scene.ny
    hot = np.zeros(scene.nz, bool)
    cells = [c for c, _, _ in scene.sources] + [c for pc, _, _ in scene.ports for c in pc]
    for c in cells:
        k = int(c) // slab
        hot[max(0, k - scene.fp32_halo):k + scene.fp32_halo + 1] = True
    return float(hot.mean())


def read_traffic(workload):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary of this workload, if present (line 'traffic_bytes_per_launch:')."""
    path = os.path.join(ROOT, "profiles", f"pmc_k_step_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read())
    return float(m.group(1)) if m else None


# Every KERNEL_EVERY-th step launch of a timed region carries a start and an
# end HIP event for the kernel average. Events on a launch cost host time (~4
# us per launch at 256^3), so a sample perturbs the timed steps less than
# events on all of them; 7 is prime to the 16-step energy period, so the
# energy steps' launches enter the sample at their own rate.
KERNEL_EVERY = 7


def gpu_run(scene, warmup, steps, tag, plain=False, kernel_every=None):
    """Time `steps` steps after `warmup`, every kernel_every-th step launch
    (default KERNEL_EVERY) carrying HIP events for the kernel average.
    plain=True times `steps` more steps with no events at all."""
    kernel_every = KERNEL_EVERY if kernel_every is None else kernel_every
    from emsolver_amd import tlm
    sc = scene.with_steps(warmup + steps * (2 if plain else 1))
    t_setup = time.perf_counter()
    dev = tlm.open_scene(sc)
    log(f"[bench] {tag}: set-up {time.perf_counter() - t_setup:.1f}s, shape {dev.launch_shape()}")
    if warmup:
        dev.run(0, warmup)
    dev.sync()
    t0 = time.perf_counter()
    _, ms_kernel = dev.time_steps(warmup, steps, kernel_every=kernel_every)
    dev.sync()
    wall = time.perf_counter() - t0
    cells = sc.ncell
    out = {
        "wall_s": wall,
        "gcells": cells * steps / wall / 1e9,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms_avg": ms_kernel,
        "kernel_launches_timed": len(range(0, steps, kernel_every)),
        "kernel_every": kernel_every,
    }
    if plain:
        t1 = time.perf_counter()
        dev.time_steps(warmup + steps, steps, kernel_every=0)
        dev.sync()
        out["plain_gcells"] = cells * steps / (time.perf_counter() - t1) / 1e9
    b = alg_bytes_per_cell(sc)
    out["alg_bytes_per_cell"] = b
    out["achieved_gbs"] = cells * b / (out["kernel_ms_avg"] * 1e-3) / 1e9
    # a touch of the results, so a broken run cannot pass silently
    if dev.nprobe:
        e = dev.read_probes(0, warmup + steps)
        out["probe_absmax"] = float(np.abs(e).max())
    dev.close()
    return out


GRAPH_TICKS = 32  # ticks per captured HIP graph in the batched and served lines


def profile_batched(count=256, ticks=48):
    """For counter profilers: the batched line's 256 grids in one batch,
    autotuned, then `ticks` plain (not graph-captured) ticks; a progress line
    per tick block. Reports no rate."""
    from emsolver_amd import scenes, tlm
    scs = scenes.batch_dipoles(count=count, seed=3)
    devs = [tlm.open_scene(sc) for sc in scs]
    with tlm.TlmBatch() as b:
        b.set_tuning(graph_ticks=0)
        for d, sc in zip(devs, scs):
            b.add(d, sc.total_steps)
        tuning = b.autotune(4)
        b.sync()
        done = 0
        while done < ticks:
            m = min(8, ticks - done)
            b.run(m)
            b.sync()
            done += m
            log(f"[bench] profile batched_{count}: {done} of {ticks} ticks")
        for d in devs:
            b.remove(d)
    for d in devs:
        d.close()
    log(f"[bench] profile batched_{count}: {ticks} ticks, tuning {tuning}, "
        f"{sum(sc.ncell for sc in scs)} cells per tick")


def batched_run(count=256, sample_every=4):
    """256 independent dipole grids (axes log-uniform in [32,128], steps
    U[500,2000]) in one batch: autotuned launch shape, graph-captured blocks
    of ticks, run to completion. Baseline: the same grids one at a time as
    plain contexts (every `sample_every`-th grid, timed alone)."""
    from emsolver_amd import scenes, tlm
    scs = scenes.batch_dipoles(count=count, seed=3)
    work = sum(sc.ncell * sc.total_steps for sc in scs)
    t = time.perf_counter()
    devs = [tlm.open_scene(sc) for sc in scs]
    log(f"[bench] batched_{count}: {sum(sc.ncell for sc in scs) / 1e6:.1f} M cells, "
        f"{work / 1e9:.1f} G cell-updates, set-up {time.perf_counter() - t:.1f}s")
    with tlm.TlmBatch() as b:
        b.set_tuning(graph_ticks=GRAPH_TICKS)
        for d, sc in zip(devs, scs):
            b.add(d, sc.total_steps)
        t0 = time.perf_counter()
        tuning = b.autotune(4)
        b.sync()
        left = max(b.remaining().values())
        while left > 0:
            m = min(256, left)
            b.run(m)
            left -= m
        b.sync()
        wall = time.perf_counter() - t0
        probe_absmax = 0.0
        for d, sc in zip(devs, scs):
            b.remove(d)
        for d, sc in zip(devs[:8], scs[:8]):
            probe_absmax = max(probe_absmax, float(np.abs(d.read_ports(0, sc.total_steps)).max()))
    for d in devs:
        d.close()
    # one at a time
    alone_work, alone_wall = 0, 0.0
    for sc in scs[::sample_every]:
        d = tlm.open_scene(sc)
        d.sync()
        t1 = time.perf_counter()
        n = 0
        while n < sc.total_steps:
            m = min(256, sc.total_steps - n)
            d.run(n, m)
            n += m
        d.sync()
        alone_wall += time.perf_counter() - t1
        alone_work += sc.ncell * sc.total_steps
        d.close()
    out = {
        "grids": count,
        "gcell_s": round(work / wall / 1e9, 3),
        "wall_s": round(wall, 3),
        "cell_updates": work,
        "tuning": tuning,
        "one_at_a_time_gcell_s": round(alone_work / alone_wall / 1e9, 3),
        "one_at_a_time_sample": f"every {sample_every}th grid ({len(scs[::sample_every])} grids)",
        "port_absmax": probe_absmax,
    }
    out["speedup_vs_one_at_a_time"] = round(out["gcell_s"] / out["one_at_a_time_gcell_s"], 3)
    # the mix's algorithmic bytes per cell-update; the rate above covers the
    # whole run, the tail where few members are left included
    bpc = sum(sc.ncell * scene_alg_bytes(sc) for sc in scs) / sum(sc.ncell for sc in scs)
    out["alg_bytes_per_cell"] = round(bpc, 3)
    out["roofline_frac"] = round(out["gcell_s"] * bpc / HBM_PEAK_GBS, 4)
    path = os.path.join(ROOT, "profiles", f"k_step_batch_{count}.md")
    if os.path.exists(path):
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_tick:\s*([0-9.eE+]+)", f.read())
        if m:
            # HBM bytes of one tick with every member active (rocprofv3 PMC)
            out["traffic_bytes_per_tick"] = float(m.group(1))
            out["alg_bytes_per_tick"] = round(bpc * sum(sc.ncell for sc in scs))
    log(f"[bench] batched_{count}: {out['gcell_s']:.2f} Gcell/s aggregate in {wall:.1f}s "
        f"(tuning {tuning}); one at a time {out['one_at_a_time_gcell_s']:.2f} Gcell/s")
    return out


def scene_alg_bytes(sc):
    """B_alg per cell-update of a scene (host map or device set-up
    operations): 2*12*s_p + 1 + 24*f_diel, s_p counted per slab for the
    16-bit formats (4 B in the FP32 band round the feed or sources)."""
    from emsolver_amd import scenes
    if sc.mat is not None:
        f_diel = sc.diel_fraction()
    else:
        m = np.zeros((sc.nz, sc.ny, sc.nx), np.uint8)
        for op in sc.ops:
            if op[0] == "cells_box":
                (i0, j0, k0), (i1, j1, k1), mid = op[1], op[2], op[3]
                m[k0:k1, j0:j1, i0:i1] = mid
        f_diel = float((m >= scenes.MAT_DIEL0).mean())
    sp = PULSE_BYTES[sc.pulse]
    if sc.pulse != "f32":
        if sc.dipole is not None and not sc.sources and sc.fp32_halo >= 0:
            _, _, g0, gap = sc.dipole
            lo, hi = max(0, g0 - sc.fp32_halo), min(sc.nz - 1, g0 + gap - 1 + sc.fp32_halo)
            sp = 2 + 2 * (hi - lo + 1) / sc.nz
        else:
            sp = 2 + 2 * fp32_slab_fraction(sc)
    return 2 * 12 * sp + 1 + 2 * 3 * 4 * f_diel


def served_run(count=200, rate=50.0, baseline_sample=20, seed=5, max_cells=int(2.5e8),
               device_setup=False, copy_gbs=None):
    """The served mixed-size stream (Poisson arrivals at `rate`/s, axes
    log-uniform in [32,384], steps U[500,4000]; requests that tolerate 1e-3
    in FP16 where the storage policy's error model allows it, all others in
    FP32) through the resident serving loop.
    device_setup: each request's scene is voxelised on the device at
    admission instead of built and uploaded from the host. Baseline:
    `baseline_sample` of the same requests solved one at a time as plain FP32
    contexts."""
    from emsolver_amd import serve, tlm
    reqs = serve.served_stream(count=count, rate=rate, seed=seed, device_setup=device_setup)
    srv = serve.TlmServer(block_ticks=32, autotune_ticks=2, keep_results=False, progress_s=20.0,
                          max_cells=max_cells, tuning={"graph_ticks": GRAPH_TICKS})
    done = srv.serve(reqs)
    summ = serve.summarize(done)
    summ.update({"rate_per_s": rate, "mean_members": round(srv.stats["mean_members"], 2),
                 "scene_setup": "device voxelisers" if device_setup else "host map + upload",
                 "tuning": srv.stats["tuning"], "admission": f"FIFO, free HBM, <= {max_cells} cells in flight"})
    alone_work, alone_wall = 0, 0.0
    for r in reqs[:baseline_sample]:
        sc = r.build()
        sc.pulse = "f32"
        d = tlm.open_scene(sc)
        d.sync()
        t1 = time.perf_counter()
        n = 0
        while n < sc.total_steps:
            m = min(256, sc.total_steps - n)
            d.run(n, m)
            n += m
        d.sync()
        alone_wall += time.perf_counter() - t1
        alone_work += sc.ncell * sc.total_steps
        d.close()
    summ["one_at_a_time_fp32_gcell_s"] = round(alone_work / alone_wall / 1e9, 3)
    summ["one_at_a_time_sample"] = f"first {baseline_sample} requests, FP32, alone"
    summ["speedup_vs_one_at_a_time"] = round(summ["gcell_s"] / summ["one_at_a_time_fp32_gcell_s"], 3)
    # the HBM roofline of this mix: every request's cell-updates at its own
    # algorithmic bytes (FP16 requests: 16-bit slabs outside the FP32 band)
    # moved at 8 TB/s; no schedule of the same requests can beat it
    work = byts = f16_work = 0.0
    for r in reqs:
        sc = r.build()
        w = sc.ncell * sc.total_steps
        work += w
        byts += w * scene_alg_bytes(sc)
        f16_work += w if sc.pulse != "f32" else 0.0
    ceil = work / (byts / (HBM_PEAK_GBS * 1e9)) / 1e9
    summ["hbm_ceiling_gcell_s"] = round(ceil, 3)
    summ["mix_alg_bytes_per_cell"] = round(byts / work, 3)
    summ["ceiling_vs_one_at_a_time"] = round(ceil / summ["one_at_a_time_fp32_gcell_s"], 3)
    summ["frac_of_ceiling"] = round(summ["gcell_s"] / ceil, 4)
    if copy_gbs:
        # the same at the bandwidth a plain device copy reaches on this box
        cc = work / (byts / (copy_gbs * 1e9)) / 1e9
        summ["copy_ceiling_gcell_s"] = round(cc, 3)
        summ["frac_of_copy_ceiling"] = round(summ["gcell_s"] / cc, 4)
    summ["target_vs_one_at_a_time"] = 5.0
    summ["f16_share"] = round(summ["f16_requests"] / summ["requests"], 4)
    summ["f16_work_share"] = round(f16_work / work, 4)
    summ["storage_policy"] = (f"tolerance 1e-3 -> FP16 only where serve.plan_storage's error model "
                              f"allows it (FP32 band holds the dipole and probes with "
                              f"{serve.F16_MARGIN} slabs to spare, axes <= {serve.F16_MAX_AXIS}, "
                              f">= {serve.F16_MIN_SHARE} of the slabs 16-bit), else FP32")
    for k in ("p50_latency_s", "p99_latency_s", "mean_latency_s", "gcell_s", "makespan_s"):
        summ[k] = round(summ[k], 4)
    log(f"[bench] served_{count}: p50 {summ['p50_latency_s']:.2f}s p99 {summ['p99_latency_s']:.2f}s, "
        f"{summ['gcell_s']:.2f} Gcell/s aggregate (mean {summ['mean_members']} members); "
        f"one at a time {summ['one_at_a_time_fp32_gcell_s']:.2f} Gcell/s; HBM ceiling of the mix "
        f"{ceil:.1f} Gcell/s = {summ['ceiling_vs_one_at_a_time']:.2f}x one at a time (target 5x), "
        f"reached {summ['frac_of_ceiling']:.3f}; FP16 {summ['f16_share']:.3f} of requests")
    return summ


def crop_scene(sc, lo, n):
    """The n^3 subdomain [lo, lo+n)^3 of a cubic scene: materials and sources
    copied, probes re-seeded inside (matched outer faces on the crop)."""
    from emsolver_amd import scenes
    m = sc.mat.reshape(sc.nz, sc.ny, sc.nx)[lo:lo + n, lo:lo + n, lo:lo + n].copy().reshape(-1)
    out = scenes.Scene(n, n, n, m, sc.epsr, sc.sigma, total_steps=sc.total_steps,
                       name=f"{sc.name}_crop{n}")
    for c, p, a in sc.sources:
        i, j, k = sc.coords(c)
        if all(lo <= x < lo + n for x in (i, j, k)):
            out.sources.append((out.cell(i - lo, j - lo, k - lo), p, a))
    rng = np.random.default_rng(11)
    probes = []
    while len(probes) < len(sc.probes):
        c = out.cell(*(int(x) for x in rng.integers(1, n - 1, 3)))
        if m[c] != scenes.MAT_PEC and c not in probes:
            probes.append(c)
    out.probes = np.array(probes, np.int64)
    return out


def precision_report(steps=4000, crop=128):
    """Error of the 16-bit pulse storage against the FP64 oracle: the 512^3
    64-box scene cropped to its central crop^3 subdomain (source inside), all
    16 probes, `steps` steps; and a dipole over the dielectric background for
    the port (S-parameter) records. Max relative error per probe / port:
    max_t |x - x64| / max_t |x64|, and the relative error of the input
    impedance at f0 = 1 GHz."""
    import dataclasses
    from emsolver_amd import scenes, tlm
    from tests import helpers
    big = scenes.dielectric_boxes(n=512, nboxes=64, steps=steps, nprobes=16, nfreq=0, seed=1)
    sc = crop_scene(big, 256 - crop // 2, crop)
    t = time.perf_counter()
    r64 = helpers.oracle_run(sc, fp64=True)
    log(f"[bench] precision: FP64 oracle {crop}^3 x {steps} in {time.perf_counter() - t:.1f}s")
    rep = {"scene": f"diel_boxes_512 (64 boxes) cropped to the central {crop}^3, {steps} steps, "
                    f"16 probes (E, 3 components)", "probes": {}}
    for pulse in ("f32", "f16", "bf16"):
        g = tlm.run_scene(dataclasses.replace(sc, pulse=pulse), state=False)
        e = np.abs(g.probes - r64.probes).max(axis=(0, 2)) / np.abs(r64.probes).max(axis=(0, 2))
        rep["probes"][pulse] = {"max": float(e.max()), "median": float(np.median(e)),
                                "per_probe": [float(x) for x in e]}
        log(f"[bench] precision {pulse}: probe max rel err max {e.max():.2e} median {np.median(e):.2e}")
    dip = scenes.dipole_scene(96, 96, 96, length_frac=0.5, steps=2000, background_boxes=8, seed=7)
    dip.faces = False
    dip.probes = np.array([dip.cell(20, 30, 40), dip.cell(70, 60, 50), dip.cell(48, 20, 48)], np.int64)
    d64 = helpers.oracle_run(dip, fp64=True)
    dt = float(tlm.timestep(dip.dl))

    def zin(ports):
        n = np.arange(ports.shape[0])
        w = np.exp(-2j * np.pi * 1e9 * n * dt)
        return (ports[:, 0, 0] * w).sum() / (ports[:, 0, 1] * w).sum()

    z64 = zin(d64.ports)
    rep["dipole"] = {"scene": "96^3 centre-fed dipole over 8 dielectric boxes, 2000 steps",
                     "z64_ohm": [float(z64.real), float(z64.imag)]}
    for pulse in ("f32", "f16", "bf16"):
        g = tlm.run_scene(dataclasses.replace(dip, pulse=pulse), state=False)
        ve = np.abs(g.ports - d64.ports).max(axis=0)[0] / np.abs(d64.ports).max(axis=0)[0]
        pe = np.abs(g.probes - d64.probes).max(axis=(0, 2)) / np.abs(d64.probes).max(axis=(0, 2))
        z = zin(g.ports)
        rep["dipole"][pulse] = {"port_v": float(ve[0]), "port_i": float(ve[1]),
                                "probe_max": float(pe.max()),
                                "zin_rel": float(abs(z - z64) / abs(z64))}
        log(f"[bench] precision dipole {pulse}: port V {ve[0]:.2e} I {ve[1]:.2e} "
            f"Zin {abs(z - z64) / abs(z64):.2e} probes {pe.max():.2e}")
    return rep


def setup_run(n=256, subdiv=5):
    """Scene -> voxel set-up of mesh_scene(n) (dielectric sphere of
    20*4^subdiv triangles, as an imported CAD mesh would be, box and rotated
    box, PEC plate, thick bent wire, fed thin dipole): on the device
    (voxelisers, cell codes, PEC faces, Huygens faces) against the host path
    a reference user takes (the voxelisers on one CPU thread, the oracle's
    restatement of Geometry.cpp, then the host upload and host face
    builds)."""
    from emsolver_amd import scenes, tlm
    from tests import helpers
    sc = scenes.mesh_scene(n, steps=4, subdiv=subdiv)
    ntri = sum(len(op[2]) for op in sc.ops if op[0] == "mesh")
    yh, gh = tlm.stub_admittances(sc.epsr, sc.sigma, sc.dl)
    best = None
    for _ in range(3):
        with tlm.TlmDevice(sc.nx, sc.ny, sc.nz, sc.dl) as dev:
            dev.sync()
            t0 = time.perf_counter()
            tlm.apply_ops(dev, sc)
            dev.upload_mat(yh, gh)
            nf = dev.pec_faces_mat()
            nh = dev.huygens_mat(sc.huy_offset)
            dev.sync()
            dt = time.perf_counter() - t0
            dmat = dev.mat_read()
        best = dt if best is None else min(best, dt)
    t0 = time.perf_counter()
    host = helpers.oracle_build(sc)           # the voxelisers on one CPU thread
    t_vox = time.perf_counter() - t0
    assert np.array_equal(host.mat, dmat), "device map differs from the CPU voxelisers"
    with tlm.TlmDevice(sc.nx, sc.ny, sc.nz, sc.dl) as dev:
        dev.sync()
        t0 = time.perf_counter()
        dev.upload(host.mat, yh, gh)
        fc, fa, _ = tlm.pec_faces(sc.nx, sc.ny, sc.nz, host.mat)
        dev.set_faces(fc, fa)
        hc, ha, _ = tlm.huygens_faces(sc.nx, sc.ny, sc.nz, host.mat, sc.huy_offset)
        dev.set_huygens(hc, ha)
        dev.sync()
        t_up = time.perf_counter() - t0
    assert fc.size == nf and hc.size == nh
    out = {"grid": [n, n, n], "triangles": ntri, "faces": int(nf), "huygens_faces": int(nh),
           "device_setup_ms": round(best * 1e3, 2),
           "cpu_voxelize_ms": round(t_vox * 1e3, 1),
           "host_upload_and_faces_ms": round(t_up * 1e3, 1),
           "speedup": round((t_vox + t_up) / best, 1),
           "cpu_sample": "oracle voxelisers, 1 thread (TMainForm::voxelizeScene is serial)"}
    log(f"[bench] scene_setup_{n}: device {out['device_setup_ms']} ms vs CPU voxelise "
        f"{out['cpu_voxelize_ms']} ms + host upload/faces {out['host_upload_and_faces_ms']} ms")
    return out


def ntff_run(n=256, offset=4, n_theta=37, n_phi=73, cpu_every=32):
    """Near-to-far-field transform of the Huygens box of an n^3 grid (box
    inset `offset` cells: 6 (n - 2 offset)^2 faces, seeded DFT values) on the
    default 37 x 73 grid: device wall time (host face arrays in, U and the
    directivity out) against the oracle's restatement of
    ComputeFarFieldFromHuygens on one CPU thread over every cpu_every-th
    face, scaled to all faces (the sums are linear in the face count)."""
    from emsolver_amd import tlm
    from oracle import oracle_tlm as ot
    mat = np.zeros(n ** 3, np.uint8)
    cells, axes, _ = tlm.huygens_faces(n, n, n, mat, offset)
    dl = float(np.float32(0.0149896))
    pos, sg = ot.huygens_geometry((n, n, n), (0.0, 0.0, 0.0), dl, cells, axes)
    rng = np.random.default_rng(2)
    dft = rng.normal(size=(cells.size, 8)).astype(np.float32)
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        r = tlm.far_field(pos, axes, sg, dft, 1e9, n_theta, n_phi)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    sub = slice(0, None, cpu_every)
    t0 = time.perf_counter()
    ot.far_field(pos[sub], axes[sub], sg[sub], dft[sub], 1e9, n_theta, n_phi)
    t_cpu = (time.perf_counter() - t0) * cpu_every
    pairs = cells.size * n_theta * n_phi
    out = {"faces": int(cells.size), "directions": n_theta * n_phi,
           "device_ms": round(best * 1e3, 2),
           "face_directions_per_s": round(pairs / best, 1),
           "cpu_1thread_ms_est": round(t_cpu * 1e3, 1),
           "cpu_sample": f"oracle float sums over every {cpu_every}th face, scaled",
           "speedup_vs_1thread": round(t_cpu / best, 1),
           "directivity": round(r.directivity, 5)}
    log(f"[bench] ntff_{n}: {cells.size} faces x {n_theta * n_phi} directions: device "
        f"{out['device_ms']} ms, CPU 1 thread ~{out['cpu_1thread_ms_est']:.0f} ms")
    return out


def free_space_64_run(with_cpu=True):
    """The reference's CPU-runnable case: 64^3 free space, matched
    boundaries, centre Ez source, Ez probe 10 cells off-centre, 1000 steps.
    Its ~25 MB of state stays in the Infinity Cache, so its bandwidth is
    reported but it is no HBM roofline case. With with_cpu the oracle runs
    the same 1000 steps (the CPU baseline of this case) and the GPU probe
    series is compared with it bit for bit."""
    from emsolver_amd import scenes, tlm
    sc = scenes.free_space_cube(n=64, steps=1000, probe_offset=10)
    dev = tlm.open_scene(sc)
    dev.sync()
    t0 = time.perf_counter()
    dev.time_steps(0, sc.total_steps, kernel_every=0)
    dev.sync()
    wall = time.perf_counter() - t0
    g = tlm.collect(dev, sc, sc.total_steps, state=False)
    dev.close()
    # kernel average from a second run of the same scene with events on every
    # launch (they cost ~20 us of host time per launch at this size)
    with tlm.open_scene(sc) as dev:
        _, ker = dev.time_steps(0, sc.total_steps, kernel_every=1)
    b = alg_bytes_per_cell(sc)
    out = {
        "grid": [64, 64, 64],
        "steps": sc.total_steps,
        "gcell_s": round(sc.ncell * sc.total_steps / wall / 1e9, 3),
        "kernel_us_avg": round(ker * 1e3, 2),
        "achieved_gbs": round(sc.ncell * b / (ker * 1e-3) / 1e9, 1),
        "note": "state (25 MB) resident in the Infinity Cache: not an HBM roofline case; "
                "gcell_s from 1000 steps with no launch events, kernel_us_avg from a second "
                "run with events on every launch",
    }
    log(f"[bench] free_space_64: {out['gcell_s']:.2f} Gcell/s wall, kernel {out['kernel_us_avg']:.1f} us")
    if with_cpu:
        from tests import helpers
        o = helpers.oracle_open(sc)
        t0 = time.perf_counter()
        o.run_steps(sc.total_steps)
        dt = time.perf_counter() - t0
        o.finalize()
        r = helpers.oracle_collect(o, sc)
        o.close()
        from oracle.oracle_tlm import load_library
        out["cpu_gcell_s"] = round(sc.ncell * sc.total_steps / dt / 1e9, 4)
        out["cpu_threads"] = load_library().ot_default_threads()
        out["cpu_sample"] = "all 1000 steps, CPU restatement of TlmSolver's step"
        out["probe_series_bitexact"] = bool(np.array_equal(g.probes, r.probes))
        log(f"[bench] free_space_64 CPU {out['cpu_gcell_s']:.3f} Gcell/s, probe series "
            f"bit-exact: {out['probe_series_bitexact']}")
    return out


def cpu_baseline(scene, budget_s=45.0, window=100):
    """The oracle on the same scene: a fixed window of `window` steps after
    two warm-up steps (fewer if the window would exceed budget_s)."""
    from tests import helpers
    o = helpers.oracle_open(scene.with_steps(10**6))
    t0 = time.perf_counter()
    o.run_steps(2)
    per = (time.perf_counter() - t0) / 2
    n = int(max(2, min(window, budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    o.run_steps(n)
    dt = time.perf_counter() - t0
    from oracle.oracle_tlm import load_library
    threads = load_library().ot_default_threads()
    o.close()
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": scene.ncell * n / dt / 1e9,
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{scene.name}: steps 2..{n + 1} ({n} steps, {dt:.1f}s) of the same scene with "
                  f"all monitors, CPU restatement of TlmSolver's step with its ThreadPool rule "
                  f"min(16, hw-1) = {threads} threads; host {os.cpu_count()} logical CPUs "
                  f"({model})",
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512,
                    help="grid edge of the dielectric-box scene (64 boxes at 512, else 8)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the extra single-grid lines (256^3, 16-bit, free space, precision, "
                         "set-up, far field)")
    ap.add_argument("--cpu-budget", type=float, default=45.0,
                    help="cap on the CPU baseline's 100-step window, seconds")
    ap.add_argument("--precision-steps", type=int, default=4000,
                    help="steps of the 16-bit error study (the workload's 4000)")
    ap.add_argument("--extra-steps", type=int, default=200,
                    help="timed steps of each extra 512^3 line")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and served lines")
    ap.add_argument("--lines", default="",
                    help="comma list of extra lines to run (diel_boxes_256, diel_boxes_512_f16, "
                         "diel_boxes_512_bf16, free_space_512_f32, precision, free_space_64, "
                         "scene_setup_256, scene_setup_512, ntff_256, batched_256, served); "
                         "default all")
    ap.add_argument("--no-graph", action="store_true",
                    help="plain launches instead of captured HIP graphs in the batched lines "
                         "(for tracing tools that do not follow graph launches)")
    ap.add_argument("--precision-report", action="store_true",
                    help="only report the 16-bit storage error against the FP64 oracle (JSON)")
    ap.add_argument("--serve-only", action="store_true",
                    help="only run the served stream (--serve-count requests) and print its JSON")
    ap.add_argument("--serve-host-setup", action="store_true",
                    help="served requests as host material maps uploaded at admission (default: "
                         "set-up operations voxelised on the device)")
    ap.add_argument("--profile-batched", type=int, default=0, metavar="TICKS",
                    help="only run TICKS plain ticks of the batched line's 256 grids, for "
                         "counter profilers (no JSON line)")
    ap.add_argument("--serve-count", type=int, default=2000,
                    help="requests of the served line (the workload's full stream: 2000)")
    ap.add_argument("--serve-rate", type=float, default=50.0,
                    help="Poisson arrival rate of the served line, requests/s (the workload's is 50)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding of the grid is a "
              "design note (z-slab halo exchange), not an implemented path. Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from emsolver_amd import scenes, tlm
    if tlm.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    global GRAPH_TICKS
    if args.no_graph:
        GRAPH_TICKS = 0
    if args.precision_report:
        print(json.dumps(precision_report(steps=args.precision_steps)), flush=True)
        return
    if args.profile_batched > 0:
        profile_batched(ticks=args.profile_batched)
        return
    if args.serve_only:
        print(json.dumps(served_run(args.serve_count, rate=args.serve_rate,
                                    device_setup=not args.serve_host_setup)),
              flush=True)
        return
    dev_name = tlm.device_name(0)
    log(f"[bench] device {dev_name}")

    n = args.size
    nboxes = 64 if n >= 512 else 8
    workload = f"diel_boxes_{n}"
    scene = scenes.dielectric_boxes(n=n, nboxes=nboxes, steps=args.warmup + args.steps, nprobes=16,
                                    nfreq=32, seed=1)
    log(f"[bench] {workload}: {scene.ncell} cells, {nboxes} boxes, f_diel "
        f"{scene.diel_fraction():.4f}, warmup {args.warmup}, steps {args.steps}")
    # at 512^3 a launch takes ~2.5 ms: events on every launch cost nothing there
    main_run = gpu_run(scene, args.warmup, args.steps, workload, plain=True,
                       kernel_every=1 if scene.ncell >= 512 ** 3 else KERNEL_EVERY)
    log(f"[bench] {workload}: {main_run['gcells']:.2f} Gcell/s wall, kernel "
        f"{main_run['kernel_ms_avg']:.3f} ms, {main_run['achieved_gbs']:.0f} GB/s algorithmic; "
        f"{main_run['plain_gcells']:.2f} Gcell/s without launch events")

    extra = {"uninstrumented": {
        "gcell_s": round(main_run["plain_gcells"], 3),
        "steps": args.steps,
        "note": f"the next steps of the same run with no HIP events on the launches; in the "
                f"timed steps every {main_run['kernel_every']}th step launch carries a start "
                f"and an end event for the kernel average"}}
    lib = tlm._native.load()
    import ctypes as C
    gbps, msc = C.c_float(0), C.c_float(0)
    if lib.tlm_copy_bandwidth(0, 2 << 30, 20, C.byref(gbps), C.byref(msc)) == 0:
        extra["device_copy_gbs"] = round(gbps.value, 1)
        log(f"[bench] device copy {gbps.value:.0f} GB/s")
    lines = set(args.lines.split(",")) if args.lines else None

    def want(name):
        return lines is None or name in lines

    def roofline_line(r, sc):
        return {
            "gcell_s": round(r["gcells"], 3),
            "kernel_ms_avg": round(r["kernel_ms_avg"], 4),
            "kernel_launches_timed": r["kernel_launches_timed"],
            "achieved_gbs": round(r["achieved_gbs"], 1),
            "roofline_frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "alg_bytes_per_cell": round(r["alg_bytes_per_cell"], 3),
            "diel_fraction": round(sc.diel_fraction(), 5),
        }

    if not args.no_extra and want("diel_boxes_256") and n != 256:
        sc = scenes.dielectric_boxes(n=256, nboxes=8, steps=args.extra_steps + 10, nprobes=16,
                                     nfreq=32, seed=1)
        r = gpu_run(sc, 10, args.extra_steps, "diel_boxes_256")
        extra["diel_boxes_256"] = roofline_line(r, sc)
        extra["diel_boxes_256"]["steps"] = args.extra_steps
        extra["diel_boxes_256"]["traffic"] = read_traffic("diel_boxes_256")
        log(f"[bench] diel_boxes_256: {r['gcells']:.2f} Gcell/s, kernel {r['kernel_ms_avg']:.3f} ms, "
            f"{r['achieved_gbs']:.0f} GB/s = {r['achieved_gbs'] / HBM_PEAK_GBS:.3f} of peak")
    if not args.no_extra:
        for pulse in [p for p in ("f16", "bf16") if want(f"diel_boxes_512_{p}")]:
            sc16 = scenes.dielectric_boxes(n=512, nboxes=64, steps=args.extra_steps + 10,
                                           nprobes=16, nfreq=32, seed=1)
            sc16.pulse = pulse
            r = gpu_run(sc16, 10, args.extra_steps, f"diel_boxes_512_{pulse}", kernel_every=1)
            extra[f"diel_boxes_512_{pulse}"] = roofline_line(r, sc16)
            extra[f"diel_boxes_512_{pulse}"].update({
                "fp32_slab_fraction": round(fp32_slab_fraction(sc16), 4),
                "steps": args.extra_steps,
                "traffic": read_traffic(f"diel_boxes_512_{pulse}")})
            log(f"[bench] diel_boxes_512_{pulse}: {r['gcells']:.2f} Gcell/s, kernel "
                f"{r['kernel_ms_avg']:.3f} ms, {r['achieved_gbs']:.0f} GB/s = "
                f"{r['achieved_gbs'] / HBM_PEAK_GBS:.3f} of peak")
    if not args.no_extra and want("free_space_512_f32"):
        big = scenes.free_space_cube(n=512, steps=args.extra_steps + 10, probe_offset=10)
        r = gpu_run(big, 10, args.extra_steps, "free_space_512", kernel_every=1)
        extra["free_space_512_f32"] = roofline_line(r, big)
        extra["free_space_512_f32"]["steps"] = args.extra_steps
        log(f"[bench] free_space_512: {r['gcells']:.2f} Gcell/s, kernel {r['kernel_ms_avg']:.3f} ms, "
            f"{r['achieved_gbs']:.0f} GB/s = {r['achieved_gbs'] / HBM_PEAK_GBS:.3f} of peak")
    if not args.no_extra and want("precision"):
        extra["precision_512_crop128"] = precision_report(steps=args.precision_steps)

    if not args.no_extra and want("free_space_64"):
        extra["free_space_64"] = free_space_64_run(with_cpu=not args.no_cpu)
    for sn in (256, 512):
        if not args.no_extra and want(f"scene_setup_{sn}"):
            extra[f"scene_setup_{sn}"] = setup_run(sn)
    if not args.no_extra and want("ntff_256"):
        extra["ntff_256"] = ntff_run(256)
    if not args.no_batch and want("batched_256"):
        extra["batched_256"] = batched_run(256)
    if not args.no_batch and want("served"):
        extra[f"served_{args.serve_count}"] = served_run(args.serve_count, rate=args.serve_rate,
                                                         device_setup=not args.serve_host_setup,
                                                         copy_gbs=extra.get("device_copy_gbs"))

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline on {workload} (budget {args.cpu_budget}s)")
        cpu = cpu_baseline(scene, args.cpu_budget)
        log(f"[bench] CPU baseline {cpu['value']:.4f} Gcell/s on {cpu['cores']} threads")

    ach = main_run["achieved_gbs"]
    line = {
        "metric": METRIC,
        "value": round(main_run["gcells"], 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{workload}: {n}^3 SCN TLM, {nboxes} seeded dielectric boxes "
                        "(eps_r U[1,10], sigma U[0,0.05] S/m), PEC ground plane z=0, centre Ez "
                        "Gaussian-derivative soft source (w=20, delay 60 steps), 16 E probes, "
                        "FP64 probe DFT at 32 frequencies, Js + energy monitors, FP32 pulses",
            "grid": [n, n, n],
            "diel_fraction": round(scene.diel_fraction(), 5),
            "device": dev_name,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(workload),
            "kernel": "tlm::k_step (fused scatter+connect)",
            "kernel_ms_avg": round(main_run["kernel_ms_avg"], 5),
            "kernel_launches_timed": main_run["kernel_launches_timed"],
            "alg_bytes_per_launch": int(round(scene.ncell * main_run["alg_bytes_per_cell"])),
        },
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
