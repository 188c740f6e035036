"""The 16-bit pulse storage scheme of the oracle (FP16 / BF16 codes times a
per-slab power of two, FP32 slabs near soft sources).

The reference has no reduced-precision mode; what pins the restatement:
  * the binary16 rounding is checked against numpy's IEEE conversion and the
    bfloat16 rounding against a float64 nearest-even construction;
  * the scale planner's exponent rule is checked on its defining inequality;
  * power-of-two scaling makes the whole 16-bit run exactly homogeneous:
    sources 2^20 or 2^-30 times larger give results exactly 2^20 / 2^-30
    times larger (plain binary16 would overflow / flush to zero there);
  * the physics error against the FP64 oracle stays within stated bounds,
    and the FP32 halo around the source is what brings FP16 to ~1e-3.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from emsolver_amd import scenes
from oracle import oracle_tlm as O
from tests import helpers


def _probe_rel_err(r, ref):
    """max |p - p_ref| / max |p_ref| per probe (over steps and components)."""
    return np.abs(r.probes - ref.probes).max(axis=(0, 2)) / np.abs(ref.probes).max(axis=(0, 2))


def test_f16_rounding_matches_ieee():
    rng = np.random.default_rng(0)
    n = 400_000
    e = rng.integers(-30, 17, n)
    x = ((rng.random(n) + 1.0) * 2.0 ** e * rng.choice([-1, 1], n)).astype(np.float32)
    # exact ties halfway between neighbouring binary16 values (normal and subnormal)
    h = rng.integers(0, 0x7bff, 50_000).astype(np.uint16)
    lo = h.view(np.float16).astype(np.float64)
    hi = (h + 1).astype(np.uint16).view(np.float16).astype(np.float64)
    ties = ((lo + hi) / 2).astype(np.float32)
    special = np.float32([0.0, -0.0, 65504.0, 65519.996, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                          1e-9, 6.1e-5, 6.2e-5])
    x = np.concatenate([x, ties, -ties, special])
    got = O.encode_pulses(1, x)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    # decoding is exact: every finite code round-trips
    codes = np.arange(0x7c00, dtype=np.uint16)
    vals = codes.view(np.float16).astype(np.float32)
    assert np.array_equal(O.encode_pulses(1, vals), codes)


def test_bf16_rounding_is_nearest_even():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(300_000) * 10.0 ** rng.integers(-20, 20, 300_000)).astype(np.float32)
    u = x.view(np.uint32)
    # ties: low 16 bits exactly 0x8000
    ties = ((u[:20000] & 0xFFFF0000) | 0x8000).view(np.float32)
    x = np.concatenate([x, ties])
    got = O.encode_pulses(2, x)
    # float64 construction: the two bfloat16 neighbours, pick the nearer, ties to even
    t = (x.view(np.uint32) & 0xFFFF0000).view(np.float32).astype(np.float64)
    step = np.abs(t) * 0 + 2.0 ** (np.floor(np.log2(np.abs(t) + 1e-300)) - 7)
    up = t + np.sign(x) * step
    d_lo = np.abs(x.astype(np.float64) - t)
    d_hi = np.abs(up - x.astype(np.float64))
    lo_code = (x.view(np.uint32) >> 16).astype(np.uint16)
    want = np.where(d_lo < d_hi, lo_code, lo_code + 1)
    tie = d_lo == d_hi
    want = np.where(tie, np.where(lo_code % 2 == 0, lo_code, lo_code + 1), want).astype(np.uint16)
    assert np.array_equal(got, want)


def test_plan_exponent_rule():
    rng = np.random.default_rng(2)
    b = (rng.random(20000) * 2.0 ** rng.integers(-80, 80, 20000)).astype(np.float32)
    b = b[b > 0]
    for v in b[:5000]:
        E = O.plan_exponent(float(v))
        s = float(v) * 2.0 ** -E
        assert s < 2.0 ** 15
        if -100 < E < 100:  # E is clamped to [-100, 100]
            assert 2.0 ** 14 <= s
    assert O.plan_exponent(0.0) == 0
    assert O.plan_exponent(1.0) == -14
    assert O.plan_exponent(1e-38) == -100


def _small_boxes(steps=200, pulse="f16", halo=8, n=20):
    sc = scenes.dielectric_boxes(n=n, nboxes=3, steps=steps, nprobes=4, nfreq=0, seed=7)
    return dataclasses.replace(sc, pulse=pulse, fp32_halo=halo)


@pytest.mark.parametrize("pulse", ["f16", "bf16"])
@pytest.mark.parametrize("halo", [-1, 3])
def test_power_of_two_scaling_is_exact(pulse, halo):
    base = _small_boxes(steps=150, pulse=pulse, halo=halo)
    r0 = helpers.oracle_run(base)
    assert np.abs(r0.probes).max() > 0
    for k in (20, -30):
        sc = dataclasses.replace(base, sources=[(c, p, a * 2.0 ** k) for c, p, a in base.sources])
        r = helpers.oracle_run(sc)
        assert np.array_equal(r.probes, r0.probes * 2.0 ** k), k
        assert np.array_equal(r.V, r0.V * 2.0 ** k), k


def test_scales_cover_the_state():
    sc = _small_boxes(steps=180, pulse="f16", halo=-1)
    o = helpers.oracle_open(sc)
    for _ in range(6):
        o.run_steps(30)
        E, M = o.slab_exponents()
        V, _ = o.state()
        slab = sc.nx * sc.ny
        vmax = np.abs(V.reshape(sc.nz, slab * 12)).max(axis=1)
        assert np.all(np.isfinite(V))
        assert np.all(vmax * 2.0 ** -E.astype(np.float64) <= 65504.0)
        # the scale is not wasted: the largest value of a live slab codes above 4
        live = vmax > 0
        assert np.all(vmax[live] * 2.0 ** -E[live].astype(np.float64) > 2.0 ** 2)
    o.close()


# measured on these scenes (600 steps): FP16 with the default 8-slab FP32 halo
# 1e-4 (free space), <= 9e-4 (dielectric boxes), <= 1.5e-3 (dipole probes),
# <= 6e-5 (dipole port V/I); BF16 about 10x more. The bounds below leave ~1.5x.
ERROR_CASES = [
    ("free", "f16", 8, 2e-4, None),
    ("free", "bf16", 8, 2e-3, None),
    ("boxes", "f16", 8, 1.5e-3, None),
    ("dipole", "f16", 8, 2.5e-3, 1e-4),
    ("dipole", "bf16", 8, 2.5e-2, 1.5e-3),
]


def _error_scene(which, steps=600):
    if which == "free":
        return scenes.free_space_cube(n=32, steps=steps, probe_offset=8)
    if which == "boxes":
        return scenes.dielectric_boxes(n=32, nboxes=3, steps=steps, nprobes=4, nfreq=0, seed=3)
    sc = scenes.dipole_scene(28, 28, 40, length_frac=0.5, steps=steps, background_boxes=3, seed=2)
    rng = np.random.default_rng(5)
    sc.probes = np.array([sc.cell(int(rng.integers(2, 26)), int(rng.integers(2, 26)),
                                  int(rng.integers(2, 38))) for _ in range(4)], np.int64)
    return sc


@pytest.mark.slow
@pytest.mark.parametrize("which,pulse,halo,pbound,portbound", ERROR_CASES)
def test_error_against_fp64(which, pulse, halo, pbound, portbound):
    sc = _error_scene(which)
    r64 = helpers.oracle_run(sc, fp64=True)
    r = helpers.oracle_run(dataclasses.replace(sc, pulse=pulse, fp32_halo=halo))
    e = _probe_rel_err(r, r64)
    assert e.max() <= pbound, e
    if portbound is not None:
        ep = np.abs(r.ports - r64.ports).max(axis=0) / np.abs(r64.ports).max(axis=0)
        assert ep.max() <= portbound, ep
    # the FP32 halo is what pays: without it FP16 is an order of magnitude worse near a soft source
    if which == "free" and pulse == "f16":
        r_nohalo = helpers.oracle_run(dataclasses.replace(sc, pulse=pulse, fp32_halo=-1))
        assert _probe_rel_err(r_nohalo, r64).max() > 10 * e.max()


def test_fp32_storage_matches_fp64_closely():
    """The FP32 path against the FP64 oracle: within 1e-5 max relative probe
    error after the run."""
    sc = _error_scene("boxes", steps=300)
    e = _probe_rel_err(helpers.oracle_run(sc), helpers.oracle_run(sc, fp64=True))
    assert e.max() < 1e-5, e
