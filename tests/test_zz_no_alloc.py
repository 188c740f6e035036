"""tlm_far_field_ctx_noalloc, the far-field entry point that allocates no
device memory (its workspace is the library's fixed device array), against
tlm_far_field_ctx through the C ABI: two calls in a row, the second with a
smaller input (a smaller context with fewer Huygens faces, fewer
directions), each bit for bit equal to the allocating entry point; inputs
beyond the fixed workspace are refused."""
import dataclasses

import numpy as np
import pytest

from emsolver_amd import scenes, tlm

pytestmark = pytest.mark.gpu


def _run_to_end(sc):
    dev = tlm.open_scene(sc)
    n = 0
    while n < sc.total_steps:
        m = min(256, sc.total_steps - n)
        dev.run(n, m)
        n += m
    dev.sync()
    return dev


def _same(a, b):
    assert np.array_equal(a.U, b.U)
    assert a.directivity == b.directivity and a.u_max == b.u_max
    assert a.peak_theta_deg == b.peak_theta_deg and a.peak_phi_deg == b.peak_phi_deg


def test_noalloc_far_field_equals_allocating_entry_point():
    big = scenes.selftest_dipole(scenes.WF_GAUSS_SINE)
    small = dataclasses.replace(scenes.dielectric_boxes(n=36, nboxes=3, steps=120, nprobes=2,
                                                        nfreq=0, seed=4),
                                huy_offset=3, settle_steps=20)
    print(f"[noalloc] {big.name} {big.nx}x{big.ny}x{big.nz}, then {small.name}", flush=True)
    with _run_to_end(big) as dev:
        samples = tlm.collect(dev, big, big.total_steps, state=False).huy_samples
        assert samples > 0
        ref = tlm.far_field_ctx(dev, big.origin, samples, big.f0)
        got = tlm.far_field_ctx(dev, big.origin, samples, big.f0, noalloc=True)
        _same(got, ref)
        assert 1.1 < got.directivity < 2.6
    with _run_to_end(small) as dev:
        samples = tlm.collect(dev, small, small.total_steps, state=False).huy_samples
        assert samples > 0
        ref = tlm.far_field_ctx(dev, small.origin, samples, small.f0, n_theta=19, n_phi=37)
        got = tlm.far_field_ctx(dev, small.origin, samples, small.f0, n_theta=19, n_phi=37,
                                noalloc=True)
        _same(got, ref)
        # beyond the fixed workspace: refused, nothing run
        with pytest.raises(tlm.TlmError, match="workspace"):
            tlm.far_field_ctx(dev, small.origin, samples, small.f0, n_theta=1500, n_phi=1500,
                              noalloc=True)
        # the entry point still works after the refusal
        _same(tlm.far_field_ctx(dev, small.origin, samples, small.f0, noalloc=True),
              tlm.far_field_ctx(dev, small.origin, samples, small.f0))
