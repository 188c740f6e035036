"""Batched execution: many grids advanced by the same launches give every
grid the same bits as the oracle (and so as a run alone), under every
launch shape, with and without HIP graphs, with members joining and leaving
between blocks of ticks."""
import dataclasses

import numpy as np
import pytest

from emsolver_amd import scenes, tlm
from tests import helpers
from tests.test_gpu_parity import _random_scene

pytestmark = pytest.mark.gpu


def _mixed_scenes():
    out = scenes.batch_dipoles(count=4, lo=12, hi=40, steps=(30, 90), seed=11, nprobes=2)
    r = _random_scene(21, 9, 13, seed=4, rho=0.37, steps=40)
    r.faces = False
    r.energy_every = 0
    out.append(r)
    out.append(dataclasses.replace(out[0], pulse="f16", fp32_halo=2, name="f16_member"))
    out.append(dataclasses.replace(out[1], pulse="bf16", fp32_halo=-1, name="bf16_member"))
    nd = _random_scene(17, 11, 8, seed=6, diel=False, steps=25)
    nd.faces = False
    nd.energy_every = 0
    out.append(nd)
    return out


def _oracle(sc):
    return helpers.oracle_run(sc)


def _compare(g, o, sc):
    assert np.abs(o.V).max() > 0
    helpers.assert_bitexact(f"{sc.name} V", g.V, o.V)
    if o.Vs is not None:
        helpers.assert_bitexact(f"{sc.name} Vs", g.Vs, o.Vs)
    if len(sc.probes):
        assert np.abs(o.probes).max() > 0
        helpers.assert_bitexact(f"{sc.name} probes", g.probes, o.probes)
    if sc.ports:
        helpers.assert_bitexact(f"{sc.name} ports", g.ports, o.ports)


BATCH_TUNINGS = [
    {"planes": 1, "block_waves": 4, "grid_blocks": 0, "graph_ticks": 0},
    {"planes": 1, "block_waves": 4, "grid_blocks": 0, "graph_ticks": 16},
    {"planes": 3, "block_waves": 2, "grid_blocks": 7, "graph_ticks": 5},
    {"planes": 0, "block_waves": 1, "grid_blocks": 300, "graph_ticks": 0},
]


@pytest.mark.parametrize("tuning", BATCH_TUNINGS,
                         ids=lambda t: "z{planes}_w{block_waves}_g{grid_blocks}_L{graph_ticks}".format(**t))
def test_batch_bitexact(tuning):
    scs = _mixed_scenes()
    print(f"[batch] {len(scs)} members tuning {tuning}", flush=True)
    res = tlm.run_scenes_batched(scs, tuning=tuning, block=17)
    for g, sc in zip(res, scs):
        _compare(g, _oracle(sc), sc)


def test_batch_thin_members():
    """Members one or two cells thick next to an ordinary one."""
    scs = []
    for i, shape in enumerate([(1, 1, 1), (4, 5, 1), (1, 1, 70), (65, 1, 3), (2, 2, 2)]):
        n = shape[0] * shape[1] * shape[2]
        sc = _random_scene(*shape, seed=n + 13, rho=0.37, pec_frac=0.0 if n < 8 else 0.1,
                           steps=20 + 3 * i)
        sc.faces = False
        sc.energy_every = 0
        scs.append(sc)
    scs.append(dataclasses.replace(scs[3], pulse="f16", fp32_halo=1, name="thin_f16"))
    scs.append(scenes.batch_dipoles(count=1, lo=12, hi=20, steps=(30, 31), seed=3, nprobes=2)[0])
    print(f"[batch] {len(scs)} thin members", flush=True)
    res = tlm.run_scenes_batched(scs, block=9)
    for g, sc in zip(res, scs):
        _compare(g, _oracle(sc), sc)


def test_batch_join_and_leave():
    """Continuous batching: members join at different ticks, finished ones
    leave, new ones take their place; every result is the oracle's."""
    scs = _mixed_scenes()
    devs = {}
    with tlm.TlmBatch() as b:
        b.set_tuning(planes=2, block_waves=4, grid_blocks=0, graph_ticks=8)
        pending = list(range(len(scs)))
        done = {}
        # three members first, the rest join as others finish or every block
        for i in pending[:3]:
            devs[i] = tlm.open_scene(scs[i])
            b.add(devs[i], scs[i].total_steps)
        pending = pending[3:]
        blocks = 0
        while b.members or pending:
            b.run(13)
            b.sync()
            blocks += 1
            for d, left in list(b.remaining().items()):
                if left == 0:
                    i = next(k for k, v in devs.items() if v is d)
                    b.remove(d)
                    done[i] = tlm.collect(d, scs[i], scs[i].total_steps)
                    d.close()
            if pending and (blocks % 2 == 0 or len(b.members) < 2):
                i = pending.pop(0)
                devs[i] = tlm.open_scene(scs[i])
                b.add(devs[i], scs[i].total_steps)
    assert sorted(done) == list(range(len(scs)))
    for i, sc in enumerate(scs):
        _compare(done[i], _oracle(sc), sc)


def test_batch_member_runs_alone_before_and_after():
    """A context can run alone, join a batch for some steps, leave and go on
    alone; the result is still the oracle's."""
    sc = _mixed_scenes()[4]
    dev = tlm.open_scene(sc)
    dev.run(0, 7)
    with tlm.TlmBatch() as b:
        b.add(dev, 20)
        b.run(25)
        b.sync()
        b.remove(dev)
    dev.run(27, sc.total_steps - 27)
    dev.sync()
    g = tlm.collect(dev, sc, sc.total_steps)
    dev.close()
    _compare(g, _oracle(sc), sc)


def test_autotune_keeps_results():
    scs = _mixed_scenes()[:5]
    res = tlm.run_scenes_batched(scs, autotune=1, block=11)
    for g, sc in zip(res, scs):
        _compare(g, _oracle(sc), sc)


def test_batch_rules():
    sc = _mixed_scenes()[0]
    with tlm.TlmBatch() as b:
        bad = dataclasses.replace(sc, faces=True)
        with tlm.open_scene(bad) as d:
            with pytest.raises(tlm.TlmError, match="probes and ports"):
                b.add(d, 5)
        # energy records would stop while batched: refused, not dropped
        with tlm.open_scene(dataclasses.replace(sc, energy_every=16)) as d:
            with pytest.raises(tlm.TlmError, match="energy"):
                b.add(d, 5)
        with tlm.open_scene(sc) as d:
            b.add(d, 5)
            with pytest.raises(tlm.TlmError, match="batch member"):
                d.run(0, 1)
            with pytest.raises(tlm.TlmError, match="already"):
                b.add(d, 5)
            b.remove(d)
