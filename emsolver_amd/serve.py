"""Resident serving loop: TLM simulations of different sizes and step counts
arrive over time, are admitted against free HBM into one running batch
(continuous batching: members join and leave between blocks of ticks while
the others keep running), and complete with their probe and port records.

The device work is the batched engine of the C ABI (``tlm_batch_*``): a block
of ticks is one captured HIP graph. While a block runs, the host reads the
records of the members that left after the previous block and sets up the
next admitted requests (their contexts run their set-up on their own
streams), so the device does not wait for that host work. A member's result is the same bits alone
and in the mix (each grid's arithmetic does not depend on its neighbours in
the batch); tests/test_gpu_serve.py checks exactly that.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import threading
import time

import numpy as np

from . import scenes, tlm


@dataclasses.dataclass
class Request:
    rid: int
    scene: object          # scenes.Scene (built lazily by `build` when None)
    arrival: float         # seconds after the start of the stream
    tol: float = 1e-5      # tolerance the client asks for; 1e-3 allows FP16 storage
    build: object = None   # callable returning the Scene (generated off the serving loop)


@dataclasses.dataclass
class Completion:
    rid: int
    arrival: float
    admitted: float
    finished: float
    cells: int
    steps: int
    pulse: str
    probes: np.ndarray | None = None
    ports: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


# Storage policy: the error model of FP16 pulse storage for a request that
# tolerates 1e-3 of probe error against the FP64 oracle (max over the run of
# |x - x64| / max |x64| per probe). Measured with the oracle's bit-exact
# restatement of the 16-bit scheme on served-stream requests (dipole over
# 8 dielectric boxes, 500-4000 steps):
#   * a probe read from an FP16 slab: 1.3e-3 to 1.8e-3 (fails at any size);
#   * probes and the whole dipole inside the FP32 band: <= 5.5e-4 for grids
#     up to 80 cells per axis (probes >= 2 slabs inside the band, FP16 slab
#     shares up to 0.52), 7.8e-4 at 99 x 135 x 122 (share 0.25); the error
#     grows with the grid's extent.
# So a 1e-3 request is stored in FP16 only if its FP32 band (source slab +-
# halo) holds the dipole and every probe with F16_MARGIN slabs to spare, no
# axis exceeds F16_MAX_AXIS (the sizes the model was checked on), and at
# least F16_MIN_SHARE of the slabs stay 16-bit (below that the 16-bit step,
# slower per byte than the FP32 one, would not pay). Everything else is FP32.
F16_TOL = 1e-3
F16_MARGIN = 4
F16_MAX_AXIS = 112
F16_MIN_SHARE = 0.4


def plan_storage(tol: float, nz: int, dims, src_slabs, keep_slabs):
    """(pulse, fp32_halo, fp16_slab_share) of a request: src_slabs are the z
    slabs of its soft sources and ports (the FP32 band is their range +-
    halo), keep_slabs the slabs that must lie inside the band (its PEC
    structure, its probes)."""
    if tol < F16_TOL or not src_slabs:
        return "f32", 8, 0.0
    lo_s, hi_s = min(src_slabs), max(src_slabs)
    reach = max([0] + [lo_s - k for k in keep_slabs] + [k - hi_s for k in keep_slabs])
    halo = reach + F16_MARGIN
    band = min(nz - 1, hi_s + halo) - max(0, lo_s - halo) + 1
    share = 1.0 - band / nz
    if max(dims) > F16_MAX_AXIS or share < F16_MIN_SHARE:
        return "f32", 8, share
    return "f16", halo, share


class TlmServer:
    """One device, one batch, requests admitted by free HBM."""

    def __init__(self, device=0, block_ticks=32, hbm_reserve=4 << 30, max_members=1024,
                 tuning=None, autotune_ticks=0, keep_results=True, progress_s=0.0,
                 max_cells=0, retune_s=60.0):
        self.device = device
        self.block_ticks = int(block_ticks)
        self.hbm_reserve = int(hbm_reserve)
        self.max_members = int(max_members)
        self.tuning = tuning
        self.autotune_ticks = int(autotune_ticks)
        # the launch shape is tuned once the batch holds a representative
        # load (half the cells in flight, or 8 members without a cell cap)
        # and again every retune_s seconds, as the mix of members changes
        self.retune_s = float(retune_s)
        self.keep_results = keep_results
        self.progress_s = float(progress_s)  # > 0: a progress line on stderr this often
        # > 0: admit while the cells in flight stay below this (first come,
        # first served; enough cells to keep HBM streaming, few enough that a
        # request's ticks stay short)
        self.max_cells = int(max_cells)
        self.stats = {}

    def _need_bytes(self, sc) -> int:
        b = tlm.estimate_bytes(sc.nx, sc.ny, sc.nz, sc.pulse, sc.has_diel)
        if sc.ops is not None:
            b += sc.ncell  # the device material map of a device-built scene
        if sc.pulse != "f32" and sc.fp32_halo >= 0:
            # FP32 slabs near the sources take 8 bytes per pair instead of 4
            b += int(2 * 12 * sc.nx * sc.ny * min(sc.nz, 2 * sc.fp32_halo + 2) * 2)
        steps = sc.total_steps
        b += steps * (len(sc.probes) * 12 + len(sc.ports) * 8) + (1 << 20)
        return b

    def serve(self, requests, realtime=True, prefetch=8):
        """Serve `requests` (sorted by arrival). With realtime=False arrivals
        are ignored (every request is available at once, admission still by
        memory and max_members). Returns the completions in finishing order."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        # scenes are generated ahead of the loop on a worker thread
        ready = {}
        cond = threading.Condition()

        def producer():
            for r in reqs:
                sc = r.scene if r.scene is not None else r.build()
                with cond:
                    while len(ready) >= prefetch:
                        cond.wait()
                    ready[r.rid] = sc
                    cond.notify_all()

        th = threading.Thread(target=producer, daemon=True)
        th.start()

        def take(rid):
            with cond:
                while rid not in ready:
                    cond.wait()
                sc = ready.pop(rid)
                cond.notify_all()
                return sc

        batch = tlm.TlmBatch(self.device)
        if self.tuning:
            batch.set_tuning(**self.tuning)
        t0 = time.perf_counter()
        now = (lambda: time.perf_counter() - t0) if realtime else (lambda: 0.0)
        arrivals = collections.deque(reqs)
        queue = collections.deque()
        live = {}       # TlmDevice -> (request, scene, admitted time): in the batch
        pending = []    # [(TlmDevice, request, scene, admitted)]: set up, joining next block
        finished = []   # [(TlmDevice, request, scene, admitted)]: left the batch, to read
        done = []
        tuned = False
        retried = []    # requests solved again in FP32 after a 16-bit range error
        ticks = 0
        member_ticks = 0
        last_report = time.perf_counter()

        def admit():
            """Set up the queue's head requests that fit (FIFO, free HBM, cells
            in flight); their contexts join the batch at the next block."""
            while queue and len(live) + len(pending) < self.max_members:
                r = queue[0]
                sc = take(r.rid)
                need = self._need_bytes(sc)
                free, _ = tlm.device_memory(self.device)
                cells = sum(v[1].ncell for v in live.values()) + sum(p[2].ncell for p in pending)
                over = self.max_cells > 0 and cells + sc.ncell > self.max_cells
                if (live or pending) and (need + self.hbm_reserve > free or over):
                    with cond:  # put it back; try again after members leave
                        ready[r.rid] = sc
                        cond.notify_all()
                    return
                queue.popleft()
                pending.append((tlm.open_scene(sc, self.device), r, sc, now()))

        def complete():
            """Read the records of the members that left and free them."""
            for dev, r, sc, adm in finished:
                c = Completion(r.rid, r.arrival, adm, 0.0, sc.ncell, sc.total_steps, sc.pulse)
                if self.keep_results:
                    if dev.nprobe:
                        c.probes = dev.read_probes(0, sc.total_steps)
                    if dev.nport:
                        c.ports = dev.read_ports(0, sc.total_steps)
                dev.close()
                c.finished = now() if realtime else time.perf_counter() - t0
                done.append(c)
            finished.clear()

        while arrivals or queue or live or pending or finished:
            if self.progress_s > 0 and time.perf_counter() - last_report >= self.progress_s:
                last_report = time.perf_counter()
                import sys
                print(f"[serve] t={last_report - t0:.0f}s done={len(done)} live={len(live)} "
                      f"queued={len(queue)} to_arrive={len(arrivals)}", file=sys.stderr, flush=True)
            t = now()
            while arrivals and arrivals[0].arrival <= t:
                queue.append(arrivals.popleft())
            # the contexts set up during the last block join now
            for dev, r, sc, adm in pending:
                batch.add(dev, sc.total_steps)
                live[dev] = (r, sc, adm)
            pending.clear()
            if live:
                if self.autotune_ticks:
                    cells = sum(v[1].ncell for v in live.values())
                    loaded = cells >= 0.5 * self.max_cells if self.max_cells > 0 \
                        else len(live) >= 8
                    due = tuned is False or (self.retune_s > 0 and
                                             time.perf_counter() - tuned >= self.retune_s)
                    if loaded and due:
                        batch.autotune(self.autotune_ticks)
                        tuned = time.perf_counter()
                batch.run(self.block_ticks)          # queued; the host goes on below
                ticks += self.block_ticks
                member_ticks += self.block_ticks * len(live)
            # while the block runs: results out, new requests set up
            complete()
            admit()
            if not live:
                if not pending and arrivals and realtime:
                    time.sleep(max(0.0, min(0.05, arrivals[0].arrival - now())))
                continue
            batch.sync()
            for dev, left in batch.remaining().items():
                if left > 0:
                    continue
                r, sc, adm = live.pop(dev)
                try:
                    batch.remove(dev)
                except tlm.TlmError as e:
                    if e.code != tlm.ERR_RANGE:
                        raise
                    # its 16-bit storage outgrew a slab scale: its results are
                    # invalid (the other members are not affected); it is
                    # solved again, in FP32, ahead of the queue
                    dev.close()
                    retried.append(r.rid)
                    with cond:
                        ready[r.rid] = dataclasses.replace(sc, pulse="f32")
                        cond.notify_all()
                    queue.appendleft(r)
                    continue
                finished.append((dev, r, sc, adm))
        self.stats = {"ticks": ticks, "mean_members": member_ticks / max(1, ticks),
                      "wall_s": time.perf_counter() - t0, "tuning": batch.get_tuning(),
                      "retried_f32": retried}
        batch.close()
        return done


def summarize(done, wall_s=None) -> dict:
    """p50 / p99 latency and aggregate Gcell-updates/s of a served stream."""
    lat = np.array([c.latency for c in done])
    work = float(sum(c.cells * c.steps for c in done))
    t_first = min(c.arrival for c in done)
    t_last = max(c.finished for c in done)
    span = wall_s if wall_s is not None else (t_last - t_first)
    return {
        "requests": len(done),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "gcell_s": work / span / 1e9,
        "cell_updates": work,
        "makespan_s": span,
        "f16_requests": int(sum(c.pulse == "f16" for c in done)),
    }


def served_stream(count=2000, rate=50.0, lo=32, hi=384, steps=(500, 4000), seed=5,
                  f16_share=0.5, nprobes=2, device_setup=False, policy=True):
    """The served workload: Poisson arrivals at `rate`/s, each axis log-uniform
    in [lo, hi], the dielectric-box background with a centre-fed dipole of
    seeded length, steps uniform in [steps], per-request tolerance 1e-3 or
    1e-5 with a seeded split; storage by plan_storage (policy=False: FP16
    with the default FP32 halo for every 1e-3 request, for tests of the
    batch with both formats). Scenes are built on demand (Request.build): as
    host material maps, or with device_setup as set-up operations that the
    device voxelises at admission (the same grids)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        dims = tuple(int(round(math.exp(rng.uniform(math.log(lo), math.log(hi))))) for _ in range(3))
        frac = float(rng.uniform(0.3, 0.7))
        st = int(rng.integers(steps[0], steps[1] + 1))
        tol = 1e-3 if rng.random() < f16_share else 1e-5
        sseed = int(rng.integers(1 << 31))

        def build(dims=dims, frac=frac, st=st, tol=tol, sseed=sseed, rid=rid):
            return request_scene(dims, frac, st, tol, sseed, nprobes, f"req{rid}", device_setup,
                                 policy)

        out.append(Request(rid, None, t, tol, build))
    return out


def request_scene(dims, length_frac, steps, tol, seed, nprobes=2, name="req",
                  device_setup=False, policy=True):
    """One served request: the centre-fed dipole over 8 dielectric boxes,
    `nprobes` seeded E probes off the dipole, stored as plan_storage decides
    for the tolerance `tol` (policy=False: FP16 with the default FP32 halo
    whenever tol allows 1e-3)."""
    nx, ny, nz = dims
    sc = scenes.dipole_scene(nx, ny, nz, length_frac=length_frac, steps=steps,
                             background_boxes=8, seed=seed, device_setup=device_setup)
    sc.faces = False
    sc.energy_every = 0  # batch members record probes and ports
    rng = np.random.default_rng(seed + 1)
    probes = []
    while len(probes) < nprobes:
        i, j, k = (int(rng.integers(1, nx - 1)), int(rng.integers(1, ny - 1)),
                   int(rng.integers(1, nz - 1)))
        c = sc.cell(i, j, k)
        # off the dipole's column (its only PEC cells)
        if (i, j) != (nx // 2, ny // 2) and c not in probes:
            probes.append(c)
    sc.probes = np.array(probes, np.int64)
    sc.name = name
    # the band: the feed gap's slabs (the port), holding the wire column and
    # the probes (a device-built scene has no host map to read them from)
    k0, k1, g0, gap = sc.dipole
    keep = [k0, k1] + [int(c) // (nx * ny) for c in sc.probes]
    if policy:
        sc.pulse, sc.fp32_halo, _ = plan_storage(tol, nz, dims, list(range(g0, g0 + gap)), keep)
    else:
        sc.pulse = "f16" if tol >= F16_TOL else "f32"
    return sc
