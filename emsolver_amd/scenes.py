"""Synthetic TLM workloads (grids, materials, sources, monitors).

A :class:`Scene` holds exactly what the reference's ``TlmSolver`` receives
through ``setup`` / ``addPort`` / ``addProbe`` (TlmSolver.h:208-260): the
voxel grid and material map, the material table, the time-step config and
the sources and monitors. The generators build the benchmark scenes:

* :func:`free_space_cube`  - 64^3 air cube, matched boundaries, soft Ez
  source at the centre, probe 10 cells off-centre, 1000 steps.
* :func:`dielectric_boxes` - 256^3 (or any size) with seeded axis-aligned
  dielectric boxes (eps_r in [1,10], sigma in [0,0.05] S/m), PEC ground
  plane at z=0, 16 seeded probes and a 32-frequency FP64 probe DFT.
* :func:`dipole_scene`     - a centre-fed z wire dipole (PEC cells, carved
  AIR gap fed as a port), optionally over the random-box background.
* :func:`selftest_dipole`  - the reference self-test dipole (SelfTest.cpp:
  98-178: 1 GHz, 20 cells/lambda, pad 10, 0.475 lambda wire).

Materials: 0 air, 1 PEC, id >= 2 dielectric with epsr[id-2], sigma[id-2]
(TlmSolver.h:57-65).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

C0 = 299792458.0
F0 = 1e9
CELLS_PER_LAMBDA = 20
DL = float(np.float32(C0 / F0 / CELLS_PER_LAMBDA))   # 14.9896 mm, (float)(lam/cpl)

WF_CW_RAMPED, WF_GAUSS_SINE, WF_GAUSS_PULSE, WF_GAUSS_DERIV = 0, 1, 2, 3
MAT_AIR, MAT_PEC, MAT_DIEL0 = 0, 1, 2


@dataclasses.dataclass
class Scene:
    nx: int
    ny: int
    nz: int
    mat: np.ndarray
    epsr: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.float32))
    sigma: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.float32))
    dl: float = DL
    origin: tuple = (0.0, 0.0, 0.0)
    rho: float = 0.0
    f0: float = F0
    waveform: int = WF_GAUSS_DERIV
    gd_width: float = 20.0
    gd_delay: float = 60.0
    total_steps: int = 1000
    settle_steps: int = 0
    huy_offset: int = 0
    sources: list = dataclasses.field(default_factory=list)   # (cell, pol_axis, amp)
    ports: list = dataclasses.field(default_factory=list)     # (cells, pol_axis, amp)
    probes: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int64))
    dft_freqs: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    energy_every: int = 16
    plane_axis: int = -1          # |E| cut plane (TlmSolver::setFieldPlane), -1 off
    plane_index: int = 0
    record_every: int = 0         # playback frames (TlmConfig::recordEvery), 0 off
    faces: bool = True            # PEC surface-current monitor (always on in the reference)
    pulse: str = "f32"            # pulse storage: "f32", "f16" or "bf16"
    fp32_halo: int = 8            # 16-bit storage: FP32 slabs each side of a source slab (< 0 none)
    V0: np.ndarray | None = None  # optional initial pulses (12 per cell, AoS)
    Vs0: np.ndarray | None = None
    name: str = "scene"

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    def cell(self, i, j, k) -> int:
        """VoxelGridSpec::cellIndex, Geometry.h:91-94."""
        return (k * self.ny + j) * self.nx + i

    def coords(self, cell):
        i = cell % self.nx
        j = (cell // self.nx) % self.ny
        k = cell // (self.nx * self.ny)
        return i, j, k

    def diel_fraction(self) -> float:
        return float((self.mat >= MAT_DIEL0).mean())

    def with_steps(self, steps):
        return dataclasses.replace(self, total_steps=int(steps))


def free_space_cube(n=64, steps=1000, probe_offset=10) -> Scene:
    """64^3 free-space cube, matched boundaries, Ez Gaussian-derivative soft
    source at the centre cell, Ez probe `probe_offset` cells off-centre."""
    mat = np.zeros(n * n * n, np.uint8)
    s = Scene(n, n, n, mat, total_steps=steps, name=f"free_space_{n}")
    c = n // 2
    s.sources = [(s.cell(c, c, c), 2, 1.0)]
    s.probes = np.array([s.cell(c + probe_offset, c, c)], np.int64)
    return s


def random_boxes(mat, nx, ny, nz, nboxes, rng, side=(16, 96), first_id=MAT_DIEL0):
    """Stamp `nboxes` seeded axis-aligned boxes (later boxes overwrite),
    material ids first_id.. first_id+nboxes-1. Returns (epsr, sigma)."""
    m = mat.reshape(nz, ny, nx)
    epsr = rng.uniform(1.0, 10.0, nboxes).astype(np.float32)
    sigma = rng.uniform(0.0, 0.05, nboxes).astype(np.float32)
    dims = (nx, ny, nz)
    for b in range(nboxes):
        lo, hi = [], []
        for a in range(3):
            sd = int(rng.integers(side[0], side[1] + 1))
            o = int(rng.integers(0, dims[a]))
            lo.append(o)
            hi.append(min(dims[a], o + sd))
        m[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = first_id + b
    return epsr, sigma


def dielectric_boxes(n=256, nboxes=8, steps=3000, nprobes=16, nfreq=32, seed=1,
                     ground_plane=True, side=None) -> Scene:
    """Random dielectric boxes with a PEC ground plane at z = 0, a centre Ez
    soft source, seeded probes and a 32-frequency FP64 probe DFT."""
    rng = np.random.default_rng(seed)
    mat = np.zeros(n * n * n, np.uint8)
    if side is None:
        side = (max(2, n // 16), max(3, (3 * n) // 8))     # 16..96 cells at 256^3
    epsr, sigma = random_boxes(mat, n, n, n, nboxes, rng, side=side)
    if ground_plane:
        mat.reshape(n, n, n)[0, :, :] = MAT_PEC
    s = Scene(n, n, n, mat, epsr, sigma, total_steps=steps, name=f"diel_boxes_{n}")
    c = n // 2
    s.sources = [(s.cell(c, c, c), 2, 1.0)]
    probes = []
    while len(probes) < nprobes:
        i, j, k = (int(x) for x in rng.integers(1, n - 1, 3))
        cell = s.cell(i, j, k)
        if mat[cell] != MAT_PEC and cell not in probes:
            probes.append(cell)
    s.probes = np.array(probes, np.int64)
    s.dft_freqs = 50e6 * (np.arange(nfreq) + 1.0)
    return s


def dipole_scene(nx, ny, nz, length_frac=0.5, steps=1000, background_boxes=0, seed=0,
                 gap=1, axis_cells=None) -> Scene:
    """Centre-fed z wire dipole: a one-cell PEC column of length
    length_frac*nz through the centre, with a `gap`-cell AIR feed gap at the
    centre fed as a port (addPort(gap, z, 1), SelfTest.cpp:154-176)."""
    rng = np.random.default_rng(seed)
    mat = np.zeros(nx * ny * nz, np.uint8)
    epsr = np.zeros(0, np.float32)
    sigma = np.zeros(0, np.float32)
    if background_boxes:
        side = (max(2, min(nx, ny, nz) // 16), max(3, (3 * min(nx, ny, nz)) // 8))
        epsr, sigma = random_boxes(mat, nx, ny, nz, background_boxes, rng, side=side)
    s = Scene(nx, ny, nz, mat, epsr, sigma, total_steps=steps, name=f"dipole_{nx}x{ny}x{nz}")
    ci, cj, ck = nx // 2, ny // 2, nz // 2
    half = max(gap + 1, int(round(length_frac * nz / 2)))
    m = mat.reshape(nz, ny, nx)
    k0 = max(1, ck - half)
    k1 = min(nz - 2, ck + half)
    m[k0:k1 + 1, cj, ci] = MAT_PEC
    gap_cells = []
    g0 = ck - (gap - 1) // 2
    for k in range(g0, g0 + gap):
        m[k, cj, ci] = MAT_AIR
        gap_cells.append(s.cell(ci, cj, k))
    s.ports = [(gap_cells, 2, 1.0)]
    return s


def selftest_dipole(waveform=WF_GAUSS_SINE, extra_steps=0, diel_sphere=False) -> Scene:
    """The self-test dipole of the reference (SelfTest.cpp:98-178):
    0.475-lambda z wire dipole with a lambda/25 gap, 1 GHz, 20 cells/lambda,
    pad 10, matched boundaries, optional eps_r=3 sphere beside it,
    settle = 2(nx+ny+nz) + 3*spp, total = settle + 12*spp + extra."""
    f0 = 1e9
    lam = C0 / f0
    cpl, pad = 20, 10
    dl = np.float32(lam / cpl)
    L = np.float32(0.475 * lam)
    g = np.float32(lam / 25.0)
    # wire bounds: x = y = 0, z in [-L/2, L/2]
    lo = np.array([0.0, 0.0, -L / 2], np.float32)
    hi = np.array([0.0, 0.0, L / 2], np.float32)
    if diel_sphere:
        hi[0] = np.float32(0.55 * lam)
    size = hi - lo
    nn = [int(math.ceil(float(size[a]) / float(dl))) + 1 + 2 * pad for a in range(3)]
    nx, ny, nz = nn
    cen = (lo + hi) * np.float32(0.5)
    origin = np.array([cen[a] - np.float32(nn[a]) * dl * np.float32(0.5) for a in range(3)],
                      np.float32)
    mat = np.zeros(nx * ny * nz, np.uint8)
    m = mat.reshape(nz, ny, nx)
    epsr = np.zeros(0, np.float32)
    sigma = np.zeros(0, np.float32)

    def cell_of(p):
        return [int(math.floor((float(p[a]) - float(origin[a])) / float(dl))) for a in range(3)]

    if diel_sphere:
        epsr = np.array([3.0], np.float32)
        sigma = np.array([0.0], np.float32)
        sc = np.array([0.42 * lam, 0.0, 0.0])
        r = 0.12 * lam
        ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        cx = origin[0] + (ii + 0.5) * dl
        cy = origin[1] + (jj + 0.5) * dl
        cz = origin[2] + (kk + 0.5) * dl
        inside = (cx - sc[0]) ** 2 + (cy - sc[1]) ** 2 + (cz - sc[2]) ** 2 <= r * r
        m[kk[inside], jj[inside], ii[inside]] = MAT_DIEL0
    # wire arms (VoxelizeWire with radius 0, Geometry.cpp:608-642)
    for za, zb in ((g / 2, L / 2), (-g / 2, -L / 2)):
        length = abs(float(zb - za))
        nst = max(1, int(math.ceil(length / (float(dl) * 0.25))))
        for q in range(nst + 1):
            z = za + (zb - za) * np.float32(q / nst)
            i, j, k = cell_of((0.0, 0.0, z))
            if 0 <= i < nx and 0 <= j < ny and 0 <= k < nz:
                m[k, j, i] = MAT_PEC
    # carve and collect the feed gap (SelfTest.cpp:141-158)
    a = np.array([0, 0, -g / 2], np.float32)
    b = np.array([0, 0, g / 2], np.float32)
    d = b - a
    a2 = a + d * np.float32(0.15)
    b2 = b - d * np.float32(0.15)
    cells = set()
    for q in range(17):
        pp = a2 + (b2 - a2) * np.float32(q / 16)
        i, j, k = cell_of(pp)
        if 0 <= i < nx and 0 <= j < ny and 0 <= k < nz:
            m[k, j, i] = MAT_AIR
            cells.add((k * ny + j) * nx + i)
    spp = 2 * cpl
    settle = 2 * (nx + ny + nz) + 3 * spp
    total = settle + 12 * spp + extra_steps
    s = Scene(nx, ny, nz, mat, epsr, sigma, dl=float(dl), origin=tuple(float(x) for x in origin),
              rho=0.0, f0=f0, waveform=waveform, total_steps=total, settle_steps=settle,
              huy_offset=pad // 2, name="selftest_dipole")
    s.ports = [(sorted(cells), 2, 1.0)]
    return s


def batch_dipoles(count=256, lo=32, hi=128, steps=(500, 2000), seed=3, nprobes=2,
                  length=(0.3, 0.7), background_boxes=8, pulse="f32") -> list:
    """Independent grids for the batched run: each axis log-uniform in
    [lo, hi] cells, a centre-fed z wire dipole of length U[length]*nz over a
    seeded random-dielectric background (the dielectric-box generator scaled
    to the grid), steps uniform in [steps], `nprobes` seeded E probes, port
    V/I recorded. No surface-current monitor (batched members record probes
    and ports)."""
    rng = np.random.default_rng(seed)
    out = []
    for g in range(count):
        nx, ny, nz = (int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
                      for _ in range(3))
        frac = float(rng.uniform(*length))
        st = int(rng.integers(steps[0], steps[1] + 1))
        sc = dipole_scene(nx, ny, nz, length_frac=frac, steps=st,
                          background_boxes=background_boxes, seed=int(rng.integers(1 << 31)))
        sc.faces = False
        sc.pulse = pulse
        probes = []
        while len(probes) < nprobes:
            i, j, k = int(rng.integers(1, nx - 1)), int(rng.integers(1, ny - 1)), int(
                rng.integers(1, nz - 1))
            c = sc.cell(i, j, k)
            if sc.mat[c] != MAT_PEC and c not in probes:
                probes.append(c)
        sc.probes = np.array(probes, np.int64)
        sc.name = f"batch{g}_{nx}x{ny}x{nz}"
        out.append(sc)
    return out
