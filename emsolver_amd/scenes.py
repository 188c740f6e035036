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
* :func:`mesh_scene`       - a scene of objects in the reference's form
  (triangle meshes, wires, a fed gap) voxelised on the device in
  TMainForm::voxelizeScene order (MainUnit.cpp:431-459).

A scene is either a host material map (``mat``) or a list of set-up
operations (``ops``, with ``mat`` None) that the device runs on the
context's material map (:func:`emsolver_amd.tlm.open_scene`):
  ("fill", id) | ("cells_box", lo, hi, id)
  | ("mesh", verts, idx, offset, id, solid) | ("wire", pts, radius, offset, id)
  | ("feed", a, b, offset, amp)
Feed operations run after all the others (the gap is carved last, as in
MainUnit.cpp:636-660) and each becomes a port on the carved cells.

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
    ops: list | None = None       # device set-up operations (mat is None then)
    dipole: tuple | None = None   # dipole_scene: (k0, k1, g0, gap) of the wire column and feed gap

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def has_diel(self) -> bool:
        """Any dielectric cell (a device-built scene: any dielectric object)."""
        if self.mat is not None:
            return bool((self.mat >= MAT_DIEL0).any())
        return any(op[0] in ("fill", "cells_box", "mesh", "wire") and op[-1 if op[0] != "mesh"
                   else 4] >= MAT_DIEL0 for op in self.ops)

    def cell(self, i, j, k) -> int:
        """VoxelGridSpec::cellIndex, Geometry.h:91-94."""
        return (k * self.ny + j) * self.nx + i

    def coords(self, cell):
        i = cell % self.nx
        j = (cell // self.nx) % self.ny
        k = cell // (self.nx * self.ny)
        return i, j, k

    def diel_fraction(self) -> float:
        if self.mat is None:
            raise ValueError("device-built scene: the material map lives on the device")
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


def add_plane_wave(scene: Scene, prop_axis: int, plane_index: int, pol_axis: int, amp: float):
    """Soft sources on the AIR cells of the plane co[prop_axis] = plane_index,
    in cell order, as TlmSolver::addPlaneWave (TlmSolver.cpp:394-407)."""
    idx = np.arange(scene.ncell, dtype=np.int64)
    co = (idx % scene.nx, (idx // scene.nx) % scene.ny, idx // (scene.nx * scene.ny))
    cells = idx[(co[prop_axis] == plane_index) & (scene.mat == MAT_AIR)]
    scene.sources.extend((int(c), int(pol_axis), float(amp)) for c in cells)
    return len(cells)


def box_list(nx, ny, nz, nboxes, rng, side=(16, 96), first_id=MAT_DIEL0):
    """`nboxes` seeded axis-aligned cell boxes [(lo, hi, id)] (later boxes
    overwrite), ids first_id .. first_id+nboxes-1, with (epsr, sigma)."""
    epsr = rng.uniform(1.0, 10.0, nboxes).astype(np.float32)
    sigma = rng.uniform(0.0, 0.05, nboxes).astype(np.float32)
    dims = (nx, ny, nz)
    boxes = []
    for b in range(nboxes):
        lo, hi = [], []
        for a in range(3):
            sd = int(rng.integers(side[0], side[1] + 1))
            o = int(rng.integers(0, dims[a]))
            lo.append(o)
            hi.append(min(dims[a], o + sd))
        boxes.append((tuple(lo), tuple(hi), first_id + b))
    return boxes, epsr, sigma


def stamp_box(mat, nx, ny, nz, lo, hi, mat_id):
    """Cells lo <= (i,j,k) < hi of a host map set to mat_id."""
    m = mat.reshape(nz, ny, nx)
    m[max(0, lo[2]):hi[2], max(0, lo[1]):hi[1], max(0, lo[0]):hi[0]] = mat_id


def random_boxes(mat, nx, ny, nz, nboxes, rng, side=(16, 96), first_id=MAT_DIEL0):
    """Stamp `nboxes` seeded axis-aligned boxes (later boxes overwrite),
    material ids first_id.. first_id+nboxes-1. Returns (epsr, sigma)."""
    boxes, epsr, sigma = box_list(nx, ny, nz, nboxes, rng, side, first_id)
    for lo, hi, mid in boxes:
        stamp_box(mat, nx, ny, nz, lo, hi, mid)
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
                 gap=1, device_setup=False) -> Scene:
    """Centre-fed z wire dipole: a one-cell PEC column of length
    length_frac*nz through the centre, with a `gap`-cell AIR feed gap at the
    centre fed as a port (addPort(gap, z, 1), SelfTest.cpp:154-176).
    device_setup: the same scene as set-up operations for the device (the
    boxes and the column as cell boxes, the gap as a feed carve)."""
    rng = np.random.default_rng(seed)
    epsr = np.zeros(0, np.float32)
    sigma = np.zeros(0, np.float32)
    boxes = []
    if background_boxes:
        side = (max(2, min(nx, ny, nz) // 16), max(3, (3 * min(nx, ny, nz)) // 8))
        boxes, epsr, sigma = box_list(nx, ny, nz, background_boxes, rng, side=side)
    ci, cj, ck = nx // 2, ny // 2, nz // 2
    half = max(gap + 1, int(round(length_frac * nz / 2)))
    k0 = max(1, ck - half)
    k1 = min(nz - 2, ck + half)
    g0 = ck - (gap - 1) // 2
    s = Scene(nx, ny, nz, None, epsr, sigma, total_steps=steps, name=f"dipole_{nx}x{ny}x{nz}",
              dipole=(k0, k1, g0, gap))
    if device_setup:
        if gap > 6:
            raise ValueError("a feed carve reaches every gap cell for gaps of up to 6 cells")
        ops = [("cells_box", lo, hi, mid) for lo, hi, mid in boxes]
        ops.append(("cells_box", (ci, cj, k0), (ci + 1, cj + 1, k1 + 1), MAT_PEC))
        # a feed segment over the gap cells' centres: its inner 70% samples
        # exactly those cells (z polarised)
        dl = np.float32(s.dl)
        za = (np.float32(g0) + np.float32(0.5)) * dl - dl * np.float32(0.5)
        zb = (np.float32(g0 + gap - 1) + np.float32(0.5)) * dl + dl * np.float32(0.5)
        x = (np.float32(ci) + np.float32(0.5)) * dl
        y = (np.float32(cj) + np.float32(0.5)) * dl
        ops.append(("feed", (x, y, za), (x, y, zb), (0.0, 0.0, 0.0), 1.0))
        s.ops = ops
        return s
    mat = np.zeros(nx * ny * nz, np.uint8)
    for lo, hi, mid in boxes:
        stamp_box(mat, nx, ny, nz, lo, hi, mid)
    m = mat.reshape(nz, ny, nx)
    m[k0:k1 + 1, cj, ci] = MAT_PEC
    gap_cells = []
    for k in range(g0, g0 + gap):
        m[k, cj, ci] = MAT_AIR
        gap_cells.append(s.cell(ci, cj, k))
    s.mat = mat
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
    V/I recorded. No surface-current or energy monitor (batched members
    record probes and ports)."""
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
        sc.energy_every = 0
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


# --- objects in the reference's form (triangle meshes, wires, feeds) --------

def box_mesh(lo, hi):
    """TriMesh::addBox (Geometry.cpp:135-147) as 8 shared vertices and 12
    triangles: the -z, +z, -y, +y, -x, +x quads, each split (a,b,c), (a,c,d)
    as addQuad does (Geometry.cpp:121-125)."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    v = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [lo[0], hi[1], lo[2]],
                  [hi[0], hi[1], lo[2]], [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]],
                  [lo[0], hi[1], hi[2]], [hi[0], hi[1], hi[2]]], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    idx = []
    for a, b, c, d in quads:
        idx += [(a, b, c), (a, c, d)]
    return v, np.array(idx, np.int32)


def plate_mesh(c, u, v):
    """TriMesh::addPlate (Geometry.cpp:128-132): the quad c -+ u/2 -+ v/2."""
    c, u, v = (np.asarray(x, np.float32) for x in (c, u, v))
    hu, hv = u * np.float32(0.5), v * np.float32(0.5)
    verts = np.array([c - hu - hv, c + hu - hv, c + hu + hv, c - hu + hv], np.float32)
    return verts, np.array([(0, 1, 2), (0, 2, 3)], np.int32)


def icosphere(radius, subdiv=3, center=(0.0, 0.0, 0.0)):
    """Sphere mesh: an icosahedron subdivided `subdiv` times (20*4^subdiv
    triangles), vertices pushed onto the sphere; float32 output."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts) * float(radius) + np.asarray(center, np.float64)
    return v.astype(np.float32), np.array(faces, np.int32)


def rotation(ax, ay, az):
    cx, sx, cy, sy, cz, sz = (math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay),
                              math.cos(az), math.sin(az))
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def mesh_scene(n=96, steps=600, subdiv=3, wire_radius_cells=0.3, huy_offset=4, nprobes=4,
               seed=11, pulse="f32") -> Scene:
    """Objects as the reference's scene holds them, voxelised on the device
    in TMainForm::voxelizeScene order (MainUnit.cpp:431-459): dielectric
    solids first (a sphere, an axis-aligned box, a rotated box), then the PEC
    objects as surfaces and wires (a reflector plate, a centre-fed z wire
    dipole of radius wire_radius_cells cells, a bent wire 0.8 cells thick),
    then the dipole's feed gap (MainUnit.cpp:636-660). An n^3 grid at 20
    cells per wavelength, origin at the grid corner. (A dipole radius of half
    a cell or more stamps PEC round the carved gap cell too, which then
    radiates nothing: the reference carves only the sampled cells.)"""
    rng = np.random.default_rng(seed)
    dl = np.float32(DL)
    L = float(n) * float(dl)
    f32 = lambda *x: tuple(float(np.float32(v)) for v in x)   # noqa: E731
    ops = []
    epsr = np.array([4.0, 2.5, 6.0], np.float32)
    sigma = np.array([0.01, 0.0, 0.02], np.float32)
    sv, si = icosphere(0.16 * L, subdiv)
    ops.append(("mesh", sv, si, f32(0.30 * L, 0.62 * L, 0.55 * L), MAT_DIEL0, True))
    bv, bi = box_mesh(f32(0.55 * L, 0.20 * L, 0.15 * L), f32(0.72 * L, 0.45 * L, 0.40 * L))
    ops.append(("mesh", bv, bi, (0.0, 0.0, 0.0), MAT_DIEL0 + 1, True))
    rv, ri = box_mesh(f32(-0.08 * L, -0.05 * L, -0.10 * L), f32(0.08 * L, 0.05 * L, 0.10 * L))
    rv = (rv.astype(np.float64) @ rotation(0.4, -0.7, 0.3).T).astype(np.float32)
    ops.append(("mesh", rv, ri, f32(0.70 * L, 0.70 * L, 0.72 * L), MAT_DIEL0 + 2, True))
    pv, pi = plate_mesh(f32(0.86 * L, 0.5 * L, 0.5 * L), f32(0, 0.5 * L, 0), f32(0, 0, 0.6 * L))
    ops.append(("mesh", pv, pi, (0.0, 0.0, 0.0), MAT_PEC, False))
    cx = float(np.float32((n // 2 + 0.5) * float(dl)))
    half = 0.22 * L
    zc = float(np.float32((n // 2 + 0.5) * float(dl)))
    ops.append(("wire", np.array([[cx, cx, zc - half], [cx, cx, zc + half]], np.float32),
                float(np.float32(wire_radius_cells * float(dl))), (0.0, 0.0, 0.0), MAT_PEC))
    bent = np.array([[0.15 * L, 0.15 * L, 0.2 * L], [0.25 * L, 0.2 * L, 0.45 * L],
                     [0.2 * L, 0.35 * L, 0.7 * L]], np.float32)
    ops.append(("wire", bent, float(np.float32(0.8 * float(dl))), (0.0, 0.0, 0.0), MAT_PEC))
    g = float(np.float32(1.2 * float(dl)))
    ops.append(("feed", (cx, cx, zc - g / 2), (cx, cx, zc + g / 2), (0.0, 0.0, 0.0), 1.0))
    s = Scene(n, n, n, None, epsr, sigma, dl=float(dl), total_steps=steps, huy_offset=huy_offset,
              pulse=pulse, name=f"mesh_scene_{n}")
    s.ops = ops
    probes = []
    while len(probes) < nprobes:
        i, j, k = (int(x) for x in rng.integers(n // 8, n - n // 8, 3))
        probes.append(s.cell(i, j, k))
    s.probes = np.array(probes, np.int64)
    return s
