"""Python handle on the gfx950 TLM step (ctypes over include/GpuTlmHip.h).

``TlmDevice`` wraps one ``tlm_ctx``: upload, per-step tables, monitors, run,
readback. ``run_scene`` drives a whole :class:`emsolver_amd.scenes.Scene`
the way ``TlmSolver::run`` drives ``RunGpuTlm`` (TlmSolver.cpp:439-454): it
builds the injection list (sources, then the AIR cells of every port,
TlmSolver.cpp:362-374 and 501-505), the PEC face list, the per-step
waveform and phase tables, runs every step on the device and reads the
monitors back.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _native
from ._native import ptr

C0 = np.float32(299792458.0)      # TlmSolver.cpp:15
ETA0 = np.float32(376.730313)     # TlmSolver.cpp:16

PULSE_TYPES = {"f32": 0, "f16": 1, "bf16": 2}


# return codes of the C ABI that callers act on
ERR_RANGE = -7   # TLM_ERR_RANGE: 16-bit storage outgrew a slab scale


class TlmError(RuntimeError):
    """A failed C ABI call; `code` is its TLM_ERR_* return code (None where
    the error was found on the Python side)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def device_count() -> int:
    return _native.load().tlm_device_count()


def device_name(dev: int = 0) -> str:
    return _native.load().tlm_device_name(dev).decode()


def scatter_unitarity_error(epsr: float = 1.0) -> float:
    """ScatterUnitarityError / DielScatterUnitarityError on the device scatter."""
    return float(_native.load().tlm_scatter_unitarity_error(float(epsr)))


def timestep(dl: float) -> np.float32:
    """dt = dl / (2 c0) in float, TlmSolver.cpp:268."""
    return np.float32(dl) / (np.float32(2.0) * C0)


def stub_admittances(epsr, sigma, dl):
    """yhat = 4(max(1,epsr)-1), ghat = max(0,sigma)*dl*eta0 in float,
    TlmSolver.cpp:272-276."""
    epsr = np.asarray(epsr, np.float32)
    sigma = np.asarray(sigma, np.float32)
    yh = np.float32(4.0) * (np.maximum(np.float32(1.0), epsr) - np.float32(1.0))
    gh = (np.maximum(np.float32(0.0), sigma) * np.float32(dl)) * ETA0
    return yh.astype(np.float32), gh.astype(np.float32)


def schedule_tables(waveform: int, f0: float, dt: float, nsteps: int, gd_width=20.0,
                    gd_delay=60.0):
    lib = _native.load()
    sv = np.empty(nsteps, np.float32)
    cph = np.empty(nsteps, np.float32)
    sph = np.empty(nsteps, np.float32)
    lib.tlm_schedule_tables(waveform, f0, float(dt), 0, nsteps, gd_width, gd_delay,
                            ptr(sv, C.c_float), ptr(cph, C.c_float), ptr(sph, C.c_float))
    return sv, cph, sph


def dft_tables(freqs, dt: float, nsteps: int):
    lib = _native.load()
    f = np.ascontiguousarray(freqs, np.float64)
    ct = np.empty((nsteps, f.size))
    st = np.empty((nsteps, f.size))
    lib.tlm_dft_tables(ptr(f, C.c_double), f.size, float(dt), nsteps, ptr(ct, C.c_double),
                       ptr(st, C.c_double))
    return ct, st


def pec_faces(nx, ny, nz, mat):
    """BuildPecSurfaceFaces order (Geometry.cpp:402-436): air cell, axis, sign."""
    lib = _native.load()
    mat = np.ascontiguousarray(mat, np.uint8).reshape(-1)
    n = lib.tlm_build_pec_faces(nx, ny, nz, ptr(mat, C.c_uint8), None, None, None, 0)
    cells = np.empty(n, np.int64)
    axes = np.empty(n, np.int32)
    signs = np.empty(n, np.int32)
    lib.tlm_build_pec_faces(nx, ny, nz, ptr(mat, C.c_uint8), ptr(cells, C.c_int64),
                            ptr(axes, C.c_int32), ptr(signs, C.c_int32), n)
    return cells, axes, signs


def huygens_faces(nx, ny, nz, mat, offset):
    lib = _native.load()
    mat = np.ascontiguousarray(mat, np.uint8).reshape(-1)
    n = lib.tlm_build_huygens(nx, ny, nz, ptr(mat, C.c_uint8), offset, None, None, None, 0)
    cells = np.empty(n, np.int64)
    axes = np.empty(n, np.int32)
    sg = np.empty(n, np.int32)
    lib.tlm_build_huygens(nx, ny, nz, ptr(mat, C.c_uint8), offset, ptr(cells, C.c_int64),
                          ptr(axes, C.c_int32), ptr(sg, C.c_int32), n)
    return cells, axes, sg


class TlmDevice:
    """One device context (``tlm_ctx``)."""

    def __init__(self, nx, ny, nz, dl, rho=0.0, pulse="f32", device=0, fp32_halo=8):
        """pulse: "f32", "f16" or "bf16" storage; fp32_halo: for the 16-bit
        formats, z-slabs on each side of a source slab kept in FP32 (< 0: none)."""
        self.lib = _native.load()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n = self.nx * self.ny * self.nz
        self.pulse = pulse
        halo = -1 if fp32_halo < 0 else int(fp32_halo)
        if halo == 0:
            raise ValueError("fp32_halo must be >= 1, or negative for none")
        d = _native.GridDesc(self.nx, self.ny, self.nz, float(dl), float(rho),
                             PULSE_TYPES[pulse], int(device), halo)
        h = C.c_void_p()
        rc = self.lib.tlm_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise TlmError(f"tlm_create failed ({rc}): {self.lib.tlm_last_error(None).decode()}")
        self.h = h
        self.nprobe = 0
        self.nport = 0
        self.nface = 0
        self.nhuy = 0
        self.nfreq = 0
        self.has_stubs = False

    def _chk(self, rc, what):
        if rc != 0:
            raise TlmError(f"{what} failed ({rc}): {self.lib.tlm_last_error(self.h).decode()}", rc)

    def close(self):
        if getattr(self, "h", None):
            self.lib.tlm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up ----
    def upload(self, mat, yhat=(), ghat=(), V=None, Vs=None):
        mat = np.ascontiguousarray(mat, np.uint8).reshape(-1)
        if mat.size != self.n:
            raise TlmError("material map must hold nx*ny*nz bytes")
        yh = np.ascontiguousarray(yhat, np.float32).reshape(-1)
        gh = np.ascontiguousarray(ghat, np.float32).reshape(-1)
        Va = None if V is None else np.ascontiguousarray(V, np.float32).reshape(-1)
        Vsa = None if Vs is None else np.ascontiguousarray(Vs, np.float32).reshape(-1)
        if Va is not None and Va.size != 12 * self.n:
            raise TlmError("V must hold 12 floats per cell")
        if Vsa is not None and Vsa.size != 3 * self.n:
            raise TlmError("Vs must hold 3 floats per cell")
        self._chk(self.lib.tlm_upload(self.h, ptr(mat, C.c_uint8), ptr(yh, C.c_float),
                                      ptr(gh, C.c_float), yh.size, ptr(Va, C.c_float),
                                      ptr(Vsa, C.c_float)), "tlm_upload")
        self.has_stubs = bool((mat >= 2).any())

    # ---- voxel set-up on the device (the context's material map) ----
    def _vgrid(self, origin, dl):
        g = _native.VoxelGrid()
        for q in range(3):
            g.origin[q] = float(np.float32(origin[q]))
        g.dl = float(np.float32(dl))
        return g

    @staticmethod
    def _f3(v):
        return np.ascontiguousarray(np.asarray(v, np.float32).reshape(3))

    def mat_fill(self, mat_id=0):
        self._chk(self.lib.tlm_mat_fill(self.h, int(mat_id)), "tlm_mat_fill")

    def mat_cells_box(self, lo, hi, mat_id):
        a = np.ascontiguousarray(lo, np.int32).reshape(3)
        b = np.ascontiguousarray(hi, np.int32).reshape(3)
        self._chk(self.lib.tlm_mat_cells_box(self.h, ptr(a, C.c_int32), ptr(b, C.c_int32),
                                             int(mat_id)), "tlm_mat_cells_box")

    def voxelize_mesh(self, origin, dl, verts, idx, offset, mat_id, solid):
        v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        ix = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
        off = self._f3(offset)
        g = self._vgrid(origin, dl)
        self._chk(self.lib.tlm_voxelize_mesh(self.h, C.byref(g), ptr(v, C.c_float), len(v),
                                             ptr(ix, C.c_int32), len(ix), ptr(off, C.c_float),
                                             int(mat_id), int(bool(solid))), "tlm_voxelize_mesh")

    def voxelize_wire(self, origin, dl, pts, radius, offset, mat_id):
        p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        off = self._f3(offset)
        g = self._vgrid(origin, dl)
        self._chk(self.lib.tlm_voxelize_wire(self.h, C.byref(g), ptr(p, C.c_float), len(p),
                                             float(np.float32(radius)), ptr(off, C.c_float),
                                             int(mat_id)), "tlm_voxelize_wire")

    def carve_feed(self, origin, dl, a, b, offset=(0, 0, 0)):
        """Returns (sorted unique cells carved to AIR, dominant axis)."""
        g = self._vgrid(origin, dl)
        aa, bb, off = self._f3(a), self._f3(b), self._f3(offset)
        pol = C.c_int32(0)
        cap = 4096
        cells = np.zeros(cap, np.int64)
        n = self.lib.tlm_carve_feed(self.h, C.byref(g), ptr(aa, C.c_float), ptr(bb, C.c_float),
                                    ptr(off, C.c_float), ptr(cells, C.c_int64), cap, C.byref(pol))
        if n < 0:
            self._chk(n, "tlm_carve_feed")
        if n > cap:
            raise TlmError("feed gap longer than 4096 cells")
        return cells[:n].copy(), int(pol.value)

    def mat_read(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint8)
        self._chk(self.lib.tlm_mat_read(self.h, ptr(out, C.c_uint8)), "tlm_mat_read")
        return out

    def mat_write(self, mat):
        m = np.ascontiguousarray(mat, np.uint8).reshape(-1)
        if m.size != self.n:
            raise TlmError("material map must hold nx*ny*nz bytes")
        self._chk(self.lib.tlm_mat_write(self.h, ptr(m, C.c_uint8)), "tlm_mat_write")

    def upload_mat(self, yhat=(), ghat=()):
        """tlm_upload of the device material map with a zero state."""
        yh = np.ascontiguousarray(yhat, np.float32).reshape(-1)
        gh = np.ascontiguousarray(ghat, np.float32).reshape(-1)
        self._chk(self.lib.tlm_upload_mat(self.h, ptr(yh, C.c_float), ptr(gh, C.c_float),
                                          yh.size), "tlm_upload_mat")

    def pec_faces_mat(self) -> int:
        n = int(self.lib.tlm_pec_faces_mat(self.h))
        if n < 0:
            self._chk(n, "tlm_pec_faces_mat")
        self.nface = n
        return n

    def huygens_mat(self, offset) -> int:
        n = int(self.lib.tlm_huygens_mat(self.h, int(offset)))
        if n < 0:
            self._chk(n, "tlm_huygens_mat")
        self.nhuy = n
        return n

    def _read_list(self, fn, name):
        n = int(fn(self.h, None, None, 0))
        if n < 0:
            self._chk(n, name)
        cells = np.zeros(n, np.int64)
        axes = np.zeros(n, np.int32)
        if n:
            fn(self.h, ptr(cells, C.c_int64), ptr(axes, C.c_int32), n)
        return cells, axes

    def read_faces(self):
        return self._read_list(self.lib.tlm_read_faces, "tlm_read_faces")

    def read_huygens_faces(self):
        return self._read_list(self.lib.tlm_read_huygens_faces, "tlm_read_huygens_faces")

    def set_sources(self, cells, pols, amps):
        c = np.ascontiguousarray(cells, np.int64)
        p = np.ascontiguousarray(pols, np.int32)
        a = np.ascontiguousarray(amps, np.float32)
        self._chk(self.lib.tlm_set_sources(self.h, ptr(c, C.c_int64), ptr(p, C.c_int32),
                                           ptr(a, C.c_float), c.size), "tlm_set_sources")

    def set_schedule(self, sv, cph, sph, settle):
        sv = np.ascontiguousarray(sv, np.float32)
        cph = np.ascontiguousarray(cph, np.float32)
        sph = np.ascontiguousarray(sph, np.float32)
        self._chk(self.lib.tlm_set_schedule(self.h, ptr(sv, C.c_float), ptr(cph, C.c_float),
                                            ptr(sph, C.c_float), sv.size, int(settle)),
                  "tlm_set_schedule")
        self.nsteps = sv.size

    def set_probes(self, cells, inv_dl):
        c = np.ascontiguousarray(cells, np.int64)
        self._chk(self.lib.tlm_set_probes(self.h, ptr(c, C.c_int64), c.size, float(inv_dl)),
                  "tlm_set_probes")
        self.nprobe = c.size

    def set_probe_dft(self, ctab, stab):
        ct = np.ascontiguousarray(ctab, np.float64)
        st = np.ascontiguousarray(stab, np.float64)
        nf = ct.shape[1] if ct.ndim == 2 else 0
        self._chk(self.lib.tlm_set_probe_dft(self.h, ptr(ct, C.c_double), ptr(st, C.c_double),
                                             nf), "tlm_set_probe_dft")
        self.nfreq = nf

    def set_ports(self, ports):
        """ports: list of (cells, pol_axis)."""
        cells = [np.asarray(c, np.int64).reshape(-1) for c, _ in ports]
        off = np.zeros(len(ports) + 1, np.int32)
        for i, c in enumerate(cells):
            off[i + 1] = off[i] + c.size
        flat = np.ascontiguousarray(np.concatenate(cells) if cells else np.zeros(0, np.int64))
        pol = np.ascontiguousarray([p for _, p in ports], np.int32)
        self._chk(self.lib.tlm_set_ports(self.h, ptr(flat, C.c_int64), ptr(off, C.c_int32),
                                         ptr(pol, C.c_int32), len(ports)), "tlm_set_ports")
        self.nport = len(ports)

    def set_faces(self, cells, axes):
        c = np.ascontiguousarray(cells, np.int64)
        a = np.ascontiguousarray(axes, np.int32)
        self._chk(self.lib.tlm_set_faces(self.h, ptr(c, C.c_int64), ptr(a, C.c_int32), c.size),
                  "tlm_set_faces")
        self.nface = c.size

    def set_huygens(self, cells, axes):
        c = np.ascontiguousarray(cells, np.int64)
        a = np.ascontiguousarray(axes, np.int32)
        self._chk(self.lib.tlm_set_huygens(self.h, ptr(c, C.c_int64), ptr(a, C.c_int32), c.size),
                  "tlm_set_huygens")
        self.nhuy = c.size

    def set_plane(self, axis, index):
        self._chk(self.lib.tlm_set_plane(self.h, int(axis), int(index)), "tlm_set_plane")
        self.plane_axis = int(axis)

    def set_frames(self, every, max_frames):
        self._chk(self.lib.tlm_set_frames(self.h, int(every), int(max_frames)), "tlm_set_frames")

    def read_plane(self):
        n1, n2, st = C.c_int(), C.c_int(), C.c_int()
        self._chk(self.lib.tlm_read_plane(self.h, None, C.byref(n1), C.byref(n2), C.byref(st)),
                  "tlm_read_plane")
        out = np.empty((n2.value, n1.value), np.float32)
        if out.size:
            self._chk(self.lib.tlm_read_plane(self.h, ptr(out, C.c_float), C.byref(n1),
                                              C.byref(n2), C.byref(st)), "tlm_read_plane")
        return out, st.value

    def read_frames(self):
        cnt = self.lib.tlm_read_frames(self.h, None, None, None, 0)
        if cnt < 0:
            self._chk(cnt, "tlm_read_frames")
        n1, n2, st = C.c_int(), C.c_int(), C.c_int()
        self.lib.tlm_read_plane(self.h, None, C.byref(n1), C.byref(n2), C.byref(st))
        steps = np.empty(cnt, np.int32)
        js = np.empty((cnt, self.nface), np.float32)
        planes = np.empty((cnt, n2.value, n1.value), np.float32) if n1.value else None
        self.lib.tlm_read_frames(self.h, ptr(steps, C.c_int32), ptr(js, C.c_float),
                                 ptr(planes, C.c_float), cnt)
        return steps, js, planes

    def set_energy_every(self, every):
        self._chk(self.lib.tlm_set_energy_every(self.h, int(every)), "tlm_set_energy_every")

    # ---- execution ----
    def run(self, step0, nsteps):
        self._chk(self.lib.tlm_run(self.h, int(step0), int(nsteps)), "tlm_run")

    def sync(self):
        self._chk(self.lib.tlm_sync(self.h), "tlm_sync")

    def time_steps(self, step0, nsteps, kernel_every=1):
        """Run and time steps [step0, step0+nsteps): (ms of the whole block,
        average ms of one fused step launch over the launches of every
        kernel_every-th step, or None with kernel_every=0)."""
        tot = C.c_float(0)
        ker = C.c_float(0)
        nk = C.c_int(0)
        k = int(kernel_every)
        self._chk(self.lib.tlm_time_steps(self.h, int(step0), int(nsteps), k, C.byref(tot),
                                          C.byref(ker) if k > 0 else None, C.byref(nk)),
                  "tlm_time_steps")
        return tot.value, (ker.value / nk.value if k > 0 and nk.value > 0 else None)

    def set_tuning(self, **changes):
        """Sets the tlm_tuning fields named (kernel, planes, block_waves,
        nontemporal, grid_blocks, fold, occupancy); the others keep their
        current values. Results are bit-identical for every setting."""
        cur = self.get_tuning()
        unknown = set(changes) - set(cur)
        if unknown:
            raise TypeError(f"set_tuning: unknown fields {sorted(unknown)}")
        cur.update(changes)
        t = _native.Tuning(*(int(cur[f]) for f, _ in _native.Tuning._fields_))
        self._chk(self.lib.tlm_set_tuning(self.h, C.byref(t)), "tlm_set_tuning")

    def get_tuning(self):
        t = _native.Tuning()
        self._chk(self.lib.tlm_get_tuning(self.h, C.byref(t)), "tlm_get_tuning")
        return {"kernel": t.kernel, "planes": t.planes, "block_waves": t.block_waves,
                "nontemporal": t.nontemporal, "grid_blocks": t.grid_blocks, "fold": t.fold,
                "occupancy": t.occupancy}

    def launch_shape(self):
        r, z, g, t = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._chk(self.lib.tlm_launch_shape(self.h, C.byref(r), C.byref(z), C.byref(g),
                                            C.byref(t)), "tlm_launch_shape")
        return {"strips_per_slab": r.value, "slabs_per_task": z.value, "groups": g.value,
                "threads_per_group": t.value}

    # ---- readback ----
    def read_state(self):
        V = np.empty(12 * self.n, np.float32)
        Vs = np.empty(3 * self.n, np.float32)
        self._chk(self.lib.tlm_read_state(self.h, ptr(V, C.c_float), ptr(Vs, C.c_float)),
                  "tlm_read_state")
        return V, Vs

    def read_scales(self):
        """16-bit formats: per z-slab (scale 2^E of the current step, largest
        |value| sent in the last step, storage format); None otherwise."""
        sc = np.zeros(self.nz, np.float32)
        mx = np.zeros(self.nz, np.float32)
        fm = np.zeros(self.nz, np.int32)
        n = self.lib.tlm_read_scales(self.h, ptr(sc, C.c_float), ptr(mx, C.c_float),
                                     ptr(fm, C.c_int32))
        if n < 0:
            self._chk(n, "tlm_read_scales")
        return (sc, mx, fm) if n > 0 else None

    def read_probes(self, step0, nsteps):
        e = np.empty((nsteps, self.nprobe, 3), np.float32)
        self._chk(self.lib.tlm_read_probes(self.h, step0, nsteps, ptr(e, C.c_float)),
                  "tlm_read_probes")
        return e

    def read_probe_dft(self):
        re = np.empty((self.nprobe, 3, self.nfreq))
        im = np.empty((self.nprobe, 3, self.nfreq))
        self._chk(self.lib.tlm_read_probe_dft(self.h, ptr(re, C.c_double), ptr(im, C.c_double)),
                  "tlm_read_probe_dft")
        return re, im

    def read_ports(self, step0, nsteps):
        vi = np.empty((nsteps, self.nport, 2), np.float32)
        self._chk(self.lib.tlm_read_ports(self.h, step0, nsteps, ptr(vi, C.c_float)),
                  "tlm_read_ports")
        return vi

    def read_js(self):
        inst = np.empty(self.nface, np.float32)
        d4 = np.empty((self.nface, 4), np.float32)
        self._chk(self.lib.tlm_read_js(self.h, ptr(inst, C.c_float), ptr(d4, C.c_float)),
                  "tlm_read_js")
        return inst, d4

    def read_huygens(self):
        d8 = np.empty((self.nhuy, 8), np.float32)
        self._chk(self.lib.tlm_read_huygens(self.h, ptr(d8, C.c_float)), "tlm_read_huygens")
        return d8

    def read_energy(self):
        n = self.lib.tlm_read_energy(self.h, None, None, 0)
        if n < 0:
            self._chk(n, "tlm_read_energy")
        st = np.empty(n, np.int32)
        v = np.empty(n, np.float64)
        self.lib.tlm_read_energy(self.h, ptr(st, C.c_int32), ptr(v, C.c_double), n)
        return st, v


@dataclasses.dataclass
class SceneResult:
    steps: int
    probes: np.ndarray | None = None      # [steps][nprobe][3]
    probe_dft: tuple | None = None        # (re, im) [nprobe][3][nfreq]
    ports: np.ndarray | None = None       # [steps][nport][2]
    js_inst: np.ndarray | None = None
    js_dft4: np.ndarray | None = None
    dft_samples: int = 0
    huygens: np.ndarray | None = None
    huy_samples: int = 0
    energy_steps: np.ndarray | None = None
    energy: np.ndarray | None = None
    V: np.ndarray | None = None
    Vs: np.ndarray | None = None
    plane: np.ndarray | None = None
    plane_step: int = -1
    frames: tuple | None = None            # (steps, js [f][nface], planes [f][n2][n1])
    kernel_ms: float | None = None
    total_ms: float | None = None


def apply_ops(dev: TlmDevice, scene) -> list:
    """Run a scene's set-up operations on the device material map (feed
    carves last); returns the feed ports [(cells, pol, amp)]."""
    org, dl = scene.origin, scene.dl
    feeds = []
    for op in scene.ops:
        kind = op[0]
        if kind == "fill":
            dev.mat_fill(op[1])
        elif kind == "cells_box":
            dev.mat_cells_box(op[1], op[2], op[3])
        elif kind == "mesh":
            dev.voxelize_mesh(org, dl, op[1], op[2], op[3], op[4], op[5])
        elif kind == "wire":
            dev.voxelize_wire(org, dl, op[1], op[2], op[3], op[4])
        elif kind != "feed":
            raise ValueError(f"unknown set-up operation {kind!r}")
    for op in scene.ops:
        if op[0] == "feed":
            cells, pol = dev.carve_feed(org, dl, op[1], op[2], op[3])
            if cells.size:
                feeds.append(([int(c) for c in cells], pol, float(op[4])))
    return feeds


def open_scene(scene, device=0) -> TlmDevice:
    """Create a context and load a Scene into it (no steps run). A scene of
    set-up operations (scene.ops) is voxelised on the device; a host map is
    uploaded."""
    dev = TlmDevice(scene.nx, scene.ny, scene.nz, scene.dl, scene.rho, pulse=scene.pulse,
                    device=device, fp32_halo=scene.fp32_halo)
    yh, gh = stub_admittances(scene.epsr, scene.sigma, scene.dl)
    if scene.ops is not None:
        if scene.ports or scene.V0 is not None:
            raise ValueError("a device-built scene takes its ports from feed operations and "
                             "starts from a zero state")
        ports = apply_ops(dev, scene)       # carved cells are AIR by construction
        dev.upload_mat(yh, gh)
        mat = None
    else:
        dev.upload(scene.mat, yh, gh, scene.V0, scene.Vs0)
        mat = scene.mat.reshape(-1)
        # addPort keeps the AIR cells only
        ports = [([int(c) for c in pc if 0 <= c < mat.size and mat[c] == 0], p, a)
                 for pc, p, a in scene.ports]
    dt = timestep(scene.dl)
    total = scene.total_steps
    sv, cph, sph = schedule_tables(scene.waveform, scene.f0, dt, total, scene.gd_width,
                                   scene.gd_delay)
    dev.set_schedule(sv, cph, sph, scene.settle_steps)
    dev.set_energy_every(scene.energy_every)
    # injection order: sources, then the cells of each port
    cells, pols, amps = [], [], []
    for c, p, a in scene.sources:
        cells.append(c)
        pols.append(p)
        amps.append(a)
    port_list = []
    for pc, p, a in ports:
        port_list.append((pc, p))
        for c in pc:
            cells.append(c)
            pols.append(p)
            amps.append(a)
    dev.set_sources(cells, pols, amps)
    if port_list:
        dev.set_ports(port_list)
    if scene.faces:
        if mat is None:
            dev.pec_faces_mat()
        else:
            fc, fa, _ = pec_faces(scene.nx, scene.ny, scene.nz, mat)
            if fc.size:
                dev.set_faces(fc, fa)
    if scene.huy_offset >= 2:
        if mat is None:
            dev.huygens_mat(scene.huy_offset)
        else:
            hc, ha, _ = huygens_faces(scene.nx, scene.ny, scene.nz, mat, scene.huy_offset)
            if hc.size:
                dev.set_huygens(hc, ha)
    if len(scene.probes):
        dev.set_probes(scene.probes, np.float32(-1.0) / np.float32(scene.dl))
        if len(scene.dft_freqs):
            ct, st = dft_tables(scene.dft_freqs, float(dt), total)
            dev.set_probe_dft(ct, st)
    if scene.plane_axis >= 0:
        dev.set_plane(scene.plane_axis, scene.plane_index)
    if scene.record_every > 0:
        dev.set_frames(scene.record_every, total // scene.record_every + 1)
    dev.dt = dt
    return dev


def collect(dev: TlmDevice, scene, steps: int, state=True) -> SceneResult:
    res = SceneResult(steps=steps)
    if dev.nprobe:
        res.probes = dev.read_probes(0, steps)
        if dev.nfreq:
            res.probe_dft = dev.read_probe_dft()
    if dev.nport:
        res.ports = dev.read_ports(0, steps)
    if dev.nface:
        res.js_inst, res.js_dft4 = dev.read_js()
    res.dft_samples = max(0, steps - scene.settle_steps)
    if dev.nhuy:
        res.huygens = dev.read_huygens()
        res.huy_samples = res.dft_samples
    res.energy_steps, res.energy = dev.read_energy()
    if scene.plane_axis >= 0:
        res.plane, res.plane_step = dev.read_plane()
    if scene.record_every > 0:
        res.frames = dev.read_frames()
    if state:
        res.V, res.Vs = dev.read_state()
    return res


def run_scene(scene, steps=None, device=0, state=True, block=256) -> SceneResult:
    """Run a Scene on the GPU from step 0 and return every monitor."""
    steps = scene.total_steps if steps is None else steps
    with open_scene(scene, device) as dev:
        n = 0
        while n < steps:
            m = min(block, steps - n)
            dev.run(n, m)
            n += m
        dev.sync()
        return collect(dev, scene, steps, state)


def device_encode(fmt, values, inv_scale=1.0, device=0):
    """16-bit codes of values * inv_scale rounded by the device conversion
    (fmt "f16" or "bf16")."""
    lib = _native.load()
    v = np.ascontiguousarray(values, np.float32).reshape(-1)
    out = np.zeros(v.size, np.uint16)
    rc = lib.tlm_device_encode(int(device), PULSE_TYPES[fmt], ptr(v, C.c_float),
                               C.c_float(inv_scale), ptr(out, C.c_uint16), v.size)
    if rc != 0:
        raise TlmError(f"tlm_device_encode failed ({rc})")
    return out


def device_memory(device=0):
    """(free, total) bytes of device memory."""
    lib = _native.load()
    f, t = C.c_size_t(0), C.c_size_t(0)
    rc = lib.tlm_device_memory(int(device), C.byref(f), C.byref(t))
    if rc != 0:
        raise TlmError(f"tlm_device_memory failed ({rc})")
    return f.value, t.value


def estimate_bytes(nx, ny, nz, pulse="f32", stubs=True) -> int:
    """Device bytes a context of this shape allocates (monitors excluded)."""
    lib = _native.load()
    d = _native.GridDesc(int(nx), int(ny), int(nz), 0.01, 0.0, PULSE_TYPES[pulse], 0, 0)
    return int(lib.tlm_estimate_bytes(C.byref(d), int(bool(stubs))))


class TlmBatch:
    """Many contexts advanced by the same launches (``tlm_batch``)."""

    def __init__(self, device=0):
        self.lib = _native.load()
        h = C.c_void_p()
        rc = self.lib.tlm_batch_create(int(device), C.byref(h))
        if rc != 0:
            raise TlmError(f"tlm_batch_create failed ({rc}): "
                           f"{self.lib.tlm_batch_last_error(None).decode()}")
        self.h = h
        self.members = {}

    def _chk(self, rc, what):
        if rc != 0:
            raise TlmError(f"{what} failed ({rc}): {self.lib.tlm_batch_last_error(self.h).decode()}",
                           rc)

    def add(self, dev: TlmDevice, nsteps: int):
        self._chk(self.lib.tlm_batch_add(self.h, dev.h, int(nsteps)), "tlm_batch_add")
        self.members[dev.h.value] = dev

    def remove(self, dev: TlmDevice):
        """Takes dev out of the batch (also when the call reports an error:
        TlmError with code ERR_RANGE means dev's own 16-bit results are
        invalid, the batch and its other members go on)."""
        rc = self.lib.tlm_batch_remove(self.h, dev.h)
        self.members.pop(dev.h.value, None)
        self._chk(rc, "tlm_batch_remove")

    def run(self, nticks: int):
        self._chk(self.lib.tlm_batch_run(self.h, int(nticks)), "tlm_batch_run")

    def sync(self):
        self._chk(self.lib.tlm_batch_sync(self.h), "tlm_batch_sync")

    def remaining(self):
        """{TlmDevice: steps still to run} as of the ticks queued."""
        n = self.lib.tlm_batch_members(self.h, None, None, 0)
        hs = (C.c_void_p * max(n, 1))()
        rem = np.zeros(max(n, 1), np.int32)
        self.lib.tlm_batch_members(self.h, hs, ptr(rem, C.c_int32), n)
        return {self.members[hs[i]]: int(rem[i]) for i in range(n)}

    def set_tuning(self, **changes):
        """Sets the tlm_batch_tuning fields named (planes, block_waves,
        grid_blocks, graph_ticks, planes16); the others keep their current
        values."""
        cur = self.get_tuning()
        unknown = set(changes) - set(cur)
        if unknown:
            raise TypeError(f"set_tuning: unknown fields {sorted(unknown)}")
        cur.update(changes)
        t = _native.BatchTuning(*(int(cur[f]) for f, _ in _native.BatchTuning._fields_))
        self._chk(self.lib.tlm_batch_set_tuning(self.h, C.byref(t)), "tlm_batch_set_tuning")

    def get_tuning(self) -> dict:
        t = _native.BatchTuning()
        self._chk(self.lib.tlm_batch_get_tuning(self.h, C.byref(t)), "tlm_batch_get_tuning")
        return {f: getattr(t, f) for f, _ in t._fields_}

    def time(self, nticks: int):
        """(ms, cell updates) of nticks ticks timed with HIP events."""
        ms = C.c_float(0)
        cu = C.c_int64(0)
        self._chk(self.lib.tlm_batch_time(self.h, int(nticks), C.byref(ms), C.byref(cu)),
                  "tlm_batch_time")
        return ms.value, cu.value

    def autotune(self, ticks_per_candidate=4) -> dict:
        t = _native.BatchTuning()
        self._chk(self.lib.tlm_batch_autotune(self.h, int(ticks_per_candidate), C.byref(t)),
                  "tlm_batch_autotune")
        return {f: getattr(t, f) for f, _ in t._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.tlm_batch_destroy(self.h)
            self.h = None
            self.members = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_scenes_batched(scene_list, device=0, tuning=None, autotune=0, block=64):
    """Run several scenes from step 0 in one batch; returns their SceneResults
    (probes, ports, final state) in order. Batch members record no energy, so
    the scenes run with energy_every = 0."""
    scene_list = [dataclasses.replace(sc, energy_every=0) for sc in scene_list]
    devs = [open_scene(sc, device) for sc in scene_list]
    with TlmBatch(device) as b:
        if tuning:
            b.set_tuning(**tuning)
        for d, sc in zip(devs, scene_list):
            b.add(d, sc.total_steps)
        left = max(sc.total_steps for sc in scene_list)
        if autotune:
            b.autotune(autotune)
            b.sync()
            left = max(b.remaining().values())
        while left > 0:
            m = min(block, left)
            b.run(m)
            left -= m
        b.sync()
        for d in devs:
            b.remove(d)
    out = [collect(d, sc, sc.total_steps) for d, sc in zip(devs, scene_list)]
    for d in devs:
        d.close()
    return out


@dataclasses.dataclass
class FarField:
    U: np.ndarray            # [n_theta, n_phi]
    u_max: float
    directivity: float
    peak_theta_deg: float
    peak_phi_deg: float


def far_field(pos, axis, sign, dft8, f0, n_theta=37, n_phi=73, samples=1, device=0) -> FarField:
    """ComputeFarFieldFromHuygens on the device from host face arrays."""
    lib = _native.load()
    p = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    a = np.ascontiguousarray(axis, np.int32).reshape(-1)
    s = np.ascontiguousarray(sign, np.int32).reshape(-1)
    d = np.ascontiguousarray(dft8, np.float32).reshape(-1, 8)
    U = np.zeros(n_theta * n_phi, np.float32)
    sm = _native.FarFieldSummary()
    rc = lib.tlm_far_field(int(device), len(p), ptr(p, C.c_float), ptr(a, C.c_int32),
                           ptr(s, C.c_int32), ptr(d, C.c_float), int(samples), float(f0),
                           n_theta, n_phi, ptr(U, C.c_float), C.byref(sm))
    if rc != 0:
        raise TlmError(f"tlm_far_field failed ({rc}): {lib.tlm_last_error(None).decode()}")
    return FarField(U.reshape(n_theta, n_phi), sm.u_max, sm.directivity, sm.peak_theta_deg,
                    sm.peak_phi_deg)


def far_field_ctx(dev: TlmDevice, origin, samples, f0, n_theta=37, n_phi=73,
                  noalloc=False) -> FarField:
    """The same from a context's resident Huygens accumulators; noalloc=True
    calls the variant whose workspace is the library's fixed device array
    (tlm_far_field_ctx_noalloc) instead of allocations of its own."""
    o = np.ascontiguousarray(np.asarray(origin, np.float32).reshape(3))
    U = np.zeros(n_theta * n_phi, np.float32)
    sm = _native.FarFieldSummary()
    name = "tlm_far_field_ctx_noalloc" if noalloc else "tlm_far_field_ctx"
    dev._chk(getattr(dev.lib, name)(dev.h, ptr(o, C.c_float), int(samples), float(f0),
                                    n_theta, n_phi, ptr(U, C.c_float), C.byref(sm)), name)
    return FarField(U.reshape(n_theta, n_phi), sm.u_max, sm.directivity, sm.peak_theta_deg,
                    sm.peak_phi_deg)
