"""ctypes wrapper of oracle/liboracle_tlm.so (the CPU restatement).

Test infrastructure only (see oracle/__init__.py). The restated algorithm
follows TlmSolver.cpp of the reference; the C++ source cites file:line for
every function.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_tlm.so")

WF_CW_RAMPED, WF_GAUSS_SINE, WF_GAUSS_PULSE, WF_GAUSS_DERIV = 0, 1, 2, 3


class _Cfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dl", C.c_float), ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float),
        ("f0", C.c_float), ("waveform", C.c_int32),
        ("total_steps", C.c_int32), ("settle_steps", C.c_int32),
        ("rho", C.c_float), ("huy_offset", C.c_int32), ("record_every", C.c_int32),
        ("gd_width", C.c_float), ("gd_delay", C.c_float),
        ("threads", C.c_int32), ("fp64", C.c_int32), ("pulse_format", C.c_int32),
        ("fp32_halo", C.c_int32),
    ]


@dataclasses.dataclass
class OracleConfig:
    nx: int
    ny: int
    nz: int
    dl: float = 0.0149896229
    origin: tuple = (0.0, 0.0, 0.0)
    f0: float = 1e9
    waveform: int = WF_CW_RAMPED
    total_steps: int = 100
    settle_steps: int = 0
    rho: float = 0.0
    huy_offset: int = 0
    record_every: int = 0
    gd_width: float = 20.0
    gd_delay: float = 60.0
    threads: int = 0
    fp64: bool = False
    pulse_format: int = 0  # 0 FP32, 1 FP16, 2 BF16 storage (float oracle only)
    fp32_halo: int = 8     # 16-bit formats: slabs within this distance of a source stay FP32


_lib = None


def build_library(force: bool = False) -> str:
    """Compile liboracle_tlm.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build_library()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    i64p = C.POINTER(C.c_int64)
    i32p = C.POINTER(C.c_int32)
    fp = C.POINTER(C.c_float)
    sig = {
        "ot_default_threads": (C.c_int, []),
        "ot_waveform_value": (C.c_float, [C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, C.c_float]),
        "ot_create": (P, [C.POINTER(_Cfg), C.POINTER(C.c_uint8), fp, fp, C.c_int]),
        "ot_destroy": (None, [P]),
        "ot_add_port": (C.c_int, [P, i64p, C.c_int, C.c_int, C.c_float]),
        "ot_add_plane_wave": (None, [P, C.c_int, C.c_int, C.c_int, C.c_float]),
        "ot_add_source": (None, [P, C.c_int64, C.c_int, C.c_float]),
        "ot_add_probe": (C.c_int, [P, C.c_float, C.c_float, C.c_float]),
        "ot_add_probe_cell": (C.c_int, [P, C.c_int64]),
        "ot_set_probe_dft": (None, [P, dp, C.c_int]),
        "ot_set_field_plane": (None, [P, C.c_int, C.c_int]),
        "ot_set_state": (None, [P, dp, dp]),
        "ot_run": (None, [P]),
        "ot_run_steps": (None, [P, C.c_int]),
        "ot_scatter_connect": (None, [P]),
        "ot_finalize": (None, [P]),
        "ot_cur_step": (C.c_int, [P]),
        "ot_has_stubs": (C.c_int, [P]),
        "ot_dt": (C.c_float, [P]),
        "ot_get_state": (None, [P, dp, dp]),
        "ot_num_faces": (C.c_int, [P]),
        "ot_get_faces": (None, [P, i64p, i32p, i32p, fp]),
        "ot_num_huygens": (C.c_int, [P]),
        "ot_get_huygens_faces": (None, [P, i64p, i32p, i32p]),
        "ot_get_js": (None, [P, dp, dp, dp, dp, C.POINTER(C.c_int)]),
        "ot_get_huygens": (None, [P, dp, C.POINTER(C.c_int)]),
        "ot_num_ports": (C.c_int, [P]),
        "ot_port_cells": (C.c_int, [P, C.c_int, i64p]),
        "ot_port_records": (C.c_int, [P, C.c_int, dp, dp]),
        "ot_num_sources": (C.c_int, [P]),
        "ot_get_sources": (None, [P, i64p, i32p, fp]),
        "ot_probe_cell": (C.c_int, [P, C.c_int]),
        "ot_probe_series": (C.c_int, [P, C.c_int, dp]),
        "ot_probe_dft": (None, [P, C.c_int, dp, dp]),
        "ot_energy": (C.c_int, [P, i32p, fp, dp]),
        "ot_plane": (C.c_int, [P, dp, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "ot_num_frames": (C.c_int, [P]),
        "ot_frame": (C.c_int, [P, C.c_int, dp, dp]),
        "ot_scatter_cell": (None, [dp, dp, C.c_double, C.c_double, C.c_int, C.c_int, dp]),
        "ot_slab_exponents": (C.c_int, [P, i32p, fp]),
        "ot_f16_rne": (C.c_uint16, [C.c_float]),
        "ot_f16_to_f32": (C.c_float, [C.c_uint16]),
        "ot_bf16_rne": (C.c_uint16, [C.c_float]),
        "ot_plan_exponent": (C.c_int, [C.c_float]),
        "ot_encode": (None, [C.c_int, fp, C.c_float, C.POINTER(C.c_uint16), C.c_int64]),
        "ot_scatter_unitarity_error": (C.c_float, []),
        "ot_diel_scatter_unitarity_error": (C.c_float, [C.c_float]),
        "ot_voxelize_mesh": (None, [C.c_int, C.c_int, C.c_int, fp, C.c_float, fp, C.c_int, i32p,
                                    C.c_int, fp, C.c_uint8, C.c_int, C.POINTER(C.c_uint8)]),
        "ot_voxelize_wire": (None, [C.c_int, C.c_int, C.c_int, fp, C.c_float, fp, C.c_int,
                                    C.c_float, fp, C.c_uint8, C.POINTER(C.c_uint8)]),
        "ot_tri_box_overlap": (C.c_int, [fp, fp, fp]),
        "ot_ray_x_tri": (C.c_int, [fp, fp, fp]),
        "ot_box_mesh": (None, [fp, fp, fp, i32p]),
        "ot_carve_feed": (C.c_int, [C.c_int, C.c_int, C.c_int, fp, C.c_float, fp, fp, fp,
                                    C.POINTER(C.c_uint8), i64p, C.c_int, i32p]),
        "ot_far_field": (C.c_int, [C.c_int, fp, i32p, i32p, fp, C.c_int, C.c_float, C.c_int,
                                   C.c_int, C.c_int, fp, fp]),
        "ot_huygens_geometry": (None, [C.c_int, C.c_int, C.c_int, fp, C.c_float, C.c_int, i64p,
                                       i32p, fp, i32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class OracleTlm:
    """One CPU TLM solver instance (TlmSolver semantics)."""

    def __init__(self, cfg: OracleConfig, mat: np.ndarray, epsr=(), sigma=()):
        self.lib = load_library()
        self.cfg = cfg
        n = cfg.nx * cfg.ny * cfg.nz
        mat = np.ascontiguousarray(mat, dtype=np.uint8).reshape(-1)
        assert mat.size == n, "material array must hold nx*ny*nz entries"
        self.mat = mat
        self.n = n
        epsr = np.ascontiguousarray(epsr, dtype=np.float32).reshape(-1)
        sigma = np.ascontiguousarray(sigma, dtype=np.float32).reshape(-1)
        assert epsr.size == sigma.size
        c = _Cfg(cfg.nx, cfg.ny, cfg.nz, cfg.dl, *cfg.origin, cfg.f0, cfg.waveform,
                 cfg.total_steps, cfg.settle_steps, cfg.rho, cfg.huy_offset,
                 cfg.record_every, cfg.gd_width, cfg.gd_delay, cfg.threads, int(cfg.fp64),
                 int(cfg.pulse_format), int(cfg.fp32_halo))
        self.h = self.lib.ot_create(C.byref(c), _ptr(mat, C.c_uint8), _ptr(epsr, C.c_float),
                                    _ptr(sigma, C.c_float), int(epsr.size))

    def close(self):
        if getattr(self, "h", None):
            self.lib.ot_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- scene ----
    def add_port(self, cells, pol=2, amp=1.0) -> int:
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        return self.lib.ot_add_port(self.h, _ptr(cells, C.c_int64), int(cells.size), pol, amp)

    def add_plane_wave(self, prop_axis, plane_index, pol_axis, amp=1.0):
        self.lib.ot_add_plane_wave(self.h, prop_axis, plane_index, pol_axis, amp)

    def add_source(self, cell, pol, amp=1.0):
        self.lib.ot_add_source(self.h, int(cell), pol, amp)

    def add_probe(self, x, y, z) -> int:
        return self.lib.ot_add_probe(self.h, x, y, z)

    def add_probe_cell(self, cell) -> int:
        return self.lib.ot_add_probe_cell(self.h, int(cell))

    def set_probe_dft(self, freqs):
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        self.lib.ot_set_probe_dft(self.h, _ptr(f, C.c_double), int(f.size))
        self.nfreq = int(f.size)

    def set_field_plane(self, axis, index):
        self.lib.ot_set_field_plane(self.h, axis, index)

    def set_state(self, V, Vs=None):
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        assert V.size == 12 * self.n
        Vsa = None
        if Vs is not None and self.has_stubs:
            Vsa = np.ascontiguousarray(Vs, dtype=np.float64).reshape(-1)
            assert Vsa.size == 3 * self.n
        self.lib.ot_set_state(self.h, _ptr(V, C.c_double), _ptr(Vsa, C.c_double))

    # ---- run ----
    def run(self):
        self.lib.ot_run(self.h)

    def run_steps(self, count):
        self.lib.ot_run_steps(self.h, int(count))

    def scatter_connect(self):
        self.lib.ot_scatter_connect(self.h)

    def finalize(self):
        self.lib.ot_finalize(self.h)

    # ---- results ----
    @property
    def cur_step(self):
        return self.lib.ot_cur_step(self.h)

    @property
    def has_stubs(self):
        return bool(self.lib.ot_has_stubs(self.h))

    @property
    def dt(self):
        return self.lib.ot_dt(self.h)

    def state(self):
        V = np.empty(12 * self.n, np.float64)
        Vs = np.empty(3 * self.n, np.float64) if self.has_stubs else None
        self.lib.ot_get_state(self.h, _ptr(V, C.c_double), _ptr(Vs, C.c_double))
        return V, Vs

    def faces(self):
        nf = self.lib.ot_num_faces(self.h)
        cell = np.empty(nf, np.int64)
        axis = np.empty(nf, np.int32)
        sign = np.empty(nf, np.int32)
        cen = np.empty((nf, 3), np.float32)
        self.lib.ot_get_faces(self.h, _ptr(cell, C.c_int64), _ptr(axis, C.c_int32),
                              _ptr(sign, C.c_int32), _ptr(cen, C.c_float))
        return cell, axis, sign, cen

    def huygens_faces(self):
        nh = self.lib.ot_num_huygens(self.h)
        cell = np.empty(nh, np.int64)
        axis = np.empty(nh, np.int32)
        sgn = np.empty(nh, np.int32)
        self.lib.ot_get_huygens_faces(self.h, _ptr(cell, C.c_int64), _ptr(axis, C.c_int32),
                                      _ptr(sgn, C.c_int32))
        return cell, axis, sgn

    def js(self):
        nf = self.lib.ot_num_faces(self.h)
        inst = np.empty(nf)
        re = np.empty(2 * nf)
        im = np.empty(2 * nf)
        mag = np.empty(nf)
        ns = C.c_int(0)
        self.lib.ot_get_js(self.h, _ptr(inst, C.c_double), _ptr(re, C.c_double),
                           _ptr(im, C.c_double), _ptr(mag, C.c_double), C.byref(ns))
        return inst, re, im, mag, ns.value

    def huygens(self):
        nh = self.lib.ot_num_huygens(self.h)
        d = np.empty(8 * nh)
        ns = C.c_int(0)
        self.lib.ot_get_huygens(self.h, _ptr(d, C.c_double), C.byref(ns))
        return d, ns.value

    def port_cells(self, p):
        n = self.lib.ot_port_cells(self.h, p, None)
        c = np.empty(n, np.int64)
        self.lib.ot_port_cells(self.h, p, _ptr(c, C.c_int64))
        return c

    def port_records(self, p):
        n = self.lib.ot_port_records(self.h, p, None, None)
        v = np.empty(n)
        i = np.empty(n)
        self.lib.ot_port_records(self.h, p, _ptr(v, C.c_double), _ptr(i, C.c_double))
        return v, i

    def sources(self):
        n = self.lib.ot_num_sources(self.h)
        cell = np.empty(n, np.int64)
        pol = np.empty(n, np.int32)
        amp = np.empty(n, np.float32)
        self.lib.ot_get_sources(self.h, _ptr(cell, C.c_int64), _ptr(pol, C.c_int32),
                                _ptr(amp, C.c_float))
        return cell, pol, amp

    def probe_cell(self, p):
        return self.lib.ot_probe_cell(self.h, p)

    def probe_series(self, p):
        n = self.lib.ot_probe_series(self.h, p, None)
        e = np.empty(3 * n)
        self.lib.ot_probe_series(self.h, p, _ptr(e, C.c_double))
        return e.reshape(n, 3)

    def probe_dft(self, p):
        nf = getattr(self, "nfreq", 0)
        re = np.empty(3 * nf)
        im = np.empty(3 * nf)
        self.lib.ot_probe_dft(self.h, p, _ptr(re, C.c_double), _ptr(im, C.c_double))
        return re.reshape(3, nf), im.reshape(3, nf)

    def energy(self):
        n = self.lib.ot_energy(self.h, None, None, None)
        st = np.empty(n, np.int32)
        v = np.empty(n, np.float32)
        d = np.empty(n, np.float64)
        self.lib.ot_energy(self.h, _ptr(st, C.c_int32), _ptr(v, C.c_float), _ptr(d, C.c_double))
        return st, v, d

    def plane(self):
        n1 = C.c_int(0)
        n2 = C.c_int(0)
        n = self.lib.ot_plane(self.h, None, C.byref(n1), C.byref(n2))
        b = np.empty(n)
        self.lib.ot_plane(self.h, _ptr(b, C.c_double), C.byref(n1), C.byref(n2))
        return b.reshape(n2.value, n1.value) if n else b

    def frames_full(self):
        """(steps, js [f][nface], planes [f][n2][n1] or None)."""
        nf = self.lib.ot_num_faces(self.h)
        cnt = self.lib.ot_num_frames(self.h)
        n1 = C.c_int(0)
        n2 = C.c_int(0)
        self.lib.ot_plane(self.h, None, C.byref(n1), C.byref(n2))
        steps = np.empty(cnt, np.int32)
        js = np.empty((cnt, nf))
        planes = np.empty((cnt, n2.value, n1.value)) if n1.value else None
        for i in range(cnt):
            pl = np.empty(n1.value * n2.value) if planes is not None else None
            steps[i] = self.lib.ot_frame(self.h, i, _ptr(js[i], C.c_double), _ptr(pl, C.c_double))
            if planes is not None:
                planes[i] = pl.reshape(n2.value, n1.value)
        return steps, js, planes

    def slab_exponents(self):
        """(E_n per slab, M_{n-1} per slab) of the 16-bit storage mode."""
        E = np.zeros(self.cfg.nz, np.int32)
        M = np.zeros(self.cfg.nz, np.float32)
        n = self.lib.ot_slab_exponents(self.h, _ptr(E, C.c_int32), _ptr(M, C.c_float))
        return (E, M) if n else (None, None)

    def frames(self):
        out = []
        nf = self.lib.ot_num_faces(self.h)
        for i in range(self.lib.ot_num_frames(self.h)):
            js = np.empty(nf)
            step = self.lib.ot_frame(self.h, i, _ptr(js, C.c_double), None)
            out.append((step, js))
        return out


def waveform_value(wtype, f0, dt, n, gd_width=20.0, gd_delay=60.0) -> float:
    return load_library().ot_waveform_value(wtype, f0, dt, n, gd_width, gd_delay)


def scatter_cell(vin, vs=None, yh=0.0, gh=0.0, diel=False, fp64=False):
    lib = load_library()
    vin = np.ascontiguousarray(vin, dtype=np.float64)
    vout = np.empty(12)
    vsa = np.ascontiguousarray(vs if vs is not None else np.zeros(3), dtype=np.float64).copy()
    lib.ot_scatter_cell(_ptr(vin, C.c_double), _ptr(vsa, C.c_double), yh, gh, int(diel),
                        int(fp64), _ptr(vout, C.c_double))
    return vout, vsa


def scatter_unitarity_error() -> float:
    return load_library().ot_scatter_unitarity_error()


def diel_scatter_unitarity_error(epsr: float) -> float:
    return load_library().ot_diel_scatter_unitarity_error(epsr)


def encode_pulses(fmt: int, values, inv_scale: float = 1.0) -> np.ndarray:
    """16-bit codes of values * inv_scale (fmt 1 binary16, 2 bfloat16), rounded
    to nearest even by the oracle's software conversion."""
    lib = load_library()
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros(v.size, np.uint16)
    lib.ot_encode(int(fmt), _ptr(v, C.c_float), C.c_float(inv_scale), _ptr(out, C.c_uint16), v.size)
    return out.reshape(v.shape)


def plan_exponent(bound: float) -> int:
    return int(load_library().ot_plan_exponent(C.c_float(bound)))


# --- voxel set-up (Geometry.cpp:306-642) -----------------------------------

def _f3(v):
    return np.ascontiguousarray(np.asarray(v, np.float32).reshape(3))


def voxelize_mesh(mat, dims, origin, dl, verts, idx, offset, mat_id, solid):
    """VoxelizeMesh / VoxelizeMeshSolid into `mat` (uint8, cellIndex order,
    updated in place)."""
    nx, ny, nz = dims
    assert mat.dtype == np.uint8 and mat.size == nx * ny * nz and mat.flags.c_contiguous
    v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
    ix = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
    o, off = _f3(origin), _f3(offset)
    load_library().ot_voxelize_mesh(nx, ny, nz, _ptr(o, C.c_float), C.c_float(dl),
                                    _ptr(v, C.c_float), len(v), _ptr(ix, C.c_int32), len(ix),
                                    _ptr(off, C.c_float), int(mat_id), int(bool(solid)),
                                    _ptr(mat, C.c_uint8))
    return mat


def voxelize_wire(mat, dims, origin, dl, pts, radius, offset, mat_id):
    """VoxelizeWire into `mat` (updated in place)."""
    nx, ny, nz = dims
    assert mat.dtype == np.uint8 and mat.size == nx * ny * nz and mat.flags.c_contiguous
    p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    o, off = _f3(origin), _f3(offset)
    load_library().ot_voxelize_wire(nx, ny, nz, _ptr(o, C.c_float), C.c_float(dl),
                                    _ptr(p, C.c_float), len(p), C.c_float(radius),
                                    _ptr(off, C.c_float), int(mat_id), _ptr(mat, C.c_uint8))
    return mat


def tri_box_overlap(center, half, tri) -> bool:
    c, h = _f3(center), _f3(half)
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    return bool(load_library().ot_tri_box_overlap(_ptr(c, C.c_float), _ptr(h, C.c_float),
                                                  _ptr(t, C.c_float)))


def ray_x_tri(o, tri):
    """+x ray from o against a triangle: the hit x or None."""
    oo = _f3(o)
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    hx = C.c_float(0)
    ok = load_library().ot_ray_x_tri(_ptr(oo, C.c_float), _ptr(t, C.c_float), C.byref(hx))
    return float(hx.value) if ok else None


def box_mesh(lo, hi):
    """TriMesh::addBox as 8 shared vertices and 12 triangles."""
    v = np.zeros((8, 3), np.float32)
    ix = np.zeros((12, 3), np.int32)
    load_library().ot_box_mesh(_ptr(_f3(lo), C.c_float), _ptr(_f3(hi), C.c_float),
                               _ptr(v, C.c_float), _ptr(ix, C.c_int32))
    return v, ix


def carve_feed(mat, dims, origin, dl, a, b, offset=(0, 0, 0)):
    """Feed-gap carve (MainUnit.cpp:636-660) into `mat`; returns (cells, pol)."""
    nx, ny, nz = dims
    cap = 4096
    cells = np.zeros(cap, np.int64)
    pol = C.c_int32(0)
    o, aa, bb, off = _f3(origin), _f3(a), _f3(b), _f3(offset)
    n = load_library().ot_carve_feed(nx, ny, nz, _ptr(o, C.c_float), C.c_float(dl),
                                     _ptr(aa, C.c_float), _ptr(bb, C.c_float),
                                     _ptr(off, C.c_float), _ptr(mat, C.c_uint8),
                                     _ptr(cells, C.c_int64), cap, C.byref(pol))
    return cells[:min(n, cap)].copy(), int(pol.value)


# --- near-to-far-field transform (TlmSolver.cpp:805-916) --------------------

def far_field(pos, axis, sign, dft8, f0, n_theta=37, n_phi=73, samples=1, fp64=False):
    """ComputeFarFieldFromHuygens: (U[n_theta, n_phi], uMax, directivity,
    peak theta deg, peak phi deg), or None without data."""
    p = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    a = np.ascontiguousarray(axis, np.int32).reshape(-1)
    s = np.ascontiguousarray(sign, np.int32).reshape(-1)
    d = np.ascontiguousarray(dft8, np.float32).reshape(-1, 8)
    U = np.zeros(n_theta * n_phi, np.float32)
    summ = np.zeros(4, np.float32)
    rc = load_library().ot_far_field(len(p), _ptr(p, C.c_float), _ptr(a, C.c_int32),
                                     _ptr(s, C.c_int32), _ptr(d, C.c_float), int(samples),
                                     C.c_float(f0), n_theta, n_phi, int(bool(fp64)),
                                     _ptr(U, C.c_float), _ptr(summ, C.c_float))
    if rc != 0:
        return None
    return U.reshape(n_theta, n_phi), float(summ[0]), float(summ[1]), float(summ[2]), \
        float(summ[3])


def huygens_geometry(dims, origin, dl, cells, axis):
    """HuyFace pos and outSign of Huygens cells (TlmSolver.cpp:323-334)."""
    nx, ny, nz = dims
    c = np.ascontiguousarray(cells, np.int64)
    a = np.ascontiguousarray(axis, np.int32)
    pos = np.zeros((c.size, 3), np.float32)
    sg = np.zeros(c.size, np.int32)
    load_library().ot_huygens_geometry(nx, ny, nz, _ptr(_f3(origin), C.c_float), C.c_float(dl),
                                       c.size, _ptr(c, C.c_int64), _ptr(a, C.c_int32),
                                       _ptr(pos, C.c_float), _ptr(sg, C.c_int32))
    return pos, sg
