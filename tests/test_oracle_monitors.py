"""Pins for the oracle's monitors (TlmSolver.cpp:616-790) from their
physical definitions rather than the code: probe E = -V_node/dl, port
current from Ampere's loop through the four transverse neighbours, surface
current from the two tangential H circulations, Huygens E/H phasors,
energy = sum V^2, running DFTs = post-hoc DFTs of the recorded series, and
the face lists from a brute-force scan of the material map.
"""
import numpy as np
import pytest

from emsolver_amd import scenes
from oracle import oracle_tlm as ot
from tests import helpers
from tests.test_oracle_scatter import EAXIS, circulation, node_voltage

DL = scenes.DL
POL = [[1, 2], [0, 2], [0, 1]]


def _open(nx, ny, nz, mat, steps=40, settle=0, huy=0, fp64=True, rho=0.0, **kw):
    cfg = ot.OracleConfig(nx, ny, nz, dl=DL, total_steps=steps, settle_steps=settle,
                          huy_offset=huy, fp64=fp64, rho=rho, waveform=ot.WF_GAUSS_DERIV, **kw)
    return ot.OracleTlm(cfg, mat)


def _states(o, nsteps):
    """State before each of the next nsteps steps (no sources => post-injection)."""
    out = []
    for _ in range(nsteps):
        out.append(o.state()[0].reshape(-1, 12).copy())
        o.run_steps(1)
    return out


def test_probe_is_minus_node_voltage_over_dl():
    rng = np.random.default_rng(1)
    n = 7
    mat = np.zeros(n ** 3, np.uint8)
    o = _open(n, n, n, mat, fp64=False)
    cells = [3, 100, 200]
    for c in cells:
        o.add_probe_cell(c)
    V = rng.standard_normal(12 * n ** 3).astype(np.float32)
    o.set_state(V)
    o.run_steps(1)
    v = V.reshape(-1, 12)
    inv = np.float32(-1.0) / np.float32(DL)
    for i, c in enumerate(cells):
        e = o.probe_series(i)[0]
        for j in range(3):
            idx = [p for p in range(12) if EAXIS[p] == j]     # increasing port order
            s = np.float32(0)
            s = (np.float32(v[c, idx[0]]) + np.float32(v[c, idx[1]]))
            s = np.float32(s + v[c, idx[2]])
            s = np.float32(s + v[c, idx[3]])
            want = np.float32(inv * np.float32(np.float32(0.5) * s))
            assert np.float32(e[j]) == want
            assert abs(e[j] - (-node_voltage(v[c].astype(np.float64), j) / DL)) < 1e-5 * (1 + abs(e[j]))


def test_port_voltage_and_ampere_current():
    rng = np.random.default_rng(2)
    nx, ny, nz = 9, 8, 10
    mat = np.zeros(nx * ny * nz, np.uint8)
    o = _open(nx, ny, nz, mat, steps=3)
    cell = lambda i, j, k: (k * ny + j) * nx + i
    gap = [cell(4, 4, 4), cell(4, 4, 5), cell(4, 4, 6)]
    o.add_port(gap, 2, 1.0)                  # z port
    V = rng.standard_normal(12 * nx * ny * nz)
    o.set_state(V)
    sv0 = ot.waveform_value(ot.WF_GAUSS_DERIV, 1e9, o.dt, 0)
    o.run_steps(1)
    v, i = o.port_records(0)
    st = V.reshape(-1, 12).copy()
    q = 0.5 * 1.0 * np.float64(np.float32(sv0))
    for c in gap:                            # port cells were injected before the monitor
        for a in (0, 1):
            for s in (0, 1):
                st[c, a * 4 + s * 2 + POL[a].index(2)] += q
    want_v = sum(node_voltage(st[c], 2) for c in gap)
    assert abs(v[0] - want_v) < 1e-12 * (1 + abs(want_v))
    # Ampere: loop of side 2 dl through the four x/y neighbours of the middle
    # cell, counter-clockwise seen from +z; H_k = -ZI_k/(eta0 dl)
    m = gap[1]
    px, mx, py, my = m + 1, m - 1, m + nx, m - nx
    loop_h = (-circulation(st[px], 1) + circulation(st[py], 0)
              + circulation(st[mx], 1) - circulation(st[my], 0))   # sum of H_t*(eta0 dl)
    want_i = 2.0 * loop_h / 376.730313
    assert abs(i[0] - want_i) < 1e-7 * (1 + abs(want_i))


def test_js_and_huygens_accumulators():
    rng = np.random.default_rng(3)
    nx, ny, nz = 14, 13, 12
    m = np.zeros((nz, ny, nx), np.uint8)
    m[5:8, 6, 3:9] = 1                        # a PEC plate
    mat = m.reshape(-1)
    settle, steps = 3, 9
    o = _open(nx, ny, nz, mat, steps=steps, settle=settle, huy=2)
    V = rng.standard_normal(12 * mat.size)
    V.reshape(-1, 12)[mat == 1] = 0.0
    o.set_state(V)
    states = _states(o, steps)
    dt = o.dt
    fcell, fax, fsign, _ = o.faces()
    inst, re, im, mag, ns = o.js()
    assert ns == steps - settle
    hre = np.zeros((fcell.size, 2))
    him = np.zeros((fcell.size, 2))
    for n, st in enumerate(states):
        ph = np.float32(2.0 * np.float32(np.pi)) * np.float32(1e9) * np.float32(n * np.float32(dt))
        cph, sph = float(np.cos(np.float32(ph))), float(np.sin(np.float32(ph)))
        for f in range(fcell.size):
            h = [circulation(st[fcell[f]], POL[fax[f]][t]) for t in range(2)]
            if n == steps - 1:
                assert abs(inst[f] - np.hypot(*h)) < 1e-9 * (1 + inst[f])
            if n >= settle:
                for t in range(2):
                    hre[f, t] += h[t] * cph
                    him[f, t] -= h[t] * sph
    assert np.allclose(re, hre.reshape(-1), atol=1e-6)
    assert np.allclose(im, him.reshape(-1), atol=1e-6)
    o.finalize()
    _, _, _, mag, _ = o.js()
    want = 2.0 / ns * np.sqrt((hre ** 2).sum(axis=1) + (him ** 2).sum(axis=1))
    assert np.allclose(mag, want, rtol=1e-6, atol=1e-9)
    # Huygens: E_t = -V_t, eta0 H_t = -ZI_t on each box face
    hc, hax, _ = o.huygens_faces()
    d, hs = o.huygens()
    d = d.reshape(-1, 8)
    acc = np.zeros_like(d)
    for n, st in enumerate(states):
        if n < settle:
            continue
        ph = np.float32(2.0 * np.float32(np.pi)) * np.float32(1e9) * np.float32(n * np.float32(dt))
        cph, sph = float(np.cos(np.float32(ph))), float(np.sin(np.float32(ph)))
        for f in range(hc.size):
            t1, t2 = POL[hax[f]]
            q = [-node_voltage(st[hc[f]], t1), -node_voltage(st[hc[f]], t2),
                 -circulation(st[hc[f]], t1), -circulation(st[hc[f]], t2)]
            for u in range(4):
                acc[f, 2 * u] += q[u] * cph
                acc[f, 2 * u + 1] -= q[u] * sph
    assert hs == steps - settle and hc.size > 0
    assert np.allclose(d, acc, atol=1e-6)


def test_energy_history_is_sum_of_squares():
    sc = scenes.free_space_cube(n=12, steps=50, probe_offset=3)
    o = helpers.oracle_open(sc, fp64=True)
    want = []
    for n in range(50):
        if n % 16 == 0:
            # post-injection state: inject by hand from the waveform
            V = o.state()[0].reshape(-1, 12).copy()
            sv = ot.waveform_value(sc.waveform, sc.f0, o.dt, n)
            c = sc.sources[0][0]
            for a in (0, 1):
                for s in (0, 1):
                    V[c, a * 4 + s * 2 + POL[a].index(2)] += 0.5 * np.float64(np.float32(sv))
            want.append((V ** 2).sum())
        o.run_steps(1)
    st, vals, d = o.energy()
    assert list(st) == [0, 16, 32, 48]
    assert np.allclose(d, want, rtol=1e-13, atol=0)
    assert np.array_equal(vals, np.asarray(d, np.float32))


def test_probe_dft_equals_posthoc_dft():
    sc = scenes.dielectric_boxes(n=16, nboxes=3, steps=300, nprobes=3, nfreq=5, seed=4)
    o = helpers.oracle_open(sc)
    o.run_steps(300)
    dt = np.float64(o.dt)
    n = np.arange(300)
    for p in range(3):
        x = o.probe_series(p)
        re, im = o.probe_dft(p)
        for q, f in enumerate(sc.dft_freqs):
            ph = 2 * np.pi * f * (n * dt)
            for j in range(3):
                assert abs(re[j, q] - np.sum(x[:, j] * np.cos(ph))) <= 1e-9 * np.abs(x[:, j]).sum()
                assert abs(im[j, q] + np.sum(x[:, j] * np.sin(ph))) <= 1e-9 * np.abs(x[:, j]).sum()


def test_field_plane_magnitude():
    rng = np.random.default_rng(5)
    nx, ny, nz = 6, 5, 7
    m = np.zeros((nz, ny, nx), np.uint8)
    m[3, 2, 1] = 1
    o = _open(nx, ny, nz, m.reshape(-1), steps=3)
    o.set_field_plane(2, 3)
    V = rng.standard_normal(12 * m.size)
    V.reshape(-1, 12)[m.reshape(-1) == 1] = 0.0
    o.set_state(V)
    o.run_steps(1)
    pl = o.plane()
    v = V.reshape(nz, ny, nx, 12)[3]
    want = np.sqrt(sum(np.vectorize(lambda a: a)(np.array([[node_voltage(v[j, i], q) for i in range(nx)]
                                                          for j in range(ny)])) ** 2 for q in range(3)))
    want[2, 1] = 0.0
    assert pl.shape == (ny, nx)
    assert np.allclose(pl, want, atol=1e-12)


def _brute_faces(nx, ny, nz, mat):
    m = mat.reshape(nz, ny, nx)
    out = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if m[k, j, i] != 1:
                    continue
                for a, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    for s in (-1, 1):
                        ii, jj, kk = i + s * di, j + s * dj, k + s * dk
                        if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz and m[kk, jj, ii] != 1:
                            out.append(((kk * ny + jj) * nx + ii, a, s))
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_face_lists(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = 11, 9, 10
    mat = np.zeros(nx * ny * nz, np.uint8)
    mat[rng.random(mat.size) < 0.2] = 1
    mat[(rng.random(mat.size) < 0.2) & (mat == 0)] = 2
    o = ot.OracleTlm(ot.OracleConfig(nx, ny, nz, huy_offset=2), mat, [3.0], [0.0])
    c, a, s, _ = o.faces()
    assert list(zip(c.tolist(), a.tolist(), s.tolist())) == _brute_faces(nx, ny, nz, mat)
    hc, hax, hsg = o.huygens_faces()
    # Huygens box: inset 2, AIR cells only, x faces, then y, then z
    box = []
    lo, hi = 2, (nx - 3, ny - 3, nz - 3)
    cell = lambda i, j, k: (k * ny + j) * nx + i
    for k in range(lo, hi[2] + 1):
        for j in range(lo, hi[1] + 1):
            box += [(cell(lo, j, k), 0, -1), (cell(hi[0], j, k), 0, 1)]
    for k in range(lo, hi[2] + 1):
        for i in range(lo, hi[0] + 1):
            box += [(cell(i, lo, k), 1, -1), (cell(i, hi[1], k), 1, 1)]
    for j in range(lo, hi[1] + 1):
        for i in range(lo, hi[0] + 1):
            box += [(cell(i, j, lo), 2, -1), (cell(i, j, hi[2]), 2, 1)]
    box = [b for b in box if mat[b[0]] == 0]
    assert list(zip(hc.tolist(), hax.tolist(), hsg.tolist())) == box
