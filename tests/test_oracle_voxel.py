"""Pins of the oracle's voxel set-up (oracle/voxel_oracle.cpp, restating
Geometry.cpp:306-642 and MainUnit.cpp:636-660) by what the geometry fixes:
closed forms for box solids, box surfaces, wire tubes and feed gaps, a
linear-programming decision of triangle/box overlap, analytic ray hits, and
inside/outside bounds for spheres and rotated boxes. Every case keeps its
geometry a margin away from the ties that FP32 rounding could decide either
way. CPU only."""
import numpy as np
import pytest
from scipy.optimize import linprog

from emsolver_amd import scenes
from oracle import oracle_tlm as ot


def _grid(rng):
    dims = tuple(int(x) for x in rng.integers(10, 22, 3))
    origin = tuple(float(np.float32(x)) for x in rng.uniform(-1.0, 1.0, 3))
    dl = float(np.float32(rng.uniform(0.05, 0.2)))
    return dims, origin, dl


def _off_grid(rng, origin, dl, cells):
    """A coordinate strictly inside cell `cells` (0.2..0.8 of the way)."""
    return float(np.float32(origin + (cells + rng.uniform(0.2, 0.8)) * dl))


def _box_case(rng):
    dims, origin, dl = _grid(rng)
    lo, hi, ia, ib = [], [], [], []
    for a in range(3):
        c0 = int(rng.integers(-3, dims[a] - 2))
        c1 = c0 + int(rng.integers(1, dims[a]))
        lo.append(_off_grid(rng, origin[a], dl, c0))
        hi.append(_off_grid(rng, origin[a], dl, c1))
        ia.append(c0)
        ib.append(c1)
    return dims, origin, dl, lo, hi, ia, ib


def _diagonal_clear(dims, origin, dl, lo, hi, ia, ib):
    """No row centre within 1e-3 of a face diagonal of the -x / +x quads
    (a ray through a shared edge reports both triangles)."""
    ny, nz = dims[1], dims[2]
    y = origin[1] + (np.arange(ny) + 0.5) * dl
    z = origin[2] + (np.arange(nz) + 0.5) * dl
    fy = (y[:, None] - lo[1]) / (hi[1] - lo[1])
    fz = (z[None, :] - lo[2]) / (hi[2] - lo[2])
    return np.abs(fy - fz).min() > 1e-3


def _expected_box(dims, ia, ib, shell):
    nx, ny, nz = dims
    m = np.zeros((nz, ny, nx), bool)
    k = np.arange(nz)[:, None, None]
    j = np.arange(ny)[None, :, None]
    i = np.arange(nx)[None, None, :]
    inside = (i >= ia[0]) & (i <= ib[0]) & (j >= ia[1]) & (j <= ib[1]) & (k >= ia[2]) & \
        (k <= ib[2])
    m |= inside
    if shell:
        strict = (i > ia[0]) & (i < ib[0]) & (j > ia[1]) & (j < ib[1]) & (k > ia[2]) & \
            (k < ib[2])
        m &= ~strict
    return m.reshape(-1)


@pytest.mark.parametrize("solid", [True, False])
def test_box_mesh_closed_form(solid):
    """VoxelizeMeshSolid of an addBox mesh marks exactly the cells meeting
    the box; VoxelizeMesh exactly those meeting its surface."""
    rng = np.random.default_rng(3 + solid)
    done = 0
    while done < 25:
        dims, origin, dl, lo, hi, ia, ib = _box_case(rng)
        if not _diagonal_clear(dims, origin, dl, lo, hi, ia, ib):
            continue
        v, ix = ot.box_mesh(lo, hi)
        mat = np.zeros(int(np.prod(dims)), np.uint8)
        ot.voxelize_mesh(mat, dims, origin, dl, v, ix, (0, 0, 0), 7, solid)
        exp = _expected_box(dims, ia, ib, shell=not solid)
        assert np.array_equal(mat == 7, exp), (dims, origin, dl, lo, hi)
        assert set(np.unique(mat)) <= {0, 7}
        done += 1


def test_box_mesh_offset_equals_translated_mesh():
    """The object position is added to every vertex (Geometry.cpp:374-376)."""
    rng = np.random.default_rng(5)
    dims, origin, dl, lo, hi, ia, ib = _box_case(rng)
    shift = np.float32([0.5 * dl, -1.25 * dl, 2.0 * dl])
    v, ix = ot.box_mesh(np.float32(lo) - shift, np.float32(hi) - shift)
    a = np.zeros(int(np.prod(dims)), np.uint8)
    ot.voxelize_mesh(a, dims, origin, dl, v, ix, shift, 3, True)
    v2 = (v + shift).astype(np.float32)
    b = np.zeros_like(a)
    ot.voxelize_mesh(b, dims, origin, dl, v2, ix, (0, 0, 0), 3, True)
    assert np.array_equal(a, b)


def test_box_mesh_layout_matches_scene_builder():
    """scenes.box_mesh (product) and the oracle's addBox agree."""
    v0, i0 = ot.box_mesh((0.1, -0.2, 0.3), (0.7, 0.4, 1.1))
    v1, i1 = scenes.box_mesh((0.1, -0.2, 0.3), (0.7, 0.4, 1.1))
    assert np.array_equal(v0, v1) and np.array_equal(i0, i1)


def _lp_margin(center, half, tri):
    """max t such that a point of the triangle lies inside the box shrunk
    by t (negative: the triangle stays that far outside)."""
    A, B, Cc = (np.asarray(p, np.float64) for p in tri)
    Au, Av = B - A, Cc - A
    rows, rhs = [], []
    for q in range(3):
        # A + u Au + v Av <= c + h - t  ->  u Au + v Av + t <= c + h - A
        rows.append([Au[q], Av[q], 1.0])
        rhs.append(center[q] + half[q] - A[q])
        # -(A + u Au + v Av) <= -(c - h) - t
        rows.append([-Au[q], -Av[q], 1.0])
        rhs.append(-(center[q] - half[q]) + A[q])
    rows.append([1.0, 1.0, 0.0])
    rhs.append(1.0)
    r = linprog([0, 0, -1.0], A_ub=rows, b_ub=rhs, bounds=[(0, 1), (0, 1), (None, None)],
                method="highs")
    assert r.status == 0
    return -r.fun


def test_tri_box_overlap_matches_linear_program():
    rng = np.random.default_rng(11)
    agree = {True: 0, False: 0}
    for _ in range(1500):
        center = rng.uniform(-0.3, 0.3, 3).astype(np.float32)
        half = rng.uniform(0.1, 0.6, 3).astype(np.float32)
        tri = rng.uniform(-1.5, 1.5, (3, 3)).astype(np.float32)
        m = _lp_margin(center.astype(np.float64), half.astype(np.float64), tri)
        if abs(m) < 1e-3:
            continue
        got = ot.tri_box_overlap(center, half, tri)
        assert got == (m > 0), (center, half, tri, m)
        agree[got] += 1
    assert agree[True] > 200 and agree[False] > 200


def test_tri_box_overlap_special_cases():
    h = (0.5, 0.5, 0.5)
    # a triangle inside the box, one containing the box's cross-section, one far away
    assert ot.tri_box_overlap((0, 0, 0), h, [[-0.1, -0.1, 0], [0.1, -0.1, 0], [0, 0.1, 0.1]])
    assert ot.tri_box_overlap((0, 0, 0), h, [[-5, -5, 0.2], [5, -5, 0.2], [0, 5, 0.2]])
    assert not ot.tri_box_overlap((0, 0, 0), h, [[3, 3, 3], [4, 3, 3], [3, 4, 3]])
    # the plane x+y+z = s cuts the corner (0.5,0.5,0.5) off for s < 1.5 and
    # misses the box beyond (the triangle's bounding box overlaps either way)
    assert ot.tri_box_overlap((0, 0, 0), h, [[1.4, 0, 0], [0, 1.4, 0], [0, 0, 1.4]])
    assert not ot.tri_box_overlap((0, 0, 0), h, [[1.6, 0, 0], [0, 1.6, 0], [0, 0, 1.6]])


def test_ray_x_tri_matches_analytic_hit():
    rng = np.random.default_rng(12)
    hits = misses = 0
    for _ in range(3000):
        tri = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        o = np.float32([rng.uniform(-3, -1.5), rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8)])
        A, B, Cc = tri.astype(np.float64)
        # o + s e_x = A + u (B - A) + v (C - A)
        M = np.stack([np.array([1.0, 0, 0]), A - B, A - Cc], axis=1)
        if abs(np.linalg.det(M)) < 1e-3:
            continue
        s, u, v = np.linalg.solve(M, A - o.astype(np.float64))
        margin = min(u, v, 1 - u - v)
        if abs(margin) < 1e-4 or abs(s) < 1e-4:
            continue
        hx = ot.ray_x_tri(o, tri)
        if margin > 0 and s > 0:
            assert hx is not None and abs(hx - (o[0] + s)) < 1e-5 * (1 + abs(o[0]) + s)
            hits += 1
        else:
            assert hx is None
            misses += 1
    assert hits > 200 and misses > 200
    # a triangle containing the x direction is never hit (det = 0)
    assert ot.ray_x_tri((-1, 0, 0), [[0, -1, -1], [1, -1, -1], [0, 1, 1]]) is None


def test_sphere_solid_bounds():
    """VoxelizeMeshSolid of an icosphere: every cell whose centre is well
    inside the polyhedron is set, none whose box lies wholly outside the
    circumscribed sphere, and the cells on the surface are set."""
    n, dl = 40, 0.05
    origin = (-1.0, -1.0, -1.0)
    R = 0.7
    cen = np.float32([0.03, -0.02, 0.05])
    v, ix = scenes.icosphere(R, 3)
    mat = np.zeros(n ** 3, np.uint8)
    ot.voxelize_mesh(mat, (n, n, n), origin, dl, v, ix, cen, 4, True)
    c = (np.arange(n) + 0.5) * dl - 1.0
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt((X - cen[0]) ** 2 + (Y - cen[1]) ** 2 + (Z - cen[2]) ** 2).reshape(-1)
    m = mat == 4
    assert m[r < 0.97 * R].all()
    assert not m[r > R + dl * np.sqrt(3) / 2 + 1e-4].any()
    # the surface pass marks the cells that the sphere's surface crosses
    assert m[np.abs(r - R) < 0.2 * dl].mean() > 0.9


def test_rotated_box_solid_bounds():
    rng = np.random.default_rng(21)
    n, dl = 36, 0.05
    origin = (-0.9, -0.9, -0.9)
    half = np.array([0.35, 0.2, 0.5])
    Rm = scenes.rotation(0.4, -0.7, 0.3)
    v, ix = scenes.box_mesh(-half, half)
    v = (v.astype(np.float64) @ Rm.T).astype(np.float32)
    cen = np.float32(rng.uniform(-0.05, 0.05, 3))
    mat = np.zeros(n ** 3, np.uint8)
    ot.voxelize_mesh(mat, (n, n, n), origin, dl, v, ix, cen, 9, True)
    c = (np.arange(n) + 0.5) * dl - 0.9
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    P = np.stack([X, Y, Z], -1).reshape(-1, 3) - cen.astype(np.float64)
    local = P @ Rm          # box frame
    inside = (np.abs(local) < half - 0.01).all(axis=1)
    d_out = np.linalg.norm(np.maximum(np.abs(local) - half, 0), axis=1)
    m = mat == 9
    assert m[inside].all()
    assert not m[d_out > dl * np.sqrt(3) / 2 + 1e-4].any()


@pytest.mark.parametrize("r_cells,T", [(0.0, None), (0.3, None), (0.8, 1.3), (1.7, 2.2)])
def test_wire_tube_closed_form(r_cells, T):
    """A z wire between two cell centres: radius below half a cell marks the
    column; larger radii mark the cells within r + dl/2 of the nearest
    sample (every cell centre on the axis is a sample), with caps."""
    nx, ny, nz = 15, 14, 24
    dl = 0.1
    origin = (-0.3, 0.2, -1.0)
    ic, jc, ka, kb = 7, 6, 4, 18
    x = np.float32(origin[0] + (ic + 0.5) * dl)
    y = np.float32(origin[1] + (jc + 0.5) * dl)
    za = np.float32(origin[2] + (ka + 0.5) * dl)
    zb = np.float32(origin[2] + (kb + 0.5) * dl)
    pts = np.float32([[x, y, za], [x, y, zb]])
    mat = np.zeros(nx * ny * nz, np.uint8)
    ot.voxelize_wire(mat, (nx, ny, nz), origin, dl, pts, r_cells * dl, (0, 0, 0), 1)
    exp = np.zeros((nz, ny, nx), bool)
    if T is None:
        exp[ka:kb + 1, jc, ic] = True
    else:
        rc = int(np.ceil(r_cells - 0.5))
        for di in range(-rc, rc + 1):
            for dj in range(-rc, rc + 1):
                rho2 = di * di + dj * dj
                if rho2 <= T * T:
                    exp[ka:kb + 1, jc + dj, ic + di] = True
                for m in range(1, rc + 1):
                    if rho2 + m * m <= T * T:
                        exp[ka - m, jc + dj, ic + di] = True
                        exp[kb + m, jc + dj, ic + di] = True
    assert np.array_equal(mat.reshape(nz, ny, nx) == 1, exp)


def test_thin_diagonal_wire_is_a_connected_chain_near_the_segment():
    nx = ny = nz = 20
    dl = 0.1
    origin = (0.0, 0.0, 0.0)
    a = np.float32([0.23, 0.31, 0.17])
    b = np.float32([1.71, 1.13, 1.64])
    mat = np.zeros(nx * ny * nz, np.uint8)
    ot.voxelize_wire(mat, (nx, ny, nz), origin, dl, np.stack([a, b]), 0.0, (0, 0, 0), 1)
    ks, js, is_ = np.nonzero(mat.reshape(nz, ny, nx))
    cells = set(zip(is_, js, ks))
    assert (int(a[0] / dl), int(a[1] / dl), int(a[2] / dl)) in cells
    assert (int(b[0] / dl), int(b[1] / dl), int(b[2] / dl)) in cells
    # every cell within a cell diagonal of the segment
    ab = (b - a).astype(np.float64)
    for i, j, k in cells:
        p = (np.array([i, j, k]) + 0.5) * dl
        t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0, 1)
        assert np.linalg.norm(p - (a + t * ab)) <= dl * np.sqrt(3) / 2 + 1e-6
    # 26-connected
    seen, todo = set(), [next(iter(cells))]
    while todo:
        c = todo.pop()
        if c in seen:
            continue
        seen.add(c)
        for d in np.ndindex(3, 3, 3):
            nb = (c[0] + d[0] - 1, c[1] + d[1] - 1, c[2] + d[2] - 1)
            if nb in cells and nb not in seen:
                todo.append(nb)
    assert seen == cells


def test_feed_carve_closed_form():
    nx, ny, nz = 12, 12, 16
    dl = 0.1
    origin = (0.0, 0.0, 0.0)
    mat = np.ones(nx * ny * nz, np.uint8)
    # a z gap across cells 6..8 (faces at 0.6 and 0.9): inner 70% is [0.645, 0.855]
    cells, pol = ot.carve_feed(mat, (nx, ny, nz), origin, dl, (0.55, 0.45, 0.6),
                               (0.55, 0.45, 0.9))
    exp = [(k * ny + 4) * nx + 5 for k in (6, 7, 8)]
    assert pol == 2 and list(cells) == exp
    assert (mat[exp] == 0).all() and (np.delete(mat, exp) == 1).all()
    # dominant axis y, offset applied
    mat = np.ones(nx * ny * nz, np.uint8)
    cells, pol = ot.carve_feed(mat, (nx, ny, nz), origin, dl, (0.0, 0.0, 0.0),
                               (0.02, 0.31, -0.01), offset=(0.55, 0.2, 0.45))
    assert pol == 1
    js = sorted({(c // nx) % ny for c in cells})
    assert js == [2, 3, 4] and all(c % nx == 5 for c in cells)
