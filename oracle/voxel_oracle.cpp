//---------------------------------------------------------------------------
// voxel_oracle.cpp - CPU restatement of the reference's voxel set-up: the
// step's input side (material map from geometry).
//
// TEST INFRASTRUCTURE ONLY: the checker of the device voxeliser. Only tests/
// load it (through liboracle_tlm.so).
//
// Restates, with the same FP32 operation order:
//   VoxelGridSpec cellIndex/cellCenter/cellOf  Geometry.h:85-113
//   Vec3 dot / cross / length, Aabb        Vec3.h:28-33, 54-60
//   AxisTest, TriBoxOverlap                Geometry.cpp:306-363
//   VoxelizeMesh (surface)                 Geometry.cpp:369-397
//   RayXTri                                Geometry.cpp:525-548
//   VoxelizeMeshSolid (parity fill)        Geometry.cpp:551-610
//   VoxelizeWire                           Geometry.cpp:608-642
//   TriMesh::addBox                        Geometry.cpp:135-147
//   feed-gap carve                         MainUnit.cpp:636-660
// Pinned in tests/test_oracle_voxel.py by closed forms (box solids, wire
// tubes), brute-force sampling (triangle/box overlap) and analytic ray hits.
//---------------------------------------------------------------------------
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct V3
{
    float x = 0, y = 0, z = 0;
    V3() {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    V3 operator+(const V3 &v) const { return V3(x + v.x, y + v.y, z + v.z); }
    V3 operator-(const V3 &v) const { return V3(x - v.x, y - v.y, z - v.z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    void set(int i, float v)
    {
        if (i == 0)
            x = v;
        else if (i == 1)
            y = v;
        else
            z = v;
    }
    float dot(const V3 &v) const { return x * v.x + y * v.y + z * v.z; }
    V3 cross(const V3 &v) const { return V3(y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x); }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
};

// Aabb (Vec3.h:54-60): grows from +-1e30 with the vmin / vmax selects.
struct Box
{
    V3 lo{1e30f, 1e30f, 1e30f}, hi{-1e30f, -1e30f, -1e30f};
    void grow(const V3 &p)
    {
        lo = V3(lo.x < p.x ? lo.x : p.x, lo.y < p.y ? lo.y : p.y, lo.z < p.z ? lo.z : p.z);
        hi = V3(hi.x > p.x ? hi.x : p.x, hi.y > p.y ? hi.y : p.y, hi.z > p.z ? hi.z : p.z);
    }
    bool valid() const { return lo.x <= hi.x; }
};

struct Grid
{
    int nx, ny, nz;
    V3 origin;
    float dl;
    size_t cellIndex(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
    bool inGrid(int i, int j, int k) const
    {
        return i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz;
    }
    V3 cellCenter(int i, int j, int k) const
    {
        return V3(origin.x + (i + 0.5f) * dl, origin.y + (j + 0.5f) * dl, origin.z + (k + 0.5f) * dl);
    }
    void cellOf(const V3 &p, int &i, int &j, int &k) const
    {
        i = (int)std::floor((p.x - origin.x) / dl);
        j = (int)std::floor((p.y - origin.y) / dl);
        k = (int)std::floor((p.z - origin.z) / dl);
    }
};

// Geometry.cpp:306-315
bool AxisTest(float a, float b, float fa, float fb, float v0a, float v0b, float v2a, float v2b,
              float ha, float hb)
{
    float p0 = a * v0a - b * v0b;
    float p2 = a * v2a - b * v2b;
    float mn = p0 < p2 ? p0 : p2, mx = p0 < p2 ? p2 : p0;
    float rad = fa * ha + fb * hb;
    return !(mn > rad || mx < -rad);
}

// Geometry.cpp:318-363 (Akenine-Moller separating axes)
bool TriBoxOverlap(const V3 &c, const V3 &h, const V3 &tv0, const V3 &tv1, const V3 &tv2)
{
    V3 v0 = tv0 - c, v1 = tv1 - c, v2 = tv2 - c;
    V3 e0 = v1 - v0, e1 = v2 - v1, e2 = v0 - v2;
    float fex = std::fabs(e0.x), fey = std::fabs(e0.y), fez = std::fabs(e0.z);
    if (!AxisTest(e0.z, e0.y, fez, fey, v0.y, v0.z, v2.y, v2.z, h.y, h.z)) return false;
    if (!AxisTest(e0.z, e0.x, fez, fex, v0.x, v0.z, v2.x, v2.z, h.x, h.z)) return false;
    if (!AxisTest(e0.y, e0.x, fey, fex, v1.x, v1.y, v2.x, v2.y, h.x, h.y)) return false;
    fex = std::fabs(e1.x);
    fey = std::fabs(e1.y);
    fez = std::fabs(e1.z);
    if (!AxisTest(e1.z, e1.y, fez, fey, v0.y, v0.z, v2.y, v2.z, h.y, h.z)) return false;
    if (!AxisTest(e1.z, e1.x, fez, fex, v0.x, v0.z, v2.x, v2.z, h.x, h.z)) return false;
    if (!AxisTest(e1.y, e1.x, fey, fex, v0.x, v0.y, v1.x, v1.y, h.x, h.y)) return false;
    fex = std::fabs(e2.x);
    fey = std::fabs(e2.y);
    fez = std::fabs(e2.z);
    if (!AxisTest(e2.z, e2.y, fez, fey, v0.y, v0.z, v1.y, v1.z, h.y, h.z)) return false;
    if (!AxisTest(e2.z, e2.x, fez, fex, v0.x, v0.z, v1.x, v1.z, h.x, h.z)) return false;
    if (!AxisTest(e2.y, e2.x, fey, fex, v1.x, v1.y, v2.x, v2.y, h.x, h.y)) return false;
    auto minMax3 = [](float a, float b, float cc, float &mn, float &mx) {
        mn = mx = a;
        if (b < mn) mn = b;
        if (b > mx) mx = b;
        if (cc < mn) mn = cc;
        if (cc > mx) mx = cc;
    };
    float mn, mx;
    minMax3(v0.x, v1.x, v2.x, mn, mx);
    if (mn > h.x || mx < -h.x) return false;
    minMax3(v0.y, v1.y, v2.y, mn, mx);
    if (mn > h.y || mx < -h.y) return false;
    minMax3(v0.z, v1.z, v2.z, mn, mx);
    if (mn > h.z || mx < -h.z) return false;
    V3 n = e0.cross(e1);
    V3 vmnC, vmxC;
    for (int q = 0; q < 3; ++q)
    {
        float v0q = v0[q], hq = h[q];
        if (n[q] > 0.0f)
        {
            vmnC.set(q, -hq - v0q);
            vmxC.set(q, hq - v0q);
        }
        else
        {
            vmnC.set(q, hq - v0q);
            vmxC.set(q, -hq - v0q);
        }
    }
    if (n.dot(vmnC) > 0.0f) return false;
    if (n.dot(vmxC) >= 0.0f) return true;
    return false;
}

// Geometry.cpp:369-397
void VoxelizeMesh(const std::vector<V3> &verts, const int *idx, int ntri, const V3 &offset,
                  const Grid &g, uint8_t *mat, uint8_t matId)
{
    const V3 half(g.dl * 0.5f, g.dl * 0.5f, g.dl * 0.5f);
    for (int t = 0; t < ntri; ++t)
    {
        V3 a = verts[idx[t * 3]] + offset;
        V3 b = verts[idx[t * 3 + 1]] + offset;
        V3 c = verts[idx[t * 3 + 2]] + offset;
        Box tb;
        tb.grow(a);
        tb.grow(b);
        tb.grow(c);
        int i0, j0, k0, i1, j1, k1;
        g.cellOf(tb.lo, i0, j0, k0);
        g.cellOf(tb.hi, i1, j1, k1);
        i0 = std::max(i0, 0);
        j0 = std::max(j0, 0);
        k0 = std::max(k0, 0);
        i1 = std::min(i1, g.nx - 1);
        j1 = std::min(j1, g.ny - 1);
        k1 = std::min(k1, g.nz - 1);
        for (int k = k0; k <= k1; ++k)
            for (int j = j0; j <= j1; ++j)
                for (int i = i0; i <= i1; ++i)
                {
                    size_t ci = g.cellIndex(i, j, k);
                    if (mat[ci] == matId)
                        continue;
                    if (TriBoxOverlap(g.cellCenter(i, j, k), half, a, b, c))
                        mat[ci] = matId;
                }
    }
}

// Geometry.cpp:525-548
bool RayXTri(const V3 &o, const V3 &a, const V3 &b, const V3 &c, float &hitX)
{
    const V3 dir(1, 0, 0);
    V3 e1 = b - a, e2 = c - a;
    V3 p = dir.cross(e2);
    float det = e1.dot(p);
    if (std::fabs(det) < 1e-12f)
        return false;
    float inv = 1.0f / det;
    V3 tv = o - a;
    float u = tv.dot(p) * inv;
    if (u < 0.0f || u > 1.0f)
        return false;
    V3 q = tv.cross(e1);
    float v = dir.dot(q) * inv;
    if (v < 0.0f || u + v > 1.0f)
        return false;
    float t = e2.dot(q) * inv;
    if (t <= 0.0f)
        return false;
    hitX = o.x + t;
    return true;
}

// Geometry.cpp:551-610
void VoxelizeMeshSolid(const std::vector<V3> &verts, const int *idx, int ntri, const V3 &offset,
                       const Grid &g, uint8_t *mat, uint8_t matId)
{
    VoxelizeMesh(verts, idx, ntri, offset, g, mat, matId);
    Box mb;
    for (const V3 &v : verts)
        mb.grow(v + offset);
    if (!mb.valid())
        return;
    const V3 lo = mb.lo;
    int i0, j0, k0, i1, j1, k1;
    g.cellOf(mb.lo, i0, j0, k0);
    g.cellOf(mb.hi, i1, j1, k1);
    i0 = std::max(i0, 0);
    j0 = std::max(j0, 0);
    k0 = std::max(k0, 0);
    i1 = std::min(i1, g.nx - 1);
    j1 = std::min(j1, g.ny - 1);
    k1 = std::min(k1, g.nz - 1);
    std::vector<float> hits;
    for (int k = k0; k <= k1; ++k)
        for (int j = j0; j <= j1; ++j)
        {
            V3 o = g.cellCenter(i0, j, k);
            o.x = lo.x - g.dl;
            hits.clear();
            for (int t = 0; t < ntri; ++t)
            {
                float hx;
                if (RayXTri(o, verts[idx[t * 3]] + offset, verts[idx[t * 3 + 1]] + offset,
                            verts[idx[t * 3 + 2]] + offset, hx))
                    hits.push_back(hx);
            }
            if (hits.size() < 2)
                continue;
            std::sort(hits.begin(), hits.end());
            for (size_t h = 0; h + 1 < hits.size(); h += 2)
            {
                float xa = hits[h], xb = hits[h + 1];
                int ia, ib, d1, d2;
                g.cellOf(V3(xa, 0, 0), ia, d1, d2);
                g.cellOf(V3(xb, 0, 0), ib, d1, d2);
                ia = std::max(ia, i0);
                ib = std::min(ib, i1);
                for (int i = ia; i <= ib; ++i)
                {
                    float cx = g.origin.x + (i + 0.5f) * g.dl;
                    if (cx >= xa && cx <= xb)
                        mat[g.cellIndex(i, j, k)] = matId;
                }
            }
        }
}

// Geometry.cpp:608-642
void VoxelizeWire(const std::vector<V3> &pts, float radius, const V3 &offset, const Grid &g,
                  uint8_t *mat, uint8_t matId)
{
    auto stamp = [&](const V3 &p) {
        int rc = (int)std::ceil(radius / g.dl - 0.5f);
        if (rc < 0)
            rc = 0;
        int ic, jc, kc;
        g.cellOf(p, ic, jc, kc);
        for (int k = kc - rc; k <= kc + rc; ++k)
            for (int j = jc - rc; j <= jc + rc; ++j)
                for (int i = ic - rc; i <= ic + rc; ++i)
                {
                    if (!g.inGrid(i, j, k))
                        continue;
                    if (rc > 0)
                    {
                        V3 d = g.cellCenter(i, j, k) - p;
                        if (d.length() > radius + g.dl * 0.5f)
                            continue;
                    }
                    mat[g.cellIndex(i, j, k)] = matId;
                }
    };
    for (size_t s = 0; s + 1 < pts.size(); ++s)
    {
        V3 a = pts[s] + offset, b = pts[s + 1] + offset;
        float len = (b - a).length();
        int steps = std::max(1, (int)std::ceil(len / (g.dl * 0.25f)));
        for (int i = 0; i <= steps; ++i)
            stamp(a + (b - a) * ((float)i / steps));
    }
}

std::vector<V3> to_v3(const float *xyz, int n)
{
    std::vector<V3> v(n);
    for (int i = 0; i < n; ++i)
        v[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return v;
}

} // namespace

extern "C" {

// mat: nx*ny*nz material ids, updated in place; origin / offset: 3 floats.
void ot_voxelize_mesh(int nx, int ny, int nz, const float *origin, float dl, const float *verts,
                      int nverts, const int32_t *idx, int ntri, const float *offset, uint8_t mat_id,
                      int solid, uint8_t *mat)
{
    Grid g{nx, ny, nz, V3(origin[0], origin[1], origin[2]), dl};
    std::vector<V3> v = to_v3(verts, nverts);
    std::vector<int> ix(idx, idx + 3 * ntri);
    V3 off(offset[0], offset[1], offset[2]);
    if (solid)
        VoxelizeMeshSolid(v, ix.data(), ntri, off, g, mat, mat_id);
    else
        VoxelizeMesh(v, ix.data(), ntri, off, g, mat, mat_id);
}

void ot_voxelize_wire(int nx, int ny, int nz, const float *origin, float dl, const float *pts,
                      int npts, float radius, const float *offset, uint8_t mat_id, uint8_t *mat)
{
    Grid g{nx, ny, nz, V3(origin[0], origin[1], origin[2]), dl};
    VoxelizeWire(to_v3(pts, npts), radius, V3(offset[0], offset[1], offset[2]), g, mat, mat_id);
}

int ot_tri_box_overlap(const float *center, const float *half, const float *tri)
{
    return TriBoxOverlap(V3(center[0], center[1], center[2]), V3(half[0], half[1], half[2]),
                         V3(tri[0], tri[1], tri[2]), V3(tri[3], tri[4], tri[5]),
                         V3(tri[6], tri[7], tri[8]));
}

int ot_ray_x_tri(const float *o, const float *tri, float *hit_x)
{
    return RayXTri(V3(o[0], o[1], o[2]), V3(tri[0], tri[1], tri[2]), V3(tri[3], tri[4], tri[5]),
                   V3(tri[6], tri[7], tri[8]), *hit_x);
}

// Feed gap (MainUnit.cpp:636-660): cells sampled along the inner 70% of
// a->b (both + offset) carved to AIR; sorted unique cells, dominant axis.
int ot_carve_feed(int nx, int ny, int nz, const float *origin, float dl, const float *fa,
                  const float *fb, const float *offset, uint8_t *mat, int64_t *cells, int cap,
                  int32_t *pol)
{
    Grid g{nx, ny, nz, V3(origin[0], origin[1], origin[2]), dl};
    const V3 off(offset[0], offset[1], offset[2]);
    V3 a = V3(fa[0], fa[1], fa[2]) + off;
    V3 b = V3(fb[0], fb[1], fb[2]) + off;
    V3 d = b - a;
    V3 a2 = a + d * 0.15f, b2 = b - d * 0.15f;
    float len = (b2 - a2).length();
    int steps = std::max(1, (int)std::ceil(len / (dl * 0.25f)));
    std::vector<int64_t> out;
    for (int s = 0; s <= steps; ++s)
    {
        V3 pp = a2 + (b2 - a2) * ((float)s / steps);
        int ci, cj, ck;
        g.cellOf(pp, ci, cj, ck);
        if (!g.inGrid(ci, cj, ck))
            continue;
        size_t cell = g.cellIndex(ci, cj, ck);
        mat[cell] = 0;
        out.push_back((int64_t)cell);
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    int p = 0;
    if (std::fabs(d.y) > std::fabs(d[p]))
        p = 1;
    if (std::fabs(d.z) > std::fabs(d[p]))
        p = 2;
    *pol = p;
    for (size_t q = 0; q < out.size() && (int)q < cap; ++q)
        cells[q] = out[q];
    return (int)out.size();
}

// TriMesh::addBox (Geometry.cpp:135-147): 8 vertices, 12 triangles.
void ot_box_mesh(const float *lo, const float *hi, float *verts, int32_t *idx)
{
    // the reference adds each quad as two triangles with fresh vertices; the
    // geometry is the same as this shared-vertex form
    const float P[8][3] = {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]}, {lo[0], hi[1], lo[2]},
                           {hi[0], hi[1], lo[2]}, {lo[0], lo[1], hi[2]}, {hi[0], lo[1], hi[2]},
                           {lo[0], hi[1], hi[2]}, {hi[0], hi[1], hi[2]}};
    for (int i = 0; i < 8; ++i)
        for (int c = 0; c < 3; ++c)
            verts[3 * i + c] = P[i][c];
    // quads (a, b, c, d) -> (a, b, c), (a, c, d), in the reference's face order
    const int Q[6][4] = {{0, 2, 3, 1}, {4, 5, 7, 6}, {0, 1, 5, 4},
                         {2, 6, 7, 3}, {0, 4, 6, 2}, {1, 3, 7, 5}};
    for (int f = 0; f < 6; ++f)
    {
        const int t0[3] = {Q[f][0], Q[f][1], Q[f][2]}, t1[3] = {Q[f][0], Q[f][2], Q[f][3]};
        for (int c = 0; c < 3; ++c)
        {
            idx[6 * f + c] = t0[c];
            idx[6 * f + 3 + c] = t1[c];
        }
    }
}

} // extern "C"
