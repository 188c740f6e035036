// TlmVoxel.hip - scene -> voxel set-up on the device (the step's input side).
//
// The reference voxelises a scene on one host thread before every run
// (TMainForm::voxelizeScene, MainUnit.cpp:431-459) and then walks the whole
// material map again to build the solver's state. Here the material map is a
// device buffer of the context, written in call order by:
//   - a cell-range fill (synthetic scenes);
//   - VoxelizeMesh (Geometry.cpp:369-397): every (triangle, cell of its
//     bounding box) pair is one thread, a triangle/box separating-axis test;
//   - VoxelizeMeshSolid (Geometry.cpp:551-606): the surface pass, then one
//     thread per (j,k) row casts the +x ray against all triangles (staged
//     through LDS), sorts its hits and fills the cells between hit pairs;
//   - VoxelizeWire (Geometry.cpp:608-642): one thread per (sample, cell of
//     the stamp cube);
//   - the feed-gap carve (MainUnit.cpp:636-660).
// Writes within one call all store the same id, so their order does not
// matter and the map equals the reference's sequential one bit for bit: the
// arithmetic (cellOf, cellCenter, the tests) is the reference's FP32
// expression order, built without contraction, with IEEE division and square
// root. tlm_upload_mat then turns the map into the cell codes and air-block
// flags on the device, and the PEC surface faces (Geometry.cpp:402-436) and
// Huygens faces (TlmSolver.cpp:313-359) are ordered compactions of the map.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function" // the step kernels of the shared header
#include "tlm_context.h"
#pragma clang diagnostic pop

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace tlmi;

namespace {

// ---- FP32 vector arithmetic in the reference's operation order (Vec3.h) ----

struct F3
{
    float x, y, z;
};

__host__ __device__ inline F3 mk(float x, float y, float z)
{
    F3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
__host__ __device__ inline F3 add(F3 a, F3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline F3 sub(F3 a, F3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline F3 mul(F3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline F3 cross(F3 a, F3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline float comp(F3 a, int q) { return q == 0 ? a.x : (q == 1 ? a.y : a.z); }
__host__ __device__ inline void setc(F3 &a, int q, float v)
{
    if (q == 0)
        a.x = v;
    else if (q == 1)
        a.y = v;
    else
        a.z = v;
}

// Aabb::grow (Vec3.h:54-60): vmin / vmax selects from +-1e30.
struct Box3
{
    F3 lo, hi;
    __host__ __device__ void reset()
    {
        lo = mk(1e30f, 1e30f, 1e30f);
        hi = mk(-1e30f, -1e30f, -1e30f);
    }
    __host__ __device__ void grow(F3 p)
    {
        lo = mk(lo.x < p.x ? lo.x : p.x, lo.y < p.y ? lo.y : p.y, lo.z < p.z ? lo.z : p.z);
        hi = mk(hi.x > p.x ? hi.x : p.x, hi.y > p.y ? hi.y : p.y, hi.z > p.z ? hi.z : p.z);
    }
};

// VoxelGridSpec (Geometry.h:85-113).
struct VGrid
{
    int nx, ny, nz;
    F3 o;
    float dl;
    __host__ __device__ F3 center(int i, int j, int k) const
    {
        return mk(o.x + (i + 0.5f) * dl, o.y + (j + 0.5f) * dl, o.z + (k + 0.5f) * dl);
    }
    __host__ __device__ void cell_of(F3 p, int &i, int &j, int &k) const
    {
        i = (int)floorf((p.x - o.x) / dl);
        j = (int)floorf((p.y - o.y) / dl);
        k = (int)floorf((p.z - o.z) / dl);
    }
    __host__ __device__ bool in_grid(int i, int j, int k) const
    {
        return i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz;
    }
    __host__ __device__ long long index(int i, int j, int k) const
    {
        return ((long long)k * ny + j) * nx + i;
    }
};

// AxisTest and TriBoxOverlap (Geometry.cpp:306-363), Akenine-Moller.
__device__ inline bool axis_test(float a, float b, float fa, float fb, float v0a, float v0b,
                                 float v2a, float v2b, float ha, float hb)
{
    const float p0 = a * v0a - b * v0b;
    const float p2 = a * v2a - b * v2b;
    const float mn = p0 < p2 ? p0 : p2, mx = p0 < p2 ? p2 : p0;
    const float rad = fa * ha + fb * hb;
    return !(mn > rad || mx < -rad);
}

__device__ inline void min_max3(float a, float b, float c, float &mn, float &mx)
{
    mn = mx = a;
    if (b < mn)
        mn = b;
    if (b > mx)
        mx = b;
    if (c < mn)
        mn = c;
    if (c > mx)
        mx = c;
}

__device__ bool tri_box_overlap(F3 c, F3 h, F3 t0, F3 t1, F3 t2)
{
    const F3 v0 = sub(t0, c), v1 = sub(t1, c), v2 = sub(t2, c);
    const F3 e0 = sub(v1, v0), e1 = sub(v2, v1), e2 = sub(v0, v2);
    float fex = fabsf(e0.x), fey = fabsf(e0.y), fez = fabsf(e0.z);
    if (!axis_test(e0.z, e0.y, fez, fey, v0.y, v0.z, v2.y, v2.z, h.y, h.z)) return false;
    if (!axis_test(e0.z, e0.x, fez, fex, v0.x, v0.z, v2.x, v2.z, h.x, h.z)) return false;
    if (!axis_test(e0.y, e0.x, fey, fex, v1.x, v1.y, v2.x, v2.y, h.x, h.y)) return false;
    fex = fabsf(e1.x);
    fey = fabsf(e1.y);
    fez = fabsf(e1.z);
    if (!axis_test(e1.z, e1.y, fez, fey, v0.y, v0.z, v2.y, v2.z, h.y, h.z)) return false;
    if (!axis_test(e1.z, e1.x, fez, fex, v0.x, v0.z, v2.x, v2.z, h.x, h.z)) return false;
    if (!axis_test(e1.y, e1.x, fey, fex, v0.x, v0.y, v1.x, v1.y, h.x, h.y)) return false;
    fex = fabsf(e2.x);
    fey = fabsf(e2.y);
    fez = fabsf(e2.z);
    if (!axis_test(e2.z, e2.y, fez, fey, v0.y, v0.z, v1.y, v1.z, h.y, h.z)) return false;
    if (!axis_test(e2.z, e2.x, fez, fex, v0.x, v0.z, v1.x, v1.z, h.x, h.z)) return false;
    if (!axis_test(e2.y, e2.x, fey, fex, v1.x, v1.y, v2.x, v2.y, h.x, h.y)) return false;
    float mn, mx;
    min_max3(v0.x, v1.x, v2.x, mn, mx);
    if (mn > h.x || mx < -h.x) return false;
    min_max3(v0.y, v1.y, v2.y, mn, mx);
    if (mn > h.y || mx < -h.y) return false;
    min_max3(v0.z, v1.z, v2.z, mn, mx);
    if (mn > h.z || mx < -h.z) return false;
    const F3 n = cross(e0, e1);
    F3 vmn = mk(0.f, 0.f, 0.f), vmx = mk(0.f, 0.f, 0.f);
    for (int q = 0; q < 3; ++q)
    {
        const float v0q = comp(v0, q), hq = comp(h, q);
        if (comp(n, q) > 0.0f)
        {
            setc(vmn, q, -hq - v0q);
            setc(vmx, q, hq - v0q);
        }
        else
        {
            setc(vmn, q, hq - v0q);
            setc(vmx, q, -hq - v0q);
        }
    }
    if (dot(n, vmn) > 0.0f)
        return false;
    return dot(n, vmx) >= 0.0f;
}

// RayXTri (Geometry.cpp:525-548): Moller-Trumbore, ray +x from o.
__device__ bool ray_x_tri(F3 o, F3 a, F3 b, F3 c, float &hit_x)
{
    const F3 dir = mk(1.f, 0.f, 0.f);
    const F3 e1 = sub(b, a), e2 = sub(c, a);
    const F3 p = cross(dir, e2);
    const float det = dot(e1, p);
    if (fabsf(det) < 1e-12f)
        return false;
    const float inv = 1.0f / det;
    const F3 tv = sub(o, a);
    const float u = dot(tv, p) * inv;
    if (u < 0.0f || u > 1.0f)
        return false;
    const F3 q = cross(tv, e1);
    const float v = dot(dir, q) * inv;
    if (v < 0.0f || u + v > 1.0f)
        return false;
    const float t = dot(e2, q) * inv;
    if (t <= 0.0f)
        return false;
    hit_x = o.x + t;
    return true;
}

__device__ inline F3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

// Largest t with off[t] <= q (off ascending, off[0] = 0).
__device__ inline int upper_index(const long long *off, int n, long long q)
{
    int lo = 0, hi = n; // off[lo] <= q < off[hi] (off[n] = +inf)
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= q)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

constexpr int kThreads = 256;
constexpr int kPer = 16; // elements per thread in the scans and compactions
constexpr int kTile = kThreads * kPer;

// ---- exclusive int64 scan (tiles of 4096, recursive over the tile sums) ----

__device__ long long block_exclusive(long long v, long long *lds, long long &total)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1)
    {
        const long long add_v = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += add_v;
        __syncthreads();
    }
    const long long incl = lds[t];
    total = lds[kThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kThreads) k_tile_sums(const long long *in, long long n,
                                                        long long *sums)
{
    __shared__ long long lds[kThreads];
    const long long b = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
    long long s = 0;
    for (int r = 0; r < kPer; ++r)
        if (b + r < n)
            s += in[b + r];
    long long tot;
    block_exclusive(s, lds, tot);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = tot;
}

// out[i] = base(tile) + exclusive prefix within the tile; *total written by
// the last tile.
__global__ void __launch_bounds__(kThreads) k_scan_tiles(const long long *in, long long n,
                                                         const long long *tile_base,
                                                         long long *out, long long *total)
{
    __shared__ long long lds[kThreads];
    const long long b = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
    long long s = 0;
    for (int r = 0; r < kPer; ++r)
        if (b + r < n)
            s += in[b + r];
    long long tot;
    long long run = block_exclusive(s, lds, tot) + (tile_base ? tile_base[blockIdx.x] : 0);
    for (int r = 0; r < kPer; ++r)
        if (b + r < n)
        {
            out[b + r] = run;
            run += in[b + r];
        }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *total = (tile_base ? tile_base[blockIdx.x] : 0) + tot;
}

// Device buffers of one call, freed at scope exit.
struct Scratch
{
    std::vector<void *> ptrs;
    ~Scratch()
    {
        for (void *p : ptrs)
            (void)hipFree(p);
    }
    template <class T> int get(tlm_ctx *c, T **p, size_t count)
    {
        *p = nullptr;
        hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess)
        {
            *p = nullptr;
            return hip_fail(c, e, "hipMalloc");
        }
        ptrs.push_back(*p);
        return TLM_OK;
    }
};

inline unsigned grid_for(long long n, long long cap = 1 << 20)
{
    return (unsigned)std::max(1LL, std::min(cap, (n + kThreads - 1) / kThreads));
}

// Exclusive scan of in[n] into out[n] on the context's stream; the sum is
// left in *d_total (device).
int scan_i64(tlm_ctx *c, Scratch &s, const long long *in, long long *out, long long n,
             long long *d_total)
{
    const long long nt = (n + kTile - 1) / kTile;
    if (nt <= 1)
    {
        k_scan_tiles<<<1, kThreads, 0, c->stream>>>(in, n, nullptr, out, d_total);
        HIPCHK(c, hipGetLastError());
        return TLM_OK;
    }
    long long *sums, *base, *dummy;
    int rc;
    if ((rc = s.get(c, &sums, nt)) || (rc = s.get(c, &base, nt)) || (rc = s.get(c, &dummy, 1)))
        return rc;
    k_tile_sums<<<(unsigned)nt, kThreads, 0, c->stream>>>(in, n, sums);
    HIPCHK(c, hipGetLastError());
    if ((rc = scan_i64(c, s, sums, base, nt, dummy)))
        return rc;
    k_scan_tiles<<<(unsigned)nt, kThreads, 0, c->stream>>>(in, n, base, out, d_total);
    HIPCHK(c, hipGetLastError());
    return TLM_OK;
}

int read_total(tlm_ctx *c, const long long *d_total, long long *total)
{
    HIPCHK(c, hipMemcpyAsync(total, d_total, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLM_OK;
}

// ---- material map kernels ----

__global__ void k_cells_box(uint8_t *mat, VGrid g, int i0, int j0, int k0, int ni, int nj,
                            long long n, uint8_t id)
{
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x)
    {
        const int i = i0 + (int)(q % ni);
        const long long r = q / ni;
        const int j = j0 + (int)(r % nj);
        const int k = k0 + (int)(r / nj);
        mat[g.index(i, j, k)] = id;
    }
}

// Per triangle: corners with the offset, clamped cell bounding box
// (VoxelizeMesh, Geometry.cpp:376-386) and its cell count.
__global__ void k_tri_prep(const float *verts, const int *idx, int ntri, F3 off, VGrid g,
                           float *tri, int *bb, long long *cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri)
        return;
    const F3 a = add(ld3(verts + 3 * (long long)idx[3 * t]), off);
    const F3 b = add(ld3(verts + 3 * (long long)idx[3 * t + 1]), off);
    const F3 c = add(ld3(verts + 3 * (long long)idx[3 * t + 2]), off);
    Box3 tb;
    tb.reset();
    tb.grow(a);
    tb.grow(b);
    tb.grow(c);
    int i0, j0, k0, i1, j1, k1;
    g.cell_of(tb.lo, i0, j0, k0);
    g.cell_of(tb.hi, i1, j1, k1);
    i0 = max(i0, 0);
    j0 = max(j0, 0);
    k0 = max(k0, 0);
    i1 = min(i1, g.nx - 1);
    j1 = min(j1, g.ny - 1);
    k1 = min(k1, g.nz - 1);
    float *o = tri + 9 * (long long)t;
    o[0] = a.x, o[1] = a.y, o[2] = a.z;
    o[3] = b.x, o[4] = b.y, o[5] = b.z;
    o[6] = c.x, o[7] = c.y, o[8] = c.z;
    int *q = bb + 6 * (long long)t;
    q[0] = i0, q[1] = j0, q[2] = k0, q[3] = i1, q[4] = j1, q[5] = k1;
    cnt[t] = (i1 >= i0 && j1 >= j0 && k1 >= k0)
                 ? (long long)(i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1)
                 : 0;
}

// One thread per (triangle, cell of its box): the surface pass.
__global__ void k_tri_cells(const float *tri, const int *bb, const long long *off, int ntri,
                            long long total, VGrid g, uint8_t *mat, uint8_t id)
{
    const F3 half = mk(g.dl * 0.5f, g.dl * 0.5f, g.dl * 0.5f);
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x)
    {
        const int t = upper_index(off, ntri, q);
        const int *b = bb + 6 * (long long)t;
        const int ni = b[3] - b[0] + 1, nj = b[4] - b[1] + 1;
        long long r = q - off[t];
        const int i = b[0] + (int)(r % ni);
        r /= ni;
        const int j = b[1] + (int)(r % nj);
        const int k = b[2] + (int)(r / nj);
        const long long ci = g.index(i, j, k);
        if (mat[ci] == id)
            continue;
        const float *p = tri + 9 * (long long)t;
        if (tri_box_overlap(g.center(i, j, k), half, ld3(p), ld3(p + 3), ld3(p + 6)))
            mat[ci] = id;
    }
}

// Rows (j,k) of the solid fill: j = j0 + row % nj, k = k0 + row / nj; the ray
// starts at (xs, cellCenter y, z) (Geometry.cpp:574-576).
struct Rows
{
    int i0, i1, j0, k0, nj;
    long long n;
    long long nh; // hits of all rows (fill pass)
    float xs;
};

// Sorts a row's hits (ascending, as std::sort) and fills the cells whose
// centre lies between the hits of each pair (Geometry.cpp:586-604).
__device__ void fill_row(float *h, long long nh, long long row, const Rows &rw, const VGrid &g,
                         uint8_t *mat, uint8_t id)
{
    if (nh < 2)
        return;
    for (long long a = 1; a < nh; ++a) // insertion sort
    {
        const float v = h[a];
        long long b = a - 1;
        while (b >= 0 && h[b] > v)
        {
            h[b + 1] = h[b];
            --b;
        }
        h[b + 1] = v;
    }
    const int j = rw.j0 + (int)(row % rw.nj);
    const int k = rw.k0 + (int)(row / rw.nj);
    for (long long p = 0; p + 1 < nh; p += 2)
    {
        const float xa = h[p], xb = h[p + 1];
        int ia, ib, d1, d2;
        g.cell_of(mk(xa, 0.f, 0.f), ia, d1, d2);
        g.cell_of(mk(xb, 0.f, 0.f), ib, d1, d2);
        ia = max(ia, rw.i0);
        ib = min(ib, rw.i1);
        for (int i = ia; i <= ib; ++i)
        {
            const float cx = g.o.x + (i + 0.5f) * g.dl;
            if (cx >= xa && cx <= xb)
                mat[g.index(i, j, k)] = id;
        }
    }
}

__device__ inline F3 row_origin(long long row, const Rows &rw, const VGrid &g)
{
    const int j = rw.j0 + (int)(row % rw.nj);
    const int k = rw.k0 + (int)(row / rw.nj);
    F3 o = g.center(rw.i0, j, k);
    o.x = rw.xs;
    return o;
}

constexpr int kRowCap = 16; // hit slots per row of the one-pass fill

// One pass over (block of 256 rows) x (chunk of triangles): each hit takes
// the next slot of its row (the order is restored by the sort); hits past
// kRowCap are counted but not kept.
__global__ void __launch_bounds__(kThreads)
    k_ray_slots(const float *tri, int ntri, int per_chunk, Rows rw, VGrid g, int *cnt,
                float *slots)
{
    __shared__ float st[kThreads * 9];
    const long long row = (long long)blockIdx.x * kThreads + threadIdx.x;
    const bool live = row < rw.n;
    const F3 o = live ? row_origin(row, rw, g) : mk(0.f, 0.f, 0.f);
    const int t_beg = blockIdx.y * per_chunk;
    const int t_end = min(ntri, t_beg + per_chunk);
    for (int t0 = t_beg; t0 < t_end; t0 += kThreads)
    {
        const int m = min(kThreads, t_end - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < m * 9; e += kThreads)
            st[e] = tri[9 * (long long)t0 + e];
        __syncthreads();
        if (!live)
            continue;
        for (int t = 0; t < m; ++t)
        {
            float hx;
            if (ray_x_tri(o, ld3(st + 9 * t), ld3(st + 9 * t + 3), ld3(st + 9 * t + 6), hx))
            {
                const int slot = atomicAdd(&cnt[row], 1);
                if (slot < kRowCap)
                    slots[row * kRowCap + slot] = hx;
            }
        }
    }
}

// Fill of the rows whose hits all fit their slots; *overflow = 1 if a row's
// did not (those rows are left to the two-pass fill).
__global__ void k_ray_fill(const int *cnt, float *slots, Rows rw, VGrid g, uint8_t *mat,
                           uint8_t id, int *overflow)
{
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rw.n)
        return;
    const int n = cnt[row];
    if (n > kRowCap)
    {
        *overflow = 1;
        return;
    }
    fill_row(slots + row * kRowCap, n, row, rw, g, mat, id);
}

// Two-pass fill for any hit count: counts (HITS = false), then the hits
// written at the row's offset, sorted and filled (HITS = true); one row per
// thread, the triangles pass through LDS 256 at a time.
template <bool HITS>
__global__ void __launch_bounds__(kThreads)
    k_rows(const float *tri, int ntri, Rows rw, VGrid g, long long *cnt, const long long *off,
           float *hits, uint8_t *mat, uint8_t id)
{
    __shared__ float st[kThreads * 9];
    const long long row = (long long)blockIdx.x * kThreads + threadIdx.x;
    const bool live = row < rw.n;
    const F3 o = live ? row_origin(row, rw, g) : mk(0.f, 0.f, 0.f);
    long long nh = 0;
    float *h = HITS && live ? hits + off[row] : nullptr;
    // the count pass gave this row's room; never write past it
    const long long room = HITS && live ? (row + 1 < rw.n ? off[row + 1] : rw.nh) - off[row] : 0;
    for (int t0 = 0; t0 < ntri; t0 += kThreads)
    {
        const int m = min(kThreads, ntri - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < m * 9; e += kThreads)
            st[e] = tri[9 * (long long)t0 + e];
        __syncthreads();
        if (!live)
            continue;
        for (int t = 0; t < m; ++t)
        {
            float hx;
            if (ray_x_tri(o, ld3(st + 9 * t), ld3(st + 9 * t + 3), ld3(st + 9 * t + 6), hx))
            {
                if (HITS && nh < room)
                    h[nh] = hx;
                ++nh;
            }
        }
    }
    if (!live)
        return;
    if (!HITS)
    {
        cnt[row] = nh;
        return;
    }
    fill_row(h, min(nh, room), row, rw, g, mat, id);
}

// One thread per (sample, cell of the (2rc+1)^3 stamp cube) of a wire
// (VoxelizeWire, Geometry.cpp:608-642). seg: per segment a (3), b (3);
// steps per segment; soff: first sample of each segment.
__global__ void k_wire(const float *seg, const int *steps, const long long *soff, int nseg,
                       long long total, int rc, float radius, VGrid g, uint8_t *mat, uint8_t id)
{
    const int D = 2 * rc + 1;
    const long long cube = (long long)D * D * D;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x)
    {
        const long long smp = q / cube;
        const int cc = (int)(q - smp * cube);
        const int s = upper_index(soff, nseg, smp);
        const int i = (int)(smp - soff[s]);
        const F3 a = ld3(seg + 6 * (long long)s), b = ld3(seg + 6 * (long long)s + 3);
        const F3 p = add(a, mul(sub(b, a), (float)i / steps[s]));
        int ic, jc, kc;
        g.cell_of(p, ic, jc, kc);
        const int ci = ic + cc % D - rc;
        const int cj = jc + (cc / D) % D - rc;
        const int ck = kc + cc / (D * D) - rc;
        if (!g.in_grid(ci, cj, ck))
            continue;
        if (rc > 0)
        {
            const F3 d = sub(g.center(ci, cj, ck), p);
            if (sqrtf(d.x * d.x + d.y * d.y + d.z * d.z) > radius + g.dl * 0.5f)
                continue;
        }
        mat[g.index(ci, cj, ck)] = id;
    }
}

__global__ void k_set_cells(uint8_t *mat, const long long *cells, int n, uint8_t id)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        mat[cells[t]] = id;
}

// Cell codes of the padded layout: material id | neighbour-PEC bits; the
// slab padding is coded PEC (the host build in tlm_upload, on the device).
// Block r < ny*nz codes row (j, k) = (r % ny, r / ny); block ny*nz + k the
// padding of slab k.
__global__ void k_codes(const uint8_t *mat, int nx, int ny, int nz, long long nxy, long long Pp,
                        uint16_t *code)
{
    const int r = blockIdx.x;
    if (r >= ny * nz)
    {
        const long long base = (long long)(r - ny * nz) * Pp;
        for (long long p = nxy + threadIdx.x; p < Pp; p += blockDim.x)
            code[base + p] = 1;
        return;
    }
    const int k = r / ny, j = r - k * ny;
    const uint8_t *row = mat + (long long)k * nxy + (long long)j * nx;
    uint16_t *out = code + (long long)k * Pp + (long long)j * nx;
    for (int i = threadIdx.x; i < nx; i += blockDim.x)
    {
        unsigned v = row[i];
        if (i > 0 && row[i - 1] == 1)
            v |= tlm::NB_XM;
        if (i < nx - 1 && row[i + 1] == 1)
            v |= tlm::NB_XP;
        if (j > 0 && row[i - nx] == 1)
            v |= tlm::NB_YM;
        if (j < ny - 1 && row[i + nx] == 1)
            v |= tlm::NB_YP;
        if (k > 0 && row[i - nxy] == 1)
            v |= tlm::NB_ZM;
        if (k < nz - 1 && row[i + nxy] == 1)
            v |= tlm::NB_ZP;
        out[i] = (uint16_t)v;
    }
}

// Per 64-cell block: bit 0 if all codes are 0 (air with no PEC neighbour),
// bit 2 if a cell is dielectric (material id >= 2 in the low byte of its
// code), bits 3 and 4 if one of cells 0-31 / 32-63 is, bit 5 (kBlockClean)
// if no code is PEC (padding included) or has a neighbour-PEC bit.
__global__ void k_uair(const uint16_t *code, long long nblk, uint8_t *uair)
{
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk)
        return;
    const uint4 *q = reinterpret_cast<const uint4 *>(code + b * tlm::kWave);
    unsigned any = 0, half = 0;
    bool clean = true;
    for (int r = 0; r < tlm::kWave / 8; ++r)
    {
        const uint4 v = q[r];
        any |= v.x | v.y | v.z | v.w;
        bool diel = false;
        for (unsigned w : {v.x, v.y, v.z, v.w})
        {
            diel = diel || (w & 0xFFu) >= 2u || ((w >> 16) & 0xFFu) >= 2u;
            clean = clean && tlm::code_is_clean(w & 0xFFFFu) && tlm::code_is_clean(w >> 16);
        }
        if (diel)
            half |= 8u << (r >> 2); // uint4 r holds cells 8r .. 8r+7
    }
    uair[b] = (uint8_t)((any == 0 ? 1u : 0u) | (half ? 4u : 0u) | half |
                        (clean ? tlm::kBlockClean : 0u));
}

// Largest dielectric index (id - 2) of the map, -1 if none (16 cells per
// load; the map is a 256-byte aligned allocation).
__global__ void k_max_diel(const uint8_t *mat, long long n, int *out)
{
    unsigned m = 0; // largest id seen
    const long long n16 = n / 16;
    const uint4 *v = reinterpret_cast<const uint4 *>(mat);
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n16;
         q += (long long)gridDim.x * blockDim.x)
    {
        const uint4 w = v[q];
        const unsigned x[4] = {w.x, w.y, w.z, w.w};
        for (int e = 0; e < 4; ++e)
            for (int b = 0; b < 32; b += 8)
                m = max(m, (x[e] >> b) & 0xffu);
    }
    if (blockIdx.x == 0)
        for (long long q = n16 * 16 + threadIdx.x; q < n; q += blockDim.x)
            m = max(m, (unsigned)mat[q]);
    int d = m >= 2 ? (int)m - 2 : -1;
    d = __reduce_max_sync(~0ull, d);
    if ((threadIdx.x & (tlm::kWave - 1)) == 0 && d >= 0)
        atomicMax(out, d);
}

// ---- ordered face lists: PEC surface faces and Huygens faces ----

// PEC faces of a PEC cell (i,j,k): bit axis*2 + (sign > 0) set where the
// neighbour across that face exists and is not PEC, the face order of
// BuildPecSurfaceFaces (Geometry.cpp:407-433).
__device__ inline unsigned pec_face_bits(const uint8_t *mat, long long c, int i, int j, int k,
                                         int nx, int ny, int nz, long long nxy)
{
    unsigned b = 0;
    if (i > 0 && mat[c - 1] != 1)
        b |= 1u;
    if (i < nx - 1 && mat[c + 1] != 1)
        b |= 2u;
    if (j > 0 && mat[c - nx] != 1)
        b |= 4u;
    if (j < ny - 1 && mat[c + nx] != 1)
        b |= 8u;
    if (k > 0 && mat[c - nxy] != 1)
        b |= 16u;
    if (k < nz - 1 && mat[c + nxy] != 1)
        b |= 32u;
    return b;
}

// PEC surface faces in cellIndex order: per thread kPer consecutive cells
// (one 16-byte load of the map). EMIT = false: face count per tile;
// EMIT = true: the faces (padded air cell, axis) at the tile's offset.
template <bool EMIT>
__global__ void __launch_bounds__(kThreads)
    k_pec_faces(const uint8_t *mat, int nx, int ny, int nz, long long nxy, long long n,
                long long *tcnt, const long long *toff, long long Pp, long long *cells, int *axes)
{
    static_assert(kPer == 16, "one uint4 of the map per thread");
    __shared__ long long lds[kThreads];
    const long long b = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
    uint8_t m[kPer];
    if (b + kPer <= n)
    {
        const uint4 w = *reinterpret_cast<const uint4 *>(mat + b);
        const unsigned x[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int r = 0; r < kPer; ++r)
            m[r] = (uint8_t)(x[r >> 2] >> (8 * (r & 3)));
    }
    else
    {
#pragma unroll
        for (int r = 0; r < kPer; ++r)
            m[r] = b + r < n ? mat[b + r] : 0;
    }
    bool any = false;
#pragma unroll
    for (int r = 0; r < kPer; ++r)
        any |= m[r] == 1;
    unsigned bits[kPer];
    long long cnt = 0;
    int i0 = 0, j0 = 0, k0 = 0;
    if (any)
    {
        k0 = (int)(b / nxy);
        const long long rem = b - (long long)k0 * nxy;
        j0 = (int)(rem / nx);
        i0 = (int)(rem - (long long)j0 * nx);
    }
    {
        int i = i0, j = j0, k = k0;
#pragma unroll
        for (int r = 0; r < kPer; ++r)
        {
            bits[r] = 0;
            if (m[r] == 1)
            {
                bits[r] = pec_face_bits(mat, b + r, i, j, k, nx, ny, nz, nxy);
                cnt += __builtin_popcount(bits[r]);
            }
            if (++i == nx)
            {
                i = 0;
                if (++j == ny)
                {
                    j = 0;
                    ++k;
                }
            }
        }
    }
    long long tot;
    long long w = block_exclusive(cnt, lds, tot);
    if (!EMIT)
    {
        if (threadIdx.x == 0)
            tcnt[blockIdx.x] = tot;
        return;
    }
    if (!any)
        return;
    w += toff[blockIdx.x];
    int i = i0, j = j0, k = k0;
#pragma unroll
    for (int r = 0; r < kPer; ++r)
    {
        if (bits[r])
        {
            const long long rowoff = (long long)j * nx + i;
            for (int f = 0; f < 6; ++f)
            {
                if (!((bits[r] >> f) & 1u))
                    continue;
                const int a = f >> 1;
                const int sg = (f & 1) ? 1 : -1;
                const long long kk = k + (a == 2 ? sg : 0);
                const long long ro = rowoff + (a == 0 ? sg : (a == 1 ? sg * (long long)nx : 0));
                cells[w] = kk * Pp + ro; // padded device address of the air cell
                axes[w] = a;
                ++w;
            }
        }
        if (++i == nx)
        {
            i = 0;
            if (++j == ny)
            {
                j = 0;
                ++k;
            }
        }
    }
}

// The three face lists of TlmSolver::buildHuygens (TlmSolver.cpp:337-358),
// concatenated: x faces per (k,j), y faces per (k,i), z faces per (j,i),
// each as (low side, high side); AIR cells only.
struct HuyFaces
{
    const uint8_t *mat;
    int nx, ny;
    int x0, x1, y0, y1, z0, z1;
    long long nX, nY; // candidates in the x and y lists
    __device__ bool operator()(long long q, long long &cell, int &axis) const
    {
        const int NX = x1 - x0 + 1, NY = y1 - y0 + 1;
        int i, j, k, a;
        if (q < nX)
        {
            const long long pr = q >> 1;
            k = z0 + (int)(pr / NY);
            j = y0 + (int)(pr % NY);
            i = (q & 1) ? x1 : x0;
            a = 0;
        }
        else if (q < nX + nY)
        {
            const long long qq = q - nX, pr = qq >> 1;
            k = z0 + (int)(pr / NX);
            i = x0 + (int)(pr % NX);
            j = (qq & 1) ? y1 : y0;
            a = 1;
        }
        else
        {
            const long long qq = q - nX - nY, pr = qq >> 1;
            j = y0 + (int)(pr / NX);
            i = x0 + (int)(pr % NX);
            k = (qq & 1) ? z1 : z0;
            a = 2;
        }
        const long long c = ((long long)k * ny + j) * nx + i;
        if (mat[c] != 0)
            return false;
        cell = c;
        axis = a;
        return true;
    }
};

template <class Pred>
__global__ void __launch_bounds__(kThreads) k_compact_count(Pred pred, long long nq, long long *tcnt)
{
    __shared__ long long lds[kThreads];
    const long long b = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
    long long s = 0;
    for (int r = 0; r < kPer; ++r)
    {
        long long cell;
        int axis;
        if (b + r < nq && pred(b + r, cell, axis))
            ++s;
    }
    long long tot;
    block_exclusive(s, lds, tot);
    if (threadIdx.x == 0)
        tcnt[blockIdx.x] = tot;
}

template <class Pred>
__global__ void __launch_bounds__(kThreads)
    k_compact_emit(Pred pred, long long nq, const long long *toff, long long nxy, long long Pp,
                   long long *cells, int *axes)
{
    __shared__ long long lds[kThreads];
    const long long b = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kPer;
    long long s = 0;
    for (int r = 0; r < kPer; ++r)
    {
        long long cell;
        int axis;
        if (b + r < nq && pred(b + r, cell, axis))
            ++s;
    }
    long long tot;
    long long w = block_exclusive(s, lds, tot) + toff[blockIdx.x];
    for (int r = 0; r < kPer; ++r)
    {
        long long cell;
        int axis;
        if (b + r < nq && pred(b + r, cell, axis))
        {
            const long long k = cell / nxy;
            cells[w] = k * Pp + (cell - k * nxy); // padded device address
            axes[w] = axis;
            ++w;
        }
    }
}

// Runs the compaction; on success *pcells / *paxes are new device arrays
// (owned by the context) of *n entries.
template <class Pred>
int compact(tlm_ctx *c, const Pred &pred, long long nq, long long **pcells, int **paxes,
            long long *n)
{
    Scratch s;
    const long long nt = std::max(1LL, (nq + kTile - 1) / kTile);
    long long *tcnt, *toff, *dtot;
    int rc;
    if ((rc = s.get(c, &tcnt, nt)) || (rc = s.get(c, &toff, nt)) || (rc = s.get(c, &dtot, 1)))
        return rc;
    k_compact_count<<<(unsigned)nt, kThreads, 0, c->stream>>>(pred, nq, tcnt);
    HIPCHK(c, hipGetLastError());
    if ((rc = scan_i64(c, s, tcnt, toff, nt, dtot)) || (rc = read_total(c, dtot, n)))
        return rc;
    if (*n > 0x7fffffffLL)
        return fail(c, TLM_ERR_UNSUPPORTED, "more than 2^31 faces");
    if ((rc = dalloc(c, pcells, (size_t)*n)) || (rc = dalloc(c, paxes, (size_t)*n)))
        return rc;
    k_compact_emit<<<(unsigned)nt, kThreads, 0, c->stream>>>(pred, nq, toff, c->nxy, c->Pp,
                                                             *pcells, *paxes);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLM_OK;
}

// ---- host side ----

int ready(tlm_ctx *c, const char *what)
{
    if (!c)
        return fail(nullptr, TLM_ERR_ARG, std::string(what) + ": null context");
    if (c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    HIPCHK(c, hipSetDevice(c->dev));
    if (!c->mat)
    {
        int rc = dalloc(c, &c->mat, (size_t)c->nx * c->ny * c->nz);
        if (rc)
            return rc;
        HIPCHK(c, hipMemsetAsync(c->mat, 0, (size_t)c->nx * c->ny * c->nz, c->stream));
    }
    return TLM_OK;
}

VGrid vgrid(const tlm_ctx *c, const tlm_voxel_grid *g)
{
    VGrid v;
    v.nx = c->nx;
    v.ny = c->ny;
    v.nz = c->nz;
    v.o = mk(g->origin[0], g->origin[1], g->origin[2]);
    v.dl = g->dl;
    return v;
}

// Coordinates whose cell index would leave the int range are refused (the
// reference's float-to-int conversion is not defined there).
bool coords_ok(const VGrid &g, F3 p)
{
    const float lim = 1.0e9f;
    for (int q = 0; q < 3; ++q)
    {
        const float v = (comp(p, q) - comp(g.o, q)) / g.dl;
        if (!(v > -lim && v < lim))
            return false;
    }
    return true;
}

bool grid_ok(const tlm_voxel_grid *g)
{
    return g && std::isfinite(g->dl) && g->dl > 0.f && std::isfinite(g->origin[0]) &&
           std::isfinite(g->origin[1]) && std::isfinite(g->origin[2]);
}

} // namespace

extern "C" {

int tlm_mat_fill(tlm_ctx *c, uint8_t id)
{
    int rc = ready(c, "tlm_mat_fill");
    if (rc)
        return rc;
    HIPCHK(c, hipMemsetAsync(c->mat, id, (size_t)c->nx * c->ny * c->nz, c->stream));
    return TLM_OK;
}

int tlm_mat_cells_box(tlm_ctx *c, const int32_t lo[3], const int32_t hi[3], uint8_t id)
{
    int rc = ready(c, "tlm_mat_cells_box");
    if (rc)
        return rc;
    if (!lo || !hi)
        return fail(c, TLM_ERR_ARG, "tlm_mat_cells_box: bad arguments");
    const int nn[3] = {c->nx, c->ny, c->nz};
    int a[3], b[3];
    for (int q = 0; q < 3; ++q)
    {
        a[q] = std::max(0, lo[q]);
        b[q] = std::min(nn[q], hi[q]);
        if (b[q] <= a[q])
            return TLM_OK;
    }
    const long long n = (long long)(b[0] - a[0]) * (b[1] - a[1]) * (b[2] - a[2]);
    VGrid g;
    g.nx = c->nx;
    g.ny = c->ny;
    g.nz = c->nz;
    k_cells_box<<<grid_for(n), kThreads, 0, c->stream>>>(c->mat, g, a[0], a[1], a[2], b[0] - a[0],
                                                          b[1] - a[1], n, id);
    HIPCHK(c, hipGetLastError());
    return TLM_OK;
}

int tlm_voxelize_mesh(tlm_ctx *c, const tlm_voxel_grid *gd, const float *verts, int32_t nverts,
                      const int32_t *idx, int32_t ntri, const float offset[3], uint8_t id,
                      int solid)
{
    int rc = ready(c, "tlm_voxelize_mesh");
    if (rc)
        return rc;
    if (!grid_ok(gd) || nverts < 0 || ntri < 0 || (nverts > 0 && !verts) || (ntri > 0 && !idx) ||
        !offset)
        return fail(c, TLM_ERR_ARG, "tlm_voxelize_mesh: bad arguments");
    const VGrid g = vgrid(c, gd);
    const F3 off = mk(offset[0], offset[1], offset[2]);
    for (long long e = 0; e < 3LL * ntri; ++e)
        if (idx[e] < 0 || idx[e] >= nverts)
            return fail(c, TLM_ERR_ARG, "triangle vertex index out of range");
    Box3 mb; // Geometry.cpp:559-563
    mb.reset();
    for (int v = 0; v < nverts; ++v)
    {
        const F3 p = add(mk(verts[3 * v], verts[3 * v + 1], verts[3 * v + 2]), off);
        if (!coords_ok(g, p))
            return fail(c, TLM_ERR_ARG, "vertex coordinates out of range");
        mb.grow(p);
    }
    Scratch s;
    float *dv, *tri;
    int *di, *bb;
    long long *cnt, *off_t, *dtot;
    if ((rc = s.get(c, &dv, 3 * (size_t)nverts)) || (rc = s.get(c, &di, 3 * (size_t)ntri)) ||
        (rc = s.get(c, &tri, 9 * (size_t)ntri)) || (rc = s.get(c, &bb, 6 * (size_t)ntri)) ||
        (rc = s.get(c, &cnt, ntri)) || (rc = s.get(c, &off_t, ntri)) || (rc = s.get(c, &dtot, 1)))
        return rc;
    if (nverts)
        HIPCHK(c, hipMemcpyAsync(dv, verts, 3 * (size_t)nverts * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream));
    if (ntri)
        HIPCHK(c, hipMemcpyAsync(di, idx, 3 * (size_t)ntri * sizeof(int),
                                 hipMemcpyHostToDevice, c->stream));
    if (ntri > 0)
    {
        // surface pass (Geometry.cpp:369-397)
        k_tri_prep<<<(ntri + kThreads - 1) / kThreads, kThreads, 0, c->stream>>>(
            dv, di, ntri, off, g, tri, bb, cnt);
        HIPCHK(c, hipGetLastError());
        long long total = 0;
        if ((rc = scan_i64(c, s, cnt, off_t, ntri, dtot)) || (rc = read_total(c, dtot, &total)))
            return rc;
        if (total > 0)
        {
            k_tri_cells<<<grid_for(total), kThreads, 0, c->stream>>>(tri, bb, off_t, ntri, total,
                                                                      g, c->mat, id);
            HIPCHK(c, hipGetLastError());
        }
    }
    if (!solid || !(mb.lo.x <= mb.hi.x))
    {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TLM_OK;
    }
    // parity fill along +x rows (Geometry.cpp:564-606)
    int i0, j0, k0, i1, j1, k1;
    g.cell_of(mb.lo, i0, j0, k0);
    g.cell_of(mb.hi, i1, j1, k1);
    i0 = std::max(i0, 0);
    j0 = std::max(j0, 0);
    k0 = std::max(k0, 0);
    i1 = std::min(i1, g.nx - 1);
    j1 = std::min(j1, g.ny - 1);
    k1 = std::min(k1, g.nz - 1);
    if (ntri < 2 || i1 < i0 || j1 < j0 || k1 < k0)
    {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return TLM_OK;
    }
    Rows rw;
    rw.i0 = i0;
    rw.i1 = i1;
    rw.j0 = j0;
    rw.k0 = k0;
    rw.nj = j1 - j0 + 1;
    rw.n = (long long)rw.nj * (k1 - k0 + 1);
    rw.xs = mb.lo.x - g.dl;
    rw.nh = 0;
    const unsigned nb = (unsigned)((rw.n + kThreads - 1) / kThreads);
    {
        // one pass with kRowCap hit slots per row; enough triangle chunks
        // that the grid fills the chip
        int *scnt, *oflow;
        float *slots;
        if ((rc = s.get(c, &scnt, rw.n)) || (rc = s.get(c, &oflow, 1)) ||
            (rc = s.get(c, &slots, (size_t)rw.n * kRowCap)))
            return rc;
        HIPCHK(c, hipMemsetAsync(scnt, 0, (size_t)rw.n * sizeof(int), c->stream));
        HIPCHK(c, hipMemsetAsync(oflow, 0, sizeof(int), c->stream));
        const int max_chunks = (ntri + kThreads - 1) / kThreads;
        const int nchunk = (int)std::max(1LL, std::min<long long>(max_chunks,
                                                                  (4096 + nb - 1) / nb));
        const int per_chunk = (ntri + nchunk - 1) / nchunk;
        k_ray_slots<<<dim3(nb, (unsigned)nchunk), kThreads, 0, c->stream>>>(tri, ntri, per_chunk,
                                                                           rw, g, scnt, slots);
        HIPCHK(c, hipGetLastError());
        k_ray_fill<<<nb, kThreads, 0, c->stream>>>(scnt, slots, rw, g, c->mat, id, oflow);
        HIPCHK(c, hipGetLastError());
        int over = 0;
        HIPCHK(c, hipMemcpyAsync(&over, oflow, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (!over)
            return TLM_OK;
    }
    // some row crossed the surface more than kRowCap times: the exact
    // two-pass fill (rows already filled are filled again with the same id)
    long long *rcnt, *roff, *htot;
    if ((rc = s.get(c, &rcnt, rw.n)) || (rc = s.get(c, &roff, rw.n)) || (rc = s.get(c, &htot, 1)))
        return rc;
    k_rows<false><<<nb, kThreads, 0, c->stream>>>(tri, ntri, rw, g, rcnt, nullptr, nullptr,
                                                  nullptr, id);
    HIPCHK(c, hipGetLastError());
    long long nh = 0;
    if ((rc = scan_i64(c, s, rcnt, roff, rw.n, htot)) || (rc = read_total(c, htot, &nh)))
        return rc;
    if (nh > 0)
    {
        float *hits;
        if ((rc = s.get(c, &hits, nh)))
            return rc;
        rw.nh = nh;
        k_rows<true><<<nb, kThreads, 0, c->stream>>>(tri, ntri, rw, g, nullptr, roff, hits,
                                                     c->mat, id);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLM_OK;
}

int tlm_voxelize_wire(tlm_ctx *c, const tlm_voxel_grid *gd, const float *pts, int32_t npts,
                      float radius, const float offset[3], uint8_t id)
{
    int rc = ready(c, "tlm_voxelize_wire");
    if (rc)
        return rc;
    if (!grid_ok(gd) || npts < 0 || (npts > 0 && !pts) || !offset || !std::isfinite(radius))
        return fail(c, TLM_ERR_ARG, "tlm_voxelize_wire: bad arguments");
    if (npts < 2)
        return TLM_OK;
    const VGrid g = vgrid(c, gd);
    const F3 off = mk(offset[0], offset[1], offset[2]);
    // per segment: ends, sample count (Geometry.cpp:635-640), first sample
    const int nseg = npts - 1;
    std::vector<float> seg(6 * (size_t)nseg);
    std::vector<int> steps(nseg);
    std::vector<long long> soff(nseg);
    long long total_s = 0;
    for (int s = 0; s < nseg; ++s)
    {
        const F3 a = add(mk(pts[3 * s], pts[3 * s + 1], pts[3 * s + 2]), off);
        const F3 b = add(mk(pts[3 * s + 3], pts[3 * s + 4], pts[3 * s + 5]), off);
        if (!coords_ok(g, a) || !coords_ok(g, b))
            return fail(c, TLM_ERR_ARG, "wire coordinates out of range");
        const F3 d = sub(b, a);
        const float len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        const float ns = std::ceil(len / (g.dl * 0.25f));
        if (!(ns < 1.0e8f))
            return fail(c, TLM_ERR_ARG, "wire segment too long for the cell size");
        steps[s] = std::max(1, (int)ns);
        float *o = &seg[6 * (size_t)s];
        o[0] = a.x, o[1] = a.y, o[2] = a.z, o[3] = b.x, o[4] = b.y, o[5] = b.z;
        soff[s] = total_s;
        total_s += steps[s] + 1;
    }
    int rcell = (int)std::ceil(radius / g.dl - 0.5f);
    if (rcell < 0)
        rcell = 0;
    if (rcell > 512)
        return fail(c, TLM_ERR_UNSUPPORTED, "wire radius above 512 cells");
    const long long D = 2LL * rcell + 1;
    const long long total = total_s * D * D * D;
    Scratch s;
    float *dseg;
    int *dsteps;
    long long *dsoff;
    if ((rc = s.get(c, &dseg, seg.size())) || (rc = s.get(c, &dsteps, steps.size())) ||
        (rc = s.get(c, &dsoff, soff.size())))
        return rc;
    HIPCHK(c, hipMemcpyAsync(dseg, seg.data(), seg.size() * sizeof(float), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(dsteps, steps.data(), steps.size() * sizeof(int),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dsoff, soff.data(), soff.size() * sizeof(long long),
                             hipMemcpyHostToDevice, c->stream));
    k_wire<<<grid_for(total), kThreads, 0, c->stream>>>(dseg, dsteps, dsoff, nseg, total, rcell,
                                                         radius, g, c->mat, id);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLM_OK;
}

int tlm_carve_feed(tlm_ctx *c, const tlm_voxel_grid *gd, const float a_[3], const float b_[3],
                   const float offset[3], int64_t *cells, int32_t cap, int32_t *pol)
{
    int rc = ready(c, "tlm_carve_feed");
    if (rc)
        return rc;
    if (!grid_ok(gd) || !a_ || !b_ || !offset || cap < 0 || (cap > 0 && !cells))
        return fail(c, TLM_ERR_ARG, "tlm_carve_feed: bad arguments");
    const VGrid g = vgrid(c, gd);
    const F3 off = mk(offset[0], offset[1], offset[2]);
    const F3 a = add(mk(a_[0], a_[1], a_[2]), off);
    const F3 b = add(mk(b_[0], b_[1], b_[2]), off);
    if (!coords_ok(g, a) || !coords_ok(g, b))
        return fail(c, TLM_ERR_ARG, "feed coordinates out of range");
    const F3 d = sub(b, a);
    const F3 a2 = add(a, mul(d, 0.15f)), b2 = sub(b, mul(d, 0.15f));
    const F3 e = sub(b2, a2);
    const float len = std::sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
    const int steps = std::max(1, (int)std::ceil(len / (g.dl * 0.25f)));
    std::vector<long long> out;
    for (int s = 0; s <= steps; ++s)
    {
        const F3 pp = add(a2, mul(e, (float)s / steps));
        int ci, cj, ck;
        g.cell_of(pp, ci, cj, ck);
        if (!g.in_grid(ci, cj, ck))
            continue;
        out.push_back(g.index(ci, cj, ck));
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    int p = 0;
    if (std::fabs(d.y) > std::fabs(comp(d, p)))
        p = 1;
    if (std::fabs(d.z) > std::fabs(comp(d, p)))
        p = 2;
    if (pol)
        *pol = p;
    if (!out.empty())
    {
        Scratch s;
        long long *dc;
        if ((rc = s.get(c, &dc, out.size())))
            return rc;
        HIPCHK(c, hipMemcpyAsync(dc, out.data(), out.size() * sizeof(long long),
                                 hipMemcpyHostToDevice, c->stream));
        k_set_cells<<<((int)out.size() + kThreads - 1) / kThreads, kThreads, 0, c->stream>>>(
            c->mat, dc, (int)out.size(), 0);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    for (size_t q = 0; q < out.size() && (int)q < cap; ++q)
        cells[q] = out[q];
    return (int)out.size();
}

int tlm_mat_read(tlm_ctx *c, uint8_t *mat)
{
    int rc = ready(c, "tlm_mat_read");
    if (rc)
        return rc;
    if (!mat)
        return fail(c, TLM_ERR_ARG, "tlm_mat_read: null output");
    HIPCHK(c, hipMemcpyAsync(mat, c->mat, (size_t)c->nx * c->ny * c->nz, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLM_OK;
}

int tlm_mat_write(tlm_ctx *c, const uint8_t *mat)
{
    int rc = ready(c, "tlm_mat_write");
    if (rc)
        return rc;
    if (!mat)
        return fail(c, TLM_ERR_ARG, "tlm_mat_write: null input");
    HIPCHK(c, hipMemcpyAsync(c->mat, mat, (size_t)c->nx * c->ny * c->nz, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return TLM_OK;
}

int tlm_upload_mat(tlm_ctx *c, const float *yhat, const float *ghat, int nmat)
{
    int rc = ready(c, "tlm_upload_mat");
    if (rc)
        return rc;
    if (nmat < 0 || (nmat > 0 && (!yhat || !ghat)))
        return fail(c, TLM_ERR_ARG, "tlm_upload_mat: bad arguments");
    const long long ncell = (long long)c->nx * c->ny * c->nz;
    Scratch s;
    int *dmax;
    if ((rc = s.get(c, &dmax, 1)))
        return rc;
    const int m1 = -1;
    HIPCHK(c, hipMemcpyAsync(dmax, &m1, sizeof(int), hipMemcpyHostToDevice, c->stream));
    k_max_diel<<<grid_for(ncell, 4096), kThreads, 0, c->stream>>>(c->mat, ncell, dmax);
    HIPCHK(c, hipGetLastError());
    int max_id = -1;
    HIPCHK(c, hipMemcpyAsync(&max_id, dmax, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (max_id >= nmat)
        return fail(c, TLM_ERR_ARG, "material id without a yhat/ghat entry");
    k_codes<<<(unsigned)(c->ny * c->nz + c->nz), 128, 0, c->stream>>>(c->mat, c->nx, c->ny, c->nz,
                                                                       c->nxy, c->Pp, c->code);
    HIPCHK(c, hipGetLastError());
    const long long nblk = c->P / tlm::kWave;
    // padded to a multiple of 4 bytes for the step's dword loads of the flags
    if ((rc = upload(c, &c->uair, (const uint8_t *)nullptr, (size_t)((nblk + 3) / 4 * 4))))
        return rc;
    c->fold_marked = false;
    k_uair<<<(unsigned)((nblk + kThreads - 1) / kThreads), kThreads, 0, c->stream>>>(c->code, nblk,
                                                                                     c->uair);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return install_material(c, max_id >= 0, yhat, ghat, nmat, nullptr, nullptr);
}

int64_t tlm_pec_faces_mat(tlm_ctx *c)
{
    int rc = ready(c, "tlm_pec_faces_mat");
    if (rc)
        return rc;
    long long *fc = nullptr;
    int *fa = nullptr;
    long long n = 0;
    {
        const long long ncell = (long long)c->nx * c->ny * c->nz;
        const long long nt = (ncell + kTile - 1) / kTile;
        Scratch s;
        long long *tcnt, *toff, *dtot;
        if ((rc = s.get(c, &tcnt, nt)) || (rc = s.get(c, &toff, nt)) || (rc = s.get(c, &dtot, 1)))
            return rc;
        k_pec_faces<false><<<(unsigned)nt, kThreads, 0, c->stream>>>(
            c->mat, c->nx, c->ny, c->nz, c->nxy, ncell, tcnt, nullptr, c->Pp, nullptr, nullptr);
        HIPCHK(c, hipGetLastError());
        if ((rc = scan_i64(c, s, tcnt, toff, nt, dtot)) || (rc = read_total(c, dtot, &n)))
            return rc;
        if (n > 0x7fffffffLL)
            return fail(c, TLM_ERR_UNSUPPORTED, "more than 2^31 faces");
        if ((rc = dalloc(c, &fc, (size_t)n)) || (rc = dalloc(c, &fa, (size_t)n)))
        {
            dfree(c, fc);
            dfree(c, fa);
            return rc;
        }
        k_pec_faces<true><<<(unsigned)nt, kThreads, 0, c->stream>>>(
            c->mat, c->nx, c->ny, c->nz, c->nxy, ncell, nullptr, toff, c->Pp, fc, fa);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    // install as tlm_set_faces does
    dfree(c, c->fcell);
    dfree(c, c->faxis);
    c->fcell = fc;
    c->faxis = fa;
    dfree(c, c->js_inst);
    dfree(c, c->js_dft4);
    if ((rc = dalloc(c, &c->js_inst, (size_t)std::max(1LL, n))) ||
        (rc = dalloc(c, &c->js_dft4, 4 * (size_t)std::max(1LL, n))))
        return rc;
    HIPCHK(c, hipMemset(c->js_inst, 0, std::max(1LL, n) * sizeof(float)));
    HIPCHK(c, hipMemset(c->js_dft4, 0, 4 * (size_t)std::max(1LL, n) * sizeof(float)));
    c->nf = (int)n;
    return n;
}

int64_t tlm_huygens_mat(tlm_ctx *c, int o)
{
    int rc = ready(c, "tlm_huygens_mat");
    if (rc)
        return rc;
    long long *hc = nullptr;
    int *ha = nullptr;
    long long n = 0;
    const bool on = o >= 2 && 2 * o + 2 < c->nx && 2 * o + 2 < c->ny && 2 * o + 2 < c->nz;
    if (on)
    {
        HuyFaces pred;
        pred.mat = c->mat;
        pred.nx = c->nx;
        pred.ny = c->ny;
        pred.x0 = o;
        pred.x1 = c->nx - 1 - o;
        pred.y0 = o;
        pred.y1 = c->ny - 1 - o;
        pred.z0 = o;
        pred.z1 = c->nz - 1 - o;
        const long long NX = pred.x1 - pred.x0 + 1, NY = pred.y1 - pred.y0 + 1,
                        NZ = pred.z1 - pred.z0 + 1;
        pred.nX = 2 * NZ * NY;
        pred.nY = 2 * NZ * NX;
        const long long nq = pred.nX + pred.nY + 2 * NY * NX;
        if ((rc = compact(c, pred, nq, &hc, &ha, &n)))
        {
            dfree(c, hc);
            dfree(c, ha);
            return rc;
        }
    }
    else if ((rc = dalloc(c, &hc, 1)) || (rc = dalloc(c, &ha, 1)))
        return rc;
    // install as tlm_set_huygens does
    dfree(c, c->hcell);
    dfree(c, c->haxis);
    c->hcell = hc;
    c->haxis = ha;
    dfree(c, c->huy8);
    if ((rc = dalloc(c, &c->huy8, 8 * (size_t)std::max(1LL, n))))
        return rc;
    HIPCHK(c, hipMemset(c->huy8, 0, 8 * (size_t)std::max(1LL, n) * sizeof(float)));
    c->nh = (int)n;
    return n;
}

static int64_t read_list(tlm_ctx *c, const long long *dcell, const int *daxis, int n,
                         int64_t *cells, int32_t *axes, int64_t cap)
{
    if (!c)
        return fail(nullptr, TLM_ERR_ARG, "null context");
    HIPCHK(c, hipSetDevice(c->dev));
    const long long m = std::min<long long>(n, std::max<int64_t>(cap, 0));
    if (m > 0)
    {
        std::vector<long long> pc((size_t)m);
        std::vector<int> pa((size_t)m);
        HIPCHK(c, hipMemcpy(pc.data(), dcell, (size_t)m * sizeof(long long), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(pa.data(), daxis, (size_t)m * sizeof(int), hipMemcpyDeviceToHost));
        for (long long q = 0; q < m; ++q)
        {
            const long long k = pc[q] / c->Pp;
            if (cells)
                cells[q] = k * c->nxy + (pc[q] - k * c->Pp);
            if (axes)
                axes[q] = pa[q];
        }
    }
    return n;
}

int64_t tlm_read_faces(tlm_ctx *c, int64_t *cells, int32_t *axes, int64_t cap)
{
    return read_list(c, c ? c->fcell : nullptr, c ? c->faxis : nullptr, c ? c->nf : 0, cells, axes,
                     cap);
}

int64_t tlm_read_huygens_faces(tlm_ctx *c, int64_t *cells, int32_t *axes, int64_t cap)
{
    return read_list(c, c ? c->hcell : nullptr, c ? c->haxis : nullptr, c ? c->nh : 0, cells, axes,
                     cap);
}

} // extern "C"
