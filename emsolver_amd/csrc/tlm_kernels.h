// tlm_kernels.h - device kernels of the TLM time step.
//
// Layout in HBM (set up in GpuTlmHip.hip). Cells are addressed in a
// slab-padded space: z-slab k starts at k*Pp with Pp = roundup(nx*ny, 64),
// so the cell (i, j, k) is k*Pp + j*nx + i and every wavefront's 64
// consecutive cells share one aligned block. The padding is never a
// neighbour.
//   * FP32 pulses: a slab is Pp/64 blocks of 3 KB, block b holding the 12
//     ports of its 64 cells port after port (port p of lane l at byte
//     p*256 + l*4 of the block), so the 24 loads and stores of a wave touch
//     two contiguous 3 KB runs.
//   * 16-bit pulses: 6 pair planes (ports 2q, 2q+1 of a cell side by side),
//     each slab stored as 16-bit codes (4 bytes per cell and pair) or, near
//     soft sources, as FP32 (8 bytes).
//   * two such pulse buffers (A = incident pulses of step n, B = of n+1).
//   * stubs: 3 planes of P floats, updated in place, dielectric cells only.
//   * cell code: 16 bits per cell address: material id in bits 0-7 (0 air,
//     1 PEC, >= 2 dielectric), then one bit per face telling whether the
//     neighbour across it is PEC (bit 8 x-, 9 x+, 10 y-, 11 y+, 12 z-,
//     13 z+; 0 on outer faces). Built once at upload, so a cell needs a
//     single 2-byte load and never reads its neighbours' materials.
//
// One step is (TlmSolver::step, TlmSolver.cpp:485-513):
//   k_inject  : soft sources add 0.5*amp*sv(n) into A        (:487-505)
//               (FP32 grids with few source cells: k_step adds them itself
//               where it loads the pulses, fold_sources)
//   k_step    : scatter every cell from A and deliver each reflected
//               pulse to the slot it reaches after connect, writing B
//               (scatter :516-548 and connect :551-613 fused); optionally
//               the FP64 sum of A^2 for the energy monitor (:766-789);
//               extra workgroups of the same launch take the Js /
//               Huygens / port / probe records from A (monitor_item,
//               :616-713), which the step itself only reads.
#pragma once

#include "scn_node.h"
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace tlm {

constexpr float ETA0 = 376.730313f; // TlmSolver.cpp:16
constexpr int kWave = 64;
constexpr int kNumXcd = 8;          // MI355X: 8 accelerator complex dies, one L2 each

// Pulse storage formats. FP32 keeps every slab in floats; FP16 / BF16 store
// each z-slab either in floats (slabs near soft sources) or as 16-bit codes
// times a per-slab power of two 2^E (decode: code * 2^E, encode: round to
// nearest even of v * 2^-E). The scatter arithmetic is FP32 throughout.
enum : int
{
    PF_F32 = 0,
    PF_F16 = 1,
    PF_BF16 = 2
};

template <int PF> __device__ __forceinline__ float dec16(unsigned h)
{
    if (PF == PF_F16)
        return (float)__builtin_bit_cast(_Float16, (unsigned short)h);
    return __uint_as_float(h << 16);
}
template <int PF> __device__ __forceinline__ unsigned short enc16(float v)
{
    if (PF == PF_F16)
        return __builtin_bit_cast(unsigned short, (_Float16)v); // v_cvt_f16_f32, round to nearest even
    unsigned u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u); // bfloat16 round to nearest even
    return (unsigned short)(u >> 16);
}

// Soft sources folded into the step (FP32 contexts with at most
// kFoldMax source cells): the step and its monitor workgroups add step n's
// sources to the incident pulses they load, in registers, instead of a
// separate injection launch writing them into A first. ninj = 0: none.
constexpr int kFoldMax = 32;
struct Fold
{
    const long long *cell = nullptr; // distinct source cells (padded index), ascending
    const int *beg = nullptr;        // sources of cell u: [beg[u], beg[u+1]), list order
    const int *pol = nullptr;
    const float *amp = nullptr;
    int ninj = 0;
    float sv = 0.0f;                 // waveform value of the step
};

// Adds the folded sources of cell c to its 12 incident pulses v, in list
// order: the same float additions k_inject would store (TlmSolver.cpp:
// 487-505), so every later use sees the same values.
__device__ __forceinline__ void fold_sources(const Fold &f, long long c, float v[12])
{
    for (int u = 0; u < f.ninj; ++u)
    {
        if (f.cell[u] != c)
            continue;
        for (int s = f.beg[u]; s < f.beg[u + 1]; ++s)
        {
            const int pol = f.pol[s];
            const float qv = 0.5f * f.amp[s] * f.sv;
#pragma unroll
            for (int pp = 0; pp < 3; ++pp)
            {
                if (pp != pol)
                    continue;
#pragma unroll
                for (int axis = 0; axis < 3; ++axis)
                {
                    if (axis == pp)
                        continue;
                    v[scn::port_of(axis, 0, pp)] = v[scn::port_of(axis, 0, pp)] + qv;
                    v[scn::port_of(axis, 1, pp)] = v[scn::port_of(axis, 1, pp)] + qv;
                }
            }
        }
    }
}

// Division by a launch constant d (1 <= d < 2^31) of any x < 2^31 as a
// multiply-high (Granlund and Montgomery): with l = ceil(log2 d) and
// m = ceil(2^(31+l) / d) < 2^32, floor(x / d) = (x * m) >> (31 + l). The
// wave's task and cell indices are split this way, with a few scalar or
// vector instructions instead of a division sequence.
struct Magic
{
    unsigned m;
    int s;
};
__host__ inline Magic magic_div(unsigned d)
{
    int l = 0;
    while ((1ull << l) < d)
        ++l;
    const unsigned long long num = 1ull << (31 + l);
    return {(unsigned)((num + d - 1) / d), 31 + l};
}
__device__ __forceinline__ unsigned mdiv(unsigned x, Magic q)
{
    return (unsigned)(((unsigned long long)x * q.m) >> q.s);
}

// Workgroup -> task-group mapping of a launch of G step workgroups over
// `groups` task groups, prepared on the host (see xcd_range below):
// nr ranges (one per XCD, or G if fewer), per / extra groups per range,
// gq / gr workgroups per range.
struct XcdMap
{
    int nr, per, extra, gq, gr;
    // chunk > 0: XCD x takes chunks x, x + 8, x + 16, ... of `chunk` groups
    // each (a z-slab's groups, or a z-chunk's for the 16-bit march), so
    // every XCD gets an even sample of the grid's depth; the launch then
    // has 8 * cpx workgroups, one group each (those past `groups` idle)
    int chunk, cpx, groups;
    Magic mchunk;
};
__host__ inline XcdMap xcd_map(int G, int groups)
{
    const int nr = G < 8 ? G : 8;
    return {nr, groups / nr, groups % nr, G / nr, G % nr, 0, 0, groups, {0u, 0}};
}
// The interleaved map for a launch of one workgroup per group: returns the
// map; its launch size is 8 * cpx.
__host__ inline XcdMap xcd_chunk_map(int groups, int chunk)
{
    chunk = chunk < 1 ? 1 : chunk;
    const int nch = (groups + chunk - 1) / chunk;
    const int cpx = (nch + 7) / 8 * chunk;
    XcdMap m = xcd_map(8 * cpx, groups);
    m.chunk = chunk;
    m.cpx = cpx;
    m.mchunk = magic_div((unsigned)chunk);
    return m;
}

struct StepArgs
{
    const char *A;        // incident pulses of step n: 12 planes of PB bytes
    char *B;              // incident pulses of step n+1
    float *Vs;            // stubs: 3P floats in 64-cell blocks (stub_block)
    const uint16_t *code; // cell code (material + neighbour-PEC bits)
    const uint8_t *uair;  // per 64-cell block (slab-major): bit 0 all its codes are 0,
                          // bit 1 it holds a folded source cell, bit 2 a dielectric
                          // cell, bits 3 / 4 a dielectric cell among cells 0-31 /
                          // 32-63, bit 5 (kBlockClean) no PEC cell, no padding and no
                          // PEC neighbour bit; padded to a multiple of 4 bytes
                          // (block_flag)
    const float4 *props;  // per dielectric id: (yh, gh, den, 0)
    double *epart;        // energy partial per wave task
    // 16-bit formats only
    const uint2 *slab;    // per slab: (byte offset inside a plane, 1 if 16-bit storage)
    const float *dsc;     // per slab: 2^E_n (decode A)
    const float *esc;     // per slab: 2^-E_{n+1} (encode B)
    float *mpart;         // per (slab, strip): largest |value| the wave's cells send
    int nx, ny, nz;
    int nxy;             // nx*ny
    long long Pp;        // slab pitch: roundup(nx*ny, 64) cells
    long long P;         // cells per plane = Pp*nz (stub planes)
    long long PB;        // bytes per pulse plane
    int h0, h1;          // 16-bit formats: slabs [h0, h1) are stored in FP32
    int strips;          // 64-cell strips per slab: Pp/64
    int zc, nzc;         // planes per task, number of z chunks
    int tasks;           // strips * nzc
    int groups;          // task groups (one per workgroup's worth of waves)
    float rho;
    int nmat;           // dielectric materials (props entries)
    Magic mstrips, mnx; // division by strips and by nx
    XcdMap xm;          // k_step's workgroup -> task-group mapping
    Fold fold; // FP32: soft sources added where the pulses are loaded
};

// Loads and stores through a uniform plane base (SGPR pair) and a 32-bit
// byte offset per lane (one VGPR), the global_load saddr+voffset form.
__device__ __forceinline__ float ldo(const char *base, unsigned boff)
{
    return *reinterpret_cast<const float *>(base + boff);
}
template <bool NT> __device__ __forceinline__ void sto(char *base, unsigned boff, float v)
{
    float *p = reinterpret_cast<float *>(base + boff);
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
__device__ __forceinline__ unsigned ldc(const uint16_t *code, unsigned off)
{
    return code[off];
}
enum : unsigned
{
    NB_XM = 1u << 8,
    NB_XP = 1u << 9,
    NB_YM = 1u << 10,
    NB_YP = 1u << 11,
    NB_ZM = 1u << 12,
    NB_ZP = 1u << 13
};

// Block flag bit 5: every cell code of the 64-cell block has a material id
// other than PEC (1) and no neighbour-PEC bit. The slab padding is coded PEC,
// so such a block has all 64 cells inside the grid.
constexpr unsigned kBlockClean = 1u << 5;
__host__ __device__ inline bool code_is_clean(unsigned c)
{
    return (c & 0xFFu) != 1u && (c & 0x3F00u) == 0u;
}

// Flag byte b of the per-block flag array, read as the aligned dword holding
// it through the constant address space: the index is wave-uniform and the
// array is not written while a step runs, so this is one scalar load whose
// result feeds a scalar branch (the array is padded to a multiple of 4
// bytes). A byte load would be a vector load, and the wave would wait on the
// vector memory counter for it before it can issue the cell-code load. With
// the scalar load the 512^3 steps run 1 % faster (2.762 against 2.795 ms
// with 40 % stub cells, 2.268 against 2.288 ms in plain air).
__device__ __forceinline__ unsigned block_flag(const uint8_t *flags, size_t b)
{
    typedef const unsigned __attribute__((address_space(4))) *cword;
    const unsigned w = *(cword)(const void *)(flags + (b & ~(size_t)3));
    return (w >> ((unsigned)(b & 3u) * 8u)) & 0xFFu;
}

// A float at a wave-uniform address that no kernel of the step writes (the
// per-slab scales), read through the constant address space so that it is a
// scalar load: a vector load there would make the wave wait for all of its
// earlier stores before it could use the value (stores share the vector
// memory counter).
__device__ __forceinline__ float ld_uniform(const float *p)
{
    return *(const float __attribute__((address_space(4))) *)(const void *)p;
}

// Workgroups are dispatched to the 8 XCDs round robin. Each XCD gets one
// contiguous range of task groups, walked by its own workgroups in a
// grid-stride loop, so neighbouring strips share an L2 (which keeps the
// cache lines written from both sides of a strip seam on one die). With as
// many workgroups as groups this is a plain renumbering; with fewer it is a
// persistent grid.
struct XcdRange
{
    int first, end, step;
};
// The same from the host-prepared map: no division when there are 8 ranges.
__device__ __forceinline__ XcdRange xcd_range_map(int b, const XcdMap &m)
{
    if (m.chunk > 0)
    {
        // workgroup b on XCD b % 8 takes local group l = b / 8 of that XCD:
        // chunk q * 8 + x, offset l - q * chunk
        const int x = b & (kNumXcd - 1), l = b >> 3;
        const int q = (int)mdiv((unsigned)l, m.mchunk);
        const int g = (q * kNumXcd + x) * m.chunk + (l - q * m.chunk);
        return {g, g < m.groups ? g + 1 : g, 1};
    }
    int xcd, local;
    if (m.nr == kNumXcd)
    {
        xcd = b & (kNumXcd - 1);
        local = b >> 3;
    }
    else
    {
        xcd = b % m.nr;
        local = b / m.nr;
    }
    const int gper = m.gq + (xcd < m.gr ? 1 : 0);
    const int start = xcd * m.per + min(xcd, m.extra);
    const int count = m.per + (xcd < m.extra ? 1 : 0);
    return {start + local, start + count, gper};
}
__device__ __forceinline__ XcdRange xcd_range(int b, int G, int ngroups)
{
    // ranges = min(8, G) so that every range has at least one workgroup
    const int nr = min(kNumXcd, G);
    const int xcd = b % nr, local = b / nr;
    const int gper = G / nr + (xcd < G % nr ? 1 : 0); // workgroups on this range
    const int per = ngroups / nr, extra = ngroups % nr;
    const int start = xcd * per + min(xcd, extra);
    const int count = per + (xcd < extra ? 1 : 0);
    return {start + local, start + count, gper};
}

// FP32 pulse layout: each z-slab is a run of 64-cell blocks, a block holding
// the 12 ports of its 64 cells one after another (port p of block lane l at
// p*256 + l*4 bytes, 3 KB per block). A wavefront's 12 loads and 12 stores
// for its strip touch one contiguous 3 KB block, and a slab is Pp*48 bytes.
// Byte offset of in-slab cell g inside its slab:
__device__ __forceinline__ unsigned blk_off(unsigned g)
{
    return (g >> 6) * 3072u + (g & 63u) * 4u;
}

// Stubs: 64-cell blocks like the pulses, block b (slab-padded cell index
// >> 6) holding its cells' stubs 0, 1, 2 one after another (stub j of block
// lane l at float b*192 + j*64 + l), so a wave's stubs are one 768-byte run.
constexpr int kStubBlock = 3 * kWave;
__device__ __forceinline__ float *stub_block(float *Vs, long long block)
{
    return Vs + (size_t)block * kStubBlock;
}

// Lane constants of a strip task, kept in bits 16.. of the cell-code word
// (the code itself takes bits 0..13).
enum : unsigned
{
    LF_BX0 = 1u << 16, // i == 0: x- face on the outer boundary
    LF_BX1 = 1u << 17, // i == nx-1
    LF_BY0 = 1u << 18, // j == 0
    LF_BY1 = 1u << 19, // j == ny-1
    LF_L = 1u << 20,   // x- neighbour in this wave (its pulse comes by shuffle)
    LF_R = 1u << 21    // x+ neighbour in this wave
};
// a | b where the compiler cannot see that b is loop-invariant: the lane
// flags then enter every slab's tests through the loaded code word, and the
// masks derived from them live for one slab instead of being held in
// (spilled) scalar registers across the z-march.
__device__ __forceinline__ unsigned or_opaque(unsigned a, unsigned b)
{
    unsigned r;
    asm("v_or_b32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Sum of the squares of a cell's 12 pulses in FP64 (each square is exact in
// double): four independent partial sums, so the additions do not form one
// dependent chain of twelve.
__device__ __forceinline__ double cell_energy(const float v[12])
{
    double e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        e[q] = (double)v[q] * (double)v[q];
#pragma unroll
    for (int p = 4; p < 12; ++p)
        e[p & 3] = fma((double)v[p], (double)v[p], e[p & 3]);
    return (e[0] + e[1]) + (e[2] + e[3]);
}

// Fused scatter + connect. A wavefront owns a strip of 64 consecutive cells
// of one z-slab (lane = cell) and marches over zc slabs.
//
// Every slot of B is written exactly once:
//   x faces: by its owner, with the neighbour's reflected pulse taken from
//            the neighbouring lane; where a row continues into the next
//            strip, the two seam cells push into each other's slot.
//   y faces: the sender pushes its reflected pulse into the neighbour row's
//            slot.
//   z faces: by the owner inside the z-march (the previous slab's z+
//            pulses stay in registers and its z+ slots are stored one slab
//            late); across z-chunk edges the sender pushes.
// The owner writes its own slot instead whenever the link is broken: own
// cell PEC (value kept, times rho on an outer face), outer face (rho * out),
// neighbour PEC (-out). This is connect(), TlmSolver.cpp:563-583.
// One wave task: a 64-cell strip of slabs [k0, k1).
template <bool STUBS, bool ENERGY, bool NT, int PF, bool FOLD = false>
__device__ __forceinline__ void strip_task(const StepArgs &a, int task, int lane, double &esum)
{
    static_assert(PF == PF_F32, "the 16-bit formats use strip_task16");
    const int tz = (int)mdiv((unsigned)task, a.mstrips), t = task - tz * a.strips;
    const int f = t * kWave + lane; // cell inside the slab
    const bool act = f < a.nxy;
    const int j = (int)mdiv((unsigned)f, a.mnx), i = f - j * a.nx;
    const int k0 = tz * a.zc, k1 = min(a.nz, k0 + a.zc);
    const float rho = a.rho;
    const size_t slabB = (size_t)a.Pp * 48u;
    // in-slab byte offsets of this cell and of its x / y neighbours
    const unsigned lo = (unsigned)t * 3072u + (unsigned)lane * 4u;
    const unsigned loL = blk_off((unsigned)f - 1u), loR = blk_off((unsigned)f + 1u);
    const unsigned loYm = blk_off((unsigned)(f - a.nx)), loYp = blk_off((unsigned)(f + a.nx));
    // x neighbour inside the wave; otherwise (strip seam) the two cells push
    const bool hasL = act && lane > 0 && i > 0;
    const bool hasR = act && lane < kWave - 1 && i < a.nx - 1;
    const bool bx0 = i == 0, bx1 = i == a.nx - 1, by0 = j == 0, by1 = j == a.ny - 1;

    float zpo0 = 0.f, zpo1 = 0.f; // previous slab: reflected z+ pulses
    bool pecPrev = false;
    unsigned mo = (unsigned)(k0 * a.Pp + f); // cell index in the slab-padded space (codes, stubs)
    const char *Ak = a.A + (size_t)k0 * slabB;
    char *Bk = a.B + (size_t)k0 * slabB;

    for (int k = k0; k < k1; ++k, mo += (unsigned)a.Pp, Ak += slabB, Bk += slabB)
    {
        float o[12];
        bool pec = true;
        unsigned code = 0;
        // blocks of plain air cells with no PEC neighbour skip the code load
        // bit 0: plain-air block (no code load); bit 1: block holds a folded source
        const unsigned fl = block_flag(a.uair, (size_t)k * a.strips + t);
        const bool uair = (fl & 1u) != 0u;
        const bool dblk = STUBS && (fl & 4u) != 0u;
        // the lane's half of the block holds a dielectric cell: its stub
        // lines are loaded and stored whole; an air half's are not touched
        const bool dhalf = STUBS && ((fl >> (3 + (lane >> 5))) & 1u) != 0u;
        const bool vtab = dblk && a.nmat <= kWave;
        float tyh = 0.f, tden = 4.f;
        if (vtab && lane < a.nmat)
        {
            const float4 pr = a.props[lane];
            tyh = pr.x;
            tden = pr.z;
        }
        float v[12];
        float st[3] = {0.f, 0.f, 0.f};
        float *S0 = stub_block(a.Vs, (long long)k * a.strips + t), *S1 = S0 + kWave,
              *S2 = S0 + 2 * kWave;
        const unsigned bs = (unsigned)lane * 4u;
        if (act)
        {
            if (!STUBS && !uair)
                code = ldc(a.code, mo);
#pragma unroll
            for (int p = 0; p < 12; ++p)
                v[p] = ldo(Ak + p * 256, lo);
            if (STUBS && !uair)
                code = ldc(a.code, mo);
            if (dhalf)
            {
                st[0] = ldo((const char *)S0, bs);
                st[1] = ldo((const char *)S1, bs);
                st[2] = ldo((const char *)S2, bs);
            }
        }
        const unsigned m = code & 0xFFu;
        float yhv = 0.f, denv = 4.f;
        if (vtab)
        {
            const int idx = m >= 2u ? (int)(m - 2u) : 0;
            yhv = __shfl(tyh, idx, kWave);
            denv = __shfl(tden, idx, kWave);
        }
        if (act)
        {
            if (FOLD && (fl & 2u))
                fold_sources(a.fold, (long long)mo, v);
            pec = (m == 1u);
            if (ENERGY)
                esum += cell_energy(v);
            if (STUBS && m >= 2u)
            {
                if (!vtab)
                {
                    const float4 pr = a.props[m - 2u];
                    yhv = pr.x;
                    denv = pr.z;
                }
                scn::scatter_diel(v, st, yhv, denv, o);
            }
            else
                scn::scatter_air(v, o);
            if (dhalf)
            {
                // every lane of a half holding a dielectric cell stores its
                // stubs (unchanged outside dielectric cells): the stub lines
                // are written whole
                sto<false>((char *)S0, bs, st[0]);
                sto<false>((char *)S1, bs, st[1]);
                sto<false>((char *)S2, bs, st[2]);
            }
            if (pec)
            {
                // a PEC cell keeps its incident pulses: o carries them
#pragma unroll
                for (int p = 0; p < 12; ++p)
                    o[p] = v[p];
            }
        }
        else
        {
#pragma unroll
            for (int p = 0; p < 12; ++p)
                o[p] = 0.f;
        }
        // every lane takes part in the shuffles
        const float L0 = __shfl_up(o[2], 1, kWave), L1 = __shfl_up(o[3], 1, kWave);
        const float R0 = __shfl_down(o[0], 1, kWave), R1 = __shfl_down(o[1], 1, kWave);
        if (act)
        {
            // ---- x faces: ports 0,1 (x-) and 2,3 (x+) ----
            {
                const bool pecL = (code & NB_XM) != 0u, pecR = (code & NB_XP) != 0u;
#pragma unroll
                for (int pol = 0; pol < 2; ++pol)
                {
                    const float oo = o[0 + pol];
                    if (pec)
                        sto<NT>(Bk + (0 + pol) * 256, lo, bx0 ? rho * oo : oo);
                    else if (bx0)
                        sto<NT>(Bk + (0 + pol) * 256, lo, rho * oo);
                    else if (pecL)
                        sto<NT>(Bk + (0 + pol) * 256, lo, -oo);
                    else if (hasL)
                        sto<NT>(Bk + (0 + pol) * 256, lo, pol ? L1 : L0);
                    else // seam with a live neighbour: it fills my slot, I fill its
                        sto<NT>(Bk + (2 + pol) * 256, loL, oo);

                    const float op = o[2 + pol];
                    if (pec)
                        sto<NT>(Bk + (2 + pol) * 256, lo, bx1 ? rho * op : op);
                    else if (bx1)
                        sto<NT>(Bk + (2 + pol) * 256, lo, rho * op);
                    else if (pecR)
                        sto<NT>(Bk + (2 + pol) * 256, lo, -op);
                    else if (hasR)
                        sto<NT>(Bk + (2 + pol) * 256, lo, pol ? R1 : R0);
                    else
                        sto<NT>(Bk + (0 + pol) * 256, loR, op);
                }
            }
            // ---- y faces: ports 4,5 (y-) and 6,7 (y+), sender push ----
            {
                const bool pecLo = (code & NB_YM) != 0u, pecHi = (code & NB_YP) != 0u;
#pragma unroll
                for (int pol = 0; pol < 2; ++pol)
                {
                    const float ol = o[4 + pol], oh = o[6 + pol];
                    if (pec)
                    {
                        sto<NT>(Bk + (4 + pol) * 256, lo, by0 ? rho * ol : ol);
                        sto<NT>(Bk + (6 + pol) * 256, lo, by1 ? rho * oh : oh);
                    }
                    else
                    {
                        if (by0)
                            sto<NT>(Bk + (4 + pol) * 256, lo, rho * ol);
                        else if (pecLo)
                            sto<NT>(Bk + (4 + pol) * 256, lo, -ol);
                        else
                            sto<NT>(Bk + (6 + pol) * 256, loYm, ol);
                        if (by1)
                            sto<NT>(Bk + (6 + pol) * 256, lo, rho * oh);
                        else if (pecHi)
                            sto<NT>(Bk + (6 + pol) * 256, lo, -oh);
                        else
                            sto<NT>(Bk + (4 + pol) * 256, loYp, oh);
                    }
                }
            }
            // ---- z faces: ports 8,9 (z-) and 10,11 (z+) ----
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                const float ol = o[8 + pol];
                if (k > k0)
                {
                    // pull: slab k-1 is in registers
                    const float zp = pol ? zpo1 : zpo0;
                    if (pec)
                        sto<NT>(Bk + (8 + pol) * 256, lo, ol);
                    else
                        sto<NT>(Bk + (8 + pol) * 256, lo, pecPrev ? -ol : zp);
                    // slab k-1's z+ slot: its kept pulse if it is PEC (zp holds
                    // it), else what this cell sends down
                    sto<NT>(Bk - slabB + (10 + pol) * 256, lo, pecPrev ? zp : (pec ? -zp : ol));
                }
                else if (k == 0)
                    sto<NT>(Bk + (8 + pol) * 256, lo, rho * ol);
                else
                {
                    const bool pecZ = (code & NB_ZM) != 0u;
                    if (pec)
                        sto<NT>(Bk + (8 + pol) * 256, lo, ol);
                    else if (pecZ)
                        sto<NT>(Bk + (8 + pol) * 256, lo, -ol);
                    else
                        sto<NT>(Bk - slabB + (10 + pol) * 256, lo, ol);
                }
                if (k == k1 - 1)
                {
                    const float oh = o[10 + pol];
                    if (k == a.nz - 1)
                        sto<NT>(Bk + (10 + pol) * 256, lo, rho * oh);
                    else
                    {
                        const bool pecZ = (code & NB_ZP) != 0u;
                        if (pec)
                            sto<NT>(Bk + (10 + pol) * 256, lo, oh);
                        else if (pecZ)
                            sto<NT>(Bk + (10 + pol) * 256, lo, -oh);
                        else
                            sto<NT>(Bk + slabB + (8 + pol) * 256, lo, oh);
                    }
                }
            }
        }
        zpo0 = o[10];
        zpo1 = o[11];
        pecPrev = pec;
    }
}

// ---------------------------------------------------------------------------
// 16-bit formats. Pulses are grouped in 6 pairs, pair q holding ports 2q and
// 2q+1 (the two polarisations of one face). A slab is a run of 64-cell
// blocks, a block holding its cells' 6 pairs pair after pair: in a 16-bit
// slab a pair is one dword (two codes, 1.5 KB per block), in an FP32 slab a
// float2 (3 KB per block). The FP32 slabs form one band [h0, h1), so slab k
// starts at Pp*24*(k + FP32 slabs below k) bytes. Every store of the step
// writes both polarisations of one face of one cell with the same value
// rule, so a lane moves one dword (or 8 bytes) per face.
// ---------------------------------------------------------------------------
struct Slab16
{
    size_t base;     // byte offset of the slab in the buffer
    unsigned lo, es; // lane's cell offset inside the slab, bytes per cell and pair
    bool h16;
    float ds, isc;
};
// Byte offset of slab k in a 16-bit buffer (FP32 slabs [h0, h1) take twice
// the bytes of a 16-bit slab).
__host__ __device__ inline size_t slab16_base(long long Pp, long long k, int h0, int h1)
{
    const long long hot = k < h0 ? 0 : (k < h1 ? k - h0 : h1 - h0);
    return (size_t)Pp * 24u * (size_t)(k + hot);
}
// Offset of in-slab cell g for es bytes per pair: 64-cell blocks of 6 pairs.
__device__ __forceinline__ unsigned blk16(unsigned g, unsigned es)
{
    return (g >> 6) * (384u * es) + (g & 63u) * es;
}
__device__ __forceinline__ Slab16 slab16(const StepArgs &a, int k, int f)
{
    // offset and format are arithmetic, so the pulse loads never wait for a
    // table lookup
    const bool h = __builtin_amdgcn_readfirstlane((int)(k < a.h0 || k >= a.h1)) != 0; // per wave
    const unsigned es = h ? 4u : 8u;
    const size_t base = slab16_base(a.Pp, k, a.h0, a.h1);
    return {base, blk16((unsigned)f, es), es, h, ld_uniform(a.dsc + k), ld_uniform(a.esc + k)};
}
template <int PF, bool H16>
__device__ __forceinline__ void ldpair(const char *p, unsigned off, const Slab16 &s, float &x,
                                       float &y)
{
    if (H16)
    {
        const unsigned u = *reinterpret_cast<const unsigned *>(p + off);
        x = dec16<PF>(u & 0xffffu) * s.ds;
        y = dec16<PF>(u >> 16) * s.ds;
    }
    else
    {
        const float2 w = *reinterpret_cast<const float2 *>(p + off);
        x = w.x;
        y = w.y;
    }
}
template <int PF, bool H16>
__device__ __forceinline__ void stpair(char *p, unsigned off, const Slab16 &s, float x, float y)
{
    if (H16)
        *reinterpret_cast<unsigned *>(p + off) =
            (unsigned)enc16<PF>(x * s.isc) | ((unsigned)enc16<PF>(y * s.isc) << 16);
    else
        *reinterpret_cast<float2 *>(p + off) = make_float2(x, y);
}
// pair q of another slab's cell (format known at run time, wave-uniform)
template <int PF>
__device__ __forceinline__ void stpair_at(char *buf, const Slab16 &s, int q, float x, float y)
{
    char *p = buf + s.base + (size_t)q * 64u * s.es;
    if (s.h16)
        stpair<PF, true>(p, s.lo, s, x, y);
    else
        stpair<PF, false>(p, s.lo, s, x, y);
}

// The fused step of strip_task with pair blocks and per-slab formats (same
// delivery rules, same FP32 arithmetic, values rounded where stored). Every
// per-lane choice of value or slot is a select, not a branch.
template <bool STUBS, bool ENERGY, int PF>
__device__ __forceinline__ void strip_task16(const StepArgs &a, int task, int lane, double &esum)
{
    const int tz = (int)mdiv((unsigned)task, a.mstrips), t = task - tz * a.strips;
    const int f = t * kWave + lane;
    const bool act = f < a.nxy;
    const int j = (int)mdiv((unsigned)f, a.mnx), i = f - j * a.nx;
    const int k0 = tz * a.zc, k1 = min(a.nz, k0 + a.zc);
    const float rho = a.rho;
    const unsigned lf = (i == 0 ? LF_BX0 : 0u) | (i == a.nx - 1 ? LF_BX1 : 0u) |
                        (j == 0 ? LF_BY0 : 0u) | (j == a.ny - 1 ? LF_BY1 : 0u) |
                        (act && lane > 0 && i > 0 ? LF_L : 0u) |
                        (act && lane < kWave - 1 && i < a.nx - 1 ? LF_R : 0u);
    float zpo0 = 0.f, zpo1 = 0.f;
    bool pecPrev = false;
    unsigned mo = (unsigned)(k0 * a.Pp + f);
    Slab16 sp = {0, 0u, 8u, false, 1.0f, 1.0f};
    // material table in two VGPRs of every lane (<= 64 materials), loaded once
    // for the task's slabs: a dielectric cell picks its (yh, den) with a lane
    // shuffle instead of a dependent load
    const bool vtab = STUBS && a.nmat <= kWave;
    float tyh = 0.f, tden = 4.f;
    if (vtab && lane < a.nmat)
    {
        const float4 pr = a.props[lane];
        tyh = pr.x;
        tden = pr.z;
    }

    for (int k = k0; k < k1; ++k, mo += (unsigned)a.Pp)
    {
        const Slab16 s = slab16(a, k, f);
        const unsigned fl = block_flag(a.uair, (size_t)k * a.strips + t);
        // a block holding a dielectric cell loads its stubs with its pulses
        const bool dblk = STUBS && (fl & 4u) != 0u;
        // the lane's half of the block holds a dielectric cell: its stub
        // lines are loaded and stored whole; an air half's are not touched
        const bool dhalf = STUBS && ((fl >> (3 + (lane >> 5))) & 1u) != 0u;
        auto body = [&](auto h16_tag, auto clean_tag) {
            constexpr bool H = decltype(h16_tag)::value;
            // a clean block (kBlockClean): every lane holds a grid cell, none
            // is PEC and none has a PEC neighbour, so the PEC rules of every
            // face drop out (the same values and slots as the general rules
            // give such a block)
            constexpr bool CL = decltype(clean_tag)::value;
            const bool live = CL || act;
            constexpr unsigned ES = H ? 4u : 8u;
            constexpr unsigned PQ = 64u * ES; // bytes between pairs of a block
            const char *Ak = a.A + s.base;
            char *Bk = a.B + s.base;
            const unsigned bo = s.lo;
            float o[12];
            float mc = 0.f;
            unsigned code = 0;
            float v[12];
            float st[3] = {0.f, 0.f, 0.f};
            float *S0 = stub_block(a.Vs, (long long)k * a.strips + t), *S1 = S0 + kWave,
                  *S2 = S0 + 2 * kWave;
            const unsigned bs = (unsigned)lane * 4u;
            if (live)
            {
#pragma unroll
                for (int q = 0; q < 6; ++q)
                    ldpair<PF, H>(Ak + q * PQ, bo, s, v[2 * q], v[2 * q + 1]);
                code = ldc(a.code, mo);
                if (dhalf)
                {
                    st[0] = ldo((const char *)S0, bs);
                    st[1] = ldo((const char *)S1, bs);
                    st[2] = ldo((const char *)S2, bs);
                }
            }
            const unsigned cf = or_opaque(code, lf);
            const unsigned m = cf & 0xFFu;
            const bool pec = !CL && (!act || m == 1u);
            float yhv = 0.f, denv = 4.f;
            if (vtab && dblk)
            {
                const int idx = m >= 2u ? (int)(m - 2u) : 0;
                yhv = __shfl(tyh, idx, kWave);
                denv = __shfl(tden, idx, kWave);
            }
            if (live)
            {
                if (ENERGY)
                    esum += cell_energy(v);
                if (STUBS && m >= 2u)
                {
                    if (!vtab)
                    {
                        const float4 pr = a.props[m - 2u];
                        yhv = pr.x;
                        denv = pr.z;
                    }
                    scn::scatter_diel(v, st, yhv, denv, o);
                    mc = fmaxf(fmaxf(fabsf(st[0]), fabsf(st[1])), fabsf(st[2]));
                }
                else
                    scn::scatter_air(v, o);
                if (dhalf)
                {
                    // every lane of a half holding a dielectric cell stores its
                    // stubs (unchanged outside dielectric cells): the stub
                    // lines are written whole
                    sto<false>((char *)S0, bs, st[0]);
                    sto<false>((char *)S1, bs, st[1]);
                    sto<false>((char *)S2, bs, st[2]);
                }
            }
            else
            {
#pragma unroll
                for (int p = 0; p < 12; ++p)
                    v[p] = 0.f;
            }
            // a PEC cell keeps its incident pulses: o carries them (an idle
            // lane counts as PEC and sends zeros)
#pragma unroll
            for (int p = 0; p < 12; ++p)
            {
                const float vp = v[p];
                o[p] = pec ? vp : o[p];
                mc = fmaxf(mc, fabsf(o[p]));
            }
            const float L0 = __shfl_up(o[2], 1, kWave), L1 = __shfl_up(o[3], 1, kWave);
            const float R0 = __shfl_down(o[0], 1, kWave), R1 = __shfl_down(o[1], 1, kWave);
            if (live)
            {
                // x and y faces: the value and the slot (this cell's own face,
                // or the neighbour's opposite face) are selected per lane, so
                // every face is one encode and one store
                auto face = [&](int p, unsigned bbit, unsigned pbit, unsigned nbit, float n0,
                                float n1, unsigned own_q, unsigned nb_q, unsigned nb_g) {
                    const bool bnd = (cf & bbit) != 0u, pecN = !CL && (cf & pbit) != 0u;
                    const bool mine = nbit != 0u && (cf & nbit) != 0u;
                    const float o0 = o[p], o1 = o[p + 1];
                    const float r0 = rho * o0, r1 = rho * o1, m0 = -o0, m1 = -o1;
                    const bool take = !pec && mine, neg = !pec && pecN;
                    float x = take ? n0 : o0, y = take ? n1 : o1;
                    x = neg ? m0 : x;
                    y = neg ? m1 : y;
                    x = bnd ? r0 : x;
                    y = bnd ? r1 : y;
                    const bool own = pec || bnd || pecN || mine;
                    const unsigned ownoff = own_q * PQ + bo, nboff = nb_q * PQ + blk16(nb_g, ES);
                    stpair<PF, H>(Bk, own ? ownoff : nboff, s, x, y);
                };
                // pair 0 (x-: ports 0,1), pair 1 (x+: ports 2,3): the left /
                // right neighbour in this wave takes the pulse from a shuffle
                face(0, LF_BX0, NB_XM, LF_L, L0, L1, 0u, 1u, (unsigned)f - 1u);
                face(2, LF_BX1, NB_XP, LF_R, R0, R1, 1u, 0u, (unsigned)f + 1u);
                // pair 2 (y-: ports 4,5), pair 3 (y+: ports 6,7): always pushed
                face(4, LF_BY0, NB_YM, 0u, 0.f, 0.f, 2u, 3u, (unsigned)(f - a.nx));
                face(6, LF_BY1, NB_YP, 0u, 0.f, 0.f, 3u, 2u, (unsigned)(f + a.nx));
                // ---- z faces: pair 4 (z-: ports 8,9), pair 5 (z+: ports 10,11) ----
                if (k > k0)
                {
                    // pull: slab k-1's reflected z+ pulses are in registers
                    const float n8 = -o[8], n9 = -o[9], nz0 = -zpo0, nz1 = -zpo1;
                    // (in a clean block the cell below is not PEC either: it
                    // would have set this cell's NB_ZM bit)
                    const bool pp = !CL && pecPrev;
                    float x = pp ? n8 : zpo0, y = pp ? n9 : zpo1;
                    x = pec ? o[8] : x;
                    y = pec ? o[9] : y;
                    stpair<PF, H>(Bk + 4 * PQ, bo, s, x, y);
                    // slab k-1's z+ slot: its kept pulses if it is PEC (then
                    // zpo holds them), else what this cell sends down
                    float u = pec ? nz0 : o[8], w = pec ? nz1 : o[9];
                    u = pp ? zpo0 : u;
                    w = pp ? zpo1 : w;
                    stpair_at<PF>(a.B, sp, 5, u, w);
                }
                else if (k == 0)
                    stpair<PF, H>(Bk + 4 * PQ, bo, s, rho * o[8], rho * o[9]);
                else
                {
                    const bool pecZ = !CL && (cf & NB_ZM) != 0u;
                    if (pec || pecZ)
                        stpair<PF, H>(Bk + 4 * PQ, bo, s, pec ? o[8] : -o[8], pec ? o[9] : -o[9]);
                    else
                        stpair_at<PF>(a.B, slab16(a, k - 1, f), 5, o[8], o[9]);
                }
                if (k == k1 - 1)
                {
                    if (k == a.nz - 1)
                        stpair<PF, H>(Bk + 5 * PQ, bo, s, rho * o[10], rho * o[11]);
                    else
                    {
                        const bool pecZ = !CL && (cf & NB_ZP) != 0u;
                        if (pec || pecZ)
                            stpair<PF, H>(Bk + 5 * PQ, bo, s, pec ? o[10] : -o[10],
                                          pec ? o[11] : -o[11]);
                        else
                            stpair_at<PF>(a.B, slab16(a, k + 1, f), 4, o[10], o[11]);
                    }
                }
            }
            // non-negative floats order like their bit patterns: a DPP max
            const unsigned wm = __reduce_max_sync(~0ull, __float_as_uint(mc));
            if (lane == 0)
                a.mpart[(size_t)k * a.strips + t] = __uint_as_float(wm);
            zpo0 = o[10];
            zpo1 = o[11];
            pecPrev = pec;
        };
        // the FP32 band (a few slabs near the sources) takes the general body
        if (s.h16 && (fl & kBlockClean) != 0u)
            body(std::true_type{}, std::true_type{});
        else if (s.h16)
            body(std::true_type{}, std::false_type{});
        else
            body(std::false_type{}, std::false_type{});
        sp = s;
    }
}


// Per-slab maximum of the wave records of one step (one workgroup per slab).
static __global__ void k_slab_max(const float *mpart, int strips, float *mreg)
{
    __shared__ float s[256];
    const float *row = mpart + (size_t)blockIdx.x * strips;
    float m = 0.f;
    for (int t = threadIdx.x; t < strips; t += blockDim.x)
        m = fmaxf(m, row[t]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        mreg[blockIdx.x] = s[0];
}

// Pulse access by padded cell index for the small kernels (injection,
// monitors, conversions), any format.
struct PulseView
{
    char *base;
    long long PB;     // bytes per plane
    long long Pp;     // slab pitch (cells)
    int pf;           // PF_*
    const uint2 *slab; // 16-bit formats: per slab (unused, 16-bit flag)
    int h0, h1;        // 16-bit formats: the FP32 slabs [h0, h1)
    const float *sc;  // per slab 2^E (current step)
    const float *isc; // per slab 2^-E
    Fold fold;        // monitors of an FP32 step with folded sources
};
// FP32 contexts: 64-cell blocks of 12 ports per slab. 16-bit contexts: 64-cell
// blocks of 6 pairs (pair p/2, half p&1), two codes per cell and pair in a
// 16-bit slab, a float2 in an FP32 slab.
__device__ __forceinline__ float pv_get(const PulseView &v, int p, long long c)
{
    const long long k = c / v.Pp, f = c - k * v.Pp;
    if (v.pf == PF_F32)
        return *reinterpret_cast<const float *>(v.base + k * v.Pp * 48 + p * 256 +
                                                blk_off((unsigned)f));
    const uint2 s = v.slab[k];
    const unsigned es = s.y ? 4u : 8u;
    const char *q = v.base + slab16_base(v.Pp, k, v.h0, v.h1) + (size_t)(p >> 1) * 64u * es + blk16((unsigned)f, es);
    if (!s.y)
        return reinterpret_cast<const float *>(q)[p & 1];
    const unsigned h = reinterpret_cast<const unsigned short *>(q)[p & 1];
    return (v.pf == PF_F16 ? dec16<PF_F16>(h) : dec16<PF_BF16>(h)) * v.sc[k];
}
// Stores x; returns |x * 2^-E| for a 16-bit slot (0 for FP32) so callers can
// check the range.
__device__ __forceinline__ float pv_set(const PulseView &v, int p, long long c, float x)
{
    const long long k = c / v.Pp, f = c - k * v.Pp;
    if (v.pf == PF_F32)
    {
        *reinterpret_cast<float *>(v.base + k * v.Pp * 48 + p * 256 + blk_off((unsigned)f)) = x;
        return 0.f;
    }
    const uint2 s = v.slab[k];
    const unsigned es = s.y ? 4u : 8u;
    char *q = v.base + slab16_base(v.Pp, k, v.h0, v.h1) + (size_t)(p >> 1) * 64u * es + blk16((unsigned)f, es);
    if (!s.y)
    {
        reinterpret_cast<float *>(q)[p & 1] = x;
        return 0.f;
    }
    const float y = x * v.isc[k];
    reinterpret_cast<unsigned short *>(q)[p & 1] =
        v.pf == PF_F16 ? enc16<PF_F16>(y) : enc16<PF_BF16>(y);
    return fabsf(y);
}

// Stores a 16-bit code into a slot of a 16-bit slab.
__device__ __forceinline__ void pv_set_code(const PulseView &v, int p, long long c, unsigned short h)
{
    const long long k = c / v.Pp, f = c - k * v.Pp;
    char *q = v.base + slab16_base(v.Pp, k, v.h0, v.h1) + (size_t)(p >> 1) * 64u * 4u + blk16((unsigned)f, 4u);
    reinterpret_cast<unsigned short *>(q)[p & 1] = h;
}

// Slab exponent for a bound: bound * 2^-E < 2^15, E in [-100, 100]
// (identical integer steps on the host and in the oracle).
__host__ __device__ inline int plan_exponent(float bound)
{
    if (!(bound > 0.0f))
        return 0;
    unsigned u;
    __builtin_memcpy(&u, &bound, 4);
    const int be = (int)((u >> 23) & 0xffu);
    if (be == 0)
        return -100;
    int e = be - 126 - 15;
    return e < -100 ? -100 : (e > 100 ? 100 : e);
}
__host__ __device__ inline float pow2_float(int e)
{
    unsigned u = (unsigned)(e + 127) << 23;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// Scale planning of the 16-bit formats, run before step n:
//   bound_k = (rp*6) * (rp*max_{|d|<=2} M[k+d] + max_{|d|<=1} Q[k+d]*|sv_n|)
//             + Q_k*|sv_{n+1}|
// bounds every value stored into slab k by step n (one step multiplies the
// max norm by at most 5; a pulse moves one slab per step), M = the per-slab
// maximum the cells sent at step n-1, Q = 0.5 * sum |amp| of the slab's soft
// sources. Also checks that the values step n-1 stored fit their codes.
struct PlanArgs
{
    int nz;
    const float *mreg, *Q;
    const uint2 *slab;
    const float *sv;
    int n, nsteps;
    float rp, fmax;
    const float *isc_cur;
    float *sc_next, *isc_next;
    int *flag;
};
__device__ __forceinline__ void plan_slab(const PlanArgs &p, int k)
{
    float Mn = 0.f, Qn = 0.f, M1 = 0.f;
    for (int d = -2; d <= 2; ++d)
        if (k + d >= 0 && k + d < p.nz)
        {
            Mn = fmaxf(Mn, p.mreg[k + d]);
            if (d >= -1 && d <= 1)
            {
                Qn = fmaxf(Qn, p.Q[k + d]);
                M1 = fmaxf(M1, p.mreg[k + d]);
            }
        }
    const float svn = fabsf(p.sv[p.n]);
    const float svn1 = (p.n + 1 < p.nsteps) ? fabsf(p.sv[p.n + 1]) : 0.0f;
    const float b = (p.rp * 6.0f) * ((p.rp * Mn) + (Qn * svn)) + (p.Q[k] * svn1);
    const int E = plan_exponent(b);
    p.sc_next[k] = pow2_float(E);
    p.isc_next[k] = pow2_float(-E);
    if (p.slab[k].y && (p.rp * M1) * p.isc_cur[k] > p.fmax)
        p.flag[0] = 1;
}

// Adds the sources [s0, s1) of one cell in list order (0.5*amp*sv on the four
// ports polarised along pol); returns the largest |code| a 16-bit slot took.
// The cell's 12 slots are loaded together, the additions done in registers
// and the touched slots stored once, so the thread waits for memory twice,
// not once per addition. A 16-bit slot is rounded after each addition, the
// value a store and a new load of the slot would give, so every slot takes
// the same sequence of values as with one read-modify-write per addition.
__device__ __forceinline__ float inject_cell(const PulseView &pv, long long c, int s0, int s1,
                                             const int *spol, const float *samp, float sv)
{
    const long long k = c / pv.Pp;
    const bool h16 = pv.pf != PF_F32 && pv.slab[k].y != 0;
    const float sc = h16 ? pv.sc[k] : 1.0f, isc = h16 ? pv.isc[k] : 1.0f;
    float v[12];
#pragma unroll
    for (int p = 0; p < 12; ++p)
        v[p] = pv_get(pv, p, c);
    unsigned short hc[12] = {};
    float over = 0.f;
    unsigned touched = 0u;
    for (int s = s0; s < s1; ++s)
    {
        const int pol = spol[s];
        const float qv = 0.5f * samp[s] * sv;
#pragma unroll
        for (int pp = 0; pp < 3; ++pp)
        {
            if (pp != pol)
                continue;
#pragma unroll
            for (int axis = 0; axis < 3; ++axis)
            {
                if (axis == pp)
                    continue;
#pragma unroll
                for (int side = 0; side < 2; ++side)
                {
                    const int p = scn::port_of(axis, side, pp);
                    float x = v[p] + qv;
                    if (h16)
                    {
                        const float y = x * isc;
                        const unsigned short h =
                            pv.pf == PF_F16 ? enc16<PF_F16>(y) : enc16<PF_BF16>(y);
                        over = fmaxf(over, fabsf(y));
                        hc[p] = h;
                        x = (pv.pf == PF_F16 ? dec16<PF_F16>(h) : dec16<PF_BF16>(h)) * sc;
                    }
                    v[p] = x;
                    touched |= 1u << p;
                }
            }
        }
    }
    // a 16-bit slot takes the code of its last addition as it is (its
    // decoded value times 2^-E need not give the code back where the decode
    // leaves the float range)
#pragma unroll
    for (int p = 0; p < 12; ++p)
        if (touched & (1u << p))
        {
            if (h16)
                pv_set_code(pv, p, c, hc[p]);
            else
                pv_set(pv, p, c, v[p]);
        }
    return over;
}

// Soft-source injection (TlmSolver.cpp:487-505). One thread per distinct
// cell adds its sources in list order, so every slot sees the same
// sequence of additions as the reference (a 16-bit slot is rounded after
// each addition). Threads past the sources plan the slab scales of the
// next step (16-bit formats, pl.nz > 0).
static __global__ void k_inject(PulseView pv, const long long *ucell, const int *ubeg, const int *spol,
                         const float *samp, int nuniq, const float *sv_tab, int n, PlanArgs pl)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nuniq)
    {
        u -= nuniq;
        if (u < pl.nz)
            plan_slab(pl, u);
        return;
    }
    const float over = inject_cell(pv, ucell[u], ubeg[u], ubeg[u + 1], spol, samp, sv_tab[n]);
    if (pl.flag && over > pl.fmax)
        pl.flag[0] = 1;
}

__device__ __forceinline__ void gather12(const PulseView &pv, long long c, float v[12])
{
#pragma unroll
    for (int p = 0; p < 12; ++p)
        v[p] = pv_get(pv, p, c);
    if (pv.fold.ninj > 0)
        fold_sources(pv.fold, c, v);
}

// Marks the blocks holding folded source cells (flag bit 1); one thread, so
// two cells of one block cannot race on the flag byte.
static __global__ void k_mark_fold(const long long *cell, int n, uint8_t *flags)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    for (int u = 0; u < n; ++u)
        flags[cell[u] / kWave] |= 2u;
}

// Port record of step n (TlmSolver.cpp:677-703): sequential float sum of
// V_j over the port cells and I from Ampere's loop round the middle cell.
__device__ __forceinline__ void monitor_port(const PulseView &pv, int id, int n, int np,
                                             const long long *pcells, const int *poff,
                                             const int *ppol, const long long *ploop, float *prec)
{
    const int pol = ppol[id];
    float vsum = 0.0f, ival = 0.0f;
    const int b = poff[id], e = poff[id + 1];
    for (int s = b; s < e; ++s)
    {
        float v[12];
        gather12(pv, pcells[s], v);
        vsum += scn::vj(v, pol);
    }
    if (e > b)
    {
        // Ampere loop round the middle cell (TlmSolver.cpp:687-699)
        const int t1 = (pol + 1) % 3, t2 = (pol + 2) % 3; // CYC_I, CYC_J
        const long long *L = ploop + 4 * (size_t)id;
        float v[12], zsum = 0.0f;
        gather12(pv, L[0], v);
        zsum += scn::zi(v, t2);
        gather12(pv, L[1], v);
        zsum -= scn::zi(v, t2);
        gather12(pv, L[2], v);
        zsum -= scn::zi(v, t1);
        gather12(pv, L[3], v);
        zsum += scn::zi(v, t1);
        ival = -2.0f * zsum / ETA0;
    }
    float *rec = prec + ((size_t)n * np + id) * 2;
    rec[0] = vsum;
    rec[1] = ival;
}

// E probe record of step n: (-1/dl) * V_j (TlmSolver.cpp:705-713).
__device__ __forceinline__ void monitor_probe(const PulseView &pv, int id, int n, int npr,
                                              const long long *prcell, float inv_dl, float *prrec)
{
    float v[12];
    gather12(pv, prcell[id], v);
    float *rec = prrec + ((size_t)n * npr + id) * 3;
    rec[0] = inv_dl * scn::vj(v, 0);
    rec[1] = inv_dl * scn::vj(v, 1);
    rec[2] = inv_dl * scn::vj(v, 2);
}

struct MonitorArgs
{
    PulseView pv;
    int n, settle;
    const float *cph, *sph;
    // Js faces
    int nf;
    const long long *fcell;
    const int *faxis;
    float *js_inst, *js_dft4;
    // Huygens faces
    int nh;
    const long long *hcell;
    const int *haxis;
    float *huy8;
    // ports
    int np;
    const long long *pcells;
    const int *poff, *ppol;
    const long long *ploop; // 4 per port: c+s1, c-s1, c+s2, c-s2
    float *prec;            // [step][np][2]
    // probes
    int npr;
    const long long *prcell;
    float inv_dl;
    float *prrec; // [step][npr][3]
    // |E| cut plane (0 threads on steps that do not take it)
    int pn, pa, pidx, pn1, pn2;
    int nx, ny;
    long long Pp;
    const uint16_t *code;
    float *plane;
};

// Monitors on the post-injection incident pulses (TlmSolver::monitors,
// TlmSolver.cpp:616-713), one thread per face / port / probe.
__device__ __forceinline__ void monitor_item(const MonitorArgs &a, int id)
{
    const bool doDft = a.n >= a.settle;
    if (id < a.nf)
    {
        float v[12];
        gather12(a.pv, a.fcell[id], v);
        const int ax = a.faxis[id];
        const float h1 = scn::zi(v, scn::pol_axis(ax, 0)), h2 = scn::zi(v, scn::pol_axis(ax, 1));
        a.js_inst[id] = sqrtf(h1 * h1 + h2 * h2);
        if (doDft)
        {
            const float cph = a.cph[a.n], sph = a.sph[a.n];
            float *d = a.js_dft4 + 4 * (size_t)id;
            d[0] += h1 * cph;
            d[1] -= h1 * sph;
            d[2] += h2 * cph;
            d[3] -= h2 * sph;
        }
        return;
    }
    id -= a.nf;
    if (id < a.nh)
    {
        if (!doDft)
            return;
        float v[12];
        gather12(a.pv, a.hcell[id], v);
        const int ax = a.haxis[id];
        const int t1 = scn::pol_axis(ax, 0), t2 = scn::pol_axis(ax, 1);
        const float e1 = -scn::vj(v, t1), e2 = -scn::vj(v, t2);
        const float m1 = -scn::zi(v, t1), m2 = -scn::zi(v, t2);
        const float cph = a.cph[a.n], sph = a.sph[a.n];
        float *d = a.huy8 + 8 * (size_t)id;
        d[0] += e1 * cph;
        d[1] -= e1 * sph;
        d[2] += e2 * cph;
        d[3] -= e2 * sph;
        d[4] += m1 * cph;
        d[5] -= m1 * sph;
        d[6] += m2 * cph;
        d[7] -= m2 * sph;
        return;
    }
    id -= a.nh;
    if (id < a.np)
    {
        monitor_port(a.pv, id, a.n, a.np, a.pcells, a.poff, a.ppol, a.ploop, a.prec);
        return;
    }
    id -= a.np;
    if (id < a.npr)
    {
        monitor_probe(a.pv, id, a.n, a.npr, a.prcell, a.inv_dl, a.prrec);
        return;
    }
    id -= a.npr;
    if (id < a.pn)
    {
        // |E| = sqrt(sum_j V_j^2) on the plane, PEC cells 0 (TlmSolver.cpp:715-745)
        const int q1 = id % a.pn1, q2 = id / a.pn1;
        const int a1 = (a.pa == 0) ? 1 : 0, a2 = (a.pa == 2) ? 1 : 2;
        int co[3];
        co[a.pa] = a.pidx;
        co[a1] = q1;
        co[a2] = q2;
        const long long c = (long long)co[2] * a.Pp + (long long)co[1] * a.nx + co[0];
        float e = 0.0f;
        if ((a.code[c] & 0xFFu) != 1u)
        {
            float v[12];
            gather12(a.pv, c, v);
            float e2 = 0.0f;
            for (int j = 0; j < 3; ++j)
            {
                const float vj = scn::vj(v, j);
                e2 += vj * vj;
            }
            e = sqrtf(e2);
        }
        a.plane[id] = e;
    }
}

// The fused step: scatter + connect of every cell (strip_task / strip_task16)
// and, in workgroups [mon_from, gridDim.x), the monitor records.
template <bool STUBS, bool ENERGY, bool NT, int PF, bool FOLD = false>
__global__ void __launch_bounds__(256) k_step(StepArgs a, MonitorArgs m, int mon_from)
{
    // workgroups from mon_from on take the monitor records of this step
    // (they read the incident pulses A, which the step only reads, and run
    // alongside the step's workgroups instead of in a launch of their own)
    if ((int)blockIdx.x >= mon_from)
    {
        monitor_item(m, ((int)blockIdx.x - mon_from) * (int)blockDim.x + (int)threadIdx.x);
        return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wpb = blockDim.x / kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const XcdRange gr = xcd_range_map(blockIdx.x, a.xm);
    double esum = 0.0;
    for (int g = gr.first; g < gr.end; g += gr.step)
    {
        const int task = g * wpb + wave;
        if (task < a.tasks)
        {
            if constexpr (PF == PF_F32)
                strip_task<STUBS, ENERGY, NT, PF, FOLD>(a, task, lane, esum);
            else
                strip_task16<STUBS, ENERGY, PF>(a, task, lane, esum);
        }
    }
    if (ENERGY)
    {
        for (int off = kWave / 2; off > 0; off >>= 1)
            esum += __shfl_down(esum, off, kWave);
        if (lane == 0)
            a.epart[blockIdx.x * wpb + wave] = esum;
    }
}

// First stage of the energy sum: block b adds partials [b*chunk, (b+1)*chunk)
// in a fixed order (so the result does not depend on scheduling).
static __global__ void k_energy_chunks(const double *part, long long nparts, long long chunk, double *out)
{
    __shared__ double s[256];
    const long long lo = blockIdx.x * chunk, hi = min(nparts, lo + chunk);
    double acc = 0.0;
    for (long long i = lo + threadIdx.x; i < hi; i += blockDim.x)
        acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[blockIdx.x] = s[0];
}

// Sum of the energy partials of one step in a fixed order.
static __global__ void k_energy_reduce(const double *part, int nparts, double *out)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = s[0];
}

// Running FP64 DFT of the probe records over steps [n0, n1), in step order.
static __global__ void k_probe_dft(const float *rec, int npr, const double *ctab, const double *stab,
                            int nfreq, int n0, int n1, double *re, double *im)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= npr * 3 * nfreq)
        return;
    const int qf = id % nfreq, pc = id / nfreq;
    double ar = re[id], ai = im[id];
    for (int n = n0; n < n1; ++n)
    {
        const double x = (double)rec[(size_t)n * npr * 3 + pc];
        ar += x * ctab[(size_t)n * nfreq + qf];
        ai -= x * stab[(size_t)n * nfreq + qf];
    }
    re[id] = ar;
    im[id] = ai;
}

// AoS (12 or 3 floats per cell, unpadded) <-> SoA padded planes, for the
// cells [c0, c0 + count) of the unpadded index space. Stubs are always FP32.
// Stubs (3 per cell) between host cell order and the device's 64-cell
// stub blocks.
static __global__ void k_stubs_in(const float *aos, float *Vs, long long nxy, long long Pp,
                                  long long c0, long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long k = cell / nxy;
    const long long pc = k * Pp + (cell - k * nxy);
    float *b = stub_block(Vs, pc >> 6) + (pc & (kWave - 1));
    for (int j = 0; j < 3; ++j)
        b[j * kWave] = aos[id * 3 + j];
}

static __global__ void k_stubs_out(const float *Vs, float *aos, long long nxy, long long Pp,
                                   long long c0, long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long k = cell / nxy;
    const long long pc = k * Pp + (cell - k * nxy);
    const float *b = stub_block(const_cast<float *>(Vs), pc >> 6) + (pc & (kWave - 1));
    for (int j = 0; j < 3; ++j)
        aos[id * 3 + j] = b[j * kWave];
}

// Pulses (12 per cell) in any storage format.
static __global__ void k_pulses_in(const float *aos, PulseView pv, long long nxy, long long c0,
                            long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long k = cell / nxy;
    const long long pc = k * pv.Pp + (cell - k * nxy);
    for (int p = 0; p < 12; ++p)
        pv_set(pv, p, pc, aos[id * 12 + p]);
}

static __global__ void k_pulses_out(PulseView pv, float *aos, long long nxy, long long c0, long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long k = cell / nxy;
    const long long pc = k * pv.Pp + (cell - k * nxy);
    for (int p = 0; p < 12; ++p)
        aos[id * 12 + p] = pv_get(pv, p, pc);
}

// Device rounding of the 16-bit formats (v * inv_scale), for checking the
// conversion instructions against the IEEE definition.
static __global__ void k_encode(int pf, const float *in, float inv_scale, unsigned short *out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float y = in[i] * inv_scale;
    out[i] = pf == PF_F16 ? enc16<PF_F16>(y) : enc16<PF_BF16>(y);
}

// Scattering matrix columns of one node (air or stub-loaded) through the
// device scatter: column c is the response to a unit pulse on port c
// (c < 12) or stub c-12.
static __global__ void k_scatter_columns(float yh, float den, int diel, float *S /*[15][15] row-major*/)
{
    const int c = threadIdx.x;
    if (c >= 15)
        return;
    float vin[12] = {0}, vs[3] = {0}, vout[12];
    if (c < 12)
        vin[c] = 1.0f;
    else
        vs[c - 12] = 1.0f;
    if (diel)
        scn::scatter_diel(vin, vs, yh, den, vout);
    else
        scn::scatter_air(vin, vout);
    for (int r = 0; r < 12; ++r)
        S[r * 15 + c] = vout[r];
    for (int r = 0; r < 3; ++r)
        S[(12 + r) * 15 + c] = vs[r];
}

// Plain grid-stride copy: the reference kernel for the measured device-copy
// bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters.
static __global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n4)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

} // namespace tlm
