// tlm_kernels.h - device kernels of the TLM time step (FP32 pulses).
//
// Layout in HBM (set up in GpuTlmHip.hip):
//   * pulses: 12 planes of P floats each (structure of arrays). Inside a
//     plane, z-slab k starts at k*Pp with Pp = roundup(nx*ny, 64), so every
//     wavefront's 64 consecutive cells are one aligned 256-byte run; the
//     cell (i, j, k) lives at k*Pp + j*nx + i. Port p of that cell is at
//     p*P + k*Pp + j*nx + i. The slab padding is never read as a neighbour.
//   * two such buffers (A = incident pulses of step n, B = of step n+1).
//   * stubs: 3 planes of P floats, updated in place, dielectric cells only.
//   * cell code: 16 bits per cell address: material id in bits 0-7 (0 air,
//     1 PEC, >= 2 dielectric), then one bit per face telling whether the
//     neighbour across it is PEC (bit 8 x-, 9 x+, 10 y-, 11 y+, 12 z-,
//     13 z+; 0 on outer faces). Built once at upload, so a cell needs a
//     single 2-byte load and never reads its neighbours' materials.
//
// One step is (TlmSolver::step, TlmSolver.cpp:485-513):
//   k_inject  : soft sources add 0.5*amp*sv(n) into A        (:487-505)
//   k_monitor : Js / Huygens / port / probe records from A     (:616-713)
//   k_step    : scatter every cell from A and deliver each reflected
//               pulse to the slot it reaches after connect, writing B
//               (scatter :516-548 and connect :551-613 fused); optionally
//               the FP64 sum of A^2 for the energy monitor (:766-789).
#pragma once

#include "scn_node.h"
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tlm {

constexpr float ETA0 = 376.730313f; // TlmSolver.cpp:16
constexpr int kWave = 64;
constexpr int kNumXcd = 8;          // MI355X: 8 accelerator complex dies, one L2 each

struct StepArgs
{
    const float *A;
    float *B;
    float *Vs;
    const uint16_t *code; // cell code (material + neighbour-PEC bits)
    const float4 *props;  // per dielectric id: (yh, gh, den, 0)
    double *epart;        // energy partial per wave task
    int nx, ny, nz;
    int nxy;             // nx*ny
    long long Pp;        // slab pitch: roundup(nx*ny, 64)
    long long P;         // floats per pulse plane = Pp*nz
    int strips;          // 64-cell strips per slab: Pp/64
    int zc, nzc;         // planes per task, number of z chunks
    int tasks;           // strips * nzc
    int groups;          // task groups (one per workgroup's worth of waves)
    float rho;
};

// Loads and stores through a uniform plane base (SGPR pair) and a 32-bit
// byte offset per lane (one VGPR), the global_load saddr+voffset form.
__device__ __forceinline__ float ldo(const float *base, unsigned boff)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + boff);
}
template <bool NT> __device__ __forceinline__ void sto(float *base, unsigned boff, float v)
{
    float *p = reinterpret_cast<float *>(reinterpret_cast<char *>(base) + boff);
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
template <bool STUBS, bool ENERGY, bool NT>
__device__ __forceinline__ void strip_task(const StepArgs &a, int task, int lane, double &esum)
{
    const int tz = task / a.strips, t = task - tz * a.strips;
    const int f = t * kWave + lane; // cell inside the slab
    const bool act = f < a.nxy;
    const int j = f / a.nx, i = f - j * a.nx;
    const int k0 = tz * a.zc, k1 = min(a.nz, k0 + a.zc);
    const float rho = a.rho;
    // byte strides (all offsets of one plane fit in 32 bits)
    const unsigned bz = (unsigned)a.Pp * 4u, by = (unsigned)a.nx * 4u;
    const float *A[12];
    float *B[12];
#pragma unroll
    for (int p = 0; p < 12; ++p)
    {
        A[p] = a.A + p * a.P;
        B[p] = a.B + p * a.P;
    }
    // x neighbour inside the wave; otherwise (strip seam) the two cells push
    const bool hasL = act && lane > 0 && i > 0;
    const bool hasR = act && lane < kWave - 1 && i < a.nx - 1;
    const bool bx0 = i == 0, bx1 = i == a.nx - 1, by0 = j == 0, by1 = j == a.ny - 1;

    float zpo0 = 0.f, zpo1 = 0.f; // previous slab: reflected z+ pulses
    bool pecPrev = false;
    unsigned mo = (unsigned)(k0 * a.Pp + f); // cell index in the slab-padded space
    unsigned bo = mo * 4u;                   // its byte offset inside a plane

    for (int k = k0; k < k1; ++k, mo += (unsigned)a.Pp, bo += bz)
    {
        float o[12];
        bool pec = true;
        unsigned code = 0;
        if (act)
        {
            code = ldc(a.code, mo);
            const unsigned m = code & 0xFFu;
            pec = (m == 1u);
            float v[12];
#pragma unroll
            for (int p = 0; p < 12; ++p)
                v[p] = ldo(A[p], bo);
            if (ENERGY)
            {
#pragma unroll
                for (int p = 0; p < 12; ++p)
                    esum += (double)v[p] * (double)v[p];
            }
            if (STUBS && m >= 2u)
            {
                float *S0 = a.Vs, *S1 = a.Vs + a.P, *S2 = a.Vs + 2 * a.P;
                float st[3] = {ldo(S0, bo), ldo(S1, bo), ldo(S2, bo)};
                const float4 pr = a.props[m - 2u];
                scn::scatter_diel(v, st, pr.x, pr.z, o);
                sto<false>(S0, bo, st[0]);
                sto<false>(S1, bo, st[1]);
                sto<false>(S2, bo, st[2]);
            }
            else
                scn::scatter_air(v, o);
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
                    {
                        const float old = ldo(A[0 + pol], bo);
                        sto<NT>(B[0 + pol], bo, bx0 ? rho * old : old);
                    }
                    else if (bx0)
                        sto<NT>(B[0 + pol], bo, rho * oo);
                    else if (pecL)
                        sto<NT>(B[0 + pol], bo, -oo);
                    else if (hasL)
                        sto<NT>(B[0 + pol], bo, pol ? L1 : L0);
                    else // seam with a live neighbour: it fills my slot, I fill its
                        sto<NT>(B[2 + pol], bo - 4u, oo);

                    const float op = o[2 + pol];
                    if (pec)
                    {
                        const float old = ldo(A[2 + pol], bo);
                        sto<NT>(B[2 + pol], bo, bx1 ? rho * old : old);
                    }
                    else if (bx1)
                        sto<NT>(B[2 + pol], bo, rho * op);
                    else if (pecR)
                        sto<NT>(B[2 + pol], bo, -op);
                    else if (hasR)
                        sto<NT>(B[2 + pol], bo, pol ? R1 : R0);
                    else
                        sto<NT>(B[0 + pol], bo + 4u, op);
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
                        const float vl = ldo(A[4 + pol], bo), vh = ldo(A[6 + pol], bo);
                        sto<NT>(B[4 + pol], bo, by0 ? rho * vl : vl);
                        sto<NT>(B[6 + pol], bo, by1 ? rho * vh : vh);
                    }
                    else
                    {
                        if (by0)
                            sto<NT>(B[4 + pol], bo, rho * ol);
                        else if (pecLo)
                            sto<NT>(B[4 + pol], bo, -ol);
                        else
                            sto<NT>(B[6 + pol], bo - by, ol);
                        if (by1)
                            sto<NT>(B[6 + pol], bo, rho * oh);
                        else if (pecHi)
                            sto<NT>(B[6 + pol], bo, -oh);
                        else
                            sto<NT>(B[4 + pol], bo + by, oh);
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
                        sto<NT>(B[8 + pol], bo, ldo(A[8 + pol], bo));
                    else
                        sto<NT>(B[8 + pol], bo, pecPrev ? -ol : zp);
                    // slab k-1's z+ slot
                    if (pecPrev)
                        sto<NT>(B[10 + pol], bo - bz, ldo(A[10 + pol], bo - bz));
                    else
                        sto<NT>(B[10 + pol], bo - bz, pec ? -zp : ol);
                }
                else if (k == 0)
                {
                    sto<NT>(B[8 + pol], bo, rho * (pec ? ldo(A[8 + pol], bo) : ol));
                }
                else
                {
                    const bool pecLo = (code & NB_ZM) != 0u;
                    if (pec)
                        sto<NT>(B[8 + pol], bo, ldo(A[8 + pol], bo));
                    else if (pecLo)
                        sto<NT>(B[8 + pol], bo, -ol);
                    else
                        sto<NT>(B[10 + pol], bo - bz, ol);
                }
                if (k == k1 - 1)
                {
                    const float oh = o[10 + pol];
                    if (k == a.nz - 1)
                        sto<NT>(B[10 + pol], bo, rho * (pec ? ldo(A[10 + pol], bo) : oh));
                    else
                    {
                        const bool pecHi = (code & NB_ZP) != 0u;
                        if (pec)
                            sto<NT>(B[10 + pol], bo, ldo(A[10 + pol], bo));
                        else if (pecHi)
                            sto<NT>(B[10 + pol], bo, -oh);
                        else
                            sto<NT>(B[8 + pol], bo + bz, oh);
                    }
                }
            }
        }
        zpo0 = o[10];
        zpo1 = o[11];
        pecPrev = pec;
    }
}

template <bool STUBS, bool ENERGY, bool NT>
__global__ void __launch_bounds__(256) k_step(StepArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wpb = blockDim.x / kWave;
    const int wave = threadIdx.x / kWave;
    const XcdRange gr = xcd_range(blockIdx.x, gridDim.x, a.groups);
    double esum = 0.0;
    for (int g = gr.first; g < gr.end; g += gr.step)
    {
        const int task = g * wpb + wave;
        if (task < a.tasks)
            strip_task<STUBS, ENERGY, NT>(a, task, lane, esum);
    }
    if (ENERGY)
    {
        for (int off = kWave / 2; off > 0; off >>= 1)
            esum += __shfl_down(esum, off, kWave);
        if (lane == 0)
            a.epart[blockIdx.x * wpb + wave] = esum;
    }
}


// Soft-source injection (TlmSolver.cpp:487-505). One thread per distinct
// cell adds its sources in list order, so every slot sees the same
// sequence of additions as the reference.
__global__ void k_inject(float *A, long long P, const long long *ucell, const int *ubeg,
                         const int *spol, const float *samp, int nuniq, const float *sv_tab, int n)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nuniq)
        return;
    const float sv = sv_tab[n];
    const long long c = ucell[u];
    for (int s = ubeg[u]; s < ubeg[u + 1]; ++s)
    {
        const int pol = spol[s];
        const float qv = 0.5f * samp[s] * sv;
        for (int axis = 0; axis < 3; ++axis)
        {
            if (axis == pol)
                continue;
            const int p0 = scn::port_of(axis, 0, pol), p1 = scn::port_of(axis, 1, pol);
            A[p0 * P + c] += qv;
            A[p1 * P + c] += qv;
        }
    }
}

__device__ __forceinline__ void gather12(const float *A, long long P, long long c, float v[12])
{
#pragma unroll
    for (int p = 0; p < 12; ++p)
        v[p] = A[p * P + c];
}

struct MonitorArgs
{
    const float *A;
    long long P;
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
__global__ void k_monitor(MonitorArgs a)
{
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    const bool doDft = a.n >= a.settle;
    if (id < a.nf)
    {
        float v[12];
        gather12(a.A, a.P, a.fcell[id], v);
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
        gather12(a.A, a.P, a.hcell[id], v);
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
        const int pol = a.ppol[id];
        float vsum = 0.0f, ival = 0.0f;
        const int b = a.poff[id], e = a.poff[id + 1];
        for (int s = b; s < e; ++s)
        {
            float v[12];
            gather12(a.A, a.P, a.pcells[s], v);
            vsum += scn::vj(v, pol);
        }
        if (e > b)
        {
            // Ampere loop round the middle cell (TlmSolver.cpp:687-699)
            const int t1 = (pol + 1) % 3, t2 = (pol + 2) % 3; // CYC_I, CYC_J
            const long long *L = a.ploop + 4 * (size_t)id;
            float v[12], zsum = 0.0f;
            gather12(a.A, a.P, L[0], v);
            zsum += scn::zi(v, t2);
            gather12(a.A, a.P, L[1], v);
            zsum -= scn::zi(v, t2);
            gather12(a.A, a.P, L[2], v);
            zsum -= scn::zi(v, t1);
            gather12(a.A, a.P, L[3], v);
            zsum += scn::zi(v, t1);
            ival = -2.0f * zsum / ETA0;
        }
        float *rec = a.prec + ((size_t)a.n * a.np + id) * 2;
        rec[0] = vsum;
        rec[1] = ival;
        return;
    }
    id -= a.np;
    if (id < a.npr)
    {
        float v[12];
        gather12(a.A, a.P, a.prcell[id], v);
        float *rec = a.prrec + ((size_t)a.n * a.npr + id) * 3;
        rec[0] = a.inv_dl * scn::vj(v, 0);
        rec[1] = a.inv_dl * scn::vj(v, 1);
        rec[2] = a.inv_dl * scn::vj(v, 2);
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
            gather12(a.A, a.P, c, v);
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

// First stage of the energy sum: block b adds partials [b*chunk, (b+1)*chunk)
// in a fixed order (so the result does not depend on scheduling).
__global__ void k_energy_chunks(const double *part, long long nparts, long long chunk, double *out)
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
__global__ void k_energy_reduce(const double *part, int nparts, double *out)
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
__global__ void k_probe_dft(const float *rec, int npr, const double *ctab, const double *stab,
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
// cells [c0, c0 + count) of the unpadded index space.
__global__ void k_aos_to_soa(const float *aos, float *soa, long long P, int ncomp, long long nxy,
                             long long Pp, long long c0, long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long k = cell / nxy;
    const long long pc = k * Pp + (cell - k * nxy);
    for (int p = 0; p < ncomp; ++p)
        soa[p * P + pc] = aos[id * ncomp + p];
}

__global__ void k_soa_to_aos(const float *soa, float *aos, long long P, int ncomp, long long nxy,
                             long long Pp, long long c0, long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long k = cell / nxy;
    const long long pc = k * Pp + (cell - k * nxy);
    for (int p = 0; p < ncomp; ++p)
        aos[id * ncomp + p] = soa[p * P + pc];
}

// Scattering matrix columns of one node (air or stub-loaded) through the
// device scatter: column c is the response to a unit pulse on port c
// (c < 12) or stub c-12.
__global__ void k_scatter_columns(float yh, float den, int diel, float *S /*[15][15] row-major*/)
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
__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n4)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

} // namespace tlm
