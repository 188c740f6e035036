// tlm_kernels.h - device kernels of the TLM time step (FP32 pulses).
//
// Layout in HBM (see the context set-up in GpuTlmHip.hip):
//   * pulses: 12 planes of P floats each (structure of arrays). Rows are
//     padded to a pitch px = roundup(nx, 4) so every 4-cell quad is one
//     16-byte aligned float4; padded cells carry material PEC and stay 0.
//     Plane p holds port p of every cell, cell c = (k*ny + j)*px + i.
//   * two such buffers (A = incident pulses of step n, B = of step n+1).
//   * stubs: 3 planes of P floats, updated in place, dielectric cells only.
//   * material: 1 byte per padded cell.
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
constexpr int kMaxThreads = 256;

struct StepArgs
{
    const float *A;
    float *B;
    float *Vs;
    const uint8_t *mat;
    const float4 *props; // per dielectric id: (yh, gh, den, 0)
    double *epart;       // energy partial per workgroup
    int nx, ny, nz;
    int px;              // padded row pitch (cells)
    int Q;               // quads per row = px/4
    int R;               // rows per workgroup
    int zc;              // planes per workgroup (z-march length)
    long long P;         // floats per pulse plane = px*ny*nz
    float rho;
};

// Bits e (0..3) of a material quad that hold material `id`.
__device__ __forceinline__ unsigned pec_bits(uint32_t m)
{
    unsigned b = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e)
        b |= (((m >> (8 * e)) & 0xFFu) == 1u ? 1u : 0u) << e;
    return b;
}
__device__ __forceinline__ unsigned diel_bits(uint32_t m)
{
    unsigned b = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e)
        b |= (((m >> (8 * e)) & 0xFFu) >= 2u ? 1u : 0u) << e;
    return b;
}

__device__ __forceinline__ void ld4(const float *p, float (&r)[4])
{
    float4 x = *reinterpret_cast<const float4 *>(p);
    r[0] = x.x;
    r[1] = x.y;
    r[2] = x.z;
    r[3] = x.w;
}
__device__ __forceinline__ void st4(float *p, const float (&r)[4])
{
    *reinterpret_cast<float4 *>(p) = make_float4(r[0], r[1], r[2], r[3]);
}
// Store the elements of r whose bit is set in mask (a full float4 when all
// four are set, the common case).
__device__ __forceinline__ void st4_masked(float *p, const float (&r)[4], unsigned mask)
{
    if (mask == 0xFu)
        st4(p, r);
    else if (mask)
    {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (mask & (1u << e))
                p[e] = r[e];
    }
}

// Fused scatter + connect over a tile of R full rows x zc planes, marching
// in z. Each thread owns one 4-cell quad of a row.
//
// Every slot of B is written exactly once:
//   x faces: by its owner; the neighbour's reflected pulse comes through LDS
//            (rows are whole inside a workgroup).
//   z faces: by its owner inside the z-march (the previous plane's z+
//            pulses stay in registers, its z+ slots are stored one plane
//            late); across z-chunk edges the sender pushes into the
//            neighbour's slot.
//   y faces: the sender pushes its reflected pulse into the neighbour row's
//            slot (aligned full-row float4 stores).
// The owner writes its own slot instead whenever the link is broken: own
// cell PEC (value kept, times rho on an outer face), outer face (rho * out),
// neighbour PEC (-out). This is connect(), TlmSolver.cpp:563-583.
template <bool STUBS, bool ENERGY>
__global__ void __launch_bounds__(kMaxThreads) k_step(StepArgs a)
{
    __shared__ float4 xo[2][kMaxThreads]; // x- pulses of cell 0, x+ pulses of cell 3
    __shared__ uint32_t xm[2][kMaxThreads];
    __shared__ double ered[kMaxThreads / 64];

    const int t = threadIdx.x;
    const int r = t / a.Q;
    const int q = t - r * a.Q;
    const int j = blockIdx.x * a.R + r;
    const bool act = (r < a.R) && (j < a.ny);
    const int k0 = blockIdx.y * a.zc;
    const int k1 = min(a.nz, k0 + a.zc);
    const int i0 = 4 * q;
    const long long sy = a.px;
    const long long sz = (long long)a.px * a.ny;
    const long long P = a.P;
    const float rho = a.rho;

    float zpo[2][4], zpv[2][4]; // previous plane: reflected z+ pulses and incident z+ pulses
    uint32_t mprev = 0;
    double esum = 0.0;

    for (int k = k0; k < k1; ++k)
    {
        const int buf = (k - k0) & 1;
        const long long c = ((long long)k * a.ny + j) * a.px + i0;
        float v[12][4], o[12][4];
        uint32_t m = 0x01010101u;
        unsigned pm = 0xFu;

        if (act)
        {
            m = *reinterpret_cast<const uint32_t *>(a.mat + c);
            pm = pec_bits(m);
#pragma unroll
            for (int p = 0; p < 12; ++p)
                ld4(a.A + p * P + c, v[p]);
            if (ENERGY)
            {
#pragma unroll
                for (int p = 0; p < 12; ++p)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        esum += (double)v[p][e] * (double)v[p][e];
            }
            unsigned dm = STUBS ? diel_bits(m) : 0u;
            float vs[3][4];
            if (STUBS && dm)
            {
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    ld4(a.Vs + s * P + c, vs[s]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                float vin[12], vout[12];
#pragma unroll
                for (int p = 0; p < 12; ++p)
                    vin[p] = v[p][e];
                if (STUBS && (dm & (1u << e)))
                {
                    float4 pr = a.props[((m >> (8 * e)) & 0xFFu) - 2u];
                    float st[3] = {vs[0][e], vs[1][e], vs[2][e]};
                    scn::scatter_diel(vin, st, pr.x, pr.z, vout);
                    vs[0][e] = st[0];
                    vs[1][e] = st[1];
                    vs[2][e] = st[2];
                }
                else
                    scn::scatter_air(vin, vout);
#pragma unroll
                for (int p = 0; p < 12; ++p)
                    o[p][e] = vout[p];
            }
            if (STUBS && dm)
            {
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    st4(a.Vs + s * P + c, vs[s]);
            }
            xo[buf][t] = make_float4(o[0][0], o[1][0], o[2][3], o[3][3]);
            xm[buf][t] = m;
        }
        __syncthreads();
        if (!act)
            continue;

        // ---- x faces (ports 0,1 = x-, 2,3 = x+): owner pull ----
        {
            float4 L = make_float4(0.f, 0.f, 0.f, 0.f), Rt = L;
            unsigned pecL = 1u, pecR = 1u;
            if (q > 0)
            {
                L = xo[buf][t - 1];
                pecL = ((xm[buf][t - 1] >> 24) & 0xFFu) == 1u;
            }
            if (q < a.Q - 1)
            {
                Rt = xo[buf][t + 1];
                pecR = (xm[buf][t + 1] & 0xFFu) == 1u;
            }
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float rm[4], rp[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                {
                    const int i = i0 + e;
                    const bool pec = (pm >> e) & 1u;
                    // x- slot: neighbour i-1 sends its x+ pulse
                    {
                        const bool bnd = (i == 0);
                        const float nbv = (e == 0) ? (pol ? L.w : L.z) : o[2 + pol][e - 1];
                        const bool nbp = (e == 0) ? (pecL != 0u) : (((pm >> (e - 1)) & 1u) != 0u);
                        const float oo = o[0 + pol][e], vv = v[0 + pol][e];
                        rm[e] = pec ? (bnd ? rho * vv : vv) : (bnd ? rho * oo : (nbp ? -oo : nbv));
                    }
                    // x+ slot: neighbour i+1 sends its x- pulse
                    {
                        const bool bnd = (i == a.nx - 1);
                        const float nbv = (e == 3) ? (pol ? Rt.y : Rt.x) : o[0 + pol][e + 1];
                        const bool nbp = (e == 3) ? (pecR != 0u) : (((pm >> (e + 1)) & 1u) != 0u);
                        const float oo = o[2 + pol][e], vv = v[2 + pol][e];
                        rp[e] = pec ? (bnd ? rho * vv : vv) : (bnd ? rho * oo : (nbp ? -oo : nbv));
                    }
                }
                st4(a.B + (0 + pol) * P + c, rm);
                st4(a.B + (2 + pol) * P + c, rp);
            }
        }

        // ---- y faces (ports 4,5 = y-, 6,7 = y+): sender push ----
        {
            const bool bLo = (j == 0), bHi = (j == a.ny - 1);
            const unsigned pLo = bLo ? 0u : pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c - sy));
            const unsigned pHi = bHi ? 0u : pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c + sy));
            const unsigned ownLo = bLo ? 0xFu : (pm | pLo);
            const unsigned pushLo = bLo ? 0u : (~pm & ~pLo & 0xFu);
            const unsigned ownHi = bHi ? 0xFu : (pm | pHi);
            const unsigned pushHi = bHi ? 0u : (~pm & ~pHi & 0xFu);
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float rl[4], rh[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                {
                    const bool pec = (pm >> e) & 1u;
                    rl[e] = pec ? (bLo ? rho * v[4 + pol][e] : v[4 + pol][e])
                                : (bLo ? rho * o[4 + pol][e] : -o[4 + pol][e]);
                    rh[e] = pec ? (bHi ? rho * v[6 + pol][e] : v[6 + pol][e])
                                : (bHi ? rho * o[6 + pol][e] : -o[6 + pol][e]);
                }
                st4_masked(a.B + (4 + pol) * P + c, rl, ownLo);
                st4_masked(a.B + (6 + pol) * P + c - sy, o[4 + pol], pushLo);
                st4_masked(a.B + (6 + pol) * P + c, rh, ownHi);
                st4_masked(a.B + (4 + pol) * P + c + sy, o[6 + pol], pushHi);
            }
        }

        // ---- z faces (ports 8,9 = z-, 10,11 = z+) ----
        if (k > k0)
        {
            const unsigned pp = pec_bits(mprev);
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float rz[4], rzp[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                {
                    const bool pec = (pm >> e) & 1u, pecB = (pp >> e) & 1u;
                    // this plane's z- slot: plane k-1 sends its z+ pulse
                    rz[e] = pec ? v[8 + pol][e] : (pecB ? -o[8 + pol][e] : zpo[pol][e]);
                    // plane k-1's z+ slot: this plane sends its z- pulse
                    rzp[e] = pecB ? zpv[pol][e] : (pec ? -zpo[pol][e] : o[8 + pol][e]);
                }
                st4(a.B + (8 + pol) * P + c, rz);
                st4(a.B + (10 + pol) * P + c - sz, rzp);
            }
        }
        else if (k == 0)
        {
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float rz[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    rz[e] = ((pm >> e) & 1u) ? rho * v[8 + pol][e] : rho * o[8 + pol][e];
                st4(a.B + (8 + pol) * P + c, rz);
            }
        }
        else
        {
            const unsigned pLo = pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c - sz));
            const unsigned own = pm | pLo, push = ~pm & ~pLo & 0xFu;
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float rz[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    rz[e] = ((pm >> e) & 1u) ? v[8 + pol][e] : -o[8 + pol][e];
                st4_masked(a.B + (8 + pol) * P + c, rz, own);
                st4_masked(a.B + (10 + pol) * P + c - sz, o[8 + pol], push);
            }
        }
        if (k == k1 - 1)
        {
            if (k == a.nz - 1)
            {
#pragma unroll
                for (int pol = 0; pol < 2; ++pol)
                {
                    float rz[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        rz[e] = ((pm >> e) & 1u) ? rho * v[10 + pol][e] : rho * o[10 + pol][e];
                    st4(a.B + (10 + pol) * P + c, rz);
                }
            }
            else
            {
                const unsigned pHi = pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c + sz));
                const unsigned own = pm | pHi, push = ~pm & ~pHi & 0xFu;
#pragma unroll
                for (int pol = 0; pol < 2; ++pol)
                {
                    float rz[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        rz[e] = ((pm >> e) & 1u) ? v[10 + pol][e] : -o[10 + pol][e];
                    st4_masked(a.B + (10 + pol) * P + c, rz, own);
                    st4_masked(a.B + (8 + pol) * P + c + sz, o[10 + pol], push);
                }
            }
        }
#pragma unroll
        for (int pol = 0; pol < 2; ++pol)
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                zpo[pol][e] = o[10 + pol][e];
                zpv[pol][e] = v[10 + pol][e];
            }
        mprev = m;
    }

    if (ENERGY)
    {
        // fixed-order reduction: wave tree, then the waves in order
        for (int off = 32; off > 0; off >>= 1)
            esum += __shfl_down(esum, off, 64);
        if ((t & 63) == 0)
            ered[t >> 6] = esum;
        __syncthreads();
        if (t == 0)
        {
            double s = 0.0;
            for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w)
                s += ered[w];
            a.epart[blockIdx.x + (size_t)blockIdx.y * gridDim.x] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Wave-independent variant of the fused step: no LDS and no barrier.
//
// Each wavefront owns a strip of rows (qw quads per row-strip, rw strips per
// wave) and marches over zc planes. x neighbours inside the wave are read
// with lane shuffles; where a row is split between two waves (rows wider
// than 256 cells) the two seam cells push their x pulse into each other's
// slot, exactly like the y and z-chunk links. Incident pulses are not kept
// live after the scatter: the rare PEC cells reload their (unchanged) value.
// ---------------------------------------------------------------------------
struct WaveArgs
{
    const float *A;
    float *B;
    float *Vs;
    const uint8_t *mat;
    const float4 *props;
    double *epart; // energy partial per wave
    int nx, ny, nz, px, Q;
    int qw;        // quads per wave row-strip: min(Q, 64)
    int rw;        // row-strips per wave: 64 / qw
    int nseg;      // strips per row: ceil(Q / qw)
    int tasks_xy;  // wave tasks per plane range: ceil(ny / rw) * nseg
    int zc, nzc;   // planes per task, number of z chunks
    long long P;
    float rho;
};

typedef float f4v __attribute__((ext_vector_type(4)));

template <bool NT> __device__ __forceinline__ void st4x(float *p, const float (&r)[4])
{
    if (NT)
    {
        f4v x = {r[0], r[1], r[2], r[3]};
        __builtin_nontemporal_store(x, reinterpret_cast<f4v *>(p));
    }
    else
        st4(p, r);
}
template <bool NT> __device__ __forceinline__ void st4x_masked(float *p, const float (&r)[4],
                                                               unsigned mask)
{
    if (mask == 0xFu)
        st4x<NT>(p, r);
    else if (mask)
    {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (mask & (1u << e))
                p[e] = r[e];
    }
}
__device__ __forceinline__ void ld4_if(const float *p, float (&r)[4], bool cond)
{
    if (cond)
        ld4(p, r);
    else
        r[0] = r[1] = r[2] = r[3] = 0.0f;
}
__device__ __forceinline__ bool byte_is_pec(uint32_t m, int e)
{
    return ((m >> (8 * e)) & 0xFFu) == 1u;
}

template <bool STUBS, bool ENERGY, bool NT>
__global__ void __launch_bounds__(kMaxThreads) k_step_wave(WaveArgs a)
{
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int tx = gw % a.tasks_xy, tz = gw / a.tasks_xy;
    if (tz >= a.nzc)
        return; // whole wave
    const int r = lane / a.qw, ql = lane - r * a.qw;
    const int jg = tx / a.nseg, s = tx - jg * a.nseg;
    const int j = jg * a.rw + r, q = s * a.qw + ql;
    const bool act = (r < a.rw) && (j < a.ny) && (q < a.Q);
    const bool hasL = ql > 0, hasR = (ql < a.qw - 1) && (q < a.Q - 1);
    const bool seamL = act && ql == 0 && q > 0;
    const bool seamR = act && ql == a.qw - 1 && q < a.Q - 1;
    const int k0 = tz * a.zc, k1 = min(a.nz, k0 + a.zc);
    const int i0 = 4 * q;
    const long long sy = a.px, sz = (long long)a.px * a.ny, P = a.P;
    const float rho = a.rho;

    float zpo[2][4]; // previous plane's reflected z+ pulses
    uint32_t mprev = 0x01010101u;
    double esum = 0.0;

    for (int k = k0; k < k1; ++k)
    {
        const long long c = ((long long)k * a.ny + j) * a.px + i0;
        uint32_t m = 0x01010101u;
        float o[12][4];
        if (act)
        {
            m = *reinterpret_cast<const uint32_t *>(a.mat + c);
            float v[12][4];
#pragma unroll
            for (int p = 0; p < 12; ++p)
                ld4(a.A + p * P + c, v[p]);
            if (ENERGY)
            {
#pragma unroll
                for (int p = 0; p < 12; ++p)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        esum += (double)v[p][e] * (double)v[p][e];
            }
            const unsigned dm = STUBS ? diel_bits(m) : 0u;
            float vs[3][4];
            if (STUBS && dm)
            {
#pragma unroll
                for (int t = 0; t < 3; ++t)
                    ld4(a.Vs + t * P + c, vs[t]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                float vin[12], vout[12];
#pragma unroll
                for (int p = 0; p < 12; ++p)
                    vin[p] = v[p][e];
                if (STUBS && (dm & (1u << e)))
                {
                    float4 pr = a.props[((m >> (8 * e)) & 0xFFu) - 2u];
                    float st[3] = {vs[0][e], vs[1][e], vs[2][e]};
                    scn::scatter_diel(vin, st, pr.x, pr.z, vout);
                    vs[0][e] = st[0];
                    vs[1][e] = st[1];
                    vs[2][e] = st[2];
                }
                else
                    scn::scatter_air(vin, vout);
#pragma unroll
                for (int p = 0; p < 12; ++p)
                    o[p][e] = vout[p];
            }
            if (STUBS && dm)
            {
#pragma unroll
                for (int t = 0; t < 3; ++t)
                    st4(a.Vs + t * P + c, vs[t]);
            }
        }
        else
        {
#pragma unroll
            for (int p = 0; p < 12; ++p)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    o[p][e] = 0.0f;
        }
        // x neighbours inside the wave (every lane takes part in the shuffles)
        const float Lz = __shfl_up(o[2][3], 1, 64), Lw = __shfl_up(o[3][3], 1, 64);
        const float Rx = __shfl_down(o[0][0], 1, 64), Ry = __shfl_down(o[1][0], 1, 64);
        const uint32_t mL = __shfl_up(m, 1, 64), mR = __shfl_down(m, 1, 64);
        if (!act)
            continue;
        const unsigned pm = pec_bits(m);

        // ---- x faces (ports 0,1 = x-, 2,3 = x+) ----
        {
            bool pecL = true, pecR = true;
            if (hasL)
                pecL = byte_is_pec(mL, 3);
            else if (seamL)
                pecL = a.mat[c - 1] == 1;
            if (hasR)
                pecR = byte_is_pec(mR, 0);
            else if (seamR)
                pecR = a.mat[c + 4] == 1;
            const bool pec0 = pm & 1u, pec3 = (pm >> 3) & 1u;
            const unsigned maskM = seamL ? (0xEu | ((pec0 || pecL) ? 1u : 0u)) : 0xFu;
            const unsigned maskP = seamR ? (0x7u | ((pec3 || pecR) ? 8u : 0u)) : 0xFu;
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float vm[4], vp[4], rm[4], rp[4];
                ld4_if(a.A + (0 + pol) * P + c, vm, pm != 0u);
                ld4_if(a.A + (2 + pol) * P + c, vp, pm != 0u);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                {
                    const int i = i0 + e;
                    const bool pec = (pm >> e) & 1u;
                    {
                        const bool bnd = (i == 0);
                        const float nbv = (e == 0) ? (pol ? Lw : Lz) : o[2 + pol][e - 1];
                        const bool nbp = (e == 0) ? pecL : (((pm >> (e - 1)) & 1u) != 0u);
                        const float oo = o[0 + pol][e];
                        rm[e] = pec ? (bnd ? rho * vm[e] : vm[e])
                                    : (bnd ? rho * oo : (nbp ? -oo : nbv));
                    }
                    {
                        const bool bnd = (i == a.nx - 1);
                        const float nbv = (e == 3) ? (pol ? Ry : Rx) : o[0 + pol][e + 1];
                        const bool nbp = (e == 3) ? pecR : (((pm >> (e + 1)) & 1u) != 0u);
                        const float oo = o[2 + pol][e];
                        rp[e] = pec ? (bnd ? rho * vp[e] : vp[e])
                                    : (bnd ? rho * oo : (nbp ? -oo : nbv));
                    }
                }
                st4x_masked<NT>(a.B + (0 + pol) * P + c, rm, maskM);
                st4x_masked<NT>(a.B + (2 + pol) * P + c, rp, maskP);
                if (seamL && !pec0 && !pecL)
                    a.B[(2 + pol) * P + c - 1] = o[0 + pol][0];
                if (seamR && !pec3 && !pecR)
                    a.B[(0 + pol) * P + c + 4] = o[2 + pol][3];
            }
        }

        // ---- y faces (ports 4,5 = y-, 6,7 = y+): sender push ----
        {
            const bool bLo = (j == 0), bHi = (j == a.ny - 1);
            const unsigned pLo = bLo ? 0u : pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c - sy));
            const unsigned pHi = bHi ? 0u : pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c + sy));
            const unsigned ownLo = bLo ? 0xFu : (pm | pLo);
            const unsigned pushLo = bLo ? 0u : (~pm & ~pLo & 0xFu);
            const unsigned ownHi = bHi ? 0xFu : (pm | pHi);
            const unsigned pushHi = bHi ? 0u : (~pm & ~pHi & 0xFu);
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float vl[4], vh[4], rl[4], rh[4];
                ld4_if(a.A + (4 + pol) * P + c, vl, pm != 0u);
                ld4_if(a.A + (6 + pol) * P + c, vh, pm != 0u);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                {
                    const bool pec = (pm >> e) & 1u;
                    rl[e] = pec ? (bLo ? rho * vl[e] : vl[e]) : (bLo ? rho * o[4 + pol][e] : -o[4 + pol][e]);
                    rh[e] = pec ? (bHi ? rho * vh[e] : vh[e]) : (bHi ? rho * o[6 + pol][e] : -o[6 + pol][e]);
                }
                st4x_masked<NT>(a.B + (4 + pol) * P + c, rl, ownLo);
                st4x_masked<NT>(a.B + (6 + pol) * P + c - sy, o[4 + pol], pushLo);
                st4x_masked<NT>(a.B + (6 + pol) * P + c, rh, ownHi);
                st4x_masked<NT>(a.B + (4 + pol) * P + c + sy, o[6 + pol], pushHi);
            }
        }

        // ---- z faces (ports 8,9 = z-, 10,11 = z+) ----
        if (k > k0)
        {
            const unsigned pp = pec_bits(mprev);
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float vz[4], vzp[4], rz[4], rzp[4];
                ld4_if(a.A + (8 + pol) * P + c, vz, pm != 0u);
                ld4_if(a.A + (10 + pol) * P + c - sz, vzp, pp != 0u);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                {
                    const bool pec = (pm >> e) & 1u, pecB = (pp >> e) & 1u;
                    rz[e] = pec ? vz[e] : (pecB ? -o[8 + pol][e] : zpo[pol][e]);
                    rzp[e] = pecB ? vzp[e] : (pec ? -zpo[pol][e] : o[8 + pol][e]);
                }
                st4x<NT>(a.B + (8 + pol) * P + c, rz);
                st4x<NT>(a.B + (10 + pol) * P + c - sz, rzp);
            }
        }
        else if (k == 0)
        {
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float vz[4], rz[4];
                ld4_if(a.A + (8 + pol) * P + c, vz, pm != 0u);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    rz[e] = ((pm >> e) & 1u) ? rho * vz[e] : rho * o[8 + pol][e];
                st4x<NT>(a.B + (8 + pol) * P + c, rz);
            }
        }
        else
        {
            const unsigned pLo = pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c - sz));
            const unsigned own = pm | pLo, push = ~pm & ~pLo & 0xFu;
#pragma unroll
            for (int pol = 0; pol < 2; ++pol)
            {
                float vz[4], rz[4];
                ld4_if(a.A + (8 + pol) * P + c, vz, pm != 0u);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    rz[e] = ((pm >> e) & 1u) ? vz[e] : -o[8 + pol][e];
                st4x_masked<NT>(a.B + (8 + pol) * P + c, rz, own);
                st4x_masked<NT>(a.B + (10 + pol) * P + c - sz, o[8 + pol], push);
            }
        }
        if (k == k1 - 1)
        {
            if (k == a.nz - 1)
            {
#pragma unroll
                for (int pol = 0; pol < 2; ++pol)
                {
                    float vz[4], rz[4];
                    ld4_if(a.A + (10 + pol) * P + c, vz, pm != 0u);
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        rz[e] = ((pm >> e) & 1u) ? rho * vz[e] : rho * o[10 + pol][e];
                    st4x<NT>(a.B + (10 + pol) * P + c, rz);
                }
            }
            else
            {
                const unsigned pHi = pec_bits(*reinterpret_cast<const uint32_t *>(a.mat + c + sz));
                const unsigned own = pm | pHi, push = ~pm & ~pHi & 0xFu;
#pragma unroll
                for (int pol = 0; pol < 2; ++pol)
                {
                    float vz[4], rz[4];
                    ld4_if(a.A + (10 + pol) * P + c, vz, pm != 0u);
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        rz[e] = ((pm >> e) & 1u) ? vz[e] : -o[10 + pol][e];
                    st4x_masked<NT>(a.B + (10 + pol) * P + c, rz, own);
                    st4x_masked<NT>(a.B + (8 + pol) * P + c + sz, o[10 + pol], push);
                }
            }
        }
#pragma unroll
        for (int pol = 0; pol < 2; ++pol)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                zpo[pol][e] = o[10 + pol][e];
        mprev = m;
    }

    if (ENERGY)
    {
        for (int off = 32; off > 0; off >>= 1)
            esum += __shfl_down(esum, off, 64);
        if (lane == 0)
            a.epart[gw] = esum;
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
    }
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
__global__ void k_aos_to_soa(const float *aos, float *soa, long long P, int ncomp, int nx, int ny,
                             int px, long long c0, long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long i = cell % nx, jk = cell / nx;
    const long long j = jk % ny, k = jk / ny;
    const long long pc = (k * ny + j) * px + i;
    for (int p = 0; p < ncomp; ++p)
        soa[p * P + pc] = aos[id * ncomp + p];
}

__global__ void k_soa_to_aos(const float *soa, float *aos, long long P, int ncomp, int nx, int ny,
                             int px, long long c0, long long count)
{
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count)
        return;
    const long long cell = c0 + id;
    const long long i = cell % nx, jk = cell / nx;
    const long long j = jk % ny, k = jk / ny;
    const long long pc = (k * ny + j) * px + i;
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
