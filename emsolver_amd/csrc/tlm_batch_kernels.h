// tlm_batch_kernels.h - kernels of the batched step: many independent grids
// advanced by the same launches.
//
// Each member grid keeps its own context buffers (pulses, stubs, cell codes,
// monitors). A tick advances every active member by one step: one
// injection launch for all members' soft sources (with the 16-bit scale
// planning), one fused step launch per member kind (stubs x storage format)
// over the concatenated task groups of its members (the first of them also
// taking every member's probe and port records in extra workgroups), one
// slab-maximum launch for the 16-bit members and a one-thread launch that
// advances the device tick. Every launch reads the
// tick from device memory, so a block of ticks can be captured once in a
// HIP graph and replayed.
#pragma once

#include "tlm_kernels.h"

namespace tlm {

struct BatchMember
{
    StepArgs st;        // geometry and pointers; A/B and the scales are picked per tick
    char *buf[2];       // pulse buffers: buf[0] holds the incident pulses of step n0
    long long tick0;    // batch tick at which the member runs its step n0
    int n0, nend;       // the member runs steps [n0, nend) in the batch
    int pf;             // PF_*
    int sbase;          // 16-bit: step whose scales are in buffer 0
    float *scl[2], *iscl[2];
    // soft sources
    const long long *ucell;
    const int *ubeg, *spol;
    const float *samp;
    const float *sv;    // schedule (nsched steps)
    int nsched;
    // 16-bit scale planning
    float *mreg;
    const float *qinj;
    int *flag;
    float rp, fmax;
    // probes and ports
    const long long *prcell;
    float *prrec;
    int npr;
    float inv_dl;
    const long long *pcells, *ploop;
    const int *poff, *ppol;
    float *prec;
    int np;
    int group0;         // first task group of the member in its kind's table
};

// Step of member m at batch tick `tick`; >= nend once it has finished.
__device__ __forceinline__ int member_step(const BatchMember &m, long long tick)
{
    return m.n0 + (int)(tick - m.tick0);
}

__device__ __forceinline__ PulseView member_view(const BatchMember &m, int n)
{
    PulseView v;
    v.base = m.buf[(n - m.n0) & 1];
    v.PB = m.st.PB;
    v.Pp = m.st.Pp;
    v.pf = m.pf;
    v.slab = m.st.slab;
    v.h0 = m.st.h0;
    v.h1 = m.st.h1;
    const int cur = (n - m.sbase) & 1;
    v.sc = m.pf != PF_F32 ? m.scl[cur] : nullptr;
    v.isc = m.pf != PF_F32 ? m.iscl[cur] : nullptr;
    return v;
}

// Injection of every member's soft sources (entries (member, unique cell)),
// then the 16-bit scale planning (entries (member, slab)).
static __global__ void k_inject_batch(const BatchMember *mem, const int2 *ent, int ninj, int nplan,
                               const long long *tick_p)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= ninj + nplan)
        return;
    const int2 e = ent[id];
    const BatchMember &m = mem[e.x];
    const int n = member_step(m, *tick_p);
    if (n >= m.nend)
        return;
    if (id < ninj)
    {
        const PulseView pv = member_view(m, n);
        const float over =
            inject_cell(pv, m.ucell[e.y], m.ubeg[e.y], m.ubeg[e.y + 1], m.spol, m.samp, m.sv[n]);
        if (m.pf != PF_F32 && over > m.fmax)
            m.flag[0] = 1;
        return;
    }
    const int cur = (n - m.sbase) & 1;
    PlanArgs pl;
    pl.nz = m.st.nz;
    pl.mreg = m.mreg;
    pl.Q = m.qinj;
    pl.slab = m.st.slab;
    pl.sv = m.sv;
    pl.n = n;
    pl.nsteps = m.nsched;
    pl.rp = m.rp;
    pl.fmax = m.fmax;
    pl.isc_cur = m.iscl[cur];
    pl.sc_next = m.scl[cur ^ 1];
    pl.isc_next = m.iscl[cur ^ 1];
    pl.flag = m.flag;
    plan_slab(pl, e.y);
}

// Probe and port records of every member (entries (member, index): index <
// npr is a probe, the rest ports).
__device__ __forceinline__ void monitor_batch_item(const BatchMember *mem, const int2 *ent, int nent,
                                                   const long long *tick_p, int id)
{
    if (id >= nent)
        return;
    const int2 e = ent[id];
    const BatchMember &m = mem[e.x];
    const int n = member_step(m, *tick_p);
    if (n >= m.nend)
        return;
    const PulseView pv = member_view(m, n);
    if (e.y < m.npr)
        monitor_probe(pv, e.y, n, m.npr, m.prcell, m.inv_dl, m.prrec);
    else
        monitor_port(pv, e.y - m.npr, n, m.np, m.pcells, m.poff, m.ppol, m.ploop, m.prec);
}

// Fused scatter + connect for the members of one kind. gmem[g] is the member
// of task group g; the workgroups walk the groups in XCD-contiguous ranges
// (a persistent grid when there are fewer workgroups than groups).
// Workgroups from mon_from on take the tick's probe and port records of all
// members (they read the incident pulses, which the step only reads).
template <bool STUBS, int PF>
__global__ void __launch_bounds__(256) k_step_batch(const BatchMember *mem, const int *gmem,
                                                    int ngroups, const long long *tick_p,
                                                    const int2 *ment, int nment, int mon_from)
{
    if ((int)blockIdx.x >= mon_from)
    {
        monitor_batch_item(mem, ment, nment, tick_p,
                           ((int)blockIdx.x - mon_from) * (int)blockDim.x + (int)threadIdx.x);
        return;
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wpb = blockDim.x / kWave;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const long long tick = *tick_p;
    const XcdRange gr = xcd_range(blockIdx.x, mon_from, ngroups);
    double esum = 0.0;
    for (int g = gr.first; g < gr.end; g += gr.step)
    {
        const int mi = gmem[g];
        const BatchMember &m = mem[mi];
        const int n = member_step(m, tick);
        if (n >= m.nend)
            continue;
        const int task = (g - m.group0) * wpb + wave;
        if (task >= m.st.tasks)
            continue;
        StepArgs a = m.st;
        const int s = (n - m.n0) & 1;
        a.A = m.buf[s];
        a.B = m.buf[s ^ 1];
        if (PF != PF_F32)
        {
            const int cur = (n - m.sbase) & 1;
            a.dsc = m.scl[cur];
            a.esc = m.iscl[cur ^ 1];
        }
        if constexpr (PF == PF_F32)
            strip_task<STUBS, false, false, PF>(a, task, lane, esum);
        else
            strip_task16<STUBS, false, PF>(a, task, lane, esum);
    }
}

// Per-slab maxima of the 16-bit members (one workgroup per (member, slab)).
static __global__ void k_slab_max_batch(const BatchMember *mem, const int2 *ent, const long long *tick_p)
{
    const int2 e = ent[blockIdx.x];
    const BatchMember &m = mem[e.x];
    const int n = member_step(m, *tick_p);
    if (n >= m.nend)
        return;
    __shared__ float s[256];
    const float *row = m.st.mpart + (size_t)e.y * m.st.strips;
    float v = 0.f;
    for (int t = threadIdx.x; t < m.st.strips; t += blockDim.x)
        v = fmaxf(v, row[t]);
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        m.mreg[e.y] = s[0];
}

// End of a tick.
static __global__ void k_tick(long long *tick_p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *tick_p += 1;
}

} // namespace tlm
