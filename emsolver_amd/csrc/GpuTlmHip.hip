// GpuTlmHip.hip - C ABI of the gfx950 TLM step (include/GpuTlmHip.h).
//
// Replaces the OpenGL set-up and per-step dispatch of GpuTlm::run
// (GpuTlm.cpp:532-884 in the reference) with HIP: one stream per context,
// device-resident per-step tables, two launches per step (inject, then the
// fused scatter/connect whose extra workgroups take the monitor records; one
// where the step adds the sources itself) and no host synchronisation inside
// a run.
#include "tlm_context.h"

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local std::string g_create_error = "no error";

struct DevBuf
{
    void *p = nullptr;
    size_t bytes = 0;
};

} // namespace


namespace tlmi {

int fail(tlm_ctx *c, int code, const std::string &msg)
{
    if (c)
        c->err = msg;
    else
        g_create_error = msg;
    return code;
}



// Launch shape: a wave task is one 64-cell strip of a z-slab marched over
// zc slabs (tuning.planes; 0 picks zc for about 64K tasks). Task groups of
// block_waves tasks go to workgroups; tuning.grid_blocks > 0 caps the grid
// (persistent workgroups walking their XCD's range of groups).
int step_occupancy(const tlm_ctx *c);

void choose_shape(tlm_ctx *c)
{
    c->strips = (int)(c->Pp / tlm::kWave);
    if (c->tune.planes > 0)
        c->zc = std::min(c->tune.planes, c->nz);
    else
    {
        const long long target = 65536;
        int nch = (int)std::max(1LL, std::min<long long>(c->nz, target / c->strips));
        c->zc = (c->nz + nch - 1) / nch;
    }
    c->nzc = (c->nz + c->zc - 1) / c->zc;
    c->tasks = c->strips * c->nzc;
    // waves per workgroup: 4, or 2 when the step runs under the occupancy cap
    // (ms a launch at 4 / 2 / 1: 512^3 64 boxes 2.535 / 2.527 / 2.522, 384^3
    // 1.069 / 1.057 / 1.059, 512^3 free space 2.180 / 2.180 / 2.183; uncapped
    // 256^3 8 boxes 0.282 / 0.282 / 0.285)
    const int auto_waves = c->pf == TLM_PULSE_F32 && step_occupancy(c) > 0 ? 2 : 4;
    c->wpb = std::max(1, std::min(4, c->tune.block_waves > 0 ? c->tune.block_waves : auto_waves));
    c->groups = (c->tasks + c->wpb - 1) / c->wpb;
    c->blocks = c->tune.grid_blocks > 0 ? std::min(c->groups, c->tune.grid_blocks) : c->groups;
    // One workgroup per group: XCD x takes chunks x, x + 8, ... of half a
    // z-slab's (or z-chunk's) groups, so that each XCD's share holds an even
    // sample of the grid's depth and no XCD is left with most of its
    // dielectric blocks (the step ends with the last XCD); a chunk's groups
    // stay on one XCD, so the slots that two waves share (strip seams, rows)
    // meet in one L2. A persistent grid keeps contiguous ranges.
    c->xchunk = 0;
    if (c->tune.grid_blocks <= 0)
    {
        c->xchunk = std::max(1, std::min(c->strips / c->wpb / 2, (c->groups + 7) / 8));
        c->blocks = 8 * tlm::xcd_chunk_map(c->groups, c->xchunk).cpx;
    }
}

size_t energy_parts(const tlm_ctx *c)
{
    return (size_t)c->blocks * c->wpb;
}

// The fused step launch; nmon > 0 appends the workgroups that take the
// step's monitor records. With events e0 / e1 the launch carries them as its
// own start / end timestamps (hipExtLaunchKernel): a marker recorded on the
// stream between two kernels would add a gap of ~5 us before and after the
// step, which hipEventRecord on either side of it measured.
// Dynamic LDS per workgroup that holds the step kernel to `occ` waves per
// SIMD (workgroups of wpb waves, 160 KB of LDS per CU): the kernel uses no
// LDS, the share only limits how many of its workgroups a CU takes. On
// large FP32 grids 5 waves per SIMD keep HBM as busy as more do and contend
// less (512^3 steps: plain air 2.219 against 2.275 ms, 64 dielectric boxes
// 2.684 against 2.745 ms); at 256^3 a cap costs 1 % and 16-bit steps gain
// nothing, so those run uncapped.
constexpr int kLdsPerCu = 160 * 1024;
int lds_share(int occ, int wpb)
{
    if (occ <= 0)
        return 0;
    const int wg = std::max(1, occ * 4 / wpb); // workgroups per CU
    return (kLdsPerCu / wg) & ~1023;           // (wg + 1) shares no longer fit
}

constexpr long long kOccCapCells = 32LL << 20;
int step_occupancy(const tlm_ctx *c)
{
    if (c->tune.occupancy >= 0)
        return c->tune.occupancy;
    return c->pf == TLM_PULSE_F32 && (long long)c->nx * c->ny * c->nz >= kOccCapCells ? 5 : 0;
}

template <bool S, bool E, int PF>
void launch_fused(tlm_ctx *c, const tlm::StepArgs &a, const tlm::MonitorArgs &m, int nmon,
                  hipEvent_t e0, hipEvent_t e1)
{
    const int bt = c->wpb * tlm::kWave;
    const int mon_blocks = nmon > 0 ? (nmon + bt - 1) / bt : 0;
    dim3 grid(c->blocks + mon_blocks), block(bt);
    const int shm = lds_share(step_occupancy(c), c->wpb);
    // automatic: non-temporal pulse stores when the grid has stub cells
    // (faster there by 1-3 %; slower by 3-4 % in plain air)
    const bool nt = c->tune.nontemporal == 1 || (c->tune.nontemporal == 2 && c->stubs);
    auto go = [&](auto kern) {
        if (e0)
            hipExtLaunchKernelGGL(kern, grid, block, shm, c->stream, e0, e1, 0, a, m, c->blocks);
        else
            hipLaunchKernelGGL(kern, grid, block, shm, c->stream, a, m, c->blocks);
    };
    if (PF != tlm::PF_F32)
        go(tlm::k_step<S, E, false, PF>);
    else if (a.fold.ninj > 0)
        nt ? go(tlm::k_step<S, E, true, PF, true>) : go(tlm::k_step<S, E, false, PF, true>);
    else
        nt ? go(tlm::k_step<S, E, true, PF>) : go(tlm::k_step<S, E, false, PF>);
}

template <int PF>
void launch_pf(tlm_ctx *c, const tlm::StepArgs &a, bool energy, const tlm::MonitorArgs &m, int nmon,
               hipEvent_t e0, hipEvent_t e1)
{
    if (c->stubs)
        energy ? launch_fused<true, true, PF>(c, a, m, nmon, e0, e1)
               : launch_fused<true, false, PF>(c, a, m, nmon, e0, e1);
    else
        energy ? launch_fused<false, true, PF>(c, a, m, nmon, e0, e1)
               : launch_fused<false, false, PF>(c, a, m, nmon, e0, e1);
}


// FP32 contexts with a few source cells may add them inside the step
// instead of k_inject. That saves the injection launch: 64^3 free space 7.2
// against 10.6 us a step, 128^3 dielectric boxes 40.9 against 44.0 us. The
// stub kernel compiled with the folded code takes fewer registers, which
// lets more of its waves in: uncapped that is slower on mid-size grids
// (256^3 boxes: 0.2859 against 0.2826 ms a launch), under the occupancy cap
// of large grids it is faster (384^3 boxes 1.019 against 1.026 ms, 512^3
// 2.558 against 2.569 ms, the injection launch saved on top). So by default
// grids with stub cells fold up to kFoldAutoCells cells or when the step
// runs capped, grids without stub cells at any size (512^3 free space:
// equal). tuning.fold: 0 never, 1 always, 2 by that rule.
constexpr long long kFoldAutoCells = 4LL << 20;
bool fold_active(const tlm_ctx *c)
{
    if (c->pf != TLM_PULSE_F32 || c->nuniq <= 0 || c->nuniq > tlm::kFoldMax)
        return false;
    if (c->tune.fold == 1)
        return true;
    return c->tune.fold == 2 && (!c->stubs || (long long)c->nx * c->ny * c->nz <= kFoldAutoCells ||
                                 step_occupancy(c) > 0);
}

tlm::Fold fold_of(const tlm_ctx *c, int n)
{
    tlm::Fold f;
    if (fold_active(c))
    {
        f.cell = c->ucell;
        f.beg = c->ubeg;
        f.pol = c->spol;
        f.amp = c->samp;
        f.ninj = c->nuniq;
        f.sv = c->sv_host[n];
    }
    return f;
}

int launch_step(tlm_ctx *c, int n, const tlm::MonitorArgs &m, int nmon, hipEvent_t e0, hipEvent_t e1)
{
    const bool energy = c->every > 0 && (n % c->every) == 0;
    tlm::StepArgs a;
    a.A = c->A;
    a.B = c->B;
    a.Vs = c->Vs;
    a.code = c->code;
    a.uair = c->uair;
    a.props = c->props;
    a.epart = c->epart;
    a.slab = c->slab;
    a.dsc = nullptr;
    a.esc = nullptr;
    a.mpart = c->mpart;
    if (c->pf != TLM_PULSE_F32)
    {
        const int cur = scale_buf(c, n);
        a.dsc = c->scl[cur];
        a.esc = c->iscl[cur ^ 1];
    }
    a.nx = c->nx;
    a.ny = c->ny;
    a.nz = c->nz;
    a.nxy = (int)c->nxy;
    a.Pp = c->Pp;
    a.P = c->P;
    a.PB = c->PB;
    a.h0 = c->h0;
    a.h1 = c->h1;
    a.strips = c->strips;
    a.zc = c->zc;
    a.nzc = c->nzc;
    a.tasks = c->tasks;
    a.groups = c->groups;
    a.rho = c->rho;
    a.nmat = c->nmat;
    a.mstrips = tlm::magic_div((unsigned)c->strips);
    a.mnx = tlm::magic_div((unsigned)c->nx);
    a.xm = c->xchunk > 0 ? tlm::xcd_chunk_map(c->groups, c->xchunk) : tlm::xcd_map(c->blocks, c->groups);
    a.fold = fold_of(c, n);
    if (c->pf == TLM_PULSE_F16)
        launch_pf<tlm::PF_F16>(c, a, energy, m, nmon, e0, e1);
    else if (c->pf == TLM_PULSE_BF16)
        launch_pf<tlm::PF_BF16>(c, a, energy, m, nmon, e0, e1);
    else
        launch_pf<tlm::PF_F32>(c, a, energy, m, nmon, e0, e1);
    if (c->pf != TLM_PULSE_F32)
        hipLaunchKernelGGL(tlm::k_slab_max, dim3(c->nz), dim3(256), 0, c->stream, c->mpart,
                           c->strips, c->mreg);
    return TLM_OK;
}

void queue_probe_dft(tlm_ctx *c, int n0, int n1, hipStream_t s)
{
    if (c->npr <= 0 || c->nfreq <= 0 || n1 <= n0)
        return;
    const int tot = c->npr * 3 * c->nfreq, b = 128;
    hipLaunchKernelGGL(tlm::k_probe_dft, dim3((tot + b - 1) / b), dim3(b), 0, s, c->prrec, c->npr,
                       c->ctab, c->stab, c->nfreq, n0, n1, c->dft_re, c->dft_im);
    c->dft_done = n1;
}

tlm::PulseView pulse_view(const tlm_ctx *c, int n)
{
    tlm::PulseView v;
    v.base = c->A;
    v.PB = c->PB;
    v.Pp = c->Pp;
    v.pf = c->pf;
    v.slab = c->slab;
    v.h0 = c->h0;
    v.h1 = c->h1;
    v.sc = nullptr;
    v.isc = nullptr;
    if (c->pf != TLM_PULSE_F32)
    {
        const int cur = scale_buf(c, n);
        v.sc = c->scl[cur];
        v.isc = c->iscl[cur];
    }
    return v;
}


// kernel_every > 0: the step launches i = 0, k, 2k, ... of the block carry
// timing events; their summed duration goes to *ms_kernel, their count to
// *nkernel.
int queue_steps(tlm_ctx *c, int step0, int nsteps, int kernel_every, float *ms_kernel, int *nkernel)
{
    if (c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member (tlm_batch_remove it first)");
    if (!c->uploaded)
        return fail(c, TLM_ERR_STATE, "tlm_run before tlm_upload");
    if (nsteps < 0 || step0 < 0)
        return fail(c, TLM_ERR_ARG, "negative step range");
    if (nsteps == 0)
        return TLM_OK;
    if (!c->sv || step0 + nsteps > c->nsteps)
        return fail(c, TLM_ERR_STATE, "step range exceeds the schedule set by tlm_set_schedule");
    HIPCHK(c, hipSetDevice(c->dev));
    if (c->pf != TLM_PULSE_F32)
    {
        if (!c->materialized)
        {
            int rc = materialize(c, step0);
            if (rc)
                return rc;
        }
        else if (step0 != c->next_step)
            return fail(c, TLM_ERR_STATE, "16-bit pulse storage runs contiguous step blocks");
    }
    const bool fold = fold_active(c);
    if (fold && !c->fold_marked)
    {
        hipLaunchKernelGGL(tlm::k_mark_fold, dim3(1), dim3(64), 0, c->stream, c->ucell, c->nuniq,
                           c->uair);
        c->fold_marked = true;
    }
    const int nmon0 = c->nf + c->nh + c->np + c->npr;
    const bool time_kernel = kernel_every > 0;
    const int ntimed = time_kernel ? (nsteps + kernel_every - 1) / kernel_every : 0;
    if (time_kernel)
    {
        // the sampled launches carry a start and an end event; no host wait
        // inside the loop
        while ((int)c->kev.size() < 2 * ntimed)
        {
            hipEvent_t ev;
            HIPCHK(c, hipEventCreate(&ev));
            c->kev.push_back(ev);
        }
    }
    for (int n = step0; n < step0 + nsteps; ++n)
    {
        const tlm::PulseView pv = pulse_view(c, n);
        tlm::PlanArgs pl = {};
        if (c->pf != TLM_PULSE_F32)
        {
            const int cur = scale_buf(c, n);
            pl.nz = c->nz;
            pl.mreg = c->mreg;
            pl.Q = c->qinj;
            pl.slab = c->slab;
            pl.sv = c->sv;
            pl.n = n;
            pl.nsteps = c->nsteps;
            pl.rp = std::max(1.0f, std::fabs(c->rho));
            pl.fmax = format_max(c->pf);
            pl.isc_cur = c->iscl[cur];
            pl.sc_next = c->scl[cur ^ 1];
            pl.isc_next = c->iscl[cur ^ 1];
            pl.flag = c->flag;
        }
        if ((fold ? 0 : c->nuniq) + pl.nz > 0)
        {
            int b = 128, g = (c->nuniq + pl.nz + b - 1) / b;
            hipLaunchKernelGGL(tlm::k_inject, dim3(g), dim3(b), 0, c->stream, pv, c->ucell,
                               c->ubeg, c->spol, c->samp, c->nuniq, c->sv, n, pl);
        }
        const bool plane_now = c->pa >= 0 && (n % 2) == 0;
        const int pn = plane_now ? c->pn1 * c->pn2 : 0;
        const int nmon = nmon0 + pn;
        tlm::MonitorArgs m = {};
        if (nmon > 0)
        {
            m.pn = pn;
            m.pa = c->pa;
            m.pidx = c->pidx;
            m.pn1 = c->pn1;
            m.pn2 = c->pn2;
            m.nx = c->nx;
            m.ny = c->ny;
            m.Pp = c->Pp;
            m.code = c->code;
            m.plane = c->plane;
            m.pv = pv;
            m.pv.fold = fold_of(c, n);
            m.n = n;
            m.settle = c->settle;
            m.cph = c->cph;
            m.sph = c->sph;
            m.nf = c->nf;
            m.fcell = c->fcell;
            m.faxis = c->faxis;
            m.js_inst = c->js_inst;
            m.js_dft4 = c->js_dft4;
            m.nh = c->nh;
            m.hcell = c->hcell;
            m.haxis = c->haxis;
            m.huy8 = c->huy8;
            m.np = c->np;
            m.pcells = c->pcells;
            m.poff = c->poff;
            m.ppol = c->ppol;
            m.ploop = c->ploop;
            m.prec = c->prec;
            m.npr = c->npr;
            m.prcell = c->prcell;
            m.inv_dl = c->inv_dl;
            m.prrec = c->prrec;
            if (plane_now)
                c->plane_step = n;
        }
        // the monitor records of step n run in the step launch's extra workgroups
        const int ti = n - step0;
        if (time_kernel && ti % kernel_every == 0)
            launch_step(c, n, m, nmon, c->kev[2 * (ti / kernel_every)], c->kev[2 * (ti / kernel_every) + 1]);
        else
            launch_step(c, n, m, nmon, nullptr, nullptr);
        if (c->frame_every > 0 && (n % c->frame_every) == 0 &&
            (int)c->frame_steps.size() < c->frame_max)
        {
            // TlmSolver.cpp:747-763: the frame keeps |Js| of this step and the latest plane
            const size_t f = c->frame_steps.size();
            if (c->nf > 0)
                HIPCHK(c, hipMemcpyAsync(c->frame_js + f * c->nf, c->js_inst, c->nf * sizeof(float),
                                         hipMemcpyDeviceToDevice, c->stream));
            const size_t np2 = (size_t)c->pn1 * c->pn2;
            if (c->pa >= 0 && c->plane_step >= 0)
                HIPCHK(c, hipMemcpyAsync(c->frame_plane + f * np2, c->plane, np2 * sizeof(float),
                                         hipMemcpyDeviceToDevice, c->stream));
            c->frame_steps.push_back(n);
            c->frame_plane_step.push_back(c->pa >= 0 ? c->plane_step : -1);
        }
        if (c->every > 0 && (n % c->every) == 0)
        {
            const int idx = n / c->every;
            const long long np = (long long)energy_parts(c);
            if (np > 8192)
            {
                // two stages: 1024 fixed chunks, then one block
                const long long chunk = (np + 1023) / 1024;
                const int nb = (int)((np + chunk - 1) / chunk);
                hipLaunchKernelGGL(tlm::k_energy_chunks, dim3(nb), dim3(256), 0, c->stream,
                                   c->epart, np, chunk, c->epart2);
                hipLaunchKernelGGL(tlm::k_energy_reduce, dim3(1), dim3(256), 0, c->stream,
                                   c->epart2, nb, c->energy + idx);
            }
            else
                hipLaunchKernelGGL(tlm::k_energy_reduce, dim3(1), dim3(256), 0, c->stream,
                                   c->epart, (int)np, c->energy + idx);
            if ((int)c->energy_steps.size() <= idx)
                c->energy_steps.resize(idx + 1, -1);
            c->energy_steps[idx] = n;
        }
        std::swap(c->A, c->B);
    }
    HIPCHK(c, hipGetLastError());
    // running probe DFT over the newly recorded steps, in step order
    if (c->npr > 0 && c->nfreq > 0)
    {
        if (step0 != c->dft_done)
            return fail(c, TLM_ERR_STATE, "probe DFT needs contiguous step blocks from step 0");
        queue_probe_dft(c, step0, step0 + nsteps, c->stream);
    }
    c->steps_run_hi = std::max(c->steps_run_hi, step0 + nsteps);
    c->next_step = step0 + nsteps;
    if (time_kernel && ms_kernel)
    {
        float kern_ms = 0.f;
        HIPCHK(c, hipEventSynchronize(c->kev[2 * ntimed - 1]));
        for (int s = 0; s < ntimed; ++s)
        {
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->kev[2 * s], c->kev[2 * s + 1]));
            kern_ms += ms;
        }
        *ms_kernel = kern_ms;
        if (nkernel)
            *nkernel = ntimed;
    }
    HIPCHK(c, hipGetLastError());
    return TLM_OK;
}

// Chunked conversion between host cell order and the device layout through
// a bounded staging buffer: pulses (ncomp 12, any storage format, through
// pv) or FP32 stubs (ncomp 3, 64-cell stub blocks at soa).
int convert(tlm_ctx *c, const float *host_in, float *host_out, const tlm::PulseView *pv,
            float *soa, int ncomp)
{
    const long long ncell = (long long)c->nx * c->ny * c->nz;
    const long long chunk = std::min<long long>(ncell, (long long)8 << 20); // 8 M cells
    float *stage = nullptr;
    int rc = dalloc(c, &stage, (size_t)chunk * ncomp);
    if (rc)
        return rc;
    for (long long c0 = 0; c0 < ncell; c0 += chunk)
    {
        long long cnt = std::min(chunk, ncell - c0);
        int b = 256;
        long long g = (cnt + b - 1) / b;
        if (host_in)
        {
            hipError_t e = hipMemcpyAsync(stage, host_in + c0 * ncomp, cnt * ncomp * sizeof(float),
                                          hipMemcpyHostToDevice, c->stream);
            if (e != hipSuccess)
            {
                dfree(c, stage);
                return hip_fail(c, e, "hipMemcpyAsync");
            }
            if (pv)
                hipLaunchKernelGGL(tlm::k_pulses_in, dim3((unsigned)g), dim3(b), 0, c->stream, stage,
                                   *pv, c->nxy, c0, cnt);
            else
                hipLaunchKernelGGL(tlm::k_stubs_in, dim3((unsigned)g), dim3(b), 0, c->stream, stage,
                                   soa, c->nxy, c->Pp, c0, cnt);
        }
        else
        {
            if (pv)
                hipLaunchKernelGGL(tlm::k_pulses_out, dim3((unsigned)g), dim3(b), 0, c->stream, *pv,
                                   stage, c->nxy, c0, cnt);
            else
                hipLaunchKernelGGL(tlm::k_stubs_out, dim3((unsigned)g), dim3(b), 0, c->stream, soa,
                                   stage, c->nxy, c->Pp, c0, cnt);
            hipError_t e = hipMemcpyAsync(host_out + c0 * ncomp, stage, cnt * ncomp * sizeof(float),
                                          hipMemcpyDeviceToHost, c->stream);
            if (e != hipSuccess)
            {
                dfree(c, stage);
                return hip_fail(c, e, "hipMemcpyAsync");
            }
        }
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
        {
            dfree(c, stage);
            return hip_fail(c, e, "hipStreamSynchronize");
        }
    }
    dfree(c, stage);
    return TLM_OK;
}

// 16-bit formats, first run after tlm_upload: decide which slabs keep FP32
// (within `halo` slabs of a soft source), lay the slabs out, choose the
// first scales from the uploaded state and the first injection
// (E_k = exponent of max|state in slab k| + Q_k*|sv(step0)|), and store the
// state rounded to those scales.
int materialize(tlm_ctx *c, int step0)
{
    const int nz = c->nz;
    const long long slab = c->nxy;
    std::vector<float> Q(nz, 0.0f);
    std::vector<char> hot(nz, 0);
    int klo = nz, khi = -1;
    for (size_t s = 0; s < c->src_cells.size(); ++s)
    {
        const int k = (int)(c->src_cells[s] / slab);
        Q[k] = Q[k] + 0.5f * std::fabs(c->src_amp[s]); // list order, as the oracle adds them
        klo = std::min(klo, k);
        khi = std::max(khi, k);
    }
    // one band of FP32 slabs: halo below the lowest source slab to halo above
    // the highest, so the step computes slab offsets without a table lookup
    c->h0 = c->h1 = 0;
    if (c->halo >= 0 && khi >= 0)
    {
        c->h0 = std::max(0, klo - c->halo);
        c->h1 = std::min(nz - 1, khi + c->halo) + 1;
        for (int k = c->h0; k < c->h1; ++k)
            hot[k] = 1;
    }
    std::vector<uint2> sl(nz);
    c->nhot = 0;
    for (int k = 0; k < nz; ++k)
    {
        // 64-cell blocks of 6 pairs: 24 bytes per cell in a 16-bit slab, 48 in an FP32 one
        sl[k] = make_uint2(0u, hot[k] ? 0u : 1u);
        c->nhot += hot[k];
    }
    if ((unsigned long long)c->Pp * 48ULL > (1ULL << 32) - (1ULL << 20))
        return fail(c, TLM_ERR_UNSUPPORTED, "16-bit storage: a slab exceeds 32-bit byte offsets");
    c->PB = (long long)tlm::slab16_base(c->Pp, nz, c->h0, c->h1); // bytes of one buffer
    // slab maxima of the uploaded state
    std::vector<float> M(nz, 0.0f);
    if (!c->pendV.empty())
        for (int k = 0; k < nz; ++k)
        {
            float m = 0.0f;
            const float *v = c->pendV.data() + (size_t)k * slab * 12;
            for (long long i = 0; i < slab * 12; ++i)
                m = std::max(m, std::fabs(v[i]));
            if (!c->pendVs.empty())
            {
                const float *w = c->pendVs.data() + (size_t)k * slab * 3;
                for (long long i = 0; i < slab * 3; ++i)
                    m = std::max(m, std::fabs(w[i]));
            }
            M[k] = m;
        }
    float sv0 = 0.0f;
    HIPCHK(c, hipMemcpy(&sv0, c->sv + step0, sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> sc(nz), isc(nz);
    for (int k = 0; k < nz; ++k)
    {
        const int E = tlm::plan_exponent(M[k] + Q[k] * std::fabs(sv0));
        sc[k] = tlm::pow2_float(E);
        isc[k] = tlm::pow2_float(-E);
    }
    for (char **buf : {&c->A, &c->B})
    {
        dfree(c, *buf);
        int rc = dalloc(c, buf, (size_t)c->PB);
        if (rc)
            return rc;
    }
    int rc;
    if ((rc = upload(c, &c->slab, sl.data(), sl.size())) ||
        (rc = upload(c, &c->scl[0], sc.data(), sc.size())) ||
        (rc = upload(c, &c->iscl[0], isc.data(), isc.size())) ||
        (rc = upload(c, &c->scl[1], sc.data(), sc.size())) ||
        (rc = upload(c, &c->iscl[1], isc.data(), isc.size())) ||
        (rc = upload(c, &c->mreg, M.data(), M.size())) ||
        (rc = upload(c, &c->qinj, Q.data(), Q.size())) ||
        (rc = upload<float>(c, &c->mpart, nullptr, (size_t)c->strips * nz)) ||
        (rc = upload<int>(c, &c->flag, nullptr, 1)))
        return rc;
    HIPCHK(c, hipMemsetAsync(c->flag, 0, sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->A, 0, (size_t)c->PB, c->stream));
    HIPCHK(c, hipMemsetAsync(c->B, 0, (size_t)c->PB, c->stream));
    c->base = step0;
    c->materialized = true;
    if (!c->pendV.empty())
    {
        const tlm::PulseView pv = pulse_view(c, step0);
        if ((rc = convert(c, c->pendV.data(), nullptr, &pv, nullptr, 12)))
            return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->pendV.clear();
    c->pendV.shrink_to_fit();
    c->pendVs.clear();
    c->pendVs.shrink_to_fit();
    return TLM_OK;
}

} // namespace tlmi

using namespace tlmi;

extern "C" {

int tlm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

const char *tlm_device_name(int device)
{
    static thread_local char name[256];
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess)
        return "unknown adapter";
    std::snprintf(name, sizeof(name), "%s (%s, %d CUs)", p.name[0] ? p.name : "AMD GPU",
                  p.gcnArchName, p.multiProcessorCount);
    return name;
}

int tlm_create(const tlm_grid_desc *d, tlm_ctx **out)
{
    if (!out)
        return fail(nullptr, TLM_ERR_ARG, "tlm_create: out is NULL");
    *out = nullptr;
    if (!d)
        return fail(nullptr, TLM_ERR_ARG, "tlm_create: desc is NULL");
    if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0)
        return fail(nullptr, TLM_ERR_ARG, "empty grid");
    if (d->pulse_type != TLM_PULSE_F32 && d->pulse_type != TLM_PULSE_F16 &&
        d->pulse_type != TLM_PULSE_BF16)
        return fail(nullptr, TLM_ERR_ARG, "unknown pulse storage type");
    {
        // one pulse plane must be addressable with 32-bit byte offsets
        const long long pp = ((long long)d->nx * d->ny + 63) / 64 * 64;
        if (pp * (long long)d->nz * 4 > (1LL << 32) - (1LL << 20))
            return fail(nullptr, TLM_ERR_UNSUPPORTED,
                        "grid above ~1.07e9 cells: a pulse plane exceeds 32-bit byte offsets");
        // the FP32 step addresses a slab's 12 ports (48 bytes per cell) with
        // 32-bit byte offsets from the slab base
        if (pp * 48 > (1LL << 32) - (1LL << 20))
            return fail(nullptr, TLM_ERR_UNSUPPORTED,
                        "z-slab above ~89e6 cells: its pulses exceed 32-bit byte offsets");
    }
    if (!(d->dl > 0.0f))
        return fail(nullptr, TLM_ERR_ARG, "cell size must be positive");
    int ndev = tlm_device_count();
    if (ndev <= 0)
        return fail(nullptr, TLM_ERR_NO_DEVICE, "no HIP device available");
    if (d->device < 0 || d->device >= ndev)
        return fail(nullptr, TLM_ERR_ARG, "device ordinal out of range");

    tlm_ctx *c = new tlm_ctx;
    c->dev = d->device;
    c->nx = d->nx;
    c->ny = d->ny;
    c->nz = d->nz;
    c->nxy = (long long)c->nx * c->ny;
    c->Pp = (c->nxy + tlm::kWave - 1) / tlm::kWave * tlm::kWave;
    c->P = c->Pp * c->nz;
    c->dl = d->dl;
    c->rho = d->rho;
    c->pf = d->pulse_type;
    // 16-bit pair planes: slabs per wave task, measured (ms a launch at 3 / 4
    // slabs): 64-box scenes 512^3 1.591 / 1.606, 384^3 0.681 / 0.687, 320^3
    // 0.350 / 0.342, 8-box 256^3 0.193 / 0.183
    if (c->pf != TLM_PULSE_F32)
        c->tune.planes = (long long)c->nx * c->ny * c->nz >= (48LL << 20) ? 3 : 4;
    c->halo = d->fp32_halo < 0 ? -1 : (d->fp32_halo == 0 ? 8 : d->fp32_halo);
    c->PB = c->P * 4;
    c->inv_dl = -1.0f / d->dl;
    choose_shape(c);
    auto bail = [&](int rc) {
        g_create_error = c->err;
        tlm_destroy(c);
        return rc;
    };
    hipError_t e = hipSetDevice(c->dev);
    if (e != hipSuccess)
        return bail(hip_fail(c, e, "hipSetDevice"));
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(c, e, "hipStreamCreate"));
    for (hipEvent_t *ev : {&c->ev0, &c->ev1, &c->evk0, &c->evk1})
        if ((e = hipEventCreate(ev)) != hipSuccess)
            return bail(hip_fail(c, e, "hipEventCreate"));
    int rc;
    c->epart_cap = energy_parts(c);
    if (c->pf == TLM_PULSE_F32 &&
        ((rc = dalloc(c, &c->A, (size_t)c->PB * 12)) || (rc = dalloc(c, &c->B, (size_t)c->PB * 12))))
        return bail(rc);
    if ((rc = dalloc(c, &c->code, (size_t)c->P)) || (rc = dalloc(c, &c->epart, c->epart_cap)) ||
        (rc = dalloc(c, &c->epart2, 1024)))
        return bail(rc);
    *out = c;
    return TLM_OK;
}

void tlm_destroy(tlm_ctx *c)
{
    if (!c)
        return;
    if (c->batch)
        (void)tlm_batch_remove(c->batch, c);
    (void)hipSetDevice(c->dev);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    for (void *p : c->owned)
        (void)hipFree(p);
    c->owned.clear();
    for (hipEvent_t ev : {c->ev0, c->ev1, c->evk0, c->evk1})
        if (ev)
            (void)hipEventDestroy(ev);
    for (hipEvent_t ev : c->kev)
        (void)hipEventDestroy(ev);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *tlm_last_error(const tlm_ctx *c)
{
    return c ? c->err.c_str() : g_create_error.c_str();
}

int tlm_upload(tlm_ctx *c, const uint8_t *mat, const float *yhat, const float *ghat, int nmat,
               const float *V, const float *Vs)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || !mat || nmat < 0 || (nmat > 0 && (!yhat || !ghat)))
        return fail(c, TLM_ERR_ARG, "tlm_upload: bad arguments");
    HIPCHK(c, hipSetDevice(c->dev));
    const long long ncell = (long long)c->nx * c->ny * c->nz;
    bool any_diel = false;
    int max_id = -1;
    for (long long i = 0; i < ncell; ++i)
        if (mat[i] >= 2)
        {
            any_diel = true;
            max_id = std::max(max_id, (int)mat[i] - 2);
        }
    if (max_id >= nmat)
        return fail(c, TLM_ERR_ARG, "material id without a yhat/ghat entry");
    // slab-padded cell codes: material id | neighbour-PEC bits; padding is
    // marked PEC and is never anybody's neighbour
    std::vector<uint16_t> code((size_t)c->P, 1);
    {
        const int nx = c->nx, ny = c->ny, nz = c->nz;
        const long long sy = nx, sz = c->nxy;
        for (long long k = 0; k < nz; ++k)
            for (long long j = 0; j < ny; ++j)
            {
                const uint8_t *row = mat + k * sz + j * sy;
                uint16_t *out = &code[(size_t)(k * c->Pp + j * sy)];
                for (long long i = 0; i < nx; ++i)
                {
                    unsigned v = row[i];
                    if (i > 0 && row[i - 1] == 1)
                        v |= tlm::NB_XM;
                    if (i < nx - 1 && row[i + 1] == 1)
                        v |= tlm::NB_XP;
                    if (j > 0 && row[i - sy] == 1)
                        v |= tlm::NB_YM;
                    if (j < ny - 1 && row[i + sy] == 1)
                        v |= tlm::NB_YP;
                    if (k > 0 && row[i - sz] == 1)
                        v |= tlm::NB_ZM;
                    if (k < nz - 1 && row[i + sz] == 1)
                        v |= tlm::NB_ZP;
                    out[i] = (uint16_t)v;
                }
            }
    }
    HIPCHK(c, hipMemcpy(c->code, code.data(), code.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    {
        // bit 0: blocks whose 64 codes are all 0 (the slab padding is coded
        // PEC, so a block with padding never qualifies); bit 2: blocks with a
        // dielectric cell (they load their stubs with their pulses); bits 3
        // and 4: cells 0-31 / 32-63 of the block hold a dielectric cell;
        // bit 5 (kBlockClean): no cell is PEC or padding or has a PEC neighbour
        const size_t nblk = (size_t)c->P / tlm::kWave;
        std::vector<uint8_t> ua((nblk + 3) / 4 * 4, 0); // padded for block_flag's dword loads
        for (size_t b = 0; b < nblk; ++b)
        {
            const uint16_t *q = &code[b * tlm::kWave];
            uint16_t any = 0;
            unsigned half = 0;
            bool clean = true;
            for (int l = 0; l < tlm::kWave; ++l)
            {
                any |= q[l];
                if ((q[l] & 0xFFu) >= 2u)
                    half |= 8u << (l >> 5);
                clean = clean && tlm::code_is_clean(q[l]);
            }
            ua[b] = (uint8_t)((any == 0 ? 1u : 0u) | (half ? 4u : 0u) | half |
                              (clean ? tlm::kBlockClean : 0u));
        }
        int rc0 = upload(c, &c->uair, ua.data(), ua.size());
        c->fold_marked = false;
        if (rc0)
            return rc0;
    }
    return install_material(c, any_diel, yhat, ghat, nmat, V, Vs);
}

} // extern "C"

namespace tlmi {

int install_material(tlm_ctx *c, bool any_diel, const float *yhat, const float *ghat, int nmat,
                     const float *V, const float *Vs)
{
    const long long ncell = (long long)c->nx * c->ny * c->nz;
    // per-material (yh, gh, den), den = 4 + yh + gh as in TlmSolver.cpp:90
    std::vector<float4> props(std::max(1, nmat));
    for (int i = 0; i < nmat; ++i)
    {
        float den = 4.0f + yhat[i] + ghat[i];
        props[i] = make_float4(yhat[i], ghat[i], den, 0.f);
    }
    int rc = upload(c, &c->props, props.data(), props.size());
    if (rc)
        return rc;
    c->nmat = nmat;
    c->stubs = any_diel;
    if (c->pf == TLM_PULSE_F32)
    {
        HIPCHK(c, hipMemsetAsync(c->A, 0, (size_t)c->PB * 12, c->stream));
        HIPCHK(c, hipMemsetAsync(c->B, 0, (size_t)c->PB * 12, c->stream));
    }
    dfree(c, c->Vs);
    if (c->stubs)
    {
        if ((rc = dalloc(c, &c->Vs, (size_t)c->P * 3)))
            return rc;
        HIPCHK(c, hipMemsetAsync(c->Vs, 0, (size_t)c->P * 3 * sizeof(float), c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->pf == TLM_PULSE_F32)
    {
        const tlm::PulseView pv = pulse_view(c, 0);
        if (V && (rc = convert(c, V, nullptr, &pv, nullptr, 12)))
            return rc;
    }
    else
    {
        // the 16-bit scales depend on the sources and the first step: the
        // state is stored at the first tlm_run
        c->materialized = false;
        c->pendV.clear();
        c->pendVs.clear();
        if (V)
        {
            c->pendV.assign(V, V + (size_t)ncell * 12);
            if (Vs && c->stubs)
                c->pendVs.assign(Vs, Vs + (size_t)ncell * 3);
        }
    }
    if (Vs && c->stubs && (rc = convert(c, Vs, nullptr, nullptr, c->Vs, 3)))
        return rc;
    c->uploaded = true;
    c->energy_steps.clear();
    c->steps_run_hi = 0;
    c->dft_done = 0;
    return TLM_OK;
}

} // namespace tlmi

extern "C" {

int tlm_set_sources(tlm_ctx *c, const int64_t *cells, const int32_t *pol, const float *amp, int n)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || n < 0 || (n > 0 && (!cells || !pol || !amp)))
        return fail(c, TLM_ERR_ARG, "tlm_set_sources: bad arguments");
    if (c->pf != TLM_PULSE_F32 && c->materialized)
        return fail(c, TLM_ERR_STATE,
                    "16-bit pulse storage: sources are fixed once the first step has run");
    // group by cell, keeping list order within a cell
    std::map<long long, std::vector<int>> by;
    for (int s = 0; s < n; ++s)
    {
        if (!cell_ok(c, cells[s]) || pol[s] < 0 || pol[s] > 2)
            return fail(c, TLM_ERR_ARG, "source cell or polarisation out of range");
        by[cells[s]].push_back(s);
    }
    std::vector<long long> uc;
    std::vector<int> ub, sp;
    std::vector<float> sa;
    for (auto &kv : by)
    {
        uc.push_back(padded_cell(c, kv.first));
        ub.push_back((int)sp.size());
        for (int s : kv.second)
        {
            sp.push_back(pol[s]);
            sa.push_back(amp[s]);
        }
    }
    ub.push_back((int)sp.size());
    c->src_cells.assign(cells, cells + n);
    c->src_amp.assign(amp, amp + n);
    int rc;
    if ((rc = upload(c, &c->ucell, uc.data(), uc.size())) ||
        (rc = upload(c, &c->ubeg, ub.data(), ub.size())) ||
        (rc = upload(c, &c->spol, sp.data(), sp.size())) ||
        (rc = upload(c, &c->samp, sa.data(), sa.size())))
        return rc;
    c->nuniq = (int)uc.size();
    c->fold_marked = false;
    return TLM_OK;
}

int tlm_set_schedule(tlm_ctx *c, const float *sv, const float *cph, const float *sph, int nsteps,
                     int settle)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || nsteps <= 0 || !sv || !cph || !sph)
        return fail(c, TLM_ERR_ARG, "tlm_set_schedule: bad arguments");
    int rc;
    if ((rc = upload(c, &c->sv, sv, nsteps)) || (rc = upload(c, &c->cph, cph, nsteps)) ||
        (rc = upload(c, &c->sph, sph, nsteps)))
        return rc;
    c->sv_host.assign(sv, sv + nsteps);
    c->nsteps = nsteps;
    c->settle = settle;
    c->nfreq = 0; // DFT tables are sized by the schedule: set them again
    c->dft_done = 0;
    const int ne = c->every > 0 ? nsteps / c->every + 1 : 1;
    dfree(c, c->energy);
    if ((rc = dalloc(c, &c->energy, ne)))
        return rc;
    c->energy_steps.clear();
    // record buffers that scale with the schedule length
    if (c->np > 0)
    {
        dfree(c, c->prec);
        if ((rc = dalloc(c, &c->prec, (size_t)nsteps * c->np * 2)))
            return rc;
        HIPCHK(c, hipMemset(c->prec, 0, (size_t)nsteps * c->np * 2 * sizeof(float)));
    }
    if (c->npr > 0)
    {
        dfree(c, c->prrec);
        if ((rc = dalloc(c, &c->prrec, (size_t)nsteps * c->npr * 3)))
            return rc;
        HIPCHK(c, hipMemset(c->prrec, 0, (size_t)nsteps * c->npr * 3 * sizeof(float)));
    }
    return TLM_OK;
}

int tlm_set_energy_every(tlm_ctx *c, int every)
{
    if (!c || every < 0)
        return fail(c, TLM_ERR_ARG, "tlm_set_energy_every: bad arguments");
    c->every = every;
    if (c->nsteps > 0)
    {
        dfree(c, c->energy);
        int rc = dalloc(c, &c->energy, every > 0 ? c->nsteps / every + 1 : 1);
        if (rc)
            return rc;
    }
    c->energy_steps.clear();
    return TLM_OK;
}

int tlm_set_plane(tlm_ctx *c, int axis, int index)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || axis > 2)
        return fail(c, TLM_ERR_ARG, "tlm_set_plane: bad arguments");
    dfree(c, c->plane);
    c->pa = axis < 0 ? -1 : axis;
    c->plane_step = -1;
    if (c->pa < 0)
        return TLM_OK;
    const int nn[3] = {c->nx, c->ny, c->nz};
    const int a1 = (axis == 0) ? 1 : 0, a2 = (axis == 2) ? 1 : 2;
    c->pn1 = nn[a1];
    c->pn2 = nn[a2];
    c->pidx = std::min(std::max(index, 0), nn[axis] - 1);
    int rc = dalloc(c, &c->plane, (size_t)c->pn1 * c->pn2);
    if (rc)
        return rc;
    HIPCHK(c, hipMemset(c->plane, 0, (size_t)c->pn1 * c->pn2 * sizeof(float)));
    return TLM_OK;
}

int tlm_set_frames(tlm_ctx *c, int every, int max_frames)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || every < 0 || max_frames < 0)
        return fail(c, TLM_ERR_ARG, "tlm_set_frames: bad arguments");
    dfree(c, c->frame_js);
    dfree(c, c->frame_plane);
    c->frame_steps.clear();
    c->frame_plane_step.clear();
    c->frame_every = (max_frames > 0) ? every : 0;
    c->frame_max = max_frames;
    if (c->frame_every == 0)
        return TLM_OK;
    int rc;
    if ((rc = dalloc(c, &c->frame_js, (size_t)max_frames * std::max(1, c->nf))))
        return rc;
    if (c->pa >= 0 &&
        (rc = dalloc(c, &c->frame_plane, (size_t)max_frames * c->pn1 * c->pn2)))
        return rc;
    return TLM_OK;
}

int tlm_read_plane(tlm_ctx *c, float *out, int *n1, int *n2, int *step)
{
    if (!c)
        return TLM_ERR_ARG;
    if (n1)
        *n1 = c->pa >= 0 ? c->pn1 : 0;
    if (n2)
        *n2 = c->pa >= 0 ? c->pn2 : 0;
    if (step)
        *step = c->pa >= 0 ? c->plane_step : -1;
    if (out && c->pa >= 0)
    {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(out, c->plane, (size_t)c->pn1 * c->pn2 * sizeof(float),
                            hipMemcpyDeviceToHost));
    }
    return TLM_OK;
}

int tlm_read_frames(tlm_ctx *c, int32_t *steps, float *js, float *planes, int cap)
{
    if (!c || cap < 0)
        return fail(c, TLM_ERR_ARG, "tlm_read_frames: bad arguments");
    const int cnt = (int)c->frame_steps.size();
    const int m = std::min(cap, cnt);
    if (m > 0)
    {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (steps)
            for (int f = 0; f < m; ++f)
                steps[f] = c->frame_steps[f];
        if (js && c->nf > 0)
            HIPCHK(c, hipMemcpy(js, c->frame_js, (size_t)m * c->nf * sizeof(float),
                                hipMemcpyDeviceToHost));
        if (planes && c->pa >= 0 && c->frame_plane)
            HIPCHK(c, hipMemcpy(planes, c->frame_plane, (size_t)m * c->pn1 * c->pn2 * sizeof(float),
                                hipMemcpyDeviceToHost));
    }
    return cnt;
}

int tlm_set_probes(tlm_ctx *c, const int64_t *cells, int n, float inv_dl)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || n < 0 || (n > 0 && !cells))
        return fail(c, TLM_ERR_ARG, "tlm_set_probes: bad arguments");
    if (c->nsteps <= 0)
        return fail(c, TLM_ERR_STATE, "tlm_set_probes before tlm_set_schedule");
    std::vector<long long> pc(n);
    for (int i = 0; i < n; ++i)
    {
        if (!cell_ok(c, cells[i]))
            return fail(c, TLM_ERR_ARG, "probe cell out of range");
        pc[i] = padded_cell(c, cells[i]);
    }
    int rc;
    if ((rc = upload(c, &c->prcell, pc.data(), pc.size())))
        return rc;
    c->npr = n;
    c->inv_dl = inv_dl;
    dfree(c, c->prrec);
    if ((rc = dalloc(c, &c->prrec, (size_t)c->nsteps * std::max(1, n) * 3)))
        return rc;
    HIPCHK(c, hipMemset(c->prrec, 0, (size_t)c->nsteps * std::max(1, n) * 3 * sizeof(float)));
    c->nfreq = 0;
    return TLM_OK;
}

int tlm_set_probe_dft(tlm_ctx *c, const double *ctab, const double *stab, int nfreq)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || nfreq < 0 || (nfreq > 0 && (!ctab || !stab)))
        return fail(c, TLM_ERR_ARG, "tlm_set_probe_dft: bad arguments");
    if (c->nsteps <= 0)
        return fail(c, TLM_ERR_STATE, "tlm_set_probe_dft before tlm_set_schedule");
    int rc;
    const size_t tab = (size_t)c->nsteps * nfreq;
    if ((rc = upload(c, &c->ctab, ctab, tab)) || (rc = upload(c, &c->stab, stab, tab)))
        return rc;
    const size_t acc = (size_t)std::max(1, c->npr) * 3 * std::max(1, nfreq);
    dfree(c, c->dft_re);
    dfree(c, c->dft_im);
    if ((rc = dalloc(c, &c->dft_re, acc)) || (rc = dalloc(c, &c->dft_im, acc)))
        return rc;
    HIPCHK(c, hipMemset(c->dft_re, 0, acc * sizeof(double)));
    HIPCHK(c, hipMemset(c->dft_im, 0, acc * sizeof(double)));
    c->nfreq = nfreq;
    c->dft_done = 0;
    return TLM_OK;
}

int tlm_set_ports(tlm_ctx *c, const int64_t *cells, const int32_t *off, const int32_t *pol, int np)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || np < 0 || (np > 0 && (!off || !pol)))
        return fail(c, TLM_ERR_ARG, "tlm_set_ports: bad arguments");
    if (c->nsteps <= 0)
        return fail(c, TLM_ERR_STATE, "tlm_set_ports before tlm_set_schedule");
    std::vector<long long> pc;
    std::vector<long long> loop(4 * (size_t)std::max(1, np), 0);
    std::vector<int> po(np + 1), pp(np);
    const long long stride[3] = {1, c->nx, (long long)c->nx * c->ny};
    for (int p = 0; p < np; ++p)
    {
        if (pol[p] < 0 || pol[p] > 2 || off[p + 1] < off[p])
            return fail(c, TLM_ERR_ARG, "bad port description");
        po[p] = (int)pc.size();
        pp[p] = pol[p];
        for (int s = off[p]; s < off[p + 1]; ++s)
        {
            if (!cell_ok(c, cells[s]))
                return fail(c, TLM_ERR_ARG, "port cell out of range");
            pc.push_back(padded_cell(c, cells[s]));
        }
        const int cnt = off[p + 1] - off[p];
        if (cnt > 0)
        {
            // Ampere loop neighbours by linear index, as TlmSolver.cpp:687-698
            long long mid = cells[off[p] + cnt / 2];
            int t1 = (pol[p] + 1) % 3, t2 = (pol[p] + 2) % 3;
            long long nb[4] = {mid + stride[t1], mid - stride[t1], mid + stride[t2], mid - stride[t2]};
            for (int q = 0; q < 4; ++q)
            {
                if (!cell_ok(c, nb[q]))
                    return fail(c, TLM_ERR_ARG, "port current loop leaves the grid");
                loop[4 * p + q] = padded_cell(c, nb[q]);
            }
        }
    }
    po[np] = (int)pc.size();
    int rc;
    if ((rc = upload(c, &c->pcells, pc.data(), pc.size())) ||
        (rc = upload(c, &c->poff, po.data(), po.size())) ||
        (rc = upload(c, &c->ppol, pp.data(), pp.size())) ||
        (rc = upload(c, &c->ploop, loop.data(), loop.size())))
        return rc;
    c->np = np;
    dfree(c, c->prec);
    if ((rc = dalloc(c, &c->prec, (size_t)c->nsteps * std::max(1, np) * 2)))
        return rc;
    HIPCHK(c, hipMemset(c->prec, 0, (size_t)c->nsteps * std::max(1, np) * 2 * sizeof(float)));
    return TLM_OK;
}

int tlm_set_faces(tlm_ctx *c, const int64_t *cells, const int32_t *axes, int n)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || n < 0 || (n > 0 && (!cells || !axes)))
        return fail(c, TLM_ERR_ARG, "tlm_set_faces: bad arguments");
    std::vector<long long> fc(n);
    std::vector<int> fa(n);
    for (int i = 0; i < n; ++i)
    {
        if (!cell_ok(c, cells[i]) || axes[i] < 0 || axes[i] > 2)
            return fail(c, TLM_ERR_ARG, "face cell or axis out of range");
        fc[i] = padded_cell(c, cells[i]);
        fa[i] = axes[i];
    }
    int rc;
    if ((rc = upload(c, &c->fcell, fc.data(), fc.size())) ||
        (rc = upload(c, &c->faxis, fa.data(), fa.size())))
        return rc;
    dfree(c, c->js_inst);
    dfree(c, c->js_dft4);
    if ((rc = dalloc(c, &c->js_inst, std::max(1, n))) ||
        (rc = dalloc(c, &c->js_dft4, 4 * (size_t)std::max(1, n))))
        return rc;
    HIPCHK(c, hipMemset(c->js_inst, 0, std::max(1, n) * sizeof(float)));
    HIPCHK(c, hipMemset(c->js_dft4, 0, 4 * (size_t)std::max(1, n) * sizeof(float)));
    c->nf = n;
    return TLM_OK;
}

int tlm_set_huygens(tlm_ctx *c, const int64_t *cells, const int32_t *axes, int n)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || n < 0 || (n > 0 && (!cells || !axes)))
        return fail(c, TLM_ERR_ARG, "tlm_set_huygens: bad arguments");
    std::vector<long long> hc(n);
    std::vector<int> ha(n);
    for (int i = 0; i < n; ++i)
    {
        if (!cell_ok(c, cells[i]) || axes[i] < 0 || axes[i] > 2)
            return fail(c, TLM_ERR_ARG, "Huygens cell or axis out of range");
        hc[i] = padded_cell(c, cells[i]);
        ha[i] = axes[i];
    }
    int rc;
    if ((rc = upload(c, &c->hcell, hc.data(), hc.size())) ||
        (rc = upload(c, &c->haxis, ha.data(), ha.size())))
        return rc;
    dfree(c, c->huy8);
    if ((rc = dalloc(c, &c->huy8, 8 * (size_t)std::max(1, n))))
        return rc;
    HIPCHK(c, hipMemset(c->huy8, 0, 8 * (size_t)std::max(1, n) * sizeof(float)));
    c->nh = n;
    return TLM_OK;
}

int tlm_run(tlm_ctx *c, int step0, int nsteps)
{
    if (!c)
        return TLM_ERR_ARG;
    return queue_steps(c, step0, nsteps, 0, nullptr, nullptr);
}

int tlm_sync(tlm_ctx *c)
{
    if (!c)
        return TLM_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    if (c->flag)
    {
        int f = 0;
        HIPCHK(c, hipMemcpy(&f, c->flag, sizeof(int), hipMemcpyDeviceToHost));
        if (f)
            return fail(c, TLM_ERR_RANGE,
                        "16-bit pulse storage: a value exceeded the range its slab scale allows");
    }
    return TLM_OK;
}

int tlm_read_scales(tlm_ctx *c, float *scale, float *smax, int32_t *fmt)
{
    if (!c)
        return TLM_ERR_ARG;
    if (c->pf == TLM_PULSE_F32 || !c->materialized)
        return 0;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (scale)
        HIPCHK(c, hipMemcpy(scale, c->scl[scale_buf(c, c->next_step)], c->nz * sizeof(float),
                            hipMemcpyDeviceToHost));
    if (smax)
        HIPCHK(c, hipMemcpy(smax, c->mreg, c->nz * sizeof(float), hipMemcpyDeviceToHost));
    if (fmt)
    {
        std::vector<uint2> sl(c->nz);
        HIPCHK(c, hipMemcpy(sl.data(), c->slab, c->nz * sizeof(uint2), hipMemcpyDeviceToHost));
        for (int k = 0; k < c->nz; ++k)
            fmt[k] = sl[k].y ? c->pf : TLM_PULSE_F32;
    }
    return c->nz;
}

int tlm_device_memory(int device, size_t *free_bytes, size_t *total_bytes)
{
    if (device < 0 || device >= tlm_device_count())
        return TLM_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return TLM_ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess)
        return TLM_ERR_HIP;
    if (free_bytes)
        *free_bytes = f;
    if (total_bytes)
        *total_bytes = t;
    return TLM_OK;
}

int64_t tlm_estimate_bytes(const tlm_grid_desc *d, int has_stubs)
{
    if (!d || d->nx <= 0 || d->ny <= 0 || d->nz <= 0)
        return TLM_ERR_ARG;
    const long long Pp = ((long long)d->nx * d->ny + 63) / 64 * 64, P = Pp * d->nz;
    const long long sp = d->pulse_type == TLM_PULSE_F32 ? 4 : 2; // FP32 halo slabs not counted
    long long bytes = 2 * 12 * P * sp + 2 * P;                   // pulses A, B and cell codes
    if (has_stubs)
        bytes += 3 * P * 4;
    bytes += (P / 64) * 8 + (1 << 20); // energy / slab-maximum partials, small tables
    return bytes;
}

int tlm_device_encode(int device, int fmt, const float *in, float inv_scale, uint16_t *out,
                      int64_t n)
{
    if ((fmt != TLM_PULSE_F16 && fmt != TLM_PULSE_BF16) || n < 0 || (n && (!in || !out)))
        return TLM_ERR_ARG;
    if (device < 0 || device >= tlm_device_count())
        return TLM_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return TLM_ERR_HIP;
    if (n == 0)
        return TLM_OK;
    float *din = nullptr;
    unsigned short *dout = nullptr;
    if (hipMalloc((void **)&din, n * sizeof(float)) != hipSuccess)
        return TLM_ERR_NOMEM;
    if (hipMalloc((void **)&dout, n * sizeof(unsigned short)) != hipSuccess)
    {
        (void)hipFree(din);
        return TLM_ERR_NOMEM;
    }
    int rc = TLM_OK;
    if (hipMemcpy(din, in, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        rc = TLM_ERR_HIP;
    else
    {
        const int b = 256;
        hipLaunchKernelGGL(tlm::k_encode, dim3((unsigned)((n + b - 1) / b)), dim3(b), 0, 0, fmt, din,
                           inv_scale, dout, (long long)n);
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, dout, n * sizeof(unsigned short), hipMemcpyDeviceToHost) != hipSuccess)
            rc = TLM_ERR_HIP;
    }
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}

int tlm_time_steps(tlm_ctx *c, int step0, int nsteps, int kernel_every, float *ms_total,
                   float *ms_kernel, int *nkernel)
{
    if (!c)
        return TLM_ERR_ARG;
    if (kernel_every < 0)
        return fail(c, TLM_ERR_ARG, "tlm_time_steps: kernel_every < 0");
    if (ms_kernel)
        *ms_kernel = 0.f;
    if (nkernel)
        *nkernel = 0;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int rc = queue_steps(c, step0, nsteps, ms_kernel ? kernel_every : 0, ms_kernel, nkernel);
    if (rc)
        return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (ms_total)
        *ms_total = ms;
    return TLM_OK;
}

int tlm_launch_shape(const tlm_ctx *c, int *rows, int *planes, int *groups, int *threads)
{
    if (!c)
        return TLM_ERR_ARG;
    if (rows)
        *rows = c->strips;
    if (planes)
        *planes = c->zc;
    if (groups)
        *groups = c->blocks;
    if (threads)
        *threads = c->wpb * tlm::kWave;
    return TLM_OK;
}

int tlm_set_tuning(tlm_ctx *c, const tlm_tuning *t)
{
    if (c && c->batch)
        return fail(c, TLM_ERR_STATE, "context is a batch member");
    if (!c || !t || t->kernel != 0 || t->planes < 0 || t->block_waves < 0 ||
        t->block_waves > 4 || t->nontemporal < 0 || t->nontemporal > 2 || t->grid_blocks < 0 ||
        t->fold < 0 || t->fold > 2 || t->occupancy < -1 || (t->occupancy > 0 && t->occupancy < 3) ||
        t->occupancy > 8)
        return fail(c, TLM_ERR_ARG, "tlm_set_tuning: bad arguments");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->tune = *t;
    choose_shape(c);
    const size_t need = energy_parts(c);
    if (need > c->epart_cap)
    {
        dfree(c, c->epart);
        int rc = dalloc(c, &c->epart, need);
        if (rc)
            return rc;
        c->epart_cap = need;
    }
    return TLM_OK;
}

int tlm_get_tuning(const tlm_ctx *c, tlm_tuning *t)
{
    if (!c || !t)
        return TLM_ERR_ARG;
    *t = c->tune;
    return TLM_OK;
}

int tlm_read_state(tlm_ctx *c, float *V, float *Vs)
{
    if (!c || !V)
        return fail(c, TLM_ERR_ARG, "tlm_read_state: bad arguments");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t ncell = (size_t)c->nx * c->ny * c->nz;
    if (c->pf != TLM_PULSE_F32 && !c->materialized)
    {
        // not run yet: the uploaded state
        if (c->pendV.empty())
            std::memset(V, 0, sizeof(float) * 12 * ncell);
        else
            std::memcpy(V, c->pendV.data(), sizeof(float) * 12 * ncell);
    }
    else
    {
        const tlm::PulseView pv = pulse_view(c, c->next_step);
        int rc = convert(c, nullptr, V, &pv, nullptr, 12);
        if (rc)
            return rc;
    }
    if (Vs)
    {
        if (c->stubs)
            return convert(c, nullptr, Vs, nullptr, c->Vs, 3);
        std::memset(Vs, 0, sizeof(float) * 3 * (size_t)c->nx * c->ny * c->nz);
    }
    return TLM_OK;
}

int tlm_read_probes(tlm_ctx *c, int step0, int nsteps, float *e)
{
    if (!c || step0 < 0 || nsteps < 0 || step0 + nsteps > c->nsteps || (!e && nsteps))
        return fail(c, TLM_ERR_ARG, "tlm_read_probes: bad arguments");
    if (c->npr == 0 || nsteps == 0)
        return TLM_OK;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(e, c->prrec + (size_t)step0 * c->npr * 3,
                        (size_t)nsteps * c->npr * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return TLM_OK;
}

int tlm_read_probe_dft(tlm_ctx *c, double *re, double *im)
{
    if (!c || !re || !im)
        return fail(c, TLM_ERR_ARG, "tlm_read_probe_dft: bad arguments");
    if (c->npr == 0 || c->nfreq == 0)
        return TLM_OK;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->npr * 3 * c->nfreq;
    HIPCHK(c, hipMemcpy(re, c->dft_re, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(im, c->dft_im, n * sizeof(double), hipMemcpyDeviceToHost));
    return TLM_OK;
}

int tlm_read_ports(tlm_ctx *c, int step0, int nsteps, float *vi)
{
    if (!c || step0 < 0 || nsteps < 0 || step0 + nsteps > c->nsteps || (!vi && nsteps))
        return fail(c, TLM_ERR_ARG, "tlm_read_ports: bad arguments");
    if (c->np == 0 || nsteps == 0)
        return TLM_OK;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(vi, c->prec + (size_t)step0 * c->np * 2,
                        (size_t)nsteps * c->np * 2 * sizeof(float), hipMemcpyDeviceToHost));
    return TLM_OK;
}

int tlm_read_js(tlm_ctx *c, float *inst, float *dft4)
{
    if (!c)
        return TLM_ERR_ARG;
    if (c->nf == 0)
        return TLM_OK;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (inst)
        HIPCHK(c, hipMemcpy(inst, c->js_inst, (size_t)c->nf * sizeof(float), hipMemcpyDeviceToHost));
    if (dft4)
        HIPCHK(c, hipMemcpy(dft4, c->js_dft4, (size_t)c->nf * 4 * sizeof(float),
                            hipMemcpyDeviceToHost));
    return TLM_OK;
}

int tlm_read_huygens(tlm_ctx *c, float *dft8)
{
    if (!c || !dft8)
        return fail(c, TLM_ERR_ARG, "tlm_read_huygens: bad arguments");
    if (c->nh == 0)
        return TLM_OK;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(dft8, c->huy8, (size_t)c->nh * 8 * sizeof(float), hipMemcpyDeviceToHost));
    return TLM_OK;
}

int tlm_read_energy(tlm_ctx *c, int32_t *steps, double *vals, int cap)
{
    if (!c || cap < 0)
        return fail(c, TLM_ERR_ARG, "tlm_read_energy: bad arguments");
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<int> idx;
    for (size_t i = 0; i < c->energy_steps.size(); ++i)
        if (c->energy_steps[i] >= 0)
            idx.push_back((int)i);
    const int cnt = (int)idx.size();
    if (cap > 0 && (steps || vals))
    {
        std::vector<double> all(c->energy_steps.size());
        if (!all.empty())
            HIPCHK(c, hipMemcpy(all.data(), c->energy, all.size() * sizeof(double),
                                hipMemcpyDeviceToHost));
        for (int q = 0; q < std::min(cap, cnt); ++q)
        {
            if (steps)
                steps[q] = c->energy_steps[idx[q]];
            if (vals)
                vals[q] = all[idx[q]];
        }
    }
    return cnt;
}

int tlm_copy_bandwidth(int device, size_t bytes, int iters, float *gbps, float *ms_per_copy)
{
    if (bytes < 16 || iters <= 0)
        return TLM_ERR_ARG;
    if (tlm_device_count() <= device || device < 0)
        return TLM_ERR_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return TLM_ERR_HIP;
    const size_t n4 = bytes / 16;
    float4 *src = nullptr, *dst = nullptr;
    if (hipMalloc((void **)&src, n4 * 16) != hipSuccess)
        return TLM_ERR_NOMEM;
    if (hipMalloc((void **)&dst, n4 * 16) != hipSuccess)
    {
        (void)hipFree(src);
        return TLM_ERR_NOMEM;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemset(src, 0, n4 * 16);
    const int b = 256;
    const unsigned g = (unsigned)std::min<size_t>((n4 + b - 1) / b, 1u << 20);
    hipLaunchKernelGGL(tlm::k_copy, dim3(g), dim3(b), 0, 0, src, dst, (long long)n4);
    (void)hipEventRecord(e0, 0);
    for (int i = 0; i < iters; ++i)
        hipLaunchKernelGGL(tlm::k_copy, dim3(g), dim3(b), 0, 0, src, dst, (long long)n4);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess)
        return TLM_ERR_HIP;
    if (ms_per_copy)
        *ms_per_copy = ms / iters;
    if (gbps)
        *gbps = (float)(2.0 * n4 * 16 * iters / (ms * 1e-3) / 1e9);
    return TLM_OK;
}

float tlm_scatter_unitarity_error(float epsr)
{
    if (tlm_device_count() <= 0)
        return (float)TLM_ERR_NO_DEVICE;
    const bool diel = epsr > 1.0f;
    const int NP = diel ? 15 : 12;
    float yh = diel ? 4.0f * (epsr - 1.0f) : 0.0f;
    float den = 4.0f + yh + 0.0f;
    float *dS = nullptr;
    if (hipMalloc((void **)&dS, 15 * 15 * sizeof(float)) != hipSuccess)
        return (float)TLM_ERR_NOMEM;
    hipLaunchKernelGGL(tlm::k_scatter_columns, dim3(1), dim3(64), 0, 0, yh, den, diel ? 1 : 0, dS);
    float S[15 * 15];
    hipError_t e = hipMemcpy(S, dS, sizeof(S), hipMemcpyDeviceToHost);
    (void)hipFree(dS);
    if (e != hipSuccess)
        return (float)TLM_ERR_HIP;
    // S^T W S against W with W = (1 x 12, yh x 3), as TlmSolver.cpp:141-169
    float W[15];
    for (int i = 0; i < 15; ++i)
        W[i] = i < 12 ? 1.0f : yh;
    float maxErr = 0.0f;
    for (int a = 0; a < NP; ++a)
        for (int b = 0; b < NP; ++b)
        {
            float dot = 0.0f;
            for (int r = 0; r < NP; ++r)
                dot += W[r] * S[r * 15 + a] * S[r * 15 + b];
            float want = (a == b) ? W[a] : 0.0f;
            maxErr = std::max(maxErr, std::fabs(dot - want));
        }
    return maxErr;
}

int64_t tlm_build_pec_faces(int nx, int ny, int nz, const uint8_t *mat, int64_t *cells,
                            int32_t *axes, int32_t *signs, int64_t cap)
{
    if (!mat || nx <= 0 || ny <= 0 || nz <= 0)
        return 0;
    const long long dirs[3] = {1, nx, (long long)nx * ny};
    const int nn[3] = {nx, ny, nz};
    int64_t cnt = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
            {
                long long c = ((long long)k * ny + j) * nx + i;
                if (mat[c] != 1)
                    continue;
                const int co[3] = {i, j, k};
                for (int a = 0; a < 3; ++a)
                    for (int s = -1; s <= 1; s += 2)
                    {
                        int qq = co[a] + s;
                        if (qq < 0 || qq >= nn[a])
                            continue;
                        long long nb = c + s * dirs[a];
                        if (mat[nb] == 1)
                            continue;
                        if (cnt < cap)
                        {
                            if (cells)
                                cells[cnt] = nb;
                            if (axes)
                                axes[cnt] = a;
                            if (signs)
                                signs[cnt] = s;
                        }
                        ++cnt;
                    }
            }
    return cnt;
}

int64_t tlm_build_huygens(int nx, int ny, int nz, const uint8_t *mat, int o, int64_t *cells,
                          int32_t *axes, int32_t *out_signs, int64_t cap)
{
    if (!mat || o < 2)
        return 0;
    const int nn[3] = {nx, ny, nz};
    if (2 * o + 2 >= nn[0] || 2 * o + 2 >= nn[1] || 2 * o + 2 >= nn[2])
        return 0;
    int64_t cnt = 0;
    auto add = [&](int i, int j, int k, int axis, int sg) {
        long long c = ((long long)k * ny + j) * nx + i;
        if (mat[c] != 0)
            return;
        if (cnt < cap)
        {
            if (cells)
                cells[cnt] = c;
            if (axes)
                axes[cnt] = axis;
            if (out_signs)
                out_signs[cnt] = sg;
        }
        ++cnt;
    };
    int x0 = o, x1 = nx - 1 - o, y0 = o, y1 = ny - 1 - o, z0 = o, z1 = nz - 1 - o;
    for (int k = z0; k <= z1; ++k)
        for (int j = y0; j <= y1; ++j)
        {
            add(x0, j, k, 0, -1);
            add(x1, j, k, 0, +1);
        }
    for (int k = z0; k <= z1; ++k)
        for (int i = x0; i <= x1; ++i)
        {
            add(i, y0, k, 1, -1);
            add(i, y1, k, 1, +1);
        }
    for (int j = y0; j <= y1; ++j)
        for (int i = x0; i <= x1; ++i)
        {
            add(i, j, z0, 2, -1);
            add(i, j, z1, 2, +1);
        }
    return cnt;
}

void tlm_schedule_tables(int type, float f0, float dt, int step0, int nsteps, float gdw, float gdd,
                         float *sv, float *cph, float *sph)
{
    for (int q = 0; q < nsteps; ++q)
    {
        const int n = step0 + q;
        if (sv)
        {
            float val;
            if (type == TLM_WF_GAUSS_DERIV)
            {
                double u = ((double)n - (double)gdd) / (double)gdw;
                val = (float)(-std::sqrt(2.0 * std::exp(1.0)) * u * std::exp(-u * u));
            }
            else
            {
                float t = n * dt;
                float w = 2.0f * (float)M_PI * f0;
                if (type == TLM_WF_CW_RAMPED)
                {
                    float Tr = 3.0f / f0;
                    float r = (t < Tr) ? 0.5f * (1.0f - std::cos((float)M_PI * t / Tr)) : 1.0f;
                    val = r * std::sin(w * t);
                }
                else if (type == TLM_WF_GAUSS_SINE)
                {
                    float tau = 2.0f / f0, t0 = 3.0f * tau;
                    float u = (t - t0) / tau;
                    val = std::exp(-u * u) * std::sin(w * (t - t0));
                }
                else
                {
                    float tau = 0.5f / f0, t0 = 3.0f * tau;
                    float u = (t - t0) / tau;
                    val = std::exp(-u * u);
                }
            }
            sv[q] = val;
        }
        float phase = 2.0f * (float)M_PI * f0 * (n * dt);
        if (cph)
            cph[q] = std::cos(phase);
        if (sph)
            sph[q] = std::sin(phase);
    }
}

void tlm_dft_tables(const double *freqs, int nfreq, double dt, int nsteps, double *ctab,
                    double *stab)
{
    for (int n = 0; n < nsteps; ++n)
        for (int q = 0; q < nfreq; ++q)
        {
            double ph = 2.0 * M_PI * freqs[q] * ((double)n * dt);
            ctab[(size_t)n * nfreq + q] = std::cos(ph);
            stab[(size_t)n * nfreq + q] = std::sin(ph);
        }
}

} // extern "C"
