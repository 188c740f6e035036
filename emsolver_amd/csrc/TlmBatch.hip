// TlmBatch.hip - batched execution of many independent TLM grids
// (include/GpuTlmHip.h, tlm_batch_*). No counterpart in the reference, whose
// GPU path runs one grid per OpenGL context.
//
// Members are ordinary contexts (tlm_create / tlm_upload / tlm_set_*), so a
// grid gives the same bits alone and in a batch: the per-cell arithmetic and
// every monitor write are the single-grid code paths, only the launch that
// drives them differs. A tick advances every active member by one step (see
// tlm_batch_kernels.h); blocks of ticks are captured in a HIP graph.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function" // single-grid kernels of the shared header
#include "tlm_batch_kernels.h"
#include "tlm_context.h"
#pragma clang diagnostic pop

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_batch_create_error = "no error";

constexpr int kKinds = 6; // (stubs 0/1) x (PF_F32, PF_F16, PF_BF16)

struct Member
{
    tlm_ctx *c = nullptr;
    int n0 = 0, nend = 0;
    long long tick0 = 0;
    char *buf[2] = {nullptr, nullptr};
    long long cells = 0;
    bool range_err = false; // 16-bit storage: a value outgrew its slab scale
};

} // namespace

struct tlm_batch
{
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err = "no error";
    std::vector<Member> mem;
    long long tick = 0; // ticks queued so far (the device counter after they ran)
    long long *d_tick = nullptr;
    tlm_batch_tuning tune = {1, 4, 0, 32, 4};
    bool dirty = true;
    // device tables, rebuilt when the membership or the tuning changes
    tlm::BatchMember *d_mem = nullptr;
    size_t mem_cap = 0;
    int *d_gmem[kKinds] = {};
    size_t gmem_cap[kKinds] = {};
    int ngroups[kKinds] = {};
    int2 *d_inj = nullptr, *d_mon = nullptr, *d_smax = nullptr;
    size_t inj_cap = 0, mon_cap = 0, smax_cap = 0;
    int ninj = 0, nplan = 0, nmon = 0, nsmax = 0;
    int wpb = 4;
    // captured block of ticks
    hipGraphExec_t gexec = nullptr;
    int glen = 0;
};

namespace {

int bfail(tlm_batch *b, int code, const std::string &msg)
{
    if (b)
        b->err = msg;
    else
        g_batch_create_error = msg;
    return code;
}

int bhip(tlm_batch *b, hipError_t e, const char *what)
{
    return bfail(b, e == hipErrorOutOfMemory ? TLM_ERR_NOMEM : TLM_ERR_HIP,
                 std::string(what) + ": " + hipGetErrorString(e));
}

#define BCHK(b, call)                                                                              \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return bhip((b), e_, #call);                                                           \
    } while (0)

// Grows a device buffer to hold n elements (contents not kept).
template <class T> int grow(tlm_batch *b, T **p, size_t *cap, size_t n)
{
    if (n <= *cap && *p)
        return TLM_OK;
    if (*p)
        (void)hipFree(*p);
    *p = nullptr;
    const size_t want = std::max<size_t>(n, 1) + std::max<size_t>(n, 1) / 2;
    BCHK(b, hipMalloc((void **)p, want * sizeof(T)));
    *cap = want;
    return TLM_OK;
}

void drop_graph(tlm_batch *b)
{
    if (b->gexec)
        (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->glen = 0;
}

int kind_of(const tlm_ctx *c)
{
    return (c->stubs ? 3 : 0) + c->pf;
}

// Current step of a member (clamped to its end).
int member_now(const tlm_batch *b, const Member &m)
{
    return (int)std::min<long long>(m.nend, m.n0 + (b->tick - m.tick0));
}

// Rebuilds the member descriptors, the group tables and the monitor entry
// lists. The stream is idle when this runs (tables are read by queued work).
int rebuild(tlm_batch *b)
{
    BCHK(b, hipStreamSynchronize(b->stream));
    drop_graph(b);
    b->wpb = std::max(1, std::min(4, b->tune.block_waves > 0 ? b->tune.block_waves : 4));
    const int wpb = b->wpb;
    std::vector<tlm::BatchMember> hm(b->mem.size());
    std::vector<int> gm[kKinds];
    std::vector<int2> inj, plan, mon, smax;
    for (size_t i = 0; i < b->mem.size(); ++i)
    {
        const Member &M = b->mem[i];
        tlm_ctx *c = M.c;
        tlm::BatchMember &d = hm[i];
        std::memset(&d, 0, sizeof(d));
        const int pl = c->pf == TLM_PULSE_F32 ? b->tune.planes : b->tune.planes16;
        const int zc = pl > 0 ? std::min(pl, c->nz) : c->nz;
        const int nzc = (c->nz + zc - 1) / zc;
        const int strips = (int)(c->Pp / tlm::kWave);
        const int tasks = strips * nzc;
        const int groups = (tasks + wpb - 1) / wpb;
        tlm::StepArgs &a = d.st;
        a.A = nullptr;
        a.B = nullptr;
        a.Vs = c->Vs;
        a.code = c->code;
        a.uair = c->uair;
        a.props = c->props;
        a.epart = nullptr;
        a.slab = c->slab;
        a.dsc = nullptr;
        a.esc = nullptr;
        a.mpart = c->mpart;
        a.nx = c->nx;
        a.ny = c->ny;
        a.nz = c->nz;
        a.nxy = (int)c->nxy;
        a.Pp = c->Pp;
        a.P = c->P;
        a.PB = c->PB;
        a.h0 = c->h0;
        a.h1 = c->h1;
        a.strips = strips;
        a.zc = zc;
        a.nzc = nzc;
        a.tasks = tasks;
        a.groups = groups;
        a.rho = c->rho;
        a.nmat = c->nmat;
        a.mstrips = tlm::magic_div((unsigned)strips);
        a.mnx = tlm::magic_div((unsigned)c->nx);
        d.buf[0] = M.buf[0];
        d.buf[1] = M.buf[1];
        d.tick0 = M.tick0;
        d.n0 = M.n0;
        d.nend = M.nend;
        d.pf = c->pf;
        d.sbase = c->base;
        for (int q = 0; q < 2; ++q)
        {
            d.scl[q] = c->scl[q];
            d.iscl[q] = c->iscl[q];
        }
        d.ucell = c->ucell;
        d.ubeg = c->ubeg;
        d.spol = c->spol;
        d.samp = c->samp;
        d.sv = c->sv;
        d.nsched = c->nsteps;
        d.mreg = c->mreg;
        d.qinj = c->qinj;
        d.flag = c->flag;
        d.rp = std::max(1.0f, std::fabs(c->rho));
        d.fmax = tlmi::format_max(c->pf);
        d.prcell = c->prcell;
        d.prrec = c->prrec;
        d.npr = c->npr;
        d.inv_dl = c->inv_dl;
        d.pcells = c->pcells;
        d.ploop = c->ploop;
        d.poff = c->poff;
        d.ppol = c->ppol;
        d.prec = c->prec;
        d.np = c->np;
        const int k = kind_of(c);
        d.group0 = (int)gm[k].size();
        gm[k].insert(gm[k].end(), groups, (int)i);
        for (int u = 0; u < c->nuniq; ++u)
            inj.push_back(make_int2((int)i, u));
        if (c->pf != TLM_PULSE_F32)
            for (int z = 0; z < c->nz; ++z)
            {
                plan.push_back(make_int2((int)i, z));
                smax.push_back(make_int2((int)i, z));
            }
        for (int p = 0; p < c->npr + c->np; ++p)
            mon.push_back(make_int2((int)i, p));
    }
    b->ninj = (int)inj.size();
    b->nplan = (int)plan.size();
    inj.insert(inj.end(), plan.begin(), plan.end());
    b->nmon = (int)mon.size();
    b->nsmax = (int)smax.size();
    int rc;
    if ((rc = grow(b, &b->d_mem, &b->mem_cap, hm.size())) ||
        (rc = grow(b, &b->d_inj, &b->inj_cap, inj.size())) ||
        (rc = grow(b, &b->d_mon, &b->mon_cap, mon.size())) ||
        (rc = grow(b, &b->d_smax, &b->smax_cap, smax.size())))
        return rc;
    if (!hm.empty())
        BCHK(b, hipMemcpy(b->d_mem, hm.data(), hm.size() * sizeof(hm[0]), hipMemcpyHostToDevice));
    if (!inj.empty())
        BCHK(b, hipMemcpy(b->d_inj, inj.data(), inj.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!mon.empty())
        BCHK(b, hipMemcpy(b->d_mon, mon.data(), mon.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!smax.empty())
        BCHK(b, hipMemcpy(b->d_smax, smax.data(), smax.size() * sizeof(int2),
                          hipMemcpyHostToDevice));
    for (int k = 0; k < kKinds; ++k)
    {
        b->ngroups[k] = (int)gm[k].size();
        if (gm[k].empty())
            continue;
        if ((rc = grow(b, &b->d_gmem[k], &b->gmem_cap[k], gm[k].size())))
            return rc;
        BCHK(b, hipMemcpy(b->d_gmem[k], gm[k].data(), gm[k].size() * sizeof(int),
                          hipMemcpyHostToDevice));
    }
    b->dirty = false;
    return TLM_OK;
}

// One kind's step launch; with_mon appends the workgroups that take the
// tick's probe and port records.
template <bool S, int PF> void launch_kind(tlm_batch *b, int k, bool with_mon)
{
    const int ng = b->ngroups[k];
    const int blocks = b->tune.grid_blocks > 0 ? std::min(ng, b->tune.grid_blocks) : ng;
    const int bt = b->wpb * tlm::kWave;
    const int mon_blocks = with_mon && b->nmon > 0 ? (b->nmon + bt - 1) / bt : 0;
    hipLaunchKernelGGL((tlm::k_step_batch<S, PF>), dim3(blocks + mon_blocks), dim3(bt), 0,
                       b->stream, b->d_mem, b->d_gmem[k], ng, b->d_tick, b->d_mon, b->nmon,
                       blocks);
}

void queue_tick(tlm_batch *b)
{
    if (b->ninj + b->nplan > 0)
    {
        const int n = b->ninj + b->nplan, t = 128;
        hipLaunchKernelGGL(tlm::k_inject_batch, dim3((n + t - 1) / t), dim3(t), 0, b->stream,
                           b->d_mem, b->d_inj, b->ninj, b->nplan, b->d_tick);
    }
    // the records ride with the first kind that has members
    bool mon = true;
    auto kind = [&](int k, auto launch) {
        if (!b->ngroups[k])
            return;
        launch(k, mon);
        mon = false;
    };
    kind(0, [&](int k, bool m) { launch_kind<false, tlm::PF_F32>(b, k, m); });
    kind(1, [&](int k, bool m) { launch_kind<false, tlm::PF_F16>(b, k, m); });
    kind(2, [&](int k, bool m) { launch_kind<false, tlm::PF_BF16>(b, k, m); });
    kind(3, [&](int k, bool m) { launch_kind<true, tlm::PF_F32>(b, k, m); });
    kind(4, [&](int k, bool m) { launch_kind<true, tlm::PF_F16>(b, k, m); });
    kind(5, [&](int k, bool m) { launch_kind<true, tlm::PF_BF16>(b, k, m); });
    if (mon && b->nmon > 0)
    {
        // no member with step groups (cannot happen with active members):
        // the records alone, through an empty step grid
        launch_kind<false, tlm::PF_F32>(b, 0, true);
    }
    if (b->nsmax > 0)
        hipLaunchKernelGGL(tlm::k_slab_max_batch, dim3(b->nsmax), dim3(256), 0, b->stream,
                           b->d_mem, b->d_smax, b->d_tick);
    hipLaunchKernelGGL(tlm::k_tick, dim3(1), dim3(64), 0, b->stream, b->d_tick);
}

int queue_ticks(tlm_batch *b, int nticks)
{
    if (b->dirty)
    {
        int rc = rebuild(b);
        if (rc)
            return rc;
    }
    int left = nticks;
    const int L = b->tune.graph_ticks;
    if (L > 0 && left >= L)
    {
        if (!b->gexec || b->glen != L)
        {
            drop_graph(b);
            hipGraph_t g = nullptr;
            BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            for (int t = 0; t < L; ++t)
                queue_tick(b);
            BCHK(b, hipStreamEndCapture(b->stream, &g));
            hipError_t e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess)
            {
                b->gexec = nullptr;
                return bhip(b, e, "hipGraphInstantiate");
            }
            b->glen = L;
        }
        while (left >= L)
        {
            BCHK(b, hipGraphLaunch(b->gexec, b->stream));
            left -= L;
        }
    }
    for (; left > 0; --left)
        queue_tick(b);
    BCHK(b, hipGetLastError());
    b->tick += nticks;
    return TLM_OK;
}

// Waits for the queued ticks and brings every member's host-side state up
// to date (current pulse buffer, step counters, running probe DFT). A 16-bit
// member whose range flag is set is marked (its results are invalid); the
// others go on, and tlm_batch_remove reports it for that member only.
int sync_members(tlm_batch *b)
{
    BCHK(b, hipSetDevice(b->dev));
    BCHK(b, hipStreamSynchronize(b->stream));
    for (Member &m : b->mem)
    {
        tlm_ctx *c = m.c;
        const int now = member_now(b, m);
        const int s = (now - m.n0) & 1;
        c->A = m.buf[s];
        c->B = m.buf[s ^ 1];
        c->next_step = now;
        c->steps_run_hi = std::max(c->steps_run_hi, now);
        if (c->npr > 0 && c->nfreq > 0 && c->dft_done < now)
            tlmi::queue_probe_dft(c, c->dft_done, now, b->stream);
        if (c->flag && !m.range_err)
        {
            int f = 0;
            BCHK(b, hipMemcpy(&f, c->flag, sizeof(int), hipMemcpyDeviceToHost));
            m.range_err = f != 0;
        }
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    return TLM_OK;
}

long long active_cells(const tlm_batch *b)
{
    long long n = 0;
    for (const Member &m : b->mem)
        if (member_now(b, m) < m.nend)
            n += m.cells;
    return n;
}

} // namespace

extern "C" {

int tlm_batch_create(int device, tlm_batch **out)
{
    if (!out)
        return bfail(nullptr, TLM_ERR_ARG, "tlm_batch_create: out is NULL");
    *out = nullptr;
    if (device < 0 || device >= tlm_device_count())
        return bfail(nullptr, TLM_ERR_NO_DEVICE, "no such HIP device");
    tlm_batch *b = new tlm_batch;
    b->dev = device;
    auto bail = [&](hipError_t e, const char *what) {
        bhip(b, e, what);
        g_batch_create_error = b->err;
        tlm_batch_destroy(b);
        return e == hipErrorOutOfMemory ? TLM_ERR_NOMEM : TLM_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess)
        return bail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess)
        return bail(e, "hipEventCreate");
    if ((e = hipMalloc((void **)&b->d_tick, sizeof(long long))) != hipSuccess)
        return bail(e, "hipMalloc");
    if ((e = hipMemset(b->d_tick, 0, sizeof(long long))) != hipSuccess)
        return bail(e, "hipMemset");
    *out = b;
    return TLM_OK;
}

void tlm_batch_destroy(tlm_batch *b)
{
    if (!b)
        return;
    (void)hipSetDevice(b->dev);
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    for (Member &m : b->mem)
        m.c->batch = nullptr;
    b->mem.clear();
    drop_graph(b);
    for (void *p : {(void *)b->d_tick, (void *)b->d_mem, (void *)b->d_inj, (void *)b->d_mon,
                    (void *)b->d_smax})
        if (p)
            (void)hipFree(p);
    for (int k = 0; k < kKinds; ++k)
        if (b->d_gmem[k])
            (void)hipFree(b->d_gmem[k]);
    for (hipEvent_t ev : {b->ev0, b->ev1})
        if (ev)
            (void)hipEventDestroy(ev);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *tlm_batch_last_error(const tlm_batch *b)
{
    return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int tlm_batch_add(tlm_batch *b, tlm_ctx *c, int nsteps)
{
    if (!b || !c || nsteps < 0)
        return bfail(b, TLM_ERR_ARG, "tlm_batch_add: bad arguments");
    if (c->batch)
        return bfail(b, TLM_ERR_STATE, "context is already a batch member");
    if (c->dev != b->dev)
        return bfail(b, TLM_ERR_ARG, "context and batch are on different devices");
    if (!c->uploaded || !c->sv)
        return bfail(b, TLM_ERR_STATE, "batch member needs tlm_upload and tlm_set_schedule first");
    if (c->nf > 0 || c->nh > 0 || c->pa >= 0 || c->frame_every > 0)
        return bfail(b, TLM_ERR_UNSUPPORTED,
                     "batch members record probes and ports (no Js, Huygens, plane or frames)");
    if (c->every > 0)
        return bfail(b, TLM_ERR_UNSUPPORTED,
                     "batch members record no energy: tlm_set_energy_every(ctx, 0) before joining");
    const int n0 = c->next_step;
    if (n0 + nsteps > c->nsteps)
        return bfail(b, TLM_ERR_ARG, "steps beyond the member's schedule");
    if (c->npr > 0 && c->nfreq > 0 && c->dft_done != n0)
        return bfail(b, TLM_ERR_STATE, "probe DFT needs contiguous steps");
    BCHK(b, hipSetDevice(b->dev));
    BCHK(b, hipStreamSynchronize(c->stream));
    if (c->pf != TLM_PULSE_F32 && !c->materialized)
    {
        int rc = tlmi::materialize(c, n0);
        if (rc)
            return bfail(b, rc, c->err);
    }
    Member m;
    m.c = c;
    m.n0 = n0;
    m.nend = n0 + nsteps;
    m.tick0 = b->tick;
    m.buf[0] = c->A;
    m.buf[1] = c->B;
    m.cells = (long long)c->nx * c->ny * c->nz;
    b->mem.push_back(m);
    c->batch = b;
    b->dirty = true;
    return TLM_OK;
}

int tlm_batch_remove(tlm_batch *b, tlm_ctx *c)
{
    if (!b || !c)
        return bfail(b, TLM_ERR_ARG, "tlm_batch_remove: bad arguments");
    auto it = std::find_if(b->mem.begin(), b->mem.end(), [&](const Member &m) { return m.c == c; });
    if (it == b->mem.end())
        return bfail(b, TLM_ERR_ARG, "context is not a member of this batch");
    int rc = sync_members(b);
    const bool range_err = it->range_err;
    b->mem.erase(it);
    c->batch = nullptr;
    b->dirty = true;
    if (rc == TLM_OK && range_err)
    {
        tlmi::fail(c, TLM_ERR_RANGE, "16-bit pulse storage: a value exceeded the range its slab scale allows");
        return bfail(b, TLM_ERR_RANGE,
                     "this member's 16-bit pulse storage exceeded its slab scale: its results are "
                     "invalid (the other members are not affected)");
    }
    return rc;
}

int tlm_batch_run(tlm_batch *b, int nticks)
{
    if (!b || nticks < 0)
        return bfail(b, TLM_ERR_ARG, "tlm_batch_run: bad arguments");
    if (nticks == 0 || b->mem.empty())
        return TLM_OK;
    BCHK(b, hipSetDevice(b->dev));
    return queue_ticks(b, nticks);
}

int tlm_batch_sync(tlm_batch *b)
{
    if (!b)
        return TLM_ERR_ARG;
    return sync_members(b);
}

int tlm_batch_members(tlm_batch *b, tlm_ctx **members, int32_t *remaining, int cap)
{
    if (!b || cap < 0)
        return TLM_ERR_ARG;
    const int n = (int)b->mem.size();
    for (int i = 0; i < n && i < cap; ++i)
    {
        if (members)
            members[i] = b->mem[i].c;
        if (remaining)
            remaining[i] = b->mem[i].nend - member_now(b, b->mem[i]);
    }
    return n;
}

int tlm_batch_set_tuning(tlm_batch *b, const tlm_batch_tuning *t)
{
    if (!b || !t || t->planes < 0 || t->block_waves < 0 || t->block_waves > 4 ||
        t->grid_blocks < 0 || t->graph_ticks < 0 || t->planes16 < 0)
        return bfail(b, TLM_ERR_ARG, "tlm_batch_set_tuning: bad tuning");
    b->tune = *t;
    b->dirty = true;
    return TLM_OK;
}

int tlm_batch_get_tuning(const tlm_batch *b, tlm_batch_tuning *t)
{
    if (!b || !t)
        return TLM_ERR_ARG;
    *t = b->tune;
    return TLM_OK;
}

int tlm_batch_time(tlm_batch *b, int nticks, float *ms, int64_t *cell_updates)
{
    if (!b || nticks <= 0)
        return bfail(b, TLM_ERR_ARG, "tlm_batch_time: bad arguments");
    BCHK(b, hipSetDevice(b->dev));
    if (b->dirty)
    {
        int rc = rebuild(b);
        if (rc)
            return rc;
    }
    long long cu = 0;
    for (const Member &m : b->mem)
    {
        const long long now = member_now(b, m);
        cu += m.cells * std::max<long long>(0, std::min<long long>(m.nend - now, nticks));
    }
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    int rc = queue_ticks(b, nticks);
    if (rc)
        return rc;
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    BCHK(b, hipEventSynchronize(b->ev1));
    float t = 0.f;
    BCHK(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
    if (ms)
        *ms = t;
    if (cell_updates)
        *cell_updates = cu;
    return TLM_OK;
}

// Online autotuner: every candidate (slabs per wave task x waves per
// workgroup x grid) runs real ticks of the batch (the launch shape never
// changes results), timed with events; the fastest per active cell wins.
int tlm_batch_autotune(tlm_batch *b, int ticks_per_candidate, tlm_batch_tuning *best)
{
    if (!b || ticks_per_candidate <= 0)
        return bfail(b, TLM_ERR_ARG, "tlm_batch_autotune: bad arguments");
    if (b->mem.empty())
        return bfail(b, TLM_ERR_STATE, "tlm_batch_autotune: empty batch");
    const tlm_batch_tuning keep = b->tune;
    bool have16 = false, have32 = false;
    for (const Member &m : b->mem)
        (m.c->pf == TLM_PULSE_F32 ? have32 : have16) = true;
    std::vector<tlm_batch_tuning> cands;
    for (int p : {1, 2, 4, 8})
        for (int w : {2, 4})
            for (int g : {0, 2048})
                for (int p16 : {2, 4, 8})
                {
                    if ((!have32 && p != 1) || (!have16 && p16 != 4))
                        continue;
                    cands.push_back({p, w, g, 0, p16});
                }
    tlm_batch_tuning top = keep;
    double top_ns = 1e300;
    for (const tlm_batch_tuning &t : cands)
    {
        if (active_cells(b) == 0)
            break;
        b->tune = t;
        b->dirty = true;
        float ms = 0.f;
        int64_t cu = 0;
        int rc = tlm_batch_time(b, 1, &ms, &cu); // warm-up (table rebuild)
        if (!rc)
            rc = tlm_batch_time(b, ticks_per_candidate, &ms, &cu);
        if (rc)
        {
            b->tune = keep;
            b->dirty = true;
            return rc;
        }
        const double ns = cu > 0 ? ms * 1e6 / (double)cu : 1e300;
        if (ns < top_ns)
        {
            top_ns = ns;
            top = t;
        }
    }
    top.graph_ticks = keep.graph_ticks;
    b->tune = top;
    b->dirty = true;
    if (best)
        *best = top;
    return TLM_OK;
}

} // extern "C"
