// tlm_context.h - the device context behind the C ABI (tlm_ctx) and the
// helpers shared by the single-grid path (GpuTlmHip.hip) and the batched
// path (TlmBatch.hip). Internal: not part of the installed interface.
#pragma once

#include "../../include/GpuTlmHip.h"
#include "tlm_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

struct tlm_ctx
{
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
    std::vector<hipEvent_t> kev; // per-launch timing events (2 per step), created on demand
    int nx = 0, ny = 0, nz = 0;
    long long nxy = 0, Pp = 0, P = 0; // slab pitch roundup(nx*ny, 64), plane stride Pp*nz
    float dl = 0.f, rho = 0.f;
    // launch shape of the fused kernel
    tlm_tuning tune = {0, 1, 4, 0, 0, 2, -1}; // folding and occupancy chosen per grid
    int strips = 0, zc = 0, nzc = 0, tasks = 0, groups = 0, blocks = 0, wpb = 4;
    int xchunk = 0; // > 0: groups per chunk of the XCD-interleaved map (choose_shape)
    size_t epart_cap = 0;
    // state
    int pf = TLM_PULSE_F32; // pulse storage format
    int halo = 8;           // 16-bit formats: FP32 slabs on each side of a source slab
    long long PB = 0;       // bytes per pulse plane
    char *A = nullptr, *B = nullptr;
    float *Vs = nullptr;
    uint16_t *code = nullptr; // cell code: material + neighbour-PEC bits
    uint8_t *uair = nullptr;  // per 64-cell block: all codes 0 (air, no PEC neighbour);
                              // allocated to a multiple of 4 bytes
    float4 *props = nullptr;
    int nmat = 0;
    bool stubs = false, uploaded = false;
    double *epart = nullptr, *epart2 = nullptr; // energy partials (per wave, per chunk)
    // sources
    long long *ucell = nullptr;
    int *ubeg = nullptr, *spol = nullptr;
    float *samp = nullptr;
    int nuniq = 0;
    bool fold_marked = false; // FP32: blocks of the source cells carry flag bit 1
    // schedule
    float *sv = nullptr, *cph = nullptr, *sph = nullptr;
    std::vector<float> sv_host; // the waveform values, for the folded sources
    int nsteps = 0, settle = 0;
    // monitors
    long long *fcell = nullptr;
    int *faxis = nullptr;
    float *js_inst = nullptr, *js_dft4 = nullptr;
    int nf = 0;
    long long *hcell = nullptr;
    int *haxis = nullptr;
    float *huy8 = nullptr;
    int nh = 0;
    long long *pcells = nullptr, *ploop = nullptr;
    int *poff = nullptr, *ppol = nullptr;
    float *prec = nullptr;
    int np = 0;
    long long *prcell = nullptr;
    float inv_dl = 0.f;
    float *prrec = nullptr;
    int npr = 0;
    double *ctab = nullptr, *stab = nullptr, *dft_re = nullptr, *dft_im = nullptr;
    int nfreq = 0;
    int dft_done = 0; // probe DFT accumulated over steps [0, dft_done)
    // cut plane and playback frames
    int pa = -1, pidx = 0, pn1 = 0, pn2 = 0, plane_step = -1;
    float *plane = nullptr;
    int frame_every = 0, frame_max = 0;
    float *frame_js = nullptr, *frame_plane = nullptr;
    std::vector<int> frame_steps;
    std::vector<int> frame_plane_step;
    // energy
    int every = 16;
    double *energy = nullptr;
    std::vector<int> energy_steps;
    int steps_run_hi = 0; // highest step index + 1 queued so far
    // 16-bit formats: per-slab storage and scales, built at the first run
    std::vector<long long> src_cells; // soft sources in list order (unpadded cells)
    std::vector<float> src_amp;
    bool materialized = false;
    std::vector<float> pendV, pendVs; // uploaded state waiting for its scales
    uint2 *slab = nullptr;            // per slab: (byte offset in a plane, 16-bit flag)
    float *scl[2] = {nullptr, nullptr}, *iscl[2] = {nullptr, nullptr};
    float *mreg = nullptr, *mpart = nullptr, *qinj = nullptr;
    int *flag = nullptr;
    int base = 0;      // step whose scales live in buffer 0
    int next_step = 0; // next step a 16-bit run must start at
    int nhot = 0;      // FP32 slabs
    int h0 = 0, h1 = 0; // the FP32 slabs are [h0, h1)
    std::string err = "no error";
    std::vector<void *> owned;
    tlm_batch *batch = nullptr; // set while the context is a member of a batch
    uint8_t *mat = nullptr;     // device material map (cellIndex order), built on demand
};

namespace tlmi {

// Records the message on the context (or the thread's create error) and
// returns the code.
int fail(tlm_ctx *c, int code, const std::string &msg);

inline int hip_fail(tlm_ctx *c, hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(c, e == hipErrorOutOfMemory ? TLM_ERR_NOMEM : TLM_ERR_HIP, m);
}

#define HIPCHK(ctx, call)                                                                          \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return hip_fail((ctx), e_, #call);                                                     \
    } while (0)

template <class T> inline int dalloc(tlm_ctx *c, T **p, size_t count)
{
    *p = nullptr;
    if (count == 0)
        count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess)
    {
        *p = nullptr;
        return hip_fail(c, e, "hipMalloc");
    }
    c->owned.push_back(*p);
    return TLM_OK;
}

template <class T> inline void dfree(tlm_ctx *c, T *&p)
{
    if (!p)
        return;
    auto it = std::find(c->owned.begin(), c->owned.end(), (void *)p);
    if (it != c->owned.end())
        c->owned.erase(it);
    (void)hipFree(p);
    p = nullptr;
}

template <class T> inline int upload(tlm_ctx *c, T **dst, const T *src, size_t count)
{
    dfree(c, *dst);
    int rc = dalloc(c, dst, count);
    if (rc)
        return rc;
    if (count && src)
        HIPCHK(c, hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return TLM_OK;
}

// Device address of cell (k*ny + j)*nx + i: k*Pp + j*nx + i.
inline long long padded_cell(const tlm_ctx *c, long long cell)
{
    long long k = cell / c->nxy;
    return k * c->Pp + (cell - k * c->nxy);
}

inline bool cell_ok(const tlm_ctx *c, long long cell)
{
    return cell >= 0 && cell < (long long)c->nx * c->ny * c->nz;
}


// Scale buffer (0 or 1) of step n for the 16-bit formats.
inline int scale_buf(const tlm_ctx *c, int n)
{
    return (n - c->base) & 1;
}

inline float format_max(int pf)
{
    return pf == TLM_PULSE_F16 ? 65504.0f : 3.0e38f;
}

// Pulses of step n as seen by the small kernels.
tlm::PulseView pulse_view(const tlm_ctx *c, int n);
// 16-bit formats: lay out the slabs and store the uploaded state (first run).
int materialize(tlm_ctx *c, int step0);
// Second half of an upload, once the cell codes and block flags are on the
// device: material table, zeroed or uploaded state, stub buffer.
int install_material(tlm_ctx *c, bool any_diel, const float *yhat, const float *ghat, int nmat,
                     const float *V, const float *Vs);
// Running FP64 probe DFT over steps [n0, n1) (queued on stream s).
void queue_probe_dft(tlm_ctx *c, int n0, int n1, hipStream_t s);

} // namespace tlmi
