// TlmFarField.hip - near-to-far-field transform from the Huygens box DFTs
// on the device (ComputeFarFieldFromHuygens, TlmSolver.cpp:805-916).
//
// The reference evaluates, for every (theta, phi) of the output grid, a
// float sum over all Huygens faces of the face's equivalent currents times
// exp(+j k0 r.r'), on the host thread pool: O(directions x faces) sin/cos.
// Here:
//   k_ff_faces  one thread per face: position (cellCenter - domain centre,
//               TlmSolver.cpp:323-334), outward sign, and the currents
//               J' = n x (eta0 H), M = -(n x E) in world components
//               (TlmSolver.cpp:823-843), 16 floats per face;
//   k_ff_sums   grid (8-direction group) x (face range): each thread takes
//               faces and accumulates, for its group's 8 directions, the
//               vector sums S_J = sum J e^{j phase}, S_M = sum M e^{j phase}
//               (12 floats per direction); wave shuffles, then LDS, reduce
//               the block's partial sums;
//   k_ff_finish one thread per direction: the face-range partials summed in
//               double in a fixed order, projected on theta-hat / phi-hat:
//               N_theta = S_J . theta-hat etc., and
//               U = |L_phi + N_theta|^2 + |L_theta - N_phi|^2.
// The direction vectors are the reference's float values (computed on the
// host with the C library's sin/cos, TlmSolver.cpp:852-861) and the phase is
// the reference's float expression; the projections are taken after the sum
// instead of per face (the same linear map), so U agrees with the reference
// to float rounding, not bit for bit. The directivity and peak search run on
// the host over U exactly as TlmSolver.cpp:893-913.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function" // the step kernels of the shared header
#include "tlm_context.h"
#pragma clang diagnostic pop

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

using namespace tlmi;

namespace {

constexpr int kFfThreads = 256;
constexpr int kDirs = 8; // directions per block
constexpr int kAcc = 12; // S_J (re x,y,z; im x,y,z), S_M (same)

// Face record: px py pz pad | jr(3) ji(3) | mr(3) mi(3)
struct FfFace
{
    float4 a, b, c, d;
};

__device__ __constant__ int kTan[3][2] = {{1, 2}, {0, 2}, {0, 1}}; // POL, TlmSolver.cpp:23

// Currents of one face from its 8 DFT floats (E1, E2, eta0 H1, eta0 H2 as
// re, im), axis and outward sign (TlmSolver.cpp:823-843).
__device__ FfFace make_face(float px, float py, float pz, int ax, float s, const float *a)
{
    const int t1 = kTan[ax][0], t2 = kTan[ax][1];
    const float sa = (ax == 1) ? -1.0f : 1.0f;
    float jr[3] = {0.f, 0.f, 0.f}, ji[3] = {0.f, 0.f, 0.f};
    float mr[3] = {0.f, 0.f, 0.f}, mi[3] = {0.f, 0.f, 0.f};
    jr[t2] = s * sa * a[4];
    ji[t2] = s * sa * a[5];
    jr[t1] = -s * sa * a[6];
    ji[t1] = -s * sa * a[7];
    mr[t2] = -s * sa * a[0];
    mi[t2] = -s * sa * a[1];
    mr[t1] = s * sa * a[2];
    mi[t1] = s * sa * a[3];
    FfFace f;
    f.a = make_float4(px, py, pz, 0.f);
    f.b = make_float4(jr[0], jr[1], jr[2], ji[0]);
    f.c = make_float4(ji[1], ji[2], mr[0], mr[1]);
    f.d = make_float4(mr[2], mi[0], mi[1], mi[2]);
    return f;
}

// Faces given by the caller: positions, axes, outward signs.
__global__ void k_ff_faces_host(int nf, const float *pos, const int *axis, const int *sign,
                                const float *dft, FfFace *out)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    out[f] = make_face(pos[3 * f], pos[3 * f + 1], pos[3 * f + 2], axis[f], (float)sign[f],
                       dft + 8 * (long long)f);
}

// Faces of a context: padded cells, axes and the resident accumulators.
__global__ void k_ff_faces_ctx(int nf, const long long *pcell, const int *axis, const float *dft,
                               int nx, int ny, int nz, long long Pp, float ox, float oy, float oz,
                               float dl, FfFace *out)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    const long long p = pcell[f];
    const long long k = p / Pp;
    const long long r = p - k * Pp;
    const int j = (int)(r / nx);
    const int i = (int)(r - (long long)j * nx);
    // HuyFace::pos = cellCenter - (origin + (n dl) * 0.5) (TlmSolver.cpp:323, 334)
    const float cx = ox + (nx * dl) * 0.5f, cy = oy + (ny * dl) * 0.5f, cz = oz + (nz * dl) * 0.5f;
    const float px = (ox + (i + 0.5f) * dl) - cx;
    const float py = (oy + (j + 0.5f) * dl) - cy;
    const float pz = (oz + ((int)k + 0.5f) * dl) - cz;
    const int a = axis[f];
    const int co = a == 0 ? i : (a == 1 ? j : (int)k);
    const int nn = a == 0 ? nx : (a == 1 ? ny : nz);
    const float s = (2 * co + 1 < nn) ? -1.0f : 1.0f; // low side of the box: outward -e_a
    out[f] = make_face(px, py, pz, a, s, dft + 8 * (long long)f);
}

__device__ inline float wave_sum(float v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Partial sums of faces [fr*per, min(nf, (fr+1)*per)) for the 8 directions
// of group blockIdx.x: part[(dir * nfr + fr) * 12 + q].
__global__ void __launch_bounds__(kFfThreads)
    k_ff_sums(const FfFace *faces, int nf, int per, int nfr, const float *rhat, int ndir, float k0,
              float *part)
{
    __shared__ float rh[kDirs][3];
    __shared__ float wsum[kFfThreads / 64][kDirs * kAcc];
    const int d0 = blockIdx.x * kDirs;
    if (threadIdx.x < kDirs * 3)
    {
        const int d = threadIdx.x / 3, q = threadIdx.x % 3;
        rh[d][q] = d0 + d < ndir ? rhat[3 * (d0 + d) + q] : 0.f;
    }
    __syncthreads();
    float acc[kDirs][kAcc];
#pragma unroll
    for (int d = 0; d < kDirs; ++d)
#pragma unroll
        for (int q = 0; q < kAcc; ++q)
            acc[d][q] = 0.f;
    const int fr = blockIdx.y;
    const int f1 = min(nf, (fr + 1) * per);
    for (int f = fr * per + threadIdx.x; f < f1; f += kFfThreads)
    {
        const FfFace F = faces[f];
        const float jr[3] = {F.b.x, F.b.y, F.b.z}, ji[3] = {F.b.w, F.c.x, F.c.y};
        const float mr[3] = {F.c.z, F.c.w, F.d.x}, mi[3] = {F.d.y, F.d.z, F.d.w};
#pragma unroll
        for (int d = 0; d < kDirs; ++d)
        {
            // the reference's phase expression (TlmSolver.cpp:868-869)
            const float pha = k0 * (rh[d][0] * F.a.x + rh[d][1] * F.a.y + rh[d][2] * F.a.z);
            float s, c;
            sincosf(pha, &s, &c);
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                acc[d][q] += jr[q] * c - ji[q] * s;
                acc[d][3 + q] += jr[q] * s + ji[q] * c;
                acc[d][6 + q] += mr[q] * c - mi[q] * s;
                acc[d][9 + q] += mr[q] * s + mi[q] * c;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < kDirs; ++d)
#pragma unroll
        for (int q = 0; q < kAcc; ++q)
        {
            const float v = wave_sum(acc[d][q]);
            if (lane == 0)
                wsum[wave][d * kAcc + q] = v;
        }
    __syncthreads();
    if (threadIdx.x < kDirs * kAcc)
    {
        const int d = threadIdx.x / kAcc, q = threadIdx.x % kAcc;
        float v = 0.f;
        for (int w = 0; w < kFfThreads / 64; ++w)
            v += wsum[w][threadIdx.x];
        if (d0 + d < ndir)
            part[((long long)(d0 + d) * nfr + fr) * kAcc + q] = v;
    }
}

// U per direction from the partial sums (dirs: rhat, that, phat; 9 floats).
__global__ void k_ff_finish(const float *part, int nfr, const float *dirs, int ndir, float *U)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndir)
        return;
    double S[kAcc];
    for (int q = 0; q < kAcc; ++q)
        S[q] = 0.0;
    for (int fr = 0; fr < nfr; ++fr)
        for (int q = 0; q < kAcc; ++q)
            S[q] += part[((long long)d * nfr + fr) * kAcc + q];
    const float *t = dirs + 9 * d + 3, *p = dirs + 9 * d + 6;
    // radiation vectors projected (J' -> N, M -> L), TlmSolver.cpp:872-889
    const double NtR = S[0] * t[0] + S[1] * t[1] + S[2] * t[2];
    const double NtI = S[3] * t[0] + S[4] * t[1] + S[5] * t[2];
    const double NpR = S[0] * p[0] + S[1] * p[1];
    const double NpI = S[3] * p[0] + S[4] * p[1];
    const double LtR = S[6] * t[0] + S[7] * t[1] + S[8] * t[2];
    const double LtI = S[9] * t[0] + S[10] * t[1] + S[11] * t[2];
    const double LpR = S[6] * p[0] + S[7] * p[1];
    const double LpI = S[9] * p[0] + S[10] * p[1];
    const double aR = LpR + NtR, aI = LpI + NtI;
    const double bR = LtR - NpR, bI = LtI - NpI;
    U[d] = (float)(aR * aR + aI * aI + bR * bR + bI * bI);
}

struct DevBufs
{
    std::vector<void *> p;
    ~DevBufs()
    {
        for (void *q : p)
            (void)hipFree(q);
    }
    template <class T> hipError_t get(T **x, size_t n)
    {
        hipError_t e = hipMalloc((void **)x, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess)
            p.push_back(*x);
        else
            *x = nullptr;
        return e;
    }
};

// Buffers carved in order from a fixed device workspace (256-byte aligned
// pieces); nothing is allocated or freed.
struct Carve
{
    unsigned char *base = nullptr;
    size_t cap = 0, off = 0;
    static size_t piece(size_t bytes)
    {
        return (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    }
    template <class T> hipError_t get(T **x, size_t n)
    {
        const size_t b = piece(n * sizeof(T));
        if (off + b > cap)
        {
            *x = nullptr;
            return hipErrorOutOfMemory;
        }
        *x = reinterpret_cast<T *>(base + off);
        off += b;
        return hipSuccess;
    }
};

// Face ranges of k_ff_sums: enough that the grid fills the chip, each range
// at least a block of faces.
int face_ranges(int nf, int ndir)
{
    const int ngrp = (ndir + kDirs - 1) / kDirs;
    return std::max(1, std::min((nf + kFfThreads - 1) / kFfThreads, (2048 + ngrp - 1) / ngrp));
}

// Workspace of one transform carved by Carve: face records, the direction
// tables, the partial sums and U.
size_t far_field_workspace(int nf, int ndir)
{
    const size_t nfr = (size_t)face_ranges(nf, ndir);
    return Carve::piece((size_t)nf * sizeof(FfFace)) + Carve::piece(9 * (size_t)ndir * sizeof(float)) +
           Carve::piece(3 * (size_t)ndir * sizeof(float)) +
           Carve::piece((size_t)ndir * nfr * kAcc * sizeof(float)) + Carve::piece((size_t)ndir * sizeof(float));
}

#define FFCHK(call)                                                                                \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return hip_fail(c, e_, #call);                                                         \
    } while (0)

// Directions, sums, U and the summary for prepared faces (on stream st);
// device buffers from b (hipMalloc'ed DevBufs or a Carve of a workspace).
template <class Bufs>
int far_field_from_faces(tlm_ctx *c, hipStream_t st, Bufs &b, const FfFace *faces, int nf,
                         float f0, int nT, int nP, float *U, tlm_far_field_summary *out)
{
    const float PI = (float)M_PI;
    const float k0 = 2.0f * PI * f0 / 299792458.0f; // TlmSolver.cpp:813
    const int ndir = nT * nP;
    // direction vectors as the reference evaluates them (TlmSolver.cpp:852-861)
    std::vector<float> dirs(9 * (size_t)ndir), rh(3 * (size_t)ndir);
    for (int ti = 0; ti < nT; ++ti)
    {
        const float th = PI * ti / (nT - 1);
        const float s = std::sin(th), ct = std::cos(th);
        for (int pi = 0; pi < nP; ++pi)
        {
            const float ph = 2.0f * PI * pi / (nP - 1);
            const float cp = std::cos(ph), sp = std::sin(ph);
            const float v[9] = {s * cp, s * sp, ct, ct * cp, ct * sp, -s, -sp, cp, 0.0f};
            const size_t d = (size_t)ti * nP + pi;
            for (int q = 0; q < 9; ++q)
                dirs[9 * d + q] = v[q];
            for (int q = 0; q < 3; ++q)
                rh[3 * d + q] = v[q];
        }
    }
    const int ngrp = (ndir + kDirs - 1) / kDirs;
    const int nfr = face_ranges(nf, ndir);
    const int per = (nf + nfr - 1) / nfr;
    float *d_dirs, *d_rh, *d_part, *d_U;
    FFCHK(b.get(&d_dirs, dirs.size()));
    FFCHK(b.get(&d_rh, rh.size()));
    FFCHK(b.get(&d_part, (size_t)ndir * nfr * kAcc));
    FFCHK(b.get(&d_U, (size_t)ndir));
    FFCHK(hipMemcpyAsync(d_dirs, dirs.data(), dirs.size() * sizeof(float), hipMemcpyHostToDevice, st));
    FFCHK(hipMemcpyAsync(d_rh, rh.data(), rh.size() * sizeof(float), hipMemcpyHostToDevice, st));
    k_ff_sums<<<dim3((unsigned)ngrp, (unsigned)nfr), kFfThreads, 0, st>>>(faces, nf, per, nfr, d_rh,
                                                                          ndir, k0, d_part);
    FFCHK(hipGetLastError());
    k_ff_finish<<<(ndir + 255) / 256, 256, 0, st>>>(d_part, nfr, d_dirs, ndir, d_U);
    FFCHK(hipGetLastError());
    std::vector<float> u((size_t)ndir);
    FFCHK(hipMemcpyAsync(u.data(), d_U, u.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    FFCHK(hipStreamSynchronize(st));
    // directivity and peak (TlmSolver.cpp:893-913)
    float uMax = 0.0f;
    int iMax = 0;
    double prad = 0.0;
    const float dTh = PI / (nT - 1), dPh = 2.0f * PI / (nP - 1);
    for (int ti = 0; ti < nT; ++ti)
    {
        const float th = PI * ti / (nT - 1);
        const float wTh = (ti == 0 || ti == nT - 1) ? 0.5f : 1.0f;
        for (int pi = 0; pi < nP; ++pi)
        {
            const float wPh = (pi == 0 || pi == nP - 1) ? 0.5f : 1.0f;
            const float v = u[(size_t)ti * nP + pi];
            if (v > uMax)
            {
                uMax = v;
                iMax = ti * nP + pi;
            }
            prad += (double)v * std::sin(th) * wTh * wPh * dTh * dPh;
        }
    }
    if (U)
        std::copy(u.begin(), u.end(), U);
    if (out)
    {
        out->u_max = uMax;
        out->directivity = prad > 0 ? (float)(4.0 * M_PI * uMax / prad) : 0.0f;
        out->peak_theta_deg = 180.0f * (iMax / nP) / (nT - 1);
        out->peak_phi_deg = 360.0f * (iMax % nP) / (nP - 1);
    }
    return TLM_OK;
}

// ---- allocation-free variant ----
// Its workspace: a file-scope device array of kScratchBytes, one instance per
// GPU the library runs on (held until the process ends). Host code reaches
// it only through the address hipGetSymbolAddress gives on each device.
constexpr size_t kScratchBytes = (size_t)128 << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// Device address of the workspace on device dev, which must be current;
// looked up once per device.
int scratch_on(tlm_ctx *c, int dev, unsigned char **out)
{
    if (dev < 0 || dev >= kMaxDevices)
        return fail(c, TLM_ERR_UNSUPPORTED, "allocation-free far field: device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev])
    {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[dev];
    return TLM_OK;
}

} // namespace

extern "C" {

int tlm_far_field(int device, int nf, const float *pos, const int32_t *axis,
                  const int32_t *out_sign, const float *dft8, int samples, float f0, int n_theta,
                  int n_phi, float *U, tlm_far_field_summary *out)
{
    tlm_ctx *c = nullptr; // errors go to tlm_last_error(NULL)
    if (nf < 0 || n_theta < 2 || n_phi < 2 || (nf > 0 && (!pos || !axis || !out_sign || !dft8)))
        return fail(nullptr, TLM_ERR_ARG, "tlm_far_field: bad arguments");
    if (nf == 0 || samples == 0)
        return fail(nullptr, TLM_ERR_STATE, "tlm_far_field: no Huygens data");
    for (int f = 0; f < nf; ++f)
        if (axis[f] < 0 || axis[f] > 2 || (out_sign[f] != 1 && out_sign[f] != -1))
            return fail(nullptr, TLM_ERR_ARG, "tlm_far_field: face axis or sign out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, TLM_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev)
        return fail(nullptr, TLM_ERR_ARG, "tlm_far_field: device out of range");
    FFCHK(hipSetDevice(device));
    hipStream_t st;
    FFCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc;
    {
        DevBufs b;
        float *d_pos = nullptr, *d_dft = nullptr;
        int *d_ax = nullptr, *d_sg = nullptr;
        FfFace *d_faces = nullptr;
        hipError_t e = hipSuccess;
        if ((e = b.get(&d_pos, 3 * (size_t)nf)) != hipSuccess || (e = b.get(&d_dft, 8 * (size_t)nf)) != hipSuccess ||
            (e = b.get(&d_ax, nf)) != hipSuccess || (e = b.get(&d_sg, nf)) != hipSuccess ||
            (e = b.get(&d_faces, nf)) != hipSuccess)
        {
            (void)hipStreamDestroy(st);
            return hip_fail(nullptr, e, "hipMalloc");
        }
        (void)hipMemcpyAsync(d_pos, pos, 3 * (size_t)nf * sizeof(float), hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(d_dft, dft8, 8 * (size_t)nf * sizeof(float), hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(d_ax, axis, (size_t)nf * sizeof(int), hipMemcpyHostToDevice, st);
        (void)hipMemcpyAsync(d_sg, out_sign, (size_t)nf * sizeof(int), hipMemcpyHostToDevice, st);
        k_ff_faces_host<<<(nf + 255) / 256, 256, 0, st>>>(nf, d_pos, d_ax, d_sg, d_dft, d_faces);
        rc = far_field_from_faces(nullptr, st, b, d_faces, nf, f0, n_theta, n_phi, U, out);
    }
    (void)hipStreamDestroy(st);
    return rc;
}

int tlm_far_field_ctx(tlm_ctx *c, const float origin[3], int samples, float f0, int n_theta,
                      int n_phi, float *U, tlm_far_field_summary *out)
{
    if (!c || !origin || n_theta < 2 || n_phi < 2)
        return fail(c, TLM_ERR_ARG, "tlm_far_field_ctx: bad arguments");
    if (c->nh == 0 || samples == 0)
        return fail(c, TLM_ERR_STATE, "tlm_far_field_ctx: no Huygens data");
    FFCHK(hipSetDevice(c->dev));
    DevBufs b;
    FfFace *d_faces;
    FFCHK(b.get(&d_faces, (size_t)c->nh));
    k_ff_faces_ctx<<<(c->nh + 255) / 256, 256, 0, c->stream>>>(c->nh, c->hcell, c->haxis, c->huy8,
                                                                c->nx, c->ny, c->nz, c->Pp, origin[0],
                                                                origin[1], origin[2], c->dl, d_faces);
    FFCHK(hipGetLastError());
    return far_field_from_faces(c, c->stream, b, d_faces, c->nh, f0, n_theta, n_phi, U, out);
}

int tlm_far_field_ctx_noalloc(tlm_ctx *c, const float origin[3], int samples, float f0, int n_theta,
                              int n_phi, float *U, tlm_far_field_summary *out)
{
    if (!c || !origin || n_theta < 2 || n_phi < 2)
        return fail(c, TLM_ERR_ARG, "tlm_far_field_ctx_noalloc: bad arguments");
    if (c->nh == 0 || samples == 0)
        return fail(c, TLM_ERR_STATE, "tlm_far_field_ctx_noalloc: no Huygens data");
    const size_t need = far_field_workspace(c->nh, n_theta * n_phi);
    if (need > kScratchBytes)
        return fail(c, TLM_ERR_UNSUPPORTED,
                    "tlm_far_field_ctx_noalloc: " + std::to_string(c->nh) + " faces x " +
                        std::to_string(n_theta * n_phi) + " directions need " + std::to_string(need) +
                        " bytes of workspace, above its fixed " + std::to_string(kScratchBytes) +
                        " (tlm_far_field_ctx allocates what it needs)");
    FFCHK(hipSetDevice(c->dev));
    Carve b;
    int rc = scratch_on(c, c->dev, &b.base);
    if (rc)
        return rc;
    b.cap = kScratchBytes;
    FfFace *d_faces;
    FFCHK(b.get(&d_faces, (size_t)c->nh));
    k_ff_faces_ctx<<<(c->nh + 255) / 256, 256, 0, c->stream>>>(c->nh, c->hcell, c->haxis, c->huy8,
                                                                c->nx, c->ny, c->nz, c->Pp, origin[0],
                                                                origin[1], origin[2], c->dl, d_faces);
    FFCHK(hipGetLastError());
    return far_field_from_faces(c, c->stream, b, d_faces, c->nh, f0, n_theta, n_phi, U, out);
}

} // extern "C"
