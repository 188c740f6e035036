/*
 * GpuTlmHip.h - C ABI of the MI355X (gfx950) SCN TLM time step.
 *
 * This is the thin extern "C" layer that replaces the OpenGL compute-shader
 * dispatch behind the reference's GPU entry point
 *     bool RunGpuTlm(TlmSolver &s, std::string &msg);     (GpuTlm.h:15)
 * The host glue emsolver_amd/host/GpuTlm.cpp implements RunGpuTlm on top of
 * these functions; nothing here uses C++ or torch types.
 *
 * Conventions
 *   - Return codes: TLM_OK (0) on success, a negative TLM_ERR_* otherwise;
 *     tlm_last_error() gives the message (the reference's bool + msg).
 *   - Cells are linear indices (k*ny + j)*nx + i (VoxelGridSpec::cellIndex,
 *     Geometry.h:91-94). Host arrays are caller owned. Pulses cross the
 *     boundary in the reference's array-of-structures order (12 floats per
 *     cell, port = axis*4 + side*2 + pol, TlmSolver.cpp:20-29; 3 stub floats
 *     per cell); the device keeps them as structure-of-arrays planes.
 *   - One context is used by one host thread at a time; separate contexts
 *     may be driven from separate threads. Work is queued on the context's
 *     own HIP stream; tlm_sync() waits for it.
 */
#ifndef GPU_TLM_HIP_H
#define GPU_TLM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TLM_OK = 0,
    TLM_ERR_ARG = -1,         /* bad argument / shape */
    TLM_ERR_NO_DEVICE = -2,   /* no HIP device (RunGpuTlm returns false) */
    TLM_ERR_NOMEM = -3,       /* device allocation failed (GpuTlm.cpp:664-668) */
    TLM_ERR_HIP = -4,         /* other HIP runtime error */
    TLM_ERR_STATE = -5,       /* call out of order (e.g. run before upload) */
    TLM_ERR_UNSUPPORTED = -6, /* configuration this build does not handle */
    TLM_ERR_RANGE = -7        /* 16-bit storage: a value outgrew its planned slab scale */
};

/* Link-pulse storage type. Arithmetic is always FP32.
 * TLM_PULSE_F16 / TLM_PULSE_BF16 (no counterpart in the reference): each
 * z-slab holds its pulses either as FP32 (slabs within fp32_halo slabs of a
 * soft source, where the strong reactive near field makes 16-bit rounding
 * costly) or as IEEE binary16 / bfloat16 codes times a per-slab power of two
 * 2^E. E is planned one step ahead from the largest value each slab sent the
 * step before (one step grows the max norm at most 5x and moves a pulse one
 * slab), so no value can overflow; tlm_sync reports TLM_ERR_RANGE if one
 * ever did. A value is rounded (to nearest even) once, where it is stored.
 * Stubs, probe records, DFT and energy accumulators keep FP32 / FP64. */
enum { TLM_PULSE_F32 = 0, TLM_PULSE_F16 = 1, TLM_PULSE_BF16 = 2 };

typedef struct tlm_grid_desc {
    int32_t nx, ny, nz;  /* VoxelGridSpec extents (Geometry.h:87) */
    float dl;            /* cell size in metres */
    float rho;           /* outer boundary reflection, TlmConfig::boundaryRho (TlmSolver.h:119) */
    int32_t pulse_type;  /* TLM_PULSE_* */
    int32_t device;      /* HIP device ordinal */
    int32_t fp32_halo;   /* 16-bit formats: FP32 z-slabs on each side of a source slab;
                            0 = default (8), < 0 = none */
} tlm_grid_desc;

typedef struct tlm_ctx tlm_ctx;
typedef struct tlm_batch tlm_batch;

/* Number of visible HIP devices (0 without a GPU). */
int tlm_device_count(void);

/* Adapter name of a device; replaces GlRendererName() (GpuTlm.cpp:463-467). */
const char *tlm_device_name(int device);

/* Creates a context: picks the device, allocates the ping-pong pulse planes
 * and the material map. Replaces the GL context/buffer set-up of
 * GpuTlm::run (GpuTlm.cpp:533-668). On failure *out is NULL and the message
 * is available from tlm_last_error(NULL). */
int tlm_create(const tlm_grid_desc *desc, tlm_ctx **out);

/* Frees every device allocation of the context (RAII GpuCtx, GpuTlm.cpp:384-407). */
void tlm_destroy(tlm_ctx *ctx);

/* Last error message of the context (or of the last failed tlm_create when
 * ctx is NULL). Never NULL. */
const char *tlm_last_error(const tlm_ctx *ctx);

/* Uploads the material map, the per-dielectric stub admittances and the
 * initial pulses (GpuTlm.cpp:564-585). mat has nx*ny*nz bytes: 0 air,
 * 1 PEC, id >= 2 dielectric with yhat[id-2], ghat[id-2] (TlmSolver.cpp:
 * 272-276). V_aos12 (12 floats per cell) and Vs_aos3 (3 per cell) may be
 * NULL for zeros. Stubs exist iff a cell has id >= 2 (TlmSolver.cpp:277-283). */
int tlm_upload(tlm_ctx *ctx, const uint8_t *mat, const float *yhat, const float *ghat, int nmat,
               const float *V_aos12, const float *Vs_aos3);

/* Soft sources in injection order: TlmSolver::sources, then every port's
 * cells (TlmSolver.cpp:501-505; the GL source list GpuTlm.cpp:588-597).
 * Each adds q = 0.5*amp*sv(n) to the 4 ports polarised along pol_axis. */
int tlm_set_sources(tlm_ctx *ctx, const int64_t *cells, const int32_t *pol_axis, const float *amp,
                    int n);

/* Per-step host scalars for steps [0, nsteps): sv(n) = waveformValue(n)
 * (TlmSolver.cpp:410-436) and the f0 phasor cph/sph (TlmSolver.cpp:619-620;
 * GpuTlm.cpp:716-719). Steps n >= settle_steps accumulate the Js and
 * Huygens DFTs. Allocates the per-step record buffers, so it is called
 * before the monitor set-up functions below. */
int tlm_set_schedule(tlm_ctx *ctx, const float *sv, const float *cph, const float *sph, int nsteps,
                     int settle_steps);

/* E-field probes, recorded every step after injection:
 * E_j = inv_dl * VjOf(cell, j) with inv_dl = -1/dl (TlmSolver.cpp:705-713). */
int tlm_set_probes(tlm_ctx *ctx, const int64_t *cells, int n, float inv_dl);

/* Running FP64 DFT of every probe component: cos_tab/sin_tab hold
 * [nsteps][nfreq] phasors; re += E*cos, im -= E*sin, summed in step order. */
int tlm_set_probe_dft(tlm_ctx *ctx, const double *cos_tab, const double *sin_tab, int nfreq);

/* Lumped ports: port p owns cells[cell_offsets[p] .. cell_offsets[p+1]).
 * Records V = sequential float sum of VjOf over the cells and
 * I = -2 * ZIloop / eta0 around the middle cell (TlmSolver.cpp:677-703). */
int tlm_set_ports(tlm_ctx *ctx, const int64_t *cells, const int32_t *cell_offsets,
                  const int32_t *pol_axis, int nports);

/* PEC surface faces for the Js monitor, in BuildPecSurfaceFaces order
 * (Geometry.cpp:402-436): air cell and normal axis (TlmSolver.cpp:627-649). */
int tlm_set_faces(tlm_ctx *ctx, const int64_t *air_cells, const int32_t *axes, int n);

/* Huygens box faces (TlmSolver.cpp:313-359, 651-675). */
int tlm_set_huygens(tlm_ctx *ctx, const int64_t *cells, const int32_t *axes, int n);

/* Energy sum(V^2) in FP64 every `every` steps (16 in TlmSolver.cpp:767);
 * 0 disables. */
int tlm_set_energy_every(tlm_ctx *ctx, int every);

/* |E| cut plane normal to `axis` at `index` (clamped into the grid), taken
 * on the post-injection pulses of every even step, PEC cells 0
 * (TlmSolver.cpp:715-745). axis < 0 disables. */
int tlm_set_plane(tlm_ctx *ctx, int axis, int index);

/* Playback frames every `every` steps (TlmSolver.cpp:747-763): each holds
 * the instantaneous |Js| of every face and the latest cut plane. At most
 * max_frames are kept; 0 disables. */
int tlm_set_frames(tlm_ctx *ctx, int every, int max_frames);

/* Queues steps [step0, step0 + nsteps): inject -> monitors -> fused
 * scatter/connect (TlmSolver::step, TlmSolver.cpp:485-513). Asynchronous. */
int tlm_run(tlm_ctx *ctx, int step0, int nsteps);

/* Waits for the queued work; reports asynchronous HIP errors and, for the
 * 16-bit formats, TLM_ERR_RANGE. A 16-bit context runs contiguous step
 * blocks (each tlm_run starts where the previous one ended). */
int tlm_sync(tlm_ctx *ctx);

/* 16-bit formats (no counterpart in the reference): per z-slab, the decode
 * scale 2^E of the current step, the largest |value| the slab's cells sent
 * in the last step, and the slab's storage format (TLM_PULSE_*). Any pointer
 * may be NULL. Returns nz, or 0 for FP32 contexts and before the first run. */
int tlm_read_scales(tlm_ctx *ctx, float *scale, float *smax, int32_t *fmt);

/* Rounds in[i] * inv_scale to binary16 (fmt TLM_PULSE_F16) or bfloat16
 * (TLM_PULSE_BF16) with the device conversion the step kernel uses, for
 * checking it against the IEEE round-to-nearest-even definition. */
int tlm_device_encode(int device, int fmt, const float *in, float inv_scale, uint16_t *out,
                      int64_t n);

/* Current pulses back in reference AoS order (for snapshot parity; the GL
 * path never read V back). Vs_aos3 may be NULL. Synchronous. */
int tlm_read_state(tlm_ctx *ctx, float *V_aos12, float *Vs_aos3);

/* Probe records of steps [step0, step0+nsteps): [nsteps][nprobes][3]. */
int tlm_read_probes(tlm_ctx *ctx, int step0, int nsteps, float *e);

/* FP64 probe DFT accumulators: [nprobes][3][nfreq] each. */
int tlm_read_probe_dft(tlm_ctx *ctx, double *re, double *im);

/* Port records of steps [step0, step0+nsteps): [nsteps][nports][2] (V, I),
 * the GL portRec layout (GpuTlm.cpp:321-322). */
int tlm_read_ports(tlm_ctx *ctx, int step0, int nsteps, float *vi);

/* Js monitor: instantaneous |Js| per face and the float DFT accumulators
 * [nfaces][4] = re1, im1, re2, im2 (GpuTlm.cpp:242-248). Either may be NULL. */
int tlm_read_js(tlm_ctx *ctx, float *inst, float *dft4);

/* Huygens accumulators, [nfaces][8] (TlmSolver.cpp:667-671). */
int tlm_read_huygens(tlm_ctx *ctx, float *dft8);

/* Energy samples recorded so far (steps and FP64 sums); returns the count,
 * writing at most cap entries. */
int tlm_read_energy(tlm_ctx *ctx, int32_t *steps, double *vals, int cap);

/* Latest cut plane: n1*n2 floats (row q2, column q1 as planeBuf,
 * TlmSolver.cpp:722-743). *step is the step it was taken at, -1 if none. */
int tlm_read_plane(tlm_ctx *ctx, float *out, int *n1, int *n2, int *step);

/* Recorded frames: returns the count; fills at most cap frames: steps[f],
 * js[f * nfaces ...], planes[f * n1*n2 ...] (either array may be NULL). */
int tlm_read_frames(tlm_ctx *ctx, int32_t *steps, float *js, float *planes, int cap);

/* Runs steps [step0, step0+nsteps) and measures them with HIP events on the
 * context stream: *ms_total spans the whole queue of the steps. With
 * kernel_every = k > 0 (and ms_kernel not NULL) the fused scatter/connect
 * launches of the steps step0, step0+k, step0+2k, ... carry their own start
 * and end events: *ms_kernel is their summed duration and *nkernel (may be
 * NULL) their count. Events on a launch cost host time (~4 us per launch at
 * 256^3, ~20 us at 64^3), so a sample (k > 1) perturbs the run less than
 * k = 1; k = 0 measures no launch. */
int tlm_time_steps(tlm_ctx *ctx, int step0, int nsteps, int kernel_every, float *ms_total,
                   float *ms_kernel, int *nkernel);

/* Device-copy bandwidth: copies `bytes` between two device buffers `iters`
 * times with a plain grid-stride float4 kernel (k_copy) and reports
 * read+write GB/s and the average time of one copy. The roofline reference
 * for the HBM-bound step; no counterpart in the reference. */
int tlm_copy_bandwidth(int device, size_t bytes, int iters, float *gbps, float *ms_per_copy);

/* Tuning of the fused scatter/connect kernel (no effect on results, which
 * are bit-identical for every setting). No counterpart in the reference. */
typedef struct tlm_tuning {
    int32_t kernel;      /* 0: one cell per lane, 64-cell strips per wavefront (only variant) */
    int32_t planes;      /* z slabs marched per wave task; 0 = automatic. Default 1 with FP32
                            pulses; with 16-bit pulses 3 on grids of at least 48 M cells,
                            else 4 */
    int32_t block_waves; /* wavefronts per workgroup, 1..4; 0 = automatic: 2 for FP32 steps
                            under the occupancy cap (below), else 4 */
    int32_t nontemporal; /* non-temporal stores of the next-step pulses: 0 no (default), 1 yes,
                            2 when the grid has dielectric cells */
    int32_t grid_blocks; /* > 0: persistent grid of that many workgroups; 0 = one per task group */
    int32_t fold;        /* FP32 with at most 32 source cells: add the soft sources inside the
                            step instead of a separate injection launch. 0 no, 1 yes, 2 (default)
                            for grids without stub cells, and for grids with them of at most
                            4 M cells or whose step runs under the occupancy cap (below) */
    int32_t occupancy;   /* most waves per SIMD the step kernel may hold (3..8), enforced by a
                            share of the CU's LDS per workgroup; 0 = what its registers allow.
                            Default -1: 5 for FP32 grids of at least 32 M cells, else none */
} tlm_tuning;

int tlm_set_tuning(tlm_ctx *ctx, const tlm_tuning *t);
int tlm_get_tuning(const tlm_ctx *ctx, tlm_tuning *t);

/* Launch shape chosen for the fused kernel: strips per slab, slabs per wave
 * task, workgroup count, threads per workgroup. For reporting. */
int tlm_launch_shape(const tlm_ctx *ctx, int *rows_per_group, int *planes_per_group,
                     int *groups, int *threads_per_group);

/* ScatterUnitarityError / DielScatterUnitarityError (TlmSolver.cpp:116-169)
 * evaluated with the device scatter: returns max |S^T W S - W| for the air
 * node (epsr <= 1) or the stub node with yhat = 4(epsr-1), or a negative
 * TLM_ERR_* code cast to float when no device is present. */
float tlm_scatter_unitarity_error(float epsr);

/* Host helpers shared by the glue code (no device needed). */

/* BuildPecSurfaceFaces (Geometry.cpp:402-436): fills at most cap faces,
 * returns the total count. air_cells/axes/signs may be NULL to count. */
int64_t tlm_build_pec_faces(int nx, int ny, int nz, const uint8_t *mat, int64_t *air_cells,
                            int32_t *axes, int32_t *signs, int64_t cap);

/* TlmSolver::buildHuygens (TlmSolver.cpp:313-359): box inset `offset`
 * cells, AIR cells only. Returns the count (0 when offset < 2). */
int64_t tlm_build_huygens(int nx, int ny, int nz, const uint8_t *mat, int offset,
                          int64_t *cells, int32_t *axes, int32_t *out_signs, int64_t cap);

/* FP64 DFT phasor tables for tlm_set_probe_dft: ph = 2*pi*f*(n*dt),
 * cos_tab[n*nfreq+q] = cos(ph), sin_tab[...] = sin(ph). */
void tlm_dft_tables(const double *freqs, int nfreq, double dt, int nsteps, double *cos_tab,
                    double *sin_tab);

/* Waveform types of TlmConfig (TlmSolver.h:70) plus the Gaussian-derivative
 * pulse sv = -sqrt(2e) u exp(-u^2), u = (n - delay)/width (in steps). */
enum { TLM_WF_CW_RAMPED = 0, TLM_WF_GAUSS_SINE = 1, TLM_WF_GAUSS_PULSE = 2, TLM_WF_GAUSS_DERIV = 3 };

/* Per-step host scalars for tlm_set_schedule: sv(n) = waveformValue(n)
 * (TlmSolver.cpp:410-436) and cph/sph of the f0 phase 2*pi*f0*(n*dt)
 * (TlmSolver.cpp:619-620), all in float like the reference, for steps
 * step0 .. step0+nsteps-1 (out[n - step0]). Any output pointer may be NULL. */
void tlm_schedule_tables(int waveform, float f0, float dt, int step0, int nsteps, float gd_width,
                         float gd_delay, float *sv, float *cph, float *sph);

/* ---------------------------------------------------------------------
 * Batched execution (no counterpart in the reference, whose GPU path runs
 * one grid per OpenGL context): many independent contexts advanced by the
 * same launches, one tick = one step of every active member. Members are
 * ordinary contexts set up with the calls above (probes and ports only: no
 * Js, Huygens, plane or frame monitors; no energy records while batched).
 * A member gives the same bits alone and in a batch. While a context is a
 * member, tlm_run and the tlm_set_* / tlm_upload calls on it are refused;
 * its readers work after tlm_batch_sync.
 * ------------------------------------------------------------------- */
typedef struct tlm_batch_tuning {
    int32_t planes;      /* z slabs per wave task (cells per lane); 0 = all of nz */
    int32_t block_waves; /* wavefronts (64-cell strips) per workgroup, 1..4; 0 = 4 */
    int32_t grid_blocks; /* > 0: persistent grid of that many workgroups; 0 = one per group */
    int32_t graph_ticks; /* ticks per captured HIP graph (default 32); 0 = plain launches */
    int32_t planes16;    /* z slabs per wave task of the 16-bit members (default 4); 0 = all */
} tlm_batch_tuning;

int tlm_batch_create(int device, tlm_batch **out);
void tlm_batch_destroy(tlm_batch *b);
const char *tlm_batch_last_error(const tlm_batch *b);

/* ctx joins at its next step and runs nsteps more steps (within its
 * schedule), starting with the next tick. Refused (TLM_ERR_UNSUPPORTED) for
 * a context with Js, Huygens, plane or frame monitors or energy records
 * (tlm_set_energy_every(ctx, 0) first). */
int tlm_batch_add(tlm_batch *b, tlm_ctx *ctx, int nsteps);
/* Waits for queued ticks; ctx leaves the batch and is a plain context again.
 * TLM_ERR_RANGE (ctx removed all the same) if ctx's 16-bit storage exceeded
 * a slab scale while batched: its results are invalid, the other members'
 * are not affected. */
int tlm_batch_remove(tlm_batch *b, tlm_ctx *ctx);
/* Queues nticks ticks (asynchronous); finished members idle. */
int tlm_batch_run(tlm_batch *b, int nticks);
/* Waits for the queued ticks and updates every member (step counters,
 * running probe DFT, the per-member 16-bit range flag that tlm_batch_remove
 * reports). */
int tlm_batch_sync(tlm_batch *b);
/* Members and their remaining steps (as of the ticks queued); returns the
 * member count. */
int tlm_batch_members(tlm_batch *b, tlm_ctx **members, int32_t *remaining, int cap);
int tlm_batch_set_tuning(tlm_batch *b, const tlm_batch_tuning *t);
int tlm_batch_get_tuning(const tlm_batch *b, tlm_batch_tuning *t);
/* Runs nticks ticks timed with HIP events; cell_updates = cells x steps of
 * the members active in them. */
int tlm_batch_time(tlm_batch *b, int nticks, float *ms, int64_t *cell_updates);
/* Online autotuner: times real ticks of the batch under each candidate
 * launch shape (slabs per wave task, waves per workgroup, grid) and keeps
 * the fastest per cell update; results do not depend on the shape. */
int tlm_batch_autotune(tlm_batch *b, int ticks_per_candidate, tlm_batch_tuning *best);

/* Free and total device memory (admission control of a serving loop). */
int tlm_device_memory(int device, size_t *free_bytes, size_t *total_bytes);
/* Device bytes a context of this shape allocates (pulses, stubs if any
 * dielectric, cell codes and per-step work buffers; monitors excluded). */
int64_t tlm_estimate_bytes(const tlm_grid_desc *desc, int has_stubs);

/* ---------------------------------------------------------------------
 * Scene -> voxel set-up on the device: the step's input side.
 *
 * A context owns a device material map (uint8 per cell, cellIndex order,
 * MAT_AIR = 0, MAT_PEC = 1, dielectric ids >= 2), created all AIR on the
 * first call below. The voxelisers write it in call order, with the
 * reference's results bit for bit; the map then becomes the grid's cell
 * codes through tlm_upload_mat, with no host pass over the cells. The
 * reference does this on one host thread before every run
 * (TMainForm::voxelizeScene, MainUnit.cpp:431-459: dielectrics first, then
 * PEC, so PEC wins overlaps; the feed gap is carved after, :636-660).
 * Refused while the context is a batch member.
 * ------------------------------------------------------------------- */
typedef struct tlm_voxel_grid {
    float origin[3]; /* VoxelGridSpec::origin (Geometry.h:88): corner of cell (0,0,0) */
    float dl;        /* VoxelGridSpec::dl; the grid extents are the context's */
} tlm_voxel_grid;

/* mat.assign(n, mat_id) (MainUnit.cpp:435). */
int tlm_mat_fill(tlm_ctx *ctx, uint8_t mat_id);
/* Cells lo <= (i,j,k) < hi (clipped to the grid) set to mat_id. No
 * counterpart in the reference: the synthetic dielectric-box and dipole
 * scenes are cell ranges. */
int tlm_mat_cells_box(tlm_ctx *ctx, const int32_t lo[3], const int32_t hi[3], uint8_t mat_id);
/* VoxelizeMesh (Geometry.cpp:369-397; solid = 0) or VoxelizeMeshSolid
 * (Geometry.cpp:551-606; solid = 1) of the triangle mesh (verts: 3 floats
 * per vertex, idx: 3 vertex indices per triangle) translated by offset. */
int tlm_voxelize_mesh(tlm_ctx *ctx, const tlm_voxel_grid *g, const float *verts, int32_t nverts,
                      const int32_t *idx, int32_t ntri, const float offset[3], uint8_t mat_id,
                      int solid);
/* VoxelizeWire (Geometry.cpp:608-642): polyline pts (3 floats each) of the
 * given radius, translated by offset. */
int tlm_voxelize_wire(tlm_ctx *ctx, const tlm_voxel_grid *g, const float *pts, int32_t npts,
                      float radius, const float offset[3], uint8_t mat_id);
/* Feed gap of a fed object (MainUnit.cpp:636-660): the cells sampled along
 * the inner 70% of (a+offset)->(b+offset) are carved to AIR; returns their
 * count (sorted, unique, at most cap written to cells) and the dominant axis
 * of b-a in *pol, or a negative TLM_ERR_*. */
int tlm_carve_feed(tlm_ctx *ctx, const tlm_voxel_grid *g, const float a[3], const float b[3],
                   const float offset[3], int64_t *cells, int32_t cap, int32_t *pol);
/* Copies between the device map and a host array of nx*ny*nz bytes. */
int tlm_mat_read(tlm_ctx *ctx, uint8_t *mat);
int tlm_mat_write(tlm_ctx *ctx, const uint8_t *mat);
/* tlm_upload with the device map as mat and a zero state (TlmSolver::setup,
 * TlmSolver.cpp:259-304): cell codes, air-block flags and the dielectric
 * check are computed on the device. */
int tlm_upload_mat(tlm_ctx *ctx, const float *yhat, const float *ghat, int nmat);
/* BuildPecSurfaceFaces (Geometry.cpp:402-436) on the device map, installed
 * as the surface-current faces (as tlm_set_faces). Returns the face count
 * or a negative TLM_ERR_*. */
int64_t tlm_pec_faces_mat(tlm_ctx *ctx);
/* TlmSolver::buildHuygens (TlmSolver.cpp:313-359) on the device map,
 * installed as the Huygens faces (as tlm_set_huygens). Returns the count. */
int64_t tlm_huygens_mat(tlm_ctx *ctx, int offset);
/* Installed face lists (unpadded cells, axes); return the count. */
int64_t tlm_read_faces(tlm_ctx *ctx, int64_t *cells, int32_t *axes, int64_t cap);
int64_t tlm_read_huygens_faces(tlm_ctx *ctx, int64_t *cells, int32_t *axes, int64_t cap);

/* ---------------------------------------------------------------------
 * Near-to-far-field transform (ComputeFarFieldFromHuygens,
 * TlmSolver.cpp:805-916) on the device: the O(directions x faces) phase sums
 * of the Huygens box currents at f0, the intensity U on an
 * n_theta x n_phi grid (theta-major, theta in [0, pi], phi in [0, 2 pi]),
 * and the directivity / peak of FarFieldData (TlmSolver.cpp:893-913).
 * U agrees with the reference's float sums to rounding (the projections
 * are applied after the sums and the partial sums are added in double),
 * not bit for bit.
 * ------------------------------------------------------------------- */
typedef struct tlm_far_field_summary {
    float u_max;          /* FarFieldData::uMax */
    float directivity;    /* 4 pi uMax / trapezoid integral of U sin(theta) */
    float peak_theta_deg; /* FarFieldData::peakThetaDeg */
    float peak_phi_deg;   /* FarFieldData::peakPhiDeg */
} tlm_far_field_summary;

/* From host arrays, as the reference's function takes them: per face the
 * HuyFace pos (3 floats), axis and outSign, and the 8 DFT floats of huyDft
 * (E1, E2, eta0 H1, eta0 H2 as re, im). U may be NULL. TLM_ERR_STATE when
 * nfaces or samples is 0 (the reference returns false). */
int tlm_far_field(int device, int nfaces, const float *pos, const int32_t *axis,
                  const int32_t *out_sign, const float *dft8, int samples, float f0, int n_theta,
                  int n_phi, float *U, tlm_far_field_summary *out);
/* From a context's own Huygens faces and accumulators (no readback of the
 * DFTs): positions cellCenter - domain centre for the given grid origin
 * (TlmSolver.cpp:323-334), outward signs from the side of the grid the face
 * is on. samples = the DFT sample count (0: no data). */
int tlm_far_field_ctx(tlm_ctx *ctx, const float origin[3], int samples, float f0, int n_theta,
                      int n_phi, float *U, tlm_far_field_summary *out);

/* tlm_far_field_ctx without device allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * transform's workspace (64-byte face records, direction tables, partial sums,
 * U) comes from a fixed 128 MiB device array that the library holds on each
 * GPU it runs on until the process ends: up to about 2 million Huygens faces
 * on the default 37 x 73 grid. Larger inputs are refused with
 * TLM_ERR_UNSUPPORTED. Calls on one GPU share that array: calls that overlap
 * on one GPU (from different threads or streams) race. Same results as
 * tlm_far_field_ctx, bit for bit. */
int tlm_far_field_ctx_noalloc(tlm_ctx *ctx, const float origin[3], int samples, float f0,
                              int n_theta, int n_phi, float *U, tlm_far_field_summary *out);

#ifdef __cplusplus
}
#endif
#endif
