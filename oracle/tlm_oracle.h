/*
 * tlm_oracle.h - C API of the CPU restatement of the reference TLM step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 */
#ifndef TLM_ORACLE_H
#define TLM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* WaveformType, TlmSolver.h:70, plus the Gaussian-derivative pulse */
enum { OT_WF_CW_RAMPED = 0, OT_WF_GAUSS_SINE = 1, OT_WF_GAUSS_PULSE = 2, OT_WF_GAUSS_DERIV = 3 };

typedef struct ot_config {
    int32_t nx, ny, nz;
    float dl, ox, oy, oz;        /* VoxelGridSpec, Geometry.h:85-113 */
    float f0;                    /* TlmConfig, TlmSolver.h:113-123 */
    int32_t waveform;
    int32_t total_steps, settle_steps;
    float rho;
    int32_t huy_offset, record_every;
    float gd_width, gd_delay;    /* Gaussian-derivative width / delay, in steps */
    int32_t threads;             /* <= 0: min(16, hw-1) like TlmSolver.cpp:247-249 */
    int32_t fp64;                /* 0: float (parity oracle), 1: double (precision oracle) */
    int32_t pulse_format;        /* float only: 0 FP32, 1 FP16, 2 BF16 pulse storage with
                                    per-slab power-of-two scales (no reference counterpart) */
    int32_t fp32_halo;           /* 16-bit formats: z-slabs within this distance of a source
                                    slab keep FP32 storage; < 0: none */
} ot_config;

typedef struct ot_handle ot_handle;

int ot_default_threads(void);
float ot_waveform_value(int type, float f0, float dt, int n, float gd_width, float gd_delay);

ot_handle *ot_create(const ot_config *cfg, const uint8_t *mat, const float *epsr,
                     const float *sigma, int nmat);
void ot_destroy(ot_handle *h);

int ot_add_port(ot_handle *h, const int64_t *cells, int n, int pol, float amp);
void ot_add_plane_wave(ot_handle *h, int prop_axis, int plane_index, int pol_axis, float amp);
void ot_add_source(ot_handle *h, int64_t cell, int pol, float amp);
int ot_add_probe(ot_handle *h, float x, float y, float z);
int ot_add_probe_cell(ot_handle *h, int64_t cell);
void ot_set_probe_dft(ot_handle *h, const double *freqs, int nf);
void ot_set_field_plane(ot_handle *h, int axis, int index);
void ot_set_state(ot_handle *h, const double *V, const double *Vs);

void ot_run(ot_handle *h);
void ot_run_steps(ot_handle *h, int count);
void ot_scatter_connect(ot_handle *h);
void ot_finalize(ot_handle *h);

int ot_slab_exponents(ot_handle *h, int32_t *E, float *M);
uint16_t ot_f16_rne(float f);
float ot_f16_to_f32(uint16_t h);
uint16_t ot_bf16_rne(float f);
int ot_plan_exponent(float bound);
void ot_encode(int fmt, const float *in, float inv_scale, uint16_t *out, int64_t n);
int ot_cur_step(ot_handle *h);
int ot_has_stubs(ot_handle *h);
float ot_dt(ot_handle *h);
void ot_get_state(ot_handle *h, double *V, double *Vs);
int ot_num_faces(ot_handle *h);
void ot_get_faces(ot_handle *h, int64_t *air_cell, int32_t *axis, int32_t *sign, float *center);
int ot_num_huygens(ot_handle *h);
void ot_get_huygens_faces(ot_handle *h, int64_t *cell, int32_t *axis, int32_t *out_sign);
void ot_get_js(ot_handle *h, double *inst, double *dft_re, double *dft_im, double *mag,
               int *samples);
void ot_get_huygens(ot_handle *h, double *dft, int *samples);
int ot_num_ports(ot_handle *h);
int ot_port_cells(ot_handle *h, int p, int64_t *cells);
int ot_port_records(ot_handle *h, int p, double *v, double *i);
int ot_num_sources(ot_handle *h);
void ot_get_sources(ot_handle *h, int64_t *cell, int32_t *pol, float *amp);
int ot_probe_cell(ot_handle *h, int p);
int ot_probe_series(ot_handle *h, int p, double *e);
void ot_probe_dft(ot_handle *h, int p, double *re, double *im);
int ot_energy(ot_handle *h, int32_t *steps, float *vals, double *dvals);
int ot_plane(ot_handle *h, double *buf, int *n1, int *n2);
int ot_num_frames(ot_handle *h);
int ot_frame(ot_handle *h, int idx, double *js, double *plane);

void ot_scatter_cell(const double *vin, double *vs, double yh, double gh, int diel, int fp64,
                     double *vout);
float ot_scatter_unitarity_error(void);
float ot_diel_scatter_unitarity_error(float epsr);

#ifdef __cplusplus
}
#endif
#endif
