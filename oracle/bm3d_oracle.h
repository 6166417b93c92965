/*
 * CPU oracle for BM3Dv2 -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's BM3Dv2 estimation shader
 * (src/bm3d.comp), its aggregation shader (src/bm3d_agg.comp) and the host
 * window semantics of src/bm3d.cpp (ref-block grid, temporal window clamping,
 * agg_z slice table).  It exists so the tests, the smoke check and the CPU
 * baseline leg of the benchmark have something independent to compare the
 * HIP path against.  Nothing in the product links or calls it.
 *
 * Two precisions are compiled from one body (bm3d_oracle_body.h):
 *   bm3d_oracle_f32_*  -- float arithmetic, the "plain form" parity target
 *   bm3d_oracle_f64_*  -- the same algorithm in double (precision report)
 *
 * Floating-point contraction is explicit (the file is compiled with
 * -ffp-contract=off) and selected per call by fp_mode:
 *   0  every GLSL fma() is a separately rounded multiply and add
 *   1  GLSL fma() is fused, every other a*b+c is rounded twice (default)
 *   2  GLSL fma() is fused and the contractible a*b+c patterns of the shader
 *      (the column SSD accumulation, the extractor, the Wiener denominator
 *      and the Wiener weight accumulation) are fused too
 */
#ifndef VSFEEL_BM3D_ORACLE_H
#define VSFEEL_BM3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bm3d_oracle_params {
    int32_t width;        /* luma plane width  (>= 8) */
    int32_t height;       /* luma plane height (>= 8) */
    int32_t pitch;        /* elements between rows of every plane given */
    int32_t block_step;   /* [1, 8] */
    int32_t bm_range;     /* > 0 */
    int32_t radius;       /* [0, 4] */
    int32_t ps_num;       /* [1, 8] */
    int32_t ps_range;     /* > 0 */
    float sigma;          /* scaled threshold sigma (SIGMA_Y of the shader) */
    float extractor;      /* 0 or 2^extractor_exp */
    int32_t final_pass;   /* 1: Wiener pass (search/reference clip + input clip) */
    int32_t fp_mode;      /* 0, 1 or 2 (see above) */
    int32_t nthreads;     /* worker threads for the ref-block loop (>= 1) */
} bm3d_oracle_params;

/* Number of active reference blocks of one plane, and the grid behind it. */
int32_t bm3d_oracle_ref_grid(const bm3d_oracle_params *p, int32_t *nx, int32_t *ny);

/*
 * One estimation dispatch (bm3d.comp main over the whole grid) for the frame
 * at window centre z = radius.  search[z] / input[z] for z in [0, 2r] are the
 * already clamped window planes (input == search outside the final pass; the
 * caller may pass input = NULL then).  res receives (2r+1) * 2 planes of
 * height * pitch values each, laid out [z][wdst, weight]; it is accumulated
 * into, not cleared.  groups (optional) receives for every active ref block,
 * in row-major (by, bx) order, 8 entries of (x, y, z).
 */
int32_t bm3d_oracle_f32_estimate(const bm3d_oracle_params *p,
                                 const float *const *search,
                                 const float *const *input,
                                 float *res, int32_t *groups);
int32_t bm3d_oracle_f64_estimate(const bm3d_oracle_params *p,
                                 const float *const *search,
                                 const float *const *input,
                                 double *res, int32_t *groups);

/*
 * Whole-clip driver: for every output frame n in [first, first+count) it
 * aggregates the estimate stacks of frames clamp(n-r+i) at slices
 * agg_z(i, n, nframes, r) exactly like the reference host code.  clip (and ref
 * in the final pass, which is the clip block matching runs on) holds nframes
 * planes of height*pitch floats; out receives count planes.
 */
int32_t bm3d_oracle_f32_clip(const bm3d_oracle_params *p, int32_t nframes,
                             const float *clip, const float *ref,
                             int32_t first, int32_t count, float *out);
int32_t bm3d_oracle_f64_clip(const bm3d_oracle_params *p, int32_t nframes,
                             const float *clip, const float *ref,
                             int32_t first, int32_t count, double *out);

/* Pieces exposed for pinning the restatement against closed forms. */
void bm3d_oracle_f32_dct8(float *v, int32_t inverse, int32_t fp_mode);
void bm3d_oracle_f64_dct8(double *v, int32_t inverse, int32_t fp_mode);
/* 8x8x8 group transform: patch[m][row][col] in, denoised patch out, returns
 * the adaptive weight; ref_patch != NULL selects the Wiener form. */
float bm3d_oracle_f32_group(float *patch, const float *ref_patch, float sigma,
                            int32_t fp_mode, int32_t *kept);
double bm3d_oracle_f64_group(double *patch, const double *ref_patch, double sigma,
                             int32_t fp_mode, int32_t *kept);
/* Distances of the candidates of one search window, in scan order
 * (row-major, seq = (cy-top)*rw + (cx-left)), for the sliding (radius < 3)
 * or the exact (radius >= 3) metric.  Returns the number of candidates. */
int32_t bm3d_oracle_f32_window_dist(const float *plane, int32_t pitch,
                                    int32_t width, int32_t height,
                                    int32_t x, int32_t y, int32_t range,
                                    int32_t cx0, int32_t cy0,
                                    int32_t exact, int32_t fp_mode, float *dist);
int32_t bm3d_oracle_agg_z(int32_t i, int32_t n, int32_t nframes, int32_t radius);

#ifdef __cplusplus
}
#endif

#endif
