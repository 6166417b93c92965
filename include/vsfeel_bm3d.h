/*
 * vsfeel_bm3d.h -- C ABI of the MI355X (gfx950) BM3Dv2 denoiser.
 *
 * Drop-in boundary for the BM3Dv2 filter of TheFeelTrain/vapoursynth-feel.
 * The reference registers BM3Dv2 in src/vsfeel.cpp:350-364 ->
 * src/bm3d.cpp:1558-1574 and implements it as BM3DCreate (src/bm3d.cpp:1043),
 * BM3DGetFrame (:698) and BM3DFree (:1030) over the Vulkan shaders
 * src/bm3d.comp and src/bm3d_agg.comp.  Each entry point below names the
 * reference interface it replaces.  Plain C types only: no torch, no HIP
 * types (streams are passed as void *, a hipStream_t or NULL for the default
 * stream).  Every call returns VSFEEL_BM3D_OK (0) or an error code, and writes
 * a message in the reference's "BM3D: ..." style into err (when err_len > 0).
 *
 * Thread safety: one handle may be used from several threads; calls on one
 * handle are serialised internally.  Device-side calls on one handle are
 * ordered across the streams they are issued on.
 */
#ifndef VSFEEL_BM3D_H
#define VSFEEL_BM3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSFEEL_BM3D_OK 0
#define VSFEEL_BM3D_EINVAL 1   /* rejected argument (BM3DCreate's set_error) */
#define VSFEEL_BM3D_EDEVICE 2  /* HIP runtime error */
#define VSFEEL_BM3D_ERANGE 3   /* frame index / size outside the clip */

/* VapourSynth-style format of a clip: the fields BM3DCreate inspects
 * (src/bm3d.cpp:1067-1088, :1257-1275). */
typedef struct vsfeel_bm3d_format {
    int32_t color_family;    /* 1 Gray, 2 RGB, 3 YUV (VapourSynth cfGray/cfRGB/cfYUV) */
    int32_t sample_type;     /* 0 integer, 1 float (stInteger/stFloat) */
    int32_t bits_per_sample;
    int32_t sub_sampling_w;
    int32_t sub_sampling_h;
    int32_t width;           /* luma plane */
    int32_t height;
    int32_t num_frames;
} vsfeel_bm3d_format;

/*
 * Arguments of core.vsfeel.BM3Dv2, with the registration string of
 * src/bm3d.cpp:1561-1571:
 *   clip:vnode; ref:vnode:opt; sigma:float[]:opt; block_step:int[]:opt;
 *   bm_range:int[]:opt; radius:int:opt; ps_num:int[]:opt; ps_range:int[]:opt;
 *   num_streams:int:opt; extractor_exp:int:opt;
 * Each optional argument is (pointer, count); count 0 means "not given".
 * Array arguments are read at indices 0..2 like mapGetFloat/mapGetInt
 * (later planes inherit the previous value); only plane 0's value is used.
 */
typedef struct vsfeel_bm3d_args {
    vsfeel_bm3d_format clip;
    int32_t has_ref;
    vsfeel_bm3d_format ref;
    const double *sigma;          int32_t num_sigma;
    const int64_t *block_step;    int32_t num_block_step;
    const int64_t *bm_range;      int32_t num_bm_range;
    const int64_t *radius;        int32_t num_radius;
    const int64_t *ps_num;        int32_t num_ps_num;
    const int64_t *ps_range;      int32_t num_ps_range;
    const int64_t *num_streams;   int32_t num_num_streams;
    const int64_t *extractor_exp; int32_t num_extractor_exp;
    /* Not in the reference's signature: rounding of the shaders' multiply-adds
     * (0 all separate, 1 GLSL fma() fused [default], 2 every a*b+c fused) and
     * the HIP device ordinal (the reference always uses device 0). */
    int32_t fp_mode;
    int32_t device;
    /* Not in the reference either: the arithmetic of the block-matching
     * distances.  0: the reference's FP32 sums (plain form, default).
     * 1 / 2: precision form (radius 0-2), distances of the pixels rounded
     * once to FP16 / BF16 as norms - 2 dot on MFMA with FP32 accumulation,
     * a shortlist of twice the entries a search keeps rescored with the
     * FP32 sums: the groups equal the plain form's wherever its matches lie
     * in the shortlist. */
    int32_t precision;
} vsfeel_bm3d_args;

/* Derived parameters of an instance (what BM3DCreate stores in BM3DData,
 * src/bm3d.cpp:74-130). */
typedef struct vsfeel_bm3d_info {
    int32_t width, height, num_frames;
    int32_t chroma;           /* YUV input: chroma passes through */
    int32_t final_pass;       /* a "ref" clip was given (Wiener pass) */
    float sigma;              /* plane-0 sigma after the reference's scaling */
    float extractor;          /* 0 or 2^extractor_exp */
    int32_t block_step, bm_range, radius, ps_num, ps_range, num_streams;
    int32_t temporal_width;   /* 2 * radius + 1 */
    int32_t ref_blocks_x, ref_blocks_y;  /* active reference-block grid */
    int32_t fp_mode;
    int32_t device;
    int32_t pitch;            /* device plane pitch (elements) of the internal buffers */
    int32_t precision;        /* vsfeel_bm3d_args.precision */
} vsfeel_bm3d_info;

typedef struct vsfeel_bm3d vsfeel_bm3d;

/*
 * Replaces BM3DCreate (src/bm3d.cpp:1043-1552): validates the arguments in
 * the reference's order with its messages ("BM3D: \"radius\" must be in range
 * [0, 4]", ...), scales sigma, sizes the frame caches and allocates the
 * device buffers.  *out is NULL on error.
 */
int vsfeel_bm3d_create(const vsfeel_bm3d_args *args, vsfeel_bm3d **out,
                       char *err, size_t err_len);

/* Replaces BM3DFree (src/bm3d.cpp:1030-1037): waits for the device and frees. */
void vsfeel_bm3d_free(vsfeel_bm3d *h);

int vsfeel_bm3d_get_info(const vsfeel_bm3d *h, vsfeel_bm3d_info *info);

/*
 * The frames output frame n needs: [clamp(n-2r), clamp(n+2r)], the request
 * window of BM3DGetFrame's arInitial branch (src/bm3d.cpp:704-712).
 */
int vsfeel_bm3d_frame_window(const vsfeel_bm3d *h, int32_t n, int32_t *first, int32_t *last);

/*
 * Replaces BM3DGetFrame's arAllFramesReady branch (src/bm3d.cpp:713-1025)
 * for the luma plane.  clip_planes[i] / ref_planes[i] are host pointers to
 * the luma planes of frames first+i of vsfeel_bm3d_frame_window(n), with row
 * strides in bytes (ref_* only when the instance has a ref clip, else NULL).
 * Frames whose data is already cached on the device are not read.  The
 * denoised luma of frame n is written to dst (host, stride in bytes).
 * Chroma pass-through and frame allocation stay with the caller (the VS
 * shim copies planes 1 and 2 like newVideoFrame2, src/bm3d.cpp:715-726).
 */
int vsfeel_bm3d_get_frame(vsfeel_bm3d *h, int32_t n,
                          const float *const *clip_planes, const ptrdiff_t *clip_strides,
                          const float *const *ref_planes, const ptrdiff_t *ref_strides,
                          float *dst, ptrdiff_t dst_stride,
                          char *err, size_t err_len);

/* A clip resident in device memory: num_frames luma planes. */
typedef struct vsfeel_bm3d_device_clip {
    const float *data;        /* frame f starts at data + f * frame_stride */
    int64_t frame_stride;     /* elements */
    int32_t pitch;            /* elements between rows (>= width) */
    const float *ref;         /* ref clip (final pass) with the same layout, or NULL */
} vsfeel_bm3d_device_clip;

/*
 * Device-resident counterpart of BM3DGetFrame for a batch of output frames
 * [first, first + count): the whole per-frame launch sequence (slot clear,
 * block matching, collaborative filtering, aggregation) is enqueued on
 * `stream`; estimate stacks are reused across the batch the way the
 * reference's res cache reuses them across frames (src/bm3d.cpp:380-456).
 * out + (n - first) * out_frame_stride receives frame n.  Returns after
 * enqueueing; synchronise the stream before reading out.
 */
int vsfeel_bm3d_process_device(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip,
                               int32_t first, int32_t count,
                               float *out, int32_t out_pitch, int64_t out_frame_stride,
                               void *stream, char *err, size_t err_len);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vsfeel_bm3d_process_device without a handle, for the
 * arguments of vsfeel_bm3d_create (radius 0, no ref clip).  It allocates
 * nothing: its group table, tile tables and tile accumulators live in a
 * static device array of 64 MiB that each GPU holds until the process ends
 * (enough for 1920x1080 at block_step 4, bm_range 16); larger frames are
 * rejected with VSFEEL_BM3D_EINVAL.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (concurrent streams or threads)
 * race; calls ordered on one stream do not.  Frames [first, first + count)
 * of the clip are denoised into out like vsfeel_bm3d_process_device, with
 * the same kernels and the same results.  Returns after enqueueing on stream.
 */
int vsfeel_bm3d_process_device_static(const vsfeel_bm3d_args *args, const vsfeel_bm3d_device_clip *clip,
                                      int32_t first, int32_t count,
                                      float *out, int32_t out_pitch, int64_t out_frame_stride,
                                      void *stream, char *err, size_t err_len);

/*
 * Debug: the groups one estimation dispatch of frame m forms (no
 * counterpart buffer in the reference, whose shader keeps them in shared
 * memory, src/bm3d.comp:657-661).  groups receives ref_blocks_y *
 * ref_blocks_x * 8 (x, y, z) int32 triplets in raster order of the ref
 * blocks, z being the window position (radius = the frame itself); dists
 * (optional) the 8 distances (-1 for a block forced in by the self-block rule).
 * Host buffers; synchronous.
 */
int vsfeel_bm3d_debug_groups(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip, int32_t m,
                             int32_t *groups, float *dists, char *err, size_t err_len);

/*
 * Debug: the estimate stack of frame m (one dispatch of the estimation
 * shader, src/bm3d.comp:445-718): (2r+1) x {weighted sum, weight} planes of
 * height x res_pitch floats, written to the device buffer res.  Synchronous.
 */
int vsfeel_bm3d_debug_estimate(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip, int32_t m,
                               float *res, int32_t res_pitch, char *err, size_t err_len);

/*
 * Differentiable form (no counterpart in the reference; built on the same
 * kernels).  With the block-match groups and the threshold masks held fixed,
 * the basic (hard-threshold) filter at radius 0 is out = A x / D with A a
 * fixed symmetric linear map (per group: gather, 3-D DCT, mask, inverse,
 * weight, scatter) and D the weight sums, so d(loss)/dx = A (g / D).
 *
 * Saved state per output frame i of a batch (device buffers, laid out
 * back to back for the count frames):
 *   groups  [count][ref_blocks_y * ref_blocks_x][8] uint32 (packed x | y<<14 | z<<28)
 *   masks   [count][ref blocks][8] uint64 keep masks (lane u, bit w*8+v)
 *   weights [count][ref blocks] float
 *   denom   [count][height][denom_pitch] float weight sums D
 */
typedef struct vsfeel_bm3d_saved {
    uint32_t *groups;
    uint64_t *masks;
    float *weights;
    float *denom;
    int32_t denom_pitch;
} vsfeel_bm3d_saved;

/* Forward pass that also fills `saved` (radius 0, basic pass only). */
int vsfeel_bm3d_forward_saved(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip,
                              int32_t first, int32_t count,
                              float *out, int32_t out_pitch, int64_t out_frame_stride,
                              const vsfeel_bm3d_saved *saved, void *stream,
                              char *err, size_t err_len);

/* Backward pass: grad_in = A (grad_out / D) per frame, from `saved`. */
int vsfeel_bm3d_backward(vsfeel_bm3d *h, const float *grad_out, int32_t grad_pitch,
                         int64_t grad_frame_stride, int32_t count,
                         const vsfeel_bm3d_saved *saved,
                         float *grad_in, int32_t gin_pitch, int64_t gin_frame_stride,
                         void *stream, char *err, size_t err_len);

/*
 * Per-kernel GPU timing, the counterpart of the reference's BM3D_GPUTRACE
 * timestamp queries (src/bm3d.cpp:1492-1501, :978-994): when enabled, HIP
 * events are recorded around every launch on the launch stream.
 * vsfeel_bm3d_set_timing resets the counters; vsfeel_bm3d_kernel_times waits
 * for the recorded launches and returns the summed milliseconds and launch
 * counts per kernel (arrays of 9): [0] slot clear, [1] block matching (at
 * radius 1-2 its spatial search), [2] collaborative filtering + scatter,
 * [3] aggregation, [4] tile reduction, [5] weight box sums of the other
 * frames' slices (radius > 0), [6] sums of the binned other-frame blocks
 * (radius > 0), [7] plane sets of the precision form, [8] the temporal
 * predictive search after the spatial one (radius 1-2).
 */
int vsfeel_bm3d_set_timing(vsfeel_bm3d *h, int32_t enable);
int vsfeel_bm3d_kernel_times(vsfeel_bm3d *h, double *ms, int64_t *count);

/*
 * Graph launches; `enable` is a bit set (default 1 at radius 0, 0 at
 * radius > 0 where process_device runs two lanes).  Bit 0:
 * vsfeel_bm3d_process_device captures its launch sequence (per output frame:
 * slot clears, matching, filtering, the other frames' sums and weights at
 * radius > 0, tile reduction, aggregation) into a HIP graph on the
 * instance's stream, and later calls with identical arguments replay it,
 * ordered after and before the caller's stream by events.  (At radius > 0
 * the replayed graph ran the two lanes' cross-linked branches with less
 * overlap than direct launches: 980 against 1024 frames/s at 1080p radius
 * 2, hence the default there.)  Bit 1:
 * vsfeel_bm3d_get_frame launches a request's kernels as one graph per
 * output slot and recompute pattern (the number of estimates the request
 * runs), captured on an idle stream and updated in place
 * (hipGraphExecUpdate) with the request's ring slots and frame; off by
 * default: measured no faster than the request's direct launches, whose
 * host cost is small next to its copies, while capture and update run
 * under the handle's mutex.  The counterpart of the
 * reference's per-frame command-buffer recording (src/bm3d.cpp:486-623,
 * :899-967).  Timing (vsfeel_bm3d_set_timing) launches directly instead.
 * Values outside 0-3: VSFEEL_BM3D_EINVAL.
 */
int vsfeel_bm3d_set_graphs(vsfeel_bm3d *h, int32_t enable);

/*
 * Test hook, with no counterpart in the reference: the next
 * vsfeel_bm3d_get_frame request that reaches the stage fails there with
 * VSFEEL_BM3D_EDEVICE and a "BM3D: ..." message, as a failed upload or
 * submission fails a getFrame (src/bm3d.cpp:747-764), without touching the
 * device: stage 1 after the request has staged its frames, before any
 * upload; stage 2 after its uploads are enqueued, before its kernels; 0
 * disarms.  Concurrent requests and the next one on the handle complete.
 */
int vsfeel_bm3d_fail_next_request(vsfeel_bm3d *h, int32_t stage);

#ifdef __cplusplus
}
#endif

#endif /* VSFEEL_BM3D_H */
