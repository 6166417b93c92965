// Test-only stand-in for libvsfeel_bm3d.so: the four entry points the
// VapourSynth shim calls, recording their arguments into g_stub.  get_frame
// writes n + 0.25 into the destination rows (width floats each).
#include <algorithm>
#include <cstdio>

#include "stub_abi.h"

StubLog g_stub;

struct vsfeel_bm3d {
    vsfeel_bm3d_format clip;
    int radius;
};

namespace {

template <class T, class S>
std::vector<T> take(const S *p, int n)
{
    return p ? std::vector<T>(p, p + n) : std::vector<T>();
}

void put_err(char *err, size_t len, const char *msg)
{
    if (err && len) {
        std::snprintf(err, len, "%s", msg);
    }
}

}  // namespace

extern "C" {

int vsfeel_bm3d_create(const vsfeel_bm3d_args *a, vsfeel_bm3d **out, char *err, size_t err_len)
{
    ++g_stub.creates;
    g_stub.clip = a->clip;
    g_stub.ref = a->ref;
    g_stub.has_ref = a->has_ref;
    g_stub.fp_mode = a->fp_mode;
    g_stub.device = a->device;
    g_stub.precision = a->precision;
    g_stub.sigma = take<double>(a->sigma, a->num_sigma);
    g_stub.block_step = take<long long>(a->block_step, a->num_block_step);
    g_stub.bm_range = take<long long>(a->bm_range, a->num_bm_range);
    g_stub.radius = take<long long>(a->radius, a->num_radius);
    g_stub.ps_num = take<long long>(a->ps_num, a->num_ps_num);
    g_stub.ps_range = take<long long>(a->ps_range, a->num_ps_range);
    g_stub.num_streams = take<long long>(a->num_streams, a->num_num_streams);
    g_stub.extractor_exp = take<long long>(a->extractor_exp, a->num_extractor_exp);
    if (g_stub.fail_create) {
        put_err(err, err_len, "BM3D: stub create failure");
        return VSFEEL_BM3D_EINVAL;
    }
    *out = new vsfeel_bm3d{ a->clip, a->num_radius > 0 ? int(a->radius[0]) : 0 };
    ++g_stub.live_handles;
    return VSFEEL_BM3D_OK;
}

void vsfeel_bm3d_free(vsfeel_bm3d *h)
{
    ++g_stub.frees;
    if (h) {
        --g_stub.live_handles;
        delete h;
    }
}

int vsfeel_bm3d_frame_window(const vsfeel_bm3d *h, int32_t n, int32_t *first, int32_t *last)
{
    if (n < 0 || n >= h->clip.num_frames) {
        return VSFEEL_BM3D_ERANGE;
    }
    *first = std::clamp(n - 2 * h->radius, 0, h->clip.num_frames - 1);
    *last = std::clamp(n + 2 * h->radius, 0, h->clip.num_frames - 1);
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_get_frame(vsfeel_bm3d *h, int32_t n, const float *const *clip_planes,
                          const ptrdiff_t *clip_strides, const float *const *ref_planes,
                          const ptrdiff_t *ref_strides, float *dst, ptrdiff_t dst_stride, char *err,
                          size_t err_len)
{
    int32_t first = 0, last = 0;
    vsfeel_bm3d_frame_window(h, n, &first, &last);
    const int count = last - first + 1;
    g_stub.n = n;
    g_stub.planes.assign(clip_planes, clip_planes + count);
    g_stub.strides.assign(clip_strides, clip_strides + count);
    g_stub.ref_planes = ref_planes ? std::vector<const float *>(ref_planes, ref_planes + count)
                                   : std::vector<const float *>();
    g_stub.ref_strides = ref_strides ? std::vector<ptrdiff_t>(ref_strides, ref_strides + count)
                                     : std::vector<ptrdiff_t>();
    g_stub.dst = dst;
    g_stub.dst_stride = dst_stride;
    if (g_stub.fail_get) {
        put_err(err, err_len, "BM3D: stub launch failure");
        return VSFEEL_BM3D_EDEVICE;
    }
    for (int y = 0; y < h->clip.height; ++y) {
        float *row = reinterpret_cast<float *>(reinterpret_cast<char *>(dst) + y * dst_stride);
        std::fill(row, row + h->clip.width, float(n) + 0.25f);
    }
    return VSFEEL_BM3D_OK;
}

}  // extern "C"
