// VapourSynth plugin shim: registers core.vsfeel.BM3Dv2 with the reference's
// identifier, namespace and argument string (src/vsfeel.cpp:350-364,
// src/bm3d.cpp:1558-1574) and forwards BM3DCreate / BM3DGetFrame / BM3DFree
// to the C ABI of include/vsfeel_bm3d.h.  Built only where VapourSynth4.h is
// available (make VS_INCLUDE=/path/to/vapoursynth/include plugin).

#include <VapourSynth4.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "vsfeel_bm3d.h"

namespace {

struct Bm3dNode {
    VSNode *node = nullptr;
    VSNode *ref_node = nullptr;
    const VSVideoInfo *vi = nullptr;
    vsfeel_bm3d *handle = nullptr;
};

vsfeel_bm3d_format to_format(const VSVideoInfo *vi)
{
    vsfeel_bm3d_format f {};
    f.color_family = vi->format.colorFamily;
    f.sample_type = vi->format.sampleType;
    f.bits_per_sample = vi->format.bitsPerSample;
    f.sub_sampling_w = vi->format.subSamplingW;
    f.sub_sampling_h = vi->format.subSamplingH;
    f.width = vi->width;
    f.height = vi->height;
    f.num_frames = vi->numFrames;
    return f;
}

const VSFrame *VS_CC bm3d_get_frame(int n, int activation_reason, void *instance_data, void **frame_data,
                                    VSFrameContext *frame_ctx, VSCore *core, const VSAPI *vsapi)
{
    (void)frame_data;
    auto *d = static_cast<Bm3dNode *>(instance_data);
    int32_t first = 0, last = 0;
    if (vsfeel_bm3d_frame_window(d->handle, n, &first, &last) != VSFEEL_BM3D_OK) {
        vsapi->setFilterError("BM3D: frame out of range", frame_ctx);
        return nullptr;
    }
    if (activation_reason == arInitial) {
        // the window of src/bm3d.cpp:704-712
        for (int f = first; f <= last; ++f) {
            vsapi->requestFrameFilter(f, d->node, frame_ctx);
            if (d->ref_node) {
                vsapi->requestFrameFilter(f, d->ref_node, frame_ctx);
            }
        }
        return nullptr;
    }
    if (activation_reason != arAllFramesReady) {
        return nullptr;
    }
    const int count = last - first + 1;
    std::vector<const VSFrame *> frames(count), refs(count, nullptr);
    std::vector<const float *> planes(count), ref_planes(count, nullptr);
    std::vector<ptrdiff_t> strides(count), ref_strides(count, 0);
    for (int i = 0; i < count; ++i) {
        frames[i] = vsapi->getFrameFilter(first + i, d->node, frame_ctx);
        planes[i] = reinterpret_cast<const float *>(vsapi->getReadPtr(frames[i], 0));
        strides[i] = vsapi->getStride(frames[i], 0);
        if (d->ref_node) {
            refs[i] = vsapi->getFrameFilter(first + i, d->ref_node, frame_ctx);
            ref_planes[i] = reinterpret_cast<const float *>(vsapi->getReadPtr(refs[i], 0));
            ref_strides[i] = vsapi->getStride(refs[i], 0);
        }
    }
    const VSFrame *src = frames[n - first];
    VSFrame *dst;
    if (d->vi->format.colorFamily == cfYUV) {
        // chroma planes and frame props copied from the source (src/bm3d.cpp:715-726)
        const int pl[] = { 0, 1, 2 };
        const VSFrame *fr[] = { nullptr, src, src };
        dst = vsapi->newVideoFrame2(&d->vi->format, d->vi->width, d->vi->height, fr, pl, src, core);
    } else {
        dst = vsapi->newVideoFrame(&d->vi->format, d->vi->width, d->vi->height, nullptr, core);
    }
    char err[512];
    const int rc = vsfeel_bm3d_get_frame(d->handle, n, planes.data(), strides.data(),
                                         d->ref_node ? ref_planes.data() : nullptr,
                                         d->ref_node ? ref_strides.data() : nullptr,
                                         reinterpret_cast<float *>(vsapi->getWritePtr(dst, 0)),
                                         vsapi->getStride(dst, 0), err, sizeof(err));
    for (int i = 0; i < count; ++i) {
        vsapi->freeFrame(frames[i]);
        if (refs[i]) {
            vsapi->freeFrame(refs[i]);
        }
    }
    if (rc != VSFEEL_BM3D_OK) {
        vsapi->setFilterError(err, frame_ctx);
        vsapi->freeFrame(dst);
        return nullptr;
    }
    return dst;
}

void VS_CC bm3d_free(void *instance_data, VSCore *core, const VSAPI *vsapi)
{
    (void)core;
    auto *d = static_cast<Bm3dNode *>(instance_data);
    vsfeel_bm3d_free(d->handle);
    vsapi->freeNode(d->node);
    if (d->ref_node) {
        vsapi->freeNode(d->ref_node);
    }
    delete d;
}

// the (pointer, count) pair of an optional array argument
template <class T>
void array_arg(const VSMap *in, const VSAPI *vsapi, const char *key, std::vector<T> &store, const T **ptr,
               int32_t *num)
{
    const int n = vsapi->mapNumElements(in, key);
    store.clear();
    for (int i = 0; i < n; ++i) {
        int e = 0;
        if constexpr (std::is_same_v<T, double>) {
            store.push_back(vsapi->mapGetFloat(in, key, i, &e));
        } else {
            store.push_back(vsapi->mapGetInt(in, key, i, &e));
        }
    }
    *ptr = store.empty() ? nullptr : store.data();
    *num = int32_t(store.size());
}

void VS_CC bm3d_create(const VSMap *in, VSMap *out, void *user_data, VSCore *core, const VSAPI *vsapi)
{
    (void)user_data;
    auto *d = new Bm3dNode();
    d->node = vsapi->mapGetNode(in, "clip", 0, nullptr);
    d->vi = vsapi->getVideoInfo(d->node);
    int e = 0;
    d->ref_node = vsapi->mapGetNode(in, "ref", 0, &e);
    if (e) {
        d->ref_node = nullptr;
    }

    vsfeel_bm3d_args a {};
    a.clip = to_format(d->vi);
    if (d->ref_node) {
        a.has_ref = 1;
        a.ref = to_format(vsapi->getVideoInfo(d->ref_node));
    }
    std::vector<double> sigma;
    std::vector<int64_t> block_step, bm_range, radius, ps_num, ps_range, num_streams, extractor_exp;
    array_arg(in, vsapi, "sigma", sigma, &a.sigma, &a.num_sigma);
    array_arg(in, vsapi, "block_step", block_step, &a.block_step, &a.num_block_step);
    array_arg(in, vsapi, "bm_range", bm_range, &a.bm_range, &a.num_bm_range);
    array_arg(in, vsapi, "radius", radius, &a.radius, &a.num_radius);
    array_arg(in, vsapi, "ps_num", ps_num, &a.ps_num, &a.num_ps_num);
    array_arg(in, vsapi, "ps_range", ps_range, &a.ps_range, &a.num_ps_range);
    array_arg(in, vsapi, "num_streams", num_streams, &a.num_streams, &a.num_num_streams);
    array_arg(in, vsapi, "extractor_exp", extractor_exp, &a.extractor_exp, &a.num_extractor_exp);
    a.fp_mode = 1;
    a.device = 0;
    a.precision = 0;  // the plain FP32 form: the reference's distances

    char err[512];
    if (vsfeel_bm3d_create(&a, &d->handle, err, sizeof(err)) != VSFEEL_BM3D_OK) {
        vsapi->mapSetError(out, err);
        vsapi->freeNode(d->node);
        if (d->ref_node) {
            vsapi->freeNode(d->ref_node);
        }
        delete d;
        return;
    }
    VSFilterDependency deps[2] = { { d->node, rpStrictSpatial }, { d->ref_node, rpStrictSpatial } };
    vsapi->createVideoFilter(out, "BM3D", d->vi, bm3d_get_frame, bm3d_free, fmParallel, deps,
                             d->ref_node ? 2 : 1, d, core);
}

}  // namespace

VS_EXTERNAL_API(void) VapourSynthPluginInit2(VSPlugin *plugin, const VSPLUGINAPI *vspapi)
{
    vspapi->configPlugin("com.thefeeltrain.vsfeel", "vsfeel", "MI355X BM3Dv2", VS_MAKE_VERSION(1, 0),
                         VAPOURSYNTH_API_VERSION, 0, plugin);
    vspapi->registerFunction("BM3Dv2",
                             "clip:vnode;"
                             "ref:vnode:opt;"
                             "sigma:float[]:opt;"
                             "block_step:int[]:opt;"
                             "bm_range:int[]:opt;"
                             "radius:int:opt;"
                             "ps_num:int[]:opt;"
                             "ps_range:int[]:opt;"
                             "num_streams:int:opt;"
                             "extractor_exp:int:opt;",
                             "clip:vnode;", bm3d_create, nullptr, plugin);
}
