// Test-only fake VapourSynth core for the shim in
// vapoursynth-feel_amd/csrc/vsfeel_plugin.cpp: plugin registration, the
// create function with argument maps, frame requests and filter free,
// printing what the shim did as "key: value" lines for tests/test_vs_shim.py.
//
//   fake_core register           configPlugin / registerFunction arguments
//   fake_core create <case>      the create function on a prepared map
//   fake_core frames <case>      (stub library) getFrame requests and free
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "VapourSynth4.h"
#ifdef VSFEEL_STUB
#include "stub_abi.h"
#endif

extern "C" void VapourSynthPluginInit2(VSPlugin *plugin, const VSPLUGINAPI *vspapi);

struct VSFrame {
    int planes = 1;
    int w[3] = {}, h[3] = {};
    ptrdiff_t stride[3] = {};
    std::vector<unsigned char> data[3];
    const VSFrame *props = nullptr;  // frame whose properties it carries
};

struct VSNode {
    VSVideoInfo vi{};
    float base = 0.0f;  // frame f, plane p holds base + f + p / 8 + x / 1024
    int frees = 0;
};

struct VSMap {
    std::map<std::string, std::vector<double>> floats;
    std::map<std::string, std::vector<long long>> ints;
    std::map<std::string, std::vector<VSNode *>> nodes;
    std::string error;
    // what createVideoFilter received
    std::string filter;
    int mode = -1, num_deps = -1, dep_pattern[2] = { -1, -1 };
    VSNode *deps[2] = {};
    VSFilterGetFrame get_frame = nullptr;
    VSFilterFree free_fn = nullptr;
    void *instance = nullptr;
};

struct VSFrameContext {
    std::vector<std::pair<int, VSNode *>> requested;
    std::string error;
};

struct VSPlugin {
    std::string identifier, ns, name, function, args, ret;
    int plugin_version = 0, api_version = 0;
    VSPublicFunction create = nullptr;
};

namespace {

int g_live_frames = 0;
std::vector<const VSFrame *> g_made;  // frames the core handed out, for pointer checks

VSFrame *make_frame(const VSVideoFormat *f, int width, int height)
{
    auto *fr = new VSFrame();
    fr->planes = f->colorFamily == cfYUV ? 3 : 1;
    for (int p = 0; p < fr->planes; ++p) {
        fr->w[p] = p ? width >> f->subSamplingW : width;
        fr->h[p] = p ? height >> f->subSamplingH : height;
        fr->stride[p] = (fr->w[p] * 4 + 63) / 64 * 64;  // rows rounded to 64 bytes
        fr->data[p].assign(size_t(fr->stride[p]) * fr->h[p], 0xAB);
    }
    ++g_live_frames;
    g_made.push_back(fr);
    return fr;
}

float *row(const VSFrame *f, int p, int y)
{
    return reinterpret_cast<float *>(const_cast<unsigned char *>(f->data[p].data()) + y * f->stride[p]);
}

void VS_CC create_video_filter(VSMap *out, const char *name, const VSVideoInfo *, VSFilterGetFrame get_frame,
                               VSFilterFree free_fn, int mode, const VSFilterDependency *deps, int num_deps,
                               void *instance, VSCore *)
{
    out->filter = name;
    out->mode = mode;
    out->num_deps = num_deps;
    for (int i = 0; i < num_deps && i < 2; ++i) {
        out->deps[i] = deps[i].source;
        out->dep_pattern[i] = deps[i].requestPattern;
    }
    out->get_frame = get_frame;
    out->free_fn = free_fn;
    out->instance = instance;
}

VSFrame *VS_CC new_video_frame(const VSVideoFormat *f, int width, int height, const VSFrame *prop_src, VSCore *)
{
    VSFrame *fr = make_frame(f, width, height);
    fr->props = prop_src;
    return fr;
}

VSFrame *VS_CC new_video_frame2(const VSVideoFormat *f, int width, int height, const VSFrame **plane_src,
                                const int *planes, const VSFrame *prop_src, VSCore *)
{
    VSFrame *fr = make_frame(f, width, height);
    fr->props = prop_src;
    for (int p = 0; p < fr->planes; ++p) {
        if (plane_src[p]) {
            fr->data[p] = plane_src[p]->data[planes[p]];
        }
    }
    return fr;
}

void VS_CC free_frame(const VSFrame *f)
{
    if (f) {
        --g_live_frames;
        delete f;
    }
}

ptrdiff_t VS_CC get_stride(const VSFrame *f, int p) { return f->stride[p]; }
const uint8_t *VS_CC get_read_ptr(const VSFrame *f, int p) { return f->data[p].data(); }
uint8_t *VS_CC get_write_ptr(VSFrame *f, int p) { return f->data[p].data(); }
void VS_CC free_node(VSNode *n) { ++n->frees; }
const VSVideoInfo *VS_CC get_video_info(VSNode *n) { return &n->vi; }

const VSFrame *VS_CC get_frame_filter(int n, VSNode *node, VSFrameContext *)
{
    VSFrame *fr = make_frame(&node->vi.format, node->vi.width, node->vi.height);
    for (int p = 0; p < fr->planes; ++p) {
        for (int y = 0; y < fr->h[p]; ++y) {
            for (int x = 0; x < fr->w[p]; ++x) {
                row(fr, p, y)[x] = node->base + n + p / 8.0f + x / 1024.0f;
            }
        }
    }
    return fr;
}

void VS_CC request_frame_filter(int n, VSNode *node, VSFrameContext *ctx) { ctx->requested.push_back({ n, node }); }
void VS_CC set_filter_error(const char *msg, VSFrameContext *ctx) { ctx->error = msg; }
void VS_CC map_set_error(VSMap *m, const char *msg) { m->error = msg; }

int VS_CC map_num_elements(const VSMap *m, const char *key)
{
    if (m->floats.count(key)) {
        return int(m->floats.at(key).size());
    }
    if (m->ints.count(key)) {
        return int(m->ints.at(key).size());
    }
    if (m->nodes.count(key)) {
        return int(m->nodes.at(key).size());
    }
    return -1;  // API4: the key is absent
}

int64_t VS_CC map_get_int(const VSMap *m, const char *key, int i, int *err)
{
    auto it = m->ints.find(key);
    const bool ok = it != m->ints.end() && i < int(it->second.size());
    if (err) {
        *err = ok ? 0 : 1;
    }
    return ok ? it->second[size_t(i)] : 0;
}

double VS_CC map_get_float(const VSMap *m, const char *key, int i, int *err)
{
    auto it = m->floats.find(key);
    const bool ok = it != m->floats.end() && i < int(it->second.size());
    if (err) {
        *err = ok ? 0 : 1;
    }
    return ok ? it->second[size_t(i)] : 0.0;
}

VSNode *VS_CC map_get_node(const VSMap *m, const char *key, int i, int *err)
{
    auto it = m->nodes.find(key);
    const bool ok = it != m->nodes.end() && i < int(it->second.size());
    if (err) {
        *err = ok ? 0 : 1;
    }
    return ok ? it->second[size_t(i)] : nullptr;
}

int VS_CC config_plugin(const char *identifier, const char *ns, const char *name, int version, int api, int,
                        VSPlugin *plugin)
{
    plugin->identifier = identifier;
    plugin->ns = ns;
    plugin->name = name;
    plugin->plugin_version = version;
    plugin->api_version = api;
    return 1;
}

int VS_CC register_function(const char *name, const char *args, const char *ret, VSPublicFunction fn, void *,
                            VSPlugin *plugin)
{
    plugin->function = name;
    plugin->args = args;
    plugin->ret = ret;
    plugin->create = fn;
    return 1;
}

const VSAPI kApi = { create_video_filter, new_video_frame, new_video_frame2, free_frame,     get_stride,
                     get_read_ptr,        get_write_ptr,   free_node,        get_video_info, get_frame_filter,
                     request_frame_filter, set_filter_error, map_set_error, map_num_elements, map_get_int,
                     map_get_float,       map_get_node };
const VSPLUGINAPI kPluginApi = { config_plugin, register_function };

VSNode gray_node(int w, int h, int frames, float base)
{
    VSNode n;
    n.vi.format = { cfGray, stFloat, 32, 4, 0, 0, 1 };
    n.vi.width = w;
    n.vi.height = h;
    n.vi.numFrames = frames;
    n.vi.fpsNum = 24;
    n.vi.fpsDen = 1;
    n.base = base;
    return n;
}

[[maybe_unused]] VSNode yuv_node(int w, int h, int frames, float base)
{
    VSNode n = gray_node(w, h, frames, base);
    n.vi.format = { cfYUV, stFloat, 32, 4, 1, 1, 3 };
    return n;
}

VSPlugin load_plugin()
{
    VSPlugin p;
    VapourSynthPluginInit2(&p, &kPluginApi);
    return p;
}

// the create function on map `in`; prints the error or the filter and what
// happened to the nodes
void create_and_report(VSPlugin &p, VSMap &in, VSMap &out, VSNode *clip, VSNode *ref)
{
    p.create(&in, &out, nullptr, nullptr, &kApi);
    std::printf("error: %s\n", out.error.c_str());
    std::printf("filter: %s\n", out.filter.c_str());
    std::printf("clip_frees: %d\n", clip->frees);
    std::printf("ref_frees: %d\n", ref ? ref->frees : -1);
}

}  // namespace

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "";
    const std::string which = argc > 2 ? argv[2] : "";
    VSPlugin p = load_plugin();
    if (mode == "register") {
        std::printf("identifier: %s\n", p.identifier.c_str());
        std::printf("namespace: %s\n", p.ns.c_str());
        std::printf("function: %s\n", p.function.c_str());
        std::printf("args: %s\n", p.args.c_str());
        std::printf("return: %s\n", p.ret.c_str());
        std::printf("api_version: %d\n", p.api_version);
        return 0;
    }
    VSNode clip = gray_node(40, 32, 6, 0.0f), ref = gray_node(40, 32, 6, 0.5f);
    VSMap in, out;
    if (mode == "create") {
        in.nodes["clip"] = { &clip };
        VSNode *r = nullptr;
        if (which == "bad_sigma") {
            in.floats["sigma"] = { 3.0, -0.5 };
        } else if (which == "bad_ref") {
            ref = gray_node(41, 32, 6, 0.5f);
            r = &ref;
        } else if (which == "bad_radius_ref") {
            in.ints["radius"] = { 9 };
            r = &ref;
        } else if (which == "valid_ref") {
            in.floats["sigma"] = { 3.0, 2.0, 1.0 };
            in.ints["radius"] = { 1 };
            r = &ref;
        }
        if (r) {
            in.nodes["ref"] = { r };
        }
        create_and_report(p, in, out, &clip, r);
        if (!out.filter.empty()) {
            out.free_fn(out.instance, nullptr, &kApi);
        }
        return 0;
    }
#ifdef VSFEEL_STUB
    if (mode == "frames") {
        const bool yuv = which == "yuv" || which == "fail";
        clip = yuv ? yuv_node(40, 32, 6, 0.0f) : gray_node(40, 32, 6, 0.0f);
        ref = yuv ? yuv_node(40, 32, 6, 0.5f) : gray_node(40, 32, 6, 0.5f);
        in.nodes["clip"] = { &clip };
        in.nodes["ref"] = { &ref };
        in.floats["sigma"] = { 3.0, 2.0, 1.0 };
        in.ints["block_step"] = { 4 };
        in.ints["bm_range"] = { 16, 9 };
        in.ints["radius"] = { 1 };
        in.ints["ps_num"] = { 2 };
        in.ints["ps_range"] = { 7 };
        in.ints["num_streams"] = { 3 };
        in.ints["extractor_exp"] = { 2 };
        g_stub.fail_get = which == "fail";
        create_and_report(p, in, out, &clip, &ref);
        std::printf("mode_parallel: %d\n", out.mode == fmParallel);
        std::printf("num_deps: %d\n", out.num_deps);
        std::printf("deps_nodes: %d\n", out.deps[0] == &clip && out.deps[1] == &ref);
        std::printf("deps_strict_spatial: %d\n",
                    out.dep_pattern[0] == rpStrictSpatial && out.dep_pattern[1] == rpStrictSpatial);
        std::printf("stub_sigma: %g %g %g\n", g_stub.sigma[0], g_stub.sigma[1], g_stub.sigma[2]);
        std::printf("stub_ints: %lld %lld,%lld %lld %lld %lld %lld %lld\n", g_stub.block_step[0],
                    g_stub.bm_range[0], g_stub.bm_range[1], g_stub.radius[0], g_stub.ps_num[0], g_stub.ps_range[0],
                    g_stub.num_streams[0], g_stub.extractor_exp[0]);
        std::printf("stub_format: %d %d %d %d %d %d %d %d\n", g_stub.clip.color_family, g_stub.clip.sample_type,
                    g_stub.clip.bits_per_sample, g_stub.clip.sub_sampling_w, g_stub.clip.width, g_stub.clip.height,
                    g_stub.clip.num_frames, g_stub.has_ref);
        std::printf("stub_modes: %d %d %d\n", g_stub.fp_mode, g_stub.device, g_stub.precision);

        const int n = 1;  // window frames 0-3 at radius 1
        VSFrameContext ctx;
        void *fd = nullptr;
        const VSFrame *none = out.get_frame(n, arInitial, out.instance, &fd, &ctx, nullptr, &kApi);
        std::string req;
        for (auto &q : ctx.requested) {
            req += std::to_string(q.first) + (q.second == &clip ? "c " : q.second == &ref ? "r " : "? ");
        }
        std::printf("initial_returns_null: %d\n", none == nullptr);
        std::printf("requested: %s\n", req.c_str());
        const size_t made_before = g_made.size();
        const VSFrame *dst = out.get_frame(n, arAllFramesReady, out.instance, &fd, &ctx, nullptr, &kApi);
        std::printf("frame_error: %s\n", ctx.error.c_str());
        std::printf("stub_n: %d\n", g_stub.n);
        // the planes handed to get_frame are the luma planes of the window's
        // frames (made by get_frame_filter, clip then ref per frame)
        bool planes_ok = g_stub.planes.size() == 4 && g_stub.ref_planes.size() == 4;
        for (size_t i = 0; planes_ok && i < 4; ++i) {
            const VSFrame *src = g_made[made_before + 2 * i];
            const VSFrame *rsrc = g_made[made_before + 2 * i + 1];
            planes_ok = g_stub.planes[i] == reinterpret_cast<const float *>(src->data[0].data()) &&
                        g_stub.strides[i] == src->stride[0] &&
                        g_stub.ref_planes[i] == reinterpret_cast<const float *>(rsrc->data[0].data()) &&
                        g_stub.ref_strides[i] == rsrc->stride[0] && row(src, 0, 3)[5] == (0 + i) + 5 / 1024.0f &&
                        row(rsrc, 0, 3)[5] == 0.5f + i + 5 / 1024.0f;
        }
        std::printf("planes_ok: %d\n", planes_ok);
        if (dst) {
            bool luma = g_stub.dst == reinterpret_cast<const float *>(dst->data[0].data()) &&
                        g_stub.dst_stride == dst->stride[0];
            bool pad = true;
            for (int y = 0; y < 32; ++y) {
                luma = luma && row(dst, 0, y)[0] == 1.25f && row(dst, 0, y)[39] == 1.25f;
                pad = pad && dst->data[0][size_t(y) * dst->stride[0] + 40 * 4] == 0xAB;
            }
            std::printf("dst_luma_ok: %d\n", luma);
            std::printf("dst_padding_untouched: %d\n", pad);
            bool chroma = dst->planes == (yuv ? 3 : 1);
            for (int pl = 1; pl < dst->planes; ++pl) {
                chroma = chroma && row(dst, pl, 7)[3] == 1 + pl / 8.0f + 3 / 1024.0f;
            }
            std::printf("chroma_copied: %d\n", chroma);
            // YUV: the source frame's properties (newVideoFrame2); Gray: none
            const VSFrame *want_props = yuv ? g_made[made_before + 2 * n] : nullptr;
            std::printf("props_ok: %d\n", dst->props == want_props);
            free_frame(dst);
        } else {
            std::printf("dst: null\n");
        }
        std::printf("live_frames_after_request: %d\n", g_live_frames);
        out.free_fn(out.instance, nullptr, &kApi);
        std::printf("stub_frees: %d\n", g_stub.frees);
        std::printf("live_handles: %d\n", g_stub.live_handles);
        std::printf("clip_frees_after_free: %d\n", clip.frees);
        std::printf("ref_frees_after_free: %d\n", ref.frees);
        return 0;
    }
#endif
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
