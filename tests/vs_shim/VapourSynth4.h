/*
 * Test-only stand-in for VapourSynth's API4 header, written for compiling
 * vapoursynth-feel_amd/csrc/vsfeel_plugin.cpp against the fake core in
 * fake_core.cpp.  It declares only the types, constants and VSAPI /
 * VSPLUGINAPI members the shim touches, with API4's names and signatures;
 * the member order and the constants' values are not the real header's, so
 * a plugin built against it loads into nothing but the fake core.
 */
#ifndef VSFEEL_TEST_VAPOURSYNTH4_H
#define VSFEEL_TEST_VAPOURSYNTH4_H

#include <cstddef>
#include <cstdint>

#define VS_CC
#define VS_EXTERNAL_API(ret) extern "C" __attribute__((visibility("default"))) ret VS_CC
#define VS_MAKE_VERSION(major, minor) (((major) << 16) | (minor))
#define VAPOURSYNTH_API_VERSION VS_MAKE_VERSION(4, 0)

struct VSFrame;
struct VSNode;
struct VSCore;
struct VSPlugin;
struct VSMap;
struct VSFrameContext;
struct VSAPI;

enum VSColorFamily { cfUndefined = 0, cfGray = 1, cfRGB = 2, cfYUV = 3 };
enum VSSampleType { stInteger = 0, stFloat = 1 };
enum VSActivationReason { arInitial = 0, arAllFramesReady = 1, arError = -1 };
enum VSFilterMode { fmParallel = 0, fmParallelRequests = 1, fmUnordered = 2, fmFrameState = 3 };
enum VSRequestPattern { rpGeneral = 0, rpNoFrameReuse = 1, rpStrictSpatial = 2 };

struct VSVideoFormat {
    int colorFamily;
    int sampleType;
    int bitsPerSample;
    int bytesPerSample;
    int subSamplingW;
    int subSamplingH;
    int numPlanes;
};

struct VSVideoInfo {
    VSVideoFormat format;
    int64_t fpsNum;
    int64_t fpsDen;
    int width;
    int height;
    int numFrames;
};

struct VSFilterDependency {
    VSNode *source;
    int requestPattern;
};

typedef const VSFrame *(VS_CC *VSFilterGetFrame)(int n, int activationReason, void *instanceData, void **frameData,
                                                 VSFrameContext *frameCtx, VSCore *core, const VSAPI *vsapi);
typedef void(VS_CC *VSFilterFree)(void *instanceData, VSCore *core, const VSAPI *vsapi);
typedef void(VS_CC *VSPublicFunction)(const VSMap *in, VSMap *out, void *userData, VSCore *core,
                                      const VSAPI *vsapi);

struct VSAPI {
    void(VS_CC *createVideoFilter)(VSMap *out, const char *name, const VSVideoInfo *vi, VSFilterGetFrame getFrame,
                                   VSFilterFree free, int filterMode, const VSFilterDependency *dependencies,
                                   int numDeps, void *instanceData, VSCore *core);
    VSFrame *(VS_CC *newVideoFrame)(const VSVideoFormat *format, int width, int height, const VSFrame *propSrc,
                                    VSCore *core);
    VSFrame *(VS_CC *newVideoFrame2)(const VSVideoFormat *format, int width, int height, const VSFrame **planeSrc,
                                     const int *planes, const VSFrame *propSrc, VSCore *core);
    void(VS_CC *freeFrame)(const VSFrame *f);
    ptrdiff_t(VS_CC *getStride)(const VSFrame *f, int plane);
    const uint8_t *(VS_CC *getReadPtr)(const VSFrame *f, int plane);
    uint8_t *(VS_CC *getWritePtr)(VSFrame *f, int plane);
    void(VS_CC *freeNode)(VSNode *node);
    const VSVideoInfo *(VS_CC *getVideoInfo)(VSNode *node);
    const VSFrame *(VS_CC *getFrameFilter)(int n, VSNode *node, VSFrameContext *frameCtx);
    void(VS_CC *requestFrameFilter)(int n, VSNode *node, VSFrameContext *frameCtx);
    void(VS_CC *setFilterError)(const char *errorMessage, VSFrameContext *frameCtx);
    void(VS_CC *mapSetError)(VSMap *map, const char *errorMessage);
    int(VS_CC *mapNumElements)(const VSMap *map, const char *key);
    int64_t(VS_CC *mapGetInt)(const VSMap *map, const char *key, int index, int *error);
    double(VS_CC *mapGetFloat)(const VSMap *map, const char *key, int index, int *error);
    VSNode *(VS_CC *mapGetNode)(const VSMap *map, const char *key, int index, int *error);
};

struct VSPLUGINAPI {
    int(VS_CC *configPlugin)(const char *identifier, const char *pluginNamespace, const char *name,
                             int pluginVersion, int apiVersion, int flags, VSPlugin *plugin);
    int(VS_CC *registerFunction)(const char *name, const char *args, const char *returnType,
                                 VSPublicFunction argsFunc, void *functionData, VSPlugin *plugin);
};

#endif
