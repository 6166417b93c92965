// Test-only record of the C-ABI calls the VapourSynth shim makes, kept by
// stub_abi.cpp (a stand-in for libvsfeel_bm3d.so with no device).
#ifndef VSFEEL_TEST_STUB_ABI_H
#define VSFEEL_TEST_STUB_ABI_H

#include <cstddef>
#include <string>
#include <vector>

#include "vsfeel_bm3d.h"

struct StubLog {
    int creates = 0, frees = 0, live_handles = 0;
    bool fail_create = false, fail_get = false;
    // the create arguments
    vsfeel_bm3d_format clip{}, ref{};
    int has_ref = 0, fp_mode = -1, device = -1, precision = -1;
    std::vector<double> sigma;
    std::vector<long long> block_step, bm_range, radius, ps_num, ps_range, num_streams, extractor_exp;
    // the last get_frame call
    int n = -1;
    std::vector<const float *> planes, ref_planes;
    std::vector<ptrdiff_t> strides, ref_strides;
    float *dst = nullptr;
    ptrdiff_t dst_stride = 0;
};

extern StubLog g_stub;

#endif
