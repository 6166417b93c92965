// Shared definitions of the BM3Dv2 HIP kernels and their host launcher.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsfeel {

constexpr int kMaxRadius = 4;
constexpr int kMaxWindow = 2 * kMaxRadius + 1;

// Group table entry: top-left pixel (x, y) of a block (14 bits each) and the
// window position z of its plane (4 bits).  Frames are limited to 16383 px.
constexpr int kMaxDim = 16383;
__host__ __device__ inline uint32_t pack_entry(uint32_t x, uint32_t y, uint32_t z)
{
    return x | (y << 14) | (z << 28);
}
__host__ __device__ inline uint32_t entry_x(uint32_t e) { return e & 0x3FFFu; }
__host__ __device__ inline uint32_t entry_y(uint32_t e) { return (e >> 14) & 0x3FFFu; }
__host__ __device__ inline uint32_t entry_z(uint32_t e) { return e >> 28; }

// Block matching: one estimate's search, for every active ref block.
struct MatchArgs {
    const float *plane[kMaxWindow];  // search clip, window positions z = 0..2r
    uint32_t *groups;                // out: [nref][8] packed entries
    float *dists;                    // optional out: [nref][8] group distances
    int width, height, pitch;
    int step, bm_range, radius, ps_num, ps_range;
    int nx, ny;                      // active ref-block grid
};

// Collaborative filtering + weighted scatter of every group.
struct EstimateArgs {
    const float *search[kMaxWindow]; // block-matching / Wiener reference clip
    const float *input[kMaxWindow];  // clip being denoised (== search outside the final pass)
    const uint32_t *groups;
    float *res;                      // [2r+1][2][height*res_pitch] accumulators
    int height, pitch;               // pitch of the search / input planes
    int res_pitch;
    int ngroups;
    float sigma;
    float extractor;
};

// Aggregation of the 2r+1 slices that land on one output frame.
struct AggArgs {
    const float *slice[kMaxWindow];  // wdst planes; the weight plane follows at +plane_elems
    float *dst;
    long long plane_elems;           // height * pitch of the res planes
    int width, height, pitch;        // res geometry
    int dst_pitch;
    int tw;
};

enum FpMode : int {
    kFpNone = 0,      // GLSL fma() rounded as a multiply and an add
    kFpExplicit = 1,  // GLSL fma() fused, nothing else contracted
    kFpAll = 2,       // GLSL fma() and every contractible a*b+c fused
};

}  // namespace vsfeel
