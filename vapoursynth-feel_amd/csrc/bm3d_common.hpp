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
    // spatial top-8 per ref block (distance bits, x | y << 16): written by
    // bm3d_match_rows when the temporal search follows, read by bm3d_match
    uint2 *spatial;
    // bm3d_match_rows geometry: ref blocks per workgroup run, and its LDS
    // layout (column-major source region, per-block distance rows)
    int run_blocks, runs_x;
    int region_cols, region_rows, dist_cols;
    // temporal search after bm3d_match_rows: distance rows per wavefront
    // (ps_num * (2 ps_range + 1)) for the threshold selection, 0: list
    int t_rows;
    // nonzero: the temporal search runs both directions in one wavefront
    // (bm3d_match_predictive; needs ps_num * (2 ps_range + 1) <= 32)
    int pred_pair;
    // precision form (operand != kOperandFp32): MFMA distances of the
    // spatial search from the reference frame's plane set (bm3d_planes,
    // layout below), runs of mfma_run_blocks ref blocks per workgroup
    // (mfma_runs_x per grid row)
    int operand;
    const uint16_t *pset;
    int pset_py;
    long long pset_elems;
    int mfma_run_blocks, mfma_runs_x;
};

// Arithmetic of the block-matching distances: the reference's FP32 sums
// (plain form), or MFMA dot products of pixels rounded once to FP16 / BF16
// (precision form; shortlists rescored with the FP32 sums).
enum OperandType : int {
    kOperandFp32 = 0,
    kOperandFp16 = 1,
    kOperandBf16 = 2,
};

// Plane set of one search frame for the precision form (bm3d_planes), with
// py = plane_py(height) rows per column (rows past the frame hold zeros):
//   s       the pixels minus kPlaneCentre, rounded once to the operand type
//           (FP16 or BF16), row-major: (x, y) at y * width + x
//   h       sum of s over columns x .. x+7 (float, exact for these operands)
//           as hi + lo in the operand type (x <= width - 8), column-major:
//           (x, y) at x * py + y
//   box     sum of s^2 over the 8x8 block at (x, y) (float; x <= width - 8,
//           y <= height - 8), column-major
// Order: hhi, hlo, s (uint16), box (float), each width * py elements.
constexpr float kPlaneCentre = 0.5f;

__host__ __device__ inline int plane_py(int height) { return (height + 24 + 7) & ~7; }

__host__ __device__ inline long long plane_set_bytes(int width, int height)
{
    return (long long)width * plane_py(height) * (3 * 2 + 4);
}

// LDS of bm3d_match_mfma for a run region of `cols` source columns and a
// window of `rh` rows: box (rows padded to 16), per wavefront 512 bytes
// (operand rows, shortlist), h hi + lo (column-major, rows padded to 8) and
// s (row-major, rows of an even number of elements), each 7 rows more than
// the window's.  The column strides of box and h are odd multiples of 4
// dwords: the 8 lanes of one pass of a 16-byte LDS read (consecutive
// columns) then fall on distinct banks.
constexpr int kMfmaWaveBytes = 512;
__host__ __device__ inline int mfma_odd4(int dw) { return ((dw >> 2) & 1) ? dw : dw + 4; }
__host__ __device__ inline int mfma_region_rows(int rh) { return (rh + 7 + 7) & ~7; }
__host__ __device__ inline int mfma_box_rows(int rh) { return (rh + 15) & ~15; }
__host__ __device__ inline int mfma_h_stride(int rh) { return 2 * mfma_odd4(mfma_region_rows(rh) / 2); }
__host__ __device__ inline int mfma_box_stride(int rh) { return mfma_odd4(mfma_box_rows(rh)); }
__host__ __device__ inline int mfma_s_stride(int cols) { return (cols + 1) & ~1; }
__host__ __device__ inline long long mfma_rows_lds_bytes(int cols, int rh)
{
    const long long nhc = cols - 7;
    return nhc * mfma_box_stride(rh) * 4 + 4LL * kMfmaWaveBytes + 2 * nhc * mfma_h_stride(rh) * 2LL +
           (rh + 7LL) * mfma_s_stride(cols) * 2;
}

// Dynamic LDS of bm3d_match with the temporal threshold selection: per
// wavefront (4) a 64-entry stage (8-byte keys, 4-byte positions) and t_rows
// distance rows of 2 ps_range + 1 floats.
__host__ __device__ inline long long temporal_lds_bytes(int t_rows, int ps_range)
{
    return t_rows > 0 ? 4LL * (64 * 12 + 4LL * t_rows * (2LL * ps_range + 1)) : 0;
}

// LDS bytes of bm3d_match_predictive (one wavefront per workgroup): 2 x 32
// survivor keys and positions, the reference block, the 2 ps_num predictor
// windows of a step ((2 ps_range + 8)^2 floats each, column-major) and a
// distance row per lane
// rows a lane of bm3d_match_predictive loads per batch
constexpr int kPredBatch = 24;

// window stride (floats) of bm3d_match_predictive's staged windows: 16 banks
// modulo 32 past the window's wc x wc floats, so that the row scan's two
// windows of a half wave (a window row per lane) fall on disjoint banks
__host__ __device__ inline int predictive_win_stride(int ps_range)
{
    const int wc = 2 * ps_range + 8;
    return wc * wc + ((48 - (wc * wc) % 32) % 32);
}

__host__ __device__ inline long long predictive_lds_bytes(int ps_num, int ps_range)
{
    const long long wh = 2LL * ps_range + 1;
    return 64LL * 8 + 64LL * 4 + 64LL * 4 + 4LL * 2 * ps_num * predictive_win_stride(ps_range) + 4LL * 64 * wh;
}

// LDS bytes of bm3d_match_rows: survivor stages of its 4 wavefronts (64 keys
// + positions each), per window row its 3 smallest distances and the
// positions of the first two (16 bytes), the source region, the blocks, their
// distance rows (region_rows - 7 window rows at most) and per block an
// overflow flag and its threshold.
// column stride of the spatial kernel's column-major source region: odd, so
// that the staging stores (one column per lane) fall in distinct LDS banks
__host__ __device__ inline int region_col_stride(int region_rows) { return region_rows | 1; }

// Run geometry of bm3d_match_rows for a frame at least as large as the run's
// region (the window's 2 bm_range + 1 rows and columns, and the run's
// columns, inside the frame): 256 window rows per workgroup, so
// 256 / (2 bm_range + 1) blocks per run.  Zero for bm_range <= 0.
struct RowsGeometry {
    int run_blocks, region_cols, region_rows, dist_cols;
};
__host__ __device__ constexpr RowsGeometry rows_geometry(int bm_range, int step)
{
    return bm_range > 0 ? RowsGeometry{256 / (2 * bm_range + 1), (256 / (2 * bm_range + 1) - 1) * step + 8 + 2 * bm_range,
                                       2 * bm_range + 8, 2 * bm_range + 1}
                        : RowsGeometry{0, 0, 0, 0};
}

__host__ __device__ inline long long rows_lds_bytes(int region_cols, int region_rows, int run_blocks, int dist_cols)
{
    // staged positions, per-block flags and thresholds, 3 row keys per
    // window row, the source region, 16-bit distance rows
    const long long rh = region_rows - 7;
    return 256LL * 4 + 8LL * run_blocks + 12LL * run_blocks * rh +
           4LL * region_cols * region_col_stride(region_rows) + 2LL * run_blocks * rh * dist_cols;
}

// Collaborative filtering + weighted scatter of every group.
struct EstimateArgs {
    const float *search[kMaxWindow]; // block-matching / Wiener reference clip
    const float *input[kMaxWindow];  // clip being denoised (== search outside the final pass)
    const uint32_t *groups;
    float *res;                      // [2r+1][2][height*res_pitch] accumulators
    int width, height, pitch;        // pitch of the search / input planes
    int res_pitch;
    int ngroups;
    float sigma;
    float extractor;
    int radius;
    // tiled accumulation (scratch != nullptr): tiles of tile_bx x tile_by ref
    // blocks, each accumulating its own frame's contributions over a region of
    // at most region_w x region_h pixels
    float *scratch;                  // [tiles][2][region_h][region_w]
    int nx, ny, step, halo;
    int tile_bx, tile_by, tiles_x, tiles_y;
    int region_w, region_h;
    // differentiable form: per group keep masks (8 words: column frequency
    // lane u, bit w*8+v) and weights, written (kFilterSave) or applied
    // (kFilterApply)
    uint64_t *masks;
    float *weights;
    // radius > 0: the slices of the other frames receive each block's weight
    // once, at its origin, in these planes ([2r][height][origin_pitch], the
    // reference frame's slice skipped); bm3d_weight_box expands them
    float *origins;
    int origin_pitch;
    // radius > 0, tiled: the other frames' blocks are binned instead of
    // added with atomics.  A block's 64 weighted values go to a record slot
    // of its workgroup (rec, 64 floats per slot); the
    // record's id is appended to the list of every 32x32 cell of its slice
    // that the block covers (cell_list, cell_cap (id, entry) pairs per cell, cell_count
    // appended so far).  bm3d_gather sums the lists into the slices.  A
    // block whose cell list is full adds its pixels in that cell atomically.
    // A workgroup counts its appends per cell of its window (cwin_x x
    // cwin_y cells per slice around its blocks, bin_halo pixels beyond
    // them) in LDS and reserves each cell's share of the list at once.
    float *rec;                      // [tiles][rec_per_tile][64]
    uint32_t *cell_count;            // [2r][ncells]
    uint2 *cell_list;                // [2r][ncells][cell_cap] (record id, entry)
    int rec_per_tile, cells_x, ncells, cell_cap;
    int cwin_x, cwin_y, bin_halo;
};

// LDS bytes the binning adds to a tiled estimate workgroup (after its
// accumulator): every record slot's entry, and per window cell its count,
// list base and fill
inline size_t bin_lds_bytes(int rec_per_tile, int radius, int cwin_x, int cwin_y)
{
    return sizeof(uint32_t) * (size_t(rec_per_tile) + 3 * size_t(2 * radius) * cwin_x * cwin_y);
}

constexpr int kCellShift = 5;        // binning cells of 32 x 32 pixels

// Sums of the binned other-frame blocks into the value planes of an
// estimate stack's other slices (one workgroup per cell and slice).
struct GatherArgs {
    const float *rec;
    const uint32_t *cell_count;
    const uint2 *cell_list;
    float *res;                      // the estimate stack, [2r+1][2][height][res_pitch]
    int width, height, res_pitch;
    int cells_x, ncells, cell_cap, radius;
};

// 8x8 box sums of the origin weight planes into the weight planes of the
// other frames' slices of one estimate stack.
struct BoxArgs {
    const float *origins;            // [2r][height][origin_pitch]
    float *res;                      // the estimate stack, [2r+1][2][height][res_pitch]
    int width, height, origin_pitch, res_pitch;
    int radius;
};

enum FilterMode : int {
    kFilterPlain = 0,  // threshold the input's own spectrum
    kFilterSave = 1,   // ... and store the keep masks and weights
    kFilterApply = 2,  // apply stored masks and weights (a fixed linear map)
};

// Pixel span [lo, hi] that the groups of tile t (along one axis) can reach in
// their own frame: the tile's ref blocks, widened by the search range.
__host__ __device__ inline void tile_span(int t, int tile_blocks, int nblocks, int step, int size, int halo,
                                          int *lo, int *hi)
{
    const int b0 = t * tile_blocks;
    const int b1 = (b0 + tile_blocks < nblocks ? b0 + tile_blocks : nblocks) - 1;
    const int p0 = b0 * step < size - 8 ? b0 * step : size - 8;
    const int p1 = b1 * step < size - 8 ? b1 * step : size - 8;
    *lo = p0 - halo > 0 ? p0 - halo : 0;
    *hi = (p1 + halo < size - 8 ? p1 + halo : size - 8) + 7;
}

// Reduction of the tile accumulators into the dense slice of the frame itself.
struct ReduceArgs {
    const float *scratch;
    float *res;                      // wdst plane of the slice; the weight plane follows at +height*res_pitch
    const int2 *col_tiles;           // per column: first and last tile covering it
    const int2 *row_tiles;           // per row
    const int *tile_x0;              // per tile column: region origin
    const int *tile_y0;              // per tile row
    int width, height, res_pitch;
    int tiles_x, region_w, region_h;
    int add_existing;                // add to the slice instead of overwriting it
    int span3;                       // every pixel lies in at most 3 x 3 tiles, scratch < 2^32 elements
    // radius 0: the frame itself instead of the slice (the aggregation of a
    // single slice, float(double(wdst) / double(weight))), or nullptr
    float *dst;
    int dst_pitch;
};

// Aggregation of the 2r+1 slices that land on one output frame.
struct AggArgs {
    const float *slice[kMaxWindow];  // wdst planes; the weight plane follows at +plane_elems
    float *dst;
    long long plane_elems;           // height * pitch of the res planes
    int width, height, pitch;        // res geometry
    int dst_pitch;
    int tw;
    float *denom;                    // optional: the weight sums (the output's denominator)
    int denom_pitch;
    int numerator_only;              // write the weighted sums without dividing
};

enum FpMode : int {
    kFpNone = 0,      // GLSL fma() rounded as a multiply and an add
    kFpExplicit = 1,  // GLSL fma() fused, nothing else contracted
    kFpAll = 2,       // GLSL fma() and every contractible a*b+c fused
};

}  // namespace vsfeel
