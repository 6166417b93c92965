// Host side of the BM3Dv2 library: argument validation (BM3DCreate), frame
// caches and per-frame launch sequences (BM3DGetFrame), behind the C ABI of
// include/vsfeel_bm3d.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <condition_variable>
#include <mutex>
#include <string>
#include <vector>

#include "bm3d_common.hpp"
#include "vsfeel_bm3d.h"

namespace vsfeel {

// block matching: radius 1-2 with a row-parallel spatial search runs in two
// passes (the spatial search storing its top-8, then the temporal search:
// temporal_pass), otherwise in one (temporal_pass false)
bool match_two_pass(const MatchArgs &a);
hipError_t launch_match(const MatchArgs &a, int fp_mode, hipStream_t s, bool temporal_pass);
hipError_t launch_estimate(const EstimateArgs &a, int fp_mode, bool final_pass, int mode, hipStream_t s);
hipError_t launch_divide(const float *num, int num_pitch, const float *den, int den_pitch, float *out, int out_pitch,
                         int width, int height, hipStream_t s);
hipError_t launch_aggregate(const AggArgs &a, hipStream_t s);
hipError_t launch_fill_zero(float *p, long long n, hipStream_t s);
hipError_t launch_tile_reduce(const ReduceArgs &a, hipStream_t s);
hipError_t launch_weight_box(const BoxArgs &a, hipStream_t s);
hipError_t launch_gather(const GatherArgs &a, hipStream_t s);
hipError_t launch_planes(const float *src, int pitch, int width, int height, uint16_t *set, int operand,
                         hipStream_t s);
hipError_t launch_tile_tables(int *tab, int width, int height, int tiles_x, int tiles_y, int tile_bx, int tile_by,
                              int nx, int ny, int step, int halo, hipStream_t s);

namespace {

// sigma factors of BM3DCreate (src/bm3d.cpp:1105-1107): 0x3f021bb6 for the
// basic pass, 0x3e40c0c1 for the final (Wiener) pass
float bits_to_float(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// vsh::int64ToIntS: saturate to int
int sat_int(int64_t v)
{
    if (v > INT_MAX) {
        return INT_MAX;
    }
    if (v < INT_MIN) {
        return INT_MIN;
    }
    return int(v);
}

int clampi(int v, int lo, int hi) { return std::min(std::max(v, lo), hi); }

// slice of frame clamp(n-r+i)'s stack that lands on frame n (src/bm3d.cpp:369-371)
int agg_z(int i, int n, int nframes, int radius)
{
    return std::min(std::max(2 * radius - i, n - nframes + 1 + radius), n + radius);
}

enum KernelKind {
    kKernelFill = 0,
    kKernelMatch,
    kKernelEstimate,
    kKernelAggregate,
    kKernelReduce,
    kKernelWeightBox,
    kKernelGather,
    kKernelPlanes,
    kKernelTemporal,
    kKernelKinds
};

// stages of a host-buffer request that the test hook can fail
enum FailStage { kFailNone = 0, kFailUpload = 1, kFailLaunch = 2 };

// LDS bytes a tile accumulator may take (one 512-thread workgroup per CU,
// 16 bytes per region pixel), and the groups one pass of it filters
constexpr int kTileLdsBudget = 112 * 1024;
constexpr int kTileGroupsPerPass = 64;

// Compute units of the MI355X: the tile choice below is fixed by it, not by
// the device the instance runs on, so the tiling (and the last bit of the
// float partial sums) never depends on the GPU.
constexpr int kComputeUnits = 256;

// Tile shape of the tiled estimate: the most ref blocks per workgroup whose
// accumulator fits the LDS budget, preferring whole passes of 64 groups.
// A frame that this leaves with fewer tiles than compute units (a 512x512
// frame at block_step 4: 88 tiles) gets the shape that minimises the
// launch's rounds of tiles times a tile's cost instead (256 tiles there),
// so that the grid fills the GPU.
bool choose_tiles(int width, int height, int step, int range, int *tbx, int *tby, int *rw, int *rh)
{
    const int nx = (width - 8 + step + step - 1) / step, ny = (height - 8 + step + step - 1) / step;
    auto search = [&](bool whole_launch) {
        double best = 1e30;
        bool found = false;
        for (int by = 1; by <= 32; ++by) {
            for (int bx = 1; bx <= 32; ++bx) {
                const int w = std::min((bx - 1) * step + 8 + 2 * range, width);
                const int h = std::min((by - 1) * step + 8 + 2 * range, height);
                if (2LL * w * h * 8 > kTileLdsBudget) {
                    continue;
                }
                const int groups = bx * by;
                const int passes = (groups + kTileGroupsPerPass - 1) / kTileGroupsPerPass;
                // per group: idle lanes plus a small charge for the scratch
                // traffic; per launch: rounds of tiles times a tile's cost
                const double tile = double(passes * kTileGroupsPerPass) + 0.05 * double(w) * h / (double(step) * step);
                const long long tiles = (long long)((nx + bx - 1) / bx) * ((ny + by - 1) / by);
                const double cost = whole_launch ? double((tiles + kComputeUnits - 1) / kComputeUnits) * tile
                                                 : tile / groups;
                if (cost < best - 1e-12) {
                    best = cost;
                    *tbx = bx;
                    *tby = by;
                    *rw = w;
                    *rh = h;
                    found = true;
                }
            }
        }
        return found;
    };
    if (!search(false)) {
        return false;
    }
    if ((long long)((nx + *tbx - 1) / *tbx) * ((ny + *tby - 1) / *tby) < kComputeUnits) {
        search(true);
    }
    return true;
}

}  // namespace

// Saved state of the differentiable form for one frame (or none).
struct SavedState {
    uint32_t *groups = nullptr;
    uint64_t *masks = nullptr;
    float *weights = nullptr;
};

struct Instance {
    // parameters (BM3DData, src/bm3d.cpp:74-130)
    int width = 0, height = 0, nframes = 0;
    bool chroma = false, final_pass = false, process = true;
    float sigma = 0, sigma_u = 0, sigma_v = 0, extractor = 0;
    int block_step = 8, bm_range = 9, radius = 0, ps_num = 2, ps_range = 4, num_streams = 4;
    int fp_mode = kFpExplicit, device = 0;
    // distances of the block matching (OperandType): FP32 (plain form) or
    // MFMA FP16 / BF16 (precision form, with a plane set per window position)
    int operand = kOperandFp32;
    uint16_t *psets = nullptr;  // the reference frame's plane set (precision form)
    long long pset_bytes = 0;
    int mfma_run_blocks = 0, mfma_runs_x = 0;
    int tw = 1, src_ring = 1, res_cap = 1, clips = 1;
    int nx = 0, ny = 0, nref = 0;
    int pitch = 0;     // elements, plane pitch of the internal buffers
    long long pe = 0;  // height * pitch

    float *src_buf = nullptr;    // src_ring * clips planes (ref clip first in final mode)
    float *res_buf = nullptr;    // res_cap stacks of tw * 2 planes
    uint32_t *groups = nullptr;  // nref * 8
    float *dists = nullptr;      // nref * 8 (debug)
    // Lanes (tiled, every window 8 candidates or more).  An estimate writes
    // only its result slot (or, at radius 0, its output frame) and its
    // intermediate buffers: group table, tile accumulators, plane set and at
    // radius > 0 the spatial top-8, origin planes and binning records.  With
    // a set of those per lane, process_device runs consecutive estimates
    // round-robin on several streams, so one frame's filtering overlaps the
    // next one's block matching.  Radius 0: kLanes lanes, lane 0 on the
    // launch stream, lane i > 0 on stream_l[i - 1] (2864 frames/s at 1080p
    // against 2826 with 2 lanes, 2653 with 1 and 2828 with 4, where the
    // streams outnumber the 4 hardware queues).  Radius > 0: kTemporalLanes
    // lanes, lane i on stream_l[i], the aggregations on the launch stream,
    // each after the latest estimate of every lane; an estimate waits for
    // the last aggregation that read the result slot it overwrites.
    // (three lanes at radius > 0: 1047 against 1037 frames/s at 1080p
    // radius 2 but 365 against 367 on 4K radius 1, for another set of
    // binning records)
    static constexpr int kLanes = 3, kTemporalLanes = 2, kMaxLanes = 3;
    static_assert(kLanes <= kMaxLanes && kTemporalLanes <= kMaxLanes - 1,
                  "radius > 0 puts every lane on a side stream");
    uint32_t *groups_l[kMaxLanes - 1] = {};
    float *scratch_l[kMaxLanes - 1] = {};
    uint16_t *psets_l[kMaxLanes - 1] = {};
    uint2 *spatial_l[kMaxLanes - 1] = {};
    float *origins_l[kMaxLanes - 1] = {};
    float *rec_l[kMaxLanes - 1] = {};
    uint32_t *cell_count_l[kMaxLanes - 1] = {};
    uint2 *cell_list_l[kMaxLanes - 1] = {};
    float *dst_l[kMaxLanes - 1] = {};  // radius 0: the backward pass's g / D plane
    hipStream_t stream_l[kMaxLanes - 1] = {};
    hipEvent_t lane_fork = nullptr, lane_join[kMaxLanes - 1] = {};
    hipEvent_t lane_done[kMaxLanes] = {};    // radius > 0: a lane's latest estimate
    std::vector<hipEvent_t> slot_read;       // radius > 0: last aggregation reading a result slot
    int lanes = 1;   // lanes of process_device (1: everything on the launch stream)
    int lane = 0;    // the buffer set of the estimate being enqueued
    uint2 *spatial = nullptr;    // nref * 8 spatial top-8 handed to the temporal search

    // row-parallel spatial search (bm3d_match_rows): ref blocks per
    // workgroup run and the LDS layout; run_blocks == 0 selects the
    // per-wavefront search instead
    int run_blocks = 0, runs_x = 0, region_cols = 0, region_rows = 0, dist_cols = 0;
    int pred_pair = 0;           // temporal search: both directions per wavefront
    int t_rows = 0;  // temporal threshold selection: distance rows per wavefront (0: list)
    float *dst_dev = nullptr;    // one plane

    // tiled accumulation of the frame's own slice (see bm3d_estimate_tiled)
    bool tiled = false;
    bool short_windows = false;  // some spatial window has fewer than 8 candidates
    int tile_bx = 0, tile_by = 0, tiles_x = 0, tiles_y = 0, region_w = 0, region_h = 0;
    float *scratch = nullptr;    // tiles * 2 * region_h * region_w
    float *origins = nullptr;    // radius > 0: 2r origin weight planes of height * pitch
    int *tile_tables = nullptr;  // col_tiles[W] (int2), row_tiles[H] (int2), tile_x0[tiles_x], tile_y0[tiles_y]
    bool reduce_span3 = false;   // every pixel lies in at most 3 x 3 tiles (bm3d_tile_reduce's unrolled form)
    // radius > 0, tiled: the other frames' blocks binned by 32x32 cell (see
    // EstimateArgs::rec), summed by bm3d_gather
    float *rec = nullptr;
    uint32_t *cell_count = nullptr;
    uint2 *cell_list = nullptr;
    int rec_per_tile = 0, cells_x = 0, ncells = 0, cell_cap = 0;
    int cwin_x = 0, cwin_y = 0, bin_halo = 0;
    long long cell_count_elems = 0;  // cleared per estimate (a multiple of 64)
    hipStream_t stream = nullptr;
    hipEvent_t last_done = nullptr;  // end of the previous device-side call
    bool have_last = false;

    std::vector<int> src_frame;  // frame held by each src ring slot (-1: none)
    std::vector<int> res_frame;  // frame whose estimate stack each res slot holds
    std::mutex mu;

    // process_device replays its launch sequence from a HIP graph while the
    // call's arguments repeat (captured on the first call with new ones)
    struct GraphKey {
        const float *data, *ref;
        int64_t frame_stride, out_frame_stride;
        int32_t pitch, first, count, out_pitch;
        float *out;
        bool operator==(const GraphKey &o) const
        {
            return data == o.data && ref == o.ref && frame_stride == o.frame_stride &&
                   out_frame_stride == o.out_frame_stride && pitch == o.pitch && first == o.first &&
                   count == o.count && out_pitch == o.out_pitch && out == o.out;
        }
    };
    bool graphs = true;            // process_device replays its graph
    bool request_graphs = false;   // get_frame launches from per-slot graphs
    hipGraphExec_t graph_exec = nullptr;
    GraphKey graph_key{};
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;  // orders the instance stream with a null-stream caller

    // The host-buffer path (allocated on first use).  Requests overlap: a
    // request takes one of num_streams output slots (a device plane, a
    // pinned download plane, pinned staging planes, its graphs) under the
    // mutex, copies the frames the ring lacks into the slot's staging planes
    // without it (src/bm3d.cpp:772-814 stages into the stream's own mapped
    // buffer the same way), then under the mutex again enqueues its uploads
    // on copy_in, its kernels on `stream` and its download on copy_out,
    // waits for its download and copies the frame out without it.  Each
    // source ring slot has the event after the last kernels that read it,
    // which an upload into the slot waits for.  All planes have the
    // library's pitch.
    hipStream_t copy_in = nullptr, copy_out = nullptr;
    hipStream_t capture = nullptr;       // idle stream the per-request graphs are captured on
    hipEvent_t up_ev = nullptr;          // a request's uploads done (copy_in)
    std::vector<hipEvent_t> slot_ev;     // per src ring slot: after its last readers (stream)
    std::vector<char> slot_ev_set;
    struct OutSlot {
        float *dev = nullptr;            // the aggregated frame
        float *pin = nullptr;            // its pinned copy
        hipEvent_t done = nullptr;       // kernels done (stream)
        hipEvent_t down = nullptr;       // download done (copy_out)
        bool busy = false, used = false;
        // pinned staging planes, `clips` per staged frame ([ref][source] like
        // a ring slot), grown to the most frames one request has staged; the
        // slot's previous request waited for its download, which follows its
        // uploads, so the planes are free while the slot is held
        std::vector<float *> stage;
        // graph of the request's kernels by the number of estimates it runs
        // (its recompute pattern), updated in place per request: a slot's
        // graph never runs while a request holds the slot
        std::vector<hipGraphExec_t> graph;
    };
    std::vector<OutSlot> outs;
    int out_next = 0;
    std::condition_variable out_cv;
    // test hook: the next request that reaches the stage fails there
    // (kFailUpload, kFailLaunch), without touching the device
    std::atomic<int> fail_next{0};

    hipError_t pin_alloc()
    {
        if (!outs.empty()) {
            return hipSuccess;
        }
        const size_t bytes = size_t(pe) * sizeof(float);
        hipError_t e = hipStreamCreateWithFlags(&copy_in, hipStreamNonBlocking);
        if (e == hipSuccess) {
            e = hipStreamCreateWithFlags(&copy_out, hipStreamNonBlocking);
        }
        if (e == hipSuccess) {
            e = hipStreamCreateWithFlags(&capture, hipStreamNonBlocking);
        }
        if (e == hipSuccess) {
            e = hipEventCreateWithFlags(&up_ev, hipEventDisableTiming);
        }
        slot_ev.assign(size_t(src_ring), nullptr);
        slot_ev_set.assign(size_t(src_ring), 0);
        for (int k = 0; k < src_ring && e == hipSuccess; ++k) {
            e = hipEventCreateWithFlags(&slot_ev[k], hipEventDisableTiming);
        }
        std::vector<OutSlot> o(size_t(std::max(num_streams, 1)));
        for (OutSlot &q : o) {
            q.graph.assign(size_t(tw) + 1, nullptr);
            if (e == hipSuccess) {
                e = hipMalloc(&q.dev, bytes);
            }
            if (e == hipSuccess) {
                e = hipHostMalloc(reinterpret_cast<void **>(&q.pin), bytes, hipHostMallocDefault);
            }
            if (e == hipSuccess) {
                e = hipEventCreateWithFlags(&q.done, hipEventDisableTiming);
            }
            if (e == hipSuccess) {
                e = hipEventCreateWithFlags(&q.down, hipEventDisableTiming);
            }
        }
        outs.swap(o);  // on failure the destructor frees what was made
        if (e != hipSuccess) {
            free_outs();
        }
        return e;
    }

    void drop_slot_graphs(OutSlot &q)
    {
        for (hipGraphExec_t &g : q.graph) {
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        }
    }

    void free_outs()
    {
        for (OutSlot &q : outs) {
            drop_slot_graphs(q);
            (void)hipFree(q.dev);
            if (q.pin) {
                (void)hipHostFree(q.pin);
            }
            for (float *p : q.stage) {
                (void)hipHostFree(p);
            }
            if (q.done) {
                (void)hipEventDestroy(q.done);
            }
            if (q.down) {
                (void)hipEventDestroy(q.down);
            }
        }
        outs.clear();
    }

    // copies the host planes of one frame (row strides in bytes) into
    // staging planes k of slot q, growing them as needed
    hipError_t stage(OutSlot &q, int k, const float *src, ptrdiff_t stride, const float *ref, ptrdiff_t ref_stride)
    {
        const size_t need = size_t(k + 1) * size_t(clips);
        while (q.stage.size() < need) {
            float *p = nullptr;
            const hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&p), size_t(pe) * sizeof(float),
                                               hipHostMallocDefault);
            if (e != hipSuccess) {
                return e;
            }
            q.stage.push_back(p);
        }
        const size_t row_bytes = size_t(width) * sizeof(float);
        for (int c = 0; c < clips; ++c) {
            const char *from = reinterpret_cast<const char *>(c == clips - 1 ? src : ref);
            const ptrdiff_t st = c == clips - 1 ? stride : ref_stride;
            float *to = q.stage[size_t(k) * clips + c];
            for (int y = 0; y < height; ++y) {
                std::memcpy(to + (size_t)y * pitch, from + (ptrdiff_t)y * st, row_bytes);
            }
        }
        return hipSuccess;
    }

    // optional per-kernel timing: HIP events recorded around every launch on
    // the launch stream, harvested on request
    struct Timed {
        hipEvent_t a, b;
        int kind;
    };
    bool timing = false;
    std::vector<Timed> timed;
    std::vector<hipEvent_t> event_pool;
    double kernel_ms[kKernelKinds] = {};
    long long kernel_count[kKernelKinds] = {};

    hipEvent_t take_event()
    {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            return nullptr;
        }
        return e;
    }

    hipError_t harvest()
    {
        hipError_t rc = hipSuccess;
        for (const Timed &t : timed) {
            float ms = 0.0f;
            hipError_t e = hipEventSynchronize(t.b);
            if (e == hipSuccess) {
                e = hipEventElapsedTime(&ms, t.a, t.b);
            }
            if (e == hipSuccess) {
                kernel_ms[t.kind] += ms;
                kernel_count[t.kind] += 1;
            } else {
                rc = e;
            }
            event_pool.push_back(t.a);
            event_pool.push_back(t.b);
        }
        timed.clear();
        return rc;
    }

    template <class F>
    hipError_t timed_launch(int kind, hipStream_t s, F &&launch)
    {
        if (!timing) {
            return launch();
        }
        if (timed.size() >= 4096) {
            hipError_t e = harvest();
            if (e != hipSuccess) {
                return e;
            }
        }
        hipEvent_t a = take_event(), b = take_event();
        if (!a || !b) {
            return hipErrorOutOfMemory;
        }
        hipError_t e = hipEventRecord(a, s);
        if (e == hipSuccess) {
            e = launch();
        }
        if (e == hipSuccess) {
            e = hipEventRecord(b, s);
        }
        timed.push_back({ a, b, kind });
        return e;
    }

    ~Instance()
    {
        if (device < 0) {
            return;  // creation failed before any device object existed
        }
        (void)hipSetDevice(device);
        for (hipStream_t q : {stream, copy_in, copy_out}) {
            if (q) {
                (void)hipStreamSynchronize(q);
            }
        }
        for (hipStream_t q : stream_l) {
            if (q) {
                (void)hipStreamSynchronize(q);
            }
        }
        (void)hipFree(src_buf);
        (void)hipFree(res_buf);
        (void)hipFree(groups);
        (void)hipFree(dists);
        (void)hipFree(spatial);
        (void)hipFree(dst_dev);
        (void)hipFree(scratch);
        (void)hipFree(origins);
        (void)hipFree(tile_tables);
        (void)hipFree(rec);
        (void)hipFree(cell_count);
        (void)hipFree(cell_list);
        (void)hipFree(psets);
        for (int i = 0; i < kMaxLanes - 1; ++i) {
            (void)hipFree(groups_l[i]);
            (void)hipFree(scratch_l[i]);
            (void)hipFree(psets_l[i]);
            (void)hipFree(spatial_l[i]);
            (void)hipFree(origins_l[i]);
            (void)hipFree(rec_l[i]);
            (void)hipFree(cell_count_l[i]);
            (void)hipFree(cell_list_l[i]);
            (void)hipFree(dst_l[i]);
            if (lane_join[i]) {
                (void)hipEventDestroy(lane_join[i]);
            }
            if (stream_l[i]) {
                (void)hipStreamDestroy(stream_l[i]);
            }
        }
        if (lane_fork) {
            (void)hipEventDestroy(lane_fork);
        }
        for (hipEvent_t e : lane_done) {
            if (e) {
                (void)hipEventDestroy(e);
            }
        }
        for (hipEvent_t e : slot_read) {
            if (e) {
                (void)hipEventDestroy(e);
            }
        }
        if (last_done) {
            (void)hipEventDestroy(last_done);
        }
        if (graph_exec) {
            (void)hipGraphExecDestroy(graph_exec);
        }
        free_outs();
        for (hipEvent_t e : slot_ev) {
            if (e) {
                (void)hipEventDestroy(e);
            }
        }
        if (up_ev) {
            (void)hipEventDestroy(up_ev);
        }
        for (hipStream_t q : {copy_in, copy_out, capture}) {
            if (q) {
                (void)hipStreamDestroy(q);
            }
        }
        if (fork_ev) {
            (void)hipEventDestroy(fork_ev);
        }
        if (join_ev) {
            (void)hipEventDestroy(join_ev);
        }
        (void)harvest();
        for (hipEvent_t e : event_pool) {
            (void)hipEventDestroy(e);
        }
        if (stream) {
            (void)hipStreamDestroy(stream);
        }
    }

    float *src_slot(int slot, int clip) const { return src_buf + ((long long)slot * clips + clip) * pe; }
    float *res_slot(int slot) const { return res_buf + (long long)slot * tw * 2 * pe; }
    uint32_t *lane_groups() const { return lane ? groups_l[lane - 1] : groups; }
    float *lane_scratch() const { return lane ? scratch_l[lane - 1] : scratch; }
    uint16_t *lane_psets() const { return lane ? psets_l[lane - 1] : psets; }
    uint2 *lane_spatial() const { return lane ? spatial_l[lane - 1] : spatial; }
    float *lane_origins() const { return lane ? origins_l[lane - 1] : origins; }
    float *lane_rec() const { return lane ? rec_l[lane - 1] : rec; }
    uint32_t *lane_cell_count() const { return lane ? cell_count_l[lane - 1] : cell_count; }
    uint2 *lane_cell_list() const { return lane ? cell_list_l[lane - 1] : cell_list; }
    float *lane_dst() const { return lane ? dst_l[lane - 1] : dst_dev; }

    // ------------------------------------------------------------------
    // one estimation dispatch (record_bm3d_kernels, src/bm3d.cpp:558-602):
    // clear the stack, block matching, collaborative filtering
    // ------------------------------------------------------------------
    MatchArgs match_args(const float *const *search, int in_pitch, float *dists_out) const
    {
        MatchArgs m {};
        for (int z = 0; z < tw; ++z) {
            m.plane[z] = search[z];
        }
        m.groups = lane_groups();
        m.dists = dists_out;
        m.width = width;
        m.height = height;
        m.pitch = in_pitch;
        m.step = block_step;
        m.bm_range = bm_range;
        m.radius = radius;
        m.ps_num = ps_num;
        m.ps_range = ps_range;
        m.nx = nx;
        m.ny = ny;
        m.spatial = lane_spatial();
        m.run_blocks = run_blocks;
        m.runs_x = runs_x;
        m.region_cols = region_cols;
        m.region_rows = region_rows;
        m.dist_cols = dist_cols;
        m.t_rows = t_rows;
        m.pred_pair = pred_pair;
        m.operand = operand;
        if (operand != kOperandFp32) {
            m.pset = lane_psets();
            m.pset_py = plane_py(height);
            m.pset_elems = (long long)width * m.pset_py;
            m.mfma_run_blocks = mfma_run_blocks;
            m.mfma_runs_x = mfma_runs_x;
        }
        return m;
    }

    // block matching of one estimate: in the precision form first the plane
    // set of the reference frame (the spatial search's operands), then the
    // search
    hipError_t match(const MatchArgs &m, const float *const *search, int in_pitch, hipStream_t s)
    {
        if (operand != kOperandFp32) {
            const hipError_t e = timed_launch(kKernelPlanes, s, [&] {
                return launch_planes(search[radius], in_pitch, width, height, lane_psets(), operand, s);
            });
            if (e != hipSuccess) {
                return e;
            }
        }
        hipError_t e = timed_launch(kKernelMatch, s, [&] { return launch_match(m, fp_mode, s, false); });
        if (e == hipSuccess && match_two_pass(m)) {
            e = timed_launch(kKernelTemporal, s, [&] { return launch_match(m, fp_mode, s, true); });
        }
        return e;
    }

    using Saved = SavedState;

    // mode kFilterPlain / kFilterSave: block matching into the group table
    // (saved.groups when given), then filtering; kFilterApply: no matching,
    // the saved groups, masks and weights are applied to `input`
    // fused_dst (radius 0, tiled, plain mode): the tile reduction writes the
    // aggregated frame there instead of the slice; fuses() says when it does
    bool fuses() const { return tiled && radius == 0; }
    hipError_t estimate(const float *const *search, const float *const *input, int in_pitch,
                        float *stack, int stack_pitch, hipStream_t s, float *dists_out = nullptr,
                        int mode = kFilterPlain, const Saved &saved = Saved(), float *fused_dst = nullptr,
                        int fused_pitch = 0)
    {
        float *const origins = lane_origins();
        float *const rec = lane_rec();
        uint32_t *const cell_count = lane_cell_count();
        uint2 *const cell_list = lane_cell_list();
        hipError_t e = hipSuccess;
        if (!tiled || radius > 0 || short_windows) {
            // the tile reduction overwrites the frame's own slice; the other
            // slices are accumulated with atomics and must start at zero
            e = timed_launch(kKernelFill, s, [&] {
                return launch_fill_zero(stack, (long long)tw * 2 * height * stack_pitch, s);
            });
            if (e != hipSuccess) {
                return e;
            }
        }
        if (origins) {
            e = timed_launch(kKernelFill, s, [&] {
                return launch_fill_zero(origins, 2LL * radius * height * pitch, s);
            });
            if (e != hipSuccess) {
                return e;
            }
        }
        if (rec) {
            e = timed_launch(kKernelFill, s, [&] {
                return launch_fill_zero(reinterpret_cast<float *>(cell_count), cell_count_elems, s);
            });
            if (e != hipSuccess) {
                return e;
            }
        }
        uint32_t *table = saved.groups ? saved.groups : lane_groups();
        if (mode != kFilterApply) {
            MatchArgs m = match_args(search, in_pitch, dists_out);
            m.groups = table;
            e = match(m, search, in_pitch, s);
            if (e != hipSuccess) {
                return e;
            }
        }
        EstimateArgs a {};
        for (int z = 0; z < tw; ++z) {
            a.search[z] = search[z];
            a.input[z] = input[z];
        }
        a.groups = table;
        a.masks = saved.masks;
        a.weights = saved.weights;
        a.res = stack;
        a.height = height;
        a.pitch = in_pitch;
        a.res_pitch = stack_pitch;
        a.ngroups = nref;
        a.sigma = sigma;
        a.extractor = extractor;
        a.width = width;
        a.radius = radius;
        a.nx = nx;
        a.ny = ny;
        a.step = block_step;
        a.halo = bm_range;
        a.origins = origins;
        a.origin_pitch = pitch;
        if (tiled) {
            a.scratch = lane_scratch();
            a.tile_bx = tile_bx;
            a.tile_by = tile_by;
            a.tiles_x = tiles_x;
            a.tiles_y = tiles_y;
            a.region_w = region_w;
            a.region_h = region_h;
            a.rec = rec;
            a.cell_count = cell_count;
            a.cell_list = cell_list;
            a.rec_per_tile = rec_per_tile;
            a.cells_x = cells_x;
            a.ncells = ncells;
            a.cell_cap = cell_cap;
            a.cwin_x = cwin_x;
            a.cwin_y = cwin_y;
            a.bin_halo = bin_halo;
        }
        e = timed_launch(kKernelEstimate, s, [&] { return launch_estimate(a, fp_mode, final_pass, mode, s); });
        if (e == hipSuccess && rec && tiled) {
            GatherArgs ga {};
            ga.rec = rec;
            ga.cell_count = cell_count;
            ga.cell_list = cell_list;
            ga.res = stack;
            ga.width = width;
            ga.height = height;
            ga.res_pitch = stack_pitch;
            ga.cells_x = cells_x;
            ga.ncells = ncells;
            ga.cell_cap = cell_cap;
            ga.radius = radius;
            e = timed_launch(kKernelGather, s, [&] { return launch_gather(ga, s); });
        }
        if (e == hipSuccess && origins) {
            BoxArgs b {};
            b.origins = origins;
            b.res = stack;
            b.width = width;
            b.height = height;
            b.origin_pitch = pitch;
            b.res_pitch = stack_pitch;
            b.radius = radius;
            e = timed_launch(kKernelWeightBox, s, [&] { return launch_weight_box(b, s); });
        }
        if (e != hipSuccess || !tiled) {
            return e;
        }
        ReduceArgs ra {};
        ra.scratch = lane_scratch();
        ra.res = stack + (long long)radius * 2 * height * stack_pitch;
        ra.col_tiles = reinterpret_cast<const int2 *>(tile_tables);
        ra.row_tiles = reinterpret_cast<const int2 *>(tile_tables + 2 * width);
        ra.tile_x0 = tile_tables + 2 * width + 2 * height;
        ra.tile_y0 = ra.tile_x0 + tiles_x;
        ra.width = width;
        ra.height = height;
        ra.res_pitch = stack_pitch;
        ra.tiles_x = tiles_x;
        ra.region_w = region_w;
        ra.region_h = region_h;
        ra.add_existing = short_windows ? 1 : 0;
        ra.span3 = reduce_span3 ? 1 : 0;
        if (fused_dst && fuses()) {
            ra.dst = fused_dst;
            ra.dst_pitch = fused_pitch;
        }
        return timed_launch(kKernelReduce, s, [&] { return launch_tile_reduce(ra, s); });
    }

    // aggregation of frame n from the res slots (record_bm3d_agg, src/bm3d.cpp:628-696)
    hipError_t aggregate(int n, const std::vector<int> &slot_of, float *dst, int dst_pitch, hipStream_t s,
                         float *denom = nullptr, int denom_pitch = 0, bool numerator_only = false)
    {
        AggArgs a {};
        a.denom = denom;
        a.denom_pitch = denom_pitch;
        a.numerator_only = numerator_only ? 1 : 0;
        for (int i = 0; i < tw; ++i) {
            const int z = radius == 0 ? 0 : agg_z(i, n, nframes, radius);
            a.slice[i] = res_slot(slot_of[i]) + (long long)z * 2 * pe;
        }
        a.dst = dst;
        a.plane_elems = pe;
        a.width = width;
        a.height = height;
        a.pitch = pitch;
        a.dst_pitch = dst_pitch;
        a.tw = tw;
        return timed_launch(kKernelAggregate, s, [&] { return launch_aggregate(a, s); });
    }

    // Frames of a radius-0 batch (forward_saved, backward) are independent:
    // frame i runs on lane i mod n with the lane's own result slot (slot i
    // mod n), buffers and stream; side streams are forked from the launch
    // stream when first used and joined back by finish().  Under per-kernel
    // timing everything stays on the launch stream.
    struct FrameLanes {
        Instance &d;
        hipStream_t q;
        int n = 1;
        bool forked[kMaxLanes] = {};
        FrameLanes(Instance &inst, hipStream_t s) : d(inst), q(s)
        {
            n = d.timing || d.radius != 0 ? 1 : std::min(d.lanes, d.res_cap);
        }
        ~FrameLanes() { d.lane = 0; }
        hipError_t start() { return n > 1 ? hipEventRecord(d.lane_fork, q) : hipSuccess; }
        // selects frame i's lane; its stream in *qs
        hipError_t select(int i, hipStream_t *qs)
        {
            d.lane = i % n;
            *qs = d.lane ? d.stream_l[d.lane - 1] : q;
            if (d.lane && !forked[d.lane]) {
                forked[d.lane] = true;
                return hipStreamWaitEvent(*qs, d.lane_fork, 0);
            }
            return hipSuccess;
        }
        hipError_t finish()
        {
            d.lane = 0;
            hipError_t e = hipSuccess;
            for (int j = 1; j < n; ++j) {
                if (forked[j]) {
                    hipError_t ej = hipEventRecord(d.lane_join[j - 1], d.stream_l[j - 1]);
                    if (ej == hipSuccess) {
                        ej = hipStreamWaitEvent(q, d.lane_join[j - 1], 0);
                    }
                    if (e == hipSuccess) {
                        e = ej;
                    }
                }
            }
            return e;
        }
    };

    // order this call's device work after the previous call's
    hipError_t begin_on(hipStream_t s)
    {
        if (have_last) {
            return hipStreamWaitEvent(s, last_done, 0);
        }
        return hipSuccess;
    }

    hipError_t end_on(hipStream_t s)
    {
        hipError_t e = hipEventRecord(last_done, s);
        have_last = e == hipSuccess;
        return e;
    }
};

namespace {

void set_err(char *err, size_t len, const std::string &msg)
{
    if (err && len > 0) {
        std::snprintf(err, len, "%s", msg.c_str());
    }
}

int fail(char *err, size_t len, int code, const std::string &msg)
{
    set_err(err, len, "BM3D: " + msg);
    return code;
}

int hip_fail(char *err, size_t len, hipError_t e, const char *what)
{
    return fail(err, len, VSFEEL_BM3D_EDEVICE, std::string(what) + " failed: " + hipGetErrorString(e));
}

#define VSFEEL_HIP(call, what)                                   \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) {                                  \
            return hip_fail(err, err_len, e_, what);             \
        }                                                        \
    } while (0)

// BM3DCreate's parameter parsing (src/bm3d.cpp:1062-1192) and plane setup
// (:1256-1275); returns an empty string or the message without the prefix
std::string parse_args(const vsfeel_bm3d_args &a, Instance &d)
{
    const vsfeel_bm3d_format &f = a.clip;
    if (a.has_ref) {
        const vsfeel_bm3d_format &r = a.ref;
        if (r.color_family != f.color_family || r.sample_type != f.sample_type ||
            r.bits_per_sample != f.bits_per_sample || r.sub_sampling_w != f.sub_sampling_w ||
            r.sub_sampling_h != f.sub_sampling_h) {
            return "\"ref\" must be of the same format as \"clip\"";
        }
        if (r.width != f.width || r.height != f.height) {
            return "\"ref\" must be of the same dimensions as \"clip\"";
        }
        if (r.num_frames != f.num_frames) {
            return "\"ref\" must be of the same number of frames as \"clip\"";
        }
        d.final_pass = true;
    }
    if (f.width <= 0 || f.height <= 0 || f.sample_type != 1 || f.bits_per_sample != 32) {
        return "only constant format 32 bit float input supported";
    }

    float sigma[3];
    for (int i = 0; i < 3; ++i) {
        if (i < a.num_sigma && a.sigma) {
            sigma[i] = float(a.sigma[i]);
            if (sigma[i] < 0.0f) {
                return "\"sigma\" must be non-negative";
            }
        } else {
            sigma[i] = i == 0 ? 3.0f : sigma[i - 1];
        }
    }
    for (int i = 0; i < 3; ++i) {
        d.process = sigma[i] >= FLT_EPSILON;
    }
    const float factor = d.final_pass ? bits_to_float(0x3e40c0c1u) : bits_to_float(0x3f021bb6u);
    for (float &s : sigma) {
        s *= factor;
    }
    d.sigma = sigma[0];
    d.sigma_u = sigma[1];
    d.sigma_v = sigma[2];

    auto int_array = [](const int64_t *v, int n, int def, int *out, auto bad, const char *msg) -> std::string {
        for (int i = 0; i < 3; ++i) {
            if (i < n && v) {
                out[i] = sat_int(v[i]);
                if (bad(out[i])) {
                    return msg;
                }
            } else {
                out[i] = i == 0 ? def : out[i - 1];
            }
        }
        return "";
    };
    int tmp[3];
    std::string e;
    e = int_array(a.block_step, a.num_block_step, 8, tmp, [](int v) { return v <= 0 || v > 8; },
                  "\"block_step\" must be in range [1, 8]");
    if (!e.empty()) {
        return e;
    }
    d.block_step = tmp[0];
    e = int_array(a.bm_range, a.num_bm_range, 9, tmp, [](int v) { return v <= 0; },
                  "\"bm_range\" must be positive");
    if (!e.empty()) {
        return e;
    }
    d.bm_range = tmp[0];
    d.radius = (a.num_radius > 0 && a.radius) ? sat_int(a.radius[0]) : 0;
    if (d.radius < 0 || d.radius > 4) {
        return "\"radius\" must be in range [0, 4]";
    }
    d.tw = 2 * d.radius + 1;
    e = int_array(a.ps_num, a.num_ps_num, 2, tmp, [](int v) { return v <= 0 || v > 8; },
                  "\"ps_num\" must be in range [1, 8]");
    if (!e.empty()) {
        return e;
    }
    d.ps_num = tmp[0];
    e = int_array(a.ps_range, a.num_ps_range, 4, tmp, [](int v) { return v <= 0; },
                  "\"ps_range\" must be positive");
    if (!e.empty()) {
        return e;
    }
    d.ps_range = tmp[0];
    d.num_streams = (a.num_num_streams > 0 && a.num_streams) ? sat_int(a.num_streams[0]) : 4;
    if (d.num_streams <= 0) {
        return "\"num_streams\" must be positive";
    }
    // cache capacities of src/bm3d.cpp:1185-1186
    d.src_ring = d.radius == 0 ? d.num_streams : 4 * d.radius + d.num_streams;
    d.res_cap = d.radius == 0 ? d.num_streams : d.tw + d.num_streams + 2 * d.radius;
    const int ex = (a.num_extractor_exp > 0 && a.extractor_exp) ? sat_int(a.extractor_exp[0]) : 0;
    d.extractor = ex != 0 ? std::ldexp(1.0f, ex) : 0.0f;
    d.nframes = f.num_frames;

    if (f.color_family == 1) {
        d.chroma = false;
    } else if (f.color_family == 3) {
        d.chroma = true;
    } else {
        // the reference's message carries the prefix twice (src/bm3d.cpp:1274)
        return "BM3D: only Gray and YUV input are currently supported";
    }
    // not checked by the reference, whose shader then reads outside the plane
    if (f.width < 8 || f.height < 8) {
        return "the luma plane must be at least 8x8";
    }
    if (f.width > kMaxDim || f.height > kMaxDim) {
        return "the luma plane must be at most 16383x16383";
    }
    if (f.num_frames <= 0) {
        return "the clip must have a known, positive number of frames";
    }
    d.fp_mode = a.fp_mode;
    if (d.fp_mode < 0 || d.fp_mode > 2) {
        return "\"fp_mode\" must be 0, 1 or 2";
    }
    d.operand = a.precision;
    if (d.operand < kOperandFp32 || d.operand > kOperandBf16) {
        return "\"precision\" must be 0 (FP32), 1 (FP16 MFMA) or 2 (BF16 MFMA)";
    }
    if (d.operand != kOperandFp32 && d.radius >= 3) {
        return "\"precision\": the MFMA distances serve the sliding metric of radius 0-2";
    }
    d.device = a.device;
    d.width = f.width;
    d.height = f.height;
    d.clips = d.final_pass ? 2 : 1;
    d.nx = (d.width - 8 + d.block_step + d.block_step - 1) / d.block_step;
    d.ny = (d.height - 8 + d.block_step + d.block_step - 1) / d.block_step;
    d.nref = d.nx * d.ny;
    d.pitch = (d.width + 63) & ~63;
    d.pe = (long long)d.height * d.pitch;
    {
        // the smallest spatial window sits at a corner of the frame
        const int rw = std::min(d.bm_range, d.width - 8) + 1;
        const int rh = std::min(d.bm_range, d.height - 8) + 1;
        d.short_windows = (long long)rw * rh < 8;
    }
    if (d.radius < 3) {
        // one window row per thread of a 256-thread workgroup; the source
        // region, the blocks and their distance rows must fit 64 KB of LDS
        const int rh = std::min(2 * d.bm_range + 1, d.height - 7);
        // (the 16-lane selection keeps a 64-bit mask of the window rows)
        const int nb = rh <= 64 ? 256 / rh : 0;
        if (nb > 0) {
            const int cols = std::min((nb - 1) * d.block_step + 8 + 2 * d.bm_range, d.width);
            const int dcols = std::min(2 * d.bm_range + 1, d.width - 7);
            const long long bytes = rows_lds_bytes(cols, rh + 7, nb, dcols);
            if (bytes <= 64 * 1024) {
                d.run_blocks = nb;
                d.runs_x = (d.nx + nb - 1) / nb;
                d.region_cols = cols;
                d.region_rows = rh + 7;
                d.dist_cols = dcols;
            }
        }
        if (d.run_blocks > 0 && d.radius > 0) {
            // the predictive search's rows of one step, one per lane
            const int wh = 2 * d.ps_range + 1;
            const int rows = d.ps_num * wh;
            if (rows <= 64 && temporal_lds_bytes(rows, d.ps_range) <= 48 * 1024) {
                d.t_rows = rows;
            }
            // both directions in one wavefront: a window row per lane of a
            // half, and a window column per lane of a half while staging
            // (2 ps_range + 8 columns)
            if (d.radius <= 2 && rows <= 32 && 2 * d.ps_range + 8 <= 32 &&
                predictive_lds_bytes(d.ps_num, d.ps_range) <= 64 * 1024) {
                d.pred_pair = 1;
            }
        }
    }
    if (d.operand != kOperandFp32) {
        // precision form: runs of up to 8 ref blocks whose region fits 64 KB
        // of LDS; the window's positions must fit the 16-bit seq of its keys
        const int rh = std::min(2 * d.bm_range + 1, d.height - 7);
        const int rw = std::min(2 * d.bm_range + 1, d.width - 7);
        for (int nb = 8; nb >= 1 && d.mfma_run_blocks == 0; --nb) {
            const int cols = std::min((nb - 1) * d.block_step + 8 + 2 * d.bm_range, d.width);
            if (mfma_rows_lds_bytes(cols, rh) <= 64 * 1024) {
                d.mfma_run_blocks = nb;
                d.mfma_runs_x = (d.nx + nb - 1) / nb;
            }
        }
        if (d.mfma_run_blocks == 0 || d.run_blocks == 0 || (long long)rw * rh >= 65536) {
            return "\"precision\": bm_range too large for the MFMA search";
        }
    }
    d.tiled = choose_tiles(d.width, d.height, d.block_step, d.bm_range, &d.tile_bx, &d.tile_by, &d.region_w,
                           &d.region_h);
    if (d.tiled) {
        d.tiles_x = (d.nx + d.tile_bx - 1) / d.tile_bx;
        d.tiles_y = (d.ny + d.tile_by - 1) / d.tile_by;
    }
    return "";
}

}  // namespace
}  // namespace vsfeel

struct vsfeel_bm3d {
    vsfeel::Instance inst;
};

using vsfeel::Instance;

// Device memory of vsfeel_bm3d_process_device_static (the entry point that
// allocates nothing): its group table, the tile reduction's tables and the
// tile accumulators.  Each GPU the library runs on holds its own copy until
// the process ends; every call on that GPU reuses it.  Host code uses the
// per-device address from hipGetSymbolAddress, never this name.
static constexpr size_t kStaticScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kStaticMaxDevices = 64;
std::mutex g_static_mu;
unsigned char *g_static_ptr[kStaticMaxDevices] = {};

// the device address of `scratch` on the current device `dev`, looked up once
hipError_t static_scratch(int dev, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_static_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            return e;
        }
        g_static_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_static_ptr[dev];
    return hipSuccess;
}

// the most tiles covering one pixel column (or row) of a plane of `size`
// pixels cut into `tiles` tiles of `tile_blocks` ref blocks
int max_tile_span(int size, int tiles, int tile_blocks, int nblocks, int step, int halo)
{
    int span = 0;
    for (int p = 0; p < size; ++p) {
        int first = -1, last = -2;
        for (int t = 0; t < tiles; ++t) {
            int lo, hi;
            vsfeel::tile_span(t, tile_blocks, nblocks, step, size, halo, &lo, &hi);
            if (lo <= p && p <= hi) {
                first = first < 0 ? t : first;
                last = t;
            }
        }
        span = std::max(span, last - first + 1);
    }
    return span;
}

size_t round256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

extern "C" {

int vsfeel_bm3d_create(const vsfeel_bm3d_args *args, vsfeel_bm3d **out, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (!out || !args) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null argument");
    }
    *out = nullptr;
    auto *h = new vsfeel_bm3d();
    Instance &d = h->inst;
    d.device = -1;
    const std::string msg = parse_args(*args, d);
    if (!msg.empty()) {
        delete h;
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, msg);
    }
    auto dev_fail = [&](hipError_t e, const char *what) {
        const int rc = hip_fail(err, err_len, e, what);
        delete h;
        return rc;
    };
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= d.device || d.device < 0) {
        const int rc = fail(err, err_len, VSFEEL_BM3D_EDEVICE,
                            e != hipSuccess ? std::string("no HIP device: ") + hipGetErrorString(e)
                                            : std::string("HIP device ") + std::to_string(d.device) +
                                                  " is not available");
        d.device = -1;
        delete h;
        return rc;
    }
    if ((e = hipSetDevice(d.device)) != hipSuccess) {
        return dev_fail(e, "hipSetDevice");
    }
    if ((e = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking)) != hipSuccess) {
        return dev_fail(e, "hipStreamCreate");
    }
    if ((e = hipEventCreateWithFlags(&d.last_done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&d.fork_ev, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&d.join_ev, hipEventDisableTiming)) != hipSuccess) {
        return dev_fail(e, "hipEventCreate");
    }
    const size_t plane = size_t(d.pe) * sizeof(float);
    if ((e = hipMalloc(&d.src_buf, plane * d.src_ring * d.clips)) != hipSuccess ||
        (e = hipMalloc(&d.res_buf, plane * d.res_cap * d.tw * 2)) != hipSuccess ||
        (e = hipMalloc(&d.groups, size_t(d.nref) * 8 * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&d.dists, size_t(d.nref) * 8 * sizeof(float))) != hipSuccess ||
        (d.radius > 0 && (e = hipMalloc(&d.spatial, size_t(d.nref) * 8 * sizeof(uint2))) != hipSuccess) ||
        (d.radius > 0 && (e = hipMalloc(&d.origins, plane * 2 * d.radius)) != hipSuccess) ||
        (e = hipMalloc(&d.dst_dev, plane)) != hipSuccess) {
        return dev_fail(e, "hipMalloc");
    }
    if (d.operand != kOperandFp32) {
        d.pset_bytes = (plane_set_bytes(d.width, d.height) + 255) & ~255LL;
        if ((e = hipMalloc(&d.psets, size_t(d.pset_bytes))) != hipSuccess) {
            return dev_fail(e, "hipMalloc");
        }
    }
    if (d.tiled) {
        // which tiles cover each column / row, and every tile's region origin
        std::vector<int> tab(size_t(2) * d.width + 2 * d.height + d.tiles_x + d.tiles_y);
        std::vector<int> x_lo(d.tiles_x), x_hi(d.tiles_x), y_lo(d.tiles_y), y_hi(d.tiles_y);
        for (int t = 0; t < d.tiles_x; ++t) {
            tile_span(t, d.tile_bx, d.nx, d.block_step, d.width, d.bm_range, &x_lo[t], &x_hi[t]);
        }
        for (int t = 0; t < d.tiles_y; ++t) {
            tile_span(t, d.tile_by, d.ny, d.block_step, d.height, d.bm_range, &y_lo[t], &y_hi[t]);
        }
        auto ranges = [](int n, const std::vector<int> &lo, const std::vector<int> &hi, int *out) {
            for (int p = 0; p < n; ++p) {
                int first = -1, last = -2;
                for (size_t t = 0; t < lo.size(); ++t) {
                    if (lo[t] <= p && p <= hi[t]) {
                        if (first < 0) {
                            first = int(t);
                        }
                        last = int(t);
                    }
                }
                out[2 * p] = first < 0 ? 0 : first;
                out[2 * p + 1] = last;  // an uncovered pixel gets an empty range
            }
        };
        ranges(d.width, x_lo, x_hi, tab.data());
        ranges(d.height, y_lo, y_hi, tab.data() + 2 * d.width);
        std::copy(x_lo.begin(), x_lo.end(), tab.begin() + 2 * d.width + 2 * d.height);
        std::copy(y_lo.begin(), y_lo.end(), tab.begin() + 2 * d.width + 2 * d.height + d.tiles_x);
        const size_t scratch_elems = size_t(d.tiles_x) * d.tiles_y * 2 * d.region_h * d.region_w;
        int span_x = 0, span_y = 0;
        for (int p = 0; p < d.width; ++p) {
            span_x = std::max(span_x, tab[2 * p + 1] - tab[2 * p] + 1);
        }
        for (int p = 0; p < d.height; ++p) {
            span_y = std::max(span_y, tab[2 * d.width + 2 * p + 1] - tab[2 * d.width + 2 * p] + 1);
        }
        d.reduce_span3 = span_x <= 3 && span_y <= 3 && scratch_elems < (size_t(1) << 32);
        if ((e = hipMalloc(&d.scratch, scratch_elems * sizeof(float))) != hipSuccess ||
            (e = hipMalloc(&d.tile_tables, tab.size() * sizeof(int))) != hipSuccess) {
            return dev_fail(e, "hipMalloc");
        }
        if ((e = hipMemcpy(d.tile_tables, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice)) !=
            hipSuccess) {
            return dev_fail(e, "hipMemcpy");
        }
        if (d.radius > 0 && 2 * d.radius * d.ps_num >= 8) {
            // binning pays where a group takes at least 8 temporal
            // insertions (at 4K, radius 1, ps_num 2 the gather cost more
            // than the atomics it replaced); a group holds at most 7
            // other-frame blocks (the self block is always a member);
            // binning is skipped where its records would take more than 4 GiB
            const long long tiles = (long long)d.tiles_x * d.tiles_y;
            const long long per_tile = (long long)d.tile_bx * d.tile_by * 7;
            const long long rec_bytes = tiles * per_tile * 256;
            const int cw = 1 << kCellShift;
            const int cx = (d.width + cw - 1) / cw, cy = (d.height + cw - 1) / cw;
            // a workgroup's window: its blocks' pixels and the temporal
            // search's reach beyond them (spatial range plus ps_range per step)
            const int halo = d.bm_range + d.radius * d.ps_range;
            const int wx = ((d.tile_bx - 1) * d.block_step + 2 * halo + 7) / cw + 2;
            const int wy = ((d.tile_by - 1) * d.block_step + 2 * halo + 7) / cw + 2;
            const size_t lds = size_t(2) * d.region_h * d.region_w * 8 + bin_lds_bytes(int(per_tile), d.radius, wx, wy);
            if (rec_bytes <= (4LL << 30) && tiles * per_tile < (1LL << 32) && lds + 20 * 1024 <= 160 * 1024) {
                d.cwin_x = wx;
                d.cwin_y = wy;
                d.bin_halo = halo;
                d.rec_per_tile = int(per_tile);
                d.cells_x = cx;
                d.ncells = cx * cy;
                d.cell_cap = 512;
                if (const char *cap = std::getenv("VSFEEL_BM3D_CELL_CAP")) {
                    // (tests: a small capacity sends most blocks down the atomic fallback)
                    d.cell_cap = std::min(std::max(std::atoi(cap), 1), 4096);
                }
                d.cell_count_elems = (2LL * d.radius * d.ncells + 63) / 64 * 64;
                if ((e = hipMalloc(&d.rec, size_t(tiles * per_tile) * 64 * sizeof(float))) != hipSuccess ||
                    (e = hipMalloc(&d.cell_count, size_t(d.cell_count_elems) * sizeof(uint32_t))) != hipSuccess ||
                    (e = hipMalloc(&d.cell_list, size_t(2LL * d.radius * d.ncells * d.cell_cap) * sizeof(uint2))) !=
                        hipSuccess) {
                    return dev_fail(e, "hipMalloc");
                }
            }
        }
        // VSFEEL_BM3D_LANES caps the lanes (1: one stream and no extra
        // buffer sets, e.g. for per-kernel profiles of kernels running alone)
        int max_lanes = Instance::kMaxLanes;
        if (const char *l = std::getenv("VSFEEL_BM3D_LANES")) {
            max_lanes = std::min(std::max(std::atoi(l), 1), int(Instance::kMaxLanes));
        }
        if (!d.short_windows && max_lanes > 1) {
            // the lanes' own buffer sets (lane 0 has the instance's)
            d.lanes = std::min(d.radius == 0 ? Instance::kLanes : Instance::kTemporalLanes, max_lanes);
            // at radius > 0 the replayed graph overlapped the lanes' cross-
            // linked branches less than direct launches do (980 against 1024
            // frames/s at 1080p radius 2): launch directly by default
            d.graphs = d.radius == 0;
            const size_t rec_bytes = size_t((long long)d.tiles_x * d.tiles_y * d.rec_per_tile) * 64 * sizeof(float);
            if ((e = hipEventCreateWithFlags(&d.lane_fork, hipEventDisableTiming)) != hipSuccess) {
                return dev_fail(e, "hipEventCreate");
            }
            // side streams: lanes 1.. at radius 0, every lane at radius > 0
            for (int i = 0; i < (d.radius == 0 ? d.lanes - 1 : d.lanes); ++i) {
                if ((e = hipStreamCreateWithFlags(&d.stream_l[i], hipStreamNonBlocking)) != hipSuccess ||
                    (e = hipEventCreateWithFlags(&d.lane_join[i], hipEventDisableTiming)) != hipSuccess) {
                    return dev_fail(e, "hipStreamCreate");
                }
            }
            for (int i = 0; i < d.lanes; ++i) {
                if ((e = hipEventCreateWithFlags(&d.lane_done[i], hipEventDisableTiming)) != hipSuccess) {
                    return dev_fail(e, "hipEventCreate");
                }
            }
            if (d.radius > 0) {
                d.slot_read.assign(d.res_cap, nullptr);
                for (hipEvent_t &ev : d.slot_read) {
                    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
                        return dev_fail(e, "hipEventCreate");
                    }
                }
            }
            for (int i = 0; i < d.lanes - 1; ++i) {
                if ((e = hipMalloc(&d.groups_l[i], size_t(d.nref) * 8 * sizeof(uint32_t))) != hipSuccess ||
                    (e = hipMalloc(&d.scratch_l[i], scratch_elems * sizeof(float))) != hipSuccess ||
                    (d.operand != kOperandFp32 &&
                     (e = hipMalloc(&d.psets_l[i], size_t(d.pset_bytes))) != hipSuccess) ||
                    (d.spatial && (e = hipMalloc(&d.spatial_l[i], size_t(d.nref) * 8 * sizeof(uint2))) != hipSuccess) ||
                    (d.origins && (e = hipMalloc(&d.origins_l[i], plane * 2 * d.radius)) != hipSuccess) ||
                    (d.rec && (e = hipMalloc(&d.rec_l[i], rec_bytes)) != hipSuccess) ||
                    (d.rec && (e = hipMalloc(&d.cell_count_l[i], size_t(d.cell_count_elems) * sizeof(uint32_t))) !=
                                  hipSuccess) ||
                    (d.rec && (e = hipMalloc(&d.cell_list_l[i],
                                             size_t(2LL * d.radius * d.ncells * d.cell_cap) * sizeof(uint2))) !=
                                  hipSuccess) ||
                    (d.radius == 0 && (e = hipMalloc(&d.dst_l[i], plane)) != hipSuccess)) {
                    return dev_fail(e, "hipMalloc");
                }
            }
        }
    }
    d.src_frame.assign(d.src_ring, -1);
    d.res_frame.assign(d.res_cap, -1);
    *out = h;
    return VSFEEL_BM3D_OK;
}

void vsfeel_bm3d_free(vsfeel_bm3d *h) { delete h; }

int vsfeel_bm3d_get_info(const vsfeel_bm3d *h, vsfeel_bm3d_info *info)
{
    if (!h || !info) {
        return VSFEEL_BM3D_EINVAL;
    }
    const Instance &d = h->inst;
    info->width = d.width;
    info->height = d.height;
    info->num_frames = d.nframes;
    info->chroma = d.chroma;
    info->final_pass = d.final_pass;
    info->sigma = d.sigma;
    info->extractor = d.extractor;
    info->block_step = d.block_step;
    info->bm_range = d.bm_range;
    info->radius = d.radius;
    info->ps_num = d.ps_num;
    info->ps_range = d.ps_range;
    info->num_streams = d.num_streams;
    info->temporal_width = d.tw;
    info->ref_blocks_x = d.nx;
    info->ref_blocks_y = d.ny;
    info->fp_mode = d.fp_mode;
    info->device = d.device;
    info->pitch = d.pitch;
    info->precision = d.operand;
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_frame_window(const vsfeel_bm3d *h, int32_t n, int32_t *first, int32_t *last)
{
    if (!h || !first || !last) {
        return VSFEEL_BM3D_EINVAL;
    }
    const Instance &d = h->inst;
    if (n < 0 || n >= d.nframes) {
        return VSFEEL_BM3D_ERANGE;
    }
    *first = vsfeel::clampi(n - 2 * d.radius, 0, d.nframes - 1);
    *last = vsfeel::clampi(n + 2 * d.radius, 0, d.nframes - 1);
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_get_frame(vsfeel_bm3d *h, int32_t n, const float *const *clip_planes,
                          const ptrdiff_t *clip_strides, const float *const *ref_planes,
                          const ptrdiff_t *ref_strides, float *dst, ptrdiff_t dst_stride, char *err,
                          size_t err_len)
{
    using namespace vsfeel;
    if (!h) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null handle");
    }
    Instance &d = h->inst;
    if (n < 0 || n >= d.nframes) {
        return fail(err, err_len, VSFEEL_BM3D_ERANGE, "frame " + std::to_string(n) + " is out of range");
    }
    if (!clip_planes || !clip_strides || !dst || (d.final_pass && (!ref_planes || !ref_strides))) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "missing frame planes");
    }
    std::unique_lock<std::mutex> lock(d.mu);
    VSFEEL_HIP(hipSetDevice(d.device), "hipSetDevice");
    VSFEEL_HIP(d.pin_alloc(), "hipHostMalloc");
    // an output slot: the request owns it until its frame is copied out
    d.out_cv.wait(lock, [&] { return !d.outs[size_t(d.out_next)].busy; });
    const int j = d.out_next;
    d.out_next = (d.out_next + 1) % int(d.outs.size());
    Instance::OutSlot &o = d.outs[size_t(j)];
    o.busy = true;
    struct Release {
        Instance &d;
        Instance::OutSlot &o;
        std::unique_lock<std::mutex> &lock;
        ~Release()
        {
            if (!lock.owns_lock()) {
                lock.lock();
            }
            o.busy = false;
            d.out_cv.notify_all();
        }
    } release{d, o, lock};

    const int r = d.radius, nf = d.nframes;
    const int lo = clampi(n - 2 * r, 0, nf - 1), hi = clampi(n + 2 * r, 0, nf - 1);
    auto missing = [&](int f) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "missing plane of frame " + std::to_string(f));
    };

    // stage the window frames the ring lacks now into the slot's pinned
    // planes, without the mutex; at radius 0 the frame is uploaded for every
    // request (the reference's per-stream slot)
    int stage_of[4 * kMaxRadius + 1];
    int staged = 0;
    for (int f = lo; f <= hi; ++f) {
        const bool need = r == 0 || d.src_frame[f % d.src_ring] != f;
        stage_of[f - lo] = need ? staged++ : -1;
    }
    lock.unlock();
    for (int f = lo; f <= hi; ++f) {
        const int i = f - lo;
        if (stage_of[i] < 0) {
            continue;
        }
        if (!clip_planes[i] || (d.final_pass && !ref_planes[i])) {
            return missing(f);
        }
        VSFEEL_HIP(d.stage(o, stage_of[i], clip_planes[i], clip_strides[i], d.final_pass ? ref_planes[i] : nullptr,
                           d.final_pass ? ref_strides[i] : 0),
                   "hipHostMalloc");
    }
    if (d.fail_next.load() == kFailUpload && d.fail_next.exchange(kFailNone) == kFailUpload) {
        return fail(err, err_len, VSFEEL_BM3D_EDEVICE, "upload of frame " + std::to_string(n) + " failed (test hook)");
    }
    lock.lock();

    // a request alone (every earlier one has copied its frame out, so their
    // work is done) keeps its copies on the kernel stream: no cross-stream
    // waits to pay when nothing can overlap
    int in_flight = 0;
    for (const Instance::OutSlot &q : d.outs) {
        in_flight += q.busy ? 1 : 0;
    }
    const bool alone = in_flight == 1;
    hipStream_t s = d.stream;
    hipStream_t cin = alone ? s : d.copy_in, cout = alone ? s : d.copy_out;
    VSFEEL_HIP(d.begin_on(s), "hipStreamWaitEvent");

    // upload the window frames the ring does not hold (src/bm3d.cpp:772-814)
    // on the copy stream, each after the kernels that last read its slot;
    // requests that ran while this one staged may have filled a slot (its
    // staged copy goes unused) or evicted a frame (staged now, under the
    // mutex).  Every frame is staged before the first upload is enqueued,
    // so an input error leaves the ring as it was.
    bool upload[4 * kMaxRadius + 1];
    for (int f = lo; f <= hi; ++f) {
        const int i = f - lo;
        upload[i] = r == 0 || d.src_frame[f % d.src_ring] != f;
        if (upload[i] && stage_of[i] < 0) {
            if (!clip_planes[i] || (d.final_pass && !ref_planes[i])) {
                return missing(f);
            }
            stage_of[i] = staged++;
            VSFEEL_HIP(d.stage(o, stage_of[i], clip_planes[i], clip_strides[i],
                               d.final_pass ? ref_planes[i] : nullptr, d.final_pass ? ref_strides[i] : 0),
                       "hipHostMalloc");
        }
    }
    for (int f = lo; f <= hi; ++f) {
        const int slot = f % d.src_ring;
        const int i = f - lo;
        if (!upload[i]) {
            continue;
        }
        if (!alone && d.slot_ev_set[size_t(slot)]) {
            VSFEEL_HIP(hipStreamWaitEvent(cin, d.slot_ev[size_t(slot)], 0), "hipStreamWaitEvent");
        }
        d.src_frame[slot] = -1;  // until the upload is enqueued
        // in the final pass the slot holds [ref][source] like the reference's ring
        for (int c = 0; c < d.clips; ++c) {
            VSFEEL_HIP(hipMemcpyAsync(d.src_slot(slot, c), o.stage[size_t(stage_of[i]) * d.clips + c],
                                      size_t(d.pe) * sizeof(float), hipMemcpyHostToDevice, cin),
                       "hipMemcpyAsync");
        }
        d.src_frame[slot] = f;
    }
    if (!alone) {
        VSFEEL_HIP(hipEventRecord(d.up_ev, cin), "hipEventRecord");
        VSFEEL_HIP(hipStreamWaitEvent(s, d.up_ev, 0), "hipStreamWaitEvent");
        // the slot's previous frame may still be on its way out (the fused
        // reduction or the aggregation writes it)
        if (o.used) {
            VSFEEL_HIP(hipStreamWaitEvent(s, o.down, 0), "hipStreamWaitEvent");
        }
    }

    // estimate stacks of the window frames the cache lacks (src/bm3d.cpp:558-602)
    std::vector<int> slot_of(d.tw);
    int est_i[kMaxWindow], est_m[kMaxWindow];
    int nest = 0;
    for (int i = 0; i < d.tw; ++i) {
        const int m = clampi(n - r + i, 0, nf - 1);
        const int slot = r == 0 ? 0 : m % d.res_cap;
        slot_of[i] = slot;
        if (r > 0 && d.res_frame[slot] == m) {
            continue;
        }
        // the same frame twice in the window (clamped at the clip's ends)
        bool dup = false;
        for (int k = 0; k < nest; ++k) {
            dup = dup || est_m[k] == m;
        }
        if (!dup) {
            est_i[nest] = i;
            est_m[nest] = m;
            ++nest;
        }
    }
    auto enqueue = [&](hipStream_t q) -> hipError_t {
        for (int k = 0; k < nest; ++k) {
            const int m = est_m[k];
            const float *search[kMaxWindow];
            const float *input[kMaxWindow];
            for (int z = 0; z < d.tw; ++z) {
                const int ring = clampi(m - r + z, 0, nf - 1) % d.src_ring;
                search[z] = d.src_slot(ring, 0);
                input[z] = d.src_slot(ring, d.clips - 1);
            }
            const hipError_t e = d.estimate(search, input, d.pitch, d.res_slot(slot_of[est_i[k]]), d.pitch, q,
                                            nullptr, kFilterPlain, Instance::Saved(), o.dev, d.pitch);
            if (e != hipSuccess) {
                return e;
            }
        }
        return d.fuses() ? hipSuccess : d.aggregate(n, slot_of, o.dev, d.pitch, q);
    };
    if (d.fail_next.load() == kFailLaunch && d.fail_next.exchange(kFailNone) == kFailLaunch) {
        // the uploads stay enqueued: their slots are ordered before later
        // uploads into them as after kernels
        for (int f = lo; f <= hi; ++f) {
            const int slot = f % d.src_ring;
            VSFEEL_HIP(hipEventRecord(d.slot_ev[size_t(slot)], s), "hipEventRecord");
            d.slot_ev_set[size_t(slot)] = 1;
        }
        VSFEEL_HIP(d.end_on(s), "hipEventRecord");
        return fail(err, err_len, VSFEEL_BM3D_EDEVICE, "launch for frame " + std::to_string(n) + " failed (test hook)");
    }
    if (d.request_graphs && !d.timing) {
        // the kernels from the slot's graph for this recompute pattern: captured
        // on the idle capture stream, then updated into the slot's executable
        // graph (whose last launch has finished: the slot's previous request
        // waited for its download)
        VSFEEL_HIP(hipStreamBeginCapture(d.capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const hipError_t e = enqueue(d.capture);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(d.capture, &g);
        if (e != hipSuccess || ec != hipSuccess) {
            if (g) {
                (void)hipGraphDestroy(g);
            }
            VSFEEL_HIP(e, "launch during graph capture");
            VSFEEL_HIP(ec, "hipStreamEndCapture");
        }
        hipGraphExec_t &x = o.graph[size_t(nest)];
        if (x) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
            if (hipGraphExecUpdate(x, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(x);
                x = nullptr;
            }
        }
        hipError_t ei = hipSuccess;
        if (!x) {
            ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        }
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            x = nullptr;
            VSFEEL_HIP(ei, "hipGraphInstantiate");
        }
        VSFEEL_HIP(hipGraphLaunch(x, s), "hipGraphLaunch");
    } else {
        VSFEEL_HIP(enqueue(s), "launch");
    }
    for (int k = 0; k < nest; ++k) {
        d.res_frame[size_t(slot_of[est_i[k]])] = r == 0 ? -1 : est_m[k];
    }
    VSFEEL_HIP(hipEventRecord(o.done, s), "hipEventRecord");
    for (int f = lo; f <= hi; ++f) {
        const int slot = f % d.src_ring;
        VSFEEL_HIP(hipEventRecord(d.slot_ev[size_t(slot)], s), "hipEventRecord");
        d.slot_ev_set[size_t(slot)] = 1;
    }
    VSFEEL_HIP(d.end_on(s), "hipEventRecord");
    if (!alone) {
        VSFEEL_HIP(hipStreamWaitEvent(cout, o.done, 0), "hipStreamWaitEvent");
    }
    VSFEEL_HIP(hipMemcpyAsync(o.pin, o.dev, size_t(d.pe) * sizeof(float), hipMemcpyDeviceToHost, cout),
               "hipMemcpyAsync");
    VSFEEL_HIP(hipEventRecord(o.down, cout), "hipEventRecord");
    o.used = true;

    // the frame leaves the slot without the lock: later requests proceed
    lock.unlock();
    VSFEEL_HIP(hipEventSynchronize(o.down), "hipEventSynchronize");
    const size_t row_bytes = size_t(d.width) * sizeof(float);
    for (int y = 0; y < d.height; ++y) {
        std::memcpy(reinterpret_cast<char *>(dst) + (ptrdiff_t)y * dst_stride, o.pin + (size_t)y * d.pitch,
                    row_bytes);
    }
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_fail_next_request(vsfeel_bm3d *h, int32_t stage)
{
    if (!h || stage < vsfeel::kFailNone || stage > vsfeel::kFailLaunch) {
        return VSFEEL_BM3D_EINVAL;
    }
    h->inst.fail_next.store(stage);
    return VSFEEL_BM3D_OK;
}

static int check_device_clip(const Instance &d, const vsfeel_bm3d_device_clip *clip, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (!clip || !clip->data) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "missing device clip");
    }
    if (clip->pitch < d.width || clip->frame_stride < (int64_t)clip->pitch * (d.height - 1) + d.width) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "device clip pitch or frame stride too small");
    }
    // the search kernels address a plane with 32-bit byte offsets
    if ((int64_t)clip->pitch * d.height * 4 >= (int64_t(1) << 31)) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "device clip pitch too large (a plane must stay below 2 GiB)");
    }
    if (d.final_pass && !clip->ref) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "the final pass needs the ref clip");
    }
    return VSFEEL_BM3D_OK;
}

static void device_window(const Instance &d, const vsfeel_bm3d_device_clip *clip, int m, const float **search,
                          const float **input)
{
    for (int z = 0; z < d.tw; ++z) {
        const int f = vsfeel::clampi(m - d.radius + z, 0, d.nframes - 1);
        const float *c = clip->data + (int64_t)f * clip->frame_stride;
        input[z] = c;
        search[z] = d.final_pass ? clip->ref + (int64_t)f * clip->frame_stride : c;
    }
}

int vsfeel_bm3d_process_device(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip, int32_t first, int32_t count,
                               float *out, int32_t out_pitch, int64_t out_frame_stride, void *stream, char *err,
                               size_t err_len)
{
    using namespace vsfeel;
    if (!h) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null handle");
    }
    Instance &d = h->inst;
    int rc = check_device_clip(d, clip, err, err_len);
    if (rc != VSFEEL_BM3D_OK) {
        return rc;
    }
    if (first < 0 || count < 0 || first + count > d.nframes) {
        return fail(err, err_len, VSFEEL_BM3D_ERANGE, "frame range outside the clip");
    }
    if (!out || out_pitch < d.width) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "bad output buffer");
    }
    std::lock_guard<std::mutex> lock(d.mu);
    VSFEEL_HIP(hipSetDevice(d.device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    VSFEEL_HIP(d.begin_on(s), "hipStreamWaitEvent");
    const int r = d.radius, nf = d.nframes;
    // the whole call's launch sequence; estimate stacks are only reused
    // inside one call: the caller owns the clip memory and may rewrite it
    // between calls.  The k-th estimate of the call runs on lane k mod
    // nlanes (see Instance::lanes; everything on the launch stream under
    // per-kernel timing, which times each kernel alone).  Side streams are
    // forked from the launch stream when first used and joined back at the
    // end.
    const int nlanes = d.timing ? 1 : d.lanes;
    const bool fused = d.fuses();
    auto enqueue = [&](hipStream_t q) -> hipError_t {
        struct LaneReset {
            Instance &d;
            ~LaneReset() { d.lane = 0; }
        } lane_reset{d};
        std::vector<int> held(d.res_cap, -1);
        std::vector<int> slot_of(d.tw);
        std::vector<char> slot_read(d.res_cap, 0);  // slot_read[slot] recorded in this call
        bool forked[Instance::kMaxLanes] = {}, fresh[Instance::kMaxLanes] = {};
        // lane j's index in stream_l (-1: the launch stream)
        auto side = [&](int j) { return nlanes == 1 ? -1 : fused ? j - 1 : j; };
        auto lane_stream = [&](int j) { return side(j) < 0 ? q : d.stream_l[side(j)]; };
        hipError_t e = hipSuccess;
        if (nlanes > 1) {
            e = hipEventRecord(d.lane_fork, q);
        }
        int k = 0;
        for (int n = first; n < first + count && e == hipSuccess; ++n) {
            for (int i = 0; i < d.tw && e == hipSuccess; ++i) {
                const int m = clampi(n - r + i, 0, nf - 1);
                const int slot = m % d.res_cap;
                slot_of[i] = slot;
                if (held[slot] == m) {
                    continue;
                }
                d.lane = k++ % nlanes;
                const hipStream_t qs = lane_stream(d.lane);
                if (qs != q && !forked[d.lane]) {
                    e = hipStreamWaitEvent(qs, d.lane_fork, 0);
                    forked[d.lane] = true;
                }
                if (e == hipSuccess && nlanes > 1 && !fused && slot_read[slot]) {
                    e = hipStreamWaitEvent(qs, d.slot_read[slot], 0);
                }
                if (e != hipSuccess) {
                    break;
                }
                const float *search[kMaxWindow];
                const float *input[kMaxWindow];
                device_window(d, clip, m, search, input);
                // radius 0: the tile reduction writes the frame (fused aggregation)
                e = d.estimate(search, input, clip->pitch, d.res_slot(slot), d.pitch, qs, nullptr, kFilterPlain,
                               Instance::Saved(), out + (int64_t)(n - first) * out_frame_stride, out_pitch);
                if (e == hipSuccess && nlanes > 1 && !fused) {
                    e = hipEventRecord(d.lane_done[d.lane], qs);
                    fresh[d.lane] = true;
                }
                held[slot] = m;
            }
            if (fused || e != hipSuccess) {
                continue;
            }
            for (int j = 0; j < nlanes && e == hipSuccess; ++j) {
                if (fresh[j]) {
                    e = hipStreamWaitEvent(q, d.lane_done[j], 0);
                    fresh[j] = false;
                }
            }
            if (e == hipSuccess) {
                e = d.aggregate(n, slot_of, out + (int64_t)(n - first) * out_frame_stride, out_pitch, q);
            }
            for (int i = 0; i < d.tw && e == hipSuccess && nlanes > 1; ++i) {
                e = hipEventRecord(d.slot_read[slot_of[i]], q);
                slot_read[slot_of[i]] = 1;
            }
        }
        // join every side stream, also after an error (a capture must end
        // with its forks joined)
        for (int j = 0; j < nlanes; ++j) {
            if (forked[j]) {
                const int li = side(j);
                hipError_t ej = hipEventRecord(d.lane_join[li], d.stream_l[li]);
                if (ej == hipSuccess) {
                    ej = hipStreamWaitEvent(q, d.lane_join[li], 0);
                }
                if (e == hipSuccess) {
                    e = ej;
                }
            }
        }
        return e;
    };
    if (d.graphs && !d.timing && count > 0) {
        // replay from a graph captured on the instance stream, ordered after
        // and before the caller's stream by events
        const Instance::GraphKey key{clip->data, clip->ref, clip->frame_stride, out_frame_stride,
                                     clip->pitch, first, count, out_pitch, out};
        if (!d.graph_exec || !(d.graph_key == key)) {
            // the previous graph, or host-buffer requests, may still run on
            // the instance stream: capture on an idle one
            VSFEEL_HIP(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
            for (hipStream_t q : d.stream_l) {
                if (q) {
                    VSFEEL_HIP(hipStreamSynchronize(q), "hipStreamSynchronize");
                }
            }
            if (d.graph_exec) {
                (void)hipGraphExecDestroy(d.graph_exec);
                d.graph_exec = nullptr;
            }
            VSFEEL_HIP(hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            const hipError_t e = enqueue(d.stream);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(d.stream, &g);
            if (e != hipSuccess || ec != hipSuccess) {
                if (g) {
                    (void)hipGraphDestroy(g);
                }
                VSFEEL_HIP(e, "launch during graph capture");
                VSFEEL_HIP(ec, "hipStreamEndCapture");
            }
            const hipError_t ei = hipGraphInstantiate(&d.graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                d.graph_exec = nullptr;
                VSFEEL_HIP(ei, "hipGraphInstantiate");
            }
            d.graph_key = key;
        }
        VSFEEL_HIP(hipEventRecord(d.fork_ev, s), "hipEventRecord");
        VSFEEL_HIP(hipStreamWaitEvent(d.stream, d.fork_ev, 0), "hipStreamWaitEvent");
        VSFEEL_HIP(hipGraphLaunch(d.graph_exec, d.stream), "hipGraphLaunch");
        VSFEEL_HIP(hipEventRecord(d.join_ev, d.stream), "hipEventRecord");
        VSFEEL_HIP(hipStreamWaitEvent(s, d.join_ev, 0), "hipStreamWaitEvent");
    } else {
        VSFEEL_HIP(enqueue(s), "launch");
    }
    // the res slots now hold this call's stacks, not the host path's frames
    std::fill(d.res_frame.begin(), d.res_frame.end(), -1);
    VSFEEL_HIP(d.end_on(s), "hipEventRecord");
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_process_device_static(const vsfeel_bm3d_args *args, const vsfeel_bm3d_device_clip *clip,
                                      int32_t first, int32_t count, float *out, int32_t out_pitch,
                                      int64_t out_frame_stride, void *stream, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (!args) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null argument");
    }
    // the parameters only: no device object of the instance is created, and
    // device < 0 keeps its destructor from touching the device
    Instance d;
    d.device = -1;
    const std::string msg = parse_args(*args, d);
    const int dev = d.device;
    d.device = -1;
    if (!msg.empty()) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, msg);
    }
    if (d.radius != 0 || d.final_pass || !d.tiled || d.short_windows || d.operand != kOperandFp32) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL,
                    "the allocation-free entry point supports the plain form at radius 0 without a ref clip, on "
                    "frames whose every search window holds at least 8 candidates");
    }
    if (dev < 0 || dev >= kStaticMaxDevices) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "device ordinal out of range");
    }
    // the scratch layout: tile tables, group table, tile accumulators
    const size_t tab_bytes = round256(size_t(2 * d.width + 2 * d.height + d.tiles_x + d.tiles_y) * sizeof(int));
    const size_t group_bytes = round256(size_t(d.nref) * 8 * sizeof(uint32_t));
    const size_t acc_elems = size_t(d.tiles_x) * d.tiles_y * 2 * d.region_h * d.region_w;
    const size_t need = tab_bytes + group_bytes + acc_elems * sizeof(float);
    if (need > kStaticScratchBytes) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL,
                    "a " + std::to_string(d.width) + "x" + std::to_string(d.height) + " frame at these parameters needs " +
                        std::to_string(need) + " bytes of scratch, more than the " +
                        std::to_string(kStaticScratchBytes) + " the allocation-free entry point holds");
    }
    int rc = check_device_clip(d, clip, err, err_len);
    if (rc != VSFEEL_BM3D_OK) {
        return rc;
    }
    if (first < 0 || count < 0 || first + count > d.nframes) {
        return fail(err, err_len, VSFEEL_BM3D_ERANGE, "frame range outside the clip");
    }
    if (!out || out_pitch < d.width) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "bad output buffer");
    }
    VSFEEL_HIP(hipSetDevice(dev), "hipSetDevice");
    unsigned char *base = nullptr;
    VSFEEL_HIP(static_scratch(dev, &base), "hipGetSymbolAddress");
    d.tile_tables = reinterpret_cast<int *>(base);
    d.groups = reinterpret_cast<uint32_t *>(base + tab_bytes);
    d.scratch = reinterpret_cast<float *>(base + tab_bytes + group_bytes);
    d.reduce_span3 = max_tile_span(d.width, d.tiles_x, d.tile_bx, d.nx, d.block_step, d.bm_range) <= 3 &&
                     max_tile_span(d.height, d.tiles_y, d.tile_by, d.ny, d.block_step, d.bm_range) <= 3 &&
                     acc_elems < (size_t(1) << 32);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = launch_tile_tables(d.tile_tables, d.width, d.height, d.tiles_x, d.tiles_y, d.tile_bx, d.tile_by,
                                      d.nx, d.ny, d.block_step, d.bm_range, s);
    for (int n = first; n < first + count && e == hipSuccess; ++n) {
        const float *search[kMaxWindow];
        const float *input[kMaxWindow];
        device_window(d, clip, n, search, input);
        // radius 0: the tile reduction writes the frame (fused aggregation);
        // the frame's own slice is never touched, so no stack is passed
        e = d.estimate(search, input, clip->pitch, nullptr, d.pitch, s, nullptr, kFilterPlain, Instance::Saved(),
                       out + (int64_t)(n - first) * out_frame_stride, out_pitch);
    }
    d.tile_tables = nullptr;
    d.groups = nullptr;
    d.scratch = nullptr;
    VSFEEL_HIP(e, "launch");
    return VSFEEL_BM3D_OK;
}

static int check_saved(const Instance &d, const vsfeel_bm3d_saved *saved, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (d.radius != 0 || d.final_pass) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL,
                    "the differentiable form supports radius 0 and the basic (hard-threshold) pass");
    }
    if (!saved || !saved->groups || !saved->masks || !saved->weights || !saved->denom ||
        saved->denom_pitch < d.width) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "incomplete saved state");
    }
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_forward_saved(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip, int32_t first, int32_t count,
                              float *out, int32_t out_pitch, int64_t out_frame_stride, const vsfeel_bm3d_saved *saved,
                              void *stream, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (!h) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null handle");
    }
    Instance &d = h->inst;
    int rc = check_device_clip(d, clip, err, err_len);
    if (rc == VSFEEL_BM3D_OK) {
        rc = check_saved(d, saved, err, err_len);
    }
    if (rc != VSFEEL_BM3D_OK) {
        return rc;
    }
    if (first < 0 || count < 0 || first + count > d.nframes) {
        return fail(err, err_len, VSFEEL_BM3D_ERANGE, "frame range outside the clip");
    }
    if (!out || out_pitch < d.width) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "bad output buffer");
    }
    std::lock_guard<std::mutex> lock(d.mu);
    VSFEEL_HIP(hipSetDevice(d.device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    VSFEEL_HIP(d.begin_on(s), "hipStreamWaitEvent");
    // (the result slots are overwritten: the host path's cache forgets them)
    std::fill(d.res_frame.begin(), d.res_frame.end(), -1);
    Instance::FrameLanes fl(d, s);
    VSFEEL_HIP(fl.start(), "hipEventRecord");
    for (int i = 0; i < count; ++i) {
        const int n = first + i;
        hipStream_t qs = s;
        hipError_t e = fl.select(i, &qs);
        const std::vector<int> slot_of(1, d.lane);
        const float *search[kMaxWindow];
        const float *input[kMaxWindow];
        device_window(d, clip, n, search, input);
        Instance::Saved sv;
        sv.groups = saved->groups + (int64_t)i * d.nref * 8;
        sv.masks = saved->masks + (int64_t)i * d.nref * 8;
        sv.weights = saved->weights + (int64_t)i * d.nref;
        if (e == hipSuccess) {
            e = d.estimate(search, input, clip->pitch, d.res_slot(d.lane), d.pitch, qs, nullptr, kFilterSave, sv);
        }
        if (e == hipSuccess) {
            e = d.aggregate(n, slot_of, out + (int64_t)i * out_frame_stride, out_pitch, qs,
                            saved->denom + (int64_t)i * d.height * saved->denom_pitch, saved->denom_pitch);
        }
        if (e != hipSuccess) {
            (void)fl.finish();
            VSFEEL_HIP(e, "forward launch");
        }
    }
    VSFEEL_HIP(fl.finish(), "hipStreamWaitEvent");
    VSFEEL_HIP(d.end_on(s), "hipEventRecord");
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_backward(vsfeel_bm3d *h, const float *grad_out, int32_t grad_pitch, int64_t grad_frame_stride,
                         int32_t count, const vsfeel_bm3d_saved *saved, float *grad_in, int32_t gin_pitch,
                         int64_t gin_frame_stride, void *stream, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (!h) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null handle");
    }
    Instance &d = h->inst;
    int rc = check_saved(d, saved, err, err_len);
    if (rc != VSFEEL_BM3D_OK) {
        return rc;
    }
    if (!grad_out || !grad_in || grad_pitch < d.width || gin_pitch < d.width || count < 0) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "bad gradient buffers");
    }
    std::lock_guard<std::mutex> lock(d.mu);
    VSFEEL_HIP(hipSetDevice(d.device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    VSFEEL_HIP(d.begin_on(s), "hipStreamWaitEvent");
    // (the result slots are overwritten: the host path's cache forgets them)
    std::fill(d.res_frame.begin(), d.res_frame.end(), -1);
    Instance::FrameLanes fl(d, s);
    VSFEEL_HIP(fl.start(), "hipEventRecord");
    for (int i = 0; i < count; ++i) {
        // y = g / D, then grad = A y with the saved groups, masks and weights
        hipStream_t qs = s;
        hipError_t e = fl.select(i, &qs);
        const std::vector<int> slot_of(1, d.lane);
        float *const yp = d.lane_dst();
        if (e == hipSuccess) {
            e = launch_divide(grad_out + (int64_t)i * grad_frame_stride, grad_pitch,
                              saved->denom + (int64_t)i * d.height * saved->denom_pitch, saved->denom_pitch, yp,
                              d.pitch, d.width, d.height, qs);
        }
        const float *y[kMaxWindow] = { yp };
        Instance::Saved sv;
        sv.groups = saved->groups + (int64_t)i * d.nref * 8;
        sv.masks = saved->masks + (int64_t)i * d.nref * 8;
        sv.weights = saved->weights + (int64_t)i * d.nref;
        if (e == hipSuccess) {
            e = d.estimate(y, y, d.pitch, d.res_slot(d.lane), d.pitch, qs, nullptr, kFilterApply, sv);
        }
        if (e == hipSuccess) {
            e = d.aggregate(0, slot_of, grad_in + (int64_t)i * gin_frame_stride, gin_pitch, qs, nullptr, 0, true);
        }
        if (e != hipSuccess) {
            (void)fl.finish();
            VSFEEL_HIP(e, "backward launch");
        }
    }
    VSFEEL_HIP(fl.finish(), "hipStreamWaitEvent");
    VSFEEL_HIP(d.end_on(s), "hipEventRecord");
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_debug_groups(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip, int32_t m, int32_t *groups,
                             float *dists, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (!h) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null handle");
    }
    Instance &d = h->inst;
    int rc = check_device_clip(d, clip, err, err_len);
    if (rc != VSFEEL_BM3D_OK) {
        return rc;
    }
    if (m < 0 || m >= d.nframes) {
        return fail(err, err_len, VSFEEL_BM3D_ERANGE, "frame out of range");
    }
    std::lock_guard<std::mutex> lock(d.mu);
    VSFEEL_HIP(hipSetDevice(d.device), "hipSetDevice");
    hipStream_t s = d.stream;
    VSFEEL_HIP(d.begin_on(s), "hipStreamWaitEvent");
    const float *search[kMaxWindow];
    const float *input[kMaxWindow];
    device_window(d, clip, m, search, input);
    const MatchArgs a = d.match_args(search, clip->pitch, d.dists);
    VSFEEL_HIP(d.match(a, search, clip->pitch, s), "match launch");
    std::vector<uint32_t> packed(size_t(d.nref) * 8);
    VSFEEL_HIP(hipMemcpyAsync(packed.data(), d.groups, packed.size() * 4, hipMemcpyDeviceToHost, s), "hipMemcpy");
    if (dists) {
        VSFEEL_HIP(hipMemcpyAsync(dists, d.dists, packed.size() * 4, hipMemcpyDeviceToHost, s), "hipMemcpy");
    }
    VSFEEL_HIP(d.end_on(s), "hipEventRecord");
    VSFEEL_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (groups) {
        for (size_t i = 0; i < packed.size(); ++i) {
            groups[i * 3 + 0] = int32_t(entry_x(packed[i]));
            groups[i * 3 + 1] = int32_t(entry_y(packed[i]));
            groups[i * 3 + 2] = int32_t(entry_z(packed[i]));
        }
    }
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_debug_estimate(vsfeel_bm3d *h, const vsfeel_bm3d_device_clip *clip, int32_t m, float *res,
                               int32_t res_pitch, char *err, size_t err_len)
{
    using namespace vsfeel;
    if (!h) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "null handle");
    }
    Instance &d = h->inst;
    int rc = check_device_clip(d, clip, err, err_len);
    if (rc != VSFEEL_BM3D_OK) {
        return rc;
    }
    if (m < 0 || m >= d.nframes) {
        return fail(err, err_len, VSFEEL_BM3D_ERANGE, "frame out of range");
    }
    if (!res || res_pitch < d.width || (res_pitch % 4) != 0) {
        return fail(err, err_len, VSFEEL_BM3D_EINVAL, "bad estimate buffer (pitch must be a multiple of 4)");
    }
    std::lock_guard<std::mutex> lock(d.mu);
    VSFEEL_HIP(hipSetDevice(d.device), "hipSetDevice");
    hipStream_t s = d.stream;
    VSFEEL_HIP(d.begin_on(s), "hipStreamWaitEvent");
    const float *search[kMaxWindow];
    const float *input[kMaxWindow];
    device_window(d, clip, m, search, input);
    VSFEEL_HIP(d.estimate(search, input, clip->pitch, res, res_pitch, s), "estimate launch");
    VSFEEL_HIP(d.end_on(s), "hipEventRecord");
    VSFEEL_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_set_timing(vsfeel_bm3d *h, int32_t enable)
{
    if (!h) {
        return VSFEEL_BM3D_EINVAL;
    }
    Instance &d = h->inst;
    std::lock_guard<std::mutex> lock(d.mu);
    if (d.device >= 0) {
        (void)hipSetDevice(d.device);
    }
    (void)d.harvest();
    for (int k = 0; k < vsfeel::kKernelKinds; ++k) {
        d.kernel_ms[k] = 0.0;
        d.kernel_count[k] = 0;
    }
    d.timing = enable != 0;
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_set_graphs(vsfeel_bm3d *h, int32_t enable)
{
    if (!h) {
        return VSFEEL_BM3D_EINVAL;
    }
    if (enable < 0 || enable > 3) {
        return VSFEEL_BM3D_EINVAL;
    }
    Instance &d = h->inst;
    std::lock_guard<std::mutex> lock(d.mu);
    d.graphs = (enable & 1) != 0;
    d.request_graphs = (enable & 2) != 0;
    if (d.device >= 0) {
        (void)hipSetDevice(d.device);
    }
    if (d.stream && ((!d.graphs && d.graph_exec) || !d.request_graphs)) {
        // (requests that hold a slot launched its graph before this)
        (void)hipStreamSynchronize(d.stream);
    }
    if (!d.graphs && d.graph_exec) {
        (void)hipGraphExecDestroy(d.graph_exec);
        d.graph_exec = nullptr;
    }
    if (!d.request_graphs) {
        for (Instance::OutSlot &q : d.outs) {
            d.drop_slot_graphs(q);
        }
    }
    return VSFEEL_BM3D_OK;
}

int vsfeel_bm3d_kernel_times(vsfeel_bm3d *h, double *ms, int64_t *count)
{
    if (!h || !ms || !count) {
        return VSFEEL_BM3D_EINVAL;
    }
    Instance &d = h->inst;
    std::lock_guard<std::mutex> lock(d.mu);
    if (d.device >= 0) {
        (void)hipSetDevice(d.device);
    }
    const hipError_t e = d.harvest();
    for (int k = 0; k < vsfeel::kKernelKinds; ++k) {
        ms[k] = d.kernel_ms[k];
        count[k] = d.kernel_count[k];
    }
    return e == hipSuccess ? VSFEEL_BM3D_OK : VSFEEL_BM3D_EDEVICE;
}

}  // extern "C"
