// Device helpers shared by the BM3Dv2 kernels (bm3d_kernels.hip, bm3d_mfma.hip):
// wavefront lane moves (DPP, permlane), the top-8 list and group list of one
// search, the 16-lane bitonic sort of (distance, seq) keys, the self-block
// rule, and the reference's 8-row column SSD.
#pragma once

#include <float.h>

#include "bm3d_common.hpp"

namespace vsfeel {

namespace {

constexpr uint64_t kPlaceholderKey = (uint64_t(0x7F7FFFFFu) << 32) | 0x7FFFFFFFu;  // (FLT_MAX, INT_MAX)
constexpr uint64_t kNoKey = ~uint64_t(0);
constexpr int kSortThreshold = 8;  // survivors above this are merged by a sort

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// XCD-aware workgroup order: the hardware hands consecutive workgroups to
// the 8 XCDs in turn, so give every XCD a contiguous run of the work (its L2
// then holds one band of the frame)
__device__ __forceinline__ int xcd_block_id()
{
    int bid = blockIdx.x;
    const int nwg = gridDim.x;
    const int nfull = (nwg / 8) * 8;
    if (bid < nfull) {
        bid = (bid % 8) * (nwg / 8) + bid / 8;
    }
    return bid;
}

// Orders LDS traffic of one wavefront: LDS operations of a wave execute in
// issue order, so only the compiler has to be kept from moving them across.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t make_key(float d, int seq)
{
    return (uint64_t(__float_as_uint(d)) << 32) | uint32_t(seq);
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m)
{
    const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), m));
    const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), m));
    return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src)
{
    const uint32_t lo = uint32_t(__shfl(int(uint32_t(v)), src));
    const uint32_t hi = uint32_t(__shfl(int(uint32_t(v >> 32)), src));
    return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int l)
{
    return uint32_t(__builtin_amdgcn_readlane(int(v), l));
}

// DPP lane moves (full-rate VALU, no LDS round trip).
// row_shr:1 -- lane i takes lane i-1 of its 16-lane row (row's lane 0 keeps its own)
__device__ __forceinline__ uint32_t dpp_row_shr1(uint32_t v)
{
    return uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0x111, 0xF, 0xF, false));
}

// wave_shl:1 -- lane i takes lane i+1 across the whole wavefront (lane 63 keeps its own)
__device__ __forceinline__ float dpp_wave_shl1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l)
{
    return (uint64_t(readlane_u32(uint32_t(v >> 32), l)) << 32) | readlane_u32(uint32_t(v), l);
}

// ---------------------------------------------------------------------------
// Top-8 list held by lanes 0..7 of a wavefront, ascending by the key
// (distance bits, seq).  Distances are non-negative, so their IEEE bits order
// like the values; seq is unique per candidate of one search, so the key
// order is exactly the reference's (distance, then smaller seq) merge order.
// ---------------------------------------------------------------------------
struct TopList {
    uint64_t key;
    uint32_t xy;  // x | y << 16
};

__device__ __forceinline__ void list_reset(TopList &L)
{
    L.key = kPlaceholderKey;
    L.xy = 0;
}

__device__ __forceinline__ uint64_t list_threshold(const TopList &L) { return readlane_u64(L.key, 7); }

// insert one wave-uniform candidate that is below the current threshold
__device__ __forceinline__ void list_insert(TopList &L, uint64_t key, uint32_t xy, int lane)
{
    const uint64_t below = __ballot(lane < 8 && L.key < key);
    const int pos = __popcll(below & 0xFFull);
    const uint64_t up_key = (uint64_t(dpp_row_shr1(uint32_t(L.key >> 32))) << 32) | dpp_row_shr1(uint32_t(L.key));
    const uint32_t up_xy = dpp_row_shr1(L.xy);
    if (lane == pos) {
        L.key = key;
        L.xy = xy;
    } else if (lane > pos && lane < 8) {
        L.key = up_key;
        L.xy = up_xy;
    }
}

// merge one candidate per lane (kNoKey where there is none) into the list:
// bitonic sort of the 64 lanes, then the 8 lowest of list and batch
__device__ void list_merge_batch(TopList &L, uint64_t key, uint32_t xy, int lane)
{
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t ok = shfl_xor64(key, j);
            const uint32_t oxy = uint32_t(__shfl_xor(int(xy), j));
            const bool asc = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            const bool take = (lower == asc) ? (ok < key) : (ok > key);
            if (take) {
                key = ok;
                xy = oxy;
            }
        }
    }
    const int rev = (7 - lane) & 63;
    const uint64_t bk = shfl64(key, rev);
    const uint32_t bxy = uint32_t(__shfl(int(xy), rev));
    uint64_t ck = L.key;
    uint32_t cxy = L.xy;
    if (bk < ck) {
        ck = bk;
        cxy = bxy;
    }
#pragma unroll
    for (int j = 4; j > 0; j >>= 1) {
        const uint64_t ok = shfl_xor64(ck, j);
        const uint32_t oxy = uint32_t(__shfl_xor(int(cxy), j));
        const bool lower = (lane & j) == 0;
        const bool take = lower ? (ok < ck) : (ok > ck);
        if (take) {
            ck = ok;
            cxy = oxy;
        }
    }
    if (lane < 8) {
        L.key = ck;
        L.xy = cxy;
    }
}

// offer one candidate per lane (valid where `cand`) to the list
__device__ __forceinline__ void list_offer(TopList &L, bool cand, float d, int seq, uint32_t xy, int lane)
{
    const bool valid = cand && (d < FLT_MAX);  // the shader's strict '<' against FLT_MAX
    const uint64_t key = valid ? make_key(d, seq) : kNoKey;
    const uint64_t thr = list_threshold(L);
    const bool surv = key < thr;
    uint64_t mask = __ballot(surv);
    if (mask == 0) {
        return;
    }
    if (__popcll(mask) > kSortThreshold) {
        list_merge_batch(L, surv ? key : kNoKey, xy, lane);
        return;
    }
    while (mask) {
        const int src = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1;
        const uint64_t k = readlane_u64(key, src);
        const uint32_t kxy = readlane_u32(xy, src);
        if (k < list_threshold(L)) {
            list_insert(L, k, kxy, lane);
        }
    }
}

// squared error of one 8-row column, accumulated like col_ssd of the shader
// (even and odd rows in two partial sums)
template <int CM>
__device__ __forceinline__ float col_ssd(const float *c, const float *s)
{
    // the shader's sums start at 0; 0 + v*v is v*v exactly (a square is
    // never -0), so the first term initialises them
    const float d0 = c[0] - s[0];
    const float d1 = c[1] - s[1];
    float e0 = d0 * d0, e1 = d1 * d1;
#pragma unroll
    for (int i = 2; i < 8; i += 2) {
        const float v0 = c[i] - s[i];
        const float v1 = c[i + 1] - s[i + 1];
        if (CM == kFpAll) {
            e0 = __builtin_fmaf(v0, v0, e0);
            e1 = __builtin_fmaf(v1, v1, e1);
        } else {
            e0 = e0 + v0 * v0;
            e1 = e1 + v1 * v1;
        }
    }
    return e0 + e1;
}

// ---------------------------------------------------------------------------
// One search window.  Lanes map to window columns: lane j of a row segment
// computes one column SSD, with the reference-block column min(j, 7) for the
// sliding metric (radius < 3), and a row's candidate k is the tree sum of the
// column values of lanes k..k+7 -- exactly what the shader's shifted column
// sums add up to.  Short rows share the wave (several row segments per pass);
// long rows are cut into chunks of 57 candidates.  Each lane walks down its
// column with an 8-row register window.  The candidates go to the list L,
// or (TO_ROWS) their distances to drows[row * drow_stride + k].
// ---------------------------------------------------------------------------
template <int CM, bool EXACT, bool TO_ROWS = false>
__device__ void scan_window(TopList &L, const float *__restrict__ plane, int pitch, int height,
                            const float *cur_lds, const float *cur_reg, int left, int right,
                            int top, int bottom, int seq_base, int lane, int skip_row = -1,
                            float *drows = nullptr, int drow_stride = 0)
{
    const int rw = right - left + 1;
    const int rh = bottom - top + 1;
    for (int k0 = 0; k0 < rw; k0 += 57) {
        const int nc = min(rw - k0, 57);
        const int seg_len = nc + 7;
        const int per_pass = 64 / seg_len;
        const int seg = lane / seg_len;
        const int kk = lane - seg * seg_len;
        const bool lane_on = seg < per_pass;
        const int j = k0 + kk;
        const int X = lane_on ? left + j : left;
        float c[8];
        if (!EXACT) {
            const int a = min(j, 7);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                c[i] = cur_lds[a * 8 + i];
            }
        }
        const float *col_ptr = plane + X;
        int cy = top + (lane_on ? seg : 0);
        float s[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            s[i] = col_ptr[(long long)min(cy + i, height - 1) * pitch];
        }
        const int npass = (rh + per_pass - 1) / per_pass;
        for (int pass = 0; pass < npass; ++pass) {
            if (pass > 0) {
                for (int q = 0; q < per_pass; ++q) {
#pragma unroll
                    for (int i = 0; i < 7; ++i) {
                        s[i] = s[i + 1];
                    }
                    s[7] = col_ptr[(long long)min(cy + 8 + q, height - 1) * pitch];
                }
                cy += per_pass;
            }
            float d;
            if (!EXACT) {
                const float col = col_ssd<CM>(c, s);
                const float p1 = col + dpp_wave_shl1(col);
                const float p2 = p1 + dpp_wave_shl1(dpp_wave_shl1(p1));
                d = p2 + dpp_wave_shl1(dpp_wave_shl1(dpp_wave_shl1(dpp_wave_shl1(p2))));
            } else {
                float S[8];
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) {
                    S[cc] = col_ssd<CM>(cur_reg + cc * 8, s);
                }
                const float t01 = S[0] + __shfl_down(S[1], 1);
                const float t23 = __shfl_down(S[2], 2) + __shfl_down(S[3], 3);
                const float t45 = __shfl_down(S[4], 4) + __shfl_down(S[5], 5);
                const float t67 = __shfl_down(S[6], 6) + __shfl_down(S[7], 7);
                d = (t01 + t23) + (t45 + t67);
            }
            const int row = cy - top;
            const bool cand = lane_on && kk < nc && row < rh && cy != skip_row;
            if (TO_ROWS) {
                if (cand) {
                    drows[row * drow_stride + j] = d;
                }
            } else {
                const int seq = seq_base + row * rw + j;
                const uint32_t xy = uint32_t(left + j) | (uint32_t(cy) << 16);
                list_offer(L, cand, d, seq, xy, lane);
            }
        }
    }
}

// Group list (spatial top-8 plus temporal inserts) held by lanes 0..7.
struct GroupList {
    float e;
    uint32_t xy;
    int z;
};

// ginsert8 of the shader: strict '<' on the distance, after equal ones
__device__ __forceinline__ void group_insert(GroupList &G, float fe, uint32_t fxy, int z, int lane)
{
    const float e7 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(G.e), 7));
    if (!(fe < e7)) {
        return;
    }
    const uint64_t before = __ballot(lane < 8 && G.e <= fe);
    const int pos = __popcll(before & 0xFFull);
    // lanes 1..7 take their left neighbour's entry (DPP row_shr:1)
    const float up_e = __uint_as_float(dpp_row_shr1(__float_as_uint(G.e)));
    const uint32_t up_xy = dpp_row_shr1(G.xy);
    const int up_z = int(dpp_row_shr1(uint32_t(G.z)));
    if (lane == pos) {
        G.e = fe;
        G.xy = fxy;
        G.z = z;
    } else if (lane > pos && lane < 8) {
        G.e = up_e;
        G.xy = up_xy;
        G.z = up_z;
    }
}

// self-block rule (src/bm3d.comp:639-663) and the group table write
__device__ __forceinline__ void finish_group(const MatchArgs &A, GroupList &G, int blk, int x, int y, int lane)
{
    const uint32_t self_xy = uint32_t(x) | (uint32_t(y) << 16);
    const uint64_t present = __ballot(lane < 8 && G.xy == self_xy && G.z == A.radius);
    if (present == 0) {
        const float up_e = __uint_as_float(dpp_row_shr1(__float_as_uint(G.e)));
        const uint32_t up_xy = dpp_row_shr1(G.xy);
        const int up_z = int(dpp_row_shr1(uint32_t(G.z)));
        if (lane == 0) {
            G.e = -1.0f;
            G.xy = self_xy;
            G.z = A.radius;
        } else if (lane < 8) {
            G.e = up_e;
            G.xy = up_xy;
            G.z = up_z;
        }
    }
    if (lane < 8) {
        const uint32_t gx = G.xy & 0xFFFFu, gy = G.xy >> 16;
        A.groups[(long long)blk * 8 + lane] = pack_entry(gx, gy, uint32_t(G.z));
        if (A.dists) {
            A.dists[(long long)blk * 8 + lane] = G.e;
        }
    }
}

// finish_group for a list held by lanes sub 0..7 of a 16-lane row (the
// spatial search at radius 0, every member in the frame itself)
__device__ __forceinline__ void finish_group16(const MatchArgs &A, float e, uint32_t xy, int blk, int x, int y,
                                               int sub)
{
    const uint32_t self_xy = uint32_t(x) | (uint32_t(y) << 16);
    const int shift = (lane_id() >> 4) * 16;
    const uint32_t seg = uint32_t(__ballot(sub < 8 && xy == self_xy) >> shift) & 0xFFu;
    if (seg == 0) {
        const uint32_t up_xy = dpp_row_shr1(xy);
        const float up_e = __uint_as_float(dpp_row_shr1(__float_as_uint(e)));
        if (sub == 0) {
            e = -1.0f;
            xy = self_xy;
        } else if (sub < 8) {
            e = up_e;
            xy = up_xy;
        }
    }
    if (sub < 8) {
        A.groups[(long long)blk * 8 + sub] = pack_entry(xy & 0xFFFFu, xy >> 16, uint32_t(A.radius));
        if (A.dists) {
            A.dists[(long long)blk * 8 + sub] = e;
        }
    }
}

// DPP move of v; lanes without a source lane in their row keep `old`
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v)
{
    return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(v), CTRL, 0xF, 0xF, false));
}

// minimum over the wavefront: DPP row shifts leave each 16-lane row's
// minimum in its lane 15, then the four rows are combined
__device__ __forceinline__ float wave_min(float v)
{
    const uint32_t inf = __float_as_uint(INFINITY);
    v = fminf(v, __uint_as_float(dpp_u32<0x111>(inf, __float_as_uint(v))));  // row_shr:1
    v = fminf(v, __uint_as_float(dpp_u32<0x112>(inf, __float_as_uint(v))));  // row_shr:2
    v = fminf(v, __uint_as_float(dpp_u32<0x114>(inf, __float_as_uint(v))));  // row_shr:4
    v = fminf(v, __uint_as_float(dpp_u32<0x118>(inf, __float_as_uint(v))));  // row_shr:8
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 47));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    return fminf(fminf(r0, r1), fminf(r2, r3));
}

// lane value of the partner lane ^ J inside a 16-lane row
// DPP move with bound_ctrl: a lane without a source lane reads 0 (no old
// value to carry, so no copy before the move)
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v)
{
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), CTRL, 0xF, 0xF, true));
}

template <int J>
__device__ __forceinline__ uint32_t dpp_xor(uint32_t v, int lane)
{
    if (J == 1) {
        return dpp_mov<0xB1>(v);  // quad_perm [1,0,3,2]
    } else if (J == 2) {
        return dpp_mov<0x4E>(v);  // quad_perm [2,3,0,1]
    } else if (J == 4) {
        const uint32_t up = dpp_mov<0x104>(v);    // row_shl:4 (lane i reads i+4)
        const uint32_t down = dpp_mov<0x114>(v);  // row_shr:4 (lane i reads i-4)
        return (lane & 4) ? down : up;
    } else {
        return dpp_mov<0x128>(v);  // row_ror:8 (lane i reads i^8)
    }
}

// minimum of an unsigned value over the 16 lanes of each row
__device__ __forceinline__ uint32_t row_min16_u32(uint32_t v, int lane)
{
    v = min(v, dpp_xor<1>(v, lane));
    v = min(v, dpp_xor<2>(v, lane));
    v = min(v, dpp_xor<4>(v, lane));
    v = min(v, dpp_xor<8>(v, lane));
    return v;
}

// inclusive prefix sum over the 16 lanes of each row (row_shr with zero fill)
__device__ __forceinline__ int row_prefix_sum16(int v)
{
    v += int(dpp_mov<0x111>(uint32_t(v)));  // row_shr:1
    v += int(dpp_mov<0x112>(uint32_t(v)));  // row_shr:2
    v += int(dpp_mov<0x114>(uint32_t(v)));  // row_shr:4
    v += int(dpp_mov<0x118>(uint32_t(v)));  // row_shr:8
    return v;
}

// minimum over the 32 lanes of each half of the wavefront: the row minimum,
// then v_permlane16_swap, which leaves every lane its own row's value and
// the other row's of the pair in its two results
__device__ __forceinline__ uint32_t half_min32_u32(uint32_t v, int lane)
{
    v = row_min16_u32(v, lane);
    const auto sw = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return min(uint32_t(sw[0]), uint32_t(sw[1]));
}

// inclusive prefix sum over the 32 lanes of each half of the wavefront: the
// odd row adds the even row's total (row_bcast:15 hands lane 15 of each row
// to the next row)
__device__ __forceinline__ int half_prefix_sum32(int v, int lane)
{
    v = row_prefix_sum16(v);
    const int below = int(__builtin_amdgcn_update_dpp(0u, uint32_t(v), 0x142, 0xA, 0xF, false));
    return (lane & 16) ? v + below : v;
}

// minimum over the whole wavefront in every lane: the half minimum, then
// v_permlane32_swap hands each lane the other half's
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v, int lane)
{
    v = half_min32_u32(v, lane);
    const auto sw = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return min(uint32_t(sw[0]), uint32_t(sw[1]));
}

// Merges one candidate per lane (kNoKey where there is none) into the list
// (lanes 0..7), keeping its k smallest keys in lanes 0..k-1 (k <= 8; lanes
// k..7 then hold placeholders): k rounds of a wavefront minimum over each
// lane's smaller remaining key, the winning lane (the lowest among equal
// keys) moving on to its other one.  No LDS; for the few keys a tie-heavy
// search needs, cheaper than sorting all 64.
__device__ void list_merge_rounds(TopList &L, uint64_t key, uint32_t xy, int lane, int k)
{
    uint64_t lo = key, hi = lane < 8 ? L.key : kNoKey;
    uint32_t lxy = xy, hxy = lane < 8 ? L.xy : 0u;
    if (hi < lo) {
        const uint64_t t = lo;
        lo = hi;
        hi = t;
        const uint32_t tx = lxy;
        lxy = hxy;
        hxy = tx;
    }
    uint64_t nk = kPlaceholderKey;
    uint32_t nxy = 0u;
    for (int r = 0; r < k; ++r) {
        const uint32_t mh = wave_min_u32(uint32_t(lo >> 32), lane);
        const uint32_t ml = wave_min_u32(uint32_t(lo >> 32) == mh ? uint32_t(lo) : ~0u, lane);
        const uint64_t win = __ballot(uint32_t(lo >> 32) == mh && uint32_t(lo) == ml);
        const int wl = __ffsll((unsigned long long)win) - 1;
        const uint32_t wxy = uint32_t(__builtin_amdgcn_readlane(int(lxy), wl));
        if (lane == r) {
            nk = (uint64_t(mh) << 32) | ml;
            nxy = wxy;
        }
        if (lane == wl) {
            lo = hi;
            lxy = hxy;
            hi = kNoKey;
        }
    }
    if (lane < 8) {
        L.key = nk < kPlaceholderKey ? nk : kPlaceholderKey;
        L.xy = nk < kPlaceholderKey ? nxy : 0u;
    }
}

template <int K, int J>
__device__ __forceinline__ void sort16_step(uint64_t &key, uint32_t &xy, int lane)
{
    const uint32_t ohi = dpp_xor<J>(uint32_t(key >> 32), lane);
    const uint32_t olo = dpp_xor<J>(uint32_t(key), lane);
    const uint32_t oxy = dpp_xor<J>(xy, lane);
    const uint64_t ok = (uint64_t(ohi) << 32) | olo;
    const bool asc = ((lane & 15) & K) == 0;
    const bool lower = (lane & J) == 0;
    const bool lt = ok < key, gt = ok > key;
    const bool take = (lt && lower == asc) || (gt && lower != asc);
    // selects, not a branch: three v_cndmask under one lane mask
    key = (uint64_t(take ? ohi : uint32_t(key >> 32)) << 32) | (take ? olo : uint32_t(key));
    xy = take ? oxy : xy;
}

// bitonic sort, ascending, of the 16 keys of each 16-lane row
__device__ __forceinline__ void sort16_dpp(uint64_t &key, uint32_t &xy, int lane)
{
    sort16_step<2, 1>(key, xy, lane);
    sort16_step<4, 2>(key, xy, lane);
    sort16_step<4, 1>(key, xy, lane);
    sort16_step<8, 4>(key, xy, lane);
    sort16_step<8, 2>(key, xy, lane);
    sort16_step<8, 1>(key, xy, lane);
    sort16_step<16, 8>(key, xy, lane);
    sort16_step<16, 4>(key, xy, lane);
    sort16_step<16, 2>(key, xy, lane);
    sort16_step<16, 1>(key, xy, lane);
}

// Row keys of the spatial selection: the bits of a distance (a non-negative
// float, so its bits order as the distance does) with the low bits replaced
// by the candidate's position in its row (pm = the position mask).  Keys of
// one row are distinct; a key bounds its distance's bits from below by
// key & ~pm and from above by key | pm.
__device__ __forceinline__ uint32_t row_key(float d, uint32_t k, uint32_t pm)
{
    return (__float_as_uint(d) & ~pm) | k;
}

__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c)
{
    return max(min(a, b), min(max(a, b), c));
}

// position mask for rows of up to n candidates
__device__ __forceinline__ uint32_t pos_mask(int n)
{
    return n <= 2 ? 1u : (2u << (31 - __clz(n - 1))) - 1u;
}

constexpr uint32_t kFltMaxBits = 0x7F7FFFFFu;  // distances at or above FLT_MAX never count

// distance row entries: the upper 16 bits of a distance (a lower bound of
// its bits) or the distance itself
__device__ __forceinline__ void store_dist(uint16_t *p, float d) { *p = uint16_t(__float_as_uint(d) >> 16); }
__device__ __forceinline__ void store_dist(float *p, float d) { *p = d; }

// floor(n / d) for 0 <= n < 2^16 and d >= 1, given inv = 1.0f / d: (n + 1/2)
// / d lies at least 1/(2d) from an integer, while the float product is off
// by at most 2^-23 (n + 1/2) / d < 2^-7 / d, so truncation gives the exact
// quotient
__device__ __forceinline__ int div_small(int n, float inv) { return int((float(n) + 0.5f) * inv); }

// whether a distance (by its bits) is at or below the threshold bits and
// below FLT_MAX (the shader's strict '<' never takes FLT_MAX; NaN never counts)
__device__ __forceinline__ bool survives(float d, uint32_t thr)
{
    const uint32_t b = __float_as_uint(d);
    return b <= thr && b < kFltMaxBits;
}

// The temporal steps' entries into the group in the shader's order
// (direction -1 for t = 1..r, then +1; entries i < ps_num of each step,
// src/bm3d.comp:624-628).  Inserting entries one by one with ginsert8's
// strict '<' leaves the 8 smallest of the group and the entries by
// (distance, rank), where the group's members rank first in their order and
// the entries after them in insertion order; so 8 entries at a time are
// merged by one 16-lane sort of (distance bits, rank, z) keys: lanes 0..7
// the group, lanes 8..15 the entries.  fkey[t - 1] / fxy[t - 1] hold step
// t's sorted (distance, seq) keys and positions of direction h in lanes
// h * 32 + i (radius r <= 2).
__device__ __forceinline__ void insert_steps(GroupList &G, const uint64_t (&fkey)[2], const uint32_t (&fxy)[2], int r,
                                             int PN, int lane)
{
    const int n_ins = 2 * r * PN;
    for (int c0 = 0; c0 < n_ins; c0 += 8) {
        const int e_idx = c0 + (lane - 8);
        const bool entry = lane >= 8 && lane < 16 && e_idx < n_ins;
        const int ei = entry ? e_idx : 0;
        const int ht = div_small(ei, 1.0f / float(PN)), i = ei - ht * PN;
        const int h = div_small(ht, 1.0f / float(r)), t = ht - h * r + 1;
        const int src = h * 32 + i;
        const uint32_t e1 = uint32_t(__shfl(int(uint32_t(fkey[0] >> 32)), src));
        const uint32_t e2 = uint32_t(__shfl(int(uint32_t(fkey[1] >> 32)), src));
        const uint32_t x1 = uint32_t(__shfl(int(fxy[0]), src));
        const uint32_t x2 = uint32_t(__shfl(int(fxy[1]), src));
        uint64_t key = kNoKey;
        uint32_t kxy = 0u;
        if (lane < 8) {
            key = (uint64_t(__float_as_uint(G.e)) << 32) | (uint32_t(lane) << 8) | uint32_t(G.z);
            kxy = G.xy;
        } else if (entry) {
            const uint32_t z = uint32_t(h ? r + t : r - t);
            key = (uint64_t(t == 1 ? e1 : e2) << 32) | (uint32_t(lane) << 8) | z;
            kxy = t == 1 ? x1 : x2;
        }
        sort16_dpp(key, kxy, lane);
        if (lane < 8) {
            G.e = __uint_as_float(uint32_t(key >> 32));
            G.xy = kxy;
            G.z = int(key & 0xFFu);
        }
    }
}

}  // namespace

}  // namespace vsfeel
