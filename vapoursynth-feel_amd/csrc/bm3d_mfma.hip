// Precision form of the BM3Dv2 block matching on gfx950 matrix cores.
//
//   bm3d_planes      per search frame: its pixels minus kPlaneCentre rounded
//                    once to FP16 or BF16 (s), their 8-column sums as hi + lo
//                    and their 8x8 sums of squares (the plane set of
//                    bm3d_common.hpp)
//   bm3d_match_mfma  spatial search of the sliding metric (radius < 3), a run
//                    of ref blocks per workgroup, one block per wavefront at a
//                    time
//
// The distance of candidate k of a window row is the reference's sum of 8
// column SSDs, column c against reference column min(k + c, 7)
// (src/bm3d.comp:268-282, :497-511).  With the pixels rounded once -- s for
// the frame, R_j for reference column j: R_7 = c7 (the block's own rounded
// column 7), R_j = c7 + d_j for j < 7 with d_j the rounded difference of the
// block's rounded column j and c7 -- the distance of the rounded blocks is
//
//     D~(k) = |R_k|^2 + box(cx, cy) - 2 P(k)
//     P(k)  = sum_i c7[i] h(cx, cy + i) + sum_{j=k}^{6} sum_i d_j[i] s(left + j, cy + i)
//
// with |R_k|^2 = sum_c |R_min(k+c,7)|^2 (the second sum of P only for k < 7).
// Both sums of P are 16x16 MFMA tiles, window rows as M and window columns
// as N: the first with a Toeplitz A (row m holds c7 shifted to the row's
// offset) against 8-row segments of h (hi and lo), the second with Toeplitz
// rows of every d_j against 8-row segments of s, masked to j >= k.  Products
// of the rounded operands are exact and accumulate in FP32, so D~ is, up to
// FP32 rounding, the distance of the rounded blocks.  The 16 smallest D~ (2k
// for the top-8) are rescored with the plain form's FP32 column sums and
// summation tree and sorted by (distance, seq): the group equals the plain
// form's wherever the plain top-8 lies in that shortlist.
//
// Inside a tile, MFMA output row m = 4g + r (lane group g, accumulator
// register r) is window row g + 4r, so the four values a lane holds lie four
// rows apart: a lane's 3 smallest keys then rarely miss a member of a
// cluster of good matches.

#include "bm3d_device.hpp"

namespace vsfeel {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// D += A B for one 16x16x32 tile: lane l holds A row l % 16 and B column
// l % 16 for K 8 (l / 16) .. +7 (8 operands, the lower index in the lower
// half of each dword); D row 4 (l / 16) + r, column l % 16 in register r
template <int OP>
__device__ __forceinline__ f32x4 mfma16(const u32x4 &a, const u32x4 &b, const f32x4 &c);

template <>
__device__ __forceinline__ f32x4 mfma16<kOperandFp16>(const u32x4 &a, const u32x4 &b, const f32x4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                  0, 0, 0);
}

template <>
__device__ __forceinline__ f32x4 mfma16<kOperandBf16>(const u32x4 &a, const u32x4 &b, const f32x4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// rounding to the operand type (to nearest, ties to even) and back
template <int OP>
__device__ __forceinline__ uint32_t to_op(float f);

template <>
__device__ __forceinline__ uint32_t to_op<kOperandFp16>(float f)
{
    return uint32_t(__builtin_bit_cast(uint16_t, static_cast<_Float16>(f)));
}

template <>
__device__ __forceinline__ uint32_t to_op<kOperandBf16>(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

template <int OP>
__device__ __forceinline__ float from_op(uint32_t h);

template <>
__device__ __forceinline__ float from_op<kOperandFp16>(uint32_t h)
{
    return float(__builtin_bit_cast(_Float16, uint16_t(h)));
}

template <>
__device__ __forceinline__ float from_op<kOperandBf16>(uint32_t h)
{
    return __uint_as_float(h << 16);
}

// 8 operands from consecutive 16-bit LDS elements (any even byte address)
__device__ __forceinline__ u32x4 load8(const uint16_t *p)
{
    u32x4 v;
    v[0] = uint32_t(p[0]) | (uint32_t(p[1]) << 16);
    v[1] = uint32_t(p[2]) | (uint32_t(p[3]) << 16);
    v[2] = uint32_t(p[4]) | (uint32_t(p[5]) << 16);
    v[3] = uint32_t(p[6]) | (uint32_t(p[7]) << 16);
    return v;
}

// ---------------------------------------------------------------------------
// Plane set of one search frame (layout in bm3d_common.hpp).  A workgroup
// makes a 32-column x 64-row tile: the rounded pixels of the tile and its
// 7-pixel margins in LDS (as floats), the 8-column sums of their squares,
// then per element (a lane per row, a wavefront per 8 columns) s, h (hi, lo)
// and box, written to both row copies.  Rows past the frame get zeros.
// ---------------------------------------------------------------------------
constexpr int kPlTileX = 32, kPlTileY = 64;

template <int OP>
__global__ __launch_bounds__(256) void bm3d_planes(const float *__restrict__ src, int pitch, int width, int height,
                                                   uint16_t *__restrict__ set, int py)
{
    __shared__ float sv[kPlTileX + 7][kPlTileY + 8];
    __shared__ float sq[kPlTileX][kPlTileY + 8];
    const int x0 = blockIdx.x * kPlTileX, y0 = blockIdx.y * kPlTileY;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kRows = kPlTileY + 7;
    for (int i = threadIdx.x; i < (kPlTileX + 7) * kRows; i += 256) {
        const int c = i / kRows, r = i - c * kRows;
        const int x = x0 + c, y = y0 + r;
        float v = 0.0f;
        if (x < width && y < height) {
            v = from_op<OP>(to_op<OP>(src[(long long)y * pitch + x] - kPlaneCentre));
        }
        sv[c][r] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kPlTileX * kRows; i += 256) {
        const int c = i / kRows, r = i - c * kRows;
        float a = 0.0f;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float v = sv[c + u][r];
            a = __builtin_fmaf(v, v, a);
        }
        sq[c][r] = a;
    }
    __syncthreads();
    const long long E = (long long)width * py;
    uint16_t *s0 = set, *s1 = set + E, *hh0 = set + 2 * E, *hh1 = set + 3 * E, *hl0 = set + 4 * E,
             *hl1 = set + 5 * E;
    float *box = reinterpret_cast<float *>(set + 6 * E);
    const int y = y0 + lane;
    if (y >= py) {
        return;
    }
    for (int cc = 0; cc < kPlTileX / 4; ++cc) {
        const int c = wave * (kPlTileX / 4) + cc;
        const int x = x0 + c;
        if (x >= width) {
            break;
        }
        const float v = sv[c][lane];
        float h = 0.0f, b = 0.0f;
        if (x <= width - 8 && y < height) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                h += sv[c + u][lane];
            }
        }
        if (x <= width - 8 && y <= height - 8) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                b += sq[c][lane + i];
            }
        }
        const uint32_t hs = to_op<OP>(v);
        const uint32_t hi = to_op<OP>(h);
        const uint32_t lo = to_op<OP>(h - from_op<OP>(hi));
        const long long o = (long long)x * py + y;
        s0[o] = uint16_t(hs);
        hh0[o] = uint16_t(hi);
        hl0[o] = uint16_t(lo);
        box[o] = b;
        if (y >= 1) {
            s1[o - 1] = uint16_t(hs);
            hh1[o - 1] = uint16_t(hi);
            hl1[o - 1] = uint16_t(lo);
        }
    }
}

// ---------------------------------------------------------------------------
// Per-block operands.  pad row j (j < 7: d_j, j = 7: c7) holds the column at
// elements 16..23 and zeros around it, so that 8 consecutive elements from
// 16 + o - rho are a Toeplitz row segment.
// ---------------------------------------------------------------------------
constexpr int kPadLen = 48;
constexpr int kCorrSteps = 6;  // K = 7 columns x 3 segments of 8 rows, 4 segments per MFMA

struct BlockOperands {
    u32x4 a_main;              // c7 Toeplitz rows for the lane's tile row
    u32x4 a_corr[kCorrSteps];  // d_j Toeplitz rows, segment 4 step + g
    float nr_first;            // |R_k|^2 for the lane's column k = n of the first column tile
    float nr7;                 // |R_k|^2 for k >= 7 (8 |c7|^2)
};

// ref: the block's rounded columns (column c at ref[c * cstride], 8 rows)
template <int OP>
__device__ __forceinline__ BlockOperands block_operands(const uint16_t *ref, int cstride, uint16_t *pad, int lane)
{
    const int g = lane >> 4, n = lane & 15;
    const int mrow = lane & 15;
    const int rho = (mrow >> 2) + 4 * (mrow & 3);
    {
        const int j = lane >> 3, u = lane & 7;
        const float c7 = from_op<OP>(ref[7 * cstride + u]);
        const float cj = from_op<OP>(ref[j * cstride + u]);
        const uint32_t v = j == 7 ? to_op<OP>(c7) : to_op<OP>(cj - c7);
        for (int e = lane; e < 8 * kPadLen; e += 64) {
            pad[e] = 0;
        }
        wave_lds_sync();
        pad[j * kPadLen + 16 + u] = uint16_t(v);
        wave_lds_sync();
    }
    BlockOperands o;
    // |R_j|^2 in lane j (R_j = c7 + d_j exactly in FP32)
    float e_own = 0.0f;
    {
        const int j = lane & 7;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float c7 = from_op<OP>(pad[7 * kPadLen + 16 + u]);
            const float r = j == 7 ? c7 : c7 + from_op<OP>(pad[j * kPadLen + 16 + u]);
            e_own = __builtin_fmaf(r, r, e_own);
        }
    }
    float ev[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        ev[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e_own), j));
    }
    o.nr7 = 8.0f * ev[7];
    o.nr_first = o.nr7;
    float suffix = 0.0f;
#pragma unroll
    for (int j = 6; j >= 0; --j) {
        suffix += ev[j];
        if (n == j) {
            o.nr_first = suffix + float(j + 1) * ev[7];
        }
    }
    o.a_main = load8(pad + 7 * kPadLen + 16 + 8 * g - rho);
#pragma unroll
    for (int st = 0; st < kCorrSteps; ++st) {
        const int c = 4 * st + g;
        if (c < 21) {
            const int jj = c / 3, tq = c - 3 * jj;
            o.a_corr[st] = load8(pad + jj * kPadLen + 16 + 8 * tq - rho);
        } else {
            o.a_corr[st] = u32x4{0u, 0u, 0u, 0u};
        }
    }
    return o;
}

// The keys of one tile's 4 accumulator registers into the lane's ascending
// 3 smallest: D~ = box + |R_k|^2 - 2 P (clamped at 0), key = its bits with
// the low bits (mask pm) replaced by seq = seq_base + row * rw + col; candidates outside
// the window never enter.  Registers whose every row lies past the window
// are skipped (wave-uniform).
__device__ __forceinline__ void tile_keys(const f32x4 &acc, const f32x4 &box, float nr, int m0, int g, int col,
                                          int rw, int rh, uint32_t seq_base, uint32_t pm, uint32_t &m1, uint32_t &m2,
                                          uint32_t &m3)
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (m0 + 4 * r >= rh) {
            break;
        }
        const int row = m0 + g + 4 * r;
        const float d = fmaxf(__builtin_fmaf(acc[r], -2.0f, box[r] + nr), 0.0f);
        uint32_t key = (__float_as_uint(d) & ~pm) | (seq_base + uint32_t(row * rw + col));
        key = (row < rh && col < rw) ? key : ~0u;
        m3 = med3_u32(m2, m3, key);
        m2 = med3_u32(m1, m2, key);
        m1 = min(m1, key);
    }
}

// The want smallest keys of the lanes' lists (keys are distinct; ~0u never
// counts): bisection on the key value, then their seqs (key & pm) to
// stage[0 .. want).  Returns want = min(limit, valid keys).
__device__ __forceinline__ int shortlist(uint32_t m1, uint32_t m2, uint32_t m3, int limit, uint32_t pm,
                                         uint32_t *stage)
{
    auto count_le = [&](uint32_t t) -> int {
        return __popcll(__ballot(m1 <= t)) + __popcll(__ballot(m2 <= t)) + __popcll(__ballot(m3 <= t));
    };
    const int want = min(limit, count_le(0xFFFFFFFEu));
    uint32_t lo = 0u, hi = 0xFFFFFFFEu;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (count_le(mid) >= want) {
            hi = mid;
        } else {
            lo = mid + 1u;
        }
    }
    int base = 0;
    const uint32_t e[3] = {m1, m2, m3};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const bool sel = e[q] <= lo;
        const uint64_t bal = __ballot(sel);
        const int p = base + int(__builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0)));
        if (sel && p < want) {
            stage[p] = e[q] & pm;
        }
        base += __popcll(bal);
    }
    wave_lds_sync();
    return want;
}

// lane i takes lane i + S of its 16-lane row (0 past the row's end)
template <int S>
__device__ __forceinline__ float dpp_shl(float v)
{
    return __uint_as_float(dpp_mov<0x100 + S>(__float_as_uint(v)));
}

// Exact FP32 distance of one candidate by 8 lanes (the plain form's column
// SSDs and summation tree, src/bm3d.comp:268-282): lane j = l % 8 takes the
// SSD of candidate column j (cand: its top-left pixel) against reference
// column min(k + j, 7) (ref: the reference block's top-left pixel; k the
// candidate's position in its window row), then the candidate's 8 lanes add
// them in the reference's tree; the distance is valid in lane j = 0.
template <int CM>
__device__ __forceinline__ float exact_dist8(const float *ref, const float *cand, int pitch, int k, int jcol)
{
    const int a = min(k + jcol, 7);
    float cc[8], sv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        cc[i] = ref[(long long)i * pitch + a];
        sv[i] = cand[(long long)i * pitch + jcol];
    }
    float v = col_ssd<CM>(cc, sv);
    v = v + dpp_shl<1>(v);
    v = v + dpp_shl<2>(v);
    v = v + dpp_shl<4>(v);
    return v;
}

// The key of a rescored candidate: (distance bits, seq), or the placeholder
// where the distance is not below FLT_MAX (the shader's strict '<' never
// takes it)
__device__ __forceinline__ void rescored_key(float d, uint32_t seq, int cx, int cy, unsigned long long &key,
                                             uint32_t &xy)
{
    const bool valid = __float_as_uint(d) < kFltMaxBits;
    key = valid ? make_key(d, int(seq)) : kPlaceholderKey;
    xy = valid ? (uint32_t(cx) | (uint32_t(cy) << 16)) : 0u;
}

constexpr int kMfmaWaves = 4;

// ---------------------------------------------------------------------------
// Spatial search of the precision form (radius < 3).  A workgroup takes a run
// of mfma_run_blocks consecutive ref blocks of one grid row; their windows
// share the rows [top, bottom], so the run's region of h (hi, lo) and s is
// staged once in LDS, column-major with segments of 8 rows starting at the
// window's top row (16-byte aligned), and box with its rows in the tiles'
// lane order.  Each wavefront takes one block at a time: its Toeplitz
// operands, the window's tiles (rows and columns by 16; the correction for
// k < 7 in the first column tile), each lane's 3 smallest keys, the
// shortlist of 16, the exact rescoring and a 16-lane sort; then the group
// (radius 0, with the self-block rule) or the spatial top-8 for the
// temporal search.
// ---------------------------------------------------------------------------
template <int CM, int OP>
__global__ __launch_bounds__(256) void bm3d_match_mfma(MatchArgs A)
{
    extern __shared__ __align__(16) unsigned long long mlds[];
    const int NB = A.mfma_run_blocks;
    const int W = A.width, H = A.height, R = A.bm_range, s = A.step;
    const int run = xcd_block_id();
    const int by = run / A.mfma_runs_x;
    const int bx0 = (run - by * A.mfma_runs_x) * NB;
    const int nb = min(NB, A.nx - bx0);
    const int y = min(by * s, H - 8);
    const int top = max(y - R, 0), bot = min(y + R, H - 8);
    const int rh = bot - top + 1;
    const int x_first = min(bx0 * s, W - 8), x_last = min((bx0 + nb - 1) * s, W - 8);
    const int rx0 = max(x_first - R, 0);
    const int rx1 = min(x_last + R, W - 8) + 7;
    const int ncols = rx1 - rx0 + 1, nhc = ncols - 7;
    const int CR = mfma_region_rows(rh), CRB = mfma_box_rows(rh);
    float *boxr = reinterpret_cast<float *>(mlds);
    unsigned char *wbase = reinterpret_cast<unsigned char *>(boxr + nhc * CRB);
    uint16_t *hhi = reinterpret_cast<uint16_t *>(wbase + kMfmaWaves * kMfmaWaveBytes);
    uint16_t *hlo = hhi + nhc * CR;
    uint16_t *sr = hlo + nhc * CR;
    const int z = A.radius;
    const uint16_t *set = A.pset[z];
    const long long E = A.pset_elems;
    const int py = A.pset_py;
    const int par = top & 1;
    {
        // h (hi, lo) and s: dword pairs of rows from the row copy whose
        // element top - par holds row top
        const int dw = CR / 2;
        const uint32_t *gs = reinterpret_cast<const uint32_t *>(set + par * E);
        const uint32_t *gh = reinterpret_cast<const uint32_t *>(set + (2 + par) * E);
        const uint32_t *gl = reinterpret_cast<const uint32_t *>(set + (4 + par) * E);
        uint32_t *ls = reinterpret_cast<uint32_t *>(sr);
        uint32_t *lh = reinterpret_cast<uint32_t *>(hhi);
        uint32_t *ll = reinterpret_cast<uint32_t *>(hlo);
        for (int i = threadIdx.x; i < ncols * dw; i += 256) {
            const int c = i / dw, q = i - c * dw;
            const long long o = ((long long)(rx0 + c) * py + top - par) / 2 + q;
            ls[i] = gs[o];
            if (c < nhc) {
                lh[i] = gh[o];
                ll[i] = gl[o];
            }
        }
        // box: region row 16 b + 4 g + r holds window row 16 b + g + 4 r
        const float *gb = reinterpret_cast<const float *>(set + 6 * E);
        for (int i = threadIdx.x; i < nhc * CRB; i += 256) {
            const int c = i / CRB, p = i - c * CRB;
            const int row = (p & ~15) + ((p >> 2) & 3) + 4 * (p & 3);
            boxr[i] = row < rh ? gb[(long long)(rx0 + c) * py + top + row] : 0.0f;
        }
    }
    __syncthreads();

    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    unsigned char *wl = wbase + wave * kMfmaWaveBytes;
    uint16_t *pad = reinterpret_cast<uint16_t *>(wl);
    uint32_t *st_seq = reinterpret_cast<uint32_t *>(wl + 8 * kPadLen * 2);
    unsigned long long *st_key = reinterpret_cast<unsigned long long *>(wl + 8 * kPadLen * 2 + 64);
    uint32_t *st_xy = reinterpret_cast<uint32_t *>(wl + 8 * kPadLen * 2 + 64 + 128);
    const int g = lane >> 4, n = lane & 15;
    const float *plane = A.plane[z];
    for (int b = wave; b < nb; b += kMfmaWaves) {
        const int bx = bx0 + b;
        const int xb = min(bx * s, W - 8);
        const int left = max(xb - R, 0), right = min(xb + R, W - 8);
        const int rw = right - left + 1;
        const BlockOperands o = block_operands<OP>(sr + (xb - rx0) * CR + (y - top), CR, pad, lane);
        const uint32_t pm = pos_mask(rw * rh);
        uint32_t m1 = ~0u, m2 = ~0u, m3 = ~0u;
        for (int m0 = 0; m0 < rh; m0 += 16) {
            const int rbase = min(m0 + 8 * g, CR - 8);
            for (int n0 = 0; n0 < rw; n0 += 16) {
                const int hc = min(left + n0 + n - rx0, nhc - 1);
                const u32x4 bh = *reinterpret_cast<const u32x4 *>(hhi + hc * CR + rbase);
                const u32x4 bl = *reinterpret_cast<const u32x4 *>(hlo + hc * CR + rbase);
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma16<OP>(o.a_main, bh, acc);
                acc = mfma16<OP>(o.a_main, bl, acc);
                if (n0 == 0) {
#pragma unroll
                    for (int st = 0; st < kCorrSteps; ++st) {
                        const int c = 4 * st + g;
                        const int jj = c / 3, tq = c - 3 * jj;
                        u32x4 bs = {0u, 0u, 0u, 0u};
                        if (c < 21 && jj >= n) {
                            bs = *reinterpret_cast<const u32x4 *>(sr + (left + jj - rx0) * CR + min(m0 + 8 * tq, CR - 8));
                        }
                        acc = mfma16<OP>(o.a_corr[st], bs, acc);
                    }
                }
                const f32x4 bx4 = *reinterpret_cast<const f32x4 *>(boxr + hc * CRB + m0 + 4 * g);
                tile_keys(acc, bx4, n0 == 0 ? o.nr_first : o.nr7, m0, g, n0 + n, rw, rh, 0u, pm, m1, m2, m3);
            }
        }
        const int want = shortlist(m1, m2, m3, 16, pm, st_seq);
        {
            // rescoring, 8 candidates per pass
            const float inv_rw = 1.0f / float(rw);
            const float *refp = plane + (long long)y * A.pitch + xb;
            for (int p0 = 0; p0 < want; p0 += 8) {
                const int ci = p0 + (lane >> 3);
                const bool on = ci < want;
                const int sq = on ? int(st_seq[ci]) : 0;
                const int row = div_small(sq, inv_rw);
                const int k = sq - row * rw;
                const int cx = left + k, cy = top + row;
                const float d = exact_dist8<CM>(refp, plane + (long long)cy * A.pitch + cx, A.pitch, k, lane & 7);
                if (on && (lane & 7) == 0) {
                    rescored_key(d, uint32_t(sq), cx, cy, st_key[ci], st_xy[ci]);
                }
            }
            wave_lds_sync();
        }
        uint64_t key = lane < want ? st_key[lane] : kNoKey;
        uint32_t kxy = lane < want ? st_xy[lane] : 0u;
        wave_lds_sync();
        sort16_dpp(key, kxy, lane);
        if (!(key < kPlaceholderKey)) {
            key = kPlaceholderKey;
            kxy = 0u;
        }
        const int blk = by * A.nx + bx;
        if (A.spatial) {
            if (lane < 8) {
                A.spatial[(long long)blk * 8 + lane] = make_uint2(uint32_t(key >> 32), kxy);
            }
        } else {
            finish_group16(A, __uint_as_float(uint32_t(key >> 32)), kxy, blk, xb, y, lane < 16 ? lane : 16);
        }
    }
}

// 8 operands (16 bytes) from a plane of the plane set at an even element:
// a 4-byte aligned load
__device__ __forceinline__ u32x4 load16g(const uint16_t *p)
{
    u32x4 v;
    __builtin_memcpy(&v, reinterpret_cast<const uint32_t *>(p), 16);
    return v;
}

// The keys of one temporal window (left, top, rw x rh candidates, seq base
// seq_base) into the lane's 3 smallest, its operands read from the frame's
// plane set in global memory: the row copy of the window's top row parity
// makes every 8-row segment start at an even element.
template <int OP>
__device__ __forceinline__ void window_keys(const BlockOperands &o, const uint16_t *set, long long E, int py, int W,
                                            int left, int top, int rw, int rh, uint32_t seq_base, uint32_t pm,
                                            int lane, uint32_t &m1, uint32_t &m2, uint32_t &m3)
{
    const int g = lane >> 4, n = lane & 15;
    const int par = top & 1;
    const uint16_t *sc = set + par * E;
    const uint16_t *hh = set + (2 + par) * E;
    const uint16_t *hl = set + (4 + par) * E;
    const float *box = reinterpret_cast<const float *>(set + 6 * E);
    for (int m0 = 0; m0 < rh; m0 += 16) {
        for (int n0 = 0; n0 < rw; n0 += 16) {
            const int hx = min(left + n0 + n, W - 8);
            const long long ob = (long long)hx * py + top + m0 + 8 * g - par;
            const u32x4 bh = load16g(hh + ob);
            const u32x4 bl = load16g(hl + ob);
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma16<OP>(o.a_main, bh, acc);
            acc = mfma16<OP>(o.a_main, bl, acc);
            if (n0 == 0) {
#pragma unroll
                for (int st = 0; st < kCorrSteps; ++st) {
                    const int c = 4 * st + g;
                    const int jj = c / 3, tq = c - 3 * jj;
                    u32x4 bs = {0u, 0u, 0u, 0u};
                    if (c < 21 && jj >= n) {
                        bs = load16g(sc + (long long)(left + jj) * py + top + m0 + 8 * tq - par);
                    }
                    acc = mfma16<OP>(o.a_corr[st], bs, acc);
                }
            }
            f32x4 bx4;
            const float *bp = box + (long long)hx * py + top + m0 + g;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                bx4[r] = bp[4 * r];
            }
            tile_keys(acc, bx4, n0 == 0 ? o.nr_first : o.nr7, m0, g, n0 + n, rw, rh, seq_base, pm, m1, m2, m3);
        }
    }
}

// ---------------------------------------------------------------------------
// Temporal predictive search of the precision form (radius 1-2,
// src/bm3d.comp:540-637) after bm3d_match_mfma stored the spatial top-8: one
// wavefront per ref block.  Each step t and direction h (-1, then +1) tiles
// the ps_num predictor windows of frame r -+ t (seq base i (2 ps_range + 1)^2
// for window i) into the lane's 3 smallest keys of that direction; the
// shortlist of 2 ps_num per direction is rescored exactly and sorted (both
// directions at once, one 16-lane row each); its first ps_num entries are the
// next step's predictors and the group's insertions, added afterwards in
// the shader's order like the plain form's.
// ---------------------------------------------------------------------------
template <int CM, int OP>
__global__ __launch_bounds__(64) void bm3d_match_predictive_mfma(MatchArgs A)
{
    __shared__ uint16_t pad[8 * kPadLen];
    __shared__ uint32_t st_seq[2][16];
    __shared__ unsigned long long st_key[2][16];
    __shared__ uint32_t st_xy[2][16];
    const int lane = lane_id();
    const int blk = xcd_block_id();
    if (blk >= A.nx * A.ny) {
        return;
    }
    const int by = blk / A.nx;
    const int bx = blk - by * A.nx;
    const int W = A.width, H = A.height, pitch = A.pitch, r = A.radius;
    const int x = min(bx * A.step, W - 8);
    const int y = min(by * A.step, H - 8);
    const int PR = A.ps_range, PN = A.ps_num;
    const int wh = 2 * PR + 1, win = wh * wh;
    const int SL = min(2 * PN, 16);
    const long long E = A.pset_elems;
    const int py = A.pset_py;
    const uint32_t pm = pos_mask(PN * win);
    const BlockOperands o = block_operands<OP>(A.pset[r] + (long long)x * py + y, py, pad, lane);
    const float *refp = A.plane[r] + (long long)y * pitch + x;

    // the spatial top-8 starts the group and predicts (lanes h * 16 + i:
    // predictor i of direction h)
    GroupList G;
    uint32_t pred;
    {
        const uint2 e = A.spatial[(long long)blk * 8 + (lane & 7)];
        G.e = __uint_as_float(e.x);
        G.xy = e.y;
        G.z = r;
        pred = e.y;
    }
    uint64_t fkey[2] = {kNoKey, kNoKey};
    uint32_t fxy[2] = {0u, 0u};
    for (int t = 1; t <= r; ++t) {
        int want[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int z = h ? r + t : r - t;
            uint32_t m1 = ~0u, m2 = ~0u, m3 = ~0u;
            for (int p = 0; p < PN; ++p) {
                const uint32_t q = uint32_t(__shfl(int(pred), h * 16 + p));
                const int cx = int(q & 0xFFFFu), cy = int(q >> 16);
                const int left = max(cx - PR, 0), top = max(cy - PR, 0);
                const int rw = min(cx + PR, W - 8) - left + 1;
                const int rh = min(cy + PR, H - 8) - top + 1;
                window_keys<OP>(o, A.pset[z], E, py, W, left, top, rw, rh, uint32_t(p * win), pm, lane, m1, m2, m3);
            }
            want[h] = shortlist(m1, m2, m3, SL, pm, st_seq[h]);
        }
        // rescoring, 8 candidates per pass: slot ci = h * SL + q
        const float inv_win = 1.0f / float(win);
        for (int p0 = 0; p0 < 2 * SL; p0 += 8) {
            const int ci = p0 + (lane >> 3);
            const int h = ci >= SL ? 1 : 0, q = ci - h * SL;
            const bool on = ci < 2 * SL && q < want[h];
            const int sq = on ? int(st_seq[h][q]) : 0;
            const int w = div_small(sq, inv_win), local = sq - w * win;
            const uint32_t pq = uint32_t(__shfl(int(pred), h * 16 + w));
            const int pcx = int(pq & 0xFFFFu), pcy = int(pq >> 16);
            const int left = max(pcx - PR, 0), top = max(pcy - PR, 0);
            const int rw = min(pcx + PR, W - 8) - left + 1;
            const int row = div_small(local, 1.0f / float(rw));
            const int k = local - row * rw;
            const int cx = left + k, cy = top + row;
            const float *plane = A.plane[h ? r + t : r - t];
            const float d = exact_dist8<CM>(refp, plane + (long long)cy * pitch + cx, pitch, k, lane & 7);
            if (on && (lane & 7) == 0) {
                rescored_key(d, uint32_t(sq), cx, cy, st_key[h][q], st_xy[h][q]);
            }
        }
        wave_lds_sync();
        // both directions sorted at once: lanes h * 16 + q
        const int hs = (lane >> 4) & 1, qs = lane & 15;
        uint64_t key = lane < 32 && qs < want[hs] ? st_key[hs][qs] : kNoKey;
        uint32_t kxy = lane < 32 && qs < want[hs] ? st_xy[hs][qs] : 0u;
        wave_lds_sync();
        sort16_dpp(key, kxy, lane);
        if (!(key < kPlaceholderKey)) {
            key = kPlaceholderKey;  // fewer candidates than the step keeps
            kxy = 0u;
        }
        pred = kxy;
        // lanes h * 32 + i: step t's entry i of direction h
        const int src = (lane >> 5) * 16 + (lane & 15);
        fkey[t - 1] = (uint64_t(uint32_t(__shfl(int(uint32_t(key >> 32)), src))) << 32) |
                      uint32_t(__shfl(int(uint32_t(key)), src));
        fxy[t - 1] = uint32_t(__shfl(int(kxy), src));
    }
    insert_steps(G, fkey, fxy, r, PN, lane);
    finish_group(A, G, blk, x, y, lane);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_planes(const float *src, int pitch, int width, int height, uint16_t *set, int operand,
                         hipStream_t s)
{
    const int py = plane_py(height);
    const dim3 grid((width + kPlTileX - 1) / kPlTileX, (py + kPlTileY - 1) / kPlTileY), block(256);
    if (operand == kOperandBf16) {
        hipLaunchKernelGGL((bm3d_planes<kOperandBf16>), grid, block, 0, s, src, pitch, width, height, set, py);
    } else {
        hipLaunchKernelGGL((bm3d_planes<kOperandFp16>), grid, block, 0, s, src, pitch, width, height, set, py);
    }
    return hipGetLastError();
}

hipError_t launch_match_mfma(const MatchArgs &a, int fp_mode, hipStream_t s)
{
    const int rh = 2 * a.bm_range + 1 < a.height - 7 ? 2 * a.bm_range + 1 : a.height - 7;
    const int span = (a.mfma_run_blocks - 1) * a.step + 8 + 2 * a.bm_range;
    const int cols = span < a.width ? span : a.width;
    const size_t lds = size_t(mfma_rows_lds_bytes(cols, rh));
    const dim3 grid(a.ny * a.mfma_runs_x), block(256);
    MatchArgs m = a;
    if (a.radius == 0) {
        m.spatial = nullptr;
    }
#define VSFEEL_MM(CM, OP) hipLaunchKernelGGL((bm3d_match_mfma<CM, OP>), grid, block, lds, s, m)
    if (fp_mode == kFpAll) {
        if (a.operand == kOperandBf16) {
            VSFEEL_MM(kFpAll, kOperandBf16);
        } else {
            VSFEEL_MM(kFpAll, kOperandFp16);
        }
    } else {
        if (a.operand == kOperandBf16) {
            VSFEEL_MM(kFpExplicit, kOperandBf16);
        } else {
            VSFEEL_MM(kFpExplicit, kOperandFp16);
        }
    }
#undef VSFEEL_MM
    return hipGetLastError();
}

hipError_t launch_predictive_mfma(const MatchArgs &a, int fp_mode, hipStream_t s)
{
    const dim3 grid(a.nx * a.ny), block(64);
#define VSFEEL_PM(CM, OP) hipLaunchKernelGGL((bm3d_match_predictive_mfma<CM, OP>), grid, block, 0, s, a)
    if (fp_mode == kFpAll) {
        if (a.operand == kOperandBf16) {
            VSFEEL_PM(kFpAll, kOperandBf16);
        } else {
            VSFEEL_PM(kFpAll, kOperandFp16);
        }
    } else {
        if (a.operand == kOperandBf16) {
            VSFEEL_PM(kFpExplicit, kOperandBf16);
        } else {
            VSFEEL_PM(kFpExplicit, kOperandFp16);
        }
    }
#undef VSFEEL_PM
    return hipGetLastError();
}

}  // namespace vsfeel
