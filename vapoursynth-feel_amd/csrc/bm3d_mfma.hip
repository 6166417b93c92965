// Precision form of the BM3Dv2 block matching on gfx950 matrix cores.
//
//   bm3d_planes                 per search frame: its pixels minus kPlaneCentre
//                               rounded once to FP16 or BF16 (s), their
//                               8-column sums as hi + lo and their 8x8 sums
//                               of squares (the plane set of bm3d_common.hpp)
//   bm3d_match_mfma             spatial search of the sliding metric (radius
//                               < 3), a run of ref blocks per workgroup, one
//                               block per wavefront at a time
//
// At radius 1-2 the temporal predictive search that follows keeps the plain
// form's FP32 distances (bm3d_match_predictive): its windows (ps_num windows
// of (2 ps_range + 1)^2 candidates per step) are too small for the MFMA
// operands to be staged once and shared, and gathering them per window from
// the plane set costs more than the FP32 search.
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
// offset) against 8-row segments of h (hi and lo); the second with the 8 x 8
// block of s right of the window's left edge at the row's offset as A (K =
// 8 rows x 8 columns) against B[(i, j)][k] = d_j[i] for j >= k (j < 7).
// Products of the rounded operands are exact and accumulate in FP32, so D~
// is, up to FP32 rounding, the distance of the rounded blocks.  The 16 smallest D~ (2k
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

// 8 operands from 8 consecutive 16-bit elements of LDS or global memory at
// any even byte address (this target's loads need no more alignment)
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
// 7-pixel margins in LDS (as floats) and the 8-column sums of their squares;
// then h (hi, lo) and box column-major (a lane per row, a wavefront per 8
// columns) and s row-major (a lane per column).  Rows past the frame get
// zeros.
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
    uint16_t *hhi = set, *hlo = set + E, *srm = set + 2 * E;
    float *box = reinterpret_cast<float *>(set + 3 * E);
    {
        // s, row-major: lanes 0..31 a column each, a wavefront per 16 rows
        const int c = lane & 31, x = x0 + c;
        for (int rr = (lane >> 5); rr < kPlTileY / 4; rr += 2) {
            const int r = wave * (kPlTileY / 4) + rr, y = y0 + r;
            if (x < width && y < py) {
                srm[(long long)y * width + x] = uint16_t(to_op<OP>(sv[c][r]));
            }
        }
    }
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
        const uint32_t hi = to_op<OP>(h);
        const uint32_t lo = to_op<OP>(h - from_op<OP>(hi));
        const long long o = (long long)x * py + y;
        hhi[o] = uint16_t(hi);
        hlo[o] = uint16_t(lo);
        box[o] = b;
    }
}

// ---------------------------------------------------------------------------
// Per-block operands.  The wavefront's scratch holds c7 at elements 16..23
// of a zero row (so that 8 elements from 16 + o - rho are a Toeplitz row
// segment) and d_j[i] transposed (row i: j = 0..6, then 0).
// ---------------------------------------------------------------------------
constexpr int kPadLen = 48;

struct BlockOperands {
    u32x4 a_main;    // -2 c7 Toeplitz row of the lane's tile row
    u32x4 b_d[2];    // -2 d_j[i] for j >= the lane's column k (K steps of 4 rows i)
    float nr_first;  // |R_k|^2 for the lane's column k = n of the first column tile
    float nr7;       // |R_k|^2 for k >= 7 (8 |c7|^2)
};

// lane i takes lane i + S of its 16-lane row (0 past the row's end)
template <int S>
__device__ __forceinline__ float dpp_shl(float v)
{
    return __uint_as_float(dpp_mov<0x100 + S>(__float_as_uint(v)));
}

// ref: the block's rounded pixels, row-major (pixel (c, i) at ref[i * rs + c]).
// The MFMA operands carry the factor -2 of -2 P (exact: a power of two), so
// that tiles accumulate onto box + |R_k|^2 and end at D~.
template <int OP>
__device__ __forceinline__ BlockOperands block_operands(const uint16_t *ref, int rs, uint16_t *scratch, int lane)
{
    const int g = lane >> 4, n = lane & 15;
    const int mrow = lane & 15;
    const int rho = (mrow >> 2) + 4 * (mrow & 3);
    uint16_t *c7pad = scratch;                 // [kPadLen]
    uint16_t *dT = scratch + kPadLen;          // [8 rows i][8 columns j]
    const int j = lane >> 3, u = lane & 7;     // lane: column j, row u of the block
    const uint32_t c7h = ref[u * rs + 7];
    const float c7 = from_op<OP>(c7h);
    const uint32_t d = j == 7 ? 0u : to_op<OP>(from_op<OP>(ref[u * rs + j]) - c7);
    // |R_j|^2 with R_j = c7 + d_j (exact in FP32) summed over the 8 rows
    const float rj = c7 + from_op<OP>(d);
    float e = rj * rj;
    e = e + dpp_shl<1>(e);
    e = e + dpp_shl<2>(e);
    e = e + dpp_shl<4>(e);
    if (lane < kPadLen) {
        c7pad[lane] = 0;
    }
    wave_lds_sync();
    if (j == 7) {
        c7pad[16 + u] = uint16_t(to_op<OP>(-2.0f * c7));
    }
    dT[u * 8 + j] = uint16_t(to_op<OP>(-2.0f * from_op<OP>(d)));
    wave_lds_sync();
    BlockOperands o;
    float ev[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        ev[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e), 8 * q));
    }
    o.nr7 = 8.0f * ev[7];
    o.nr_first = o.nr7;
    float suffix = 0.0f;
#pragma unroll
    for (int q = 6; q >= 0; --q) {
        suffix += ev[q];
        if (n == q) {
            o.nr_first = suffix + float(q + 1) * ev[7];
        }
    }
    o.a_main = load8(c7pad + 16 + 8 * g - rho);
    // B[(i, j)][n] = d_j[i] where j >= n: row i = 4 st + g of dT, the halves
    // of columns below n cleared
    u32x4 keep;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        keep[q] = (2 * q >= n ? 0x0000FFFFu : 0u) | (2 * q + 1 >= n ? 0xFFFF0000u : 0u);
    }
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        o.b_d[st] = load8(dT + (4 * st + g) * 8) & keep;
    }
    return o;
}

// The keys of one tile's 4 accumulator registers (D~ of the lane's 4
// candidates: rows m0 + g + 4 r, column col) into the lane's ascending T
// smallest (T = 2 or 3): the bits of D~ clamped at 0 with the low bits (mask
// pm) replaced by seq = seq_base + row * rw + col; candidates outside the
// window (edge tiles only) never enter.  Registers whose every row lies past
// the window are skipped (rh is wave-uniform).
template <int T>
__device__ __forceinline__ void tile_keys(const f32x4 &acc, int m0, int g, int col, int rw, int rh,
                                          uint32_t seq_base, uint32_t pm, uint32_t &m1, uint32_t &m2, uint32_t &m3)
{
    const bool edge = m0 + 16 > rh || col - (col & 15) + 16 > rw;
    const uint32_t col_out = col < rw ? 0u : ~0u;
    uint32_t seq = seq_base + uint32_t((m0 + g) * rw + col);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (m0 + 4 * r >= rh) {
            break;
        }
        // negative D~ (rounding near 0) clamps to +0 as a signed integer
        uint32_t key = (uint32_t(max(__float_as_int(acc[r]), 0)) & ~pm) | seq;
        if (edge) {
            key |= (m0 + g + 4 * r < rh ? col_out : ~0u);
        }
        if (T == 3) {
            m3 = med3_u32(m2, m3, key);
        }
        m2 = med3_u32(m1, m2, key);
        m1 = min(m1, key);
        seq += uint32_t(4 * rw);
    }
}

// The want smallest keys of the lanes' 3-entry lists (keys are distinct;
// ~0u never counts): bisection on the key value, then their seqs (key & pm)
// to stage[0 .. want).  Returns want = min(limit, valid keys).
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

// Exact FP32 distance of one candidate by 8 lanes (the plain form's column
// SSDs and summation tree, src/bm3d.comp:268-282): lane j = l % 8 takes the
// SSD of candidate column j against reference column min(k + j, 7) (k the
// candidate's position in its window row), then the candidate's 8 lanes add
// them in the reference's tree; the distance is valid in lane j = 0.  ref /
// cand: the planes (wave-uniform pointers) and the element offsets of the
// reference block's and the candidate's top-left pixels (planes below 2 GiB).
template <int CM>
__device__ __forceinline__ float exact_dist8(const float *ref, uint32_t ref_off, const float *cand, uint32_t cand_off,
                                             uint32_t pitch, int k, int jcol)
{
    const int a = min(k + jcol, 7);
    float cc[8], sv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        cc[i] = ref[ref_off + uint32_t(i) * pitch + uint32_t(a)];
        sv[i] = cand[cand_off + uint32_t(i) * pitch + uint32_t(jcol)];
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
// share the rows [top, bottom], so the run's region of h (hi, lo; column-
// major, 8-row segments from the window's top row, 16-byte aligned), box
// (column-major, rows in the tiles' lane order) and s (row-major) is staged
// once in LDS.  Each wavefront takes one block at a time: its operands; per
// band of 16 window rows the correction's A (8 x 8 blocks of s) and up to 3
// column tiles at once (loads, then MFMAs, then keys); each lane's 3
// smallest keys; the shortlist of 16, the exact rescoring and a 16-lane
// sort; then the group (radius 0, with the self-block rule) or the spatial
// top-8 for the temporal search.
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
    const int CR = mfma_region_rows(rh), CRB = mfma_box_rows(rh), RSR = rh + 7;
    const int HS = mfma_h_stride(rh), BS = mfma_box_stride(rh), SS = mfma_s_stride(ncols);
    float *boxr = reinterpret_cast<float *>(mlds);
    unsigned char *wbase = reinterpret_cast<unsigned char *>(boxr + nhc * BS);
    uint16_t *hhi = reinterpret_cast<uint16_t *>(wbase + kMfmaWaves * kMfmaWaveBytes);
    uint16_t *hlo = hhi + nhc * HS;
    uint16_t *srr = hlo + nhc * HS;  // [RSR][SS]
    const int z = A.radius;
    const uint16_t *set = A.pset;
    const long long E = A.pset_elems;
    const int py = A.pset_py;
    {
        // (the loops step their 2-D index by the block size without a
        // division per element: (major, minor) += (256 / n, 256 % n))
        const uint32_t uE = uint32_t(E), upy = uint32_t(py);
        // h (hi, lo): a column's CR rows from the window's top row, as
        // dwords (unaligned loads)
        {
            const int dw = CR / 2, sc = 256 / dw, sq = 256 - sc * dw;
            uint32_t *lh = reinterpret_cast<uint32_t *>(hhi);
            uint32_t *ll = reinterpret_cast<uint32_t *>(hlo);
            int c = int(threadIdx.x) / dw, q = int(threadIdx.x) - c * dw;
            while (c < nhc) {
                const uint32_t o = uint32_t(rx0 + c) * upy + uint32_t(top + 2 * q);
                uint32_t vh, vl;
                __builtin_memcpy(&vh, set + o, 4);
                __builtin_memcpy(&vl, set + uE + o, 4);
                lh[c * (HS / 2) + q] = vh;
                ll[c * (HS / 2) + q] = vl;
                q += sq;
                c += sc;
                if (q >= dw) {
                    q -= dw;
                    ++c;
                }
            }
        }
        // s: rows top .. top + RSR - 1, columns rx0 .. rx1 (the padding
        // column of an odd count takes the next element of the plane set)
        {
            const uint16_t *gs = set + 2 * uE;
            const int sdw = SS / 2, sr = 256 / sdw, sq = 256 - sr * sdw;
            uint32_t *ls = reinterpret_cast<uint32_t *>(srr);
            int r = int(threadIdx.x) / sdw, q = int(threadIdx.x) - r * sdw;
            while (r < RSR) {
                uint32_t v;
                __builtin_memcpy(&v, gs + uint32_t(top + r) * uint32_t(W) + uint32_t(rx0 + 2 * q), 4);
                ls[r * sdw + q] = v;
                q += sq;
                r += sr;
                if (q >= sdw) {
                    q -= sdw;
                    ++r;
                }
            }
        }
        // box: region rows 16 b + 4 g .. + 3 hold window rows 16 b + g + 4 r
        // (r = 0..3), 16 bytes per thread and step
        {
            const float *gb = reinterpret_cast<const float *>(set + 3 * uE);
            const int nu = CRB / 4, sc = 256 / nu, su = 256 - sc * nu;
            int c = int(threadIdx.x) / nu, u = int(threadIdx.x) - c * nu;
            while (c < nhc) {
                const int row0 = 16 * (u >> 2) + (u & 3);
                const uint32_t col = uint32_t(rx0 + c) * upy + uint32_t(top + row0);
                f32x4 v;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    v[r] = row0 + 4 * r < rh ? gb[col + 4u * r] : 0.0f;
                }
                *reinterpret_cast<f32x4 *>(boxr + c * BS + 4 * u) = v;
                u += su;
                c += sc;
                if (u >= nu) {
                    u -= nu;
                    ++c;
                }
            }
        }
    }
    __syncthreads();

    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    unsigned char *wl = wbase + wave * kMfmaWaveBytes;
    uint16_t *scratch = reinterpret_cast<uint16_t *>(wl);                          // 224 bytes
    uint32_t *st_seq = reinterpret_cast<uint32_t *>(wl + 256);                     // [16]
    unsigned long long *st_key = reinterpret_cast<unsigned long long *>(wl + 320); // [16]
    uint32_t *st_xy = reinterpret_cast<uint32_t *>(wl + 448);                      // [16]
    const int g = lane >> 4, n = lane & 15;
    const int mrow = lane & 15;
    const int rho = (mrow >> 2) + 4 * (mrow & 3);
    const float *plane = A.plane[z];
    for (int b = wave; b < nb; b += kMfmaWaves) {
        const int bx = bx0 + b;
        const int xb = min(bx * s, W - 8);
        const int left = max(xb - R, 0), right = min(xb + R, W - 8);
        const int rw = right - left + 1;
        const BlockOperands o = block_operands<OP>(srr + (y - top) * SS + (xb - rx0), SS, scratch, lane);
        const uint32_t pm = pos_mask(rw * rh);
        uint32_t m1 = ~0u, m2 = ~0u, m3 = ~0u;
        const uint16_t *s_left = srr + (left - rx0);
        for (int m0 = 0; m0 < rh; m0 += 16) {
            // the correction's A: 8 x 8 blocks of s at the lane's tile row
            u32x4 a_s[2];
#pragma unroll
            for (int st = 0; st < 2; ++st) {
                a_s[st] = load8(s_left + min(m0 + rho + 4 * st + g, RSR - 1) * SS);
            }
            const int rbase = min(m0 + 8 * g, CR - 8);
            for (int n00 = 0; n00 < rw; n00 += 48) {
                u32x4 bh[3], bl[3];
                f32x4 bx4[3], acc[3];
                int hc[3];
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    hc[t] = min(left + n00 + 16 * t + n - rx0, nhc - 1);
                    bh[t] = *reinterpret_cast<const u32x4 *>(hhi + hc[t] * HS + rbase);
                    bl[t] = *reinterpret_cast<const u32x4 *>(hlo + hc[t] * HS + rbase);
                    bx4[t] = *reinterpret_cast<const f32x4 *>(boxr + hc[t] * BS + m0 + 4 * g);
                }
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    // D~ = box + |R_k|^2 - 2 P: the operands carry the -2
                    const float nr = n00 + 16 * t == 0 ? o.nr_first : o.nr7;
                    acc[t] = f32x4{bx4[t][0] + nr, bx4[t][1] + nr, bx4[t][2] + nr, bx4[t][3] + nr};
                    acc[t] = mfma16<OP>(o.a_main, bh[t], acc[t]);
                }
                if (n00 == 0) {
                    acc[0] = mfma16<OP>(a_s[0], o.b_d[0], acc[0]);
                    acc[0] = mfma16<OP>(a_s[1], o.b_d[1], acc[0]);
                }
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    acc[t] = mfma16<OP>(o.a_main, bl[t], acc[t]);
                }
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int n0 = n00 + 16 * t;
                    if (n0 < rw) {
                        tile_keys<3>(acc[t], m0, g, n0 + n, rw, rh, 0u, pm, m1, m2, m3);
                    }
                }
            }
        }
        const int want = shortlist(m1, m2, m3, 16, pm, st_seq);
        {
            // rescoring, 8 candidates per pass
            const float inv_rw = 1.0f / float(rw);
            const uint32_t pitch = uint32_t(A.pitch);
            const uint32_t ref_off = uint32_t(y) * pitch + uint32_t(xb);
            for (int p0 = 0; p0 < want; p0 += 8) {
                const int ci = p0 + (lane >> 3);
                const bool on = ci < want;
                const int sq = on ? int(st_seq[ci]) : 0;
                const int row = div_small(sq, inv_rw);
                const int k = sq - row * rw;
                const int cx = left + k, cy = top + row;
                const float d = exact_dist8<CM>(plane, ref_off, plane, uint32_t(cy) * pitch + uint32_t(cx), pitch, k,
                                                lane & 7);
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

}  // namespace vsfeel
