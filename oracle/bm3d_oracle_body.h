/*
 * Body of the BM3Dv2 CPU oracle (test infrastructure only), included twice
 * by bm3d_oracle.c with REAL = float / double and FN(name) adding the
 * precision prefix.  Every function cites the part of the reference it
 * restates.  The file is compiled with -ffp-contract=off: every rounding step
 * below is the one written, and fused multiply-adds appear only through
 * FMA_X (a GLSL fma() of the shader) or FMA_C (a contractible a*b+c that
 * fp_mode 2 fuses).
 */

/* GLSL fma(): fused in fp_mode >= 1, src/bm3d.comp:119-180 */
static inline REAL FN(fma_x)(REAL a, REAL b, REAL c, int mode)
{
    if (mode >= 1) {
        return RFMA(a, b, c);
    }
    REAL t = a * b;
    return t + c;
}

/* contractible a*b+c: fused only in fp_mode 2 */
static inline REAL FN(fma_c)(REAL a, REAL b, REAL c, int mode)
{
    if (mode >= 2) {
        return RFMA(a, b, c);
    }
    REAL t = a * b;
    return t + c;
}

/*
 * 8-point forward DCT-II butterfly of src/bm3d.comp:91-135 (an unnormalised
 * FFTW-style codelet: the result equals 4 * the orthonormal DCT-II).  The
 * operation sequence is kept so that every intermediate rounds the same way.
 */
static void FN(dct8_fwd)(REAL *v, int mode)
{
    const REAL c_tan_pi_8 = RLIT(0.414213562373095048801688724209698078569671875);
    const REAL c_2cos_pi_8 = RLIT(1.847759065022573512256366378793576573644833252);
    const REAL c_tan_pi_16 = RLIT(0.198912367379658006911597622644676228597850501);
    const REAL c_2cos_pi_16 = RLIT(1.961570560806460898252364472268478073947867462);
    const REAL c_sqrt2 = RLIT(1.414213562373095048801688724209698078569671875);
    const REAL c_tan_3pi_16 = RLIT(0.668178637919298919997757686523080761552472251);
    const REAL c_2cos_3pi_16 = RLIT(1.662939224605090474157576755235811513477121624);
    const REAL c_inv_sqrt2 = RLIT(0.707106781186547524400844362104849039284835938);

    const REAL d07 = v[0] - v[7], s07 = v[0] + v[7];
    const REAL d43 = v[4] - v[3], s43 = v[4] + v[3];
    const REAL d25 = v[2] - v[5], s25 = v[2] + v[5];
    const REAL d16 = v[1] - v[6], s16 = v[1] + v[6];
    const REAL a = d25 + d16;
    const REAL f = d25 - d16;
    const REAL b = FN(fma_x)(-c_inv_sqrt2, a, d07, mode);
    const REAL g = FN(fma_x)(-c_inv_sqrt2, f, d43, mode);
    const REAL o3 = c_2cos_3pi_16 * FN(fma_x)(c_tan_3pi_16, g, b, mode);
    const REAL o5 = -(c_2cos_3pi_16 * FN(fma_x)(-c_tan_3pi_16, b, g, mode));
    const REAL p = s07 + s43;
    const REAL q = s25 + s16;
    const REAL o4 = c_sqrt2 * (p - q);
    const REAL o0 = c_sqrt2 * (p + q);
    const REAL h = FN(fma_x)(c_inv_sqrt2, a, d07, mode);
    const REAL i = FN(fma_x)(c_inv_sqrt2, f, d43, mode);
    const REAL o1 = c_2cos_pi_16 * FN(fma_x)(-c_tan_pi_16, i, h, mode);
    const REAL o7 = c_2cos_pi_16 * FN(fma_x)(c_tan_pi_16, h, i, mode);
    const REAL l = s07 - s43;
    const REAL o = s25 - s16;
    const REAL o2 = c_2cos_pi_8 * FN(fma_x)(-c_tan_pi_8, o, l, mode);
    const REAL o6 = c_2cos_pi_8 * FN(fma_x)(c_tan_pi_8, l, o, mode);
    v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3;
    v[4] = o4; v[5] = o5; v[6] = o6; v[7] = o7;
}

/* 8-point inverse butterfly of src/bm3d.comp:137-181 (4 * orthonormal DCT-III) */
static void FN(dct8_inv)(REAL *v, int mode)
{
    const REAL c_2cos_3pi_16 = RLIT(1.662939224605090474157576755235811513477121624);
    const REAL c_tan_3pi_16 = RLIT(0.668178637919298919997757686523080761552472251);
    const REAL c_2cos_pi_16 = RLIT(1.961570560806460898252364472268478073947867462);
    const REAL c_tan_pi_16 = RLIT(0.198912367379658006911597622644676228597850501);
    const REAL c_2cos_pi_8 = RLIT(1.847759065022573512256366378793576573644833252);
    const REAL c_inv_sqrt2 = RLIT(0.707106781186547524400844362104849039284835938);
    const REAL c_tan_pi_8 = RLIT(0.414213562373095048801688724209698078569671875);
    const REAL c_sqrt2 = RLIT(1.414213562373095048801688724209698078569671875);

    const REAL dc = v[0] * c_sqrt2;
    const REAL e_plus = FN(fma_x)(c_sqrt2, v[4], dc, mode);
    const REAL e_minus = FN(fma_x)(-c_sqrt2, v[4], dc, mode);
    const REAL r26 = FN(fma_x)(c_tan_pi_8, v[6], v[2], mode);
    const REAL r62 = FN(fma_x)(c_tan_pi_8, v[2], -v[6], mode);
    const REAL s53 = v[5] + v[3];
    const REAL d35 = v[3] - v[5];
    const REAL c1 = FN(fma_x)(c_inv_sqrt2, s53, v[1], mode);
    const REAL n7 = FN(fma_x)(-c_inv_sqrt2, d35, v[7], mode);
    const REAL f7 = FN(fma_x)(c_inv_sqrt2, d35, v[7], mode);
    const REAL m1 = FN(fma_x)(-c_inv_sqrt2, s53, v[1], mode);
    const REAL t7 = FN(fma_x)(c_2cos_pi_8, r26, e_plus, mode);
    const REAL g = FN(fma_x)(c_tan_pi_16, f7, c1, mode);
    const REAL o7 = FN(fma_x)(-c_2cos_pi_16, g, t7, mode);
    const REAL o0 = FN(fma_x)(c_2cos_pi_16, g, t7, mode);
    const REAL tp = FN(fma_x)(-c_2cos_pi_8, r62, e_minus, mode);
    const REAL tq = FN(fma_x)(c_tan_3pi_16, m1, n7, mode);
    const REAL o5 = FN(fma_x)(-c_2cos_3pi_16, tq, tp, mode);
    const REAL o2 = FN(fma_x)(c_2cos_3pi_16, tq, tp, mode);
    const REAL th = FN(fma_x)(-c_2cos_pi_8, r26, e_plus, mode);
    const REAL ti = FN(fma_x)(-c_tan_pi_16, c1, f7, mode);
    const REAL o3 = FN(fma_x)(-c_2cos_pi_16, ti, th, mode);
    const REAL o4 = FN(fma_x)(c_2cos_pi_16, ti, th, mode);
    const REAL tl = FN(fma_x)(c_2cos_pi_8, r62, e_minus, mode);
    const REAL to = FN(fma_x)(-c_tan_3pi_16, n7, m1, mode);
    const REAL o6 = FN(fma_x)(-c_2cos_3pi_16, to, tl, mode);
    const REAL o1 = FN(fma_x)(c_2cos_3pi_16, to, tl, mode);
    v[0] = o0; v[1] = o1; v[2] = o2; v[3] = o3;
    v[4] = o4; v[5] = o5; v[6] = o6; v[7] = o7;
}

void FN(dct8)(REAL *v, int32_t inverse, int32_t fp_mode)
{
    if (inverse) {
        FN(dct8_inv)(v, fp_mode);
    } else {
        FN(dct8_fwd)(v, fp_mode);
    }
}

/* apply the 1-D transform along one axis of a [8][8][8] cube */
static void FN(cube_axis)(REAL *c, int axis, int inverse, int mode)
{
    const int stride = axis == 0 ? 64 : (axis == 1 ? 8 : 1);
    for (int a = 0; a < 8; ++a) {
        for (int b = 0; b < 8; ++b) {
            int base;
            if (axis == 0) {
                base = a * 8 + b;
            } else if (axis == 1) {
                base = a * 64 + b;
            } else {
                base = a * 64 + b * 8;
            }
            REAL v[8];
            for (int k = 0; k < 8; ++k) {
                v[k] = c[base + k * stride];
            }
            if (inverse) {
                FN(dct8_inv)(v, mode);
            } else {
                FN(dct8_fwd)(v, mode);
            }
            for (int k = 0; k < 8; ++k) {
                c[base + k * stride] = v[k];
            }
        }
    }
}

/*
 * Collaborative filtering of one group, cube[m][row][col].  Forward order
 * rows (vertical) -> columns -> group axis and inverse order vertical ->
 * horizontal -> group axis, as collaborative_hard / collaborative_wiener run
 * them through transform_pack8 / transpose_pack8 / transform_pack8_1d
 * (src/bm3d.comp:183-236, :384-399, :422-443).  Hard threshold: keep when
 * |c| >= sigma, the DC term always (src/bm3d.comp:358-382).  Wiener:
 * c * r^2 / (r^2 + sigma^2), DC coefficient 1 (src/bm3d.comp:403-420).  Kept
 * values are scaled by 1/4096.  The weight sum follows the shader's order:
 * per column-frequency lane u, partial sums ks[i % 4] over i = w*8 + v, then
 * ((ks0+ks1)+(ks2+ks3)), then the 8-lane tree of reduce_group8 (:238-246).
 */
REAL FN(group)(REAL *cube, const REAL *ref_cube, REAL sigma, int32_t mode, int32_t *kept)
{
    REAL rc[512];
    FN(cube_axis)(cube, 1, 0, mode);
    FN(cube_axis)(cube, 2, 0, mode);
    FN(cube_axis)(cube, 0, 0, mode);
    if (ref_cube) {
        memcpy(rc, ref_cube, sizeof(rc));
        FN(cube_axis)(rc, 1, 0, mode);
        FN(cube_axis)(rc, 2, 0, mode);
        FN(cube_axis)(rc, 0, 0, mode);
    }
    REAL lane_k[8];
    int32_t count = 0;
    const REAL scale = (REAL)(1.0f / 4096.0f);
    for (int u = 0; u < 8; ++u) {
        REAL ks[4] = { 0, 0, 0, 0 };
        for (int i = 0; i < 64; ++i) {
            const int w = i >> 3, v = i & 7;
            const int idx = w * 64 + v * 8 + u;
            const REAL val = cube[idx];
            const int dc = (i == 0) && (u == 0);
            if (!ref_cube) {
                const REAL thr = dc ? (REAL)0 : sigma;
                const int flag = FABS(val) >= thr;
                ks[i % 4] += flag ? (REAL)1 : (REAL)0;
                count += flag;
                cube[idx] = flag ? val * scale : (REAL)0;
            } else {
                const REAL r = rc[idx];
                const REAL r2 = r * r;
                const REAL s2 = sigma * sigma;
                const REAL den = FN(fma_c)(r, r, s2, mode);
                REAL coeff = r2 / den;
                if (dc) {
                    coeff = (REAL)1;
                }
                const REAL filtered = val * coeff;
                ks[i % 4] = FN(fma_c)(coeff, coeff, ks[i % 4], mode);
                cube[idx] = filtered * scale;
            }
        }
        lane_k[u] = (ks[0] + ks[1]) + (ks[2] + ks[3]);
    }
    const REAL k = ((lane_k[0] + lane_k[1]) + (lane_k[2] + lane_k[3])) +
                   ((lane_k[4] + lane_k[5]) + (lane_k[6] + lane_k[7]));
    if (kept) {
        *kept = count;
    }
    FN(cube_axis)(cube, 1, 1, mode);
    FN(cube_axis)(cube, 2, 1, mode);
    FN(cube_axis)(cube, 0, 1, mode);
    return (REAL)1 / k;
}

/* per-lane sorted top-8 list of src/bm3d.comp:284-302 */
typedef struct {
    REAL e[8];
    int32_t x[8], y[8], s[8];
} FN(list8);

static void FN(list_init)(FN(list8) *l)
{
    for (int k = 0; k < 8; ++k) {
        l->e[k] = (REAL)FLT_MAX;
        l->x[k] = 0;
        l->y[k] = 0;
        l->s[k] = INT32_MAX;
    }
}

/* insert8 / ginsert8 (src/bm3d.comp:284-322): strict '<' on the distance,
 * the new entry goes before the first larger one; ties keep the older one */
static void FN(insert8)(REAL *e, int32_t *xs, int32_t *ys, int32_t *ss,
                        REAL err, int32_t cx, int32_t cy, int32_t cs)
{
    if (!(err < e[7])) {
        return;
    }
    int pos = 8;
    for (int k = 7; k >= 0; --k) {
        if (err < e[k]) {
            pos = k;
        }
    }
    for (int k = 7; k > pos; --k) {
        e[k] = e[k - 1];
        xs[k] = xs[k - 1];
        ys[k] = ys[k - 1];
        ss[k] = ss[k - 1];
    }
    e[pos] = err;
    xs[pos] = cx;
    ys[pos] = cy;
    ss[pos] = cs;
}

/* merge_group64 (src/bm3d.comp:324-356): 8-way merge of the 8 lane lists by
 * (distance, then smaller seq); equal keys go to the lower lane */
static void FN(merge64)(const FN(list8) *lanes, REAL *me, int32_t *mx, int32_t *my, int32_t *ms)
{
    int head[8] = { 0 };
    for (int k = 0; k < 8; ++k) {
        int bl = 0, bh = 0, found = 0;
        REAL be = (REAL)FLT_MAX;
        int32_t bs = INT32_MAX;
        for (int l = 0; l < 8; ++l) {
            const int hv = head[l];
            const int valid = hv < 8;
            const REAL e = valid ? lanes[l].e[hv] : (REAL)FLT_MAX;
            const int32_t sq = valid ? lanes[l].s[hv] : INT32_MAX;
            const int better = valid && (!found || e < be || (e == be && sq < bs));
            if (better) {
                bl = l;
                bh = hv;
                be = e;
                bs = sq;
            }
            found = found || valid;
        }
        me[k] = be;
        mx[k] = lanes[bl].x[bh];
        my[k] = lanes[bl].y[bh];
        ms[k] = bs;
        head[bl] += 1;
    }
}

/* squared error of one 8-row column: reference column c of cur against the
 * plane column X starting at row cy (col_ssd, src/bm3d.comp:268-275) */
static inline REAL FN(col_ssd)(const REAL *cur, const float *plane, int64_t pitch,
                               int cy, int X, int c, int mode)
{
    REAL err[2] = { 0, 0 };
    for (int i = 0; i < 8; ++i) {
        const REAL val = cur[c * 8 + i] - (REAL)plane[(int64_t)(cy + i) * pitch + X];
        err[i % 2] = FN(fma_c)(val, val, err[i % 2], mode);
    }
    return err[0] + err[1];
}

/* reduce_cols (src/bm3d.comp:280-282) */
static inline REAL FN(reduce_cols)(const REAL *col)
{
    return ((col[0] + col[1]) + (col[2] + col[3])) + ((col[4] + col[5]) + (col[6] + col[7]));
}

/* true 8x8 SSD in the column/tree order of ssd_err (src/bm3d.comp:253-264) */
static inline REAL FN(ssd_exact)(const REAL *cur, const float *plane, int64_t pitch,
                                 int cx, int cy, int mode)
{
    REAL col[8];
    for (int c = 0; c < 8; ++c) {
        col[c] = FN(col_ssd)(cur, plane, pitch, cy, cx + c, c, mode);
    }
    return FN(reduce_cols)(col);
}

/*
 * One search window scanned by the 8 lanes of a shader group into their
 * private lists: the sliding column scan of radius < 3
 * (src/bm3d.comp:497-511 spatial, :582-596 temporal) or the strided exact
 * scan of radius >= 3 (:512-526, :597-611).
 */
static void FN(scan_window)(FN(list8) *lanes, const REAL *cur, const float *plane,
                            int64_t pitch, int left, int right, int top, int bottom,
                            int seq_base, int exact, int mode)
{
    const int rw = right - left + 1;
    for (int lane = 0; lane < 8; ++lane) {
        FN(list8) *L = &lanes[lane];
        if (!exact) {
            for (int cy = top + lane; cy <= bottom; cy += 8) {
                REAL col[8];
                for (int c = 0; c < 8; ++c) {
                    col[c] = FN(col_ssd)(cur, plane, pitch, cy, left + c, c, mode);
                }
                FN(insert8)(L->e, L->x, L->y, L->s, FN(reduce_cols)(col), left, cy,
                            seq_base + (cy - top) * rw);
                for (int cx = left + 1; cx <= right; ++cx) {
                    for (int c = 0; c < 7; ++c) {
                        col[c] = col[c + 1];
                    }
                    col[7] = FN(col_ssd)(cur, plane, pitch, cy, cx + 7, 7, mode);
                    FN(insert8)(L->e, L->x, L->y, L->s, FN(reduce_cols)(col), cx, cy,
                                seq_base + (cy - top) * rw + (cx - left));
                }
            }
        } else {
            const int total = (bottom - top + 1) * rw;
            int row_i = top + lane / rw;
            int col_i = left + lane % rw;
            for (int s = lane; s < total; s += 8) {
                const REAL err = FN(ssd_exact)(cur, plane, pitch, col_i, row_i, mode);
                FN(insert8)(L->e, L->x, L->y, L->s, err, col_i, row_i, seq_base + s);
                col_i += 8;
                while (col_i > right) {
                    col_i -= rw;
                    row_i += 1;
                }
            }
        }
    }
}

int32_t FN(window_dist)(const float *plane, int32_t pitch, int32_t width, int32_t height,
                        int32_t x, int32_t y, int32_t range, int32_t cx0, int32_t cy0,
                        int32_t exact, int32_t mode, float *dist)
{
    REAL cur[64];
    for (int c = 0; c < 8; ++c) {
        for (int i = 0; i < 8; ++i) {
            cur[c * 8 + i] = plane[(int64_t)(y + i) * pitch + x + c];
        }
    }
    const int left = cx0 - range > 0 ? cx0 - range : 0;
    const int right = cx0 + range < width - 8 ? cx0 + range : width - 8;
    const int top = cy0 - range > 0 ? cy0 - range : 0;
    const int bottom = cy0 + range < height - 8 ? cy0 + range : height - 8;
    const int rw = right - left + 1;
    int n = 0;
    for (int cy = top; cy <= bottom; ++cy) {
        REAL col[8];
        for (int c = 0; c < 8; ++c) {
            col[c] = FN(col_ssd)(cur, plane, pitch, cy, left + c, c, mode);
        }
        for (int cx = left; cx <= right; ++cx) {
            if (cx > left) {
                for (int c = 0; c < 7; ++c) {
                    col[c] = col[c + 1];
                }
                col[7] = FN(col_ssd)(cur, plane, pitch, cy, cx + 7, 7, mode);
            }
            const REAL d = exact ? FN(ssd_exact)(cur, plane, pitch, cx, cy, mode)
                                 : FN(reduce_cols)(col);
            dist[(cy - top) * rw + (cx - left)] = (float)d;
            ++n;
        }
    }
    return n;
}

typedef struct {
    const bm3d_oracle_params *p;
    const float *const *search;
    const float *const *input;
    RES_T *res;          /* shared output (single thread) or private copy */
    int32_t *groups;
    int by_begin, by_end;
    int nx;
} FN(job);

/*
 * One shader group: the body of main() in src/bm3d.comp:445-718 for the ref
 * block at grid position (bx, by).
 */
static void FN(block)(const bm3d_oracle_params *p, const float *const *search,
                      const float *const *input, RES_T *res, int32_t *groups,
                      int bx, int by, int nx)
{
    const int W = p->width, H = p->height, step = p->block_step;
    const int64_t pitch = p->pitch;
    const int64_t pe = (int64_t)H * pitch;
    const int r = p->radius, mode = p->fp_mode;
    const int exact = r >= 3;

    const int x0 = bx * step;
    int y = by * step;
    int x = x0 < W - 8 ? x0 : W - 8;
    y = y < H - 8 ? y : H - 8;

    /* cur[c*8+i] (src/bm3d.comp:460-466) */
    const float *centre = search[r];
    REAL cur[64];
    for (int c = 0; c < 8; ++c) {
        for (int i = 0; i < 8; ++i) {
            cur[c * 8 + i] = centre[(int64_t)(y + i) * pitch + x + c];
        }
    }

    REAL ge[8];
    int32_t gx[8], gy[8], gz[8];
    {
        FN(list8) lanes[8];
        for (int l = 0; l < 8; ++l) {
            FN(list_init)(&lanes[l]);
        }
        const int R = p->bm_range;
        const int left = x - R > 0 ? x - R : 0;
        const int right = x + R < W - 8 ? x + R : W - 8;
        const int top = y - R > 0 ? y - R : 0;
        const int bottom = y + R < H - 8 ? y + R : H - 8;
        FN(scan_window)(lanes, cur, centre, pitch, left, right, top, bottom, 0, exact, mode);
        int32_t gs[8];
        FN(merge64)(lanes, ge, gx, gy, gs);
        for (int k = 0; k < 8; ++k) {
            gz[k] = r;
        }
    }

    /* temporal predictive search (src/bm3d.comp:540-637) */
    if (r > 0) {
        int32_t sx[8], sy[8];
        memcpy(sx, gx, sizeof(sx));
        memcpy(sy, gy, sizeof(sy));
        const int PR = p->ps_range;
        const int win = (2 * PR + 1) * (2 * PR + 1);
        for (int dir = -1; dir <= 1; dir += 2) {
            int32_t px[8], py[8];
            memcpy(px, sx, sizeof(px));
            memcpy(py, sy, sizeof(py));
            for (int t = 1; t <= r; ++t) {
                const int z = r + dir * t;
                const float *plane = search[z];
                FN(list8) lanes[8];
                for (int l = 0; l < 8; ++l) {
                    FN(list_init)(&lanes[l]);
                }
                for (int i = 0; i < p->ps_num; ++i) {
                    const int left = px[i] - PR > 0 ? px[i] - PR : 0;
                    const int right = px[i] + PR < W - 8 ? px[i] + PR : W - 8;
                    const int top = py[i] - PR > 0 ? py[i] - PR : 0;
                    const int bottom = py[i] + PR < H - 8 ? py[i] + PR : H - 8;
                    FN(scan_window)(lanes, cur, plane, pitch, left, right, top, bottom,
                                    i * win, exact, mode);
                }
                REAL fe[8];
                int32_t fx[8], fy[8], fs[8];
                FN(merge64)(lanes, fe, fx, fy, fs);
                for (int i = 0; i < p->ps_num; ++i) {
                    FN(insert8)(ge, gx, gy, gz, fe[i], fx[i], fy[i], z);
                }
                memcpy(px, fx, sizeof(px));
                memcpy(py, fy, sizeof(py));
            }
        }
    }

    /* the reference block itself is always in the group (src/bm3d.comp:639-663) */
    {
        int match = 0;
        for (int k = 0; k < 8; ++k) {
            match += (gx[k] == x && gy[k] == y && gz[k] == r);
        }
        if (match == 0) {
            for (int k = 7; k >= 1; --k) {
                gx[k] = gx[k - 1];
                gy[k] = gy[k - 1];
                gz[k] = gz[k - 1];
            }
            gx[0] = x;
            gy[0] = y;
            gz[0] = r;
        }
    }

    if (groups) {
        int32_t *g = groups + ((int64_t)by * nx + bx) * 24;
        for (int k = 0; k < 8; ++k) {
            g[k * 3 + 0] = gx[k];
            g[k * 3 + 1] = gy[k];
            g[k * 3 + 2] = gz[k];
        }
    }

    /* gather (src/bm3d.comp:677-689), cube[m][row][col] */
    REAL cube[512], ref_cube[512];
    for (int m = 0; m < 8; ++m) {
        const float *src = p->final_pass ? input[gz[m]] : search[gz[m]];
        const float *refp = search[gz[m]];
        for (int j = 0; j < 8; ++j) {
            for (int c = 0; c < 8; ++c) {
                const int64_t o = (int64_t)(gy[m] + j) * pitch + gx[m] + c;
                cube[m * 64 + j * 8 + c] = src[o];
                ref_cube[m * 64 + j * 8 + c] = refp[o];
            }
        }
    }
    const REAL w = FN(group)(cube, p->final_pass ? ref_cube : NULL, (REAL)p->sigma, mode, NULL);

    /* weighted scatter with the extractor (src/bm3d.comp:697-715) */
    const REAL E = (REAL)p->extractor;
    const REAL wq = (w + E) - E;
    for (int m = 0; m < 8; ++m) {
        RES_T *wd = res + (int64_t)gz[m] * 2 * pe + (int64_t)gy[m] * pitch + gx[m];
        RES_T *wt = wd + pe;
        for (int j = 0; j < 8; ++j) {
            for (int c = 0; c < 8; ++c) {
                const REAL v = FN(fma_c)(w, cube[m * 64 + j * 8 + c], E, mode) - E;
                wd[(int64_t)j * pitch + c] += (RES_T)v;
                wt[(int64_t)j * pitch + c] += (RES_T)wq;
            }
        }
    }
}

static void *FN(worker)(void *arg)
{
    FN(job) *J = (FN(job) *)arg;
    for (int by = J->by_begin; by < J->by_end; ++by) {
        for (int bx = 0; bx < J->nx; ++bx) {
            FN(block)(J->p, J->search, J->input, J->res, J->groups, bx, by, J->nx);
        }
    }
    return NULL;
}

int32_t FN(estimate)(const bm3d_oracle_params *p, const float *const *search,
                     const float *const *input, RES_T *res, int32_t *groups)
{
    int32_t nx, ny;
    if (bm3d_oracle_ref_grid(p, &nx, &ny) <= 0) {
        return -1;
    }
    if (!input) {
        input = search;
    }
    const int tw = 2 * p->radius + 1;
    const int64_t plane_elems = (int64_t)tw * 2 * p->height * p->pitch;
    int nt = p->nthreads < 1 ? 1 : p->nthreads;
    if (nt > ny) {
        nt = ny;
    }
    if (nt == 1) {
        FN(job) J = { p, search, input, res, groups, 0, ny, nx };
        FN(worker)(&J);
        return 0;
    }
    /* each thread accumulates into its own stack; the stacks are summed in
     * thread order afterwards (the reference sums with unordered atomics) */
    FN(job) *jobs = (FN(job) *)calloc((size_t)nt, sizeof(FN(job)));
    pthread_t *th = (pthread_t *)calloc((size_t)nt, sizeof(pthread_t));
    RES_T **priv = (RES_T **)calloc((size_t)nt, sizeof(RES_T *));
    for (int t = 0; t < nt; ++t) {
        priv[t] = (RES_T *)calloc((size_t)plane_elems, sizeof(RES_T));
        jobs[t].p = p;
        jobs[t].search = search;
        jobs[t].input = input;
        jobs[t].res = priv[t];
        jobs[t].groups = groups;
        jobs[t].by_begin = (int)((int64_t)ny * t / nt);
        jobs[t].by_end = (int)((int64_t)ny * (t + 1) / nt);
        jobs[t].nx = nx;
        pthread_create(&th[t], NULL, FN(worker), &jobs[t]);
    }
    for (int t = 0; t < nt; ++t) {
        pthread_join(th[t], NULL);
    }
    for (int t = 0; t < nt; ++t) {
        for (int64_t i = 0; i < plane_elems; ++i) {
            res[i] += priv[t][i];
        }
        free(priv[t]);
    }
    free(priv);
    free(th);
    free(jobs);
    return 0;
}

/*
 * Whole clip: estimate stacks per frame (one dispatch of the shader each),
 * aggregated per output frame over the tw slices of agg_z in double
 * (src/bm3d.cpp:369-371, :656-671; src/bm3d_agg.comp:36-51).
 */
int32_t FN(clip)(const bm3d_oracle_params *p, int32_t nframes, const float *clip,
                 const float *ref, int32_t first, int32_t count, RES_T *out)
{
    const int r = p->radius, tw = 2 * r + 1;
    const int64_t pe = (int64_t)p->height * p->pitch;
    if (nframes <= 0 || first < 0 || count < 0 || first + count > nframes) {
        return -1;
    }
    const float *search_clip = p->final_pass ? ref : clip;
    /* estimate stacks are needed for frames clamp(first-r) .. clamp(first+count-1+r) */
    int lo = first - r < 0 ? 0 : first - r;
    int hi = first + count - 1 + r > nframes - 1 ? nframes - 1 : first + count - 1 + r;
    if (count == 0) {
        return 0;
    }
    const int nstacks = hi - lo + 1;
    RES_T *stacks = (RES_T *)calloc((size_t)nstacks * tw * 2 * pe, sizeof(RES_T));
    if (!stacks) {
        return -2;
    }
    for (int m = lo; m <= hi; ++m) {
        const float *sw[9], *iw[9];
        for (int z = 0; z < tw; ++z) {
            int f = m - r + z;
            f = f < 0 ? 0 : (f > nframes - 1 ? nframes - 1 : f);
            sw[z] = search_clip + (int64_t)f * pe;
            iw[z] = clip + (int64_t)f * pe;
        }
        FN(estimate)(p, sw, iw, stacks + (int64_t)(m - lo) * tw * 2 * pe, NULL);
    }
    for (int n = first; n < first + count; ++n) {
        const RES_T *slice_wd[9];
        for (int i = 0; i < tw; ++i) {
            int m = n - r + i;
            m = m < 0 ? 0 : (m > nframes - 1 ? nframes - 1 : m);
            const int z = r == 0 ? 0 : bm3d_oracle_agg_z(i, n, nframes, r);
            slice_wd[i] = stacks + (int64_t)(m - lo) * tw * 2 * pe + (int64_t)z * 2 * pe;
        }
        RES_T *o = out + (int64_t)(n - first) * pe;
        for (int yy = 0; yy < p->height; ++yy) {
            for (int xx = 0; xx < p->width; ++xx) {
                const int64_t q = (int64_t)yy * p->pitch + xx;
                double acc = 0.0, acw = 0.0;
                for (int i = 0; i < tw; ++i) {
                    acc += (double)slice_wd[i][q];
                    acw += (double)slice_wd[i][q + pe];
                }
                o[q] = (RES_T)(acc / acw);
            }
        }
    }
    free(stacks);
    return 0;
}
