/*
 * BM3Dv2 CPU oracle -- test infrastructure only (see bm3d_oracle.h).  Only
 * tests/, the smoke check of __graft_entry__.py and the cpu_baseline leg of
 * bench.py load it, and only as the checker; the product never does.
 */
#include "bm3d_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/*
 * Active reference-block grid of one plane (src/bm3d.comp:451-458 with the
 * dispatch size of src/bm3d.cpp:1369-1370): column bx is active while
 * bx*step < W-8+step, row by exists while by*step < H-8+step.
 */
int32_t bm3d_oracle_ref_grid(const bm3d_oracle_params *p, int32_t *nx, int32_t *ny)
{
    if (p->width < 8 || p->height < 8 || p->block_step < 1 || p->block_step > 8) {
        return -1;
    }
    const int s = p->block_step;
    const int gx = (p->width - 8 + s + s - 1) / s;
    const int gy = (p->height - 8 + s + s - 1) / s;
    if (nx) {
        *nx = gx;
    }
    if (ny) {
        *ny = gy;
    }
    return gx * gy;
}

/* slice of frame clamp(n-r+i)'s stack that lands on frame n (src/bm3d.cpp:369-371) */
int32_t bm3d_oracle_agg_z(int32_t i, int32_t n, int32_t nframes, int32_t radius)
{
    int32_t a = 2 * radius - i;
    const int32_t b = n - nframes + 1 + radius;
    a = a > b ? a : b;
    const int32_t c = n + radius;
    return a < c ? a : c;
}

#define REAL float
#define RES_T float
/* the shader's float constants (decimal -> float) */
#define RLIT(x) (x##f)
#define RFMA fmaf
#define FABS fabsf
#define FN(name) bm3d_oracle_f32_##name
#include "bm3d_oracle_body.h"
#undef REAL
#undef RES_T
#undef RFMA
#undef FABS
#undef FN
#undef RLIT

#define REAL double
#define RES_T double
/* exact double constants: the float64 oracle is the algorithm without float rounding */
#define RLIT(x) (x)
#define RFMA fma
#define FABS fabs
#define FN(name) bm3d_oracle_f64_##name
#include "bm3d_oracle_body.h"
#undef REAL
#undef RES_T
#undef RFMA
#undef FABS
#undef FN
#undef RLIT
