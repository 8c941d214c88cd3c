/*
 * ddfan_oracle.c -- CPU restatement of the reference reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (aapm_lowdosectimagerecon_competition_amd) never links or calls it.
 *
 * What it restates (file:line of the reference, AAPM_script/):
 *   orc_detector_arc     DD2MutiSlices.m:55-76            (a1)
 *   orc_cossin           utilities.cuh:143-148,218-227    (a2)
 *   orc_boundaries       utilities.cuh:268-287            (a3)
 *   orc_proj_*           MultiSlicesProjGPU.cu:26-51,100-209,211-364 (a4-a7)
 *   orc_back_*           MultiSlicesBackGPU.cu:12-66,115-184,187-258 (a8-a9)
 *   orc_helical_to_fan   HelicalToFanFunc_mex.cpp:27-34,41-185      (a11)
 *   orc_ossart_*         OSSART_AAPM.m:49-117                       (a13)
 *   orc_recon_mask       reconscript.m:28-38                        (a14)
 *
 * Arithmetic contract.
 *   - Every discrete decision (slab-branch choice, texel index
 *     floor(coord-0.5), rebin row floor(dv), bounds tests) is evaluated on the
 *     FP32 expression of the reference with no FMA contraction and IEEE
 *     division / square root.  The build compiles this file with
 *     -ffp-contract=off.  Where the reference used a fast-math intrinsic
 *     (rsqrtf, asinf, approximate divide) the FP32 index path uses the
 *     correctly-rounded operation (1/sqrtf, a/b) and the Cephes single-precision
 *     arcsine (orc_asinf32, written from S. L. Moshier's published Cephes
 *     method), because the CUDA fast-math internals are not reproducible.
 *   - "f64" variants evaluate every continuous quantity (coordinates, weights,
 *     lengths, summed-area tables, sums) in FP64 from the same FP32 host
 *     tables.  "f32" variants evaluate everything in FP32 in the reference's
 *     operation order with an exact (not 8-bit-weight) linear interpolation;
 *     the f32 projector/backprojector are the multithreaded CPU baseline.
 *   - Texture semantics (MultiSlicesProjGPU.cu:294-309): unnormalised,
 *     texel-centred, clamp-to-edge linear filtering.  A fetch at coordinate c
 *     interpolates texels i=floor(c-0.5) and i+1 with weight c-0.5-i, indices
 *     clamped to [0, N].  The summed-area-table border entry 0 is 0 (the
 *     reference leaves it uninitialised, MultiSlicesProjGPU.cu:232-233).
 *   - Masked voxels of the backprojector are 0 (the reference leaves them
 *     uninitialised, MultiSlicesBackGPU.cu:131-133).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_TWOPI 6.283185307179586 /* utilities.cuh:43 */

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */
typedef void *(*orc_worker_fn)(void *);

typedef struct {
    int tid, nthreads;
    void *ctx;
} orc_task;

static void orc_parallel(int nthreads, orc_worker_fn fn, void *ctx) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    orc_task tk[256];
    for (int t = 0; t < nthreads; ++t) {
        tk[t].tid = t;
        tk[t].nthreads = nthreads;
        tk[t].ctx = ctx;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, &tk[t]);
    fn(&tk[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* Texel index of an unnormalised linear-filter fetch: floor(c - 0.5) in FP32,
 * limited to [-1, N] (every value outside fetches the clamped edge texel). */
static inline int orc_texel(float c, int N) {
    float f = floorf(c - 0.5f);
    if (!(f >= -1.0f)) f = -1.0f; /* also maps NaN to -1 */
    if (f > (float)N) f = (float)N;
    return (int)f;
}

static inline int orc_clampi(int i, int N) { return i < 0 ? 0 : (i > N ? N : i); }

/* FP64 fetch: texel pair chosen by the FP32 index i, weight from the FP64
 * coordinate c64. */
static inline double orc_fetch64(const double *S, int N, int i, double c64) {
    int i0 = orc_clampi(i, N), i1 = orc_clampi(i + 1, N);
    double t = c64 - 0.5 - (double)i;
    return S[i0] + t * (S[i1] - S[i0]);
}

/* FP32 fetch with an exact FP32 lerp. */
static inline float orc_fetch32(const float *S, int N, float c) {
    int i = orc_texel(c, N);
    int i0 = orc_clampi(i, N), i1 = orc_clampi(i + 1, N);
    float t = (c - 0.5f) - (float)i;
    return S[i0] + t * (S[i1] - S[i0]);
}

/* Single-precision arcsine, Cephes method (S. L. Moshier, "Cephes Math
 * Library", asinf.c): odd minimax polynomial on [0, 0.5] and the identity
 * asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) above 0.5.  Horner steps use fmaf so
 * the result is identical on every IEEE-754 platform with a correctly rounded
 * fmaf/sqrtf. */
float orc_asinf32(float xx) {
    float a = fabsf(xx), x, z;
    int flag;
    if (a > 1.0f) return NAN;
    if (a < 1.0e-4f) return xx;
    if (a > 0.5f) {
        z = 0.5f * (1.0f - a);
        x = sqrtf(z);
        flag = 1;
    } else {
        x = a;
        z = x * x;
        flag = 0;
    }
    float p = fmaf(4.2163199048e-2f, z, 2.4181311049e-2f);
    p = fmaf(p, z, 4.5470025998e-2f);
    p = fmaf(p, z, 7.4953002686e-2f);
    p = fmaf(p, z, 1.6666752422e-1f);
    z = fmaf(p * z, x, x);
    if (flag) {
        z = z + z;
        z = 1.5707963267948966f - z;
    }
    return xx < 0.0f ? -z : z;
}

/* ------------------------------------------------------------------ */
/* a1-a3: host geometry tables                                         */
/* ------------------------------------------------------------------ */

/* DD2MutiSlices.m:61-76 in MATLAB single arithmetic (platform libm). */
void orc_detector_arc(int DNU, float detCntIdx, float detCellWidth, float sdd, float sid,
                      float *xds, float *yds) {
    float stepTheta = atanf(detCellWidth / 2.0f / sdd) * 2.0f;
    for (int ii = 1; ii <= DNU; ++ii) {
        float beta = ((float)ii - detCntIdx) * stepTheta;
        xds[ii - 1] = sinf(beta) * sdd;
        yds[ii - 1] = sid - cosf(beta) * sdd;
    }
}

/* utilities.cuh:143-148 regularizeAngle<float>: the comparisons and updates
 * promote to double (TWOPI is a double literal) and round back to float. */
float orc_regularize(float c) {
    while ((double)c >= ORC_TWOPI) c = (float)((double)c - ORC_TWOPI);
    while (c < 0.0f) c = (float)((double)c + ORC_TWOPI);
    return c;
}

/* utilities.cuh:218-227 Constant_MultiSlice_host: (cosf, sinf) on the host. */
void orc_cossin(const float *hangs, long n, float *cs) {
    for (long i = 0; i < n; ++i) {
        float a = orc_regularize(hangs[i]);
        cs[2 * i] = cosf(a);
        cs[2 * i + 1] = sinf(a);
    }
}

/* utilities.cuh:268-287 DD3Boundaries<float>: double arithmetic, float store. */
void orc_boundaries(int nb, const float *c, float *b) {
    if (nb >= 3) {
        for (int i = 0; i < nb; ++i) {
            if (i == 0)
                b[0] = (float)(1.5 * (double)c[0] - 0.5 * (double)c[1]);
            else if (i == nb - 1)
                b[i] = (float)(1.5 * (double)c[i - 1] - 0.5 * (double)c[i - 2]);
            else
                b[i] = (float)(0.5 * (double)c[i - 1] + 0.5 * (double)c[i]);
        }
    } else {
        b[0] = (float)((double)c[0] - 0.5);
        b[1] = (float)((double)c[0] + 0.5);
    }
}

/* reconscript.m:28-38 reconstruction mask, mask[y*XN + x]. */
void orc_recon_mask(int XN, int YN, uint8_t *mask) {
    for (int jj = 1; jj <= YN; ++jj)
        for (int ii = 1; ii <= XN; ++ii) {
            double a = ((double)ii - 0.5 - (double)XN / 2.0) / ((double)XN / 2.0);
            double b = ((double)jj - 0.5 - (double)YN / 2.0) / ((double)YN / 2.0);
            mask[(jj - 1) * XN + (ii - 1)] = sqrt(a * a + b * b) < 1.3 ? 1 : 0;
        }
}

/* Backprojector host scalars (MultiSlicesBackGPU.cu:211,226-227). */
void orc_back_scalars(float x0, float y0, const float *xds, const float *yds, float *S2D_out,
                      float *dbeta_out) {
    const float S2D = hypotf(xds[0] - x0, yds[0] - y0);
    float dts = (float)sqrt((double)(powf(xds[1] - xds[0], 2) + powf(yds[1] - yds[0], 2)));
    float dbeta = atanf((float)((double)(dts / S2D) * 0.5)) * 2.0f;
    *S2D_out = S2D;
    *dbeta_out = dbeta;
}

/* ------------------------------------------------------------------ */
/* a5 / a8: summed-area tables                                           */
/* ------------------------------------------------------------------ */
/* Volume layout vol[(y*XN+x)*SLN+z].  SAT_X[z][y][k] = sum_{x<k} m*vol (k=0..XN),
 * SAT_Y[z][x][k] = sum_{y<k} m*vol (k=0..YN); mask multiplies by its byte value
 * (MultiSlicesProjGPU.cu:202-209). */
static void orc_vol_sat64(const double *vol, const uint8_t *mask, int XN, int YN, int SLN,
                          double *SX, double *SY) {
    for (int z = 0; z < SLN; ++z) {
        for (int y = 0; y < YN; ++y) {
            double *row = SX + ((size_t)z * YN + y) * (XN + 1);
            double acc = 0.0;
            row[0] = 0.0;
            for (int x = 0; x < XN; ++x) {
                acc += vol[((size_t)y * XN + x) * SLN + z] * (double)mask[y * XN + x];
                row[x + 1] = acc;
            }
        }
        for (int x = 0; x < XN; ++x) {
            double *row = SY + ((size_t)z * XN + x) * (YN + 1);
            double acc = 0.0;
            row[0] = 0.0;
            for (int y = 0; y < YN; ++y) {
                acc += vol[((size_t)y * XN + x) * SLN + z] * (double)mask[y * XN + x];
                row[y + 1] = acc;
            }
        }
    }
}

/* FP32 sequential prefix, as horizontalIntegral (MultiSlicesProjGPU.cu:42-51). */
static void orc_vol_sat32(const float *vol, const uint8_t *mask, int XN, int YN, int SLN, float *SX,
                          float *SY) {
    for (int z = 0; z < SLN; ++z) {
        for (int y = 0; y < YN; ++y) {
            float *row = SX + ((size_t)z * YN + y) * (XN + 1);
            row[0] = 0.0f;
            for (int x = 0; x < XN; ++x)
                row[x + 1] = row[x] + vol[((size_t)y * XN + x) * SLN + z] * (float)mask[y * XN + x];
        }
        for (int x = 0; x < XN; ++x) {
            float *row = SY + ((size_t)z * XN + x) * (YN + 1);
            row[0] = 0.0f;
            for (int y = 0; y < YN; ++y)
                row[y + 1] = row[y] + vol[((size_t)y * XN + x) * SLN + z] * (float)mask[y * XN + x];
        }
    }
}

/* Exposed for tests: FP64 SATs of a float volume (layouts as above). */
void orc_vol_sat_f64(const double *vol, const uint8_t *mask, int XN, int YN, int SLN, double *SX,
                     double *SY) {
    orc_vol_sat64(vol, mask, XN, YN, SLN, SX, SY);
}
void orc_vol_sat_f32(const float *vol, const uint8_t *mask, int XN, int YN, int SLN, float *SX,
                     float *SY) {
    orc_vol_sat32(vol, mask, XN, YN, SLN, SX, SY);
}

/* ------------------------------------------------------------------ */
/* a7: DD projector                                                      */
/* ------------------------------------------------------------------ */
typedef struct {
    float x0, y0;
    int DNU, PN, XN, YN, SLN;
    const float *xds, *yds, *bx, *by, *cs;
    float dx, ocx, ocy, invdx;
    const double *SX64, *SY64; /* f64 */
    const float *SX32, *SY32;  /* f32 */
    double *prj64;
    float *prj32;
    int32_t *trace; /* optional [v][u][z][ii][2] texel indices (f64 run only) */
} orc_proj_ctx;

/* Work item = one (view, slice) pair: the branch is fixed for the item, and
 * every slab's table row is read for all DNU detectors before the next slab,
 * so a row stays in the core's cache.  Each detector still sums its slabs in
 * ascending order, the same operations in the same order as one ray at a
 * time, so the results do not depend on this loop order. */
typedef struct {
    float dLx, dLy, dRx, dRy, fL, fR;
    double dLx64, dLy64, dRx64, dRy64, fL64, fR64, adir; /* |normalised ray component| */
    double summ;
} orc_ray64;

static void *orc_proj_worker64(void *arg) {
    orc_task *tk = (orc_task *)arg;
    orc_proj_ctx *C = (orc_proj_ctx *)tk->ctx;
    const int SLN = C->SLN, DNU = C->DNU, XN = C->XN, YN = C->YN;
    const float x0 = C->x0, y0 = C->y0, dx = C->dx, ocx = C->ocx, ocy = C->ocy, invdx = C->invdx;
    const int Smax = XN > YN ? XN : YN;
    orc_ray64 *R = (orc_ray64 *)malloc(sizeof(orc_ray64) * (size_t)DNU);
    if (!R) return NULL;
    for (long vz = tk->tid; vz < (long)C->PN * SLN; vz += tk->nthreads) {
        const int v = (int)(vz / SLN), z = (int)(vz % SLN);
        const float c = C->cs[2 * ((size_t)v * SLN + z)];
        const float s = C->cs[2 * ((size_t)v * SLN + z) + 1];
        /* FP32 index path (MultiSlicesProjGPU.cu:122-137) */
        const float srcx = x0 * c - y0 * s, srcy = x0 * s + y0 * c;
        /* FP64 continuous path */
        const double c64 = c, s64 = s;
        const double srcx64 = (double)x0 * c64 - (double)y0 * s64;
        const double srcy64 = (double)x0 * s64 + (double)y0 * c64;
        const int xslab = fabsf(c) <= fabsf(s); /* MultiSlicesProjGPU.cu:152 */
        for (int u = 0; u < DNU; ++u) {
            orc_ray64 *r = &R[u];
            const float dLx = C->bx[u] * c - C->by[u] * s, dLy = C->bx[u] * s + C->by[u] * c;
            const float dRx = C->bx[u + 1] * c - C->by[u + 1] * s;
            const float dRy = C->bx[u + 1] * s + C->by[u + 1] * c;
            const double dLx64 = (double)C->bx[u] * c64 - (double)C->by[u] * s64;
            const double dLy64 = (double)C->bx[u] * s64 + (double)C->by[u] * c64;
            const double dRx64 = (double)C->bx[u + 1] * c64 - (double)C->by[u + 1] * s64;
            const double dRy64 = (double)C->bx[u + 1] * s64 + (double)C->by[u + 1] * c64;
            double ddx = ((double)C->xds[u] * c64 - (double)C->yds[u] * s64) - srcx64;
            double ddy = ((double)C->xds[u] * s64 + (double)C->yds[u] * c64) - srcy64;
            const double dl = sqrt(ddx * ddx + ddy * ddy);
            ddx /= dl;
            ddy /= dl;
            r->dLx = dLx; r->dLy = dLy; r->dRx = dRx; r->dRy = dRy;
            r->dLx64 = dLx64; r->dLy64 = dLy64; r->dRx64 = dRx64; r->dRy64 = dRy64;
            r->summ = 0.0;
            if (xslab) { /* x-slabs, SAT along y (MultiSlicesProjGPU.cu:152-173) */
                r->fL = (dLy - srcy) / (dLx - srcx);
                r->fR = (dRy - srcy) / (dRx - srcx);
                r->fL64 = (dLy64 - srcy64) / (dLx64 - srcx64);
                r->fR64 = (dRy64 - srcy64) / (dRx64 - srcx64);
                r->adir = fabs(ddx);
            } else { /* y-slabs, SAT along x (MultiSlicesProjGPU.cu:176-196) */
                r->fL = (dLx - srcx) / (dLy - srcy);
                r->fR = (dRx - srcx) / (dRy - srcy);
                r->fL64 = (dLx64 - srcx64) / (dLy64 - srcy64);
                r->fR64 = (dRx64 - srcx64) / (dRy64 - srcy64);
                r->adir = fabs(ddy);
            }
        }
        if (xslab) {
            const double *S = C->SY64 + (size_t)z * XN * (YN + 1);
            for (int ii = 0; ii < XN; ++ii) {
                const float obj = ((float)ii - ocx) * dx;
                const double obj64 = ((double)ii - (double)ocx) * (double)dx;
                const double *row = S + (size_t)ii * (YN + 1);
                for (int u = 0; u < DNU; ++u) {
                    orc_ray64 *r = &R[u];
                    const float rL = (obj - r->dLx) * r->fL + r->dLy;
                    const float rR = (obj - r->dRx) * r->fR + r->dRy;
                    const int iL = orc_texel(rL * invdx + ocy + 1.0f, YN);
                    const int iR = orc_texel(rR * invdx + ocy + 1.0f, YN);
                    const double rL64 = (obj64 - r->dLx64) * r->fL64 + r->dLy64;
                    const double rR64 = (obj64 - r->dRx64) * r->fR64 + r->dRy64;
                    const double TL = orc_fetch64(row, YN, iL, rL64 / (double)dx + (double)ocy + 1.0);
                    const double TR = orc_fetch64(row, YN, iR, rR64 / (double)dx + (double)ocy + 1.0);
                    r->summ += (TR - TL) / (rR64 - rL64);
                    if (C->trace) {
                        int32_t *tr = C->trace + ((((size_t)v * DNU + u) * SLN + z) * (size_t)Smax) * 2;
                        tr[2 * ii] = iL;
                        tr[2 * ii + 1] = iR;
                    }
                }
            }
        } else {
            const double *S = C->SX64 + (size_t)z * YN * (XN + 1);
            for (int ii = 0; ii < YN; ++ii) {
                const float obj = ((float)ii - ocy) * dx;
                const double obj64 = ((double)ii - (double)ocy) * (double)dx;
                const double *row = S + (size_t)ii * (XN + 1);
                for (int u = 0; u < DNU; ++u) {
                    orc_ray64 *r = &R[u];
                    const float rL = (obj - r->dLy) * r->fL + r->dLx;
                    const float rR = (obj - r->dRy) * r->fR + r->dRx;
                    const int iL = orc_texel(rL * invdx + ocx + 1.0f, XN);
                    const int iR = orc_texel(rR * invdx + ocx + 1.0f, XN);
                    const double rL64 = (obj64 - r->dLy64) * r->fL64 + r->dLx64;
                    const double rR64 = (obj64 - r->dRy64) * r->fR64 + r->dRx64;
                    const double TL = orc_fetch64(row, XN, iL, rL64 / (double)dx + (double)ocx + 1.0);
                    const double TR = orc_fetch64(row, XN, iR, rR64 / (double)dx + (double)ocx + 1.0);
                    r->summ += (TR - TL) / (rR64 - rL64);
                    if (C->trace) {
                        int32_t *tr = C->trace + ((((size_t)v * DNU + u) * SLN + z) * (size_t)Smax) * 2;
                        tr[2 * ii] = iL;
                        tr[2 * ii + 1] = iR;
                    }
                }
            }
        }
        for (int u = 0; u < DNU; ++u)
            C->prj64[((size_t)v * DNU + u) * SLN + z] = R[u].summ * (double)dx / R[u].adir;
    }
    free(R);
    return NULL;
}

typedef struct {
    float dLx, dLy, dRx, dRy, fL, fR, cv, summ;
} orc_ray32;

static void *orc_proj_worker32(void *arg) {
    orc_task *tk = (orc_task *)arg;
    orc_proj_ctx *C = (orc_proj_ctx *)tk->ctx;
    const int SLN = C->SLN, DNU = C->DNU, XN = C->XN, YN = C->YN;
    const float x0 = C->x0, y0 = C->y0, dx = C->dx, ocx = C->ocx, ocy = C->ocy, invdx = C->invdx;
    orc_ray32 *R = (orc_ray32 *)malloc(sizeof(orc_ray32) * (size_t)DNU);
    if (!R) return NULL;
    for (long vz = tk->tid; vz < (long)C->PN * SLN; vz += tk->nthreads) {
        const int v = (int)(vz / SLN), z = (int)(vz % SLN);
        const float c = C->cs[2 * ((size_t)v * SLN + z)];
        const float s = C->cs[2 * ((size_t)v * SLN + z) + 1];
        const float srcx = x0 * c - y0 * s, srcy = x0 * s + y0 * c;
        const int xslab = fabsf(c) <= fabsf(s);
        for (int u = 0; u < DNU; ++u) {
            orc_ray32 *r = &R[u];
            const float dLx = C->bx[u] * c - C->by[u] * s, dLy = C->bx[u] * s + C->by[u] * c;
            const float dRx = C->bx[u + 1] * c - C->by[u + 1] * s;
            const float dRy = C->bx[u + 1] * s + C->by[u + 1] * c;
            float ddx = (C->xds[u] * c - C->yds[u] * s) - srcx;
            float ddy = (C->xds[u] * s + C->yds[u] * c) - srcy;
            const float dl = sqrtf(ddx * ddx + ddy * ddy);
            ddx = ddx / dl;
            ddy = ddy / dl;
            r->dLx = dLx; r->dLy = dLy; r->dRx = dRx; r->dRy = dRy;
            r->summ = 0.0f;
            if (xslab) {
                r->fL = (dLy - srcy) / (dLx - srcx);
                r->fR = (dRy - srcy) / (dRx - srcx);
                r->cv = dx / fabsf(ddx);
            } else {
                r->fL = (dLx - srcx) / (dLy - srcy);
                r->fR = (dRx - srcx) / (dRy - srcy);
                r->cv = dx / fabsf(ddy);
            }
        }
        if (xslab) {
            const float *S = C->SY32 + (size_t)z * XN * (YN + 1);
            for (int ii = 0; ii < XN; ++ii) {
                const float obj = ((float)ii - ocx) * dx;
                const float *row = S + (size_t)ii * (YN + 1);
                for (int u = 0; u < DNU; ++u) {
                    orc_ray32 *r = &R[u];
                    const float rL = (obj - r->dLx) * r->fL + r->dLy;
                    const float rR = (obj - r->dRx) * r->fR + r->dRy;
                    const float len = rR - rL;
                    r->summ += (orc_fetch32(row, YN, rR * invdx + ocy + 1.0f) -
                                orc_fetch32(row, YN, rL * invdx + ocy + 1.0f)) / len;
                }
            }
        } else {
            const float *S = C->SX32 + (size_t)z * YN * (XN + 1);
            for (int ii = 0; ii < YN; ++ii) {
                const float obj = ((float)ii - ocy) * dx;
                const float *row = S + (size_t)ii * (XN + 1);
                for (int u = 0; u < DNU; ++u) {
                    orc_ray32 *r = &R[u];
                    const float rL = (obj - r->dLy) * r->fL + r->dLx;
                    const float rR = (obj - r->dRy) * r->fR + r->dRx;
                    const float len = rR - rL;
                    r->summ += (orc_fetch32(row, XN, rR * invdx + ocx + 1.0f) -
                                orc_fetch32(row, XN, rL * invdx + ocx + 1.0f)) / len;
                }
            }
        }
        for (int u = 0; u < DNU; ++u)
            C->prj32[((size_t)v * DNU + u) * SLN + z] = R[u].summ * R[u].cv;
    }
    free(R);
    return NULL;
}

static void orc_proj_setup(orc_proj_ctx *C, float x0, float y0, int DNU, const float *xds,
                           const float *yds, float imgXCenter, float imgYCenter, int PN, int XN,
                           int YN, int SLN, float dx, float *bx, float *by) {
    memset(C, 0, sizeof(*C));
    C->x0 = x0;
    C->y0 = y0;
    C->DNU = DNU;
    C->PN = PN;
    C->XN = XN;
    C->YN = YN;
    C->SLN = SLN;
    C->xds = xds;
    C->yds = yds;
    orc_boundaries(DNU + 1, xds, bx);
    orc_boundaries(DNU + 1, yds, by);
    C->bx = bx;
    C->by = by;
    C->dx = dx;
    /* MultiSlicesProjGPU.cu:225-226 and :149 (float arithmetic) */
    C->ocx = (XN - 1.0f) * 0.5f - imgXCenter / dx;
    C->ocy = (YN - 1.0f) * 0.5f - imgYCenter / dx;
    C->invdx = 1.0f / dx;
}

/* FP64 projector over a double volume.  hangs[v*SLN+z], prj[(v*DNU+u)*SLN+z].
 * trace (optional): int32 [PN][DNU][SLN][max(XN,YN)][2] FP32 texel indices. */
int orc_proj_f64(float x0, float y0, int DNU, const float *xds, const float *yds, float imgXCenter,
                 float imgYCenter, const float *hangs, int PN, int XN, int YN, int SLN,
                 const double *vol, double *prj, float dx, const uint8_t *mask, int32_t *trace,
                 int nthreads) {
    if (DNU < 2 || PN < 0 || XN < 1 || YN < 1 || SLN < 1) return -1;
    orc_proj_ctx C;
    float *bx = (float *)malloc(sizeof(float) * (DNU + 1));
    float *by = (float *)malloc(sizeof(float) * (DNU + 1));
    float *cs = (float *)malloc(sizeof(float) * 2 * (size_t)PN * SLN + 8);
    double *SX = (double *)malloc(sizeof(double) * (size_t)SLN * YN * (XN + 1));
    double *SY = (double *)malloc(sizeof(double) * (size_t)SLN * XN * (YN + 1));
    if (!bx || !by || !cs || !SX || !SY) {
        free(bx); free(by); free(cs); free(SX); free(SY);
        return -2;
    }
    orc_proj_setup(&C, x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, PN, XN, YN, SLN, dx, bx, by);
    orc_cossin(hangs, (long)PN * SLN, cs);
    C.cs = cs;
    orc_vol_sat64(vol, mask, XN, YN, SLN, SX, SY);
    C.SX64 = SX;
    C.SY64 = SY;
    C.prj64 = prj;
    C.trace = trace;
    orc_parallel(nthreads, orc_proj_worker64, &C);
    free(bx); free(by); free(cs); free(SX); free(SY);
    return 0;
}

/* FP32 transliteration of the projector (CPU baseline). */
int orc_proj_f32(float x0, float y0, int DNU, const float *xds, const float *yds, float imgXCenter,
                 float imgYCenter, const float *hangs, int PN, int XN, int YN, int SLN,
                 const float *vol, float *prj, float dx, const uint8_t *mask, int nthreads) {
    if (DNU < 2 || PN < 0 || XN < 1 || YN < 1 || SLN < 1) return -1;
    orc_proj_ctx C;
    float *bx = (float *)malloc(sizeof(float) * (DNU + 1));
    float *by = (float *)malloc(sizeof(float) * (DNU + 1));
    float *cs = (float *)malloc(sizeof(float) * 2 * (size_t)PN * SLN + 8);
    float *SX = (float *)malloc(sizeof(float) * (size_t)SLN * YN * (XN + 1));
    float *SY = (float *)malloc(sizeof(float) * (size_t)SLN * XN * (YN + 1));
    if (!bx || !by || !cs || !SX || !SY) {
        free(bx); free(by); free(cs); free(SX); free(SY);
        return -2;
    }
    orc_proj_setup(&C, x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, PN, XN, YN, SLN, dx, bx, by);
    orc_cossin(hangs, (long)PN * SLN, cs);
    C.cs = cs;
    orc_vol_sat32(vol, mask, XN, YN, SLN, SX, SY);
    C.SX32 = SX;
    C.SY32 = SY;
    C.prj32 = prj;
    orc_parallel(nthreads, orc_proj_worker32, &C);
    free(bx); free(by); free(cs); free(SX); free(SY);
    return 0;
}

/* ------------------------------------------------------------------ */
/* a8-a9: projection SAT and DD backprojector                            */
/* ------------------------------------------------------------------ */
typedef struct {
    float x0, y0, dx, ocx, ocy, detCntIdx, dbeta;
    double ocx64, ocy64;
    int DNU, PN, XN, YN, SLN;
    const float *cs;
    const uint8_t *mask;
    const double *P64; /* [z][v][k] k=0..DNU */
    const float *P32;
    double *vol64;
    float *vol32;
    int32_t *trace; /* optional [y][x][z][v][2] */
} orc_back_ctx;

/* Work item = one image row y; per slice the views run in the outer loop and
 * the row's voxels in the inner one, so each view's table row is read for
 * the whole image row while it is in cache.  Every voxel still sums its
 * views in ascending order with the same operations. */
static void *orc_back_worker64(void *arg) {
    orc_task *tk = (orc_task *)arg;
    orc_back_ctx *C = (orc_back_ctx *)tk->ctx;
    const int SLN = C->SLN, DNU = C->DNU, XN = C->XN, PN = C->PN;
    const float x0 = C->x0, y0 = C->y0, dx = C->dx;
    /* kernel constants (MultiSlicesBackGPU.cu:141,150) */
    const float inv_sid = 1.0f / sqrtf(x0 * x0 + y0 * y0);
    const float invdb = (float)(1.0 / (double)C->dbeta);
    const double inv_sid64 = 1.0 / sqrt((double)x0 * x0 + (double)y0 * y0);
    const double invdb64 = 1.0 / (double)C->dbeta;
    double *acc = (double *)malloc(sizeof(double) * (size_t)XN);
    if (!acc) return NULL;
    for (int y = tk->tid; y < C->YN; y += tk->nthreads) {
        const float vy = ((float)y - C->ocy) * dx;
        const double vy64 = ((double)y - C->ocy64) * dx;
        for (int z = 0; z < SLN; ++z) {
            for (int x = 0; x < XN; ++x) acc[x] = 0.0;
            for (int v = 0; v < PN; ++v) {
                const float c = C->cs[2 * ((size_t)v * SLN + z)];
                const float s = C->cs[2 * ((size_t)v * SLN + z) + 1];
                /* FP32 index path (MultiSlicesBackGPU.cu:155-176) */
                const float srcx = x0 * c - y0 * s, srcy = x0 * s + y0 * c;
                const int bxsel = fabsf(srcx) > fabsf(srcy);
                /* FP64 continuous path */
                const double c64 = c, s64 = s;
                const double srcx64 = (double)x0 * c64 - (double)y0 * s64;
                const double srcy64 = (double)x0 * s64 + (double)y0 * c64;
                const double *row = C->P64 + ((size_t)z * PN + v) * (DNU + 1);
                for (int x = 0; x < XN; ++x) {
                    if (C->mask[y * XN + x] == 0) continue;
                    const float vx = ((float)x - C->ocx) * dx;
                    const double vx64 = ((double)x - C->ocx64) * dx;
                    const float ddx = vx - srcx, ddy = vy - srcy;
                    const float l2 = ddx * ddx + ddy * ddy;
                    const float l = 1.0f / sqrtf(l2);
                    const float alpha = orc_asinf32(((srcy * ddx - srcx * ddy) * inv_sid) * l);
                    const float ddv = bxsel ? ddx : ddy;
                    const float dA = ((ddv / l2) * dx) * 0.5f;
                    const float idxL = ((alpha - dA) * invdb + C->detCntIdx) + 1.0f;
                    const float idxR = ((alpha + dA) * invdb + C->detCntIdx) + 1.0f;
                    const int iL = orc_texel(idxL, DNU), iR = orc_texel(idxR, DNU);
                    const double ex = vx64 - srcx64, ey = vy64 - srcy64;
                    const double l264 = ex * ex + ey * ey;
                    const double al64 = asin((srcy64 * ex - srcx64 * ey) * inv_sid64 / sqrt(l264));
                    const double ddv64 = bxsel ? ex : ey;
                    const double dA64 = ddv64 / l264 * (double)dx * 0.5;
                    const double cosVal = (double)dx / ddv64 * sqrt(l264);
                    const double TL = orc_fetch64(row, DNU, iL, (al64 - dA64) * invdb64 + C->detCntIdx + 1.0);
                    const double TR = orc_fetch64(row, DNU, iR, (al64 + dA64) * invdb64 + C->detCntIdx + 1.0);
                    acc[x] += (TR - TL) * cosVal;
                    if (C->trace) {
                        const size_t o = ((size_t)y * XN + x) * SLN + z;
                        int32_t *tr = C->trace + (o * PN + v) * 2;
                        tr[0] = iL;
                        tr[1] = iR;
                    }
                }
            }
            for (int x = 0; x < XN; ++x)
                C->vol64[((size_t)y * XN + x) * SLN + z] = C->mask[y * XN + x] == 0 ? 0.0 : acc[x];
        }
    }
    free(acc);
    return NULL;
}

static void *orc_back_worker32(void *arg) {
    orc_task *tk = (orc_task *)arg;
    orc_back_ctx *C = (orc_back_ctx *)tk->ctx;
    const int SLN = C->SLN, DNU = C->DNU, XN = C->XN, PN = C->PN;
    const float x0 = C->x0, y0 = C->y0, dx = C->dx;
    const float inv_sid = 1.0f / sqrtf(x0 * x0 + y0 * y0);
    const float invdb = (float)(1.0 / (double)C->dbeta);
    float *acc = (float *)malloc(sizeof(float) * (size_t)XN);
    if (!acc) return NULL;
    for (int y = tk->tid; y < C->YN; y += tk->nthreads) {
        const float vy = ((float)y - C->ocy) * dx;
        for (int z = 0; z < SLN; ++z) {
            for (int x = 0; x < XN; ++x) acc[x] = 0.0f;
            for (int v = 0; v < PN; ++v) {
                const float c = C->cs[2 * ((size_t)v * SLN + z)];
                const float s = C->cs[2 * ((size_t)v * SLN + z) + 1];
                const float srcx = x0 * c - y0 * s, srcy = x0 * s + y0 * c;
                const int bxsel = fabsf(srcx) > fabsf(srcy);
                const float *row = C->P32 + ((size_t)z * PN + v) * (DNU + 1);
                for (int x = 0; x < XN; ++x) {
                    if (C->mask[y * XN + x] == 0) continue;
                    const float vx = ((float)x - C->ocx) * dx;
                    const float ddx = vx - srcx, ddy = vy - srcy;
                    const float l2 = ddx * ddx + ddy * ddy;
                    const float l = 1.0f / sqrtf(l2);
                    const float alpha = orc_asinf32(((srcy * ddx - srcx * ddy) * inv_sid) * l);
                    const float ddv = bxsel ? ddx : ddy;
                    const float dA = ((ddv / l2) * dx) * 0.5f;
                    const float cosVal = (dx / ddv) * sqrtf(l2);
                    const float idxL = ((alpha - dA) * invdb + C->detCntIdx) + 1.0f;
                    const float idxR = ((alpha + dA) * invdb + C->detCntIdx) + 1.0f;
                    acc[x] += (orc_fetch32(row, DNU, idxR) - orc_fetch32(row, DNU, idxL)) * cosVal;
                }
            }
            for (int x = 0; x < XN; ++x)
                C->vol32[((size_t)y * XN + x) * SLN + z] = C->mask[y * XN + x] == 0 ? 0.0f : acc[x];
        }
    }
    free(acc);
    return NULL;
}

static void orc_back_setup(orc_back_ctx *C, float x0, float y0, int DNU, const float *xds,
                           const float *yds, float detCntIdx, float imgXCenter, float imgYCenter,
                           int PN, int XN, int YN, int SLN, float dx, const uint8_t *mask) {
    memset(C, 0, sizeof(*C));
    C->x0 = x0;
    C->y0 = y0;
    C->dx = dx;
    C->DNU = DNU;
    C->PN = PN;
    C->XN = XN;
    C->YN = YN;
    C->SLN = SLN;
    C->detCntIdx = detCntIdx;
    C->mask = mask;
    /* MultiSlicesBackGPU.cu:198-200: double arithmetic, float store */
    C->ocx = (float)((XN - 1.0) * 0.5 - (double)(imgXCenter / dx));
    C->ocy = (float)((YN - 1.0) * 0.5 - (double)(imgYCenter / dx));
    C->ocx64 = C->ocx;
    C->ocy64 = C->ocy;
    float S2D;
    orc_back_scalars(x0, y0, xds, yds, &S2D, &C->dbeta);
}

/* FP64 backprojector of a double sinogram prj[(v*DNU+u)*SLN+z] into
 * vol[(y*XN+x)*SLN+z].  trace (optional): int32 [YN][XN][SLN][PN][2]. */
int orc_back_f64(float x0, float y0, int DNU, const float *xds, const float *yds, float detCntIdx,
                 float imgXCenter, float imgYCenter, const float *hangs, int PN, int XN, int YN,
                 int SLN, double *vol, const double *prj, float dx, const uint8_t *mask,
                 int32_t *trace, int nthreads) {
    if (DNU < 2 || PN < 0 || XN < 1 || YN < 1 || SLN < 1) return -1;
    orc_back_ctx C;
    orc_back_setup(&C, x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, PN, XN, YN, SLN,
                   dx, mask);
    float *cs = (float *)malloc(sizeof(float) * 2 * (size_t)PN * SLN + 8);
    double *P = (double *)malloc(sizeof(double) * (size_t)SLN * PN * (DNU + 1) + 8);
    if (!cs || !P) {
        free(cs); free(P);
        return -2;
    }
    orc_cossin(hangs, (long)PN * SLN, cs);
    /* projection SAT along the detector (MultiSlicesBackGPU.cu:12-32) */
    for (int z = 0; z < SLN; ++z)
        for (int v = 0; v < PN; ++v) {
            double *row = P + ((size_t)z * PN + v) * (DNU + 1);
            double acc = 0.0;
            row[0] = 0.0;
            for (int u = 0; u < DNU; ++u) {
                acc += prj[((size_t)v * DNU + u) * SLN + z];
                row[u + 1] = acc;
            }
        }
    C.cs = cs;
    C.P64 = P;
    C.vol64 = vol;
    C.trace = trace;
    orc_parallel(nthreads, orc_back_worker64, &C);
    free(cs);
    free(P);
    return 0;
}

/* FP32 transliteration of the backprojector (CPU baseline). */
int orc_back_f32(float x0, float y0, int DNU, const float *xds, const float *yds, float detCntIdx,
                 float imgXCenter, float imgYCenter, const float *hangs, int PN, int XN, int YN,
                 int SLN, float *vol, const float *prj, float dx, const uint8_t *mask,
                 int nthreads) {
    if (DNU < 2 || PN < 0 || XN < 1 || YN < 1 || SLN < 1) return -1;
    orc_back_ctx C;
    orc_back_setup(&C, x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, PN, XN, YN, SLN,
                   dx, mask);
    float *cs = (float *)malloc(sizeof(float) * 2 * (size_t)PN * SLN + 8);
    float *P = (float *)malloc(sizeof(float) * (size_t)SLN * PN * (DNU + 1) + 8);
    if (!cs || !P) {
        free(cs); free(P);
        return -2;
    }
    orc_cossin(hangs, (long)PN * SLN, cs);
    for (int z = 0; z < SLN; ++z)
        for (int v = 0; v < PN; ++v) {
            float *row = P + ((size_t)z * PN + v) * (DNU + 1);
            row[0] = 0.0f;
            for (int u = 0; u < DNU; ++u) row[u + 1] = row[u] + prj[((size_t)v * DNU + u) * SLN + z];
        }
    C.cs = cs;
    C.P32 = P;
    C.vol32 = vol;
    orc_parallel(nthreads, orc_back_worker32, &C);
    free(cs);
    free(P);
    return 0;
}

/* ------------------------------------------------------------------ */
/* a11: helical-to-fan rebinning                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    float *Proj, *Views;
    uint8_t *cut;   /* optional [SLN][DefTimes] */
    int32_t *idv;   /* optional [SLN][DefTimes][W] */
    int32_t *start; /* optional [SLN] */
    const float *proj, *zPos;
    long long TotalViews;
    int SLN, W, H, DefTimes, fp64, status;
    float SD, SO, BVAngle, PerDetW, PerDetH, DetCenterW, pitch;
    float PerDetA, delta, DetCenterHreal, h, deltaZ;
} orc_rebin_ctx;

/* AAPM::mod<float> (HelicalToFanFunc_mex.cpp:27-34): repeated float subtraction. */
float orc_mod_f(float lambda, float regV) {
    float v = lambda;
    while (v > regV) v -= regV;
    return v;
}

void orc_rebin_params(float SD, float SO, int DetHeight, float PerDetW, float PerDetH, int DefTimes,
                      float pitch, float *out5) {
    /* HelicalToFanFunc_mex.cpp:160-164 */
    out5[0] = 2.0f * atanf(PerDetW / SD * 0.5f);
    out5[1] = (float)(2.0 * 3.141592653589793 / DefTimes);
    out5[2] = (float)((DetHeight - 1.0) * 0.5 / PerDetH);
    out5[3] = pitch * PerDetH * (float)DetHeight * SO / SD;
    out5[4] = out5[3] / (float)DefTimes;
}

int orc_rebin_start(float z, float h, float deltaZ) { return (int)floorf((z - h * 0.5f) / deltaZ); }

static void *orc_rebin_worker(void *arg) {
    orc_task *tk = (orc_task *)arg;
    orc_rebin_ctx *C = (orc_rebin_ctx *)tk->ctx;
    const int W = C->W, H = C->H, DT = C->DefTimes;
    float *PData = (float *)malloc(sizeof(float) * (size_t)DT * W);
    int *mark = (int *)malloc(sizeof(int) * (size_t)DT * W);
    float *VA = (float *)malloc(sizeof(float) * DT);
    const float twopi_f = (float)(2 * 3.141592653589793);
    for (int iii = tk->tid; iii < C->SLN; iii += tk->nthreads) {
        const float z = C->zPos[iii];
        memset(PData, 0, sizeof(float) * (size_t)DT * W);
        memset(mark, 0, sizeof(int) * (size_t)DT * W);
        memset(VA, 0, sizeof(float) * DT);
        const int start = orc_rebin_start(z, C->h, C->deltaZ);
        if (C->start) C->start[iii] = start;
        if (C->idv) memset(C->idv + (size_t)iii * DT * W, 0, sizeof(int32_t) * (size_t)DT * W);
        /* Loop bounds of :73/:76 under size_t comparison: nothing is processed
         * when start < 0; raw views max(1,start)..start+DefTimes-1 otherwise. */
        if (start >= 0) {
            const long long last = (long long)start + DT - 1;
            for (long long i = start < 1 ? 1 : start; i <= last; ++i) {
                const float *data = C->proj + (size_t)(i - 1) * W * H;
                const float zz = (float)(i - 1) * C->deltaZ;
                const float lambda = (float)(i - 1) * C->delta + C->BVAngle;
                const int ibeta = (int)(i - start + 1);
                VA[ibeta - 1] = orc_mod_f(lambda, twopi_f);
                for (int j = 1; j <= W; ++j) {
                    const float ang = ((float)j - C->DetCenterW) * C->PerDetA;
                    const float SM = C->SO * cosf(ang);
                    const float tanAngV = (z - zz) / SM;
                    const float v = C->SD * tanAngV;
                    const float dv = (float)((double)((C->DetCenterHreal + v) / C->PerDetH) + 1.0);
                    const int idv = (int)floorf(dv);
                    float val = 0.0f;
                    int tm = 0;
                    if (idv >= 1 && idv <= H - 1) {
                        const float d0 = data[(size_t)(j - 1) * H + (idv - 1)];
                        const float d1 = data[(size_t)(j - 1) * H + idv];
                        if (C->fp64) {
                            const double ang64 = ((double)j - (double)C->DetCenterW) * (double)C->PerDetA;
                            const double SM64 = (double)C->SO * cos(ang64);
                            const double zz64 = (double)(i - 1) * (double)C->deltaZ;
                            const double v64 = (double)C->SD * (((double)z - zz64) / SM64);
                            const double dv64 = ((double)C->DetCenterHreal + v64) / (double)C->PerDetH + 1.0;
                            const double t64 = dv64 - (double)idv;
                            const double cosA = (double)C->SD / sqrt((double)C->SD * C->SD + v64 * v64);
                            val = (float)(cosA * ((double)d0 * (1.0 - t64) + (double)d1 * t64));
                        } else {
                            const float t = dv - (float)idv;
                            const float cosAngV = C->SD / sqrtf(C->SD * C->SD + v * v);
                            val = cosAngV * (d0 * (1.0f - t) + d1 * t);
                        }
                        tm = 1;
                    }
                    PData[(size_t)(ibeta - 1) * W + (j - 1)] = val;
                    mark[(size_t)(ibeta - 1) * W + (j - 1)] = tm;
                    if (C->idv) C->idv[((size_t)iii * DT + (ibeta - 1)) * W + (j - 1)] = idv;
                }
            }
        }
        /* :113-123 rows holding any unmarked channel are zeroed, kept */
        for (int a = 0; a < DT; ++a) {
            int isCut = 0;
            for (int j = 0; j < W; ++j)
                if (mark[(size_t)a * W + j] == 0) {
                    isCut = 1;
                    break;
                }
            if (isCut) memset(PData + (size_t)a * W, 0, sizeof(float) * W);
            if (C->cut) C->cut[(size_t)iii * DT + a] = (uint8_t)isCut;
        }
        memcpy(C->Proj + (size_t)iii * DT * W, PData, sizeof(float) * (size_t)DT * W);
        memcpy(C->Views + (size_t)iii * DT, VA, sizeof(float) * DT);
    }
    free(PData);
    free(mark);
    free(VA);
    return NULL;
}

/* Proj[(s*DefTimes+v)*W+j], Views[s*DefTimes+v] from raw proj[(view*W+ch)*H+row].
 * Returns -3 when a slice would read past TotalViews (the reference reads
 * out of bounds there). */
int orc_helical_to_fan(float *Proj, float *Views, uint8_t *cut, int32_t *idv, int32_t *start,
                       const float *proj, long long TotalViews, const float *zPos, int SLN, float SD,
                       float SO, float BVAngle, int DetWidth, int DetHeight, float PerDetW,
                       float PerDetH, int DefTimes, float DetCenterW, float pitch, int fp64,
                       int nthreads) {
    if (SLN < 0 || DetWidth < 1 || DetHeight < 2 || DefTimes < 1) return -1;
    orc_rebin_ctx C;
    memset(&C, 0, sizeof(C));
    float p5[5];
    orc_rebin_params(SD, SO, DetHeight, PerDetW, PerDetH, DefTimes, pitch, p5);
    C.PerDetA = p5[0];
    C.delta = p5[1];
    C.DetCenterHreal = p5[2];
    C.h = p5[3];
    C.deltaZ = p5[4];
    for (int s = 0; s < SLN; ++s) {
        int st = orc_rebin_start(zPos[s], C.h, C.deltaZ);
        if (st >= 0 && (long long)st + DefTimes - 1 > TotalViews) return -3;
    }
    C.Proj = Proj;
    C.Views = Views;
    C.cut = cut;
    C.idv = idv;
    C.start = start;
    C.proj = proj;
    C.zPos = zPos;
    C.TotalViews = TotalViews;
    C.SLN = SLN;
    C.W = DetWidth;
    C.H = DetHeight;
    C.DefTimes = DefTimes;
    C.fp64 = fp64;
    C.SD = SD;
    C.SO = SO;
    C.BVAngle = BVAngle;
    C.PerDetW = PerDetW;
    C.PerDetH = PerDetH;
    C.DetCenterW = DetCenterW;
    C.pitch = pitch;
    orc_parallel(nthreads, orc_rebin_worker, &C);
    return 0;
}

/* ------------------------------------------------------------------ */
/* a13: OS-SART (+FISTA) driver                                          */
/* ------------------------------------------------------------------ */
/* Subset k (0-based) holds views k, k+numOS, ... (OSSART_AAPM.m:60). */
int orc_subset_views(int numViews, int numOS, int k, int32_t *out) {
    int n = 0;
    for (int v = k; v < numViews; v += numOS) out[n++] = v;
    return n;
}

/* FP64 OS-SART.  meas[(v*DNU+u)*SLN+z] over numViews views with angles
 * hangs[v*SLN+z]; x in/out [(y*XN+x)*SLN+z].  Weights from slice 0 are
 * replicated to every slice (OSSART_AAPM.m:61-64). */
int orc_ossart_f64(float x0, float y0, int DNU, const float *xds, const float *yds, float detCntIdx,
                   float imgXCenter, float imgYCenter, const float *hangs, int numViews, int XN,
                   int YN, int SLN, const float *meas, double *x, float dx, const uint8_t *mask,
                   int numIter, int numOS, int useFISTA, int nthreads) {
    if (numOS < 1 || numOS > numViews) return -1;
    const size_t nvox = (size_t)XN * YN * SLN;
    const int maxPN = (numViews + numOS - 1) / numOS;
    int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * maxPN);
    float *hsub = (float *)malloc(sizeof(float) * (size_t)maxPN * SLN);
    float *hsub0 = (float *)malloc(sizeof(float) * (size_t)maxPN);
    double *rowSum = (double *)malloc(sizeof(double) * (size_t)numOS * maxPN * DNU);
    double *colSum = (double *)malloc(sizeof(double) * (size_t)numOS * XN * YN);
    double *ones_v = (double *)malloc(sizeof(double) * (size_t)XN * YN);
    double *ones_p = (double *)malloc(sizeof(double) * (size_t)maxPN * DNU);
    double *prj = (double *)malloc(sizeof(double) * (size_t)maxPN * DNU * SLN);
    double *upd = (double *)malloc(sizeof(double) * nvox);
    double *xprev = (double *)malloc(sizeof(double) * nvox);
    if (!idx || !hsub || !hsub0 || !rowSum || !colSum || !ones_v || !ones_p || !prj || !upd || !xprev) {
        free(idx); free(hsub); free(hsub0); free(rowSum); free(colSum); free(ones_v); free(ones_p);
        free(prj); free(upd); free(xprev);
        return -2;
    }
    for (size_t i = 0; i < (size_t)XN * YN; ++i) ones_v[i] = 1.0;
    for (size_t i = 0; i < (size_t)maxPN * DNU; ++i) ones_p[i] = 1.0;
    /* weights from slice 0 (SLN = 1 with slice-0 angles) */
    for (int k = 0; k < numOS; ++k) {
        int n = orc_subset_views(numViews, numOS, k, idx);
        for (int j = 0; j < n; ++j) hsub0[j] = hangs[(size_t)idx[j] * SLN + 0];
        orc_proj_f64(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, hsub0, n, XN, YN, 1, ones_v,
                     rowSum + (size_t)k * maxPN * DNU, dx, mask, NULL, nthreads);
        orc_back_f64(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, hsub0, n, XN, YN, 1,
                     colSum + (size_t)k * XN * YN, ones_p, dx, mask, NULL, nthreads);
    }
    double t0 = 1.0;
    for (int it = 0; it < numIter; ++it) {
        if (useFISTA) memcpy(xprev, x, sizeof(double) * nvox);
        for (int k = 0; k < numOS; ++k) {
            int n = orc_subset_views(numViews, numOS, k, idx);
            for (int j = 0; j < n; ++j)
                for (int z = 0; z < SLN; ++z) hsub[(size_t)j * SLN + z] = hangs[(size_t)idx[j] * SLN + z];
            orc_proj_f64(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, hsub, n, XN, YN, SLN, x, prj,
                         dx, mask, NULL, nthreads);
            const double *rs = rowSum + (size_t)k * maxPN * DNU;
            for (int j = 0; j < n; ++j)
                for (int u = 0; u < DNU; ++u) {
                    const double w = rs[(size_t)j * DNU + u];
                    if (w == 0.0) continue; /* OSSART_AAPM.m:88 keeps prj */
                    for (int z = 0; z < SLN; ++z) {
                        const size_t o = ((size_t)j * DNU + u) * SLN + z;
                        const double b = meas[((size_t)idx[j] * DNU + u) * SLN + z];
                        prj[o] = (b - prj[o]) / w;
                    }
                }
            orc_back_f64(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, hsub, n, XN, YN,
                         SLN, upd, prj, dx, mask, NULL, nthreads);
            const double *cw = colSum + (size_t)k * XN * YN;
            for (size_t p = 0; p < (size_t)XN * YN; ++p) {
                const double w = cw[p];
                for (int z = 0; z < SLN; ++z) {
                    const size_t o = p * SLN + z;
                    x[o] += (w != 0.0) ? upd[o] / w : upd[o];
                }
            }
        }
        if (useFISTA) { /* OSSART_AAPM.m:111-115 */
            const double t1 = (1.0 + sqrt(1.0 + 4.0 * t0 * t0)) / 2.0;
            const double mom = (t0 - 1.0) / t1;
            for (size_t o = 0; o < nvox; ++o) x[o] = x[o] + mom * (x[o] - xprev[o]);
            t0 = t1;
        }
    }
    free(idx); free(hsub); free(hsub0); free(rowSum); free(colSum); free(ones_v); free(ones_p);
    free(prj); free(upd); free(xprev);
    return 0;
}

/* FP32 OS-SART over the FP32 transliterated operators (CPU baseline of the
 * single-slice end-to-end case).  MATLAB single arithmetic: the FISTA
 * momentum is a double scalar converted to single before the array update. */
int orc_ossart_f32(float x0, float y0, int DNU, const float *xds, const float *yds, float detCntIdx,
                   float imgXCenter, float imgYCenter, const float *hangs, int numViews, int XN,
                   int YN, int SLN, const float *meas, float *x, float dx, const uint8_t *mask,
                   int numIter, int numOS, int useFISTA, int nthreads) {
    if (numOS < 1 || numOS > numViews) return -1;
    const size_t nvox = (size_t)XN * YN * SLN;
    const int maxPN = (numViews + numOS - 1) / numOS;
    int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * maxPN);
    float *hsub = (float *)malloc(sizeof(float) * (size_t)maxPN * SLN);
    float *hsub0 = (float *)malloc(sizeof(float) * (size_t)maxPN);
    float *rowSum = (float *)malloc(sizeof(float) * (size_t)numOS * maxPN * DNU);
    float *colSum = (float *)malloc(sizeof(float) * (size_t)numOS * XN * YN);
    float *ones_v = (float *)malloc(sizeof(float) * (size_t)XN * YN);
    float *ones_p = (float *)malloc(sizeof(float) * (size_t)maxPN * DNU);
    float *prj = (float *)malloc(sizeof(float) * (size_t)maxPN * DNU * SLN);
    float *upd = (float *)malloc(sizeof(float) * nvox);
    float *xprev = (float *)malloc(sizeof(float) * nvox);
    if (!idx || !hsub || !hsub0 || !rowSum || !colSum || !ones_v || !ones_p || !prj || !upd || !xprev) {
        free(idx); free(hsub); free(hsub0); free(rowSum); free(colSum); free(ones_v); free(ones_p);
        free(prj); free(upd); free(xprev);
        return -2;
    }
    for (size_t i = 0; i < (size_t)XN * YN; ++i) ones_v[i] = 1.0f;
    for (size_t i = 0; i < (size_t)maxPN * DNU; ++i) ones_p[i] = 1.0f;
    for (int k = 0; k < numOS; ++k) {
        int n = orc_subset_views(numViews, numOS, k, idx);
        for (int j = 0; j < n; ++j) hsub0[j] = hangs[(size_t)idx[j] * SLN + 0];
        orc_proj_f32(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, hsub0, n, XN, YN, 1, ones_v,
                     rowSum + (size_t)k * maxPN * DNU, dx, mask, nthreads);
        orc_back_f32(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, hsub0, n, XN, YN, 1,
                     colSum + (size_t)k * XN * YN, ones_p, dx, mask, nthreads);
    }
    double t0 = 1.0;
    for (int it = 0; it < numIter; ++it) {
        if (useFISTA) memcpy(xprev, x, sizeof(float) * nvox);
        for (int k = 0; k < numOS; ++k) {
            int n = orc_subset_views(numViews, numOS, k, idx);
            for (int j = 0; j < n; ++j)
                for (int z = 0; z < SLN; ++z) hsub[(size_t)j * SLN + z] = hangs[(size_t)idx[j] * SLN + z];
            orc_proj_f32(x0, y0, DNU, xds, yds, imgXCenter, imgYCenter, hsub, n, XN, YN, SLN, x, prj,
                         dx, mask, nthreads);
            const float *rs = rowSum + (size_t)k * maxPN * DNU;
            for (int j = 0; j < n; ++j)
                for (int u = 0; u < DNU; ++u) {
                    const float w = rs[(size_t)j * DNU + u];
                    if (w == 0.0f) continue;
                    for (int z = 0; z < SLN; ++z) {
                        const size_t o = ((size_t)j * DNU + u) * SLN + z;
                        prj[o] = (meas[((size_t)idx[j] * DNU + u) * SLN + z] - prj[o]) / w;
                    }
                }
            orc_back_f32(x0, y0, DNU, xds, yds, detCntIdx, imgXCenter, imgYCenter, hsub, n, XN, YN,
                         SLN, upd, prj, dx, mask, nthreads);
            const float *cw = colSum + (size_t)k * XN * YN;
            for (size_t p = 0; p < (size_t)XN * YN; ++p) {
                const float w = cw[p];
                for (int z = 0; z < SLN; ++z) {
                    const size_t o = p * SLN + z;
                    x[o] = x[o] + ((w != 0.0f) ? upd[o] / w : upd[o]);
                }
            }
        }
        if (useFISTA) {
            const double t1 = (1.0 + sqrt(1.0 + 4.0 * t0 * t0)) / 2.0;
            const float mom = (float)((t0 - 1.0) / t1);
            for (size_t o = 0; o < nvox; ++o) x[o] = x[o] + mom * (x[o] - xprev[o]);
            t0 = t1;
        }
    }
    free(idx); free(hsub); free(hsub0); free(rowSum); free(colSum); free(ones_v); free(ones_p);
    free(prj); free(upd); free(xprev);
    return 0;
}
