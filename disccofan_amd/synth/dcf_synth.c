/*
 * dcf_synth.c -- seeded synthetic images for the component-tree workloads.
 *
 * Every generator is a pure function of (shape, seed, counts): per-pixel noise
 * comes from a counter-based SplitMix64 hash of (seed, pixel index), and the
 * shape lists (disks, ellipses, spheres, blobs) from a sequential SplitMix64
 * stream, so the output does not depend on the OpenMP thread count.
 * Layout is row-major, x fastest: index = x + y*W + z*W*H.
 *
 * Workloads (sizes are the caller's choice; the benchmark sizes are
 *   1024^2 u8, 8192^2 u16, 512^3 u8, 4096^2 f32->u16 and a 256-image batch):
 *   dcf_synth_disks   uniform noise, 5x5 integer box blur (round half up),
 *                     random disks with grey offsets, clipped (u8 or u16)
 *   dcf_synth_vnoise  6-octave value noise on [0,65535] + random rotated
 *                     ellipses with grey offsets, clipped (u16)
 *   dcf_synth_spheres uniform noise, separable Gaussian blur (sigma 2, radius 6),
 *                     random spheres with grey offsets, clipped (u8)
 *   dcf_synth_blobs   lognormal(0,1) background + Gaussian blobs with
 *                     log-uniform amplitude (f32), and its FP64 quantisation to
 *                     65536 levels q = floor(65535*(v-min)/(max-min)) (u16)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define DCF_SYNTH_API __attribute__((visibility("default")))

static inline uint64_t smix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* sequential stream */
typedef struct { uint64_t s; } rng_t;
static inline uint64_t rng_next(rng_t *r) { r->s += 0x9E3779B97F4A7C15ull; return smix(r->s); }
static inline double rng_u01(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
static inline double rng_uni(rng_t *r, double a, double b) { return a + (b - a) * rng_u01(r); }
static inline int64_t rng_int(rng_t *r, int64_t lo, int64_t hi) /* inclusive */
{
    return lo + (int64_t)(rng_next(r) % (uint64_t)(hi - lo + 1));
}

/* counter-based per-pixel hash */
static inline uint64_t phash(uint64_t seed, uint64_t i) { return smix(seed * 0xD1B54A32D192ED03ull ^ smix(i)); }

static inline int64_t clampi(int64_t v, int64_t lo, int64_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ---------------------------------------------------------------------------------------------
 * disks: noise -> 5x5 box blur -> disks (radius U[3,40], offset U[-60,60] scaled by maxval/255)
 * ------------------------------------------------------------------------------------------- */
DCF_SYNTH_API int dcf_synth_disks(int64_t W, int64_t H, int is16, uint64_t seed, int ndisks, void *out)
{
    if (W <= 0 || H <= 0 || !out) return -1;
    const int64_t maxv = is16 ? 65535 : 255;
    const int64_t n = W * H;
    int32_t *noise = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    int32_t *acc = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    if (!noise || !acc) { free(noise); free(acc); return -2; }

#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) noise[i] = (int32_t)(phash(seed, (uint64_t)i) % (uint64_t)(maxv + 1));

    /* 5x5 box blur, clamp-to-edge, integer mean rounded half up */
#pragma omp parallel for schedule(static)
    for (int64_t y = 0; y < H; y++) {
        for (int64_t x = 0; x < W; x++) {
            int64_t s = 0;
            for (int dy = -2; dy <= 2; dy++) {
                int64_t yy = clampi(y + dy, 0, H - 1);
                for (int dx = -2; dx <= 2; dx++) {
                    int64_t xx = clampi(x + dx, 0, W - 1);
                    s += noise[yy * W + xx];
                }
            }
            acc[y * W + x] = (int32_t)((s + 12) / 25);
        }
    }

    rng_t r = { seed ^ 0xD15C0000D15C0000ull };
    typedef struct { double cx, cy, rad; int32_t off; } disk_t;
    disk_t *dk = (disk_t *)malloc(sizeof(disk_t) * (size_t)(ndisks > 0 ? ndisks : 1));
    if (!dk) { free(noise); free(acc); return -2; }
    for (int k = 0; k < ndisks; k++) {
        dk[k].cx = rng_uni(&r, 0, (double)W);
        dk[k].cy = rng_uni(&r, 0, (double)H);
        dk[k].rad = rng_uni(&r, 3.0, 40.0);
        dk[k].off = (int32_t)(rng_int(&r, -60, 60) * (maxv / 255));
    }
    /* rows are independent: every thread owns whole rows and applies all disks in order */
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t y = 0; y < H; y++) {
        double yc = (double)y + 0.5;
        for (int k = 0; k < ndisks; k++) {
            double dy = yc - dk[k].cy;
            double r2 = dk[k].rad * dk[k].rad - dy * dy;
            if (r2 < 0) continue;
            double hw = sqrt(r2);
            int64_t x0 = (int64_t)ceil(dk[k].cx - hw - 0.5), x1 = (int64_t)floor(dk[k].cx + hw - 0.5);
            x0 = clampi(x0, 0, W); x1 = clampi(x1, -1, W - 1);
            for (int64_t x = x0; x <= x1; x++) acc[y * W + x] += dk[k].off;
        }
    }
    if (is16) {
        uint16_t *o = (uint16_t *)out;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) o[i] = (uint16_t)clampi(acc[i], 0, maxv);
    } else {
        uint8_t *o = (uint8_t *)out;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) o[i] = (uint8_t)clampi(acc[i], 0, maxv);
    }
    free(dk); free(noise); free(acc);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * value noise: 6 octaves, coarsest lattice cell = max(W,H)/8, amplitudes 1/2^o, smoothstep
 * bilinear interpolation; scaled to [0,65535]; + ellipses (semi-axes U[2,200] scaled by
 * max(W,H)/8192, theta U[0,pi), offset U[-4000,4000]); clipped
 * ------------------------------------------------------------------------------------------- */
static inline double lattice(uint64_t seed, int oct, int64_t ix, int64_t iy)
{
    uint64_t k = ((uint64_t)(uint32_t)ix << 32) ^ (uint64_t)(uint32_t)iy;
    return (double)(phash(seed + 0x1000u * (uint64_t)(oct + 1), k) >> 11) * 0x1.0p-53;
}

DCF_SYNTH_API int dcf_synth_vnoise(int64_t W, int64_t H, uint64_t seed, int nell, uint16_t *out)
{
    if (W <= 0 || H <= 0 || !out) return -1;
    const int64_t n = W * H;
    const double big = (double)(W > H ? W : H);
    int32_t *acc = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    if (!acc) return -2;
    double cell0 = big / 8.0;
    if (cell0 < 2.0) cell0 = 2.0;
    double ampsum = 0;
    for (int o = 0; o < 6; o++) ampsum += ldexp(1.0, -o);

#pragma omp parallel for schedule(static)
    for (int64_t y = 0; y < H; y++) {
        for (int64_t x = 0; x < W; x++) {
            double v = 0;
            for (int o = 0; o < 6; o++) {
                double cell = cell0 / (double)(1 << o);
                if (cell < 1.0) cell = 1.0;
                double fx = ((double)x + 0.5) / cell, fy = ((double)y + 0.5) / cell;
                int64_t ix = (int64_t)floor(fx), iy = (int64_t)floor(fy);
                double tx = fx - (double)ix, ty = fy - (double)iy;
                tx = tx * tx * (3 - 2 * tx);
                ty = ty * ty * (3 - 2 * ty);
                double a = lattice(seed, o, ix, iy), b = lattice(seed, o, ix + 1, iy);
                double c = lattice(seed, o, ix, iy + 1), d = lattice(seed, o, ix + 1, iy + 1);
                double top = a + (b - a) * tx, bot = c + (d - c) * tx;
                v += ldexp(1.0, -o) * (top + (bot - top) * ty);
            }
            acc[y * W + x] = (int32_t)floor(65535.0 * v / ampsum);
        }
    }

    rng_t r = { seed ^ 0xE11195E5ull };
    typedef struct { double cx, cy, a, b, c, s; int32_t off; } ell_t;
    ell_t *el = (ell_t *)malloc(sizeof(ell_t) * (size_t)(nell > 0 ? nell : 1));
    if (!el) { free(acc); return -2; }
    double sc = big / 8192.0;
    for (int k = 0; k < nell; k++) {
        el[k].cx = rng_uni(&r, 0, (double)W);
        el[k].cy = rng_uni(&r, 0, (double)H);
        el[k].a = rng_uni(&r, 2.0, 200.0) * sc;
        el[k].b = rng_uni(&r, 2.0, 200.0) * sc;
        if (el[k].a < 1.0) el[k].a = 1.0;
        if (el[k].b < 1.0) el[k].b = 1.0;
        double th = rng_uni(&r, 0, M_PI);
        el[k].c = cos(th);
        el[k].s = sin(th);
        el[k].off = (int32_t)rng_int(&r, -4000, 4000);
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t y = 0; y < H; y++) {
        double yc = (double)y + 0.5;
        for (int k = 0; k < nell; k++) {
            double R = el[k].a > el[k].b ? el[k].a : el[k].b;
            double dy = yc - el[k].cy;
            if (dy < -R || dy > R) continue;
            int64_t x0 = clampi((int64_t)floor(el[k].cx - R), 0, W - 1);
            int64_t x1 = clampi((int64_t)ceil(el[k].cx + R), 0, W - 1);
            for (int64_t x = x0; x <= x1; x++) {
                double dx = (double)x + 0.5 - el[k].cx;
                double u = (dx * el[k].c + dy * el[k].s) / el[k].a;
                double w = (-dx * el[k].s + dy * el[k].c) / el[k].b;
                if (u * u + w * w <= 1.0) acc[y * W + x] += el[k].off;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = (uint16_t)clampi(acc[i], 0, 65535);
    free(el); free(acc);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * spheres: u8 noise -> separable Gaussian blur (sigma 2, radius 6, clamp-to-edge, float) ->
 * spheres (radius U[2,30] scaled by max dim/512, offset U[-50,50]); clipped
 * ------------------------------------------------------------------------------------------- */
DCF_SYNTH_API int dcf_synth_spheres(int64_t W, int64_t H, int64_t D, uint64_t seed, int nsph, uint8_t *out)
{
    if (W <= 0 || H <= 0 || D <= 0 || !out) return -1;
    const int64_t n = W * H * D;
    float *a = (float *)malloc((size_t)n * sizeof(float));
    float *b = (float *)malloc((size_t)n * sizeof(float));
    if (!a || !b) { free(a); free(b); return -2; }
    float wk[13];
    double ws = 0;
    for (int k = -6; k <= 6; k++) { wk[k + 6] = (float)exp(-(double)(k * k) / 8.0); ws += wk[k + 6]; }
    for (int k = 0; k < 13; k++) wk[k] = (float)(wk[k] / ws);

#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) a[i] = (float)(phash(seed, (uint64_t)i) & 255u);

    const int64_t sx = 1, sy = W, sz = W * H;
    const int64_t dims[3] = { W, H, D };
    const int64_t strides[3] = { sx, sy, sz };
    float *src = a, *dst = b;
    for (int ax = 0; ax < 3; ax++) {
        const int64_t len = dims[ax], st = strides[ax];
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            int64_t c = (i / st) % len;
            int64_t base = i - c * st;
            float s = 0.f;
            for (int k = -6; k <= 6; k++) s += wk[k + 6] * src[base + clampi(c + k, 0, len - 1) * st];
            dst[i] = s;
        }
        float *t = src; src = dst; dst = t;
    }
    int32_t *acc = (int32_t *)dst; /* reuse the other buffer */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) acc[i] = (int32_t)lrintf(src[i]);

    rng_t r = { seed ^ 0x5FE4E5ull };
    typedef struct { double cx, cy, cz, rad; int32_t off; } sph_t;
    sph_t *sp = (sph_t *)malloc(sizeof(sph_t) * (size_t)(nsph > 0 ? nsph : 1));
    if (!sp) { free(a); free(b); return -2; }
    double big = (double)(W > H ? (W > D ? W : D) : (H > D ? H : D));
    double sc = big / 512.0;
    for (int k = 0; k < nsph; k++) {
        sp[k].cx = rng_uni(&r, 0, (double)W);
        sp[k].cy = rng_uni(&r, 0, (double)H);
        sp[k].cz = rng_uni(&r, 0, (double)D);
        sp[k].rad = rng_uni(&r, 2.0, 30.0) * sc;
        if (sp[k].rad < 1.0) sp[k].rad = 1.0;
        sp[k].off = (int32_t)rng_int(&r, -50, 50);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t z = 0; z < D; z++) {
        double zc = (double)z + 0.5;
        for (int k = 0; k < nsph; k++) {
            double dz = zc - sp[k].cz;
            double rz2 = sp[k].rad * sp[k].rad - dz * dz;
            if (rz2 < 0) continue;
            double rz = sqrt(rz2);
            int64_t y0 = clampi((int64_t)floor(sp[k].cy - rz), 0, H - 1);
            int64_t y1 = clampi((int64_t)ceil(sp[k].cy + rz), 0, H - 1);
            for (int64_t y = y0; y <= y1; y++) {
                double dy = (double)y + 0.5 - sp[k].cy;
                double r2 = rz2 - dy * dy;
                if (r2 < 0) continue;
                double hw = sqrt(r2);
                int64_t x0 = (int64_t)ceil(sp[k].cx - hw - 0.5), x1 = (int64_t)floor(sp[k].cx + hw - 0.5);
                x0 = clampi(x0, 0, W); x1 = clampi(x1, -1, W - 1);
                for (int64_t x = x0; x <= x1; x++) acc[z * sz + y * W + x] += sp[k].off;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = (uint8_t)clampi(acc[i], 0, 255);
    free(sp); free(a); free(b);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * blobs: lognormal(0,1) background + Gaussian blobs (sigma U[1,15] scaled by max(W,H)/4096,
 * amplitude log-uniform [1,1000], footprint radius 3 sigma).  raw (f32, optional) and the
 * FP64 quantisation q = floor(65535 (v-min)/(max-min)) (u16, optional).
 * ------------------------------------------------------------------------------------------- */
DCF_SYNTH_API int dcf_synth_blobs(int64_t W, int64_t H, uint64_t seed, int nblobs, float *raw, uint16_t *quant)
{
    if (W <= 0 || H <= 0) return -1;
    const int64_t n = W * H;
    double *v = (double *)malloc((size_t)n * sizeof(double));
    if (!v) return -2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        uint64_t h1 = phash(seed, 2 * (uint64_t)i), h2 = phash(seed, 2 * (uint64_t)i + 1);
        double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53, u2 = (double)(h2 >> 11) * 0x1.0p-53;
        double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
        v[i] = exp(z);
    }
    rng_t r = { seed ^ 0xB10B5ull };
    typedef struct { double cx, cy, sig, amp; } blob_t;
    blob_t *bl = (blob_t *)malloc(sizeof(blob_t) * (size_t)(nblobs > 0 ? nblobs : 1));
    if (!bl) { free(v); return -2; }
    double sc = (double)(W > H ? W : H) / 4096.0;
    for (int k = 0; k < nblobs; k++) {
        bl[k].cx = rng_uni(&r, 0, (double)W);
        bl[k].cy = rng_uni(&r, 0, (double)H);
        bl[k].sig = rng_uni(&r, 1.0, 15.0) * sc;
        if (bl[k].sig < 0.5) bl[k].sig = 0.5;
        bl[k].amp = exp(rng_uni(&r, 0.0, log(1000.0)));
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t y = 0; y < H; y++) {
        double yc = (double)y + 0.5;
        for (int k = 0; k < nblobs; k++) {
            double R = 3.0 * bl[k].sig, dy = yc - bl[k].cy;
            if (dy < -R || dy > R) continue;
            int64_t x0 = clampi((int64_t)floor(bl[k].cx - R), 0, W - 1);
            int64_t x1 = clampi((int64_t)ceil(bl[k].cx + R), 0, W - 1);
            double inv = 1.0 / (2.0 * bl[k].sig * bl[k].sig);
            for (int64_t x = x0; x <= x1; x++) {
                double dx = (double)x + 0.5 - bl[k].cx;
                v[y * W + x] += bl[k].amp * exp(-(dx * dx + dy * dy) * inv);
            }
        }
    }
    double mn = v[0], mx = v[0];
    for (int64_t i = 0; i < n; i++) { if (v[i] < mn) mn = v[i]; if (v[i] > mx) mx = v[i]; }
    double range = mx > mn ? mx - mn : 1.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        if (raw) raw[i] = (float)v[i];
        if (quant) {
            double q = floor(65535.0 * (v[i] - mn) / range);
            quant[i] = (uint16_t)(q < 0 ? 0 : (q > 65535 ? 65535 : q));
        }
    }
    free(bl); free(v);
    return 0;
}

/* per-image shapes and seeds of the uneven batch: W,H ~ U[lo,hi], even index u8, odd u16 */
DCF_SYNTH_API int dcf_synth_batch_shapes(int count, int64_t lo, int64_t hi, uint64_t seed,
                                         int64_t *W, int64_t *H, int *is16, uint64_t *seeds, double *lambdas,
                                         double lam_lo, double lam_hi)
{
    if (count < 0 || lo <= 0 || hi < lo) return -1;
    rng_t r = { seed ^ 0xBA7C4ull };
    for (int i = 0; i < count; i++) {
        W[i] = rng_int(&r, lo, hi);
        H[i] = rng_int(&r, lo, hi);
        is16[i] = i & 1;
        seeds[i] = rng_next(&r);
        lambdas[i] = rng_uni(&r, lam_lo, lam_hi);
    }
    return 0;
}
