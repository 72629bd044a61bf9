/*
 * dcf_oracle.c -- CPU restatement of disccofan's component-tree path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product path is
 * compared against, and the timed CPU baseline.  Nothing in the product library
 * (disccofan_amd/) links, loads or calls it.
 *
 * What it restates (reference file:line, paths relative to the reference repository):
 *   grey key            transform_float                 src/sources/flood.c:362-371
 *   ranks               LSD radix, 4 x 8-bit digits      src/sources/flood.c:378-502
 *   hierarchical queue  Teeninga bit-queue over ranks    src/sources/queue.c:117-167
 *   flooding            tree_flood/remaining/check_nb    src/sources/flood.c:222-247,556-690
 *   canonical form      compression loop, get_levelroot  src/sources/flood.c:127-141,693-722
 *   slab fuse (omp)     fuse_sections / merge_nodes      src/sources/flood.c:79-165,253-353
 *   re-accumulation     flood_attributes                 src/sources/flood.c:184-219
 *   attributes          area/extent/mean/wmean/inertia   src/sources/attributes.c:47-874,1201-1248
 *   min-tree inversion  preprocessing                    src/sources/preprocessing.c:7-30,84-91
 *   filters             direct/min/max/subtractive/check src/sources/filtering.c:18-261
 *   extract / invert    perform_filter_or_extract        src/sources/operation.c:10-29
 *   spectra             find_scale, 1D/2D single-process src/sources/pattern.c:7-24,80-105,170-202
 *   lambda files        lambda_vector_read               src/sources/lambdavec.c:35-84
 *   connectivity map    check_connectivity               src/sources/io.c:64-79
 *
 * Eigen-based attributes (14 elong, 15 flat, 16 spars) use a closed-form symmetric
 * 2x2 / 3x3 eigen-solver instead of EISPACK rs() (attributes.c:654,679); their parity
 * is within FP64 rounding, not bit-exact.
 *
 * How each part is pinned (tests/):
 *   test_oracle_golden.py  the reference's own fixtures: area and non-compactness check images,
 *                          direct-rule filters, GT_tree.txt (parents, moments, flux), lambda files,
 *                          thread-count invariance of the slab fuse
 *   test_oracle_brute.py   parents and areas against connected components of every level set (2D,
 *                          3D, all connectivities, integer and float greys); min-tree = max-tree of
 *                          the inverse; the direct rule against its definition; closed forms of the
 *                          min, max and subtractive scans for every attribute; the eigenvalue
 *                          attributes against LAPACK; 2D spectrum marginals; find_scale
 * No function is left unpinned.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))
#define BOTTOM (-1)

/* ----------------------------------------------------------------------------------------- */
/* attribute records (types.h:251-286)                                                       */
/* ----------------------------------------------------------------------------------------- */
typedef struct { uint64_t area; } a_area_t;
typedef struct { uint64_t area, minX, minY, minZ, maxX, maxY, maxZ; } a_ext_t;
typedef struct { uint64_t area, sumX, sumY, sumZ; } a_mean_t;
typedef struct { uint64_t area; double sumG, sumXd, sumYd, sumZd; } a_wmean_t;
typedef struct {
    uint64_t area;
    double sval, sval2, sx, sy, sz, sx2, sy2, sz2, sxy, syz, sxz, sxd, syd, szd;
} a_inert_t;

static int group_of(int idx)
{
    if (idx == 0) return 0;
    if (idx <= 5) return 1;
    if (idx <= 8) return 2;
    if (idx <= 12) return 3;
    return 4;
}

static size_t group_size(int grp)
{
    switch (grp) {
    case 0: return sizeof(a_area_t);
    case 1: return sizeof(a_ext_t);
    case 2: return sizeof(a_mean_t);
    case 3: return sizeof(a_wmean_t);
    default: return sizeof(a_inert_t);
    }
}

typedef struct orc_tree {
    int64_t dims[3];
    int64_t n;
    int conn;
    int tree_min;
    float gmax, gmin;
    float pixdim[3];
    float *g;          /* grey values the tree is built on (inverted for a min-tree) */
    int64_t *parent;   /* canonical parent array */
    int grp;
    size_t asz;
    unsigned char *attr;
    int nthreads;
} orc_tree;

#define AREC(t, i) ((void *)((t)->attr + (size_t)(i) * (t)->asz))

static void attr_init(int grp, void *a, int init, uint64_t x, uint64_t y, uint64_t z, float gv)
{
    switch (grp) {
    case 0: ((a_area_t *)a)->area = (uint64_t)init; break;
    case 1: {
        a_ext_t *e = (a_ext_t *)a;
        e->area = (uint64_t)init;
        if (init) { e->minX = x; e->maxX = x + 1; e->minY = y; e->maxY = y + 1; e->minZ = z; e->maxZ = z + 1; }
        else { e->minX = e->minY = e->minZ = ULONG_MAX; e->maxX = e->maxY = e->maxZ = 0; }
    } break;
    case 2: {
        a_mean_t *m = (a_mean_t *)a;
        m->area = (uint64_t)init;
        m->sumX = init ? x : 0; m->sumY = init ? y : 0; m->sumZ = init ? z : 0;
    } break;
    case 3: {
        a_wmean_t *w = (a_wmean_t *)a;
        w->area = (uint64_t)init;
        w->sumG = (double)init * (double)gv;
        w->sumXd = init ? (double)gv * (double)x : 0.;
        w->sumYd = init ? (double)gv * (double)y : 0.;
        w->sumZd = init ? (double)gv * (double)z : 0.;
    } break;
    default: {
        a_inert_t *q = (a_inert_t *)a;
        q->area = (uint64_t)init;
        q->sx = init ? (double)x : 0.; q->sy = init ? (double)y : 0.; q->sz = init ? (double)z : 0.;
        q->sx2 = init ? (double)x * x : 0.; q->sy2 = init ? (double)y * y : 0.; q->sz2 = init ? (double)z * z : 0.;
        q->sxy = init ? (double)x * y : 0.; q->syz = init ? (double)y * z : 0.; q->sxz = init ? (double)z * x : 0.;
        q->sval = init ? (double)gv : 0; q->sval2 = init ? (double)gv * gv : 0;
        q->sxd = init ? (double)gv * x : 0.; q->syd = init ? (double)gv * y : 0.; q->szd = init ? (double)gv * z : 0.;
    } break;
    }
}

/* parent += child  (merge_*_data) */
static void attr_merge(int grp, void *pa, const void *ch)
{
    switch (grp) {
    case 0: ((a_area_t *)pa)->area += ((const a_area_t *)ch)->area; break;
    case 1: {
        a_ext_t *p = (a_ext_t *)pa; const a_ext_t *c = (const a_ext_t *)ch;
        p->area += c->area;
        if (c->minX < p->minX) p->minX = c->minX;
        if (c->minY < p->minY) p->minY = c->minY;
        if (c->minZ < p->minZ) p->minZ = c->minZ;
        if (c->maxX > p->maxX) p->maxX = c->maxX;
        if (c->maxY > p->maxY) p->maxY = c->maxY;
        if (c->maxZ > p->maxZ) p->maxZ = c->maxZ;
    } break;
    case 2: {
        a_mean_t *p = (a_mean_t *)pa; const a_mean_t *c = (const a_mean_t *)ch;
        p->area += c->area; p->sumX += c->sumX; p->sumY += c->sumY; p->sumZ += c->sumZ;
    } break;
    case 3: {
        a_wmean_t *p = (a_wmean_t *)pa; const a_wmean_t *c = (const a_wmean_t *)ch;
        p->area += c->area; p->sumG += c->sumG; p->sumXd += c->sumXd; p->sumYd += c->sumYd; p->sumZd += c->sumZd;
    } break;
    default: {
        a_inert_t *p = (a_inert_t *)pa; const a_inert_t *c = (const a_inert_t *)ch;
        p->area += c->area;
        p->sx += c->sx; p->sy += c->sy; p->sz += c->sz;
        p->sx2 += c->sx2; p->sy2 += c->sy2; p->sz2 += c->sz2;
        p->sxy += c->sxy; p->syz += c->syz; p->sxz += c->sxz;
        p->sval += c->sval; p->sval2 += c->sval2;
        p->sxd += c->sxd; p->syd += c->syd; p->szd += c->szd;
    } break;
    }
}

/* closed-form eigenvalues of a symmetric matrix, ascending (replaces EISPACK rs) */
static void eig2(double a, double b, double d, double ev[2])
{
    double tr = 0.5 * (a + d), df = 0.5 * (a - d);
    double r = sqrt(df * df + b * b);
    ev[0] = tr - r; ev[1] = tr + r;
}

static void eig3(const double m[6] /* xx yy zz xy yz xz */, double ev[3])
{
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5];
    double p1 = d * d + e * e + f * f;
    if (p1 == 0.0) {
        double t[3] = { a, b, c };
        for (int i = 0; i < 3; i++)
            for (int j = i + 1; j < 3; j++)
                if (t[j] < t[i]) { double s = t[i]; t[i] = t[j]; t[j] = s; }
        ev[0] = t[0]; ev[1] = t[1]; ev[2] = t[2];
        return;
    }
    double q = (a + b + c) / 3.0;
    double p2 = (a - q) * (a - q) + (b - q) * (b - q) + (c - q) * (c - q) + 2.0 * p1;
    double p = sqrt(p2 / 6.0);
    double B00 = (a - q) / p, B11 = (b - q) / p, B22 = (c - q) / p, B01 = d / p, B12 = e / p, B02 = f / p;
    double detB = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
    double r = detB / 2.0;
    double phi = r <= -1 ? M_PI / 3.0 : (r >= 1 ? 0.0 : acos(r) / 3.0);
    double e1 = q + 2.0 * p * cos(phi);
    double e3 = q + 2.0 * p * cos(phi + 2.0 * M_PI / 3.0);
    double e2 = 3.0 * q - e1 - e3;
    ev[0] = e3; ev[1] = e2; ev[2] = e1;
}

/* attribute value of one record; attributes.c area_attribute .. inertia_ncom_attribute */
static double attr_value_rec(const orc_tree *t, int idx, const void *a)
{
    const double p0 = t->pixdim[0], p1 = t->pixdim[1], p2 = t->pixdim[2];
    switch (idx) {
    case 0: return t->pixdim[0] * t->pixdim[1] * t->pixdim[2] * ((const a_area_t *)a)->area;
    case 1: {
        const a_ext_t *e = (const a_ext_t *)a;
        double v = (double)(e->maxX - e->minX) * (double)(e->maxY - e->minY) * (double)(e->maxZ - e->minZ);
        v *= t->pixdim[0] * t->pixdim[1] * t->pixdim[2];
        return v;
    }
    case 2: {
        const a_ext_t *e = (const a_ext_t *)a;
        double mnx = p0 * (double)e->minX, mny = p1 * (double)e->minY, mnz = p2 * (double)e->minZ;
        double mxx = p0 * (double)e->maxX, mxy = p1 * (double)e->maxY, mxz = p2 * (double)e->maxZ;
        return (mxx - mnx) * (mxx - mnx) + (mxy - mny) * (mxy - mny) + (mxz - mnz) * (mxz - mnz);
    }
    case 3: { const a_ext_t *e = (const a_ext_t *)a; return t->pixdim[0] * (double)(e->maxX - e->minX); }
    case 4: { const a_ext_t *e = (const a_ext_t *)a; return t->pixdim[1] * (double)(e->maxY - e->minY); }
    case 5: { const a_ext_t *e = (const a_ext_t *)a; return t->pixdim[2] * (double)(e->maxZ - e->minZ); }
    case 6: { const a_mean_t *m = (const a_mean_t *)a; return t->pixdim[0] * (double)m->sumX / (double)m->area; }
    case 7: { const a_mean_t *m = (const a_mean_t *)a; return t->pixdim[1] * (double)m->sumY / (double)m->area; }
    case 8: { const a_mean_t *m = (const a_mean_t *)a; return t->pixdim[2] * (double)m->sumZ / (double)m->area; }
    case 9: { const a_wmean_t *w = (const a_wmean_t *)a; return (double)t->pixdim[0] * w->sumXd / w->sumG; }
    case 10: { const a_wmean_t *w = (const a_wmean_t *)a; return (double)t->pixdim[1] * w->sumYd / w->sumG; }
    case 11: { const a_wmean_t *w = (const a_wmean_t *)a; return (double)t->pixdim[2] * w->sumZd / w->sumG; }
    case 12: return ((const a_wmean_t *)a)->sumG;
    default: break;
    }
    /* inertia group: inertia_arr (attributes.c:607-626) then per-attribute formulas */
    const a_inert_t *q = (const a_inert_t *)a;
    double in[15];
    in[0] = (double)t->pixdim[0] * t->pixdim[1] * t->pixdim[2] * q->area;
    in[1] = t->pixdim[0] * q->sx; in[2] = t->pixdim[1] * q->sy; in[3] = t->pixdim[2] * q->sz;
    in[4] = t->pixdim[0] * t->pixdim[0] * q->sx2;
    in[5] = t->pixdim[1] * t->pixdim[1] * q->sy2;
    in[6] = t->pixdim[2] * t->pixdim[2] * q->sz2;
    in[7] = t->pixdim[0] * t->pixdim[1] * q->sxy;
    in[8] = t->pixdim[1] * t->pixdim[2] * q->syz;
    in[9] = t->pixdim[0] * t->pixdim[2] * q->sxz;
    in[10] = q->sval; in[11] = q->sval2;
    in[12] = t->pixdim[0] * q->sxd; in[13] = t->pixdim[1] * q->syd; in[14] = t->pixdim[2] * q->szd;
    double corr = in[0] / 12;
    double mxx = in[4] - in[1] * in[1] / in[0] + t->pixdim[0] * corr;
    double myy = in[5] - in[2] * in[2] / in[0] + t->pixdim[1] * corr;
    if (idx == 13 || idx == 17) {
        if (in[3] == 0) {
            double tr = mxx + myy;
            return idx == 13 ? tr : tr / pow(in[0], 2.0);
        }
        double mzz = in[6] - in[3] * in[3] / in[0] + t->pixdim[2] * corr;
        double tr = mxx + myy + mzz;
        return idx == 13 ? tr : tr / pow(in[0], 5.0 / 3.0);
    }
    double mxy = in[7] - in[1] * in[2] / in[0];
    if (in[3] == 0) {
        if (idx == 15) return -1;
        double ev[2];
        eig2(mxx, mxy, myy, ev);
        if (idx == 14) return ev[1] / ev[0];
        double ax0 = sqrt(4 * ev[1] / in[0]), ax1 = sqrt(4 * ev[0] / in[0]);
        return M_PI * ax0 * ax1 / in[0];
    }
    double m6[6] = { mxx, myy, in[6] - in[3] * in[3] / in[0] + t->pixdim[2] * corr, mxy, in[8] - in[2] * in[3] / in[0],
                     in[9] - in[1] * in[3] / in[0] };
    double ev[3];
    eig3(m6, ev);
    if (idx == 14) return ev[2] / ev[1];
    if (idx == 15) return ev[1] / ev[0];
    double a0 = sqrt(20 * ev[2] / in[0]), a1 = sqrt(20 * ev[1] / in[0]), a2 = sqrt(20 * ev[0] / in[0]);
    return M_PI * a0 * a1 * a2 / (6 * in[0]);
}

/* the "area" function of an attribute's vtable entry (attributes.c:1224-1248) */
static double attr_area_rec(const orc_tree *t, int idx, const void *a)
{
    int grp = group_of(idx);
    uint64_t ar = *(const uint64_t *)a; /* every record starts with its area count */
    switch (grp) {
    case 0: return t->pixdim[0] * t->pixdim[1] * t->pixdim[2] * ar;            /* float product (area_attribute) */
    case 2: return t->pixdim[0] * t->pixdim[1] * t->pixdim[2] * (double)ar;    /* mean_area_attribute */
    case 4: return (double)t->pixdim[0] * t->pixdim[1] * t->pixdim[2] * ar;    /* inertia_area_attribute */
    default: return (double)ar; /* extent_area_attribute, wmean_area_attribute: no pixdim */
    }
}

/* ----------------------------------------------------------------------------------------- */
/* grey keys and ranks                                                                       */
/* ----------------------------------------------------------------------------------------- */
static inline uint32_t fkey(float v)
{
    uint32_t u;
    memcpy(&u, &v, sizeof u);
    return (u & 0x80000000u) ? 0xFFFFFFFFu - u : (u | 0x80000000u);
}

/* rank -> local index over [lwb,upb): stable sort of (key, index) by 4 LSD 8-bit passes */
static uint32_t *ranks_inverse(const float *g, int64_t lwb, int64_t upb)
{
    int64_t m = upb - lwb;
    uint32_t *k0 = (uint32_t *)malloc((size_t)m * 4), *k1 = (uint32_t *)malloc((size_t)m * 4);
    uint32_t *i0 = (uint32_t *)malloc((size_t)m * 4), *i1 = (uint32_t *)malloc((size_t)m * 4);
    if (!k0 || !k1 || !i0 || !i1) { free(k0); free(k1); free(i0); free(i1); return NULL; }
    uint64_t hist[4][256];
    memset(hist, 0, sizeof hist);
    for (int64_t i = 0; i < m; i++) {
        uint32_t k = fkey(g[lwb + i]);
        k0[i] = k; i0[i] = (uint32_t)i;
        for (int d = 0; d < 4; d++) hist[d][(k >> (8 * d)) & 255]++;
    }
    for (int d = 0; d < 4; d++) {
        uint64_t s = 0;
        for (int b = 0; b < 256; b++) { uint64_t c = hist[d][b]; hist[d][b] = s; s += c; }
    }
    for (int d = 0; d < 4; d++) {
        for (int64_t i = 0; i < m; i++) {
            uint32_t k = k0[i];
            uint64_t pos = hist[d][(k >> (8 * d)) & 255]++;
            k1[pos] = k; i1[pos] = i0[i];
        }
        uint32_t *t = k0; k0 = k1; k1 = t;
        t = i0; i0 = i1; i1 = t;
    }
    free(k0); free(k1); free(i1);
    return i0;
}

/* ----------------------------------------------------------------------------------------- */
/* hierarchical bit queue over ranks (max-priority)                                          */
/* ----------------------------------------------------------------------------------------- */
typedef struct {
    uint64_t *words;
    int64_t off[8];  /* word offset of each level, level 0 = single top word */
    int nlev;
    int64_t top;     /* highest queued rank, -1 when empty */
} bitq_t;

static int bitq_init(bitq_t *q, int64_t nranks)
{
    int64_t cnt[8];
    int L = 0;
    int64_t w = (nranks + 63) / 64;
    if (w < 1) w = 1;
    for (;;) {
        cnt[L++] = w;
        if (w == 1) break;
        w = (w + 63) / 64;
    }
    q->nlev = L;
    int64_t total = 0;
    for (int l = 0; l < L; l++) { q->off[l] = total; total += cnt[L - 1 - l]; }
    q->words = (uint64_t *)calloc((size_t)total, 8);
    q->top = -1;
    return q->words ? 0 : -1;
}

static inline void bitq_insert(bitq_t *q, int64_t r)
{
    int64_t idx = r;
    for (int l = q->nlev - 1; l >= 0; l--) {
        uint64_t *wd = &q->words[q->off[l] + (idx >> 6)];
        uint64_t before = *wd;
        *wd = before | (1ull << (idx & 63));
        if (before) break;
        idx >>= 6;
    }
    if (r > q->top) q->top = r;
}

static inline void bitq_pop(bitq_t *q)
{
    int64_t idx = q->top;
    for (int l = q->nlev - 1; l >= 0; l--) {
        uint64_t *wd = &q->words[q->off[l] + (idx >> 6)];
        *wd &= ~(1ull << (idx & 63));
        if (*wd) break;
        idx >>= 6;
    }
    if (q->words[0] == 0) { q->top = -1; return; }
    int64_t i = 0;
    for (int l = 0; l < q->nlev; l++) i = i * 64 + (63 - __builtin_clzll(q->words[q->off[l] + i]));
    q->top = i;
}

/* ----------------------------------------------------------------------------------------- */
/* flooding of one slab [lwb,upb) into the rank tree                                          */
/* ----------------------------------------------------------------------------------------- */
static int nb_offsets(const int64_t dims[3], int conn, int dx[26], int dy[26], int dz[26])
{
    static const int b[6][3] = { { -1, 0, 0 }, { 1, 0, 0 }, { 0, -1, 0 }, { 0, 1, 0 }, { 0, 0, -1 }, { 0, 0, 1 } };
    static const int d8[4][3] = { { -1, -1, 0 }, { 1, -1, 0 }, { -1, 1, 0 }, { 1, 1, 0 } };
    int c = 0;
    for (int i = 0; i < 6; i++) { dx[c] = b[i][0]; dy[c] = b[i][1]; dz[c] = b[i][2]; c++; }
    if (conn >= 8)
        for (int i = 0; i < 4; i++) { dx[c] = d8[i][0]; dy[c] = d8[i][1]; dz[c] = d8[i][2]; c++; }
    if (conn == 26) {
        for (int sz = -1; sz <= 1; sz += 2)
            for (int sy = -1; sy <= 1; sy++)
                for (int sx = -1; sx <= 1; sx++) {
                    if (sx == 0 && sy == 0) continue;
                    dx[c] = sx; dy[c] = sy; dz[c] = sz; c++;
                }
    }
    (void)dims;
    return c;
}

static int flood_slab(orc_tree *t, int64_t lwb, int64_t upb, int with_attr)
{
    const int64_t W = t->dims[0], H = t->dims[1], D = t->dims[2], WH = W * H;
    const int64_t m = upb - lwb;
    uint32_t *rinv = ranks_inverse(t->g, lwb, upb);
    if (!rinv) return -1;
    uint32_t *rank = (uint32_t *)malloc((size_t)m * 4);
    uint64_t *vis = (uint64_t *)calloc((size_t)((m + 63) / 64), 8);
    bitq_t q;
    if (!rank || !vis || bitq_init(&q, m) != 0) { free(rinv); free(rank); free(vis); return -1; }
    for (int64_t r = 0; r < m; r++) rank[rinv[r]] = (uint32_t)r;
    int dx[26], dy[26], dz[26];
    int nn = nb_offsets(t->dims, t->conn, dx, dy, dz);

    int64_t cur = lwb;
    int64_t crank = rank[0];
    vis[0] |= 1ull;
    int64_t x = cur % W, y = (cur % WH) / W, z = cur / WH;
    for (;;) {
        int climbed;
        do {
            climbed = 0;
            for (int k = 0; k < nn; k++) {
                int64_t nx = x + dx[k], ny = y + dy[k], nz = z + dz[k];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) continue;
                int64_t nbi = nx + ny * W + nz * WH;
                if (nbi < lwb || nbi >= upb) continue;
                int64_t li = nbi - lwb;
                if (vis[li >> 6] & (1ull << (li & 63))) continue;
                vis[li >> 6] |= 1ull << (li & 63);
                int64_t nr = rank[li];
                if (crank > nr) { bitq_insert(&q, nr); continue; }
                bitq_insert(&q, crank);
                cur = nbi; crank = nr; x = nx; y = ny; z = nz;
                climbed = 1;
                break;
            }
        } while (climbed);
        if (q.top < 0) break;
        int64_t prank = q.top;
        int64_t par = (int64_t)rinv[prank] + lwb;
        bitq_pop(&q);
        if (with_attr) attr_merge(t->grp, AREC(t, par), AREC(t, cur));
        t->parent[cur] = par;
        cur = par; crank = prank;
        x = cur % W; y = (cur % WH) / W; z = cur / WH;
    }
    t->parent[cur] = BOTTOM;
    free(q.words); free(rinv); free(rank); free(vis);
    return 0;
}

static int64_t levelroot(orc_tree *t, int64_t x)
{
    if (x == BOTTOM) return BOTTOM;
    int64_t r = x;
    float gv = t->g[x];
    while (t->parent[r] != BOTTOM && t->g[t->parent[r]] == gv) r = t->parent[r];
    while (x != r) { int64_t y = t->parent[x]; t->parent[x] = r; x = y; }
    return r;
}

static int64_t get_parent(orc_tree *t, int64_t x) { return levelroot(t, t->parent[x]); }

static int is_levelroot(const orc_tree *t, int64_t x)
{
    return t->parent[x] == BOTTOM || t->g[x] != t->g[t->parent[x]];
}

static void compress_range(orc_tree *t, int64_t lwb, int64_t upb)
{
    for (int64_t x = lwb; x < upb; x++) {
        int64_t r = x;
        float gv = t->g[x];
        while (t->parent[r] != BOTTOM && t->g[t->parent[r]] == gv) r = t->parent[r];
        int64_t i = x;
        while (i != r) { int64_t y = t->parent[i]; t->parent[i] = r; i = y; }
        t->parent[r] = levelroot(t, t->parent[r]);
    }
}

/* Wilkinson merge of the ancestor chains of x and y (merge_nodes, flood.c:292-353) */
static void merge_nodes(orc_tree *t, int64_t x, int64_t y, int with_attr)
{
    unsigned char cor[sizeof(a_inert_t)], copa[sizeof(a_inert_t)];
    int have_cor = 0, have_copa = 0;
    if (t->g[y] > t->g[x]) { int64_t h = x; x = y; y = h; }
    x = levelroot(t, x);
    y = levelroot(t, y);
    while (x != y && y != BOTTOM) {
        int64_t z = get_parent(t, x);
        if (z != BOTTOM && t->g[z] >= t->g[y]) {
            if (have_cor && with_attr) attr_merge(t->grp, AREC(t, x), cor);
            x = z;
        } else {
            if (with_attr) {
                if (have_cor) {
                    memcpy(copa, AREC(t, x), t->asz);
                    attr_merge(t->grp, copa, cor);
                } else {
                    memcpy(copa, AREC(t, x), t->asz);
                }
                have_copa = 1;
                memcpy(cor, AREC(t, x), t->asz);
                have_cor = 1;
                memcpy(AREC(t, x), copa, t->asz);
            }
            t->parent[x] = y;
            x = y;
            y = z;
        }
    }
    if (y == BOTTOM && have_cor && with_attr) {
        while (x != BOTTOM) {
            attr_merge(t->grp, AREC(t, x), cor);
            x = get_parent(t, x);
        }
    }
    (void)have_copa;
}

static int64_t slab_bound(const orc_tree *t, int id, int nt)
{
    if (t->dims[2] == 1) return t->dims[0] * ((id * t->dims[1]) / nt);
    return t->dims[0] * t->dims[1] * ((id * t->dims[2]) / nt);
}

/* flood the two-row (or two-plane) band across a slab border and merge (fuse_sections) */
static int fuse_band(orc_tree *t, int64_t mdb, int with_attr)
{
    const int64_t inc = t->dims[2] == 1 ? t->dims[0] : t->dims[0] * t->dims[1];
    const int64_t lwb = mdb - inc, upb = mdb + inc;
    /* flood the band, merging each popped (child, parent) pair into the slab trees */
    const int64_t W = t->dims[0], H = t->dims[1], D = t->dims[2], WH = W * H;
    const int64_t m = upb - lwb;
    uint32_t *rinv = ranks_inverse(t->g, lwb, upb);
    if (!rinv) return -1;
    uint32_t *rank = (uint32_t *)malloc((size_t)m * 4);
    uint64_t *vis = (uint64_t *)calloc((size_t)((m + 63) / 64), 8);
    bitq_t q;
    if (!rank || !vis || bitq_init(&q, m) != 0) { free(rinv); free(rank); free(vis); return -1; }
    for (int64_t r = 0; r < m; r++) rank[rinv[r]] = (uint32_t)r;
    int dx[26], dy[26], dz[26];
    int nn = nb_offsets(t->dims, t->conn, dx, dy, dz);
    int64_t cur = lwb, crank = rank[0];
    vis[0] |= 1ull;
    int64_t x = cur % W, y = (cur % WH) / W, z = cur / WH;
    for (;;) {
        int climbed;
        do {
            climbed = 0;
            for (int k = 0; k < nn; k++) {
                int64_t nx = x + dx[k], ny = y + dy[k], nz = z + dz[k];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) continue;
                int64_t nbi = nx + ny * W + nz * WH;
                if (nbi < lwb || nbi >= upb) continue;
                int64_t li = nbi - lwb;
                if (vis[li >> 6] & (1ull << (li & 63))) continue;
                vis[li >> 6] |= 1ull << (li & 63);
                int64_t nr = rank[li];
                if (crank > nr) { bitq_insert(&q, nr); continue; }
                bitq_insert(&q, crank);
                cur = nbi; crank = nr; x = nx; y = ny; z = nz;
                climbed = 1;
                break;
            }
        } while (climbed);
        if (q.top < 0) break;
        int64_t prank = q.top;
        int64_t par = (int64_t)rinv[prank] + lwb;
        bitq_pop(&q);
        merge_nodes(t, cur, par, with_attr);
        cur = par; crank = prank;
        x = cur % W; y = (cur % WH) / W; z = cur / WH;
    }
    free(q.words); free(rinv); free(rank); free(vis);
    return 0;
}

static void map_connectivity(const int64_t dims[3], int *conn)
{
    if (dims[2] > 1 && *conn == 4) *conn = 6;
    else if (*conn == 8 && dims[2] > 1) *conn = 26;
    else if (*conn == 26 && dims[2] == 1) *conn = 8;
}

static void init_attributes(orc_tree *t)
{
    const int64_t W = t->dims[0], WH = t->dims[0] * t->dims[1];
#pragma omp parallel for schedule(static) num_threads(t->nthreads)
    for (int64_t i = 0; i < t->n; i++)
        attr_init(t->grp, AREC(t, i), 1, (uint64_t)(i % W), (uint64_t)((i % WH) / W), (uint64_t)(i / WH), t->g[i]);
}

/*
 * After a slab fuse the representative of a node need not be its minimum index (merge_nodes
 * links equal-grey levelroots in chain order).  Move every node to its minimum-index pixel so
 * the parent array is the canonical one a single-slab flood gives; the attribute record moves
 * with the representative.
 */
static int recanonicalize(orc_tree *t)
{
    int64_t *rep = (int64_t *)malloc((size_t)t->n * 8);
    int64_t *mn = (int64_t *)malloc((size_t)t->n * 8);
    if (!rep || !mn) { free(rep); free(mn); return -1; }
    for (int64_t i = 0; i < t->n; i++) mn[i] = INT64_MAX;
    for (int64_t i = 0; i < t->n; i++) {
        rep[i] = levelroot(t, i);
        if (i < mn[rep[i]]) mn[rep[i]] = i;
    }
    for (int64_t i = 0; i < t->n; i++)
        if (rep[i] == i && mn[i] != i) memcpy(AREC(t, mn[i]), AREC(t, i), t->asz);
    int64_t *np = (int64_t *)malloc((size_t)t->n * 8);
    if (!np) { free(rep); free(mn); return -1; }
    for (int64_t i = 0; i < t->n; i++) {
        int64_t r = rep[i], m = mn[r];
        if (i != m) { np[i] = m; continue; }
        int64_t pr = t->parent[r];
        np[i] = pr == BOTTOM ? BOTTOM : mn[rep[pr]];
    }
    memcpy(t->parent, np, (size_t)t->n * 8);
    free(np); free(rep); free(mn);
    return 0;
}

/* ----------------------------------------------------------------------------------------- */
/* public API                                                                                */
/* ----------------------------------------------------------------------------------------- */
ORC_API void orc_free(orc_tree *t)
{
    if (!t) return;
    free(t->g); free(t->parent); free(t->attr); free(t);
}

ORC_API orc_tree *orc_build(const float *img, const int64_t dims[3], int conn, int tree_min, int attr_idx,
                            const float *pixdim, int nthreads)
{
    if (!img || !dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return NULL;
    if (attr_idx < 0 || attr_idx >= 18) return NULL;
    if (conn != 4 && conn != 6 && conn != 8 && conn != 26) return NULL;
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    if (!t) return NULL;
    memcpy(t->dims, dims, sizeof t->dims);
    t->n = dims[0] * dims[1] * dims[2];
    t->conn = conn;
    map_connectivity(t->dims, &t->conn);
    t->tree_min = tree_min;
    t->pixdim[0] = pixdim ? pixdim[0] : 1.f;
    t->pixdim[1] = pixdim ? pixdim[1] : 1.f;
    t->pixdim[2] = pixdim ? pixdim[2] : 1.f;
    {
        /* every slab needs at least two rows (planes) for the fuse band */
        int64_t slabs = (t->dims[2] == 1 ? t->dims[1] : t->dims[2]) / 2;
        int64_t nt = nthreads > 0 ? nthreads : 1;
        if (nt > slabs) nt = slabs > 0 ? slabs : 1;
        t->nthreads = (int)nt;
    }
    t->grp = group_of(attr_idx);
    t->asz = group_size(t->grp);
    t->g = (float *)malloc((size_t)t->n * sizeof(float));
    t->parent = (int64_t *)malloc((size_t)t->n * sizeof(int64_t));
    t->attr = (unsigned char *)malloc((size_t)t->n * t->asz);
    if (!t->g || !t->parent || !t->attr) { orc_free(t); return NULL; }
    memcpy(t->g, img, (size_t)t->n * sizeof(float));
    /* find_extremas + min-tree inversion (preprocessing.c:7-30, 84-91) */
    float mn = t->g[0], mx = t->g[0];
    for (int64_t i = 0; i < t->n; i++) { if (t->g[i] < mn) mn = t->g[i]; if (t->g[i] > mx) mx = t->g[i]; }
    t->gmin = mn; t->gmax = mx;
    if (tree_min)
        for (int64_t i = 0; i < t->n; i++) t->g[i] = mx - t->g[i];
    init_attributes(t);

    const int nt = t->nthreads;
    int err = 0;
    /* one slab per thread (flood.c:79-141), then the binary-tree fuse schedule (flood.c:144-165) */
#pragma omp parallel num_threads(nt) reduction(| : err)
    {
#ifdef _OPENMP
        int id = omp_get_thread_num();
#else
        int id = 0;
#endif
        int64_t lwb = slab_bound(t, id, nt), upb = slab_bound(t, id + 1, nt);
        if (upb > lwb) {
            if (flood_slab(t, lwb, upb, 1) != 0) err |= 1;
            compress_range(t, lwb, upb);
        }
    }
    if (err) { orc_free(t); return NULL; }
    for (int step = 1; step < nt; step *= 2) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(| : err)
        for (int id = 0; id < nt; id += 2 * step) {
            if (id + step < nt && fuse_band(t, slab_bound(t, id + step, nt), 1) != 0) err |= 1;
        }
    }
    if (err) { orc_free(t); return NULL; }
    if (nt > 1 && recanonicalize(t) != 0) { orc_free(t); return NULL; }
    return t;
}

/* flood_attributes (flood.c:184-219): re-accumulate a (possibly new) attribute group */
ORC_API int orc_reaccumulate(orc_tree *t, int attr_idx)
{
    if (!t || attr_idx < 0 || attr_idx >= 18) return -1;
    int grp = group_of(attr_idx);
    if (grp == t->grp) return 0;
    free(t->attr);
    t->grp = grp;
    t->asz = group_size(grp);
    t->attr = (unsigned char *)malloc((size_t)t->n * t->asz);
    uint32_t *ranks = ranks_inverse(t->g, 0, t->n);
    unsigned char *visited = (unsigned char *)calloc((size_t)t->n, 1);
    if (!t->attr || !ranks || !visited) { free(ranks); free(visited); return -1; }
    init_attributes(t);
    for (int64_t i = t->n - 1; i >= 0; i--) {
        int64_t p = ranks[i];
        visited[p] = 1;
        int64_t par = get_parent(t, p);
        if (par == BOTTOM) continue;
        attr_merge(t->grp, AREC(t, par), AREC(t, p));
        if (t->g[p] == t->g[par] && visited[par]) {
            par = get_parent(t, par);
            if (par != BOTTOM) attr_merge(t->grp, AREC(t, par), AREC(t, p));
        }
    }
    free(visited); free(ranks);
    return 0;
}

ORC_API int64_t orc_size(const orc_tree *t) { return t ? t->n : 0; }
ORC_API int orc_connectivity(const orc_tree *t) { return t ? t->conn : 0; }
ORC_API float orc_gmax(const orc_tree *t) { return t ? t->gmax : 0.f; }
ORC_API int orc_group(const orc_tree *t) { return t ? t->grp : -1; }

ORC_API void orc_parent(const orc_tree *t, int64_t *out) { memcpy(out, t->parent, (size_t)t->n * 8); }
ORC_API void orc_grey(const orc_tree *t, float *out) { memcpy(out, t->g, (size_t)t->n * 4); }

/* attribute of every pixel's record (meaningful at levelroots) */
ORC_API int orc_attr_all(const orc_tree *t, int attr_idx, double *out, double *area_out)
{
    if (!t || group_of(attr_idx) != t->grp) return -1;
    for (int64_t i = 0; i < t->n; i++) {
        if (out) out[i] = attr_value_rec(t, attr_idx, AREC(t, i));
        if (area_out) area_out[i] = attr_area_rec(t, attr_idx, AREC(t, i));
    }
    return 0;
}

/* tree_attribute_check (filtering.c:47-52) */
ORC_API int orc_check(orc_tree *t, int attr_idx, float *out)
{
    if (!t || group_of(attr_idx) != t->grp) return -1;
    for (int64_t v = 0; v < t->n; v++) out[v] = (float)attr_value_rec(t, attr_idx, AREC(t, levelroot(t, v)));
    return 0;
}

#define ATTR(w) attr_value_rec(t, attr_idx, AREC(t, (w)))

static void filter_direct(orc_tree *t, float *out, unsigned char *reached, int attr_idx, double lambda)
{
    for (int64_t v = 0; v < t->n; v++) {
        if (reached[v]) continue;
        int64_t w = v, par = t->parent[w];
        while (par != BOTTOM && !reached[w] && (t->g[w] == t->g[par] || ATTR(w) < lambda)) { w = par; par = t->parent[w]; }
        float val;
        if (reached[w]) val = out[w];
        else if (ATTR(w) >= lambda) val = t->g[w];
        else val = 0;
        int64_t u = v;
        while (u != w) { out[u] = val; reached[u] = 1; u = t->parent[u]; }
        out[w] = val; reached[w] = 1;
    }
}

static void filter_min(orc_tree *t, float *out, unsigned char *reached, int attr_idx, double lambda)
{
    for (int64_t v = 0; v < t->n; v++) {
        if (reached[v]) continue;
        int64_t w = levelroot(t, v), r = w, par = get_parent(t, w);
        while (par != BOTTOM && !reached[w]) {
            if (ATTR(w) < lambda) r = w;
            w = par;
            par = get_parent(t, w);
        }
        float val;
        if (reached[w] && out[w] != t->g[w]) { val = out[w]; r = w; }
        else if (ATTR(r) >= lambda) val = t->g[r];
        else if (ATTR(r) < lambda && t->parent[r] != BOTTOM) val = t->g[t->parent[r]];
        else val = 0;
        int64_t u = v;
        while (u != r) { reached[u] = 1; out[u] = val; u = t->parent[u]; }
        reached[r] = 1; out[r] = val;
    }
}

static void filter_max(orc_tree *t, float *out, unsigned char *reached, int attr_idx, double lambda)
{
    for (int64_t v = 0; v < t->n; v++) {
        if (reached[v]) continue;
        int64_t w = v, par = t->parent[w];
        while (par != BOTTOM && !reached[w] && (t->g[w] == t->g[par] || ATTR(w) < lambda)) { w = par; par = t->parent[w]; }
        float val;
        if (reached[w]) val = out[w];
        else if (ATTR(w) >= lambda) {
            val = t->g[w];
            int64_t u = par;
            while (u != BOTTOM && out[u] != t->g[u]) { out[u] = t->g[u]; reached[u] = 1; u = t->parent[u]; }
        } else val = 0;
        int64_t u = v;
        while (u != w) { out[u] = val; reached[u] = 1; u = t->parent[u]; }
        out[w] = val; reached[w] = 1;
    }
    /* serial post-pass of tree_filtering (filtering.c:33-43) */
    for (int64_t i = 0; i < t->n; i++) {
        if (reached[i]) continue;
        int64_t im = i;
        while (im != BOTTOM && out[im] != t->g[im]) { out[im] = t->g[im]; im = t->parent[im]; }
    }
}

static int filter_subtractive(orc_tree *t, float *out, unsigned char *reached, int attr_idx, double lambda)
{
    int64_t *levpath = (int64_t *)malloc((size_t)(t->n + 2) * 8);
    float *filt = (float *)malloc((size_t)(t->n + 2) * 4);
    if (!levpath || !filt) { free(levpath); free(filt); return -1; }
    for (int64_t v = 0; v < t->n; v++) {
        if (reached[v]) continue;
        int64_t w = levelroot(t, v), i = 0;
        while (w != BOTTOM && !reached[w]) { levpath[i++] = w; w = get_parent(t, w); }
        levpath[i] = w;
        if (w != BOTTOM) filt[i] = out[w];
        for (int64_t j = i - 1; j >= 0; j--) {
            int64_t u = levpath[j];
            if (w != BOTTOM) {
                if (ATTR(u) < lambda) filt[j] = filt[j + 1];
                else filt[j] = t->g[u] + filt[j + 1] - t->g[w];
            } else if (ATTR(u) >= lambda) filt[j] = t->g[u];
            else filt[j] = 0;
            w = u;
        }
        int64_t u = v;
        for (int64_t j = 0; j <= i; j++) {
            while (u != levpath[j]) { out[u] = filt[j]; reached[u] = 1; u = t->parent[u]; }
            if (u != BOTTOM) { out[u] = filt[j]; reached[u] = 1; }
            if (u != BOTTOM) u = t->parent[u];
        }
    }
    free(levpath); free(filt);
    return 0;
}

/* tree_filtering + perform_filter_or_extract (operation.c:10-29); out is zero-initialised */
ORC_API int orc_filter(orc_tree *t, int attr_idx, int decision, double lambda, int extract, float *out)
{
    if (!t || group_of(attr_idx) != t->grp || decision < 0 || decision > 3) return -1;
    unsigned char *reached = (unsigned char *)calloc((size_t)t->n, 1);
    if (!reached) return -1;
    memset(out, 0, (size_t)t->n * sizeof(float));
    int rc = 0;
    switch (decision) {
    case 0: filter_direct(t, out, reached, attr_idx, lambda); break;
    case 1: filter_min(t, out, reached, attr_idx, lambda); break;
    case 2: filter_max(t, out, reached, attr_idx, lambda); break;
    default: rc = filter_subtractive(t, out, reached, attr_idx, lambda); break;
    }
    free(reached);
    if (rc) return rc;
    if (extract)
        for (int64_t i = 0; i < t->n; i++) out[i] = t->g[i] - out[i];
    if (t->tree_min)
        for (int64_t i = 0; i < t->n; i++) out[i] = t->gmax - out[i];
    return 0;
}

/* find_scale (pattern.c:7-24) */
static int find_scale(const float *lv, int n, double a)
{
    int upper = n - 1, lower = 0, mid;
    if (a >= (double)lv[upper]) return upper;
    mid = (upper + lower) / 2;
    while (mid != lower) {
        if (a >= (double)lv[mid]) lower = mid;
        else upper = mid;
        mid = (upper + lower) / 2;
    }
    return lower;
}

ORC_API int orc_find_scale(const float *lv, int n, double a) { return find_scale(lv, n, a); }

/* tree_pattern_spectrum single-process branch (pattern.c:80-105), background = 1 */
ORC_API int orc_spectrum1d(const orc_tree *t, int attr_idx, const float *lv, int n, double *spec)
{
    if (!t || n <= 0 || group_of(attr_idx) != t->grp) return -1;
    memset(spec, 0, (size_t)n * sizeof(double));
    for (int64_t v = 0; v < t->n; v++) {
        if (!is_levelroot(t, v)) continue;
        float dh = t->parent[v] != BOTTOM ? (t->g[v] - t->g[t->parent[v]]) : t->g[v];
        if (dh == 0) continue;
        int s = find_scale(lv, n, attr_value_rec(t, attr_idx, AREC(t, v)));
        if (s < n) spec[s] += dh * attr_area_rec(t, attr_idx, AREC(t, v));
    }
    return 0;
}

/*
 * tree_pattern_spectrum2d single-process branch (pattern.c:170-191) with the wiring of
 * perform_pattern_spectrum (operation.c:93-100): attr2 == 0 with attr1 != 0 means the
 * second axis is the area function of attr1's vtable entry.
 */
ORC_API int orc_spectrum2d(const orc_tree *t, int attr1, const float *l1, int n1, int attr2, const float *l2, int n2,
                           double *spec)
{
    if (!t || n1 <= 0 || n2 <= 0 || group_of(attr1) != t->grp) return -1;
    if (attr1 == 0 && attr2 != 0) return -2; /* operation.c:93-95 refuses this order */
    int second_is_area = (attr2 == 0 && attr1 != 0);
    if (!second_is_area && group_of(attr2) != t->grp) return -1;
    memset(spec, 0, (size_t)n1 * n2 * sizeof(double));
    for (int64_t v = 0; v < t->n; v++) {
        if (!is_levelroot(t, v)) continue;
        float dh = t->parent[v] != BOTTOM ? (t->g[v] - t->g[t->parent[v]]) : t->g[v];
        if (dh == 0) continue;
        const void *a = AREC(t, v);
        int s1 = find_scale(l1, n1, attr_value_rec(t, attr1, a));
        double a2 = second_is_area ? attr_area_rec(t, attr1, a) : attr_value_rec(t, attr2, a);
        int s2 = find_scale(l2, n2, a2);
        spec[s1 + s2 * n1] += dh * attr_area_rec(t, attr1, a);
    }
    return 0;
}

/*
 * lambda_vector_read (lambdavec.c:35-84): the header is matched the way the reference's
 * fscanf literal matches it, lambda_0 = 0 is prepended, every value is divided by imscale.
 * Returns the number of lambdas written (count + 1), or -1 when the file cannot be opened.
 */
ORC_API int orc_lvec_read(const char *path, double imscale, float *out, int cap)
{
    static const char *hdr = "#LambdaVector file. Keep this exact line in all similar lambda vector files\n";
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    /* literal match: whitespace in the pattern skips any run of whitespace in the input */
    const char *p = hdr;
    int c = fgetc(f);
    while (*p) {
        if (*p == ' ' || *p == '\n') {
            while (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f') c = fgetc(f);
            p++;
            continue;
        }
        if (c != (unsigned char)*p) break;
        c = fgetc(f);
        p++;
    }
    if (c != EOF) ungetc(c, f);
    long pos = ftell(f);
    float v;
    int size = 0;
    while (fscanf(f, "%f", &v) == 1) size++;
    if (out && cap > 0) out[0] = 0;
    fseek(f, pos, SEEK_SET);
    for (int i = 0; i < size; i++) {
        if (fscanf(f, "%f", &v) != 1) break;
        v /= imscale;
        if (out && i + 1 < cap) out[i + 1] = v;
    }
    fclose(f);
    return size + 1;
}

/* depth statistics of the canonical tree (levelroot nodes only) */
ORC_API int orc_tree_stats(orc_tree *t, int64_t *n_nodes, int64_t *max_depth, int64_t *n_leaves)
{
    int64_t *depth = (int64_t *)malloc((size_t)t->n * 8);
    unsigned char *haschild = (unsigned char *)calloc((size_t)t->n, 1);
    if (!depth || !haschild) { free(depth); free(haschild); return -1; }
    uint32_t *ranks = ranks_inverse(t->g, 0, t->n); /* ascending (g, index): parents first */
    if (!ranks) { free(depth); free(haschild); return -1; }
    int64_t K = 0, D = 0, L = 0;
    for (int64_t i = 0; i < t->n; i++) {
        int64_t p = ranks[i];
        if (!is_levelroot(t, p)) continue;
        K++;
        int64_t par = t->parent[p];
        depth[p] = par == BOTTOM ? 1 : depth[par] + 1;
        if (depth[p] > D) D = depth[p];
        if (par != BOTTOM) haschild[par] = 1;
    }
    for (int64_t i = 0; i < t->n; i++) if (is_levelroot(t, i) && !haschild[i]) L++;
    free(ranks); free(depth); free(haschild);
    *n_nodes = K; *max_depth = D; *n_leaves = L;
    return 0;
}
