/*
 * dcf_build.hip -- component-tree construction and attribute accumulation on gfx950.
 *
 * Replaces build_local_tree_generic / tree_flood / fuse_sections (src/sources/flood.c:58-353)
 * and flood_attributes (flood.c:184-219).  The reference floods one slab per OpenMP thread with
 * a hierarchical bit queue and fuses slabs pairwise; here the image is cut into tiles that each
 * map to one workgroup:
 *
 *   k_tile_build   one tile in LDS: the intra-tile edges are merged with the lock-free zip of
 *                  dcf_zip.h (ds_cmpst_b64 on LDS link words) in a hierarchical seam order, then
 *                  the tile is canonicalised and its link words are written to HBM once.
 *   k_seam         the edges that cross tile borders are zipped in HBM (64-bit CAS at agent
 *                  scope), one thread per tile face walking its seam in order: the on-GPU
 *                  counterpart of the reference's boundary-tree merge (src/sources/boundary.c:427-583).
 *   k_canon        levelroot of every pixel with path halving, canonical parent array and
 *                  levelroot flags (the compression loop of flood.c:127-141).
 *   scan           node numbering (hipcub exclusive sum of the flags)
 *   k_nodes        node parent / grey, pixel -> node map
 *   k_own_*        per-node own sums (run-aggregated atomics)
 *   k_leaf_*       subtree sums by pointer doubling: after round r every node holds the sum over
 *                  descendants at distance < 2^r, so ceil(log2(depth+1)) rounds finish the
 *                  leaffix with no per-level launches (trees of u16 images are ~50k deep).
 */
#include <hipcub/hipcub.hpp>

#include "dcf_internal.h"
#include "dcf_zip.h"

/* ------------------------------------------------------------------------------------------------ */
/* pixel access                                                                                     */
/* ------------------------------------------------------------------------------------------------ */
__device__ __forceinline__ uint32_t dcf_fkey(float v)
{
    uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? 0xFFFFFFFFu - u : (u | 0x80000000u);
}

/* grey key of pixel p (the order the tree is built on) */
__device__ __forceinline__ uint32_t px_key(const ImgView &iv, int64_t p)
{
    if (iv.dtype == DCF_U8) {
        uint32_t v = ((const uint8_t *)iv.img)[p];
        return iv.invert ? (uint32_t)(*iv.gmax) - v : v;
    }
    if (iv.dtype == DCF_U16) {
        uint32_t v = ((const uint16_t *)iv.img)[p];
        return iv.invert ? (uint32_t)(*iv.gmax) - v : v;
    }
    float v = ((const float *)iv.img)[p];
    return dcf_fkey(iv.invert ? *iv.gmax - v : v);
}

/* grey value of pixel p as the reference's float `value` */
__device__ __forceinline__ float px_grey(const ImgView &iv, int64_t p)
{
    float v;
    if (iv.dtype == DCF_U8) v = (float)((const uint8_t *)iv.img)[p];
    else if (iv.dtype == DCF_U16) v = (float)((const uint16_t *)iv.img)[p];
    else v = ((const float *)iv.img)[p];
    return iv.invert ? *iv.gmax - v : v;
}

/* ------------------------------------------------------------------------------------------------ */
/* memory policies for the zip                                                                      */
/* ------------------------------------------------------------------------------------------------ */
struct GlobalMem {
    uint64_t *w;
    __device__ uint64_t load(uint32_t i) { return __hip_atomic_load(&w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ void store(uint32_t i, uint64_t v) { __hip_atomic_store(&w[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ uint64_t cas(uint32_t i, uint64_t e, uint64_t d)
    {
        uint64_t exp = e;
        __hip_atomic_compare_exchange_strong(&w[i], &exp, d, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return exp;
    }
};

struct LdsMem {
    uint64_t *w;
    __device__ uint64_t load(uint32_t i) { return __hip_atomic_load(&w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
    __device__ void store(uint32_t i, uint64_t v) { __hip_atomic_store(&w[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
    __device__ uint64_t cas(uint32_t i, uint64_t e, uint64_t d)
    {
        uint64_t exp = e;
        __hip_atomic_compare_exchange_strong(&w[i], &exp, d, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return exp;
    }
};

/* last ancestor of x with key kx; non-levelroot links on the way are halved (all writes point to
 * a same-key ancestor, so concurrent finders never break a chain) */
template <class Mem>
__device__ __forceinline__ uint32_t find_levelroot(Mem &m, uint32_t x, uint32_t kx)
{
    uint64_t w = m.load(x);
    for (;;) {
        uint32_t y = dcf_link_idx(w);
        if (y == DCF_NONE || dcf_link_key(w) != kx) return x;
        uint64_t wy = m.load(y);
        uint32_t gp = dcf_link_idx(wy);
        if (gp == DCF_NONE || dcf_link_key(wy) != kx) return y;
        m.store(x, wy);
        x = gp;
        w = m.load(x);
    }
}

/* forward neighbour offsets: one direction of every edge of the connectivity */
struct Fwd {
    int n;
    int8_t dx[13], dy[13], dz[13];
};

static Fwd make_fwd(int conn, int is3d)
{
    Fwd f;
    f.n = 0;
    auto add = [&](int x, int y, int z) { f.dx[f.n] = (int8_t)x; f.dy[f.n] = (int8_t)y; f.dz[f.n] = (int8_t)z; f.n++; };
    if (!is3d) {
        add(1, 0, 0); add(0, 1, 0);
        if (conn == 8) { add(1, 1, 0); add(-1, 1, 0); }
    } else if (conn == 6) {
        add(1, 0, 0); add(0, 1, 0); add(0, 0, 1);
    } else {
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    if (z > 0 || (z == 0 && y > 0) || (z == 0 && y == 0 && x > 0)) add(x, y, z);
                }
    }
    return f;
}

/* ------------------------------------------------------------------------------------------------ */
/* tile build                                                                                       */
/* ------------------------------------------------------------------------------------------------ */
#define TILE_MAX 4096
#define TILE_THREADS 256

struct TileGeom {
    int W, H, D;
    int TX, TY, TZ;
    int ntx, nty, ntz;
};

/*
 * Merge order inside a tile.  Zipping all edges at once is correct but makes every edge of a
 * seam walk the full chain of its two blocks (O(edges x depth) work on smooth 16-bit tiles, whose
 * trees are single chains thousands of nodes deep).  Instead the tile is merged hierarchically,
 * one thread per seam walking its edges in seam order: the first zip of a seam interleaves the
 * two chains, every later zip starts next to the previous one and meets the merged chain within
 * a few nodes.  Rows first (x edges), then row blocks of 1, 2, 4 ... rows (y seams), then plane
 * blocks (z seams); each level ends with a barrier.
 */
__global__ __launch_bounds__(TILE_THREADS) void k_tile_build(ImgView iv, uint64_t *__restrict__ link, TileGeom g, Fwd fwd)
{
    __shared__ uint64_t slink[TILE_MAX];
    __shared__ uint32_t skey[TILE_MAX];
    __shared__ uint16_t slr[TILE_MAX];

    const int b = blockIdx.x;
    const int tx = b % g.ntx, ty = (b / g.ntx) % g.nty, tz = b / (g.ntx * g.nty);
    const int x0 = tx * g.TX, y0 = ty * g.TY, z0 = tz * g.TZ;
    const int ex = min(g.TX, g.W - x0), ey = min(g.TY, g.H - y0), ez = min(g.TZ, g.D - z0);
    const int np = ex * ey * ez;
    const int exy = ex * ey;
    const int64_t WH = (int64_t)g.W * g.H;
    const int64_t base = (int64_t)x0 + (int64_t)y0 * g.W + (int64_t)z0 * WH;

    for (int l = threadIdx.x; l < np; l += TILE_THREADS) {
        int lx = l % ex, ly = (l / ex) % ey, lz = l / exy;
        int64_t p = base + lx + (int64_t)ly * g.W + (int64_t)lz * WH;
        skey[l] = px_key(iv, p);
        slink[l] = dcf_make_link(0, DCF_NONE);
    }
    __syncthreads();
    LdsMem m{ slink };

    /* rows: one thread per row, sequential x edges */
    for (int r = threadIdx.x; r < ey * ez; r += TILE_THREADS) {
        int l0 = r * ex;
        for (int x = 1; x < ex; x++) dcf_zip(m, (uint32_t)(l0 + x - 1), skey[l0 + x - 1], (uint32_t)(l0 + x), skey[l0 + x]);
    }
    __syncthreads();

    /* y seams inside every plane: level lv joins row blocks [a, a+2^lv) and [a+2^lv, a+2^(lv+1)) */
    for (int lv = 0; (1 << lv) < ey; lv++) {
        const int step = 2 << lv, half = 1 << lv;
        const int per_plane = (ey - half + step - 1) / step;
        for (int s = threadIdx.x; s < per_plane * ez; s += TILE_THREADS) {
            int lz = s / per_plane, ys = (s % per_plane) * step + half; /* first row of the upper block */
            int lrow = lz * exy + (ys - 1) * ex;
            for (int x = 0; x < ex; x++) {
                int l = lrow + x;
                for (int k = 0; k < fwd.n; k++) {
                    if (fwd.dz[k] != 0 || fwd.dy[k] != 1) continue;
                    int nx = x + fwd.dx[k];
                    if (nx < 0 || nx >= ex) continue;
                    int nl = l + ex + fwd.dx[k];
                    dcf_zip(m, (uint32_t)l, skey[l], (uint32_t)nl, skey[nl]);
                }
            }
        }
        __syncthreads();
    }

    /* z seams: level lv joins plane blocks; one thread walks the whole face */
    for (int lv = 0; (1 << lv) < ez; lv++) {
        const int step = 2 << lv, half = 1 << lv;
        const int nseam = (ez - half + step - 1) / step;
        for (int s = threadIdx.x; s < nseam; s += TILE_THREADS) {
            int zs = s * step + half;
            int lpl = (zs - 1) * exy;
            for (int y = 0; y < ey; y++)
                for (int x = 0; x < ex; x++) {
                    int l = lpl + y * ex + x;
                    for (int k = 0; k < fwd.n; k++) {
                        if (fwd.dz[k] != 1) continue;
                        int nx = x + fwd.dx[k], ny = y + fwd.dy[k];
                        if (nx < 0 || nx >= ex || ny < 0 || ny >= ey) continue;
                        int nl = l + exy + fwd.dy[k] * ex + fwd.dx[k];
                        dcf_zip(m, (uint32_t)l, skey[l], (uint32_t)nl, skey[nl]);
                    }
                }
        }
        __syncthreads();
    }

    for (int l = threadIdx.x; l < np; l += TILE_THREADS) slr[l] = (uint16_t)find_levelroot(m, (uint32_t)l, skey[l]);
    __syncthreads();

    for (int l = threadIdx.x; l < np; l += TILE_THREADS) {
        int lx = l % ex, ly = (l / ex) % ey, lz = l / exy;
        int64_t p = base + lx + (int64_t)ly * g.W + (int64_t)lz * WH;
        uint64_t out;
        int r = slr[l];
        if (r != l) {
            int rx = r % ex, ry = (r / ex) % ey, rz = r / exy;
            out = dcf_make_link(skey[l], (uint32_t)(base + rx + (int64_t)ry * g.W + (int64_t)rz * WH));
        } else {
            uint64_t w = slink[l];
            uint32_t z = dcf_link_idx(w);
            if (z == DCF_NONE) {
                out = dcf_make_link(0, DCF_NONE);
            } else {
                int q = slr[z];
                int qx = q % ex, qy = (q / ex) % ey, qz = q / exy;
                out = dcf_make_link(dcf_link_key(w), (uint32_t)(base + qx + (int64_t)qy * g.W + (int64_t)qz * WH));
            }
        }
        link[p] = out;
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* seam merge: the edges that cross tile borders                                                   */
/* ------------------------------------------------------------------------------------------------ */
/*
 * One thread per (tile, face) walks the face in order and zips every edge leaving the tile through
 * it, so the edges of one seam are zipped one after the other (same argument as inside a tile).
 * An edge belongs to the face of the first axis on which it changes tile (x, then y, then z), so
 * every cross-tile edge is zipped exactly once.
 */
enum { FACE_X_LAST = 0, FACE_X_FIRST, FACE_Y_LAST, FACE_Y_FIRST, FACE_Z_LAST, NFACES };

__device__ __forceinline__ int edge_face(int dx, int dy, int dz, bool cx, bool cy, bool cz)
{
    if (cx) return dx > 0 ? FACE_X_LAST : FACE_X_FIRST;
    if (cy) return dy > 0 ? FACE_Y_LAST : FACE_Y_FIRST;
    if (cz) return FACE_Z_LAST;
    return -1;
}

__global__ __launch_bounds__(64) void k_seam(ImgView iv, uint64_t *__restrict__ link, TileGeom g, Fwd fwd, int64_t ntiles)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles * NFACES) return;
    const int face = (int)(t % NFACES);
    const int64_t tile = t / NFACES;
    const int tx = (int)(tile % g.ntx), ty = (int)((tile / g.ntx) % g.nty), tz = (int)(tile / ((int64_t)g.ntx * g.nty));
    const int x0 = tx * g.TX, y0 = ty * g.TY, z0 = tz * g.TZ;
    const int ex = min(g.TX, g.W - x0), ey = min(g.TY, g.H - y0), ez = min(g.TZ, g.D - z0);
    const int64_t WH = (int64_t)g.W * g.H;
    /* the face as a 2D range (u, v) -> (lx, ly, lz) */
    int nu, nv;
    if (face == FACE_X_LAST || face == FACE_X_FIRST) { nu = ey; nv = ez; }
    else if (face == FACE_Y_LAST || face == FACE_Y_FIRST) { nu = ex; nv = ez; }
    else { nu = ex; nv = ey; }
    GlobalMem m{ link };
    for (int v = 0; v < nv; v++)
        for (int u = 0; u < nu; u++) {
            int lx, ly, lz;
            if (face == FACE_X_LAST) { lx = ex - 1; ly = u; lz = v; }
            else if (face == FACE_X_FIRST) { lx = 0; ly = u; lz = v; }
            else if (face == FACE_Y_LAST) { lx = u; ly = ey - 1; lz = v; }
            else if (face == FACE_Y_FIRST) { lx = u; ly = 0; lz = v; }
            else { lx = u; ly = v; lz = ez - 1; }
            int x = x0 + lx, y = y0 + ly, z = z0 + lz;
            int64_t p = (int64_t)x + (int64_t)y * g.W + (int64_t)z * WH;
            uint32_t kp = 0;
            bool have = false;
            for (int k = 0; k < fwd.n; k++) {
                int nx = x + fwd.dx[k], ny = y + fwd.dy[k], nz = z + fwd.dz[k];
                if (nx < 0 || nx >= g.W || ny < 0 || ny >= g.H || nz < 0 || nz >= g.D) continue;
                bool cx = nx / g.TX != tx, cy = ny / g.TY != ty, cz = nz / g.TZ != tz;
                if (edge_face(fwd.dx[k], fwd.dy[k], fwd.dz[k], cx, cy, cz) != face) continue;
                int64_t q = (int64_t)nx + (int64_t)ny * g.W + (int64_t)nz * WH;
                if (!have) { kp = px_key(iv, p); have = true; }
                dcf_zip(m, (uint32_t)p, kp, (uint32_t)q, px_key(iv, q));
            }
        }
}

/* ------------------------------------------------------------------------------------------------ */
/* canonical parent array + levelroot flags                                                         */
/* ------------------------------------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void k_canon(ImgView iv, uint64_t *__restrict__ link, int32_t *__restrict__ parent,
                                               uint32_t *__restrict__ flag, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    GlobalMem m{ link };
    uint32_t kp = px_key(iv, p);
    uint32_t r = find_levelroot(m, (uint32_t)p, kp);
    if (r != (uint32_t)p) {
        parent[p] = (int32_t)r;
        flag[p] = 0;
        return;
    }
    uint64_t w = m.load((uint32_t)p);
    uint32_t z = dcf_link_idx(w);
    parent[p] = z == DCF_NONE ? -1 : (int32_t)find_levelroot(m, z, dcf_link_key(w));
    flag[p] = 1;
}

/* pixel -> node map (in place over the exclusive scan of the flags) and node parent / grey */
__global__ __launch_bounds__(256) void k_nodes(ImgView iv, const int32_t *__restrict__ parent, const uint32_t *__restrict__ flag,
                                               uint32_t *__restrict__ p2n, int32_t *__restrict__ npar, float *__restrict__ ngrey,
                                               int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int32_t par = parent[p];
    if (!flag[p]) {
        p2n[p] = p2n[par]; /* par is a levelroot: its entry is never rewritten */
        return;
    }
    uint32_t k = p2n[p];
    npar[k] = par < 0 ? -1 : (int32_t)p2n[par];
    ngrey[k] = px_grey(iv, p);
}

/* levelroot flags from a canonical parent array */
__global__ __launch_bounds__(256) void k_flags(ImgView iv, const int32_t *__restrict__ parent, uint32_t *__restrict__ flag, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int32_t r = parent[p];
    flag[p] = (r < 0 || px_key(iv, r) != px_key(iv, p)) ? 1u : 0u;
}

/* node-indexed attribute records from the pixel-indexed ones of the levelroots */
__global__ __launch_bounds__(256) void k_gather_attr(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ p2n,
                                                     DevAttr P, DevAttr A, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !flag[p]) return;
    uint32_t k = p2n[p];
    A.area[k] = P.area[p];
    for (int i = 0; i < A.nsum; i++) A.s[i][k] = P.s[i][p];
    if (A.has_ext)
        for (int i = 0; i < 3; i++) { A.mn[i][k] = P.mn[i][p]; A.mx[i][k] = P.mx[i][p]; }
}

/* ------------------------------------------------------------------------------------------------ */
/* attribute accumulation                                                                           */
/* ------------------------------------------------------------------------------------------------ */
#define RUN 8

static __global__ __launch_bounds__(256) void k_fill_u32(uint32_t *a, uint32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

static __global__ __launch_bounds__(256) void k_fill_i64(int64_t *a, int64_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* own pixel count (and own sums) per node; each thread folds a run of RUN pixels */
__global__ __launch_bounds__(256) void k_own(ImgView iv, const uint32_t *__restrict__ p2n, DevAttr A, int grp, int W, int H,
                                             int64_t n)
{
    int64_t p0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * RUN;
    if (p0 >= n) return;
    int64_t p1 = p0 + RUN < n ? p0 + RUN : n;
    const int64_t WH = (int64_t)W * H;
    uint32_t cur = p2n[p0];
    uint32_t cnt = 0;
    int64_t s[DCF_MAX_SUMS];
    uint32_t mn[3], mx[3];
    for (int i = 0; i < DCF_MAX_SUMS; i++) s[i] = 0;
    for (int i = 0; i < 3; i++) { mn[i] = 0xFFFFFFFFu; mx[i] = 0; }
    auto flush = [&](uint32_t k) {
        atomicAdd(&A.area[k], cnt);
        for (int i = 0; i < A.nsum; i++)
            if (s[i]) atomicAdd((unsigned long long *)&A.s[i][k], (unsigned long long)s[i]);
        if (A.has_ext) {
            for (int i = 0; i < 3; i++) { atomicMin(&A.mn[i][k], mn[i]); atomicMax(&A.mx[i][k], mx[i]); }
        }
    };
    for (int64_t p = p0; p < p1; p++) {
        uint32_t k = p2n[p];
        if (k != cur) {
            flush(cur);
            cur = k; cnt = 0;
            for (int i = 0; i < DCF_MAX_SUMS; i++) s[i] = 0;
            for (int i = 0; i < 3; i++) { mn[i] = 0xFFFFFFFFu; mx[i] = 0; }
        }
        cnt++;
        if (grp == 0) continue;
        int64_t x = p % W, y = (p / W) % H, z = p / WH;
        if (grp == 1) {
            uint32_t c[3] = { (uint32_t)x, (uint32_t)y, (uint32_t)z };
            for (int i = 0; i < 3; i++) { mn[i] = min(mn[i], c[i]); mx[i] = max(mx[i], c[i] + 1); }
        } else if (grp == 2) {
            s[0] += x; s[1] += y; s[2] += z;
        } else if (grp == 3) {
            int64_t gv = (int64_t)px_grey(iv, p);
            s[0] += gv; s[1] += gv * x; s[2] += gv * y; s[3] += gv * z;
        } else {
            s[0] += x; s[1] += y; s[2] += z;
            s[3] += x * x; s[4] += y * y; s[5] += z * z;
            s[6] += x * y; s[7] += y * z; s[8] += x * z;
        }
    }
    flush(cur);
}

/* one doubling round: dst = src + sum over nodes m with anc(m) = k of src[m]; anc' = anc(anc) */
__global__ __launch_bounds__(256) void k_leaf_copy(DevAttr src, DevAttr dst, const uint32_t *__restrict__ active, int round,
                                                   int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    dst.area[k] = src.area[k];
    for (int i = 0; i < src.nsum; i++) dst.s[i][k] = src.s[i][k];
    if (src.has_ext)
        for (int i = 0; i < 3; i++) { dst.mn[i][k] = src.mn[i][k]; dst.mx[i][k] = src.mx[i][k]; }
}

__global__ __launch_bounds__(256) void k_leaf_round(DevAttr src, DevAttr dst, const int32_t *__restrict__ anc,
                                                    int32_t *__restrict__ anc_out, uint32_t *__restrict__ active, int round,
                                                    int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t a2 = -1;
    if (k < K) {
        int32_t a = anc[k];
        if (a >= 0) {
            atomicAdd(&dst.area[a], src.area[k]);
            for (int i = 0; i < src.nsum; i++) {
                int64_t v = src.s[i][k];
                if (v) atomicAdd((unsigned long long *)&dst.s[i][a], (unsigned long long)v);
            }
            if (src.has_ext)
                for (int i = 0; i < 3; i++) { atomicMin(&dst.mn[i][a], src.mn[i][k]); atomicMax(&dst.mx[i][a], src.mx[i][k]); }
            a2 = anc[a];
        }
        anc_out[k] = a2;
    }
    /* count nodes that still have an ancestor at the next distance */
    unsigned long long b = __ballot(k < K && a2 >= 0);
    /* "another round is needed": a flag, set once (an atomic add per wave on one word serialised
     * hundreds of thousands of updates per round) */
    if ((threadIdx.x & (warpSize - 1)) == 0 && b &&
        __hip_atomic_load(&active[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        __hip_atomic_store(&active[round], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* ------------------------------------------------------------------------------------------------ */
/* host side                                                                                        */
/* ------------------------------------------------------------------------------------------------ */
static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static ImgView view_of(dcf_tree *t)
{
    ImgView iv;
    iv.img = t->img;
    iv.dtype = t->dtype;
    iv.invert = t->tree_min;
    iv.gmax = t->gmax;
    return iv;
}

/* node numbering from levelroot flags: p2n (scan), K, node parent / grey; consumes flag */
static int compact_nodes(dcf_tree *t, uint32_t *flag)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t n = t->n;
    ImgView iv = view_of(t);
    size_t need = 0;
    DCF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, flag, t->p2n, (int)n, ctx->stream));
    void *scan_tmp = dcf_alloc(ctx, need > 0 ? need : 16);
    if (!scan_tmp) return dcf_set_error(ctx, DCF_ENOMEM, "scan scratch");
    {
        hipEvent_t ea = nullptr, eb = nullptr;
        if (ctx->profiling && (ctx->prof_only.empty() || strstr("scan_nodes", ctx->prof_only.c_str())))
            dcf_prof_begin(ctx, "scan_nodes", &ea, &eb);
        hipError_t se = hipcub::DeviceScan::ExclusiveSum(scan_tmp, need, flag, t->p2n, (int)n, ctx->stream);
        if (ea) dcf_prof_end(ctx, "scan_nodes", ea, eb);
        dcf_release(ctx, scan_tmp);
        if (se != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "scan: %s", hipGetErrorString(se));
    }
    uint32_t last[2];
    DCF_HIP(ctx, hipMemcpyAsync(&last[0], t->p2n + (n - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipMemcpyAsync(&last[1], flag + (n - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    t->K = (int64_t)last[0] + last[1];
    dcf_release(ctx, t->npar);
    dcf_release(ctx, t->ngrey);
    t->npar = (int32_t *)dcf_alloc(ctx, (size_t)(t->K > 0 ? t->K : 1) * 4);
    t->ngrey = (float *)dcf_alloc(ctx, (size_t)(t->K > 0 ? t->K : 1) * 4);
    if (!t->npar || !t->ngrey) return dcf_set_error(ctx, DCF_ENOMEM, "node arrays");
    DCF_LAUNCH(ctx, "k_nodes", k_nodes, dim3(nblk(n, 256)), dim3(256), 0, iv, t->parent, flag, t->p2n, t->npar, t->ngrey, n);
    /* depth bound for the doubling loops: the number of distinct keys */
    t->max_rounds = t->dtype == DCF_U8 ? 9 : (t->dtype == DCF_U16 ? 17 : 33);
    return DCF_OK;
}

static int alloc_attr(dcf_ctx *ctx, DevAttr &a, int grp, int64_t K);
static int leaffix_group(dcf_tree *t, int grp);

/* u8/u16: flooding build (parent + pixel-indexed attributes), then node arrays and node attributes */
static int flood_and_gather(dcf_tree *t, int grp, bool renumber)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t n = t->n;
    int rc = dcf_flood_build(t, grp);
    if (rc) { dcf_free_attr(ctx, t->attr_pix); return rc; }
    uint32_t *flag = (uint32_t *)dcf_alloc(ctx, (size_t)n * 4);
    if (!flag) { dcf_free_attr(ctx, t->attr_pix); return dcf_set_error(ctx, DCF_ENOMEM, "levelroot flags"); }
    ImgView iv = view_of(t);
    DCF_LAUNCH(ctx, "k_flags", k_flags, dim3(nblk(n, 256)), dim3(256), 0, iv, t->parent, flag, n);
    if (renumber && (rc = compact_nodes(t, flag)) != DCF_OK) {
        dcf_release(ctx, flag);
        dcf_free_attr(ctx, t->attr_pix);
        return rc;
    }
    dcf_free_attr(ctx, t->attr);
    if ((rc = alloc_attr(ctx, t->attr, grp, t->K)) != DCF_OK) {
        dcf_release(ctx, flag);
        dcf_free_attr(ctx, t->attr_pix);
        return rc;
    }
    DCF_LAUNCH(ctx, "k_gather_attr", k_gather_attr, dim3(nblk(n, 256)), dim3(256), 0, flag, t->p2n, t->attr_pix, t->attr, n);
    dcf_release(ctx, flag);
    dcf_free_attr(ctx, t->attr_pix);
    t->grp = grp;
    return DCF_OK;
}

int dcf_build_tree(dcf_tree *t, int grp)
{
    if (t->dtype != DCF_F32) return flood_and_gather(t, grp, true);

    /* float images: tile zips + seam zips + canonicalisation, then the doubling leaffix */
    dcf_ctx *ctx = t->ctx;
    const int64_t n = t->n;
    const int is3d = t->dims[2] > 1;
    if (!t->link) {
        t->link = (uint64_t *)dcf_alloc(ctx, (size_t)n * 8);
        if (!t->link) return dcf_set_error(ctx, DCF_ENOMEM, "link words");
    }
    TileGeom g;
    g.W = (int)t->dims[0]; g.H = (int)t->dims[1]; g.D = (int)t->dims[2];
    if (is3d) { g.TX = ctx->tile3_x; g.TY = ctx->tile3_y; g.TZ = ctx->tile3_z; }
    else { g.TX = ctx->tile_x; g.TY = ctx->tile_y; g.TZ = 1; }
    if (g.TX < 1 || g.TY < 1 || g.TZ < 1 || (int64_t)g.TX * g.TY * g.TZ > TILE_MAX)
        return dcf_set_error(ctx, DCF_EINVAL, "tile %dx%dx%d exceeds %d pixels", g.TX, g.TY, g.TZ, TILE_MAX);
    g.ntx = (g.W + g.TX - 1) / g.TX;
    g.nty = (g.H + g.TY - 1) / g.TY;
    g.ntz = (g.D + g.TZ - 1) / g.TZ;
    int64_t ntiles = (int64_t)g.ntx * g.nty * g.ntz;
    Fwd fwd = make_fwd(t->conn, is3d);
    ImgView iv = view_of(t);

    DCF_LAUNCH(ctx, "k_tile_build", k_tile_build, dim3((unsigned)ntiles), dim3(TILE_THREADS), 0, iv, t->link, g, fwd);
    DCF_LAUNCH(ctx, "k_seam", k_seam, dim3(nblk(ntiles * NFACES, 64)), dim3(64), 0, iv, t->link, g, fwd, ntiles);

    uint32_t *flag = (uint32_t *)dcf_alloc(ctx, (size_t)n * 4);
    if (!flag) return dcf_set_error(ctx, DCF_ENOMEM, "levelroot flags");
    DCF_LAUNCH(ctx, "k_canon", k_canon, dim3(nblk(n, 256)), dim3(256), 0, iv, t->link, t->parent, flag, n);
    int rc = compact_nodes(t, flag);
    dcf_release(ctx, flag);
    if (rc) return rc;
    return leaffix_group(t, grp);
}

int dcf_accumulate_group(dcf_tree *t, int grp)
{
    if (t->dtype != DCF_F32) return flood_and_gather(t, grp, false);
    return leaffix_group(t, grp);
}

void dcf_free_attr(dcf_ctx *ctx, DevAttr &a)
{
    dcf_release(ctx, a.area);
    for (int i = 0; i < DCF_MAX_SUMS; i++) dcf_release(ctx, a.s[i]);
    for (int i = 0; i < 3; i++) {
        dcf_release(ctx, a.mn[i]);
        dcf_release(ctx, a.mx[i]);
    }
    a = DevAttr();
}

static int alloc_attr(dcf_ctx *ctx, DevAttr &a, int grp, int64_t K)
{
    a = DevAttr();
    a.nsum = grp == 2 ? 3 : (grp == 3 ? 4 : (grp == 4 ? 9 : 0));
    a.has_ext = grp == 1;
    size_t k = (size_t)(K > 0 ? K : 1);
    bool ok = (a.area = (uint32_t *)dcf_alloc(ctx, k * 4)) != nullptr;
    for (int i = 0; ok && i < a.nsum; i++) ok = (a.s[i] = (int64_t *)dcf_alloc(ctx, k * 8)) != nullptr;
    for (int i = 0; ok && a.has_ext && i < 3; i++) {
        ok = (a.mn[i] = (uint32_t *)dcf_alloc(ctx, k * 4)) != nullptr;
        ok = ok && (a.mx[i] = (uint32_t *)dcf_alloc(ctx, k * 4)) != nullptr;
    }
    if (!ok) {
        dcf_free_attr(ctx, a);
        return dcf_set_error(ctx, DCF_ENOMEM, "attribute arrays");
    }
    return DCF_OK;
}

static int leaffix_group(dcf_tree *t, int grp)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K, n = t->n;
    dcf_free_attr(ctx, t->attr);
    DevAttr A, B;
    int rc = alloc_attr(ctx, A, grp, K);
    if (rc) return rc;
    rc = alloc_attr(ctx, B, grp, K);
    if (rc) { dcf_free_attr(ctx, A); return rc; }

    DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(nblk(K, 256)), dim3(256), 0, A.area, 0u, K);
    for (int i = 0; i < A.nsum; i++) DCF_LAUNCH(ctx, "k_fill", k_fill_i64, dim3(nblk(K, 256)), dim3(256), 0, A.s[i], (int64_t)0, K);
    if (A.has_ext)
        for (int i = 0; i < 3; i++) {
            DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(nblk(K, 256)), dim3(256), 0, A.mn[i], 0xFFFFFFFFu, K);
            DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(nblk(K, 256)), dim3(256), 0, A.mx[i], 0u, K);
        }
    ImgView iv = view_of(t);
    DCF_LAUNCH(ctx, "k_own", k_own, dim3(nblk((n + RUN - 1) / RUN, 256)), dim3(256), 0, iv, t->p2n, A, grp, (int)t->dims[0],
               (int)t->dims[1], n);

    int32_t *anc0 = (int32_t *)dcf_alloc(ctx, (size_t)(K > 0 ? K : 1) * 4);
    int32_t *anc1 = (int32_t *)dcf_alloc(ctx, (size_t)(K > 0 ? K : 1) * 4);
    if (!anc0 || !anc1) {
        dcf_release(ctx, anc0); dcf_release(ctx, anc1);
        dcf_free_attr(ctx, A); dcf_free_attr(ctx, B);
        return dcf_set_error(ctx, DCF_ENOMEM, "ancestor arrays");
    }
    DCF_HIP(ctx, hipMemcpyAsync(anc0, t->npar, (size_t)K * 4, hipMemcpyDeviceToDevice, ctx->stream));
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(1), dim3(64), 0, t->counters, 0u, (int64_t)64);
    DevAttr *src = &A, *dst = &B;
    int32_t *ain = anc0, *aout = anc1;
    for (int r = 0; r < t->max_rounds; r++) {
        DCF_LAUNCH(ctx, "k_leaf_copy", k_leaf_copy, dim3(nblk(K, 256)), dim3(256), 0, *src, *dst, t->counters, r, K);
        DCF_LAUNCH(ctx, "k_leaf_round", k_leaf_round, dim3(nblk(K, 256)), dim3(256), 0, *src, *dst, ain, aout, t->counters, r, K);
        DevAttr *ts = src; src = dst; dst = ts;
        int32_t *ta = ain; ain = aout; aout = ta;
    }
    dcf_release(ctx, anc0);
    dcf_release(ctx, anc1);
    /* rounds skip themselves on the device once a round reports no remaining ancestors, so the rounds
     * that ran are 0..R-1 with R-1 the first round whose counter is 0; the sums are in the buffer
     * written by round R-1: B when R is odd, A when R is even */
    uint32_t act[64];
    DCF_HIP(ctx, hipMemcpyAsync(act, t->counters, sizeof act, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int ran = 0;
    for (int r = 0; r < t->max_rounds; r++) {
        ran = r + 1;
        if (act[r] == 0) break;
    }
    if (ran == t->max_rounds && act[t->max_rounds - 1] != 0) {
        dcf_free_attr(ctx, A); dcf_free_attr(ctx, B);
        return dcf_set_error(ctx, DCF_EDEVICE, "leaffix did not converge in %d rounds", t->max_rounds);
    }
    if (ran & 1) { t->attr = B; dcf_free_attr(ctx, A); }
    else { t->attr = A; dcf_free_attr(ctx, B); }
    t->grp = grp;
    return DCF_OK;
}
