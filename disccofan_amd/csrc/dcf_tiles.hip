/*
 * dcf_tiles.hip -- tiled build: the image cut into tiles with a one-pixel overlap, every tile's tree
 * built on its own, and the tiles' boundary trees merged on the GPU into the tree of the whole image.
 *
 * The reference distributes an image over its MPI processes this way: near-equal tiles, the first
 * ones a pixel larger when the size does not divide, each extended by one pixel into every neighbour
 * (io.c:860-882); every process builds its tile's tree (build_local_tree), extracts the boundary tree
 * of its faces, the overlap pixels' levelroots and all their ancestors (create_boundary, add_side,
 * add_ancestors: boundary.c:1445-1707), merges boundary trees pairwise by zipping ancestor chains by
 * grey, equal greys becoming one node through border_lr links (merge_branch, boundary.c:427-583;
 * merge_boundary_trees :623-724), pushes the merged parents back (update_boundary_trees :1201-1317) and
 * corrects its local tree (correct_local_tree :1326-1399).  On one GPU, the tiles of one image:
 *
 *   tiles     every tile gathered into its own buffer and built by the single-image builder; its
 *             attribute sums take the tile's origin (coordinates of the whole image) and a min-tree
 *             the whole image's grey maximum, so node keys and sums are comparable across tiles
 *   boundary  the nodes that hold a pixel shared with another tile, and all their ancestors (pointer
 *             doubling), numbered across the tiles: the boundary nodes
 *   pairs     a pixel that a tile holds but another tile owns is one pixel in two trees: its two
 *             nodes (same grey) are one node of the whole tree.  Pixels of one band (same holder,
 *             same owner) with a neighbour of higher (grey, -index) in the band are skipped: the zip
 *             of that neighbour's nodes merges their chains too
 *   zip       merge_branch on the GPU: one thread per pair zips the two ancestor chains by grey, with
 *             compare-and-swap on the boundary nodes' parent entries (lock-free); two nodes of equal
 *             grey are linked towards the smaller levelroot index, so every class of equal-grey nodes
 *             ends with its canonical levelroot (the minimum raster index of its level pixels) on top
 *   classes   each such class is one node of the whole tree.  Its attributes are subtree sums, over
 *             the merged boundary forest, of every node's base: its level pixels the tile owns plus
 *             its non-boundary children (whose subtrees lie inside the tile), summed along an Euler
 *             tour of the classes (list ranking by pointer jumping, prefix sums)
 *   output    the canonical parent array of the whole image, every pixel read from its owner tile,
 *             then node arrays and attributes as dcf_tree_build gives them
 * Every attribute group merges exactly: sums through prefix sums over the tour, the extent group's
 * minima and maxima through range queries over it.
 */
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <memory>
#include <type_traits>
#include <vector>

#include "dcf_internal.h"

#define TM_NONE 0xFFFFFFFFu
/* a zip or class walk that takes more steps than this reports a failed merge instead of spinning */
#define TM_SPIN_LIMIT (1u << 24)

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

/* ------------------------------------------------------------------------------------------------ */
/* geometry                                                                                         */
/* ------------------------------------------------------------------------------------------------ */
/* the tiling of the whole image: per axis g tiles, the first rem of them cnt + 1 pixels long, the
 * others cnt (io.c:864-873) */
struct GridGeo {
    int dim[3];
    int g[3];
    int cnt[3], rem[3];
};

__host__ __device__ __forceinline__ int own_lo(const GridGeo &G, int a, int r)
{
    return r * G.cnt[a] + (r < G.rem[a] ? r : G.rem[a]);
}

/* the tile (along axis a) that owns coordinate c */
__host__ __device__ __forceinline__ int owner_along(const GridGeo &G, int a, int c)
{
    const int big = G.rem[a] * (G.cnt[a] + 1);
    return c < big ? c / (G.cnt[a] + 1) : G.rem[a] + (c - big) / G.cnt[a];
}

/* one tile's tree on the device */
struct TileDev {
    int ox, oy, oz;      /* origin of the extended tile in the whole image */
    int W, H, D;         /* extended tile dimensions */
    int r[3];            /* tile coordinates in the grid */
    int lo[3], hi[3];    /* the pixels it owns: [lo, hi) per axis, whole-image coordinates */
    int64_t n, K;
    const int32_t *par;  /* tile-local canonical parent */
    const uint32_t *p2n; /* pixel -> node */
    const int32_t *npar; /* node -> parent node */
    uint32_t *n2p;       /* node -> levelroot pixel (tile-local index) */
    uint32_t *bid;       /* node -> boundary node id, or TM_NONE */
    DevAttr attr;        /* node-indexed attributes (whole-image coordinates) */
};

__device__ __forceinline__ int64_t tile_global(const TileDev &t, const GridGeo &G, int64_t pl)
{
    const int64_t row = pl / t.W;
    const int x = (int)(pl - row * t.W);
    const int z = (int)(row / t.H);
    const int y = (int)(row - (int64_t)z * t.H);
    return (int64_t)(t.ox + x) + (int64_t)(t.oy + y) * G.dim[0] + (int64_t)(t.oz + z) * G.dim[0] * G.dim[1];
}

__device__ __forceinline__ int tile_index(const GridGeo &G, int rx, int ry, int rz) { return rx + G.g[0] * (ry + G.g[1] * rz); }

/* grey key of pixel p of the whole image, as the tiles' trees order them (px_key of the build: a
 * min-tree inverts against the whole image's grey maximum) */
__device__ __forceinline__ uint32_t tm_fkey(float v)
{
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? 0xFFFFFFFFu - u : (u | 0x80000000u);
}

__device__ __forceinline__ uint32_t tm_key(const ImgView &iv, int64_t p)
{
    if (iv.dtype == DCF_U8) {
        const uint32_t v = ((const uint8_t *)iv.img)[p];
        return iv.invert ? (uint32_t)(*iv.gmax) - v : v;
    }
    if (iv.dtype == DCF_U16) {
        const uint32_t v = ((const uint16_t *)iv.img)[p];
        return iv.invert ? (uint32_t)(*iv.gmax) - v : v;
    }
    const float v = ((const float *)iv.img)[p];
    return tm_fkey(iv.invert ? *iv.gmax - v : v);
}

__global__ __launch_bounds__(256) void k_tm_fill32(uint32_t *a, uint32_t v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_tm_fill64(int64_t *a, int64_t v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* ------------------------------------------------------------------------------------------------ */
/* tiles                                                                                            */
/* ------------------------------------------------------------------------------------------------ */
/* the extended tile's pixels, copied out of the whole image (esz bytes per pixel) */
__global__ __launch_bounds__(256) void k_tm_gather(const uint8_t *__restrict__ img, int esz, GridGeo G, int ox, int oy, int oz, int W,
                                                   int H, int64_t n, uint8_t *__restrict__ dst)
{
    const int64_t pl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pl >= n) return;
    const int64_t row = pl / W;
    const int x = (int)(pl - row * W);
    const int z = (int)(row / H);
    const int y = (int)(row - (int64_t)z * H);
    const int64_t p = (int64_t)(ox + x) + (int64_t)(oy + y) * G.dim[0] + (int64_t)(oz + z) * G.dim[0] * G.dim[1];
    for (int b = 0; b < esz; b++) dst[pl * esz + b] = img[p * esz + b];
}

/* node -> levelroot pixel; every node of the tile's tree starts unmarked */
__global__ __launch_bounds__(256) void k_tm_n2p(TileDev t, uint32_t *__restrict__ mark)
{
    const int64_t pl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pl >= t.n) return;
    const uint32_t k = t.p2n[pl];
    const int32_t r = t.par[pl];
    if (r < 0 || t.p2n[r] != k) {
        t.n2p[k] = (uint32_t)pl;
        mark[k] = 0;
    }
}

/* the nodes of pixels shared with another tile: a pixel in the one-pixel band on either side of a
 * face with a neighbour tile (it lies in both tiles' extended boxes) */
__global__ __launch_bounds__(256) void k_tm_mark_shared(TileDev t, GridGeo G, uint32_t *__restrict__ mark)
{
    const int64_t pl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pl >= t.n) return;
    const int64_t row = pl / t.W;
    const int c[3] = { t.ox + (int)(pl - row * t.W), t.oy + (int)(row % t.H), t.oz + (int)(row / t.H) };
    bool shared = false;
    for (int a = 0; a < 3; a++) {
        if (t.lo[a] > 0 && c[a] <= t.lo[a]) shared = true;
        if (t.hi[a] < G.dim[a] && c[a] >= t.hi[a] - 1) shared = true;
    }
    if (shared) mark[t.p2n[pl]] = 1u;
}

/* one round of ancestor marking by pointer doubling: a marked node marks its ancestor 2^r steps up
 * (after round r every ancestor within 2^(r+1) steps of a first mark is marked) */
__global__ __launch_bounds__(256) void k_tm_anc_round(uint32_t *__restrict__ mark, const int32_t *__restrict__ jin,
                                                      int32_t *__restrict__ jout, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t j = jin[k];
    if (j >= 0 && mark[k]) mark[j] = 1u;
    jout[k] = j >= 0 ? jin[j] : -1;
}

/* boundary node ids from the exclusive scan of the marks (scan in bid) */
__global__ __launch_bounds__(256) void k_tm_bid(const uint32_t *__restrict__ mark, uint32_t *__restrict__ bid, uint32_t base, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    bid[k] = mark[k] ? base + bid[k] : TM_NONE;
}

/* the boundary nodes of one tile: levelroot index in the whole image, and key << 32 | parent boundary node */
__global__ __launch_bounds__(256) void k_tm_bnodes(TileDev t, GridGeo G, ImgView iv, uint32_t *__restrict__ blr,
                                                   unsigned long long *__restrict__ bw)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= t.K) return;
    const uint32_t v = t.bid[k];
    if (v == TM_NONE) return;
    const int64_t p = tile_global(t, G, t.n2p[k]);
    const uint32_t key = tm_key(iv, p);
    blr[v] = (uint32_t)p;
    const int32_t q = t.npar[k];
    bw[v] = ((unsigned long long)key << 32) | (q >= 0 ? t.bid[q] : TM_NONE);
}

/* ------------------------------------------------------------------------------------------------ */
/* base attributes of the boundary nodes                                                            */
/* ------------------------------------------------------------------------------------------------ */
template <int G> struct TmShape {
    static constexpr int NS = G == 2 ? 3 : (G == 3 ? 4 : (G == 4 ? 9 : 0));
    static constexpr bool EXT = G == 1;
};

__device__ __forceinline__ uint32_t tm_wmin(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t tm_wmax(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
    return v;
}

/* adds (area, sums, extents) of every valid lane to record r: one atomic per field when the wave's valid
 * lanes share one record (a node's many owned pixels), else one per lane */
template <int G>
__device__ __forceinline__ void tm_add(const DevAttr &A, bool valid, uint32_t r, uint32_t area, const int64_t (&s)[TmShape<G>::NS > 0 ? TmShape<G>::NS : 1],
                                       const uint32_t (&mn)[3], const uint32_t (&mx)[3])
{
    constexpr int NS = TmShape<G>::NS;
    const uint64_t vm = __ballot(valid);
    if (!vm) return;
    const int leader = __ffsll((unsigned long long)vm) - 1;
    const uint32_t rl = __shfl(r, leader, 64);
    if (__ballot(valid && r != rl) == 0) {
        uint64_t a = valid ? area : 0u;
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t lo = __shfl_xor((uint32_t)a, o, 64), hi = __shfl_xor((uint32_t)(a >> 32), o, 64);
            a += ((uint64_t)hi << 32) | lo;
        }
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&A.area[rl], (uint32_t)a);
#pragma unroll
        for (int i = 0; i < NS; i++) {
            uint64_t v = valid ? (uint64_t)s[i] : 0u;
            for (int o = 32; o > 0; o >>= 1) {
                const uint32_t lo = __shfl_xor((uint32_t)v, o, 64), hi = __shfl_xor((uint32_t)(v >> 32), o, 64);
                v += ((uint64_t)hi << 32) | lo;
            }
            if ((int)(threadIdx.x & 63) == leader && v) atomicAdd((unsigned long long *)&A.s[i][rl], (unsigned long long)v);
        }
        if constexpr (TmShape<G>::EXT)
            for (int i = 0; i < 3; i++) {
                const uint32_t lo = tm_wmin(valid ? mn[i] : 0xFFFFFFFFu), hi = tm_wmax(valid ? mx[i] : 0u);
                if ((int)(threadIdx.x & 63) == leader) {
                    atomicMin(&A.mn[i][rl], lo);
                    atomicMax(&A.mx[i][rl], hi);
                }
            }
        return;
    }
    if (!valid) return;
    atomicAdd(&A.area[r], area);
#pragma unroll
    for (int i = 0; i < NS; i++)
        if (s[i]) atomicAdd((unsigned long long *)&A.s[i][r], (unsigned long long)s[i]);
    if constexpr (TmShape<G>::EXT)
        for (int i = 0; i < 3; i++) {
            atomicMin(&A.mn[i][r], mn[i]);
            atomicMax(&A.mx[i][r], mx[i]);
        }
}

/* the owned pixels of the tile whose node is a boundary node add their contribution to it (the
 * reference initialises the other overlap pixels empty, init_attrib_array attributes.c:1201-1221) */
template <int G>
__global__ __launch_bounds__(256) void k_tm_base_pix(TileDev t, GridGeo Gg, ImgView iv, DevAttr BA)
{
    constexpr int NS = TmShape<G>::NS;
    const int64_t pl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    uint32_t v = 0;
    int64_t s[NS > 0 ? NS : 1] = {};
    uint32_t mn[3] = { 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu }, mx[3] = { 0u, 0u, 0u };
    if (pl < t.n) {
        const int64_t row = pl / t.W;
        const int x = t.ox + (int)(pl - row * t.W), y = t.oy + (int)(row % t.H), z = t.oz + (int)(row / t.H);
        const bool owned = x >= t.lo[0] && x < t.hi[0] && y >= t.lo[1] && y < t.hi[1] && z >= t.lo[2] && z < t.hi[2];
        v = owned ? t.bid[t.p2n[pl]] : TM_NONE;
        valid = v != TM_NONE;
        if (valid) {
            if constexpr (G == 1) {
                mn[0] = (uint32_t)x; mn[1] = (uint32_t)y; mn[2] = (uint32_t)z;
                mx[0] = (uint32_t)x + 1; mx[1] = (uint32_t)y + 1; mx[2] = (uint32_t)z + 1;
            } else if constexpr (G == 2) {
                s[0] = x; s[1] = y; s[2] = z;
            } else if constexpr (G == 3) {
                const int64_t p = (int64_t)x + (int64_t)y * Gg.dim[0] + (int64_t)z * Gg.dim[0] * Gg.dim[1];
                const int64_t gv = (int64_t)tm_key(iv, p);
                s[0] = gv; s[1] = gv * x; s[2] = gv * y; s[3] = gv * z;
            } else if constexpr (G == 4) {
                s[0] = x; s[1] = y; s[2] = z;
                s[3] = (int64_t)x * x; s[4] = (int64_t)y * y; s[5] = (int64_t)z * z;
                s[6] = (int64_t)x * y; s[7] = (int64_t)y * z; s[8] = (int64_t)x * z;
            }
        }
    }
    tm_add<G>(BA, valid, v, 1u, s, mn, mx);
}

/* non-boundary nodes under a boundary node add their complete records to it (their subtrees lie in
 * the tile and hold only owned pixels) */
template <int G>
__global__ __launch_bounds__(256) void k_tm_base_nodes(TileDev t, DevAttr BA)
{
    constexpr int NS = TmShape<G>::NS;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    uint32_t v = 0, area = 0;
    int64_t s[NS > 0 ? NS : 1] = {};
    uint32_t mn[3] = { 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu }, mx[3] = { 0u, 0u, 0u };
    if (k < t.K && t.bid[k] == TM_NONE) {
        const int32_t q = t.npar[k];
        v = q >= 0 ? t.bid[q] : TM_NONE;
        valid = v != TM_NONE;
        if (valid) {
            area = t.attr.area[k];
#pragma unroll
            for (int i = 0; i < NS; i++) s[i] = t.attr.s[i][k];
            if constexpr (TmShape<G>::EXT)
                for (int i = 0; i < 3; i++) { mn[i] = t.attr.mn[i][k]; mx[i] = t.attr.mx[i][k]; }
        }
    }
    tm_add<G>(BA, valid, v, area, s, mn, mx);
}

/* ------------------------------------------------------------------------------------------------ */
/* pairs and zip                                                                                    */
/* ------------------------------------------------------------------------------------------------ */
struct TmNbr {
    int n;
    int8_t dx[26], dy[26], dz[26];
};

/* every pixel that tile ti holds and another tile owns: the pair (its node here, its node in the
 * owner), unless a neighbour in the same band (held by ti, owned by the same tile) has a higher
 * (key, -index) */
__global__ __launch_bounds__(256) void k_tm_pairs(const TileDev *__restrict__ tiles, int ti, GridGeo G, ImgView iv, TmNbr nb,
                                                  uint2 *__restrict__ pairs, uint32_t *__restrict__ npairs, uint32_t cap,
                                                  uint32_t *__restrict__ fail)
{
    const TileDev &t = tiles[ti];
    const int64_t pl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pl >= t.n) return;
    const int64_t row = pl / t.W;
    const int x = t.ox + (int)(pl - row * t.W), y = t.oy + (int)(row % t.H), z = t.oz + (int)(row / t.H);
    const bool owned = x >= t.lo[0] && x < t.hi[0] && y >= t.lo[1] && y < t.hi[1] && z >= t.lo[2] && z < t.hi[2];
    if (owned) return;
    const int oj = tile_index(G, owner_along(G, 0, x), owner_along(G, 1, y), owner_along(G, 2, z));
    const int64_t WH = (int64_t)G.dim[0] * G.dim[1];
    const int64_t p = (int64_t)x + (int64_t)y * G.dim[0] + (int64_t)z * WH;
    const uint32_t kp = tm_key(iv, p);
    for (int k = 0; k < nb.n; k++) {
        const int qx = x + nb.dx[k], qy = y + nb.dy[k], qz = z + nb.dz[k];
        /* in ti's extended box, not owned by ti, owned by the same tile */
        if (qx < t.ox || qx >= t.ox + t.W || qy < t.oy || qy >= t.oy + t.H || qz < t.oz || qz >= t.oz + t.D) continue;
        if (qx >= t.lo[0] && qx < t.hi[0] && qy >= t.lo[1] && qy < t.hi[1] && qz >= t.lo[2] && qz < t.hi[2]) continue;
        if (tile_index(G, owner_along(G, 0, qx), owner_along(G, 1, qy), owner_along(G, 2, qz)) != oj) continue;
        const int64_t q = (int64_t)qx + (int64_t)qy * G.dim[0] + (int64_t)qz * WH;
        const uint32_t kq = tm_key(iv, q);
        if (kq > kp || (kq == kp && q < p)) return; /* covered by q's zip */
    }
    const TileDev &o = tiles[oj];
    const int64_t po = (int64_t)(x - o.ox) + (int64_t)(y - o.oy) * o.W + (int64_t)(z - o.oz) * o.W * o.H;
    const uint32_t a = t.bid[t.p2n[pl]], b = o.bid[o.p2n[po]];
    if (a == TM_NONE || b == TM_NONE) { /* a shared pixel's node is always a boundary node */
        atomicOr(fail, 4u);
        return;
    }
    const uint32_t slot = atomicAdd(npairs, 1u);
    if (slot < cap) pairs[slot] = make_uint2(a, b);
}

/* boundary node order: x above y when its key is higher, or equal with a larger (levelroot, id) --
 * the class of equal-key nodes keeps the smallest levelroot index on top.  Every node's key and parent
 * entry share one 64-bit word (key << 32 | parent), so one load brings both and the parent moves by a
 * 64-bit compare-and-swap. */
struct ZipView {
    unsigned long long *bw;
    const uint32_t *blr;
    uint32_t *fail;
};

__device__ __forceinline__ unsigned long long zp_word(const ZipView &z, uint32_t x)
{
    return __hip_atomic_load(&z.bw[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t zw_key(unsigned long long w) { return (uint32_t)(w >> 32); }
__device__ __forceinline__ uint32_t zw_par(unsigned long long w) { return (uint32_t)w; }

__device__ __forceinline__ bool zp_above(const ZipView &z, uint32_t x, uint32_t kx, uint32_t y, uint32_t ky)
{
    if (kx != ky) return kx > ky;
    const uint32_t lx = z.blr[x], ly = z.blr[y];
    return lx != ly ? lx > ly : x > y;
}

/* the top of x's class of equal-key nodes (its levelroot in the merged tree so far), with its word */
__device__ __forceinline__ uint32_t zp_lr(const ZipView &z, uint32_t x, unsigned long long &wx, uint32_t &steps)
{
    wx = zp_word(z, x);
    const uint32_t k = zw_key(wx);
    for (;;) {
        const uint32_t p = zw_par(wx);
        if (p == TM_NONE || ++steps > TM_SPIN_LIMIT) return x;
        const unsigned long long wp = zp_word(z, p);
        if (zw_key(wp) != k) return x;
        x = p;
        wx = wp;
    }
}

/* merge_branch: the ancestor chains of a and b become one chain ordered by key, equal keys forming one
 * class.  A node's parent entry only ever moves to a node between it and its old parent (compare-and-swap
 * on its word), so concurrent zips keep a forest whose chains descend in (key, levelroot, id); a zip stops
 * where its two chains have already met. */
__device__ void zip_pair(const ZipView &z, uint32_t a0, uint32_t b0)
{
    uint32_t steps = 0;
    unsigned long long wa, wb;
    uint32_t a = zp_lr(z, a0, wa, steps), b = zp_lr(z, b0, wb, steps);
    while (a != b) {
        if (++steps > TM_SPIN_LIMIT) {
            atomicOr(z.fail, 1u);
            return;
        }
        if (zp_above(z, b, zw_key(wb), a, zw_key(wa))) {
            const uint32_t t = a;
            a = b;
            b = t;
            const unsigned long long tw = wa;
            wa = wb;
            wb = tw;
        }
        /* a above b: b's class belongs on a's chain, at or below a */
        const uint32_t raw = zw_par(wa);
        uint32_t pa = TM_NONE;
        unsigned long long wpa = 0;
        if (raw != TM_NONE) {
            pa = zp_lr(z, raw, wpa, steps);
            if (pa == b) return;
            if (zp_above(z, pa, zw_key(wpa), b, zw_key(wb))) {
                a = pa; /* climb */
                wa = wpa;
                continue;
            }
        }
        /* b goes between a and its parent */
        unsigned long long exp = wa;
        const unsigned long long want = (wa & 0xFFFFFFFF00000000ull) | b;
        if (__hip_atomic_compare_exchange_strong(&z.bw[a], &exp, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            if (raw == TM_NONE) return;
            a = b;
            wa = wb;
            b = pa;
            wb = wpa;
        } else {
            a = zp_lr(z, a, wa, steps);
            b = zp_lr(z, b, wb, steps);
        }
    }
}

/* the first ancestor of x in the tiles' own trees (lift: 2^j-th ancestors, nl levels) whose key is below t */
__device__ __forceinline__ uint32_t first_below(const ZipView &z, const uint32_t *__restrict__ lift, int nl, uint32_t V, uint32_t x,
                                                uint32_t t)
{
    for (int j = nl - 1; j >= 0; j--) {
        const uint32_t y = lift[(size_t)j * V + x];
        if (y != TM_NONE && (uint32_t)(z.bw[y] >> 32) >= t) x = y;
    }
    return lift[x];
}

/* thread (pair i, segment s): s = 0 zips the pair's two nodes; s > 0 zips the first ancestors of each
 * below the key K (S - s) / S, K the pair's key.  Both zips state facts of the merged tree (every ancestor
 * of the pixel's nodes lies on one chain), and the segments' zips run concurrently, so a chain thousands of
 * nodes deep is merged in about 1/S of the steps. */
__global__ __launch_bounds__(256) void k_tm_zip(ZipView z, const uint2 *__restrict__ pairs, uint32_t np, uint32_t S,
                                                const uint32_t *__restrict__ lift, int nl, uint32_t V)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)np * S) return;
    const uint32_t i = (uint32_t)(t / S), seg = (uint32_t)(t % S);
    uint32_t a = pairs[i].x, b = pairs[i].y;
    if (seg) {
        const uint32_t K = (uint32_t)(z.bw[a] >> 32);
        const uint32_t th = (uint32_t)((uint64_t)K * (S - seg) / S);
        a = first_below(z, lift, nl, V, a, th);
        b = first_below(z, lift, nl, V, b, th);
        if (a == TM_NONE || b == TM_NONE) return;
    }
    zip_pair(z, a, b);
}

/* lift level 0: every boundary node's parent in its tile's tree (before any zip) */
__global__ __launch_bounds__(256) void k_tm_lift0(const unsigned long long *__restrict__ bw, uint32_t *__restrict__ lift, int64_t V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) lift[v] = (uint32_t)bw[v];
}

__global__ __launch_bounds__(256) void k_tm_lift(uint32_t *__restrict__ lift, int j, int64_t V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t u = lift[(size_t)(j - 1) * V + v];
    lift[(size_t)j * V + v] = u == TM_NONE ? TM_NONE : lift[(size_t)(j - 1) * V + u];
}

/* ------------------------------------------------------------------------------------------------ */
/* classes and their subtree sums                                                                   */
/* ------------------------------------------------------------------------------------------------ */
/* cls[v] = top of v's class; flag the tops */
__global__ __launch_bounds__(256) void k_tm_cls(ZipView z, uint32_t *__restrict__ cls, uint32_t *__restrict__ top,
                                                uint32_t *__restrict__ bpar, int64_t V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint32_t steps = 0;
    unsigned long long w;
    const uint32_t c = zp_lr(z, (uint32_t)v, w, steps);
    if (steps > TM_SPIN_LIMIT) atomicOr(z.fail, 2u);
    cls[v] = c;
    top[v] = c == (uint32_t)v ? 1u : 0u;
    bpar[v] = zw_par(zp_word(z, (uint32_t)v));
}

/* every boundary node's base into its class top (in place: tops receive, the others give) */
template <int G>
__global__ __launch_bounds__(256) void k_tm_class_base(const uint32_t *__restrict__ cls, DevAttr BA, int64_t V)
{
    constexpr int NS = TmShape<G>::NS;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint32_t c = cls[v];
    if (c == (uint32_t)v) return;
    atomicAdd(&BA.area[c], BA.area[v]);
#pragma unroll
    for (int i = 0; i < NS; i++)
        if (BA.s[i][v]) atomicAdd((unsigned long long *)&BA.s[i][c], (unsigned long long)BA.s[i][v]);
    if constexpr (TmShape<G>::EXT)
        for (int i = 0; i < 3; i++) {
            atomicMin(&BA.mn[i][c], BA.mn[i][v]);
            atomicMax(&BA.mx[i][c], BA.mx[i][v]);
        }
}

/* extents of the classes' subtrees: a sparse table of minima (maxima) over the tour positions, the
 * class bases at their entries and the neutral value at the exits.  f = i (0..2) for min i, 3 + i for
 * max i. */
__global__ __launch_bounds__(256) void k_tm_place_ext(const uint32_t *__restrict__ rank, const uint32_t *__restrict__ top_of,
                                                      DevAttr BA, int f, uint32_t *__restrict__ st, uint32_t E)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E || rank[e] >= E) return;
    const uint32_t pos = E - 1 - rank[e];
    const uint32_t u = e >> 1;
    const bool is_max = f >= 3;
    uint32_t x = is_max ? 0u : 0xFFFFFFFFu;
    if (!(e & 1) && u + 1 < E / 2) {
        const uint32_t v = top_of[u];
        x = is_max ? BA.mx[f - 3][v] : BA.mn[f][v];
    }
    st[pos] = x;
}

__global__ __launch_bounds__(256) void k_tm_st_level(uint32_t *__restrict__ st, uint32_t E, int j, int is_max)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t half = 1u << (j - 1);
    if ((uint64_t)i + 2 * half > E) return;
    const uint32_t *prev = st + (size_t)(j - 1) * E;
    const uint32_t a = prev[i], b = prev[i + half];
    st[(size_t)j * E + i] = is_max ? max(a, b) : min(a, b);
}

__global__ __launch_bounds__(256) void k_tm_st_query(const uint32_t *__restrict__ rank, const uint32_t *__restrict__ st, DevAttr MA,
                                                     int f, uint32_t C)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= C) return;
    const uint32_t E = 2 * (C + 1);
    if (rank[2 * u] >= E || rank[2 * u + 1] >= E) return;
    const uint32_t pin = E - 1 - rank[2 * u], pout = E - 1 - rank[2 * u + 1];
    const uint32_t len = pout - pin + 1;
    const int j = 31 - __clz(len);
    const uint32_t *lv = st + (size_t)j * E;
    const uint32_t a = lv[pin], b = lv[pout + 1 - (1u << j)];
    if (f >= 3) MA.mx[f - 3][u] = max(a, b);
    else MA.mn[f][u] = min(a, b);
}

/* per class (numbered by the exclusive scan ci of the tops): its parent class number (C for a root,
 * the virtual root of the forest), as the sort key of the child lists */
__global__ __launch_bounds__(256) void k_tm_class_par(const uint32_t *__restrict__ bpar, const uint32_t *__restrict__ cls,
                                                      const uint32_t *__restrict__ top, const uint32_t *__restrict__ ci,
                                                      uint32_t *__restrict__ key, uint32_t *__restrict__ val,
                                                      uint32_t *__restrict__ pci, uint32_t C, int64_t V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || !top[v]) return;
    const uint32_t u = ci[v];
    const uint32_t p = bpar[v];
    const uint32_t pu = p == TM_NONE ? C : ci[cls[p]];
    key[u] = pu;
    val[u] = u;
    pci[u] = pu;
}

/* first child and next sibling from the child lists sorted by parent */
__global__ __launch_bounds__(256) void k_tm_children(const uint32_t *__restrict__ skey, const uint32_t *__restrict__ sval,
                                                     uint32_t *__restrict__ fc, uint32_t *__restrict__ ns, uint32_t C)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= C) return;
    const uint32_t p = skey[s], c = sval[s];
    if (s == 0 || skey[s - 1] != p) fc[p] = c;
    ns[c] = (s + 1 < C && skey[s + 1] == p) ? sval[s + 1] : TM_NONE;
}

/* Euler tour of the class forest under the virtual root C: entry 2u enters u, 2u + 1 leaves it */
__global__ __launch_bounds__(256) void k_tm_tour(const uint32_t *__restrict__ fc, const uint32_t *__restrict__ ns,
                                                 const uint32_t *__restrict__ pci, uint32_t *__restrict__ nxt,
                                                 uint32_t *__restrict__ rank, uint32_t C)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * (C + 1)) return;
    const uint32_t u = e >> 1;
    uint32_t s;
    if (!(e & 1)) s = fc[u] != TM_NONE ? 2 * fc[u] : e + 1;
    else if (u == C) s = TM_NONE;
    else s = ns[u] != TM_NONE ? 2 * ns[u] : 2 * pci[u] + 1;
    nxt[e] = s;
    rank[e] = s == TM_NONE ? 0u : 1u;
}

/* list ranking: rank = steps to the end of the tour */
__global__ __launch_bounds__(256) void k_tm_rank(const uint32_t *__restrict__ nin, const uint32_t *__restrict__ rin,
                                                 uint32_t *__restrict__ nout, uint32_t *__restrict__ rout, uint32_t E)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint32_t s = nin[e];
    rout[e] = s == TM_NONE ? rin[e] : rin[e] + rin[s];
    nout[e] = s == TM_NONE ? TM_NONE : nin[s];
}

/* the class bases at the tour positions of their entries (0 at the exits), one field at a time
 * (f = 0 area, f = 1 + i sum i) */
__global__ __launch_bounds__(256) void k_tm_place(const uint32_t *__restrict__ rank, const uint32_t *__restrict__ top_of,
                                                  DevAttr BA, int f, int64_t *__restrict__ val, uint32_t E, uint32_t *__restrict__ fail)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    if (rank[e] >= E) { /* not one list: the class forest is broken */
        atomicOr(fail, 8u);
        return;
    }
    const uint32_t pos = E - 1 - rank[e];
    const uint32_t u = e >> 1;
    int64_t x = 0;
    if (!(e & 1) && u + 1 < E / 2) {
        const uint32_t v = top_of[u];
        x = f == 0 ? (int64_t)BA.area[v] : BA.s[f - 1][v];
    }
    val[pos] = x;
}

/* class u's subtree sum = prefix(exit) - prefix(entry) + its own base */
__global__ __launch_bounds__(256) void k_tm_subtree(const uint32_t *__restrict__ rank, const int64_t *__restrict__ pre,
                                                    const uint32_t *__restrict__ top_of, DevAttr BA, DevAttr MA, int f, uint32_t C)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= C) return;
    const uint32_t E = 2 * (C + 1);
    if (rank[2 * u] >= E || rank[2 * u + 1] >= E) return; /* reported by k_tm_place */
    const uint32_t pin = E - 1 - rank[2 * u], pout = E - 1 - rank[2 * u + 1];
    const uint32_t v = top_of[u];
    const int64_t own = f == 0 ? (int64_t)BA.area[v] : BA.s[f - 1][v];
    const int64_t tot = pre[pout] - pre[pin] + own;
    if (f == 0) MA.area[u] = (uint32_t)tot;
    else MA.s[f - 1][u] = tot;
}

__global__ __launch_bounds__(256) void k_tm_top_of(const uint32_t *__restrict__ top, const uint32_t *__restrict__ ci,
                                                   uint32_t *__restrict__ top_of, int64_t V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V && top[v]) top_of[ci[v]] = (uint32_t)v;
}

/* ------------------------------------------------------------------------------------------------ */
/* output                                                                                           */
/* ------------------------------------------------------------------------------------------------ */
/* the owner tile of pixel p and p's index in it */
__device__ __forceinline__ const TileDev &tm_owner(const TileDev *tiles, const GridGeo &G, int64_t p, int64_t &pl)
{
    const int64_t WH = (int64_t)G.dim[0] * G.dim[1];
    const int z = (int)(p / WH);
    const int64_t r = p - (int64_t)z * WH;
    const int y = (int)(r / G.dim[0]);
    const int x = (int)(r - (int64_t)y * G.dim[0]);
    const TileDev &t = tiles[tile_index(G, owner_along(G, 0, x), owner_along(G, 1, y), owner_along(G, 2, z))];
    pl = (int64_t)(x - t.ox) + (int64_t)(y - t.oy) * t.W + (int64_t)(z - t.oz) * t.W * t.H;
    return t;
}

/* the canonical parent of every pixel of the whole image */
__global__ __launch_bounds__(256) void k_tm_parent(const TileDev *__restrict__ tiles, GridGeo G, const uint32_t *__restrict__ cls,
                                                   const uint32_t *__restrict__ bpar, const uint32_t *__restrict__ blr,
                                                   int32_t *__restrict__ par, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int64_t pl;
    const TileDev &t = tm_owner(tiles, G, p, pl);
    const uint32_t k = t.p2n[pl];
    const uint32_t b = t.bid[k];
    int32_t out;
    if (b == TM_NONE) {
        const uint32_t lr = t.n2p[k];
        if ((int64_t)lr != pl) {
            out = (int32_t)tile_global(t, G, lr);
        } else {
            const int32_t kp = t.npar[k];
            if (kp < 0) out = -1;
            else if (t.bid[kp] == TM_NONE) out = (int32_t)tile_global(t, G, t.n2p[kp]);
            else out = (int32_t)blr[cls[t.bid[kp]]];
        }
    } else {
        const uint32_t L = cls[b];
        const uint32_t c = blr[L];
        if ((int64_t)c != p) out = (int32_t)c;
        else out = bpar[L] == TM_NONE ? -1 : (int32_t)blr[cls[bpar[L]]];
    }
    par[p] = out;
}

/* node-indexed attributes of the whole tree: a non-boundary node's record from its tile, a class's
 * merged record */
template <int G>
__global__ __launch_bounds__(256) void k_tm_attr_out(const TileDev *__restrict__ tiles, GridGeo Gg, const uint32_t *__restrict__ flag,
                                                     const uint32_t *__restrict__ p2n, const uint32_t *__restrict__ cls,
                                                     const uint32_t *__restrict__ ci, DevAttr MA, DevAttr A, int64_t n)
{
    constexpr int NS = TmShape<G>::NS;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !flag[p]) return;
    int64_t pl;
    const TileDev &t = tm_owner(tiles, Gg, p, pl);
    const uint32_t k = t.p2n[pl];
    const uint32_t b = t.bid[k];
    const uint32_t node = p2n[p];
    if (b == TM_NONE) {
        A.area[node] = t.attr.area[k];
#pragma unroll
        for (int i = 0; i < NS; i++) A.s[i][node] = t.attr.s[i][k];
        if constexpr (TmShape<G>::EXT)
            for (int i = 0; i < 3; i++) { A.mn[i][node] = t.attr.mn[i][k]; A.mx[i][node] = t.attr.mx[i][k]; }
    } else {
        const uint32_t u = ci[cls[b]];
        A.area[node] = MA.area[u];
#pragma unroll
        for (int i = 0; i < NS; i++) A.s[i][node] = MA.s[i][u];
        if constexpr (TmShape<G>::EXT)
            for (int i = 0; i < 3; i++) { A.mn[i][node] = MA.mn[i][u]; A.mx[i][node] = MA.mx[i][u]; }
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* host side                                                                                        */
/* ------------------------------------------------------------------------------------------------ */
namespace {

struct Blocks {
    dcf_ctx *ctx;
    std::vector<void *> p;
    template <class T> T *get(size_t count)
    {
        void *q = dcf_alloc(ctx, (count > 0 ? count : 1) * sizeof(T));
        if (q) p.push_back(q);
        return (T *)q;
    }
    ~Blocks()
    {
        for (void *q : p) dcf_release(ctx, q);
    }
};

struct TileTree {
    dcf_tree t;
    uint32_t *n2p = nullptr, *bid = nullptr, *mark = nullptr;
};

/* a tile's pixels, parent array and node numbering come from the caller's context; what its build
 * allocates (node arrays, attributes) from the context that built it (t.ctx: a worker when the tiles are
 * built concurrently), and each block goes back to the pool it came from */
struct TileTrees {
    dcf_ctx *ctx;
    std::vector<std::unique_ptr<TileTree>> v;
    ~TileTrees()
    {
        for (auto &tt : v) {
            dcf_tree &t = tt->t;
            dcf_ctx *b = t.ctx ? t.ctx : ctx;
            dcf_release(ctx, t.img);
            dcf_release(ctx, t.gmax);
            dcf_release(ctx, t.parent);
            dcf_release(ctx, t.p2n);
            dcf_release(ctx, t.counters);
            dcf_release(b, t.npar);
            dcf_release(b, t.ngrey);
            dcf_free_attr(b, t.attr);
            dcf_free_attr(b, t.attr_pix);
            dcf_release(ctx, tt->n2p);
            dcf_release(ctx, tt->bid);
            dcf_release(ctx, tt->mark);
        }
    }
};

TmNbr tm_nbr(int conn, int is3d)
{
    TmNbr f;
    f.n = 0;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                if (!x && !y && !z) continue;
                if (!is3d && z) continue;
                const int l1 = abs(x) + abs(y) + abs(z);
                if (!(is3d ? (conn == 26 || l1 == 1) : (conn == 8 || l1 == 1))) continue;
                f.dx[f.n] = (int8_t)x; f.dy[f.n] = (int8_t)y; f.dz[f.n] = (int8_t)z;
                f.n++;
            }
    return f;
}

template <class F> int tm_group(int grp, F &&f)
{
    switch (grp) {
    case 0: return f(std::integral_constant<int, 0>());
    case 1: return f(std::integral_constant<int, 1>());
    case 2: return f(std::integral_constant<int, 2>());
    case 3: return f(std::integral_constant<int, 3>());
    default: return f(std::integral_constant<int, 4>());
    }
}

} // namespace

int dcf_tiled_build(dcf_tree *T, const int grid[3], int grp, int schedule)
{
    (void)schedule;
    dcf_ctx *ctx = T->ctx;
    const int64_t n = T->n;
    GridGeo G;
    for (int a = 0; a < 3; a++) {
        G.dim[a] = (int)T->dims[a];
        G.g[a] = grid[a];
        G.cnt[a] = G.dim[a] / grid[a];
        G.rem[a] = G.dim[a] % grid[a];
    }
    const int is3d = T->dims[2] > 1;
    const int ntiles = grid[0] * grid[1] * grid[2];
    const size_t esz = T->dtype == DCF_U8 ? 1 : (T->dtype == DCF_U16 ? 2 : 4);
    ImgView iv;
    iv.img = T->img;
    iv.dtype = T->dtype;
    iv.invert = T->tree_min;
    iv.gmax = T->gmax;
    Blocks bl{ ctx, {} };
    TileTrees tt{ ctx, {} };
    std::vector<TileDev> td((size_t)ntiles);
    std::vector<uint32_t> vbase((size_t)ntiles + 1, 0);
    int rc = DCF_OK;

    /* ---- every tile's own tree ---- */
    for (int rz = 0; rz < grid[2]; rz++)
        for (int ry = 0; ry < grid[1]; ry++)
            for (int rx = 0; rx < grid[0]; rx++) {
                const int i = rx + grid[0] * (ry + grid[1] * rz);
                const int r[3] = { rx, ry, rz };
                TileDev &d = td[(size_t)i];
                int e0[3], e1[3];
                for (int a = 0; a < 3; a++) {
                    d.r[a] = r[a];
                    d.lo[a] = own_lo(G, a, r[a]);
                    d.hi[a] = own_lo(G, a, r[a] + 1);
                    /* the one-pixel extension into every neighbour (io.c:874-881) */
                    e0[a] = d.lo[a] > 0 ? d.lo[a] - 1 : 0;
                    e1[a] = d.hi[a] < G.dim[a] ? d.hi[a] + 1 : G.dim[a];
                }
                d.ox = e0[0]; d.oy = e0[1]; d.oz = e0[2];
                d.W = e1[0] - e0[0]; d.H = e1[1] - e0[1]; d.D = e1[2] - e0[2];
                d.n = (int64_t)d.W * d.H * d.D;
                tt.v.emplace_back(new TileTree());
                dcf_tree &t = tt.v.back()->t;
                t.ctx = ctx;
                t.dims[0] = (uint64_t)d.W; t.dims[1] = (uint64_t)d.H; t.dims[2] = (uint64_t)d.D;
                t.n = d.n;
                t.conn = T->conn;
                t.tree_min = T->tree_min;
                t.dtype = T->dtype;
                memcpy(t.pixdim, T->pixdim, sizeof t.pixdim);
                t.off[0] = d.ox; t.off[1] = d.oy; t.off[2] = d.oz;
                t.img = dcf_alloc(ctx, (size_t)d.n * esz);
                t.own_img = 1;
                t.gmax = (float *)dcf_alloc(ctx, 16);
                t.counters = (uint32_t *)dcf_alloc(ctx, 64 * 4);
                t.parent = (int32_t *)dcf_alloc(ctx, (size_t)d.n * 4);
                t.p2n = (uint32_t *)dcf_alloc(ctx, (size_t)d.n * 4);
                if (!t.img || !t.gmax || !t.counters || !t.parent || !t.p2n) return dcf_set_error(ctx, DCF_ENOMEM, "tile %d arrays", i);
            }
    /* one tile's pixels gathered and its tree built on context c (its stream) */
    auto build_tile = [&](dcf_ctx *c, int i) -> int {
        const TileDev &d = td[(size_t)i];
        dcf_tree &t = tt.v[(size_t)i]->t;
        t.ctx = c;
        DCF_LAUNCH(c, "k_tm_gather", k_tm_gather, dim3(nblk(d.n, 256)), dim3(256), 0, (const uint8_t *)T->img, (int)esz, G,
                   d.ox, d.oy, d.oz, d.W, d.H, d.n, (uint8_t *)t.img);
        /* the whole image's grey extremes: a min-tree tile inverts against the image's maximum */
        DCF_HIP(c, hipMemcpyAsync(t.gmax, T->gmax, 16, hipMemcpyDeviceToDevice, c->stream));
        return dcf_build_tree(&t, grp);
    };
    /* the tiles are independent: built concurrently on the worker contexts (each on its own stream,
     * after the work queued so far; the call returns with every tile built) */
    if (ntiles >= 2 && ctx->max_workers >= 2) {
        if ((rc = dcf_run_on_workers(ctx, ntiles, build_tile)) != DCF_OK) return rc;
    } else {
        for (int i = 0; i < ntiles; i++)
            if ((rc = build_tile(ctx, i)) != DCF_OK) return rc;
    }

    /* ---- every tile's boundary nodes and their numbering ---- */
    for (int i = 0; i < ntiles; i++) {
                TileDev &d = td[(size_t)i];
                TileTree &tr = *tt.v[(size_t)i];
                dcf_tree &t = tr.t;
                /* the tile's pixels are no longer needed: only its tree */
                dcf_release(ctx, t.img);
                t.img = nullptr;
                d.K = t.K;
                d.par = t.parent;
                d.p2n = t.p2n;
                d.npar = t.npar;
                d.attr = t.attr;
                tr.n2p = (uint32_t *)dcf_alloc(ctx, (size_t)(t.K > 0 ? t.K : 1) * 4);
                tr.bid = (uint32_t *)dcf_alloc(ctx, (size_t)(t.K > 0 ? t.K : 1) * 4);
                tr.mark = (uint32_t *)dcf_alloc(ctx, (size_t)(t.K > 0 ? t.K : 1) * 4);
                int32_t *j0 = bl.get<int32_t>((size_t)t.K), *j1 = bl.get<int32_t>((size_t)t.K);
                if (!tr.n2p || !tr.bid || !tr.mark || !j0 || !j1) return dcf_set_error(ctx, DCF_ENOMEM, "tile %d boundary arrays", i);
                d.n2p = tr.n2p;
                d.bid = tr.bid;
                /* boundary nodes: nodes of shared pixels, then all their ancestors */
                DCF_LAUNCH(ctx, "k_tm_n2p", k_tm_n2p, dim3(nblk(d.n, 256)), dim3(256), 0, d, tr.mark);
                DCF_LAUNCH(ctx, "k_tm_mark", k_tm_mark_shared, dim3(nblk(d.n, 256)), dim3(256), 0, d, G, tr.mark);
                DCF_HIP(ctx, hipMemcpyAsync(j0, t.npar, (size_t)t.K * 4, hipMemcpyDeviceToDevice, ctx->stream));
                int rounds = 1;
                while (rounds < 32 && ((int64_t)1 << rounds) < t.K) rounds++;
                for (int k = 0; k <= rounds; k++) {
                    DCF_LAUNCH(ctx, "k_tm_anc", k_tm_anc_round, dim3(nblk(t.K, 256)), dim3(256), 0, tr.mark, (const int32_t *)j0, j1, t.K);
                    std::swap(j0, j1);
                }
                size_t need = 0;
                DCF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, tr.mark, tr.bid, (int)t.K, ctx->stream));
                void *tmp = bl.get<char>(need);
                if (!tmp) return dcf_set_error(ctx, DCF_ENOMEM, "scan scratch");
                DCF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, need, tr.mark, tr.bid, (int)t.K, ctx->stream));
                uint32_t last[2];
                DCF_HIP(ctx, hipMemcpyAsync(&last[0], tr.bid + (t.K - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
                DCF_HIP(ctx, hipMemcpyAsync(&last[1], tr.mark + (t.K - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
                DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
                const uint64_t vi = (uint64_t)last[0] + last[1];
                if ((uint64_t)vbase[(size_t)i] + vi >= (uint64_t)TM_NONE) return dcf_set_error(ctx, DCF_EUNSUPPORTED, "too many boundary nodes");
                DCF_LAUNCH(ctx, "k_tm_bid", k_tm_bid, dim3(nblk(t.K, 256)), dim3(256), 0, (const uint32_t *)tr.mark, tr.bid,
                           vbase[(size_t)i], t.K);
                vbase[(size_t)i + 1] = vbase[(size_t)i] + (uint32_t)vi;
    }
    const int64_t V = vbase[(size_t)ntiles];

    /* ---- boundary nodes of all tiles ---- */
    uint32_t *blr = bl.get<uint32_t>((size_t)V), *bpar = bl.get<uint32_t>((size_t)V);
    unsigned long long *bw = bl.get<unsigned long long>((size_t)V);
    uint32_t *fail = bl.get<uint32_t>(4);
    TileDev *dtiles = bl.get<TileDev>((size_t)ntiles);
    if (!blr || !bpar || !bw || !fail || !dtiles) return dcf_set_error(ctx, DCF_ENOMEM, "boundary nodes");
    DCF_HIP(ctx, hipMemcpyAsync(dtiles, td.data(), td.size() * sizeof(TileDev), hipMemcpyHostToDevice, ctx->stream));
    DCF_LAUNCH(ctx, "k_fill", k_tm_fill32, dim3(1), dim3(64), 0, fail, 0u, (int64_t)4);
    DevAttr BA;
    if ((rc = dcf_alloc_attr(ctx, BA, grp, V)) != DCF_OK) return rc;
    struct AttrGuard {
        dcf_ctx *c;
        DevAttr *a;
        ~AttrGuard() { dcf_free_attr(c, *a); }
    } ba_guard{ ctx, &BA };
    const int64_t V1 = V > 0 ? V : 1;
    DCF_LAUNCH(ctx, "k_fill", k_tm_fill32, dim3(nblk(V1, 256)), dim3(256), 0, BA.area, 0u, V1);
    for (int i = 0; i < BA.nsum; i++) DCF_LAUNCH(ctx, "k_fill", k_tm_fill64, dim3(nblk(V1, 256)), dim3(256), 0, BA.s[i], (int64_t)0, V1);
    for (int i = 0; BA.has_ext && i < 3; i++) {
        DCF_LAUNCH(ctx, "k_fill", k_tm_fill32, dim3(nblk(V1, 256)), dim3(256), 0, BA.mn[i], 0xFFFFFFFFu, V1);
        DCF_LAUNCH(ctx, "k_fill", k_tm_fill32, dim3(nblk(V1, 256)), dim3(256), 0, BA.mx[i], 0u, V1);
    }
    for (int i = 0; i < ntiles; i++) {
        const TileDev &d = td[(size_t)i];
        DCF_LAUNCH(ctx, "k_tm_bnodes", k_tm_bnodes, dim3(nblk(d.K, 256)), dim3(256), 0, d, G, iv, blr, bw);
        rc = tm_group(grp, [&](auto g) -> int {
            constexpr int GV = decltype(g)::value;
            DCF_LAUNCH(ctx, "k_tm_base", k_tm_base_pix<GV>, dim3(nblk(d.n, 256)), dim3(256), 0, d, G, iv, BA);
            DCF_LAUNCH(ctx, "k_tm_base", k_tm_base_nodes<GV>, dim3(nblk(d.K, 256)), dim3(256), 0, d, BA);
            return DCF_OK;
        });
        if (rc) return rc;
    }

    /* ---- identification pairs and the zip ---- */
    int64_t held = 0;
    for (const TileDev &d : td) held += d.n;
    const uint32_t cap = (uint32_t)std::min<int64_t>(held - n + 1, (int64_t)TM_NONE - 1);
    uint2 *pairs = bl.get<uint2>(cap);
    uint32_t *npairs = bl.get<uint32_t>(4);
    if (!pairs || !npairs) return dcf_set_error(ctx, DCF_ENOMEM, "pairs");
    DCF_LAUNCH(ctx, "k_fill", k_tm_fill32, dim3(1), dim3(64), 0, npairs, 0u, (int64_t)4);
    const TmNbr nb = tm_nbr(T->conn, is3d);
    for (int i = 0; i < ntiles; i++)
        DCF_LAUNCH(ctx, "k_tm_pairs", k_tm_pairs, dim3(nblk(td[(size_t)i].n, 256)), dim3(256), 0, (const TileDev *)dtiles, i, G, iv, nb,
                   pairs, npairs, cap, fail);
    uint32_t np = 0;
    DCF_HIP(ctx, hipMemcpyAsync(&np, npairs, 4, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (np > cap) return dcf_set_error(ctx, DCF_EDEVICE, "pair list overflow (%u > %u)", np, cap);
    ZipView zv{ bw, blr, fail };
    if (np) {
        /* segment zips along the chains, from 2^j-th ancestors in the tiles' own trees */
        uint32_t S = (uint32_t)std::max(1, ctx->zip_segments);
        int nl = 1;
        while (nl < 32 && ((int64_t)1 << nl) < V) nl++;
        uint32_t *lift = S > 1 ? bl.get<uint32_t>((size_t)V * nl) : nullptr;
        if (!lift) S = 1;
        if ((uint64_t)np * S >= ((uint64_t)1 << 31)) S = 1;
        if (S > 1) {
            DCF_LAUNCH(ctx, "k_tm_lift", k_tm_lift0, dim3(nblk(V, 256)), dim3(256), 0, (const unsigned long long *)bw, lift, V);
            for (int j = 1; j < nl; j++) DCF_LAUNCH(ctx, "k_tm_lift", k_tm_lift, dim3(nblk(V, 256)), dim3(256), 0, lift, j, V);
        }
        DCF_LAUNCH(ctx, "k_tm_zip", k_tm_zip, dim3(nblk((int64_t)np * S, 256)), dim3(256), 0, zv, (const uint2 *)pairs, np, S,
                   (const uint32_t *)lift, nl, (uint32_t)V);
    }

    /* ---- classes, their parents and subtree sums ---- */
    uint32_t *cls = bl.get<uint32_t>((size_t)V), *top = bl.get<uint32_t>((size_t)V), *ci = bl.get<uint32_t>((size_t)V);
    if (!cls || !top || !ci) return dcf_set_error(ctx, DCF_ENOMEM, "classes");
    if (V) DCF_LAUNCH(ctx, "k_tm_cls", k_tm_cls, dim3(nblk(V, 256)), dim3(256), 0, zv, cls, top, bpar, V);
    uint32_t C = 0;
    if (V) {
        size_t need = 0;
        DCF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, top, ci, (int)V, ctx->stream));
        void *tmp = bl.get<char>(need);
        if (!tmp) return dcf_set_error(ctx, DCF_ENOMEM, "scan scratch");
        DCF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, need, top, ci, (int)V, ctx->stream));
        uint32_t last[2], fl = 0;
        DCF_HIP(ctx, hipMemcpyAsync(&last[0], ci + (V - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipMemcpyAsync(&last[1], top + (V - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipMemcpyAsync(&fl, fail, 4, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (fl) return dcf_set_error(ctx, DCF_EDEVICE, "boundary-tree merge did not converge (%u)", fl);
        C = last[0] + last[1];
    }
    DevAttr MA;
    if ((rc = dcf_alloc_attr(ctx, MA, grp, C)) != DCF_OK) return rc;
    AttrGuard ma_guard{ ctx, &MA };
    if (C) {
        rc = tm_group(grp, [&](auto g) -> int {
            DCF_LAUNCH(ctx, "k_tm_class_base", k_tm_class_base<decltype(g)::value>, dim3(nblk(V, 256)), dim3(256), 0, (const uint32_t *)cls,
                       BA, V);
            return DCF_OK;
        });
        if (rc) return rc;
        const uint32_t E = 2 * (C + 1);
        uint32_t *key = bl.get<uint32_t>(C), *val = bl.get<uint32_t>(C), *skey = bl.get<uint32_t>(C), *sval = bl.get<uint32_t>(C);
        uint32_t *pci = bl.get<uint32_t>(C), *fc = bl.get<uint32_t>((size_t)C + 1), *ns = bl.get<uint32_t>(C);
        uint32_t *top_of = bl.get<uint32_t>(C);
        uint32_t *n0 = bl.get<uint32_t>(E), *n1 = bl.get<uint32_t>(E), *r0 = bl.get<uint32_t>(E), *r1 = bl.get<uint32_t>(E);
        int64_t *vals = bl.get<int64_t>(E), *pre = bl.get<int64_t>(E);
        if (!key || !val || !skey || !sval || !pci || !fc || !ns || !top_of || !n0 || !n1 || !r0 || !r1 || !vals || !pre)
            return dcf_set_error(ctx, DCF_ENOMEM, "class tour");
        DCF_LAUNCH(ctx, "k_tm_class_par", k_tm_class_par, dim3(nblk(V, 256)), dim3(256), 0, (const uint32_t *)bpar,
                   (const uint32_t *)cls, (const uint32_t *)top, (const uint32_t *)ci, key, val, pci, C, V);
        DCF_LAUNCH(ctx, "k_tm_top_of", k_tm_top_of, dim3(nblk(V, 256)), dim3(256), 0, (const uint32_t *)top, (const uint32_t *)ci,
                   top_of, V);
        size_t need_sort = 0, need_scan = 0;
        DCF_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, need_sort, key, skey, val, sval, (int)C, 0, 32, ctx->stream));
        DCF_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, need_scan, vals, pre, (int)E, ctx->stream));
        void *tmp = bl.get<char>(std::max(need_sort, need_scan));
        if (!tmp) return dcf_set_error(ctx, DCF_ENOMEM, "class sort scratch");
        DCF_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, need_sort, key, skey, val, sval, (int)C, 0, 32, ctx->stream));
        DCF_LAUNCH(ctx, "k_fill", k_tm_fill32, dim3(nblk((int64_t)C + 1, 256)), dim3(256), 0, fc, TM_NONE, (int64_t)C + 1);
        DCF_LAUNCH(ctx, "k_tm_children", k_tm_children, dim3(nblk(C, 256)), dim3(256), 0, (const uint32_t *)skey, (const uint32_t *)sval,
                   fc, ns, C);
        DCF_LAUNCH(ctx, "k_tm_tour", k_tm_tour, dim3(nblk(E, 256)), dim3(256), 0, (const uint32_t *)fc, (const uint32_t *)ns,
                   (const uint32_t *)pci, n0, r0, C);
        int rounds = 1;
        while (rounds < 33 && ((uint64_t)1 << rounds) < E) rounds++;
        for (int k = 0; k < rounds; k++) {
            DCF_LAUNCH(ctx, "k_tm_rank", k_tm_rank, dim3(nblk(E, 256)), dim3(256), 0, (const uint32_t *)n0, (const uint32_t *)r0, n1, r1, E);
            std::swap(n0, n1);
            std::swap(r0, r1);
        }
        for (int f = 0; f <= BA.nsum; f++) {
            DCF_LAUNCH(ctx, "k_tm_place", k_tm_place, dim3(nblk(E, 256)), dim3(256), 0, (const uint32_t *)r0, (const uint32_t *)top_of, BA, f,
                       vals, E, fail);
            DCF_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp, need_scan, vals, pre, (int)E, ctx->stream));
            DCF_LAUNCH(ctx, "k_tm_subtree", k_tm_subtree, dim3(nblk(C, 256)), dim3(256), 0, (const uint32_t *)r0, (const int64_t *)pre,
                       (const uint32_t *)top_of, BA, MA, f, C);
        }
        if (BA.has_ext) {
            /* extents: range minima / maxima over each class's span of the tour */
            int L = 1;
            while (((uint64_t)1 << L) <= E) L++;
            uint32_t *st = bl.get<uint32_t>((size_t)E * L);
            if (!st) return dcf_set_error(ctx, DCF_ENOMEM, "extent table");
            for (int f = 0; f < 6; f++) {
                DCF_LAUNCH(ctx, "k_tm_place_ext", k_tm_place_ext, dim3(nblk(E, 256)), dim3(256), 0, (const uint32_t *)r0,
                           (const uint32_t *)top_of, BA, f, st, E);
                for (int j = 1; j < L; j++)
                    DCF_LAUNCH(ctx, "k_tm_st_level", k_tm_st_level, dim3(nblk(E, 256)), dim3(256), 0, st, E, j, f >= 3 ? 1 : 0);
                DCF_LAUNCH(ctx, "k_tm_st_query", k_tm_st_query, dim3(nblk(C, 256)), dim3(256), 0, (const uint32_t *)r0,
                           (const uint32_t *)st, MA, f, C);
            }
        }
    }

    {
        uint32_t fl = 0;
        DCF_HIP(ctx, hipMemcpyAsync(&fl, fail, 4, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (fl) return dcf_set_error(ctx, DCF_EDEVICE, "boundary-tree merge failed (%u)", fl);
    }

    /* ---- the whole tree ---- */
    DCF_LAUNCH(ctx, "k_tm_parent", k_tm_parent, dim3(nblk(n, 256)), dim3(256), 0, (const TileDev *)dtiles, G, (const uint32_t *)cls,
               (const uint32_t *)bpar, (const uint32_t *)blr, T->parent, n);
    uint32_t *flag = nullptr;
    if ((rc = dcf_tree_nodes_from_parent(T, &flag)) != DCF_OK) return rc;
    struct FlagGuard {
        dcf_ctx *c;
        uint32_t *f;
        ~FlagGuard() { dcf_release(c, f); }
    } flag_guard{ ctx, flag };
    dcf_free_attr(ctx, T->attr);
    if ((rc = dcf_alloc_attr(ctx, T->attr, grp, T->K)) != DCF_OK) return rc;
    rc = tm_group(grp, [&](auto g) -> int {
        DCF_LAUNCH(ctx, "k_tm_attr_out", k_tm_attr_out<decltype(g)::value>, dim3(nblk(n, 256)), dim3(256), 0, (const TileDev *)dtiles, G,
                   (const uint32_t *)flag, (const uint32_t *)T->p2n, (const uint32_t *)cls, (const uint32_t *)ci, MA, T->attr, n);
        return DCF_OK;
    });
    if (rc) return rc;
    T->grp = grp;
    /* the kernels above read the tile trees and the blocks released on return: wait for them */
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DCF_OK;
}
