/*
 * dcf_flood.hip -- level-synchronous union-find construction of the canonical max-tree with its
 * attributes, for u8, u16 and (through dense ranks of their float keys) f32 images.
 *
 * The reference floods pixels one at a time in decreasing (grey, index) order and merges the
 * attribute of every finished pixel into its parent (tree_flood, src/sources/flood.c:222-247;
 * the same order as flood_attributes, flood.c:184-219).  Here all pixels of one grey level are
 * processed together:
 *
 *   k_union   every pixel of level k is united with its already-processed neighbours (level > k,
 *             or level k with a larger index) in a lock-free union-find.  A set of level-k pixels
 *             that reaches an old set (a root from a level > k) is linked under the old root, so a
 *             thread only ever CASes its own set's root entry; between new sets the smaller index
 *             wins, so a set with no old root is rooted at its minimum pixel, the canonical
 *             levelroot.  Old roots linked under another root are pushed on the `absorbed` list.
 *   k_final_pixels  every pixel of level k finds its root.  A root of level k is canonical: the
 *             pixel takes it as parent and adds its attribute contribution to it.  A set rooted at
 *             an old root r gets its canonical levelroot c by atomicMin; the contributions of its
 *             pixels and the complete records of the old sets absorbed into it are kept pending
 *             at r, and their parents are left pending on the set's current node nrep[r].
 *   k_fix     every such set gets its node c: nrep[r] hangs below c, and c's record becomes
 *             nrep[r]'s record plus what is pending at r.  r stays the set's union-find root (a
 *             set's root changes only when the set is merged into an older one), so paths do not
 *             grow with the levels a set gains.  After the pass, k_held_resolve resolves the
 *             pending parents.
 * Attributes are therefore accumulated level by level, exactly like the reference's merges during
 * flooding, with no separate leaffix.
 *
 * The critical path is the number of levels, not the depth of the tree (u16 trees of smooth
 * images are ~50k nodes deep).  Keys are taken in 8-bit digits, one pass of 256 levels per digit
 * (u8 one pass, u16 two, float ranks up to four).  For 16-bit keys:
 *   pass 1  levels = key >> 8: the coarse tree (coarse levelroot cl(p), coarse parent, and the
 *           exact attributes of every coarse node region);
 *   pass 2  levels = key & 255 inside every coarse node N: its level region LR(N) = {p : cl(p) = N}
 *           is refined, and every coarse child C of N enters N's union-find as ONE super-vertex
 *           S_C (priority above all of LR(N), attributes = the totals of C's region).  A pixel of
 *           LR(N) adjacent to a pixel q of a child region is united with S_C, C being the
 *           ancestor of cl(q) whose coarse parent is N (binary lifting on the coarse tree).
 *           When S_C is absorbed at fine level d, the fine node it joins is the parent of C's
 *           lowest fine node (the final root of LR(C)).  The child is found by depth: it is cl(q)'s
 *           ancestor at depth(N) + 1.  Since a region's sets hold only its own pixels and its
 *           children's super-vertices, the regions refine independently: a large image's pass 2
 *           runs as several level loops (regions dealt by node id), each on a stream of its own.
 * A further pass repeats pass 2 one digit lower, with the tree of the previous pass (levels =
 * key prefixes) as the coarse tree.  Pixels with equal keys are merged by the union-find itself
 * (no chains), so flat zones cost nothing extra.
 *
 * Batches of small images can instead be built whole in LDS, one workgroup per image
 * (k_small_build, near the end of the kernels).
 */
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "dcf_internal.h"

#define NLEV 256

__device__ __forceinline__ uint32_t fl_key(const ImgView &iv, int64_t p)
{
    if (iv.dtype == DCF_KEY32) return ((const uint32_t *)iv.img)[p];
    if (iv.dtype == DCF_U8) {
        uint32_t v = ((const uint8_t *)iv.img)[p];
        return iv.invert ? (uint32_t)(*iv.gmax) - v : v;
    }
    uint32_t v = ((const uint16_t *)iv.img)[p];
    return iv.invert ? (uint32_t)(*iv.gmax) - v : v;
}

__device__ __forceinline__ float fl_grey(const ImgView &iv, int64_t p) { return (float)fl_key(iv, p); }

struct Nbr {
    int n;
    int8_t dx[26], dy[26], dz[26];
};

struct Geo {
    int W, H, D;
    int64_t n;
};

/* ------------------------------------------------------------------------------------------------ */
/* union-find                                                                                       */
/* ------------------------------------------------------------------------------------------------ */
/* uf[x] is the parent of x, or for a root UF_ROOT | level(x), so that a find returns the root's level
 * with the load that detects the root (no image read after it).  Element ids stay below 2^31. */
#define UF_ROOT 0x80000000u
#define UF_LVL 0x1FFu

/* the union-find entries: element x's entry is p[x * s].  s = 1 in a first pass; in a refining pass
 * every element has a 16-byte record {entry, key, coarse node, 0}, so that one line of a neighbour
 * row holds all three for the k_union neighbour scan (s = 4). */
struct UFV {
    uint32_t *p;
    uint32_t s;
};

/* workgroup-scope load: may be served by the CU's L1 and be stale.  Every value an entry ever holds
 * is an ancestor of the element (links only turn a root into a child, compression only moves a
 * pointer up), so a stale read still leads towards the root; a stale root tag is caught by the
 * link's CAS, which is performed in L2 and returns the fresh parent.  The root of a big component
 * is read by every thread that touches it, and this keeps those reads in the CUs' L1s. */
__device__ __forceinline__ uint32_t ldc(const UFV &u, uint32_t x)
{
    return __hip_atomic_load(&u.p[(size_t)x * u.s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void st(const UFV &u, uint32_t x, uint32_t v)
{
    __hip_atomic_store(&u.p[(size_t)x * u.s], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* root of x with path splitting; tag = the root's entry.  Only non-root entries are rewritten, and
 * only ever to an ancestor, so these stores cannot race with a link (links CAS root tags). */
__device__ __forceinline__ uint32_t uf_find(const UFV &uf, uint32_t x, uint32_t &tag)
{
    uint32_t p = ldc(uf, x);
    if (p & UF_ROOT) {
        tag = p;
        return x;
    }
    for (;;) {
        const uint32_t gp = ldc(uf, p);
        if (gp & UF_ROOT) {
            tag = gp;
            return p;
        }
        st(uf, x, gp); /* splitting */
        x = p;
        p = gp;
    }
}

/* the root reached from x whose (fresh) parent is known to be p */
__device__ __forceinline__ uint32_t uf_find_from(const UFV &uf, uint32_t x, uint32_t p, uint32_t &tag)
{
    for (;;) {
        const uint32_t gp = ldc(uf, p);
        if (gp & UF_ROOT) {
            tag = gp;
            return p;
        }
        st(uf, x, gp);
        x = p;
        p = gp;
    }
}

/* block-local list of the old roots linked by this block's threads, flushed with one global
 * atomic per block */
struct AbsBuf {
    uint32_t *buf;
    uint32_t *n;
};

/* unite the sets of roots a (tag ta) and b (tag tb) during the level `lev`.  A set that reaches a
 * completed (old) set is linked under that set's root, so the threads of a band of level-`lev`
 * pixels along a big old component each CAS their own root's entry and never a shared one; old
 * roots merge among themselves by (level, id), current-level roots by id.  A linked old root is
 * recorded.  The node of a set that contains old roots (its levelroot, the minimum-index pixel of
 * level `lev`) is found after the unions (k_final_pixels, k_fix); the set keeps its old root.
 * Returns the surviving root (possibly stale) and its tag. */
__device__ __forceinline__ uint32_t uf_unite(const UFV &uf, uint32_t a, uint32_t ta, uint32_t b, uint32_t tb, uint32_t lev,
                                             const AbsBuf &ab, uint32_t &tout)
{
    for (;;) {
        if (a == b) {
            tout = ta;
            return a;
        }
        uint32_t la = ta & UF_LVL, lb = tb & UF_LVL;
        /* elements of the current level go under completed (older) sets; old sets merge by
         * (level, id) and current-level sets by id */
        const bool an = la == lev, bn = lb == lev;
        bool a_hi = an != bn ? an : (la > lb || (la == lb && a > b));
        uint32_t hi = a_hi ? a : b, lo = a_hi ? b : a;
        uint32_t th = a_hi ? ta : tb, tl = a_hi ? tb : ta;
        uint32_t exp = th;
        if (__hip_atomic_compare_exchange_strong(&uf.p[(size_t)hi * uf.s], &exp, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            if ((th & UF_LVL) > lev) ab.buf[atomicAdd(ab.n, 1u)] = hi;
            tout = tl;
            return lo;
        }
        /* hi has been linked meanwhile: exp is its fresh parent; lo may be stale as well */
        if (a_hi) {
            a = uf_find_from(uf, a, exp, ta);
            b = uf_find(uf, b, tb);
        } else {
            b = uf_find_from(uf, b, exp, tb);
            a = uf_find(uf, a, ta);
        }
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* counting sort of the pixels by level digit                                                       */
/* ------------------------------------------------------------------------------------------------ */
/* Refining passes may run as K independent level loops (1 <= K <= FINE_LOOPS_MAX): a coarse region's sets
 * hold its own pixels and its children's super-vertices only, so the regions refine independently, and a
 * pixel belongs to loop cnode[p] % K.  Its bin in the level lists is then loop * NLEV + level digit. */
#define FINE_LOOPS_MAX 4
__device__ __forceinline__ uint32_t lvl_bin(uint32_t d, int64_t p, const uint32_t *cnode, uint32_t K)
{
    return K > 1 ? (cnode[p] % K) * NLEV + d : d;
}

/* only != nullptr: count only the pixels with only[p] == p (zone representatives) */
__global__ __launch_bounds__(256) void k_lvl_hist(ImgView iv, int pass, int shift, int64_t n, const uint32_t *__restrict__ only,
                                                  uint32_t *__restrict__ hist, const uint32_t *__restrict__ cnode, uint32_t K)
{
    __shared__ uint32_t sh[NLEV * FINE_LOOPS_MAX];
    const uint32_t nbins = NLEV * K;
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) sh[b] = 0;
    __syncthreads();
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (only && only[p] != (uint32_t)p) continue;
        uint32_t k = fl_key(iv, p);
        uint32_t d = (k >> shift) & 255u;
        atomicAdd(&sh[lvl_bin(d, p, cnode, K)], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x)
        if (sh[b]) atomicAdd(&hist[b], sh[b]);
}

struct TileGeo {
    int W, H, D;
    int TW, TH, TD;
    int ntx, nty;
    const BatchImg *bimg; /* a batch: every image has its own tiles, from its tile0 on */
    int nbatch;
};

/* tile b: its origin in its image, the image's dimensions and first pixel */
struct TileAt {
    int x0, y0, z0;
    int W, H, D;
    int64_t off;
};

__device__ __forceinline__ TileAt tile_at(const TileGeo &tg, int b)
{
    TileAt t;
    int ntx = tg.ntx, nty = tg.nty;
    if (tg.bimg) {
        int lo = 0, hi = tg.nbatch - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((int)tg.bimg[mid].tile0 <= b) lo = mid;
            else hi = mid - 1;
        }
        const BatchImg im = tg.bimg[lo];
        t.W = (int)im.W; t.H = (int)im.H; t.D = (int)im.D;
        t.off = (int64_t)im.off;
        b -= (int)im.tile0;
        ntx = (t.W + tg.TW - 1) / tg.TW;
        nty = (t.H + tg.TH - 1) / tg.TH;
    } else {
        t.W = tg.W; t.H = tg.H; t.D = tg.D;
        t.off = 0;
    }
    t.x0 = (b % ntx) * tg.TW;
    t.y0 = ((b / ntx) % nty) * tg.TH;
    t.z0 = (b / (ntx * nty)) * tg.TD;
    return t;
}

/* each block owns one tile of the image (64x64 in 2D, 16x16x16 in 3D; in a batch, one of its image's
 * zone tiles); one global atomic per (block, level) reserves the block's slots, so every level's list
 * is ordered by tile and the pixels a block of a level kernel handles lie close together */
struct ScatterGeo {
    int batch;          /* 0: one image's tiles (in groups), 1: a batch's tiles */
    const BatchImg *bimg;
    int nbatch;
    int W, H, D;
    int TW, TH, TD;     /* tile shape */
    int GX, GY, GZ;     /* tiles per group: a group's tiles are consecutive in the list */
    int ngx, ngy;       /* groups per row / column */
    int ntx, nty, ntz;  /* tiles per row / column / slab */
};


/* Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8), and each XCD has its own
 * L2.  This renumbering gives every XCD one contiguous eighth of the level list, i.e. one compact
 * part of the image, so neighbouring pixels share that XCD's L2. */
__device__ __forceinline__ uint32_t xcd_map(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
    return x * q + min(x, r) + k;
}

__device__ __forceinline__ uint32_t xcd_block() { return xcd_map(blockIdx.x, gridDim.x); }

/* ------------------------------------------------------------------------------------------------ */
/* per-level kernels                                                                                */
/* ------------------------------------------------------------------------------------------------ */
struct LevelArgs {
    ImgView iv;
    Geo g;
    Nbr nb;
    int pass, shift;
    UFV uf;
    uint32_t *nrep;        /* per set root: its set's current node (levelroot pixel), NONE for a super-vertex
                              whose set has no node in this pass yet */
    uint32_t *absorbed;
    uint32_t *nabs;        /* [0] running count */
    uint32_t *canon;       /* per element: minimum pixel of the current level in its set (sets rooted at an old root) */
    uint32_t *fix;         /* old set roots that get a node, all levels of the pass back to back */
    uint32_t *fcount;      /* [0] running count of fix, [1 + i] where level i's entries start */
    /* pass 2 only */
    const uint32_t *cnode; /* coarse node id of every pixel */
    const int32_t *lift;   /* nlift x Kc: 2^j-th coarse ancestor (node id), -1 above the root */
    int nlift;             /* 2^nlift exceeds the depth of the coarse tree */
    const uint32_t *cnkey; /* coarse key of every coarse node */
    const uint2 *cpd;       /* every coarse node's parent (x, -1 at a root) and depth (y, a root: 0), side by side so a
                             * lookup's first step fetches one line */
    /* batched form: image table (nullptr for one image) */
    const BatchImg *bimg;
    int nbatch;
    /* zone pass: the level lists hold zone representatives, whose records carry their zone's totals */
    int zone;
    /* coordinate offset of the image in a larger one (a tile of a tiled build): the attribute sums use
     * x + ox, y + oy, z + oz, as the reference adds its tile offsets (attributes.c:1219) */
    int ox, oy, oz;
};

/* local coordinates of pixel p and the dimensions of its image (pixel ids are below 2^30, so the
 * local index fits 32-bit division) */
struct PixGeo {
    int x, y, z;
    int W, H, D;
};

__device__ __forceinline__ PixGeo pix_geo(const LevelArgs &a, uint32_t p)
{
    PixGeo g;
    uint32_t local = p;
    if (a.bimg) {
        const BatchImg b = a.bimg[batch_locate(a.bimg, a.nbatch, p)];
        local = p - (uint32_t)b.off;
        g.W = (int)b.W; g.H = (int)b.H; g.D = (int)b.D;
    } else {
        g.W = a.g.W; g.H = a.g.H; g.D = a.g.D;
    }
    const uint32_t row = local / (uint32_t)g.W;
    g.x = (int)(local - row * (uint32_t)g.W);
    g.z = (int)(row / (uint32_t)g.H);
    g.y = (int)(row - (uint32_t)g.z * (uint32_t)g.H);
    return g;
}

/* the child of coarse node N (id nid, depth dn) on the path up from its descendant e: e itself when
 * its coarse parent is N, else e's ancestor one level below N, reached by jumping up depth(e) - dn - 1
 * levels in one 2^j step per set bit (binary lifting) */
__device__ __forceinline__ uint32_t child_toward(const LevelArgs &a, uint32_t e, uint32_t nid, uint32_t dn, uint32_t kc_count)
{
    const uint2 pd = a.cpd[e];
    const int32_t up = (int32_t)pd.x;
    const uint32_t de = pd.y;
    if (up == (int32_t)nid) return e;
    e = (uint32_t)up;
    uint32_t rest = de - dn - 2; /* levels still to climb above up */
    for (int j = 0; rest && j < a.nlift; j++, rest >>= 1)
        if (rest & 1u) {
            const int32_t u = a.lift[(int64_t)j * kc_count + e];
            if (u >= 0) e = (uint32_t)u;
        }
    return e;
}

/* depth of every coarse node: the sum of the 2^j jumps that stay inside the tree, largest first
 * (2^nlift exceeds the depth) */
__global__ __launch_bounds__(256) void k_cdepth(const int32_t *__restrict__ lift, int nlift, int64_t kc, uint2 *__restrict__ cpd)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kc) return;
    uint32_t e = (uint32_t)c, d = 0;
    for (int j = nlift - 1; j >= 0; j--) {
        const int32_t u = lift[(int64_t)j * kc + e];
        if (u >= 0) {
            d += 1u << j;
            e = (uint32_t)u;
        }
    }
    cpd[c] = make_uint2((uint32_t)lift[c], d);
}

#define CH 8 /* neighbours gathered together */

/* union targets of pixel p at its level lev: every already-processed neighbour (level digit above
 * lev, or equal with a larger index); in a refining pass a neighbour in a coarse child region stands
 * for that child's super-vertex (duplicates of one super-vertex are dropped).  NONE marks no target. */
template <int NB>
__device__ __forceinline__ void pixel_targets(const LevelArgs &a, const PixGeo &pg, uint32_t p, uint32_t lev, uint32_t kc_count,
                                              uint32_t (&t)[NB])
{
    const uint32_t NONE = 0xFFFFFFFFu;
    const int WH = pg.W * pg.H;
    const bool rec = a.uf.s == 4; /* key and coarse node in the union-find record */
    const uint32_t kp = rec ? a.uf.p[(size_t)p * 4 + 1] : fl_key(a.iv, p);
    const uint32_t cp = a.pass == 2 ? kp >> (a.shift + 8) : 0u;
    const uint32_t nid = a.pass == 2 ? (rec ? a.uf.p[(size_t)p * 4 + 2] : a.cnode[p]) : 0u;
    /* every neighbour's key (and in a refining pass its coarse node) is loaded before any is looked at, so the
     * loads are in flight together; a neighbour outside the image reads the pixel itself and is skipped */
    uint32_t qs[NB], kqs[NB], cns[NB];
    bool inb[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const int nx = pg.x + a.nb.dx[k], ny = pg.y + a.nb.dy[k], nz = pg.z + a.nb.dz[k];
        inb[k] = nx >= 0 && nx < pg.W && ny >= 0 && ny < pg.H && nz >= 0 && nz < pg.D;
        qs[k] = inb[k] ? (uint32_t)((int)p + a.nb.dx[k] + a.nb.dy[k] * pg.W + a.nb.dz[k] * WH) : p;
        kqs[k] = rec ? a.uf.p[(size_t)qs[k] * 4 + 1] : fl_key(a.iv, qs[k]);
        cns[k] = a.pass == 2 ? (rec ? a.uf.p[(size_t)qs[k] * 4 + 2] : a.cnode[qs[k]]) : 0u;
    }
    /* neighbours in one child region mostly share their coarse node: its child of nid is looked up once */
    uint32_t dn = NONE, last_node = NONE, last_sv = NONE;
#pragma unroll
    for (int k = 0; k < NB; k++) {
        t[k] = NONE;
        if (!inb[k]) continue;
        const uint32_t q = qs[k];
        const uint32_t kq = kqs[k];
        const uint32_t dq = (kq >> a.shift) & 255u;
        if (a.pass == 1) {
            if (dq > lev || (dq == lev && q > p)) t[k] = q;
        } else {
            const uint32_t cq = kq >> (a.shift + 8);
            if (cq == cp) {
                if (dq > lev || (dq == lev && q > p)) t[k] = q;
            } else if (cq > cp) {
                const uint32_t cq_node = cns[k];
                if (cq_node != last_node) {
                    if (dn == NONE) dn = a.cpd[nid].y;
                    last_node = cq_node;
                    last_sv = (uint32_t)a.g.n + child_toward(a, cq_node, nid, dn, kc_count);
                }
                const uint32_t sv = last_sv;
                bool dup = false;
#pragma unroll
                for (int m = 0; m < k; m++) dup = dup || t[m] == sv;
                if (!dup) t[k] = sv;
            }
        }
    }
}

/* p's set (root rp, tag tp) is united with the set of every target in cur: the targets' roots are
 * found together (independent loads in flight, path splitting), duplicates dropped */
/* the first entries of a chunk of targets, loaded before they are needed (their finds and CASes
 * tolerate entries that have changed since, as they must for any concurrent update) */
__device__ __forceinline__ void chunk_entries(const LevelArgs &a, const uint32_t (&cur)[CH], uint32_t (&val)[CH])
{
    const uint32_t NONE = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < CH; j++) val[j] = cur[j] != NONE ? ldc(a.uf, cur[j]) : UF_ROOT;
}

template <bool PRE = false>
__device__ __forceinline__ uint32_t unite_chunk(const LevelArgs &a, const uint32_t (&cur)[CH], uint32_t rp, uint32_t &tp, uint32_t lev,
                                                const AbsBuf &ab, const uint32_t *pre = nullptr)
{
    const uint32_t NONE = 0xFFFFFFFFu;
    uint32_t val[CH], node[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) {
        node[j] = cur[j];
        if constexpr (PRE) val[j] = pre[j];
        else val[j] = cur[j] != NONE ? ldc(a.uf, cur[j]) : UF_ROOT;
    }
    bool more = true;
    while (more) {
        more = false;
#pragma unroll
        for (int j = 0; j < CH; j++)
            if (!(val[j] & UF_ROOT)) {
                const uint32_t pj = val[j];
                const uint32_t gp = ldc(a.uf, pj);
                if (!(gp & UF_ROOT)) st(a.uf, node[j], gp);
                node[j] = pj;
                val[j] = gp;
                more |= !(gp & UF_ROOT);
            }
    }
#pragma unroll
    for (int j = 0; j < CH; j++) {
        bool live = cur[j] != NONE;
#pragma unroll
        for (int m = 0; m < j; m++) live = live && !(cur[m] != NONE && node[m] == node[j]);
        if (live) rp = uf_unite(a.uf, rp, tp, node[j], val[j], lev, ab, tp);
    }
    return rp;
}

/* the block's absorbed old roots, appended to the level's list */
__device__ __forceinline__ void flush_absorbed(const LevelArgs &a, const uint32_t *s_abs, uint32_t &s_n, uint32_t &s_base)
{
    __syncthreads();
    if (threadIdx.x == 0 && s_n) s_base = atomicAdd(a.nabs, s_n);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < s_n; j += blockDim.x) a.absorbed[s_base + j] = s_abs[j];
}

/* every pixel p of the level is united with each already-processed neighbour (pixel_targets,
 * evaluated here from the keys) */
template <int NB>
__global__ __launch_bounds__(256) void k_union(LevelArgs a, const uint32_t *__restrict__ lvl, uint32_t cnt, uint32_t lev,
                                               uint32_t kc_count)
{
    __shared__ uint32_t s_abs[256 * NB];
    __shared__ uint32_t s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const AbsBuf ab{ s_abs, &s_n };
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t NONE = 0xFFFFFFFFu;
    if (i < cnt) {
        const uint32_t p = lvl[i];
        uint32_t t[NB];
        pixel_targets<NB>(a, pix_geo(a, p), p, lev, kc_count, t);
        uint32_t tp = 0, rp = uf_find(a.uf, p, tp);
#pragma unroll
        for (int k0 = 0; k0 < NB; k0 += CH) {
            uint32_t cur[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) cur[j] = k0 + j < NB ? t[k0 + j] : NONE;
            rp = unite_chunk(a, cur, rp, tp, lev, ab);
        }
    }
    flush_absorbed(a, s_abs, s_n, s_base);
}

/* the same with the targets read from the level's target slots (NB per list position, written by
 * k_lvl_scatter<NB>): no key, coarse-node or lifting loads at the level */
template <int NB>
__global__ __launch_bounds__(256) void k_union_tgt(LevelArgs a, const uint32_t *__restrict__ lvl, const uint32_t *__restrict__ tgt,
                                                   uint32_t cnt, uint32_t lev)
{
    __shared__ uint32_t s_abs[256 * NB];
    __shared__ uint32_t s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const AbsBuf ab{ s_abs, &s_n };
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    const uint32_t NONE = 0xFFFFFFFFu;
    if (i < cnt) {
        const uint32_t p = lvl[i];
        uint32_t t[NB];
        const uint32_t *src = tgt + (size_t)i * NB;
        if (NB % 4 == 0) {
#pragma unroll
            for (int k = 0; k < NB; k += 4) {
                const uint4 v = *(const uint4 *)(src + k);
                t[k] = v.x; t[k + 1] = v.y; t[k + 2] = v.z; t[k + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < NB; k += 2) {
                const uint2 v = *(const uint2 *)(src + k);
                t[k] = v.x; t[k + 1] = v.y;
            }
        }
        /* the first chunk's entries are loaded together with p's own */
        uint32_t cur0[CH], pre[CH];
#pragma unroll
        for (int j = 0; j < CH; j++) cur0[j] = j < NB ? t[j] : NONE;
        chunk_entries(a, cur0, pre);
        uint32_t tp = 0, rp = uf_find(a.uf, p, tp);
        rp = unite_chunk<true>(a, cur0, rp, tp, lev, ab, pre);
#pragma unroll
        for (int k0 = CH; k0 < NB; k0 += CH) {
            uint32_t cur[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) cur[j] = k0 + j < NB ? t[k0 + j] : NONE;
            rp = unite_chunk(a, cur, rp, tp, lev, ab);
        }
    }
    flush_absorbed(a, s_abs, s_n, s_base);
}

/* NB > 0: also the union targets of every listed pixel, NB slots at its list position (tgt) */
template <int NB>
__global__ __launch_bounds__(256) void k_lvl_scatter(ImgView iv, int pass, int shift, int64_t n, ScatterGeo sg,
                                                     const uint32_t *__restrict__ only, uint32_t *__restrict__ cursor,
                                                     uint32_t *__restrict__ order, LevelArgs a, uint32_t kc_count,
                                                     uint32_t *__restrict__ tgt, uint32_t K)
{
    __shared__ uint32_t cnt[NLEV * FINE_LOOPS_MAX], base[NLEV * FINE_LOOPS_MAX];
    const uint32_t nbins = NLEV * K;
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) cnt[b] = 0;
    __syncthreads();
    const int tsz = sg.TW * sg.TH * sg.TD;
    int64_t p0 = 0;
    int x0 = 0, y0 = 0, z0 = 0;
    int iW = sg.W, iH = sg.H, iD = sg.D;
    if (sg.batch) {
        TileGeo tg;
        tg.TW = sg.TW; tg.TH = sg.TH; tg.TD = sg.TD;
        tg.bimg = sg.bimg;
        tg.nbatch = sg.nbatch;
        const TileAt ta = tile_at(tg, blockIdx.x);
        x0 = ta.x0; y0 = ta.y0; z0 = ta.z0;
        iW = ta.W; iH = ta.H; iD = ta.D;
        p0 = ta.off;
    } else {
        const int gsz = sg.GX * sg.GY * sg.GZ;
        const int g = blockIdx.x / gsz, w = blockIdx.x % gsz;
        const int gx = g % sg.ngx, gy = (g / sg.ngx) % sg.ngy, gz = g / (sg.ngx * sg.ngy);
        const int tx = gx * sg.GX + w % sg.GX, ty = gy * sg.GY + (w / sg.GX) % sg.GY, tz = gz * sg.GZ + w / (sg.GX * sg.GY);
        x0 = tx * sg.TW; y0 = ty * sg.TH; z0 = tz * sg.TD;
        /* tiles of a partial group beyond the image: the block has no pixel */
        if (tx >= sg.ntx || ty >= sg.nty || tz >= sg.ntz) x0 = sg.W;
    }
    /* pixel of local slot l of this block, or -1 outside the image */
    auto pix = [&](int l) -> int64_t {
        const int x = x0 + l % sg.TW, y = y0 + (l / sg.TW) % sg.TH, z = z0 + l / (sg.TW * sg.TH);
        if (x >= iW || y >= iH || z >= iD) return -1;
        const int64_t p = p0 + (int64_t)x + (int64_t)y * iW + (int64_t)z * iW * iH;
        return (only && only[p] != (uint32_t)p) ? -1 : p;
    };
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        const int64_t p = pix(l);
        if (p < 0) continue;
        uint32_t k = fl_key(iv, p);
        atomicAdd(&cnt[lvl_bin((k >> shift) & 255u, p, a.cnode, K)], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) {
        if (cnt[b]) base[b] = atomicAdd(&cursor[b], cnt[b]);
        cnt[b] = 0;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        const int64_t p = pix(l);
        if (p < 0) continue;
        uint32_t k = fl_key(iv, p);
        uint32_t d = (k >> shift) & 255u;
        const uint32_t bin = lvl_bin(d, p, a.cnode, K);
        uint32_t r = atomicAdd(&cnt[bin], 1u);
        order[base[bin] + r] = (uint32_t)p;
        if constexpr (NB > 0) {
            uint32_t t[NB];
            /* the block's tile knows its image: no lookup in the image table */
            const int l = (int)(((int64_t)p - p0) % ((int64_t)iW * iH));
            const int zz = (int)(((int64_t)p - p0) / ((int64_t)iW * iH));
            const PixGeo pg{ l % iW, l / iW, zz, iW, iH, iD };
            pixel_targets<NB>(a, pg, (uint32_t)p, d, kc_count, t);
            uint32_t *dst = tgt + (size_t)(base[bin] + r) * NB;
            if (NB % 4 == 0) {
#pragma unroll
                for (int k = 0; k < NB; k += 4) *(uint4 *)(dst + k) = make_uint4(t[k], t[k + 1], t[k + 2], t[k + 3]);
            } else {
#pragma unroll
                for (int k = 0; k < NB; k += 2) *(uint2 *)(dst + k) = make_uint2(t[k], t[k + 1]);
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* zone graph of a pass (coarse pass of u16 images, u8 planes)                                      */
/* ------------------------------------------------------------------------------------------------ */
/* A tile zone is a connected set of pixels of one digit inside one tile.  Its pixels are linked to
 * its representative (its minimum pixel, the minimum raster index of the zone) before the pass, so
 * the union-find starts with every zone as one set; the unions a level needs are then the distinct
 * (zone, neighbouring zone) pairs, de-duplicated per tile in LDS, instead of one attempt per pixel
 * and neighbour.  The canonical levelroot of a set is its minimum pixel of the level, which is its
 * minimum zone representative of the level, so the level kernels work unchanged. */
#define ZT_MAX ZONE_TILE_MAX
#define ZH_CAP 4096
#define ZH_EMPTY 0xFFFFFFFFFFFFFFFFull


__device__ __forceinline__ uint32_t lds_ld(const uint32_t *a, uint32_t i)
{
    return __hip_atomic_load(&a[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t lds_find(uint32_t *par, uint32_t x)
{
    uint32_t p = lds_ld(par, x);
    while (p != x) {
        const uint32_t gp = lds_ld(par, p);
        if (gp != p) __hip_atomic_store(&par[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ void lds_union(uint32_t *par, uint32_t a, uint32_t b)
{
    for (;;) {
        a = lds_find(par, a);
        b = lds_find(par, b);
        if (a == b) return;
        const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
        uint32_t exp = hi;
        if (__hip_atomic_compare_exchange_strong(&par[hi], &exp, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP))
            return;
    }
}

/* zones of every tile: zrep[p] = the zone's minimum pixel; uf[p] -> zrep for the other pixels */
__global__ __launch_bounds__(256) void k_tile_zones(ImgView iv, int shift, TileGeo tg, Nbr nb, uint32_t *__restrict__ zrep,
                                                    UFV uf, uint32_t *__restrict__ zarea)
{
    __shared__ uint32_t par[ZT_MAX];
    __shared__ uint32_t zc[ZT_MAX];
    __shared__ uint16_t dg[ZT_MAX];
    const TileAt ta = tile_at(tg, blockIdx.x);
    const int x0 = ta.x0, y0 = ta.y0, z0 = ta.z0;
    const int64_t W = ta.W, WH = (int64_t)ta.W * ta.H;
    const int tsz = tg.TW * tg.TH * tg.TD;
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        const int x = x0 + l % tg.TW, y = y0 + (l / tg.TW) % tg.TH, z = z0 + l / (tg.TW * tg.TH);
        const bool in = x < ta.W && y < ta.H && z < ta.D;
        dg[l] = in ? (uint16_t)(fl_key(iv, ta.off + x + y * W + z * WH) >> shift) : 0xFFFFu;
        par[l] = (uint32_t)l;
        zc[l] = 0;
    }
    __syncthreads();
    /* 2D 8-connectivity: a diagonal neighbour of the same digit is reached through the pixel below (or, for the
     * lower right one, the pixel to the right) when that one has the digit too, by unions made anyway (the
     * right and lower neighbours are always linked), so only the other diagonal unions are needed */
    const bool nb8 = nb.n == 8;
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        if (dg[l] == 0xFFFFu) continue;
        const int lx = l % tg.TW, ly = (l / tg.TW) % tg.TH, lz = l / (tg.TW * tg.TH);
        if (nb8) {
            const uint16_t d0 = dg[l];
            const bool hr = lx + 1 < tg.TW, hd = ly + 1 < tg.TH;
            const bool sr = hr && dg[l + 1] == d0, sd = hd && dg[l + tg.TW] == d0;
            if (sr) lds_union(par, (uint32_t)l, (uint32_t)(l + 1));
            if (sd) lds_union(par, (uint32_t)l, (uint32_t)(l + tg.TW));
            if (hd && lx > 0 && !sd && dg[l + tg.TW - 1] == d0) lds_union(par, (uint32_t)l, (uint32_t)(l + tg.TW - 1));
            if (hd && hr && !sd && !sr && dg[l + tg.TW + 1] == d0) lds_union(par, (uint32_t)l, (uint32_t)(l + tg.TW + 1));
            continue;
        }
        for (int k = 0; k < nb.n; k++) {
            /* the forward half of the neighbourhood: every pair once */
            const int dx = nb.dx[k], dy = nb.dy[k], dz = nb.dz[k];
            if (!(dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0))))) continue;
            const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
            if (nx < 0 || nx >= tg.TW || ny < 0 || ny >= tg.TH || nz < 0 || nz >= tg.TD) continue;
            const int l2 = nx + ny * tg.TW + nz * tg.TW * tg.TH;
            if (dg[l2] == dg[l]) lds_union(par, (uint32_t)l, (uint32_t)l2);
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        if (dg[l] == 0xFFFFu) continue;
        const int r = (int)lds_find(par, (uint32_t)l);
        const int x = x0 + l % tg.TW, y = y0 + (l / tg.TW) % tg.TH, z = z0 + l / (tg.TW * tg.TH);
        const int rx = x0 + r % tg.TW, ry = y0 + (r / tg.TW) % tg.TH, rz = z0 + r / (tg.TW * tg.TH);
        const uint32_t p = (uint32_t)(ta.off + x + y * W + z * WH);
        const uint32_t rp = (uint32_t)(ta.off + rx + ry * W + rz * WH);
        zrep[p] = rp;
        if (p != rp) uf.p[(size_t)p * uf.s] = rp;
        atomicAdd(&zc[r], 1u);
    }
    __syncthreads();
    /* every zone lies in this tile: its pixel count goes to its representative's record */
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        if (dg[l] == 0xFFFFu || par[l] != (uint32_t)l) continue;
        const int x = x0 + l % tg.TW, y = y0 + (l / tg.TW) % tg.TH, z = z0 + l / (tg.TW * tg.TH);
        zarea[ta.off + x + y * W + z * WH] = zc[l];
    }
}

/* the neighbour pairs of equal digit inside every tile (forward half of the neighbourhood): each zone
 * merge needs one, so a tiling has at least n - pairs zones.  Counted before labelling, it tells when
 * the zones are sure to exceed half the pixels and the labelling can be skipped. */
__global__ __launch_bounds__(256) void k_zone_pairs(ImgView iv, int shift, TileGeo tg, Nbr nb, unsigned long long *count)
{
    __shared__ unsigned long long part[4];
    const TileAt ta = tile_at(tg, blockIdx.x);
    const int tsz = tg.TW * tg.TH * tg.TD;
    const int64_t W = ta.W, WH = (int64_t)ta.W * ta.H;
    unsigned long long c = 0;
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        const int lx = l % tg.TW, ly = (l / tg.TW) % tg.TH, lz = l / (tg.TW * tg.TH);
        const int x = ta.x0 + lx, y = ta.y0 + ly, z = ta.z0 + lz;
        if (x >= ta.W || y >= ta.H || z >= ta.D) continue;
        const uint32_t d = fl_key(iv, ta.off + x + y * W + z * WH) >> shift;
        for (int k = 0; k < nb.n; k++) {
            const int dx = nb.dx[k], dy = nb.dy[k], dz = nb.dz[k];
            if (!(dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0))))) continue;
            const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
            if (nx < 0 || nx >= tg.TW || ny < 0 || ny >= tg.TH || nz < 0 || nz >= tg.TD) continue;
            const int gx = ta.x0 + nx, gy = ta.y0 + ny, gz = ta.z0 + nz;
            if (gx >= ta.W || gy >= ta.H || gz >= ta.D) continue;
            c += (fl_key(iv, ta.off + gx + gy * W + gz * WH) >> shift) == d;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)c, o, 64), hi = __shfl_xor((uint32_t)(c >> 32), o, 64);
        c += ((unsigned long long)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(count, part[0] + part[1] + part[2] + part[3]);
}

__device__ __forceinline__ uint32_t zh_hash(unsigned long long k)
{
    k ^= k >> 31;
    k *= 0x9E3779B97F4A7C15ull;
    return (uint32_t)(k >> 40);
}

/* distinct (zone, processed neighbour zone) pairs of every tile, appended to the bucket of the
 * zone's level: a neighbour q of p (zone rp, digit d) is processed when its digit is higher, or
 * equal in another zone (each such pair once, from the zone of smaller representative).  A level's
 * bucket holds at most NB edges per pixel of the level, which bounds every append.  NMAX: the
 * neighbourhood's size rounded up to 4, 6, 8 or 26; a pixel loads every neighbour's digit and zone
 * before it looks at any, so the loads are in flight together instead of one round trip each. */
template <int NMAX>
__global__ __launch_bounds__(256) void k_zone_edges(ImgView iv, int shift, TileGeo tg, Nbr nb, const uint32_t *__restrict__ zrep,
                                                    const uint64_t *__restrict__ bucket, uint32_t *__restrict__ cursor,
                                                    uint2 *__restrict__ edges)
{
    __shared__ unsigned long long tab[ZH_CAP];
    __shared__ uint32_t cnt[NLEV], base[NLEV];
    for (int i = threadIdx.x; i < ZH_CAP; i += blockDim.x) tab[i] = ZH_EMPTY;
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const TileAt ta = tile_at(tg, blockIdx.x);
    const int x0 = ta.x0, y0 = ta.y0, z0 = ta.z0;
    const int tsz = tg.TW * tg.TH * tg.TD;
    const int64_t WH = (int64_t)ta.W * ta.H;
    for (int l = threadIdx.x; l < tsz; l += blockDim.x) {
        const int x = x0 + l % tg.TW, y = y0 + (l / tg.TW) % tg.TH, z = z0 + l / (tg.TW * tg.TH);
        if (x >= ta.W || y >= ta.H || z >= ta.D) continue;
        const int64_t p = ta.off + x + (int64_t)y * ta.W + (int64_t)z * WH;
        const uint32_t dp = fl_key(iv, p) >> shift;
        const uint32_t rp = zrep[p];
        uint32_t dq[NMAX], rq[NMAX];
#pragma unroll
        for (int k = 0; k < NMAX; k++) {
            /* a neighbour outside the image (or beyond the neighbourhood) reads the pixel itself, whose
             * pair is skipped below (same digit, same zone) */
            int64_t q = p;
            if (k < nb.n) {
                const int nx = x + nb.dx[k], ny = y + nb.dy[k], nz = z + nb.dz[k];
                if (nx >= 0 && nx < ta.W && ny >= 0 && ny < ta.H && nz >= 0 && nz < ta.D)
                    q = ta.off + nx + (int64_t)ny * ta.W + (int64_t)nz * WH;
            }
            dq[k] = fl_key(iv, q) >> shift;
            rq[k] = zrep[q];
        }
#pragma unroll
        for (int k = 0; k < NMAX; k++) {
            if (dq[k] < dp || (dq[k] == dp && rq[k] <= rp)) continue;
            const unsigned long long key = ((unsigned long long)rp << 32) | rq[k];
            uint32_t h = zh_hash(key) & (ZH_CAP - 1);
            bool placed = false;
            for (int probe = 0; probe < 64; probe++) {
                const unsigned long long prev = atomicCAS(&tab[h], ZH_EMPTY, key);
                if (prev == ZH_EMPTY) {
                    atomicAdd(&cnt[dp], 1u);
                    placed = true;
                    break;
                }
                if (prev == key) {
                    placed = true;
                    break;
                }
                h = (h + 1) & (ZH_CAP - 1);
            }
            if (!placed) {
                /* table crowded: append this pair directly */
                const uint32_t pos = atomicAdd(&cursor[dp], 1u);
                edges[bucket[dp] + pos] = make_uint2(rp, rq[k]);
            }
        }
    }
    __syncthreads();
    const uint32_t c = cnt[threadIdx.x];
    base[threadIdx.x] = c ? atomicAdd(&cursor[threadIdx.x], c) : 0u;
    cnt[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < ZH_CAP; i += blockDim.x) {
        const unsigned long long key = tab[i];
        if (key == ZH_EMPTY) continue;
        const uint32_t rp = (uint32_t)(key >> 32), rq = (uint32_t)key;
        const uint32_t d = fl_key(iv, rp) >> shift;
        const uint32_t j = atomicAdd(&cnt[d], 1u);
        edges[bucket[d] + base[d] + j] = make_uint2(rp, rq);
    }
}

/* the unions of one level from its zone edges (k_union's linking rules) */
__global__ __launch_bounds__(256) void k_union_edges(LevelArgs a, const uint2 *__restrict__ edges, uint32_t cnt, uint32_t lev)
{
    __shared__ uint32_t s_abs[256 * 2];
    __shared__ uint32_t s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const AbsBuf ab{ s_abs, &s_n };
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i < cnt) {
        const uint2 e = edges[i];
        uint32_t ta, tb;
        const uint32_t ra = uf_find(a.uf, e.x, ta);
        const uint32_t rb = uf_find(a.uf, e.y, tb);
        uint32_t tout;
        (void)uf_unite(a.uf, ra, ta, rb, tb, lev, ab, tout);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_n) s_base = atomicAdd(a.nabs, s_n);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < s_n; j += blockDim.x) a.absorbed[s_base + j] = s_abs[j];
}

/* ------------------------------------------------------------------------------------------------ */
/* wave-aggregated attribute updates                                                                */
/* ------------------------------------------------------------------------------------------------ */
/* All pixels of one level that fall in the same component add into the same root record; plain
 * atomics would serialise on that address (a 512^3 level is one component of ~500k voxels).  A
 * wavefront first groups its lanes by target root (ballot on the leader's root), reduces each
 * group with cross-lane shuffles and issues one atomic per group and field.  All 64 lanes must
 * call these functions (inactive lanes pass valid = false). */
__device__ __forceinline__ uint64_t wsum64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, o, 64);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), o, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ __forceinline__ uint32_t wmin32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t wmax32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor(v, o, 64));
    return v;
}

/* the sums and extents an attribute group accumulates (its record layout, as alloc_attr_n) */
template <int G> struct GroupShape {
    static constexpr int NS = G == 2 ? 3 : (G == 3 ? 4 : (G == 4 ? 9 : 0));
    static constexpr bool EXT = G == 1;
};

/* one lane's contribution to a record of group G: count, the group's sums, its minima and maxima
 * (sized at compile time, so it lives in registers) */
template <int G> struct Contrib {
    static constexpr int NS = GroupShape<G>::NS;
    static constexpr bool EXT = GroupShape<G>::EXT;
    uint32_t area;
    int64_t s[NS > 0 ? NS : 1];
    uint32_t mn[3], mx[3];
};

template <int G> __device__ __forceinline__ void contrib_own(const LevelArgs &a, uint32_t p, Contrib<G> &c)
{
    c.area = 1;
    if constexpr (G == 0) {
        return;
    } else {
        const PixGeo pg = pix_geo(a, p);
        const int x = pg.x + a.ox, y = pg.y + a.oy, z = pg.z + a.oz;
        if constexpr (G == 1) {
            c.mn[0] = (uint32_t)x; c.mn[1] = (uint32_t)y; c.mn[2] = (uint32_t)z;
            c.mx[0] = (uint32_t)x + 1; c.mx[1] = (uint32_t)y + 1; c.mx[2] = (uint32_t)z + 1;
        } else if constexpr (G == 2) {
            c.s[0] = x; c.s[1] = y; c.s[2] = z;
        } else if constexpr (G == 3) {
            int64_t gv = (int64_t)fl_key(a.iv, p);
            c.s[0] = gv; c.s[1] = gv * x; c.s[2] = gv * y; c.s[3] = gv * z;
        } else {
            c.s[0] = x; c.s[1] = y; c.s[2] = z;
            c.s[3] = (int64_t)x * x; c.s[4] = (int64_t)y * y; c.s[5] = (int64_t)z * z;
            c.s[6] = (int64_t)x * y; c.s[7] = (int64_t)y * z; c.s[8] = (int64_t)x * z;
        }
    }
}

template <int G> __device__ __forceinline__ void contrib_rec(const DevAttr &B, uint32_t h, Contrib<G> &c)
{
    c.area = B.area[h];
#pragma unroll
    for (int i = 0; i < Contrib<G>::NS; i++) c.s[i] = B.s[i][h];
    if constexpr (Contrib<G>::EXT)
        for (int i = 0; i < 3; i++) { c.mn[i] = B.mn[i][h]; c.mx[i] = B.mx[i][h]; }
}

template <int G> __device__ __forceinline__ void contrib_zero(Contrib<G> &c)
{
    c.area = 0;
#pragma unroll
    for (int i = 0; i < Contrib<G>::NS; i++) c.s[i] = 0;
    if constexpr (Contrib<G>::EXT)
        for (int i = 0; i < 3; i++) { c.mn[i] = 0xFFFFFFFFu; c.mx[i] = 0; }
}

/* the first lane to lower canon[r] from empty queues r for k_fix */
__device__ __forceinline__ void canon_min(const LevelArgs &a, uint32_t r, uint32_t p)
{
    if (atomicMin(&a.canon[r], p) == 0xFFFFFFFFu) a.fix[atomicAdd(a.fcount, 1u)] = r;
}

/* adds every valid lane's contribution to record r of A: one reduction when the wave has one root,
 * one atomic per lane when no two neighbouring lanes share a root, else segmented reductions over the
 * runs of neighbouring lanes with one root (a fixed cost, whatever the number of roots) and one atomic
 * per run and field.  CANON: also lowers canon[r] to the lane's pixel p (the canonical levelroot
 * candidates of the sets rooted at old roots) in the same pass */
template <int G, bool CANON = false>
__device__ void wave_add(const DevAttr &A, bool valid, uint32_t r, const Contrib<G> &c, const LevelArgs *a = nullptr,
                         uint32_t p = 0)
{
    constexpr int NS = Contrib<G>::NS;
    const int lane = threadIdx.x & 63;
    const uint64_t vmask = __ballot(valid);
    if (!vmask) return;
    {
        /* one root for the whole wave (big flat level sets): one reduction, one atomic per field */
        const int leader = __ffsll((unsigned long long)vmask) - 1;
        const uint32_t rl = __shfl(r, leader, 64);
        if (__ballot(valid && r != rl) == 0) {
            const uint64_t ar = wsum64(valid ? c.area : 0);
            if (lane == leader) atomicAdd(&A.area[rl], (uint32_t)ar);
#pragma unroll
            for (int i = 0; i < NS; i++) {
                const uint64_t v = wsum64(valid ? (uint64_t)c.s[i] : 0);
                if (lane == leader && v) atomicAdd((unsigned long long *)&A.s[i][rl], (unsigned long long)v);
            }
            if constexpr (Contrib<G>::EXT)
                for (int i = 0; i < 3; i++) {
                    const uint32_t lo = wmin32(valid ? c.mn[i] : 0xFFFFFFFFu), hi = wmax32(valid ? c.mx[i] : 0u);
                    if (lane == leader) { atomicMin(&A.mn[i][rl], lo); atomicMax(&A.mx[i][rl], hi); }
                }
            if constexpr (CANON) {
                const uint32_t pm = wmin32(valid ? p : 0xFFFFFFFFu);
                if (lane == leader) canon_min(*a, rl, pm);
            }
            return;
        }
    }
    /* runs of consecutive valid lanes with one root (the level lists are tile-ordered, so a root's
     * pixels tend to sit in neighbouring lanes) */
    const uint32_t rprev = __shfl_up(r, 1, 64);
    const bool cont = valid && lane > 0 && ((vmask >> (lane - 1)) & 1ull) && rprev == r;
    const uint64_t cmask = __ballot(cont);
    if (!cmask) {
        /* no two neighbouring lanes share a root: every lane updates its own */
        if (valid) {
            atomicAdd(&A.area[r], c.area);
#pragma unroll
            for (int i = 0; i < NS; i++)
                if (c.s[i]) atomicAdd((unsigned long long *)&A.s[i][r], (unsigned long long)c.s[i]);
            if constexpr (Contrib<G>::EXT)
                for (int i = 0; i < 3; i++) { atomicMin(&A.mn[i][r], c.mn[i]); atomicMax(&A.mx[i][r], c.mx[i]); }
            if constexpr (CANON) canon_min(*a, r, p);
        }
        return;
    }
    /* segmented suffix reductions: each run's first lane ends with the run's totals (lanes that are
     * not valid carry neutral values and never start a run) */
    const bool head = valid && !cont;
    const uint64_t hmask = __ballot(head);
    const uint32_t rid = (uint32_t)__popcll(hmask & (lane == 63 ? ~0ull : ((2ull << lane) - 1)));
    uint32_t area = valid ? c.area : 0u;
    int64_t sm[NS > 0 ? NS : 1];
#pragma unroll
    for (int i = 0; i < NS; i++) sm[i] = valid ? c.s[i] : 0;
    uint32_t mn[3], mx[3];
    if constexpr (Contrib<G>::EXT)
        for (int i = 0; i < 3; i++) { mn[i] = valid ? c.mn[i] : 0xFFFFFFFFu; mx[i] = valid ? c.mx[i] : 0u; }
    uint32_t pm = valid ? p : 0xFFFFFFFFu;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t rid2 = __shfl_down(rid, off, 64);
        const bool same = lane + off < 64 && rid2 == rid;
        const uint32_t a2 = __shfl_down(area, off, 64);
        if (same) area += a2;
#pragma unroll
        for (int i = 0; i < NS; i++) {
            const uint32_t lo = __shfl_down((uint32_t)sm[i], off, 64), hi = __shfl_down((uint32_t)((uint64_t)sm[i] >> 32), off, 64);
            if (same) sm[i] += (int64_t)(((uint64_t)hi << 32) | lo);
        }
        if constexpr (Contrib<G>::EXT)
            for (int i = 0; i < 3; i++) {
                const uint32_t m1 = __shfl_down(mn[i], off, 64), m2 = __shfl_down(mx[i], off, 64);
                if (same) { mn[i] = min(mn[i], m1); mx[i] = max(mx[i], m2); }
            }
        if constexpr (CANON) {
            const uint32_t p2 = __shfl_down(pm, off, 64);
            if (same) pm = min(pm, p2);
        }
    }
    if (head) {
        atomicAdd(&A.area[r], area);
#pragma unroll
        for (int i = 0; i < NS; i++)
            if (sm[i]) atomicAdd((unsigned long long *)&A.s[i][r], (unsigned long long)sm[i]);
        if constexpr (Contrib<G>::EXT)
            for (int i = 0; i < 3; i++) { atomicMin(&A.mn[i][r], mn[i]); atomicMax(&A.mx[i][r], mx[i]); }
        if constexpr (CANON) canon_min(*a, r, pm);
    }
}

/* zone pass: the coordinate / grey sums and extents of every zone into its representative's record
 * (the pixel counts come from k_tile_zones) */
template <int G>
__global__ __launch_bounds__(256) void k_zone_totals(LevelArgs a, const uint32_t *__restrict__ zrep, DevAttr A, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    uint32_t r = 0;
    Contrib<G> c;
    contrib_zero(c);
    if (valid) {
        r = zrep[i];
        contrib_own(a, (uint32_t)i, c);
        c.area = 0;
    }
    wave_add(A, valid, r, c);
}

/* zone pass, after the levels: a pixel that is not its zone's representative r shares r's node: its
 * levelroot is r when r is one (r's parent lies at a lower digit, or r is the root), else r's parent */
__global__ __launch_bounds__(256) void k_zone_parents(ImgView iv, int pass, int shift, const uint32_t *__restrict__ zrep,
                                                      int32_t *__restrict__ par, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t r = zrep[p];
    if (r == (uint32_t)p) return;
    const int32_t q = par[r];
    auto digit = [&](int64_t x) { return (fl_key(iv, x) >> shift) & 255u; };
    par[p] = (q >= 0 && digit(q) == digit(r)) ? q : (int32_t)r;
}

/* Block-level variant for kernels run with FB threads per block: when every valid lane of the block
 * targets the same root (the pixels of one big set), the block reduces its contributions through
 * LDS and issues one atomic per field; otherwise each wave aggregates on its own.  Same-address
 * atomics from thousands of waves were the cost of the final passes on levels that are one big
 * component.  Every thread of the block must call these. */
#define FB 256
/* a parent left pending during a pass: the set rooted at r, whose canonical levelroot is not known yet */
#define HELD(r) (-2 - (int32_t)(r))
#define HELD_ROOT(v) ((uint32_t)(-2 - (v)))
#define FW (FB / 64)
#define R_EMPTY 0xFFFFFFFFu
#define R_MIXED 0xFFFFFFFEu

struct BlockScratch {
    uint32_t root[FW];
    uint32_t pmin[FW];
    uint64_t red[FW][1 + DCF_MAX_SUMS + 6];
};

/* the root shared by all valid lanes of the block, or R_EMPTY / R_MIXED */
__device__ __forceinline__ uint32_t block_common_root(BlockScratch &bs, bool valid, uint32_t r)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t vm = __ballot(valid);
    uint32_t code = R_EMPTY;
    if (vm) {
        const uint32_t rl = __shfl(r, __ffsll((unsigned long long)vm) - 1, 64);
        code = __ballot(valid && r != rl) ? R_MIXED : rl;
    }
    if (lane == 0) bs.root[w] = code;
    __syncthreads();
    uint32_t R = R_EMPTY;
    for (int k = 0; k < FW; k++) {
        const uint32_t ck = bs.root[k];
        if (ck == R_EMPTY) continue;
        if (ck == R_MIXED || (R != R_EMPTY && ck != R)) return R_MIXED;
        R = ck;
    }
    return R;
}

/* a block's totals for one root (thread 0), kept across several chunks of a level while the root stays
 * the same, then added with one atomic per field */
template <int G> struct BlkCarry {
    static constexpr int NS = Contrib<G>::NS;
    uint32_t R;
    uint64_t area;
    uint64_t s[NS > 0 ? NS : 1];
    uint32_t mn[3], mx[3];
    uint32_t pmin;
};

template <int G> __device__ __forceinline__ void carry_zero(BlkCarry<G> &k, uint32_t R)
{
    k.R = R;
    k.area = 0;
#pragma unroll
    for (int i = 0; i < BlkCarry<G>::NS; i++) k.s[i] = 0;
    for (int i = 0; i < 3; i++) { k.mn[i] = 0xFFFFFFFFu; k.mx[i] = 0u; }
    k.pmin = 0xFFFFFFFFu;
}

template <int G, bool CANON>
__device__ __forceinline__ void carry_flush(const DevAttr &A, const BlkCarry<G> &k, const LevelArgs *a)
{
    if (k.R == R_EMPTY) return;
    atomicAdd(&A.area[k.R], (uint32_t)k.area);
    if constexpr (CANON) canon_min(*a, k.R, k.pmin);
#pragma unroll
    for (int i = 0; i < BlkCarry<G>::NS; i++)
        if (k.s[i]) atomicAdd((unsigned long long *)&A.s[i][k.R], (unsigned long long)k.s[i]);
    if constexpr (Contrib<G>::EXT)
        for (int i = 0; i < 3; i++) {
            atomicMin(&A.mn[i][k.R], k.mn[i]);
            atomicMax(&A.mx[i][k.R], k.mx[i]);
        }
}

/* block-wide wave_add: when every valid lane of the block shares one root R, the block's totals go to
 * thread 0, which adds them to its carry (CARRY: flushing the carry first when it holds another root)
 * or straight to R's record with one atomic per field; otherwise each wave aggregates on its own */
template <int G, bool CANON = false, bool CARRY = false>
__device__ void block_add(const DevAttr &A, BlockScratch &bs, bool valid, uint32_t r, const Contrib<G> &c,
                          const LevelArgs *a = nullptr, uint32_t p = 0, BlkCarry<G> *carry = nullptr)
{
    constexpr int NS = Contrib<G>::NS;
    const uint32_t R = block_common_root(bs, valid, r);
    if (R == R_EMPTY) return;
    if (R == R_MIXED) {
        wave_add<G, CANON>(A, valid, r, c, a, p);
        return;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t area = wsum64(valid ? c.area : 0);
    if (lane == 0) bs.red[w][0] = area;
    if constexpr (CANON) {
        const uint32_t pm = wmin32(valid ? p : 0xFFFFFFFFu);
        if (lane == 0) bs.pmin[w] = pm;
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
        const uint64_t v = wsum64(valid ? (uint64_t)c.s[i] : 0);
        if (lane == 0) bs.red[w][1 + i] = v;
    }
    if constexpr (Contrib<G>::EXT)
        for (int i = 0; i < 3; i++) {
            const uint32_t lo = wmin32(valid ? c.mn[i] : 0xFFFFFFFFu), hi = wmax32(valid ? c.mx[i] : 0u);
            if (lane == 0) { bs.red[w][1 + DCF_MAX_SUMS + i] = lo; bs.red[w][1 + DCF_MAX_SUMS + 3 + i] = hi; }
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        BlkCarry<G> t;
        carry_zero(t, R);
        for (int k = 0; k < FW; k++) t.area += bs.red[k][0];
        if constexpr (CANON)
            for (int k = 0; k < FW; k++) t.pmin = min(t.pmin, bs.pmin[k]);
#pragma unroll
        for (int i = 0; i < NS; i++)
            for (int k = 0; k < FW; k++) t.s[i] += bs.red[k][1 + i];
        if constexpr (Contrib<G>::EXT)
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < FW; k++) {
                    t.mn[i] = min(t.mn[i], (uint32_t)bs.red[k][1 + DCF_MAX_SUMS + i]);
                    t.mx[i] = max(t.mx[i], (uint32_t)bs.red[k][1 + DCF_MAX_SUMS + 3 + i]);
                }
        if constexpr (CARRY) {
            BlkCarry<G> &k = *carry;
            if (k.R != R) {
                carry_flush<G, CANON>(A, k, a);
                k = t;
            } else {
                k.area += t.area;
                k.pmin = min(k.pmin, t.pmin);
#pragma unroll
                for (int i = 0; i < NS; i++) k.s[i] += t.s[i];
                for (int i = 0; i < 3; i++) { k.mn[i] = min(k.mn[i], t.mn[i]); k.mx[i] = max(k.mx[i], t.mx[i]); }
            }
        } else {
            carry_flush<G, CANON>(A, t, a);
        }
    }
}

/* where a pending parent of r's set points until the pass ends: the set's current node, whose parent
 * k_fix sets to the set's new node; for a super-vertex root with no node yet, the super-vertex itself,
 * whose join k_fix records in spar */
__device__ __forceinline__ uint32_t held_target(const LevelArgs &a, uint32_t r)
{
    const uint32_t x = a.nrep[r];
    return x != 0xFFFFFFFFu ? x : r;
}

/* after the unions of a level, every pixel p of the level finds its set's root r.  A root of this
 * level is the canonical levelroot: p takes it as parent and adds its own attribute contribution.
 * A set still rooted at an old root r gets its canonical levelroot computed here (minimum pixel of
 * the level, one atomicMin per wave and set; the first to touch a set queues r for k_fix); p's
 * contribution is kept pending at r (P) and its parent is left pending (HELD). */
template <int G, bool BLK>
__global__ __launch_bounds__(FB) void k_final_pixels(LevelArgs a, const uint32_t *__restrict__ lvl, uint32_t cnt, uint32_t lev,
                                                     DevAttr A, DevAttr S, DevAttr P, int32_t *__restrict__ par,
                                                     int32_t *__restrict__ spar, uint32_t *__restrict__ marks, uint32_t chunk,
                                                     uint32_t nmain)
{
    __shared__ BlockScratch bs;
    /* blocks [0, nmain) take the level's elements, the others the roots absorbed at the level: the two
     * chains of dependent loads run side by side instead of one after the other in each thread */
    const bool main_part = blockIdx.x < nmain;
    /* BLK: the block takes `chunk` consecutive FB-element chunks of the level list (one when the grid
     * covers the list); the totals of a root shared by the whole block are carried from chunk to chunk
     * (thread 0) and added once the root changes, so a level that is one big set sends one atomic per
     * field and block instead of one per chunk */
    BlkCarry<G> cf, co;
    carry_zero(cf, R_EMPTY);
    carry_zero(co, R_EMPTY);
    const uint32_t L = xcd_map(blockIdx.x, nmain);
    const uint32_t c0 = main_part ? (BLK ? L * chunk : L) : 0u, c1 = main_part ? (BLK ? c0 + chunk : c0 + 1) : 0u;
    for (uint32_t ch = c0; ch < c1 && ch * FB < cnt; ch++) {
        const uint32_t i = ch * FB + threadIdx.x;
        const bool valid = i < cnt;
        uint32_t p = 0, r = 0;
        bool fresh = false, old = false;
        Contrib<G> c;
        contrib_zero(c);
        if (valid) {
            p = lvl[i];
            uint32_t tag;
            r = uf_find(a.uf, p, tag);
            fresh = (tag & UF_LVL) == lev;
            old = !fresh;
            if (fresh) {
                if (r != p) par[p] = (int32_t)r;
                if (!a.zone) contrib_own(a, p, c);
                else if (r != p) contrib_rec(A, p, c); /* a root already holds its own zone's totals */
            } else {
                par[p] = HELD(held_target(a, r));
                if (!a.zone) contrib_own(a, p, c);
                else contrib_rec(A, p, c);
            }
        }
        /* BLK: a block whose pixels share one root (big flat level sets, as in a first pass over pixels)
         * adds once; otherwise (zone representatives, a refining pass's scattered pixels) the waves
         * aggregate on their own */
        if constexpr (BLK) {
            block_add<G, false, true>(A, bs, fresh, r, c, nullptr, 0, &cf);
            __syncthreads();
            /* the held pixels: contributions pending at their old root, canonical levelroot candidates */
            block_add<G, true, true>(P, bs, old, r, c, &a, p, &co);
            __syncthreads();
        } else {
            wave_add(A, fresh, r, c);
            wave_add<G, true>(P, old, r, c, &a, p);
        }
    }
    if (BLK && threadIdx.x == 0) {
        carry_flush<G, false>(A, cf, nullptr);
        carry_flush<G, true>(P, co, &a);
    }
    if (main_part) return;
    /* the old roots linked under another root at this level: parent = the set's canonical levelroot
     * (resolved after the pass), complete record pending at the set's root */
    const uint32_t start = marks[0], end = *a.nabs;
    if (blockIdx.x == nmain && threadIdx.x == 0) marks[1] = end;
    const uint32_t stride = (gridDim.x - nmain) * blockDim.x;
    /* every lane of a wave runs the same number of iterations so wave_add sees the full wave */
    for (uint32_t base = start + (blockIdx.x - nmain) * blockDim.x; base < end; base += stride) {
        const uint32_t j = base + threadIdx.x;
        const bool valid = j < end;
        uint32_t r = 0;
        Contrib<G> c;
        contrib_zero(c);
        if (valid) {
            const uint32_t h = a.absorbed[j];
            uint32_t tag;
            r = uf_find(a.uf, h, tag);
            /* h's set's current node becomes a child of the node r's set gets at this level */
            const uint32_t hn = a.nrep[h], to = HELD(held_target(a, r));
            if (hn != 0xFFFFFFFFu) {
                par[hn] = to;
                contrib_rec(A, hn, c);
            } else {
                const uint32_t cc = h - (uint32_t)a.g.n;
                spar[cc] = to;
                contrib_rec(S, cc, c);
            }
        }
        wave_add(P, valid, r, c);
    }
}

/* every queued set r gets its node of this level, c = its minimum pixel of the level: the set's
 * current node (or, for a super-vertex root with no node yet, the super-vertex's join) hangs below c,
 * and c's record is that node's (or the super-vertex's) complete record plus everything pending at r
 * (the set's pixels of this level, the old sets absorbed into it).  r stays the set's root (only its
 * tag moves to this level), so its members' paths do not grow with every level the set gains.
 * canon[r] and P[r] are cleared for the set's next level. */
template <int G>
__global__ __launch_bounds__(256) void k_fix(LevelArgs a, uint32_t ord, uint32_t lev, DevAttr A, DevAttr S, DevAttr P,
                                             int32_t *__restrict__ par, int32_t *__restrict__ spar)
{
    constexpr int NS = GroupShape<G>::NS;
    const uint32_t start = a.fcount[1 + ord], end = a.fcount[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.fcount[2 + ord] = end;
    for (uint32_t j = start + blockIdx.x * blockDim.x + threadIdx.x; j < end; j += gridDim.x * blockDim.x) {
        const uint32_t r = a.fix[j];
        const uint32_t c = a.canon[r];
        const uint32_t x = a.nrep[r];
        const bool sv = x == 0xFFFFFFFFu;
        const DevAttr &R = sv ? S : A;
        const uint32_t ri = sv ? r - (uint32_t)a.g.n : x;
        if (sv) spar[ri] = (int32_t)c;
        else par[x] = (int32_t)c;
        A.area[c] = R.area[ri] + P.area[r];
        P.area[r] = 0;
#pragma unroll
        for (int k = 0; k < NS; k++) {
            A.s[k][c] = R.s[k][ri] + P.s[k][r];
            P.s[k][r] = 0;
        }
        if constexpr (GroupShape<G>::EXT)
            for (int k = 0; k < 3; k++) {
                A.mn[k][c] = min(R.mn[k][ri], P.mn[k][r]);
                A.mx[k][c] = max(R.mx[k][ri], P.mx[k][r]);
                P.mn[k][r] = 0xFFFFFFFFu;
                P.mx[k][r] = 0u;
            }
        a.nrep[r] = c;
        a.canon[r] = 0xFFFFFFFFu;
        /* r's tag becomes this level: among old roots the most recently grown set wins the links (no
         * CAS touches r during this kernel, and no thread of this level reads r's tag after it) */
        st(a.uf, r, UF_ROOT | lev);
    }
}

/* after a pass: a parent (or super-vertex join) left pending as HELD(x) is the parent k_fix gave x:
 * par[x] for a node x, spar for a super-vertex x; the node that is its own such parent has no parent
 * in this pass */
__global__ __launch_bounds__(256) void k_held_resolve(int32_t *__restrict__ par, int64_t n, int32_t *__restrict__ spar, int64_t kc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n + kc) return;
    int32_t *slot = i < n ? &par[i] : &spar[i - n];
    const int32_t v = *slot;
    if (v > -2) return;
    const uint32_t x = HELD_ROOT(v);
    const int32_t c = x < (uint32_t)n ? par[x] : spar[x - n];
    *slot = (i < n && c == (int32_t)i) ? -1 : c;
}

/* ------------------------------------------------------------------------------------------------ */
/* between the passes                                                                               */
/* ------------------------------------------------------------------------------------------------ */
/* every element a singleton root: pixels tagged with their level digit, super-vertices (ids n and
 * above) with NLEV */
__global__ __launch_bounds__(256) void k_uf_init(ImgView iv, int pass, int shift, int64_t n, int64_t total, UFV uf)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint32_t lv = NLEV;
    if (i < n) {
        uint32_t k = fl_key(iv, i);
        lv = (k >> shift) & 255u;
    }
    uf.p[(size_t)i * uf.s] = UF_ROOT | lv;
}

/* the same for a refining pass with records: {singleton root tag, key, coarse node, 0} per pixel,
 * {root tag of a super-vertex, 0, 0, 0} for the coarse nodes' super-vertices */
__global__ __launch_bounds__(256) void k_rec_init(ImgView iv, int shift, const uint32_t *__restrict__ cnode, int64_t n, int64_t total,
                                                  uint4 *__restrict__ rec)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint4 v = make_uint4(UF_ROOT | NLEV, 0u, 0u, 0u);
    if (i < n) {
        const uint32_t k = fl_key(iv, i);
        v = make_uint4(UF_ROOT | ((k >> shift) & 255u), k, cnode[i], 0u);
    }
    rec[i] = v;
}

__global__ __launch_bounds__(256) void k_iota_or_none(uint32_t *a, int64_t n, int64_t total)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) a[i] = i < n ? (uint32_t)i : 0xFFFFFFFFu;
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *a, int32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_u32f(uint32_t *a, uint32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_i64f(int64_t *a, int64_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* coarse node id of every pixel.  par1 of a coarse levelroot is its coarse parent (a pixel of
 * lower coarse key) or -1; of any other pixel it is its levelroot (same coarse key); flag marks the
 * levelroots and cidx numbers them. */
__global__ __launch_bounds__(256) void k_coarse_node_of(const int32_t *__restrict__ par1, const uint32_t *__restrict__ flag,
                                                        const uint32_t *__restrict__ cidx, uint32_t *__restrict__ cnode, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    cnode[p] = cidx[flag[p] ? p : (int64_t)par1[p]];
}

/* levelroots of the previous pass's tree, whose levels are the key prefixes key >> pshift */
__global__ __launch_bounds__(256) void k_coarse_flag(ImgView iv, const int32_t *__restrict__ par1, uint32_t *__restrict__ flag,
                                                     int pshift, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int32_t r = par1[p];
    flag[p] = (r < 0 || (fl_key(iv, r) >> pshift) != (fl_key(iv, p) >> pshift)) ? 1u : 0u;
}

/* per coarse node: key, parent id (lift level 0) and the region totals used by its super-vertex */
__global__ __launch_bounds__(256) void k_coarse_nodes(ImgView iv, const int32_t *__restrict__ par1,
                                                      const uint32_t *__restrict__ flag, const uint32_t *__restrict__ cidx,
                                                      uint32_t *__restrict__ cnkey, int32_t *__restrict__ lift0,
                                                      uint32_t *__restrict__ cpix, DevAttr A, DevAttr S, int pshift, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !flag[p]) return;
    uint32_t c = cidx[p];
    int32_t r = par1[p];
    cnkey[c] = fl_key(iv, p) >> pshift;
    lift0[c] = r < 0 ? -1 : (int32_t)cidx[r];
    cpix[c] = (uint32_t)p;
    S.area[c] = A.area[p];
    for (int i = 0; i < A.nsum; i++) S.s[i][c] = A.s[i][p];
    if (A.has_ext)
        for (int i = 0; i < 3; i++) { S.mn[i][c] = A.mn[i][p]; S.mx[i][c] = A.mx[i][p]; }
}

__global__ __launch_bounds__(256) void k_lift(int32_t *__restrict__ lift, int j, int64_t kc)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kc) return;
    int32_t u = lift[(j - 1) * kc + c];
    lift[j * kc + c] = u < 0 ? -1 : lift[(j - 1) * kc + u];
}

/* after pass 2: the lowest fine node of every coarse node hangs below the fine node its
 * super-vertex joined (or is the root) */
__global__ __launch_bounds__(256) void k_fine_roots(UFV uf, const uint32_t *__restrict__ nrep, const uint32_t *__restrict__ cpix,
                                                    const int32_t *__restrict__ spar, int32_t *__restrict__ par, int64_t kc)
{
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kc) return;
    uint32_t tag;
    const uint32_t fr = nrep[uf_find(uf, cpix[c], tag)];
    par[fr] = spar[c];
}


/* ------------------------------------------------------------------------------------------------ */
/* small images of a batch: each one's whole tree built in one workgroup's LDS                      */
/* ------------------------------------------------------------------------------------------------ */
/* A persistent kernel fed by a work queue: every 64-lane workgroup takes the next image of its list
 * from a counter and builds that image's canonical parent array and levelroot areas in LDS, with no
 * launch per grey level.  For one image of N <= SMAX pixels:
 *  - sort: the pixels by key, highest first, equal keys by index, highest first (bitonic, in LDS);
 *  - flood: in that order each pixel p becomes a set of its own, the lanes find the set of each of its
 *    processed neighbours, and every such set hangs below p in the tree (parent[rep] = p for the set's
 *    representative pixel, its most recent one) while the sets merge in a union-find by rank whose
 *    representative becomes p: one pixel at a time;
 *  - levelroots: the last pixel of a level component to be processed is its root in this parent
 *    forest, and with equal keys taken by falling index that is the component's minimum pixel, the
 *    canonical levelroot; every pixel finds its levelroot by pointer jumping over same-key parents,
 *    and the canonical parent of p is the levelroot of parent[p];
 *  - areas: each pixel counts 1 at its levelroot, then the levelroots add their totals to their
 *    parents, one run of equal keys after the other (a levelroot's parent lies at a lower key).
 * The result is the level loop's: the canonical parent array (-1 at the image's root) and the area at
 * every levelroot. */
/* p / d for p < 2^16 and 1 <= d <= 2^16, with m = floor((2^32 - 1) / d) + 1 (exact: the error of the
 * product is below 2^-16 and a quotient's fraction is at most 1 - 1/d) */
__device__ __forceinline__ uint32_t small_div(uint32_t p, uint32_t d, uint32_t m) { return d == 1 ? p : __umulhi(p, m); }

template <int SMAX>
__global__ __launch_bounds__(64) void k_small_build(ImgView iv, const BatchImg *__restrict__ bimg, const uint32_t *__restrict__ list,
                                                    uint32_t nlist, Nbr nb, int32_t *__restrict__ par, uint32_t *__restrict__ area,
                                                    uint32_t *__restrict__ queue)
{
    constexpr uint32_t NONE = 0xFFFFu;
    constexpr uint32_t IMASK = (uint32_t)SMAX - 1u; /* SMAX is a power of two */
    constexpr int IBITS = __builtin_ctz(SMAX);
    __shared__ uint32_t s32[SMAX]; /* sort keys; then low half: processing order, high half: areas */
    __shared__ uint16_t kk[SMAX];  /* grey keys */
    __shared__ uint16_t zp[SMAX];  /* union-find while flooding (NONE: not processed yet), then levelroots */
    __shared__ uint16_t pa[SMAX];  /* parent while flooding, then canonical parent */
    __shared__ uint16_t rep[SMAX]; /* a union-find root's representative: the set's most recent pixel */
    __shared__ uint8_t rk[SMAX];   /* union-find rank */
    const int lane = threadIdx.x;
    /* this lane's neighbour offset (lanes beyond the neighbourhood have none) */
    const bool has_nb = lane < nb.n;
    int ndx = 0, ndy = 0, ndz = 0;
    for (int k = 0; k < nb.n; k++)
        if (k == lane) {
            ndx = nb.dx[k];
            ndy = nb.dy[k];
            ndz = nb.dz[k];
        }
    for (;;) {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(queue, 1u);
        j = __shfl(j, 0, 64);
        if (j >= nlist) break;
        const BatchImg b = bimg[list[j]];
        const int W = (int)b.W, H = (int)b.H, D = (int)b.D, WH = W * H;
        const int N = WH * D;
        const uint32_t mW = 0xFFFFFFFFu / (uint32_t)W + 1u, mWH = 0xFFFFFFFFu / (uint32_t)WH + 1u;
        const int64_t off = (int64_t)b.off;
        int Np = 1;
        while (Np < N) Np <<= 1;
        for (int l = lane; l < Np; l += 64) {
            uint32_t v = 0xFFFFFFFFu;
            if (l < N) {
                const uint32_t k = fl_key(iv, off + l);
                kk[l] = (uint16_t)k;
                zp[l] = (uint16_t)NONE;
                pa[l] = (uint16_t)l;
                v = ((65535u - k) << IBITS) | (IMASK - (uint32_t)l);
            }
            s32[l] = v;
        }
        __syncthreads();
        for (int k = 2; k <= Np; k <<= 1)
            for (int h = k >> 1; h > 0; h >>= 1) {
                for (int i = lane; i < Np; i += 64) {
                    const int x = i ^ h;
                    if (x > i) {
                        const uint32_t a = s32[i], c = s32[x];
                        if ((a > c) == ((i & k) == 0)) {
                            s32[i] = c;
                            s32[x] = a;
                        }
                    }
                }
                __syncthreads();
            }
        for (int i = lane; i < N; i += 64) s32[i] = IMASK - (s32[i] & IMASK);
        __syncthreads();
        /* flood */
        for (int i = 0; i < N; i++) {
            const uint32_t p = s32[i] & 0xFFFFu;
            if (lane == 0) {
                zp[p] = (uint16_t)p;
                rep[p] = (uint16_t)p;
                rk[p] = 0;
            }
            uint32_t r = NONE;
            if (has_nb) {
                const uint32_t z = small_div(p, (uint32_t)WH, mWH), pr = p - z * (uint32_t)WH;
                const uint32_t y = small_div(pr, (uint32_t)W, mW), x = pr - y * (uint32_t)W;
                const int nx = (int)x + ndx, ny = (int)y + ndy, nz = (int)z + ndz;
                if (nx >= 0 && nx < W && ny >= 0 && ny < H && nz >= 0 && nz < D) {
                    uint32_t q = (uint32_t)(nx + ny * W + nz * WH);
                    if (zp[q] != NONE) {
                        while (zp[q] != q) {
                            const uint32_t g = zp[zp[q]];
                            zp[q] = (uint16_t)g;
                            q = g;
                        }
                        r = q;
                    }
                }
            }
            __syncthreads();
            /* the neighbours' sets, in lane order: each distinct one hangs below p and joins p's set (a set
             * merged earlier in this step is no longer a root, or is p's set itself) */
            uint32_t cur = p;
            for (int k = 0; k < nb.n; k++) {
                const uint32_t rr = (uint32_t)__shfl((int)r, k, 64);
                if (lane == 0 && rr != NONE && rr != cur && zp[rr] == rr) {
                    pa[rep[rr]] = (uint16_t)p;
                    const uint32_t ka = rk[cur], kb = rk[rr];
                    const uint32_t a = ka >= kb ? cur : rr, c = ka >= kb ? rr : cur;
                    zp[c] = (uint16_t)a;
                    if (ka == kb) rk[a] = (uint8_t)(ka + 1);
                    rep[a] = (uint16_t)p;
                    cur = a;
                }
            }
            __syncthreads();
        }
        /* levelroots: pointer jumping over same-key parents (zp is free now) */
        for (int l = lane; l < N; l += 64) {
            const int q = pa[l];
            zp[l] = (uint16_t)((q != l && kk[q] == kk[l]) ? q : l);
        }
        __syncthreads();
        for (;;) {
            bool changed = false;
            for (int l = lane; l < N; l += 64) {
                const uint32_t a = zp[l], c = zp[a];
                if (c != a) {
                    zp[l] = (uint16_t)c;
                    changed = true;
                }
            }
            __syncthreads();
            if (!__ballot(changed)) break;
        }
        /* canonical parents; area counts in the high halves of s32 */
        for (int l = lane; l < N; l += 64) {
            const int q = pa[l];
            pa[l] = (uint16_t)(q != l ? zp[q] : l);
            s32[l] &= 0xFFFFu;
        }
        __syncthreads();
        for (int l = lane; l < N; l += 64) atomicAdd(&s32[zp[l]], 1u << 16);
        __syncthreads();
        /* levelroot totals into their parents, one run of equal keys after the other */
        for (int i = 0; i < N;) {
            const uint32_t k0 = kk[s32[i] & 0xFFFFu];
            int e = i;
            for (;;) {
                const int at = e + lane;
                const bool same = at < N && kk[s32[at] & 0xFFFFu] == k0;
                const uint64_t m = __ballot(!same);
                if (m) {
                    e += __ffsll((unsigned long long)m) - 1;
                    break;
                }
                e += 64;
            }
            for (int at = i + lane; at < e; at += 64) {
                const int p = (int)(s32[at] & 0xFFFFu);
                if (zp[p] == p && pa[p] != p) atomicAdd(&s32[pa[p]], s32[p] & 0xFFFF0000u);
            }
            __syncthreads();
            i = e;
        }
        for (int l = lane; l < N; l += 64) {
            const int c = pa[l];
            par[off + l] = c == l ? -1 : (int32_t)(off + c);
            area[off + l] = s32[l] >> 16;
        }
        __syncthreads();
    }
}

/* ------------------------------------------------------------------------------------------------ */
/* host side                                                                                        */
/* ------------------------------------------------------------------------------------------------ */
static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static Nbr make_nbr(int conn, int is3d)
{
    Nbr f;
    f.n = 0;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                if (!x && !y && !z) continue;
                if (!is3d && z) continue;
                int l1 = abs(x) + abs(y) + abs(z);
                bool ok = (!is3d) ? (conn == 8 || l1 == 1) : (conn == 26 || l1 == 1);
                if (!ok) continue;
                f.dx[f.n] = (int8_t)x; f.dy[f.n] = (int8_t)y; f.dz[f.n] = (int8_t)z;
                f.n++;
            }
    return f;
}

struct Scratch {
    dcf_ctx *ctx;
    std::vector<void *> ptrs;
    void *get(size_t bytes)
    {
        void *p = dcf_alloc(ctx, bytes);
        if (p) ptrs.push_back(p);
        return p;
    }
    ~Scratch()
    {
        for (void *p : ptrs) dcf_release(ctx, p);
    }
};

static int alloc_attr_n(dcf_ctx *ctx, Scratch *sc, DevAttr &a, int grp, int64_t n)
{
    a = DevAttr();
    a.nsum = grp == 2 ? 3 : (grp == 3 ? 4 : (grp == 4 ? 9 : 0));
    a.has_ext = grp == 1;
    size_t k = (size_t)(n > 0 ? n : 1);
    auto get = [&](size_t b) -> void * { return sc ? sc->get(b) : dcf_alloc(ctx, b); };
    bool ok = (a.area = (uint32_t *)get(k * 4)) != nullptr;
    for (int i = 0; ok && i < a.nsum; i++) ok = (a.s[i] = (int64_t *)get(k * 8)) != nullptr;
    for (int i = 0; ok && a.has_ext && i < 3; i++) {
        ok = (a.mn[i] = (uint32_t *)get(k * 4)) != nullptr;
        ok = ok && (a.mx[i] = (uint32_t *)get(k * 4)) != nullptr;
    }
    if (!ok) return dcf_set_error(ctx, DCF_ENOMEM, "attribute arrays");
    return DCF_OK;
}

static int clear_attr(dcf_ctx *ctx, DevAttr &a, int64_t n)
{
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(nblk(n, 256)), dim3(256), 0, a.area, 0u, n);
    for (int i = 0; i < a.nsum; i++) DCF_LAUNCH(ctx, "k_fill", k_fill_i64f, dim3(nblk(n, 256)), dim3(256), 0, a.s[i], (int64_t)0, n);
    if (a.has_ext)
        for (int i = 0; i < 3; i++) {
            DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(nblk(n, 256)), dim3(256), 0, a.mn[i], 0xFFFFFFFFu, n);
            DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(nblk(n, 256)), dim3(256), 0, a.mx[i], 0u, n);
        }
    return DCF_OK;
}

/* one flooding pass over all levels (high to low) */
/* tgt: NB target slots per pixel for a pixel pass (nullptr: k_union evaluates the targets at each level) */
static int run_pass(dcf_ctx *ctx, LevelArgs &la, Scratch &sc, DevAttr &A, DevAttr &S, int grp, int32_t *par, int32_t *spar,
                    int64_t kc, bool zones, uint32_t *tgt, int is3d, int64_t batch_tiles)
{
    const int64_t n = la.g.n;
    /* a refining pass of a large image runs as K independent level loops, each on a stream of its own
     * (not under graph capture): level lists, absorbed roots and fix queues per loop, bin = loop * NLEV +
     * level.  Small images keep one loop: their levels are launch-bound already */
    const uint32_t K = (la.pass == 2 && !ctx->graph && ctx->fine_loops > 1 && n >= ctx->loops_min_pixels)
                           ? (uint32_t)std::min(ctx->fine_loops, FINE_LOOPS_MAX) : 1u;
    const uint32_t NB = K * NLEV;
    uint32_t *hist = (uint32_t *)sc.get((size_t)(2 * NB + NLEV) * 4);
    uint32_t *order = (uint32_t *)sc.get((size_t)n * 4);
    if (!hist || !order) return dcf_set_error(ctx, DCF_ENOMEM, "level lists");
    uint32_t *cursor = hist + NB, *hist_all = hist + 2 * NB;
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(nblk(2 * NB + NLEV, 256)), dim3(256), 0, hist, 0u, (int64_t)(2 * NB + NLEV));
    unsigned hb = nblk(n, 256);
    if (hb > 2048) hb = 2048;
    /* zone pass: label the tile zones first (pixels pre-linked to their representative, zone pixel
     * counts in the representatives' records); the level lists then hold representatives only */
    uint32_t *zrep = nullptr;
    TileGeo tg;
    int64_t ntiles = 0;
    la.zone = zones ? 1 : 0;
    if (zones) {
        tg.W = la.g.W; tg.H = la.g.H; tg.D = la.g.D;
        if (is3d) { tg.TW = ctx->zt3; tg.TH = ctx->zt3; tg.TD = ctx->zt3; }
        else { tg.TW = ctx->zt2[0]; tg.TH = ctx->zt2[1]; tg.TD = 1; }
        tg.ntx = (tg.W + tg.TW - 1) / tg.TW;
        tg.nty = (tg.H + tg.TH - 1) / tg.TH;
        tg.bimg = la.bimg;
        tg.nbatch = la.nbatch;
        ntiles = la.bimg ? batch_tiles : (int64_t)tg.ntx * tg.nty * ((tg.D + tg.TD - 1) / tg.TD);
    }
    if (zones && la.bimg) {
        /* a batch (often of noisy images): at least n - pairs zones; when that is already more than half
         * the pixels, skip the labelling (on the 8192^2 u16 image the count cost more than it saved) */
        unsigned long long *pairs = (unsigned long long *)sc.get(8);
        if (!pairs) return dcf_set_error(ctx, DCF_ENOMEM, "zone pair count");
        unsigned long long npairs = 0;
        DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(1), dim3(64), 0, (uint32_t *)pairs, 0u, (int64_t)2);
        DCF_LAUNCH(ctx, "k_zone_pairs", k_zone_pairs, dim3((unsigned)ntiles), dim3(256), 0, la.iv, la.shift, tg, la.nb, pairs);
        DCF_HIP(ctx, hipMemcpyAsync(&npairs, pairs, 8, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (((uint64_t)n - (npairs < (uint64_t)n ? npairs : (uint64_t)n)) * 2 > (uint64_t)n) {
            zones = false;
            la.zone = 0;
        }
    }
    if (zones) {
        zrep = (uint32_t *)sc.get((size_t)n * 4);
        if (!zrep) return dcf_set_error(ctx, DCF_ENOMEM, "zone representatives");
        DCF_LAUNCH(ctx, "k_tile_zones", k_tile_zones, dim3((unsigned)ntiles), dim3(256), 0, la.iv, la.shift, tg, la.nb, zrep, la.uf,
                   A.area);
        if (grp != 0)
            switch (grp) {
            case 1: DCF_LAUNCH(ctx, "k_zone_totals", k_zone_totals<1>, dim3(nblk(n, 256)), dim3(256), 0, la, (const uint32_t *)zrep, A, n); break;
            case 2: DCF_LAUNCH(ctx, "k_zone_totals", k_zone_totals<2>, dim3(nblk(n, 256)), dim3(256), 0, la, (const uint32_t *)zrep, A, n); break;
            case 3: DCF_LAUNCH(ctx, "k_zone_totals", k_zone_totals<3>, dim3(nblk(n, 256)), dim3(256), 0, la, (const uint32_t *)zrep, A, n); break;
            default: DCF_LAUNCH(ctx, "k_zone_totals", k_zone_totals<4>, dim3(nblk(n, 256)), dim3(256), 0, la, (const uint32_t *)zrep, A, n); break;
            }
        DCF_LAUNCH(ctx, "k_lvl_hist", k_lvl_hist, dim3(hb), dim3(256), 0, la.iv, la.pass, la.shift, n, (const uint32_t *)nullptr,
                   hist_all, (const uint32_t *)nullptr, 1u);
    }
    DCF_LAUNCH(ctx, "k_lvl_hist", k_lvl_hist, dim3(hb), dim3(256), 0, la.iv, la.pass, la.shift, n, (const uint32_t *)zrep, hist,
               la.cnode, K);
    uint32_t h[NLEV * FINE_LOOPS_MAX], hall[NLEV];
    DCF_HIP(ctx, hipMemcpyAsync(h, hist, (size_t)NB * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (zones) DCF_HIP(ctx, hipMemcpyAsync(hall, hist_all, sizeof hall, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (zones) {
        /* zones holding more than half the pixels save little in the level loop and crowd the edge
         * de-duplication: undo the pre-links and zone totals and take the pixel path */
        uint64_t nz = 0;
        for (int i = 0; i < NLEV; i++) nz += h[i];
        if (nz * 2 > (uint64_t)n) {
            zones = false;
            la.zone = 0;
            zrep = nullptr;
            memcpy(h, hall, sizeof hall);
            DCF_LAUNCH(ctx, "k_uf_init", k_uf_init, dim3(nblk(n, 256)), dim3(256), 0, la.iv, la.pass, la.shift, n, n, la.uf);
            int rc = clear_attr(ctx, A, n);
            if (rc) return rc;
        }
    }
    uint32_t off[NLEV * FINE_LOOPS_MAX + 1];
    off[0] = 0;
    for (uint32_t i = 0; i < NB; i++) off[i + 1] = off[i] + h[i];
    DCF_HIP(ctx, hipMemcpyAsync(cursor, off, (size_t)NB * 4, hipMemcpyHostToDevice, ctx->stream));
    ScatterGeo sg;
    sg.batch = la.bimg != nullptr;
    sg.bimg = la.bimg;
    sg.nbatch = la.nbatch;
    sg.W = la.g.W; sg.H = la.g.H; sg.D = la.g.D;
    const int *ts = sg.D > 1 ? ctx->lt3 : ctx->lt2, *gs = sg.D > 1 ? ctx->lg3 : ctx->lg2;
    sg.TW = ts[0]; sg.TH = ts[1]; sg.TD = ts[2];
    sg.GX = gs[0]; sg.GY = gs[1]; sg.GZ = gs[2];
    if (la.bimg) {
        /* a batch: every image's own zone tiles (their first indices are in the image table) */
        sg.TW = is3d ? ctx->zt3 : ctx->zt2[0];
        sg.TH = is3d ? ctx->zt3 : ctx->zt2[1];
        sg.TD = is3d ? ctx->zt3 : 1;
        sg.GX = sg.GY = sg.GZ = 1;
    }
    sg.ntx = (sg.W + sg.TW - 1) / sg.TW;
    sg.nty = (sg.H + sg.TH - 1) / sg.TH;
    sg.ntz = (sg.D + sg.TD - 1) / sg.TD;
    sg.ngx = (sg.ntx + sg.GX - 1) / sg.GX;
    sg.ngy = (sg.nty + sg.GY - 1) / sg.GY;
    const int64_t ngz = (sg.ntz + sg.GZ - 1) / sg.GZ;
    const int64_t sblocks = sg.batch ? batch_tiles : (int64_t)sg.ngx * sg.ngy * ngz * sg.GX * sg.GY * sg.GZ;
    if (zones) tgt = nullptr;
#define DCF_SCATTER(NBV)                                                                                                    \
    DCF_LAUNCH(ctx, "k_lvl_scatter", k_lvl_scatter<NBV>, dim3((unsigned)sblocks), dim3(256), 0, la.iv, la.pass, la.shift, n, sg, \
               (const uint32_t *)zrep, cursor, order, la, (uint32_t)kc, tgt, K)
    if (!tgt) DCF_SCATTER(0);
    else switch (la.nb.n) {
        case 4: DCF_SCATTER(4); break;
        case 6: DCF_SCATTER(6); break;
        case 8: DCF_SCATTER(8); break;
        default: DCF_SCATTER(26); break;
        }
#undef DCF_SCATTER
    /* zone graph: distinct zone pairs bucketed by level; a level's bucket holds at most NB edges per
     * pixel of the level */
    uint2 *edges = nullptr;
    uint64_t bucket_h[NLEV];
    uint32_t ecnt[NLEV];
    if (zones) {
        uint64_t *bucket = (uint64_t *)sc.get(NLEV * 8);
        uint32_t *ecur = (uint32_t *)sc.get(NLEV * 4);
        edges = (uint2 *)sc.get((size_t)n * la.nb.n * sizeof(uint2));
        if (!bucket || !ecur || !edges) return dcf_set_error(ctx, DCF_ENOMEM, "zone graph");
        uint64_t acc = 0;
        for (int i = 0; i < NLEV; i++) {
            bucket_h[i] = acc;
            acc += (uint64_t)la.nb.n * hall[i];
        }
        DCF_HIP(ctx, hipMemcpyAsync(bucket, bucket_h, sizeof bucket_h, hipMemcpyHostToDevice, ctx->stream));
        DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(1), dim3(256), 0, ecur, 0u, (int64_t)NLEV);
        const dim3 zg((unsigned)ntiles), zb(256);
        const uint32_t *zr = zrep;
        const uint64_t *bk = bucket;
        if (la.nb.n <= 4)
            DCF_LAUNCH(ctx, "k_zone_edges", k_zone_edges<4>, zg, zb, 0, la.iv, la.shift, tg, la.nb, zr, bk, ecur, edges);
        else if (la.nb.n <= 6)
            DCF_LAUNCH(ctx, "k_zone_edges", k_zone_edges<6>, zg, zb, 0, la.iv, la.shift, tg, la.nb, zr, bk, ecur, edges);
        else if (la.nb.n <= 8)
            DCF_LAUNCH(ctx, "k_zone_edges", k_zone_edges<8>, zg, zb, 0, la.iv, la.shift, tg, la.nb, zr, bk, ecur, edges);
        else
            DCF_LAUNCH(ctx, "k_zone_edges", k_zone_edges<26>, zg, zb, 0, la.iv, la.shift, tg, la.nb, zr, bk, ecur, edges);
        DCF_HIP(ctx, hipMemcpyAsync(ecnt, ecur, sizeof ecnt, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    /* nabs[0]: running count of absorbed roots; nabs[1 + i]: where the i-th processed level's
     * entries start (written on the device, so the level loop never waits for the host) */
    uint32_t *nabs = la.nabs;
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(2), dim3(256), 0, nabs, 0u, (int64_t)(NLEV + 2));
    /* node bookkeeping for sets rooted at old roots (one fix entry per node made from an old set) */
    const int64_t nel = n + kc;
    la.canon = (uint32_t *)sc.get((size_t)nel * 4);
    la.fix = (uint32_t *)sc.get((size_t)(nel + 16) * 4);
    la.fcount = (uint32_t *)sc.get((NLEV + 4) * 4);
    if (!la.canon || !la.fix || !la.fcount) return dcf_set_error(ctx, DCF_ENOMEM, "canonical root lists");
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(nblk(nel, 256)), dim3(256), 0, la.canon, 0xFFFFFFFFu, nel);
    /* every pixel is its own set's node until the set gains a level; super-vertices have no node */
    la.nrep = (uint32_t *)sc.get((size_t)nel * 4);
    if (!la.nrep) return dcf_set_error(ctx, DCF_ENOMEM, "set nodes");
    DCF_LAUNCH(ctx, "k_iota", k_iota_or_none, dim3(nblk(nel, 256)), dim3(256), 0, la.nrep, n, nel);
    /* records pending at the old roots of the sets of a level (each element is such a root at one
     * level at most) */
    DevAttr P;
    int rc = alloc_attr_n(ctx, &sc, P, grp, nel);
    if (rc || (rc = clear_attr(ctx, P, nel))) return rc;
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(2), dim3(256), 0, la.fcount, 0u, (int64_t)(NLEV + 4));
    /* every further loop: its own absorbed roots and fix queue, its own stream */
    LevelArgs lg[FINE_LOOPS_MAX];
    hipStream_t strm[FINE_LOOPS_MAX];
    lg[0] = la;
    strm[0] = ctx->stream;
    if (dcf_add_aux_streams(ctx, K - 1) != DCF_OK) return dcf_set_error(ctx, DCF_EDEVICE, "loop stream");
    for (uint32_t g = 1; g < K; g++) {
        lg[g] = la;
        lg[g].absorbed = (uint32_t *)sc.get((size_t)(nel + 16) * 4);
        lg[g].nabs = (uint32_t *)sc.get((NLEV + 4) * 4);
        lg[g].fix = (uint32_t *)sc.get((size_t)(nel + 16) * 4);
        lg[g].fcount = (uint32_t *)sc.get((NLEV + 4) * 4);
        if (!lg[g].absorbed || !lg[g].nabs || !lg[g].fix || !lg[g].fcount) return dcf_set_error(ctx, DCF_ENOMEM, "loop queues");
        DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(2), dim3(256), 0, lg[g].nabs, 0u, (int64_t)(NLEV + 2));
        DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(2), dim3(256), 0, lg[g].fcount, 0u, (int64_t)(NLEV + 4));
        strm[g] = ctx->aux[g - 1];
    }
    /* fork: the further loops start after everything queued so far */
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events()
        {
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } ev;
    if (K > 1) {
        hipEvent_t f = nullptr;
        if (hipEventCreateWithFlags(&f, hipEventDisableTiming) != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "event");
        ev.e.push_back(f);
        DCF_HIP(ctx, hipEventRecord(f, strm[0]));
        for (uint32_t g = 1; g < K; g++) DCF_HIP(ctx, hipStreamWaitEvent(strm[g], f, 0));
    }
    /* the level loop: its launches are fixed by the histogram already on the host.  With ctx->graph on
     * (and no per-kernel events) they are captured on the context's own stream into one HIP graph,
     * which is then launched on the working stream */
    int ordg[FINE_LOOPS_MAX] = { 0 };
    auto level_loop = [&]() -> int {
        for (int lev = NLEV - 1; lev >= 0; lev--)
        for (uint32_t g = 0; g < K; g++) {
            const uint32_t bin = g * NLEV + (uint32_t)lev;
            uint32_t cnt = h[bin];
            if (!cnt) continue;
            const uint32_t *lvl = order + off[bin];
            LevelArgs &la = lg[g];
            uint32_t *nabs = la.nabs;
            int &ord = ordg[g];
            if (K > 1) ctx->stream = strm[g];
            if (zones) {
                if (ecnt[lev])
                    DCF_LAUNCH(ctx, "k_union", k_union_edges, dim3(nblk(ecnt[lev], 256)), dim3(256), 0, la,
                               (const uint2 *)(edges + bucket_h[lev]), ecnt[lev], (uint32_t)lev);
            } else if (tgt) {
                const uint32_t *tl = tgt + (size_t)off[bin] * la.nb.n;
                switch (la.nb.n) {
                case 4: DCF_LAUNCH(ctx, "k_union", k_union_tgt<4>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, tl, cnt, (uint32_t)lev); break;
                case 6: DCF_LAUNCH(ctx, "k_union", k_union_tgt<6>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, tl, cnt, (uint32_t)lev); break;
                case 8: DCF_LAUNCH(ctx, "k_union", k_union_tgt<8>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, tl, cnt, (uint32_t)lev); break;
                default: DCF_LAUNCH(ctx, "k_union", k_union_tgt<26>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, tl, cnt, (uint32_t)lev); break;
                }
            } else switch (la.nb.n) {
            case 4: DCF_LAUNCH(ctx, "k_union", k_union<4>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, cnt, (uint32_t)lev, (uint32_t)kc); break;
            case 6: DCF_LAUNCH(ctx, "k_union", k_union<6>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, cnt, (uint32_t)lev, (uint32_t)kc); break;
            case 8: DCF_LAUNCH(ctx, "k_union", k_union<8>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, cnt, (uint32_t)lev, (uint32_t)kc); break;
            default: DCF_LAUNCH(ctx, "k_union", k_union<26>, dim3(nblk(cnt, 256)), dim3(256), 0, la, lvl, cnt, (uint32_t)lev, (uint32_t)kc); break;
            }
            /* the grid covers the level's pixels and, striding, the roots absorbed at the level */
            const bool blk = la.pass == 1 && !zones;
            unsigned rb = nblk(cnt, FB), ab = nblk((int64_t)cnt * la.nb.n / 4, FB), chunk = 1;
            /* block-wide aggregation: at most final_blocks blocks, each carrying its totals over chunks */
            if (blk && ctx->final_blocks > 0 && rb > (unsigned)ctx->final_blocks) {
                chunk = (rb + (unsigned)ctx->final_blocks - 1) / (unsigned)ctx->final_blocks;
                rb = (rb + chunk - 1) / chunk;
            }
            if (ab > 1024) ab = 1024;
            if (ab < 1) ab = 1;
            /* rb blocks for the level's elements, then ab blocks striding over the absorbed roots */
    #define DCF_FINAL(GV)                                                                                                        \
        if (blk)                                                                                                                 \
            DCF_LAUNCH(ctx, "k_final_pixels", (k_final_pixels<GV, true>), dim3(rb + ab), dim3(FB), 0, la, lvl, cnt, (uint32_t)lev, A, \
                       S, P, par, spar, nabs + 1 + ord, chunk, rb);                                                              \
        else                                                                                                                     \
            DCF_LAUNCH(ctx, "k_final_pixels", (k_final_pixels<GV, false>), dim3(rb + ab), dim3(FB), 0, la, lvl, cnt, (uint32_t)lev, A, \
                       S, P, par, spar, nabs + 1 + ord, 1u, rb)
            switch (grp) {
            case 0: DCF_FINAL(0); break;
            case 1: DCF_FINAL(1); break;
            case 2: DCF_FINAL(2); break;
            case 3: DCF_FINAL(3); break;
            default: DCF_FINAL(4); break;
            }
    #undef DCF_FINAL
            /* the sets to fix are far fewer than the level's elements: a small strided grid */
            unsigned fb = nblk(cnt, 256);
            if (fb > (unsigned)ctx->fix_blocks) fb = (unsigned)ctx->fix_blocks;
            switch (grp) {
            case 0: DCF_LAUNCH(ctx, "k_fix", k_fix<0>, dim3(fb), dim3(256), 0, la, (uint32_t)ord, (uint32_t)lev, A, S, P, par, spar); break;
            case 1: DCF_LAUNCH(ctx, "k_fix", k_fix<1>, dim3(fb), dim3(256), 0, la, (uint32_t)ord, (uint32_t)lev, A, S, P, par, spar); break;
            case 2: DCF_LAUNCH(ctx, "k_fix", k_fix<2>, dim3(fb), dim3(256), 0, la, (uint32_t)ord, (uint32_t)lev, A, S, P, par, spar); break;
            case 3: DCF_LAUNCH(ctx, "k_fix", k_fix<3>, dim3(fb), dim3(256), 0, la, (uint32_t)ord, (uint32_t)lev, A, S, P, par, spar); break;
            default: DCF_LAUNCH(ctx, "k_fix", k_fix<4>, dim3(fb), dim3(256), 0, la, (uint32_t)ord, (uint32_t)lev, A, S, P, par, spar); break;
            }
            ord++;
        }
        return DCF_OK;
    };
    const bool graph = ctx->graph && !ctx->profiling;
    if (!graph) {
        rc = level_loop();
        ctx->stream = strm[0];
        if (rc != DCF_OK) return rc;
        /* join: the pass's last kernels wait for every loop */
        for (uint32_t g = 1; g < K; g++) {
            hipEvent_t j = nullptr;
            if (hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "event");
            ev.e.push_back(j);
            DCF_HIP(ctx, hipEventRecord(j, strm[g]));
            DCF_HIP(ctx, hipStreamWaitEvent(strm[0], j, 0));
        }
    } else {
        hipStream_t work = ctx->stream;
        /* the capture stream is idle: nothing it records runs before the graph is launched on `work` */
        DCF_HIP(ctx, hipStreamBeginCapture(ctx->own, hipStreamCaptureModeThreadLocal));
        ctx->stream = ctx->own;
        rc = level_loop();
        ctx->stream = work;
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->own, &g);
        if (rc != DCF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "level loop capture: %s", hipGetErrorString(ec));
        hipGraphExec_t ge = nullptr;
        hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(ge, work);
        if (e == hipSuccess) e = hipStreamSynchronize(work);
        if (ge) (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "level loop graph: %s", hipGetErrorString(e));
    }
    DCF_LAUNCH(ctx, "k_held_resolve", k_held_resolve, dim3(nblk(n + kc, 256)), dim3(256), 0, par, n, spar, kc);
    if (zones)
        DCF_LAUNCH(ctx, "k_zone_parents", k_zone_parents, dim3(nblk(n, 256)), dim3(256), 0, la.iv, la.pass, la.shift,
                   (const uint32_t *)zrep, par, n);
    return DCF_OK;
}


/* a batch of small images (every one at most DCF_SMALL_MAX pixels, attribute group 0): the LDS build
 * (k_small_build, about 52 KB of LDS per workgroup, three per CU); the canonical parent array into
 * t->parent, the areas into t->attr_pix */
static int small_build(dcf_tree *t)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t n = t->n;
    const int nb_img = t->nbatch;
    std::vector<BatchImg> tab((size_t)nb_img);
    DCF_HIP(ctx, hipMemcpyAsync(tab.data(), t->bimg, tab.size() * sizeof(BatchImg), hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<uint32_t> list((size_t)nb_img);
    for (int i = 0; i < nb_img; i++) {
        if ((uint64_t)tab[i].W * tab[i].H * tab[i].D > DCF_SMALL_MAX)
            return dcf_set_error(ctx, DCF_EINVAL, "image %d: too large for the LDS build", i);
        list[(size_t)i] = (uint32_t)i;
    }
    Scratch sc{ ctx, {} };
    uint32_t *dl = (uint32_t *)sc.get(list.size() * 4);
    uint32_t *queue = (uint32_t *)sc.get(4);
    if (!dl || !queue) return dcf_set_error(ctx, DCF_ENOMEM, "small-image list");
    DevAttr A;
    int rc = dcf_alloc_attr(ctx, A, 0, n);
    if (rc != DCF_OK) return rc;
    t->attr_pix = A;
    DCF_HIP(ctx, hipMemcpyAsync(dl, list.data(), list.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream)); /* the host list is freed on return */
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32f, dim3(1), dim3(256), 0, queue, 0u, (int64_t)1);
    ImgView iv;
    iv.img = t->img;
    iv.dtype = t->dtype;
    iv.invert = 0;
    iv.gmax = t->gmax;
    const Nbr nb = make_nbr(t->conn, t->batch3d);
    DCF_LAUNCH(ctx, "k_small_build", k_small_build<DCF_SMALL_MAX>, dim3(std::min(nb_img, 768)), dim3(64), 0, iv, t->bimg, dl,
               (uint32_t)nb_img, nb, t->parent, A.area, queue);
    return DCF_OK;
}

int dcf_flood_build(dcf_tree *t, int grp)
{
    if (t->small_batch) {
        if (grp != 0) return dcf_set_error(t->ctx, DCF_EINVAL, "the LDS build accumulates the area only");
        return small_build(t);
    }
    dcf_ctx *ctx = t->ctx;
    const int64_t n = t->n;
    const int is3d = t->bimg ? t->batch3d : t->dims[2] > 1;
    /* union-find ids (pixels + coarse super-vertices) must stay below the root tag bit */
    if (n >= ((int64_t)1 << 30)) return dcf_set_error(ctx, DCF_EUNSUPPORTED, "images are limited to 2^30 pixels");
    Scratch sc{ ctx, {} };
    LevelArgs la;
    la.g.W = (int)t->dims[0];
    la.g.H = (int)t->dims[1];
    la.g.D = (int)t->dims[2];
    la.g.n = n;
    la.nb = make_nbr(t->conn, is3d);
    la.bimg = t->bimg;
    la.nbatch = t->nbatch;
    la.zone = 0;
    la.ox = t->off[0];
    la.oy = t->off[1];
    la.oz = t->off[2];
    la.cnode = nullptr; la.lift = nullptr; la.cnkey = nullptr; la.cpd = nullptr; la.nlift = 0;
    int rc = DCF_OK;
    /* the keys, in passes of 8-bit digits: u8 one pass, u16 two; a float image is built on the dense
     * ranks of its float keys (the same order, the same equalities), 1 to 4 passes */
    int P;
    uint32_t maxkey;
    if (t->dtype == DCF_F32) {
        uint32_t *rank = (uint32_t *)sc.get((size_t)n * 4);
        if (!rank) return dcf_set_error(ctx, DCF_ENOMEM, "grey ranks");
        uint32_t maxr = 0;
        if ((rc = dcf_rank_keys(t, rank, &maxr)) != DCF_OK) return rc;
        la.iv.img = rank;
        la.iv.dtype = DCF_KEY32;
        la.iv.invert = 0;
        la.iv.gmax = nullptr;
        P = maxr < 256u ? 1 : (maxr < 65536u ? 2 : (maxr < (1u << 24) ? 3 : 4));
        maxkey = maxr;
    } else {
        la.iv.img = t->img;
        la.iv.dtype = t->dtype;
        la.iv.invert = t->tree_min;
        la.iv.gmax = t->gmax;
        P = t->dtype == DCF_U16 ? 2 : 1;
        maxkey = t->dtype == DCF_U16 ? 65535u : 255u;
    }

    /* pixel-indexed attribute records (valid at roots) */
    DevAttr A;
    if ((rc = alloc_attr_n(ctx, &sc, A, grp, n))) return rc;
    la.uf.p = (uint32_t *)sc.get((size_t)n * 4);
    la.uf.s = 1;
    la.absorbed = (uint32_t *)sc.get((size_t)(n + 16) * 4);
    la.nabs = (uint32_t *)sc.get((NLEV + 4) * 4);
    int32_t *par = P == 1 ? t->parent : (int32_t *)sc.get((size_t)n * 4);
    if (!la.uf.p || !la.absorbed || !la.nabs || !par) return dcf_set_error(ctx, DCF_ENOMEM, "flood scratch");

    /* ---- pass 1: the top digit over all pixels ---- */
    la.pass = 1;
    la.shift = 8 * (P - 1);
    if ((rc = clear_attr(ctx, A, n))) return rc;
    DCF_LAUNCH(ctx, "k_uf_init", k_uf_init, dim3(nblk(n, 256)), dim3(256), 0, la.iv, 1, la.shift, n, n, la.uf);
    DCF_LAUNCH(ctx, "k_fill", k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, par, -1, n);
    DevAttr none;
    /* zone graph for the first pass (in a batch every image has tiles of its own) */
    const int64_t edge_bytes = n * la.nb.n * (int64_t)sizeof(uint2);
    const bool zones = edge_bytes <= ((int64_t)8 << 30) &&
                       (ctx->zones == 1 || (ctx->zones == 2 && (P > 1 || !is3d)));
    /* union targets precomputed by the level scatter, when they fit the budget: for batches (each
     * neighbour lookup of a batch locates its image), not for single images, whose neighbour keys
     * mostly hit the caches and whose target stream costs more than the lookups it saves */
    uint32_t *tgt = nullptr;
    const bool use_tgt = ctx->targets == 1 || (ctx->targets == 2 && t->bimg);
    if (use_tgt && (int64_t)n * la.nb.n * 4 <= ((int64_t)16 << 30)) {
        tgt = (uint32_t *)sc.get((size_t)n * la.nb.n * 4);
        if (!tgt) return dcf_set_error(ctx, DCF_ENOMEM, "union targets");
    }
    if ((rc = run_pass(ctx, la, sc, A, none, grp, par, nullptr, 0, zones, tgt, is3d, t->batch_tiles))) return rc;

    /* ---- passes 2..P: the next digit inside every node of the previous pass's tree ---- */
    for (int pass = 2; pass <= P; pass++) {
        const int pshift = 8 * (P - pass + 1); /* the previous pass's levels: key >> pshift */
        uint32_t *flag = (uint32_t *)sc.get((size_t)n * 4);
        uint32_t *cidx = (uint32_t *)sc.get((size_t)n * 4);
        uint32_t *cnode = (uint32_t *)sc.get((size_t)n * 4);
        if (!flag || !cidx || !cnode) return dcf_set_error(ctx, DCF_ENOMEM, "coarse arrays");
        DCF_LAUNCH(ctx, "k_coarse_flag", k_coarse_flag, dim3(nblk(n, 256)), dim3(256), 0, la.iv, (const int32_t *)par, flag, pshift, n);
        size_t need = 0;
        DCF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, flag, cidx, (int)n, ctx->stream));
        void *tmp = sc.get(need > 0 ? need : 16);
        if (!tmp) return dcf_set_error(ctx, DCF_ENOMEM, "scan scratch");
        DCF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, need, flag, cidx, (int)n, ctx->stream));
        uint32_t last[2];
        DCF_HIP(ctx, hipMemcpyAsync(&last[0], cidx + (n - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipMemcpyAsync(&last[1], flag + (n - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const int64_t kc = (int64_t)last[0] + last[1];
        if (n + kc >= ((int64_t)1 << 31)) return dcf_set_error(ctx, DCF_EUNSUPPORTED, "too many coarse nodes");
        uint32_t *cnkey = (uint32_t *)sc.get((size_t)kc * 4);
        uint32_t *cpix = (uint32_t *)sc.get((size_t)kc * 4);
        /* the coarse tree's depth is at most the number of distinct prefixes (maxkey >> pshift) + 1 */
        int nlift = 1;
        while (nlift < 32 && ((uint64_t)(maxkey >> pshift) >> nlift) != 0) nlift++;
        int32_t *lift = (int32_t *)sc.get((size_t)kc * 4 * nlift);
        int32_t *spar = (int32_t *)sc.get((size_t)kc * 4);
        if (!cnkey || !cpix || !lift || !spar) return dcf_set_error(ctx, DCF_ENOMEM, "coarse node arrays");
        DevAttr S;
        if ((rc = alloc_attr_n(ctx, &sc, S, grp, kc))) return rc;
        DCF_LAUNCH(ctx, "k_coarse_nodes", k_coarse_nodes, dim3(nblk(n, 256)), dim3(256), 0, la.iv, (const int32_t *)par, flag, cidx,
                   cnkey, lift, cpix, A, S, pshift, n);
        for (int j = 1; j < nlift; j++) DCF_LAUNCH(ctx, "k_lift", k_lift, dim3(nblk(kc, 256)), dim3(256), 0, lift, j, kc);
        uint2 *cpd = (uint2 *)sc.get((size_t)kc * 8);
        if (!cpd) return dcf_set_error(ctx, DCF_ENOMEM, "coarse depths");
        DCF_LAUNCH(ctx, "k_cdepth", k_cdepth, dim3(nblk(kc, 256)), dim3(256), 0, (const int32_t *)lift, nlift, kc, cpd);
        DCF_LAUNCH(ctx, "k_coarse_node_of", k_coarse_node_of, dim3(nblk(n, 256)), dim3(256), 0, (const int32_t *)par, flag, cidx,
                   cnode, n);
        DCF_LAUNCH(ctx, "k_fill", k_fill_i32, dim3(nblk(kc, 256)), dim3(256), 0, spar, -1, kc);

        /* union-find of the pass: 16-byte records {entry, key, coarse node, 0} (DCF_RECORDS), or entries */
        const bool recs = ctx->records != 0;
        uint32_t *uf2 = (uint32_t *)sc.get((size_t)(n + kc) * (recs ? 16 : 4));
        uint32_t *abs2 = (uint32_t *)sc.get((size_t)(n + kc + 16) * 4);
        int32_t *npar = pass == P ? t->parent : (int32_t *)sc.get((size_t)n * 4);
        if (!uf2 || !abs2 || !npar) return dcf_set_error(ctx, DCF_ENOMEM, "fine union-find");
        la.uf.p = uf2;
        la.uf.s = recs ? 4 : 1;
        la.absorbed = abs2;
        la.pass = 2;
        la.shift = 8 * (P - pass);
        la.cnode = cnode;
        la.lift = lift;
        la.cnkey = cnkey;
        la.cpd = cpd;
        la.nlift = nlift;
        if (recs)
            DCF_LAUNCH(ctx, "k_uf_init", k_rec_init, dim3(nblk(n + kc, 256)), dim3(256), 0, la.iv, la.shift, (const uint32_t *)cnode, n,
                       n + kc, (uint4 *)uf2);
        else
            DCF_LAUNCH(ctx, "k_uf_init", k_uf_init, dim3(nblk(n + kc, 256)), dim3(256), 0, la.iv, 2, la.shift, n, n + kc, la.uf);
        DCF_LAUNCH(ctx, "k_fill", k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, npar, -1, n);
        if ((rc = clear_attr(ctx, A, n))) return rc;
        if ((rc = run_pass(ctx, la, sc, A, S, grp, npar, spar, kc, false, tgt, is3d, t->batch_tiles))) return rc;
        DCF_LAUNCH(ctx, "k_fine_roots", k_fine_roots, dim3(nblk(kc, 256)), dim3(256), 0, la.uf, (const uint32_t *)la.nrep, cpix, spar,
                   npar, kc);
        par = npar;
    }

    t->attr_pix = A;
    /* hand the attribute arrays over to the tree: detach them from the scratch list */
    for (auto it = sc.ptrs.begin(); it != sc.ptrs.end();) {
        bool keep = *it == A.area;
        for (int i = 0; i < A.nsum; i++) keep = keep || *it == A.s[i];
        for (int i = 0; A.has_ext && i < 3; i++) keep = keep || *it == A.mn[i] || *it == A.mx[i];
        if (keep) it = sc.ptrs.erase(it);
        else ++it;
    }
    return DCF_OK;
}
