/*
 * dcf_build.hip -- node arrays of a component tree, and the float-image keys.
 *
 * The tree itself comes from the level-synchronous build of dcf_flood.hip (canonical parent
 * array and pixel-indexed attribute records at the levelroots).  Here:
 *
 *   k_flags        levelroot flags (the node test of the reference's compression loop,
 *                  src/sources/flood.c:127-141)
 *   scan           node numbering (hipcub exclusive sum of the flags)
 *   k_nodes        node parent / grey, pixel -> node map
 *   k_gather_attr  node-indexed attribute records
 *   k_rank_*       float images: dense ranks of the float keys (radix sort, new-key flags, scan),
 *                  the order transform_float gives them (flood.c:362-371); the build runs on the
 *                  ranks in 8-bit digit passes like a 16- or 32-bit integer image.
 */
#include <hipcub/hipcub.hpp>

#include "dcf_internal.h"

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

/* ------------------------------------------------------------------------------------------------ */
/* dense ranks of float keys                                                                        */
/* ------------------------------------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void k_rank_in(ImgView iv, uint32_t *__restrict__ key, uint32_t *__restrict__ idx, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    key[p] = px_key(iv, p);
    idx[p] = (uint32_t)p;
}

/* 1 where a sorted key differs from its predecessor */
__global__ __launch_bounds__(256) void k_rank_new(const uint32_t *__restrict__ skey, uint32_t *__restrict__ neu, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    neu[i] = (i > 0 && skey[i] != skey[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_rank_out(const uint32_t *__restrict__ sidx, const uint32_t *__restrict__ r,
                                                  uint32_t *__restrict__ rank, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rank[sidx[i]] = r[i];
}

/* rank[p] = number of distinct float keys below pixel p's: the same order and the same equalities
 * as the keys, in 0..distinct-1, so a tree built on the ranks is the tree of the float image */
int dcf_rank_keys(dcf_tree *t, uint32_t *rank, uint32_t *maxrank)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t n = t->n;
    ImgView iv = view_of(t);
    struct Held {
        dcf_ctx *ctx;
        void *p[5];
        ~Held() { for (void *q : p) dcf_release(ctx, q); }
    } h{ ctx, {} };
    uint32_t *key = (uint32_t *)(h.p[0] = dcf_alloc(ctx, (size_t)n * 4));
    uint32_t *idx = (uint32_t *)(h.p[1] = dcf_alloc(ctx, (size_t)n * 4));
    uint32_t *skey = (uint32_t *)(h.p[2] = dcf_alloc(ctx, (size_t)n * 4));
    uint32_t *sidx = (uint32_t *)(h.p[3] = dcf_alloc(ctx, (size_t)n * 4));
    if (!key || !idx || !skey || !sidx) return dcf_set_error(ctx, DCF_ENOMEM, "rank arrays");
    size_t need_sort = 0, need_scan = 0;
    DCF_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, need_sort, key, skey, idx, sidx, (int)n, 0, 32, ctx->stream));
    DCF_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, need_scan, key, idx, (int)n, ctx->stream));
    void *tmp = h.p[4] = dcf_alloc(ctx, std::max(std::max(need_sort, need_scan), (size_t)16));
    if (!tmp) return dcf_set_error(ctx, DCF_ENOMEM, "rank sort scratch");
    DCF_LAUNCH(ctx, "k_rank_in", k_rank_in, dim3(nblk(n, 256)), dim3(256), 0, iv, key, idx, n);
    DCF_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, need_sort, key, skey, idx, sidx, (int)n, 0, 32, ctx->stream));
    /* key and idx are free again: new-key flags into key, their inclusive sum (the ranks) into idx */
    DCF_LAUNCH(ctx, "k_rank_new", k_rank_new, dim3(nblk(n, 256)), dim3(256), 0, skey, key, n);
    DCF_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp, need_scan, key, idx, (int)n, ctx->stream));
    DCF_LAUNCH(ctx, "k_rank_out", k_rank_out, dim3(nblk(n, 256)), dim3(256), 0, sidx, idx, rank, n);
    DCF_HIP(ctx, hipMemcpyAsync(maxrank, idx + (n - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DCF_OK;
}

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

int dcf_build_tree(dcf_tree *t, int grp) { return flood_and_gather(t, grp, true); }

/* node arrays of a tree whose canonical parent array is already in t->parent (a tiled build's merged
 * tree): levelroot flags, node numbering, node parent / grey; the flags go to the caller (*flag_out,
 * released with dcf_release) */
int dcf_tree_nodes_from_parent(dcf_tree *t, uint32_t **flag_out)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t n = t->n;
    *flag_out = nullptr;
    uint32_t *flag = (uint32_t *)dcf_alloc(ctx, (size_t)n * 4);
    if (!flag) return dcf_set_error(ctx, DCF_ENOMEM, "levelroot flags");
    ImgView iv = view_of(t);
    DCF_LAUNCH(ctx, "k_flags", k_flags, dim3(nblk(n, 256)), dim3(256), 0, iv, t->parent, flag, n);
    /* compact_nodes only reads the flags */
    const int rc = compact_nodes(t, flag);
    if (rc != DCF_OK) {
        dcf_release(ctx, flag);
        return rc;
    }
    *flag_out = flag;
    return DCF_OK;
}

int dcf_alloc_attr(dcf_ctx *ctx, DevAttr &a, int grp, int64_t K) { return alloc_attr(ctx, a, grp, K); }

int dcf_accumulate_group(dcf_tree *t, int grp) { return flood_and_gather(t, grp, false); }

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

