/*
 * dcf_api.cpp -- the C ABI of libdcf: contexts, trees, operations and the operator driver.
 *
 * Host-side mirror of the reference's operator surface:
 *   AttribsArray / Decisions tables         src/sources/attributes.c:1224-1248, filtering.c:271-277
 *   connectivity mapping                     src/sources/io.c:64-79
 *   min-tree inversion / output re-inversion src/sources/preprocessing.c:84-91, operation.c:23-28
 *   operation tuples                         src/sources/config.c:297-433
 *   operation loop with group reuse          src/sources/disccofan.c:137-205, operation.c:110-212
 *   lambda-vector files                      src/sources/lambdavec.c:35-84
 * Errors are status codes with a per-context message (the reference aborts through MPI_Abort).
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <exception>
#include <string>
#include <vector>

#include <algorithm>
#include <cstdint>
#include <atomic>
#include <functional>
#include <memory>
#include <thread>

#include "dcf_internal.h"

/* ------------------------------------------------------------------------------------------------ */
/* tables                                                                                           */
/* ------------------------------------------------------------------------------------------------ */
static const char *kAttrName[DCF_NUM_ATTR] = {
    "Area", "Area of minimal enclosing rectangle", "Square of diagonal of minimal enclosing rectangle", "X-extent",
    "Y-extent", "Z-extent", "Mean X position", "Mean Y position", "Mean Z position", "Weighted Mean X position",
    "Weighted Mean Y position", "Weighted Mean Z position", "Total flux", "Trace of Inertia Matrice", "Elongation",
    "Flatness", "Sparseness", "Non-compactnesss"
};
static const char *kAttrShort[DCF_NUM_ATTR] = { "area", "rectarea", "rectdiag", "xext", "yext", "zext",
                                                "xmean", "ymean", "zmean", "xwmean", "ywmean", "zwmean",
                                                "flux", "inertiatrace", "elong", "flat", "spars", "ncomp" };
static const char *kDecision[4] = { "Direct", "Min", "Max", "Subtractive" };

extern "C" const char *dcf_attr_name(int i) { return (i >= 0 && i < DCF_NUM_ATTR) ? kAttrName[i] : nullptr; }
extern "C" const char *dcf_attr_short_name(int i) { return (i >= 0 && i < DCF_NUM_ATTR) ? kAttrShort[i] : nullptr; }
extern "C" const char *dcf_decision_name(int d) { return (d >= 0 && d < 4) ? kDecision[d] : nullptr; }

int dcf_group_of(int idx)
{
    if (idx < 0 || idx >= DCF_NUM_ATTR) return -1;
    if (idx == 0) return 0;
    if (idx <= 5) return 1;
    if (idx <= 8) return 2;
    if (idx <= 12) return 3;
    return 4;
}

extern "C" int dcf_attr_group(int i) { return dcf_group_of(i); }

/* ------------------------------------------------------------------------------------------------ */
/* errors, profiling, allocator                                                                     */
/* ------------------------------------------------------------------------------------------------ */
int dcf_set_error(dcf_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        try {
            ctx->err = buf;
        } catch (const std::exception &) { /* no memory for the message: the code still goes back */
        }
    }
    return code;
}

static hipEvent_t take_event(dcf_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void dcf_prof_begin(dcf_ctx *ctx, const char *name, hipEvent_t *a, hipEvent_t *b)
{
    (void)name;
    *a = take_event(ctx);
    *b = take_event(ctx);
    if (!*a || !*b) { /* one of the pair could not be created: this launch goes untimed, the other waits in the pool */
        if (*a) ctx->event_pool.push_back(*a);
        if (*b) ctx->event_pool.push_back(*b);
        *a = *b = nullptr;
        return;
    }
    (void)hipEventRecord(*a, ctx->stream);
}

void dcf_prof_end(dcf_ctx *ctx, const char *name, hipEvent_t a, hipEvent_t b)
{
    if (!a || !b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({ name, a, b });
}

/* fold recorded event pairs into the per-kernel totals (called after a synchronisation) */
static void prof_collect(dcf_ctx *ctx)
{
    for (auto &p : ctx->pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            KernelStat &s = ctx->stats[p.name];
            if (s.launches == 0) ctx->names.push_back(p.name);
            s.launches++;
            s.ms += ms;
        }
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
}

void *dcf_alloc(dcf_ctx *ctx, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    bytes = (bytes + 255) & ~(size_t)255;
    auto it = ctx->free_blocks.lower_bound(bytes);
    if (it != ctx->free_blocks.end() && it->first <= bytes * 2) {
        void *p = it->second;
        ctx->live_blocks[p] = it->first;
        ctx->free_blocks.erase(it);
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        /* drop the cache and retry once */
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &f : ctx->free_blocks) (void)hipFree(f.second);
        ctx->free_blocks.clear();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    ctx->live_blocks[p] = bytes;
    return p;
}

void dcf_release(dcf_ctx *ctx, void *p)
{
    if (!p) return;
    auto it = ctx->live_blocks.find(p);
    if (it == ctx->live_blocks.end()) return;
    ctx->free_blocks.insert({ it->second, p });
    ctx->live_blocks.erase(it);
}

/* ------------------------------------------------------------------------------------------------ */
/* context                                                                                          */
/* ------------------------------------------------------------------------------------------------ */
/* one tunable of the build, by its environment name; DCF_EINVAL for an unknown name or a value out of range */
static int set_option(dcf_ctx *c, const char *name, const char *value)
{
    if (!name || !value) return DCF_EINVAL;
    auto shape3 = [&](int v[3]) -> int {
        int a = 0, b = 0, d = 0;
        if (sscanf(value, "%dx%dx%d", &a, &b, &d) != 3 || a <= 0 || b <= 0 || d <= 0 || a > (1 << 20) || b > (1 << 20) ||
            d > (1 << 20) || (int64_t)a * b * d > (1 << 20))
            return DCF_EINVAL;
        v[0] = a; v[1] = b; v[2] = d;
        return DCF_OK;
    };
    auto integer = [&](int lo, int hi, int &out) -> int {
        char *end = nullptr;
        const long v = strtol(value, &end, 10);
        if (end == value || *end || v < lo || v > hi) return DCF_EINVAL;
        out = (int)v;
        return DCF_OK;
    };
    const std::string n(name);
    if (n == "DCF_ZONES") return integer(0, 2, c->zones);
    if (n == "DCF_TARGETS") return integer(0, 2, c->targets);
    if (n == "DCF_RECORDS") return integer(0, 1, c->records);
    if (n == "DCF_FIX_BLOCKS") return integer(1, 65536, c->fix_blocks);
    if (n == "DCF_GRAPH") return integer(0, 1, c->graph);
    if (n == "DCF_ZIP_SEGMENTS") return integer(1, 4096, c->zip_segments);
    if (n == "DCF_WORKERS") return integer(1, 16, c->max_workers);
    if (n == "DCF_BATCH_PARTS") return integer(0, 64, c->batch_parts);
    if (n == "DCF_FINAL_BLOCKS") return integer(0, 1 << 20, c->final_blocks);
    if (n == "DCF_FINE_LOOPS") return integer(1, 4, c->fine_loops);
    if (n == "DCF_WORKER_LOOPS") return integer(0, 4, c->worker_loops);
    if (n == "DCF_SMALL_PIXELS") {
        int v = 0;
        if (integer(0, DCF_SMALL_MAX, v) != DCF_OK) return DCF_EINVAL;
        c->small_pixels = v;
        return DCF_OK;
    }
    if (n == "DCF_LARGE_PIXELS") {
        int v = 0;
        if (integer(1, 0x7FFFFFFF, v) != DCF_OK) return DCF_EINVAL;
        c->large_pixels = v;
        return DCF_OK;
    }
    if (n == "DCF_ZONE_TILE2D") {
        int a = 0, b = 0;
        if (sscanf(value, "%dx%d", &a, &b) != 2 || a <= 0 || b <= 0 || a > ZONE_TILE_MAX || b > ZONE_TILE_MAX ||
            (int64_t)a * b > ZONE_TILE_MAX)
            return DCF_EINVAL;
        c->zt2[0] = a;
        c->zt2[1] = b;
        c->zt2_set = 1;
        return DCF_OK;
    }
    if (n == "DCF_ZONE_TILE3D") {
        int a = 0;
        if (integer(1, ZONE_TILE_MAX, a) != DCF_OK || (int64_t)a * a * a > ZONE_TILE_MAX) return DCF_EINVAL;
        c->zt3 = a;
        return DCF_OK;
    }
    if (n == "DCF_LEVEL_TILE2D") return shape3(c->lt2);
    if (n == "DCF_LEVEL_GROUP2D") return shape3(c->lg2);
    if (n == "DCF_LEVEL_TILE3D") return shape3(c->lt3);
    if (n == "DCF_LEVEL_GROUP3D") return shape3(c->lg3);
    return DCF_EINVAL;
}

extern "C" int dcf_ctx_set_option(dcf_ctx *ctx, const char *name, const char *value)
try {
    if (!ctx) return DCF_EINVAL;
    const int rc = set_option(ctx, name, value);
    if (rc != DCF_OK) return dcf_set_error(ctx, rc, "bad option %s=%s", name ? name : "(null)", value ? value : "(null)");
    return DCF_OK;
} catch (const std::exception &e) {
    return dcf_set_error(ctx, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

/* the streams of a context's further level loops, created up to n */
int dcf_add_aux_streams(dcf_ctx *c, size_t n)
{
    while (c->aux.size() < n) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return DCF_EDEVICE;
        c->aux.push_back(st);
    }
    return DCF_OK;
}

/* a caller's context, or a worker context (no loop streams of its own yet: dcf_run_on_workers creates them) */
static int ctx_create(int device, bool worker, dcf_ctx **out)
try {
    if (!out) return DCF_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DCF_EDEVICE;
    if (device < 0 || device >= ndev) return DCF_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return DCF_EDEVICE;
    dcf_ctx *c = new dcf_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return DCF_EDEVICE;
    }
    c->stream = c->own;
    /* concurrent worker streams run side by side only on hardware queues of their own */
    if (const char *q = getenv("GPU_MAX_HW_QUEUES")) {
        const long v = strtol(q, nullptr, 10);
        if (v > 0) c->hw_queues = (int)v;
    }
    if (c->hw_queues >= 8) c->max_workers = 4;
    /* the build's tunables: the environment gives a new context its defaults (dcf_ctx_set_option
     * changes them afterwards) */
    static const char *const kOptions[] = { "DCF_ZONES", "DCF_TARGETS", "DCF_RECORDS", "DCF_FIX_BLOCKS", "DCF_GRAPH",
                                            "DCF_ZIP_SEGMENTS", "DCF_WORKERS", "DCF_BATCH_PARTS", "DCF_FINAL_BLOCKS", "DCF_SMALL_PIXELS", "DCF_FINE_LOOPS", "DCF_WORKER_LOOPS", "DCF_LARGE_PIXELS", "DCF_ZONE_TILE2D",
                                            "DCF_ZONE_TILE3D", "DCF_LEVEL_TILE2D", "DCF_LEVEL_GROUP2D", "DCF_LEVEL_TILE3D",
                                            "DCF_LEVEL_GROUP3D" };
    for (const char *name : kOptions)
        if (const char *v = getenv(name)) (void)set_option(c, name, v);
    /* the streams of the refining pass's further loops, right after the context's own.  The queues of a
     * process are spread over the command processor's pipes in the order they are created (in turn over
     * four, as measured), and kernels of queues that share a pipe launch one after the other: streams
     * created together get pipes of their own whatever the process created before */
    if (!worker && c->fine_loops > 1 && dcf_add_aux_streams(c, (size_t)std::min(c->fine_loops, 4) - 1) != DCF_OK) {
        (void)dcf_ctx_destroy(c);
        return DCF_EDEVICE;
    }
    /* the allocation-free entry points' scratch address on this device (their calls fail with
     * DCF_EDEVICE when it is unavailable; the context's other calls do not need it) */
    (void)dcf_noalloc_init(c);
    c->err.clear();
    *out = c;
    return DCF_OK;
} catch (const std::exception &) {
    return DCF_ENOMEM;
}

extern "C" int dcf_ctx_create(int device, unsigned flags, dcf_ctx **out)
{
    (void)flags;
    return ctx_create(device, false, out);
}

static void tree_release(dcf_tree *t);

extern "C" int dcf_ctx_destroy(dcf_ctx *ctx)
{
    if (!ctx) return DCF_EINVAL;
    for (dcf_ctx *w : ctx->workers) (void)dcf_ctx_destroy(w);
    ctx->workers.clear();
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (dcf_tree *t : ctx->trees) {
        tree_release(t);
        delete t;
    }
    ctx->trees.clear();
    prof_collect(ctx);
    for (auto &e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto &f : ctx->free_blocks) (void)hipFree(f.second);
    for (auto &l : ctx->live_blocks) (void)hipFree(l.first);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    for (hipStream_t st : ctx->aux) (void)hipStreamDestroy(st);
    delete ctx;
    return DCF_OK;
}

extern "C" int dcf_ctx_set_stream(dcf_ctx *ctx, void *stream)
{
    if (!ctx) return DCF_EINVAL;
    ctx->stream = stream ? (hipStream_t)stream : ctx->own;
    return DCF_OK;
}

extern "C" int dcf_ctx_set_graph(dcf_ctx *ctx, int on)
{
    if (!ctx) return DCF_EINVAL;
    ctx->graph = on != 0;
    return DCF_OK;
}

extern "C" void *dcf_ctx_stream(dcf_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }
extern "C" const char *dcf_last_error(dcf_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int dcf_ctx_synchronize(dcf_ctx *ctx)
{
    if (!ctx) return DCF_EINVAL;
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DCF_OK;
}

extern "C" int dcf_ctx_set_profiling(dcf_ctx *ctx, int on)
{
    if (!ctx) return DCF_EINVAL;
    ctx->profiling = on ? 1 : 0;
    return DCF_OK;
}

extern "C" int dcf_ctx_profile_only(dcf_ctx *ctx, const char *name)
{
    if (!ctx) return DCF_EINVAL;
    ctx->prof_only = name ? name : "";
    return DCF_OK;
}

extern "C" int dcf_kernel_stats(dcf_ctx *ctx, const char *name, int64_t *launches, double *total_ms)
{
    if (!ctx || !name) return DCF_EINVAL;
    (void)hipStreamSynchronize(ctx->stream);
    prof_collect(ctx);
    int64_t n = 0;
    double ms = 0;
    for (auto &s : ctx->stats)
        if (strstr(s.first.c_str(), name)) { n += s.second.launches; ms += s.second.ms; }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    return DCF_OK;
}

extern "C" int dcf_kernel_stats_reset(dcf_ctx *ctx)
{
    if (!ctx) return DCF_EINVAL;
    (void)hipStreamSynchronize(ctx->stream);
    prof_collect(ctx);
    ctx->stats.clear();
    ctx->names.clear();
    return DCF_OK;
}

extern "C" const char *dcf_kernel_name(dcf_ctx *ctx, int i)
{
    if (!ctx || i < 0) return nullptr;
    (void)hipStreamSynchronize(ctx->stream);
    prof_collect(ctx);
    return i < (int)ctx->names.size() ? ctx->names[i].c_str() : nullptr;
}

/* ------------------------------------------------------------------------------------------------ */
/* trees                                                                                            */
/* ------------------------------------------------------------------------------------------------ */
static size_t dtype_size(dcf_dtype d) { return d == DCF_U8 ? 1 : (d == DCF_U16 ? 2 : 4); }

/* pixels of an image whose dimensions are each below 2^31, without wrap-around: the product of the
 * first two is checked before the third multiplies it (both partial products stay below 2^62);
 * UINT64_MAX when the count reaches 2^31 - 1 */
static uint64_t pixel_count(const uint64_t d[3])
{
    if (d[0] > 0x7FFFFFFFull || d[1] > 0x7FFFFFFFull || d[2] > 0x7FFFFFFFull) return UINT64_MAX;
    const uint64_t a = d[0] * d[1];
    if (a >= 0x7FFFFFFFull) return UINT64_MAX;
    const uint64_t n = a * d[2];
    return n >= 0x7FFFFFFFull ? UINT64_MAX : n;
}

static void tree_release(dcf_tree *t)
{
    dcf_ctx *ctx = t->ctx;
    if (t->own_img) dcf_release(ctx, t->img);
    dcf_release(ctx, t->gmax);
    dcf_release(ctx, t->parent);
    dcf_release(ctx, t->p2n);
    dcf_release(ctx, t->npar);
    dcf_release(ctx, t->ngrey);
    dcf_release(ctx, t->counters);
    dcf_free_attr(ctx, t->attr);
    dcf_free_attr(ctx, t->attr_pix);
}

/* a tree under construction: released (with its device arrays) unless handed to the caller, also
 * when an exception unwinds */
struct TreeDel {
    void operator()(dcf_tree *t) const
    {
        if (!t) return;
        t->bimg = nullptr; /* a batch's image table belongs to the batch call */
        tree_release(t);
        delete t;
    }
};
using TreePtr = std::unique_ptr<dcf_tree, TreeDel>;

/* device blocks of one call, given back to the context's cache when the call returns */
struct DevBlocks {
    dcf_ctx *ctx;
    std::vector<void *> p;
    template <class T> T *add(T *q)
    {
        if (q) p.push_back((void *)q);
        return q;
    }
    ~DevBlocks()
    {
        for (void *q : p) dcf_release(ctx, q);
    }
};

extern "C" int dcf_tree_free(dcf_tree *t)
{
    if (!t) return DCF_EINVAL;
    (void)hipSetDevice(t->ctx->device);
    t->ctx->trees.erase(t);
    tree_release(t);
    delete t;
    return DCF_OK;
}

/* the argument checks and set-up shared by the single and the tiled build: a tree holding the image
 * (uploaded unless DCF_DEVICE_PTR), its parent and pixel -> node arrays and, for a min-tree, its grey
 * extremes */
static int tree_setup(dcf_ctx *ctx, const void *img, dcf_dtype dtype, const uint64_t dims[3], int connectivity, int tree_type,
                      int attr_idx, const float pixdim[3], unsigned flags, TreePtr &t)
{
    if (!img || !dims) return dcf_set_error(ctx, DCF_EINVAL, "null argument");
    if (dtype != DCF_U8 && dtype != DCF_U16 && dtype != DCF_F32) return dcf_set_error(ctx, DCF_EINVAL, "bad dtype %d", (int)dtype);
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0) return dcf_set_error(ctx, DCF_EINVAL, "empty image");
    const uint64_t n = pixel_count(dims);
    if (n == UINT64_MAX) return dcf_set_error(ctx, DCF_EINVAL, "image of %llux%llux%llu pixels: the count exceeds 2^31-1",
                                              (unsigned long long)dims[0], (unsigned long long)dims[1], (unsigned long long)dims[2]);
    if (tree_type != DCF_MAX_TREE && tree_type != DCF_MIN_TREE) return dcf_set_error(ctx, DCF_EINVAL, "bad tree type %d", tree_type);
    if (attr_idx < 0 || attr_idx >= DCF_NUM_ATTR) return dcf_set_error(ctx, DCF_EINVAL, "bad attribute index %d", attr_idx);
    /* check_connectivity (io.c:64-79) */
    int conn = connectivity;
    bool is3d = dims[2] > 1;
    if (is3d && conn == 4) conn = 6;
    else if (conn != 4 && conn != 6 && conn != 8 && conn != 26)
        return dcf_set_error(ctx, DCF_EINVAL, "wrong connectivity %d (must be 4 or 8 for 2D, 6 or 26 for 3D)", connectivity);
    else if (conn == 8 && is3d) conn = 26;
    else if (conn == 26 && !is3d) conn = 8;
    if (!is3d && conn == 6) conn = 4;
    if (dtype == DCF_F32 && dcf_group_of(attr_idx) == 3)
        return dcf_set_error(ctx, DCF_EUNSUPPORTED, "grey-weighted attributes on float images are not implemented");
    if (hipSetDevice(ctx->device) != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "hipSetDevice");

    t.reset(new dcf_tree());
    t->ctx = ctx;
    memcpy(t->dims, dims, sizeof t->dims);
    t->n = (int64_t)n;
    t->conn = conn;
    t->tree_min = tree_type == DCF_MIN_TREE;
    t->dtype = dtype;
    if (pixdim) memcpy(t->pixdim, pixdim, sizeof t->pixdim);
    size_t isz = dtype_size(dtype);
    if (flags & DCF_DEVICE_PTR) {
        t->img = (void *)img;
        t->own_img = 0;
    } else {
        t->img = dcf_alloc(ctx, (size_t)n * isz);
        t->own_img = 1;
        if (!t->img) return dcf_set_error(ctx, DCF_ENOMEM, "image copy");
        if (hipMemcpyAsync(t->img, img, (size_t)n * isz, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            return dcf_set_error(ctx, DCF_EDEVICE, "image upload");
    }
    t->gmax = (float *)dcf_alloc(ctx, 16);
    t->counters = (uint32_t *)dcf_alloc(ctx, 64 * 4);
    t->parent = (int32_t *)dcf_alloc(ctx, (size_t)n * 4);
    t->p2n = (uint32_t *)dcf_alloc(ctx, (size_t)n * 4);
    if (!t->gmax || !t->counters || !t->parent || !t->p2n) return dcf_set_error(ctx, DCF_ENOMEM, "tree arrays");
    /* find_extremas (preprocessing.c:7-30): only a min-tree reads gmax */
    if (t->tree_min) return dcf_ops_minmax(t.get());
    return DCF_OK;
}

extern "C" int dcf_tree_build(dcf_ctx *ctx, const void *img, dcf_dtype dtype, const uint64_t dims[3], int connectivity,
                              int tree_type, int attr_idx, const float pixdim[3], unsigned flags, dcf_tree **out)
try {
    if (!ctx) return DCF_EINVAL;
    if (!out) return dcf_set_error(ctx, DCF_EINVAL, "null argument");
    *out = nullptr;
    TreePtr t;
    int rc = tree_setup(ctx, img, dtype, dims, connectivity, tree_type, attr_idx, pixdim, flags, t);
    if (rc != DCF_OK) return rc;
    if ((rc = dcf_build_tree(t.get(), dcf_group_of(attr_idx))) != DCF_OK) return rc;
    ctx->trees.insert(t.get());
    *out = t.release();
    return DCF_OK;
} catch (const std::exception &e) {
    return dcf_set_error(ctx, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

extern "C" int dcf_tree_build_tiled(dcf_ctx *ctx, const void *img, dcf_dtype dtype, const uint64_t dims[3], int connectivity,
                                    int tree_type, int attr_idx, const float pixdim[3], const int grid[3], unsigned flags,
                                    dcf_tree **out)
try {
    if (!ctx) return DCF_EINVAL;
    if (!out || !grid) return dcf_set_error(ctx, DCF_EINVAL, "null argument");
    *out = nullptr;
    if (dims) {
        for (int a = 0; a < 3; a++)
            if (grid[a] < 1 || (uint64_t)grid[a] > dims[a])
                return dcf_set_error(ctx, DCF_EINVAL, "grid %dx%dx%d does not fit the image %llux%llux%llu", grid[0], grid[1], grid[2],
                                     (unsigned long long)dims[0], (unsigned long long)dims[1], (unsigned long long)dims[2]);
        if ((int64_t)grid[0] * grid[1] * grid[2] > 4096) return dcf_set_error(ctx, DCF_EINVAL, "more than 4096 tiles");
    }
    const int grp = dcf_group_of(attr_idx);
    TreePtr t;
    int rc = tree_setup(ctx, img, dtype, dims, connectivity, tree_type, attr_idx, pixdim, flags, t);
    if (rc != DCF_OK) return rc;
    if ((rc = dcf_tiled_build(t.get(), grid, grp, 0)) != DCF_OK) return rc;
    ctx->trees.insert(t.get());
    *out = t.release();
    return DCF_OK;
} catch (const std::exception &e) {
    return dcf_set_error(ctx, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

/* ------------------------------------------------------------------------------------------------ */
/* batched form                                                                                     */
/* ------------------------------------------------------------------------------------------------ */
/* waits for the stream when it goes out of scope: host tables read by asynchronous copies are
 * declared before it, so they are freed only after those copies */
struct StreamWait {
    hipStream_t s;
    ~StreamWait() { (void)hipStreamSynchronize(s); }
};

/* one batch tree over images that are all u8 or all u16 (keys of that width: one digit pass for u8,
 * two for u16) */
static int batch_filter_same(dcf_ctx *ctx, const dcf_batch_image *images, int count, int connectivity, int tree_type,
                             int attr_idx, int decision, int extract, unsigned flags, bool small = false)
{
    if (!ctx) return DCF_EINVAL;
    if (!images || count <= 0) return dcf_set_error(ctx, DCF_EINVAL, "empty batch");
    if (tree_type != DCF_MAX_TREE && tree_type != DCF_MIN_TREE) return dcf_set_error(ctx, DCF_EINVAL, "bad tree type %d", tree_type);
    if (attr_idx < 0 || attr_idx >= DCF_NUM_ATTR) return dcf_set_error(ctx, DCF_EINVAL, "bad attribute index %d", attr_idx);
    if (decision < 0 || decision > 3) return dcf_set_error(ctx, DCF_EINVAL, "bad decision %d", decision);
    std::vector<BatchImg> tab((size_t)count);
    std::vector<double> lam((size_t)count);
    uint64_t N = 0, tiles = 0;
    int is3d = -1, all_u8 = 1;
    uint32_t any_min = 0;
    size_t max_u8 = 0;
    for (int i = 0; i < count; i++) {
        const dcf_batch_image &b = images[i];
        if (!b.img || !b.out) return dcf_set_error(ctx, DCF_EINVAL, "image %d: null pointer", i);
        if (b.dtype != DCF_U8 && b.dtype != DCF_U16) return dcf_set_error(ctx, DCF_EINVAL, "image %d: dtype must be u8 or u16", i);
        if (!b.dims[0] || !b.dims[1] || !b.dims[2]) return dcf_set_error(ctx, DCF_EINVAL, "image %d: empty", i);
        const uint64_t n = pixel_count(b.dims);
        if (n == UINT64_MAX) return dcf_set_error(ctx, DCF_EINVAL, "image %d: the pixel count exceeds 2^31-1", i);
        const int d3 = b.dims[2] > 1;
        if (is3d < 0) is3d = d3;
        else if (is3d != d3) return dcf_set_error(ctx, DCF_EINVAL, "image %d: a batch is all 2D or all 3D", i);
        tab[i].off = N;
        tab[i].W = (uint32_t)b.dims[0];
        tab[i].H = (uint32_t)b.dims[1];
        tab[i].D = (uint32_t)b.dims[2];
        tab[i].tile0 = (uint32_t)tiles;
        tab[i].inv = (b.flags & DCF_IMAGE_MIN_TREE) ? 1u : ((b.flags & DCF_IMAGE_MAX_TREE) ? 0u : (tree_type == DCF_MIN_TREE));
        tab[i].pad = 0;
        any_min |= tab[i].inv;
        const uint64_t t3 = (uint64_t)ctx->zt3, tw = (uint64_t)ctx->zt2[0], th = (uint64_t)ctx->zt2[1];
        tiles += d3 ? ((b.dims[0] + t3 - 1) / t3) * ((b.dims[1] + t3 - 1) / t3) * ((b.dims[2] + t3 - 1) / t3)
                    : ((b.dims[0] + tw - 1) / tw) * ((b.dims[1] + th - 1) / th);
        lam[i] = b.lambda;
        N += n;
        if (N >= (1ull << 30)) return dcf_set_error(ctx, DCF_EUNSUPPORTED, "a batch is limited to 2^30 pixels");
        if (b.dtype != DCF_U8) all_u8 = 0;
        else if (n > max_u8) max_u8 = (size_t)n;
    }
    /* check_connectivity (io.c:64-79) */
    int conn = connectivity;
    if (is3d && conn == 4) conn = 6;
    else if (conn != 4 && conn != 6 && conn != 8 && conn != 26)
        return dcf_set_error(ctx, DCF_EINVAL, "wrong connectivity %d (must be 4 or 8 for 2D, 6 or 26 for 3D)", connectivity);
    else if (conn == 8 && is3d) conn = 26;
    else if (conn == 26 && !is3d) conn = 8;
    if (!is3d && conn == 6) conn = 4;
    if (hipSetDevice(ctx->device) != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "hipSetDevice");

    const dcf_dtype cdt = all_u8 ? DCF_U8 : DCF_U16;
    const size_t csz = all_u8 ? 1 : 2;
    const bool dev = (flags & DCF_DEVICE_PTR) != 0;
    std::vector<BatchOut> outs((size_t)count);
    std::vector<size_t> boff((size_t)count);
    /* device blocks and the tree are released when the call returns, also when an exception unwinds;
     * wait_for_copies is destroyed first, so nothing is released (and no host table above freed) while
     * an asynchronous copy or kernel of this call may still read it */
    DevBlocks blocks{ ctx, {} };
    TreePtr t(new dcf_tree());
    StreamWait wait_for_copies{ ctx->stream };
    t->ctx = ctx;
    t->dims[0] = N;
    t->n = (int64_t)N;
    t->conn = conn;
    t->dtype = cdt;
    t->nbatch = count;
    t->batch3d = is3d;
    t->batch_tiles = (int64_t)tiles;
    t->small_batch = small ? 1 : 0;
    int rc = DCF_OK;
    t->img = dcf_alloc(ctx, (size_t)N * csz);
    t->own_img = 1;
    BatchImg *dtab = blocks.add((BatchImg *)dcf_alloc(ctx, (size_t)count * sizeof(BatchImg)));
    double *dlam = blocks.add((double *)dcf_alloc(ctx, (size_t)count * 8));
    BatchOut *douts = blocks.add((BatchOut *)dcf_alloc(ctx, (size_t)count * sizeof(BatchOut)));
    uint32_t *gm = nullptr;
    uint8_t *staging = nullptr, *obuf = nullptr;
    t->gmax = (float *)dcf_alloc(ctx, 16);
    t->counters = (uint32_t *)dcf_alloc(ctx, 64 * 4);
    t->parent = (int32_t *)dcf_alloc(ctx, (size_t)N * 4);
    t->p2n = (uint32_t *)dcf_alloc(ctx, (size_t)N * 4);
    if (!t->img || !dtab || !dlam || !douts || !t->gmax || !t->counters || !t->parent || !t->p2n)
        return dcf_set_error(ctx, DCF_ENOMEM, "batch arrays");
    t->bimg = dtab;
    if (hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(BatchImg), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dlam, lam.data(), lam.size() * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return dcf_set_error(ctx, DCF_EDEVICE, "batch table upload");
    /* the images, back to back in the batch's key type */
    if (!dev && !all_u8 && max_u8) {
        staging = blocks.add((uint8_t *)dcf_alloc(ctx, max_u8));
        if (!staging) return dcf_set_error(ctx, DCF_ENOMEM, "batch staging");
    }
    for (int i = 0; i < count && rc == DCF_OK; i++) {
        const dcf_batch_image &b = images[i];
        const int64_t n = (int64_t)(b.dims[0] * b.dims[1] * b.dims[2]);
        uint8_t *dst = (uint8_t *)t->img + tab[i].off * csz;
        if (dev) {
            rc = dcf_batch_copy_in(ctx, b.img, b.dtype, dst, cdt, n);
        } else if (b.dtype == cdt) {
            if (hipMemcpyAsync(dst, b.img, (size_t)n * csz, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
                rc = dcf_set_error(ctx, DCF_EDEVICE, "image %d upload", i);
        } else {
            if (hipMemcpyAsync(staging, b.img, (size_t)n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
                rc = dcf_set_error(ctx, DCF_EDEVICE, "image %d upload", i);
            else
                rc = dcf_batch_copy_in(ctx, staging, DCF_U8, dst, cdt, n);
        }
    }
    if (rc != DCF_OK) return rc;
    /* min-trees: every min-tree image inverted against its own grey maximum (preprocessing.c:7-30,84-91) */
    if (any_min) {
        gm = blocks.add((uint32_t *)dcf_alloc(ctx, (size_t)count * 4));
        if (!gm) return dcf_set_error(ctx, DCF_ENOMEM, "batch grey maxima");
        if ((rc = dcf_batch_invert(ctx, t->img, cdt, dtab, count, gm, (int64_t)N)) != DCF_OK) return rc;
    }
    if ((rc = dcf_build_tree(t.get(), dcf_group_of(attr_idx))) != DCF_OK) return rc;
    /* outputs: straight into device buffers, or into one staging buffer copied back per image */
    size_t ob = 0;
    for (int i = 0; i < count; i++) {
        boff[i] = ob;
        ob += (size_t)(images[i].dims[0] * images[i].dims[1] * images[i].dims[2]) * (images[i].dtype == DCF_U8 ? 1 : 2);
    }
    if (!dev) {
        obuf = blocks.add((uint8_t *)dcf_alloc(ctx, ob));
        if (!obuf) return dcf_set_error(ctx, DCF_ENOMEM, "batch output staging");
    }
    for (int i = 0; i < count; i++) {
        outs[i].ptr = dev ? images[i].out : (void *)(obuf + boff[i]);
        outs[i].dtype = images[i].dtype;
    }
    if (hipMemcpyAsync(douts, outs.data(), outs.size() * sizeof(BatchOut), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return dcf_set_error(ctx, DCF_EDEVICE, "batch output table upload");
    if ((rc = dcf_ops_batch_filter(t.get(), attr_idx, decision, dlam, extract, douts, gm, flags)) != DCF_OK) return rc;
    if (!dev)
        for (int i = 0; i < count && rc == DCF_OK; i++) {
            const size_t bytes = (size_t)(images[i].dims[0] * images[i].dims[1] * images[i].dims[2]) * (images[i].dtype == DCF_U8 ? 1 : 2);
            if (hipMemcpyAsync(images[i].out, obuf + boff[i], bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
                rc = dcf_set_error(ctx, DCF_EDEVICE, "image %d download", i);
        }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && rc == DCF_OK) rc = dcf_set_error(ctx, DCF_EDEVICE, "batch synchronize");
    return rc;
}

/* joins every started thread when it goes out of scope, also when an exception unwinds */
struct Joiner {
    std::vector<std::thread> t;
    ~Joiner()
    {
        for (auto &x : t)
            if (x.joinable()) x.join();
    }
};

/* Independent parts of one batched call, run concurrently on worker contexts: job(w, j) runs part j on worker w.
 * A part that fills the GPU level by level (a large image, or a batch tree of many images) still leaves it
 * idle between the dependent accesses of its latency-bound level kernels, and each level loop waits once
 * per digit pass on the host for its level histogram, so two loops overlap only from two host threads.
 * Each worker has its own stream and host thread and takes the next part not yet taken (parts come largest
 * first); the workers first wait for the work already queued on ctx's stream, and the call returns once
 * every worker stream is drained. */
int dcf_run_on_workers(dcf_ctx *ctx, int njobs, const std::function<int(dcf_ctx *, int)> &job)
{
    const int K = std::min(njobs, ctx->max_workers);
    /* the level loops of each of k concurrent workers: up to two for an image of at least 2^25 pixels (the
     * loops of smaller ones are launch-bound), as many as the hardware queues left over by the null stream
     * and the caller's own loops give every worker stream a queue of its own (two busy streams on one queue
     * run one after the other) */
    const int spare = ctx->hw_queues - 1 - ctx->fine_loops;
    auto loops_for = [&](int k) { return ctx->worker_loops ? ctx->worker_loops : std::max(1, std::min(2, spare / k)); };
    auto add_worker = [&]() -> int {
        dcf_ctx *w = nullptr;
        const int rc = ctx_create(ctx->device, true, &w);
        if (rc != DCF_OK) return rc;
        ctx->workers.push_back(w);
        return DCF_OK;
    };
    if (ctx->workers.empty()) {
        /* every worker's stream at once, then their further loop streams in a rotated order: queues created
         * in turn take the pipes in turn, so W workers' streams get pipes of their own, and so do two
         * workers' two loops each (worker 0 and 1, then the loop streams of workers W/2 and W/2 + 1) */
        const int W = std::max(K, ctx->max_workers);
        for (int k = 0; k < W; k++)
            if (const int rc = add_worker()) return dcf_set_error(ctx, rc, "worker context");
        const int L = loops_for(2);
        for (int j = 1; j < L; j++)
            for (int i = 0; i < W; i++)
                if (dcf_add_aux_streams(ctx->workers[(size_t)((i + j * (W / 2)) % W)], (size_t)j) != DCF_OK)
                    return dcf_set_error(ctx, DCF_EDEVICE, "loop stream");
    }
    while ((int)ctx->workers.size() < K)
        if (const int rc = add_worker()) return dcf_set_error(ctx, rc, "worker context");
    for (int k = 0; k < K; k++) {
        dcf_ctx *w = ctx->workers[(size_t)k];
        /* the caller's tunables and profiling */
        w->zones = ctx->zones; w->targets = ctx->targets; w->records = ctx->records; w->fix_blocks = ctx->fix_blocks;
        w->graph = ctx->graph; w->zip_segments = ctx->zip_segments; w->large_pixels = ctx->large_pixels;
        w->final_blocks = ctx->final_blocks; w->small_pixels = ctx->small_pixels;
        w->fine_loops = loops_for(K);
        w->loops_min_pixels = (int64_t)1 << 25;
        memcpy(w->zt2, ctx->zt2, sizeof w->zt2); w->zt2_set = ctx->zt2_set; w->zt3 = ctx->zt3;
        memcpy(w->lt2, ctx->lt2, sizeof w->lt2); memcpy(w->lg2, ctx->lg2, sizeof w->lg2);
        memcpy(w->lt3, ctx->lt3, sizeof w->lt3); memcpy(w->lg3, ctx->lg3, sizeof w->lg3);
        w->profiling = ctx->profiling;
        w->prof_only = ctx->prof_only;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "hipSetDevice");
    hipEvent_t ready = nullptr;
    if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "event");
    struct EventGuard {
        hipEvent_t e;
        ~EventGuard() { (void)hipEventDestroy(e); }
    } eg{ ready };
    if (hipEventRecord(ready, ctx->stream) != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "event record");
    for (int k = 0; k < K; k++)
        if (hipStreamWaitEvent(ctx->workers[(size_t)k]->stream, ready, 0) != hipSuccess)
            return dcf_set_error(ctx, DCF_EDEVICE, "stream wait");
    std::vector<int> rcs((size_t)K, DCF_OK);
    std::atomic<int> next{ 0 };
    {
        Joiner j;
        for (int k = 0; k < K; k++)
            j.t.emplace_back([&, k]() {
                dcf_ctx *w = ctx->workers[(size_t)k];
                (void)hipSetDevice(w->device);
                for (int i = next++; i < njobs && rcs[(size_t)k] == DCF_OK; i = next++) {
                    /* no exception may leave a thread (it would end the process) */
                    try {
                        rcs[(size_t)k] = job(w, i);
                    } catch (const std::exception &e) {
                        rcs[(size_t)k] = dcf_set_error(w, DCF_ENOMEM, "out of host memory (%s)", e.what());
                    }
                }
                if (hipStreamSynchronize(w->stream) != hipSuccess && rcs[(size_t)k] == DCF_OK) rcs[(size_t)k] = DCF_EDEVICE;
            });
    }
    int rc = DCF_OK;
    for (int k = 0; k < K; k++) {
        dcf_ctx *w = ctx->workers[(size_t)k];
        if (rcs[(size_t)k] != DCF_OK && rc == DCF_OK) {
            rc = rcs[(size_t)k];
            dcf_set_error(ctx, rc, "%s", w->err.c_str());
        }
        /* the workers' kernel timings join the caller's */
        if (w->profiling) {
            prof_collect(w);
            for (auto &kv : w->stats) {
                KernelStat &st = ctx->stats[kv.first];
                if (st.launches == 0 && kv.second.launches) ctx->names.push_back(kv.first);
                st.launches += kv.second.launches;
                st.ms += kv.second.ms;
            }
            w->stats.clear();
            w->names.clear();
        }
    }
    return rc;
}

extern "C" int dcf_batch_filter(dcf_ctx *ctx, const dcf_batch_image *images, int count, int connectivity, int tree_type,
                                int attr_idx, int decision, int extract, unsigned flags)
try {
    if (!ctx) return DCF_EINVAL;
    if (!images || count <= 0) return dcf_set_error(ctx, DCF_EINVAL, "empty batch");
    /* every image is checked here, so that errors name the caller's index */
    int is3d = -1;
    uint64_t N = 0;
    for (int i = 0; i < count; i++) {
        const dcf_batch_image &b = images[i];
        if (!b.img || !b.out) return dcf_set_error(ctx, DCF_EINVAL, "image %d: null pointer", i);
        if (b.dtype != DCF_U8 && b.dtype != DCF_U16) return dcf_set_error(ctx, DCF_EINVAL, "image %d: dtype must be u8 or u16", i);
        if (!b.dims[0] || !b.dims[1] || !b.dims[2]) return dcf_set_error(ctx, DCF_EINVAL, "image %d: empty", i);
        const uint64_t n = pixel_count(b.dims);
        if (n == UINT64_MAX) return dcf_set_error(ctx, DCF_EINVAL, "image %d: the pixel count exceeds 2^31-1", i);
        const int d3 = b.dims[2] > 1;
        if (is3d < 0) is3d = d3;
        else if (is3d != d3) return dcf_set_error(ctx, DCF_EINVAL, "image %d: a batch is all 2D or all 3D", i);
        N += n;
        if (N >= (1ull << 30)) return dcf_set_error(ctx, DCF_EUNSUPPORTED, "a batch is limited to 2^30 pixels");
        if ((b.flags & ~(DCF_IMAGE_MAX_TREE | DCF_IMAGE_MIN_TREE)) || b.flags == (DCF_IMAGE_MAX_TREE | DCF_IMAGE_MIN_TREE))
            return dcf_set_error(ctx, DCF_EINVAL, "image %d: bad flags 0x%x", i, b.flags);
    }
    if (connectivity != 4 && connectivity != 6 && connectivity != 8 && connectivity != 26)
        return dcf_set_error(ctx, DCF_EINVAL, "wrong connectivity %d (must be 4 or 8 for 2D, 6 or 26 for 3D)", connectivity);
    if (attr_idx < 0 || attr_idx >= DCF_NUM_ATTR) return dcf_set_error(ctx, DCF_EINVAL, "bad attribute index %d", attr_idx);
    if (decision < 0 || decision > 3) return dcf_set_error(ctx, DCF_EINVAL, "bad decision %d", decision);
    if (tree_type != DCF_MAX_TREE && tree_type != DCF_MIN_TREE) return dcf_set_error(ctx, DCF_EINVAL, "bad tree type %d", tree_type);
    /* the parts: each image of at least large_pixels pixels on its own (when there are two or more of them:
     * a single one gains nothing), then the u8 images and the u16 images as one batch tree each (a u8 image
     * on 16-bit keys would spend a whole first pass on one coarse level) */
    std::vector<int> large;
    for (int i = 0; i < count; i++)
        if ((int64_t)pixel_count(images[i].dims) >= ctx->large_pixels) large.push_back(i);
    if (large.size() < 2) large.clear();
    /* small images (the area group only): one part built in LDS, one workgroup per image */
    const int64_t small_max = std::min<int64_t>(ctx->small_pixels, DCF_SMALL_MAX);
    std::vector<char> is_small((size_t)count, 0);
    std::vector<dcf_batch_image> gsmall;
    uint64_t small_pix = 0;
    if (dcf_group_of(attr_idx) == 0)
        for (int i = 0; i < count; i++)
            if ((int64_t)pixel_count(images[i].dims) <= small_max) {
                is_small[(size_t)i] = 1;
                gsmall.push_back(images[i]);
                small_pix += pixel_count(images[i].dims);
            }
    /* each dtype's images are dealt into batch trees, largest image first onto the tree with the fewest
     * pixels: batch_parts trees per dtype, or by default (0), when the call has no large image, workers / dtypes
     * trees per dtype (u8 and u16 with two or three workers: one tree each; one dtype: one per worker) */
    int ndt = 0;
    for (int dt : { (int)DCF_U8, (int)DCF_U16 })
        for (int i = 0; i < count; i++)
            if (images[i].dtype == dt) {
                ndt++;
                break;
            }
    const int P = ctx->batch_parts > 0 ? ctx->batch_parts
                                       : (large.empty() ? std::max(1, ctx->max_workers / std::max(1, ndt)) : 1);
    std::vector<std::vector<dcf_batch_image>> groups;
    std::vector<uint64_t> gpix;
    for (int dt : { (int)DCF_U8, (int)DCF_U16 }) {
        std::vector<int> ids;
        size_t li = 0;
        for (int i = 0; i < count; i++) {
            if (li < large.size() && large[li] == i) {
                li++;
                continue;
            }
            if (images[i].dtype == dt && !is_small[(size_t)i]) ids.push_back(i);
        }
        if (ids.empty()) continue;
        std::stable_sort(ids.begin(), ids.end(),
                         [&](int x, int y) { return pixel_count(images[x].dims) > pixel_count(images[y].dims); });
        const size_t g0 = groups.size();
        const int np = std::min(P, (int)ids.size());
        groups.resize(g0 + (size_t)np);
        gpix.resize(g0 + (size_t)np, 0);
        for (int i : ids) {
            size_t best = g0;
            for (size_t g = g0; g < groups.size(); g++)
                if (gpix[g] < gpix[best]) best = g;
            groups[best].push_back(images[i]);
            gpix[best] += pixel_count(images[i].dims);
        }
    }
    if (!gsmall.empty()) {
        groups.push_back(gsmall);
        gpix.push_back(small_pix);
    }
    const size_t small_group = gsmall.empty() ? SIZE_MAX : groups.size() - 1;
    /* the parts, largest first: batch trees by pixel count, then the large images (also largest first) */
    std::vector<int> gorder(groups.size());
    for (size_t g = 0; g < groups.size(); g++) gorder[g] = (int)g;
    std::stable_sort(gorder.begin(), gorder.end(), [&](int x, int y) { return gpix[(size_t)x] > gpix[(size_t)y]; });
    std::stable_sort(large.begin(), large.end(),
                     [&](int x, int y) { return pixel_count(images[x].dims) > pixel_count(images[y].dims); });
    auto part = [&](dcf_ctx *c, int j) -> int {
        if (j < (int)groups.size()) {
            const size_t gi = (size_t)gorder[(size_t)j];
            const std::vector<dcf_batch_image> &g = groups[gi];
            return batch_filter_same(c, g.data(), (int)g.size(), connectivity, tree_type, attr_idx, decision, extract, flags,
                                     gi == small_group);
        }
        const dcf_batch_image &b = images[large[(size_t)j - groups.size()]];
        const int tt = (b.flags & DCF_IMAGE_MIN_TREE) ? DCF_MIN_TREE : ((b.flags & DCF_IMAGE_MAX_TREE) ? DCF_MAX_TREE : tree_type);
        dcf_tree *t = nullptr;
        int rc = dcf_tree_build(c, b.img, b.dtype, b.dims, connectivity, tt, attr_idx, nullptr, flags & DCF_DEVICE_PTR, &t);
        if (rc == DCF_OK) rc = dcf_filter(t, attr_idx, decision, b.lambda, extract, b.out, b.dtype, flags);
        if (t) dcf_tree_free(t);
        return rc;
    };
    const int njobs = (int)(groups.size() + large.size());
    if (njobs >= 2 && ctx->max_workers >= 2) return dcf_run_on_workers(ctx, njobs, part);
    for (int j = 0; j < njobs; j++) {
        const int rc = part(ctx, j);
        if (rc != DCF_OK) return rc;
    }
    return DCF_OK;
} catch (const std::exception &e) {
    return dcf_set_error(ctx, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

extern "C" int dcf_tree_accumulate(dcf_tree *t, int attr_idx)
try {
    if (!t) return DCF_EINVAL;
    int grp = dcf_group_of(attr_idx);
    if (grp < 0) return dcf_set_error(t->ctx, DCF_EINVAL, "bad attribute index %d", attr_idx);
    if (grp == t->grp && t->attr.area) return DCF_OK;
    if (t->dtype == DCF_F32 && grp == 3)
        return dcf_set_error(t->ctx, DCF_EUNSUPPORTED, "grey-weighted attributes on float images are not implemented");
    (void)hipSetDevice(t->ctx->device);
    return dcf_accumulate_group(t, grp);
} catch (const std::exception &e) {
    return dcf_set_error(t ? t->ctx : nullptr, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

extern "C" int dcf_tree_group(const dcf_tree *t) { return t ? t->grp : -1; }

extern "C" int dcf_tree_num_nodes(dcf_tree *t, int64_t *k)
{
    if (!t || !k) return DCF_EINVAL;
    *k = t->K;
    return DCF_OK;
}

extern "C" int dcf_tree_parent(dcf_tree *t, int32_t *parent, unsigned flags)
{
    if (!t || !parent) return DCF_EINVAL;
    dcf_ctx *ctx = t->ctx;
    (void)hipSetDevice(ctx->device);
    DCF_HIP(ctx, hipMemcpyAsync(parent, t->parent, (size_t)t->n * 4,
                                (flags & DCF_DEVICE_PTR) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
    if (!(flags & DCF_DEVICE_PTR)) DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DCF_OK;
}

static int need_group(dcf_tree *t, int attr_idx)
{
    if (attr_idx < 0 || attr_idx >= DCF_NUM_ATTR) return dcf_set_error(t->ctx, DCF_EINVAL, "bad attribute index %d", attr_idx);
    if (dcf_group_of(attr_idx) != t->grp)
        return dcf_set_error(t->ctx, DCF_EGROUP, "attribute %d (group %d) not accumulated (tree has group %d)", attr_idx,
                             dcf_group_of(attr_idx), t->grp);
    return DCF_OK;
}

extern "C" int dcf_filter(dcf_tree *t, int attr_idx, int decision, double lambda, int extract, void *out, dcf_dtype out_dtype,
                          unsigned flags)
try {
    if (!t || !out) return DCF_EINVAL;
    dcf_ctx *ctx = t->ctx;
    int rc = need_group(t, attr_idx);
    if (rc) return rc;
    if (decision < 0 || decision > 3) return dcf_set_error(ctx, DCF_EINVAL, "bad decision index %d", decision);
    if (out_dtype != DCF_U8 && out_dtype != DCF_U16 && out_dtype != DCF_F32) return dcf_set_error(ctx, DCF_EINVAL, "bad output dtype");
    (void)hipSetDevice(ctx->device);
    size_t bytes = (size_t)t->n * dtype_size(out_dtype);
    void *dst = out;
    if (!(flags & DCF_DEVICE_PTR)) {
        dst = dcf_alloc(ctx, bytes);
        if (!dst) return dcf_set_error(ctx, DCF_ENOMEM, "output staging");
    }
    rc = dcf_ops_filter(t, attr_idx, decision, lambda, extract, dst, out_dtype, flags);
    if (rc == DCF_OK && !(flags & DCF_DEVICE_PTR)) {
        if (hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = dcf_set_error(ctx, DCF_EDEVICE, "output download");
    }
    if (!(flags & DCF_DEVICE_PTR)) dcf_release(ctx, dst);
    return rc;
} catch (const std::exception &e) {
    return dcf_set_error(t ? t->ctx : nullptr, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

extern "C" int dcf_dmp(dcf_tree *t, int attr_idx, int decision, const float *lambdas, int n, void *out, dcf_dtype out_dtype,
                       unsigned flags)
try {
    if (!t || !out || !lambdas) return DCF_EINVAL;
    dcf_ctx *ctx = t->ctx;
    if (n < 2) return dcf_set_error(ctx, DCF_EINVAL, "a differential profile needs at least 2 lambdas (lambda_0 = 0 first)");
    int rc = need_group(t, attr_idx);
    if (rc) return rc;
    if (decision < 0 || decision > 3) return dcf_set_error(ctx, DCF_EINVAL, "bad decision index %d", decision);
    if (out_dtype != DCF_U8 && out_dtype != DCF_U16 && out_dtype != DCF_F32) return dcf_set_error(ctx, DCF_EINVAL, "bad output dtype");
    (void)hipSetDevice(ctx->device);
    const size_t one = (size_t)t->n * dtype_size(out_dtype);
    const bool dev = (flags & DCF_DEVICE_PTR) != 0;
    void *dst = dev ? nullptr : dcf_alloc(ctx, one);
    if (!dev && !dst) return dcf_set_error(ctx, DCF_ENOMEM, "output staging");
    for (int i = 1; i < n && rc == DCF_OK; i++) {
        void *o = dev ? (void *)((uint8_t *)out + (size_t)(i - 1) * one) : dst;
        /* lambda_cur is the float threshold passed on as a double (operation.c:67-68) */
        rc = dcf_ops_filter(t, attr_idx, decision, (double)lambdas[i], 1, o, out_dtype, (flags & ~DCF_DEVICE_PTR) | DCF_KEEP_INVERTED);
        if (rc == DCF_OK && !dev) {
            if (hipMemcpyAsync((uint8_t *)out + (size_t)(i - 1) * one, dst, one, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                rc = dcf_set_error(ctx, DCF_EDEVICE, "output download");
        }
    }
    if (!dev) dcf_release(ctx, dst);
    return rc;
} catch (const std::exception &e) {
    return dcf_set_error(t ? t->ctx : nullptr, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

/* write_tree (io.c:1431-1573): one row per levelroot in raster order, the reference's printf formats */
static int tree_csv_text(dcf_tree *t, int attr_idx, std::string &text)
{
    dcf_ctx *ctx = t->ctx;
    int rc = need_group(t, attr_idx);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const int grp = dcf_group_of(attr_idx);
    std::vector<uint32_t> n2p;
    std::vector<int32_t> npar;
    std::vector<float> grey;
    std::vector<double> size, cols;
    int ncol = 0;
    if ((rc = dcf_ops_tree_table(t, grp, n2p, npar, grey, size, cols, ncol)) != DCF_OK) return rc;
    text.clear();
    text += "index,x,y,z,gval,parent,flux,area,";
    if (grp == 1) text += "area_enc_rect,diag_enc_rect,x_extent,y_extent,z_extent";
    else if (grp == 2) text += "x_mean,y_mean,z_mean";
    else if (grp == 3) text += "x_wmean,y_wmean,z_wmean";
    else if (grp == 4) text += "x_mean,y_mean,z_mean,x_wmean,y_wmean,z_wmean,elong,flat,spars,ncomp";
    text += "\n";
    const uint64_t W = t->dims[0], WH = t->dims[0] * t->dims[1];
    /* the longest row: 6 longs of 20 characters, 2 "%f" and 10 "%.2f" of floats of at most 39 integer
     * digits (48 characters each with sign and fraction), commas: under 800 */
    char buf[1024];
    for (int64_t k = 0; k < t->K; k++) {
        const uint64_t i = n2p[(size_t)k];
        const long x = (long)((i % WH) % W), y = (long)((i % WH) / W), z = (long)(i / WH);
        const int32_t pk = npar[(size_t)k];
        const long par = pk < 0 ? -1L : (long)n2p[(size_t)pk];
        const float g = grey[(size_t)k];
        const float dh = pk < 0 ? g : g - grey[(size_t)pk];
        const long sz = (long)size[(size_t)k];
        int len = snprintf(buf, sizeof buf, "%ld,%ld,%ld,%ld,%f,%ld,%f,%ld,", (long)i, x, y, z, (double)g, par,
                           (double)(dh * (float)sz), sz);
        const double *c = cols.data() + (size_t)k * ncol;
        auto f = [&](int j) { return (double)(float)c[j]; };
        int more = 0;
        if (len >= 0 && len < (int)sizeof buf) {
            if (grp == 1)
                more = snprintf(buf + len, sizeof buf - len, "%.2f,%.2f,%.2f,%.2f,%.2f", f(0), f(1), f(2), f(3), f(4));
            else if (grp == 2 || grp == 3)
                more = snprintf(buf + len, sizeof buf - len, "%.2f,%.2f,%.2f", f(0), f(1), f(2));
            else if (grp == 4)
                more = snprintf(buf + len, sizeof buf - len, "%.2f,%.2f,%.2f,%.2f,%.2f,%.2f,%.3f,%.3f,%.3f,%.4f ", f(0), f(1), f(2),
                                f(3), f(4), f(5), f(6), f(7), f(8), f(9));
        }
        if (len < 0 || more < 0 || len + more >= (int)sizeof buf)
            return dcf_set_error(ctx, DCF_EINVAL, "tree row %lld does not fit its %d-byte buffer", (long long)k, (int)sizeof buf);
        len += more;
        text.append(buf, (size_t)len);
        text += "\n";
    }
    return DCF_OK;
}

extern "C" int dcf_tree_write_csv(dcf_tree *t, int attr_idx, const char *path)
try {
    if (!t || !path) return DCF_EINVAL;
    std::string text;
    int rc = tree_csv_text(t, attr_idx, text);
    if (rc) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return dcf_set_error(t->ctx, DCF_EIO, "cannot open %s", path);
    const size_t w = fwrite(text.data(), 1, text.size(), f);
    const int cl = fclose(f);
    if (w != text.size() || cl != 0) return dcf_set_error(t->ctx, DCF_EIO, "cannot write %s", path);
    return DCF_OK;
} catch (const std::exception &e) {
    return dcf_set_error(t ? t->ctx : nullptr, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

extern "C" int dcf_check(dcf_tree *t, int attr_idx, float *out, unsigned flags)
try {
    if (!t || !out) return DCF_EINVAL;
    dcf_ctx *ctx = t->ctx;
    int rc = need_group(t, attr_idx);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    size_t bytes = (size_t)t->n * 4;
    float *dst = out;
    if (!(flags & DCF_DEVICE_PTR)) {
        dst = (float *)dcf_alloc(ctx, bytes);
        if (!dst) return dcf_set_error(ctx, DCF_ENOMEM, "output staging");
    }
    rc = dcf_ops_check(t, attr_idx, dst);
    if (rc == DCF_OK && !(flags & DCF_DEVICE_PTR)) {
        if (hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = dcf_set_error(ctx, DCF_EDEVICE, "output download");
    }
    if (!(flags & DCF_DEVICE_PTR)) dcf_release(ctx, dst);
    return rc;
} catch (const std::exception &e) {
    return dcf_set_error(t ? t->ctx : nullptr, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

extern "C" int dcf_spectrum1d(dcf_tree *t, int attr_idx, const float *lambdas, int n, double *out)
try {
    if (!t || !lambdas || !out || n <= 0) return t ? dcf_set_error(t->ctx, DCF_EINVAL, "bad spectrum arguments") : DCF_EINVAL;
    if (n > DCF_MAX_SPECTRUM_BINS) return dcf_set_error(t->ctx, DCF_EINVAL, "%d bins exceed %d", n, DCF_MAX_SPECTRUM_BINS);
    int rc = need_group(t, attr_idx);
    if (rc) return rc;
    (void)hipSetDevice(t->ctx->device);
    return dcf_ops_spectrum(t, attr_idx, lambdas, n, 0, nullptr, 1, 0, out);
} catch (const std::exception &e) {
    return dcf_set_error(t ? t->ctx : nullptr, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

extern "C" int dcf_spectrum1d_noalloc(dcf_tree *t, int attr_idx, const float *lambdas, int n, double *out)
{
    if (!t) return DCF_EINVAL;
    if (!lambdas || !out || n <= 0) return dcf_set_error(t->ctx, DCF_EINVAL, "bad spectrum arguments");
    if (n > DCF_NOALLOC_MAX_BINS)
        return dcf_set_error(t->ctx, DCF_EINVAL, "%d bins exceed the allocation-free maximum of %d", n, DCF_NOALLOC_MAX_BINS);
    const int rc = need_group(t, attr_idx);
    if (rc) return rc;
    if (hipSetDevice(t->ctx->device) != hipSuccess) return dcf_set_error(t->ctx, DCF_EDEVICE, "hipSetDevice");
    return dcf_ops_spectrum_noalloc(t, attr_idx, lambdas, n, out);
}

extern "C" int dcf_spectrum2d(dcf_tree *t, int attr1, const float *l1, int n1, int attr2, const float *l2, int n2, double *out)
try {
    if (!t || !l1 || !l2 || !out || n1 <= 0 || n2 <= 0) return t ? dcf_set_error(t->ctx, DCF_EINVAL, "bad spectrum arguments") : DCF_EINVAL;
    dcf_ctx *ctx = t->ctx;
    /* the bins are indexed s1 + s2*n1 in 32-bit arithmetic on both sides: bound the product first */
    if ((int64_t)n1 * n2 > DCF_MAX_SPECTRUM_BINS)
        return dcf_set_error(ctx, DCF_EINVAL, "%d x %d bins exceed %d", n1, n2, DCF_MAX_SPECTRUM_BINS);
    if (attr1 < 0 || attr1 >= DCF_NUM_ATTR || attr2 < 0 || attr2 >= DCF_NUM_ATTR)
        return dcf_set_error(ctx, DCF_EINVAL, "bad attribute index");
    /* operation.c:93-95: the area attribute must be the second axis */
    if (attr1 == 0 && attr2 != 0)
        return dcf_set_error(ctx, DCF_EINVAL, "invert the two attributes: the one with the highest index must be first");
    int second_is_area = attr2 == 0 && attr1 != 0;
    int rc = need_group(t, attr1);
    if (rc) return rc;
    if (!second_is_area && (rc = need_group(t, attr2)) != 0) return rc;
    (void)hipSetDevice(ctx->device);
    return dcf_ops_spectrum(t, attr1, l1, n1, attr2, l2, n2, second_is_area, out);
} catch (const std::exception &e) {
    return dcf_set_error(t ? t->ctx : nullptr, DCF_ENOMEM, "out of host memory (%s)", e.what());
}

/* ------------------------------------------------------------------------------------------------ */
/* lambda files (lambdavec.c:35-84)                                                                 */
/* ------------------------------------------------------------------------------------------------ */
extern "C" int dcf_lvec_read(const char *path, double imscale, float **lambdas, int *n)
{
    static const char *hdr = "#LambdaVector file. Keep this exact line in all similar lambda vector files\n";
    if (!path || !lambdas || !n) return DCF_EINVAL;
    *lambdas = nullptr;
    *n = 0;
    FILE *f = fopen(path, "r");
    if (!f) return DCF_EIO;
    /* the header is consumed the way a scanf literal would consume it */
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
    std::vector<float> v;
    try {
        v.push_back(0.0f);
        float x;
        while (fscanf(f, "%f", &x) == 1) v.push_back(x);
    } catch (const std::exception &) {
        fclose(f);
        return DCF_ENOMEM;
    }
    fclose(f);
    if (v.size() > (size_t)DCF_MAX_SPECTRUM_BINS) return DCF_EINVAL; /* more thresholds than any spectrum takes */
    for (size_t i = 1; i < v.size(); i++) v[i] /= imscale;
    float *out = (float *)malloc(v.size() * sizeof(float));
    if (!out) return DCF_ENOMEM;
    memcpy(out, v.data(), v.size() * sizeof(float));
    *lambdas = out;
    *n = (int)v.size();
    return DCF_OK;
}

extern "C" void dcf_lvec_free(float *l) { free(l); }

/* ------------------------------------------------------------------------------------------------ */
/* operator driver: the reference's operation tuples (config.c:297-433), run like loop_operations  */
/* ------------------------------------------------------------------------------------------------ */
struct OpTuple {
    std::vector<std::string> items; /* strings unquoted */
    std::vector<int> is_str;
};

static int parse_tuples(const char *s, std::vector<OpTuple> &ops, std::string &err)
{
    size_t i = 0, L = strlen(s);
    auto skip = [&]() { while (i < L && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r' || s[i] == ',' || s[i] == ';')) i++; };
    skip();
    /* accept an optional outer pair of parentheses, as the CLI form ((tree,15)) (config.c:528) */
    while (i < L) {
        skip();
        if (i >= L) break;
        if (s[i] != '(') { err = "expected '('"; return -1; }
        size_t j = i + 1;
        while (j < L && (s[j] == ' ' || s[j] == '\t')) j++;
        if (j < L && s[j] == '(') { i = j; continue; } /* nested: strip the outer level */
        i++;
        OpTuple t;
        for (;;) {
            while (i < L && (s[i] == ' ' || s[i] == '\t')) i++;
            if (i >= L) { err = "unterminated tuple"; return -1; }
            if (s[i] == ')') { i++; break; }
            if (s[i] == ',') { i++; continue; }
            if (s[i] == '"') {
                size_t k = s[i + 1] ? i + 1 : i;
                size_t e = k;
                while (e < L && s[e] != '"') e++;
                if (e >= L) { err = "unterminated string"; return -1; }
                t.items.emplace_back(s + k, e - k);
                t.is_str.push_back(1);
                i = e + 1;
            } else {
                size_t e = i;
                while (e < L && s[e] != ',' && s[e] != ')') e++;
                std::string tok(s + i, e - i);
                while (!tok.empty() && (tok.back() == ' ' || tok.back() == '\t')) tok.pop_back();
                /* bare words (CLI form) are strings unless they parse as numbers */
                char *endp = nullptr;
                strtod(tok.c_str(), &endp);
                t.items.push_back(tok);
                t.is_str.push_back(!(endp && *endp == 0 && !tok.empty()));
                i = e;
            }
        }
        while (i < L && s[i] == ')') i++; /* close an outer level */
        ops.push_back(t);
    }
    return 0;
}

struct Op {
    std::string name;
    int attr = 0, attr2 = 0, decision = 0;
    double lambda = 0, imscale = 1, imscale2 = 1;
    std::string lvec, lvec2;
};

static int to_ops(dcf_ctx *ctx, const std::vector<OpTuple> &tp, std::vector<Op> &ops)
{
    for (auto &t : tp) {
        Op o;
        if (t.items.size() < 2) return dcf_set_error(ctx, DCF_EINVAL, "an operation needs at least 2 parameters");
        o.name = t.items[0];
        o.attr = atoi(t.items[1].c_str());
        if (o.attr < 0 || o.attr >= DCF_NUM_ATTR) return dcf_set_error(ctx, DCF_EINVAL, "bad attribute index %d", o.attr);
        if (o.name == "filter" || o.name == "extract") {
            if (t.items.size() < 3 || t.items.size() > 4) return dcf_set_error(ctx, DCF_EINVAL, "filter/extract take 3 or 4 parameters");
            o.lambda = atof(t.items[2].c_str());
            o.decision = t.items.size() > 3 ? atoi(t.items[3].c_str()) : 0;
            if (o.decision < 0 || o.decision > 3) return dcf_set_error(ctx, DCF_EINVAL, "bad decision index %d", o.decision);
        } else if (o.name == "pattern") {
            if (t.items.size() < 3 || t.items.size() > 4) return dcf_set_error(ctx, DCF_EINVAL, "pattern takes 3 or 4 parameters");
            o.lvec = t.items[2];
            o.imscale = t.items.size() > 3 ? atof(t.items[3].c_str()) : 1.0;
        } else if (o.name == "pattern2D") {
            if (t.items.size() != 7) return dcf_set_error(ctx, DCF_EINVAL, "pattern2D takes 7 parameters");
            o.lvec = t.items[2];
            o.imscale = atof(t.items[3].c_str());
            o.attr2 = atoi(t.items[4].c_str());
            o.lvec2 = t.items[5];
            o.imscale2 = atof(t.items[6].c_str());
            if (o.attr2 < 0 || o.attr2 >= DCF_NUM_ATTR) return dcf_set_error(ctx, DCF_EINVAL, "bad attribute index %d", o.attr2);
        } else if (o.name == "check") {
            if (t.items.size() > 2) return dcf_set_error(ctx, DCF_EINVAL, "check takes 2 parameters");
        } else if (o.name == "dmp") {
            /* ("dmp", attr, lvec[, imscale]); the decision stays 0 (config.c:355-379) */
            if (t.items.size() < 3 || t.items.size() > 4) return dcf_set_error(ctx, DCF_EINVAL, "dmp takes 3 or 4 parameters");
            o.lvec = t.items[2];
            o.imscale = t.items.size() > 3 ? atof(t.items[3].c_str()) : 1.0;
        } else if (o.name == "tree") {
            if (t.items.size() != 2) return dcf_set_error(ctx, DCF_EINVAL, "tree takes 2 parameters");
        } else if (o.name == "csl") {
            return dcf_set_error(ctx, DCF_EUNSUPPORTED, "operation \"%s\" is outside the GPU path", o.name.c_str());
        } else {
            return dcf_set_error(ctx, DCF_EINVAL, "operation \"%s\" not recognized", o.name.c_str());
        }
        ops.push_back(o);
    }
    return DCF_OK;
}

extern "C" int dcf_run_operations(dcf_ctx *ctx, const void *img, dcf_dtype dtype, const uint64_t dims[3], int connectivity,
                                  int tree_type, const float pixdim[3], const char *opstr, dcf_output_fn cb, void *user)
try {
    if (!ctx) return DCF_EINVAL;
    if (!opstr || !img || !dims) return dcf_set_error(ctx, DCF_EINVAL, "null argument");
    std::vector<OpTuple> tp;
    std::string perr;
    if (parse_tuples(opstr, tp, perr) != 0) return dcf_set_error(ctx, DCF_EINVAL, "operation list: %s", perr.c_str());
    std::vector<Op> ops;
    int rc = to_ops(ctx, tp, ops);
    if (rc) return rc;
    if (ops.empty()) return DCF_OK;
    dcf_tree *t = nullptr;
    int64_t n = 0; /* pixels, from the tree once the build has checked the dimensions */
    std::vector<float> fbuf;
    for (size_t i = 0; i < ops.size() && rc == DCF_OK; i++) {
        const Op &o = ops[i];
        if (i == 0) {
            rc = dcf_tree_build(ctx, img, dtype, dims, connectivity, tree_type, o.attr, pixdim, 0, &t);
            if (rc) break;
            n = t->n;
        } else if (dcf_group_of(o.attr) != dcf_group_of(ops[i - 1].attr)) {
            rc = dcf_tree_accumulate(t, o.attr); /* disccofan.c:179-195 */
            if (rc) break;
        }
        if (o.name == "filter" || o.name == "extract") {
            fbuf.resize((size_t)n);
            rc = dcf_filter(t, o.attr, o.decision, o.lambda, o.name == "extract", fbuf.data(), DCF_F32, 0);
            if (rc == DCF_OK && cb) cb(user, (int)i, 0, fbuf.data(), n);
        } else if (o.name == "check") {
            fbuf.resize((size_t)n);
            rc = dcf_check(t, o.attr, fbuf.data(), 0);
            if (rc == DCF_OK && cb) cb(user, (int)i, 1, fbuf.data(), n);
        } else if (o.name == "pattern") {
            float *lv = nullptr;
            int nl = 0;
            rc = dcf_lvec_read(o.lvec.c_str(), o.imscale, &lv, &nl);
            if (rc) { rc = dcf_set_error(ctx, rc, "couldn't read the lvec file: %s", o.lvec.c_str()); break; }
            std::vector<double> sp((size_t)nl);
            rc = dcf_spectrum1d(t, o.attr, lv, nl, sp.data());
            dcf_lvec_free(lv);
            if (rc == DCF_OK && cb) cb(user, (int)i, 2, sp.data(), nl);
        } else if (o.name == "tree") {
            std::string text;
            rc = tree_csv_text(t, o.attr, text);
            if (rc == DCF_OK && cb) cb(user, (int)i, 5, text.data(), (int64_t)text.size());
        } else if (o.name == "dmp") {
            float *lv = nullptr;
            int nl = 0;
            rc = dcf_lvec_read(o.lvec.c_str(), o.imscale, &lv, &nl);
            if (rc) { rc = dcf_set_error(ctx, rc, "couldn't read the lvec file: %s", o.lvec.c_str()); break; }
            std::vector<float> prof((size_t)n * (size_t)(nl - 1));
            rc = nl < 2 ? dcf_set_error(ctx, DCF_EINVAL, "empty lvec file: %s", o.lvec.c_str())
                        : dcf_dmp(t, o.attr, o.decision, lv, nl, prof.data(), DCF_F32, 0);
            dcf_lvec_free(lv);
            if (rc == DCF_OK && cb) cb(user, (int)i, 4, prof.data(), (int64_t)prof.size());
        } else if (o.name == "pattern2D") {
            float *l1 = nullptr, *l2 = nullptr;
            int n1 = 0, n2 = 0;
            rc = dcf_lvec_read(o.lvec.c_str(), o.imscale, &l1, &n1);
            if (rc) { rc = dcf_set_error(ctx, rc, "couldn't read the lvec file: %s", o.lvec.c_str()); break; }
            rc = dcf_lvec_read(o.lvec2.c_str(), o.imscale2, &l2, &n2);
            if (rc) { dcf_lvec_free(l1); rc = dcf_set_error(ctx, rc, "couldn't read the lvec file: %s", o.lvec2.c_str()); break; }
            if ((int64_t)n1 * n2 > DCF_MAX_SPECTRUM_BINS) {
                dcf_lvec_free(l1);
                dcf_lvec_free(l2);
                rc = dcf_set_error(ctx, DCF_EINVAL, "%d x %d bins exceed %d", n1, n2, DCF_MAX_SPECTRUM_BINS);
                break;
            }
            std::vector<double> sp((size_t)n1 * n2);
            rc = dcf_spectrum2d(t, o.attr, l1, n1, o.attr2, l2, n2, sp.data());
            dcf_lvec_free(l1);
            dcf_lvec_free(l2);
            if (rc == DCF_OK && cb) cb(user, (int)i, 3, sp.data(), (int64_t)n1 * n2);
        }
    }
    if (t) dcf_tree_free(t);
    return rc;
} catch (const std::exception &e) {
    return dcf_set_error(ctx, DCF_ENOMEM, "out of host memory (%s)", e.what());
}
