/*
 * dcf_internal.h -- device data layout and host-side launchers of libdcf (not installed).
 *
 * HBM layout of one tree over N pixels (all arrays indexed by raster index p = x + y*W + z*W*H):
 *   img        caller's image (u8/u16/f32), read in place; a min-tree reads gmax - v on the fly
 *   rank[N]    u32  (f32 images, during the build) dense rank of the pixel's float key
 *   parent[N]  i32  canonical parent (levelroot form, root = -1)
 *   p2n[N]     u32  node index of the pixel's levelroot
 * and of the K nodes (levelroots, numbered in raster order of their levelroot pixel):
 *   npar[K]    i32  parent node, -1 at the root
 *   ngrey[K]   f32  grey value of the node (exact for integer images)
 *   attribute fields (structure of arrays, exact integers): area u32, and per group
 *     group 1 extent : min x/y/z, max x/y/z (max is exclusive, as init_extent_data)
 *     group 2 mean   : sum x, y, z
 *     group 3 wmean  : sum g, g*x, g*y, g*z
 *     group 4 inertia: sum x, y, z, x^2, y^2, z^2, xy, yz, xz
 */
#ifndef DCF_INTERNAL_H
#define DCF_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <set>
#include <string>
#include <functional>
#include <vector>

#include "../../include/dcf.h"

#define DCF_MAX_SUMS 9

/* the largest image the LDS build takes (a power of two: its LDS arrays hold 13 bytes per pixel) */
#define DCF_SMALL_MAX 4096

struct DevAttr {
    uint32_t *area = nullptr;
    int64_t *s[DCF_MAX_SUMS] = {};
    uint32_t *mn[3] = {};
    uint32_t *mx[3] = {};
    int nsum = 0;
    int has_ext = 0;
};

/* what the kernels need to read a pixel's grey key and value */
/* internal ImgView type: 32-bit keys (dense grey ranks of a float image), never inverted */
#define DCF_KEY32 7

struct ImgView {
    const void *img;
    int dtype;        /* dcf_dtype, or DCF_KEY32 */
    int invert;       /* min-tree: value = gmax - v */
    const float *gmax; /* device scalar (only read when invert) */
};

struct KernelStat {
    int64_t launches = 0;
    double ms = 0;
};

/* zone tiles of the first pass by default: 64x64 in 2D, 16x16x16 in 3D (a context may set others, see
 * dcf_ctx::zt2 / zt3); a tile holds at most ZONE_TILE_MAX pixels */
#define ZONE_T2 64
#define ZONE_T3 16
#define ZONE_TILE_MAX 4096

struct dcf_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    /* a refining pass of an image of at least 2^22 pixels runs as fine_loops independent level loops
     * (DCF_FINE_LOOPS, 1..4); the streams of the further loops, created on first use */
    std::vector<hipStream_t> aux;
    int fine_loops = 3;
    /* smallest image whose refining pass is split into loops */
    int64_t loops_min_pixels = (int64_t)1 << 22;
    /* the loops of each worker context when several run at once (DCF_WORKER_LOOPS, 1..4; 0: up to two, as
     * many as give every worker stream a hardware queue of its own), for images of at least 2^25 pixels */
    int worker_loops = 0;
    /* the HIP runtime's hardware queues per process (GPU_MAX_HW_QUEUES, 4 when unset), read when the
     * context is created: streams beyond that number share queues and run one after the other */
    int hw_queues = 4;
    std::string err;
    int profiling = 0;
    std::string prof_only; /* empty: every kernel */
    struct Pending { std::string name; hipEvent_t a, b; };
    std::multimap<size_t, void *> free_blocks;
    std::map<void *, size_t> live_blocks;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, KernelStat> stats;
    std::vector<std::string> names;
    std::set<dcf_tree *> trees; /* live trees, freed by dcf_ctx_destroy */
    /* level-list tiles of the integer build (DCF_LEVEL_TILE2D / DCF_LEVEL_TILE3D, "WxHxD") and the
     * groups of tiles kept together in the list (DCF_LEVEL_GROUP2D / DCF_LEVEL_GROUP3D) */
    int lt2[3] = { 64, 64, 1 }, lg2[3] = { 1, 1, 1 };
    int lt3[3] = { 16, 16, 16 }, lg3[3] = { 1, 1, 1 };
    /* zone tiles of the first pass (DCF_ZONE_TILE2D "WxH", DCF_ZONE_TILE3D "N"; at most ZONE_TILE_MAX pixels, the
     * size of the labelling's LDS tables), which are also a batch's level-list tiles */
    int zt2[2] = { ZONE_T2, ZONE_T2 }, zt3 = ZONE_T3;
    /* 1 when the 2D zone tile was set (DCF_ZONE_TILE2D): a batch then keeps it instead of choosing its own */
    int zt2_set = 0;
    /* refining passes keep {union-find entry, key, coarse node} in one 16-byte record per element
     * (DCF_RECORDS, 0 off: the default, since the precomputed targets already carry what the records
     * would) */
    int records = 0;
    /* union targets of pixel passes precomputed by the level scatter (DCF_TARGETS): 0 off, 1 on,
     * 2 batches only */
    int targets = 1;
    /* zone graph of the first pass (DCF_ZONES): 0 off, 1 every single-image first pass,
     * 2 the coarse pass of u16 images and 2D u8 images */
    int zones = 2;
    /* grid cap of the per-level k_fix launch (DCF_FIX_BLOCKS) */
    int fix_blocks = 1024;
    /* capture each pass's level loop into a HIP graph (DCF_GRAPH, dcf_ctx_set_graph) */
    int graph = 0;
    /* tiled build: zips per identification pair, spread down the key range (DCF_ZIP_SEGMENTS; 1 = one) */
    int zip_segments = 64;
    /* device address of the allocation-free entry points' scratch array on this context's device */
    void *noalloc_scratch = nullptr;
    /* batched form: the parts of a call (each image of at least large_pixels pixels on its own, and the
     * u8 and u16 images in batch_parts batch trees each, 0: enough trees for every worker) run concurrently
     * on at most max_workers worker contexts, each with its own stream and host thread (DCF_LARGE_PIXELS,
     * DCF_BATCH_PARTS, DCF_WORKERS) */
    int64_t large_pixels = (int64_t)1 << 24;
    int max_workers = 3; /* 4 when the process has at least 8 hardware queues */
    int batch_parts = 0;
    /* first-pass pixel levels: grid cap of k_final_pixels, whose blocks then carry their totals over
     * several chunks of the level (0: one chunk per block) (DCF_FINAL_BLOCKS) */
    int final_blocks = 1024;
    /* batched form: images of at most small_pixels pixels (and at most DCF_SMALL_MAX) are built in LDS,
     * one workgroup per image, with the area attribute group (DCF_SMALL_PIXELS; 0, the default: never) */
    int64_t small_pixels = 0;
    std::vector<dcf_ctx *> workers;
};

/* batched form: one entry per image of a batch laid out back to back in one pixel space */
struct BatchImg {
    uint64_t off;     /* first pixel of the image */
    uint32_t W, H, D; /* its dimensions */
    uint32_t tile0;   /* its first tile in the batch's zone tiles (the context's zone tile shape) */
    uint32_t inv;     /* 1: a min-tree image, stored inverted against its own grey maximum */
    uint32_t pad;
};



/* output of one batch image: device pointer and element type */
struct BatchOut {
    void *ptr;
    int dtype;
};

/* the image of batch pixel p (binary search over the image offsets) */
__device__ __forceinline__ int batch_locate(const BatchImg *b, int nb, uint64_t p)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b[mid].off <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct dcf_tree {
    dcf_ctx *ctx = nullptr;
    uint64_t dims[3] = { 1, 1, 1 };
    int64_t n = 0;
    int conn = 4;
    int tree_min = 0;
    dcf_dtype dtype = DCF_U8;
    float pixdim[3] = { 1, 1, 1 };
    int grp = 0;
    /* device */
    void *img = nullptr;  /* owned copy when the caller gave host memory */
    int own_img = 0;
    float *gmax = nullptr; /* device scalars: [0] max, [1] min */
    int32_t *parent = nullptr;
    uint32_t *p2n = nullptr;
    int64_t K = 0;
    int32_t *npar = nullptr;
    float *ngrey = nullptr;
    DevAttr attr;       /* node-indexed */
    DevAttr attr_pix;   /* pixel-indexed records at levelroots, produced by the flooding build */
    uint32_t *counters = nullptr; /* 64 device words: per-round early-exit counters */
    int max_rounds = 32;
    float host_gmax = 0; /* filled lazily */
    int host_gmax_valid = 0;
    /* batched form (dcf_batch_filter): device image table, image count, 3D flag, per-node lambda */
    const BatchImg *bimg = nullptr;
    int64_t batch_tiles = 0; /* zone tiles of all images */
    int nbatch = 0;
    int batch3d = 0;
    const double *nlam = nullptr;
    /* a batch of small images (at most DCF_SMALL_MAX pixels each): built in LDS, one image per workgroup */
    int small_batch = 0;
    /* a tile of a tiled build: its origin in the whole image, added to the coordinates the attributes
     * sum (so they are those of the whole image) */
    int off[3] = { 0, 0, 0 };
};

/* internal filter flag (never set by callers): a min-tree's output stays in the inverted grey
 * space, as perform_dmp_operation writes it (operation.c:63-77) */
#define DCF_KEEP_INVERTED 0x80000000u

/* ---- error / launch helpers ------------------------------------------------------------------ */
int dcf_set_error(dcf_ctx *ctx, int code, const char *fmt, ...);
#define DCF_HIP(ctx, call)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return dcf_set_error((ctx), DCF_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

void dcf_prof_begin(dcf_ctx *ctx, const char *name, hipEvent_t *a, hipEvent_t *b);
void dcf_prof_end(dcf_ctx *ctx, const char *name, hipEvent_t a, hipEvent_t b);

#define DCF_LAUNCH(ctx, name, kern, grid, block, shmem, ...)                                       \
    do {                                                                                           \
        hipEvent_t ea_ = nullptr, eb_ = nullptr;                                                   \
        if ((ctx)->profiling && ((ctx)->prof_only.empty() || strstr((name), (ctx)->prof_only.c_str()))) \
            dcf_prof_begin((ctx), (name), &ea_, &eb_);                                             \
        hipLaunchKernelGGL(kern, (grid), (block), (shmem), (ctx)->stream, __VA_ARGS__);           \
        hipError_t le_ = hipGetLastError();                                                        \
        if (ea_) dcf_prof_end((ctx), (name), ea_, eb_);                                            \
        if (le_ != hipSuccess)                                                                     \
            return dcf_set_error((ctx), DCF_EDEVICE, "launch %s: %s", (name), hipGetErrorString(le_)); \
    } while (0)

/* ---- caching device allocator (context-owned, stream-ordered reuse) --------------------------- */
void *dcf_alloc(dcf_ctx *ctx, size_t bytes);
void dcf_release(dcf_ctx *ctx, void *p);

/* ---- attribute table (AttribsArray, attributes.c:1224-1248) ---------------------------------- */
int dcf_group_of(int attr_idx);

/* ---- launchers (dcf_build.hip) ---------------------------------------------------------------- */
int dcf_build_tree(dcf_tree *t, int grp);      /* parent, nodes and the attributes of group grp */
int dcf_accumulate_group(dcf_tree *t, int grp); /* attributes of group grp on the existing tree */
/* node arrays from t->parent (levelroot flags handed to the caller) */
int dcf_tree_nodes_from_parent(dcf_tree *t, uint32_t **flag_out);
int dcf_alloc_attr(dcf_ctx *ctx, DevAttr &a, int grp, int64_t K);

/* ---- launchers (dcf_tiles.hip) ---------------------------------------------------------------- */
/* tiled build: t holds the whole image (img, dims, conn, tree type, gmax of a min-tree, parent and p2n
 * arrays allocated); grid = tiles per axis.  Fills parent, the node arrays and group grp's attributes. */
int dcf_tiled_build(dcf_tree *t, const int grid[3], int grp, int schedule);
/* independent jobs on the context's worker contexts (own stream and host thread each), after the work
 * queued on ctx's stream; returns once every worker stream is drained (dcf_api.cpp) */
int dcf_run_on_workers(dcf_ctx *ctx, int njobs, const std::function<int(dcf_ctx *, int)> &job);
/* creates the context's further level-loop streams up to n (dcf_api.cpp) */
int dcf_add_aux_streams(dcf_ctx *c, size_t n);

/* ---- launchers (dcf_flood.hip) ---------------------------------------------------------------- */
int dcf_flood_build(dcf_tree *t, int grp);      /* canonical parent + pixel-indexed attributes */
/* f32 images: dense ranks of the tree's float keys (rank[p], and the largest rank) */
int dcf_rank_keys(dcf_tree *t, uint32_t *rank, uint32_t *maxrank);
/* batched form: per-image lambda (device, one per image), outputs (device table) */
int dcf_ops_batch_filter(dcf_tree *t, int attr_idx, int decision, const double *lam_img, int extract, const BatchOut *outs,
                         const uint32_t *gmax_img, unsigned flags);
/* write_tree: per node its levelroot pixel, parent node, grey, area function and printed columns */
int dcf_ops_tree_table(dcf_tree *t, int grp, std::vector<uint32_t> &n2p, std::vector<int32_t> &npar, std::vector<float> &grey,
                       std::vector<double> &size, std::vector<double> &cols, int &ncol);
/* batched form: widen / copy one image into the batch buffer; per-image grey max and inversion */
int dcf_batch_copy_in(dcf_ctx *ctx, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n);
int dcf_batch_invert(dcf_ctx *ctx, void *cat, int dtype, const BatchImg *bimg, int nb, uint32_t *gmax_img, int64_t n);

/* ---- launchers (dcf_ops.hip) ------------------------------------------------------------------ */
int dcf_ops_filter(dcf_tree *t, int attr_idx, int decision, double lambda, int extract, void *out_dev,
                   dcf_dtype out_dtype, unsigned flags);
int dcf_ops_check(dcf_tree *t, int attr_idx, float *out_dev);
int dcf_ops_spectrum(dcf_tree *t, int attr1, const float *l1, int n1, int attr2, const float *l2, int n2,
                     int second_is_area, double *out_host);
int dcf_ops_minmax(dcf_tree *t);
/* allocation-free spectrum: the scratch array's address for ctx's device (ctx->device current), and the
 * 1D spectrum with its buffers in that array */
int dcf_noalloc_init(dcf_ctx *ctx);
int dcf_ops_spectrum_noalloc(dcf_tree *t, int attr1, const float *l1, int n1, double *out_host);
void dcf_free_attr(dcf_ctx *ctx, DevAttr &a);

#endif
