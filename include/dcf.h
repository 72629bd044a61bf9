/*
 * dcf.h -- C ABI of the MI355X component-tree library (libdcf.so).
 *
 * A drop-in for disccofan's component-tree path: build the max-tree or min-tree of a 2D/3D
 * image, accumulate node attributes, apply the attribute filter rules and compute 1D/2D pattern
 * spectra, all on one GPU.  Plain pointers and sizes only; every call returns a status code and
 * no call keeps global state (a context owns a HIP stream and its trees).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   build_tree_with_attributes(Arguments*, Node*)          src/sources/flood.c:175
 *   flood_attributes(Node*)                                  src/sources/flood.c:184
 *   tree_filtering(Node*, value*, ulong, int, int, double)   src/sources/filtering.c:18
 *   tree_attribute_check(Node*, value*, ulong, ulong, ...)   src/sources/filtering.c:47
 *   tree_pattern_spectrum(...)                               src/sources/pattern.c:27
 *   tree_pattern_spectrum2d(...)                             src/sources/pattern.c:110
 *   lambda_vector_read(char*, double)                        src/sources/lambdavec.c:35
 *   AttribsArray[18] / Decisions[4] plugin tables            src/sources/attributes.c:1224, filtering.c:271
 * and the operator semantics of perform_operation (src/sources/operation.c:110-212) and
 * loop_operations (src/sources/disccofan.c:137-205) for the operations on this path.
 *
 * Pointers: by default image/output pointers are host memory; pass DCF_DEVICE_PTR in `flags`
 * when they are device (HBM) pointers already resident on the context's device.
 * Threading: contexts are independent; calls on one context must be serialised by the caller.
 */
#ifndef DCF_H
#define DCF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dcf_ctx dcf_ctx;
typedef struct dcf_tree dcf_tree;

typedef enum { DCF_U8 = 0, DCF_U16 = 1, DCF_F32 = 2 } dcf_dtype;

/* tree type, as the reference's tree_type string "max" / "min" (preprocessing.c:84-91) */
enum { DCF_MAX_TREE = 0, DCF_MIN_TREE = 1 };

/* decision index, as Decisions[] (filtering.c:271-277) */
enum { DCF_DIRECT = 0, DCF_MIN = 1, DCF_MAX = 2, DCF_SUBTRACTIVE = 3 };

/* status codes */
enum {
    DCF_OK = 0,
    DCF_EINVAL = -1,       /* bad argument (connectivity, attribute or decision index, shape) */
    DCF_ENOMEM = -2,       /* device or host allocation failed */
    DCF_EDEVICE = -3,      /* HIP runtime error */
    DCF_EUNSUPPORTED = -4, /* valid but not implemented on this path */
    DCF_EIO = -5,          /* file could not be read */
    DCF_EGROUP = -6        /* attribute not in the tree's accumulated group (call dcf_tree_accumulate) */
};

/* flags */
#define DCF_DEVICE_PTR 0x1u     /* image / output pointer is a device pointer */
#define DCF_ALL_RULES 0x2u      /* dcf_filter: always run the general rule kernels (no increasing-attribute shortcut) */

/* number of attributes, as NUMATTR (types.h) */
#define DCF_NUM_ATTR 18

/* ---- context ---------------------------------------------------------------------------------- */
int dcf_ctx_create(int device, unsigned flags, dcf_ctx **out);
/* frees the context's remaining trees too; their handles are invalid afterwards */
int dcf_ctx_destroy(dcf_ctx *ctx);
/* Use an existing HIP stream (hipStream_t passed as void*) for every launch of this context;
 * NULL restores the context's own stream. */
int dcf_ctx_set_stream(dcf_ctx *ctx, void *stream);
void *dcf_ctx_stream(dcf_ctx *ctx);
/* Tunables of the build (no counterpart in the reference; they change the schedule, never a result), by
 * the names of the environment variables that give new contexts their defaults: DCF_ZONES, DCF_TARGETS,
 * DCF_RECORDS, DCF_FIX_BLOCKS, DCF_FINAL_BLOCKS, DCF_GRAPH, DCF_ZIP_SEGMENTS, DCF_WORKERS, DCF_BATCH_PARTS,
 * DCF_LARGE_PIXELS, DCF_SMALL_PIXELS, DCF_FINE_LOOPS, DCF_WORKER_LOOPS, DCF_ZONE_TILE2D ("WxH"), DCF_ZONE_TILE3D ("N"), DCF_LEVEL_TILE2D / DCF_LEVEL_GROUP2D / DCF_LEVEL_TILE3D / DCF_LEVEL_GROUP3D
 * ("WxHxD").  DCF_EINVAL for
 * an unknown name or a value out of range.  The autotuner (disccofan_amd/tune.py) sets them here. */
int dcf_ctx_set_option(dcf_ctx *ctx, const char *name, const char *value);
/* Build-loop launch mode (no counterpart in the reference): with on != 0, every digit pass of a build
 * captures its level loop (three launches per grey level) into one HIP graph and launches that; off
 * (the default) launches the kernels one by one.  Results are identical.  The environment variable
 * DCF_GRAPH sets the default of new contexts. */
int dcf_ctx_set_graph(dcf_ctx *ctx, int on);
const char *dcf_last_error(dcf_ctx *ctx);
int dcf_ctx_synchronize(dcf_ctx *ctx);

/* Per-kernel timing with HIP events on the context's stream (off by default).  When on, every
 * launch is bracketed by two events; dcf_kernel_stats returns the launch count and the summed
 * event time of kernels whose name contains `name`. */
int dcf_ctx_set_profiling(dcf_ctx *ctx, int on);
/* restrict the events to kernels whose name contains `name` (NULL or "" = every kernel), so that a
 * timed run can measure one kernel without bracketing every other launch */
int dcf_ctx_profile_only(dcf_ctx *ctx, const char *name);
int dcf_kernel_stats(dcf_ctx *ctx, const char *name, int64_t *launches, double *total_ms);
int dcf_kernel_stats_reset(dcf_ctx *ctx);
/* name of the i-th distinct profiled kernel, or NULL past the end */
const char *dcf_kernel_name(dcf_ctx *ctx, int i);

/* ---- build (build_tree_with_attributes, flood.c:175) ------------------------------------------- *
 * img: dims[0]*dims[1]*dims[2] values, x fastest (index = x + y*dims[0] + z*dims[0]*dims[1]).
 * connectivity: 4 or 8 (2D), 6 or 26 (3D); mapped like check_connectivity (io.c:64-79).
 * tree_type: DCF_MAX_TREE or DCF_MIN_TREE (min-tree = max-tree of gmax - g).
 * attr_idx: the first operation's attribute; its group is accumulated (disccofan.c:145-148).
 * pixdim: physical pixel size (NULL = 1,1,1), as data_properties.pixdim.
 * Size: at most 2^30 - 1 pixels (DCF_EUNSUPPORTED from 2^30); DCF_EINVAL when a dimension is 0 or
 * the pixel count reaches 2^31 - 1, a count taken without 64-bit wrap-around of the product. */
int dcf_tree_build(dcf_ctx *ctx, const void *img, dcf_dtype dtype, const uint64_t dims[3], int connectivity,
                   int tree_type, int attr_idx, const float pixdim[3], unsigned flags, dcf_tree **out);
int dcf_tree_free(dcf_tree *tree);

/* ---- tiled build with boundary-tree merging (the reference's distributed path on one GPU) ------- *
 * Replaces the distributed build of the reference: the tile decomposition with a one-pixel overlap
 * (read_* in src/sources/io.c:860-882), build_local_tree per tile (flood.c:175), create_boundary
 * (boundary.c:1553-1707), merge_boundary_trees / merge_branch (boundary.c:427-724), the update and
 * correction of the local trees (boundary.c:1201-1399).  The image is cut into grid[0] x grid[1] x
 * grid[2] tiles (per axis near-equal, the first tiles one pixel longer when the size does not divide,
 * each extended one pixel into every neighbour), every tile's tree is built on its own (the tiles
 * concurrently on DCF_WORKERS worker contexts, like the parts of dcf_batch_filter), and their
 * boundary trees (the nodes of the overlap pixels and all their ancestors) are merged on the GPU.  The
 * result is the tree dcf_tree_build gives for the whole image: the same canonical parent array, node
 * arrays and attributes, for every filter, check and spectrum.  Arguments as dcf_tree_build; grid[a]
 * must be at least 1 and at most dims[a] (grid[2] = 1 for a 2D image), 4096 tiles at most. */
int dcf_tree_build_tiled(dcf_ctx *ctx, const void *img, dcf_dtype dtype, const uint64_t dims[3], int connectivity,
                         int tree_type, int attr_idx, const float pixdim[3], const int grid[3], unsigned flags,
                         dcf_tree **out);

/* Re-accumulate the attributes of attr_idx's group on the existing tree when the group differs
 * from the accumulated one (flood_attributes, flood.c:184; group reuse disccofan.c:179-200). */
int dcf_tree_accumulate(dcf_tree *tree, int attr_idx);
int dcf_tree_group(const dcf_tree *tree);

/* Canonical parent array: every pixel's levelroot, levelroots -> levelroot of the parent node,
 * the root -> -1; levelroot = minimum raster index of the node's pixels at its grey level. */
int dcf_tree_parent(dcf_tree *tree, int32_t *parent, unsigned flags);
/* number of nodes (levelroots) */
int dcf_tree_num_nodes(dcf_tree *tree, int64_t *num_nodes);

/* ---- filter (tree_filtering, filtering.c:18; perform_filter_or_extract, operation.c:10-29) ------ *
 * out: one value per pixel in out_dtype (DCF_F32 gives the reference's float output; DCF_U8/U16
 * are exact for integer images of that range).  extract != 0 gives g - f; a min-tree's output is
 * re-inverted gmax - f, as the reference.  The min and max rules give the reference's single-thread
 * scan result (tree_filter_min / tree_filter_max, filtering.c:55-208: the max rule's stale values,
 * the min rule's zeros below a failing root).  DCF_EUNSUPPORTED: either rule of a non-increasing
 * attribute (any attribute but area / volume) on a float max-tree with a node of grey 0 below a
 * negative root. */
int dcf_filter(dcf_tree *tree, int attr_idx, int decision, double lambda, int extract, void *out,
               dcf_dtype out_dtype, unsigned flags);

/* ---- differential profile (perform_dmp_operation, operation.c:63-77) ------------------------------ *
 * lambdas as returned by dcf_lvec_read (lambda_0 = 0 first, n values): out receives n-1 images, image
 * i = g - filter(lambda_{i+1}) with the given decision (the dmp operation uses the direct rule); for
 * a min-tree g and the filter stay in the inverted grey space, as the reference writes them. */
int dcf_dmp(dcf_tree *tree, int attr_idx, int decision, const float *lambdas, int n, void *out, dcf_dtype out_dtype,
            unsigned flags);

/* ---- check (tree_attribute_check, filtering.c:47-52): out[p] = (float) attr(levelroot(p)) ---- */
int dcf_check(dcf_tree *tree, int attr_idx, float *out, unsigned flags);

/* ---- spectra (pattern.c:27, 110), background = 1 (operation.c:82) ------------------------------ *
 * lambdas: thresholds with lambda_0 = 0 included, as returned by dcf_lvec_read.
 * 2D: bin index s1 + s2*n1; attr2 == 0 with attr1 != 0 means "area of attr1's group"
 * (operation.c:97-98); attr1 == 0 with attr2 != 0 is refused (operation.c:93-95).
 * A spectrum has at most DCF_MAX_SPECTRUM_BINS bins (n, or n1*n2); more is DCF_EINVAL. */
#define DCF_MAX_SPECTRUM_BINS (1 << 28)
int dcf_spectrum1d(dcf_tree *tree, int attr_idx, const float *lambdas, int n, double *out);
int dcf_spectrum2d(dcf_tree *tree, int attr1, const float *l1, int n1, int attr2, const float *l2, int n2,
                   double *out);

/* ---- allocation-free 1D spectrum ------------------------------------------------------------------ *
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): dcf_spectrum1d with every
 * device buffer in a fixed array of the library (28 bytes per bin, DCF_NOALLOC_MAX_BINS bins at most,
 * held on each device the library runs on until the process ends); it calls no allocator.  More than
 * DCF_NOALLOC_MAX_BINS bins is DCF_EINVAL.  Every call on one device shares that device's array, so
 * calls that overlap on one device (several host threads, several contexts) race; serialise them. */
#define DCF_NOALLOC_MAX_BINS 65536
int dcf_spectrum1d_noalloc(dcf_tree *tree, int attr_idx, const float *lambdas, int n, double *out);

/* ---- tree export (write_tree, io.c:1431-1573) --------------------------------------------------- *
 * The "tree" operation's CSV: header, then one row per node (levelroot) in raster order --
 * index,x,y,z,gval,parent,flux,area and the columns of attr_idx's group, with the reference's printf
 * formats.  The group must be the accumulated one (DCF_EGROUP otherwise). */
int dcf_tree_write_csv(dcf_tree *tree, int attr_idx, const char *path);

/* ---- lambda-vector files (lambda_vector_read, lambdavec.c:35-84) -------------------------------- *
 * On success *lambdas (malloc'd, free with dcf_lvec_free) holds count+1 values, lambda_0 = 0. */
int dcf_lvec_read(const char *path, double imscale, float **lambdas, int *n);
void dcf_lvec_free(float *lambdas);

/* ---- attribute / decision tables (AttribsArray, Decisions) -------------------------------------- */
const char *dcf_attr_name(int attr_idx);
const char *dcf_attr_short_name(int attr_idx);
int dcf_attr_group(int attr_idx);
const char *dcf_decision_name(int decision);

/* ---- batched form ------------------------------------------------------------------------------ *
 * Many images of uneven sizes, each with its own lambda: the batch analogue of a loop of
 * dcf_tree_build + dcf_filter per image (build_tree_with_attributes flood.c:175 + tree_filtering
 * filtering.c:18 per image, as loop_operations runs them for one image).  Every image gets its own
 * max-tree or min-tree (no pixel of one image touches another) and all of them are built by one
 * level loop over the whole batch, then filtered with their own lambda.  Images must all be 2D or
 * all be 3D; u8 and u16 may be mixed; out receives the image's own dtype.  The batch holds at most
 * 2^30 pixels.  flags: DCF_DEVICE_PTR (img and out are device pointers), DCF_ALL_RULES.
 * tree_type is every image's tree unless the image's own flags name one (DCF_IMAGE_MAX_TREE /
 * DCF_IMAGE_MIN_TREE): the max-tree and the min-tree of one image can share a call.
 * The call runs in parts: each image of at least DCF_LARGE_PIXELS pixels (default 2^24) on its own when
 * there are two or more of them, and the other images in batch trees, u8 and u16 apart, DCF_BATCH_PARTS
 * per dtype (default 0: workers / dtypes).  With two or more parts and DCF_WORKERS > 1 (default 4 when
 * GPU_MAX_HW_QUEUES was at least 8 at context creation, else 3) the parts run concurrently on worker
 * contexts, each with its own HIP stream and host thread; they wait for the work already queued on ctx's
 * stream, and the call returns once they are done.  The streams overlap only on hardware queues of their
 * own: run the process with GPU_MAX_HW_QUEUES=16.  Results do not depend on the split. */
typedef struct {
    const void *img;  /* dims[0]*dims[1]*dims[2] values of dtype, x fastest */
    dcf_dtype dtype;  /* DCF_U8 or DCF_U16 */
    uint64_t dims[3];
    double lambda;    /* this image's lambda */
    void *out;        /* filtered (or extracted) image, same dtype and size */
    unsigned flags;   /* 0: the call's tree type; DCF_IMAGE_MAX_TREE or DCF_IMAGE_MIN_TREE: this one */
} dcf_batch_image;
#define DCF_IMAGE_MAX_TREE 0x1u
#define DCF_IMAGE_MIN_TREE 0x2u
int dcf_batch_filter(dcf_ctx *ctx, const dcf_batch_image *images, int count, int connectivity, int tree_type, int attr_idx,
                     int decision, int extract, unsigned flags);

/* ---- operator driver (loop_operations, disccofan.c:137-205) ------------------------------------- *
 * ops: the reference's operation tuples, e.g. "(\"filter\",0,100),(\"pattern\",0,\"lvec.txt\",1)".
 * Runs them in order on one image with attribute-group reuse; outputs are written by the callback
 * (kind: 0 filtered/extracted image (float, n values), 1 check image (float), 2 1D spectrum
 * (double, count values), 3 2D spectrum (double), 4 differential profile (float, (lambdas-1) images),
 * 5 tree CSV text (char, count bytes)).  Unsupported operations return
 * DCF_EUNSUPPORTED before anything runs. */
typedef void (*dcf_output_fn)(void *user, int op_index, int kind, const void *data, int64_t count);
int dcf_run_operations(dcf_ctx *ctx, const void *img, dcf_dtype dtype, const uint64_t dims[3], int connectivity,
                       int tree_type, const float pixdim[3], const char *ops, dcf_output_fn cb, void *user);

#ifdef __cplusplus
}
#endif

#endif /* DCF_H */
