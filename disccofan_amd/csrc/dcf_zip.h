/*
 * dcf_zip.h -- the lock-free "zip" merge of two ancestor chains, shared by the LDS tile kernel,
 * the global seam kernel and the CPU concurrency test (tests/cpu_zip_test.cpp).
 *
 * Node order.  Pixels are totally ordered by ord(p) = (key(p) << 32) | p, where key is the grey
 * key (the value itself for u8/u16, the order-preserving transform of flood.c:362-371 for f32).
 * This is the reference's rank order: a stable sort of (grey, index) (flood.c:378-502).
 *
 * Link words.  Every pixel owns one 64-bit word  (key(parent) << 32) | parent,  parent = NONE
 * for a root.  Every chain p -> parent(p) -> ... is strictly decreasing in ord, so the forest is
 * a (partially compressed) rank tree: the levelroot of p is the last ancestor with key(p), and
 * it is the minimum-index pixel of p's node (the canonical form of flood.c:127-141).
 *
 * zip(x, y) merges the chains of two adjacent pixels the way merge_nodes (flood.c:292-353)
 * merges two trees across a slab border: walk down the higher chain while its next node is still
 * above the lower one, then splice the lower node in with a compare-and-swap and continue with
 * the remainder.  A link only ever moves to a node that is higher in ord and still below its
 * owner, so every CAS that succeeds keeps all chains sorted; a failed CAS returns the newer
 * link and the walk resumes from it.  Any interleaving of zips over all edges of a graph leaves
 * the unique rank tree of that graph (up to same-grey compression), which is why the tile,
 * seam and thread schedule cannot change the canonical result.
 *
 * Mem must provide:  uint64_t load(uint32_t i);  uint64_t cas(uint32_t i, uint64_t expect,
 * uint64_t desired)  (returns the value found).
 */
#ifndef DCF_ZIP_H
#define DCF_ZIP_H

#include <stdint.h>

#if defined(__HIPCC__)
#define DCF_HD __host__ __device__
#else
#define DCF_HD
#endif

#define DCF_NONE 0xFFFFFFFFu

DCF_HD static inline uint32_t dcf_link_idx(uint64_t w) { return (uint32_t)w; }
DCF_HD static inline uint32_t dcf_link_key(uint64_t w) { return (uint32_t)(w >> 32); }
DCF_HD static inline uint64_t dcf_make_link(uint32_t key, uint32_t idx) { return ((uint64_t)key << 32) | idx; }
DCF_HD static inline uint64_t dcf_ord(uint32_t key, uint32_t idx) { return ((uint64_t)key << 32) | idx; }

/* merge the chains of x and y (any order); returns the number of link words written */
template <class Mem>
DCF_HD static inline int dcf_zip(Mem &m, uint32_t x, uint32_t kx, uint32_t y, uint32_t ky)
{
    uint64_t ox = dcf_ord(kx, x), oy = dcf_ord(ky, y);
    int writes = 0;
    if (x == y) return 0;
    if (ox < oy) {
        uint64_t t = ox; ox = oy; oy = t;
        uint32_t u = x; x = y; y = u;
    }
    uint64_t w = m.load(x);
    for (;;) {
        /* invariant: ox > oy, w = a recent value of x's link word */
        uint32_t z = dcf_link_idx(w);
        uint64_t oz = 0;
        if (z != DCF_NONE) {
            oz = dcf_ord(dcf_link_key(w), z);
            if (oz >= oy) {
                if (z == y) return writes; /* chains already joined here */
                x = z; ox = oz;
                w = m.load(x);
                continue;
            }
        }
        /* z is below y (or x is a root): splice y between x and z */
        uint64_t nw = dcf_make_link((uint32_t)(oy >> 32), y);
        uint64_t seen = m.cas(x, w, nw);
        if (seen != w) { w = seen; continue; } /* x's link moved up meanwhile: re-examine */
        writes++;
        if (z == DCF_NONE) return writes;
        /* continue merging y's chain with z's */
        x = y; ox = oy;
        y = z; oy = oz;
        w = m.load(x);
    }
}

#endif /* DCF_ZIP_H */
