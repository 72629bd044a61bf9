/*
 * dcf_ops.hip -- attribute evaluation, filter rules, check images and pattern spectra on gfx950.
 *
 * Replaces tree_filtering / tree_filter_direct / tree_filter_min / tree_filter_max /
 * tree_filter_subtractive / tree_attribute_check (src/sources/filtering.c:18-261),
 * tree_pattern_spectrum[2d] (src/sources/pattern.c:27-203) and
 * the attribute functions of AttribsArray (src/sources/attributes.c:47-874).
 *
 * The reference walks parent chains per pixel with a `reached` memo.  Here every rule is first
 * solved per node (K nodes, coalesced), then one pass maps pixels to their node's value:
 *   direct       out(n) = g(first ancestor-or-self a with attr(a) >= lambda), 0 if none
 *                -> nearest-passing-ancestor by in-place pointer jumping
 *   subtractive  out(root) = pass ? g : 0;  out(n) = out(par) + pass(n) * (g(n) - g(par))
 *                -> rootfix sum by pointer doubling (ping-pong buffers)
 *   min, max    the reference's single-thread raster scan, solved per node from subtree minima
 *                (preorder range minima) and pointer doubling (max_general, min_general)
 *   For the area attribute (increasing: a parent's area is never smaller) every ancestor of a
 *   passing node passes, so the subtractive, min and max rules collapse to the direct one (the min
 *   rule under a failing root aside) and the cheaper direct kernels give the identical result
 *   (DCF_ALL_RULES forces the general kernels).
 * Attribute values are evaluated in FP64 from exact integer moments with the reference's
 * operation order, including its float evaluation of area_attribute (attributes.c:105: the
 * ulong area is converted to float before the product), so bins and thresholds match bit for bit.
 */
#include <hipcub/hipcub.hpp>
#include <mutex>
#include <vector>

#include "dcf_internal.h"

struct AttrCtx {
    DevAttr A;
    float p0, p1, p2;
};

/* closed-form eigenvalues, ascending (replaces EISPACK rs, attributes.c:654,679) */
__device__ static void eig2(double a, double b, double d, double ev[2])
{
    double tr = 0.5 * (a + d), df = 0.5 * (a - d);
    double r = sqrt(df * df + b * b);
    ev[0] = tr - r;
    ev[1] = tr + r;
}

__device__ static void eig3(const double m[6], double ev[3])
{
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5];
    double p1 = d * d + e * e + f * f;
    if (p1 == 0.0) {
        double t0 = a, t1 = b, t2 = c, s;
        if (t1 < t0) { s = t0; t0 = t1; t1 = s; }
        if (t2 < t1) { s = t1; t1 = t2; t2 = s; }
        if (t1 < t0) { s = t0; t0 = t1; t1 = s; }
        ev[0] = t0; ev[1] = t1; ev[2] = t2;
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

/* attribute `idx` of node k (AttribsArray[idx].attribute) */
__device__ static double attr_value(const AttrCtx &c, int idx, int64_t k)
{
    const DevAttr &A = c.A;
    const float p0 = c.p0, p1 = c.p1, p2 = c.p2;
    uint32_t area = A.area[k];
    switch (idx) {
    case 0: return (double)(p0 * p1 * p2 * (float)area);
    case 1: {
        double v = (double)(A.mx[0][k] - A.mn[0][k]) * (double)(A.mx[1][k] - A.mn[1][k]) * (double)(A.mx[2][k] - A.mn[2][k]);
        v *= (double)(p0 * p1 * p2);
        return v;
    }
    case 2: {
        double mnx = p0 * (double)A.mn[0][k], mny = p1 * (double)A.mn[1][k], mnz = p2 * (double)A.mn[2][k];
        double mxx = p0 * (double)A.mx[0][k], mxy = p1 * (double)A.mx[1][k], mxz = p2 * (double)A.mx[2][k];
        return (mxx - mnx) * (mxx - mnx) + (mxy - mny) * (mxy - mny) + (mxz - mnz) * (mxz - mnz);
    }
    case 3: return p0 * (double)(A.mx[0][k] - A.mn[0][k]);
    case 4: return p1 * (double)(A.mx[1][k] - A.mn[1][k]);
    case 5: return p2 * (double)(A.mx[2][k] - A.mn[2][k]);
    case 6: return p0 * (double)A.s[0][k] / (double)area;
    case 7: return p1 * (double)A.s[1][k] / (double)area;
    case 8: return p2 * (double)A.s[2][k] / (double)area;
    case 9: return (double)p0 * (double)A.s[1][k] / (double)A.s[0][k];
    case 10: return (double)p1 * (double)A.s[2][k] / (double)A.s[0][k];
    case 11: return (double)p2 * (double)A.s[3][k] / (double)A.s[0][k];
    case 12: return (double)A.s[0][k];
    default: break;
    }
    /* inertia group (inertia_arr, attributes.c:607-626) */
    double in0 = (double)p0 * p1 * p2 * area;
    double in1 = p0 * (double)A.s[0][k], in2 = p1 * (double)A.s[1][k], in3 = p2 * (double)A.s[2][k];
    double in4 = (double)(p0 * p0) * (double)A.s[3][k];
    double in5 = (double)(p1 * p1) * (double)A.s[4][k];
    double in6 = (double)(p2 * p2) * (double)A.s[5][k];
    double corr = in0 / 12;
    double mxx = in4 - in1 * in1 / in0 + p0 * corr;
    double myy = in5 - in2 * in2 / in0 + p1 * corr;
    if (idx == 13 || idx == 17) {
        if (in3 == 0) {
            double tr = mxx + myy;
            return idx == 13 ? tr : tr / (in0 * in0); /* pow(A, 2.0) is the correctly rounded square */
        }
        double mzz = in6 - in3 * in3 / in0 + p2 * corr;
        double tr = mxx + myy + mzz;
        return idx == 13 ? tr : tr / pow(in0, 5.0 / 3.0);
    }
    double in7 = (double)(p0 * p1) * (double)A.s[6][k];
    double mxy = in7 - in1 * in2 / in0;
    if (in3 == 0) {
        if (idx == 15) return -1;
        double ev[2];
        eig2(mxx, mxy, myy, ev);
        if (idx == 14) return ev[1] / ev[0];
        double ax0 = sqrt(4 * ev[1] / in0), ax1 = sqrt(4 * ev[0] / in0);
        return M_PI * ax0 * ax1 / in0;
    }
    double in8 = (double)(p1 * p2) * (double)A.s[7][k];
    double in9 = (double)(p0 * p2) * (double)A.s[8][k];
    double m6[6] = { mxx, myy, in6 - in3 * in3 / in0 + p2 * corr, mxy, in8 - in2 * in3 / in0, in9 - in1 * in3 / in0 };
    double ev[3];
    eig3(m6, ev);
    if (idx == 14) return ev[2] / ev[1];
    if (idx == 15) return ev[1] / ev[0];
    double a0 = sqrt(20 * ev[2] / in0), a1 = sqrt(20 * ev[1] / in0), a2 = sqrt(20 * ev[0] / in0);
    return M_PI * a0 * a1 * a2 / (6 * in0);
}

/* AttribsArray[idx].area of node k */
__device__ static double attr_area(const AttrCtx &c, int grp, int64_t k)
{
    uint32_t area = c.A.area[k];
    switch (grp) {
    case 0: return (double)(c.p0 * c.p1 * c.p2 * (float)area);
    case 2: return (double)(c.p0 * c.p1 * c.p2) * (double)area;
    case 4: return (double)c.p0 * c.p1 * c.p2 * area;
    default: return (double)area;
    }
}

/* find_scale (pattern.c:7-24) */
__device__ static int find_scale(const float *lv, int n, double a)
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

static __global__ __launch_bounds__(256) void k_fill_u32(uint32_t *a, uint32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* ------------------------------------------------------------------------------------------------ */
/* per-node kernels                                                                                 */
/* ------------------------------------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void k_attr_eval(AttrCtx c, int idx, double *__restrict__ out, int64_t K)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) out[k] = attr_value(c, idx, k);
}

/* direct rule: J(k) = k if attr(k) >= lambda else parent; pass flags */
/* nlam: per-node lambda of a batch (each node takes its image's lambda), or nullptr */
__global__ __launch_bounds__(256) void k_direct_init(AttrCtx c, int idx, double lambda, const double *__restrict__ nlam,
                                                     const int32_t *__restrict__ npar, int32_t *__restrict__ J,
                                                     uint8_t *__restrict__ pass, int64_t K)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    bool ok = attr_value(c, idx, k) >= (nlam ? nlam[k] : lambda);
    pass[k] = ok;
    J[k] = ok ? (int32_t)k : npar[k];
}

/* in place: J(k) <- J(J(k)) while J(k) is a failing node; values only move up the path to the
 * fixed target, so reads of concurrently updated entries are always valid */
__global__ __launch_bounds__(256) void k_direct_jump(int32_t *__restrict__ J, const uint8_t *__restrict__ pass,
                                                     uint32_t *__restrict__ active, int round, int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool more = false;
    if (k < K) {
        int32_t j = J[k];
        if (j >= 0 && !pass[j]) {
            int32_t jj = J[j];
            J[k] = jj;
            more = jj >= 0 && !pass[jj];
        }
    }
    unsigned long long b = __ballot(more);
    /* "another round is needed": a flag, set once (an atomic add per wave on one word serialised
     * hundreds of thousands of updates per round) */
    if ((threadIdx.x & (warpSize - 1)) == 0 && b &&
        __hip_atomic_load(&active[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        __hip_atomic_store(&active[round], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_direct_value(const int32_t *__restrict__ J, const float *__restrict__ ngrey,
                                                      float *__restrict__ val, int64_t K)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    int32_t j = J[k];
    val[k] = j < 0 ? 0.0f : ngrey[j];
}

/* subtractive rule: F(k) = pass ? g(k) - g(par) : 0 (root: pass ? g : 0); out = rootfix sum */
__global__ __launch_bounds__(256) void k_sub_init(AttrCtx c, int idx, double lambda, const double *__restrict__ nlam,
                                                  const int32_t *__restrict__ npar, const float *__restrict__ ngrey,
                                                  double *__restrict__ F, int32_t *__restrict__ A, int64_t K)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    bool ok = attr_value(c, idx, k) >= (nlam ? nlam[k] : lambda);
    int32_t p = npar[k];
    F[k] = ok ? (p >= 0 ? (double)ngrey[k] - (double)ngrey[p] : (double)ngrey[k]) : 0.0;
    A[k] = p;
}

__global__ __launch_bounds__(256) void k_sub_round(const double *__restrict__ F, double *__restrict__ F2,
                                                   const int32_t *__restrict__ A, int32_t *__restrict__ A2,
                                                   uint32_t *__restrict__ active, int round, int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t a2 = -1;
    if (k < K) {
        int32_t a = A[k];
        double f = F[k];
        if (a >= 0) {
            f += F[a];
            a2 = A[a];
        }
        F2[k] = f;
        A2[k] = a2;
    }
    unsigned long long b = __ballot(k < K && a2 >= 0);
    /* "another round is needed": a flag, set once (an atomic add per wave on one word serialised
     * hundreds of thousands of updates per round) */
    if ((threadIdx.x & (warpSize - 1)) == 0 && b &&
        __hip_atomic_load(&active[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        __hip_atomic_store(&active[round], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_d2f(const double *__restrict__ F, float *__restrict__ val, int64_t K)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) val[k] = (float)F[k];
}

/* ------------------------------------------------------------------------------------------------ */
/* per-pixel output                                                                                 */
/* ------------------------------------------------------------------------------------------------ */
__device__ __forceinline__ float px_grey_ops(const ImgView &iv, int64_t p)
{
    float v;
    if (iv.dtype == DCF_U8) v = (float)((const uint8_t *)iv.img)[p];
    else if (iv.dtype == DCF_U16) v = (float)((const uint16_t *)iv.img)[p];
    else v = ((const float *)iv.img)[p];
    return iv.invert ? *iv.gmax - v : v;
}

/* the rule's value at pixel p of node k: val[k], except under the max rule of a non-increasing
 * attribute (kq != nullptr), where a pixel of a node fixed before its turn, and the levelroot of a node
 * ever fixed, keep the node's grey */
__device__ __forceinline__ float pixel_value(int64_t p, uint32_t k, const float *__restrict__ val, const uint32_t *__restrict__ kq,
                                             const uint32_t *__restrict__ klr, const float *__restrict__ ngrey)
{
    if (kq) {
        const uint32_t q = kq[k];
        if (q != 0xFFFFFFFFu && ((uint64_t)p > q || (uint64_t)p == klr[k])) return ngrey[k];
    }
    return val[k];
}

/* out[p] = post(value at p), post = extract (g - f) then min-tree re-inversion (gmax - f) */
__global__ __launch_bounds__(256) void k_filter_out(ImgView iv, const uint32_t *__restrict__ p2n, const float *__restrict__ val,
                                                    const uint32_t *__restrict__ kq, const uint32_t *__restrict__ klr,
                                                    const float *__restrict__ ngrey, int extract, int reinvert, void *__restrict__ out,
                                                    int out_dtype, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = p2n[p];
    float f = pixel_value(p, k, val, kq, klr, ngrey);
    if (extract) f = px_grey_ops(iv, p) - f;
    if (reinvert && iv.invert) f = *iv.gmax - f;
    if (out_dtype == DCF_F32) ((float *)out)[p] = f;
    else if (out_dtype == DCF_U16) ((uint16_t *)out)[p] = (uint16_t)f;
    else ((uint8_t *)out)[p] = (uint8_t)f;
}

/* batched form: every node takes the lambda of its image */
__global__ __launch_bounds__(256) void k_node_lambda(const BatchImg *__restrict__ bimg, int nb, const double *__restrict__ lam,
                                                     const uint32_t *__restrict__ p2n, double *__restrict__ nlam, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    /* every pixel of a node lies in the same image: concurrent writers store the same value */
    nlam[p2n[p]] = lam[batch_locate(bimg, nb, (uint64_t)p)];
}

/* batched form: filtered value of every pixel into its image's output (extract on the possibly
 * inverted greys, then min-tree re-inversion with the image's own grey maximum, as k_filter_out) */
__global__ __launch_bounds__(256) void k_batch_out(ImgView iv, const BatchImg *__restrict__ bimg, int nb,
                                                   const BatchOut *__restrict__ outs, const uint32_t *__restrict__ gmax_img,
                                                   const uint32_t *__restrict__ p2n, const float *__restrict__ val,
                                                   const uint32_t *__restrict__ kq, const uint32_t *__restrict__ klr,
                                                   const float *__restrict__ ngrey, int extract, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = batch_locate(bimg, nb, (uint64_t)p);
    const BatchOut o = outs[i];
    const uint64_t local = (uint64_t)p - bimg[i].off;
    float f = pixel_value(p, p2n[p], val, kq, klr, ngrey);
    if (extract) f = px_grey_ops(iv, p) - f;
    if (gmax_img && bimg[i].inv) f = (float)gmax_img[i] - f;
    if (o.dtype == DCF_F32) ((float *)o.ptr)[local] = f;
    else if (o.dtype == DCF_U16) ((uint16_t *)o.ptr)[local] = (uint16_t)f;
    else ((uint8_t *)o.ptr)[local] = (uint8_t)f;
}

__global__ __launch_bounds__(256) void k_check_out(const uint32_t *__restrict__ p2n, const double *__restrict__ av,
                                                   float *__restrict__ out, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = (float)av[p2n[p]];
}

/* ------------------------------------------------------------------------------------------------ */
/* pattern spectra: LDS histogram per block, exact 64-bit integer bins for integer images          */
/* ------------------------------------------------------------------------------------------------ */
#define SPEC_LDS_BINS 2048

/* sum over the wave's lanes (every lane gets the total) */
__device__ __forceinline__ unsigned long long spec_wsum(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, o, 64), hi = __shfl_xor((uint32_t)(v >> 32), o, 64);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

__device__ __forceinline__ double spec_wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/* Every levelroot adds (g - g_parent) * area to the bin of its attribute (pattern.c:80-96, 170-191).
 * With at most SPEC_LDS_BINS bins each block keeps the histogram in LDS, and each wave first sums its
 * lanes' contributions per distinct bin (a ballot picks the lanes of the leader's bin), so one LDS atomic
 * per distinct bin and wave replaces one per node: most nodes are small and share the first bins.
 * Integer greys at unit pixel size add exact 64-bit integers, so the order of the additions does not
 * matter; float greys add doubles (their sum order differs from the reference's thread split anyway). */
__global__ __launch_bounds__(256) void k_spectrum(AttrCtx c, int grp, int attr1, const float *__restrict__ l1, int n1, int attr2,
                                                  const float *__restrict__ l2, int n2, int second_is_area,
                                                  const int32_t *__restrict__ npar, const float *__restrict__ ngrey, int exact,
                                                  unsigned long long *__restrict__ hist_i, double *__restrict__ hist_d, int64_t K)
{
    __shared__ unsigned long long sh_i[SPEC_LDS_BINS];
    __shared__ double sh_d[SPEC_LDS_BINS];
    const int nb = n1 * n2;
    const bool use_lds = nb <= SPEC_LDS_BINS;
    if (use_lds) {
        for (int i = threadIdx.x; i < nb; i += blockDim.x) { sh_i[i] = 0; sh_d[i] = 0; }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    /* every lane of a wave runs the same iterations (the ballots below need the whole wave) */
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < K; base += stride) {
        const int64_t k = base + threadIdx.x;
        int bin = -1;
        unsigned long long vi = 0;
        double vd = 0;
        if (k < K) {
            const int32_t p = npar[k];
            const float dh = p >= 0 ? ngrey[k] - ngrey[p] : ngrey[k];
            if (dh != 0) {
                bin = find_scale(l1, n1, attr_value(c, attr1, k));
                if (l2) {
                    const double a2 = second_is_area ? attr_area(c, grp, k) : attr_value(c, attr2, k);
                    bin += find_scale(l2, n2, a2) * n1;
                }
                const double ar = attr_area(c, grp, k);
                if (exact) vi = (unsigned long long)(long long)dh * (unsigned long long)(long long)ar;
                else vd = (double)dh * ar;
            }
        }
        if (!use_lds) {
            if (bin >= 0) {
                if (exact) atomicAdd(&hist_i[bin], vi);
                else atomicAdd(&hist_d[bin], vd);
            }
            continue;
        }
        uint64_t pending = __ballot(bin >= 0);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int b = __shfl(bin, leader, 64);
            const bool mine = bin == b;
            pending &= ~__ballot(mine);
            if (exact) {
                const unsigned long long s = spec_wsum(mine ? vi : 0ull);
                if (lane == leader) atomicAdd(&sh_i[b], s);
            } else {
                const double s = spec_wsum(mine ? vd : 0.0);
                if (lane == leader) atomicAdd(&sh_d[b], s);
            }
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            if (sh_i[i]) atomicAdd(&hist_i[i], sh_i[i]);
            if (sh_d[i] != 0) atomicAdd(&hist_d[i], sh_d[i]);
        }
    }
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

static AttrCtx attr_ctx(dcf_tree *t)
{
    AttrCtx c;
    c.A = t->attr;
    c.p0 = t->pixdim[0];
    c.p1 = t->pixdim[1];
    c.p2 = t->pixdim[2];
    return c;
}

/* ---- the min rule under a failing root (tree_filter_min, filtering.c:55-100) ------------------------ *
 * The reference walks every pixel, in raster order, from its levelroot up to the first node already
 * reached by an earlier walk, never testing the root.  When the root fails (an increasing attribute:
 * then every node fails) the root is reached only by the walk of its first pixel c_root, which marks it
 * 0; a walk that arrives at the root before that takes the root's grey.  The first walk into the
 * subtree of a child a of the root is that of a's first pixel first(a), and every later walk into the
 * subtree stops inside it and copies its value: every node under a gets 0 when c_root < first(a), else
 * the root's grey; the root's own pixels get 0. */

/* pass flags for the jumps to the root's children: J(k) -> the child of the root above k (-1 at a root) */
__global__ __launch_bounds__(256) void k_rc_init(const int32_t *__restrict__ npar, int32_t *__restrict__ J, uint8_t *__restrict__ pass,
                                                 int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t p = npar[k];
    const bool rc = p >= 0 && npar[p] < 0;
    pass[k] = rc;
    J[k] = rc ? (int32_t)k : p;
}

/* first[x] = the first pixel (raster order) under each child of a root, and of each root's own level */
__global__ __launch_bounds__(256) void k_first_pixel(const uint32_t *__restrict__ p2n, const int32_t *__restrict__ npar,
                                                     const int32_t *__restrict__ J, uint32_t *__restrict__ first, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < n;
    int32_t a = -1;
    if (valid) {
        const uint32_t k = p2n[p];
        a = npar[k] < 0 ? (int32_t)k : J[k];
    }
    /* consecutive pixels mostly lie under one child of the root: one atomic per wave then (the
     * wave's smallest pixel is its first valid lane's) */
    const uint64_t vm = __ballot(valid);
    if (!vm) return;
    const int leader = __ffsll((unsigned long long)vm) - 1;
    const int32_t al = __shfl(a, leader, 64);
    if (__ballot(valid && a != al) == 0) {
        if ((int)(threadIdx.x & 63) == leader) atomicMin(&first[al], (uint32_t)p);
        return;
    }
    if (valid) atomicMin(&first[a], (uint32_t)p);
}

__global__ __launch_bounds__(256) void k_min_rootfail(AttrCtx c, int idx, double lambda, const double *__restrict__ nlam,
                                                      const int32_t *__restrict__ npar, const int32_t *__restrict__ J,
                                                      const uint32_t *__restrict__ first, const float *__restrict__ ngrey,
                                                      float *__restrict__ val, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t a = npar[k] < 0 ? -1 : J[k];
    const int32_t root = a < 0 ? (int32_t)k : npar[a];
    if (attr_value(c, idx, root) >= (nlam ? nlam[root] : lambda)) return; /* the root passes: the direct rule */
    val[k] = (a < 0 || first[root] < first[a]) ? 0.0f : ngrey[root];
}

/* ---- the max rule of a non-increasing attribute (tree_filter_max, filtering.c:156-208) ------------ *
 * The reference walks the pixels in raster order.  A pixel climbs from its node through failing,
 * unreached nodes; at the first passing node w it keeps g(w) and sets out(u) = g(u) on every ancestor u
 * of w (a "fix", stopping at the first ancestor already fixed); at an already reached node it copies
 * that node's out; at a failing root it writes 0.  Every node on the climb takes the same value.
 * Solved per node, with first(x) = the first pixel of x's subtree (the walk of that pixel is the first
 * to enter x, and it reaches x) and Q(x) = min first(w) over the passing nodes w in x's subtree (the
 * moment x is fixed; none when x keeps no passing node):
 *   - the walk of first(x) through a node x that is not fixed then (x fails, Q(x) > first(x)) writes
 *     V(x) = g(a) for the parent a of x when a passes or is already fixed (Q(a) < first(x)), and V(a)
 *     otherwise (the same walk continues, or a was reached by an earlier walk with the value V(a));
 *     a failing root gives 0.  V(x) is g(parent(c)) for the nearest ancestor-or-self c of x whose
 *     parent passes or has Q(parent) < first(c): pointer jumping as for the direct rule.
 *   - the levelroot pixel of x ends with g(x) when x is ever fixed, else V(x); any other pixel v of x
 *     reads out(x) at its own turn: g(x) when x was fixed before it (Q(x) < v), else V(x).
 * The reference's out buffer starts zeroed (fresh pages), so a fix also stops at a node whose grey is
 * 0.  Only a float max-tree with negative greys has such a node below its root; that case is
 * refused. */
#define DCF_NONE32 0xFFFFFFFFu

/* a node other than a root with grey 0 */
__global__ __launch_bounds__(256) void k_zero_grey(const int32_t *__restrict__ npar, const float *__restrict__ ngrey,
                                                   uint32_t *__restrict__ flag, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K && npar[k] >= 0 && ngrey[k] == 0.0f) __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* lr[k] = the levelroot pixel of node k (the pixel whose canonical parent lies in another node) */
__global__ __launch_bounds__(256) void k_levelroot_pixel(const int32_t *__restrict__ parent, const uint32_t *__restrict__ p2n,
                                                         uint32_t *__restrict__ lr, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t r = parent[p];
    const uint32_t k = p2n[p];
    if (r < 0 || p2n[r] != k) lr[k] = (uint32_t)p;
}

/* ---- subtree minima in log depth ---------------------------------------------------------------- *
 * The pixels are laid out in preorder: a node's own pixels, then its children's subtrees in order of
 * node id.  Subtree sizes are the accumulated areas, so a node's position is the sum, up its path, of
 * its parent's own-pixel count plus the areas of its earlier siblings (a sort by parent, a scan, then
 * pointer doubling).  Sorted by position, the nodes of every subtree are one contiguous run, and a
 * minimum over a subtree is a range minimum (sparse table, O(1) per node).  No step depends on the
 * depth of the tree. */
__global__ __launch_bounds__(256) void k_child_area(const int32_t *__restrict__ npar, const uint32_t *__restrict__ area,
                                                    uint32_t *__restrict__ csum, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K && npar[k] >= 0) atomicAdd(&csum[npar[k]], area[k]);
}

/* sort keys: parent + 1 (0 for the roots, siblings under one virtual root); with heavy flags, one more
 * bit puts each node's heavy child first among its siblings */
__global__ __launch_bounds__(256) void k_parent_keys(const int32_t *__restrict__ npar, const uint8_t *__restrict__ heavy,
                                                     uint32_t *__restrict__ key, uint32_t *__restrict__ id, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const uint32_t pk = (uint32_t)(npar[k] + 1);
    key[k] = heavy ? (pk << 1) | (heavy[k] ? 0u : 1u) : pk;
    id[k] = (uint32_t)k;
}

__global__ __launch_bounds__(256) void k_gather_area(const uint32_t *__restrict__ id, const uint32_t *__restrict__ area,
                                                     uint64_t *__restrict__ out, int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K) out[i] = area[id[i]];
}

/* base[key] = the scan value at the start of the key's run of siblings */
__global__ __launch_bounds__(256) void k_seg_base(const uint32_t *__restrict__ key, int shift, const uint64_t *__restrict__ E,
                                                  uint64_t *__restrict__ base, int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K && (i == 0 || (key[i - 1] >> shift) != (key[i] >> shift))) base[key[i] >> shift] = E[i];
}

/* off(x) = own pixels of x's parent + areas of x's earlier siblings; A = the parent, for the path sum */
__global__ __launch_bounds__(256) void k_offsets(const uint32_t *__restrict__ key, int shift, const uint32_t *__restrict__ id,
                                                 const uint64_t *__restrict__ E, const uint64_t *__restrict__ base,
                                                 const uint32_t *__restrict__ area, const uint32_t *__restrict__ csum,
                                                 const int32_t *__restrict__ npar, uint32_t *__restrict__ off, int32_t *__restrict__ A,
                                                 int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K) return;
    const uint32_t kk = key[i] >> shift, x = id[i];
    const uint32_t own_parent = kk == 0 ? 0u : area[kk - 1] - csum[kk - 1];
    off[x] = own_parent + (uint32_t)(E[i] - base[kk]);
    A[x] = npar[x];
}

/* the "another round is needed" flag, set once per round */
__device__ __forceinline__ void round_flag(uint32_t *active, int round, bool more)
{
    const unsigned long long b = __ballot(more);
    if ((threadIdx.x & (warpSize - 1)) == 0 && b &&
        __hip_atomic_load(&active[round], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        __hip_atomic_store(&active[round], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* path sums: S covers [k, A(k)) */
__global__ __launch_bounds__(256) void k_psum_round(const uint32_t *__restrict__ S, uint32_t *__restrict__ S2, const int32_t *__restrict__ A,
                                                    int32_t *__restrict__ A2, uint32_t *__restrict__ active, int round, int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t a2 = -1;
    if (k < K) {
        const int32_t a = A[k];
        uint32_t v = S[k];
        if (a >= 0) {
            v += S[a];
            a2 = A[a];
        }
        S2[k] = v;
        A2[k] = a2;
    }
    round_flag(active, round, a2 >= 0);
}

__global__ __launch_bounds__(256) void k_iota_u32(uint32_t *__restrict__ id, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < K) id[k] = (uint32_t)k;
}

/* rank of every node in preorder, the length of its subtree's run, and level 0 of the table (vals
 * gathered into preorder) */
__global__ __launch_bounds__(256) void k_preorder(const uint32_t *__restrict__ spos, const uint32_t *__restrict__ order,
                                                  const uint32_t *__restrict__ area, uint32_t *__restrict__ rank, uint32_t *__restrict__ cnt,
                                                  int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K) return;
    const uint32_t x = order[i];
    const uint64_t end = (uint64_t)spos[i] + area[x];
    /* first position at or past the subtree's end */
    uint64_t lo = (uint64_t)i + 1, hi = (uint64_t)K;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)spos[mid] < end) lo = mid + 1;
        else hi = mid;
    }
    rank[x] = (uint32_t)i;
    cnt[x] = (uint32_t)(lo - (uint64_t)i);
}

__global__ __launch_bounds__(256) void k_gather_pre(const uint32_t *__restrict__ order, const uint32_t *__restrict__ v,
                                                    const uint8_t *__restrict__ pass, uint32_t *__restrict__ T0, int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K) return;
    const uint32_t x = order[i];
    T0[i] = (pass == nullptr || pass[x]) ? v[x] : DCF_NONE32;
}

/* level j of the sparse table: the minimum over [i, i + 2^j) */
__global__ __launch_bounds__(256) void k_rmq_level(const uint32_t *__restrict__ prev, uint32_t *__restrict__ cur, int64_t half, int64_t m)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) cur[i] = min(prev[i], prev[i + half]);
}

__global__ __launch_bounds__(256) void k_rmq_query(const uint32_t *__restrict__ T, const uint32_t *__restrict__ rank,
                                                   const uint32_t *__restrict__ cnt, uint32_t *__restrict__ out, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const uint32_t l = rank[k], c = cnt[k];
    const int j = 31 - __clz(c);
    const uint32_t *Tj = T + (size_t)j * (size_t)K;
    out[k] = min(Tj[l], Tj[l + c - (1u << j)]);
}

/* runs a ping-pong doubling to convergence: launch(r, in, out) for rounds r = 0.. (a round returns at
 * once after the flags say the previous one changed nothing); returns which buffer (0 / 1) holds the
 * result */
template <typename F>
static int doubling(dcf_tree *t, F launch, int *which)
{
    dcf_ctx *ctx = t->ctx;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, t->counters, 0u, (int64_t)64);
    for (int r = 0; r < t->max_rounds; r++) launch(r, r & 1);
    uint32_t act[64];
    if (hipGetLastError() != hipSuccess) return dcf_set_error(ctx, DCF_EDEVICE, "doubling launches");
    if (hipMemcpyAsync(act, t->counters, sizeof act, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return dcf_set_error(ctx, DCF_EDEVICE, "doubling rounds");
    int ran = 0;
    for (int r = 0; r < t->max_rounds; r++) { ran = r + 1; if (act[r] == 0) break; }
    if (act[t->max_rounds - 1] != 0) return dcf_set_error(ctx, DCF_EDEVICE, "tree deeper than the doubling rounds");
    *which = ran & 1; /* round r reads buffer r & 1 and writes the other */
    return DCF_OK;
}

/* the nodes in preorder of the pixel layout: order[i] = the node at position rank i, spos[i] = its
 * position (strictly increasing).  Siblings follow in order of node id, the heavy child first when
 * heavy flags are given. */
static int node_preorder(dcf_tree *t, const uint8_t *heavy, uint32_t *spos, uint32_t *order)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K;
    const dim3 g(nblk(K, 256)), b(256);
    const int shift = heavy ? 1 : 0;
    uint32_t *csum = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *k0 = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), *k1 = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *v0 = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), *v1 = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint64_t *sa = (uint64_t *)dcf_alloc(ctx, (size_t)K * 8), *E = (uint64_t *)dcf_alloc(ctx, (size_t)K * 8);
    uint64_t *base = (uint64_t *)dcf_alloc(ctx, ((size_t)K + 1) * 8);
    uint32_t *S[2] = { (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), (uint32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    int32_t *A[2] = { (int32_t *)dcf_alloc(ctx, (size_t)K * 4), (int32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    void *tmp = nullptr;
    const uint32_t *area = t->attr.area;
    int rc = DCF_OK, sw = 0;
    if (!csum || !k0 || !k1 || !v0 || !v1 || !sa || !E || !base || !S[0] || !S[1] || !A[0] || !A[1])
        rc = dcf_set_error(ctx, DCF_ENOMEM, "preorder scratch");
    /* scratch for the larger of the radix sort and the scan */
    size_t need_sort = 0, need_scan = 0;
    if (rc == DCF_OK &&
        (hipcub::DeviceRadixSort::SortPairs(nullptr, need_sort, k0, k1, v0, v1, (int)K, 0, 32, ctx->stream) != hipSuccess ||
         hipcub::DeviceScan::ExclusiveSum(nullptr, need_scan, sa, E, (int)K, ctx->stream) != hipSuccess))
        rc = dcf_set_error(ctx, DCF_EDEVICE, "preorder scratch size");
    const size_t need = need_sort > need_scan ? need_sort : need_scan;
    if (rc == DCF_OK && !(tmp = dcf_alloc(ctx, need ? need : 16))) rc = dcf_set_error(ctx, DCF_ENOMEM, "preorder sort scratch");
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_fill_u32, g, b, 0, ctx->stream, csum, 0u, K);
        hipLaunchKernelGGL(k_child_area, g, b, 0, ctx->stream, (const int32_t *)t->npar, area, csum, K);
        /* siblings together, in order of node id (the sort is stable) */
        hipLaunchKernelGGL(k_parent_keys, g, b, 0, ctx->stream, (const int32_t *)t->npar, heavy, k0, v0, K);
        size_t ns = need;
        if (hipcub::DeviceRadixSort::SortPairs(tmp, ns, k0, k1, v0, v1, (int)K, 0, 32, ctx->stream) != hipSuccess)
            rc = dcf_set_error(ctx, DCF_EDEVICE, "sort by parent");
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_gather_area, g, b, 0, ctx->stream, (const uint32_t *)v1, area, sa, K);
        size_t ns = need;
        if (hipcub::DeviceScan::ExclusiveSum(tmp, ns, sa, E, (int)K, ctx->stream) != hipSuccess)
            rc = dcf_set_error(ctx, DCF_EDEVICE, "sibling scan");
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_seg_base, g, b, 0, ctx->stream, (const uint32_t *)k1, shift, (const uint64_t *)E, base, K);
        hipLaunchKernelGGL(k_offsets, g, b, 0, ctx->stream, (const uint32_t *)k1, shift, (const uint32_t *)v1, (const uint64_t *)E,
                           (const uint64_t *)base, area, (const uint32_t *)csum, (const int32_t *)t->npar, S[0], A[0], K);
        rc = doubling(t, [&](int r, int i) {
            hipLaunchKernelGGL(k_psum_round, g, b, 0, ctx->stream, (const uint32_t *)S[i], S[1 - i], (const int32_t *)A[i], A[1 - i],
                               t->counters, r, K);
        }, &sw);
    }
    if (rc == DCF_OK) {
        /* preorder: nodes sorted by position */
        hipLaunchKernelGGL(k_iota_u32, g, b, 0, ctx->stream, v0, K);
        size_t ns = need;
        if (hipcub::DeviceRadixSort::SortPairs(tmp, ns, S[sw], spos, v0, order, (int)K, 0, 32, ctx->stream) != hipSuccess)
            rc = dcf_set_error(ctx, DCF_EDEVICE, "sort by position");
    }
    dcf_release(ctx, csum); dcf_release(ctx, k0); dcf_release(ctx, k1); dcf_release(ctx, v0); dcf_release(ctx, v1);
    dcf_release(ctx, sa); dcf_release(ctx, E); dcf_release(ctx, base); dcf_release(ctx, tmp);
    for (int i = 0; i < 2; i++) { dcf_release(ctx, S[i]); dcf_release(ctx, A[i]); }
    return rc;
}

/* lr = the levelroot pixel of every node, first = the first pixel of its subtree, Q = the smallest
 * first of a passing node in its subtree (DCF_NONE32: none); pass flags given */
static int subtree_firsts(dcf_tree *t, const uint8_t *pass, uint32_t *lr, uint32_t *first, uint32_t *Q)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K;
    const dim3 g(nblk(K, 256)), b(256);
    int levels = 1;
    while (((int64_t)1 << levels) <= K) levels++; /* levels j = 0 .. levels - 1, 2^j <= K */
    uint32_t *spos = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), *order = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *rank = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), *cnt = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *T = (uint32_t *)dcf_alloc(ctx, (size_t)levels * (size_t)K * 4);
    int rc = DCF_OK;
    if (!spos || !order || !rank || !cnt || !T) rc = dcf_set_error(ctx, DCF_ENOMEM, "subtree scratch");
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_levelroot_pixel, dim3(nblk(t->n, 256)), b, 0, ctx->stream, (const int32_t *)t->parent, (const uint32_t *)t->p2n,
                           lr, t->n);
        rc = node_preorder(t, nullptr, spos, order);
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_preorder, g, b, 0, ctx->stream, (const uint32_t *)spos, (const uint32_t *)order, t->attr.area, rank, cnt, K);
        /* first: range minima of the levelroot pixels; Q: of first over the passing nodes */
        for (int pass_q = 0; pass_q < 2 && rc == DCF_OK; pass_q++) {
            hipLaunchKernelGGL(k_gather_pre, g, b, 0, ctx->stream, (const uint32_t *)order, pass_q ? (const uint32_t *)first : (const uint32_t *)lr,
                               pass_q ? pass : (const uint8_t *)nullptr, T, K);
            for (int j = 1; j < levels; j++) {
                const int64_t half = (int64_t)1 << (j - 1), m = K - ((int64_t)1 << j) + 1;
                hipLaunchKernelGGL(k_rmq_level, dim3(nblk(m, 256)), b, 0, ctx->stream, (const uint32_t *)(T + (size_t)(j - 1) * K),
                                   T + (size_t)j * K, half, m);
            }
            hipLaunchKernelGGL(k_rmq_query, g, b, 0, ctx->stream, (const uint32_t *)T, (const uint32_t *)rank, (const uint32_t *)cnt,
                               pass_q ? Q : first, K);
            if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "subtree minima launches");
        }
    }
    dcf_release(ctx, spos); dcf_release(ctx, order); dcf_release(ctx, rank); dcf_release(ctx, cnt); dcf_release(ctx, T);
    return rc;
}

/* ---- subtractive rule on float greys ------------------------------------------------------------ *
 * tree_filter_subtractive (filtering.c:211-261) evaluates, top-down, f(u) = g(u) + f(p) - g(p) in float
 * for a passing node u with parent p (f(p) for a failing one; the root: g or 0).  Float sums do not
 * reassociate, so every path is evaluated in order.  The tree is cut into heavy paths (each node
 * continues into its child of largest area), laid out so that each path is one run of the preorder:
 * one thread walks a path with f in a register, streaming the run's greys and flags (gathered into
 * preorder beforehand) 64 at a time.  A path
 * starts from its head's parent, which lies on a path with one light edge fewer; a light child has at
 * most half its parent's area, so at most log2(n) + 1 rounds of launches cover the tree. */
__global__ __launch_bounds__(256) void k_fill_u64(uint64_t *a, uint64_t v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

/* hv[p] = max over p's children of (area, -id): the heavy child, the smallest id among equal areas */
__global__ __launch_bounds__(256) void k_heavy_key(const int32_t *__restrict__ npar, const uint32_t *__restrict__ area,
                                                   uint64_t *__restrict__ hv, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K || npar[k] < 0) return;
    atomicMax((unsigned long long *)&hv[npar[k]], ((unsigned long long)area[k] << 32) | (0xFFFFFFFFull - (unsigned long long)k));
}

/* heavy / head flags; S, A start the path count of heads (the light depth) */
__global__ __launch_bounds__(256) void k_heavy_flag(const int32_t *__restrict__ npar, const uint64_t *__restrict__ hv,
                                                    uint8_t *__restrict__ heavy, uint8_t *__restrict__ head, uint32_t *__restrict__ S,
                                                    int32_t *__restrict__ A, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t p = npar[k];
    const bool h = p >= 0 && (uint32_t)hv[p] == 0xFFFFFFFFu - (uint32_t)k;
    heavy[k] = h;
    head[k] = !h;
    S[k] = h ? 0u : 1u;
    A[k] = p;
}

__global__ __launch_bounds__(256) void k_rank_of(const uint32_t *__restrict__ order, uint32_t *__restrict__ rank, int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K) rank[order[i]] = (uint32_t)i;
}

/* the walk's inputs in preorder: grey, and flags bit 0 pass, bit 1 head of a path */
__global__ __launch_bounds__(256) void k_chain_gather(const uint32_t *__restrict__ order, const float *__restrict__ ngrey,
                                                      const uint8_t *__restrict__ pass, const uint8_t *__restrict__ head,
                                                      float *__restrict__ Gr, uint8_t *__restrict__ Fr, int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K) return;
    const uint32_t x = order[i];
    Gr[i] = ngrey[x];
    Fr[i] = (uint8_t)((pass[x] ? 1 : 0) | (head[x] ? 2 : 0));
}

/* the heavy paths whose heads have light depth round + 1 (L counts the heads from the root down); a
 * path is the run of preorder ranks from its head to the next head, read 64 at a time; values go to
 * Vr by rank */
__global__ __launch_bounds__(256) void k_sub_chain(int round, const uint32_t *__restrict__ rank, const uint8_t *__restrict__ head,
                                                   const uint32_t *__restrict__ L, const int32_t *__restrict__ npar,
                                                   const float *__restrict__ ngrey, const uint8_t *__restrict__ pass,
                                                   const float *__restrict__ Gr, const uint8_t *__restrict__ Fr, float *__restrict__ Vr,
                                                   int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K || !head[k] || L[k] != (uint32_t)round + 1u) return;
    const int32_t p = npar[k];
    const int64_t r0 = rank[k];
    const float gk = ngrey[k];
    float f;
    if (p < 0) f = pass[k] ? gk : 0.0f;
    else {
        const float fp = Vr[rank[p]];
        if (pass[k]) {
            const float t = gk + fp;
            f = t - ngrey[p];
        } else f = fp;
    }
    Vr[r0] = f;
    float gprev = gk;
    constexpr int B = 64;
    for (int64_t r = r0 + 1; r < K; r += B) {
        float gv[B];
        uint8_t fv[B];
        const int m = (int)((K - r) < B ? (K - r) : B);
#pragma unroll
        for (int j = 0; j < B; j++) {
            gv[j] = j < m ? Gr[r + j] : 0.0f;
            fv[j] = j < m ? Fr[r + j] : (uint8_t)2;
        }
        bool end = false;
#pragma unroll
        for (int j = 0; j < B; j++) {
            if (end || (fv[j] & 2)) { end = true; continue; }
            if (fv[j] & 1) {
                const float t = gv[j] + f;
                f = t - gprev;
            }
            Vr[r + j] = f;
            gprev = gv[j];
        }
        if (end) break;
    }
}

__global__ __launch_bounds__(256) void k_chain_scatter(const uint32_t *__restrict__ order, const float *__restrict__ Vr,
                                                       float *__restrict__ val, int64_t K)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K) val[order[i]] = Vr[i];
}

static int sub_float(dcf_tree *t, const AttrCtx &c, int attr_idx, double lambda, float *val)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K;
    const dim3 g(nblk(K, 256)), b(256);
    uint8_t *pass = (uint8_t *)dcf_alloc(ctx, (size_t)K), *heavy = (uint8_t *)dcf_alloc(ctx, (size_t)K), *head = (uint8_t *)dcf_alloc(ctx, (size_t)K);
    uint64_t *hv = (uint64_t *)dcf_alloc(ctx, (size_t)K * 8);
    uint32_t *spos = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), *order = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *rank = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *S[2] = { (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), (uint32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    int32_t *A[2] = { (int32_t *)dcf_alloc(ctx, (size_t)K * 4), (int32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    int rc = DCF_OK, sw = 0;
    if (!pass || !heavy || !head || !hv || !spos || !order || !rank || !S[0] || !S[1] || !A[0] || !A[1])
        rc = dcf_set_error(ctx, DCF_ENOMEM, "subtractive scratch");
    if (rc == DCF_OK) {
        /* pass flags (k_direct_init's J goes to A[1], overwritten below) */
        hipLaunchKernelGGL(k_direct_init, g, b, 0, ctx->stream, c, attr_idx, lambda, t->nlam, t->npar, A[1], pass, K);
        hipLaunchKernelGGL(k_fill_u64, g, b, 0, ctx->stream, hv, (uint64_t)0, K);
        hipLaunchKernelGGL(k_heavy_key, g, b, 0, ctx->stream, (const int32_t *)t->npar, (const uint32_t *)t->attr.area, hv, K);
        hipLaunchKernelGGL(k_heavy_flag, g, b, 0, ctx->stream, (const int32_t *)t->npar, (const uint64_t *)hv, heavy, head, S[0], A[0], K);
        rc = doubling(t, [&](int r, int i) {
            hipLaunchKernelGGL(k_psum_round, g, b, 0, ctx->stream, (const uint32_t *)S[i], S[1 - i], (const int32_t *)A[i], A[1 - i],
                               t->counters, r, K);
        }, &sw);
    }
    if (rc == DCF_OK) rc = node_preorder(t, heavy, spos, order);
    if (rc == DCF_OK) {
        /* spos and the heavy flags are free again: Vr, Gr and Fr use their memory */
        float *Vr = (float *)spos, *Gr = (float *)hv;
        uint8_t *Fr = heavy;
        hipLaunchKernelGGL(k_rank_of, g, b, 0, ctx->stream, (const uint32_t *)order, rank, K);
        hipLaunchKernelGGL(k_chain_gather, g, b, 0, ctx->stream, (const uint32_t *)order, (const float *)t->ngrey, (const uint8_t *)pass,
                           (const uint8_t *)head, Gr, Fr, K);
        for (int r = 0; r < 32; r++)
            hipLaunchKernelGGL(k_sub_chain, g, b, 0, ctx->stream, r, (const uint32_t *)rank, (const uint8_t *)head, (const uint32_t *)S[sw],
                               (const int32_t *)t->npar, (const float *)t->ngrey, (const uint8_t *)pass, (const float *)Gr,
                               (const uint8_t *)Fr, Vr, K);
        hipLaunchKernelGGL(k_chain_scatter, g, b, 0, ctx->stream, (const uint32_t *)order, (const float *)Vr, val, K);
        if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "launch k_sub_chain");
    }
    dcf_release(ctx, pass); dcf_release(ctx, heavy); dcf_release(ctx, head); dcf_release(ctx, hv); dcf_release(ctx, spos);
    dcf_release(ctx, order); dcf_release(ctx, rank);
    for (int i = 0; i < 2; i++) { dcf_release(ctx, S[i]); dcf_release(ctx, A[i]); }
    return rc;
}

/* stop flags for V: J(k) -> the nearest ancestor-or-self c whose parent passes or is fixed before first(c) */
__global__ __launch_bounds__(256) void k_max_stop(const int32_t *__restrict__ npar, const uint8_t *__restrict__ pass,
                                                  const uint32_t *__restrict__ first, const uint32_t *__restrict__ Q,
                                                  int32_t *__restrict__ J, uint8_t *__restrict__ stop, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t p = npar[k];
    const bool s = p >= 0 && (pass[p] || Q[p] < first[k]);
    stop[k] = s;
    J[k] = s ? (int32_t)k : p;
}

__global__ __launch_bounds__(256) void k_max_value(const int32_t *__restrict__ J, const int32_t *__restrict__ npar,
                                                   const float *__restrict__ ngrey, float *__restrict__ val, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t j = J[k];
    val[k] = j < 0 ? 0.0f : ngrey[npar[j]];
}

/* the reference's out buffer starts zeroed, so a node of grey 0 below the root reads as already
 * filtered to its own grey: DCF_EUNSUPPORTED for such a tree (only a float max-tree with negative
 * greys has one) */
static int refuse_zero_grey(dcf_tree *t, const char *rule)
{
    dcf_ctx *ctx = t->ctx;
    if (t->dtype != DCF_F32 || t->tree_min) return DCF_OK;
    uint32_t zero = 0;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, t->counters, 0u, (int64_t)64);
    hipLaunchKernelGGL(k_zero_grey, dim3(nblk(t->K, 256)), dim3(256), 0, ctx->stream, (const int32_t *)t->npar, (const float *)t->ngrey,
                       t->counters, t->K);
    if (hipMemcpyAsync(&zero, t->counters, 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return dcf_set_error(ctx, DCF_EDEVICE, "zero-grey check");
    if (zero)
        return dcf_set_error(ctx, DCF_EUNSUPPORTED,
                             "%s rule of a non-increasing attribute on a float max-tree with a node of grey 0 below its root", rule);
    return DCF_OK;
}

/* node values V into val, plus Q and the levelroot pixels for the per-pixel choice (*kq, *klr: the
 * caller releases them) */
static int max_general(dcf_tree *t, const AttrCtx &c, int attr_idx, double lambda, float *val, uint32_t **kq, uint32_t **klr)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K;
    *kq = *klr = nullptr;
    uint8_t *pass = (uint8_t *)dcf_alloc(ctx, (size_t)K), *stop = (uint8_t *)dcf_alloc(ctx, (size_t)K);
    int32_t *J = (int32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *first = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), *Q = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *lr = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    int rc = DCF_OK;
    if (!pass || !stop || !J || !first || !Q || !lr) rc = dcf_set_error(ctx, DCF_ENOMEM, "max rule scratch");
    if (rc == DCF_OK) rc = refuse_zero_grey(t, "max");
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_direct_init, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, c, attr_idx, lambda, t->nlam, t->npar, J, pass, K);
        rc = subtree_firsts(t, pass, lr, first, Q);
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_max_stop, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, (const int32_t *)t->npar, (const uint8_t *)pass,
                           (const uint32_t *)first, (const uint32_t *)Q, J, stop, K);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, t->counters, 0u, (int64_t)64);
        for (int r = 0; r < t->max_rounds; r++)
            hipLaunchKernelGGL(k_direct_jump, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, J, (const uint8_t *)stop, t->counters, r, K);
        hipLaunchKernelGGL(k_max_value, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, (const int32_t *)J, (const int32_t *)t->npar,
                           (const float *)t->ngrey, val, K);
        if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "max rule launches");
    }
    dcf_release(ctx, pass); dcf_release(ctx, stop); dcf_release(ctx, J); dcf_release(ctx, first);
    if (rc == DCF_OK) {
        *kq = Q;
        *klr = lr;
    } else {
        dcf_release(ctx, Q);
        dcf_release(ctx, lr);
    }
    return rc;
}

/* ---- the min rule of any attribute (tree_filter_min, filtering.c:55-100) ------------------------- *
 * A walk starts at the pixel's levelroot, climbs to the first node already reached (the root when
 * none is), never testing the root, and writes the pixels and nodes up to the highest failing node it
 * passed: g(parent) of that node.  A walk that passed no failing node writes only its own levelroot,
 * with its grey.  A walk that stops at a reached node whose out differs from its grey copies that
 * out.  The result is the textbook min rule with the root left out (a node's value is g(parent(h))
 * for the highest failing node h below the root on its path, its own grey when there is none), except
 * for zeros spread by a failing root of nonzero grey.  Such a root is written 0 by its own levelroot
 * pixel, and a walk that later stops at it, or at a node already zeroed, writes 0.
 * Which nodes are zeroed, with t(x) = first(x) for a node with a failing node h at or above it (the
 * walk of first(x) is the first to enter x, and it writes x) and t(x) = its levelroot pixel otherwise
 * (for a node that gets zeroed earlier, any time from its first zeroing chance to its true write time
 * gives its descendants the same answers, and the levelroot is one), D(x) = the strict ancestor with
 * the smallest t (the lowest on ties) and tau(x) = t(D(x)):
 *   a newly reached node takes the status of the lowest reached node above it, so every node reached
 *   on x's path after tau(x) has D(x)'s status, and x is zeroed iff D(x) is and tau(x) < first(x) (h
 *   exists) or tau(x) < its levelroot (no h).  The root starts the chains: zeroed iff it fails with a
 *   nonzero grey.
 * Each step is a path aggregate solved by pointer doubling (ping-pong buffers, per-round flags that
 * end the rounds early): h by "the topmost flagged", (tau, D) by a path minimum with argmin, the zero
 * status by an AND along the D chain. */
__global__ __launch_bounds__(256) void k_hf_init(const int32_t *__restrict__ npar, const uint8_t *__restrict__ pass,
                                                 int32_t *__restrict__ H, int32_t *__restrict__ A, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t p = npar[k];
    H[k] = (p >= 0 && !pass[k]) ? (int32_t)k : -1;
    A[k] = p;
}

/* H covers [k, A(k)): the topmost failing node there; an ancestor's answer overrides */
__global__ __launch_bounds__(256) void k_hf_round(const int32_t *__restrict__ H, int32_t *__restrict__ H2, const int32_t *__restrict__ A,
                                                  int32_t *__restrict__ A2, uint32_t *__restrict__ active, int round, int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t a2 = -1;
    if (k < K) {
        const int32_t a = A[k];
        int32_t h = H[k];
        if (a >= 0) {
            const int32_t ha = H[a];
            if (ha >= 0) h = ha;
            a2 = A[a];
        }
        H2[k] = h;
        A2[k] = a2;
    }
    round_flag(active, round, a2 >= 0);
}

/* write times t, then the path minimum N over [k, A(k)) with its argmin Dn */
__global__ __launch_bounds__(256) void k_pm_init(const int32_t *__restrict__ npar, const int32_t *__restrict__ H,
                                                 const uint32_t *__restrict__ first, const uint32_t *__restrict__ lr,
                                                 uint32_t *__restrict__ N, int32_t *__restrict__ Dn, int32_t *__restrict__ A, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t p = npar[k];
    N[k] = (p >= 0 && H[k] >= 0) ? first[k] : lr[k];
    Dn[k] = (int32_t)k;
    A[k] = p;
}

__global__ __launch_bounds__(256) void k_pm_round(const uint32_t *__restrict__ N, uint32_t *__restrict__ N2, const int32_t *__restrict__ Dn,
                                                  int32_t *__restrict__ Dn2, const int32_t *__restrict__ A, int32_t *__restrict__ A2,
                                                  uint32_t *__restrict__ active, int round, int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t a2 = -1;
    if (k < K) {
        const int32_t a = A[k];
        uint32_t nv = N[k];
        int32_t d = Dn[k];
        if (a >= 0) {
            const uint32_t na = N[a];
            if (na < nv) { /* ties keep the lower node */
                nv = na;
                d = Dn[a];
            }
            a2 = A[a];
        }
        N2[k] = nv;
        Dn2[k] = d;
        A2[k] = a2;
    }
    round_flag(active, round, a2 >= 0);
}

/* zero status: C covers the D chain from k up to (not including) P(k); the root's is its own */
__global__ __launch_bounds__(256) void k_z_init(const int32_t *__restrict__ npar, const uint8_t *__restrict__ pass,
                                                const float *__restrict__ ngrey, const uint32_t *__restrict__ first,
                                                const uint32_t *__restrict__ lr, const int32_t *__restrict__ H,
                                                const uint32_t *__restrict__ N, const int32_t *__restrict__ Dn, uint8_t *__restrict__ C,
                                                int32_t *__restrict__ P, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int32_t p = npar[k];
    if (p < 0) {
        C[k] = !pass[k] && ngrey[k] != 0.0f;
        P[k] = -1;
        return;
    }
    const uint32_t tau = N[p];
    const bool c = tau < (H[k] >= 0 ? first[k] : lr[k]);
    C[k] = c;
    P[k] = c ? Dn[p] : -1;
}

__global__ __launch_bounds__(256) void k_z_round(const uint8_t *__restrict__ C, uint8_t *__restrict__ C2, const int32_t *__restrict__ P,
                                                 int32_t *__restrict__ P2, uint32_t *__restrict__ active, int round, int64_t K)
{
    if (round > 0 && active[round - 1] == 0) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t p2 = -1;
    if (k < K) {
        const int32_t pp = P[k];
        bool c = C[k] != 0;
        if (pp >= 0) {
            c = c && C[pp] != 0;
            p2 = c ? P[pp] : -1; /* a false link settles the chain */
        }
        C2[k] = c;
        P2[k] = p2;
    }
    round_flag(active, round, p2 >= 0);
}

__global__ __launch_bounds__(256) void k_min_value(const int32_t *__restrict__ npar, const uint8_t *__restrict__ pass,
                                                   const float *__restrict__ ngrey, const int32_t *__restrict__ H,
                                                   const uint8_t *__restrict__ Z, float *__restrict__ val, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    if (npar[k] < 0) val[k] = pass[k] ? ngrey[k] : 0.0f;
    else val[k] = Z[k] ? 0.0f : (H[k] >= 0 ? ngrey[npar[H[k]]] : ngrey[k]);
}

static int min_general(dcf_tree *t, const AttrCtx &c, int attr_idx, double lambda, float *val)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K;
    const dim3 g(nblk(K, 256)), b(256);
    uint8_t *pass = (uint8_t *)dcf_alloc(ctx, (size_t)K);
    uint8_t *C[2] = { (uint8_t *)dcf_alloc(ctx, (size_t)K), (uint8_t *)dcf_alloc(ctx, (size_t)K) };
    int32_t *H[2] = { (int32_t *)dcf_alloc(ctx, (size_t)K * 4), (int32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    int32_t *A[2] = { (int32_t *)dcf_alloc(ctx, (size_t)K * 4), (int32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    int32_t *D[2] = { (int32_t *)dcf_alloc(ctx, (size_t)K * 4), (int32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    uint32_t *N[2] = { (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), (uint32_t *)dcf_alloc(ctx, (size_t)K * 4) };
    uint32_t *first = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4), *Q = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    uint32_t *lr = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    int rc = DCF_OK, hw = 0, nw = 0, zw = 0;
    if (!pass || !C[0] || !C[1] || !H[0] || !H[1] || !A[0] || !A[1] || !D[0] || !D[1] || !N[0] || !N[1] || !first || !Q || !lr)
        rc = dcf_set_error(ctx, DCF_ENOMEM, "min rule scratch");
    if (rc == DCF_OK) rc = refuse_zero_grey(t, "min");
    if (rc == DCF_OK) {
        /* pass flags (k_direct_init's J goes to scratch that is overwritten below) */
        hipLaunchKernelGGL(k_direct_init, g, b, 0, ctx->stream, c, attr_idx, lambda, t->nlam, t->npar, A[0], pass, K);
        rc = subtree_firsts(t, pass, lr, first, Q);
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_hf_init, g, b, 0, ctx->stream, (const int32_t *)t->npar, (const uint8_t *)pass, H[0], A[0], K);
        rc = doubling(t, [&](int r, int i) {
            hipLaunchKernelGGL(k_hf_round, g, b, 0, ctx->stream, (const int32_t *)H[i], H[1 - i], (const int32_t *)A[i], A[1 - i],
                               t->counters, r, K);
        }, &hw);
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_pm_init, g, b, 0, ctx->stream, (const int32_t *)t->npar, (const int32_t *)H[hw], (const uint32_t *)first,
                           (const uint32_t *)lr, N[0], D[0], A[0], K);
        rc = doubling(t, [&](int r, int i) {
            hipLaunchKernelGGL(k_pm_round, g, b, 0, ctx->stream, (const uint32_t *)N[i], N[1 - i], (const int32_t *)D[i], D[1 - i],
                               (const int32_t *)A[i], A[1 - i], t->counters, r, K);
        }, &nw);
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_z_init, g, b, 0, ctx->stream, (const int32_t *)t->npar, (const uint8_t *)pass, (const float *)t->ngrey,
                           (const uint32_t *)first, (const uint32_t *)lr, (const int32_t *)H[hw], (const uint32_t *)N[nw],
                           (const int32_t *)D[nw], C[0], A[0], K);
        rc = doubling(t, [&](int r, int i) {
            hipLaunchKernelGGL(k_z_round, g, b, 0, ctx->stream, (const uint8_t *)C[i], C[1 - i], (const int32_t *)A[i], A[1 - i],
                               t->counters, r, K);
        }, &zw);
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_min_value, g, b, 0, ctx->stream, (const int32_t *)t->npar, (const uint8_t *)pass, (const float *)t->ngrey,
                           (const int32_t *)H[hw], (const uint8_t *)C[zw], val, K);
        if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "launch k_min_value");
    }
    dcf_release(ctx, pass);
    for (int i = 0; i < 2; i++) {
        dcf_release(ctx, C[i]); dcf_release(ctx, H[i]); dcf_release(ctx, A[i]); dcf_release(ctx, D[i]); dcf_release(ctx, N[i]);
    }
    dcf_release(ctx, first); dcf_release(ctx, Q); dcf_release(ctx, lr);
    return rc;
}

static int rule_values(dcf_tree *t, int attr_idx, int decision, double lambda, unsigned flags, float *val, uint32_t **kq,
                       uint32_t **klr)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K;
    AttrCtx c = attr_ctx(t);
    *kq = *klr = nullptr;
    /* For an increasing attribute (area / volume, idx 0) a node passes only if its parent passes, so
     * the min, max and subtractive rules (filtering.c:55-261) all give the direct rule's output, except
     * the min rule under a failing root (k_min_rootfail above); the subtractive kernels can still be
     * forced with DCF_ALL_RULES.  The max and min rules of a non-increasing attribute (or of any
     * attribute under DCF_ALL_RULES) follow the reference's raster scan (max_general, min_general
     * above). */
    /* float greys: the subtractive rule's float rounding differs from the direct rule's copies */
    const bool float_grey = t->dtype == DCF_F32;
    bool general_sub = decision == DCF_SUBTRACTIVE && (float_grey || !(attr_idx == 0 && !(flags & DCF_ALL_RULES)));
    if (decision == DCF_MAX && (attr_idx != 0 || (flags & DCF_ALL_RULES))) return max_general(t, c, attr_idx, lambda, val, kq, klr);
    if (decision == DCF_MIN && (attr_idx != 0 || (flags & DCF_ALL_RULES))) return min_general(t, c, attr_idx, lambda, val);
    bool min_max_inc = (decision == DCF_MIN || decision == DCF_MAX) && attr_idx == 0;
    if (decision == DCF_DIRECT || (decision == DCF_SUBTRACTIVE && !general_sub) || min_max_inc) {
        int32_t *J = (int32_t *)dcf_alloc(ctx, (size_t)K * 4);
        uint8_t *pass = (uint8_t *)dcf_alloc(ctx, (size_t)K);
        if (!J || !pass) { dcf_release(ctx, J); dcf_release(ctx, pass); return dcf_set_error(ctx, DCF_ENOMEM, "filter scratch"); }
        DCF_LAUNCH(ctx, "k_direct_init", k_direct_init, dim3(nblk(K, 256)), dim3(256), 0, c, attr_idx, lambda, t->nlam, t->npar, J, pass, K);
        DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(1), dim3(64), 0, t->counters, 0u, (int64_t)64);
        for (int r = 0; r < t->max_rounds; r++)
            DCF_LAUNCH(ctx, "k_direct_jump", k_direct_jump, dim3(nblk(K, 256)), dim3(256), 0, J, pass, t->counters, r, K);
        DCF_LAUNCH(ctx, "k_direct_value", k_direct_value, dim3(nblk(K, 256)), dim3(256), 0, J, t->ngrey, val, K);
        int rc = DCF_OK;
        if (decision == DCF_MIN) {
            /* J, pass reused for the jumps to the roots' children */
            uint32_t *first = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
            if (!first) rc = dcf_set_error(ctx, DCF_ENOMEM, "min rule scratch");
            else {
                hipLaunchKernelGGL(k_rc_init, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, (const int32_t *)t->npar, J, pass, K);
                hipLaunchKernelGGL(k_fill_u32, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, first, 0xFFFFFFFFu, K);
                hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, t->counters, 0u, (int64_t)64);
                for (int r = 0; r < t->max_rounds; r++)
                    hipLaunchKernelGGL(k_direct_jump, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, J, (const uint8_t *)pass, t->counters,
                                       r, K);
                hipLaunchKernelGGL(k_first_pixel, dim3(nblk(t->n, 256)), dim3(256), 0, ctx->stream, (const uint32_t *)t->p2n,
                                   (const int32_t *)t->npar, (const int32_t *)J, first, t->n);
                hipLaunchKernelGGL(k_min_rootfail, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, c, attr_idx, lambda,
                                   (const double *)t->nlam, (const int32_t *)t->npar, (const int32_t *)J, (const uint32_t *)first,
                                   (const float *)t->ngrey, val, K);
                if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "min rule launches");
                dcf_release(ctx, first);
            }
        }
        dcf_release(ctx, J);
        dcf_release(ctx, pass);
        return rc;
    }
    if (decision == DCF_SUBTRACTIVE && float_grey) return sub_float(t, c, attr_idx, lambda, val);
    if (decision == DCF_SUBTRACTIVE) {
        double *F0 = (double *)dcf_alloc(ctx, (size_t)K * 8), *F1 = (double *)dcf_alloc(ctx, (size_t)K * 8);
        int32_t *A0 = (int32_t *)dcf_alloc(ctx, (size_t)K * 4), *A1 = (int32_t *)dcf_alloc(ctx, (size_t)K * 4);
        if (!F0 || !F1 || !A0 || !A1) {
            dcf_release(ctx, F0); dcf_release(ctx, F1); dcf_release(ctx, A0); dcf_release(ctx, A1);
            return dcf_set_error(ctx, DCF_ENOMEM, "filter scratch");
        }
        DCF_LAUNCH(ctx, "k_sub_init", k_sub_init, dim3(nblk(K, 256)), dim3(256), 0, c, attr_idx, lambda, t->nlam, t->npar, t->ngrey, F0, A0, K);
        DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(1), dim3(64), 0, t->counters, 0u, (int64_t)64);
        double *fi = F0, *fo = F1;
        int32_t *ai = A0, *ao = A1;
        for (int r = 0; r < t->max_rounds; r++) {
            DCF_LAUNCH(ctx, "k_sub_round", k_sub_round, dim3(nblk(K, 256)), dim3(256), 0, fi, fo, ai, ao, t->counters, r, K);
            double *tf = fi; fi = fo; fo = tf;
            int32_t *ta = ai; ai = ao; ao = ta;
        }
        uint32_t act[64];
        DCF_HIP(ctx, hipMemcpyAsync(act, t->counters, sizeof act, hipMemcpyDeviceToHost, ctx->stream));
        DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int ran = 0;
        for (int r = 0; r < t->max_rounds; r++) { ran = r + 1; if (act[r] == 0) break; }
        double *res = (ran & 1) ? F1 : F0;
        DCF_LAUNCH(ctx, "k_d2f", k_d2f, dim3(nblk(K, 256)), dim3(256), 0, res, val, K);
        dcf_release(ctx, F0); dcf_release(ctx, F1); dcf_release(ctx, A0); dcf_release(ctx, A1);
        return DCF_OK;
    }
    return dcf_set_error(ctx, DCF_EINVAL, "decision %d", decision);
}

int dcf_ops_filter(dcf_tree *t, int attr_idx, int decision, double lambda, int extract, void *out_dev, dcf_dtype out_dtype,
                   unsigned flags)
{
    dcf_ctx *ctx = t->ctx;
    float *val = (float *)dcf_alloc(ctx, (size_t)t->K * 4);
    if (!val) return dcf_set_error(ctx, DCF_ENOMEM, "filter values");
    uint32_t *kq = nullptr, *klr = nullptr;
    int rc = rule_values(t, attr_idx, decision, lambda, flags, val, &kq, &klr);
    if (rc == DCF_OK) {
        ImgView iv = view_of(t);
        hipEvent_t ea = nullptr, eb = nullptr;
        if (ctx->profiling && (ctx->prof_only.empty() || strstr("k_filter_out", ctx->prof_only.c_str())))
            dcf_prof_begin(ctx, "k_filter_out", &ea, &eb);
        const int reinvert = (flags & DCF_KEEP_INVERTED) ? 0 : 1;
        hipLaunchKernelGGL(k_filter_out, dim3(nblk(t->n, 256)), dim3(256), 0, ctx->stream, iv, t->p2n, val, (const uint32_t *)kq,
                           (const uint32_t *)klr, (const float *)t->ngrey, extract, reinvert, out_dev, (int)out_dtype, t->n);
        hipError_t le = hipGetLastError();
        if (ea) dcf_prof_end(ctx, "k_filter_out", ea, eb);
        if (le != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "launch k_filter_out: %s", hipGetErrorString(le));
    }
    dcf_release(ctx, val);
    dcf_release(ctx, kq);
    dcf_release(ctx, klr);
    return rc;
}

int dcf_ops_batch_filter(dcf_tree *t, int attr_idx, int decision, const double *lam_img, int extract, const BatchOut *outs,
                         const uint32_t *gmax_img, unsigned flags)
{
    dcf_ctx *ctx = t->ctx;
    double *nlam = (double *)dcf_alloc(ctx, (size_t)t->K * 8);
    float *val = (float *)dcf_alloc(ctx, (size_t)t->K * 4);
    if (!nlam || !val) {
        dcf_release(ctx, nlam);
        dcf_release(ctx, val);
        return dcf_set_error(ctx, DCF_ENOMEM, "batch filter values");
    }
    int rc = DCF_OK;
    hipLaunchKernelGGL(k_node_lambda, dim3(nblk(t->n, 256)), dim3(256), 0, ctx->stream, t->bimg, t->nbatch, lam_img, t->p2n, nlam,
                       t->n);
    if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "launch k_node_lambda");
    uint32_t *kq = nullptr, *klr = nullptr;
    if (rc == DCF_OK) {
        t->nlam = nlam;
        rc = rule_values(t, attr_idx, decision, 0.0, flags, val, &kq, &klr);
        t->nlam = nullptr;
    }
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_batch_out, dim3(nblk(t->n, 256)), dim3(256), 0, ctx->stream, view_of(t), t->bimg, t->nbatch, outs, gmax_img,
                           t->p2n, val, (const uint32_t *)kq, (const uint32_t *)klr, (const float *)t->ngrey, extract, t->n);
        if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "launch k_batch_out");
    }
    dcf_release(ctx, nlam);
    dcf_release(ctx, val);
    dcf_release(ctx, kq);
    dcf_release(ctx, klr);
    return rc;
}

int dcf_ops_check(dcf_tree *t, int attr_idx, float *out_dev)
{
    dcf_ctx *ctx = t->ctx;
    double *av = (double *)dcf_alloc(ctx, (size_t)t->K * 8);
    if (!av) return dcf_set_error(ctx, DCF_ENOMEM, "check values");
    AttrCtx c = attr_ctx(t);
    DCF_LAUNCH(ctx, "k_attr_eval", k_attr_eval, dim3(nblk(t->K, 256)), dim3(256), 0, c, attr_idx, av, t->K);
    DCF_LAUNCH(ctx, "k_check_out", k_check_out, dim3(nblk(t->n, 256)), dim3(256), 0, t->p2n, av, out_dev, t->n);
    dcf_release(ctx, av);
    return DCF_OK;
}

int dcf_ops_spectrum(dcf_tree *t, int attr1, const float *l1, int n1, int attr2, const float *l2, int n2, int second_is_area,
                     double *out_host)
{
    dcf_ctx *ctx = t->ctx;
    const int nb = n1 * (l2 ? n2 : 1);
    bool unit = t->pixdim[0] == 1.0f && t->pixdim[1] == 1.0f && t->pixdim[2] == 1.0f;
    int exact = (t->dtype != DCF_F32) && unit;
    size_t bytes = (size_t)nb * 16 + (size_t)(n1 + (l2 ? n2 : 0)) * 4;
    char *buf = (char *)dcf_alloc(ctx, bytes);
    if (!buf) return dcf_set_error(ctx, DCF_ENOMEM, "spectrum bins");
    unsigned long long *hi = (unsigned long long *)buf;
    double *hd = (double *)(buf + (size_t)nb * 8);
    float *dl1 = (float *)(buf + (size_t)nb * 16);
    float *dl2 = l2 ? dl1 + n1 : nullptr;
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(nblk((int64_t)nb * 4, 256)), dim3(256), 0, (uint32_t *)buf, 0u, (int64_t)nb * 4);
    DCF_HIP(ctx, hipMemcpyAsync(dl1, l1, (size_t)n1 * 4, hipMemcpyHostToDevice, ctx->stream));
    if (l2) DCF_HIP(ctx, hipMemcpyAsync(dl2, l2, (size_t)n2 * 4, hipMemcpyHostToDevice, ctx->stream));
    AttrCtx c = attr_ctx(t);
    int grp = dcf_group_of(attr1);
    unsigned blocks = nblk(t->K, 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    DCF_LAUNCH(ctx, "k_spectrum", k_spectrum, dim3(blocks), dim3(256), 0, c, grp, attr1, dl1, n1, attr2, dl2, l2 ? n2 : 1,
               second_is_area, t->npar, t->ngrey, exact, hi, hd, t->K);
    std::vector<unsigned long long> vi(nb);
    std::vector<double> vd(nb);
    DCF_HIP(ctx, hipMemcpyAsync(vi.data(), hi, (size_t)nb * 8, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipMemcpyAsync(vd.data(), hd, (size_t)nb * 8, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < nb; i++) out_host[i] = exact ? (double)vi[i] : vd[i];
    dcf_release(ctx, buf);
    return DCF_OK;
}

/* ------------------------------------------------------------------------------------------------ */
/* the allocation-free 1D spectrum: its scratch is a file-scope device array                      */
/* ------------------------------------------------------------------------------------------------ */
/* Layout per call of n bins: u64 integer bins [n], double bins [n], double output [n], float
 * thresholds [n], 28 bytes per bin.  Every call on a device shares that device's array and leaves
 * its contents behind, so calls that overlap on one device race. */
#define NOALLOC_BYTES ((size_t)DCF_NOALLOC_MAX_BINS * 28)
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

/* the array's device address per device (the symbol's host name is no device address) */
static void *g_scratch_ptr[64];
static std::mutex g_scratch_mu;

int dcf_noalloc_init(dcf_ctx *ctx)
{
    if (ctx->device < 0 || ctx->device >= 64) return dcf_set_error(ctx, DCF_EDEVICE, "device %d: no scratch slot", ctx->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void *ptr = nullptr;
        /* the caller made ctx->device current */
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return dcf_set_error(ctx, DCF_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->noalloc_scratch = g_scratch_ptr[ctx->device];
    return DCF_OK;
}

__global__ __launch_bounds__(256) void k_spec_out(int exact, const unsigned long long *__restrict__ hi, const double *__restrict__ hd,
                                                  double *__restrict__ out, int nb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) out[i] = exact ? (double)hi[i] : hd[i];
}

/* dcf_ops_spectrum for one axis, with every device buffer in the scratch array: no allocation */
int dcf_ops_spectrum_noalloc(dcf_tree *t, int attr1, const float *l1, int n1, double *out_host)
{
    dcf_ctx *ctx = t->ctx;
    if (!ctx->noalloc_scratch) return dcf_set_error(ctx, DCF_EDEVICE, "no allocation-free scratch on device %d", ctx->device);
    if (n1 <= 0 || n1 > DCF_NOALLOC_MAX_BINS)
        return dcf_set_error(ctx, DCF_EINVAL, "%d bins: the allocation-free spectrum takes 1 to %d", n1, DCF_NOALLOC_MAX_BINS);
    const int nb = n1;
    const bool unit = t->pixdim[0] == 1.0f && t->pixdim[1] == 1.0f && t->pixdim[2] == 1.0f;
    const int exact = (t->dtype != DCF_F32) && unit;
    char *buf = (char *)ctx->noalloc_scratch;
    unsigned long long *hi = (unsigned long long *)buf;
    double *hd = (double *)(buf + (size_t)nb * 8);
    double *dout = (double *)(buf + (size_t)nb * 16);
    float *dl1 = (float *)(buf + (size_t)nb * 24);
    /* the integer and double bins start at zero: a fill kernel (no memset node if a caller captures) */
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(nblk((int64_t)nb * 4, 256)), dim3(256), 0, (uint32_t *)buf, 0u, (int64_t)nb * 4);
    DCF_HIP(ctx, hipMemcpyAsync(dl1, l1, (size_t)n1 * 4, hipMemcpyHostToDevice, ctx->stream));
    AttrCtx c = attr_ctx(t);
    const int grp = dcf_group_of(attr1);
    unsigned blocks = nblk(t->K, 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    DCF_LAUNCH(ctx, "k_spectrum", k_spectrum, dim3(blocks), dim3(256), 0, c, grp, attr1, (const float *)dl1, n1, 0,
               (const float *)nullptr, 1, 0, t->npar, t->ngrey, exact, hi, hd, t->K);
    DCF_LAUNCH(ctx, "k_spec_out", k_spec_out, dim3(nblk(nb, 256)), dim3(256), 0, exact, (const unsigned long long *)hi,
               (const double *)hd, dout, nb);
    DCF_HIP(ctx, hipMemcpyAsync(out_host, dout, (size_t)nb * 8, hipMemcpyDeviceToHost, ctx->stream));
    DCF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DCF_OK;
}

/* ------------------------------------------------------------------------------------------------ */
/* global extremes (find_extremas, preprocessing.c:7-17)                                            */
/* ------------------------------------------------------------------------------------------------ */
__device__ static void atomic_max_f(float *a, float v)
{
    unsigned int *ai = (unsigned int *)a;
    unsigned int old = *ai;
    while (__uint_as_float(old) < v) {
        unsigned int prev = atomicCAS(ai, old, __float_as_uint(v));
        if (prev == old) break;
        old = prev;
    }
}

__device__ static void atomic_min_f(float *a, float v)
{
    unsigned int *ai = (unsigned int *)a;
    unsigned int old = *ai;
    while (__uint_as_float(old) > v) {
        unsigned int prev = atomicCAS(ai, old, __float_as_uint(v));
        if (prev == old) break;
        old = prev;
    }
}

__global__ __launch_bounds__(256) void k_minmax_init(const void *img, int dtype, float *mm)
{
    float v = dtype == DCF_U8 ? (float)((const uint8_t *)img)[0]
                              : (dtype == DCF_U16 ? (float)((const uint16_t *)img)[0] : ((const float *)img)[0]);
    mm[0] = v;
    mm[1] = v;
}

__global__ __launch_bounds__(256) void k_minmax(const void *img, int dtype, float *mm, int64_t n)
{
    __shared__ float smx[256], smn[256];
    float mx = -INFINITY, mn = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v = dtype == DCF_U8 ? (float)((const uint8_t *)img)[i]
                                  : (dtype == DCF_U16 ? (float)((const uint16_t *)img)[i] : ((const float *)img)[i]);
        mx = fmaxf(mx, v);
        mn = fminf(mn, v);
    }
    smx[threadIdx.x] = mx;
    smn[threadIdx.x] = mn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + s]);
            smn[threadIdx.x] = fminf(smn[threadIdx.x], smn[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomic_max_f(&mm[0], smx[0]);
        atomic_min_f(&mm[1], smn[0]);
    }
}

int dcf_ops_minmax(dcf_tree *t)
{
    dcf_ctx *ctx = t->ctx;
    DCF_LAUNCH(ctx, "k_minmax", k_minmax_init, dim3(1), dim3(1), 0, (const void *)t->img, (int)t->dtype, t->gmax);
    unsigned blocks = nblk(t->n, 256);
    if (blocks > 2048) blocks = 2048;
    DCF_LAUNCH(ctx, "k_minmax", k_minmax, dim3(blocks), dim3(256), 0, (const void *)t->img, (int)t->dtype, t->gmax, t->n);
    return DCF_OK;
}

/* ------------------------------------------------------------------------------------------------ */
/* batched form: input staging                                                                      */
/* ------------------------------------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void k_batch_in(const void *__restrict__ src, int src_dtype, void *__restrict__ dst, int dst_dtype,
                                                  int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t v = src_dtype == DCF_U8 ? ((const uint8_t *)src)[i] : ((const uint16_t *)src)[i];
    if (dst_dtype == DCF_U8) ((uint8_t *)dst)[i] = (uint8_t)v;
    else ((uint16_t *)dst)[i] = (uint16_t)v;
}

__global__ __launch_bounds__(256) void k_batch_gmax(const void *__restrict__ cat, int dtype, const BatchImg *__restrict__ bimg, int nb,
                                                    uint32_t *__restrict__ gmax_img, int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < n;
    uint32_t v = 0;
    int i = -1;
    if (valid) {
        v = dtype == DCF_U8 ? ((const uint8_t *)cat)[p] : ((const uint16_t *)cat)[p];
        i = batch_locate(bimg, nb, (uint64_t)p);
    }
    /* one atomic per wave when the wave lies in one image (lane 0 is valid whenever any lane is) */
    const int i0 = __shfl(i, 0, 64);
    if (__ballot(valid && i != i0) == 0) {
        uint32_t m = v;
        for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0 && valid) atomicMax(&gmax_img[i0], m);
    } else if (valid) {
        atomicMax(&gmax_img[i], v);
    }
}

/* min-tree of every image: its keys become gmax_i - g (preprocessing.c:84-91) */
__global__ __launch_bounds__(256) void k_batch_invert(void *__restrict__ cat, int dtype, const BatchImg *__restrict__ bimg, int nb,
                                                      const uint32_t *__restrict__ gmax_img, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = batch_locate(bimg, nb, (uint64_t)p);
    if (!bimg[i].inv) return;
    const uint32_t g = gmax_img[i];
    if (dtype == DCF_U8) ((uint8_t *)cat)[p] = (uint8_t)(g - ((uint8_t *)cat)[p]);
    else ((uint16_t *)cat)[p] = (uint16_t)(g - ((uint16_t *)cat)[p]);
}

int dcf_batch_copy_in(dcf_ctx *ctx, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n)
{
    DCF_LAUNCH(ctx, "k_batch_in", k_batch_in, dim3(nblk(n, 256)), dim3(256), 0, src, src_dtype, dst, dst_dtype, n);
    return DCF_OK;
}

int dcf_batch_invert(dcf_ctx *ctx, void *cat, int dtype, const BatchImg *bimg, int nb, uint32_t *gmax_img, int64_t n)
{
    DCF_LAUNCH(ctx, "k_fill", k_fill_u32, dim3(nblk(nb, 256)), dim3(256), 0, gmax_img, 0u, (int64_t)nb);
    DCF_LAUNCH(ctx, "k_batch_gmax", k_batch_gmax, dim3(nblk(n, 256)), dim3(256), 0, (const void *)cat, dtype, bimg, nb, gmax_img, n);
    DCF_LAUNCH(ctx, "k_batch_invert", k_batch_invert, dim3(nblk(n, 256)), dim3(256), 0, cat, dtype, bimg, nb,
               (const uint32_t *)gmax_img, n);
    return DCF_OK;
}

/* ------------------------------------------------------------------------------------------------ */
/* tree export (write_tree, io.c:1431-1573)                                                         */
/* ------------------------------------------------------------------------------------------------ */
/* node -> its levelroot pixel (nodes are numbered in raster order of their levelroots) */
__global__ __launch_bounds__(256) void k_node_pixel(ImgView iv, const int32_t *__restrict__ parent, const uint32_t *__restrict__ p2n,
                                                    uint32_t *__restrict__ n2p, int64_t n)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t r = parent[p];
    if (r < 0 || px_grey_ops(iv, r) != px_grey_ops(iv, p)) n2p[p2n[p]] = (uint32_t)p;
}

/* the group's area function and printed columns of every node: group 1 attributes 1-5, group 2
 * attributes 6-8, group 3 attributes 9-11, group 4 the entries 3-12 of inertia_attribute_arr
 * (attributes.c:628-688; its 2D / 3D choice follows the image, not the node; entries 6-8, the
 * grey-weighted means, are filled from the group-3 records by the caller) */
__global__ __launch_bounds__(256) void k_csv_cols(AttrCtx c, int grp, int is3d, double *__restrict__ size,
                                                  double *__restrict__ cols, int ncol, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    size[k] = attr_area(c, grp, k);
    double *o = cols + k * ncol;
    if (grp == 1) { for (int j = 0; j < 5; j++) o[j] = attr_value(c, 1 + j, k); return; }
    if (grp == 2) { for (int j = 0; j < 3; j++) o[j] = attr_value(c, 6 + j, k); return; }
    if (grp == 3) { for (int j = 0; j < 3; j++) o[j] = attr_value(c, 9 + j, k); return; }
    if (grp != 4) return;
    const DevAttr &A = c.A;
    const float p0 = c.p0, p1 = c.p1, p2 = c.p2;
    const uint32_t area = A.area[k];
    const double in0 = (double)p0 * p1 * p2 * area;
    const double in1 = p0 * (double)A.s[0][k], in2 = p1 * (double)A.s[1][k], in3 = p2 * (double)A.s[2][k];
    const double in4 = (double)(p0 * p0) * (double)A.s[3][k];
    const double in5 = (double)(p1 * p1) * (double)A.s[4][k];
    const double in6 = (double)(p2 * p2) * (double)A.s[5][k];
    const double in7 = (double)(p0 * p1) * (double)A.s[6][k];
    const double in8 = (double)(p1 * p2) * (double)A.s[7][k];
    const double in9 = (double)(p0 * p2) * (double)A.s[8][k];
    o[0] = in1 / in0;
    o[1] = in2 / in0;
    o[2] = in3 / in0;
    const double corr = in0 / 12;
    if (!is3d) {
        const double m0 = in4 - in1 * in1 / in0 + p0 * corr, m1 = in5 - in2 * in2 / in0 + p1 * corr, m2 = in7 - in1 * in2 / in0;
        double ev[2];
        eig2(m0, m2, m1, ev);
        o[6] = ev[1] / ev[0];
        o[7] = -1;
        o[8] = M_PI * sqrt(4 * ev[1] / in0) * sqrt(4 * ev[0] / in0) / in0;
        o[9] = (m0 + m1) / (in0 * in0); /* pow(A, 2.0) is the correctly rounded square */
    } else {
        const double m6[6] = { in4 - in1 * in1 / in0 + p0 * corr, in5 - in2 * in2 / in0 + p1 * corr,
                               in6 - in3 * in3 / in0 + p2 * corr, in7 - in1 * in2 / in0,
                               in8 - in2 * in3 / in0, in9 - in1 * in3 / in0 };
        double ev[3];
        eig3(m6, ev);
        o[6] = ev[2] / ev[1];
        o[7] = ev[1] / ev[0];
        o[8] = M_PI * sqrt(20 * ev[2] / in0) * sqrt(20 * ev[1] / in0) * sqrt(20 * ev[0] / in0) / (6 * in0);
        o[9] = (m6[0] + m6[1] + m6[2]) / pow(in0, 5.0 / 3.0);
    }
}

/* grey-weighted means (attributes 9-11) of the group-3 records into columns 3-5 of group 4 */
__global__ __launch_bounds__(256) void k_csv_wmeans(AttrCtx c3, double *__restrict__ cols, int ncol, int64_t K)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    for (int j = 0; j < 3; j++) cols[k * ncol + 3 + j] = attr_value(c3, 9 + j, k);
}

int dcf_ops_tree_table(dcf_tree *t, int grp, std::vector<uint32_t> &n2p, std::vector<int32_t> &npar, std::vector<float> &grey,
                       std::vector<double> &size, std::vector<double> &cols, int &ncol)
{
    dcf_ctx *ctx = t->ctx;
    const int64_t K = t->K;
    ncol = grp == 1 ? 5 : (grp == 2 || grp == 3 ? 3 : (grp == 4 ? 10 : 0));
    uint32_t *d_n2p = (uint32_t *)dcf_alloc(ctx, (size_t)K * 4);
    double *d_size = (double *)dcf_alloc(ctx, (size_t)K * 8);
    double *d_cols = (double *)dcf_alloc(ctx, (size_t)K * (ncol ? ncol : 1) * 8);
    int rc = DCF_OK;
    if (!d_n2p || !d_size || !d_cols) rc = dcf_set_error(ctx, DCF_ENOMEM, "tree table");
    if (rc == DCF_OK) {
        hipLaunchKernelGGL(k_node_pixel, dim3(nblk(t->n, 256)), dim3(256), 0, ctx->stream, view_of(t), t->parent, t->p2n, d_n2p, t->n);
        hipLaunchKernelGGL(k_csv_cols, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, attr_ctx(t), grp, t->dims[2] > 1 ? 1 : 0, d_size,
                           d_cols, ncol, K);
        if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "launch tree table kernels");
    }
    if (rc == DCF_OK && grp == 4) {
        /* the weighted means need the grey sums of group 3: accumulate them beside the group-4 records */
        DevAttr g4 = t->attr;
        t->attr = DevAttr();
        rc = dcf_accumulate_group(t, 3);
        if (rc == DCF_OK) {
            hipLaunchKernelGGL(k_csv_wmeans, dim3(nblk(K, 256)), dim3(256), 0, ctx->stream, attr_ctx(t), d_cols, ncol, K);
            if (hipGetLastError() != hipSuccess) rc = dcf_set_error(ctx, DCF_EDEVICE, "launch k_csv_wmeans");
        }
        dcf_free_attr(ctx, t->attr);
        t->attr = g4;
        t->grp = 4;
    }
    if (rc == DCF_OK) {
        n2p.resize((size_t)K);
        npar.resize((size_t)K);
        grey.resize((size_t)K);
        size.resize((size_t)K);
        cols.resize((size_t)K * ncol);
        if (hipMemcpyAsync(n2p.data(), d_n2p, (size_t)K * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(npar.data(), t->npar, (size_t)K * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(grey.data(), t->ngrey, (size_t)K * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipMemcpyAsync(size.data(), d_size, (size_t)K * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            (ncol && hipMemcpyAsync(cols.data(), d_cols, (size_t)K * ncol * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = dcf_set_error(ctx, DCF_EDEVICE, "tree table download");
    }
    dcf_release(ctx, d_n2p);
    dcf_release(ctx, d_size);
    dcf_release(ctx, d_cols);
    return rc;
}
