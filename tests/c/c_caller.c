/*
 * A plain C caller of libdcf.so, compiled with gcc against include/dcf.h -- the way a C host
 * program (such as the reference's operator driver) binds the library.
 *
 *   c_caller tables              print the attribute / decision tables (no GPU needed)
 *   c_caller lvec <file>         read a lambda-vector file (no GPU needed)
 *   c_caller run <W> <H>         build + area filter + spectrum of a synthetic u8 image on GPU 0
 *                                (prints "nodes K", the filtered-image checksum and the spectrum sum)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dcf.h"

static int tables(void)
{
    for (int i = 0; i < DCF_NUM_ATTR; i++) printf("%d %s %s %d\n", i, dcf_attr_short_name(i), dcf_attr_name(i), dcf_attr_group(i));
    for (int d = 0; d < 4; d++) printf("decision %d %s\n", d, dcf_decision_name(d));
    return 0;
}

static int lvec(const char *path)
{
    float *lv = NULL;
    int n = 0;
    int rc = dcf_lvec_read(path, 1.0, &lv, &n);
    if (rc != DCF_OK) {
        printf("status %d\n", rc);
        return 1;
    }
    printf("n %d\n", n);
    for (int i = 0; i < n; i++) printf("%.9g\n", lv[i]);
    dcf_lvec_free(lv);
    return 0;
}

static int run(int W, int H)
{
    dcf_ctx *ctx = NULL;
    int rc = dcf_ctx_create(0, 0, &ctx);
    if (rc != DCF_OK) {
        printf("no device (status %d)\n", rc);
        return 2;
    }
    uint8_t *img = (uint8_t *)malloc((size_t)W * H);
    uint8_t *out = (uint8_t *)malloc((size_t)W * H);
    uint64_t s = 12345;
    for (int i = 0; i < W * H; i++) {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        img[i] = (uint8_t)(s >> 59); /* 0..31 */
    }
    const uint64_t dims[3] = { (uint64_t)W, (uint64_t)H, 1 };
    dcf_tree *t = NULL;
    rc = dcf_tree_build(ctx, img, DCF_U8, dims, 4, DCF_MAX_TREE, 0, NULL, 0, &t);
    if (rc == DCF_OK) rc = dcf_filter(t, 0, DCF_DIRECT, 25.0, 0, out, DCF_U8, 0);
    int64_t K = 0;
    double spec[8];
    const float lv[8] = { 0, 1, 4, 16, 64, 256, 1024, 4096 };
    if (rc == DCF_OK) rc = dcf_tree_num_nodes(t, &K);
    if (rc == DCF_OK) rc = dcf_spectrum1d(t, 0, lv, 8, spec);
    if (rc != DCF_OK) {
        printf("status %d: %s\n", rc, dcf_last_error(ctx));
        return 1;
    }
    uint64_t sum = 0, gsum = 0;
    double ssum = 0;
    for (int i = 0; i < W * H; i++) { sum += out[i]; gsum += img[i]; }
    for (int i = 0; i < 8; i++) ssum += spec[i];
    printf("nodes %lld\nfilter_sum %llu\nspectrum_sum %.0f\ngrey_sum %llu\n", (long long)K, (unsigned long long)sum, ssum,
           (unsigned long long)gsum);
    dcf_tree_free(t);
    dcf_ctx_destroy(ctx);
    free(img);
    free(out);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && !strcmp(argv[1], "tables")) return tables();
    if (argc >= 3 && !strcmp(argv[1], "lvec")) return lvec(argv[2]);
    if (argc >= 4 && !strcmp(argv[1], "run")) return run(atoi(argv[2]), atoi(argv[3]));
    fprintf(stderr, "usage: c_caller tables | lvec <file> | run <W> <H>\n");
    return 64;
}
