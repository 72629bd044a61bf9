// CPU emulation of the GPU build schedule (tile zips -> tile canonicalisation -> seam zips ->
// canonicalisation), using the same dcf_zip.h merge with std::atomic link words and several
// threads racing on every phase.  Test infrastructure: it checks on the CPU that any interleaving
// of the lock-free merges yields the canonical tree the oracle computes.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <random>
#include <thread>
#include <vector>

#include "../../disccofan_amd/csrc/dcf_zip.h"

struct AtomMem {
    std::atomic<uint64_t> *w;
    uint64_t load(uint32_t i) { return w[i].load(std::memory_order_relaxed); }
    void store(uint32_t i, uint64_t v) { w[i].store(v, std::memory_order_relaxed); }
    uint64_t cas(uint32_t i, uint64_t e, uint64_t d)
    {
        uint64_t exp = e;
        w[i].compare_exchange_strong(exp, d, std::memory_order_relaxed);
        return exp;
    }
};

template <class Mem>
static uint32_t find_levelroot(Mem &m, uint32_t x, uint32_t kx)
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

struct Off { int dx, dy, dz; };

static std::vector<Off> fwd_offsets(int conn, bool is3d)
{
    std::vector<Off> f;
    if (!is3d) {
        f.push_back({ 1, 0, 0 }); f.push_back({ 0, 1, 0 });
        if (conn == 8) { f.push_back({ 1, 1, 0 }); f.push_back({ -1, 1, 0 }); }
    } else if (conn == 6) {
        f.push_back({ 1, 0, 0 }); f.push_back({ 0, 1, 0 }); f.push_back({ 0, 0, 1 });
    } else {
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++)
                    if (z > 0 || (z == 0 && y > 0) || (z == 0 && y == 0 && x > 0)) f.push_back({ x, y, z });
    }
    return f;
}

template <class F>
static void run_threads(int nt, F f)
{
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(f, t);
    for (auto &t : th) t.join();
}

extern "C" int zip_emulate(const uint32_t *keys, int W, int H, int D, int conn, int TX, int TY, int TZ, int nthreads,
                           uint64_t seed, int32_t *parent_out)
{
    const int64_t n = (int64_t)W * H * D, WH = (int64_t)W * H;
    const bool is3d = D > 1;
    if (!is3d) TZ = 1;
    /* connectivity mapping of the library (io.c:64-79) */
    if (is3d && conn == 4) conn = 6;
    if (is3d && conn == 8) conn = 26;
    if (!is3d && conn == 26) conn = 8;
    if (!is3d && conn == 6) conn = 4;
    auto fwd = fwd_offsets(conn, is3d);
    std::vector<std::atomic<uint64_t>> link(n);
    std::mt19937_64 rng(seed);
    const int ntx = (W + TX - 1) / TX, nty = (H + TY - 1) / TY, ntz = (D + TZ - 1) / TZ;

    // ---- phase 1: every tile in its own local link array ----
    for (int tz = 0; tz < ntz; tz++)
        for (int ty = 0; ty < nty; ty++)
            for (int tx = 0; tx < ntx; tx++) {
                int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
                int ex = std::min(TX, W - x0), ey = std::min(TY, H - y0), ez = std::min(TZ, D - z0);
                int np = ex * ey * ez;
                std::vector<std::atomic<uint64_t>> sl(np);
                std::vector<uint32_t> sk(np), lr(np);
                auto gidx = [&](int l) {
                    int lx = l % ex, ly = (l / ex) % ey, lz = l / (ex * ey);
                    return (int64_t)(x0 + lx) + (int64_t)(y0 + ly) * W + (int64_t)(z0 + lz) * WH;
                };
                for (int l = 0; l < np; l++) { sk[l] = keys[gidx(l)]; sl[l] = dcf_make_link(0, DCF_NONE); }
                std::vector<std::pair<int, int>> edges;
                for (int l = 0; l < np; l++) {
                    int lx = l % ex, ly = (l / ex) % ey, lz = l / (ex * ey);
                    for (auto &o : fwd) {
                        int nx = lx + o.dx, ny = ly + o.dy, nz = lz + o.dz;
                        if (nx < 0 || nx >= ex || ny < 0 || ny >= ey || nz < 0 || nz >= ez) continue;
                        edges.push_back({ l, nx + ny * ex + nz * ex * ey });
                    }
                }
                std::shuffle(edges.begin(), edges.end(), rng);
                AtomMem m{ sl.data() };
                run_threads(nthreads, [&](int t) {
                    for (size_t e = t; e < edges.size(); e += nthreads)
                        dcf_zip(m, (uint32_t)edges[e].first, sk[edges[e].first], (uint32_t)edges[e].second, sk[edges[e].second]);
                });
                run_threads(nthreads, [&](int t) {
                    for (int l = t; l < np; l += nthreads) lr[l] = find_levelroot(m, (uint32_t)l, sk[l]);
                });
                for (int l = 0; l < np; l++) {
                    uint64_t out;
                    if ((int)lr[l] != l) out = dcf_make_link(sk[l], (uint32_t)gidx(lr[l]));
                    else {
                        uint64_t w = sl[l].load();
                        uint32_t z = dcf_link_idx(w);
                        out = z == DCF_NONE ? dcf_make_link(0, DCF_NONE) : dcf_make_link(dcf_link_key(w), (uint32_t)gidx(lr[z]));
                    }
                    link[gidx(l)] = out;
                }
            }

    // ---- phase 2: seam edges, all in flight ----
    std::vector<std::pair<int64_t, int64_t>> seams;
    for (int64_t p = 0; p < n; p++) {
        int x = (int)(p % W), y = (int)((p / W) % H), z = (int)(p / WH);
        for (auto &o : fwd) {
            int nx = x + o.dx, ny = y + o.dy, nz = z + o.dz;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) continue;
            if (nx / TX == x / TX && ny / TY == y / TY && nz / TZ == z / TZ) continue;
            seams.push_back({ p, (int64_t)nx + (int64_t)ny * W + (int64_t)nz * WH });
        }
    }
    std::shuffle(seams.begin(), seams.end(), rng);
    AtomMem gm{ link.data() };
    run_threads(nthreads, [&](int t) {
        for (size_t e = t; e < seams.size(); e += nthreads)
            dcf_zip(gm, (uint32_t)seams[e].first, keys[seams[e].first], (uint32_t)seams[e].second, keys[seams[e].second]);
    });

    // ---- phase 3: canonical parent ----
    run_threads(nthreads, [&](int t) {
        for (int64_t p = t; p < n; p += nthreads) {
            uint32_t r = find_levelroot(gm, (uint32_t)p, keys[p]);
            if (r != (uint32_t)p) { parent_out[p] = (int32_t)r; continue; }
            uint64_t w = gm.load((uint32_t)p);
            uint32_t z = dcf_link_idx(w);
            parent_out[p] = z == DCF_NONE ? -1 : (int32_t)find_levelroot(gm, z, dcf_link_key(w));
        }
    });
    return 0;
}
