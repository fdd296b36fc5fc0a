// Plans cached per layout pair behind the one-call drop-in entry point
// g2g_transform_layouts (include/grid2grid_amd.h).
//
// The reference plans every call of transform() anew: the grid cover, the
// decomposition into pieces and the sort by rank (transform.cpp:126-142,
// utils.hpp:146-192). Here the plan of a layout pair -- interval table, piece
// table, scans -- is built on the first call and kept in HBM; later calls with
// the same layouts (same splits, owners, ranks and block addresses and strides,
// element type and operation) only launch the gather. The key is the exact
// bytes of both descriptors, so a hit never depends on a hash alone.
#include "grid2grid_amd.h"
#include "g2g_internal.hpp"

#include <algorithm>
#include <cctype>
#include <cstddef>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace g2g;

namespace {

struct cache_entry {
    uint64_t hash = 0;
    std::string key;
    std::shared_ptr<g2g_plan_s> plan;
    uint64_t last_use = 0;
};

struct plan_cache {
    std::mutex mu;
    std::vector<cache_entry> entries;
    size_t capacity = 64;
    uint64_t tick = 0;
    int64_t hits = 0, misses = 0;
};

plan_cache &cache() {
    static plan_cache *c = new plan_cache();  // never destroyed: plans may be in use at exit
    return *c;
}

// A plan leaving the cache may still be read by launches in flight on any
// stream: the device is drained before its tables are freed.
void retire_plan(g2g_plan_s *p) {
    (void)hipDeviceSynchronize();
    g2g_plan_destroy(p);
}

// g2g_block has no padding: its bytes are exactly its four fields
static_assert(sizeof(g2g_block) == 24 && offsetof(g2g_block, stride) == 8 && offsetof(g2g_block, data) == 16,
              "g2g_block layout");

// The key of a call as views of the caller's arrays, nothing copied: the
// device, element type and operation, then per layout its four counts, its
// splits, owners and blocks. Segments point into the view itself (prefix,
// counts), so a view is used where it is built.
struct key_view {
    unsigned char prefix[9];
    int32_t counts[2][4];
    const void *p[11];
    size_t n[11];
    int segs = 0;
    size_t total = 0;
    key_view(const key_view &) = delete;
    key_view &operator=(const key_view &) = delete;
    key_view(int dev, int dtype, char trans) {
        std::memcpy(prefix, &dev, 4);
        std::memcpy(prefix + 4, &dtype, 4);
        prefix[8] = (unsigned char)trans;
        add(prefix, sizeof prefix);
    }
    void add(const void *q, size_t m) {
        p[segs] = q;
        n[segs] = m;
        ++segs;
        total += m;
    }
    // false when the layout's arrays are not even well formed (g2g_plan_create
    // then reports what is wrong)
    bool add_layout(const g2g_layout *l, int which) {
        if (!l || l->n_rows_split < 1 || l->n_cols_split < 1 || !l->rows_split || !l->cols_split || !l->owners ||
            l->n_blocks < 0 || (l->n_blocks > 0 && !l->blocks))
            return false;
        int32_t *c = counts[which];
        c[0] = l->n_rows_split;
        c[1] = l->n_cols_split;
        c[2] = l->n_ranks;
        c[3] = l->n_blocks;
        add(c, sizeof counts[which]);
        add(l->rows_split, sizeof(int32_t) * (size_t)l->n_rows_split);
        add(l->cols_split, sizeof(int32_t) * (size_t)l->n_cols_split);
        add(l->owners, sizeof(int32_t) * (size_t)(l->n_rows_split - 1) * (size_t)(l->n_cols_split - 1));
        add(l->blocks, sizeof(g2g_block) * (size_t)l->n_blocks);
        return true;
    }
    // the key's bytes, for a new cache entry
    std::string bytes() const {
        std::string k;
        k.reserve(total);
        for (int s = 0; s < segs; ++s)
            if (n[s]) k.append((const char *)p[s], n[s]);
        return k;
    }
    bool equals(const std::string &k) const {
        if (k.size() != total) return false;
        size_t off = 0;
        for (int s = 0; s < segs; ++s) {
            if (n[s] && std::memcmp(k.data() + off, p[s], n[s]) != 0) return false;
            off += n[s];
        }
        return true;
    }
};

// Hash of the key, segment by segment from the caller's arrays: a shortcut for
// the search only, since a hit also compares the whole key. Large keys (one
// entry per block) hash near memory speed: three independent CRC32C lanes
// over 8-byte words where the host has SSE4.2 (one word per cycle each), four
// independent rotate-multiply lanes otherwise.
inline uint64_t mix64(uint64_t x) {
    x *= 0x9e3779b97f4a7c15ull;
    x ^= x >> 32;
    x *= 0xd6e8feb86659fd93ull;
    return x ^ (x >> 29);
}

inline uint64_t tail_word(const unsigned char *q, size_t m) {
    uint64_t t = 0;
    std::memcpy(&t, q, m);
    return t;
}

uint64_t seg_hash_mul(const unsigned char *q, size_t m, uint64_t seed) {
    uint64_t h[4] = {seed, seed ^ 0x13198a2e03707344ull, seed ^ 0xa4093822299f31d0ull, seed ^ 0x082efa98ec4e6c89ull};
    size_t i = 0;
    for (; i + 32 <= m; i += 32)
        for (int j = 0; j < 4; ++j) {
            uint64_t w;
            std::memcpy(&w, q + i + 8 * j, 8);
            const uint64_t x = h[j] ^ w;
            h[j] = (x << 29 | x >> 35) * 0x9e3779b97f4a7c15ull;
        }
    for (int j = 0; i < m; i += 8, j = (j + 1) & 3) h[j] = mix64(h[j] ^ tail_word(q + i, std::min<size_t>(8, m - i)));
    return mix64(h[0] ^ (h[1] << 1 | h[1] >> 63) ^ (h[2] << 2 | h[2] >> 62) ^ (h[3] << 3 | h[3] >> 61) ^ m);
}

#if defined(__x86_64__)
__attribute__((target("sse4.2"))) uint64_t seg_hash_crc(const unsigned char *q, size_t m, uint64_t seed) {
    unsigned long long a = seed & 0xffffffffu, b = seed >> 32, c = (seed * 0x9e3779b97f4a7c15ull) >> 32;
    size_t i = 0;
    for (; i + 24 <= m; i += 24) {
        uint64_t x, y, z;
        std::memcpy(&x, q + i, 8);
        std::memcpy(&y, q + i + 8, 8);
        std::memcpy(&z, q + i + 16, 8);
        a = __builtin_ia32_crc32di(a, x);
        b = __builtin_ia32_crc32di(b, y);
        c = __builtin_ia32_crc32di(c, z);
    }
    for (; i < m; i += 8) a = __builtin_ia32_crc32di(a, tail_word(q + i, std::min<size_t>(8, m - i)));
    return mix64(a ^ (b << 21) ^ (c << 42) ^ ((uint64_t)m << 32) ^ seed);
}
#endif

uint64_t key_hash(const key_view &v) {
#if defined(__x86_64__)
    static const bool crc = __builtin_cpu_supports("sse4.2");
#else
    static const bool crc = false;
#endif
    uint64_t h = 0x243f6a8885a308d3ull;
    for (int s = 0; s < v.segs; ++s) {
        const unsigned char *q = (const unsigned char *)v.p[s];
#if defined(__x86_64__)
        h = crc ? seg_hash_crc(q, v.n[s], h) : seg_hash_mul(q, v.n[s], h);
#else
        h = seg_hash_mul(q, v.n[s], h);
#endif
    }
    return h;
}

}  // namespace

extern "C" {

int g2g_transform_layouts(const g2g_layout *initial, const g2g_layout *final_, int dtype, char trans,
                          const void *alpha, const void *beta, void *stream) {
    if ((alpha == nullptr) != (beta == nullptr))
        return internal_fail(G2G_ERR_INVALID, "give both alpha and beta or neither");
    trans = (char)std::toupper((unsigned char)trans);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return internal_fail(G2G_ERR_HIP, "hipGetDevice failed");
    key_view key(dev, dtype, trans);
    if (!key.add_layout(initial, 0) || !key.add_layout(final_, 1)) {
        // malformed descriptors: no cache entry; the plan's checks name the problem
        g2g_plan p = nullptr;
        const int rc = g2g_plan_create(initial, final_, dtype, trans, stream, &p);
        if (rc != G2G_OK) return rc;
        g2g_plan_destroy(p);
        return internal_fail(G2G_ERR_INVALID, "malformed layout descriptor");
    }
    const uint64_t h = key_hash(key);
    plan_cache &c = cache();
    std::shared_ptr<g2g_plan_s> plan;
    {
        std::lock_guard<std::mutex> g(c.mu);
        for (cache_entry &e : c.entries)
            if (e.hash == h && key.equals(e.key)) {
                e.last_use = ++c.tick;
                plan = e.plan;
                ++c.hits;
                break;
            }
        if (!plan) ++c.misses;
    }
    if (!plan) {
        g2g_plan p = nullptr;
        const int rc = g2g_plan_create(initial, final_, dtype, trans, stream, &p);
        if (rc != G2G_OK) return rc;
        plan = std::shared_ptr<g2g_plan_s>(p, retire_plan);
        std::vector<std::shared_ptr<g2g_plan_s>> evicted;  // released outside the lock
        {
            std::lock_guard<std::mutex> g(c.mu);
            bool present = false;
            for (cache_entry &e : c.entries)
                if (e.hash == h && key.equals(e.key)) {  // another thread planned the same pair meanwhile
                    evicted.push_back(plan);
                    plan = e.plan;
                    e.last_use = ++c.tick;
                    present = true;
                    break;
                }
            if (!present) {
                cache_entry e;
                e.hash = h;
                e.key = key.bytes();
                e.plan = plan;
                e.last_use = ++c.tick;
                c.entries.push_back(std::move(e));
                while (c.entries.size() > c.capacity) {
                    auto lru = std::min_element(c.entries.begin(), c.entries.end(),
                                                [](const cache_entry &a, const cache_entry &b) {
                                                    return a.last_use < b.last_use;
                                                });
                    evicted.push_back(std::move(lru->plan));
                    c.entries.erase(lru);
                }
            }
        }
    }
    return g2g_transform(plan.get(), alpha, beta, stream);
}

int g2g_plan_cache_set_capacity(int64_t entries) {
    if (entries < 0) return internal_fail(G2G_ERR_INVALID, "capacity %lld < 0", (long long)entries);
    plan_cache &c = cache();
    std::vector<std::shared_ptr<g2g_plan_s>> evicted;
    {
        std::lock_guard<std::mutex> g(c.mu);
        c.capacity = (size_t)entries;
        std::sort(c.entries.begin(), c.entries.end(),
                  [](const cache_entry &a, const cache_entry &b) { return a.last_use > b.last_use; });
        while (c.entries.size() > c.capacity) {
            evicted.push_back(std::move(c.entries.back().plan));
            c.entries.pop_back();
        }
    }
    return G2G_OK;
}

int g2g_plan_cache_clear(void) {
    plan_cache &c = cache();
    std::vector<cache_entry> old;
    {
        std::lock_guard<std::mutex> g(c.mu);
        old.swap(c.entries);
        c.hits = c.misses = 0;
    }
    return G2G_OK;
}

int g2g_plan_cache_stats(int64_t *hits, int64_t *misses, int64_t *entries) {
    plan_cache &c = cache();
    std::lock_guard<std::mutex> g(c.mu);
    if (hits) *hits = c.hits;
    if (misses) *misses = c.misses;
    if (entries) *entries = (int64_t)c.entries.size();
    return G2G_OK;
}

}  // extern "C"
