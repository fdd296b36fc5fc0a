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

void put(std::string &k, const void *p, size_t n) { k.append((const char *)p, n); }

// g2g_block has no padding: its bytes are exactly its four fields
static_assert(sizeof(g2g_block) == 24 && offsetof(g2g_block, stride) == 8 && offsetof(g2g_block, data) == 16,
              "g2g_block layout");

// The descriptor's bytes, or false when its arrays are not even well formed
// (g2g_plan_create then reports what is wrong).
bool layout_key(const g2g_layout *l, std::string &k) {
    if (!l || l->n_rows_split < 1 || l->n_cols_split < 1 || !l->rows_split || !l->cols_split || !l->owners ||
        l->n_blocks < 0 || (l->n_blocks > 0 && !l->blocks))
        return false;
    const size_t cells = (size_t)(l->n_rows_split - 1) * (size_t)(l->n_cols_split - 1);
    k.reserve(k.size() + 16 + sizeof(int32_t) * ((size_t)l->n_rows_split + (size_t)l->n_cols_split + cells) +
              sizeof(g2g_block) * (size_t)l->n_blocks);
    put(k, &l->n_rows_split, sizeof l->n_rows_split);
    put(k, &l->n_cols_split, sizeof l->n_cols_split);
    put(k, &l->n_ranks, sizeof l->n_ranks);
    put(k, &l->n_blocks, sizeof l->n_blocks);
    put(k, l->rows_split, sizeof(int32_t) * (size_t)l->n_rows_split);
    put(k, l->cols_split, sizeof(int32_t) * (size_t)l->n_cols_split);
    put(k, l->owners, sizeof(int32_t) * cells);
    put(k, l->blocks, sizeof(g2g_block) * (size_t)l->n_blocks);
    return true;
}

// Hash of the key: four independent rotate-multiply lanes over 8-byte words
// (a large key -- one entry per block -- hashes near memory speed, not at one
// byte per multiply), mixed with the length at the end. Only a shortcut for the
// search: a hit also compares the whole key.
inline uint64_t mix64(uint64_t x) {
    x *= 0x9e3779b97f4a7c15ull;
    x ^= x >> 32;
    x *= 0xd6e8feb86659fd93ull;
    return x ^ (x >> 29);
}

uint64_t key_hash(const std::string &k) {
    const unsigned char *p = (const unsigned char *)k.data();
    const size_t n = k.size();
    uint64_t h[4] = {0x243f6a8885a308d3ull, 0x13198a2e03707344ull, 0xa4093822299f31d0ull, 0x082efa98ec4e6c89ull};
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
        for (int j = 0; j < 4; ++j) {
            uint64_t w;
            std::memcpy(&w, p + i + 8 * j, 8);
            const uint64_t x = h[j] ^ w;
            h[j] = (x << 29 | x >> 35) * 0x9e3779b97f4a7c15ull;
        }
    for (int j = 0; i < n; i += 8, ++j) {
        uint64_t w = 0;
        std::memcpy(&w, p + i, std::min<size_t>(8, n - i));
        h[j] = mix64(h[j] ^ w);
    }
    return mix64(h[0] ^ (h[1] << 1 | h[1] >> 63) ^ (h[2] << 2 | h[2] >> 62) ^ (h[3] << 3 | h[3] >> 61) ^ n);
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
    std::string key;
    put(key, &dev, sizeof dev);
    put(key, &dtype, sizeof dtype);
    put(key, &trans, sizeof trans);
    if (!layout_key(initial, key) || !layout_key(final_, key)) {
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
            if (e.hash == h && e.key == key) {
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
                if (e.hash == h && e.key == key) {  // another thread planned the same pair meanwhile
                    evicted.push_back(plan);
                    plan = e.plan;
                    e.last_use = ++c.tick;
                    present = true;
                    break;
                }
            if (!present) {
                cache_entry e;
                e.hash = h;
                e.key = std::move(key);
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
