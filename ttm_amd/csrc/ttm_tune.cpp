// Contraction-split cache of single products (see ttm_tune.hpp).
#include "ttm_tune.hpp"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>

namespace ttm_amd
{
namespace
{
std::mutex g_mu;
std::map<std::string, unsigned> g_cache;
bool g_loaded = false;
std::atomic<bool> g_any{false}; // lets untuned processes skip the lock

const char *cache_file() { return std::getenv("TTM_AMD_SPLIT_CACHE"); }

void load_locked()
{
    if (g_loaded) return;
    g_loaded = true;
    const char *path = cache_file();
    if (!path) return;
    std::ifstream in(path);
    std::string key;
    unsigned s = 0;
    while (in >> key >> s)
        if (s >= 1 && s <= 65535) g_cache[key] = s;
    if (!g_cache.empty()) g_any.store(true);
}

void store_locked()
{
    const char *path = cache_file();
    if (!path) return;
    std::ofstream out(path, std::ios::trunc);
    for (auto const &kv : g_cache) out << kv.first << ' ' << kv.second << '\n';
}

int log2_class(uint64_t x)
{
    int k = 0;
    while (x > 1) {
        x >>= 1;
        ++k;
    }
    return k;
}
} // namespace

std::string split_class_key(Kernel k, const Fold &f, bool is_double)
{
    char buf[128];
    std::snprintf(buf, sizeof buf, "%s:%s:%s:%s:m%d:nq%d:cols%d", kernel_name(k), is_double ? "f64" : "f32",
                  f.w == 1 ? "tn" : "nn", f.b_colmajor ? "bcol" : "brow", log2_class(f.m), log2_class(f.nq),
                  log2_class(f.outer * f.w));
    return buf;
}

unsigned tuned_splits(Kernel k, const Fold &f, bool is_double)
{
    static std::once_flag once;
    std::call_once(once, [] {
        std::lock_guard<std::mutex> lock(g_mu);
        load_locked();
    });
    if (!g_any.load(std::memory_order_relaxed)) return 0;
    std::string const key = split_class_key(k, f, is_double);
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_cache.find(key);
    return it == g_cache.end() ? 0 : it->second;
}

void store_tuned_splits(Kernel k, const Fold &f, bool is_double, unsigned splits)
{
    std::lock_guard<std::mutex> lock(g_mu);
    load_locked();
    g_cache[split_class_key(k, f, is_double)] = splits < 1 ? 1 : splits;
    g_any.store(true);
    store_locked();
}

void clear_tuned_splits()
{
    std::lock_guard<std::mutex> lock(g_mu);
    g_loaded = true;
    g_cache.clear();
    g_any.store(false);
}

} // namespace ttm_amd
