#include "bt_devmem.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

namespace btm {
namespace {

struct Device {
    std::multimap<size_t, void *> free;  // size -> block
    size_t cached = 0;
};

// Never destroyed: plans released while the process exits (after static
// destructors have run) still find the cache.
struct State {
    std::mutex mu;
    std::map<int, Device> devices;
    std::unordered_map<void *, std::pair<size_t, int>> live;  // block -> (size, device)
    size_t limit = size_t(32) << 30;
};
State &S()
{
    static State *s = new State;
    return *s;
}

// Sizes in steps of 512 B below 1 MiB, of 2 MiB above: repeated solves of
// similar problems meet the same sizes.
size_t round_up(size_t b)
{
    const size_t step = b < (size_t(1) << 20) ? 512 : size_t(2) << 20;
    return (std::max<size_t>(b, 1) + step - 1) / step * step;
}

// Frees cached blocks of `d` (largest first) until at most `keep` bytes stay.
size_t trim(Device &d, size_t keep)
{
    size_t freed = 0;
    while (d.cached > keep && !d.free.empty()) {
        auto it = std::prev(d.free.end());
        (void)hipFree(it->second);
        d.cached -= it->first;
        freed += it->first;
        d.free.erase(it);
    }
    return freed;
}

}  // namespace

void *alloc(size_t bytes)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    const size_t sz = round_up(bytes);
    State &st = S();
    std::lock_guard<std::mutex> g(st.mu);
    Device &d = st.devices[dev];
    // a cached block no more than twice the size (and at most 64 MiB larger)
    auto it = d.free.lower_bound(sz);
    if (it != d.free.end() && it->first <= std::min(2 * sz, sz + (size_t(64) << 20))) {
        void *p = it->second;
        const size_t s = it->first;
        d.free.erase(it);
        d.cached -= s;
        st.live[p] = {s, dev};
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        trim(d, 0);
        p = nullptr;
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    st.live[p] = {sz, dev};
    return p;
}

void release(void *p)
{
    if (!p) return;
    State &st = S();
    std::lock_guard<std::mutex> g(st.mu);
    auto it = st.live.find(p);
    if (it == st.live.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t s = it->second.first;
    Device &d = st.devices[it->second.second];
    st.live.erase(it);
    if (d.cached + s > st.limit) {
        (void)hipFree(p);
        return;
    }
    d.free.emplace(s, p);
    d.cached += s;
}

size_t empty_cache()
{
    State &st = S();
    std::lock_guard<std::mutex> g(st.mu);
    size_t freed = 0;
    for (auto &kv : st.devices) freed += trim(kv.second, 0);
    return freed;
}

void set_cache_limit(size_t bytes)
{
    State &st = S();
    std::lock_guard<std::mutex> g(st.mu);
    st.limit = bytes;
    for (auto &kv : st.devices) trim(kv.second, st.limit);
}

size_t cached_bytes()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    State &st = S();
    std::lock_guard<std::mutex> g(st.mu);
    auto it = st.devices.find(dev);
    return it == st.devices.end() ? 0 : it->second.cached;
}

}  // namespace btm
