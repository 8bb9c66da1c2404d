// Caching pool for the large device work buffers of the Krylov solvers.
//
// The continuation gives every step a fresh solver copy (DFGMRES.m:204-221),
// and each copy needs the same Arnoldi bases, deflation basis and its
// products -- gigabytes at a million unknowns.  Allocating and freeing them
// per copy costs device synchronisations and page mapping every step.  Freed
// buffers are kept here instead, per device, and handed to the next request
// of a similar size (at least the request, at most twice it).  Sizes are
// rounded to 2 MiB so that the recurring request sizes match.  When
// hipMalloc fails, the cached buffers are released and the request retried.
//
// A buffer is returned to the pool only after the device has finished with
// it: ssr_pool_free synchronises the device before caching (hipFree itself
// waits for the device, so this costs no more than the free it replaces).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <unordered_map>

#include "ssr_common.h"

namespace {

constexpr size_t kRound = (size_t)2 << 20;
constexpr size_t kMaxCached = (size_t)96 << 30;  // bytes kept per process at most

struct Pool {
    std::mutex mu;
    std::multimap<std::pair<int, size_t>, void *> free_list;  // (device, bytes) -> block
    std::unordered_map<void *, std::pair<int, size_t>> live;  // block -> (device, bytes)
    size_t cached = 0;

    void release_all_locked()
    {
        for (auto &kv : free_list) (void)hipFree(kv.second);
        free_list.clear();
        cached = 0;
    }
};

Pool &pool()
{
    static Pool *p = new Pool();  // never destroyed: buffers may be freed at exit
    return *p;
}

}  // namespace

int ssr_pool_malloc(void **out, size_t bytes)
{
    int dev = 0;
    SSR_HIP_CHECK(hipGetDevice(&dev));
    const size_t want = ((bytes ? bytes : 1) + kRound - 1) / kRound * kRound;
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.free_list.lower_bound({dev, want});
    if (it != P.free_list.end() && it->first.first == dev && it->first.second <= 2 * want) {
        void *p = it->second;
        const size_t sz = it->first.second;
        P.free_list.erase(it);
        P.cached -= sz;
        P.live[p] = {dev, sz};
        *out = p;
        return SSR_OK;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        P.release_all_locked();
        e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) {
        ssr_set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        return SSR_ERR_ALLOC;
    }
    P.live[p] = {dev, want};
    *out = p;
    return SSR_OK;
}

void ssr_pool_free(void *p)
{
    if (!p) return;
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) {
        (void)hipFree(p);  // not from the pool
        return;
    }
    const auto key = it->second;
    P.live.erase(it);
    if (P.cached + key.second > kMaxCached) {
        (void)hipFree(p);
        return;
    }
    (void)hipDeviceSynchronize();  // no queued work still reads or writes the block
    P.free_list.insert({key, p});
    P.cached += key.second;
}
