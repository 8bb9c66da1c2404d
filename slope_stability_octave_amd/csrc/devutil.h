// Device-side helpers of the setup code (setup.hip, amg_device.hip): an
// owning device buffer, radix sorts and prefix sums (hipcub), scalar reads.
// Setup only: the buffers come from hipMalloc and the calls synchronise where
// a size is read back.
#pragma once

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "ssr_common.h"

namespace ssr_dev {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
    int alloc(size_t bytes)
    {
        (void)hipFree(p);
        p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            ssr_set_error("hipMalloc(%zu) failed in the device setup", bytes);
            return SSR_ERR_ALLOC;
        }
        return SSR_OK;
    }
    template <class T>
    T *as() const
    {
        return (T *)p;
    }
    void *release()
    {
        void *q = p;
        p = nullptr;
        return q;
    }
};

inline int bits_for(uint64_t maxval)
{
    int b = 1;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

template <class T>
int exclusive_sum(const T *in, T *out, int64_t n, hipStream_t s)
{
    size_t bytes = 0;
    SSR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
    DevBuf tmp;
    int rc = tmp.alloc(bytes);
    if (rc) return rc;
    SSR_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, s));
    return SSR_OK;
}

template <class K, class V>
int sort_pairs(const K *kin, K *kout, const V *vin, V *vout, int64_t n, int end_bit, hipStream_t s)
{
    size_t bytes = 0;
    SSR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, s));
    DevBuf tmp;
    int rc = tmp.alloc(bytes);
    if (rc) return rc;
    SSR_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, s));
    return SSR_OK;
}

template <class K>
int sort_keys(const K *kin, K *kout, int64_t n, int end_bit, hipStream_t s)
{
    size_t bytes = 0;
    SSR_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, kin, kout, (int)n, 0, end_bit, s));
    DevBuf tmp;
    int rc = tmp.alloc(bytes);
    if (rc) return rc;
    SSR_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, kin, kout, (int)n, 0, end_bit, s));
    return SSR_OK;
}

template <class T>
int read_scalar(const T *dev, T *host, hipStream_t s)
{
    SSR_HIP_CHECK(hipMemcpyAsync(host, dev, sizeof(T), hipMemcpyDeviceToHost, s));
    SSR_HIP_CHECK(hipStreamSynchronize(s));
    return SSR_OK;
}

// counts (n + 1 entries, the last zero) -> exclusive prefix sums in `ptr`;
// returns the total
inline int counts_to_ptr(DevBuf &cnt, int64_t n, DevBuf &ptr, int32_t *total, hipStream_t s)
{
    int rc;
    if ((rc = ptr.alloc(sizeof(int32_t) * (n + 1))) || (rc = exclusive_sum(cnt.as<int32_t>(), ptr.as<int32_t>(), n + 1, s)) ||
        (rc = read_scalar(ptr.as<int32_t>() + n, total, s)))
        return rc;
    return SSR_OK;
}

}  // namespace ssr_dev

// 1-D launch of `kernel` over n items on stream `s` (in scope at the call)
#define SSR_LAUNCH(kernel, n, ...)                                                             \
    do {                                                                                       \
        if ((n) > 0) kernel<<<ssr_blocks((n), 256), 256, 0, s>>>(__VA_ARGS__);                 \
        SSR_HIP_CHECK(hipGetLastError());                                                      \
    } while (0)
