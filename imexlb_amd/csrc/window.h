// A device address range backed by equal chunks of physical memory, each its
// own allocation (HIP virtual memory management), so that the chunks a
// placement search did not keep can be given back one by one while the kept
// ones stay on the same physical pages.  Host code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace imx {

struct ChunkedWindow {
    char *base = nullptr;       // start of the reserved address range
    size_t reserved = 0;        // bytes reserved: chunks.size() * chunk
    size_t chunk = 0;           // bytes per chunk, a multiple of the allocation granularity
    std::vector<hipMemGenericAllocationHandle_t> chunks;  // null once given back
    std::vector<bool> mapped;
    int device = 0;

    bool empty() const { return base == nullptr; }

    // bytes of physical memory still held
    size_t held() const
    {
        size_t n = 0;
        for (bool m : mapped) n += m ? chunk : 0;
        return n;
    }
};

// The chunk size for a window of `bytes` searched in steps of `step` bytes:
// `step` itself, or a multiple of it that keeps the window within 512
// chunks; rounded up to the allocation granularity.
inline size_t window_chunk(size_t bytes, size_t step, size_t gran)
{
    size_t c = step;
    while (bytes / c > 512) c += step;
    return (c + gran - 1) / gran * gran;
}

inline void window_destroy(ChunkedWindow &w)
{
    if (w.empty()) return;
    for (size_t i = 0; i < w.chunks.size(); ++i) {
        if (w.mapped[i]) (void)hipMemUnmap(w.base + i * w.chunk, w.chunk);
        if (w.chunks[i]) (void)hipMemRelease(w.chunks[i]);
    }
    (void)hipMemAddressFree(w.base, w.reserved);
    w = ChunkedWindow{};
}

// Reserves and maps at least `bytes` on `device` in chunks of about `step`
// bytes (window_chunk).  On failure nothing stays allocated.
inline hipError_t window_create(ChunkedWindow &w, int device, size_t bytes, size_t step)
{
    w = ChunkedWindow{};
    hipMemAllocationProp prop = {};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = device;
    size_t gran = 0;
    hipError_t e = hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum);
    if (e != hipSuccess) return e;
    if (gran == 0) gran = size_t(2) << 20;
    const size_t chunk = window_chunk(bytes, step, gran);
    const size_t n = (bytes + chunk - 1) / chunk;
    void *base = nullptr;
    if ((e = hipMemAddressReserve(&base, n * chunk, 0, nullptr, 0)) != hipSuccess) return e;
    w.base = (char *)base;
    w.reserved = n * chunk;
    w.chunk = chunk;
    w.device = device;
    w.chunks.assign(n, nullptr);
    w.mapped.assign(n, false);
    hipMemAccessDesc acc = {};
    acc.location = prop.location;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    for (size_t i = 0; i < n; ++i) {
        if ((e = hipMemCreate(&w.chunks[i], chunk, &prop, 0)) != hipSuccess) break;
        if ((e = hipMemMap(w.base + i * chunk, chunk, 0, w.chunks[i], 0)) != hipSuccess) break;
        w.mapped[i] = true;
        if ((e = hipMemSetAccess(w.base + i * chunk, chunk, &acc, 1)) != hipSuccess) break;
    }
    if (e != hipSuccess) window_destroy(w);
    return e;
}

// Gives back every chunk that does not overlap [off, off + len).  The caller
// has synchronised the device: no kernel may still use those chunks.
inline void window_keep(ChunkedWindow &w, size_t off, size_t len)
{
    if (w.empty()) return;
    const size_t first = off / w.chunk, last = (off + len - 1) / w.chunk;
    for (size_t i = 0; i < w.chunks.size(); ++i) {
        if (i >= first && i <= last) continue;
        if (w.mapped[i]) (void)hipMemUnmap(w.base + i * w.chunk, w.chunk);
        if (w.chunks[i]) (void)hipMemRelease(w.chunks[i]);
        w.mapped[i] = false;
        w.chunks[i] = nullptr;
    }
}

}  // namespace imx
