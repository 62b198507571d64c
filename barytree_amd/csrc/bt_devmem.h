// Device memory of plans and their helpers, with released blocks kept for
// reuse.  A one-call solve allocates and frees every buffer of its plan;
// hipFree unmaps the memory and costs milliseconds per call at N = 16M
// (several GB of buffers), so released blocks go to a per-device cache
// that later allocations of a similar size take back.
#pragma once
#include <cstddef>

namespace btm {

// A block of at least `bytes` on the current device: a cached block of a
// close size, else a new hipMalloc (after emptying the device's cache if the
// first attempt runs out of memory).  nullptr when none can be had.
void *alloc(size_t bytes);
// Returns a block from alloc() to the cache, or frees it when the cache
// would exceed its limit.  Every use of the block on the device must be
// complete (callers synchronise first, as hipFree does by itself).
void release(void *p);
// Frees the cached blocks of every device; returns the bytes freed.
size_t empty_cache();
// Limit on the bytes cached per device (default 32 GiB); a lower limit
// frees what lies beyond it.
void set_cache_limit(size_t bytes);
size_t cached_bytes();  // cached on the current device

}  // namespace btm
