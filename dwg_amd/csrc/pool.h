// Device memory for the library's buffers, from a per-device cache of the
// blocks it has released (hipMalloc and hipFree cost tens to hundreds of
// microseconds each, and hipFree waits for the whole device).
#pragma once
#include <cstddef>

namespace dwg {

// `bytes` of device memory on the current device (rounded up: 256 B granules
// below 1 MiB, 2 MiB granules above).  A released block of at most 1.25
// times the rounded size is reused, but only after a device-wide
// synchronisation that follows its release, so that no work issued before
// the release can still touch it.  Throws Error(DWG_ERR_HIP) when the device
// is out of memory even after the cache is handed back.
void* dev_malloc(size_t bytes);

// Releases a block of dev_malloc into the cache (nullptr: nothing).
void dev_free(void* p);

// Hands every cached block of every device back to HIP (after synchronising
// each device); returns the bytes freed.
size_t dev_cache_release();

}  // namespace dwg
