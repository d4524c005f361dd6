// Stream-ordered device memory for the engine and the mode solver.
//
// Blocks come from the device's default memory pool with hipMallocAsync on the
// owner's stream and go back with hipFreeAsync on the same stream.  Unlike
// hipFree, a stream-ordered free does not wait for the whole device, so a
// simulation that ends does not stall the simulations still running on other
// streams.  The pool's release threshold is raised to its maximum once per
// process: freed blocks stay resident in HBM and later simulations reuse them
// instead of going back to the driver at every synchronisation.
#ifndef COURANT_DEVMEM_HPP
#define COURANT_DEVMEM_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cem {

inline void pool_keep_resident() {
    static const bool done = [] {
        int dev = 0;
        hipMemPool_t pool = nullptr;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            std::uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        return true;
    }();
    (void)done;
}

inline hipError_t dev_alloc(void **p, size_t bytes, hipStream_t st) {
    pool_keep_resident();
    return hipMallocAsync(p, bytes, st);
}

inline void dev_free(void *p, hipStream_t st) {
    if (p) (void)hipFreeAsync(p, st);
}

}  // namespace cem

#endif
