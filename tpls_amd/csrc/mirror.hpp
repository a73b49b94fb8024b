// Device mirrors of host arrays for the per-routine entry points.
//
// A host array attached to the mirror (tpls_mirror_attach) gets a device copy
// of the same bytes that the per-routine entry points use in place of a
// fresh upload.  Coherence is kept with page protection, per granule of the
// array's whole pages:
//   CLEAN  host and device agree; the pages are read-only, so the first host
//          write faults and turns the granule DIRTY;
//   DIRTY  the host wrote since the last upload; the pages are writable and
//          the next entry point that reads the array uploads the granule;
//   STALE  a kernel wrote the device copy; the pages are inaccessible, and
//          the first host access faults and downloads the granule.
// So data crosses the host link only where the host reads what the device
// wrote, or the device reads what the host wrote.  The partial pages at
// either end of an array (shared with neighbouring allocations) are never
// protected: they are uploaded before, and downloaded after, every device use.
//
// The host side is plain C++ (no HIP types): the device allocation and the
// copies come from a Backend, which the HIP library installs at load time.
#pragma once

#include <cstddef>

namespace tpls {
namespace mirror {

struct Backend {
    void* (*alloc)(size_t bytes);                                // nullptr on failure
    void (*release)(void* dev);
    int (*upload)(void* dev, const void* host, size_t bytes);    // synchronous; 0 = ok
    int (*download)(void* host, const void* dev, size_t bytes);  // synchronous; 0 = ok
    int (*fail)(int code, const char* msg);                      // records an error, returns code
};

// installed once, before the first attach
void set_backend(const Backend& b);

// true when any array is attached (a cheap check for the entry points)
bool any();

// The device copy of the attached array that starts at host and is exactly
// bytes long; nullptr when no attached array matches.
void* device_of(const void* host, size_t bytes);

// Make the device copy of that attached array current (uploads the DIRTY
// granules and the end pages).  0 or a library error code.
int to_device(const void* host, size_t bytes);

// A kernel wrote the device copy of that attached array: its granules become
// STALE and its end pages are downloaded.  Call after the kernels finished.
int written(const void* host, size_t bytes);

// True when no granule of that attached array is STALE (the host copy is
// current everywhere), or when the array is not attached.
bool host_current(const void* host, size_t bytes);

// Make every attached byte in [p, p + bytes) current on the host and
// accessible (writable when write is set), for code that hands host memory
// to a copy engine rather than touching it from the CPU.  0 or an error code.
int host_access(const void* p, size_t bytes, bool write);

}  // namespace mirror
}  // namespace tpls
