// vmm_export.cpp -- test helper for the graphics-interop entry points
// (ocn_external_import).  No Vulkan or OpenGL stack is available to the
// tests, so the exported memory a renderer would hand over is made with HIP's
// virtual memory API instead: a device allocation exported as a POSIX file
// descriptor (a dma-buf, the kind of descriptor AMD's Linux graphics drivers
// export for opaque-FD memory), plus a second mapping of the same physical
// memory through which the test reads what was written through the import.
// Host code only; built by tests/cpp/Makefile.
#include <hip/hip_runtime.h>

#include <cstddef>

extern "C" {

// Allocates at least `bytes` on `device`, exports it, and maps it at *va
// (read-write for that device).  Returns 0 or the failing hipError_t; on
// failure nothing is left allocated.
int vmm_alloc_export(int device, size_t bytes, int* fd, size_t* size, void** va, void** handle)
{
    hipMemAllocationProp prop{};
    prop.type                = hipMemAllocationTypePinned;
    prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
    prop.location.type       = hipMemLocationTypeDevice;
    prop.location.id         = device;
    size_t     gran          = 0;
    hipError_t e             = hipSetDevice(device);
    if (e == hipSuccess) e = hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum);
    if (e != hipSuccess) return (int)e;
    if (gran == 0) gran = 4096;
    const size_t sz = (bytes + gran - 1) / gran * gran;

    hipMemGenericAllocationHandle_t h = nullptr;
    if ((e = hipMemCreate(&h, sz, &prop, 0)) != hipSuccess) return (int)e;
    int f = -1;
    if ((e = hipMemExportToShareableHandle(&f, h, hipMemHandleTypePosixFileDescriptor, 0)) != hipSuccess) {
        (void)hipMemRelease(h);
        return (int)e;
    }
    void* p = nullptr;
    if ((e = hipMemAddressReserve(&p, sz, 0, nullptr, 0)) != hipSuccess) {
        (void)hipMemRelease(h);
        return (int)e;
    }
    if ((e = hipMemMap(p, sz, 0, h, 0)) != hipSuccess) {
        (void)hipMemAddressFree(p, sz);
        (void)hipMemRelease(h);
        return (int)e;
    }
    hipMemAccessDesc acc{};
    acc.location = prop.location;
    acc.flags    = hipMemAccessFlagsProtReadWrite;
    if ((e = hipMemSetAccess(p, sz, &acc, 1)) != hipSuccess) {
        (void)hipMemUnmap(p, sz);
        (void)hipMemAddressFree(p, sz);
        (void)hipMemRelease(h);
        return (int)e;
    }
    *fd     = f;
    *size   = sz;
    *va     = p;
    *handle = h;
    return 0;
}

// Unmaps and releases what vmm_alloc_export made (the exported fd is the
// caller's to close).
int vmm_free(void* va, size_t size, void* handle)
{
    hipError_t e  = hipMemUnmap(va, size);
    hipError_t e2 = hipMemAddressFree(va, size);
    hipError_t e3 = hipMemRelease(static_cast<hipMemGenericAllocationHandle_t>(handle));
    return e != hipSuccess ? (int)e : e2 != hipSuccess ? (int)e2 : (int)e3;
}

}  // extern "C"
