// kt_count_noalloc: kt_count on a workspace that is never allocated.
//
// The plain path's device workspace (degrees, keep bits, oriented CSR, work
// queues, scan and sort scratch) lives in a file-scope device array instead
// of the context's grow-only arena, so a call allocates nothing: no
// hipMalloc, no hipFree, no allocating library routine.  The array exists
// once per GPU the library runs on (every device holds its own copy at its
// own address) and is held until the process ends.
//
// The array's name is a host-side stand-in in host code, not a device
// address: its device address is looked up once per device with
// hipGetSymbolAddress (at context creation, with the device current) and
// only that address is handed to kernels.
#include <mutex>

#include "../../include/katric.h"
#include "kt_ctx.h"

namespace kt {

constexpr uint64_t NOALLOC_BYTES = KT_NOALLOC_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;

// the workspace (per GPU); every kernel of a call is handed scratch_ptr[d]
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
std::mutex table_mu;
void* scratch_ptr[MAX_DEVICES] = {};
// calls on one device share its array: they are serialised here (calls from
// other processes use their own copies)
std::mutex device_mu[MAX_DEVICES];
}  // namespace

// Looks up this device's copy of the array (the device must be current).
hipError_t noalloc_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(table_mu);
    if (scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    scratch_ptr[device] = p;
    return hipSuccess;
}

}  // namespace kt

using namespace kt;

extern "C" int kt_count_noalloc(kt_ctx* c, uint32_t n, const uint64_t* offsets, const uint32_t* heads,
                                uint64_t* total_out, uint64_t* per_vertex_out) {
    if (!c) return KT_EINVAL;
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = c->begin_call(true);
    if (rc) return rc;
    if (c->device < 0 || c->device >= MAX_DEVICES) return c->fail(KT_ENODEV, "device index beyond the table");
    // set up at kt_create; else on the first call on this device (current now)
    rc = c->hip_check(noalloc_init(c->device), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    void* base;
    {
        std::lock_guard<std::mutex> tl(table_mu);
        base = scratch_ptr[c->device];
    }
    std::lock_guard<std::mutex> dl(device_mu[c->device]);
    // device input and output; a graph whose workspace exceeds the array is
    // KT_ENOMEM with the sizes in kt_last_error
    return count_locked(c, n, offsets, heads, KT_DEVICE_INPUT | KT_DEVICE_OUTPUT, total_out, per_vertex_out, base,
                        NOALLOC_BYTES);
}
