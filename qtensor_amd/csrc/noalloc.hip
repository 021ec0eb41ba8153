// noalloc.hip -- qb_execute_noalloc: the batched executor for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  The caller passes the inputs
// (program, items, angles) and the output (item_out); the work-queue ticket and
// the items' global scratch slots live in a static device array instead of
// caller or library allocations.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>

#include "../../include/qtensor_amd.h"
#include "internal.h"

// Workspace of qb_execute_noalloc: the queue ticket in the first
// kNoallocQueueBytes bytes, then the scratch slots.  Every device holds its own
// copy until the process ends; every call on a device shares it and it keeps the
// last call's contents, so calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[QB_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
constexpr int64_t kNoallocQueueBytes = 256;

std::mutex g_noalloc_mutex;
// device address of `scratch` on each device, from hipGetSymbolAddress with that
// device current (the array's name is not a device address in host code)
unsigned char* g_noalloc_base[kMaxDevices] = {};

int fail_fmt(int code, const char* fmt, long long a, long long b, long long c, long long d) {
    char buf[320];
    snprintf(buf, sizeof(buf), fmt, a, b, c, d);
    return qb_internal::fail(code, buf);
}

}  // namespace

namespace qb_internal {

// Called by qb_init with `device` current: resolve this device's copy once.
int noalloc_init(int device) {
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (device < 0 || device >= kMaxDevices) return fail(QB_ERR_UNSUPPORTED, "device index beyond the static scratch table");
    if (g_noalloc_base[device]) return QB_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(QB_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
    g_noalloc_base[device] = static_cast<unsigned char*>(ptr);
    return QB_OK;
}

}  // namespace qb_internal

extern "C" int qb_execute_noalloc(const qb_exec_desc* d, void* stream) {
    const int dev = qb_internal::current_device();
    if (dev < 0) return qb_internal::fail(QB_ERR_NOT_INIT, "qb_init was not called");
    if (!d) return qb_internal::fail(QB_ERR_INVALID, "null descriptor");
    if (d->scratch || d->queue || d->n_slots)
        return qb_internal::fail(QB_ERR_INVALID, "qb_execute_noalloc takes no scratch, n_slots or queue (it uses its own)");
    if (d->dtype != QB_C64 && d->dtype != QB_C128) return qb_internal::fail(QB_ERR_INVALID, "dtype must be QB_C64 or QB_C128");
    if (d->scratch_elems < 0) return qb_internal::fail(QB_ERR_INVALID, "scratch_elems < 0");
    unsigned char* base = dev < kMaxDevices ? g_noalloc_base[dev] : nullptr;
    if (!base) return qb_internal::fail(QB_ERR_NOT_INIT, "static scratch of this device was not resolved");
    int32_t n_slots = 0;
    if (d->scratch_elems > 0) {
        // the launch keeps at most this many items resident, one slot each
        const int rc = qb_internal::resident_items(d, &n_slots);
        if (rc) return rc;
        if (d->n_items < n_slots) n_slots = d->n_items > 0 ? d->n_items : 1;
        const int64_t elem = d->dtype == QB_C128 ? 16 : 8;
        const int64_t cap = (int64_t)QB_NOALLOC_SCRATCH_BYTES - kNoallocQueueBytes;
        if ((int64_t)n_slots * d->scratch_elems > cap / elem)
            return fail_fmt(QB_ERR_UNSUPPORTED,
                            "qb_execute_noalloc: %lld resident items x %lld scratch elements x %lld bytes exceed "
                            "the static scratch of %lld bytes",
                            n_slots, (long long)d->scratch_elems, (long long)elem, (long long)cap);
    }
    return qb_internal::execute_impl(d, d->scratch_elems > 0 ? base + kNoallocQueueBytes : nullptr, n_slots,
                                     reinterpret_cast<uint32_t*>(base), stream);
}
